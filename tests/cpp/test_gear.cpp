// Unit tests of the gear layer (MemoryManager, Scheduler) -- host only.
#include <cstdio>
#include <cstdlib>
#include <string>

#include "gear/Argument.h"
#include "gear/ArgumentManager.h"
#include "gear/CpuHandler.h"
#include "gear/Scheduler.h"

static int failures = 0;
#define CHECK(c)                                                     \
  do {                                                               \
    if (!(c)) {                                                      \
      std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);       \
      ++failures;                                                    \
    }                                                                \
  } while (0)

ARGUMENT(a_in, char)
ARGUMENT(b_tmp, unsigned)
ARGUMENT(c_out, float)
ARGUMENT(d_tmp, short)

struct alg0 {
  static constexpr const char* name = "alg0";
  using Arguments = ARGUMENTS(a_in, b_tmp);
  using arguments_t = ArgumentRefManager<Arguments>;
};
struct alg1 {
  static constexpr const char* name = "alg1";
  using Arguments = ARGUMENTS(b_tmp, c_out);
  using arguments_t = ArgumentRefManager<Arguments>;
};
struct alg2 {
  static constexpr const char* name = "alg2";
  using Arguments = ARGUMENTS(c_out, d_tmp);
  using arguments_t = ArgumentRefManager<Arguments>;
};
struct alg3 {
  static constexpr const char* name = "alg3";
  using Arguments = ARGUMENTS(d_tmp);
  using arguments_t = ArgumentRefManager<Arguments>;
};
using seq_t = std::tuple<alg0, alg1, alg2, alg3>;
using out_t = std::tuple<c_out>;

static_assert(Sch::first_use<a_in, seq_t> == 0 && Sch::last_use<a_in, seq_t> == 0, "a_in");
static_assert(Sch::first_use<b_tmp, seq_t> == 0 && Sch::last_use<b_tmp, seq_t> == 1, "b_tmp");
static_assert(Sch::first_use<c_out, seq_t> == 1 && Sch::last_use<c_out, seq_t> == 2, "c_out");
static_assert(Sch::first_use<d_tmp, seq_t> == 2 && Sch::last_use<d_tmp, seq_t> == 3, "d_tmp");
static_assert(std::tuple_size_v<Scheduler<seq_t, out_t>::arguments_tuple_t> == 4, "distinct arguments");

void test_memory_manager()
{
  MemoryManager m;
  m.set_reserved_memory(4096);
  CHECK(m.reserve("x", 1) == 0);    // rounded to 256
  CHECK(m.reserve("y", 0) == 256);  // zero -> 8 B -> 256
  CHECK(m.reserve("z", 300) == 512); // 512 B
  m.free("y");
  CHECK(m.reserve("w", 200) == 256); // first fit reuses the hole
  m.free("x");
  m.free("w"); // coalesces with x's hole
  CHECK(m.reserve("v", 512) == 0);
  CHECK(m.total_memory_required == 1024);
  bool threw = false;
  try {
    m.reserve("big", 1 << 20);
  }
  catch (const StrException&) {
    threw = true;
  }
  CHECK(threw);
  threw = false;
  try {
    m.free("nope");
  }
  catch (const StrException&) {
    threw = true;
  }
  CHECK(threw);
  m.free_all();
  CHECK(m.memory_segments.size() == 1 && m.memory_segments.front().size == 4096);
}

void test_scheduler()
{
  static char arena[1 << 16];
  Scheduler<seq_t, out_t> s;
  s.initialize(false, sizeof(arena), arena);
  auto& t = s.argument_manager.arguments_tuple;
  std::get<a_in>(t).size = 1000;
  std::get<b_tmp>(t).size = 100;
  std::get<c_out>(t).size = 4000;
  std::get<d_tmp>(t).size = 10;
  s.setup<0, alg0>();
  CHECK(s.memory_manager.is_reserved("a_in") && s.memory_manager.is_reserved("b_tmp"));
  CHECK(!s.memory_manager.is_reserved("c_out"));
  s.setup<1, alg1>();
  CHECK(!s.memory_manager.is_reserved("a_in")); // last used by alg0
  CHECK(s.memory_manager.is_reserved("b_tmp") && s.memory_manager.is_reserved("c_out"));
  s.setup<2, alg2>();
  CHECK(!s.memory_manager.is_reserved("b_tmp"));
  CHECK(s.memory_manager.is_reserved("d_tmp"));
  s.setup<3, alg3>();
  CHECK(s.memory_manager.is_reserved("c_out")); // an output: never freed
  CHECK(s.argument_manager.offset<c_out>() >= (float*) arena);
  // external arguments are neither reserved nor freed
  s.reset();
  char external[16];
  auto refs = s.arguments<alg0>();
  refs.set_external<a_in>(external, sizeof(external));
  s.setup<0, alg0>();
  CHECK(!s.memory_manager.is_reserved("a_in"));
  CHECK(refs.offset<a_in>() == external);
  s.setup<1, alg1>(); // frees a_in: skipped because external
  CHECK(s.memory_manager.is_reserved("b_tmp"));
}

// CPU_ALGORITHM: a host function as a sequence step between device steps.
static int host_sum(const unsigned* values, unsigned n, unsigned* total)
{
  unsigned s = 0;
  for (unsigned i = 0; i < n; ++i) s += values[i];
  *total = s;
  return (int) n;
}
CPU_ALGORITHM(host_sum, cpu_sum_t, ARGUMENTS(b_tmp, c_out))

using cpu_seq_t = std::tuple<alg0, cpu_sum_t, alg2>;

struct SumVisitor {
  unsigned host_values[4] = {1, 2, 3, 4};
  unsigned total = 0;
  int returned = -1;
  std::string steps;
  template<typename T>
  void set_arguments_size(typename T::arguments_t arguments)
  {
    if constexpr (std::is_same_v<T, alg0>) {
      arguments.template set_size<a_in>(16);
      arguments.template set_size<b_tmp>(4);
    }
    if constexpr (std::is_same_v<T, alg2>) arguments.template set_size<d_tmp>(2);
  }
  template<typename T>
  void visit(T& state, const typename T::arguments_t& arguments)
  {
    steps += T::name;
    steps += ";";
    if constexpr (std::is_same_v<T, cpu_sum_t>) {
      // the step's device arguments are reserved around it like any other's
      CHECK(arguments.template size<b_tmp>() == 4 * sizeof(unsigned));
      returned = state.invoke(host_values, 4u, &total);
    }
  }
};

void test_cpu_algorithm()
{
  static char arena[1 << 16];
  Scheduler<cpu_seq_t, std::tuple<>> s;
  s.initialize(false, sizeof(arena), arena);
  SumVisitor v;
  Sch::NoHook hook;
  Sch::run_sequence(s, v, hook, std::tuple<> {}, std::tuple<> {});
  CHECK(v.steps == "alg0;cpu_sum_t;alg2;");
  CHECK(v.total == 10 && v.returned == 4);
  CHECK(std::string(cpu_sum_t::name) == "cpu_sum_t");
  CHECK(!s.memory_manager.is_reserved("b_tmp")); // freed after its last use (cpu_sum_t)
}

int main()
{
  test_memory_manager();
  test_scheduler();
  test_cpu_algorithm();
  if (failures) return 1;
  std::printf("ok\n");
  return 0;
}
