// Scheduler: decides at compile time, for every step of a sequence, which
// arguments are reserved (first use) and which are freed (the step after
// their last use, unless they are outputs), and applies that at run time
// (reference interface: Scheduler.cuh:32-93, SchedulerMachinery.cuh:59-140).
#pragma once

#include <cstdio>
#include <tuple>
#include <type_traits>
#include <utility>

#include "ArgumentManager.h"
#include "MemoryManager.h"

namespace Sch {

  template<typename T, typename Tuple>
  struct Contains;
  template<typename T, typename... Ts>
  struct Contains<T, std::tuple<Ts...>> : std::bool_constant<(std::is_same_v<T, Ts> || ...)> {
  };

  // Tuple of the distinct types of Tuple, first occurrences kept in order.
  template<typename Done, typename Rest>
  struct Unique;
  template<typename... D>
  struct Unique<std::tuple<D...>, std::tuple<>> {
    using t = std::tuple<D...>;
  };
  template<typename... D, typename H, typename... R>
  struct Unique<std::tuple<D...>, std::tuple<H, R...>> {
    using t = typename std::conditional_t<
      Contains<H, std::tuple<D...>>::value,
      Unique<std::tuple<D...>, std::tuple<R...>>,
      Unique<std::tuple<D..., H>, std::tuple<R...>>>::t;
  };

  template<typename Sequence>
  struct AllArguments;
  template<typename... Algorithms>
  struct AllArguments<std::tuple<Algorithms...>> {
    using t = typename Unique<
      std::tuple<>,
      decltype(std::tuple_cat(std::declval<typename Algorithms::Arguments>()...))>::t;
  };

  // First and last step (index into Sequence) that use Arg; -1 if unused.
  template<typename Arg, typename Sequence, size_t... Is>
  constexpr int first_use_impl(std::index_sequence<Is...>)
  {
    constexpr bool uses[] = {Contains<Arg, typename std::tuple_element_t<Is, Sequence>::Arguments>::value...};
    for (size_t i = 0; i < sizeof...(Is); ++i)
      if (uses[i]) return (int) i;
    return -1;
  }
  template<typename Arg, typename Sequence, size_t... Is>
  constexpr int last_use_impl(std::index_sequence<Is...>)
  {
    constexpr bool uses[] = {Contains<Arg, typename std::tuple_element_t<Is, Sequence>::Arguments>::value...};
    int r = -1;
    for (size_t i = 0; i < sizeof...(Is); ++i)
      if (uses[i]) r = (int) i;
    return r;
  }
  template<typename Arg, typename Sequence>
  constexpr int first_use = first_use_impl<Arg, Sequence>(std::make_index_sequence<std::tuple_size_v<Sequence>> {});
  template<typename Arg, typename Sequence>
  constexpr int last_use = last_use_impl<Arg, Sequence>(std::make_index_sequence<std::tuple_size_v<Sequence>> {});

} // namespace Sch

template<typename ConfiguredSequence, typename OutputArguments>
struct Scheduler {
  using arguments_tuple_t = typename Sch::AllArguments<ConfiguredSequence>::t;
  using argument_manager_t = ArgumentManager<arguments_tuple_t>;

  MemoryManager memory_manager;
  argument_manager_t argument_manager;
  ConfiguredSequence sequence_tuple;
  bool do_print = false;

  Scheduler() = default;
  Scheduler(const Scheduler&) = delete;

  void initialize(bool print, size_t reserved_bytes, char* base_pointer)
  {
    do_print = print;
    memory_manager.set_reserved_memory(reserved_bytes);
    argument_manager.set_base_pointer(base_pointer);
  }

  void reset() { memory_manager.free_all(); }

  template<typename T>
  typename T::arguments_t arguments()
  {
    return produce<T>(static_cast<typename T::Arguments*>(nullptr));
  }

  // Frees what step I no longer needs and reserves what it uses first.
  template<size_t I, typename T>
  void setup()
  {
    static_assert(std::is_same_v<std::tuple_element_t<I, ConfiguredSequence>, T>, "Scheduler index mismatch");
    free_step<I>(std::make_index_sequence<std::tuple_size_v<arguments_tuple_t>> {});
    reserve_step<I>(typename T::Arguments {});
    if (do_print) {
      std::printf("Sequence step %zu \"%s\":\n", I, T::name);
      memory_manager.print();
    }
  }

private:
  template<typename T, typename... Args>
  typename T::arguments_t produce(std::tuple<Args...>*)
  {
    return typename T::arguments_t {std::forward_as_tuple(std::get<Args>(argument_manager.arguments_tuple)...)};
  }

  template<size_t I, size_t... Js>
  void free_step(std::index_sequence<Js...>)
  {
    (free_one<I, std::tuple_element_t<Js, arguments_tuple_t>>(), ...);
  }

  template<size_t I, typename Arg>
  void free_one()
  {
    if constexpr (I > 0 && Sch::last_use<Arg, ConfiguredSequence> == (int) I - 1 &&
                  !Sch::Contains<Arg, OutputArguments>::value) {
      auto& a = std::get<Arg>(argument_manager.arguments_tuple);
      if (!a.external) memory_manager.free(Arg::name);
    }
  }

  template<size_t I, typename... Args>
  void reserve_step(std::tuple<Args...>)
  {
    (reserve_one<I, Args>(), ...);
  }

  template<size_t I, typename Arg>
  void reserve_one()
  {
    if constexpr (Sch::first_use<Arg, ConfiguredSequence> == (int) I) {
      auto& a = std::get<Arg>(argument_manager.arguments_tuple);
      if (a.external) return;
      const size_t off = memory_manager.reserve(Arg::name, a.size);
      argument_manager.template set_offset<Arg>(off);
    }
  }
};

namespace Sch {

  // Runs every algorithm of the sequence: set_arguments_size, setup, visit
  // (reference: SchedulerMachinery.cuh:274-393).  `hook(i, name)` is called
  // before step i and once more with i = number of steps at the end.
  struct NoHook {
    void operator()(size_t, const char*) const {}
  };

  template<typename Scheduler, typename Visitor, typename Hook, size_t... Is, typename... SizeArgs, typename... VisitArgs>
  void run_sequence_impl(
    Scheduler& scheduler, Visitor& visitor, Hook& hook, std::index_sequence<Is...>, std::tuple<SizeArgs...> size_args,
    std::tuple<VisitArgs...> visit_args)
  {
    using seq_t = std::remove_reference_t<decltype(scheduler.sequence_tuple)>;
    auto step = [&](auto idx) {
      constexpr size_t I = decltype(idx)::value;
      using T = std::tuple_element_t<I, seq_t>;
      std::apply(
        [&](auto&&... a) { visitor.template set_arguments_size<T>(scheduler.template arguments<T>(), a...); },
        size_args);
      scheduler.template setup<I, T>();
      hook(I, T::name);
      std::apply(
        [&](auto&&... a) {
          visitor.template visit<T>(std::get<I>(scheduler.sequence_tuple), scheduler.template arguments<T>(), a...);
        },
        visit_args);
    };
    (step(std::integral_constant<size_t, Is> {}), ...);
    hook(sizeof...(Is), "end");
  }

  // Lays the sequence out without running it: set_arguments_size and
  // setup for every step (no visit, nothing launched), so the memory
  // manager's high-water mark is the arena the run needs.
  template<typename Scheduler, typename Visitor, size_t... Is, typename... SizeArgs>
  void size_sequence_impl(
    Scheduler& scheduler, Visitor& visitor, std::index_sequence<Is...>, std::tuple<SizeArgs...> size_args)
  {
    using seq_t = std::remove_reference_t<decltype(scheduler.sequence_tuple)>;
    auto step = [&](auto idx) {
      constexpr size_t I = decltype(idx)::value;
      using T = std::tuple_element_t<I, seq_t>;
      std::apply(
        [&](auto&&... a) { visitor.template set_arguments_size<T>(scheduler.template arguments<T>(), a...); },
        size_args);
      scheduler.template setup<I, T>();
    };
    (step(std::integral_constant<size_t, Is> {}), ...);
  }

  template<typename Scheduler, typename Visitor, typename... SizeArgs>
  void size_sequence(Scheduler& scheduler, Visitor& visitor, std::tuple<SizeArgs...> size_args)
  {
    using seq_t = std::remove_reference_t<decltype(scheduler.sequence_tuple)>;
    size_sequence_impl(scheduler, visitor, std::make_index_sequence<std::tuple_size_v<seq_t>> {}, size_args);
  }

  // Calls visitor.check<T>(args...) for every step T of the sequence, in
  // order (reference: SchedulerMachinery.cuh:396-420, RunChecker).
  template<typename Sequence, typename Visitor, size_t... Is, typename... Args>
  void check_sequence_impl(const Visitor& visitor, std::index_sequence<Is...>, Args&&... args)
  {
    (visitor.template check<std::tuple_element_t<Is, Sequence>>(args...), ...);
  }

  template<typename Sequence, typename Visitor, typename... Args>
  void check_sequence(const Visitor& visitor, Args&&... args)
  {
    check_sequence_impl<Sequence>(visitor, std::make_index_sequence<std::tuple_size_v<Sequence>> {}, args...);
  }

  template<typename Scheduler, typename Visitor, typename Hook, typename... SizeArgs, typename... VisitArgs>
  void run_sequence(
    Scheduler& scheduler, Visitor& visitor, Hook& hook, std::tuple<SizeArgs...> size_args,
    std::tuple<VisitArgs...> visit_args)
  {
    using seq_t = std::remove_reference_t<decltype(scheduler.sequence_tuple)>;
    run_sequence_impl(
      scheduler, visitor, hook, std::make_index_sequence<std::tuple_size_v<seq_t>> {}, size_args, visit_args);
  }

} // namespace Sch
