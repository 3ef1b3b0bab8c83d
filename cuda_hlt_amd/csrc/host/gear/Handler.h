// ALGORITHM: binds a sequence step to its C-ABI launcher
// (reference interface: stream/gear/include/Handler.cuh:12-32).
//
// The reference stores a __global__ function and its argument tuple and
// launches with <<<>>> without checking the result (Handler.cuh:57).  Here
// the handler stores the launcher's POD argument structure: set_opts records
// the grid (the event count) and the workgroup size, set_arguments fills the
// pointer fields in declaration order, and invoke() calls the launcher and
// throws std::invalid_argument on any hipError_t.
#pragma once

#include <hip/hip_runtime_api.h>

#include <string>
#include <type_traits>

#include "Common.h"
#include "velo_hlt.h"

template<typename A, typename = void>
struct has_n_events : std::false_type {
};
template<typename A>
struct has_n_events<A, std::void_t<decltype(std::declval<A&>().n_events)>> : std::true_type {
};

template<typename ArgsT>
struct Handler {
  using launcher_t = int (*)(const ArgsT*, void*);
  launcher_t function;
  ArgsT args {};
  hipStream_t* stream = nullptr;
  const char* name = "";

  explicit Handler(launcher_t f) : function(f) {}

  void set_opts(const dim3& num_blocks, const dim3& num_threads, hipStream_t& s, unsigned /*shared_memory_size*/ = 0)
  {
    stream = &s;
    if constexpr (has_n_events<ArgsT>::value) args.n_events = num_blocks.x;
    args.tuning.block_dim = num_threads.x * num_threads.y * num_threads.z;
  }

  void set_tuning(const velo_tuning& t)
  {
    if (t.block_dim) args.tuning.block_dim = t.block_dim;
    args.tuning.lds_tile = t.lds_tile; // 0: the launcher's default (nothing kept from an earlier run)
    args.tuning.grid_dim = t.grid_dim;
    args.tuning.dev_work_counter = t.dev_work_counter;
    args.tuning.dev_work_order = t.dev_work_order;
    args.tuning.dev_count = t.dev_count;
  }

  template<typename... T>
  void set_arguments(T... params)
  {
    const velo_tuning tuning = args.tuning;
    if constexpr (has_n_events<ArgsT>::value) {
      const uint32_t n = args.n_events;
      args = ArgsT {params...};
      args.n_events = n;
    }
    else {
      args = ArgsT {params...};
    }
    args.tuning = tuning;
  }

  void invoke()
  {
    if (!stream) throw std::invalid_argument(std::string(name) + ": set_opts was not called");
    const int rc = function(&args, static_cast<void*>(*stream));
    if (rc != 0)
      throw std::invalid_argument(std::string(name) + ": " + hipGetErrorString(static_cast<hipError_t>(rc)));
  }
};

template<typename ArgsT>
Handler<ArgsT> make_handler(int (*f)(const ArgsT*, void*))
{
  return Handler<ArgsT> {f};
}

#define ALGORITHM(LAUNCHER, EXPOSED_TYPE_NAME, DEPENDENCIES)                                        \
  struct EXPOSED_TYPE_NAME {                                                                        \
    static constexpr const char* name = #EXPOSED_TYPE_NAME;                                        \
    using Arguments = DEPENDENCIES;                                                                 \
    using arguments_t = ArgumentRefManager<Arguments>;                                              \
    decltype(make_handler(LAUNCHER)) handler {LAUNCHER};                                            \
    void set_opts(const dim3& g, const dim3& b, hipStream_t& s, const unsigned shm = 0)             \
    {                                                                                               \
      handler.name = name;                                                                          \
      handler.set_opts(g, b, s, shm);                                                               \
    }                                                                                               \
    template<typename... T>                                                                         \
    void set_arguments(T... p)                                                                      \
    {                                                                                               \
      handler.set_arguments(p...);                                                                  \
    }                                                                                               \
    void invoke() { handler.invoke(); }                                                             \
  };

// PREFIX_SUM_ALGORITHM: exclusive scan with the total appended
// (reference interface: PrefixSumHandler.cuh:15-56).
#define PREFIX_SUM_ALGORITHM(EXPOSED_TYPE_NAME, DEPENDENCIES)                                       \
  struct EXPOSED_TYPE_NAME {                                                                        \
    static constexpr const char* name = #EXPOSED_TYPE_NAME;                                        \
    using Arguments = DEPENDENCIES;                                                                 \
    using arguments_t = ArgumentRefManager<Arguments>;                                              \
    static size_t aux_array_size(size_t n) { return velo_prefix_sum_aux_size((uint32_t) n); }       \
    size_t array_size = 0;                                                                          \
    size_t auxiliary_array_size = 1;                                                                \
    hipStream_t* stream = nullptr;                                                                  \
    velo_prefix_sum_args args {};                                                                   \
    void set_size(size_t n)                                                                         \
    {                                                                                               \
      array_size = n;                                                                               \
      auxiliary_array_size = aux_array_size(n);                                                     \
    }                                                                                               \
    void set_opts(hipStream_t& s) { stream = &s; }                                                  \
    /* element count read on the device (array_size is then the capacity) */                      \
    void set_device_count(const uint32_t* c) { args.tuning.dev_count = c; }                        \
    template<typename T>                                                                            \
    void set_arguments(T array, T auxiliary_array)                                                  \
    {                                                                                               \
      args.dev_array = reinterpret_cast<uint32_t*>(array);                                          \
      args.dev_auxiliary_array = reinterpret_cast<uint32_t*>(auxiliary_array);                      \
      args.array_size = (uint32_t) array_size;                                                      \
    }                                                                                               \
    void invoke()                                                                                   \
    {                                                                                               \
      const int rc = velo_prefix_sum_launch(&args, stream ? static_cast<void*>(*stream) : nullptr); \
      if (rc != 0)                                                                                  \
        throw std::invalid_argument(std::string(name) + ": " +                                     \
                                    hipGetErrorString(static_cast<hipError_t>(rc)));                \
    }                                                                                               \
  };
