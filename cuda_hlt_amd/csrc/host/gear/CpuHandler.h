// CPU_ALGORITHM: a host function as a step of a sequence
// (reference interface: stream/gear/include/CpuHandler.cuh:13-44).
//
// The step declares its device arguments like any ALGORITHM, so the
// scheduler reserves and frees them around it; its visitor copies what the
// function needs to host memory, synchronises, and calls invoke() with
// host pointers.  invoke() forwards its arguments to the function and
// returns what it returns.
#pragma once

#include <functional>
#include <utility>

#include "ArgumentManager.h"

template<typename R, typename... T>
struct CpuHandler {
  std::function<R(T...)> function;
  explicit CpuHandler(std::function<R(T...)> f) : function(std::move(f)) {}
};

template<typename R, typename... T>
CpuHandler<R, T...> make_cpu_handler(R (*f)(T...))
{
  return CpuHandler<R, T...> {f};
}

#define CPU_ALGORITHM(FUNCTION_NAME, EXPOSED_TYPE_NAME, DEPENDENCIES)      \
  struct EXPOSED_TYPE_NAME {                                               \
    static constexpr const char* name = #EXPOSED_TYPE_NAME;               \
    using Arguments = DEPENDENCIES;                                        \
    using arguments_t = ArgumentRefManager<Arguments>;                     \
    decltype(make_cpu_handler(FUNCTION_NAME)) handler {FUNCTION_NAME};     \
    template<typename... T>                                                \
    auto invoke(T&&... arguments)                                          \
    {                                                                      \
      return handler.function(std::forward<T>(arguments)...);              \
    }                                                                      \
  };
