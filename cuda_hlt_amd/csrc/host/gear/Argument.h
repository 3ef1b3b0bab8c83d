// ARGUMENT / ARGUMENTS: named, typed device buffers of a sequence
// (reference interface: stream/gear/include/Argument.cuh:14-20).
//
// An argument is reserved in the stream's arena at its first use and freed
// after its last use.  `external` marks an argument bound by the caller to
// memory it owns (e.g. raw events already resident in HBM); the memory
// manager then neither reserves nor frees it.
#pragma once

#include <cstddef>
#include <tuple>

#define ARGUMENTS(...) std::tuple<__VA_ARGS__>

#define ARGUMENT(ARGUMENT_NAME, ARGUMENT_TYPE)           \
  struct ARGUMENT_NAME {                                 \
    static constexpr const char* name = #ARGUMENT_NAME; \
    using type = ARGUMENT_TYPE;                          \
    size_t size = 0;                                     \
    char* offset = nullptr;                              \
    bool external = false;                               \
  };
