// Errors of the host layer.  The reference throws std::invalid_argument from
// cudaCheck (CudaCommon.h:10-18) and StrException from the memory manager
// (MemoryManager.cuh:74-80); both are std::exception subclasses here too.
#pragma once

#include <hip/hip_runtime_api.h>

#include <stdexcept>
#include <string>

struct StrException : public std::runtime_error {
  explicit StrException(const std::string& s) : std::runtime_error(s) {}
};

inline void hip_check(hipError_t err, const char* what)
{
  if (err != hipSuccess) throw std::invalid_argument(std::string(what) + ": " + hipGetErrorString(err));
}

#define hipCheck(stmt) hip_check((stmt), #stmt)
