// Library information entry points of include/velo_hlt.h.
#include <hip/hip_runtime_api.h>

#include "velo_hlt.h"

extern "C" {

const char* velo_hlt_version(void) { return "velo-hlt gfx950 1.0"; }

const char* velo_error_string(int code) { return hipGetErrorString(static_cast<hipError_t>(code)); }

} // extern "C"
