// Read-only device tables shared by every stream (Constants.cuh,
// Constants.cu:22-53 and :239-250, VELO part).
#pragma once

#include <cstdint>
#include <vector>

#include "../gear/Common.h"
#include "velo_hlt.h"

struct Constants {
  velo_constants dev {};
  std::vector<uint8_t> host_candidate_ks {0, 0, 1, 4, 4, 5, 5, 5, 5};
  uint8_t* dev_velo_candidate_ks = nullptr;
  uint8_t* dev_velo_sp_patterns = nullptr;
  float* dev_velo_sp_fx = nullptr;
  float* dev_velo_sp_fy = nullptr;
  float* dev_velo_module_zs = nullptr;
  char* dev_velo_geometry = nullptr;

  void reserve_and_initialize(const void* geometry, size_t bytes, const float* module_zs)
  {
    hipCheck(static_cast<hipError_t>(velo_constants_init(geometry, bytes, module_zs, &dev)));
    dev_velo_candidate_ks = dev.dev_velo_candidate_ks;
    dev_velo_sp_patterns = dev.dev_velo_sp_patterns;
    dev_velo_sp_fx = dev.dev_velo_sp_fx;
    dev_velo_sp_fy = dev.dev_velo_sp_fy;
    dev_velo_module_zs = dev.dev_velo_module_zs;
    dev_velo_geometry = dev.dev_velo_geometry;
  }

  void release() { velo_constants_free(&dev); }
};
