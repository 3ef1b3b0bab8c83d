// Host-side views of the VELO event model (VeloEventModel.cuh:9-111,
// States.cuh:17, VeloConsolidated.cuh:12-144).
#pragma once

#include <cstdint>

#include "velo_hlt.h"

namespace Velo {
#pragma pack(push, 1)
  struct TrackHits { // 54 B
    uint16_t hitsNum;
    uint16_t hits[VELO_MAX_TRACK_SIZE];
  };
  struct TrackletHits { // 6 B
    uint16_t hits[3];
  };
#pragma pack(pop)
  struct Hit { // 16 B
    float x, y, z;
    uint32_t LHCbID;
  };
  static_assert(sizeof(TrackHits) == VELO_TRACK_HITS_BYTES, "TrackHits layout");
  static_assert(sizeof(TrackletHits) == VELO_TRACKLET_BYTES, "TrackletHits layout");
} // namespace Velo

struct VeloState { // 48 B + bool in the SoA
  float x, y, tx, ty;
  float c00, c20, c22, c11, c31, c33;
  float chi2;
  float z;
  bool backward;
};
