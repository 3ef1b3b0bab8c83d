// MC truth for the checks of one run and where their results go
// (reference interface: checker/tracking/include/CheckerInvoker.h,
// checker/tracking/src/CheckerInvoker.cpp:3-44).  The truth covers input
// events start_event_offset, start_event_offset + 1, ...; the checks map the
// run's selected events to them through the host event list.
#pragma once

#include <cstddef>
#include <cstdint>

#include "velo_checker.h"

struct CheckerInvoker {
  const velo_gen_mcp* mcps = nullptr; // non-decreasing in .event
  size_t n_mcps = 0;
  const uint32_t* mc_hits = nullptr;
  uint32_t n_mc_events = 0;

  // filled by SequenceVisitor::check<consolidate_velo_tracks_t>
  velo_checker_result* velo_result = nullptr;
  bool velo_checked = false;
};
