// Library information entry points of include/velo_hlt.h.
#include <hip/hip_runtime_api.h>

#include <cstring>

#include "velo_checker.h"
#include "velo_ghost.h"
#include "velo_hlt.h"
#include "velo_mdf.h"
#include "velo_pv.h"
#include "velo_sequence.h"

extern "C" {

const char* velo_hlt_version(void) { return "velo-hlt gfx950 1.1"; }

const char* velo_error_string(int code) { return hipGetErrorString(static_cast<hipError_t>(code)); }

size_t velo_abi_struct_size(const char* name)
{
  if (!name) return 0;
  struct Entry {
    const char* name;
    size_t size;
  };
#define VELO_ABI_ENTRY(T) {#T, sizeof(T)}
  static const Entry table[] = {
    VELO_ABI_ENTRY(velo_tuning),
    VELO_ABI_ENTRY(velo_constants),
    VELO_ABI_ENTRY(velo_init_event_list_args),
    VELO_ABI_ENTRY(velo_event_order_args),
    VELO_ABI_ENTRY(velo_global_event_cut_args),
    VELO_ABI_ENTRY(velo_estimate_input_size_args),
    VELO_ABI_ENTRY(velo_prefix_sum_args),
    VELO_ABI_ENTRY(velo_masked_clustering_args),
    VELO_ABI_ENTRY(velo_calculate_phi_and_sort_args),
    VELO_ABI_ENTRY(velo_fill_candidates_args),
    VELO_ABI_ENTRY(velo_search_by_triplet_args),
    VELO_ABI_ENTRY(velo_weak_tracks_adder_args),
    VELO_ABI_ENTRY(velo_copy_and_prefix_sum_single_block_args),
    VELO_ABI_ENTRY(velo_copy_velo_track_hit_number_args),
    VELO_ABI_ENTRY(velo_kalman_fit_args),
    VELO_ABI_ENTRY(velo_consolidate_velo_tracks_args),
    VELO_ABI_ENTRY(velo_pv_track),
    VELO_ABI_ENTRY(velo_pv_vertex),
    VELO_ABI_ENTRY(velo_pv_beamline_extrapolate_args),
    VELO_ABI_ENTRY(velo_pv_beamline_histo_args),
    VELO_ABI_ENTRY(velo_pv_beamline_peak_args),
    VELO_ABI_ENTRY(velo_pv_beamline_multi_fitter_args),
    VELO_ABI_ENTRY(velo_ghost_model),
    VELO_ABI_ENTRY(velo_track_features_args),
    VELO_ABI_ENTRY(velo_ghost_infer_args),
    VELO_ABI_ENTRY(velo_ghost_grad_args),
    VELO_ABI_ENTRY(velo_ghost_adam_args),
    VELO_ABI_ENTRY(velo_checker_category),
    VELO_ABI_ENTRY(velo_checker_result),
    VELO_ABI_ENTRY(velo_mc_truth),
    VELO_ABI_ENTRY(velo_clustering_check_result),
    VELO_ABI_ENTRY(velo_mdf_events),
    VELO_ABI_ENTRY(velo_sequence_config),
    VELO_ABI_ENTRY(velo_sequence_input),
    VELO_ABI_ENTRY(velo_sequence_output),
  };
#undef VELO_ABI_ENTRY
  for (const Entry& e : table)
    if (std::strcmp(e.name, name) == 0) return e.size;
  return 0;
}

} // extern "C"
