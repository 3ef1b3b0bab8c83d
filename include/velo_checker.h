/* VELO track checker: efficiency per MC category, clones and ghosts.
 *
 * Host-only, no GPU.  Replaces the VELO part of the reference's checker:
 *   prepareVeloTracks        checker/tracking/src/PrepareTracks.cpp:14-47
 *   MCAssociator             checker/tracking/src/MCAssociator.cpp:13-60,
 *                            checker/tracking/include/MCAssociator.h:166-182
 *   TrackChecker::operator() checker/tracking/src/TrackChecker.cpp:349-405
 *   TrackEffReport           checker/tracking/src/TrackChecker.cpp:93-146
 *   TrackCheckerVelo categories
 *                            checker/tracking/src/CategoriesTrackChecker.cpp:5-230
 *
 * Semantics kept from the reference:
 *  - each LHCbID of a track is looked up in the event's (LHCbID -> MC
 *    particle) map; for an LHCbID shared by several particles the particle
 *    listed first wins (the reference sorts with std::sort, which leaves the
 *    choice among equal keys unspecified);
 *  - weight(p) = hits matched to p / hits matched to any particle; the best
 *    weight decides (ties: lowest particle index); a track with no match or a
 *    best weight below 0.7 is a ghost;
 *  - a matched track counts as found for every category its particle is in,
 *    the second and later tracks of one particle as clones;
 *  - clone rate as printed by the reference = clones / (2 * found)
 *    (TrackChecker.cpp:130, kept as is); hit efficiency and purity are the
 *    running means of TrackChecker.cpp:110-117.
 */
#ifndef VELO_CHECKER_H
#define VELO_CHECKER_H

#include <stddef.h>
#include <stdint.h>

#include "velo_gen.h"

#ifdef __cplusplus
extern "C" {
#endif

#define VELO_CHECKER_N_CATEGORIES 40
#define VELO_CHECKER_MIN_WEIGHT 0.7f

/* MC particle flags beyond velo_gen.h's (never set by the generator, which
 * has no decay chains; categories that need them accept no particle) */
#define VELO_MCP_FROM_B 0x10u
#define VELO_MCP_FROM_D 0x20u
#define VELO_MCP_FROM_S 0x40u

typedef struct {
  uint64_t accepted; /* reconstructible particles in the category */
  uint64_t found;    /* of which matched by at least one track */
  uint64_t clones;   /* further tracks matched to a found particle */
  double eff_per_event;
  double hit_eff;
  double hit_purity;
} velo_checker_category;

typedef struct {
  uint64_t events;
  uint64_t tracks;
  uint64_t ghosts;
  double ghost_per_event;
  velo_checker_category categories[VELO_CHECKER_N_CATEGORIES];
} velo_checker_result;

/* MC truth of the reference's checker input (the MC_info/tracks folder). */
typedef struct {
  velo_gen_mcp* mcps; /* particles with VELO hits, in event order; malloc'ed */
  size_t n_mcps;
  uint32_t* mc_hits;  /* their VELO LHCbIDs (mcps[i].hit_offset, .n_hits); malloc'ed */
  size_t n_mc_hits;
  uint32_t n_events;
} velo_mc_truth;

/* Reads n_events (0: all) per-event MC files of a folder from start_event
 * on, in the folder's natural order (CheckerInvoker.cpp:3-44), each file
 * parsed as MCEvent.cpp:16-103: u32 particle count, then per particle key,
 * pid, p, pt, eta, phi, eight flag bytes (isLong, isDown, hasVelo, hasUT,
 * hasSciFi, fromB, fromD, fromS), nPV, and the VELO, UT and SciFi LHCbID
 * lists (u32 count + ids each).  Only particles with VELO hits are kept,
 * with their VELO hits; flags map to VELO_MCP_* (electron: |pid| = 11,
 * backward: eta < 0).  A file whose size differs from what it declares, or
 * a request past the last file, is an error (-1, message in err). */
int velo_mc_read_folder(const char* folder, uint32_t n_events, uint32_t start_event, velo_mc_truth* out, char* err,
                        size_t err_len);
void velo_mc_free(velo_mc_truth* t);

/* Name of category i (the reference's report label), or NULL. */
const char* velo_checker_category_name(uint32_t i);

/* Check the consolidated tracks of n_events events.
 *   event_track_offsets  n_events+1 (the scanned track counts, dev_atomics_velo + n_events)
 *   track_hit_offsets    T+1 (the scanned hit counts)
 *   hit_ids              LHCbIDs of the consolidated hits
 *   accepted             per track, 0 = dropped from the check (NULL: all kept)
 *   mcps, n_mcps         MC particles, non-decreasing in .event
 *   mc_hits              LHCbIDs indexed by mcps[i].hit_offset
 * Optional per-track outputs (NULL to skip): track_mcp = index into mcps of
 * the best particle or -1, track_weight = its weight (0 when unmatched).
 * Returns 0, or -1 on a null pointer or unsorted mcps. */
int velo_checker_run(uint32_t n_events, const uint32_t* event_track_offsets, const uint32_t* track_hit_offsets,
                     const uint32_t* hit_ids, const uint8_t* accepted, const velo_gen_mcp* mcps, size_t n_mcps,
                     const uint32_t* mc_hits, velo_checker_result* result, int32_t* track_mcp,
                     float* track_weight);

/* ---- clustering checker ----
 * Replaces checker/clustering/src/ClusteringChecker.cpp:32-135 (checkClustering)
 * and the x86 clustering it takes as the expected clusters
 * (x86/velo/clustering/src/Clustering.cpp:3-270, `clustering`). */
typedef struct {
  uint64_t reconstructible; /* expected clusters */
  uint64_t reconstructed;   /* found clusters */
  uint64_t correct;         /* found clusters that match an expected one */
  uint64_t clones;          /* per distinct found cluster, its further matches among the found */
  uint64_t ghosts;          /* found clusters that match no expected one */
  float efficiency;         /* correct / reconstructible */
  float clone_fraction;     /* clones / reconstructed */
  float ghost_fraction;     /* ghosts / reconstructed */
} velo_clustering_check_result;

/* The x86 clustering of n_events raw VP events (offsets: n_events + 1 byte
 * offsets into events): per bank, isolated SPs through the SP pattern
 * table (one or two clusters), the pixels of the other SPs through an
 * 8-connected row-scan flood fill seeded in SP, pixel order; each cluster's
 * LHCb ID is its centroid pixel (integer quotients of the pixel sums).  IDs
 * per event in bank order, event e's at [id_offsets[e], id_offsets[e+1]).
 * Pixels outside the sensor's 768 columns are skipped (the reference
 * writes them past its pixel buffer).  Returns 0; 1 when more than
 * `capacity` IDs exist (*n_ids says how many, ids holds the first
 * capacity); -1 on a null pointer or a bank whose sensor is out of range. */
int velo_x86_clustering(const char* events, const uint32_t* offsets, uint32_t n_events, uint32_t* ids,
                        uint32_t* id_offsets, size_t capacity, size_t* n_ids);

/* The reference's clustering report: found vs expected cluster LHCb IDs of
 * n_events events (per event at [offsets[e], offsets[e+1])).  A found
 * cluster is correct when an expected one has its ID or, unless
 * just_check_ids, lies on the same sensor and chip within
 * allowed_distance_error pixels (row and column of the ID's low 16 bits);
 * otherwise it is a ghost.  Returns 0, or -1 on a null pointer. */
int velo_clustering_check(uint32_t n_events, const uint32_t* found_offsets, const uint32_t* found_ids,
                          const uint32_t* expected_offsets, const uint32_t* expected_ids, int just_check_ids,
                          float allowed_distance_error, velo_clustering_check_result* result);

#ifdef __cplusplus
}
#endif

#endif
