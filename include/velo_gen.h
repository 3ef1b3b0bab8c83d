/*
 * velo_gen.h -- seeded synthetic VELO event generator (C ABI).
 *
 * Produces, deterministically from a seed:
 *   - a detector geometry blob in the reference's velo_geometry.bin layout
 *     (ClusteringDefinitions.cu:23-53: float local_x[768], x_pitch[768],
 *     ltg[16*208]; 19456 bytes);
 *   - raw VP events in the reference's SuperPixel bank layout
 *     (ClusteringDefinitions.cu:3-21, cuda/velo/readme.md:9-33): per event
 *     u32 n_banks=208, u32 bank_offset[209], then per bank u32 sensor_index,
 *     u32 sp_count, u32 sp_word[sp_count]; events concatenated with a
 *     u32 offsets[N+1] array of byte offsets;
 *   - MC truth per generated particle (the information the reference reads
 *     from MC_info/tracks, MCEvent.cpp:16-103): LHCbIDs of its clusters and
 *     the flags the VELO checker categories use.
 *
 * Physics model: PVs ~ Poisson(nu) with z ~ N(0,45 mm), x,y ~ N(0,0.04 mm);
 * Poisson(tracks_per_pv) straight tracks per PV, eta ~ U(eta_min, eta_max),
 * phi ~ U(0, 2pi), a fraction mirrored backward; each sensor crossing inside
 * the active radius gives a 1-3 pixel cluster with probability hit_eff;
 * noise pixels are added as Poisson(noise_frac * signal pixels).
 */
#ifndef VELO_GEN_H
#define VELO_GEN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VELO_GEN_GEOMETRY_BYTES 19456u

typedef struct {
  uint64_t seed;
  uint32_t n_events;
  uint32_t n_threads;    /* 0: hardware concurrency */
  double nu;             /* Poisson mean of PVs per event when nu_min < 0 */
  double nu_min, nu_max; /* if nu_min >= 0: per-event nu ~ U(nu_min, nu_max) */
  double tracks_per_pv;
  double eta_min, eta_max;
  double backward_fraction;
  double hit_efficiency;
  double noise_fraction;
} velo_gen_config;

/* Particle flags (MCParticle.h semantics used by the VELO checker) */
#define VELO_MCP_HAS_VELO 0x1u
#define VELO_MCP_IS_LONG 0x2u
#define VELO_MCP_IS_ELECTRON 0x4u
#define VELO_MCP_BACKWARD 0x8u

typedef struct {
  uint32_t event;
  uint32_t key;
  int32_t pid;
  float p, pt, eta, phi;
  uint32_t flags;
  uint32_t n_pv;
  uint32_t hit_offset; /* into the flat LHCbID array */
  uint32_t n_hits;
} velo_gen_mcp;

/* Fill the default configuration: seed 0, nu 7.6, 25 tracks/PV, eta 2..5,
 * 20% backward, 98% efficiency, 0.1% noise. */
void velo_gen_default_config(velo_gen_config* cfg);

/* Geometry blob (VELO_GEN_GEOMETRY_BYTES) and the 52 module z positions. */
void velo_gen_geometry(char* out, size_t size);
void velo_gen_module_zs(float* out52);

/* Generate events.  Returns an opaque handle (NULL on failure). */
void* velo_gen_run(const velo_gen_config* cfg);
uint32_t velo_gen_n_events(const void* h);
size_t velo_gen_events_size(const void* h);
const char* velo_gen_events(const void* h);
const uint32_t* velo_gen_offsets(const void* h); /* n_events+1 */
uint32_t velo_gen_n_mcps(const void* h);
const velo_gen_mcp* velo_gen_mcps(const void* h);
uint32_t velo_gen_n_mc_hits(const void* h);
const uint32_t* velo_gen_mc_hits(const void* h);
void velo_gen_free(void* h);
/* sizeof velo_gen_config / velo_gen_mcp as the library was built (by
 * typedef name; 0 for an unknown name), for bindings to check their
 * mirrors before the library writes into them. */
size_t velo_gen_struct_size(const char* name);

#ifdef __cplusplus
}
#endif
#endif
