/* velo_mdf.h -- raw banks from MDF files into the per-type event layout.
 *
 * Replaces, for the VP banks, the reference's MDF input path:
 *   MDF::read_events   mdf/src/read_mdf.cpp:28-189  (bank selection, the
 *                      per-event blob: u32 n_banks, u32 offsets[n_banks+1]
 *                      in bytes, then per bank u32 source id + the bank's
 *                      data words)
 *   MDF::read_banks    mdf/src/read_mdf.cpp:213-337 (record header,
 *                      checksum, decompression)
 *   LHCb::genChecksum  mdf/src/raw_helpers.cpp:186-199 (flag 1: the
 *                      one-at-a-time hash of raw_helpers.cpp:8-21, over the
 *                      record from byte 16 on, bytes taken as signed char)
 *   Compression::unzip mdf/src/compression.cpp:51-130 (ROOT "ZL" blocks:
 *                      9-byte header, zlib stream)
 * The output is what VeloRawEvent reads (cuda/velo/common/src/ClusteringDefinitions.cu:3-11):
 * concatenated events plus u32 byte offsets[n_events + 1].
 */
#ifndef VELO_MDF_H
#define VELO_MDF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* LHCb raw bank types (mdf/include/raw_bank.hpp:57-110) that the reference
 * reads (mdf/include/read_mdf.hpp:20-24). */
#define VELO_MDF_BANK_MUON 13u
#define VELO_MDF_BANK_VP 63u
#define VELO_MDF_BANK_FT_CLUSTER 64u
#define VELO_MDF_BANK_UT 66u

typedef struct {
  char* events;        /* malloc'ed; free with velo_mdf_free */
  size_t events_size;  /* bytes */
  uint32_t* offsets;   /* n_events + 1 byte offsets */
  uint32_t n_events;
  uint32_t n_records;  /* MDF records read (events skipped included) */
} velo_mdf_events;

/* Reads the VP banks of up to max_events events (0: all), after skipping
 * the first start_event records, from the files in order.  With
 * verify_checksum set, a record whose stored checksum (when non-zero)
 * differs from the computed one is an error.  Records are decoded
 * (checksum, inflate, bank walk) on n_threads host threads (0 or 1: the
 * calling thread); the output does not depend on n_threads.  Returns 0, or
 * -1 with a message in err (the earliest failing record's). */
int velo_mdf_read_vp(const char* const* paths, uint32_t n_paths, uint32_t max_events, uint32_t start_event,
                     int verify_checksum, uint32_t n_threads, velo_mdf_events* out, char* err, size_t err_len);

/* The same for several bank types in one pass over the files: outs[k]
 * receives the events of bank_types[k] (each a distinct type below 256),
 * in the layout above.  An event without banks of a type holds
 * n_banks = 0 and one offset.  On an error every outs[k] is empty. */
int velo_mdf_read_banks(const char* const* paths, uint32_t n_paths, uint32_t max_events, uint32_t start_event,
                        int verify_checksum, uint32_t n_threads, const uint32_t* bank_types, uint32_t n_types,
                        velo_mdf_events* outs, char* err, size_t err_len);

void velo_mdf_free(velo_mdf_events* ev);

/* Per-event files of one bank type in a folder, into the same layout
 * (main/src/InputTools.cpp:178-211 read_folder, with list_folder and its
 * natural order :118-157).  File names are "N.bin" or "N_M.bin", ordered as
 * the number pairs (0, N) and (N, M); event i of the output is file
 * (start_event + i) mod (number of files), so a request beyond the folder
 * starts again from its first file, as the reference does.  n_events = 0
 * reads each file once.  n_records receives the number of files.  Returns
 * 0, or -1 with a message in err: missing or empty folder, a file name of
 * another form, an unreadable file.  Free with velo_mdf_free. */
int velo_read_folder(const char* folder, uint32_t n_events, uint32_t start_event, velo_mdf_events* out, char* err,
                     size_t err_len);

/* check_velopix_events (main/src/CheckVP.cpp:5-52): the number of banks,
 * over all events, whose sensor index differs from their position; -1 when
 * an event's bank table or a bank offset reaches past the event. */
int64_t velo_check_vp_events(const char* events, const uint32_t* offsets, uint32_t n_events);

/* The MDF record checksum (genChecksum flag 1) of len bytes. */
uint32_t velo_mdf_checksum(const void* data, size_t len);

#ifdef __cplusplus
}
#endif

#endif
