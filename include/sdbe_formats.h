/*
 * sdbe_formats.h -- C ABI of the input-format variants of the SWARM data
 * back-end stage (libsdbe.so).
 *
 * Conventions as in sdbe_stage.h: device pointers are plain pointers, `stream`
 * is a hipStream_t passed as void* (NULL = the default stream), every call
 * returns SDBE_OK (0) or a negative SDBE_ERR_* code and never aborts.
 *
 *   sdbe_reorder_2015 -> reorderTz_smem
 *       (software/prototyping/cuda/kernel_template.py:107-164; numpy form
 *        software/prototyping/cuda/reorder.py:284-293)
 *     The 2015-format B-engine reorder.  Input: complex64 spectra
 *     [16384 channels][n_frames B-frames][128 snapshots], channel-major.
 *     Output: complex64 [n_frames - 1][128][16385]:
 *       out[f][s][c] = in[c][f + (s >= 69)][(s + 2) mod 128]  if ((c / 4) & 1) == 0
 *       out[f][s][c] = in[c][f + (s >= 69)][s]                 otherwise
 *       out[f][s][16384] = 0
 *     i.e. channels a-d of every 8 lag the others by two snapshots, and
 *     snapshots 69..127 of an output frame come from the next B-frame (the
 *     last input frame only feeds the frame before it).
 */
#ifndef SDBE_FORMATS_H
#define SDBE_FORMATS_H

#include <stdint.h>

#include "sdbe_stage.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SDBE_REORDER2015_CHANNELS 16384
#define SDBE_REORDER2015_SNAPSHOTS 128
#define SDBE_REORDER2015_SPLIT 69   /* first snapshot taken from the next B-frame */

/* n_frames >= 2 input B-frames; d_in and d_out must not overlap. */
int sdbe_reorder_2015(const void *d_in, int32_t n_frames, void *d_out, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* SDBE_FORMATS_H */
