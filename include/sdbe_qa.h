/*
 * sdbe_qa.h -- C ABI of the FX cross-correlation checker (libsdbe.so), the
 * GPU form of the reference's verification correlator:
 *
 *   sdbe_xcorr_search -> corr_Xt_search / corr_Xt
 *       (software/verification/cross_corr.py:23-116), as verify_scan.py:231-236
 *       uses it on APHIDS output against R2DBE recordings.
 *     Windows of 32768 samples of two real series x0, x1; X = the first 16384
 *     bins of each window's DFT (numpy fft, unnormalised).  From the centre
 *     window c = n0_windows / 2 - avg / 2 of x0, for each offset
 *     iw = first_offset + o (o < n_offsets):
 *       S_0x1 = mean_{w < avg} X0[c + w] conj(X1[c + iw + w])
 *       s_0x1 = irfft(S_0x1, 32768) / sqrt(max s_0x0 max s_1x1)
 *     with s_0x0, s_1x1 the inverse real DFTs of the mean auto-powers of the
 *     same windows (their maxima are at lag 0).  Outputs per offset: s_0x1
 *     (32768 floats), its peak max |s_0x1| and its standard deviation; the
 *     reference's SNR is peak / std at the offset of the highest peak.
 *     x1 leading x0 by L samples within a window puts the peak at lag +L
 *     (lag 32768 - L for a lag of -L).
 *
 * Conventions as in sdbe_stage.h: device pointers, `stream` a hipStream_t as
 * void*, SDBE_OK or a negative SDBE_ERR_* code.
 */
#ifndef SDBE_QA_H
#define SDBE_QA_H

#include <stdint.h>

#include "sdbe_stage.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SDBE_XCORR_WINDOW 32768

typedef struct sdbe_xcorr sdbe_xcorr_t;

/* Capacity: max_windows >= n_offsets + avg - 1 (the x1 windows a search
 * uses), max_offsets >= n_offsets. */
int sdbe_xcorr_create(int device, int32_t max_windows, int32_t max_offsets, sdbe_xcorr_t **out);
void sdbe_xcorr_destroy(sdbe_xcorr_t *x);
int sdbe_xcorr_last_hip_error(const sdbe_xcorr_t *x);

/* d_x0: n0_windows * 32768 floats, d_x1: n1_windows * 32768 floats.  Requires
 * avg >= 1, c + avg <= n0_windows, c + first_offset >= 0 and
 * c + first_offset + n_offsets - 1 + avg <= n1_windows.
 * d_s: [n_offsets][32768] floats; d_peak, d_std: [n_offsets] floats. */
int sdbe_xcorr_search(sdbe_xcorr_t *x, const float *d_x0, int32_t n0_windows, const float *d_x1,
                      int32_t n1_windows, int32_t first_offset, int32_t n_offsets, int32_t avg, float *d_s,
                      float *d_peak, float *d_std, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* SDBE_QA_H */
