/*
 * sdbe_pfb.h -- C-ABI of the polyphase filterbank (PFB) path on gfx950:
 * a critically sampled real PFB of 32768-sample blocks (16385 bins) with a
 * 4-tap prototype, and its least-squares inverse.
 *
 * Reference interfaces replaced (paths relative to the Smithsonian/sdbe tree):
 *   sdbe_pfb_forward -> pfb (software/prototyping/cuda/pfb_inverse.cu:172-208):
 *                       window over 4 * 32768 samples, R2C, every 4th bin
 *   sdbe_pfb_inverse -> inverse_pfb (pfb_inverse.cu:213-488): C2R of each
 *                       spectrum to the pseudo-timestream (x 1/32768), then per
 *                       branch the minimum-norm least-squares solution
 *                       x = P^T (P P^T)^-1 yh, the method the file attributes to
 *                       R. Shaw's pfb-inverse (its own solve is compiled out,
 *                       :392-455)
 *   sdbe_pfb_create  -> the plan, band-matrix and window set-up of
 *                       inverse_pfb (pfb_inverse.cu:219-330) and main (:491-527)
 *
 * The window (prototype filter) is the caller's: the reference's Hamming
 * window (pfb_inverse.cu:51-53) or any other of 4 * 32768 coefficients.
 * Device pointers, sizes and return codes as in sdbe_stage.h.
 */
#ifndef SDBE_PFB_H
#define SDBE_PFB_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SDBE_PFB_LBLOCK 32768          /* samples per block                  */
#define SDBE_PFB_NBINS  16385          /* bins per spectrum                  */
#define SDBE_PFB_NTAP   4

typedef struct sdbe_pfb sdbe_pfb_t;

/* h_window: host float [n_taps * 32768]; n_taps must be 4.  max_spectra
 * bounds the spectra per call (scratch of about 9 bytes per sample).  The
 * plan precomputes the window's banded factor checkpoints and span
 * transfers at creation (under a millisecond for 4096 spectra).           */
int sdbe_pfb_create(int device, int32_t n_taps, const float *h_window, int32_t max_spectra,
                    sdbe_pfb_t **out);
void sdbe_pfb_destroy(sdbe_pfb_t *p);
int sdbe_pfb_last_hip_error(const sdbe_pfb_t *p);

/* d_x: float [(n_spectra + 3) * 32768] time series; d_spec: float2
 * [n_spectra][16385]; spectrum i covers samples [32768 i, 32768 (i + 4)). */
int sdbe_pfb_forward(sdbe_pfb_t *p, const float *d_x, int32_t n_spectra, void *d_spec, void *stream);

/* d_spec: float2 [n_spectra][16385] (the imaginary parts of bins 0 and
 * 16384 are ignored, as by a real inverse FFT); d_x: float
 * [(n_spectra + 3) * 32768], the minimum-norm time series whose PFB
 * spectra's inverse FFTs are those of d_spec.                              */
int sdbe_pfb_inverse(sdbe_pfb_t *p, const void *d_spec, int32_t n_spectra, float *d_x, void *stream);

#ifdef __cplusplus
}
#endif

#endif
