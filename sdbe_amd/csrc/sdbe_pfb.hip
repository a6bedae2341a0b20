// gfx950 polyphase filterbank (PFB) and its least-squares inverse behind the
// C-ABI of include/sdbe_pfb.h.
//
// Geometry: blocks of L = 32768 real samples, 16385 bins, 4 taps.
//   forward  spec_i = rfft_L( sum_p w[p L + n] x[(i + p) L + n] )
//            (equal to every 4th bin of the rfft of the windowed 4 L samples,
//            software/prototyping/cuda/pfb_inverse.cu:172-208)
//   inverse  yh_i = irfft_L(spec_i); per branch n, with P[i][j] = w[n + L (j - i)]
//            (0 <= j - i < 4): solve (P P^T) z = yh[., n], x[., n] = P^T z
//            (the minimum-norm least-squares solution; pfb_inverse.cu:213-488
//            sets up the same band matrices, its solve is compiled out)
//
// Kernels:
//   k_pfb_fold_r2c  one workgroup per spectrum: 4-tap fold, packed R2C
//                   (16384-point complex FFT in LDS + split post-processing)
//   k_pfb_c2r       one workgroup per spectrum: packed C2R to the pseudo-timestream
//   k_pfb_fwd       one thread per branch: banded Cholesky of P P^T (bandwidth 3,
//                   Toeplitz per branch) and forward substitution, FP64; the
//                   factor's state is checkpointed every 32 rows
//   k_pfb_bwd       one thread per branch: per 32-row segment, from the last,
//                   recompute the factor rows from the checkpoint, back-substitute
//                   and emit the 4-tap projection x = P^T z
// The solve is FP64: P P^T is badly conditioned for the branches whose tap
// polynomial has a zero near z = -1 (condition ~ (nblock / pi)^2 for a
// symmetric prototype near the middle branch).
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <string.h>

#include <new>
#include <vector>

#include "sdbe_device.hpp"
#include "../../include/sdbe_pfb.h"
#include "../../include/sdbe_stage.h"

using namespace sdbe;

namespace {

constexpr int LB = SDBE_PFB_LBLOCK;       // 32768
constexpr int NH = LB / 2;                // 16384-point complex FFT
constexpr int NBIN = SDBE_PFB_NBINS;      // 16385
constexpr int NT = SDBE_PFB_NTAP;         // 4
constexpr int SEG = 32;                   // factor rows per checkpoint segment
constexpr int SOLVE_WG = 64;

static_assert(NH == NCH, "the PFB reuses the 16384-point FFT");

// Forward PFB of spectrum i.  z[m] = y[2m] + i y[2m+1] of the folded block;
// Z = FFT_16384(z); X[k] = 1/2 (Z[k] + conj Z[N-k]) - 1/2 i e^{-2 pi i k/L} (Z[k] - conj Z[N-k]),
// X[N] = Re Z[0] - Im Z[0].
__global__ void __launch_bounds__(1024)
k_pfb_fold_r2c(const float *__restrict__ x, const float *__restrict__ win, float2 *__restrict__ spec,
               const float2 *__restrict__ fft_img, const float2 *__restrict__ tw32k) {
    __shared__ float lds[LDS_FFT_N];
    __shared__ FftTw ftw;
    fft_tw_load(&ftw, fft_img);
    const int i = blockIdx.x, t = threadIdx.x;
    const float2 *xb = reinterpret_cast<const float2 *>(x + (size_t)i * LB);
    const float2 *wb = reinterpret_cast<const float2 *>(win);
    float2 v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int m = t + 1024 * q;
        float2 acc = make_float2(0.f, 0.f);
#pragma unroll
        for (int p = 0; p < NT; ++p) {
            const float2 xv = xb[(size_t)p * NH + m];
            const float2 wv = wb[(size_t)p * NH + m];
            acc.x = fmaf(wv.x, xv.x, acc.x);
            acc.y = fmaf(wv.y, xv.y, acc.y);
        }
        v[q] = acc;
    }
    fft16k<-1>(v, lds, &ftw);
    __syncthreads();    // the FFT's last LDS reads are done
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int k = t + 1024 * q;
        lds_put(lds, lds_re(k), v[q]);
    }
    __syncthreads();
    float2 *out = spec + (size_t)i * NBIN;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int k = t + 1024 * q;
        const int km = (NH - k) & (NH - 1);
        const float2 zm = cconj(lds_get(lds, lds_re(km)));   // conj Z[N - k]
        const float2 E = cscale(cadd(v[q], zm), 0.5f);
        const float2 D = cscale(csub(v[q], zm), 0.5f);
        out[k] = cadd(E, mul_si<-1>(cmul(cconj(tw32k[k]), D)));
    }
    if (t == 0) out[NH] = make_float2(v[0].x - v[0].y, 0.f);
}

// Pseudo-timestream of spectrum i: Z[k] = (X[k] + conj X[N-k]) + i e^{2 pi i k/L} (X[k] - conj X[N-k])
// (imaginary parts of X[0], X[N] dropped), z = IFFT_16384(Z) / L, samples (2m, 2m+1) = (Re, Im) z[m]
// -- numpy's irfft.
__global__ void __launch_bounds__(1024)
k_pfb_c2r(const float2 *__restrict__ spec, float *__restrict__ yh, const float2 *__restrict__ fft_img,
          const float2 *__restrict__ tw32k) {
    __shared__ float lds[LDS_FFT_N];
    __shared__ FftTw ftw;
    fft_tw_load(&ftw, fft_img);
    const int i = blockIdx.x, t = threadIdx.x;
    const float2 *X = spec + (size_t)i * NBIN;
    float2 v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int k = t + 1024 * q;
        float2 a = X[k], b = X[NH - k];     // k = 0 pairs with bin N
        if (k == 0) {
            a.y = 0.f;
            b.y = 0.f;
        }
        const float2 sum = cadd(a, cconj(b)), dif = csub(a, cconj(b));
        v[q] = cadd(sum, mul_si<+1>(cmul(tw32k[k], dif)));
    }
    fft16k<+1>(v, lds, &ftw);
    float2 *dst = reinterpret_cast<float2 *>(yh + (size_t)i * LB);
    const float sc = 1.0f / (float)LB;
#pragma unroll
    for (int q = 0; q < 16; ++q) dst[t + 1024 * q] = cscale(v[q], sc);
}

// One row of the banded Cholesky factor of the Toeplitz band (a0, a1, a2, a3):
// inv0 = 1 / L(i, i), dk = L(i, i - k).  r1, r2, r3 are rows i - 1, i - 2,
// i - 3 ({1, 0, 0, 0} before the first row).  Keeping the reciprocal of the
// diagonal makes a row one reciprocal square root and fused multiply-adds: the
// rows form a dependent chain, and FP64 division and square root are
// multi-instruction sequences that would sit on it four times per row.
struct Row {
    double inv0, d1, d2, d3;
};

__device__ __forceinline__ Row chol_row(int i, double a0, double a1, double a2, double a3, const Row &r1,
                                        const Row &r2, const Row &r3) {
    const double e1 = i >= 1 ? a1 : 0.0, e2 = i >= 2 ? a2 : 0.0, e3 = i >= 3 ? a3 : 0.0;
    Row r;
    r.d3 = e3 * r3.inv0;
    r.d2 = (e2 - r.d3 * r2.d1) * r2.inv0;
    r.d1 = (e1 - r.d3 * r1.d2 - r.d2 * r1.d1) * r1.inv0;
    r.inv0 = rsqrt(a0 - r.d1 * r.d1 - r.d2 * r.d2 - r.d3 * r.d3);
    return r;
}

// Forward substitution L u = yh[., n], factor rows computed on the fly; the
// three previous rows are saved at the start of every SEG-row segment.
__global__ void __launch_bounds__(SOLVE_WG)
k_pfb_fwd(const float *__restrict__ yh, int nblock, const double *__restrict__ a, double *__restrict__ u,
          double *__restrict__ ckpt) {
    const int n = blockIdx.x * SOLVE_WG + threadIdx.x;
    const double a0 = a[n], a1 = a[LB + n], a2 = a[2 * LB + n], a3 = a[3 * LB + n];
    Row r1 = {1.0, 0.0, 0.0, 0.0}, r2 = r1, r3 = r1;
    double u1 = 0.0, u2 = 0.0, u3 = 0.0;
    for (int s0 = 0; s0 < nblock; s0 += SEG) {
        double *ck = ckpt + (size_t)(s0 / SEG) * 12 * LB + n;
        const double st[12] = {r1.inv0, r1.d1, r1.d2, r1.d3, r2.inv0, r2.d1, r2.d2, r2.d3,
                               r3.inv0, r3.d1, r3.d2, r3.d3};
#pragma unroll
        for (int k = 0; k < 12; ++k) ck[(size_t)k * LB] = st[k];
        float y[SEG];
#pragma unroll
        for (int k = 0; k < SEG; ++k) y[k] = yh[(size_t)min(s0 + k, nblock - 1) * LB + n];
#pragma unroll
        for (int k = 0; k < SEG; ++k) {
            const int i = s0 + k;
            if (i < nblock) {
                const Row r = chol_row(i, a0, a1, a2, a3, r1, r2, r3);
                const double ui = ((double)y[k] - r.d1 * u1 - r.d2 * u2 - r.d3 * u3) * r.inv0;
                u[(size_t)i * LB + n] = ui;
                r3 = r2;
                r2 = r1;
                r1 = r;
                u3 = u2;
                u2 = u1;
                u1 = ui;
            }
        }
    }
}

// Back substitution L^T z = u and projection x[j] = sum_p w[n + L p] z[j - p],
// segment by segment from the last; x[i + 3] is complete once z[i] is known.
__global__ void __launch_bounds__(SOLVE_WG)
k_pfb_bwd(const double *__restrict__ u, int nblock, const double *__restrict__ a, const double *__restrict__ ckpt,
          const float *__restrict__ win, float *__restrict__ x) {
    const int n = blockIdx.x * SOLVE_WG + threadIdx.x;
    const double a0 = a[n], a1 = a[LB + n], a2 = a[2 * LB + n], a3 = a[3 * LB + n];
    const double w0 = win[n], w1 = win[LB + n], w2 = win[2 * LB + n], w3 = win[3 * LB + n];
    Row q1 = {1.0, 0.0, 0.0, 0.0}, q2 = q1, q3 = q1;      // rows i + 1, i + 2, i + 3
    double z1 = 0.0, z2 = 0.0, z3 = 0.0;                  // z[i + 1], z[i + 2], z[i + 3]
    const int nseg = (nblock + SEG - 1) / SEG;
    for (int s = nseg - 1; s >= 0; --s) {
        const int s0 = s * SEG;
        const double *ck = ckpt + (size_t)s * 12 * LB + n;
        Row r1 = {ck[0], ck[(size_t)LB], ck[(size_t)2 * LB], ck[(size_t)3 * LB]};
        Row r2 = {ck[(size_t)4 * LB], ck[(size_t)5 * LB], ck[(size_t)6 * LB], ck[(size_t)7 * LB]};
        Row r3 = {ck[(size_t)8 * LB], ck[(size_t)9 * LB], ck[(size_t)10 * LB], ck[(size_t)11 * LB]};
        double uu[SEG];
#pragma unroll
        for (int k = 0; k < SEG; ++k) uu[k] = u[(size_t)min(s0 + k, nblock - 1) * LB + n];
        Row Lr[SEG];
#pragma unroll
        for (int k = 0; k < SEG; ++k) {
            Lr[k] = chol_row(s0 + k, a0, a1, a2, a3, r1, r2, r3);
            r3 = r2;
            r2 = r1;
            r1 = Lr[k];
        }
#pragma unroll
        for (int k = SEG - 1; k >= 0; --k) {
            const int i = s0 + k;
            if (i < nblock) {
                const double zi = (uu[k] - q1.d1 * z1 - q2.d2 * z2 - q3.d3 * z3) * Lr[k].inv0;
                x[(size_t)(i + 3) * LB + n] = (float)(w0 * z3 + w1 * z2 + w2 * z1 + w3 * zi);
                q3 = q2;
                q2 = q1;
                q1 = Lr[k];
                z3 = z2;
                z2 = z1;
                z1 = zi;
            }
        }
    }
    // z[i] = 0 for i < 0
    x[(size_t)2 * LB + n] = (float)(w0 * z3 + w1 * z2 + w2 * z1);
    x[(size_t)LB + n] = (float)(w0 * z2 + w1 * z1);
    x[n] = (float)(w0 * z1);
}

}  // namespace

struct sdbe_pfb {
    int device = 0;
    int max_spec = 0;
    int last_hip_err = 0;
    float2 *fft_img = nullptr;    // FftTw image
    float2 *tw32k = nullptr;      // e^{2 pi i k / 32768}, k < 16384
    float *win = nullptr;         // [4][32768]
    double *a = nullptr;          // [4][32768] band of P P^T per branch
    float *yh = nullptr;          // [max_spec][32768]
    double *u = nullptr;          // [max_spec][32768]
    double *ckpt = nullptr;       // [segments][12][32768]
};

namespace {
int pfb_fail(sdbe_pfb_t *p, hipError_t e) {
    if (p) p->last_hip_err = (int)e;
    return SDBE_ERR_HIP;
}
#define PFB_TRY(p, expr)                               \
    do {                                               \
        hipError_t _e = (expr);                        \
        if (_e != hipSuccess) return pfb_fail((p), _e); \
    } while (0)
}  // namespace

extern "C" {

void sdbe_pfb_destroy(sdbe_pfb_t *p) {
    if (!p) return;
    (void)hipSetDevice(p->device);
    (void)hipDeviceSynchronize();
    (void)hipFree(p->fft_img);
    (void)hipFree(p->tw32k);
    (void)hipFree(p->win);
    (void)hipFree(p->a);
    (void)hipFree(p->yh);
    (void)hipFree(p->u);
    (void)hipFree(p->ckpt);
    delete p;
}

int sdbe_pfb_last_hip_error(const sdbe_pfb_t *p) { return p ? p->last_hip_err : 0; }

int sdbe_pfb_create(int device, int32_t n_taps, const float *h_window, int32_t max_spectra, sdbe_pfb_t **out) {
    if (!out) return SDBE_ERR_ARG;
    *out = nullptr;
    if (n_taps != NT || !h_window || max_spectra < 1 || max_spectra > (1 << 16)) return SDBE_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return SDBE_ERR_NODEV;
    if (hipSetDevice(device) != hipSuccess) return SDBE_ERR_NODEV;
    sdbe_pfb_t *p = new (std::nothrow) sdbe_pfb();
    if (!p) return SDBE_ERR_NOMEM;
    p->device = device;
    p->max_spec = max_spectra;
    const size_t nseg = (size_t)(max_spectra + SEG - 1) / SEG;
    int rc = SDBE_OK;
    std::vector<float2> img(FFT_TW_N), tw(NH);
    std::vector<double> band((size_t)NT * LB);
    const double two_pi = 6.283185307179586476925286766559;
    fft_tw_fill(img.data());
    for (int k = 0; k < NH; ++k) {
        const double ang = two_pi * (double)k / (double)LB;
        tw[k] = make_float2((float)cos(ang), (float)sin(ang));
    }
    for (int n = 0; n < LB; ++n)
        for (int d = 0; d < NT; ++d) {
            double acc = 0.0;
            for (int k = 0; k + d < NT; ++k)
                acc += (double)h_window[(size_t)(k + d) * LB + n] * (double)h_window[(size_t)k * LB + n];
            band[(size_t)d * LB + n] = acc;
        }
#define PALLOC(ptr, bytes)                                              \
    do {                                                                \
        if (hipMalloc((void **)&(ptr), (bytes)) != hipSuccess) {        \
            rc = SDBE_ERR_NOMEM;                                        \
            goto fail;                                                  \
        }                                                               \
    } while (0)
    PALLOC(p->fft_img, FFT_TW_N * sizeof(float2));
    PALLOC(p->tw32k, NH * sizeof(float2));
    PALLOC(p->win, (size_t)NT * LB * sizeof(float));
    PALLOC(p->a, (size_t)NT * LB * sizeof(double));
    PALLOC(p->yh, (size_t)max_spectra * LB * sizeof(float));
    PALLOC(p->u, (size_t)max_spectra * LB * sizeof(double));
    PALLOC(p->ckpt, nseg * 12 * LB * sizeof(double));
#undef PALLOC
    if (hipMemcpy(p->fft_img, img.data(), FFT_TW_N * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->tw32k, tw.data(), NH * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->win, h_window, (size_t)NT * LB * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->a, band.data(), (size_t)NT * LB * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        rc = SDBE_ERR_HIP;
        goto fail;
    }
    *out = p;
    return SDBE_OK;
fail:
    sdbe_pfb_destroy(p);
    return rc;
}

int sdbe_pfb_forward(sdbe_pfb_t *p, const float *d_x, int32_t n_spectra, void *d_spec, void *stream) {
    if (!p || !d_x || !d_spec || n_spectra < 1) return SDBE_ERR_ARG;
    if (n_spectra > p->max_spec) return SDBE_ERR_CAPACITY;
    PFB_TRY(p, hipSetDevice(p->device));
    hipLaunchKernelGGL(k_pfb_fold_r2c, dim3(n_spectra), dim3(1024), 0, (hipStream_t)stream, d_x,
                       (const float *)p->win, (float2 *)d_spec, (const float2 *)p->fft_img,
                       (const float2 *)p->tw32k);
    PFB_TRY(p, hipGetLastError());
    return SDBE_OK;
}

int sdbe_pfb_inverse(sdbe_pfb_t *p, const void *d_spec, int32_t n_spectra, float *d_x, void *stream) {
    if (!p || !d_spec || !d_x || n_spectra < 1) return SDBE_ERR_ARG;
    if (n_spectra > p->max_spec) return SDBE_ERR_CAPACITY;
    PFB_TRY(p, hipSetDevice(p->device));
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_pfb_c2r, dim3(n_spectra), dim3(1024), 0, s, (const float2 *)d_spec, p->yh,
                       (const float2 *)p->fft_img, (const float2 *)p->tw32k);
    hipLaunchKernelGGL(k_pfb_fwd, dim3(LB / SOLVE_WG), dim3(SOLVE_WG), 0, s, (const float *)p->yh, (int)n_spectra,
                       (const double *)p->a, p->u, p->ckpt);
    hipLaunchKernelGGL(k_pfb_bwd, dim3(LB / SOLVE_WG), dim3(SOLVE_WG), 0, s, (const double *)p->u, (int)n_spectra,
                       (const double *)p->a, (const double *)p->ckpt, (const float *)p->win, d_x);
    PFB_TRY(p, hipGetLastError());
    return SDBE_OK;
}

}  // extern "C"
