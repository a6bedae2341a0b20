// FX cross-correlation checker behind include/sdbe_qa.h: the GPU form of
// corr_Xt_search / corr_Xt (software/verification/cross_corr.py:23-116).
//
// Kernels (32768-sample windows, 16384-point complex FFTs in LDS):
//   k_xc_r2c    one workgroup per window: packed R2C, bins 0..16383 (numpy fft scale)
//   k_xc_pow0   auto-power of the x0 windows, weighted for its lag-0 value
//   k_xc_cross  per (bin block, offset): mean cross-power S_0x1 and the lag-0
//               weighted auto-power of that offset's x1 windows
//   k_xc_c2r    one workgroup per offset: packed C2R of S_0x1 (numpy irfft
//               scale, bin 16384 zero) divided by sqrt(s_0x0[0] s_1x1[0]),
//               with the peak |s| and the standard deviation of s
// The maximum of an autocorrelation (the inverse real DFT of a non-negative
// power spectrum) is its lag-0 value, (P[0] + 2 sum_{k>0} P[k]) / 32768.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>

#include <new>
#include <vector>

#include "sdbe_device.hpp"
#include "../../include/sdbe_qa.h"

using namespace sdbe;

namespace {

constexpr int LW = SDBE_XCORR_WINDOW;     // 32768
constexpr int NHW = LW / 2;               // 16384
constexpr int KB = NHW / 256;             // bin blocks of k_xc_cross / k_xc_pow0
static_assert(NHW == NCH, "the correlator reuses the 16384-point FFT");

__global__ void __launch_bounds__(1024)
k_xc_r2c(const float *__restrict__ x, int w0, float2 *__restrict__ X, const float2 *__restrict__ fft_img,
         const float2 *__restrict__ tw32k) {
    __shared__ float lds[LDS_FFT_N];
    __shared__ FftTw ftw;
    fft_tw_load(&ftw, fft_img);
    const int w = blockIdx.x, t = threadIdx.x;
    const float2 *xb = reinterpret_cast<const float2 *>(x + (size_t)(w0 + w) * LW);
    float2 v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = xb[t + 1024 * q];
    fft16k<-1>(v, lds, &ftw);
    __syncthreads();    // the FFT's last LDS reads are done
#pragma unroll
    for (int q = 0; q < 16; ++q) lds_put(lds, lds_re(t + 1024 * q), v[q]);
    __syncthreads();
    float2 *out = X + (size_t)w * NHW;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int k = t + 1024 * q;
        const int km = (NHW - k) & (NHW - 1);
        const float2 zm = cconj(lds_get(lds, lds_re(km)));   // conj Z[N - k]
        const float2 E = cscale(cadd(v[q], zm), 0.5f);
        const float2 D = cscale(csub(v[q], zm), 0.5f);
        out[k] = cadd(E, mul_si<-1>(cmul(cconj(tw32k[k]), D)));
    }
}

__device__ __forceinline__ double block_sum_256(double a, double *red) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
    __syncthreads();
    return red[0] + red[1] + red[2] + red[3];
}

// part0[kb] = sum over the block's bins of (k ? 2 : 1) mean_w |X0[w][k]|^2
__global__ void __launch_bounds__(256)
k_xc_pow0(const float2 *__restrict__ X0, int avg, double *__restrict__ part0) {
    __shared__ double red[4];
    const int k = blockIdx.x * 256 + threadIdx.x;
    double p = 0.0;
    for (int w = 0; w < avg; ++w) {
        const float2 a = X0[(size_t)w * NHW + k];
        p += (double)a.x * a.x + (double)a.y * a.y;
    }
    p *= (k == 0 ? 1.0 : 2.0) / avg;
    const double s = block_sum_256(p, red);
    if (threadIdx.x == 0) part0[blockIdx.x] = s;
}

// offset o: S[o][k] = mean_w X0[w][k] conj(X1[o + w][k]); part1[o][kb] as part0
// for the windows X1[o .. o + avg)
__global__ void __launch_bounds__(256)
k_xc_cross(const float2 *__restrict__ X0, const float2 *__restrict__ X1, int avg, float2 *__restrict__ S,
           double *__restrict__ part1) {
    __shared__ double red[4];
    const int k = blockIdx.x * 256 + threadIdx.x, o = blockIdx.y;
    double sr = 0.0, si = 0.0, p = 0.0;
    for (int w = 0; w < avg; ++w) {
        const float2 a = X0[(size_t)w * NHW + k], b = X1[(size_t)(o + w) * NHW + k];
        sr += (double)a.x * b.x + (double)a.y * b.y;     // a conj(b)
        si += (double)a.y * b.x - (double)a.x * b.y;
        p += (double)b.x * b.x + (double)b.y * b.y;
    }
    S[(size_t)o * NHW + k] = make_float2((float)(sr / avg), (float)(si / avg));
    p *= (k == 0 ? 1.0 : 2.0) / avg;
    const double s = block_sum_256(p, red);
    if (threadIdx.x == 0) part1[(size_t)o * KB + blockIdx.x] = s;
}

__global__ void __launch_bounds__(1024)
k_xc_c2r(const float2 *__restrict__ S, const double *__restrict__ part0, const double *__restrict__ part1,
         float *__restrict__ s_out, float *__restrict__ peak, float *__restrict__ sdev,
         const float2 *__restrict__ fft_img, const float2 *__restrict__ tw32k) {
    __shared__ float lds[LDS_FFT_N];
    __shared__ FftTw ftw;
    __shared__ double red[3][16];
    fft_tw_load(&ftw, fft_img);
    const int o = blockIdx.x, t = threadIdx.x;
    double n0 = 0.0, n1 = 0.0;
    for (int i = 0; i < KB; ++i) {
        n0 += part0[i];
        n1 += part1[(size_t)o * KB + i];
    }
    // irfft scale 1 / 32768 and the normalisation by the lag-0 autocorrelations
    const float sc = (float)(1.0 / ((double)LW * sqrt((n0 / LW) * (n1 / LW))));
    const float2 *Xo = S + (size_t)o * NHW;
    float2 v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int k = t + 1024 * q;
        float2 a = Xo[k];
        float2 b = k == 0 ? make_float2(0.f, 0.f) : Xo[NHW - k];    // bin 16384 is zero
        if (k == 0) a.y = 0.f;                                      // bin 0: real part only
        const float2 sum = cadd(a, cconj(b)), dif = csub(a, cconj(b));
        v[q] = cadd(sum, mul_si<+1>(cmul(tw32k[k], dif)));
    }
    fft16k<+1>(v, lds, &ftw);
    float2 *dst = reinterpret_cast<float2 *>(s_out + (size_t)o * LW);
    double s1 = 0.0, s2 = 0.0;
    float mx = 0.f;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const float2 y = cscale(v[q], sc);
        dst[t + 1024 * q] = y;
        s1 += (double)y.x + (double)y.y;
        s2 += (double)y.x * y.x + (double)y.y * y.y;
        mx = fmaxf(mx, fmaxf(fabsf(y.x), fabsf(y.y)));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        s1 += __shfl_xor(s1, off);
        s2 += __shfl_xor(s2, off);
        mx = fmaxf(mx, __shfl_xor(mx, off));
    }
    if ((t & 63) == 0) {
        red[0][t >> 6] = s1;
        red[1][t >> 6] = s2;
        red[2][t >> 6] = (double)mx;
    }
    __syncthreads();
    if (t == 0) {
        double a = 0.0, b = 0.0, m = 0.0;
        for (int i = 0; i < 16; ++i) {
            a += red[0][i];
            b += red[1][i];
            m = fmax(m, red[2][i]);
        }
        const double mean = a / LW;
        double var = b / LW - mean * mean;
        if (var < 0.0) var = 0.0;
        peak[o] = (float)m;
        sdev[o] = (float)sqrt(var);
    }
}

}  // namespace

struct sdbe_xcorr {
    int device = 0;
    int max_win = 0, max_off = 0;
    int last_hip_err = 0;
    float2 *fft_img = nullptr;
    float2 *tw32k = nullptr;      // e^{2 pi i k / 32768}, k < 16384
    float2 *X0 = nullptr;         // [max_win][16384]
    float2 *X1 = nullptr;         // [max_win][16384]
    float2 *S = nullptr;          // [max_off][16384]
    double *part0 = nullptr;      // [KB]
    double *part1 = nullptr;      // [max_off][KB]
};

namespace {
int xc_fail(sdbe_xcorr_t *x, hipError_t e) {
    if (x) x->last_hip_err = (int)e;
    return SDBE_ERR_HIP;
}
#define XC_TRY(x, expr)                               \
    do {                                              \
        hipError_t _e = (expr);                       \
        if (_e != hipSuccess) return xc_fail((x), _e); \
    } while (0)
}  // namespace

extern "C" {

void sdbe_xcorr_destroy(sdbe_xcorr_t *x) {
    if (!x) return;
    (void)hipSetDevice(x->device);
    (void)hipDeviceSynchronize();
    (void)hipFree(x->fft_img);
    (void)hipFree(x->tw32k);
    (void)hipFree(x->X0);
    (void)hipFree(x->X1);
    (void)hipFree(x->S);
    (void)hipFree(x->part0);
    (void)hipFree(x->part1);
    delete x;
}

int sdbe_xcorr_last_hip_error(const sdbe_xcorr_t *x) { return x ? x->last_hip_err : 0; }

int sdbe_xcorr_create(int device, int32_t max_windows, int32_t max_offsets, sdbe_xcorr_t **out) {
    if (!out) return SDBE_ERR_ARG;
    *out = nullptr;
    if (max_windows < 1 || max_windows > 65536 || max_offsets < 1 || max_offsets > 65535) return SDBE_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return SDBE_ERR_NODEV;
    if (hipSetDevice(device) != hipSuccess) return SDBE_ERR_NODEV;
    sdbe_xcorr_t *x = new (std::nothrow) sdbe_xcorr();
    if (!x) return SDBE_ERR_NOMEM;
    x->device = device;
    x->max_win = max_windows;
    x->max_off = max_offsets;
    int rc = SDBE_OK;
    std::vector<float2> img(FFT_TW_N), tw(NHW);
    const double two_pi = 6.283185307179586476925286766559;
    fft_tw_fill(img.data());
    for (int k = 0; k < NHW; ++k) {
        const double ang = two_pi * (double)k / (double)LW;
        tw[k] = make_float2((float)cos(ang), (float)sin(ang));
    }
#define XALLOC(ptr, bytes)                                              \
    do {                                                                \
        if (hipMalloc((void **)&(ptr), (bytes)) != hipSuccess) {        \
            rc = SDBE_ERR_NOMEM;                                        \
            goto fail;                                                  \
        }                                                               \
    } while (0)
    XALLOC(x->fft_img, FFT_TW_N * sizeof(float2));
    XALLOC(x->tw32k, NHW * sizeof(float2));
    XALLOC(x->X0, (size_t)max_windows * NHW * sizeof(float2));
    XALLOC(x->X1, (size_t)max_windows * NHW * sizeof(float2));
    XALLOC(x->S, (size_t)max_offsets * NHW * sizeof(float2));
    XALLOC(x->part0, KB * sizeof(double));
    XALLOC(x->part1, (size_t)max_offsets * KB * sizeof(double));
#undef XALLOC
    if (hipMemcpy(x->fft_img, img.data(), FFT_TW_N * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(x->tw32k, tw.data(), NHW * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess) {
        rc = SDBE_ERR_HIP;
        goto fail;
    }
    *out = x;
    return SDBE_OK;
fail:
    sdbe_xcorr_destroy(x);
    return rc;
}

int sdbe_xcorr_search(sdbe_xcorr_t *x, const float *d_x0, int32_t n0_windows, const float *d_x1,
                      int32_t n1_windows, int32_t first_offset, int32_t n_offsets, int32_t avg, float *d_s,
                      float *d_peak, float *d_std, void *stream) {
    if (!x || !d_x0 || !d_x1 || !d_s || !d_peak || !d_std) return SDBE_ERR_ARG;
    if (avg < 1 || n_offsets < 1 || n0_windows < 1 || n1_windows < 1) return SDBE_ERR_ARG;
    const int64_t c = n0_windows / 2 - avg / 2;                 // cross_corr.py:97
    const int64_t lo = c + first_offset, hi = c + first_offset + (int64_t)n_offsets - 1 + avg;
    if (c < 0 || c + avg > n0_windows || lo < 0 || hi > n1_windows) return SDBE_ERR_ARG;
    if (n_offsets > x->max_off || avg > x->max_win || (int64_t)n_offsets + avg - 1 > x->max_win)
        return SDBE_ERR_CAPACITY;
    XC_TRY(x, hipSetDevice(x->device));
    hipStream_t s = (hipStream_t)stream;
    // spectra of the windows the search uses: x0[c, c + avg), x1[lo, lo + n_offsets + avg - 1)
    hipLaunchKernelGGL(k_xc_r2c, dim3(avg), dim3(1024), 0, s, d_x0, (int)c, x->X0, (const float2 *)x->fft_img,
                       (const float2 *)x->tw32k);
    hipLaunchKernelGGL(k_xc_r2c, dim3(n_offsets + avg - 1), dim3(1024), 0, s, d_x1, (int)lo, x->X1,
                       (const float2 *)x->fft_img, (const float2 *)x->tw32k);
    hipLaunchKernelGGL(k_xc_pow0, dim3(KB), dim3(256), 0, s, (const float2 *)x->X0, (int)avg, x->part0);
    hipLaunchKernelGGL(k_xc_cross, dim3(KB, n_offsets), dim3(256), 0, s, (const float2 *)x->X0,
                       (const float2 *)x->X1, (int)avg, x->S, x->part1);
    hipLaunchKernelGGL(k_xc_c2r, dim3(n_offsets), dim3(1024), 0, s, (const float2 *)x->S,
                       (const double *)x->part0, (const double *)x->part1, d_s, d_peak, d_std,
                       (const float2 *)x->fft_img, (const float2 *)x->tw32k);
    XC_TRY(x, hipGetLastError());
    return SDBE_OK;
}

}  // extern "C"
