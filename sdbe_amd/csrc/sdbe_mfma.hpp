// 16384-point FFT whose radix-16 passes run as DFT-as-GEMM on the gfx950
// matrix cores (v_mfma_f32_16x16x32_{bf16,f16}, FP32 accumulation).
//
// A complex DFT-16 in real form is a 32 x 32 matrix (rows: 16 real then 16
// imaginary outputs; K = 2 r + c over the 16 inputs r and their real / imag
// parts c) times a 32-row column per DFT.  One 16x16x32 MFMA produces 16 output
// rows for 16 DFT columns, so a wave transforms 16 DFTs with one MFMA for the
// real and one for the imaginary outputs (x3 for the split-precision mode).
//
// Operand layout of v_mfma_f32_16x16x32 on gfx950 (checked on the device):
//   A: lane l holds A[l % 16][8 (l / 16) + e],  e < 8
//   B: lane l holds B[8 (l / 16) + e][l % 16]
//   C: lane l holds C[4 (l / 16) + i][l % 16],  i < 4
// so lane (n = l % 16, kb = l / 16) supplies inputs r = 4 kb .. 4 kb + 3 of DFT
// column n and receives outputs k = 4 kb .. 4 kb + 3 of the same column.
//
// Work distribution: wave w (of 16), group g (of 4): DFT d = 64 w + 16 g + n.
// In every radix-16 pass DFT d reads elements d + 1024 r and writes its
// output k to (d / Ns) 16 Ns + (d % Ns) + Ns k (Stockham, Ns = 1, 16, 256).
// The last pass (radix 4, Ns = 4096) is the VALU pass of fft16k, so the
// output is in natural order: register q of thread t holds element t + 1024 q.
//
// Precision modes:
//   FFT_BF16X3  x = hi + lo, W = Whi + Wlo (bf16 each), Y = Whi hi + Whi lo + Wlo hi:
//               about 2^-16 relative per pass, bf16 keeps the FP32 exponent range
//   FFT_BF16    one bf16 product (2^-9 relative per pass)
//   FFT_FP16    one f16 product (2^-12 relative per pass); every DFT column is
//               scaled by a power of two so that its largest value is below 2^15
#pragma once

#include "sdbe_device.hpp"

namespace sdbe {

enum FftMode { FFT_BF16X3 = 1, FFT_BF16 = 2, FFT_FP16 = 3 };

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// cos(2 pi m / 16), m = 0..15
__device__ __forceinline__ float cos16(int m) {
    const float c[16] = {1.0f, 0.92387953251128674f, 0.70710678118654757f, 0.38268343236508978f,
                         0.0f, -0.38268343236508978f, -0.70710678118654757f, -0.92387953251128674f,
                         -1.0f, -0.92387953251128674f, -0.70710678118654757f, -0.38268343236508978f,
                         0.0f, 0.38268343236508978f, 0.70710678118654757f, 0.92387953251128674f};
    return c[m & 15];
}

// Constant DFT-16 operands of one lane for sign S (row k = l % 16).
template <int S, int P>
struct MfmaW {
    bf16x8 re_hi, re_lo, im_hi, im_lo;
    f16x8 re_h, im_h;
    __device__ __forceinline__ void init(int t) {
        const int l = t & 63, k = l & 15, kb = l >> 4;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int r = 4 * kb + (e >> 1);
            const int m = (r * k) & 15;
            const float c = cos16(m), s = (S > 0 ? 1.0f : -1.0f) * cos16(m - 4);   // sin(2 pi m/16) = cos(2 pi (m - 4)/16)
            const float are = (e & 1) ? -s : c;
            const float aim = (e & 1) ? c : s;
            if (P == FFT_FP16) {
                re_h[e] = (_Float16)are;
                im_h[e] = (_Float16)aim;
            } else {
                re_hi[e] = (__bf16)are;
                im_hi[e] = (__bf16)aim;
                re_lo[e] = (__bf16)(are - (float)re_hi[e]);
                im_lo[e] = (__bf16)(aim - (float)im_hi[e]);
            }
        }
    }
};

// 16 DFT-16s of one wave: x[j] = input 4 kb + j of column n -> output 4 kb + j.
template <int S, int P>
__device__ __forceinline__ void mfma_dft16(float2 *x, const MfmaW<S, P> &W) {
    f32x4 cr = {0.f, 0.f, 0.f, 0.f}, ci = {0.f, 0.f, 0.f, 0.f};
    if (P == FFT_FP16) {
        float m = 0.f;
#pragma unroll
        for (int j = 0; j < 4; ++j) m = fmaxf(m, fmaxf(fabsf(x[j].x), fabsf(x[j].y)));
        m = fmaxf(m, __shfl_xor(m, 16));
        m = fmaxf(m, __shfl_xor(m, 32));
        // power-of-two scale putting the column's largest value in [2^13, 2^14)
        int ex;
        (void)frexpf(m, &ex);
        const float sc = m > 0.f ? ldexpf(1.0f, 14 - ex) : 1.0f;
        f16x8 b;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            b[2 * j] = (_Float16)(x[j].x * sc);
            b[2 * j + 1] = (_Float16)(x[j].y * sc);
        }
        cr = __builtin_amdgcn_mfma_f32_16x16x32_f16(W.re_h, b, cr, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f32_16x16x32_f16(W.im_h, b, ci, 0, 0, 0);
        const float inv = 1.0f / sc;
#pragma unroll
        for (int i = 0; i < 4; ++i) x[i] = make_float2(cr[i] * inv, ci[i] * inv);
    } else {
        bf16x8 bh, bl;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            bh[2 * j] = (__bf16)x[j].x;
            bh[2 * j + 1] = (__bf16)x[j].y;
            if (P == FFT_BF16X3) {
                bl[2 * j] = (__bf16)(x[j].x - (float)bh[2 * j]);
                bl[2 * j + 1] = (__bf16)(x[j].y - (float)bh[2 * j + 1]);
            }
        }
        cr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.re_hi, bh, cr, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.im_hi, bh, ci, 0, 0, 0);
        if (P == FFT_BF16X3) {
            cr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.re_hi, bl, cr, 0, 0, 0);
            ci = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.im_hi, bl, ci, 0, 0, 0);
            cr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.re_lo, bh, cr, 0, 0, 0);
            ci = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.im_lo, bh, ci, 0, 0, 0);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) x[i] = make_float2(cr[i], ci[i]);
    }
}

// Element index of register 4 g + j of this lane in the MFMA layout of a pass
// input: DFT d = 64 w + 16 g + n, point r = 4 kb + j, element d + 1024 r.
// t is threadIdx.x (a persistent caller passes an opaque copy, see row_tid).
__device__ __forceinline__ int mfma_elem(int g, int j, int t) {
    const int w = t >> 6, l = t & 63;
    return 64 * w + 16 * g + (l & 15) + 1024 * (4 * (l >> 4) + j);
}

// LDS words of the MFMA layouts, from one opaque base per use (lds_opaque), so
// that the compiler does not keep sixteen addresses of every pass live across
// the transform.  lds_re is XOR-linear over element bits 0-9 and passes bits
// 10-13 through doubled, so with lane (w, n, kb):
//   input element 64 w + 16 g + n + 1024 (4 kb + j)  ->  (rb ^ lds_re(16 g)) + 2048 j,
//     rb = lds_re(64 w + n) | 8192 kb
//   pass output, Ns = 1:    base lds_re(16 n + 4 kb) | 2048 w,              ^ lds_re(256 g) ^ lds_re(i)
//   pass output, Ns = 16:   base lds_re(64 kb + n) | 2048 w,                ^ lds_re(256 g) ^ lds_re(16 i)
//   pass output, Ns = 256:  base lds_re(64 (w & 3) + n) | 8192 (w >> 2) | 2048 kb,
//                                                                           ^ lds_re(16 g) ^ lds_re(256 i)
// (the constants are below 2048, the bases' high parts above).
__device__ __forceinline__ int mfma_rd_base(int t) {
    const int w = t >> 6, l = t & 63;
    return lds_opaque(lds_re(64 * w + (l & 15)) | (8192 * (l >> 4)));
}
__device__ __forceinline__ int mfma_rd(int rb, int g, int j) { return (rb ^ lds_re(16 * g)) + 2048 * j; }

// Radix-16 passes on the matrix cores, then the radix-4 VALU pass.
// v: the input in MFMA layout (v[4 g + j] = element mfma_elem(g, j));
// on return v[q] = element threadIdx.x + 1024 q of the transform.
template <int S, int P>
__device__ __forceinline__ void fft16k_mfma(float2 *v, float *lds, const FftTw *tw,
                                            const MfmaW<S, P> &W, int t) {
    const int w = t >> 6, l = t & 63, n = l & 15, kb = l >> 4;
#pragma unroll
    for (int pass = 0; pass < 3; ++pass) {
        const int ns = pass == 0 ? 1 : (pass == 1 ? 16 : 256);
        if (pass > 0) {
            // twiddle w_d^r, w_d = e^{S 2 pi i (d % Ns) / (16 Ns)}, r = 4 kb + j
            if (pass == 1) {
                // d % 16 = n for every group: four twiddles per lane
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const float2 tv = ctw<S>(tw->t256[n * (4 * kb + j)]);
#pragma unroll
                    for (int g = 0; g < 4; ++g) v[4 * g + j] = cmul(v[4 * g + j], tv);
                }
            } else {
                // w^(4 kb + j) = w^(4 kb) w^j with w = e^{S 2 pi i (d % 256)/4096}
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    const int m = (64 * w + 16 * g + n) & 255;
                    const int mb = m * 4 * kb;                   // < 4096
                    const float2 w1 = ctw<S>(cmul(tw->h4k[m >> 6], tw->l4k[m & 63]));
                    float2 tj = ctw<S>(cmul(tw->h4k[mb >> 6], tw->l4k[mb & 63]));
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        v[4 * g + j] = cmul(v[4 * g + j], tj);
                        tj = cmul(tj, w1);
                    }
                }
            }
        }
#pragma unroll
        for (int g = 0; g < 4; ++g) mfma_dft16<S, P>(v + 4 * g, W);
        __syncthreads();   // previous readers of the LDS array are done
        {
            // output k = 4 kb + i of DFT d = 64 w + 16 g + n goes to element
            // (d / Ns) 16 Ns + (d % Ns) + Ns k
            const int wb = lds_opaque(pass == 0   ? (lds_re(16 * n + 4 * kb) | (2048 * w))
                                      : pass == 1 ? (lds_re(64 * kb + n) | (2048 * w))
                                                  : (lds_re(64 * (w & 3) + n) | (8192 * (w >> 2)) | (2048 * kb)));
#pragma unroll
            for (int g = 0; g < 4; ++g)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int c = pass == 0   ? (lds_re(256 * g) ^ lds_re(i))
                                  : pass == 1 ? (lds_re(256 * g) ^ lds_re(16 * i))
                                              : (lds_re(16 * g) ^ lds_re(256 * i));
                    lds_put(lds, wb ^ c, v[4 * g + i]);
                }
        }
        __syncthreads();
        if (pass < 2) {
            const int rb = mfma_rd_base(t);
#pragma unroll
            for (int g = 0; g < 4; ++g)
#pragma unroll
                for (int j = 0; j < 4; ++j) v[4 * g + j] = lds_get(lds, mfma_rd(rb, g, j));
        }
    }
    // radix 4, Ns = 4096 (as in fft16k): element t + 1024 (b + 4 r) -> v[b + 4 r]
    const int rd = lds_re(t);
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = lds_get(lds, rd + 2048 * q);
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const uint32_t j = (uint32_t)(t + 1024 * b);
        const uint32_t j2 = 2u * j;   // < 8192
        const float2 w1 = ctw<S>(cmul(tw->h16k[j >> 7], tw->l16k[j & 127]));
        const float2 w2 = ctw<S>(cmul(tw->h16k[j2 >> 7], tw->l16k[j2 & 127]));
        v[b + 4] = cmul(v[b + 4], w1);
        v[b + 8] = cmul(v[b + 8], w2);
        v[b + 12] = cmul(v[b + 12], cmul(w1, w2));
        dft4<S>(v[b], v[b + 4], v[b + 8], v[b + 12]);
    }
}

}  // namespace sdbe
