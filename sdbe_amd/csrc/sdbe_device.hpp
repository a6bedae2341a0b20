// Device helpers for the gfx950 SWARM data back-end stage: complex arithmetic,
// twiddle tables, small DFTs (3, 4, 8, 13, 16, 39) in registers, and a
// 16384-point Stockham FFT held in registers of a 1024-thread workgroup with
// LDS exchanges between radix passes.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace sdbe {

constexpr int NCH = 16384;            // complex channels per snapshot (= N/2 of the 32768 C2R)
constexpr int NBLK = 39;              // snapshots per block
constexpr int MP = NBLK * NCH;        // 638,976: complex length of the packed 1,277,952 R2C
constexpr int P20 = 1 << 20;          // complex length of the packed 2,097,152 C2R
constexpr int KHI = 64;               // P20 = KHI * NCH
constexpr int FFT_THREADS = 1024;     // threads of the 16384-point FFT workgroups
constexpr int LDS_FFT_N = 2 * NCH;     // LDS floats of the work array of the 16384-point FFT

// LDS layout of a 16384-point complex work array (one float array of
// LDS_FFT_N words): rows of 64 elements, the real parts of row i in words
// 128 i .. 128 i + 63 and the imaginary parts in the next 64 words, so both
// parts of an element share one address (the imaginary part at +256 bytes,
// within a DS instruction's offset field).  Within its row an element sits at
// an XOR swizzle of its index,
//   lds_swz(e) = e ^ ((e >> 5) & 15) ^ ((e >> 4) & 48),
// which spreads every access pattern of the FFT passes (reads of t + 1024 r,
// writes of 16 t + r, 256 (t>>4) + (t&15) + 16 r and 4096 (t>>8) + (t&255) + 256 r)
// over 32 distinct banks per half-wave and 64 per wave, so none of them has a
// bank conflict.  lds_re(e) is the word of the real part; it is linear over
// GF(2) (lds_re(a ^ b) = lds_re(a) ^ lds_re(b)) and bits 10..13 of e pass
// through: lds_re(e + 1024 r) = lds_re(e) + 2048 r for e < 1024.
__host__ __device__ constexpr int lds_swz(int e) { return e ^ ((e >> 5) & 15) ^ ((e >> 4) & 48); }
__host__ __device__ constexpr int lds_re(int e) { return (lds_swz(e) & 63) | ((e & ~63) << 1); }
__device__ __forceinline__ void lds_put(float *lds, int p, float2 x) {
    lds[p] = x.x;
    lds[p + 64] = x.y;
}
__device__ __forceinline__ float2 lds_get(const float *lds, int p) { return make_float2(lds[p], lds[p + 64]); }

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by S*i (S = +1 or -1)
template <int S> __device__ __forceinline__ float2 mul_si(float2 a) {
    return S > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}
// multiply by the constant (c + S*i*s)
template <int S> __device__ __forceinline__ float2 cmulc(float2 a, float c, float s) {
    const float ss = S > 0 ? s : -s;
    return make_float2(a.x * c - a.y * ss, a.x * ss + a.y * c);
}

// Two-level table of exp(+2*pi*i*idx/N): value = hi[idx >> lo_bits] * lo[idx & mask].
// Both levels are computed in FP64 on the host and rounded to FP32.
struct Tw2 {
    const float2 *hi;
    const float2 *lo;
    int lo_bits;
};
// The tables are read through the constant address space: no kernel writes
// them, and so workgroup-uniform lookups stay scalar loads even after the
// kernel's own global stores (a persistent kernel's next row).
typedef float F2v __attribute__((ext_vector_type(2)));
typedef const __attribute__((address_space(4))) F2v ConstF2;
template <int S> __device__ __forceinline__ float2 tw(const Tw2 &t, uint32_t idx) {
    const ConstF2 *hi = (const ConstF2 *)t.hi, *lo = (const ConstF2 *)t.lo;
    const F2v h = hi[idx >> t.lo_bits], l = lo[idx & ((1u << t.lo_bits) - 1u)];
    float2 w = cmul(make_float2(h.x, h.y), make_float2(l.x, l.y));
    return S > 0 ? w : cconj(w);
}

// The same table at an index that is a multiple of 2^lo_bits (the low factor is
// exactly 1): hi[i] for idx = i << lo_bits, no product.
template <int S> __device__ __forceinline__ float2 tw_hi(const Tw2 &t, uint32_t i) {
    const F2v h = ((const ConstF2 *)t.hi)[i];
    const float2 w = make_float2(h.x, h.y);
    return S > 0 ? w : cconj(w);
}


struct Tables {
    const float2 *fft;    // FftTw image (640 entries) for the 16384-point FFT
    Tw2 n32k;    // N = 32768        (C2R pre-twiddle of the 32768-point stage)
    Tw2 nmp;     // N = 638,976      (inter-pass twiddle of the 1,277,952 R2C)
    Tw2 nm;      // N = 1,277,952    (R2C post-processing)
    Tw2 nl;      // N = 2,097,152    (C2R pre-processing of the final stage)
    Tw2 np;      // N = 1,048,576    (inter-pass twiddle of the final stage)
    const float2 *c1col;  // [2][16384]: e^{-2 pi i c/1,277,952}, e^{2 pi i c/2,097,152}
};

// ---------------------------------------------------------------- small DFTs
// Sign S: -1 forward (exp(-i...)), +1 inverse.  Inputs/outputs in natural order.
template <int S> __device__ __forceinline__ void dft2(float2 &a, float2 &b) {
    float2 t = a;
    a = cadd(t, b);
    b = csub(t, b);
}

template <int S> __device__ __forceinline__ void dft4(float2 &a, float2 &b, float2 &c, float2 &d) {
    float2 t0 = cadd(a, c), t1 = csub(a, c), t2 = cadd(b, d), t3 = mul_si<S>(csub(b, d));
    a = cadd(t0, t2);
    c = csub(t0, t2);
    b = cadd(t1, t3);
    d = csub(t1, t3);
}

template <int S> __device__ __forceinline__ void dft3(float2 &a, float2 &b, float2 &c) {
    const float s3 = 8.660254038e-01f;
    float2 sbc = cadd(b, c), dbc = csub(b, c);
    float2 m = make_float2(a.x - 0.5f * sbc.x, a.y - 0.5f * sbc.y);
    float2 r = mul_si<S>(cscale(dbc, s3));
    a = cadd(a, sbc);
    b = cadd(m, r);
    c = csub(m, r);
}

// DFT-8: n = 4*n1 + n2, k = k1 + 2*k2
template <int S> __device__ __forceinline__ void dft8(float2 *v) {
    const float h = 7.071067812e-01f;
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) dft2<S>(v[n2], v[4 + n2]);
    v[5] = cmulc<S>(v[5], h, h);
    v[6] = mul_si<S>(v[6]);
    v[7] = cmulc<S>(v[7], -h, h);
    dft4<S>(v[0], v[1], v[2], v[3]);
    dft4<S>(v[4], v[5], v[6], v[7]);
    float2 t[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) t[k] = v[4 * (k & 1) + (k >> 1)];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = t[k];
}

// DFT-16: n = 4*n1 + n2, k = k1 + 4*k2
// DFT16 as 4 x 4: radix-4 over n2, internal twiddles W16^(n2 k1), radix-4 over
// k1.  The twiddles are fused into the second radix-4 layer: with
// W = e^{i S pi/8}, c = cos(pi/8), tau = tan(pi/8), h = sqrt(1/2):
//   W^1 = c (1 + i S tau), W^3 = i S c (1 - i S tau), W^9 = -c (1 + i S tau),
//   W^2 = h (1 + i S), W^6 = h (-1 + i S), W^4 = i S,
// so each twiddled input is (1 +- i S tau) or (+-1 + i S) times the input (two
// FMAs or two adds), and the common factor c or h multiplies the sums of the
// butterfly inside FMAs: 144 operations in place of 160.
template <int S> __device__ __forceinline__ void dft16(float2 *v) {
    constexpr float sf = (float)S;
    const float c1 = 9.238795325e-01f, h = 7.071067812e-01f, tau = 4.142135624e-01f;
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) dft4<S>(v[n2], v[4 + n2], v[8 + n2], v[12 + n2]);
    dft4<S>(v[0], v[1], v[2], v[3]);
    // the second layer's radix-4 on (a, b, c, d) -> (t0 + t2, t1 + t3, t0 - t2, t1 - t3)
    // with t0, t1 = a +- c and t2, t3 = (b + d), i S (b - d) given as factor * (U, W)
    auto out = [](float2 &x0, float2 &x1, float2 &x2, float2 &x3, float2 t0, float2 t1, float f, float2 U,
                  float2 W) {
        x0 = make_float2(fmaf(f, U.x, t0.x), fmaf(f, U.y, t0.y));
        x2 = make_float2(fmaf(-f, U.x, t0.x), fmaf(-f, U.y, t0.y));
        x1 = make_float2(fmaf(f, W.x, t1.x), fmaf(f, W.y, t1.y));
        x3 = make_float2(fmaf(-f, W.x, t1.x), fmaf(-f, W.y, t1.y));
    };
    {   // k1 = 1: v4, W v5, W^2 v6, W^3 v7
        const float2 a = v[4], x5 = v[5], x6 = v[6], x7 = v[7];
        const float2 A = make_float2(fmaf(-sf * tau, x5.y, x5.x), fmaf(sf * tau, x5.x, x5.y));   // (1 + iS tau) v5
        const float2 B = make_float2(fmaf(sf * tau, x7.y, x7.x), fmaf(-sf * tau, x7.x, x7.y));   // (1 - iS tau) v7
        const float2 C = make_float2(x6.x - sf * x6.y, x6.y + sf * x6.x);                       // (1 + iS) v6
        const float2 t0 = make_float2(fmaf(h, C.x, a.x), fmaf(h, C.y, a.y));
        const float2 t1 = make_float2(fmaf(-h, C.x, a.x), fmaf(-h, C.y, a.y));
        const float2 U = make_float2(A.x - sf * B.y, A.y + sf * B.x);     // A + iS B
        const float2 Vd = make_float2(A.x + sf * B.y, A.y - sf * B.x);    // A - iS B
        out(v[4], v[5], v[6], v[7], t0, t1, c1, U, make_float2(-sf * Vd.y, sf * Vd.x));
    }
    {   // k1 = 2: v8, W^2 v9, W^4 v10, W^6 v11
        const float2 a = v[8], c = make_float2(-sf * v[10].y, sf * v[10].x);
        const float2 P = csub(v[9], v[11]), Q = cadd(v[9], v[11]);
        const float2 R2 = make_float2(P.x - sf * Q.y, P.y + sf * Q.x);    // P + iS Q
        const float2 R3 = make_float2(Q.x - sf * P.y, Q.y + sf * P.x);    // Q + iS P
        out(v[8], v[9], v[10], v[11], cadd(a, c), csub(a, c), h, R2, make_float2(-sf * R3.y, sf * R3.x));
    }
    {   // k1 = 3: v12, W^3 v13, W^6 v14, W^9 v15
        const float2 a = v[12], x13 = v[13], x14 = v[14], x15 = v[15];
        const float2 B = make_float2(fmaf(sf * tau, x13.y, x13.x), fmaf(-sf * tau, x13.x, x13.y));  // (1 - iS tau) v13
        const float2 A = make_float2(fmaf(-sf * tau, x15.y, x15.x), fmaf(sf * tau, x15.x, x15.y));  // (1 + iS tau) v15
        const float2 C = make_float2(-x14.x - sf * x14.y, sf * x14.x - x14.y);                      // (-1 + iS) v14
        const float2 t0 = make_float2(fmaf(h, C.x, a.x), fmaf(h, C.y, a.y));
        const float2 t1 = make_float2(fmaf(-h, C.x, a.x), fmaf(-h, C.y, a.y));
        const float2 U = make_float2(-sf * B.y - A.x, sf * B.x - A.y);    // iS B - A
        const float2 Vd = make_float2(A.x - sf * B.y, A.y + sf * B.x);    // iS B + A
        out(v[12], v[13], v[14], v[15], t0, t1, c1, U, make_float2(-sf * Vd.y, sf * Vd.x));
    }
    float2 t[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) t[k] = v[4 * (k & 3) + (k >> 2)];
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = t[k];
}

// cos / sin of 2*pi*m/13 for m = 0..12
__device__ __forceinline__ float c13(int m) {
    const float c[7] = {1.0f, 8.854560257e-01f, 5.680647467e-01f, 1.205366803e-01f,
                        -3.546048870e-01f, -7.485107482e-01f, -9.709418174e-01f};
    return m <= 6 ? c[m] : c[13 - m];
}
__device__ __forceinline__ float s13(int m) {
    const float s[7] = {0.0f, 4.647231720e-01f, 8.229838659e-01f, 9.927088741e-01f,
                        9.350162427e-01f, 6.631226582e-01f, 2.393156643e-01f};
    return m <= 6 ? s[m] : -s[13 - m];
}

// DFT-13 by conjugate pairing: X[d] = x0 + A_d + S*i*B_d, X[13-d] = x0 + A_d - S*i*B_d
template <int S> __device__ __forceinline__ void dft13(float2 *x) {
    float2 sp[7], dm[7];
    float2 tot = x[0];
#pragma unroll
    for (int j = 1; j <= 6; ++j) {
        sp[j] = cadd(x[j], x[13 - j]);
        dm[j] = csub(x[j], x[13 - j]);
        tot = cadd(tot, sp[j]);
    }
    float2 x0 = x[0];
    x[0] = tot;
#pragma unroll
    for (int d = 1; d <= 6; ++d) {
        float2 A = x0, B = make_float2(0.f, 0.f);
#pragma unroll
        for (int j = 1; j <= 6; ++j) {
            const int m = (j * d) % 13;
            A.x += sp[j].x * c13(m);
            A.y += sp[j].y * c13(m);
            B.x += dm[j].x * s13(m);
            B.y += dm[j].y * s13(m);
        }
        float2 iB = mul_si<S>(B);
        x[d] = cadd(A, iB);
        x[13 - d] = csub(A, iB);
    }
}

// DFT-39 by the prime-factor (Good-Thomas) mapping, 39 = 3 x 13 with 3 and 13
// coprime: input n = (13 n1 + 3 n2) mod 39, output k = (13 k1 + 27 k2) mod 39
// (13 = 1 mod 3 and 0 mod 13, 27 = 0 mod 3 and 1 mod 13), so that
// e^{S 2 pi i n k/39} = e^{S 2 pi i n1 k1/3} e^{S 2 pi i n2 k2/13}: three 13-point
// DFTs over n2, then 13 three-point DFTs over n1, and no twiddle factors.  In
// place; output in natural order.
template <int S> __device__ __forceinline__ void dft39(float2 *x) {
    float2 y[3][13];
#pragma unroll
    for (int n1 = 0; n1 < 3; ++n1)
#pragma unroll
        for (int n2 = 0; n2 < 13; ++n2) y[n1][n2] = x[(13 * n1 + 3 * n2) % 39];
#pragma unroll
    for (int n1 = 0; n1 < 3; ++n1) dft13<S>(y[n1]);
#pragma unroll
    for (int k2 = 0; k2 < 13; ++k2) {
        dft3<S>(y[0][k2], y[1][k2], y[2][k2]);
#pragma unroll
        for (int k1 = 0; k1 < 3; ++k1) x[(13 * k1 + 27 * k2) % 39] = y[k1][k2];
    }
}

// Workgroup `lin` of a 1D grid of `total` runs on XCD lin % 8 (round-robin
// dispatch).  The returned logical index is contiguous per XCD: logical
// neighbours share that XCD's L2 and are dispatched close together in time.
__device__ __forceinline__ int xcd_contig(int lin, int total) {
    const int x = lin & 7, slot = lin >> 3, q = total >> 3, r = total & 7;
    return x * q + (x < r ? x : r) + slot;
}

// ------------------------------------------- LDS twiddles of the 16384-point FFT
// exp(+2 pi i m / N) for the three twiddled passes, computed in FP64 on the
// host (640 entries, this order) and copied to LDS once per workgroup:
//   t256[m]            N = 256,   m < 256       (radix-16 pass, Ns = 16)
//   h4k[m>>6] l4k[m&63]   N = 4096,  m < 4096   (radix-16 pass, Ns = 256)
//   h16k[m>>7] l16k[m&127] N = 16384, m < 16384 (radix-4 pass, Ns = 4096)
struct FftTw {
    float2 t256[256];
    float2 h4k[64];
    float2 l4k[64];
    float2 h16k[128];
    float2 l16k[128];
};
constexpr int FFT_TW_N = 640;

__device__ __forceinline__ void fft_tw_load(FftTw *L, const float2 *g) {
    float2 *d = reinterpret_cast<float2 *>(L);
    for (int i = threadIdx.x; i < FFT_TW_N; i += blockDim.x) d[i] = g[i];
}

// Host: the FftTw image (FFT_TW_N entries, layout of FftTw), computed in FP64.
inline void fft_tw_fill(float2 *host) {
    const double two_pi = 6.283185307179586476925286766559;
    auto e = [&](double num, double den) {
        return make_float2((float)cos(two_pi * num / den), (float)sin(two_pi * num / den));
    };
    float2 *p = host;
    for (int m = 0; m < 256; ++m) *p++ = e(m, 256);
    for (int m = 0; m < 64; ++m) *p++ = e(64.0 * m, 4096);
    for (int m = 0; m < 64; ++m) *p++ = e(m, 4096);
    for (int m = 0; m < 128; ++m) *p++ = e(128.0 * m, 16384);
    for (int m = 0; m < 128; ++m) *p++ = e(m, 16384);
}
template <int S> __device__ __forceinline__ float2 ctw(float2 w) { return S > 0 ? w : cconj(w); }

// v[r] *= w^r (r = 1..15) given w, w^2, w^4, w^8; the other powers are
// products of at most four of them.
__device__ __forceinline__ void twiddle16_pow(float2 *v, float2 w1, float2 w2, float2 w4, float2 w8) {
    const float2 w3 = cmul(w1, w2), w5 = cmul(w1, w4), w6 = cmul(w2, w4), w7 = cmul(w3, w4);
    v[1] = cmul(v[1], w1);
    v[2] = cmul(v[2], w2);
    v[3] = cmul(v[3], w3);
    v[4] = cmul(v[4], w4);
    v[5] = cmul(v[5], w5);
    v[6] = cmul(v[6], w6);
    v[7] = cmul(v[7], w7);
    v[8] = cmul(v[8], w8);
    v[9] = cmul(v[9], cmul(w8, w1));
    v[10] = cmul(v[10], cmul(w8, w2));
    v[11] = cmul(v[11], cmul(w8, w3));
    v[12] = cmul(v[12], cmul(w8, w4));
    v[13] = cmul(v[13], cmul(w8, w5));
    v[14] = cmul(v[14], cmul(w8, w6));
    v[15] = cmul(v[15], cmul(w8, w7));
}

// -------------------------------------------------- 16384-point Stockham FFT
// 1024 threads; thread t holds elements t + 1024*q (q = 0..15) on entry and on
// exit (natural order both ways).  Passes: radix 16 (Ns = 1, 16, 256) then
// radix 4 (Ns = 4096), through the LDS work array (layout of lds_re).  By the
// linearity of lds_re every access is a per-thread base XOR a compile-time
// constant:
//   reads  of t + 1024 r                     -> lds_re(t) + 2048 r
//   pass-1 writes of 16 t + r                -> lds_re(16 t) ^ r
//   pass-2 writes of (t>>4) 256 + k + 16 r   -> lds_re((t>>4) 256 + k) ^ lds_re(16 r)   (k = t & 15)
//   pass-3 writes of (t>>8) 4096 + k + 256 r -> (lds_re((t>>8) 4096 + k) ^ (lds_re(256 r) & 63)) + 512 r
//                                               (k = t & 255; the base has bits 9..12 clear)
// Each pass computes its write addresses from an opaque copy of its base
// (lds_opaque), so the compiler does not keep sixteen addresses of every pass
// live across the whole transform.
__device__ __forceinline__ int lds_opaque(int x) {
    asm("" : "+v"(x));   // not volatile: a volatile asm counts as a memory clobber and
                         // would keep later uniform loads off the scalar unit
    return x;
}
// fft16k = fft16k_p123 (passes 1-3, ending with the pass-4 inputs read from
// LDS into v: the LDS array is free after the next barrier) + fft16k_p4
// (the radix-4 pass, registers only).  Callers may issue work that needs no
// LDS between the two.
// t is threadIdx.x; a persistent caller passes an opaque copy (row_tid) so
// that the compiler recomputes the per-thread addresses in every row instead
// of hoisting them out of the row loop and spilling them.
template <int S>
__device__ __forceinline__ void fft16k_p123(float2 *v, float *lds, const FftTw *tw, int t);
template <int S>
__device__ __forceinline__ void fft16k_p4(float2 *v, const FftTw *tw, int t);

template <int S>
__device__ __forceinline__ void fft16k(float2 *v, float *lds, const FftTw *tw) {
    fft16k_p123<S>(v, lds, tw, threadIdx.x);
    fft16k_p4<S>(v, tw, threadIdx.x);
}

// threadIdx.x behind a volatile asm: not loop-invariant to the compiler
__device__ __forceinline__ int row_tid() {
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
}

// fft16k_p123 = fft16k_p12 (passes 1 and 2, ending with the pass-2 outputs in
// LDS behind a barrier) + fft16k_p3 (pass 3, then the pass-4 inputs read into v).
template <int S>
__device__ __forceinline__ void fft16k_p12(float2 *v, float *lds, const FftTw *tw, int t);
template <int S>
__device__ __forceinline__ void fft16k_p3(float2 *v, float *lds, const FftTw *tw, int t);

template <int S>
__device__ __forceinline__ void fft16k_p123(float2 *v, float *lds, const FftTw *tw, int t) {
    fft16k_p12<S>(v, lds, tw, t);
    fft16k_p3<S>(v, lds, tw, t);
}

// The same transform with a wave-local third exchange: fft16k_p12, then
// fft16k_p3w (pass 3 and the exchange into pass 4, no workgroup barrier), then
// fft16k_p4 with t' = fft16k_tord(t).  Thread t ends with element
// fft16k_tord(t) + 1024 q in v[q].
//
// Thread t runs the pass-3 butterfly m = T = fft16k_tord(t) and the pass-4
// butterflies T + 1024 rho (rho = 0..3).  Every element that wave w reads or
// writes in passes 3 and 4 has bits 4..7 equal to w: the pass-3 inputs
// m + 1024 r, the pass-3 outputs 4096 (m >> 8) + (m & 255) + 256 r, and the
// pass-4 inputs T + 1024 q.  So the wave writes its pass-3 outputs into the
// LDS words it read its inputs from, and reads its pass-4 inputs back from
// them, with no other wave touching those words.  Element o of that exchange
// sits at lds_re(lds_fold(o)), where lds_fold XORs bits 12-13 into bits 8-9.
// The writes spread a wave's lanes over element bits 0-3 and 12-13, and the
// reads over bits 0-3 and 8-9, so both reach all 64 banks (and the 32 of each
// half-wave), as the pass-3 reads of lds_re(T + 1024 r) do.
__host__ __device__ constexpr int fft16k_tord(int t) {
    return 16 * (t >> 6) + (t & 15) + 256 * ((t >> 4) & 3);
}
__host__ __device__ constexpr int lds_fold(int o) { return o ^ ((o >> 4) & 0x300); }

// Orders this wave's LDS writes before its later LDS reads of words written by
// other lanes of the wave (a wave's LDS operations execute in order).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// pass 3 twiddles of butterfly k = m & 255: w^r, w = e^{S 2 pi i k / 4096}
template <int S>
__device__ __forceinline__ void fft16k_tw3(float2 *v, const FftTw *tw, int k) {
    const int m1 = k, m2 = 2 * k, m4 = 4 * k, m8 = 8 * k;   // < 4096
    twiddle16_pow(v, ctw<S>(cmul(tw->h4k[m1 >> 6], tw->l4k[m1 & 63])),
                  ctw<S>(cmul(tw->h4k[m2 >> 6], tw->l4k[m2 & 63])),
                  ctw<S>(cmul(tw->h4k[m4 >> 6], tw->l4k[m4 & 63])),
                  ctw<S>(cmul(tw->h4k[m8 >> 6], tw->l4k[m8 & 63])));
}

template <int S>
__device__ __forceinline__ void fft16k_p3w(float2 *v, float *lds, const FftTw *tw, int t) {
    const int T = fft16k_tord(t), k = T & 255;
    const int rd = lds_re(T);
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = lds_get(lds, rd + 2048 * r);
    fft16k_tw3<S>(v, tw, k);
    dft16<S>(v);
    // lds_re passes element bits 10-13 through (as 2 (e & ~1023)), so with
    // rt = lds_re(T) < 2048 every word below is a per-thread XOR base plus a
    // constant that fits the DS offset field.
    // Output r = 4 h + rho is element 4096 b + 1024 h + 256 rho + k (b = T >> 8),
    // folded to 4096 b + 1024 h + 256 (rho ^ b) + k: word
    // (rt ^ lds_re(256 rho)) + 8192 b + 2048 h.
    {
        const int rt = lds_opaque(rd), hb = 8192 * (T >> 8);
#pragma unroll
        for (int r = 0; r < 16; ++r) lds_put(lds, ((rt ^ lds_re(256 * (r & 3))) + hb) + 2048 * (r >> 2), v[r]);
    }
    wave_lds_sync();
    // Element T + 1024 q, q = 4 g + rho, folded to (T ^ 256 g) + 1024 rho + 4096 g:
    // word (rt ^ lds_re(256 g)) + 8192 g + 2048 rho.
    {
        const int rt = lds_opaque(rd);
#pragma unroll
        for (int q = 0; q < 16; ++q)
            v[q] = lds_get(lds, ((rt ^ lds_re(256 * (q >> 2))) + 8192 * (q >> 2)) + 2048 * (q & 3));
    }
}

// The whole transform in that form, for callers that store the result: thread t
// ends holding element fft16k_tord(t) + 1024 q in v[q].
template <int S>
__device__ __forceinline__ void fft16k_t(float2 *v, float *lds, const FftTw *tw) {
    fft16k_p12<S>(v, lds, tw, threadIdx.x);
    fft16k_p3w<S>(v, lds, tw, threadIdx.x);
    fft16k_p4<S>(v, tw, fft16k_tord(threadIdx.x));
}

template <int S>
__device__ __forceinline__ void fft16k_p12(float2 *v, float *lds, const FftTw *tw, int t) {
    const int rd = lds_re(t);
    // pass 1: radix 16, Ns = 1 (no twiddle)
    dft16<S>(v);
    __syncthreads();   // previous users of the LDS array are done
    {
        const int base = lds_opaque(lds_re(16 * t));
#pragma unroll
        for (int r = 0; r < 16; ++r) lds_put(lds, base ^ r, v[r]);
    }
    __syncthreads();
    // pass 2: radix 16, Ns = 16
    {
        const int k = t & 15;
#pragma unroll
        for (int r = 0; r < 16; ++r) v[r] = lds_get(lds, rd + 2048 * r);
        twiddle16_pow(v, ctw<S>(tw->t256[k]), ctw<S>(tw->t256[2 * k]), ctw<S>(tw->t256[4 * k]),
                      ctw<S>(tw->t256[8 * k]));
        dft16<S>(v);
        __syncthreads();
        const int base = lds_opaque(lds_re((t >> 4) * 256 + k));
#pragma unroll
        for (int r = 0; r < 16; ++r) lds_put(lds, base ^ lds_re(16 * r), v[r]);
        __syncthreads();
    }
}

template <int S>
__device__ __forceinline__ void fft16k_p3(float2 *v, float *lds, const FftTw *tw, int t) {
    const int rd = lds_re(t);
    // pass 3: radix 16, Ns = 256
    {
        const int k = t & 255;
#pragma unroll
        for (int r = 0; r < 16; ++r) v[r] = lds_get(lds, rd + 2048 * r);
        fft16k_tw3<S>(v, tw, k);
        dft16<S>(v);
        __syncthreads();
        const int base = lds_opaque(lds_re((t >> 8) * 4096 + k));
#pragma unroll
        for (int r = 0; r < 16; ++r) lds_put(lds, (base ^ (lds_re(256 * r) & 63)) + 512 * r, v[r]);
        __syncthreads();
    }
    // pass 4: radix 4, Ns = 4096; butterfly j = t + 1024*b reads/writes j + 4096*r,
    // i.e. element t + 1024*(b + 4r) -> register v[b + 4r]
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = lds_get(lds, rd + 2048 * q);
}

template <int S>
__device__ __forceinline__ void fft16k_p4(float2 *v, const FftTw *tw, int t) {
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
