// Band-trimmed rational resampler of the aphids GPU thread at a SWARM rate
// preset (software/aphids/vdif_inout_gpu_thread.cu:589-709, plans :269-333;
// presets sample_rates.h:21-63): per block of D snapshots
//   SwarmC2R  : inverse real DFT of 32768 points per snapshot (k_trim_c2r),
//   SwarmR2C  : real DFTs of R = 256 D samples over the concatenation, 128 segments,
//   R2dbeC2R  : inverse real DFT of P = 256 E points from bins skip .. skip + P/2
//               of each segment's spectrum (k_trim_seg),
// then per-block statistics (k_trim_stats), the threshold policy (k_thresh) and
// 2-bit packing in natural sample order (k_trim_pack).
//
// Both real transforms of a segment use the packed form: the R-point R2C is an
// R/2-point complex FFT of z[m] = t[2m] + i t[2m+1] followed by
//   X[k] = 1/2 (Z[k] + conj Z[R/2-k]) - 1/2 i e^{-2 pi i k/R} (Z[k] - conj Z[R/2-k]),
// and the P-point C2R an P/2-point inverse FFT of
//   W[j] = (Y[j] + conj Y[P/2-j]) + i e^{2 pi i j/P} (Y[j] - conj Y[P/2-j]),
// Y[j] = X[skip + j] with the imaginary parts of Y[0] and Y[P/2] dropped (numpy
// irfft semantics, as the rest of the stage).  The R/2- and P/2-point FFTs are
// mixed-radix Stockham passes in place in one LDS buffer:
//   6/11  (D 39, E 32):   4992 = 16 8 13 3,   4096 = 16 16 16
//   8/11  (D 13, E 8):    1664 = 16 8 13,     1024 = 16 8 8
//   10/11 (D 65, E 64):   8320 = 16 8 5 13,   8192 = 16 16 16 2
//   11/11 (D 143, E 128): 18304 = 16 8 11 13, 16384 = 16 16 16 4
//
// Units: the unnormalised chain is numpy's irfft / rfft / irfft times
// 32768 P, and 65536 P inside k_trim_seg, whose R2C post-processing leaves out
// its factor 1/2; k_trim_seg removes that exact power of two, so samples,
// statistics and thresholds are in numpy units.
#pragma once

namespace trim {

constexpr int NSEG = 128;             // segments per block (D 32768 / (256 D))
// Threads of a segment workgroup, sized to the butterfly counts of its passes
// (measured): 128 at 8/11 (13 KB buffer; passes of 64 to 208 butterflies), 256
// at 6/11 (39 KB, four workgroups per CU), 512 at 10/11 (66 KB, two per CU;
// radix-16 passes of 512 and 520 butterflies), 1024 at 11/11 (146 KB), whose
// C2R is the 16384-point transform of stages B and C2.
__host__ __device__ constexpr int seg_threads(int D) {
    return D > 100 ? 1024 : (D > 40 ? 512 : (D > 20 ? 256 : 128));
}

// Host: geometry of a preset.  Returns false for an unsupported preset or a
// skip that leaves fewer than P/2 + 1 bins.
struct Geometry {
    int frac = 0, D = 0, E = 0, R = 0, P = 0, skip = 0;
};
inline bool geometry(int rate_frac, int skip_mhz, Geometry &g) {
    int D, E;
    if (rate_frac == 6) {
        D = 39;
        E = 32;
    } else if (rate_frac == 8) {
        D = 13;
        E = 8;
    } else if (rate_frac == 10) {
        D = 65;
        E = 64;
    } else if (rate_frac == 11) {
        D = 143;
        E = 128;
    } else {
        return false;
    }
    g.frac = rate_frac;
    g.D = D;
    g.E = E;
    g.R = 256 * D;
    g.P = 256 * E;
    const double swarm_rate = 4576e6 * rate_frac / 11.0;
    g.skip = (int)(skip_mhz * 1e6 / (swarm_rate / g.R));   // vdif_inout_gpu_thread.cu:269
    return skip_mhz >= 0 && g.skip + g.P / 2 <= g.R / 2;
}

// LDS position of element e of a segment buffer: an XOR swizzle within blocks
// of 16 float2 (one 128-byte bank row).  The Stockham writes of a pass with
// Ns = 1 (out[b R + k]) and Ns = 16 would otherwise put 16 lanes on one bank
// group; with the swizzle every read and write of the passes spreads over the
// 16 groups (N is a multiple of 16 for every size here).
__device__ __forceinline__ int swz16(int e) { return e ^ ((e >> 4) & 15); }
// swz16(e + c) from sb = swz16(e) for a constant c that is a multiple of 128:
// adding c leaves bits 0-3 of e alone and adds 0 or 8 (c mod 256 = 0 or 128)
// to bits 4-7 modulo 16, an XOR with 8, so swz16(e + c) = (sb ^ 8 [c mod 256 =
// 128]) + c.  With c known at compile time the access is one of at most two
// base registers plus an immediate offset.
__device__ __forceinline__ int swz_add(int sb, int c) { return (sb ^ ((c & 128) >> 4)) + c; }

// DFT of odd prime length P by conjugate pairing, in place, sign S:
// X[d] = x0 + A_d + S i B_d and X[P-d] = x0 + A_d - S i B_d with
// A_d = sum_j (x[j] + x[P-j]) cos(2 pi j d/P), B_d = sum_j (x[j] - x[P-j]) sin(2 pi j d/P).
// c[m], s[m]: cos and sin of 2 pi m/P for m <= (P-1)/2.
template <int P, int S>
__device__ __forceinline__ void dft_pairs(float2 *x, const float *c, const float *s) {
    constexpr int H = (P - 1) / 2;
    float2 sp[H + 1], dm[H + 1];
    float2 tot = x[0];
#pragma unroll
    for (int j = 1; j <= H; ++j) {
        sp[j] = cadd(x[j], x[P - j]);
        dm[j] = csub(x[j], x[P - j]);
        tot = cadd(tot, sp[j]);
    }
    const float2 x0 = x[0];
    x[0] = tot;
#pragma unroll
    for (int d = 1; d <= H; ++d) {
        float2 A = x0, B = make_float2(0.f, 0.f);
#pragma unroll
        for (int j = 1; j <= H; ++j) {
            const int m = (j * d) % P;
            const float cm = m <= H ? c[m] : c[P - m];
            const float sm = m <= H ? s[m] : -s[P - m];
            A.x = fmaf(sp[j].x, cm, A.x);
            A.y = fmaf(sp[j].y, cm, A.y);
            B.x = fmaf(dm[j].x, sm, B.x);
            B.y = fmaf(dm[j].y, sm, B.y);
        }
        const float2 iB = mul_si<S>(B);
        x[d] = cadd(A, iB);
        x[P - d] = csub(A, iB);
    }
}
template <int S>
__device__ __forceinline__ void dft5(float2 *x) {
    const float c[3] = {1.0f, 3.0901699437e-01f, -8.0901699437e-01f};
    const float s[3] = {0.0f, 9.5105651630e-01f, 5.8778525229e-01f};
    dft_pairs<5, S>(x, c, s);
}
template <int S>
__device__ __forceinline__ void dft11(float2 *x) {
    const float c[6] = {1.0f, 8.4125353283e-01f, 4.1541501300e-01f, -1.4231483827e-01f, -6.5486073395e-01f,
                        -9.5949297361e-01f};
    const float s[6] = {0.0f, 5.4064081746e-01f, 9.0963199535e-01f, 9.8982144188e-01f, 7.5574957435e-01f,
                        2.8173255684e-01f};
    dft_pairs<11, S>(x, c, s);
}

// DFT of radix R in place, natural order in and out, sign S.
template <int R, int S>
__device__ __forceinline__ void dft_r(float2 *x) {
    if constexpr (R == 2) {
        dft2<S>(x[0], x[1]);
    } else if constexpr (R == 3) {
        dft3<S>(x[0], x[1], x[2]);
    } else if constexpr (R == 4) {
        dft4<S>(x[0], x[1], x[2], x[3]);
    } else if constexpr (R == 5) {
        dft5<S>(x);
    } else if constexpr (R == 8) {
        dft8<S>(x);
    } else if constexpr (R == 11) {
        dft11<S>(x);
    } else if constexpr (R == 13) {
        dft13<S>(x);
    } else {
        static_assert(R == 16, "radix");
        dft16<S>(x);
    }
}

// One radix-R Stockham pass of an N-point FFT: butterfly b < N/R reads
// in[b + (N/R) i], multiplies input i by w^(j i) (j = b mod Ns,
// w = e^{S 2 pi i/(Ns R)} = tab[N/(Ns R)], tab[m] = e^{S 2 pi i m/N}), takes
// the DFT-R and writes output k to out[(b / Ns) Ns R + j + Ns k].
// Sizes are compile-time, so the butterfly loop unrolls and every load of a
// thread (LDS data, and the table entries w, w^2, w^4, w^8 from which the other
// powers are products of at most four) is issued before the first is used.
// Loads of a partial last iteration read butterfly N/R - 1; only its writes are
// guarded.  In place: all reads of the pass happen before the barrier that
// precedes the first write.  The buffer is swizzled (swz16) on input and on
// output, or plain on output with LIN (a last pass: with Ns >= 64 its writes
// are contiguous over a wave).  Every stride N/R of the reads here is a multiple
// of 128 (swz_add); the writes' per-output offsets are immediate for Ns a
// multiple of 128 (swz_add) and an XOR with a constant for Ns = 1 and 16.
template <int N, int R, int S, int T>
__device__ __forceinline__ void pass_reads(const float2 *in, int b, float2 (&x)[R]) {
    constexpr int NB = N / R;
    if constexpr (NB % 128 == 0) {
        const int sb = swz16(b);
#pragma unroll
        for (int i = 0; i < R; ++i) x[i] = in[swz_add(sb, NB * i)];
    } else {
#pragma unroll
        for (int i = 0; i < R; ++i) x[i] = in[swz16(b + NB * i)];
    }
}
template <int N, int R, int NS, bool LIN>
__device__ __forceinline__ void pass_writes(float2 *out, int b, const float2 (&x)[R]) {
    const int base = (b / NS) * NS * R + (b % NS);
    if constexpr (LIN) {
#pragma unroll
        for (int k = 0; k < R; ++k) out[base + NS * k] = x[k];
    } else if constexpr (NS % 128 == 0) {
        const int sb = swz16(base);
#pragma unroll
        for (int k = 0; k < R; ++k) out[swz_add(sb, NS * k)] = x[k];
    } else if constexpr (NS == 16 && (R & (R - 1)) == 0) {
        // bits 4 .. 3 + log2 R of base are zero, so adding 16 k (k < R) sets them
        // without a carry: swz16(base + 16 k) = (swz16(base) ^ k) + 16 k
        const int sb = swz16(base);
#pragma unroll
        for (int k = 0; k < R; ++k) out[(sb ^ k) + 16 * k] = x[k];
    } else if constexpr (NS == 1 && R == 16) {
        const int sb = swz16(base);   // base = 16 b: swz16(16 b + k) = swz16(16 b) ^ k
#pragma unroll
        for (int k = 0; k < R; ++k) out[sb ^ k] = x[k];
    } else {
#pragma unroll
        for (int k = 0; k < R; ++k) out[swz16(base + NS * k)] = x[k];
    }
}
// Whether this thread's wave has a butterfly in iteration it of a pass of NB
// butterflies on T threads: in a partial last iteration the waves past NB skip
// the loads and the DFT (the condition is uniform over a wave, and the barriers
// stay outside it); the lanes of a partly used wave past NB still run on a
// clamped butterfly and do not write.  The issue slots a skipping wave gives up
// go to the other workgroups on its CU; a 1024-thread workgroup fills its CU
// alone, where the barrier waits for the busiest wave anyway (measured: 0.5%
// slower at 11/11 with the skip, 3.2% faster at 6/11).
template <int NB, int T>
__device__ __forceinline__ bool wave_has_work(int it) {
    if constexpr (NB % T == 0 || NB % T > T - 64 || T >= 1024) return true;   // no skipping wave
    const int w0 = __builtin_amdgcn_readfirstlane((int)threadIdx.x & ~63);
    return it < (NB + T - 1) / T - 1 || w0 + it * T < NB;
}
template <int N, int R, int NS, int S, int T, bool LIN = false>
__device__ __forceinline__ void pass(float2 *buf, const float2 *__restrict__ tab) {
    constexpr int NB = N / R, IT = (NB + T - 1) / T, STRIDE = N / (NS * R);
    static_assert(!LIN || NS >= 64, "plain output layout needs contiguous writes");
    float2 x[IT][R];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        if (!wave_has_work<NB, T>(it)) continue;
        const int b0 = threadIdx.x + it * T;
        const int b = (NB % T == 0 || it < IT - 1) ? b0 : (b0 < NB ? b0 : NB - 1);
        pass_reads<N, R, S, T>(buf, b, x[it]);
        if constexpr (NS > 1) {
            const int m = (b % NS) * STRIDE;
            float2 pw[16];
            pw[1] = tab[m];
            if constexpr (R > 2) pw[2] = tab[2 * m];
            if constexpr (R > 4) pw[4] = tab[4 * m];
            if constexpr (R > 8) pw[8] = tab[8 * m];
#pragma unroll
            for (int i = 3; i < R; ++i) {
                if (i == 4 || i == 8) continue;
                const int hi = i >= 8 ? 8 : (i >= 4 ? 4 : 2);
                pw[i] = cmul(pw[hi], pw[i - hi]);
            }
#pragma unroll
            for (int i = 1; i < R; ++i) x[it][i] = cmul(x[it][i], pw[i]);
        }
    }
    __syncthreads();   // every read of the pass is done
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        if (!wave_has_work<NB, T>(it)) continue;
        const int b = threadIdx.x + it * T;
        dft_r<R, S>(x[it]);
        if (NB % T == 0 || it < IT - 1 || b < NB) pass_writes<N, R, NS, LIN>(buf, b, x[it]);
    }
    __syncthreads();
}

// The R2C transform sizes after their first pass (pass1_global): radix passes in
// place, T threads.  The last pass writes the plain layout (element e at e),
// which the post-processing reads.
template <int N, int S, int T>
__device__ __forceinline__ void fft_tail(float2 *a, const float2 *__restrict__ tab) {
    if constexpr (N == 4992) {
        pass<N, 8, 16, S, T>(a, tab);
        pass<N, 13, 128, S, T>(a, tab);
        pass<N, 3, 1664, S, T, true>(a, tab);
    } else if constexpr (N == 1664) {
        pass<N, 8, 16, S, T>(a, tab);
        pass<N, 13, 128, S, T, true>(a, tab);
    } else if constexpr (N == 8320) {
        pass<N, 8, 16, S, T>(a, tab);
        pass<N, 5, 128, S, T>(a, tab);
        pass<N, 13, 640, S, T, true>(a, tab);
    } else {
        static_assert(N == 18304, "size");
        pass<N, 8, 16, S, T>(a, tab);
        pass<N, 11, 128, S, T>(a, tab);
        pass<N, 13, 1408, S, T, true>(a, tab);
    }
}

// The inverse (C2R) transform sizes split as first pass, fft_mid, last pass;
// LastPass gives the last pass's radix and Ns (Ns R = N: butterfly b writes
// outputs b + Ns k).
template <int N>
struct LastPass;
template <>
struct LastPass<1024> {
    static constexpr int R = 8, NS = 128;
};
template <>
struct LastPass<4096> {
    static constexpr int R = 16, NS = 256;
};
template <>
struct LastPass<8192> {
    static constexpr int R = 2, NS = 4096;
};
template <int N, int S, int T>
__device__ __forceinline__ void fft_mid(float2 *a, const float2 *__restrict__ tab) {
    if constexpr (N == 4096) {
        pass<N, 16, 16, S, T>(a, tab);
    } else if constexpr (N == 8192) {
        pass<N, 16, 16, S, T>(a, tab);
        pass<N, 16, 256, S, T>(a, tab);
    } else {
        static_assert(N == 1024, "size");
        pass<N, 8, 16, S, T>(a, tab);
    }
}

// The last pass of the transform with its outputs left in registers: thread t
// runs butterflies b = t + T it and holds output b + Ns k in x[it][k], i.e.
// element t + T (it + (Ns / T) k), the elements it stores next.  Reads and
// twiddles as pass<N, R, NS>; nothing is written to the buffer.
template <int N, int R, int NS, int S, int T>
__device__ __forceinline__ void pass_to_regs(const float2 *buf, const float2 *__restrict__ tab,
                                             float2 (&x)[N / R / T][R]) {
    constexpr int NB = N / R, IT = NB / T, STRIDE = N / (NS * R);
    static_assert(NB % T == 0 && NS * R == N && NS % T == 0, "last pass");
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int b = threadIdx.x + it * T;
        pass_reads<N, R, S, T>(buf, b, x[it]);
        const int m = (b % NS) * STRIDE;
        float2 pw[16];
        pw[1] = tab[m];
        if constexpr (R > 2) pw[2] = tab[2 * m];
        if constexpr (R > 4) pw[4] = tab[4 * m];
        if constexpr (R > 8) pw[8] = tab[8 * m];
#pragma unroll
        for (int i = 3; i < R; ++i) {
            if (i == 4 || i == 8) continue;
            const int hi = i >= 8 ? 8 : (i >= 4 ? 4 : 2);
            pw[i] = cmul(pw[hi], pw[i - hi]);
        }
#pragma unroll
        for (int i = 1; i < R; ++i) x[it][i] = cmul(x[it][i], pw[i]);
    }
#pragma unroll
    for (int it = 0; it < IT; ++it) dft_r<R, S>(x[it]);
}

// The first pass of fft<N> reading its input x (N points in global memory)
// directly: butterfly b takes x[b + (N/16) i] into registers, which the LDS
// staging of the input would only hand back.  Writes as pass<N, 16, 1>, then a
// barrier.  The buffer must not be in use.
template <int N, int S, int T>
__device__ __forceinline__ void pass1_global(float2 *a, const float2 *__restrict__ x) {
    constexpr int NB = N / 16, IT = (NB + T - 1) / T;
    float2 v[IT][16];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        if (!wave_has_work<NB, T>(it)) continue;
        const int b0 = threadIdx.x + it * T;
        const int b = (NB % T == 0 || it < IT - 1) ? b0 : (b0 < NB ? b0 : NB - 1);
#pragma unroll
        for (int i = 0; i < 16; ++i) v[it][i] = x[b + NB * i];
    }
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        if (!wave_has_work<NB, T>(it)) continue;
        const int b = threadIdx.x + it * T;
        dft_r<16, S>(v[it]);
        if (NB % T == 0 || it < IT - 1 || b < NB) pass_writes<N, 16, 1, false>(a, b, v[it]);
    }
    __syncthreads();
}

// The first pass of fft<N> when N / 16 == T and thread t already holds its
// butterfly's inputs, elements t + (N/16) i, in v[i].  A barrier precedes the
// writes (the caller's reads of the buffer finish first) and follows them.
template <int N, int S, int T>
__device__ __forceinline__ void pass1_regs(float2 *a, float2 (&v)[16]) {
    static_assert(N / 16 == T, "one butterfly per thread");
    dft_r<16, S>(v);
    __syncthreads();
    pass_writes<N, 16, 1, false>(a, (int)threadIdx.x, v);
    __syncthreads();
}

// Device tables of a preset (host-filled in FP64, rounded to FP32).
struct Tabs {
    const float2 *zr;     // e^{-2 pi i m/(R/2)}, m < R/2        (forward R/2-point FFT)
    const float2 *xr;     // e^{-2 pi i k/R}, k <= R/2            (R2C post-processing)
    const float2 *wp;     // e^{+2 pi i j/P}, j < P/2             (C2R pre-processing)
    const float2 *zp;     // e^{+2 pi i m/(P/2)}, m < P/2         (inverse P/2-point FFT)
};

// Low-index bits of the N = 32768 two-level twiddle table (tw<S>(tab, k) =
// hi[k >> 7] lo[k & 127]).
constexpr int N32K_LO = 7;

// The snapshot C2R's input elements k = t + 1024 q (q < 16) of thread t, for
// k_trim_c2r and k_trim_c2r_pk alike (so that both give the same bits).  From
// the values (offsets removed) of channel k, (Ra, Ia), and of its mirror N2 - k,
// (Rb, Ib), given by val(q, Ra, Ia, Rb, Ib): Z[k] = sum + i e^{2 pi i k/32768}
// dif (a_combine: the sums and differences as integers first), and for k = 0
// the DC real part only (the Nyquist bin is zero).  The twiddle is the product
// tw<+1>(tb.n32k, k) forms, hi[k >> 7] lo[k & 127]: lo[t & 127] is the same
// for every q, and hi[(t >> 7) + 8 q] for the 64 threads of a wave, which read
// it through the scalar cache.
template <typename Val>
__device__ __forceinline__ void c2r_inputs(const Tw2 &n32k, float2 (&v)[16], Val val) {
    const int t = threadIdx.x;
    const int hb = __builtin_amdgcn_readfirstlane(t >> N32K_LO);
    const F2v l = ((const ConstF2 *)n32k.lo)[t & ((1 << N32K_LO) - 1)];
    const float2 lo = make_float2(l.x, l.y);
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const F2v h = ((const ConstF2 *)n32k.hi)[hb + (1024 >> N32K_LO) * q];
        const float2 T = cmul(make_float2(h.x, h.y), lo);
        int Ra, Ia, Rb, Ib;
        val(q, Ra, Ia, Rb, Ib);
        v[q] = (q == 0 && t == 0) ? a_combine(Ra, 0, Ra, 0, T) : a_combine(Ra + Rb, Ia - Ib, Ra - Rb, Ia + Ib, T);
    }
}

}  // namespace trim

// SwarmC2R of one snapshot (workgroup = snapshot s of block bpl, 1024 threads):
// the packed 32768-point C2R of stage A's input (DC imaginary part dropped,
// Nyquist bin zero, LSB flip) into the block's series T[bpl][s][32768],
// unnormalised (numpy irfft times 32768).  (A persistent form with the next
// snapshot's int8 pairs prefetched during the last pass measured 8% slower.)
template <int D>
__global__ void __launch_bounds__(1024)
k_trim_c2r(const BpDesc *__restrict__ bpd, float2 *__restrict__ T, Tables tb) {
    __shared__ float lds[LDS_FFT_N];
    __shared__ FftTw ftw;
    fft_tw_load(&ftw, tb.fft);
    const int bpl = blockIdx.x / D, s = blockIdx.x - bpl * D, t = threadIdx.x;
    const GAS uint16_t *src = gptr(reinterpret_cast<const uint16_t *>(bpd[bpl].rows)) + (size_t)s * NCH;
    const int flip = bpd[bpl].flip;
    // channel ka = k (flipped: N - 1 - k) and its mirror kb = N - k (flipped:
    // k - 1; k = 0 pairs with the zero Nyquist bin and reads nothing used)
    const int da = flip ? -1024 : 1024;
    const int ea = flip ? NCH - 1 - t : t, eb = flip ? t - 1 : NCH - t;
    float2 v[16];
    trim::c2r_inputs(tb.n32k, v, [&](int q, int &Ra, int &Ia, int &Rb, int &Ib) {
        // only k = 0 (t = 0, q = 0) has its mirror outside the row
        const int kb = q == 0 ? min(max(eb, 0), NCH - 1) : eb - da * q;
        const uint16_t ua = src[ea + da * q], ub = src[kb];
        Ra = (int8_t)(ua >> 8);
        Ia = (int8_t)(ua & 0xFF);
        Rb = (int8_t)(ub >> 8);
        Ib = (int8_t)(ub & 0xFF);
    });
    fft16k_t<+1>(v, lds, &ftw);   // element fft16k_tord(t) + 1024 q in v[q]
    float2 *dst = T + ((size_t)bpl * D + s) * NCH + fft16k_tord(t);
#pragma unroll
    for (int q = 0; q < 16; ++q) dst[1024 * q] = v[q];
}

// SwarmC2R of one snapshot read from the group's packets (the int8 rows are
// not materialised): the same arithmetic as k_trim_c2r.  Snapshot S = D blk + s
// of the group sits in B-frame S >> 7.  Its 16384 channels' payload bytes are
// staged in LDS (aliasing the FFT work array, before the FFT uses it) so that
// channel c is byte 3 - (c & 3) of staged word c >> 2 in both geometries:
// * reference (G2 false): channel group g = c >> 3 is packet (cid 2 (g >> 4) +
//   (g & 1), fid (g >> 1) & 7), payload words 2 r, 2 r + 1 for row r = S & 127
//   (vdif_inout_gpu_thread.cu:442, 464-494): one 8-byte load per group;
// * 8 x 1024 (G2 true): packet (cid c >> 10, fid (S & 127) >> 3), payload row
//   S & 7, word (c & 1023) >> 2: one 16-byte load per thread.
// Both polarisations' codes share a byte (pol 0 in bits 4-7: re 6-7, im 4-5;
// pol 1 in bits 0-3); a missing packet is staged as 0xAA bytes, which decode to
// 0 + 0j.  The workgroups are persistent (one per CU) and take the snapshots in
// XCD-contiguous order in each round, so the snapshots of a block, whose words
// share cache lines, are read through one XCD's L2.
template <int D, bool G2>
__global__ void __launch_bounds__(1024)
k_trim_c2r_pk(const BpDesc *__restrict__ bpd, float2 *__restrict__ T, Tables tb, const uint32_t *__restrict__ pk,
              const int *__restrict__ map, int n_snap) {
    __shared__ float lds[LDS_FFT_N];
    __shared__ FftTw ftw;
    fft_tw_load(&ftw, tb.fft);
    const int t = threadIdx.x, G = (int)gridDim.x;
    uint32_t *stg = reinterpret_cast<uint32_t *>(lds);   // NCH / 4 words
    // Persistent: the workgroup takes snapshots L = xcd_contig(wg) + G r (in each
    // round the snapshots on one XCD are consecutive).  While a snapshot's
    // transform runs, the packet words of the next one are in flight, from map
    // entries read one snapshot earlier still.
    auto lookup = [&](int L, int (&p)[2]) {   // packet index of each staged unit
        const int bpl = L / D, S = D * bpd[bpl].blk + (L - bpl * D), frame = S >> 7;
        if constexpr (G2) {
            p[0] = map[(frame * G2_CIDS + (t >> 6)) * G2_FIDS + ((S & 127) >> 3)];
        } else {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int g = t + 1024 * h;
                p[h] = map[(frame * 256 + 2 * (g >> 4) + (g & 1)) * 8 + ((g >> 1) & 7)];
            }
        }
    };
    auto words = [&](int L, const int (&p)[2], uint4 &w) {   // the staged units (0xAA: missing)
        const int bpl = L / D, S = D * bpd[bpl].blk + (L - bpl * D);
        w = make_uint4(0xAAAAAAAAu, 0xAAAAAAAAu, 0xAAAAAAAAu, 0xAAAAAAAAu);
        if constexpr (G2) {
            if (p[0] != MAP_NONE)
                w = *reinterpret_cast<const uint4 *>(pk + (size_t)p[0] * G2_PKT_WORDS + 8 + 256 * (S & 7) + 4 * (t & 63));
        } else {
            if (p[0] != MAP_NONE) {
                const uint2 u = *reinterpret_cast<const uint2 *>(pk + (size_t)p[0] * PKT_WORDS + 8 + 2 * (S & 127));
                w.x = u.x;
                w.y = u.y;
            }
            if (p[1] != MAP_NONE) {
                const uint2 u = *reinterpret_cast<const uint2 *>(pk + (size_t)p[1] * PKT_WORDS + 8 + 2 * (S & 127));
                w.z = u.x;
                w.w = u.y;
            }
        }
    };
    int L = xcd_contig((int)blockIdx.x, G);
    int p[2] = {MAP_NONE, MAP_NONE}, pn[2] = {MAP_NONE, MAP_NONE};
    uint4 w;
    lookup(L, p);
    words(L, p, w);
    if (L + G < n_snap) lookup(L + G, pn);
    for (;;) {   // every wave takes the same snapshots: the barriers below are uniform
        const int bpl = L / D, s = L - bpl * D;
        const int flip = bpd[bpl].flip, pol = bpd[bpl].pol;
        // the staging area is the transform's work array: every wave must be done
        // with the previous snapshot's transform (its last exchange is wave-local)
        __syncthreads();
        if constexpr (G2) {
            reinterpret_cast<uint4 *>(stg)[t] = w;
        } else {
            reinterpret_cast<uint2 *>(stg)[t] = make_uint2(w.x, w.y);
            reinterpret_cast<uint2 *>(stg)[t + 1024] = make_uint2(w.z, w.w);
        }
        __syncthreads();
        const int pbit = pol == 0 ? 4 : 0;
        // channel ka = k (flipped: N - 1 - k) and its mirror kb = N - k (flipped:
        // k - 1; k = 0 pairs with the zero Nyquist bin and reads nothing used): the
        // staged words base +- 256 q, the same byte of each
        const int da = flip ? -256 : 256;
        const int ea = flip ? NCH - 1 - t : t, eb = flip ? t - 1 : NCH - t;
        const int wa0 = ea >> 2, wb0 = eb >> 2;   // eb = -1 (t = 0, flipped): word 256 q - 1 for q >= 1
        const uint32_t sha = 8 * (3 - (ea & 3)) + pbit, shb = 8 * (3 - (eb & 3)) + pbit;
        float2 v[16];
        trim::c2r_inputs(tb.n32k, v, [&](int q, int &Ra, int &Ia, int &Rb, int &Ib) {
            // only k = 0 (t = 0, q = 0) has its mirror outside the staged row
            const int wb = q == 0 ? min(max(wb0, 0), NCH / 4 - 1) : wb0 - da * q;
            const uint32_t a = stg[wa0 + da * q], b = stg[wb];
            // codes: im in the nibble's bits 0-1, re in bits 2-3; value = code - 2
            Ra = (int)__builtin_amdgcn_ubfe(a, sha + 2u, 2u) - 2;
            Ia = (int)__builtin_amdgcn_ubfe(a, sha, 2u) - 2;
            Rb = (int)__builtin_amdgcn_ubfe(b, shb + 2u, 2u) - 2;
            Ib = (int)__builtin_amdgcn_ubfe(b, shb, 2u) - 2;
        });
        __syncthreads();   // the staged words are read; the FFT reuses the array
        const int next = L + G;
        if (next < n_snap) {   // workgroup-uniform
            words(next, pn, w);
            if (next + G < n_snap) lookup(next + G, pn);
        }
        // keep the next snapshot's loads here, ahead of the transform
        __builtin_amdgcn_sched_barrier(0);
        fft16k_t<+1>(v, lds, &ftw);   // element fft16k_tord(t) + 1024 q in v[q]
        float2 *dst = T + ((size_t)bpl * D + s) * NCH + fft16k_tord(t);
#pragma unroll
        for (int q = 0; q < 16; ++q) dst[1024 * q] = v[q];
        if (next >= n_snap) break;
        L = next;
    }
}

// SwarmR2C + R2dbeC2R of segment g of block bpl (workgroup = segment,
// trim::seg_threads(D) threads).  Output: P samples (numpy units) at
// out[bpl] + g P, where out is bpd[bpl].out (float samples) when F32, else
// stage + bpl * 128 P; FP64 sums of x and x^2 of the segment at
// part[(bpl * 128 + g) * 2].
template <int D, int E, bool F32>
__global__ void __launch_bounds__(trim::seg_threads(D))
k_trim_seg(const float *__restrict__ T, const BpDesc *__restrict__ bpd, float *__restrict__ stage,
           double *__restrict__ part, trim::Tabs tt, int skip, const float2 *__restrict__ fft_tw) {
    constexpr int R = 256 * D, P = 256 * E, NR = R / 2, NP = P / 2;
    constexpr int NBUF = NR > NP ? NR : NP;
    constexpr int NT = trim::seg_threads(D);
    // 11/11: the C2R is the 16384-point transform of stages B and C2 on 1024
    // threads (registers and a 128 KB work array inside buf, its twiddles in LDS)
    constexpr bool F16K = NP == NCH && NT == 1024;
    static_assert(!F16K || NBUF * 2 >= LDS_FFT_N, "work array");
    // one in-place buffer: 39 KB at 6/11 (four workgroups per CU), 146 KB at 11/11
    __shared__ float2 buf[NBUF];
    __shared__ double red[2][NT / 64];
    const FftTw *ftwp = nullptr;
    if constexpr (F16K) {
        __shared__ FftTw ftw;
        fft_tw_load(&ftw, fft_tw);
        ftwp = &ftw;
    }
    const int bpl = blockIdx.x / trim::NSEG, g = blockIdx.x - bpl * trim::NSEG, t = threadIdx.x;
    const float2 *z = reinterpret_cast<const float2 *>(T + (size_t)bpl * D * 2 * NCH + (size_t)g * R);
    trim::pass1_global<NR, -1, NT>(buf, z);   // the input goes to registers, not through LDS
    trim::fft_tail<NR, -1, NT>(buf, tt.zr);   // Z in natural order, plain layout
    // bins skip .. skip + NP of the R2C, then the C2R pre-processing; the NP / 256
    // values of a thread stay in registers until every read of Z is done.
    // 2 X[k] = (Z[k] + conj Z[NR-k]) - i xr[k] (Z[k] - conj Z[NR-k]) (twice the
    // R2C spectrum: the factor 1/2 goes into the output scale, a power of two).
    // Element j = t + NT q needs k = skip + j and k' = skip + NP - j: the reads
    // Z[k], Z[NR-k], Z[k'], Z[NR-k'] and the table entries xr[k], xr[k'], wp[j]
    // are per-thread bases plus multiples of NT.  Only j = 0 (t = 0, q = 0) meets
    // the ends: Z[NR-k] with k = 0 and Z[k'] with k' = NR both read Z[0].
    auto X2 = [&](int ia, int ib, float2 w) {
        const float2 a = buf[ia], b = cconj(buf[ib]);
        return cadd(cadd(a, b), mul_si<-1>(cmul(w, csub(a, b))));
    };
    constexpr int JT = NP / NT;
    static_assert(NP % NT == 0, "C2R size");
    // k = ka + NT q, NR - k = kb + NT (JT - 1 - q), k' = kc + NT (JT - 1 - q),
    // NR - k' = kd + NT q: every offset from a base is a non-negative constant
    const int ka = skip + t, kc = skip + NT - t;
    const int kb = NR - NT * (JT - 1) - ka, kd = NR - NT * (JT - 1) - kc;
    const __amdgpu_buffer_rsrc_t rxr =
        __builtin_amdgcn_make_buffer_rsrc((void *)tt.xr, 0, (NR + 1) * (int)sizeof(float2), BUF_RSRC_WORD3);
    const __amdgpu_buffer_rsrc_t rwp =
        __builtin_amdgcn_make_buffer_rsrc((void *)tt.wp, 0, NP * (int)sizeof(float2), BUF_RSRC_WORD3);
    auto ld = [](__amdgpu_buffer_rsrc_t r, int vo, int so) {
        return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 0));
    };
    float2 wv[JT];
#pragma unroll
    for (int q = 0; q < JT; ++q) {
        const int qd = JT - 1 - q;
        int ib = kb + NT * qd, ic = kc + NT * qd;
        if (q == 0) {
            ib = ib == NR ? 0 : ib;
            ic = ic == NR ? 0 : ic;
        }
        float2 y = X2(ka + NT * q, ib, ld(rxr, 8 * ka, 8 * NT * q));
        float2 ym = X2(ic, kd + NT * q, ld(rxr, 8 * kc, 8 * NT * qd));
        if (q == 0 && t == 0) {
            y.y = 0.f;         // Y[0] and Y[P/2] enter with their real parts only
            ym.y = 0.f;
        }
        const float2 sum = cadd(y, cconj(ym)), dif = csub(y, cconj(ym));
        wv[q] = cadd(sum, mul_si<+1>(cmul(ld(rwp, 8 * t, 8 * NT * q), dif)));
    }
    // samples (2m, 2m+1) = (Re, Im) w[m] / (65536 P)
    const float sc = 1.0f / (65536.0f * (float)P);
    float2 *dst = reinterpret_cast<float2 *>((F32 ? (float *)bpd[bpl].out : stage + (size_t)bpl * trim::NSEG * P) +
                                             (size_t)g * P);
    // the segment's sums of x and x^2: FP32 over the thread's 2 JT samples,
    // then FP64 across threads and segments
    float s1f = 0.f, s2f = 0.f;
    if constexpr (F16K) {
        // wv[q] is element t + 1024 q, fft16k_t's input layout; it waits at its
        // first barrier until every read of Z is done
        fft16k_t<+1>(wv, reinterpret_cast<float *>(buf), ftwp);   // element fft16k_tord(t) + 1024 q in wv[q]
        float2 *d = dst + fft16k_tord(t);
#pragma unroll
        for (int q = 0; q < JT; ++q) {
            const float2 x = cscale(wv[q], sc);
            d[1024 * q] = x;
            s1f += x.x + x.y;
            s2f = fmaf(x.x, x.x, fmaf(x.y, x.y, s2f));
        }
    } else {
        if constexpr (NP / 16 == NT) {
            // 6/11, 10/11: wv[q] (element t + NT q) is butterfly t's input q of the
            // first pass, which runs on the registers
            trim::pass1_regs<NP, +1, NT>(buf, wv);
        } else {
            static_assert(NT % 128 == 0, "swz_add offsets");
            __syncthreads();
            const int sb = trim::swz16(t);
#pragma unroll
            for (int q = 0; q < JT; ++q) buf[trim::swz_add(sb, NT * q)] = wv[q];
            __syncthreads();
            trim::pass<NP, 16, 1, +1, NT>(buf, tt.zp);
        }
        trim::fft_mid<NP, +1, NT>(buf, tt.zp);
        // the last pass leaves element t + NT q in registers; stored in q order
        // through a buffer resource (the q offsets in the scalar operand)
        constexpr int LR = trim::LastPass<NP>::R, LNS = trim::LastPass<NP>::NS, QS = LNS / NT;
        float2 o[NP / LR / NT][LR];
        trim::pass_to_regs<NP, LR, LNS, +1, NT>(buf, tt.zp, o);
        const __amdgpu_buffer_rsrc_t rd =
            __builtin_amdgcn_make_buffer_rsrc((void *)dst, 0, NP * (int)sizeof(float2), BUF_RSRC_WORD3);
#pragma unroll
        for (int q = 0; q < JT; ++q) {
            const float2 x = cscale(o[q % QS][q / QS], sc);
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32v2, x), rd, 8 * t, 8 * NT * q, 0);
            s1f += x.x + x.y;
            s2f = fmaf(x.x, x.x, fmaf(x.y, x.y, s2f));
        }
    }
    double s1 = (double)s1f, s2 = (double)s2f;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        s1 += __shfl_xor(s1, off);
        s2 += __shfl_xor(s2, off);
    }
    if ((t & 63) == 0) {
        red[0][t >> 6] = s1;
        red[1][t >> 6] = s2;
    }
    __syncthreads();
    if (t == 0) {
        double a = 0.0, b = 0.0;
        for (int i = 0; i < NT / 64; ++i) {
            a += red[0][i];
            b += red[1][i];
        }
        part[((size_t)bpl * trim::NSEG + g) * 2] = a;
        part[((size_t)bpl * trim::NSEG + g) * 2 + 1] = b;
    }
}

// Mean and sigma (ddof 1) of each block from the segment sums: ms[b] = (mean, sigma).
__global__ void __launch_bounds__(128)
k_trim_stats(const double *__restrict__ part, int P, int bp0, double2 *__restrict__ ms, double *__restrict__ stats_out) {
    __shared__ double red[2][trim::NSEG];
    const int t = threadIdx.x, b = blockIdx.x;
    red[0][t] = part[((size_t)b * trim::NSEG + t) * 2];
    red[1][t] = part[((size_t)b * trim::NSEG + t) * 2 + 1];
    __syncthreads();
    for (int s = trim::NSEG / 2; s > 0; s >>= 1) {
        if (t < s) {
            red[0][t] += red[0][t + s];
            red[1][t] += red[1][t + s];
        }
        __syncthreads();
    }
    if (t == 0) {
        const double L = (double)trim::NSEG * P;
        const double mean = red[0][0] / L;
        double var = (red[1][0] - L * mean * mean) / (L - 1.0);
        if (var < 0.0) var = 0.0;
        const double sig = sqrt(var);
        ms[b] = make_double2(mean, sig);
        if (stats_out) {
            stats_out[(size_t)(bp0 + b) * 2] = mean;
            stats_out[(size_t)(bp0 + b) * 2 + 1] = sig;
        }
    }
}

// 2-bit codes of block bpl in natural sample order (vdif_inout_gpu_thread.cu:565-583:
// v = x - off; ((|v| >= th) | 2) ^ 3 (v < 0); sample i of a word at bits 2i),
// stream byte q to frame q / 8192 of the block's output.  Thread = one 32-bit word.
__global__ void __launch_bounds__(256)
k_trim_pack(const float *__restrict__ stage, const float2 *__restrict__ thr, const BpDesc *__restrict__ bpd, int P,
            OutDesc od) {
    const int words = trim::NSEG * P / 16;
    const int wpb = words / 256;                      // workgroups per block
    const int bpl = blockIdx.x / wpb, wi = (blockIdx.x - bpl * wpb) * 256 + threadIdx.x;
    const float4 *src = reinterpret_cast<const float4 *>(stage + (size_t)bpl * trim::NSEG * P) + (size_t)wi * 4;
    const float2 to = thr[bpl];
    uint32_t word = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float4 f = src[i];
        const float x[4] = {f.x, f.y, f.z, f.w};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float v = x[j] - to.y;
            const uint32_t c = ((((fabsf(v) >= to.x) ? 1u : 0u) | 2u) ^ (v < 0.f ? 3u : 0u)) & 3u;
            word |= c << (2 * (4 * i + j));
        }
    }
    const int64_t q = (int64_t)wi * 4;
    uint8_t *dst = bpd[bpl].out + (q / OUT_FRAME_PAYLOAD) * od.frame_bytes + od.hdr_bytes + (q % OUT_FRAME_PAYLOAD);
    *reinterpret_cast<uint32_t *>(dst) = word;
}
