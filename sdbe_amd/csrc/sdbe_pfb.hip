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
//   k_pfb_factor    once per plan, one thread per branch: banded Cholesky of
//                   P P^T (bandwidth 3, Toeplitz per branch), FP64; the factor's
//                   state is checkpointed every 16 rows
//   k_pfb_fspan, k_pfb_scan, k_pfb_ffin
//                   forward substitution, one thread per (branch, 256-row span):
//                   span transfers, a per-branch scan of the incoming states,
//                   the span from its true state
//   k_pfb_bspan, k_pfb_scan, k_pfb_bfin
//                   back substitution likewise (16-row chunks from the last, factor
//                   rows recomputed from the checkpoints), the final pass emitting
//                   the 4-tap projection x = P^T z
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
constexpr int SEG = 16;                   // factor checkpoint spacing, backward chunk
constexpr int SPAN = 256;                 // rows per thread in the span-parallel solves
constexpr int CKW = 6;                    // doubles per factor checkpoint
constexpr int SOLVE_WG = 64;

static_assert(NH == NCH, "the PFB reuses the 16384-point FFT");

// Forward PFB of spectrum i.  z[m] = y[2m] + i y[2m+1] of the folded block;
// Z = FFT_16384(z); X[k] = 1/2 (Z[k] + conj Z[N-k]) - 1/2 i e^{-2 pi i k/L} (Z[k] - conj Z[N-k]),
// X[N] = Re Z[0] - Im Z[0].
// Spectra are contiguous per XCD (xcd_contig): the workgroups in flight on an
// XCD fold neighbouring spectra, whose 4-block input windows overlap by 3
// blocks, so most block reads are hits in that XCD's L2.
__global__ void __launch_bounds__(1024)
k_pfb_fold_r2c(const float *__restrict__ x, const float *__restrict__ win, float2 *__restrict__ spec,
               const float2 *__restrict__ fft_img, const float2 *__restrict__ tw32k) {
    __shared__ float lds[LDS_FFT_N];
    __shared__ FftTw ftw;
    fft_tw_load(&ftw, fft_img);
    const int i = xcd_contig((int)blockIdx.x, (int)gridDim.x), t = threadIdx.x;
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
    fft16k_t<+1>(v, lds, &ftw);   // z[fft16k_tord(t) + 1024 q] in v[q]
    float2 *dst = reinterpret_cast<float2 *>(yh + (size_t)i * LB) + fft16k_tord(t);
    const float sc = 1.0f / (float)LB;
#pragma unroll
    for (int q = 0; q < 16; ++q) dst[1024 * q] = cscale(v[q], sc);
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

template <bool EDGE = true>
__device__ __forceinline__ Row chol_row(int i, double a0, double a1, double a2, double a3, const Row &r1,
                                        const Row &r2, const Row &r3) {
    // rows 0..2 see fewer band entries; EDGE false is for rows i >= 3
    const double e1 = !EDGE || i >= 1 ? a1 : 0.0, e2 = !EDGE || i >= 2 ? a2 : 0.0, e3 = !EDGE || i >= 3 ? a3 : 0.0;
    Row r;
    r.d3 = e3 * r3.inv0;
    r.d2 = (e2 - r.d3 * r2.d1) * r2.inv0;
    r.d1 = (e1 - r.d3 * r1.d2 - r.d2 * r1.d1) * r1.inv0;
    r.inv0 = rsqrt(a0 - r.d1 * r.d1 - r.d2 * r.d2 - r.d3 * r.d3);
    return r;
}

// The banded solves run span-parallel: thread (branch n, span s) covers rows
// s0 = 256 s .. s1 - 1 (s1 = min(s0 + 256, nblock)).  A substitution is a
// linear recurrence of order 3, so a span's output is its particular solution
// (zero incoming state) plus a 3 x 3 transfer of the incoming state; per branch,
// a scan over the spans composes the transfers into each span's true incoming
// state, and a final pass recomputes the span from it.  The factor rows are
// data-independent: the factor's state at every 16th row is checkpointed once,
// at plan creation (k_pfb_factor), and rows are recomputed from it.  Span and
// checkpoint layouts put the branch index innermost, [s][k][32768], so every
// access is coalesced.

// A factor row needs, of the three rows before it, r1 = {inv0, d1, d2},
// r2 = {inv0, d1} and r3 = {inv0}: ckpt[c][6][LB] holds them for rows
// 16 c - 1, 16 c - 2, 16 c - 3.
__global__ void __launch_bounds__(SOLVE_WG)
k_pfb_factor(int nck, const double *__restrict__ a, double *__restrict__ ckpt) {
    const int n = blockIdx.x * SOLVE_WG + threadIdx.x;
    const double a0 = a[n], a1 = a[LB + n], a2 = a[2 * LB + n], a3 = a[3 * LB + n];
    Row r1 = {1.0, 0.0, 0.0, 0.0}, r2 = r1, r3 = r1;
    for (int c = 0; c < nck; ++c) {
        double *ck = ckpt + (size_t)c * CKW * LB + n;
        const double st[CKW] = {r1.inv0, r1.d1, r1.d2, r2.inv0, r2.d1, r3.inv0};
#pragma unroll
        for (int k = 0; k < CKW; ++k) ck[(size_t)k * LB] = st[k];
        for (int k = 0; k < SEG; ++k) {
            const Row r = chol_row(c * SEG + k, a0, a1, a2, a3, r1, r2, r3);
            r3 = r2;
            r2 = r1;
            r1 = r;
        }
    }
}

// The six checkpoint doubles of chunk c, and the three rows before it.
struct Ckpt {
    double v[CKW];
};

__device__ __forceinline__ Ckpt load_ckpt(const double *__restrict__ ckpt, int c, int n) {
    const double *ck = ckpt + (size_t)c * CKW * LB + n;
    Ckpt r;
#pragma unroll
    for (int k = 0; k < CKW; ++k) r.v[k] = ck[(size_t)k * LB];
    return r;
}

__device__ __forceinline__ void ckpt_rows(const Ckpt &c, Row &r1, Row &r2, Row &r3) {
    r1 = Row{c.v[0], c.v[1], c.v[2], 0.0};
    r2 = Row{c.v[3], c.v[4], 0.0, 0.0};
    r3 = Row{c.v[5], 0.0, 0.0, 0.0};
}

__device__ __forceinline__ void load_band(const double *__restrict__ a, int n, double &a0, double &a1,
                                          double &a2, double &a3) {
    a0 = a[n];
    a1 = a[LB + n];
    a2 = a[2 * LB + n];
    a3 = a[3 * LB + n];
}

// A span pass runs either the particular solution (one chain, zero incoming
// state, the data as right-hand side) or the three responses to unit incoming
// states (no data).  The responses are data-independent and, for every span
// that is full, independent of the number of rows: they are computed once per
// plan for full spans.  The last span of a call may be partial, and its
// transfer is never used (forward: its outgoing state ends the scan; backward:
// its incoming state is zero).  Output [s][3 NC][LB], index 3 j + c = state
// component c of chain j.
template <int NC>
__device__ __forceinline__ void chains_init(double (&st)[NC][3]) {
#pragma unroll
    for (int j = 0; j < NC; ++j)
#pragma unroll
        for (int c = 0; c < 3; ++c) st[j][c] = (NC == 3 && j == c) ? 1.0 : 0.0;
}

template <int NC>
__device__ __forceinline__ void chains_store(double *__restrict__ out, int sp, int n, const double (&st)[NC][3]) {
    double *o = out + (size_t)sp * 3 * NC * LB + n;
#pragma unroll
    for (int j = 0; j < NC; ++j)
#pragma unroll
        for (int c = 0; c < 3; ++c) o[(size_t)(3 * j + c) * LB] = st[j][c];
}

// Chunks are 16 rows from a multiple of 16.  GEN handles the general chunk
// (the first rows of the matrix, a partial last chunk); the others take the
// path with no row tests.
template <bool GEN, int NC, bool HOM>
__device__ __forceinline__ void fchunk(int c0, int s1, const float (&y)[SEG], double a0, double a1, double a2,
                                       double a3, Row &r1, Row &r2, Row &r3, double (&st)[NC][3]) {
#pragma unroll
    for (int k = 0; k < SEG; ++k) {
        if (!GEN || c0 + k < s1) {
            const Row r = chol_row<GEN>(c0 + k, a0, a1, a2, a3, r1, r2, r3);
#pragma unroll
            for (int j = 0; j < NC; ++j) {
                const double rhs = HOM ? 0.0 : (double)y[k];
                const double ui = (rhs - r.d1 * st[j][0] - r.d2 * st[j][1] - r.d3 * st[j][2]) * r.inv0;
                st[j][2] = st[j][1];
                st[j][1] = st[j][0];
                st[j][0] = ui;
            }
            r3 = r2;
            r2 = r1;
            r1 = r;
        }
    }
}

// Forward L u = yh[., n]: per span the final state (u[s1-1], u[s1-2], u[s1-3])
// of the particular solution (HOM false, out = p[s][3][LB]) or of the
// responses to (u[s0-1], u[s0-2], u[s0-3]) = e_j (HOM true, out = M[s][9][LB]).
template <bool HOM>
__global__ void __launch_bounds__(SOLVE_WG)
k_pfb_fspan(const float *__restrict__ yh, int nblock, const double *__restrict__ a,
            const double *__restrict__ ckpt, double *__restrict__ out) {
    constexpr int NC = HOM ? 3 : 1;
    const int n = blockIdx.x * SOLVE_WG + threadIdx.x, sp = blockIdx.y, s0 = sp * SPAN;
    const int s1 = min(s0 + SPAN, nblock);
    double a0, a1, a2, a3;
    load_band(a, n, a0, a1, a2, a3);
    Row r1, r2, r3;
    ckpt_rows(load_ckpt(ckpt, s0 / SEG, n), r1, r2, r3);
    double st[NC][3];
    chains_init<NC>(st);
    for (int c0 = s0; c0 < s1; c0 += SEG) {
        float y[SEG];
#pragma unroll
        for (int k = 0; k < SEG; ++k) y[k] = HOM ? 0.f : yh[(size_t)min(c0 + k, nblock - 1) * LB + n];
        if (c0 >= SEG && c0 + SEG <= s1)
            fchunk<false, NC, HOM>(c0, s1, y, a0, a1, a2, a3, r1, r2, r3, st);
        else
            fchunk<true, NC, HOM>(c0, s1, y, a0, a1, a2, a3, r1, r2, r3, st);
    }
    chains_store<NC>(out, sp, n, st);
}

// Per branch: the incoming state of every span, state[s][3][LB], from the
// particular states p[s][3][LB] and the transfers M[s][9][LB].
// Forward (dir +1): state_0 = 0, state_{s+1} = M_s state_s + p_s;
// backward (dir -1): from the last span down.
__global__ void __launch_bounds__(SOLVE_WG)
k_pfb_scan(const double *__restrict__ pst, const double *__restrict__ mt, int nspan, int dir,
           double *__restrict__ state) {
    const int n = blockIdx.x * SOLVE_WG + threadIdx.x;
    double x0 = 0.0, x1 = 0.0, x2 = 0.0;
    for (int i = 0; i < nspan; ++i) {
        const int sp = dir > 0 ? i : nspan - 1 - i;
        double *st = state + (size_t)sp * 3 * LB + n;
        st[0] = x0;
        st[(size_t)LB] = x1;
        st[(size_t)2 * LB] = x2;
        const double *pp = pst + (size_t)sp * 3 * LB + n;
        const double *t = mt + (size_t)sp * 9 * LB + n;
        double m[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) m[k] = t[(size_t)k * LB];
        const double y0 = pp[0] + m[0] * x0 + m[3] * x1 + m[6] * x2;
        const double y1 = pp[(size_t)LB] + m[1] * x0 + m[4] * x1 + m[7] * x2;
        const double y2 = pp[(size_t)2 * LB] + m[2] * x0 + m[5] * x1 + m[8] * x2;
        x0 = y0;
        x1 = y1;
        x2 = y2;
    }
}

template <bool GEN>
__device__ __forceinline__ void ffchunk(int c0, int s1, const float (&y)[SEG], double a0, double a1, double a2,
                                        double a3, Row &r1, Row &r2, Row &r3, double (&st)[3]) {
#pragma unroll
    for (int k = 0; k < SEG; ++k) {
        if (!GEN || c0 + k < s1) {
            const Row r = chol_row<GEN>(c0 + k, a0, a1, a2, a3, r1, r2, r3);
            const double ui = ((double)y[k] - r.d1 * st[0] - r.d2 * st[1] - r.d3 * st[2]) * r.inv0;
            st[2] = st[1];
            st[1] = st[0];
            st[0] = ui;
            r3 = r2;
            r2 = r1;
            r1 = r;
        }
    }
}

// Forward, final pass: the span from its true incoming state.  u itself is not
// stored: the state entering every 16-row chunk, uck[c][3][LB] = (u[16c-1],
// u[16c-2], u[16c-3]), is, and the backward passes recompute a chunk's u from it
// and yh along with the chunk's factor rows.
__global__ void __launch_bounds__(SOLVE_WG)
k_pfb_ffin(const float *__restrict__ yh, int nblock, const double *__restrict__ a,
           const double *__restrict__ ckpt, const double *__restrict__ state, double *__restrict__ uck) {
    const int n = blockIdx.x * SOLVE_WG + threadIdx.x, sp = blockIdx.y, s0 = sp * SPAN;
    const int s1 = min(s0 + SPAN, nblock);
    double a0, a1, a2, a3;
    load_band(a, n, a0, a1, a2, a3);
    Row r1, r2, r3;
    ckpt_rows(load_ckpt(ckpt, s0 / SEG, n), r1, r2, r3);
    const double *st = state + (size_t)sp * 3 * LB + n;
    double st3[3] = {st[0], st[(size_t)LB], st[(size_t)2 * LB]};   // u[i-1], u[i-2], u[i-3]
    for (int c0 = s0;; c0 += SEG) {
        double *uc = uck + (size_t)(c0 / SEG) * 3 * LB + n;
        uc[0] = st3[0];
        uc[(size_t)LB] = st3[1];
        uc[(size_t)2 * LB] = st3[2];
        if (c0 + SEG >= s1) break;   // the span's last chunk feeds no later chunk
        float y[SEG];
#pragma unroll
        for (int k = 0; k < SEG; ++k) y[k] = yh[(size_t)(c0 + k) * LB + n];
        if (c0 >= SEG)
            ffchunk<false>(c0, s1, y, a0, a1, a2, a3, r1, r2, r3, st3);
        else
            ffchunk<true>(c0, s1, y, a0, a1, a2, a3, r1, r2, r3, st3);
    }
}

// Factor rows 16 c + k0 .. 16 c + k0 + K - 1 into Lx[k0 ..] from checkpoint
// c's rows (k0 = 0, K = 16: the chunk; rows after it are the next chunk's first).
template <int K, bool GEN = true>
__device__ __forceinline__ void chunk_rows(const Ckpt &ck, int row0, double a0, double a1, double a2, double a3,
                                           Row *Lx) {
    Row r1, r2, r3;
    ckpt_rows(ck, r1, r2, r3);
#pragma unroll
    for (int k = 0; k < K; ++k) {
        Lx[k] = chol_row<GEN>(row0 + k, a0, a1, a2, a3, r1, r2, r3);
        r3 = r2;
        r2 = r1;
        r1 = Lx[k];
    }
}

// The chunk's u (forward substitution from the state entering the chunk, with
// the factor rows just recomputed), for the backward passes.
__device__ __forceinline__ void chunk_u(const Row (&Lx)[SEG + 3], const float (&y)[SEG], const double (&us)[3],
                                        double (&uu)[SEG]) {
    double w1 = us[0], w2 = us[1], w3 = us[2];
#pragma unroll
    for (int k = 0; k < SEG; ++k) {
        const double ui = ((double)y[k] - Lx[k].d1 * w1 - Lx[k].d2 * w2 - Lx[k].d3 * w3) * Lx[k].inv0;
        uu[k] = ui;
        w3 = w2;
        w2 = w1;
        w1 = ui;
    }
}

template <bool GEN, int NC, bool HOM>
__device__ __forceinline__ void bchunk(const Ckpt &ck, int c0, int s1, const float (&y)[SEG], const double (&us)[3],
                                       double a0, double a1, double a2, double a3, Row (&Lx)[SEG + 3],
                                       double (&st)[NC][3]) {
    chunk_rows<SEG, GEN>(ck, c0, a0, a1, a2, a3, Lx);
    double uu[SEG];
    if constexpr (!HOM) chunk_u(Lx, y, us, uu);
#pragma unroll
    for (int k = SEG - 1; k >= 0; --k) {
        if (!GEN || c0 + k < s1) {
#pragma unroll
            for (int j = 0; j < NC; ++j) {
                double rhs = 0.0;
                if constexpr (!HOM) rhs = uu[k];
                const double zi = (rhs - Lx[k + 1].d1 * st[j][0] - Lx[k + 2].d2 * st[j][1] -
                                   Lx[k + 3].d3 * st[j][2]) * Lx[k].inv0;
                st[j][2] = st[j][1];
                st[j][1] = st[j][0];
                st[j][0] = zi;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) Lx[SEG + k] = Lx[k];
}

// Backward L^T z = u: per span, in 16-row chunks from the last, the final state
// (z[s0], z[s0+1], z[s0+2]) of the particular solution (HOM false, out =
// p[s][3][LB]) or of the responses to (z[s1], z[s1+1], z[s1+2]) = e_j (HOM
// true, out = M[s][9][LB]).  Row i uses L(i+1, i), L(i+2, i), L(i+3, i):
// chunk c keeps its rows in Lx[0..15] and the three following rows in
// Lx[16..18].  The next chunk's checkpoint is loaded one chunk ahead.
template <bool HOM>
__global__ void __launch_bounds__(SOLVE_WG)
k_pfb_bspan(const float *__restrict__ yh, const double *__restrict__ uck, int nblock, const double *__restrict__ a,
            const double *__restrict__ ckpt, double *__restrict__ out) {
    constexpr int NC = HOM ? 3 : 1;
    const int n = blockIdx.x * SOLVE_WG + threadIdx.x, sp = blockIdx.y, s0 = sp * SPAN;
    const int s1 = min(s0 + SPAN, nblock);
    const int chi = (s1 - 1) / SEG, clo = s0 / SEG;
    double a0, a1, a2, a3;
    load_band(a, n, a0, a1, a2, a3);
    Row Lx[SEG + 3];
    chunk_rows<3>(load_ckpt(ckpt, chi + 1, n), (chi + 1) * SEG, a0, a1, a2, a3, Lx + SEG);
    Ckpt ck = load_ckpt(ckpt, chi, n);
    double st[NC][3];
    chains_init<NC>(st);
    for (int c = chi; c >= clo; --c) {
        const int c0 = c * SEG;
        float y[SEG];
        double us[3] = {0.0, 0.0, 0.0};
        if constexpr (!HOM) {
#pragma unroll
            for (int k = 0; k < SEG; ++k) y[k] = yh[(size_t)min(c0 + k, nblock - 1) * LB + n];
            const double *uc = uck + (size_t)c * 3 * LB + n;
            us[0] = uc[0];
            us[1] = uc[(size_t)LB];
            us[2] = uc[(size_t)2 * LB];
        }
        const Ckpt ckn = load_ckpt(ckpt, max(c - 1, clo), n);
        if (c0 >= SEG && c0 + SEG <= s1)
            bchunk<false, NC, HOM>(ck, c0, s1, y, us, a0, a1, a2, a3, Lx, st);
        else
            bchunk<true, NC, HOM>(ck, c0, s1, y, us, a0, a1, a2, a3, Lx, st);
        ck = ckn;
    }
    chains_store<NC>(out, sp, n, st);
}

template <bool GEN>
__device__ __forceinline__ void bfchunk(const Ckpt &ck, int c0, int s1, const float (&y)[SEG], const double (&us)[3],
                                        double a0, double a1, double a2, double a3, const double (&w)[4],
                                        Row (&Lx)[SEG + 3], double (&z)[3], float *__restrict__ xc) {
    chunk_rows<SEG, GEN>(ck, c0, a0, a1, a2, a3, Lx);
    double uu[SEG];
    chunk_u(Lx, y, us, uu);
#pragma unroll
    for (int k = SEG - 1; k >= 0; --k) {
        if (!GEN || c0 + k < s1) {
            const double zi = (uu[k] - Lx[k + 1].d1 * z[0] - Lx[k + 2].d2 * z[1] - Lx[k + 3].d3 * z[2]) * Lx[k].inv0;
            xc[(size_t)k * LB] = (float)(w[0] * z[2] + w[1] * z[1] + w[2] * z[0] + w[3] * zi);
            z[2] = z[1];
            z[1] = z[0];
            z[0] = zi;
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) Lx[SEG + k] = Lx[k];
}

// Backward, final pass: z of the span from its true incoming state and the
// 4-tap projection x = P^T z: x[i + 3] = w0 z[i+3] + w1 z[i+2] + w2 z[i+1] + w3 z[i]
// for the span's rows i; span 0 also emits x[0..2] (z[i] = 0 for i < 0).
__global__ void __launch_bounds__(SOLVE_WG)
k_pfb_bfin(const float *__restrict__ yh, const double *__restrict__ uck, int nblock, const double *__restrict__ a,
           const double *__restrict__ ckpt, const double *__restrict__ state, const float *__restrict__ win,
           float *__restrict__ x) {
    const int n = blockIdx.x * SOLVE_WG + threadIdx.x, sp = blockIdx.y, s0 = sp * SPAN;
    const int s1 = min(s0 + SPAN, nblock);
    const int chi = (s1 - 1) / SEG, clo = s0 / SEG;
    double a0, a1, a2, a3;
    load_band(a, n, a0, a1, a2, a3);
    const double w[4] = {win[n], win[LB + n], win[2 * LB + n], win[3 * LB + n]};
    Row Lx[SEG + 3];
    chunk_rows<3>(load_ckpt(ckpt, chi + 1, n), (chi + 1) * SEG, a0, a1, a2, a3, Lx + SEG);
    Ckpt ck = load_ckpt(ckpt, chi, n);
    const double *st = state + (size_t)sp * 3 * LB + n;
    double z[3] = {st[0], st[(size_t)LB], st[(size_t)2 * LB]};   // z[i+1], z[i+2], z[i+3]
    for (int c = chi; c >= clo; --c) {
        const int c0 = c * SEG;
        float y[SEG];
#pragma unroll
        for (int k = 0; k < SEG; ++k) y[k] = yh[(size_t)min(c0 + k, nblock - 1) * LB + n];
        const double *uc = uck + (size_t)c * 3 * LB + n;
        const double us[3] = {uc[0], uc[(size_t)LB], uc[(size_t)2 * LB]};
        const Ckpt ckn = load_ckpt(ckpt, max(c - 1, clo), n);
        float *xc = x + (size_t)(c0 + 3) * LB + n;
        if (c0 >= SEG && c0 + SEG <= s1)
            bfchunk<false>(ck, c0, s1, y, us, a0, a1, a2, a3, w, Lx, z, xc);
        else
            bfchunk<true>(ck, c0, s1, y, us, a0, a1, a2, a3, w, Lx, z, xc);
        ck = ckn;
    }
    if (sp == 0) {
        x[(size_t)2 * LB + n] = (float)(w[0] * z[2] + w[1] * z[1] + w[2] * z[0]);
        x[(size_t)LB + n] = (float)(w[0] * z[1] + w[1] * z[0]);
        x[n] = (float)(w[0] * z[0]);
    }
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
    double *uck = nullptr;        // [chunks][3][32768] forward state entering every 16-row chunk
    double *ckpt = nullptr;       // [chunks][6][32768] factor state at every 16th row
    double *pst = nullptr;        // [spans][3][32768] particular outgoing states
    double *mf = nullptr;         // [spans][9][32768] forward transfers of full spans
    double *mb = nullptr;         // [spans][9][32768] backward transfers of full spans
    double *state = nullptr;      // [spans][3][32768] incoming state of every span
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
    (void)hipFree(p->uck);
    (void)hipFree(p->ckpt);
    (void)hipFree(p->pst);
    (void)hipFree(p->mf);
    (void)hipFree(p->mb);
    (void)hipFree(p->state);
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
    // one transfer per span; checkpoints up to the chunk after the last full span
    const size_t nspan = (size_t)(max_spectra + SPAN - 1) / SPAN;
    const size_t nck = nspan * (SPAN / SEG) + 1;
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
    PALLOC(p->uck, nck * 3 * LB * sizeof(double));
    PALLOC(p->ckpt, nck * CKW * LB * sizeof(double));
    PALLOC(p->pst, nspan * 3 * LB * sizeof(double));
    PALLOC(p->mf, nspan * 9 * LB * sizeof(double));
    PALLOC(p->mb, nspan * 9 * LB * sizeof(double));
    PALLOC(p->state, nspan * 3 * LB * sizeof(double));
#undef PALLOC
    if (hipMemcpy(p->fft_img, img.data(), FFT_TW_N * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->tw32k, tw.data(), NH * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->win, h_window, (size_t)NT * LB * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->a, band.data(), (size_t)NT * LB * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        rc = SDBE_ERR_HIP;
        goto fail;
    }
    // the factor's checkpoints and the full spans' transfers depend on the window only
    hipLaunchKernelGGL(k_pfb_factor, dim3(LB / SOLVE_WG), dim3(SOLVE_WG), 0, 0, (int)nck, (const double *)p->a,
                       p->ckpt);
    hipLaunchKernelGGL(k_pfb_fspan<true>, dim3(LB / SOLVE_WG, nspan), dim3(SOLVE_WG), 0, 0, (const float *)nullptr,
                       (int)(nspan * SPAN), (const double *)p->a, (const double *)p->ckpt, p->mf);
    hipLaunchKernelGGL(k_pfb_bspan<true>, dim3(LB / SOLVE_WG, nspan), dim3(SOLVE_WG), 0, 0, (const float *)nullptr,
                       (const double *)nullptr, (int)(nspan * SPAN), (const double *)p->a, (const double *)p->ckpt,
                       p->mb);
    if (hipDeviceSynchronize() != hipSuccess) {
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
    const int nspan = (n_spectra + SPAN - 1) / SPAN;
    const dim3 grid(LB / SOLVE_WG, nspan);
    const double *ck = p->ckpt, *a = p->a;
    hipLaunchKernelGGL(k_pfb_fspan<false>, grid, dim3(SOLVE_WG), 0, s, (const float *)p->yh, (int)n_spectra, a, ck,
                       p->pst);
    hipLaunchKernelGGL(k_pfb_scan, dim3(LB / SOLVE_WG), dim3(SOLVE_WG), 0, s, (const double *)p->pst, (const double *)p->mf, nspan, 1,
                       p->state);
    hipLaunchKernelGGL(k_pfb_ffin, grid, dim3(SOLVE_WG), 0, s, (const float *)p->yh, (int)n_spectra, a, ck,
                       (const double *)p->state, p->uck);
    hipLaunchKernelGGL(k_pfb_bspan<false>, grid, dim3(SOLVE_WG), 0, s, (const float *)p->yh,
                       (const double *)p->uck, (int)n_spectra, a, ck, p->pst);
    hipLaunchKernelGGL(k_pfb_scan, dim3(LB / SOLVE_WG), dim3(SOLVE_WG), 0, s, (const double *)p->pst, (const double *)p->mb, nspan, -1,
                       p->state);
    hipLaunchKernelGGL(k_pfb_bfin, grid, dim3(SOLVE_WG), 0, s, (const float *)p->yh, (const double *)p->uck,
                       (int)n_spectra, a, ck, (const double *)p->state, (const float *)p->win, d_x);
    PFB_TRY(p, hipGetLastError());
    return SDBE_OK;
}

}  // extern "C"
