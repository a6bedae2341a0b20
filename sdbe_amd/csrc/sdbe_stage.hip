// gfx950 implementation of the SWARM data back-end GPU stage behind the C-ABI
// of include/sdbe_stage.h.
//
// Pipeline per group (39 B-frames, both polarisations, 256 block-pols):
//   k_fill_i32 + k_packet_map   packet index map [frame][cid][fid], first packet
//                               wins; invalid / out-of-window packets counted
//   k_reorder                   LDS transpose of 16 packets per (frame, cid pair)
//                               into the reference int8 [snapshot][channel][im,re]
//                               rows, missing packets -> 0 + 0j, fid bitmap
// then per chunk of block-pols:
//   k_stage_a    decode + C2R pre-processing of each 32768-point snapshot
//                transform + 39-point DFT across the block's snapshots
//   k_stage_b    per 39-DFT row: 16384-point inverse FFT, inter-pass twiddle,
//                16384-point forward FFT  => packed 1,277,952-point R2C
//   k_stage_c1   R2C post-processing, zero padding, C2R pre-processing of the
//                2,097,152-point output transform, first (64-point) pass of
//                its packed 2^20 inverse FFT; FP64 Parseval partial sums
//   k_stats      per block-pol mean / ddof=1 sigma in FP64, threshold policy
//   k_stage_c2   second (16384-point) pass, scaling, 2-bit quantisation and
//                packing straight into the VDIF payloads
//   k_headers    VDIF output headers (FRAMES mode)
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <new>
#include <vector>

#include "sdbe_device.hpp"
#include "../../include/sdbe_stage.h"

using namespace sdbe;

namespace {

constexpr int PKT_WORDS = 264;
constexpr int NFR = 39;                       // frames per group
constexpr int NSNAP_GROUP = NFR * 128;        // 4992 snapshots per group
constexpr int MAP_N = NFR * 256 * 8;          // 79,872 packet slots per group
constexpr int MAP_NONE = 0x7FFFFFFF;
// Stage C1 tiles: columns in residue classes mod 39, 16 per tile
// (c = a + 39 (16 u + j)), 14 tiles per class, 39 classes.
constexpr int C1_CLASS_TILES = 14;
constexpr int C1_TILES = NBLK * C1_CLASS_TILES;   // 546
constexpr int C1_TILES_PER_WG = 8;                 // consecutive tiles run by one stage C1 workgroup
constexpr int BC_ROW = C1_TILES * 32;            // 17472 slots per row of the first-pass output
constexpr int OUT_FRAME_PAYLOAD = 8192;
constexpr int OUT_FRAMES_BLOCK = 64;

enum { CNT_INVALID = 0, CNT_RANGE = 1, CNT_ACCEPTED = 2, CNT_UNIQUE = 3, CNT_N = 8 };

// ---------------------------------------------------------------- unpack side
__global__ void __launch_bounds__(256) k_fill_i32(int *p, int n, int val) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = val;
}

// One thread per packet: header words 0, 4, 5 give invalid, B-count, cid, fid
// (vdif_inout_gpu_thread.cu:353-369,437 restated).  Valid packets inside the
// 39-frame window claim their map slot with atomicMin -> lowest index wins.
__global__ void __launch_bounds__(256)
k_packet_map(const uint32_t *__restrict__ pk, int n, uint32_t bzero32, int *__restrict__ map,
             int *__restrict__ cnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool inval = false, range = false, acc = false;
    if (i < n) {
        const uint32_t *h = pk + (size_t)i * PKT_WORDS;
        const uint32_t w0 = h[0], w4 = h[4], w5 = h[5];
        const uint32_t cid = w5 & 0xFFu, fid = (w5 >> 16) & 0xFFu;
        const uint32_t bc32 = (w5 >> 24) + ((w4 & 0xFFFFFFu) << 8);
        const int32_t rel = (int32_t)(bc32 - bzero32);
        if (w0 >> 31) {
            inval = true;
        } else if (rel < 0 || rel >= NFR || fid >= 8u) {
            range = true;
        } else {
            acc = true;
            atomicMin(map + ((rel * 256 + (int)cid) * 8 + (int)fid), i);
        }
    }
    // one atomic per wave per counter
    const unsigned long long bi = __ballot(inval), br = __ballot(range), ba = __ballot(acc);
    if ((threadIdx.x & 63) == 0) {
        if (bi) atomicAdd(cnt + CNT_INVALID, __popcll(bi));
        if (br) atomicAdd(cnt + CNT_RANGE, __popcll(br));
        if (ba) atomicAdd(cnt + CNT_ACCEPTED, __popcll(ba));
    }
}

// Workgroup (h = cid >> 1, frame): the 16 packets (fid 0..7, cid 2h, 2h+1)
// cover channels [128h, 128h+128).  Payload byte q of a snapshot carries
// channel offset (3,2,1,0,7,6,5,4)[q]; its bits from the LSB are pol1 im,
// pol1 re, pol0 im, pol0 re, each decoded as (code & 3) - 2
// (vdif_inout_gpu_thread.cu:387-392, 464-494).  Output: the reference int8
// layout [pol][snapshot][16384][im, re] (its beng_0 / beng_1 buffers), a
// missing packet leaving 0 + 0j; plus the fid bitmap and the unique count.
__global__ void __launch_bounds__(256)
k_reorder(const uint32_t *__restrict__ pk, const int *__restrict__ map, int8_t *__restrict__ beng,
          uint8_t *__restrict__ bitmap, int *__restrict__ cnt) {
    __shared__ int pidx[16];
    __shared__ uint32_t pay[16][257];     // packet payload words, [w][payload word]
    const int h = blockIdx.x, frame = blockIdx.y, t = threadIdx.x;
    if (t < 16) {
        const int w = t, fid = w >> 1, lo = w & 1;
        pidx[w] = map[(frame * 256 + 2 * h + lo) * 8 + fid];
    }
    __syncthreads();
    if (t < 2) {   // fid bitmap byte of cid = 2h + t, and the unique count
        uint32_t byte = 0;
        int found = 0;
        for (int fid = 0; fid < 8; ++fid) {
            if (pidx[2 * fid + t] != MAP_NONE) {
                byte |= 1u << fid;
                ++found;
            }
        }
        bitmap[frame * 256 + 2 * h + t] = (uint8_t)byte;
        if (found) atomicAdd(cnt + CNT_UNIQUE, found);
    }
#pragma unroll 4
    for (int it = 0; it < 16; ++it) {      // coalesced 1 KB payload reads
        const int p = pidx[it];
        pay[it][t] = (p != MAP_NONE) ? pk[(size_t)p * PKT_WORDS + 8 + t] : 0xAAAAAAAAu;   // 0xAA: value 0
    }
    __syncthreads();
    // thread -> (pol, snapshot, channel pair): 2 pols x 128 snapshots x 64 pairs
    // = 16384 uint32 stores of (im, re, im, re) for two adjacent channels
#pragma unroll 4
    for (int it = 0; it < 64; ++it) {
        const int e = it * 256 + t;
        const int pol = e >> 13, sn = (e >> 6) & 127, cp = e & 63;
        const int ch0 = 2 * cp;                 // channel within the 128-channel block
        const int w = ch0 >> 3, off = ch0 & 7;  // packet slot, channel offset (even)
        // offsets 0..3 sit in word 0 (bytes 3..0), 4..7 in word 1 (bytes 3..0)
        const uint32_t word = pay[w][2 * sn + (off >> 2)];
        const int q0 = 3 - (off & 3);           // byte of offset off; offset off+1 is byte q0-1
        const uint32_t b0 = (word >> (8 * q0)) & 0xFFu, b1 = (word >> (8 * (q0 - 1))) & 0xFFu;
        const int sh = pol == 0 ? 4 : 0;
        const uint32_t n0 = (b0 >> sh) & 0xFu, n1 = (b1 >> sh) & 0xFu;
        const uint32_t im0 = (uint32_t)(uint8_t)(int8_t)((int)(n0 & 3u) - 2);
        const uint32_t re0 = (uint32_t)(uint8_t)(int8_t)((int)(n0 >> 2) - 2);
        const uint32_t im1 = (uint32_t)(uint8_t)(int8_t)((int)(n1 & 3u) - 2);
        const uint32_t re1 = (uint32_t)(uint8_t)(int8_t)((int)(n1 >> 2) - 2);
        const size_t row = (size_t)pol * NSNAP_GROUP + (size_t)frame * 128 + sn;
        uint32_t *dst = reinterpret_cast<uint32_t *>(beng + (row * NCH + (size_t)h * 128 + ch0) * 2);
        *dst = im0 | (re0 << 8) | (im1 << 16) | (re1 << 24);
    }
}

// ---------------------------------------------------------------- resampler
// Workgroup `lin` of a 1D grid of `total` runs on XCD lin % 8 (round-robin
// dispatch).  The returned logical index is contiguous per XCD: logical
// neighbours share that XCD's L2 and are dispatched close together in time.
__device__ __forceinline__ int xcd_contig(int lin, int total) {
    const int x = lin & 7, slot = lin >> 3, q = total >> 3, r = total & 7;
    return x * q + (x < r ? x : r) + slot;
}

struct SrcDesc {
    const int8_t *i8;        // int8 [..][16384][2 = (im, re)] rows
    int group_mode;          // 1: rows [2][4992] of a group, bp = 2 blk + pol
};

// Stage A: thread per channel k of a block-pol.  For each snapshot s the
// packed C2R input Z_s[k] = (X[k] + conj(X[N2-k])) + i e^{2 pi i k/32768}
// (X[k] - conj(X[N2-k])) (DC imaginary part and the zero Nyquist bin dropped,
// as numpy irfft), then V[k1][k] = sum_s Z_s[k] e^{-2 pi i s k1/39}.
// Workgroups are paired (w, 63 - w) on one XCD back to back, so the mirror
// channels N2 - k a workgroup reads are the ones its partner reads as k and
// are served from that XCD's L2.
__global__ void __launch_bounds__(256, 2)
k_stage_a(SrcDesc sd, int bp0, int nb, float2 *__restrict__ V, Tables tb) {
    const int L = xcd_contig((int)blockIdx.x, (NCH / 256) * nb);
    const int bpl = L >> 6, idx = L & 63;
    const int w = (idx & 1) ? 63 - (idx >> 1) : (idx >> 1);
    const int k = w * 256 + threadIdx.x;
    const int bp = bp0 + bpl;
    size_t row0;
    if (sd.group_mode) {
        const int blk = bp >> 1, pol = bp & 1;
        row0 = (size_t)pol * NSNAP_GROUP + (size_t)NBLK * blk;
    } else {
        row0 = (size_t)bp * NBLK;
    }
    const char2 *src = reinterpret_cast<const char2 *>(sd.i8) + row0 * NCH;
    const int km = (NCH - k) & (NCH - 1);      // k = 0 pairs with the zero Nyquist bin
    const float kz = k == 0 ? 0.f : 1.f;
    const float2 T = tw<+1>(tb.n32k, (uint32_t)k);
    float2 x[NBLK];
#pragma unroll
    for (int s = 0; s < NBLK; ++s) {
        const char2 va = src[(size_t)s * NCH + k];
        const char2 vb = src[(size_t)s * NCH + km];
        const float2 a = make_float2((float)va.y, kz * (float)va.x);     // DC imaginary part dropped
        const float2 b = make_float2(kz * (float)vb.y, kz * (float)vb.x);  // Nyquist bin is zero
        const float2 sum = cadd(a, cconj(b));
        const float2 dif = csub(a, cconj(b));
        x[s] = cadd(sum, mul_si<+1>(cmul(T, dif)));
    }
    dft39<-1>(x);
    float2 *dst = V + (size_t)bpl * NBLK * NCH + k;
#pragma unroll
    for (int k1 = 0; k1 < NBLK; ++k1) dst[(size_t)k1 * NCH] = x[k1];
}

// Stage B: row k1 of block-pol bpl: w = IFFT16384(V[k1]); w[n] *= e^{-2 pi i n k1/638976};
// Zf[k1 + 39 k2] = FFT16384(w)[k2]   (packed forward DFT of the 1,277,952 samples)
// Workgroups are dispatched round-robin over the 8 XCDs; with xcd_map the 39
// rows of a block-pol all run on one XCD (requires the chunk to be a multiple
// of 8).  The packed spectrum is stored row-major: Zf[k1 + 39 k2] at
// k1 * 16384 + k2.
__global__ void __launch_bounds__(1024)
k_stage_b(const float2 *__restrict__ V, float2 *__restrict__ Zf, Tables tb, int xcd_map) {
    __shared__ float lre[LDS_PAD_N];
    __shared__ float lim[LDS_PAD_N];
    __shared__ FftTw ftw;
    fft_tw_load(&ftw, tb.fft);
    const int lin = blockIdx.x, t = threadIdx.x;
    int k1, bpl;
    if (xcd_map) {
        const int slot = lin >> 3;
        bpl = (lin & 7) + 8 * (slot / NBLK);
        k1 = slot % NBLK;
    } else {
        bpl = lin / NBLK;
        k1 = lin % NBLK;
    }
    const float2 *src = V + ((size_t)bpl * NBLK + k1) * NCH;
    float2 v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = src[t + 1024 * q];
    fft16k<+1>(v, lre, lim, &ftw);
    // e^{-2 pi i (t + 1024 q) k1 / M'} = A(t) * B(q); B is uniform over the workgroup
    const float2 A = tw<-1>(tb.nmp, (uint32_t)(t * k1));
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = cmul(v[q], cmul(A, tw<-1>(tb.nmp, (uint32_t)(1024 * q * k1))));
    fft16k<-1>(v, lre, lim, &ftw);
    // row-major [k1][k2]: Zf[k1 + 39 k2] lives at k1 * 16384 + k2
    float2 *dst = Zf + ((size_t)bpl * NBLK + k1) * NCH + t;
#pragma unroll
    for (int q = 0; q < 16; ++q) dst[1024 * q] = v[q];
}

// Stage C1: workgroup = 16 column pairs (c, 16384 - c) of the [64][16384] view
// (k = c + 16384 kh) of the packed 2^20-point C2R input, all 64 rows.
//   Y[m]  = 1/2 (Zf[m] + conj Zf[M'-m]) - 1/2 i e^{-2 pi i m/M} (Zf[m] - conj Zf[M'-m]),  m <= M'
//   Zc[k] = (C[k] + conj C[P-k]) + i e^{2 pi i k/L} (C[k] - conj C[P-k]),  C = Y zero-padded
//   Bc[tx][c] = sum_kh Zc[c + 16384 kh] e^{2 pi i kh tx/64}
// (the inter-pass twiddle e^{2 pi i c tx/P} is applied by stage C2 on load)
// and the FP64 Parseval sum  sum_m w_m |Y[m]|^2 (w_0 = 1 on Re Y[0], else 2).
// Tile (a, u) takes the lower columns c = a + 39 (16 u + j), j < 16: for each
// row r the 16 values m = c + 16384 r share the packed-spectrum row
// (m mod 39 = (a + 4 r) mod 39) at consecutive k2, so every read is a
// 128-byte run of the row-major store; the mirror columns likewise.  Output
// slots are tile-major (slot = 32 tile + col) and stage C2 puts them back in
// column order.  Twiddles factor per column and per row:
// e^{-2 pi i (c + 16384 r)/M} = e^{-2 pi i c/M} e^{-2 pi i r/78},
// e^{2 pi i (c + 16384 kh)/L} = e^{2 pi i c/L} e^{2 pi i kh/128}.
__device__ __forceinline__ void c1_column(int tile, int col, int &c, bool &valid, bool &dup) {
    const int a = tile / C1_CLASS_TILES, u = tile - a * C1_CLASS_TILES;
    const int cl = a + NBLK * (16 * u + (col & 15));
    valid = cl <= NCH / 2;
    c = col < 16 ? cl : ((NCH - cl) & (NCH - 1));
    dup = (col >= 16) && (cl == 0 || cl == NCH / 2);
}

// Each workgroup runs `tpw` consecutive tiles in (block-pol, a, u) order, and
// the workgroups are contiguous per XCD: tiles u and u + 1 of a class share the
// 128-byte lines at the ends of their runs, which the second finds in L2.  The
// next tile's packed-spectrum values and column twiddles are loaded while the
// current tile is computed.
struct C1Tile {
    int bpl, tile, c;
    bool valid, dup;
};

__device__ __forceinline__ void c1_fetch(const float2 *__restrict__ Zf, const Tables &tb, int lt, int col, int rw,
                                         C1Tile &x, float2 (&z)[5], float2 &cm, float2 &cl) {
    x.bpl = lt / C1_TILES;
    x.tile = lt - x.bpl * C1_TILES;
    c1_column(x.tile, col, x.c, x.valid, x.dup);
    const int a = x.tile / C1_CLASS_TILES, u = x.tile - a * C1_CLASS_TILES, j16 = col & 15;
    const bool mirror = col >= 16 && x.c != 0;
    const float2 *zf = Zf + (size_t)x.bpl * MP;
    // m = a + 16384 r + 39 (16 u + j) sits at row (a + 16384 r) mod 39, column
    // (a + 16384 r) / 39 + 16 u + j of the row-major store; the mirror column
    // 16384 - c at m = 16384 (r + 1) - a - 39 (16 u + j).  Row 39 holds only
    // m = M' (c = 0), which the store keeps at address 0 (as m = 0); values
    // that are not used also read address 0, so the loads are unconditional and
    // all five are in flight together.
#pragma unroll
    for (int it = 0; it < 5; ++it) {
        const int r = it * 8 + rw;
        const bool ok = x.valid && r < NBLK;
        int addr;
        if (mirror) {
            const int hi = NCH * (r + 1) - a;
            addr = (hi % NBLK) * NCH + hi / NBLK - 16 * u - j16;
        } else {
            const int lo = a + NCH * r;
            addr = (lo % NBLK) * NCH + lo / NBLK + 16 * u + j16;
        }
        z[it] = zf[ok ? addr : 0];
    }
    cm = tb.c1col[x.c];
    cl = tb.c1col[NCH + x.c];
}

__global__ void __launch_bounds__(256)
k_stage_c1(const float2 *__restrict__ Zf, float2 *__restrict__ Bc, double *__restrict__ part, Tables tb, int nb,
           int tpw) {
    __shared__ float2 sz[32][41];
    __shared__ float2 sy[32][41];
    __shared__ float2 sc[32][65];
    __shared__ float2 rowm[40], rowl[64], w64[64];
    __shared__ double red[2][4];
    const int t = threadIdx.x, col = t & 31, rw = t >> 5, pcol = col ^ 16;
    const int n_tiles = C1_TILES * nb;
    const int first = xcd_contig((int)blockIdx.x, (int)gridDim.x) * tpw;
    const int last = min(first + tpw, n_tiles);
    if (t < 40) {
        rowm[t] = tw<-1>(tb.nm, (uint32_t)(NCH * t));
    } else if (t < 104) {
        rowl[t - 40] = tw<+1>(tb.nl, (uint32_t)(NCH * (t - 40)));
    } else if (t < 168) {
        const int e = t - 104;   // W64^(r j) for r = e >> 3, j = e & 7
        w64[e] = tw<+1>(tb.np, (uint32_t)((e >> 3) * (e & 7) * NCH));
    }
    C1Tile cur;
    float2 z[5], cm, cl;
    if (first < last) c1_fetch(Zf, tb, first, col, rw, cur, z, cm, cl);
    for (int lt = first; lt < last; ++lt) {
        const C1Tile x = cur;
        const int c = x.c;
        const float2 cmc = cm, clc = cl;
#pragma unroll
        for (int it = 0; it < 5; ++it) {
            const int r = it * 8 + rw;
            const bool ok = x.valid && (r < NBLK || c == 0);
            sz[col][r] = ok ? z[it] : make_float2(0.f, 0.f);
        }
        __syncthreads();
        if (lt + 1 < last) c1_fetch(Zf, tb, lt + 1, col, rw, cur, z, cm, cl);
        // Y rows 0..39 (zero above M') + Parseval
        double acc = 0.0, dc = 0.0;
#pragma unroll
        for (int it = 0; it < 5; ++it) {
            const int r = it * 8 + rw;
            float2 Y = make_float2(0.f, 0.f);
            if (x.valid && (r < NBLK || c == 0)) {
                const int pr = (c == 0) ? (39 - r) : (38 - r);
                const float2 zm = sz[col][r];
                const float2 zp = cconj(sz[pcol][pr]);
                const float2 E = cscale(cadd(zm, zp), 0.5f);
                const float2 D = cscale(csub(zm, zp), 0.5f);
                Y = cadd(E, mul_si<-1>(cmul(cmul(cmc, rowm[r]), D)));
                if (!x.dup) {
                    if (c == 0 && r == 0) {
                        acc += (double)Y.x * (double)Y.x;
                        dc = (double)Y.x;
                    } else {
                        acc += 2.0 * ((double)Y.x * (double)Y.x + (double)Y.y * (double)Y.y);
                    }
                }
            }
            sy[col][r] = Y;
        }
        __syncthreads();
        // Zc rows 0..63
#pragma unroll
        for (int it = 0; it < 8; ++it) {
            const int kh = it * 8 + rw;
            const int pr = (c == 0) ? (64 - kh) : (63 - kh);
            const float2 ck = kh <= 39 ? sy[col][kh] : make_float2(0.f, 0.f);
            const float2 cp = pr <= 39 ? cconj(sy[pcol][pr]) : make_float2(0.f, 0.f);
            sc[col][kh] = cadd(cadd(ck, cp), mul_si<+1>(cmul(cmul(clc, rowl[kh]), csub(ck, cp))));
        }
        __syncthreads();
        // 64-point inverse DFT over kh per column: 8 threads x radix 8, two passes
        // (lanes of a half-wave take different columns: odd row stride, no bank conflicts)
        {
            const int j = rw;
            float2 v[8];
#pragma unroll
            for (int r = 0; r < 8; ++r) v[r] = sc[col][j + 8 * r];
            dft8<+1>(v);
            __syncthreads();
#pragma unroll
            for (int r = 0; r < 8; ++r) sc[col][8 * j + r] = v[r];
            __syncthreads();
#pragma unroll
            for (int r = 0; r < 8; ++r) v[r] = sc[col][j + 8 * r];
#pragma unroll
            for (int r = 1; r < 8; ++r) v[r] = cmul(v[r], w64[8 * r + j]);
            dft8<+1>(v);
            __syncthreads();
#pragma unroll
            for (int r = 0; r < 8; ++r) sc[col][j + 8 * r] = v[r];
        }
        __syncthreads();
        // write the tile's 32 slots of every row tx (256-byte runs)
        float2 *bc = Bc + (size_t)x.bpl * KHI * BC_ROW + (size_t)x.tile * 32;
#pragma unroll
        for (int it = 0; it < 8; ++it) {
            const int tx = it * 8 + rw;
            if (x.valid && !x.dup) bc[(size_t)tx * BC_ROW + col] = sc[col][tx];
        }
        // Parseval partial: wave reduce then cross-wave
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            acc += __shfl_xor(acc, off);
            dc += __shfl_xor(dc, off);
        }
        if ((t & 63) == 0) {
            red[0][t >> 6] = acc;
            red[1][t >> 6] = dc;
        }
        __syncthreads();
        if (t == 0) {
            double *o = part + ((size_t)x.bpl * C1_TILES + x.tile) * 2;
            o[0] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
            o[1] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
        }
    }
}

// Statistics and thresholds of a chunk.  Units: the pipeline's samples are
// 32768 * 2^21 = 2^36 times numpy's; Y here carries a factor 32768.
//   sum x^2 = W / (32768^2 * L),  mean = Re Y[0] / (32768 * L),
//   sigma = sqrt((sum x^2 - L mean^2) / (L - 1))
struct StatsDesc {
    int policy;
    float fixed_th, fixed_off;
    int group_mode;
};

__global__ void __launch_bounds__(256)
k_stats(const double *__restrict__ part, int bp0, double2 *__restrict__ ms, double *__restrict__ stats_out) {
    __shared__ double red[2][256];
    const int t = threadIdx.x, b = blockIdx.x;
    const double L = 2097152.0, FA = 32768.0;
    double a = 0.0, d = 0.0;
    for (int i = t; i < C1_TILES; i += 256) {
        a += part[((size_t)b * C1_TILES + i) * 2];
        d += part[((size_t)b * C1_TILES + i) * 2 + 1];
    }
    red[0][t] = a;
    red[1][t] = d;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) {
            red[0][t] += red[0][t + s];
            red[1][t] += red[1][t + s];
        }
        __syncthreads();
    }
    if (t == 0) {
        const double sumsq = red[0][0] / (FA * FA * L);
        const double mean = red[1][0] / (FA * L);
        double var = (sumsq - L * mean * mean) / (L - 1.0);
        if (var < 0.0) var = 0.0;
        const double sig = sqrt(var);
        ms[b] = make_double2(mean, sig);
        if (stats_out) {
            stats_out[(size_t)(bp0 + b) * 2] = mean;
            stats_out[(size_t)(bp0 + b) * 2 + 1] = sig;
        }
    }
}

// Threshold policy over a chunk, in block order (one thread: the REFERENCE
// latch is sequential by definition).
__global__ void k_thresh(const double2 *__restrict__ ms, int nbpl, int bp0, StatsDesc sdsc,
                         float *__restrict__ latch, float2 *__restrict__ thr) {
    if (threadIdx.x != 0) return;
    for (int b = 0; b < nbpl; ++b) {
        float th, off;
        const int pol = sdsc.group_mode ? ((bp0 + b) & 1) : 0;
        if (sdsc.policy == SDBE_THRESH_PER_BLOCK) {
            th = (float)ms[b].y;
            off = (float)ms[b].x;
        } else if (sdsc.policy == SDBE_THRESH_REFERENCE) {
            if (latch[0] == 0.0f && pol == 0) {
                latch[0] = (float)ms[b].y;
                latch[1] = (float)ms[b].x;
            }
            th = latch[0];
            off = latch[1];
        } else {
            th = sdsc.fixed_th;
            off = sdsc.fixed_off;
        }
        thr[b] = make_float2(th, off);
    }
}

struct OutDesc {
    uint8_t *base;
    int64_t pol_stride;
    int32_t frame_bytes;
    int32_t hdr_bytes;
    int32_t group_mode;
    float *f32;          // resample-only output [n_bp][2^21] or null
};

// Stage C2: workgroup = (tx, block-pol).
// z[tx + 64 ty] = sum_c Bc[tx][c] e^{2 pi i c tx/2^20} e^{2 pi i c ty/16384} is the packed output:
// samples (2t, 2t+1) = (Re, Im) z[t], scaled by 2^-36 to numpy units.  Codes
// per vdif_inout_gpu_thread.cu:577 (v = x - off; ((|v| >= th) | 2) ^ 3 (v < 0));
// the 4-bit pair code (re | im << 2) of t goes to stage[bp][tx][ty].
template <bool F32>
__global__ void __launch_bounds__(1024)
k_stage_c2(const float2 *__restrict__ Bc, const float2 *__restrict__ thr, int bp0, OutDesc od,
           uint8_t *__restrict__ stage, const int *__restrict__ slot_col, Tables tb) {
    __shared__ float lre[LDS_PAD_N];
    __shared__ float lim[LDS_PAD_N];
    __shared__ FftTw ftw;
    fft_tw_load(&ftw, tb.fft);
    const int tx = blockIdx.x, bpl = blockIdx.y, t = threadIdx.x, bp = bp0 + bpl;
    const float scale = 1.4551915228366852e-11f;   // 2^-36, exact
    // the row arrives in stage C1's tile-major slot order: scatter it into LDS in
    // column order (slot -> column table, -1 for slots without a column)
    // (all loads are issued before any LDS store; slots without a column are
    // read too -- the row is allocated -- and dropped)
    const float2 *src = Bc + ((size_t)bpl * KHI + tx) * BC_ROW;
    constexpr int NSLOT = BC_ROW / 1024;          // 17 full passes ...
    constexpr int REM = BC_ROW - NSLOT * 1024;    // ... and 64 slots left
    {
        float2 val[NSLOT];
        int cc[NSLOT];
#pragma unroll
        for (int i = 0; i < NSLOT; ++i) {
            val[i] = src[t + 1024 * i];
            cc[i] = slot_col[t + 1024 * i];
        }
        float2 vr = make_float2(0.f, 0.f);
        int cr = -1;
        if (t < REM) {
            vr = src[t + 1024 * NSLOT];
            cr = slot_col[t + 1024 * NSLOT];
        }
#pragma unroll
        for (int i = 0; i < NSLOT; ++i) {
            if (cc[i] >= 0) {
                lre[cc[i] + (cc[i] >> 4)] = val[i].x;
                lim[cc[i] + (cc[i] >> 4)] = val[i].y;
            }
        }
        if (cr >= 0) {
            lre[cr + (cr >> 4)] = vr.x;
            lim[cr + (cr >> 4)] = vr.y;
        }
    }
    __syncthreads();
    float2 v[16];
    {
        const int rd = t + (t >> 4);
#pragma unroll
        for (int q = 0; q < 16; ++q) v[q] = make_float2(lre[rd + 1088 * q], lim[rd + 1088 * q]);
    }
    // inter-pass twiddle e^{2 pi i c tx / 2^20}, c = t + 1024 q: per-thread factor
    // times a workgroup-uniform one
    const float2 A = tw<+1>(tb.np, (uint32_t)(t * tx));
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = cmul(v[q], cmul(A, tw<+1>(tb.np, (uint32_t)(1024 * q * tx))));
    fft16k<+1>(v, lre, lim, &ftw);
    if (F32) {
        float2 *dst = reinterpret_cast<float2 *>(od.f32) + (size_t)bp * P20 + tx + KHI * t;
#pragma unroll
        for (int q = 0; q < 16; ++q) dst[(size_t)KHI * 1024 * q] = cscale(v[q], scale);
    } else {
        const float2 to = thr[bpl];
        uint8_t *dst = stage + ((size_t)bpl * KHI + tx) * NCH + t;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const float a = v[q].x * scale - to.y;
            const float b = v[q].y * scale - to.y;
            const uint32_t ca = ((((fabsf(a) >= to.x) ? 1u : 0u) | 2u) ^ (a < 0.f ? 3u : 0u)) & 3u;
            const uint32_t cb = ((((fabsf(b) >= to.x) ? 1u : 0u) | 2u) ^ (b < 0.f ? 3u : 0u)) & 3u;
            dst[1024 * q] = (uint8_t)(ca | (cb << 2));
        }
    }
}

// Pack: workgroup = (64 consecutive ty, block-pol).  Stream byte q = t/2 of a
// block holds samples 4q..4q+3 (sample j at bits 2j, the uint32 packing of
// vdif_inout_gpu_thread.cu:578-580): byte (i + 32 ty) = code(t = 2i + 64 ty) |
// code(t = 2i + 1 + 64 ty) << 4.  Block bytes go to 64 frames of 8192 bytes.
__global__ void __launch_bounds__(256)
k_pack(const uint8_t *__restrict__ stage, int bp0, int nb, OutDesc od) {
    __shared__ uint8_t tile[64][65];
    // neighbouring ty ranges share 128-byte lines of the code rows: keep them
    // on one XCD, back to back
    const int L = xcd_contig((int)blockIdx.x, (NCH / 64) * nb);
    const int bpl = L >> 8, ty0 = (L & 255) * 64, t = threadIdx.x, bp = bp0 + bpl;
    const uint8_t *src = stage + (size_t)bpl * KHI * NCH + ty0;
#pragma unroll
    for (int it = 0; it < 16; ++it) {
        const int e = it * 256 + t, tx = e >> 6, ty = e & 63;
        tile[ty][tx] = src[(size_t)tx * NCH + ty];
    }
    __syncthreads();
    int pol, blk;
    if (od.group_mode) {
        pol = bp & 1;
        blk = bp >> 1;
    } else {
        pol = 0;
        blk = bp;
    }
    // 64 ty x 32 bytes = 2048 contiguous stream bytes starting at 32 * ty0
    uint8_t *ob = od.base + (int64_t)pol * od.pol_stride + (int64_t)blk * OUT_FRAMES_BLOCK * od.frame_bytes +
                  od.hdr_bytes;
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        const int e = it * 256 + t;            // 4-byte word index within the 2048 bytes
        const int ty = e >> 3, i0 = (e & 7) * 4;
        uint32_t wv = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int i = i0 + b;
            wv |= (uint32_t)(tile[ty][2 * i] | (tile[ty][2 * i + 1] << 4)) << (8 * b);
        }
        const int qb = 32 * (ty0 + ty) + i0;    // stream byte
        const int f = qb >> 13, off = qb & (OUT_FRAME_PAYLOAD - 1);
        *reinterpret_cast<uint32_t *>(ob + (int64_t)f * od.frame_bytes + off) = wv;
    }
}

struct HdrDesc {
    uint32_t secs, df, ref_epoch;
    uint64_t psn;
    uint32_t station, bdc, rx;
    int n_frames_per_pol;
    int n_pols;
};

// VDIF output headers (vdif_out_databuf.h:19-52) stamped as the output thread
// does (vdif_out_net_thread.c:294-317, defaults vdif_out_databuf.c:195-227):
// frame n of the call has df = (df0 + n) mod 125000, secs = secs0 + carry,
// psn = psn0 + n, threadID 0, bps = 1, df_len = 8224/8, edv5 = 2.
__global__ void __launch_bounds__(256) k_headers(OutDesc od, HdrDesc hd) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= hd.n_frames_per_pol * hd.n_pols) return;
    const int pol = i / hd.n_frames_per_pol, n = i - pol * hd.n_frames_per_pol;
    const uint64_t dft = (uint64_t)hd.df + (uint64_t)n;
    const uint32_t df = (uint32_t)(dft % 125000u);
    const uint32_t secs = hd.secs + (uint32_t)(dft / 125000u);
    const uint64_t psn = hd.psn + (uint64_t)n;
    uint32_t w[8];
    w[0] = secs & 0x3FFFFFFFu;
    w[1] = (df & 0xFFFFFFu) | ((hd.ref_epoch & 0x3Fu) << 24);
    w[2] = (uint32_t)((OUT_FRAME_PAYLOAD + 32) / 8);
    w[3] = (hd.station & 0xFFFFu) | (1u << 26);
    w[4] = ((uint32_t)pol & 1u) | ((hd.bdc & 1u) << 1) | ((hd.rx & 1u) << 2) | (2u << 24);
    w[5] = 0u;
    w[6] = (uint32_t)(psn & 0xFFFFFFFFu);
    w[7] = (uint32_t)(psn >> 32);
    uint4 *dst = reinterpret_cast<uint4 *>(od.base + (int64_t)pol * od.pol_stride + (int64_t)n * od.frame_bytes);
    dst[0] = make_uint4(w[0], w[1], w[2], w[3]);
    dst[1] = make_uint4(w[4], w[5], w[6], w[7]);
}

}  // namespace

// ======================================================================= host
enum KernelId {
    KID_FILL, KID_MAP, KID_REORDER, KID_A, KID_B, KID_C1, KID_STATS, KID_THRESH, KID_C2,
    KID_PACK, KID_HEADERS, KID_N
};
static const char *const kKernelNames[KID_N] = {
    "k_fill_i32", "k_packet_map", "k_reorder", "k_stage_a", "k_stage_b",
    "k_stage_c1", "k_stats", "k_thresh", "k_stage_c2", "k_pack", "k_headers"};

struct ProfRec {
    int kid;
    double units;
    hipEvent_t e0, e1;
};

struct sdbe_handle {
    int device = 0;
    sdbe_config_t cfg{};
    int chunk = 0;
    int last_hip_err = 0;
    hipStream_t last_stream = nullptr;
    sdbe_stream_state_t st{};
    // device buffers
    float2 *tables = nullptr;
    Tables tb{};
    int *map = nullptr;
    int8_t *beng = nullptr;          // int8 rows [2][4992][16384][2] of the current group
    uint8_t *bitmap = nullptr;
    int *cnt = nullptr;
    float *latch = nullptr;
    float2 *V = nullptr, *Zf = nullptr, *Bc = nullptr;
    double *part = nullptr;
    float2 *thr = nullptr;
    double2 *ms = nullptr;
    uint8_t *stage = nullptr;
    int *slot_col = nullptr;         // stage C1 output slot -> column (or -1)
    // pinned host mirrors
    int *h_cnt = nullptr;
    float *h_latch = nullptr;
    int n_packets_last = 0;
    int n_bp_last = 0;
    int have_group = 0;
    // profiling
    int prof = 0;
    std::vector<ProfRec> recs;
    std::vector<hipEvent_t> pool;
};

namespace {

int hip_fail(sdbe_handle_t *h, hipError_t e) {
    if (h) h->last_hip_err = (int)e;
    return SDBE_ERR_HIP;
}
#define HIP_TRY(h, expr)                         \
    do {                                         \
        hipError_t _e = (expr);                  \
        if (_e != hipSuccess) return hip_fail((h), _e); \
    } while (0)

ProfRec *prof_begin(sdbe_handle_t *h, int kid, double units, hipStream_t s) {
    if (!h->prof) return nullptr;
    hipEvent_t ev[2];
    for (int i = 0; i < 2; ++i) {
        if (!h->pool.empty()) {
            ev[i] = h->pool.back();
            h->pool.pop_back();
        } else if (hipEventCreate(&ev[i]) != hipSuccess) {
            return nullptr;
        }
    }
    h->recs.push_back(ProfRec{kid, units, ev[0], ev[1]});
    (void)hipEventRecord(ev[0], s);
    return &h->recs.back();
}

void prof_end(sdbe_handle_t *h, int rec_index, hipStream_t s) {
    if (rec_index >= 0) (void)hipEventRecord(h->recs[rec_index].e1, s);
}

#define LAUNCH(h, kid, units, s, kern, grid, block, ...)                         \
    do {                                                                        \
        ProfRec *_r = prof_begin((h), (kid), (units), (s));                     \
        const int _ri = _r ? (int)(h)->recs.size() - 1 : -1;                    \
        hipLaunchKernelGGL(kern, grid, block, 0, s, __VA_ARGS__);               \
        prof_end((h), _ri, (s));                                                \
    } while (0)

struct TwSpec {
    int64_t N;
    int lo_bits;
    int64_t max_idx;   // exclusive upper bound of the indices used
};

size_t tw_entries(const TwSpec &s) {
    return ((size_t)1 << s.lo_bits) + (size_t)((s.max_idx >> s.lo_bits) + 2);
}

void tw_fill(float2 *host, const TwSpec &s) {
    const size_t nlo = (size_t)1 << s.lo_bits;
    const size_t nhi = (size_t)((s.max_idx >> s.lo_bits) + 2);
    const double two_pi = 6.283185307179586476925286766559;
    for (size_t j = 0; j < nlo; ++j) {
        const double a = two_pi * (double)j / (double)s.N;
        host[j] = make_float2((float)cos(a), (float)sin(a));
    }
    for (size_t i = 0; i < nhi; ++i) {
        const double a = two_pi * (double)((int64_t)i << s.lo_bits) / (double)s.N;
        host[nlo + i] = make_float2((float)cos(a), (float)sin(a));
    }
}

void fft_tw_fill(float2 *host) {
    const double two_pi = 6.283185307179586476925286766559;
    auto e = [&](double num, double den) { return make_float2((float)cos(two_pi * num / den), (float)sin(two_pi * num / den)); };
    float2 *p = host;
    for (int m = 0; m < 256; ++m) *p++ = e(m, 256);
    for (int m = 0; m < 64; ++m) *p++ = e(64.0 * m, 4096);
    for (int m = 0; m < 64; ++m) *p++ = e(m, 4096);
    for (int m = 0; m < 128; ++m) *p++ = e(128.0 * m, 16384);
    for (int m = 0; m < 128; ++m) *p++ = e(m, 16384);
}

int build_tables(sdbe_handle_t *h) {
    const TwSpec specs[6] = {
        {16384, 14, 0},                 // unused slot (the FFT image is stored separately)
        {32768, 7, 16384},              // n32k
        {MP, 10, (int64_t)MP},          // nmp: n * k1 < 16384 * 39
        {2 * (int64_t)MP, 10, (int64_t)MP + 1},   // nm: m <= M'
        {2 * (int64_t)P20, 10, (int64_t)P20},     // nl: k < 2^20
        {P20, 10, (int64_t)P20},        // np: c * tx < 2^20, r * j * 16384 < 2^20
    };
    size_t total = FFT_TW_N;
    for (int i = 1; i < 6; ++i) total += tw_entries(specs[i]);
    const size_t c1col_off = total;
    total += 2 * NCH;
    std::vector<float2> host(total);
    size_t off[6];
    size_t o = FFT_TW_N;
    fft_tw_fill(host.data());
    for (int i = 1; i < 6; ++i) {
        off[i] = o;
        tw_fill(host.data() + o, specs[i]);
        o += tw_entries(specs[i]);
    }
    // stage C1 column twiddles e^{-2 pi i c/M} and e^{2 pi i c/L}, c < 16384
    {
        const double two_pi = 6.283185307179586476925286766559;
        for (int c = 0; c < NCH; ++c) {
            const double am = -two_pi * (double)c / (2.0 * (double)MP);
            const double al = two_pi * (double)c / (2.0 * (double)P20);
            host[c1col_off + c] = make_float2((float)cos(am), (float)sin(am));
            host[c1col_off + NCH + c] = make_float2((float)cos(al), (float)sin(al));
        }
    }
    HIP_TRY(h, hipMalloc(&h->tables, total * sizeof(float2)));
    HIP_TRY(h, hipMemcpy(h->tables, host.data(), total * sizeof(float2), hipMemcpyHostToDevice));
    h->tb.fft = h->tables;
    h->tb.c1col = h->tables + c1col_off;
    Tw2 *dst[6] = {nullptr, &h->tb.n32k, &h->tb.nmp, &h->tb.nm, &h->tb.nl, &h->tb.np};
    for (int i = 1; i < 6; ++i) {
        dst[i]->lo = h->tables + off[i];
        dst[i]->hi = h->tables + off[i] + ((size_t)1 << specs[i].lo_bits);
        dst[i]->lo_bits = specs[i].lo_bits;
    }
    return SDBE_OK;
}

// Slot -> column table of the stage C1 output (same rule as c1_column); every
// column 0..16383 must own exactly one slot.
int build_slot_table(sdbe_handle_t *h) {
    std::vector<int> tab(BC_ROW, -1), seen(NCH, 0);
    for (int tile = 0; tile < C1_TILES; ++tile) {
        const int a = tile / C1_CLASS_TILES, u = tile - a * C1_CLASS_TILES;
        for (int col = 0; col < 32; ++col) {
            const int cl = a + NBLK * (16 * u + (col & 15));
            if (cl > NCH / 2) continue;
            if (col >= 16 && (cl == 0 || cl == NCH / 2)) continue;
            const int c = col < 16 ? cl : ((NCH - cl) & (NCH - 1));
            tab[tile * 32 + col] = c;
            seen[c]++;
        }
    }
    for (int c = 0; c < NCH; ++c)
        if (seen[c] != 1) return SDBE_ERR_STATE;
    HIP_TRY(h, hipMemcpy(h->slot_col, tab.data(), BC_ROW * sizeof(int), hipMemcpyHostToDevice));
    return SDBE_OK;
}

// Run stages A..C2 over block-pols [0, n_bp) in chunks.
int run_resampler(sdbe_handle_t *h, const SrcDesc &sd, int n_bp, const OutDesc &od,
                  double *d_stats, hipStream_t s) {
    StatsDesc sdsc;
    sdsc.policy = h->cfg.thresh_policy;
    sdsc.fixed_th = h->cfg.fixed_threshold;
    sdsc.fixed_off = h->cfg.fixed_offset;
    sdsc.group_mode = sd.group_mode;
    for (int bp0 = 0; bp0 < n_bp; bp0 += h->chunk) {
        const int nb = (n_bp - bp0) < h->chunk ? (n_bp - bp0) : h->chunk;
        LAUNCH(h, KID_A, nb, s, k_stage_a, dim3((NCH / 256) * nb), dim3(256), sd, bp0, nb, h->V, h->tb);
        LAUNCH(h, KID_B, nb, s, k_stage_b, dim3(NBLK * nb), dim3(1024), (const float2 *)h->V, h->Zf, h->tb,
               (nb % 8) == 0 ? 1 : 0);
        const int c1_wgs = (C1_TILES * nb + C1_TILES_PER_WG - 1) / C1_TILES_PER_WG;
        LAUNCH(h, KID_C1, nb, s, k_stage_c1, dim3(c1_wgs), dim3(256), (const float2 *)h->Zf, h->Bc, h->part, h->tb,
               nb, C1_TILES_PER_WG);
        LAUNCH(h, KID_STATS, nb, s, k_stats, dim3(nb), dim3(256), (const double *)h->part, bp0, h->ms, d_stats);
        LAUNCH(h, KID_THRESH, nb, s, k_thresh, dim3(1), dim3(64), (const double2 *)h->ms, nb, bp0, sdsc, h->latch,
               h->thr);
        if (od.f32) {
            LAUNCH(h, KID_C2, nb, s, k_stage_c2<true>, dim3(KHI, nb), dim3(1024), (const float2 *)h->Bc,
                   (const float2 *)h->thr, bp0, od, h->stage, (const int *)h->slot_col, h->tb);
        } else {
            LAUNCH(h, KID_C2, nb, s, k_stage_c2<false>, dim3(KHI, nb), dim3(1024), (const float2 *)h->Bc,
                   (const float2 *)h->thr, bp0, od, h->stage, (const int *)h->slot_col, h->tb);
            LAUNCH(h, KID_PACK, nb, s, k_pack, dim3((NCH / 64) * nb), dim3(256), (const uint8_t *)h->stage, bp0, nb,
                   od);
        }
        HIP_TRY(h, hipGetLastError());
    }
    return SDBE_OK;
}

int unpack_to_rows(sdbe_handle_t *h, const void *d_pkts, int32_t n_pkts, uint64_t b_zero,
                   uint8_t *d_bitmap, hipStream_t s) {
    LAUNCH(h, KID_FILL, 1, s, k_fill_i32, dim3((MAP_N + 255) / 256), dim3(256), h->map, MAP_N, MAP_NONE);
    LAUNCH(h, KID_FILL, 0, s, k_fill_i32, dim3(1), dim3(256), h->cnt, CNT_N, 0);
    if (n_pkts > 0)
        LAUNCH(h, KID_MAP, 1, s, k_packet_map, dim3((n_pkts + 255) / 256), dim3(256), (const uint32_t *)d_pkts,
               n_pkts, (uint32_t)(b_zero & 0xFFFFFFFFull), h->map, h->cnt);
    LAUNCH(h, KID_REORDER, 1, s, k_reorder, dim3(128, NFR), dim3(256), (const uint32_t *)d_pkts,
           (const int *)h->map, h->beng, d_bitmap ? d_bitmap : h->bitmap, h->cnt);
    HIP_TRY(h, hipGetLastError());
    return SDBE_OK;
}

void advance_state(sdbe_stream_state_t &st, int64_t n_frames) {
    const uint64_t dft = (uint64_t)st.df + (uint64_t)n_frames;
    st.secs += (uint32_t)(dft / 125000u);
    st.df = (uint32_t)(dft % 125000u);
    st.psn += (uint64_t)n_frames;
}

int finish_call(sdbe_handle_t *h, hipStream_t s) {
    HIP_TRY(h, hipMemcpyAsync(h->h_cnt, h->cnt, CNT_N * sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(h, hipMemcpyAsync(h->h_latch, h->latch, 2 * sizeof(float), hipMemcpyDeviceToHost, s));
    h->last_stream = s;
    return SDBE_OK;
}

}  // namespace

extern "C" {

int sdbe_default_config(sdbe_config_t *cfg) {
    if (!cfg) return SDBE_ERR_ARG;
    memset(cfg, 0, sizeof(*cfg));
    cfg->thresh_policy = SDBE_THRESH_REFERENCE;
    cfg->fixed_threshold = 1.0f;
    cfg->fixed_offset = 0.0f;
    cfg->out_mode = SDBE_OUT_PAYLOAD;
    cfg->chunk_blockpols = 128;
    cfg->max_spectra_blockpols = 256;
    cfg->station_id = (uint16_t)((('X' & 0xFF) << 8) | ('X' & 0xFF));
    return SDBE_OK;
}

const char *sdbe_error_string(int code) {
    switch (code) {
        case SDBE_OK: return "ok";
        case SDBE_ERR_ARG: return "invalid argument";
        case SDBE_ERR_HIP: return "HIP runtime error";
        case SDBE_ERR_NOMEM: return "out of memory";
        case SDBE_ERR_CAPACITY: return "request exceeds handle capacity";
        case SDBE_ERR_NODEV: return "no such HIP device";
        case SDBE_ERR_STATE: return "invalid state";
        default: return "unknown error";
    }
}

int sdbe_last_hip_error(const sdbe_handle_t *h) { return h ? h->last_hip_err : 0; }

void sdbe_destroy(sdbe_handle_t *h) {
    if (!h) return;
    (void)hipSetDevice(h->device);
    if (h->last_stream) (void)hipStreamSynchronize(h->last_stream);
    (void)hipDeviceSynchronize();
    (void)hipFree(h->tables);
    (void)hipFree(h->map);
    (void)hipFree(h->beng);
    (void)hipFree(h->bitmap);
    (void)hipFree(h->cnt);
    (void)hipFree(h->latch);
    (void)hipFree(h->V);
    (void)hipFree(h->Zf);
    (void)hipFree(h->Bc);
    (void)hipFree(h->part);
    (void)hipFree(h->thr);
    (void)hipFree(h->stage);
    (void)hipFree(h->slot_col);
    (void)hipFree(h->ms);
    (void)hipHostFree(h->h_cnt);
    (void)hipHostFree(h->h_latch);
    for (auto &r : h->recs) {
        (void)hipEventDestroy(r.e0);
        (void)hipEventDestroy(r.e1);
    }
    for (auto e : h->pool) (void)hipEventDestroy(e);
    delete h;
}

int sdbe_create(const sdbe_config_t *cfg_in, int device, sdbe_handle_t **out) {
    if (!out) return SDBE_ERR_ARG;
    *out = nullptr;
    sdbe_config_t cfg;
    if (cfg_in) cfg = *cfg_in;
    else sdbe_default_config(&cfg);
    if (cfg.thresh_policy < SDBE_THRESH_REFERENCE || cfg.thresh_policy > SDBE_THRESH_FIXED) return SDBE_ERR_ARG;
    if (cfg.out_mode != SDBE_OUT_PAYLOAD && cfg.out_mode != SDBE_OUT_FRAMES) return SDBE_ERR_ARG;
    if (cfg.chunk_blockpols < 2 || cfg.chunk_blockpols > 256 || (cfg.chunk_blockpols & 1)) return SDBE_ERR_ARG;
    if (cfg.max_spectra_blockpols < 1) return SDBE_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return SDBE_ERR_NODEV;
    sdbe_handle_t *h = new (std::nothrow) sdbe_handle();
    if (!h) return SDBE_ERR_NOMEM;
    h->device = device;
    h->cfg = cfg;
    h->chunk = cfg.chunk_blockpols;
    h->st.psn = 1;
    int rc = SDBE_OK;
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) {
        delete h;
        return SDBE_ERR_NODEV;
    }
    const size_t C = (size_t)h->chunk;
#define ALLOC(ptr, bytes)                              \
    do {                                               \
        if (hipMalloc((void **)&(ptr), (bytes)) != hipSuccess) { \
            rc = SDBE_ERR_NOMEM;                       \
            goto fail;                                 \
        }                                              \
    } while (0)
    ALLOC(h->map, MAP_N * sizeof(int));
    ALLOC(h->beng, 2 * (size_t)NSNAP_GROUP * NCH * 2);
    ALLOC(h->bitmap, NFR * 256);
    ALLOC(h->cnt, CNT_N * sizeof(int));
    ALLOC(h->latch, 2 * sizeof(float));
    ALLOC(h->V, C * NBLK * NCH * sizeof(float2));
    ALLOC(h->Zf, C * MP * sizeof(float2));
    ALLOC(h->Bc, C * KHI * BC_ROW * sizeof(float2));
    ALLOC(h->slot_col, BC_ROW * sizeof(int));
    ALLOC(h->part, C * C1_TILES * 2 * sizeof(double));
    ALLOC(h->thr, C * sizeof(float2));
    ALLOC(h->stage, C * KHI * NCH);
    ALLOC(h->ms, C * sizeof(double2));
#undef ALLOC
    if (hipHostMalloc((void **)&h->h_cnt, CNT_N * sizeof(int), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&h->h_latch, 2 * sizeof(float), hipHostMallocDefault) != hipSuccess) {
        rc = SDBE_ERR_NOMEM;
        goto fail;
    }
    memset(h->h_cnt, 0, CNT_N * sizeof(int));
    memset(h->h_latch, 0, 2 * sizeof(float));
    rc = build_tables(h);
    if (rc != SDBE_OK) goto fail;
    rc = build_slot_table(h);
    if (rc != SDBE_OK) goto fail;
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(256), 0, 0, h->cnt, CNT_N, 0);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(256), 0, 0, (int *)h->latch, 2, 0);
    if (hipDeviceSynchronize() != hipSuccess) {
        rc = SDBE_ERR_HIP;
        goto fail;
    }
    *out = h;
    return SDBE_OK;
fail:
    sdbe_destroy(h);
    return rc;
}

int sdbe_set_stream_state(sdbe_handle_t *h, const sdbe_stream_state_t *st) {
    if (!h || !st) return SDBE_ERR_ARG;
    h->st = *st;
    return SDBE_OK;
}

int sdbe_get_stream_state(const sdbe_handle_t *h, sdbe_stream_state_t *st) {
    if (!h || !st) return SDBE_ERR_ARG;
    *st = h->st;
    return SDBE_OK;
}

int sdbe_stream_state_from_template(const void *tmpl, int32_t n_words_per_chan, sdbe_stream_state_t *st,
                                    int32_t *skip_int32, int32_t *skip_frames) {
    if (!tmpl || !st || n_words_per_chan <= 0) return SDBE_ERR_ARG;
    uint32_t w[8];
    memcpy(w, tmpl, sizeof(w));
    const int32_t words_per_frame = OUT_FRAME_PAYLOAD / 4;
    int32_t num_usable = n_words_per_chan / words_per_frame;
    uint32_t secs = w[0] & 0x3FFFFFFFu;
    uint32_t df = w[1] & 0xFFFFFFu;
    int32_t sk_frames = 0;
    if (df > (uint32_t)(125000 - num_usable)) {
        sk_frames = 125000 - (int32_t)df;
        df = 0;
        secs++;
    }
    const uint64_t psn_tmpl = (uint64_t)w[6] | ((uint64_t)w[7] << 32);
    const uint64_t skip_ps = psn_tmpl + 3715791ull;
    const int skip_samples = (int)floor((double)skip_ps * 4096e6 / 1e12);
    const int sk_int32 = (int)floor((double)skip_samples / (8.0 / 2.0) / (double)sizeof(int32_t));
    memset(st, 0, sizeof(*st));
    st->secs = secs & 0x3FFFFFFFu;
    st->df = df;
    st->ref_epoch = (w[1] >> 24) & 0x3Fu;
    st->psn = 1;
    if (skip_int32) *skip_int32 = sk_int32;
    if (skip_frames) *skip_frames = sk_frames;
    return SDBE_OK;
}

int sdbe_unpack_group_int8(sdbe_handle_t *h, const void *d_pkts, int32_t n_pkts, uint64_t b_zero,
                           int8_t *d_beng0, int8_t *d_beng1, uint8_t *d_bitmap, void *stream) {
    if (!h || !d_beng0 || !d_beng1 || (n_pkts > 0 && !d_pkts)) return SDBE_ERR_ARG;
    if (n_pkts < 0 || n_pkts > MAP_N) return SDBE_ERR_CAPACITY;
    HIP_TRY(h, hipSetDevice(h->device));
    hipStream_t s = (hipStream_t)stream;
    int rc = unpack_to_rows(h, d_pkts, n_pkts, b_zero, d_bitmap, s);
    if (rc) return rc;
    const size_t pol_bytes = (size_t)NSNAP_GROUP * NCH * 2;
    HIP_TRY(h, hipMemcpyAsync(d_beng0, h->beng, pol_bytes, hipMemcpyDeviceToDevice, s));
    HIP_TRY(h, hipMemcpyAsync(d_beng1, h->beng + pol_bytes, pol_bytes, hipMemcpyDeviceToDevice, s));
    h->n_packets_last = n_pkts;
    h->n_bp_last = 0;
    h->have_group = 1;
    return finish_call(h, s);
}

int sdbe_process_group(sdbe_handle_t *h, const void *d_pkts, int32_t n_pkts, uint64_t b_zero, void *d_out,
                       int64_t out_pol_stride_bytes, uint8_t *d_bitmap, void *stream) {
    if (!h || !d_out || (n_pkts > 0 && !d_pkts)) return SDBE_ERR_ARG;
    if (n_pkts < 0 || n_pkts > MAP_N) return SDBE_ERR_CAPACITY;
    const int frame_bytes = h->cfg.out_mode == SDBE_OUT_FRAMES ? OUT_FRAME_PAYLOAD + 32 : OUT_FRAME_PAYLOAD;
    if (out_pol_stride_bytes < (int64_t)SDBE_OUT_FRAMES_PER_GROUP * frame_bytes) return SDBE_ERR_ARG;
    HIP_TRY(h, hipSetDevice(h->device));
    hipStream_t s = (hipStream_t)stream;
    int rc = unpack_to_rows(h, d_pkts, n_pkts, b_zero, d_bitmap, s);
    if (rc) return rc;
    SrcDesc sd = {h->beng, 1};
    OutDesc od;
    od.base = (uint8_t *)d_out;
    od.pol_stride = out_pol_stride_bytes;
    od.frame_bytes = frame_bytes;
    od.hdr_bytes = h->cfg.out_mode == SDBE_OUT_FRAMES ? 32 : 0;
    od.group_mode = 1;
    od.f32 = nullptr;
    rc = run_resampler(h, sd, 2 * SDBE_BLOCKS_PER_GROUP, od, nullptr, s);
    if (rc) return rc;
    if (h->cfg.out_mode == SDBE_OUT_FRAMES) {
        HdrDesc hd;
        hd.secs = h->st.secs;
        hd.df = h->st.df;
        hd.ref_epoch = h->st.ref_epoch;
        hd.psn = h->st.psn;
        hd.station = h->cfg.station_id;
        hd.bdc = h->cfg.bdc_sideband;
        hd.rx = h->cfg.rx_sideband;
        hd.n_frames_per_pol = SDBE_OUT_FRAMES_PER_GROUP;
        hd.n_pols = 2;
        LAUNCH(h, KID_HEADERS, 2 * SDBE_BLOCKS_PER_GROUP, s, k_headers,
               dim3((2 * SDBE_OUT_FRAMES_PER_GROUP + 255) / 256), dim3(256), od, hd);
        HIP_TRY(h, hipGetLastError());
    }
    advance_state(h->st, SDBE_OUT_FRAMES_PER_GROUP);
    h->n_packets_last = n_pkts;
    h->n_bp_last = 2 * SDBE_BLOCKS_PER_GROUP;
    h->have_group = 1;
    return finish_call(h, s);
}

int sdbe_process_spectra(sdbe_handle_t *h, const int8_t *d_beng, int32_t n_bp, void *d_out, void *stream) {
    if (!h || !d_beng || !d_out || n_bp <= 0) return SDBE_ERR_ARG;
    if (n_bp > h->cfg.max_spectra_blockpols) return SDBE_ERR_CAPACITY;
    HIP_TRY(h, hipSetDevice(h->device));
    hipStream_t s = (hipStream_t)stream;
    const int frame_bytes = h->cfg.out_mode == SDBE_OUT_FRAMES ? OUT_FRAME_PAYLOAD + 32 : OUT_FRAME_PAYLOAD;
    SrcDesc sd = {d_beng, 0};
    OutDesc od;
    od.base = (uint8_t *)d_out;
    od.pol_stride = 0;
    od.frame_bytes = frame_bytes;
    od.hdr_bytes = h->cfg.out_mode == SDBE_OUT_FRAMES ? 32 : 0;
    od.group_mode = 0;
    od.f32 = nullptr;
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(256), 0, s, h->cnt, CNT_N, 0);
    int rc = run_resampler(h, sd, n_bp, od, nullptr, s);
    if (rc) return rc;
    if (h->cfg.out_mode == SDBE_OUT_FRAMES) {
        HdrDesc hd;
        hd.secs = h->st.secs;
        hd.df = h->st.df;
        hd.ref_epoch = h->st.ref_epoch;
        hd.psn = h->st.psn;
        hd.station = h->cfg.station_id;
        hd.bdc = h->cfg.bdc_sideband;
        hd.rx = h->cfg.rx_sideband;
        hd.n_frames_per_pol = n_bp * OUT_FRAMES_BLOCK;
        hd.n_pols = 1;
        LAUNCH(h, KID_HEADERS, n_bp, s, k_headers, dim3((hd.n_frames_per_pol + 255) / 256), dim3(256), od, hd);
        HIP_TRY(h, hipGetLastError());
    }
    advance_state(h->st, (int64_t)n_bp * OUT_FRAMES_BLOCK);
    h->n_packets_last = 0;
    h->n_bp_last = n_bp;
    h->have_group = 0;
    return finish_call(h, s);
}

int sdbe_resample_spectra_f32(sdbe_handle_t *h, const int8_t *d_beng, int32_t n_bp, float *d_samples,
                              double *d_stats, void *stream) {
    if (!h || !d_beng || !d_samples || n_bp <= 0) return SDBE_ERR_ARG;
    if (n_bp > h->cfg.max_spectra_blockpols) return SDBE_ERR_CAPACITY;
    HIP_TRY(h, hipSetDevice(h->device));
    hipStream_t s = (hipStream_t)stream;
    SrcDesc sd = {d_beng, 0};
    OutDesc od;
    memset(&od, 0, sizeof(od));
    od.f32 = d_samples;
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(256), 0, s, h->cnt, CNT_N, 0);
    int rc = run_resampler(h, sd, n_bp, od, d_stats, s);
    if (rc) return rc;
    h->n_packets_last = 0;
    h->n_bp_last = n_bp;
    h->have_group = 0;
    return finish_call(h, s);
}

int sdbe_get_stats(sdbe_handle_t *h, sdbe_stats_t *o) {
    if (!h || !o) return SDBE_ERR_ARG;
    HIP_TRY(h, hipSetDevice(h->device));
    if (h->last_stream || h->have_group || h->n_bp_last) HIP_TRY(h, hipStreamSynchronize(h->last_stream));
    memset(o, 0, sizeof(*o));
    o->n_packets = h->n_packets_last;
    o->n_invalid = h->h_cnt[CNT_INVALID];
    o->n_out_of_range = h->h_cnt[CNT_RANGE];
    o->n_unique = h->h_cnt[CNT_UNIQUE];
    o->n_duplicate = h->h_cnt[CNT_ACCEPTED] - h->h_cnt[CNT_UNIQUE];
    o->n_missing = h->have_group ? MAP_N - h->h_cnt[CNT_UNIQUE] : 0;
    o->n_blockpols = h->n_bp_last;
    o->latched_threshold = h->h_latch[0];
    o->latched_offset = h->h_latch[1];
    return SDBE_OK;
}

int sdbe_profile_enable(sdbe_handle_t *h, int enable) {
    if (!h) return SDBE_ERR_ARG;
    h->prof = enable ? 1 : 0;
    return SDBE_OK;
}

int sdbe_profile_read(sdbe_handle_t *h, sdbe_kernel_time_t *out, int32_t max_records, int32_t *n_records) {
    if (!h || !n_records || (max_records > 0 && !out)) return SDBE_ERR_ARG;
    HIP_TRY(h, hipSetDevice(h->device));
    double ms[KID_N] = {0}, units[KID_N] = {0};
    int launches[KID_N] = {0};
    for (auto &r : h->recs) {
        HIP_TRY(h, hipEventSynchronize(r.e1));
        float t = 0.f;
        HIP_TRY(h, hipEventElapsedTime(&t, r.e0, r.e1));
        ms[r.kid] += t;
        units[r.kid] += r.units;
        launches[r.kid] += 1;
        h->pool.push_back(r.e0);
        h->pool.push_back(r.e1);
    }
    h->recs.clear();
    int n = 0;
    for (int k = 0; k < KID_N; ++k) {
        if (!launches[k]) continue;
        if (n < max_records) {
            memset(&out[n], 0, sizeof(out[n]));
            strncpy(out[n].name, kKernelNames[k], sizeof(out[n].name) - 1);
            out[n].launches = launches[k];
            out[n].total_ms = ms[k];
            out[n].units = units[k];
        }
        ++n;
    }
    *n_records = n;
    return SDBE_OK;
}

}  // extern "C"
