// gfx950 implementation of the SWARM data back-end GPU stage behind the C-ABI
// of include/sdbe_stage.h.
//
// Pipeline per group (39 B-frames, both polarisations, 256 block-pols):
//   k_clear_map + k_packet_map  packet index map [frame][cid][fid], first packet
//                               wins; invalid / out-of-window packets counted
//   k_reorder                   LDS transpose of 16 packets per (frame, cid pair)
//                               into the reference int8 [snapshot][channel][im,re]
//                               rows, missing packets -> 0 + 0j, fid bitmap
//                               (int8 unpack entry point, spectra calls)
//   k_group_bitmap              fid bitmap and unique count only (groups whose
//                               first kernel reads the packets)
//   (k_reorder_8x1024, k_group_bitmap_8x1024: the same for 8 x 1024 packets)
// then per chunk of block-pols:
//   k_stage_a    decode + C2R pre-processing of each 32768-point snapshot
//                transform + 39-point DFT across the block's snapshots
//                (k_stage_a_pk / k_stage_a_pk8: the same, decoding the group's
//                packets in place)
//   k_stage_b    per 39-DFT row k1 = 1..38: 16384-point inverse FFT, inter-pass
//                twiddle, 16384-point forward FFT  => packed 1,277,952-point R2C
//                (row 0, whose twiddle is 1, is 2^14 times stage A's row 0 and is
//                written by stage A)
//   k_stage_c1   R2C post-processing, zero padding, C2R pre-processing of the
//                2,097,152-point output transform, first (64-point) pass of
//                its packed 2^20 inverse FFT; FP64 Parseval partial sums
//   k_stats      per block-pol mean / ddof=1 sigma in FP64, PER_BLOCK and FIXED
//                thresholds (k_thresh: the REFERENCE latch walk, when needed)
//   k_stage_c2   second (16384-point) pass, scaling, 2-bit quantisation, rows in
//                pairs so that each stage byte holds four codes
//   k_pack       stage bytes to VDIF payloads in stream order, frame headers
// (band-trimmed rates: sdbe_trim.hpp)
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <new>
#include <vector>

#include "sdbe_device.hpp"
#include "sdbe_mfma.hpp"
#include "../../include/sdbe_stage.h"

using namespace sdbe;

namespace {

constexpr int PKT_WORDS = 264;
constexpr int NFR = 39;                       // frames per group of the 64/39 resampler
constexpr int NSNAP_GROUP = NFR * 128;        // 4992 snapshots per group
constexpr int MAP_N = NFR * 256 * 8;          // 79,872 packet slots per group
// A band-trimmed handle's group is D B-frames (BENG_FRAMES_PER_GROUP =
// DECIMATION_FACTOR, sample_rates.h:67): at most 143 (11/11).
constexpr int NFR_MAX = 143;
constexpr int MAP_NONE = 0x7FFFFFFF;
// Packet geometry SDBE_GEOM_8X1024 (the "8 snapshots x 1024 channels" packets
// of the benchmark's configuration text; no reference counterpart): a packet
// (frame, cid, fid) carries snapshots 8 fid .. 8 fid + 7 of its B-frame for
// channels 1024 cid .. 1024 cid + 1023, one byte per channel and snapshot with
// the reference's byte encoding; 256 packets per B-frame, 8224 bytes each.
constexpr int G2_CIDS = 16, G2_FIDS = 16, G2_SNAPS = 8;
constexpr int G2_PKT_WORDS = 8 + G2_SNAPS * NCH / G2_CIDS / 4;   // 2056
// Stage C1 tiles: columns in residue classes mod 39, 16 per tile
// (c = a + 39 (16 u + j)), 14 tiles per class, 39 classes.
constexpr int C1_CLASS_TILES = 14;
constexpr int C1_TILES = NBLK * C1_CLASS_TILES;   // 546
constexpr int C1_TILES_PER_WG = 8;                 // default tiles per stage C1 workgroup
constexpr int C1_INTERLEAVE = 16;                  // default stage C1 workgroups per group of adjacent tiles
constexpr int BC_ROW = C1_TILES * 32;            // 17472 slots per row of the first-pass output
constexpr int OUT_FRAME_PAYLOAD = 8192;
constexpr int OUT_FRAMES_BLOCK = 64;

enum { CNT_INVALID = 0, CNT_RANGE = 1, CNT_ACCEPTED = 2, CNT_UNIQUE = 3, CNT_N = 8 };

// Pointers read from a descriptor are generic to the compiler, which would
// then emit flat memory instructions (counted on both the vector-memory and
// LDS counters, so that every load waits for the previous ones).  These casts
// state that they address global memory.
#define GAS __attribute__((address_space(1)))
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
template <typename T>
__device__ __forceinline__ GAS T *gptr(T *p) {
    return (GAS T *)p;
}

// ---------------------------------------------------------------- unpack side
__global__ void __launch_bounds__(256) k_fill_i32(int *p, int n, int val) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = val;
}

// A group's packet map to MAP_NONE and its counters to zero, in one launch.
__global__ void __launch_bounds__(256) k_clear_map(int *__restrict__ map, int n, int *__restrict__ cnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) map[i] = MAP_NONE;
    if (i < CNT_N) cnt[i] = 0;
}

// One thread per packet: header words 0, 4, 5 give invalid, B-count, cid, fid
// (vdif_inout_gpu_thread.cu:353-369,437 restated).  Valid packets inside the
// group's window of nfr frames claim their map slot [frame][cid][fid] with
// atomicMin -> lowest index wins.  Geometry: n_cid x n_fid packets per frame
// of pkt_words words each (reference: 256 x 8 x 264; 8 x 1024: 16 x 16 x 2056).
__global__ void __launch_bounds__(256)
k_packet_map(const uint32_t *__restrict__ pk, int n, uint32_t bzero32, int nfr, int n_cid, int n_fid,
             int pkt_words, int *__restrict__ map, int *__restrict__ cnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool inval = false, range = false, acc = false;
    if (i < n) {
        const uint32_t *h = pk + (size_t)i * pkt_words;
        const uint32_t w0 = h[0], w4 = h[4], w5 = h[5];
        const uint32_t cid = w5 & 0xFFu, fid = (w5 >> 16) & 0xFFu;
        const uint32_t bc32 = (w5 >> 24) + ((w4 & 0xFFFFFFu) << 8);
        const int32_t rel = (int32_t)(bc32 - bzero32);
        if (w0 >> 31) {
            inval = true;
        } else if (rel < 0 || rel >= nfr || cid >= (uint32_t)n_cid || fid >= (uint32_t)n_fid) {
            range = true;
        } else {
            acc = true;
            atomicMin(map + ((rel * n_cid + (int)cid) * n_fid + (int)fid), i);
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
// layout [pol][snapshot][16384][im, re] (its beng_0 / beng_1 buffers, nsnap
// snapshots per polarisation), a missing packet leaving 0 + 0j; plus the fid
// bitmap and the unique count.
__global__ void __launch_bounds__(256)
k_reorder(const uint32_t *__restrict__ pk, const int *__restrict__ map, int nsnap, int8_t *__restrict__ beng,
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
        const size_t row = (size_t)pol * nsnap + (size_t)frame * 128 + sn;
        uint32_t *dst = reinterpret_cast<uint32_t *>(beng + (row * NCH + (size_t)h * 128 + ch0) * 2);
        *dst = im0 | (re0 << 8) | (im1 << 16) | (re1 << 24);
    }
}

// Geometry 8 x 1024: workgroup = packet slot (cid, fid, frame).  Snapshot row
// sn of the payload is 256 words; word w holds channels 4 w .. 4 w + 3 with
// channel 4 w + j in byte 3 - j (the reference's order within a word), each
// byte decoded as in k_reorder.  Thread t takes word t of the 8 rows and
// writes the (im, re) pairs of its 4 channels for both polarisations as
// 8-byte stores; a missing packet reads as 0xAA bytes (0 + 0j).  The slot of
// fid 0 also writes the (frame, cid) bitmap, uint16 with bit fid set per
// packet present, and adds the slots found to the unique count.
__global__ void __launch_bounds__(256)
k_reorder_8x1024(const uint32_t *__restrict__ pk, const int *__restrict__ map, int nsnap,
                 int8_t *__restrict__ beng, uint16_t *__restrict__ bitmap, int *__restrict__ cnt) {
    const int cid = blockIdx.x, fid = blockIdx.y, frame = blockIdx.z, t = threadIdx.x;
    const int *slots = map + (frame * G2_CIDS + cid) * G2_FIDS;
    const int p = slots[fid];
    if (fid == 0 && t < 64) {
        const unsigned long long b = __ballot(t < G2_FIDS && slots[t < G2_FIDS ? t : 0] != MAP_NONE);
        if (t == 0) {
            bitmap[frame * G2_CIDS + cid] = (uint16_t)b;
            if (b) atomicAdd(cnt + CNT_UNIQUE, __popcll(b));
        }
    }
    uint32_t words[G2_SNAPS];
    if (p != MAP_NONE) {   // workgroup-uniform; no packet is read for a missing slot
        const GAS uint32_t *src = gptr(pk) + (size_t)p * G2_PKT_WORDS + 8 + t;
#pragma unroll
        for (int sn = 0; sn < G2_SNAPS; ++sn) words[sn] = src[256 * sn];
    } else {
#pragma unroll
        for (int sn = 0; sn < G2_SNAPS; ++sn) words[sn] = 0xAAAAAAAAu;
    }
#pragma unroll
    for (int sn = 0; sn < G2_SNAPS; ++sn) {
        const uint32_t word = words[sn];
        uint32_t q0[2] = {0u, 0u}, q1[2] = {0u, 0u};   // pol 0, pol 1: channels 4t .. 4t + 3 as (im, re) bytes
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t b = (word >> (8 * (3 - j))) & 0xFFu;
            const uint32_t im0 = (uint32_t)(uint8_t)(int8_t)((int)((b >> 4) & 3u) - 2);
            const uint32_t re0 = (uint32_t)(uint8_t)(int8_t)((int)(b >> 6) - 2);
            const uint32_t im1 = (uint32_t)(uint8_t)(int8_t)((int)(b & 3u) - 2);
            const uint32_t re1 = (uint32_t)(uint8_t)(int8_t)((int)((b >> 2) & 3u) - 2);
            q0[j >> 1] |= (im0 | (re0 << 8)) << (16 * (j & 1));
            q1[j >> 1] |= (im1 | (re1 << 8)) << (16 * (j & 1));
        }
        const size_t row = (size_t)frame * 128 + (size_t)fid * G2_SNAPS + sn;
        const size_t ch = (size_t)cid * (NCH / G2_CIDS) + 4 * t;
        *reinterpret_cast<GAS uint64_t *>(gptr(beng) + (row * NCH + ch) * 2) = q0[0] | ((uint64_t)q0[1] << 32);
        *reinterpret_cast<GAS uint64_t *>(gptr(beng) + (((size_t)nsnap + row) * NCH + ch) * 2) =
            q1[0] | ((uint64_t)q1[1] << 32);
    }
}

// ---------------------------------------------------------------- resampler

// One block-pol of a call: where its spectra come from, where its frames go,
// how it is thresholded and how its frames are stamped.  Built on the host for
// every call and copied to the device ahead of the kernels.
struct BpDesc {
    const int8_t *rows;      // 39 rows of int8 [16384][2 = (im, re)]
    uint8_t *out;            // first of its 64 frames (header included in FRAMES mode) or float samples
    uint64_t psn;            // VDIF time of its first frame
    uint32_t secs, df;
    uint32_t w1hi;           // ref_epoch << 24
    uint32_t w3;             // stationID | (bps = 1) << 26
    uint32_t w4;             // pol | bdc << 1 | rx << 2 | (edv5 = 2) << 24
    int16_t flip, pol, policy, slot;
    float fixed_th, fixed_off;
    int32_t blk;             // block index within its group (re-framed group output)
};
static_assert(sizeof(BpDesc) == 64, "BpDesc layout");


// Word 3 of a raw buffer resource (untyped dword access, gfx9 family): loads
// through it are range-checked against the resource's size and read 0 past it.
constexpr int BUF_RSRC_WORD3 = 0x00020000;
// Cache-policy operand of the buffer intrinsics that selects nt (non-temporal)
// on gfx950.
constexpr int CPOL_NT = 2;

// Stage A's per-snapshot input x_s = sum_s + i T dif_s, with
//   a = (Ra, kz Ia), b = (kz Rb, kz Ib), sum = a + conj b, dif = a - conj b.
// sum and dif are small integers, exact in FP32 in any order of operations: the
// packet path passes 2-bit codes (value + 2) and folds the -2 offsets into the
// per-lane constants cs, cd; the int8 path passes values and zero constants.
// Both give exactly the same float inputs to the first rounding, cmul(T, dif).
// Z_s[k] = sum + i T dif from the channel pair's values X[k] = (Ra, Ia) and
// X[N2-k] = (Rb, Ib) (offsets removed), combined as integers first:
// sum = X[k] + conj X[N2-k] = (sx, sy), dif = X[k] - conj X[N2-k] = (dx, dy).
// For k = 0 the DC imaginary part is dropped and the Nyquist bin is zero (as
// numpy irfft): sum = dif = (Ra, 0).
__device__ __forceinline__ float2 a_combine(int sx, int sy, int dx, int dy, float2 T) {
    const float fx = (float)dx, fy = (float)dy;
    return make_float2(fmaf(-T.x, fy, fmaf(-T.y, fx, (float)sx)), fmaf(T.x, fx, fmaf(-T.y, fy, (float)sy)));
}

// The 39 snapshots' Z_s[k] from 2-bit codes (value = code - 2): code words wa
// (channel k) and wb (channel N2 - k) of snapshot s at word(s), field (im at bit
// sh, re at bit sh + 2).  k = 0 (one thread of the block-pol) takes its own loop.
template <typename Word>
__device__ __forceinline__ void a_inputs_codes(float2 (&x)[NBLK], int k, uint32_t sha, uint32_t shb, float2 T,
                                               Word word) {
    if (k != 0) {
#pragma unroll
        for (int s = 0; s < NBLK; ++s) {
            uint32_t wa, wb;
            word(s, wa, wb);
            const int Ia = (int)__builtin_amdgcn_ubfe(wa, sha, 2u), Ra = (int)__builtin_amdgcn_ubfe(wa, sha + 2u, 2u);
            const int Ib = (int)__builtin_amdgcn_ubfe(wb, shb, 2u), Rb = (int)__builtin_amdgcn_ubfe(wb, shb + 2u, 2u);
            x[s] = a_combine(Ra + Rb - 4, Ia - Ib, Ra - Rb, Ia + Ib - 4, T);
        }
    } else {
#pragma unroll
        for (int s = 0; s < NBLK; ++s) {
            uint32_t wa, wb;
            word(s, wa, wb);
            const int Ra = (int)__builtin_amdgcn_ubfe(wa, sha + 2u, 2u) - 2;
            x[s] = a_combine(Ra, 0, Ra, 0, T);
        }
    }
}

typedef unsigned int u32v2 __attribute__((__vector_size__(2 * sizeof(unsigned int))));

// V[bpl][k1][k] for k1 = 1..38: buffer stores, the row in the scalar offset,
// non-temporal so that the streamed output does not evict the input lines
// that the paired workgroup reads again.  Row k1 = 0 goes straight to stage
// B's output: its twiddle is 1, so stage B's inverse and forward 16384-point
// transforms cancel (FFT(IFFT(x)) = 16384 x) and Zf's row 0 is 2^14 x (exact).
__device__ __forceinline__ void a_store(float2 *__restrict__ V, float2 *__restrict__ Zf, int bpl, int k,
                                        const float2 (&x)[NBLK]) {
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(V + (size_t)bpl * NBLK * NCH), 0, NBLK * NCH * (int)sizeof(float2), BUF_RSRC_WORD3);
    const __amdgpu_buffer_rsrc_t rz = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(Zf + (size_t)bpl * NBLK * NCH), 0, NCH * (int)sizeof(float2), BUF_RSRC_WORD3);
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32v2, cscale(x[0], 16384.0f)), rz, 8 * k, 0, CPOL_NT);
#pragma unroll
    for (int k1 = 1; k1 < NBLK; ++k1)
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32v2, x[k1]), rs, 8 * k,
                                              k1 * NCH * (int)sizeof(float2), CPOL_NT);
}

// Stage A: thread per channel k of a block-pol.  For each snapshot s the
// packed C2R input Z_s[k] = (X[k] + conj(X[N2-k])) + i e^{2 pi i k/32768}
// (X[k] - conj(X[N2-k])) (DC imaginary part and the zero Nyquist bin dropped,
// as numpy irfft), then V[k1][k] = sum_s Z_s[k] e^{-2 pi i s k1/39}.
// A flipped (lower-sideband) block-pol reads channel 16383 - k in place of k
// (the Nyquist slot stays zero).
// Workgroups are paired (w, 63 - w) on one XCD back to back, so the mirror
// channels N2 - k a workgroup reads are the ones its partner reads as k and
// are served from that XCD's L2.
__global__ void __launch_bounds__(256, 4)
k_stage_a(const BpDesc *__restrict__ bpd, int nb, float2 *__restrict__ V, float2 *__restrict__ Zf, Tables tb) {
    const int L = xcd_contig((int)blockIdx.x, (NCH / 256) * nb);
    const int bpl = L >> 6, idx = L & 63;
    const int w = (idx & 1) ? 63 - (idx >> 1) : (idx >> 1);
    const int k = w * 256 + threadIdx.x;
    const GAS uint16_t *src = gptr(reinterpret_cast<const uint16_t *>(bpd[bpl].rows));
    const int flip = bpd[bpl].flip;
    const int km = (NCH - k) & (NCH - 1);      // k = 0 pairs with the zero Nyquist bin
    const int ka = flip ? NCH - 1 - k : k, kb = flip ? NCH - 1 - km : km;
    const float2 T = tw<+1>(tb.n32k, (uint32_t)k);
    float2 x[NBLK];
    if (k != 0) {
#pragma unroll
        for (int s = 0; s < NBLK; ++s) {
            const uint16_t ua = src[(size_t)s * NCH + ka], ub = src[(size_t)s * NCH + kb];
            const int Ia = (int8_t)(ua & 0xFF), Ra = (int8_t)(ua >> 8);   // (im, re) values
            const int Ib = (int8_t)(ub & 0xFF), Rb = (int8_t)(ub >> 8);
            x[s] = a_combine(Ra + Rb, Ia - Ib, Ra - Rb, Ia + Ib, T);
        }
    } else {
#pragma unroll
        for (int s = 0; s < NBLK; ++s) {
            const int Ra = (int8_t)(src[(size_t)s * NCH + ka] >> 8);
            x[s] = a_combine(Ra, 0, Ra, 0, T);
        }
    }
    dft39<-1>(x);
    a_store(V, Zf, bpl, k, x);
}

// Stage A of a whole group, reading the packets in place of the int8 rows
// (k_reorder's decode fused in).  Block-pol (blk, pol) covers group snapshots
// S0 = 39 blk .. S0 + 38, which lie in B-frames f0 = S0 / 128 and f0 + 1 (from
// s = 128 (f0 + 1) - S0 on).  Channel c sits in packet (frame, cid, fid) with
// g = c >> 3, cid = 2 (g >> 4) + (g & 1), fid = (g >> 1) & 7: payload words
// 2 sn, 2 sn + 1 of snapshot sn hold its 8 channels, channel c in word
// (c & 7) >> 2, byte 3 - (c & 3), bits from the LSB pol1 im, pol1 re, pol0 im,
// pol0 re (vdif_inout_gpu_thread.cu:442, 464-494).  A missing packet reads as
// 0xAA bytes, which decode to 0 + 0j (as in k_reorder).
// The workgroup's 256 channels k lie in 32 channel groups and their mirrors
// (or, flipped, their reversals) in 33: the 39 snapshot word pairs of those 65
// groups are staged in LDS with coalesced 16-byte loads (the packet indices
// first), then every thread decodes its two channels from LDS.  A group's
// pairs come from rows r0 = S0 - 128 f0 .. 127 of its frame-f0 packet, then
// rows 0 .. of its frame-f1 packet; with snapshot s at LDS word 2 s + 2 (r0 & 1)
// both runs start on 16-byte boundaries (1056-byte packets, 32-byte header),
// so 20 units of 4 words cover a group.  Same workgroup mapping and arithmetic
// as k_stage_a.
constexpr int APK_GA = 32, APK_GB = 33, APK_G = APK_GA + APK_GB;
constexpr int APK_U = 20;        // 16-byte units loaded per group
constexpr int APK_US = 21;       // units per group in LDS (84 words: conflict-free decode reads)

__global__ void __launch_bounds__(256, 2)
k_stage_a_pk(const BpDesc *__restrict__ bpd, int nb, float2 *__restrict__ V, float2 *__restrict__ Zf, Tables tb,
             const uint32_t *__restrict__ pk, const int *__restrict__ map) {
    __shared__ uint4 seg[APK_G * APK_US];   // [group slot][unit]: snapshot s at word 2 s + off
    __shared__ int pidx[APK_G * 2];         // [group slot][frame f0, f1] packet index
    const int L = xcd_contig((int)blockIdx.x, (NCH / 256) * nb);
    const int bpl = L >> 6, idx = L & 63;
    const int w = (idx & 1) ? 63 - (idx >> 1) : (idx >> 1);
    const int t = threadIdx.x, k = w * 256 + t;
    const int flip = bpd[bpl].flip, pol = bpd[bpl].pol;
    const int S0 = NBLK * bpd[bpl].blk, f0 = S0 >> 7, s_sw = 128 * (f0 + 1) - S0;
    const int f1 = min(f0 + 1, NFR - 1);
    // LDS word offset of snapshot 0; first payload word of the f0 run (16-byte
    // aligned); LDS word where the f1 run starts (past the group when s_sw >= 39)
    const int r0 = S0 - 128 * f0, off = 2 * (r0 & 1), a0 = 2 * r0 - off, E = 256 - a0;
    // first group of each range: k (32 groups), mirror / reversed mirror (33)
    const int ga0 = flip ? 2047 - 32 * w - 31 : 32 * w;
    const int gb0 = flip ? 32 * w - 1 : 2048 - 32 * w - 32;
    if (t < 2 * APK_G) {
        const int i = t >> 1, fr = (t & 1) ? f1 : f0;
        const int g = min(max(i < APK_GA ? ga0 + i : gb0 + (i - APK_GA), 0), NCH / 8 - 1);
        const int cid = 2 * (g >> 4) + (g & 1), fid = (g >> 1) & 7;
        pidx[t] = map[(fr * 256 + cid) * 8 + fid];
    }
    __syncthreads();
    for (int e = t; e < APK_G * APK_U; e += 256) {
        const int i = e / APK_U, j = 4 * (e - i * APK_U);
        const bool first = j < E;
        const int p = pidx[2 * i + (first ? 0 : 1)];
        const int word = first ? a0 + j : j - E;
        uint4 v = make_uint4(0xAAAAAAAAu, 0xAAAAAAAAu, 0xAAAAAAAAu, 0xAAAAAAAAu);
        if (p != MAP_NONE) v = *reinterpret_cast<const uint4 *>(pk + (size_t)p * PKT_WORDS + 8 + word);
        seg[i * APK_US + (j >> 2)] = v;
    }
    __syncthreads();
    const int km = (NCH - k) & (NCH - 1);      // k = 0 pairs with the zero Nyquist bin
    const int ka = flip ? NCH - 1 - k : k, kb = flip ? NCH - 1 - km : km;
    // word offsets into seg (as uint32), and the nibble's bit
    const uint32_t *seg32 = reinterpret_cast<const uint32_t *>(seg);
    const int sa = (ka >> 3) - ga0, sb = min(max((kb >> 3) - gb0, 0), APK_GB - 1);   // k = 0: kb unused
    const int oa = 4 * APK_US * sa + off + ((ka & 7) >> 2);
    const int ob = 4 * APK_US * (APK_GA + sb) + off + ((kb & 7) >> 2);
    const int pbit = pol == 0 ? 4 : 0;
    const int sha = 8 * (3 - (ka & 3)) + pbit, shb = 8 * (3 - (kb & 3)) + pbit;
    const float2 T = tw<+1>(tb.n32k, (uint32_t)k);
    float2 x[NBLK];
    // codes: im in the nibble's bits 0-1, re in bits 2-3
    a_inputs_codes(x, k, (uint32_t)sha, (uint32_t)shb, T, [&](int s, uint32_t &wa, uint32_t &wb) {
        wa = seg32[oa + 2 * s];
        wb = seg32[ob + 2 * s];
    });
    dft39<-1>(x);
    a_store(V, Zf, bpl, k, x);
}

// Stage A from packets of the 8 x 1024 geometry (the int8 rows are not
// materialised).  Group snapshot S = 39 blk + s sits in B-frame S >> 7, packet
// fid (S & 127) >> 3, payload row S & 7; channel c in packet cid c >> 10, row
// word (c & 1023) >> 2, byte 3 - (c & 3), decoded as in k_reorder_8x1024.  The
// workgroup's channels ka are 256 consecutive ones: one aligned 64-word run of
// a packet row per snapshot.  Their mirrors kb (N - k, or k - 1 flipped) lie in
// 65 words, staged as the 68-word run around them that starts on a 4-word
// boundary and may straddle two cids.  Payload rows start 16-byte aligned
// (8224-byte packets, 32-byte header, 1024-byte rows), so with a 16-byte
// aligned buffer the 39 x 132 words are staged in LDS with 16-byte coalesced
// loads (one packet lookup per snapshot and cid first); then every thread
// decodes its two channels from LDS.  Same workgroup mapping and arithmetic as
// k_stage_a_pk.
constexpr int AP8_QA = 16, AP8_QB = 17, AP8_Q = AP8_QA + AP8_QB;   // staged 4-word units per snapshot
constexpr int AP8_W = 4 * AP8_Q;                                     // staged words per snapshot (132)

__global__ void __launch_bounds__(256, 2)
k_stage_a_pk8(const BpDesc *__restrict__ bpd, int nb, float2 *__restrict__ V, float2 *__restrict__ Zf,
              Tables tb, const uint32_t *__restrict__ pk, const int *__restrict__ map) {
    __shared__ uint4 seg[NBLK * AP8_Q];      // [snapshot][unit]
    __shared__ int pidx[NBLK * 3];           // [snapshot][run A cid, run B first cid, run B second cid]
    const int L = xcd_contig((int)blockIdx.x, (NCH / 256) * nb);
    const int bpl = L >> 6, idx = L & 63;
    const int w = (idx & 1) ? 63 - (idx >> 1) : (idx >> 1);
    const int t = threadIdx.x, k = w * 256 + t;
    const int flip = bpd[bpl].flip, pol = bpd[bpl].pol;
    const int S0 = NBLK * bpd[bpl].blk;
    // first word (channel / 4) of each run: ka over [256 w, 256 w + 255] (flipped:
    // N - 1 - that); kb over [N - 256 w - 255, N - 256 w] (flipped: [256 w - 1,
    // 256 w + 254]), clamped into the row and aligned down to 4 words; k = 0
    // pairs with the zero Nyquist bin and reads nothing
    const int wa0 = (flip ? NCH - 256 * w - 256 : 256 * w) >> 2;
    const int wb0 = min(max((flip ? 256 * w - 1 : NCH - 256 * w - 255) >> 2, 0), NCH / 4 - 65) & ~3;
    const int ca = wa0 >> 8, cb0 = wb0 >> 8, cb1 = (wb0 + 4 * AP8_QB - 1) >> 8;
    if (t < NBLK * 3) {
        const int sn = t / 3, r = t - 3 * sn, S = S0 + sn;
        const int cid = r == 0 ? ca : (r == 1 ? cb0 : cb1);
        pidx[t] = map[((S >> 7) * G2_CIDS + cid) * G2_FIDS + ((S & 127) >> 3)];
    }
    __syncthreads();
    for (int e = t; e < NBLK * AP8_Q; e += 256) {
        const int sn = e / AP8_Q, q = e - sn * AP8_Q;
        const int W = q < AP8_QA ? wa0 + 4 * q : wb0 + 4 * (q - AP8_QA);
        const int p = pidx[3 * sn + (q < AP8_QA ? 0 : ((W >> 8) == cb0 ? 1 : 2))];
        const int row = (S0 + sn) & 7;
        uint4 v = make_uint4(0xAAAAAAAAu, 0xAAAAAAAAu, 0xAAAAAAAAu, 0xAAAAAAAAu);
        if (p != MAP_NONE)
            v = *reinterpret_cast<const uint4 *>(pk + (size_t)p * G2_PKT_WORDS + 8 + 256 * row + (W & 255));
        seg[e] = v;
    }
    __syncthreads();
    const int km = (NCH - k) & (NCH - 1);
    const int ka = flip ? NCH - 1 - k : k, kb = flip ? NCH - 1 - km : km;
    const uint32_t *seg32 = reinterpret_cast<const uint32_t *>(seg);
    // k = 0: kb unused
    const int oa = (ka >> 2) - wa0, ob = 4 * AP8_QA + min(max((kb >> 2) - wb0, 0), 4 * AP8_QB - 1);
    const int pbit = pol == 0 ? 4 : 0;
    const int sha = 8 * (3 - (ka & 3)) + pbit, shb = 8 * (3 - (kb & 3)) + pbit;
    const float2 T = tw<+1>(tb.n32k, (uint32_t)k);
    float2 x[NBLK];
    a_inputs_codes(x, k, (uint32_t)sha, (uint32_t)shb, T, [&](int s, uint32_t &wa, uint32_t &wb) {
        wa = seg32[AP8_W * s + oa];
        wb = seg32[AP8_W * s + ob];
    });
    dft39<-1>(x);
    a_store(V, Zf, bpl, k, x);
}

// The 8 x 1024 geometry's uint16 fid bitmap of every (frame, cid) and the
// count of filled packet slots (what k_reorder_8x1024 records, for the
// packet-reading stage A).  Workgroup = frame, thread = (cid, fid).
__global__ void __launch_bounds__(256)
k_group_bitmap_8x1024(const int *__restrict__ map, uint16_t *__restrict__ bitmap, int *__restrict__ cnt) {
    const int frame = blockIdx.x, t = threadIdx.x;
    const unsigned long long b = __ballot(map[frame * G2_CIDS * G2_FIDS + t] != MAP_NONE);
    const int l = t & 63;
    if ((l & 15) == 0) {
        const uint16_t m = (uint16_t)(b >> l);
        bitmap[frame * G2_CIDS + (t >> 4)] = m;
        if (m) atomicAdd(cnt + CNT_UNIQUE, __popc((unsigned)m));
    }
}

// The fid bitmap byte of every (frame, cid) of the group and the count of
// filled packet slots (what k_reorder records, for the packet-reading stage A).
__global__ void __launch_bounds__(256)
k_group_bitmap(const int *__restrict__ map, uint8_t *__restrict__ bitmap, int *__restrict__ cnt) {
    __shared__ int red[4];
    const int frame = blockIdx.x, cid = threadIdx.x;
    const int4 *m = reinterpret_cast<const int4 *>(map + (frame * 256 + cid) * 8);
    const int4 m0 = m[0], m1 = m[1];
    const int v[8] = {m0.x, m0.y, m0.z, m0.w, m1.x, m1.y, m1.z, m1.w};
    uint32_t byte = 0;
    int found = 0;
#pragma unroll
    for (int fid = 0; fid < 8; ++fid) {
        if (v[fid] != MAP_NONE) {
            byte |= 1u << fid;
            ++found;
        }
    }
    bitmap[frame * 256 + cid] = (uint8_t)byte;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) found += __shfl_xor(found, o, 64);
    if ((cid & 63) == 0) red[cid >> 6] = found;
    __syncthreads();
    if (cid == 0) {
        const int tot = red[0] + red[1] + red[2] + red[3];
        if (tot) atomicAdd(cnt + CNT_UNIQUE, tot);
    }
}


// Stage B: row k1 of block-pol bpl: w = IFFT16384(V[k1]); w[n] *= e^{-2 pi i n k1/638976};
// Zf[k1 + 39 k2] = FFT16384(w)[k2]   (packed forward DFT of the 1,277,952 samples)
// for k1 = 1..38 (stage A writes row 0, whose transforms cancel).
// The packed spectrum is stored row-major: Zf[k1 + 39 k2] at k1 * 16384 + k2.
// P = 0: FP32 VALU FFTs; P = FftMode: radix-16 passes on the matrix cores
// (fft16k_mfma).
// The workgroups are persistent (one per CU: the FFT's LDS fills a CU) and take
// row slots lin = blockIdx.x, blockIdx.x + gridDim.x, ...  With xcd_map the
// 38 rows of a block-pol all run on one XCD (workgroups are dealt round-robin
// over the 8 XCDs; requires the chunk and the grid to be multiples of 8).  The
// next row is loaded (buffer loads, no VALU address arithmetic) while the last
// pass, the conjugation and the stores of the current row run.
constexpr int B_ROWS = NBLK - 1;   // rows k1 = 1..38 of a block-pol
__device__ __forceinline__ void b_row(int lin, int xcd_map, int &bpl, int &k1) {
    if (xcd_map) {
        const int slot = lin >> 3;
        bpl = (lin & 7) + 8 * (slot / B_ROWS);
        k1 = 1 + slot % B_ROWS;
    } else {
        bpl = lin / B_ROWS;
        k1 = 1 + lin % B_ROWS;
    }
}

__device__ __forceinline__ void b_load(const float2 *__restrict__ V, int bpl, int k1, float2 (&x)[16]) {
    const int t = threadIdx.x;
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(V + ((size_t)bpl * NBLK + k1) * NCH), 0, NCH * (int)sizeof(float2), BUF_RSRC_WORD3);
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const float2 y = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(rs, 8 * t, 8192 * q, 0));
        x[q] = y;
    }
}

template <int P>
__global__ void __launch_bounds__(1024)
k_stage_b(const float2 *__restrict__ V, float2 *__restrict__ Zf, Tables tb, int xcd_map, int n_rows) {
    __shared__ float lds[LDS_FFT_N];
    __shared__ FftTw ftw;
    fft_tw_load(&ftw, tb.fft);
    int lin = blockIdx.x;
    float2 nx[16];
    if constexpr (P == 0) {
        int bpl, k1;
        b_row(lin, xcd_map, bpl, k1);
        b_load(V, bpl, k1, nx);
    }
    for (;;) {
        int k1, bpl;
        b_row(lin, xcd_map, bpl, k1);
        // workgroup-uniform: the row twiddles below are scalar loads
        k1 = __builtin_amdgcn_readfirstlane(k1);
        bpl = __builtin_amdgcn_readfirstlane(bpl);
        const int next = lin + (int)gridDim.x;
        int nbpl, nk1;
        b_row(next < n_rows ? next : lin, xcd_map, nbpl, nk1);
        nbpl = __builtin_amdgcn_readfirstlane(nbpl);
        nk1 = __builtin_amdgcn_readfirstlane(nk1);
        const int t = row_tid();
        float2 v[16];
        if constexpr (P == 0) {
#pragma unroll
            for (int q = 0; q < 16; ++q) v[q] = nx[q];
            fft16k_p123<+1>(v, lds, &ftw, t);
            fft16k_p4<+1>(v, &ftw, t);
            // e^{-2 pi i (t + 1024 q) k1 / M'} = A(t) * B(q); B is uniform over the workgroup
            const float2 A = tw<-1>(tb.nmp, (uint32_t)(t * k1));
#pragma unroll
            for (int q = 0; q < 16; ++q) v[q] = cmul(v[q], cmul(A, tw_hi<-1>(tb.nmp, (uint32_t)(q * k1))));
            // the second transform's third exchange is wave-local: the row comes
            // out in the order of fft16k_tord (element fft16k_tord(t) + 1024 q in v[q])
            fft16k_p12<-1>(v, lds, &ftw, t);
            fft16k_p3w<-1>(v, lds, &ftw, t);
            // the next row (unconditionally: the last row loads itself again, so
            // the registers are always overwritten); keep the loads here, the
            // scheduler would otherwise sink them below this row's stores
            b_load(V, nbpl, nk1, nx);
            __builtin_amdgcn_sched_barrier(0);
            fft16k_p4<-1>(v, &ftw, fft16k_tord(row_tid()));
        } else {
            // the MFMA path runs both transforms with the inverse-sign operands: the
            // forward FFT is conj(IFFT(conj(x)))
            // the matrix-core variants hold their operands across the passes and
            // load their own row here, straight into the matrix-core layout
            // (no registers to spare for a prefetch)
            MfmaW<+1, P> W;
            W.init(t);
            {
                // element 64 w + 16 g + n + 1024 (4 kb + j) of the row: one VGPR offset,
                // g in the immediate offset, j in the scalar offset
                const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
                    (void *)(V + ((size_t)bpl * NBLK + k1) * NCH), 0, NCH * (int)sizeof(float2), BUF_RSRC_WORD3);
                const int l = t & 63;
                const int vo = 8 * (64 * (t >> 6) + (l & 15) + 4096 * (l >> 4));
#pragma unroll
                for (int g = 0; g < 4; ++g)
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        v[4 * g + j] = __builtin_bit_cast(
                            float2, __builtin_amdgcn_raw_buffer_load_b64(rs, vo + 128 * g, 8192 * j, 0));
            }
            fft16k_mfma<+1, P>(v, lds, &ftw, W, t);
            // natural order -> MFMA layout through LDS; the twiddle of element
            // e = d + 1024 r is applied there: e^{-2 pi i e k1 / M'} = A(d) * B(r)
            const int rd = lds_re(t), kb = (t & 63) >> 4;
            __syncthreads();
#pragma unroll
            for (int q = 0; q < 16; ++q) lds_put(lds, rd + 2048 * q, v[q]);
            __syncthreads();
            float2 Br[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) Br[j] = tw_hi<-1>(tb.nmp, (uint32_t)((4 * kb + j) * k1));
            const int rb = mfma_rd_base(t);
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const float2 Ad = tw<-1>(tb.nmp, (uint32_t)((mfma_elem(g, 0, t) & 1023) * k1));
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    v[4 * g + j] = cconj(cmul(lds_get(lds, mfma_rd(rb, g, j)), cmul(Ad, Br[j])));
                }
            }
            fft16k_mfma<+1, P>(v, lds, &ftw, W, t);
#pragma unroll
            for (int q = 0; q < 16; ++q) v[q] = cconj(v[q]);
        }
        // row-major [k1][k2]: Zf[k1 + 39 k2] lives at k1 * 16384 + k2
        float2 *dst = Zf + ((size_t)bpl * NBLK + k1) * NCH + (P == 0 ? fft16k_tord(t) : t);
#pragma unroll
        for (int q = 0; q < 16; ++q) dst[1024 * q] = v[q];
        if (next >= n_rows) break;
        lin = next;
    }
}

// Stage C1: workgroup = 16 column pairs (c, 16384 - c) of the [64][16384] view
// (k = c + 16384 kh) of the packed 2^20-point C2R input, all 64 rows.
//   Y[m]  = (Zf[m] + conj Zf[M'-m]) - i e^{-2 pi i m/M} (Zf[m] - conj Zf[M'-m]),  m <= M'
// (twice the R2C spectrum: the factor 1/2 is left to the statistics and stage C2's
// output scale, both exact powers of two)
//   Zc[k] = (C[k] + conj C[P-k]) + i e^{2 pi i k/L} (C[k] - conj C[P-k]),  C = Y zero-padded
//   Bc[tx][c] = sum_kh Zc[c + 16384 kh] e^{2 pi i kh tx/64}
// (the inter-pass twiddle e^{2 pi i c tx/P} is applied by stage C2 on load)
// and the Parseval sum  sum_m w_m |Y[m]|^2 (w_0 = 1 on Re Y[0], else 2): five-term
// FP32 partials per thread, accumulated in FP64.
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

// Tiles in (block-pol, a, u) order: tiles u and u + 1 of a class share the
// 128-byte lines at the ends of their runs.  Workgroups are contiguous per XCD,
// and `ilv` consecutive ones form a group that covers ilv * tpw consecutive
// tiles: at step i, workgroup r of the group runs tile r + ilv i, so the group's
// workgroups (dispatched together on one XCD) read neighbouring tiles at the
// same time and fetch the shared lines once into L2.  With ilv = 1 a workgroup
// runs tpw consecutive tiles, and a shared line must survive in L2 until its
// next tile.  The next tile's packed-spectrum values and column twiddles are
// loaded while the current tile is computed.
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
    // m = a + 16384 r + 39 (16 u + j) sits at row (a + 16384 r) mod 39, column
    // (a + 16384 r) / 39 + 16 u + j of the row-major store; the mirror column
    // 16384 - c at m = 16384 (r + 1) - a - 39 (16 u + j).  With 16384 = 39 * 420 + 4:
    //   a + 16384 r          = 39 (420 r) + S,         S = a + 4 r
    //   16384 (r + 1) - a    = 39 (420 r + 419) + S,   S = 4 r + 43 - a
    // and 0 <= S <= 199, so S / 39 = (S * 1681) >> 16 exactly: no branches and
    // no divisions.  Row 39 holds only m = M' (c = 0), which the store keeps at
    // address 0 (as m = 0); values that are not used also read address 0, so the
    // loads are unconditional and all five are in flight together.
    const int s0 = (mirror ? 43 - a : a) + 4 * rw;
    const int c0 = (mirror ? 419 - 16 * u - j16 : 16 * u + j16) + 420 * rw;
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(Zf + (size_t)x.bpl * MP), 0, MP * (int)sizeof(float2), BUF_RSRC_WORD3);
#pragma unroll
    for (int it = 0; it < 5; ++it) {
        const int r = it * 8 + rw;
        const bool ok = x.valid && r < NBLK;
        const int S = s0 + 32 * it;
        const int Q = (S * 1681) >> 16;
        const int addr = (S - NBLK * Q) * NCH + Q + c0 + 3360 * it;   // 3360 = 420 * 8
        z[it] = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(rs, ok ? 8 * addr : 0, 0, 0));
    }
    cm = tb.c1col[x.c];
    cl = tb.c1col[NCH + x.c];
}

__global__ void __launch_bounds__(256, 5)
k_stage_c1(const float2 *__restrict__ Zf, float *__restrict__ Bc, double *__restrict__ part, Tables tb, int nb,
           int tpw, int ilv) {
    // sz (packed-spectrum tile) is dead once Y is formed and sc (first-pass
    // tile) is written after that barrier, so they share storage: 28 KB of LDS
    // per workgroup, five workgroups per CU.
    __shared__ float2 szc[32 * 65];
    __shared__ float2 sy[32][41];
    float2 (*sz)[41] = reinterpret_cast<float2 (*)[41]>(szc);
    float2 (*sc)[65] = reinterpret_cast<float2 (*)[65]>(szc);
    __shared__ float2 rowm[40], rowl[64], w64[64];
    __shared__ double red[2][4];
    const int t = threadIdx.x, col = t & 31, rw = t >> 5, pcol = col ^ 16;
    const int n_tiles = C1_TILES * nb;
    // the grid is a whole number of groups; tiles past n_tiles are not run
    const int wg = xcd_contig((int)blockIdx.x, (int)gridDim.x), grp = wg / ilv;
    const int first = grp * ilv * tpw + (wg - grp * ilv);
    const int n_run = first < n_tiles ? min(tpw, (n_tiles - 1 - first) / ilv + 1) : 0;
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
    // Parseval sums of the tiles run so far for the current block-pol: reduced
    // across the workgroup once per block-pol it touches, into the slot of that
    // block-pol's last tile here; its other slots get zeros, so the per-tile
    // slots still sum to the block-pol's total
    double acc_run = 0.0, dc_run = 0.0;
    if (n_run > 0) c1_fetch(Zf, tb, first, col, rw, cur, z, cm, cl);
    for (int i = 0; i < n_run; ++i) {
        const int lt = first + ilv * i;
        const bool more = i + 1 < n_run;
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
        if (more) c1_fetch(Zf, tb, lt + ilv, col, rw, cur, z, cm, cl);
        // Y rows 0..39 (zero above M') + Parseval; this thread's rows r = 8 it + rw
        // also stay in yv for its Zc rows below.  The Parseval terms of its five
        // rows are summed in FP32 (relative error ~2^-22 on a five-term partial,
        // ~1e-10 on the block's sum) before the FP64 accumulation across tiles.
        float accf = 0.f;
        double dc = 0.0;
        float2 yv[5];
#pragma unroll
        for (int it = 0; it < 5; ++it) {
            const int r = it * 8 + rw;
            float2 Y = make_float2(0.f, 0.f);
            if (x.valid && (r < NBLK || c == 0)) {
                const int pr = (c == 0) ? (39 - r) : (38 - r);
                const float2 zm = sz[col][r];
                const float2 zp = cconj(sz[pcol][pr]);
                Y = cadd(cadd(zm, zp), mul_si<-1>(cmul(cmul(cmc, rowm[r]), csub(zm, zp))));
                if (!x.dup) {
                    if (c == 0 && r == 0) {
                        accf += 0.5f * Y.x * Y.x;   // weight 1 of 2 (the factor 2 is applied below)
                        dc = (double)Y.x;
                    } else {
                        accf += Y.x * Y.x + Y.y * Y.y;
                    }
                }
            }
            yv[it] = Y;
            sy[col][r] = Y;
        }
        __syncthreads();
        // Zc rows kh = 8 it + rw (this thread's Y rows for it < 5, the partner
        // column's from LDS), which are the inputs of this thread's first radix-8
        // pass of the 64-point inverse DFT over kh: 8 threads x radix 8, two
        // passes (lanes of a half-wave take different columns: odd row stride, no
        // bank conflicts).  Thread (col, j = rw) ends holding output rows j + 8 r
        // in v[r]: the rows it stores below, straight from its registers.
        float2 v[8];
#pragma unroll
        for (int it = 0; it < 8; ++it) {
            const int kh = it * 8 + rw;
            const int pr = (c == 0) ? (64 - kh) : (63 - kh);
            const float2 ck = it < 5 ? yv[it] : make_float2(0.f, 0.f);   // kh <= 39
            const float2 cp = pr <= 39 ? cconj(sy[pcol][pr]) : make_float2(0.f, 0.f);
            v[it] = cadd(cadd(ck, cp), mul_si<+1>(cmul(cmul(clc, rowl[kh]), csub(ck, cp))));
        }
        {
            const int j = rw;
            dft8<+1>(v);
            // sc shares sz's storage, which no thread reads after the barrier above
#pragma unroll
            for (int r = 0; r < 8; ++r) sc[col][8 * j + r] = v[r];
            __syncthreads();
#pragma unroll
            for (int r = 0; r < 8; ++r) v[r] = sc[col][j + 8 * r];
#pragma unroll
            for (int r = 1; r < 8; ++r) v[r] = cmul(v[r], w64[8 * r + j]);
            dft8<+1>(v);
        }
        // write the tile's 32 slots of every row tx: Bc rows are planar (real
        // plane then imaginary plane, BC_ROW floats each): two 128-byte runs.
        // Buffer stores: one VGPR offset, row tx = 8 it + rw in the scalar offset.
        // They are non-temporal (cache policy nt): the output streams out and
        // should not evict the packed-spectrum lines the next tile reads again.
        {
            const __amdgpu_buffer_rsrc_t rb = __builtin_amdgcn_make_buffer_rsrc(
                (void *)(Bc + (size_t)x.bpl * KHI * 2 * BC_ROW), 0, KHI * 2 * BC_ROW * (int)sizeof(float),
                BUF_RSRC_WORD3);
            const int vo = 4 * (rw * 2 * BC_ROW + x.tile * 32 + col);
            if (x.valid && !x.dup) {
#pragma unroll
                for (int it = 0; it < 8; ++it) {
                    const float2 y = v[it];   // row 8 it + rw
                    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(y.x), rb, vo, 4 * (it * 8 * 2 * BC_ROW),
                                                          CPOL_NT);
                    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(y.y), rb, vo,
                                                          4 * (it * 8 * 2 * BC_ROW + BC_ROW), CPOL_NT);
                }
            }
        }
        acc_run += 2.0 * (double)accf;
        dc_run += dc;
        // the next tile writes sz, which is sc's storage: all waves are done
        // reading sc (the DFT's second pass)
        __syncthreads();
        double *o = part + ((size_t)x.bpl * C1_TILES + x.tile) * 2;
        if (!more || (lt + ilv) / C1_TILES != x.bpl) {   // workgroup-uniform
            // wave reduce then cross-wave
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                acc_run += __shfl_xor(acc_run, off);
                dc_run += __shfl_xor(dc_run, off);
            }
            if ((t & 63) == 0) {
                red[0][t >> 6] = acc_run;
                red[1][t >> 6] = dc_run;
            }
            __syncthreads();
            if (t == 0) {
                o[0] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
                o[1] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
            }
            acc_run = 0.0;
            dc_run = 0.0;
        } else if (t == 0) {
            o[0] = 0.0;
            o[1] = 0.0;
        }
    }
}

// Statistics and thresholds of a chunk.  Units: the pipeline's samples are
// 65536 * 2^21 = 2^37 times numpy's; Y here carries a factor 65536.
//   sum x^2 = W / (65536^2 * L),  mean = Re Y[0] / (65536 * L),
//   sigma = sqrt((sum x^2 - L mean^2) / (L - 1))
// PER_BLOCK and FIXED thresholds depend on the block-pol alone and are written
// here; the REFERENCE walk (k_thresh) runs only for chunks that need it.
__global__ void __launch_bounds__(256)
k_stats(const double *__restrict__ part, int bp0, double2 *__restrict__ ms, double *__restrict__ stats_out,
        const BpDesc *__restrict__ bpd, float2 *__restrict__ thr) {
    __shared__ double red[2][256];
    const int t = threadIdx.x, b = blockIdx.x;
    const double L = 2097152.0, FA = 65536.0;
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
        const BpDesc &d = bpd[b];
        if (d.policy != SDBE_THRESH_REFERENCE)
            thr[b] = d.policy == SDBE_THRESH_PER_BLOCK ? make_float2((float)sig, (float)mean)
                                                       : make_float2(d.fixed_th, d.fixed_off);
        if (stats_out) {
            stats_out[(size_t)(bp0 + b) * 2] = mean;
            stats_out[(size_t)(bp0 + b) * 2 + 1] = sig;
        }
    }
}

// Threshold policy over a chunk.  PER_BLOCK and FIXED are per block-pol and
// run in parallel.  REFERENCE is sequential in block-pol order: the first
// polarisation-0 block-pol seen while its slot's threshold is 0 latches its
// sigma and mean, which then serve both polarisations
// (vdif_inout_gpu_thread.cu:954-1005).  One thread walks the chunk for it,
// over the statistics, policy fields and latch slots staged in LDS, so that
// no step of the walk waits on global memory; the slots go back at the end.
constexpr int THRESH_MAX_CHUNK = SDBE_MAX_CHUNK;
__global__ void __launch_bounds__(256) k_thresh(const double2 *__restrict__ ms, int nbpl,
                                                const BpDesc *__restrict__ bpd, float *__restrict__ latch,
                                                float2 *__restrict__ thr) {
    __shared__ float2 sms[THRESH_MAX_CHUNK];      // (sigma, mean) as float
    __shared__ int sinfo[THRESH_MAX_CHUNK];       // pol | slot << 8, or -1: not REFERENCE
    __shared__ float2 slat[SDBE_LATCH_SLOTS];
    __shared__ int any_ref;
    const int t = threadIdx.x;
    if (t == 0) any_ref = 0;
    if (t < SDBE_LATCH_SLOTS) slat[t] = make_float2(latch[2 * t], latch[2 * t + 1]);
    __syncthreads();
    for (int b = t; b < nbpl; b += blockDim.x) {
        const BpDesc &d = bpd[b];
        const float2 v = make_float2((float)ms[b].y, (float)ms[b].x);
        sms[b] = v;
        if (d.policy == SDBE_THRESH_REFERENCE) {
            sinfo[b] = (int)d.pol | ((int)d.slot << 8);
            any_ref = 1;
        } else {
            sinfo[b] = -1;
            thr[b] = d.policy == SDBE_THRESH_PER_BLOCK ? v : make_float2(d.fixed_th, d.fixed_off);
        }
    }
    __syncthreads();
    if (t != 0 || !any_ref) return;
    for (int b = 0; b < nbpl; ++b) {
        const int info = sinfo[b];
        if (info < 0) continue;
        const int pol = info & 0xFF, slot = info >> 8;
        float2 l = slat[slot];
        if (l.x == 0.0f && pol == 0) {
            l = sms[b];
            slat[slot] = l;
            latch[2 * slot] = l.x;
            latch[2 * slot + 1] = l.y;
        }
        thr[b] = l;
    }
}

struct OutDesc {
    int32_t frame_bytes;     // 8192 (PAYLOAD) or 8224 (FRAMES)
    int32_t hdr_bytes;       // 0 or 32
    int32_t bp_headers;      // stamp the block-pols' frame headers (k_pack at 64/39, k_headers band-trimmed)
    int32_t reserved;
    // re-framed group output (output-thread time alignment): block blk's
    // stream byte q goes to group stream byte blk * 524288 + q + shift_bytes;
    // bytes below 0 are dropped, bytes at or beyond lim_bytes go to
    // lead[pol][. - lim_bytes], the head of the next group's first frame
    int64_t shift_bytes;
    int64_t lim_bytes;
    uint8_t *lead;           // [2][8192]
};

// Stage C2: row (tx, block-pol), row index KHI bpl + tx.
// z[tx + 64 ty] = sum_c Bc[tx][c] e^{2 pi i c tx/2^20} e^{2 pi i c ty/16384} is the packed output:
// samples (2t, 2t+1) = (Re, Im) z[t], scaled by 2^-37 to numpy units.  Codes
// per vdif_inout_gpu_thread.cu:577 (v = x - off; ((|v| >= th) | 2) ^ 3 (v < 0));
// the 4-bit pair code (re | im << 2) of t goes to the low nibble of
// stage[bp][tx / 2][ty] for even tx, the high nibble for odd tx.
// The workgroups are persistent (one per CU: the FFT's LDS fills a CU) and take
// the rows of pairs blockIdx.x, blockIdx.x + gridDim.x, ... in turn (pair i =
// rows 2i, 2i + 1, so that a byte's two nibbles are stored once, by the odd
// row's pass).  A row arrives in stage C1's
// tile-major slot order (planar: real plane then imaginary plane) and is
// scattered into LDS in column order (slot_col: column of each slot, -1 for
// none); the LDS word of each of a thread's slots is computed once.  The next
// row's loads are issued when the transform no longer needs LDS (after pass 3)
// and land while pass 4, the scaling and the stores of the current row run
// (unconditionally: the last row loads itself again, so the registers are
// always overwritten and are not live across the transform).
constexpr int C2_NSLOT = BC_ROW / 1024;          // 17 full slot passes ...
constexpr int C2_REM = BC_ROW - C2_NSLOT * 1024; // ... and 64 slots left
constexpr int C2_SLOT_TAB = (C2_NSLOT + 1) * 1024; // slot table entries (padded with -1)

// Loads of one row through a raw buffer resource (base and size in SGPRs, one
// VGPR offset 4 t, the slot pass in the scalar offset): no per-load address
// arithmetic on the VALU.  Every access is inside the row by construction: in
// the last, partial slot pass the lanes past the row read the pass's first
// slot (their values go to the junk row).

__device__ __forceinline__ void c2_load(const float *__restrict__ Bc, int row, float (&pr)[C2_NSLOT + 1],
                                        float (&pi)[C2_NSLOT + 1]) {
    const int t = threadIdx.x;
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(Bc + (size_t)row * 2 * BC_ROW), 0, 2 * BC_ROW * (int)sizeof(float), BUF_RSRC_WORD3);
#pragma unroll
    for (int i = 0; i < C2_NSLOT; ++i) {
        pr[i] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, 4 * t, 4096 * i, 0));
        pi[i] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, 4 * t, 4 * BC_ROW + 4096 * i, 0));
    }
    const int tl = t < C2_REM ? 4 * t : 0;
    pr[C2_NSLOT] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, tl, 4096 * C2_NSLOT, 0));
    pi[C2_NSLOT] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, tl, 4 * BC_ROW + 4096 * C2_NSLOT, 0));
}

template <bool F32, int P>
__global__ void __launch_bounds__(1024)
k_stage_c2(const float *__restrict__ Bc, const float2 *__restrict__ thr, const BpDesc *__restrict__ bpd,
           uint8_t *__restrict__ stage, const int *__restrict__ slot_col, Tables tb, int n_rows) {
    // the work array, then a junk row (words LDS_FFT_N + lane, + 64) that takes
    // the slots without a column, so the scatter has no branches
    __shared__ float lds[LDS_FFT_N + 128];
    __shared__ FftTw ftw;
    fft_tw_load(&ftw, tb.fft);
    const int t = threadIdx.x;
    const float scale = 7.275957614183426e-12f;   // 2^-37, exact
    // LDS word of each slot of this thread, two 16-bit words per register
    // (slot_col has C2_SLOT_TAB entries, -1 past the row: all loads are issued
    // before any is used)
    int cs[C2_NSLOT + 1];
#pragma unroll
    for (int i = 0; i <= C2_NSLOT; ++i) cs[i] = slot_col[t + 1024 * i];
    uint32_t pos2[(C2_NSLOT + 2) / 2];
#pragma unroll
    for (int i = 0; i <= C2_NSLOT; ++i) {
        const uint32_t p = (uint32_t)(cs[i] >= 0 ? lds_re(cs[i]) : LDS_FFT_N + (t & 63));
        if (i & 1)
            pos2[i >> 1] |= p << 16;
        else
            pos2[i >> 1] = p;
    }
    float pr[C2_NSLOT + 1], pi[C2_NSLOT + 1];
    // rows go in pairs (2 i, 2 i + 1) of one block-pol, pair blockIdx.x, + gridDim.x, ...:
    // the even row's codes wait in hold (byte q & 3 of hold[q >> 2] for element
    // te + 1024 q) and the odd row stores both as one byte
    int row = 2 * (int)blockIdx.x;
    uint32_t hold[4] = {0u, 0u, 0u, 0u};
    if constexpr (P == 0) c2_load(Bc, row, pr, pi);
    for (;;) {
        const int tx = row & (KHI - 1), bpl = row / KHI;
        // the matrix-core variants load their own row here instead of prefetching
        if constexpr (P != 0) c2_load(Bc, row, pr, pi);
        __syncthreads();   // the previous row's LDS reads are done
#pragma unroll
        for (int i = 0; i <= C2_NSLOT; ++i)
            lds_put(lds, (i & 1) ? (int)(pos2[i >> 1] >> 16) : (int)(pos2[i >> 1] & 0xffffu),
                    make_float2(pr[i], pi[i]));
        __syncthreads();
        const int next = (row & 1) ? row - 1 + 2 * (int)gridDim.x : row + 1;
        float2 v[16];
        if constexpr (P == 0) {
            const int tr = row_tid();
#pragma unroll
            for (int q = 0; q < 16; ++q) v[q] = lds_get(lds, lds_re(tr) + 2048 * q);
            // inter-pass twiddle e^{2 pi i c tx / 2^20}, c = t + 1024 q: per-thread factor
            // times a workgroup-uniform one, e^{2 pi i q tx / 1024} = the high table's entry q tx
            const float2 A = tw<+1>(tb.np, (uint32_t)(tr * tx));
#pragma unroll
            for (int q = 0; q < 16; ++q) v[q] = cmul(v[q], cmul(A, tw_hi<+1>(tb.np, (uint32_t)(q * tx))));
            // third exchange wave-local: the row comes out in the order of
            // fft16k_tord (element fft16k_tord(t) + 1024 q in v[q])
            fft16k_p12<+1>(v, lds, &ftw, tr);
            fft16k_p3w<+1>(v, lds, &ftw, tr);
            c2_load(Bc, next < n_rows ? next : row, pr, pi);
            // keep the loads here: the scheduler would otherwise sink them below
            // the stores at the end of the row, exposing their whole latency
            __builtin_amdgcn_sched_barrier(0);
            fft16k_p4<+1>(v, &ftw, fft16k_tord(tr));
        } else {
            // MFMA layout: element c = d + 1024 r, twiddle e^{2 pi i c tx/2^20} = A(d) B(r)
            const int t = row_tid(), l = t & 63, kb = l >> 4;
            float2 Br[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) Br[j] = tw_hi<+1>(tb.np, (uint32_t)((4 * kb + j) * tx));
            const int rb = mfma_rd_base(t);
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int e0 = mfma_elem(g, 0, t);
                const float2 Ad = tw<+1>(tb.np, (uint32_t)((e0 & 1023) * tx));
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    v[4 * g + j] = cmul(lds_get(lds, mfma_rd(rb, g, j)), cmul(Ad, Br[j]));
                }
            }
            MfmaW<+1, P> W;
            W.init(t);
            fft16k_mfma<+1, P>(v, lds, &ftw, W, t);
        }
        // element of v[0]: fft16k_tord(t) on the VALU path, t on the matrix-core path
        const int te = P == 0 ? fft16k_tord(t) : t;
        if (F32) {
            GAS uint64_t *dst = gptr(reinterpret_cast<uint64_t *>(bpd[bpl].out)) + tx + KHI * te;
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const float2 y = cscale(v[q], scale);
                dst[(size_t)KHI * 1024 * q] = (uint64_t)__float_as_uint(y.x) | ((uint64_t)__float_as_uint(y.y) << 32);
            }
        } else {
            const float2 to = thr[bpl];
            uint32_t code[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const float a = v[q].x * scale - to.y;
                const float b = v[q].y * scale - to.y;
                // code = (|v| >= th ? 3 : 2) ^ (v < 0 ? 3 : 0): 3 if v >= th, 2 if 0 <= v < th,
                // 1 if -th < v < 0, 0 if v <= -th; b's code in bits 2-3 (disjoint: XOR = OR)
                code[q] = ((fabsf(a) >= to.x) ? 3u : 2u) ^ (a < 0.f ? 3u : 0u) ^
                          ((fabsf(b) >= to.x) ? 12u : 8u) ^ (b < 0.f ? 12u : 0u);
            }
            if ((tx & 1) == 0) {   // workgroup-uniform
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    hold[j] = code[4 * j] | (code[4 * j + 1] << 8) | (code[4 * j + 2] << 16) | (code[4 * j + 3] << 24);
            } else {
                // stage row tx / 2 of the block-pol, 16384 bytes, as a buffer: element
                // te + 1024 q at offset te (VGPR) + 1024 q (scalar), with no 64-bit
                // address arithmetic per store; its low nibble from row tx - 1
                const __amdgpu_buffer_rsrc_t rd = __builtin_amdgcn_make_buffer_rsrc(
                    (void *)(stage + ((size_t)bpl * (KHI / 2) + (tx >> 1)) * NCH), 0, NCH, BUF_RSRC_WORD3);
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    const uint32_t m = ((hold[q >> 2] >> (8 * (q & 3))) & 0xFFu) | (code[q] << 4);
                    __builtin_amdgcn_raw_buffer_store_b8((uint8_t)m, rd, te, 1024 * q, 0);
                }
            }
        }
        if (next >= n_rows) break;
        row = next;
    }
}

// Pack: stage C2's rows r hold the 2-bit codes of samples 2p, 2p + 1
// (p = r + 64 ty) for ty < 16384; it stores rows 2i and 2i + 1 together as the
// 32 stage rows M_i of 16384 bytes per block-pol, M_i[ty] = S[2i][ty] |
// S[2i+1][ty] << 4.  Stream byte q of a block holds samples 4q..4q+3 (sample j
// at bits 2j, the uint32 packing of vdif_inout_gpu_thread.cu:578-580):
// byte (i + 32 ty) = M_i[ty].
// Thread = (block-pol, columns ty = 4c .. 4c+3): 32 row words in (a wave reads
// 256 contiguous bytes of a row per load), and byte j of word M_i is stream
// byte i + 32 (4c + j): the thread's 128 contiguous stream bytes are a 4 x 4
// byte transpose of M per output word.  Block bytes go to 64 frames of 8192 bytes;
// re-framed output (shift, lead) takes them word by word.
// VDIF output headers (vdif_out_databuf.h:19-52) stamped as the output thread
// does (vdif_out_net_thread.c:294-317, defaults vdif_out_databuf.c:195-227):
// frame n (of fpb per block-pol) of a block-pol whose first frame is (secs, df, psn) has
// df = (df + n) mod 125000, secs + carry, psn + n; threadID 0, bps = 1,
// df_len = 8224 / 8, edv5 = 2 (words 3 and 4 come prebuilt per block-pol).
// k_headers stamps them in a launch of its own; at 64/39 k_pack stamps its
// block-pols' headers while it packs their payloads.
__device__ __forceinline__ void stamp_header(const BpDesc &d, int n, int frame_bytes) {
    const uint64_t dft = (uint64_t)d.df + (uint64_t)n;
    const uint32_t df = (uint32_t)(dft % 125000u);
    const uint32_t secs = d.secs + (uint32_t)(dft / 125000u);
    const uint64_t psn = d.psn + (uint64_t)n;
    GAS u32x4 *dst = reinterpret_cast<GAS u32x4 *>(gptr(d.out) + (int64_t)n * frame_bytes);
    dst[0] = u32x4{secs & 0x3FFFFFFFu, (df & 0xFFFFFFu) | d.w1hi, (uint32_t)((OUT_FRAME_PAYLOAD + 32) / 8), d.w3};
    dst[1] = u32x4{d.w4, 0u, (uint32_t)(psn & 0xFFFFFFFFu), (uint32_t)(psn >> 32)};
}

__device__ __forceinline__ uint32_t pack_word(const uint32_t (&M)[32], int w) {
    const int j = w >> 3, i0 = (w & 7) * 4;
    return ((M[i0] >> (8 * j)) & 0xFFu) | (((M[i0 + 1] >> (8 * j)) & 0xFFu) << 8) |
           (((M[i0 + 2] >> (8 * j)) & 0xFFu) << 16) | (((M[i0 + 3] >> (8 * j)) & 0xFFu) << 24);
}

__global__ void __launch_bounds__(256)
k_pack(const uint8_t *__restrict__ stage, const BpDesc *__restrict__ bpd, int nb, OutDesc od) {
    const int bpl = blockIdx.x >> 4, c = ((blockIdx.x & 15) << 8) + threadIdx.x;   // 4096 column quads
    // the block-pol's 64 frame headers (FRAMES mode), by its first workgroup
    if (od.hdr_bytes && od.bp_headers && (blockIdx.x & 15) == 0 && threadIdx.x < OUT_FRAMES_BLOCK)
        stamp_header(bpd[bpl], (int)threadIdx.x, od.frame_bytes);
    const GAS uint32_t *src = gptr(reinterpret_cast<const uint32_t *>(stage + (size_t)bpl * (KHI / 2) * NCH)) + c;
    uint32_t M[32];
#pragma unroll
    for (int i = 0; i < 32; ++i) M[i] = src[(size_t)i * (NCH / 4)];
    GAS uint8_t *ob = gptr(bpd[bpl].out) + od.hdr_bytes;
    const int64_t qb0 = 128 * (int64_t)c;                 // stream byte of the run, relative to the block
    if (od.shift_bytes == 0 && od.lim_bytes == INT64_MAX && (((uintptr_t)ob) & 15) == 0) {
        // the workgroup's 32 KB of stream bytes through LDS, so that a wave
        // stores 1 KB contiguous per instruction (row stride 33 words)
        __shared__ uint32_t runs[256 * 33];
        const int t = threadIdx.x;
#pragma unroll
        for (int w = 0; w < 32; ++w) runs[t * 33 + w] = pack_word(M, w);
        __syncthreads();
        const int64_t wg0 = 128 * (int64_t)((blockIdx.x & 15) << 8);     // first stream byte of the workgroup
#pragma unroll
        for (int it = 0; it < 8; ++it) {
            const int e = it * 1024 + 4 * t;                                // word of the workgroup's run
            const int src = (e >> 5) * 33 + (e & 31);
            u32x4 q;
            q.x = runs[src];
            q.y = runs[src + 1];
            q.z = runs[src + 2];
            q.w = runs[src + 3];
            const int64_t qb = wg0 + 4 * e;
            *reinterpret_cast<GAS u32x4 *>(ob + (qb >> 13) * od.frame_bytes + (qb & (OUT_FRAME_PAYLOAD - 1))) = q;
        }
        return;
    }
    const int64_t blk_pos = (int64_t)bpd[bpl].blk * (OUT_FRAMES_BLOCK * OUT_FRAME_PAYLOAD);
    const int pol = bpd[bpl].pol;
#pragma unroll
    for (int w = 0; w < 32; ++w) {
        const uint32_t wv = pack_word(M, w);
        const int64_t qb = qb0 + 4 * w + od.shift_bytes;           // stream byte, relative to the block
        const int64_t gpos = blk_pos + qb;                          // ... and to the group
        if (gpos < 0) continue;
        if (gpos >= od.lim_bytes) {
            *reinterpret_cast<GAS uint32_t *>(gptr(od.lead) + pol * OUT_FRAME_PAYLOAD + (gpos - od.lim_bytes)) = wv;
            continue;
        }
        const int64_t f = qb >> 13, off = qb & (OUT_FRAME_PAYLOAD - 1);   // f may be negative: frames of earlier blocks
        *reinterpret_cast<GAS uint32_t *>(ob + f * od.frame_bytes + off) = wv;
    }
}

__global__ void __launch_bounds__(256) k_headers(const BpDesc *__restrict__ bpd, int n_bp, OutDesc od, int fpb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_bp * fpb) return;
    const int b = i / fpb, n = i - b * fpb;
    stamp_header(bpd[b], n, od.frame_bytes);
}

// Headers of n frames per polarisation of a re-framed group: frame n of pol p
// at base + p * pol_stride + n * frame_bytes, stamped from (secs, df, psn).
struct HdrLin {
    uint8_t *base;
    int64_t pol_stride;
    int32_t frame_bytes, n_frames;
    uint32_t secs, df, w1hi, w3, w4;
    uint64_t psn;
};
__global__ void __launch_bounds__(256) k_headers_lin(HdrLin hd) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 2 * hd.n_frames) return;
    const int pol = i / hd.n_frames, n = i - pol * hd.n_frames;
    const uint64_t dft = (uint64_t)hd.df + (uint64_t)n;
    const uint32_t df = (uint32_t)(dft % 125000u);
    const uint32_t secs = hd.secs + (uint32_t)(dft / 125000u);
    const uint64_t psn = hd.psn + (uint64_t)n;
    GAS u32x4 *dst = reinterpret_cast<GAS u32x4 *>(gptr(hd.base) + (int64_t)pol * hd.pol_stride +
                                                    (int64_t)n * hd.frame_bytes);
    dst[0] = u32x4{secs & 0x3FFFFFFFu, (df & 0xFFFFFFu) | hd.w1hi, (uint32_t)((OUT_FRAME_PAYLOAD + 32) / 8), hd.w3};
    dst[1] = u32x4{hd.w4 | (uint32_t)pol, 0u, (uint32_t)(psn & 0xFFFFFFFFu), (uint32_t)(psn >> 32)};
}

// The previous group's tail (n_bytes per polarisation) becomes the head of
// this group's first frame.
__global__ void __launch_bounds__(256) k_lead_copy(const uint8_t *__restrict__ lead, uint8_t *base, int64_t pol_stride,
                                                   int hdr_bytes, int n_bytes) {
    const int i = (blockIdx.x * blockDim.x + threadIdx.x) * 4;
    const int pol = blockIdx.y;
    if (i >= n_bytes) return;
    *reinterpret_cast<GAS uint32_t *>(gptr(base) + pol * pol_stride + hdr_bytes + i) =
        *reinterpret_cast<const GAS uint32_t *>(gptr(lead) + pol * OUT_FRAME_PAYLOAD + i);
}

// Descriptors of n_bp back-to-back block-pols (the table sdbe_process_spectra
// builds on the host), written on the device: block-pol b reads its rows at
// rows + b * row_stride bytes, writes its frames at out + b * out_stride, is
// polarisation 0 of no flip, and its first frame lies fpb b frames after
// (secs, df, psn), df wrapping at 125000 frames per second.
struct SpecDescArgs {
    const int8_t *rows;
    uint8_t *out;
    int64_t row_stride, out_stride;
    uint64_t psn;
    uint32_t secs, df, w1hi, w3, w4;
    int32_t policy, fpb;
    float fixed_th, fixed_off;
};
__global__ void __launch_bounds__(64) k_spectra_desc(SpecDescArgs a, int n_bp, BpDesc *__restrict__ bpd) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n_bp) return;
    const uint64_t n = (uint64_t)b * (uint64_t)a.fpb;
    const uint64_t dft = (uint64_t)a.df + n;
    BpDesc d;
    d.rows = a.rows + (int64_t)b * a.row_stride;
    d.out = a.out + (int64_t)b * a.out_stride;
    d.psn = a.psn + n;
    d.secs = a.secs + (uint32_t)(dft / 125000u);
    d.df = (uint32_t)(dft % 125000u);
    d.w1hi = a.w1hi;
    d.w3 = a.w3;
    d.w4 = a.w4;
    d.flip = 0;
    d.pol = 0;
    d.policy = (int16_t)a.policy;
    d.slot = 0;
    d.fixed_th = a.fixed_th;
    d.fixed_off = a.fixed_off;
    d.blk = 0;
    bpd[b] = d;
}

#include "sdbe_trim.hpp"

}  // namespace

// ======================================================================= host
enum KernelId {
    KID_FILL, KID_MAP, KID_REORDER, KID_A, KID_B, KID_C1, KID_STATS, KID_THRESH, KID_C2,
    KID_PACK, KID_HEADERS, KID_TRIM_C2R, KID_TRIM_SEG, KID_TRIM_PACK, KID_A_PK, KID_BITMAP,
    KID_REORDER8, KID_A_PK8, KID_BITMAP8, KID_TRIM_C2R_PK, KID_CLEAR, KID_N
};
static const char *const kKernelNames[KID_N] = {
    "k_fill_i32", "k_packet_map", "k_reorder", "k_stage_a", "k_stage_b",
    "k_stage_c1", "k_stats", "k_thresh", "k_stage_c2", "k_pack", "k_headers", "k_trim_c2r", "k_trim_seg",
    "k_trim_pack", "k_stage_a_pk", "k_group_bitmap", "k_reorder_8x1024", "k_stage_a_pk8",
    "k_group_bitmap_8x1024", "k_trim_c2r_pk", "k_clear_map"};

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
    // per-block-pol descriptors: a ring of host (pinned) / device tables, each
    // reused only after the call that last used it has finished with it
    static constexpr int RING = 4;
    int cap_bp = 0;
    BpDesc *h_bpd[RING] = {};
    BpDesc *d_bpd[RING] = {};
    hipEvent_t bpd_done[RING] = {};
    int bpd_used[RING] = {};
    int ring_i = 0;
    // graph replay of sdbe_process_blocks (own descriptor table)
    int capturing = 0;
    hipStream_t cap_stream = nullptr;
    hipGraphExec_t gexec = nullptr;
    int g_nbp = 0;
    BpDesc *g_h = nullptr, *g_d = nullptr;
    hipEvent_t g_done = nullptr;
    int g_used = 0;
    int c1_tpw = 8;
    int c1_ilv = 16;
    int n_cu = 256;                  // compute units of the device
    int prof_kid = -1;               // profile only this kernel kind (-1: all)
    // resampler: 64/39 (trim.D == 0) or the band-trimmed path of a rate preset
    trim::Geometry trim;
    trim::Tabs trim_tabs{};
    float2 *trim_tab = nullptr;
    int snaps = NBLK;                // snapshots per block-pol
    int fpb = OUT_FRAMES_BLOCK;      // output frames per block-pol
    int nfr = NFR;                   // B-frames per group (DECIMATION_FACTOR of the rate)
    int map_n = MAP_N;               // packet slots per group: nfr * packets per frame
    int geom = SDBE_GEOM_REFERENCE;  // packet geometry
    int n_cid = 256, n_fid = 8, pkt_words = PKT_WORDS;
    // re-framed group output (output-thread time alignment)
    int align_pending = 0;           // the next group starts an aligned stream
    int align_active = 0;
    int64_t align_words = 0;         // skip_int32 + 2048 skip_frames of the next aligned start
    int lead_words = 0;              // words carried from one group into the next
    uint8_t *lead = nullptr;         // [2][8192] device
    int n_frames_last = 0;
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
    if (!h->prof || h->capturing || (h->prof_kid >= 0 && kid != h->prof_kid)) return nullptr;
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

// The resampler's twiddle tables (FFT image, the two-level tables of
// N = 32768, 638,976, 1,277,952, 2,097,152 and 1,048,576, stage C1's column
// twiddles), computed in FP64 on the host and stored FP32, back to back.
constexpr TwSpec kTwSpecs[6] = {
    {16384, 14, 0},                 // unused slot (the FFT image is stored separately)
    {32768, 7, 16384},              // n32k
    {MP, 10, (int64_t)MP},          // nmp: n * k1 < 16384 * 39
    {2 * (int64_t)MP, 10, (int64_t)MP + 1},   // nm: m <= M'
    {2 * (int64_t)P20, 10, (int64_t)P20},     // nl: k < 2^20
    {P20, 10, (int64_t)P20},        // np: c * tx < 2^20, r * j * 16384 < 2^20
};
constexpr size_t tw_entries_c(const TwSpec &s) {
    return ((size_t)1 << s.lo_bits) + (size_t)((s.max_idx >> s.lo_bits) + 2);
}
constexpr size_t kTabC1Col = FFT_TW_N + tw_entries_c(kTwSpecs[1]) + tw_entries_c(kTwSpecs[2]) +
                             tw_entries_c(kTwSpecs[3]) + tw_entries_c(kTwSpecs[4]) + tw_entries_c(kTwSpecs[5]);
constexpr size_t kTabEntries = kTabC1Col + 2 * NCH;

void tables_fill_host(float2 *host) {
    size_t o = FFT_TW_N;
    fft_tw_fill(host);
    for (int i = 1; i < 6; ++i) {
        tw_fill(host + o, kTwSpecs[i]);
        o += tw_entries(kTwSpecs[i]);
    }
    // stage C1 column twiddles e^{-2 pi i c/M} and e^{2 pi i c/L}, c < 16384
    const double two_pi = 6.283185307179586476925286766559;
    for (int c = 0; c < NCH; ++c) {
        const double am = -two_pi * (double)c / (2.0 * (double)MP);
        const double al = two_pi * (double)c / (2.0 * (double)P20);
        host[kTabC1Col + c] = make_float2((float)cos(am), (float)sin(am));
        host[kTabC1Col + NCH + c] = make_float2((float)cos(al), (float)sin(al));
    }
}

// Table pointers into a device copy of the tables_fill_host image.
Tables tables_bind(const float2 *dev) {
    Tables tb{};
    tb.fft = dev;
    tb.c1col = dev + kTabC1Col;
    Tw2 *dst[6] = {nullptr, &tb.n32k, &tb.nmp, &tb.nm, &tb.nl, &tb.np};
    size_t o = FFT_TW_N;
    for (int i = 1; i < 6; ++i) {
        dst[i]->lo = dev + o;
        dst[i]->hi = dev + o + ((size_t)1 << kTwSpecs[i].lo_bits);
        dst[i]->lo_bits = kTwSpecs[i].lo_bits;
        o += tw_entries(kTwSpecs[i]);
    }
    return tb;
}

int build_tables(sdbe_handle_t *h) {
    std::vector<float2> host(kTabEntries);
    tables_fill_host(host.data());
    HIP_TRY(h, hipMalloc(&h->tables, kTabEntries * sizeof(float2)));
    HIP_TRY(h, hipMemcpy(h->tables, host.data(), kTabEntries * sizeof(float2), hipMemcpyHostToDevice));
    h->tb = tables_bind(h->tables);
    return SDBE_OK;
}

// Twiddle tables of the band-trimmed path (trim::Tabs), FP64 on the host.
int build_trim_tables(sdbe_handle_t *h) {
    const int NR = h->trim.R / 2, NP = h->trim.P / 2;
    const size_t n = (size_t)NR + (NR + 1) + NP + NP;
    std::vector<float2> host(n);
    auto e = [](double ang) { return make_float2((float)cos(ang), (float)sin(ang)); };
    const double tau = 6.283185307179586476925286766559;
    size_t o = 0;
    const size_t zr = o;
    for (int m = 0; m < NR; ++m) host[o++] = e(-tau * m / NR);
    const size_t xr = o;
    for (int k = 0; k <= NR; ++k) host[o++] = e(-tau * k / h->trim.R);
    const size_t wp = o;
    for (int j = 0; j < NP; ++j) host[o++] = e(tau * j / h->trim.P);
    const size_t zp = o;
    for (int m = 0; m < NP; ++m) host[o++] = e(tau * m / NP);
    HIP_TRY(h, hipMalloc(&h->trim_tab, n * sizeof(float2)));
    HIP_TRY(h, hipMemcpy(h->trim_tab, host.data(), n * sizeof(float2), hipMemcpyHostToDevice));
    h->trim_tabs.zr = h->trim_tab + zr;
    h->trim_tabs.xr = h->trim_tab + xr;
    h->trim_tabs.wp = h->trim_tab + wp;
    h->trim_tabs.zp = h->trim_tab + zp;
    return SDBE_OK;
}

// Slot -> column table of the stage C1 output (same rule as c1_column); every
// column 0..16383 must own exactly one slot (false otherwise).  seen: NCH ints
// of scratch.
bool slot_table_fill_host(int *tab, int *seen) {
    for (int i = 0; i < C2_SLOT_TAB; ++i) tab[i] = -1;
    for (int c = 0; c < NCH; ++c) seen[c] = 0;
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
        if (seen[c] != 1) return false;
    return true;
}

int build_slot_table(sdbe_handle_t *h) {
    std::vector<int> tab(C2_SLOT_TAB), seen(NCH);
    if (!slot_table_fill_host(tab.data(), seen.data())) return SDBE_ERR_STATE;
    HIP_TRY(h, hipMemcpy(h->slot_col, tab.data(), C2_SLOT_TAB * sizeof(int), hipMemcpyHostToDevice));
    return SDBE_OK;
}

// Run stages A..C2 over block-pols [0, n_bp) in chunks.
template <typename K>
void launch_c2_kernel(sdbe_handle_t *h, int nb, hipStream_t s, const BpDesc *bpd, K kern) {
    const int rows = KHI * nb, pairs = rows / 2;
    const int grid = pairs < h->n_cu ? pairs : h->n_cu;
    LAUNCH(h, KID_C2, nb, s, kern, dim3(grid), dim3(1024), (const float *)h->Bc, (const float2 *)h->thr, bpd,
           h->stage, (const int *)h->slot_col, h->tb, rows);
}

template <bool F32>
void launch_c2(sdbe_handle_t *h, int nb, hipStream_t s, const BpDesc *bpd) {
    switch (h->cfg.fft_precision) {
    case FFT_BF16X3: launch_c2_kernel(h, nb, s, bpd, k_stage_c2<F32, FFT_BF16X3>); break;
    case FFT_BF16: launch_c2_kernel(h, nb, s, bpd, k_stage_c2<F32, FFT_BF16>); break;
    case FFT_FP16: launch_c2_kernel(h, nb, s, bpd, k_stage_c2<F32, FFT_FP16>); break;
    default: launch_c2_kernel(h, nb, s, bpd, k_stage_c2<F32, 0>);
    }
}

// The band-trimmed resampler over n_bp blocks of D snapshots, in chunks: the
// per-snapshot C2R into V (as the concatenated series), the segment R2C / C2R
// into Zf (float samples) or the caller's buffers (f32), statistics, the
// threshold policy and the natural-order packing.
template <int D, int E>
int run_trim(sdbe_handle_t *h, const BpDesc *d_bpd, int n_bp, const OutDesc &od, bool f32, double *d_stats,
             hipStream_t s, const uint32_t *pk) {
    constexpr int P = 256 * E;
    const auto kc2r = k_trim_c2r<D>;
    const auto kc2r_pk = h->geom == SDBE_GEOM_8X1024 ? k_trim_c2r_pk<D, true> : k_trim_c2r_pk<D, false>;
    const auto kseg_f = k_trim_seg<D, E, true>;
    const auto kseg_c = k_trim_seg<D, E, false>;
    for (int bp0 = 0; bp0 < n_bp; bp0 += h->chunk) {
        const int nb = (n_bp - bp0) < h->chunk ? (n_bp - bp0) : h->chunk;
        const BpDesc *bpd = d_bpd + bp0;
        if (pk)
            LAUNCH(h, KID_TRIM_C2R_PK, nb, s, kc2r_pk, dim3(D * nb < h->n_cu ? D * nb : h->n_cu), dim3(1024), bpd,
                   h->V, h->tb, pk, (const int *)h->map, D * nb);
        else
            LAUNCH(h, KID_TRIM_C2R, nb, s, kc2r, dim3(D * nb), dim3(1024), bpd, h->V, h->tb);
        if (f32)
            LAUNCH(h, KID_TRIM_SEG, nb, s, kseg_f, dim3(trim::NSEG * nb), dim3(trim::seg_threads(D)),
                   (const float *)h->V, bpd, (float *)h->Zf, h->part, h->trim_tabs, h->trim.skip,
                   (const float2 *)h->tb.fft);
        else
            LAUNCH(h, KID_TRIM_SEG, nb, s, kseg_c, dim3(trim::NSEG * nb), dim3(trim::seg_threads(D)),
                   (const float *)h->V, bpd, (float *)h->Zf, h->part, h->trim_tabs, h->trim.skip,
                   (const float2 *)h->tb.fft);
        LAUNCH(h, KID_STATS, nb, s, k_trim_stats, dim3(nb), dim3(trim::NSEG), (const double *)h->part, P, bp0, h->ms,
               d_stats);
        if (!f32) {
            LAUNCH(h, KID_THRESH, nb, s, k_thresh, dim3(1), dim3(256), (const double2 *)h->ms, nb, bpd, h->latch,
                   h->thr);
            LAUNCH(h, KID_TRIM_PACK, nb, s, k_trim_pack, dim3(nb * (trim::NSEG * P / 16 / 256)), dim3(256),
                   (const float *)h->Zf, (const float2 *)h->thr, bpd, P, od);
        }
        HIP_TRY(h, hipGetLastError());
    }
    if (!f32 && od.hdr_bytes && od.bp_headers) {
        LAUNCH(h, KID_HEADERS, n_bp, s, k_headers, dim3((n_bp * E + 255) / 256), dim3(256), d_bpd, n_bp, od, E);
        HIP_TRY(h, hipGetLastError());
    }
    return SDBE_OK;
}

// Run stages A..C2 (+ pack) over the n_bp block-pols described by the device
// table d_bpd, in chunks.  f32: write float samples instead of codes.
// pk: the group's packets when the first kernel (stage A, or the band-trimmed
// per-snapshot C2R) reads them directly (the block-pols' blk / pol then locate
// their snapshots through the packet map), else null.
int run_resampler(sdbe_handle_t *h, const BpDesc *d_bpd, int n_bp, const OutDesc &od, bool f32,
                  double *d_stats, hipStream_t s, const uint32_t *pk = nullptr, bool ref_walk = true) {
    if (h->trim.frac == 6) return run_trim<39, 32>(h, d_bpd, n_bp, od, f32, d_stats, s, pk);
    if (h->trim.frac == 8) return run_trim<13, 8>(h, d_bpd, n_bp, od, f32, d_stats, s, pk);
    if (h->trim.frac == 10) return run_trim<65, 64>(h, d_bpd, n_bp, od, f32, d_stats, s, pk);
    if (h->trim.frac == 11) return run_trim<143, 128>(h, d_bpd, n_bp, od, f32, d_stats, s, pk);
    for (int bp0 = 0; bp0 < n_bp; bp0 += h->chunk) {
        const int nb = (n_bp - bp0) < h->chunk ? (n_bp - bp0) : h->chunk;
        const BpDesc *bpd = d_bpd + bp0;
        if (pk && h->geom == SDBE_GEOM_8X1024)
            LAUNCH(h, KID_A_PK8, nb, s, k_stage_a_pk8, dim3((NCH / 256) * nb), dim3(256), bpd, nb, h->V, h->Zf, h->tb,
                   pk, (const int *)h->map);
        else if (pk)
            LAUNCH(h, KID_A_PK, nb, s, k_stage_a_pk, dim3((NCH / 256) * nb), dim3(256), bpd, nb, h->V, h->Zf, h->tb, pk,
                   (const int *)h->map);
        else
            LAUNCH(h, KID_A, nb, s, k_stage_a, dim3((NCH / 256) * nb), dim3(256), bpd, nb, h->V, h->Zf, h->tb);
        {
            const int rows = B_ROWS * nb;
            const int grid = rows < h->n_cu ? rows : h->n_cu;
            const int xm = (nb % 8) == 0 && (grid % 8) == 0 ? 1 : 0;
            switch (h->cfg.fft_precision) {
            case FFT_BF16X3:
                LAUNCH(h, KID_B, nb, s, k_stage_b<FFT_BF16X3>, dim3(grid), dim3(1024), (const float2 *)h->V, h->Zf,
                       h->tb, xm, rows);
                break;
            case FFT_BF16:
                LAUNCH(h, KID_B, nb, s, k_stage_b<FFT_BF16>, dim3(grid), dim3(1024), (const float2 *)h->V, h->Zf,
                       h->tb, xm, rows);
                break;
            case FFT_FP16:
                LAUNCH(h, KID_B, nb, s, k_stage_b<FFT_FP16>, dim3(grid), dim3(1024), (const float2 *)h->V, h->Zf,
                       h->tb, xm, rows);
                break;
            default:
                LAUNCH(h, KID_B, nb, s, k_stage_b<0>, dim3(grid), dim3(1024), (const float2 *)h->V, h->Zf, h->tb,
                       xm, rows);
            }
        }
        // whole groups of c1_ilv workgroups, c1_tpw tiles each
        const int per_grp = h->c1_tpw * h->c1_ilv;
        const int c1_wgs = (C1_TILES * nb + per_grp - 1) / per_grp * h->c1_ilv;
        LAUNCH(h, KID_C1, nb, s, k_stage_c1, dim3(c1_wgs), dim3(256), (const float2 *)h->Zf, (float *)h->Bc, h->part, h->tb,
               nb, h->c1_tpw, h->c1_ilv);
        LAUNCH(h, KID_STATS, nb, s, k_stats, dim3(nb), dim3(256), (const double *)h->part, bp0, h->ms, d_stats, bpd,
               h->thr);
        if (ref_walk)   // some block-pol of the call has the REFERENCE policy
            LAUNCH(h, KID_THRESH, nb, s, k_thresh, dim3(1), dim3(256), (const double2 *)h->ms, nb, bpd, h->latch,
                   h->thr);
        if (f32) {
            launch_c2<true>(h, nb, s, bpd);
        } else {
            launch_c2<false>(h, nb, s, bpd);
            // with the block-pols' headers (FRAMES mode)
            LAUNCH(h, KID_PACK, nb, s, k_pack, dim3(16 * nb), dim3(256), (const uint8_t *)h->stage, bpd, nb, od);
        }
        HIP_TRY(h, hipGetLastError());
    }
    return SDBE_OK;
}

OutDesc out_desc(const sdbe_handle_t *h) {
    OutDesc od;
    memset(&od, 0, sizeof(od));
    od.hdr_bytes = h->cfg.out_mode == SDBE_OUT_FRAMES ? 32 : 0;
    od.frame_bytes = OUT_FRAME_PAYLOAD + od.hdr_bytes;
    od.bp_headers = 1;
    od.shift_bytes = 0;
    od.lim_bytes = INT64_MAX;
    od.lead = nullptr;
    return od;
}

// Fill one descriptor.  t0 is the VDIF time of the block-pol's first frame.
void set_bp(BpDesc &d, const int8_t *rows, uint8_t *out, const sdbe_stream_state_t &t0, int pol, int flip,
            int policy, int slot, float fth, float foff, uint16_t station, uint8_t bdc, uint8_t rx) {
    memset(&d, 0, sizeof(d));
    d.rows = rows;
    d.out = out;
    d.secs = t0.secs;
    d.df = t0.df;
    d.psn = t0.psn;
    d.w1hi = (t0.ref_epoch & 0x3Fu) << 24;
    d.w3 = (uint32_t)station | (1u << 26);
    d.w4 = ((uint32_t)pol & 1u) | ((uint32_t)(bdc & 1u) << 1) | ((uint32_t)(rx & 1u) << 2) | (2u << 24);
    d.flip = (int16_t)(flip ? 1 : 0);
    d.pol = (int16_t)pol;
    d.policy = (int16_t)policy;
    d.slot = (int16_t)slot;
    d.fixed_th = fth;
    d.fixed_off = foff;
}

// Next descriptor table of the ring (waits until the call that last used it
// has finished on the device).
int ring_acquire(sdbe_handle_t *h, int &slot) {
    slot = h->ring_i;
    h->ring_i = (h->ring_i + 1) % sdbe_handle::RING;
    if (h->bpd_used[slot]) HIP_TRY(h, hipEventSynchronize(h->bpd_done[slot]));
    h->bpd_used[slot] = 0;
    return SDBE_OK;
}

int ring_upload(sdbe_handle_t *h, int slot, int n_bp, hipStream_t s) {
    HIP_TRY(h, hipMemcpyAsync(h->d_bpd[slot], h->h_bpd[slot], (size_t)n_bp * sizeof(BpDesc),
                              hipMemcpyHostToDevice, s));
    return SDBE_OK;
}

int ring_release(sdbe_handle_t *h, int slot, hipStream_t s) {
    HIP_TRY(h, hipEventRecord(h->bpd_done[slot], s));
    h->bpd_used[slot] = 1;
    return SDBE_OK;
}

void advance_state(sdbe_stream_state_t &st, int64_t n_frames);

sdbe_stream_state_t state_plus(const sdbe_stream_state_t &st, int64_t n_frames) {
    sdbe_stream_state_t t = st;
    advance_state(t, n_frames);
    return t;
}

// The group's packet index map [frame][cid][fid] and counters; then either
// the int8 rows (k_reorder) or, for the packet-reading stage A, only the fid
// bitmap and the unique count.
int unpack_map(sdbe_handle_t *h, const void *d_pkts, int32_t n_pkts, uint64_t b_zero, hipStream_t s) {
    LAUNCH(h, KID_CLEAR, 1, s, k_clear_map, dim3((h->map_n + 255) / 256), dim3(256), h->map, h->map_n, h->cnt);
    if (n_pkts > 0)
        LAUNCH(h, KID_MAP, 1, s, k_packet_map, dim3((n_pkts + 255) / 256), dim3(256), (const uint32_t *)d_pkts,
               n_pkts, (uint32_t)(b_zero & 0xFFFFFFFFull), h->nfr, h->n_cid, h->n_fid, h->pkt_words, h->map,
               h->cnt);
    HIP_TRY(h, hipGetLastError());
    return SDBE_OK;
}

int unpack_to_rows(sdbe_handle_t *h, const void *d_pkts, int32_t n_pkts, uint64_t b_zero,
                   uint8_t *d_bitmap, hipStream_t s) {
    int rc = unpack_map(h, d_pkts, n_pkts, b_zero, s);
    if (rc) return rc;
    if (h->geom == SDBE_GEOM_8X1024)
        LAUNCH(h, KID_REORDER8, 1, s, k_reorder_8x1024, dim3(G2_CIDS, G2_FIDS, h->nfr), dim3(256),
               (const uint32_t *)d_pkts, (const int *)h->map, h->nfr * 128, h->beng,
               (uint16_t *)(d_bitmap ? d_bitmap : h->bitmap), h->cnt);
    else
        LAUNCH(h, KID_REORDER, 1, s, k_reorder, dim3(128, h->nfr), dim3(256), (const uint32_t *)d_pkts,
               (const int *)h->map, h->nfr * 128, h->beng, d_bitmap ? d_bitmap : h->bitmap, h->cnt);
    HIP_TRY(h, hipGetLastError());
    return SDBE_OK;
}

int unpack_bitmap(sdbe_handle_t *h, const void *d_pkts, int32_t n_pkts, uint64_t b_zero,
                  uint8_t *d_bitmap, hipStream_t s) {
    int rc = unpack_map(h, d_pkts, n_pkts, b_zero, s);
    if (rc) return rc;
    if (h->geom == SDBE_GEOM_8X1024)
        LAUNCH(h, KID_BITMAP8, 1, s, k_group_bitmap_8x1024, dim3(h->nfr), dim3(256), (const int *)h->map,
               (uint16_t *)(d_bitmap ? d_bitmap : h->bitmap), h->cnt);
    else
        LAUNCH(h, KID_BITMAP, 1, s, k_group_bitmap, dim3(h->nfr), dim3(256), (const int *)h->map,
               d_bitmap ? d_bitmap : h->bitmap, h->cnt);
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
    // the counters and latch slot 0 lie back to back on both sides: one copy
    HIP_TRY(h, hipMemcpyAsync(h->h_cnt, h->cnt, CNT_N * sizeof(int) + 2 * sizeof(float), hipMemcpyDeviceToHost, s));
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
    cfg->chunk_blockpols = 256;
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
    (void)hipFree(h->cnt);   // the latch slots too
    (void)hipFree(h->V);
    (void)hipFree(h->Zf);
    (void)hipFree(h->Bc);
    (void)hipFree(h->part);
    (void)hipFree(h->thr);
    (void)hipFree(h->stage);
    (void)hipFree(h->slot_col);
    (void)hipFree(h->trim_tab);
    (void)hipFree(h->ms);
    (void)hipHostFree(h->h_cnt);   // and h_latch
    for (int i = 0; i < sdbe_handle::RING; ++i) {
        (void)hipFree(h->d_bpd[i]);
        (void)hipHostFree(h->h_bpd[i]);
        if (h->bpd_done[i]) (void)hipEventDestroy(h->bpd_done[i]);
    }
    if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
    (void)hipFree(h->g_d);
    (void)hipFree(h->lead);
    (void)hipHostFree(h->g_h);
    if (h->g_done) (void)hipEventDestroy(h->g_done);
    if (h->cap_stream) (void)hipStreamDestroy(h->cap_stream);
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
    if (cfg.chunk_blockpols < 2 || cfg.chunk_blockpols > SDBE_MAX_CHUNK || (cfg.chunk_blockpols & 1)) return SDBE_ERR_ARG;
    if (cfg.max_spectra_blockpols < 1 || cfg.max_batch_blockpols < 0) return SDBE_ERR_ARG;
    if (cfg.c1_tiles_per_wg < 0 || cfg.c1_tiles_per_wg > 64) return SDBE_ERR_ARG;
    if (cfg.c1_interleave < 0 || cfg.c1_interleave > 64) return SDBE_ERR_ARG;
    if (cfg.packet_geometry != SDBE_GEOM_REFERENCE && cfg.packet_geometry != SDBE_GEOM_8X1024) return SDBE_ERR_ARG;
    if (cfg.fft_precision < SDBE_FFT_FP32 || cfg.fft_precision > SDBE_FFT_FP16) return SDBE_ERR_ARG;
    trim::Geometry tg;
    if (cfg.rate_frac != SDBE_RATE_64_39) {
        // the band-trimmed path runs its own FP32 transforms
        if (!trim::geometry(cfg.rate_frac, cfg.skip_mhz, tg) || cfg.fft_precision != SDBE_FFT_FP32)
            return SDBE_ERR_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return SDBE_ERR_NODEV;
    sdbe_handle_t *h = new (std::nothrow) sdbe_handle();
    if (!h) return SDBE_ERR_NOMEM;
    h->device = device;
    h->cfg = cfg;
    h->chunk = cfg.chunk_blockpols;
    h->c1_tpw = cfg.c1_tiles_per_wg ? cfg.c1_tiles_per_wg : C1_TILES_PER_WG;
    h->c1_ilv = cfg.c1_interleave ? cfg.c1_interleave : C1_INTERLEAVE;
    h->trim = tg;
    if (tg.D) {
        // a group is D B-frames of 128 snapshots: 128 blocks of D snapshots
        h->snaps = tg.D;
        h->fpb = tg.E;
        h->nfr = tg.D;
    }
    h->geom = cfg.packet_geometry;
    if (h->geom == SDBE_GEOM_8X1024) {
        h->n_cid = G2_CIDS;
        h->n_fid = G2_FIDS;
        h->pkt_words = G2_PKT_WORDS;
    }
    h->map_n = h->nfr * h->n_cid * h->n_fid;
    h->cap_bp = 2 * SDBE_BLOCKS_PER_GROUP;
    if (cfg.max_spectra_blockpols > h->cap_bp) h->cap_bp = cfg.max_spectra_blockpols;
    {
        const int mb = cfg.max_batch_blockpols ? cfg.max_batch_blockpols : 1024;
        if (mb > h->cap_bp) h->cap_bp = mb;
        if (2 * SDBE_BLOCKS_PER_GROUP > h->cap_bp) h->cap_bp = 2 * SDBE_BLOCKS_PER_GROUP;
    }
    h->st.psn = 1;
    int rc = SDBE_OK;
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) {
        delete h;
        return SDBE_ERR_NODEV;
    }
    {
        int ncu = 0;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && ncu > 0)
            h->n_cu = ncu;
    }
    const size_t C = (size_t)h->chunk;
#define ALLOC(ptr, bytes)                              \
    do {                                               \
        if (hipMalloc((void **)&(ptr), (bytes)) != hipSuccess) { \
            rc = SDBE_ERR_NOMEM;                       \
            goto fail;                                 \
        }                                              \
    } while (0)
    ALLOC(h->map, (size_t)h->map_n * sizeof(int));
    ALLOC(h->beng, 2 * (size_t)h->nfr * 128 * NCH * 2);
    ALLOC(h->bitmap, (size_t)h->nfr * 256);
    // the latch slots follow the counters in one allocation (read back together)
    ALLOC(h->cnt, CNT_N * sizeof(int) + 2 * SDBE_LATCH_SLOTS * sizeof(float));
    h->latch = reinterpret_cast<float *>(h->cnt + CNT_N);
    for (int i = 0; i < sdbe_handle::RING; ++i) ALLOC(h->d_bpd[i], (size_t)h->cap_bp * sizeof(BpDesc));
    ALLOC(h->g_d, (size_t)h->cap_bp * sizeof(BpDesc));
    ALLOC(h->lead, 2 * OUT_FRAME_PAYLOAD);
    // V: stage A's rows (64/39) or the block's concatenated C2R series
    // (band-trimmed, D snapshots of 32768 floats); Zf: stage B's packed
    // spectrum (64/39) or the band-trimmed samples (128 segments of P floats)
    ALLOC(h->V, C * (size_t)(h->snaps > NBLK ? h->snaps : NBLK) * NCH * sizeof(float2));
    ALLOC(h->Zf, C * (tg.D && (size_t)trim::NSEG * tg.P * sizeof(float) > MP * sizeof(float2)
                          ? (size_t)trim::NSEG * tg.P * sizeof(float) : MP * sizeof(float2)));
    ALLOC(h->Bc, C * KHI * BC_ROW * sizeof(float2));
    ALLOC(h->slot_col, C2_SLOT_TAB * sizeof(int));
    ALLOC(h->part, C * C1_TILES * 2 * sizeof(double));
    ALLOC(h->thr, C * sizeof(float2));
    ALLOC(h->stage, C * (KHI / 2) * NCH);
    ALLOC(h->ms, C * sizeof(double2));
#undef ALLOC
    if (hipHostMalloc((void **)&h->h_cnt, CNT_N * sizeof(int) + 2 * sizeof(float), hipHostMallocDefault) !=
        hipSuccess) {
        rc = SDBE_ERR_NOMEM;
        goto fail;
    }
    h->h_latch = reinterpret_cast<float *>(h->h_cnt + CNT_N);
    for (int i = 0; i < sdbe_handle::RING; ++i) {
        if (hipHostMalloc((void **)&h->h_bpd[i], (size_t)h->cap_bp * sizeof(BpDesc), hipHostMallocDefault) !=
                hipSuccess ||
            hipEventCreateWithFlags(&h->bpd_done[i], hipEventDisableTiming) != hipSuccess) {
            rc = SDBE_ERR_NOMEM;
            goto fail;
        }
    }
    if (hipHostMalloc((void **)&h->g_h, (size_t)h->cap_bp * sizeof(BpDesc), hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&h->g_done, hipEventDisableTiming) != hipSuccess) {
        rc = SDBE_ERR_NOMEM;
        goto fail;
    }
    memset(h->h_cnt, 0, CNT_N * sizeof(int));
    memset(h->h_latch, 0, 2 * sizeof(float));
    rc = build_tables(h);
    if (rc != SDBE_OK) goto fail;
    rc = build_slot_table(h);
    if (rc != SDBE_OK) goto fail;
    if (h->trim.D) {
        rc = build_trim_tables(h);
        if (rc != SDBE_OK) goto fail;
    }
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(256), 0, 0, h->cnt, CNT_N, 0);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(256), 0, 0, (int *)h->latch, 2 * SDBE_LATCH_SLOTS, 0);
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
    if (n_pkts < 0 || n_pkts > h->map_n) return SDBE_ERR_CAPACITY;
    HIP_TRY(h, hipSetDevice(h->device));
    hipStream_t s = (hipStream_t)stream;
    int rc = unpack_to_rows(h, d_pkts, n_pkts, b_zero, d_bitmap, s);
    if (rc) return rc;
    const size_t pol_bytes = (size_t)h->nfr * 128 * NCH * 2;
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
    if (n_pkts < 0 || n_pkts > h->map_n) return SDBE_ERR_CAPACITY;
    OutDesc od = out_desc(h);
    const int nsnap = h->nfr * 128;                      // snapshots per polarisation
    const int blocks = nsnap / h->snaps;                 // per polarisation (128)
    if (out_pol_stride_bytes < (int64_t)blocks * h->fpb * od.frame_bytes) return SDBE_ERR_ARG;
    HIP_TRY(h, hipSetDevice(h->device));
    hipStream_t s = (hipStream_t)stream;
    // The first kernel (stage A at 64/39, the per-snapshot C2R band-trimmed)
    // reads the packets in 8- or 16-byte units: with a 16-byte aligned buffer,
    // as device allocations are; otherwise the int8 rows
    const bool direct = n_pkts > 0 && (((uintptr_t)d_pkts) & 15) == 0;
    int rc = direct ? unpack_bitmap(h, d_pkts, n_pkts, b_zero, d_bitmap, s)
                    : unpack_to_rows(h, d_pkts, n_pkts, b_zero, d_bitmap, s);
    if (rc) return rc;
    int slot;
    if ((rc = ring_acquire(h, slot))) return rc;
    const int n_bp = 2 * blocks;
    BpDesc *hb = h->h_bpd[slot];
    const sdbe_config_t &c = h->cfg;
    for (int bp = 0; bp < n_bp; ++bp) {
        const int blk = bp >> 1, pol = bp & 1;
        const int8_t *rows = h->beng + ((size_t)pol * nsnap + (size_t)h->snaps * blk) * NCH * 2;
        uint8_t *out = (uint8_t *)d_out + (int64_t)pol * out_pol_stride_bytes +
                       (int64_t)blk * h->fpb * od.frame_bytes;
        set_bp(hb[bp], rows, out, state_plus(h->st, (int64_t)blk * h->fpb), pol, 0, c.thresh_policy, 0,
               c.fixed_threshold, c.fixed_offset, c.station_id, c.bdc_sideband, c.rx_sideband);
        hb[bp].blk = blk;
    }
    if ((rc = ring_upload(h, slot, n_bp, s))) return rc;
    // Re-framed output: the stream of 2048-word frames starts align_words
    // into the first aligned group; a group's last lead_words words open the
    // next group's first frame (the output thread's time alignment,
    // vdif_out_net_thread.c:203-290, done before the copy to the host).
    int n_out = blocks * h->fpb;
    if (h->align_pending || h->align_active) {
        int64_t sigma;
        if (h->align_pending) {
            const int64_t fr = (h->align_words + 2047) / 2048;
            n_out = SDBE_OUT_FRAMES_PER_GROUP - (int)fr;
            h->lead_words = (int)(fr * 2048 - h->align_words);
            sigma = -h->align_words;
            h->align_pending = 0;
            h->align_active = 1;
        } else {
            sigma = h->lead_words;
            if (h->lead_words > 0) {
                hipLaunchKernelGGL(k_lead_copy, dim3((h->lead_words + 255) / 256, 2), dim3(256), 0, s,
                                   (const uint8_t *)h->lead, (uint8_t *)d_out, out_pol_stride_bytes, od.hdr_bytes,
                                   4 * h->lead_words);
                HIP_TRY(h, hipGetLastError());
            }
        }
        od.shift_bytes = 4 * sigma;
        od.lim_bytes = (int64_t)n_out * OUT_FRAME_PAYLOAD;
        od.lead = h->lead;
        od.bp_headers = 0;
    }
    rc = run_resampler(h, h->d_bpd[slot], n_bp, od, false, nullptr, s, direct ? (const uint32_t *)d_pkts : nullptr,
                       c.thresh_policy == SDBE_THRESH_REFERENCE);
    if (rc) return rc;
    if ((rc = ring_release(h, slot, s))) return rc;
    if (!od.bp_headers && od.hdr_bytes) {
        HdrLin hd;
        hd.base = (uint8_t *)d_out;
        hd.pol_stride = out_pol_stride_bytes;
        hd.frame_bytes = od.frame_bytes;
        hd.n_frames = n_out;
        hd.secs = h->st.secs;
        hd.df = h->st.df;
        hd.psn = h->st.psn;
        hd.w1hi = (h->st.ref_epoch & 0x3Fu) << 24;
        hd.w3 = (uint32_t)c.station_id | (1u << 26);
        hd.w4 = ((uint32_t)(c.bdc_sideband & 1u) << 1) | ((uint32_t)(c.rx_sideband & 1u) << 2) | (2u << 24);
        LAUNCH(h, KID_HEADERS, 0, s, k_headers_lin, dim3((2 * n_out + 255) / 256), dim3(256), hd);
        HIP_TRY(h, hipGetLastError());
    }
    h->n_frames_last = n_out;
    advance_state(h->st, n_out);
    h->n_packets_last = n_pkts;
    h->n_bp_last = n_bp;
    h->have_group = 1;
    return finish_call(h, s);
}

namespace {
// Descriptors of n_bp independent block-pols laid out back to back
// (sdbe_process_spectra / sdbe_resample_spectra_f32).
int spectra_call(sdbe_handle_t *h, const int8_t *d_beng, int32_t n_bp, uint8_t *out, int64_t out_bp_stride,
                 bool f32, double *d_stats, hipStream_t s) {
    int slot, rc;
    if ((rc = ring_acquire(h, slot))) return rc;
    BpDesc *hb = h->h_bpd[slot];
    const sdbe_config_t &c = h->cfg;
    for (int bp = 0; bp < n_bp; ++bp)
        set_bp(hb[bp], d_beng + (size_t)bp * h->snaps * NCH * 2, out + (int64_t)bp * out_bp_stride,
               state_plus(h->st, (int64_t)bp * h->fpb), 0, 0, c.thresh_policy, 0, c.fixed_threshold,
               c.fixed_offset, c.station_id, c.bdc_sideband, c.rx_sideband);
    if ((rc = ring_upload(h, slot, n_bp, s))) return rc;
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(256), 0, s, h->cnt, CNT_N, 0);
    rc = run_resampler(h, h->d_bpd[slot], n_bp, out_desc(h), f32, d_stats, s, nullptr,
                       c.thresh_policy == SDBE_THRESH_REFERENCE);
    if (rc) return rc;
    return ring_release(h, slot, s);
}
}  // namespace

int sdbe_process_spectra(sdbe_handle_t *h, const int8_t *d_beng, int32_t n_bp, void *d_out, void *stream) {
    if (!h || !d_beng || !d_out || n_bp <= 0) return SDBE_ERR_ARG;
    if (n_bp > h->cfg.max_spectra_blockpols) return SDBE_ERR_CAPACITY;
    HIP_TRY(h, hipSetDevice(h->device));
    hipStream_t s = (hipStream_t)stream;
    const OutDesc od = out_desc(h);
    int rc = spectra_call(h, d_beng, n_bp, (uint8_t *)d_out, (int64_t)h->fpb * od.frame_bytes, false, nullptr, s);
    if (rc) return rc;
    advance_state(h->st, (int64_t)n_bp * h->fpb);
    h->n_frames_last = n_bp * h->fpb;
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
    const int64_t spb = h->trim.D ? (int64_t)trim::NSEG * h->trim.P : (int64_t)P20 * 2;   // samples per block-pol
    int rc = spectra_call(h, d_beng, n_bp, (uint8_t *)d_samples, spb * (int64_t)sizeof(float), true, d_stats, s);
    if (rc) return rc;
    h->n_packets_last = 0;
    h->n_bp_last = n_bp;
    h->have_group = 0;
    return finish_call(h, s);
}

namespace {
// Descriptors of a ragged batch: stream by stream, block by block, pol 0 then 1.
int fill_batch(sdbe_handle_t *h, const sdbe_stream_desc_t *streams, int32_t n_streams, BpDesc *hb) {
    const OutDesc od = out_desc(h);
    int n = 0;
    for (int i = 0; i < n_streams; ++i) {
        const sdbe_stream_desc_t &d = streams[i];
        for (int b = 0; b < d.n_blocks; ++b)
            for (int pol = 0; pol < 2; ++pol) {
                uint8_t *out = (uint8_t *)d.d_out + (int64_t)pol * d.out_pol_stride_bytes +
                               (int64_t)b * OUT_FRAMES_BLOCK * od.frame_bytes;
                set_bp(hb[n++], d.d_beng[pol] + (size_t)b * NBLK * NCH * 2, out,
                       state_plus(*d.state, (int64_t)b * OUT_FRAMES_BLOCK), pol, d.flip, d.thresh_policy,
                       d.latch_slot, d.fixed_threshold, d.fixed_offset, d.station_id, d.bdc_sideband,
                       d.rx_sideband);
            }
    }
    return n;
}

int check_batch(const sdbe_handle_t *h, const sdbe_stream_desc_t *streams, int32_t n_streams, int &n_bp) {
    if (!streams || n_streams <= 0) return SDBE_ERR_ARG;
    const OutDesc od = out_desc(h);
    int64_t total = 0;
    for (int i = 0; i < n_streams; ++i) {
        const sdbe_stream_desc_t &d = streams[i];
        if (d.n_blocks < 0 || !d.state || (d.n_blocks > 0 && (!d.d_beng[0] || !d.d_beng[1] || !d.d_out)))
            return SDBE_ERR_ARG;
        if (d.thresh_policy < SDBE_THRESH_REFERENCE || d.thresh_policy > SDBE_THRESH_FIXED) return SDBE_ERR_ARG;
        if (d.latch_slot < 0 || d.latch_slot >= SDBE_LATCH_SLOTS) return SDBE_ERR_ARG;
        if (d.out_pol_stride_bytes < (int64_t)d.n_blocks * OUT_FRAMES_BLOCK * od.frame_bytes) return SDBE_ERR_ARG;
        total += 2 * (int64_t)d.n_blocks;
    }
    if (total > h->cap_bp) return SDBE_ERR_CAPACITY;
    n_bp = (int)total;
    return SDBE_OK;
}

int batch_graph(sdbe_handle_t *h, int n_bp) {
    if (h->gexec && h->g_nbp == n_bp) return SDBE_OK;
    if (h->gexec) {
        HIP_TRY(h, hipEventSynchronize(h->g_done));
        HIP_TRY(h, hipGraphExecDestroy(h->gexec));
        h->gexec = nullptr;
    }
    if (!h->cap_stream) HIP_TRY(h, hipStreamCreateWithFlags(&h->cap_stream, hipStreamNonBlocking));
    hipStream_t cs = h->cap_stream;
    HIP_TRY(h, hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    h->capturing = 1;
    hipGraph_t g = nullptr;
    hipError_t e1 = hipMemcpyAsync(h->g_d, h->g_h, (size_t)n_bp * sizeof(BpDesc), hipMemcpyHostToDevice, cs);
    if (e1 == hipSuccess) {
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(256), 0, cs, h->cnt, CNT_N, 0);
        e1 = hipGetLastError();
    }
    int rc = e1 == hipSuccess ? run_resampler(h, h->g_d, n_bp, out_desc(h), false, nullptr, cs) : SDBE_OK;
    hipError_t e2 = hipStreamEndCapture(cs, &g);
    h->capturing = 0;
    if (e1 != hipSuccess) return hip_fail(h, e1);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e2 != hipSuccess) return hip_fail(h, e2);
    hipError_t e3 = hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e3 != hipSuccess) {
        h->gexec = nullptr;
        return hip_fail(h, e3);
    }
    h->g_nbp = n_bp;
    return SDBE_OK;
}
}  // namespace

int sdbe_process_blocks(sdbe_handle_t *h, const sdbe_stream_desc_t *streams, int32_t n_streams, void *stream) {
    if (h && h->trim.D) return SDBE_ERR_STATE;   // ragged batches: 64/39 resampler only
    if (!h) return SDBE_ERR_ARG;
    int n_bp = 0;
    int rc = check_batch(h, streams, n_streams, n_bp);
    if (rc) return rc;
    HIP_TRY(h, hipSetDevice(h->device));
    hipStream_t s = (hipStream_t)stream;
    if (n_bp > 0) {
        if (h->cfg.use_graph) {
            if ((rc = batch_graph(h, n_bp))) return rc;
            // the graph copies g_h in: wait until its previous launch has done so
            if (h->g_used) HIP_TRY(h, hipEventSynchronize(h->g_done));
            fill_batch(h, streams, n_streams, h->g_h);
            HIP_TRY(h, hipGraphLaunch(h->gexec, s));
            HIP_TRY(h, hipEventRecord(h->g_done, s));
            h->g_used = 1;
        } else {
            int slot;
            if ((rc = ring_acquire(h, slot))) return rc;
            fill_batch(h, streams, n_streams, h->h_bpd[slot]);
            if ((rc = ring_upload(h, slot, n_bp, s))) return rc;
            hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(256), 0, s, h->cnt, CNT_N, 0);
            bool any_ref = false;
            for (int i = 0; i < n_streams; ++i)
                any_ref |= streams[i].n_blocks > 0 && streams[i].thresh_policy == SDBE_THRESH_REFERENCE;
            if ((rc = run_resampler(h, h->d_bpd[slot], n_bp, out_desc(h), false, nullptr, s, nullptr, any_ref)))
                return rc;
            if ((rc = ring_release(h, slot, s))) return rc;
        }
    }
    for (int i = 0; i < n_streams; ++i) advance_state(*streams[i].state, (int64_t)streams[i].n_blocks * OUT_FRAMES_BLOCK);
    h->n_packets_last = 0;
    h->n_bp_last = n_bp;
    h->have_group = 0;
    return finish_call(h, s);
}

int sdbe_set_output_alignment(sdbe_handle_t *h, int32_t enable, int32_t skip_int32, int32_t skip_frames) {
    if (h && h->trim.D) return SDBE_ERR_STATE;   // re-framed output: 64/39 resampler only
    if (!h) return SDBE_ERR_ARG;
    if (!enable) {
        h->align_pending = h->align_active = 0;
        h->lead_words = 0;
        return SDBE_OK;
    }
    const int64_t words = (int64_t)skip_int32 + 2048 * (int64_t)skip_frames;
    if (skip_int32 < 0 || skip_frames < 0 || words >= (int64_t)(SDBE_OUT_FRAMES_PER_GROUP - 1) * 2048)
        return SDBE_ERR_ARG;
    h->align_words = words;
    h->align_pending = 1;
    h->align_active = 0;
    return SDBE_OK;
}

int sdbe_reset_latch(sdbe_handle_t *h, int32_t slot, void *stream) {
    if (!h || slot < 0 || slot >= SDBE_LATCH_SLOTS) return SDBE_ERR_ARG;
    HIP_TRY(h, hipSetDevice(h->device));
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, (hipStream_t)stream, (int *)(h->latch + 2 * slot), 2, 0);
    HIP_TRY(h, hipGetLastError());
    return SDBE_OK;
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
    o->n_missing = h->have_group ? h->map_n - h->h_cnt[CNT_UNIQUE] : 0;
    o->n_blockpols = h->n_bp_last;
    o->n_frames_per_pol = h->n_frames_last;
    o->latched_threshold = h->h_latch[0];
    o->latched_offset = h->h_latch[1];
    return SDBE_OK;
}

int sdbe_profile_enable(sdbe_handle_t *h, int enable) {
    if (!h) return SDBE_ERR_ARG;
    h->prof = enable ? 1 : 0;
    return SDBE_OK;
}

int sdbe_profile_select(sdbe_handle_t *h, const char *kernel) {
    if (!h) return SDBE_ERR_ARG;
    if (!kernel || !kernel[0]) {
        h->prof_kid = -1;
        return SDBE_OK;
    }
    for (int k = 0; k < KID_N; ++k)
        if (strcmp(kKernelNames[k], kernel) == 0) {
            h->prof_kid = k;
            return SDBE_OK;
        }
    return SDBE_ERR_ARG;
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

// ============================================ allocation-free spectra entry
// sdbe_process_spectra_noalloc keeps the pipeline's working memory in one
// file-scope device array, `scratch`, instead of a handle's allocations: the
// chunk buffers of NA_CHUNK block-pols, the twiddle and slot tables, the
// counters, the latch slots and the descriptors of up to
// SDBE_NOALLOC_MAX_BLOCKPOLS block-pols.  Every GPU the library runs on holds
// its own copy of the array until the process ends.  In host code `scratch`
// names a host-side stand-in, never the device memory: its device address is
// taken once per device with hipGetSymbolAddress (na_setup) and only that
// address is handed to kernels and copies.
namespace {
constexpr int NA_CHUNK = 2;
constexpr int NA_MAX_DEV = 64;
constexpr size_t na_al(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr size_t NA_V = 0;
constexpr size_t NA_ZF = NA_V + na_al((size_t)NA_CHUNK * NBLK * NCH * sizeof(float2));
constexpr size_t NA_BC = NA_ZF + na_al((size_t)NA_CHUNK * MP * sizeof(float2));
constexpr size_t NA_STAGE = NA_BC + na_al((size_t)NA_CHUNK * KHI * BC_ROW * sizeof(float2));
constexpr size_t NA_PART = NA_STAGE + na_al((size_t)NA_CHUNK * (KHI / 2) * NCH);
constexpr size_t NA_THR = NA_PART + na_al((size_t)NA_CHUNK * C1_TILES * 2 * sizeof(double));
constexpr size_t NA_MS = NA_THR + na_al((size_t)NA_CHUNK * sizeof(float2));
constexpr size_t NA_SLOT = NA_MS + na_al((size_t)NA_CHUNK * sizeof(double2));
constexpr size_t NA_TAB = NA_SLOT + na_al((size_t)C2_SLOT_TAB * sizeof(int));
constexpr size_t NA_CNT = NA_TAB + na_al(kTabEntries * sizeof(float2));
constexpr size_t NA_LATCH = NA_CNT + na_al((size_t)CNT_N * sizeof(int));
constexpr size_t NA_BPD = NA_LATCH + na_al((size_t)2 * SDBE_LATCH_SLOTS * sizeof(float));
constexpr size_t NA_BYTES = NA_BPD + na_al((size_t)SDBE_NOALLOC_MAX_BLOCKPOLS * sizeof(BpDesc));
static_assert(NA_BYTES <= SDBE_NOALLOC_SCRATCH_BYTES, "no-alloc scratch exceeds its stated size");
}  // namespace

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

namespace {
struct NaDevice {
    int ready;
    int n_cu;
    unsigned char *ptr;   // device address of `scratch` on this device
};
NaDevice g_na[NA_MAX_DEV];
std::mutex g_na_mu;
// host images of the tables, filled once (static storage: no allocation)
float2 g_na_tab_host[kTabEntries];
int g_na_slot_host[C2_SLOT_TAB];
int g_na_seen_host[NCH];

// Once per device, with that device current: the address of `scratch` and its
// tables.  Returns the device's record or sets *rc.
const NaDevice *na_setup(int dev, int *rc) {
    std::lock_guard<std::mutex> lock(g_na_mu);
    NaDevice &d = g_na[dev];
    if (d.ready) return &d;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        *rc = SDBE_ERR_HIP;
        return nullptr;
    }
    unsigned char *base = (unsigned char *)p;
    tables_fill_host(g_na_tab_host);
    if (!slot_table_fill_host(g_na_slot_host, g_na_seen_host)) {
        *rc = SDBE_ERR_STATE;
        return nullptr;
    }
    if (hipMemcpy(base + NA_TAB, g_na_tab_host, kTabEntries * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(base + NA_SLOT, g_na_slot_host, C2_SLOT_TAB * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
        *rc = SDBE_ERR_HIP;
        return nullptr;
    }
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0) ncu = 256;
    d.n_cu = ncu;
    d.ptr = base;
    d.ready = 1;
    return &d;
}
}  // namespace

extern "C" int sdbe_process_spectra_noalloc(const sdbe_config_t *cfg, sdbe_stream_state_t *state,
                                            const int8_t *d_beng, int32_t n_bp, void *d_out, void *stream) {
    if (!cfg || !state || !d_beng || !d_out || n_bp <= 0) return SDBE_ERR_ARG;
    if (n_bp > SDBE_NOALLOC_MAX_BLOCKPOLS) return SDBE_ERR_CAPACITY;
    if (cfg->thresh_policy < SDBE_THRESH_REFERENCE || cfg->thresh_policy > SDBE_THRESH_FIXED) return SDBE_ERR_ARG;
    if (cfg->out_mode != SDBE_OUT_PAYLOAD && cfg->out_mode != SDBE_OUT_FRAMES) return SDBE_ERR_ARG;
    if (cfg->c1_tiles_per_wg < 0 || cfg->c1_tiles_per_wg > 64) return SDBE_ERR_ARG;
    if (cfg->c1_interleave < 0 || cfg->c1_interleave > 64) return SDBE_ERR_ARG;
    if (cfg->fft_precision < SDBE_FFT_FP32 || cfg->fft_precision > SDBE_FFT_FP16) return SDBE_ERR_ARG;
    if (cfg->rate_frac != SDBE_RATE_64_39) return SDBE_ERR_ARG;
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NA_MAX_DEV) return SDBE_ERR_NODEV;
    int rc = SDBE_OK;
    const NaDevice *nd = na_setup(dev, &rc);
    if (!nd) return rc;
    unsigned char *base = nd->ptr;
    // a view of the pipeline's buffers inside `scratch` (a stack object: its
    // profiling vectors stay empty and it allocates nothing)
    sdbe_handle h;
    h.device = dev;
    h.cfg = *cfg;
    h.chunk = NA_CHUNK;
    h.c1_tpw = cfg->c1_tiles_per_wg ? cfg->c1_tiles_per_wg : C1_TILES_PER_WG;
    h.c1_ilv = cfg->c1_interleave ? cfg->c1_interleave : C1_INTERLEAVE;
    h.n_cu = nd->n_cu;
    h.tb = tables_bind((const float2 *)(base + NA_TAB));
    h.V = (float2 *)(base + NA_V);
    h.Zf = (float2 *)(base + NA_ZF);
    h.Bc = (float2 *)(base + NA_BC);
    h.stage = base + NA_STAGE;
    h.part = (double *)(base + NA_PART);
    h.thr = (float2 *)(base + NA_THR);
    h.ms = (double2 *)(base + NA_MS);
    h.slot_col = (int *)(base + NA_SLOT);
    h.cnt = (int *)(base + NA_CNT);
    h.latch = (float *)(base + NA_LATCH);
    BpDesc *bpd = (BpDesc *)(base + NA_BPD);
    hipStream_t s = (hipStream_t)stream;
    const OutDesc od = out_desc(&h);
    // the array keeps the previous call's contents: the counters and the
    // REFERENCE latch (slot 0) start at zero in every call
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(256), 0, s, h.cnt, CNT_N, 0);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, (int *)h.latch, 2, 0);
    SpecDescArgs a;
    a.rows = d_beng;
    a.out = (uint8_t *)d_out;
    a.row_stride = (int64_t)NBLK * NCH * 2;
    a.out_stride = (int64_t)OUT_FRAMES_BLOCK * od.frame_bytes;
    a.psn = state->psn;
    a.secs = state->secs;
    a.df = state->df;
    a.w1hi = (state->ref_epoch & 0x3Fu) << 24;
    a.w3 = (uint32_t)cfg->station_id | (1u << 26);
    a.w4 = ((uint32_t)(cfg->bdc_sideband & 1u) << 1) | ((uint32_t)(cfg->rx_sideband & 1u) << 2) | (2u << 24);
    a.policy = cfg->thresh_policy;
    a.fpb = OUT_FRAMES_BLOCK;
    a.fixed_th = cfg->fixed_threshold;
    a.fixed_off = cfg->fixed_offset;
    hipLaunchKernelGGL(k_spectra_desc, dim3((n_bp + 63) / 64), dim3(64), 0, s, a, n_bp, bpd);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return SDBE_ERR_HIP;
    rc = run_resampler(&h, bpd, n_bp, od, false, nullptr, s, nullptr, cfg->thresh_policy == SDBE_THRESH_REFERENCE);
    if (rc) return rc;
    advance_state(*state, (int64_t)n_bp * OUT_FRAMES_BLOCK);
    return SDBE_OK;
}
