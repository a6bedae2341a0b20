// Input-format variants of the stage: the 2015-format B-engine reorder.
//
// sdbe_reorder_2015 restates reorderTz_smem
// (software/prototyping/cuda/kernel_template.py:107-164) as one launch over all
// output frames.  Workgroup = (64-channel tile, 64-snapshot half, output frame);
// 256 threads move 64 x 64 complex values through an LDS tile:
//   read:  lanes run along the input snapshots of one channel (the input is
//          channel-major, snapshots contiguous), 512 B per wave-load except where
//          the a-d shift wraps or the frame changes at snapshot 69;
//   write: lanes run along the channels of one output snapshot (output rows are
//          channel-contiguous), 512 B per wave-store.
// Pure data movement: HBM-bound, 2 x 8 B per element.
#include <hip/hip_runtime.h>

#include "../../include/sdbe_formats.h"

namespace {

constexpr int NCH = SDBE_REORDER2015_CHANNELS;
constexpr int NSN = SDBE_REORDER2015_SNAPSHOTS;
constexpr int SPLIT = SDBE_REORDER2015_SPLIT;
constexpr int TILE = 64;

__global__ void __launch_bounds__(256)
k_reorder_2015(const float2 *__restrict__ in, int n_frames, float2 *__restrict__ out) {
    __shared__ float2 tile[TILE][TILE + 1];
    const int t = threadIdx.x;
    const int c0 = blockIdx.x * TILE;          // channel tile
    const int s0 = blockIdx.y * TILE;          // output snapshot half
    const int f = blockIdx.z;                  // output frame, < n_frames - 1
    // read: j = output snapshot in the tile, channel rows i = (t >> 6) + 4 k
    {
        const int j = t & (TILE - 1), s = s0 + j;
        const int fi = f + (s >= SPLIT ? 1 : 0);
#pragma unroll 4
        for (int k = 0; k < TILE / 4; ++k) {
            const int i = (t >> 6) + 4 * k, c = c0 + i;
            const int si = ((c >> 2) & 1) == 0 ? ((s + 2) & (NSN - 1)) : s;
            tile[i][j] = in[((size_t)c * n_frames + fi) * NSN + si];
        }
    }
    __syncthreads();
    // write: i = channel in the tile, snapshots j = (t >> 6) + 4 k
    {
        const int i = t & (TILE - 1);
#pragma unroll 4
        for (int k = 0; k < TILE / 4; ++k) {
            const int j = (t >> 6) + 4 * k;
            float2 *row = out + ((size_t)f * NSN + s0 + j) * (NCH + 1);
            row[c0 + i] = tile[i][j];
            if (c0 + TILE == NCH && i == 0) row[NCH] = make_float2(0.f, 0.f);
        }
    }
}

}  // namespace

extern "C" int sdbe_reorder_2015(const void *d_in, int32_t n_frames, void *d_out, void *stream) {
    if (!d_in || !d_out || n_frames < 2) return SDBE_ERR_ARG;
    // in and out must not overlap
    const char *ib = (const char *)d_in, *ob = (const char *)d_out;
    const size_t in_bytes = (size_t)NCH * n_frames * NSN * sizeof(float2);
    const size_t out_bytes = (size_t)(n_frames - 1) * NSN * (NCH + 1) * sizeof(float2);
    if (ib < ob + out_bytes && ob < ib + in_bytes) return SDBE_ERR_ARG;
    if (n_frames - 1 > 65535) return SDBE_ERR_ARG;   // grid z limit
    hipLaunchKernelGGL(k_reorder_2015, dim3(NCH / TILE, NSN / TILE, n_frames - 1), dim3(256), 0,
                       (hipStream_t)stream, (const float2 *)d_in, (int)n_frames, (float2 *)d_out);
    return hipGetLastError() == hipSuccess ? SDBE_OK : SDBE_ERR_HIP;
}
