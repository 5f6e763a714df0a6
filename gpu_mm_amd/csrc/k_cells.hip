// Cell-level map utilities on gfx950: gather and reduce of 3x64x64 map cells through an index
// map (the device side of a cell exchange between map layouts), and local -> global map copy.
//
// Reference: src_lib/cell_broadcast.cu:19-36, src_lib/cell_reduce.cu:19-44 and
// src_lib/local_map_to_global.cu:11-45, which are host loops; here they run on the device.

#include "gmm_internal.hpp"

namespace gmm {

namespace {

constexpr int CL_THREADS = 256;
constexpr int CELL_FLOATS = 3 * CELL_NPIX;        // 12288
constexpr int CELL_VEC = CELL_FLOATS / 4;         // float4 per cell

// dst cell i = src cell index_map[i]. One block per destination cell; an out-of-range index
// sets *bad and leaves the cell unwritten.
__global__ void __launch_bounds__(CL_THREADS)
cell_gather_kernel(float4 *__restrict__ dst, const float4 *__restrict__ src, const int64_t *__restrict__ index_map,
                   int64_t nsrc, uint32_t *__restrict__ bad)
{
    const int64_t j = index_map[blockIdx.x];
    if (j < 0 || j >= nsrc) {
        if (threadIdx.x == 0)
            atomicOr(bad, 1u);
        return;
    }
    const float4 *s = src + j * CELL_VEC;
    float4 *d = dst + int64_t(blockIdx.x) * CELL_VEC;
    for (int k = threadIdx.x; k < CELL_VEC; k += CL_THREADS)
        d[k] = s[k];
}

// dst cell d = sum of src cells i with index_map[i] == d, in increasing i (deterministic); zero
// if there is none. One block per destination cell, which scans the index map in LDS-sized
// tiles; out-of-range indices set *bad.
__global__ void __launch_bounds__(CL_THREADS)
cell_reduce_kernel(float4 *__restrict__ dst, int64_t ndst, const float4 *__restrict__ src,
                   const int64_t *__restrict__ index_map, int64_t nsrc, uint32_t *__restrict__ bad)
{
    constexpr int TILE = 1024;
    __shared__ int32_t s_hit[TILE];
    __shared__ int s_n;
    __shared__ int s_wcnt[CL_THREADS / 64];
    const int64_t d = blockIdx.x;
    float4 acc[CELL_VEC / CL_THREADS];
#pragma unroll
    for (int r = 0; r < CELL_VEC / CL_THREADS; r++)
        acc[r] = make_float4(0.f, 0.f, 0.f, 0.f);

    for (int64_t t0 = 0; t0 < nsrc; t0 += TILE) {
        if (threadIdx.x == 0)
            s_n = 0;
        __syncthreads();
        // Ordered compaction of this tile's matches: each wave ballots, then waves append in
        // order through a running count.
        for (int w = 0; w < TILE / CL_THREADS; w++) {
            const int64_t i = t0 + w * CL_THREADS + threadIdx.x;
            bool hit = false;
            if (i < nsrc) {
                const int64_t v = index_map[i];
                if (v < 0 || v >= ndst) {
                    if (d == 0)
                        atomicOr(bad, 1u);
                }
                else
                    hit = (v == d);
            }
            // Block-wide ordered append: per-wave ballot, waves in order.
            const uint64_t m = __ballot(hit);
            const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
            if (lane == 0)
                s_wcnt[wave] = __popcll(m);
            __syncthreads();
            int before = s_n;
            for (int k = 0; k < wave; k++)
                before += s_wcnt[k];
            if (hit)
                s_hit[before + __popcll(m & ((uint64_t(1) << lane) - 1))] = int32_t(i - t0);
            __syncthreads();
            if (threadIdx.x == 0) {
                int tot = 0;
                for (int k = 0; k < CL_THREADS / 64; k++)
                    tot += s_wcnt[k];
                s_n += tot;
            }
            __syncthreads();
        }
        const int n = s_n;
        for (int h = 0; h < n; h++) {
            const float4 *s = src + (t0 + s_hit[h]) * CELL_VEC;
#pragma unroll
            for (int r = 0; r < CELL_VEC / CL_THREADS; r++) {
                const float4 v = s[r * CL_THREADS + threadIdx.x];
                acc[r].x += v.x;
                acc[r].y += v.y;
                acc[r].z += v.z;
                acc[r].w += v.w;
            }
        }
        __syncthreads();
    }
    float4 *o = dst + d * CELL_VEC;
#pragma unroll
    for (int r = 0; r < CELL_VEC / CL_THREADS; r++)
        o[r * CL_THREADS + threadIdx.x] = acc[r];
}

// Global map (3, ny, nx) from a local map; pixels of absent cells are zero.
__global__ void __launch_bounds__(CL_THREADS)
local_to_global_kernel(float *__restrict__ dst, const float *__restrict__ src, const int64_t *__restrict__ cell_offsets,
                       int ny, int nx, int nycells, int nxcells, int64_t ystride, int64_t polstride)
{
    const int64_t n = 3 * int64_t(ny) * nx;
    for (int64_t i = int64_t(blockIdx.x) * CL_THREADS + threadIdx.x; i < n; i += int64_t(gridDim.x) * CL_THREADS) {
        const int64_t p = i / (int64_t(ny) * nx);
        const int64_t r = i - p * ny * nx;
        const int y = int(r / nx), x = int(r - int64_t(y) * nx);
        const int cy = y >> 6, cx = x >> 6;
        const int64_t off = (cy < nycells && cx < nxcells) ? cell_offsets[int64_t(cy) * nxcells + cx] : -1;
        dst[i] = (off >= 0) ? src[off + p * polstride + (y & 63) * ystride + (x & 63)] : 0.0f;
    }
}

}  // namespace

void launch_cell_broadcast(float *dst, int64_t ncells, const float *src, int64_t nsrc, const int64_t *index_map,
                           uint32_t *bad, hipStream_t stream)
{
    if (ncells <= 0)
        return;
    cell_gather_kernel<<<unsigned(ncells), CL_THREADS, 0, stream>>>(reinterpret_cast<float4 *>(dst),
                                                                    reinterpret_cast<const float4 *>(src), index_map,
                                                                    nsrc, bad);
    GMM_HIP(hipGetLastError());
}

void launch_cell_reduce(float *dst, int64_t ndst, const float *src, int64_t nsrc, const int64_t *index_map,
                        uint32_t *bad, hipStream_t stream)
{
    if (ndst <= 0)
        return;
    cell_reduce_kernel<<<unsigned(ndst), CL_THREADS, 0, stream>>>(reinterpret_cast<float4 *>(dst), ndst,
                                                                  reinterpret_cast<const float4 *>(src), index_map,
                                                                  nsrc, bad);
    GMM_HIP(hipGetLastError());
}

void launch_local_to_global(float *dst, const float *src, const gmm_lpix &lp, hipStream_t stream)
{
    const int64_t n = 3 * lp.nypix_global * lp.nxpix_global;
    const int grid = int(std::min<int64_t>((n + CL_THREADS - 1) / CL_THREADS, 16384));
    local_to_global_kernel<<<grid, CL_THREADS, 0, stream>>>(dst, src, lp.offsets_gpu, int(lp.nypix_global),
                                                            int(lp.nxpix_global), int(lp.nycells), int(lp.nxcells),
                                                            lp.ystride, lp.polstride);
    GMM_HIP(hipGetLastError());
}

}  // namespace gmm
