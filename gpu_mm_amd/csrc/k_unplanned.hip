// Plan-free map2tod / tod2map on gfx950: one lane per sample, bilinear gather straight from
// the local map in HBM (through L2), or scatter with float atomics.
// Reference: src_lib/map2tod_unplanned.cu:17-127, src_lib/tod2map_unplanned.cu:17-121.

#include "gmm_internal.hpp"

namespace gmm {

namespace {

constexpr int UP_THREADS = 256;

__device__ inline uint32_t wave_or_u(uint32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v |= __shfl_xor(v, off, 64);
    return v;
}

struct LocalMapView
{
    const int64_t *__restrict__ cell_offsets;
    int nycells, nxcells;
    int64_t ystride, polstride;

    // Offset of the T value of pixel (iy, ix), or -1 if its cell is not in the local map.
    __device__ inline int64_t offset(int iy, int ix) const
    {
        const int cy = iy >> 6, cx = ix >> 6;
        if (cy >= nycells || cx >= nxcells)
            return -1;
        const int64_t o = cell_offsets[int64_t(cy) * nxcells + cx];
        return (o < 0) ? int64_t(-1) : (o + int64_t(iy & 63) * ystride + (ix & 63));
    }
};

template<bool Periodic>
__global__ void __launch_bounds__(UP_THREADS)
unplanned_map2tod_kernel(float *__restrict__ tod, const float *__restrict__ lmap, const float *__restrict__ xp,
                         LocalMapView mv, uint32_t *__restrict__ errflags, int64_t nsamp,
                         int64_t nsamp_per_block, int ny, int nx, bool partial, bool accum)
{
    __shared__ uint32_t s_err;
    if (threadIdx.x == 0)
        s_err = 0;
    __syncthreads();

    const int64_t s0 = int64_t(blockIdx.x) * nsamp_per_block;
    const int64_t s1 = min(nsamp, s0 + nsamp_per_block);
    uint32_t err = 0;

    for (int64_t s = s0 + threadIdx.x; s < s1; s += UP_THREADS) {
        const float y = xp[s], x = xp[s + nsamp], alpha = xp[s + 2 * nsamp];
        float sn, cs;
        sincos_angle(2.0f * alpha, sn, cs);
        Pixel<float> p;
        err |= locate_pixel(y, x, ny, nx, Periodic, p);

        const int py[4] = { p.iy0, p.iy0, p.iy1, p.iy1 };
        const int px[4] = { p.ix0, p.ix1, p.ix0, p.ix1 };
        const float wy0 = 1.0f - p.dy, wx0 = 1.0f - p.dx;
        const float w[4] = { wy0 * wx0, wy0 * p.dx, p.dy * wx0, p.dy * p.dx };
        float val = 0.0f;
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const int64_t o = mv.offset(py[c], px[c]);
            float v = 0.0f;
            if (o >= 0)
                v = lmap[o] + cs * lmap[o + mv.polstride] + sn * lmap[o + 2 * mv.polstride];
            else if (!partial)
                err |= ERR_NOT_IN_PIXELIZATION;
            val += w[c] * v;
        }
        tod[s] = accum ? (tod[s] + val) : val;
    }

    err = wave_or_u(err);
    if ((threadIdx.x & 63) == 0 && err)
        atomicOr(&s_err, err);
    __syncthreads();
    if (threadIdx.x == 0)
        errflags[blockIdx.x] = s_err;
}

template<bool Periodic>
__global__ void __launch_bounds__(UP_THREADS)
unplanned_tod2map_kernel(float *__restrict__ lmap, const float *__restrict__ tod, const float *__restrict__ xp,
                         LocalMapView mv, uint32_t *__restrict__ errflags, int64_t nsamp,
                         int64_t nsamp_per_block, int ny, int nx, bool partial)
{
    __shared__ uint32_t s_err;
    if (threadIdx.x == 0)
        s_err = 0;
    __syncthreads();

    const int64_t s0 = int64_t(blockIdx.x) * nsamp_per_block;
    const int64_t s1 = min(nsamp, s0 + nsamp_per_block);
    uint32_t err = 0;

    for (int64_t s = s0 + threadIdx.x; s < s1; s += UP_THREADS) {
        const float y = xp[s], x = xp[s + nsamp], alpha = xp[s + 2 * nsamp];
        const float t = tod[s];
        float sn, cs;
        sincos_angle(2.0f * alpha, sn, cs);
        const float q = t * cs, u = t * sn;
        Pixel<float> p;
        err |= locate_pixel(y, x, ny, nx, Periodic, p);

        const int py[4] = { p.iy0, p.iy0, p.iy1, p.iy1 };
        const int px[4] = { p.ix0, p.ix1, p.ix0, p.ix1 };
        const float wy0 = 1.0f - p.dy, wx0 = 1.0f - p.dx;
        const float w[4] = { wy0 * wx0, wy0 * p.dx, p.dy * wx0, p.dy * p.dx };
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const int64_t o = mv.offset(py[c], px[c]);
            if (o >= 0) {
                unsafeAtomicAdd(lmap + o, w[c] * t);
                unsafeAtomicAdd(lmap + o + mv.polstride, w[c] * q);
                unsafeAtomicAdd(lmap + o + 2 * mv.polstride, w[c] * u);
            }
            else if (!partial)
                err |= ERR_NOT_IN_PIXELIZATION;
        }
    }

    err = wave_or_u(err);
    if ((threadIdx.x & 63) == 0 && err)
        atomicOr(&s_err, err);
    __syncthreads();
    if (threadIdx.x == 0)
        errflags[blockIdx.x] = s_err;
}

int64_t samples_per_block(int64_t nsamp, int64_t max_blocks)
{
    // Same partition rule as map2tod_unplanned.cu:95-103 (at least 8 samples per thread),
    // with 256-thread blocks.
    int64_t spb = (nsamp + max_blocks - 1) / max_blocks;
    int64_t spt = (spb + UP_THREADS - 1) / UP_THREADS;
    spt = (spt > 8) ? spt : 8;
    return int64_t(UP_THREADS) * spt;
}

}  // namespace

int64_t unplanned_nblocks(int64_t nsamp, int64_t max_blocks)
{
    int64_t spb = samples_per_block(nsamp, max_blocks);
    return (nsamp + spb - 1) / spb;
}

void launch_unplanned_map2tod(float *tod, const float *lmap, const float *xp, const OpArgs &a,
                              uint32_t *errflags, int64_t max_blocks, bool accum, hipStream_t stream)
{
    const gmm_lpix &lp = *a.lpix;
    const int64_t spb = samples_per_block(a.nsamp, max_blocks);
    const int64_t nb = (a.nsamp + spb - 1) / spb;
    LocalMapView mv{lp.offsets_gpu, int(lp.nycells), int(lp.nxcells), lp.ystride, lp.polstride};
    auto k = lp.periodic_xcoord ? unplanned_map2tod_kernel<true> : unplanned_map2tod_kernel<false>;
    k<<<unsigned(nb), UP_THREADS, 0, stream>>>(tod, lmap, xp, mv, errflags, a.nsamp, spb,
                                               int(lp.nypix_global), int(lp.nxpix_global), a.partial, accum);
    GMM_HIP(hipGetLastError());
}

void launch_unplanned_tod2map(float *lmap, const float *tod, const float *xp, const OpArgs &a,
                              uint32_t *errflags, int64_t max_blocks, hipStream_t stream)
{
    const gmm_lpix &lp = *a.lpix;
    const int64_t spb = samples_per_block(a.nsamp, max_blocks);
    const int64_t nb = (a.nsamp + spb - 1) / spb;
    LocalMapView mv{lp.offsets_gpu, int(lp.nycells), int(lp.nxcells), lp.ystride, lp.polstride};
    auto k = lp.periodic_xcoord ? unplanned_tod2map_kernel<true> : unplanned_tod2map_kernel<false>;
    k<<<unsigned(nb), UP_THREADS, 0, stream>>>(lmap, tod, xp, mv, errflags, a.nsamp, spb,
                                               int(lp.nypix_global), int(lp.nxpix_global), a.partial);
    GMM_HIP(hipGetLastError());
}

}  // namespace gmm
