// Planned map2tod / tod2map on gfx950.
//
// One block per work unit ("chunk": a run of plan entries of one cell, cut at plan time). The
// block stages the cell's 3x64x64 T/Q/U values in LDS, then its waves walk the chunk's entries
// two at a time (one 32-sample TOD line per half-wave): pointing and TOD stream in with
// coalesced 128-byte loads, the bilinear gather (map2tod) or scatter-add (tod2map) hits LDS.
//
// map2tod keeps a zero-padded 66x66 tile per plane so that corners outside the cell read 0
// without a branch; lines touching several cells are zeroed first and then accumulated with
// float atomics (reference: src_lib/map2tod.cu:15-62, 81-191). tod2map accumulates into an
// unpadded 64x64x3 LDS tile with LDS float atomics and flushes once per chunk: a plain
// read-modify-write when the chunk owns its cell, float atomics when the cell is shared
// (reference: src_lib/tod2map.cu:33-154).

#include "gmm_internal.hpp"

namespace gmm {

namespace {

constexpr int WAVE = 64;
constexpr int TP = 66;              // padded tile side (map2tod)
constexpr int TPLANE = TP * TP;     // floats per padded plane
constexpr int OP_THREADS = 512;
constexpr int OP_NW = OP_THREADS / WAVE;

__device__ inline int clamp_i(int v, int lo, int hi) { return min(max(v, lo), hi); }

__device__ inline uint32_t wave_or(uint32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v |= __shfl_xor(v, off, WAVE);
    return v;
}

__device__ inline void report_err(uint32_t err, uint32_t *err_word)
{
    err = wave_or(err);
    if ((threadIdx.x & 63) == 0 && err)
        atomicOr(err_word, err);
}

__device__ inline int64_t chunk_cell_offset(const Chunk &ch, const int64_t *__restrict__ cell_offsets,
                                            int nycells, int nxcells)
{
    const int cy = int(ch.icell >> 10), cx = int(ch.icell & 1023u);
    return (cy < nycells && cx < nxcells) ? cell_offsets[int64_t(cy) * nxcells + cx] : int64_t(-1);
}

// ------------------------------------------------------------------------------------------

template<bool Periodic>
__global__ void __launch_bounds__(OP_THREADS)
planned_map2tod_kernel(float *__restrict__ tod, const float *__restrict__ lmap, const float *__restrict__ xp,
                       const int64_t *__restrict__ cell_offsets, const uint64_t *__restrict__ plan_mt,
                       const Chunk *__restrict__ chunks, uint32_t *__restrict__ err_word,
                       int64_t nsamp, int ny, int nx, int nycells, int nxcells,
                       int64_t ystride, int64_t polstride, bool partial, bool accum)
{
    __shared__ float tile[3 * TPLANE];

    const Chunk ch = chunks[blockIdx.x];
    const int64_t off = chunk_cell_offset(ch, cell_offsets, nycells, nxcells);
    const uint32_t ncl = uint32_t(nsamp >> 5);
    uint32_t err = (off < 0 && !partial) ? ERR_NOT_IN_PIXELIZATION : 0u;

    if (off >= 0) {
        // Interior: 3 planes x 64 rows x 64 columns, one row of 64 floats per wave load.
        for (int i = threadIdx.x; i < 3 * CELL_NPIX; i += OP_THREADS) {
            int p = i >> 12, r = (i >> 6) & 63, c = i & 63;
            tile[p * TPLANE + (r + 1) * TP + (c + 1)] = lmap[off + p * polstride + r * ystride + c];
        }
        // Zero border: rows 0 and 65 (66 entries each), columns 0 and 65 of rows 1..64.
        for (int i = threadIdx.x; i < 3 * 260; i += OP_THREADS) {
            int p = i / 260, j = i % 260, idx;
            if (j < 66)
                idx = j;
            else if (j < 132)
                idx = 65 * TP + (j - 66);
            else if (j < 196)
                idx = (j - 132 + 1) * TP;
            else
                idx = (j - 196 + 1) * TP + 65;
            tile[p * TPLANE + idx] = 0.0f;
        }
    }
    __syncthreads();

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int half = lane >> 5, sub = lane & 31;
    const int ybase = int(ch.icell >> 10) << 6, xbase = int(ch.icell & 1023u) << 6;

    for (uint32_t e = ch.begin + 2 * wave; e < ch.end; e += 2 * OP_NW) {
        const uint32_t my = e + half;
        if (my >= ch.end)
            continue;
        const uint64_t mt = plan_mt[my];
        const uint32_t icl = uint32_t(mt >> ICL_SHIFT) & ICL_MASK;
        const bool mflag = (mt & MFLAG) != 0;
        if (icl >= ncl) {
            err |= ERR_INCONSISTENT_NMT;
            continue;
        }
        const int64_t s = int64_t(icl) * LINE + sub;

        if (off < 0) {
            if (!accum && !mflag)
                tod[s] = 0.0f;
            continue;
        }

        const float y = xp[s], x = xp[s + nsamp], alpha = xp[s + 2 * nsamp];
        float sn, cs;
        sincosf(2.0f * alpha, &sn, &cs);

        Pixel<float> p;
        err |= locate_pixel(y, x, ny, nx, Periodic, p);

        const int r0 = clamp_i(p.iy0 - ybase + 1, 0, TP - 1) * TP;
        const int r1 = clamp_i(p.iy1 - ybase + 1, 0, TP - 1) * TP;
        const int c0 = clamp_i(p.ix0 - xbase + 1, 0, TP - 1);
        const int c1 = clamp_i(p.ix1 - xbase + 1, 0, TP - 1);

        const float *tq = tile + TPLANE, *tu = tile + 2 * TPLANE;
        const float v00 = tile[r0 + c0] + cs * tq[r0 + c0] + sn * tu[r0 + c0];
        const float v01 = tile[r0 + c1] + cs * tq[r0 + c1] + sn * tu[r0 + c1];
        const float v10 = tile[r1 + c0] + cs * tq[r1 + c0] + sn * tu[r1 + c0];
        const float v11 = tile[r1 + c1] + cs * tq[r1 + c1] + sn * tu[r1 + c1];

        const float wy0 = 1.0f - p.dy, wx0 = 1.0f - p.dx;
        const float val = (wy0 * wx0) * v00 + (wy0 * p.dx) * v01 + (p.dy * wx0) * v10 + (p.dy * p.dx) * v11;

        if (mflag)
            unsafeAtomicAdd(tod + s, val);
        else if (accum)
            tod[s] += val;
        else
            tod[s] = val;
    }

    report_err(err, err_word);
}

// Zeroes the TOD lines that touch more than one cell (they are accumulated atomically).
__global__ void zero_lines_kernel(float *__restrict__ tod, const uint32_t *__restrict__ mlines, int64_t nml)
{
    const int64_t n = nml * LINE;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        tod[int64_t(mlines[i >> 5]) * LINE + (i & 31)] = 0.0f;
}

// ------------------------------------------------------------------------------------------

template<bool Periodic>
__global__ void __launch_bounds__(OP_THREADS)
planned_tod2map_kernel(float *__restrict__ lmap, const float *__restrict__ tod, const float *__restrict__ xp,
                       const int64_t *__restrict__ cell_offsets, const uint64_t *__restrict__ plan_mt,
                       const Chunk *__restrict__ chunks, uint32_t *__restrict__ err_word,
                       int64_t nsamp, int ny, int nx, int nycells, int nxcells,
                       int64_t ystride, int64_t polstride, bool partial)
{
    __shared__ float tile[3 * CELL_NPIX];

    const Chunk ch = chunks[blockIdx.x];
    const int64_t off = chunk_cell_offset(ch, cell_offsets, nycells, nxcells);
    const uint32_t ncl = uint32_t(nsamp >> 5);
    uint32_t err = (off < 0 && !partial) ? ERR_NOT_IN_PIXELIZATION : 0u;

    if (off < 0) {
        // Samples whose cells are not in the local map are dropped (tod2map_reference.cu:14-46
        // skips them too); only the error bit is reported.
        report_err(err, err_word);
        return;
    }

    for (int i = threadIdx.x; i < 3 * CELL_NPIX; i += OP_THREADS)
        tile[i] = 0.0f;
    __syncthreads();

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int half = lane >> 5, sub = lane & 31;
    const int ybase = int(ch.icell >> 10) << 6, xbase = int(ch.icell & 1023u) << 6;

    for (uint32_t e = ch.begin + 2 * wave; e < ch.end; e += 2 * OP_NW) {
        const uint32_t my = e + half;
        if (my >= ch.end)
            continue;
        const uint64_t mt = plan_mt[my];
        const uint32_t icl = uint32_t(mt >> ICL_SHIFT) & ICL_MASK;
        if (icl >= ncl) {
            err |= ERR_INCONSISTENT_NMT;
            continue;
        }
        const int64_t s = int64_t(icl) * LINE + sub;

        const float y = xp[s], x = xp[s + nsamp], alpha = xp[s + 2 * nsamp];
        const float t = tod[s];
        float sn, cs;
        sincosf(2.0f * alpha, &sn, &cs);
        const float q = t * cs, u = t * sn;

        Pixel<float> p;
        err |= locate_pixel(y, x, ny, nx, Periodic, p);

        const int ry[2] = { p.iy0 - ybase, p.iy1 - ybase };
        const int rx[2] = { p.ix0 - xbase, p.ix1 - xbase };
        const float wy[2] = { 1.0f - p.dy, p.dy };
        const float wx[2] = { 1.0f - p.dx, p.dx };
#pragma unroll
        for (int a = 0; a < 2; a++) {
#pragma unroll
            for (int b = 0; b < 2; b++) {
                if ((unsigned(ry[a]) < 64u) && (unsigned(rx[b]) < 64u)) {
                    const float w = wy[a] * wx[b];
                    const int k = (ry[a] << 6) | rx[b];
                    atomicAdd(tile + k, w * t);
                    atomicAdd(tile + CELL_NPIX + k, w * q);
                    atomicAdd(tile + 2 * CELL_NPIX + k, w * u);
                }
            }
        }
    }
    __syncthreads();

    const bool shared = (ch.flags & CHUNK_SHARED) != 0;
    for (int i = threadIdx.x; i < 3 * CELL_NPIX; i += OP_THREADS) {
        const int pl = i >> 12, r = (i >> 6) & 63, c = i & 63;
        float *g = lmap + off + pl * polstride + r * ystride + c;
        const float v = tile[i];
        if (shared)
            unsafeAtomicAdd(g, v);
        else
            *g += v;
    }

    report_err(err, err_word);
}

}  // namespace

// ==========================================================================================

void launch_planned_map2tod(float *tod, const float *lmap, const float *xp, const OpArgs &a,
                            const gmm_plan &plan, bool accum, hipStream_t stream)
{
    const gmm_lpix &lp = *a.lpix;
    uint32_t *err_word = plan.err + (plan.pp.n_err - 1);

    if (!accum && plan.nmlines > 0) {
        int64_t n = plan.nmlines * LINE;
        int grid = int(std::min<int64_t>((n + 255) / 256, 8192));
        zero_lines_kernel<<<grid, 256, 0, stream>>>(tod, plan.mlines, plan.nmlines);
        GMM_HIP(hipGetLastError());
    }
    if (plan.nchunks == 0)
        return;

    dim3 grid(unsigned(plan.nchunks)), block(OP_THREADS);
    auto k = lp.periodic_xcoord ? planned_map2tod_kernel<true> : planned_map2tod_kernel<false>;
    k<<<grid, block, 0, stream>>>(tod, lmap, xp, lp.offsets_gpu, plan.plan_mt, plan.chunks, err_word,
                                  a.nsamp, int(lp.nypix_global), int(lp.nxpix_global), int(lp.nycells),
                                  int(lp.nxcells), lp.ystride, lp.polstride, a.partial, accum);
    GMM_HIP(hipGetLastError());
}

void launch_planned_tod2map(float *lmap, const float *tod, const float *xp, const OpArgs &a,
                            const gmm_plan &plan, hipStream_t stream)
{
    const gmm_lpix &lp = *a.lpix;
    uint32_t *err_word = plan.err + (plan.pp.n_err - 1);
    if (plan.nchunks == 0)
        return;

    dim3 grid(unsigned(plan.nchunks)), block(OP_THREADS);
    auto k = lp.periodic_xcoord ? planned_tod2map_kernel<true> : planned_tod2map_kernel<false>;
    k<<<grid, block, 0, stream>>>(lmap, tod, xp, lp.offsets_gpu, plan.plan_mt, plan.chunks, err_word,
                                  a.nsamp, int(lp.nypix_global), int(lp.nxpix_global), int(lp.nycells),
                                  int(lp.nxcells), lp.ystride, lp.polstride, a.partial);
    GMM_HIP(hipGetLastError());
}

}  // namespace gmm
