// Planned map2tod / tod2map on gfx950.
//
// One block per work unit ("chunk": a run of plan entries of one cell, cut at plan time). The
// block stages the cell's 3x64x64 T/Q/U values in LDS (map2tod) or accumulates them there
// (tod2map). Each wave takes a contiguous slice of the chunk's entries, loads 64 plan entries
// per vector load, and walks them two at a time (one 32-sample TOD line per half-wave), with
// the pointing/TOD loads of U entry pairs issued before any of them is used.
//
// map2tod keeps a zero-padded 66x66 tile per plane so that corners outside the cell read 0
// without a branch; lines touching several cells are zeroed first and then accumulated with
// float atomics (reference: src_lib/map2tod.cu:15-62, 81-191).
//
// tod2map accumulates in LDS with 64-bit integer atomics on fixed-point values: float LDS
// atomics (ds_add_f32) run an order of magnitude slower than integer ones on this GPU. A pre-pass
// finds max|tod|; with e = its binary exponent, every contribution w*t, w*t*cos2a, w*t*sin2a is
// scaled by S = 2^(30-e), truncated to int32 (|w*t*S| < 2^31) and added exactly into an int64
// tile; the flush converts each pixel once, float(sum)/S. The per-contribution quantum is
// max|t| * 2^-30 and the sums themselves are exact, so this is more accurate than float32
// accumulation. If the TOD holds Inf/NaN the kernel falls back to float LDS atomics.
// Reference: src_lib/tod2map.cu:33-154.

#include "gmm_internal.hpp"

namespace gmm {

namespace {

constexpr int WAVE = 64;
constexpr int TP = 66;              // padded tile side (map2tod)
constexpr int TPLANE = TP * TP;     // floats per padded plane

constexpr int M2T_THREADS = 512;
constexpr int M2T_U = 2;            // entry pairs in flight per wave
constexpr int T2M_THREADS = 1024;
constexpr int T2M_U = 2;

constexpr int ABSMAX_BLOCKS = 1024;
constexpr uint32_t FINITE_LIMIT = 0x7f800000u;   // bits of +Inf

__device__ inline int clamp_i(int v, int lo, int hi) { return min(max(v, lo), hi); }

__device__ inline uint32_t wave_or(uint32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v |= __shfl_xor(v, off, WAVE);
    return v;
}

__device__ inline uint32_t wave_max(uint32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v = max(v, uint32_t(__shfl_xor(v, off, WAVE)));
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

// Slice [wb, we) of the chunk's entries handled by wave 'wave' of 'nw'.
__device__ inline void wave_slice(const Chunk &ch, int wave, int nw, uint32_t &wb, uint32_t &we)
{
    const uint32_t n = ch.end - ch.begin;
    wb = ch.begin + uint32_t((uint64_t(n) * wave) / nw);
    we = ch.begin + uint32_t((uint64_t(n) * (wave + 1)) / nw);
}

// Plan entry packed for broadcast: line index in bits [0,26), mflag in bit 31.
__device__ inline uint32_t pack_entry(uint64_t mt)
{
    return (uint32_t(mt >> ICL_SHIFT) & ICL_MASK) | ((mt & MFLAG) ? 0x80000000u : 0u);
}

// Entry j of the 64 cached in 'packed' (j wave-uniform), for the half-wave 'half':
// lanes 0-31 get entry j, lanes 32-63 get entry j+1.
__device__ inline uint32_t entry_for_half(uint32_t packed, int j, int half)
{
    const uint32_t a = __builtin_amdgcn_readlane(packed, j);
    const uint32_t b = __builtin_amdgcn_readlane(packed, min(j + 1, 63));
    return half ? b : a;
}

// ------------------------------------------------------------------------------------------

template<bool Periodic>
__global__ void __launch_bounds__(M2T_THREADS)
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
        for (int i = threadIdx.x; i < 3 * CELL_NPIX; i += M2T_THREADS) {
            int p = i >> 12, r = (i >> 6) & 63, c = i & 63;
            tile[p * TPLANE + (r + 1) * TP + (c + 1)] = lmap[off + p * polstride + r * ystride + c];
        }
        // Zero border: rows 0 and 65 (66 entries each), columns 0 and 65 of rows 1..64.
        for (int i = threadIdx.x; i < 3 * 260; i += M2T_THREADS) {
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
    const float *tq = tile + TPLANE, *tu = tile + 2 * TPLANE;

    uint32_t wb, we;
    wave_slice(ch, wave, M2T_THREADS / WAVE, wb, we);

    for (uint32_t base = wb; base < we; base += 64) {
        const uint32_t cnt = min(64u, we - base);
        const uint32_t packed = (uint32_t(lane) < cnt) ? pack_entry(plan_mt[base + lane]) : 0u;

        for (uint32_t j = 0; j < cnt; j += 2 * M2T_U) {
            float y[M2T_U], x[M2T_U], al[M2T_U];
            int64_t s[M2T_U];
            bool act[M2T_U], mfl[M2T_U];
#pragma unroll
            for (int u = 0; u < M2T_U; u++) {
                const uint32_t jj = j + 2 * u;
                const uint32_t e = entry_for_half(packed, int(min(jj, 63u)), half);
                const uint32_t icl = e & ICL_MASK;
                act[u] = (jj + half < cnt) && (icl < ncl);
                mfl[u] = (e >> 31) != 0;
                s[u] = int64_t(act[u] ? icl : 0u) * LINE + sub;
                y[u] = xp[s[u]];
                x[u] = xp[s[u] + nsamp];
                al[u] = xp[s[u] + 2 * nsamp];
                if (jj + half < cnt && icl >= ncl)
                    err |= ERR_INCONSISTENT_NMT;
            }
#pragma unroll
            for (int u = 0; u < M2T_U; u++) {
                if (!act[u])
                    continue;
                if (off < 0) {
                    if (!accum && !mfl[u])
                        tod[s[u]] = 0.0f;
                    continue;
                }
                float sn, cs;
                sincosf(2.0f * al[u], &sn, &cs);
                Pixel<float> p;
                err |= locate_pixel(y[u], x[u], ny, nx, Periodic, p);

                const int r0 = clamp_i(p.iy0 - ybase + 1, 0, TP - 1) * TP;
                const int r1 = clamp_i(p.iy1 - ybase + 1, 0, TP - 1) * TP;
                const int c0 = clamp_i(p.ix0 - xbase + 1, 0, TP - 1);
                const int c1 = clamp_i(p.ix1 - xbase + 1, 0, TP - 1);

                const float v00 = tile[r0 + c0] + cs * tq[r0 + c0] + sn * tu[r0 + c0];
                const float v01 = tile[r0 + c1] + cs * tq[r0 + c1] + sn * tu[r0 + c1];
                const float v10 = tile[r1 + c0] + cs * tq[r1 + c0] + sn * tu[r1 + c0];
                const float v11 = tile[r1 + c1] + cs * tq[r1 + c1] + sn * tu[r1 + c1];

                const float wy0 = 1.0f - p.dy, wx0 = 1.0f - p.dx;
                const float val = (wy0 * wx0) * v00 + (wy0 * p.dx) * v01 + (p.dy * wx0) * v10 + (p.dy * p.dx) * v11;

                if (mfl[u])
                    unsafeAtomicAdd(tod + s[u], val);
                else if (accum)
                    tod[s[u]] += val;
                else
                    tod[s[u]] = val;
            }
        }
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

// Per-block max of |tod| as float bits (NaN > Inf > finite in this order), ABSMAX_BLOCKS blocks.
template<bool Vec>
__global__ void __launch_bounds__(256)
tod_absmax_kernel(const float *__restrict__ tod, int64_t n, uint32_t *__restrict__ partial)
{
    __shared__ uint32_t s_w[4];
    uint32_t m = 0;
    if (Vec) {
        const float4 *t4 = reinterpret_cast<const float4 *>(tod);
        for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < (n >> 2); i += int64_t(gridDim.x) * 256) {
            const float4 v = t4[i];
            m = max(m, __float_as_uint(v.x) & 0x7fffffffu);
            m = max(m, __float_as_uint(v.y) & 0x7fffffffu);
            m = max(m, __float_as_uint(v.z) & 0x7fffffffu);
            m = max(m, __float_as_uint(v.w) & 0x7fffffffu);
        }
    }
    else {
        for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
            m = max(m, __float_as_uint(tod[i]) & 0x7fffffffu);
    }
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0)
        s_w[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0)
        partial[blockIdx.x] = max(max(s_w[0], s_w[1]), max(s_w[2], s_w[3]));
}

template<bool Periodic>
__global__ void __launch_bounds__(T2M_THREADS)
planned_tod2map_kernel(float *__restrict__ lmap, const float *__restrict__ tod, const float *__restrict__ xp,
                       const int64_t *__restrict__ cell_offsets, const uint64_t *__restrict__ plan_mt,
                       const Chunk *__restrict__ chunks, const uint32_t *__restrict__ absmax_partial,
                       uint32_t *__restrict__ err_word, int64_t nsamp, int ny, int nx, int nycells, int nxcells,
                       int64_t ystride, int64_t polstride, bool partial)
{
    __shared__ unsigned long long acc[3 * CELL_NPIX];    // 96 KiB
    __shared__ uint32_t s_red[T2M_THREADS / WAVE];
    float *facc = reinterpret_cast<float *>(acc);         // float tile for the non-finite fallback

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
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

    // Scale from max|tod| (ABSMAX_BLOCKS == T2M_THREADS partials, one per thread).
    uint32_t m = wave_max(absmax_partial[threadIdx.x]);
    if (lane == 0)
        s_red[wave] = m;
    for (int i = threadIdx.x; i < 3 * CELL_NPIX; i += T2M_THREADS)
        acc[i] = 0ull;
    __syncthreads();
    m = 0;
#pragma unroll
    for (int w = 0; w < T2M_THREADS / WAVE; w++)
        m = max(m, s_red[w]);
    const bool finite = (m < FINITE_LIMIT);
    const int e = int(m >> 23) - 127;                    // binary exponent of max|t|
    const int sexp = clamp_i(30 - e, -97, 126);
    const float scale = __uint_as_float(uint32_t(sexp + 127) << 23);
    const float inv_scale = __uint_as_float(uint32_t(127 - sexp) << 23);

    const int half = lane >> 5, sub = lane & 31;
    const int ybase = int(ch.icell >> 10) << 6, xbase = int(ch.icell & 1023u) << 6;

    uint32_t wb, we;
    wave_slice(ch, wave, T2M_THREADS / WAVE, wb, we);

    for (uint32_t base = wb; base < we; base += 64) {
        const uint32_t cnt = min(64u, we - base);
        const uint32_t packed = (uint32_t(lane) < cnt) ? pack_entry(plan_mt[base + lane]) : 0u;

        for (uint32_t j = 0; j < cnt; j += 2 * T2M_U) {
            float y[T2M_U], x[T2M_U], al[T2M_U], tv[T2M_U];
            bool act[T2M_U];
#pragma unroll
            for (int u = 0; u < T2M_U; u++) {
                const uint32_t jj = j + 2 * u;
                const uint32_t e2 = entry_for_half(packed, int(min(jj, 63u)), half);
                const uint32_t icl = e2 & ICL_MASK;
                act[u] = (jj + half < cnt) && (icl < ncl);
                if (jj + half < cnt && icl >= ncl)
                    err |= ERR_INCONSISTENT_NMT;
                const int64_t s = int64_t(act[u] ? icl : 0u) * LINE + sub;
                y[u] = xp[s];
                x[u] = xp[s + nsamp];
                al[u] = xp[s + 2 * nsamp];
                tv[u] = tod[s];
            }
#pragma unroll
            for (int u = 0; u < T2M_U; u++) {
                if (!act[u])
                    continue;
                float sn, cs;
                sincosf(2.0f * al[u], &sn, &cs);
                Pixel<float> p;
                err |= locate_pixel(y[u], x[u], ny, nx, Periodic, p);

                const int ry[2] = { p.iy0 - ybase, p.iy1 - ybase };
                const int rx[2] = { p.ix0 - xbase, p.ix1 - xbase };
                const float wy[2] = { 1.0f - p.dy, p.dy };
                const float wx[2] = { 1.0f - p.dx, p.dx };

                if (finite) {
                    const float ts = tv[u] * scale, qs = ts * cs, us = ts * sn;
#pragma unroll
                    for (int a = 0; a < 2; a++) {
#pragma unroll
                        for (int b = 0; b < 2; b++) {
                            if ((unsigned(ry[a]) < 64u) && (unsigned(rx[b]) < 64u)) {
                                const float w = wy[a] * wx[b];
                                const int k = (ry[a] << 6) | rx[b];
                                atomicAdd(acc + k, (unsigned long long)(long long)int(w * ts));
                                atomicAdd(acc + CELL_NPIX + k, (unsigned long long)(long long)int(w * qs));
                                atomicAdd(acc + 2 * CELL_NPIX + k, (unsigned long long)(long long)int(w * us));
                            }
                        }
                    }
                }
                else {
                    const float t = tv[u], q = t * cs, uu = t * sn;
#pragma unroll
                    for (int a = 0; a < 2; a++) {
#pragma unroll
                        for (int b = 0; b < 2; b++) {
                            if ((unsigned(ry[a]) < 64u) && (unsigned(rx[b]) < 64u)) {
                                const float w = wy[a] * wx[b];
                                const int k = (ry[a] << 6) | rx[b];
                                atomicAdd(facc + k, w * t);
                                atomicAdd(facc + CELL_NPIX + k, w * q);
                                atomicAdd(facc + 2 * CELL_NPIX + k, w * uu);
                            }
                        }
                    }
                }
            }
        }
    }
    __syncthreads();

    const bool shared = (ch.flags & CHUNK_SHARED) != 0;
    for (int i = threadIdx.x; i < 3 * CELL_NPIX; i += T2M_THREADS) {
        const int pl = i >> 12, r = (i >> 6) & 63, c = i & 63;
        float *g = lmap + off + pl * polstride + r * ystride + c;
        const float v = finite ? float((long long)acc[i]) * inv_scale : facc[i];
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

    dim3 grid(unsigned(plan.nchunks)), block(M2T_THREADS);
    auto k = lp.periodic_xcoord ? planned_map2tod_kernel<true> : planned_map2tod_kernel<false>;
    k<<<grid, block, 0, stream>>>(tod, lmap, xp, lp.offsets_gpu, plan.plan_mt, plan.chunks, err_word,
                                  a.nsamp, int(lp.nypix_global), int(lp.nxpix_global), int(lp.nycells),
                                  int(lp.nxcells), lp.ystride, lp.polstride, a.partial, accum);
    GMM_HIP(hipGetLastError());
}

int64_t tod2map_scratch_words() { return ABSMAX_BLOCKS; }

void launch_planned_tod2map(float *lmap, const float *tod, const float *xp, const OpArgs &a,
                            const gmm_plan &plan, hipStream_t stream)
{
    const gmm_lpix &lp = *a.lpix;
    uint32_t *err_word = plan.err + (plan.pp.n_err - 1);
    if (plan.nchunks == 0)
        return;

    static_assert(ABSMAX_BLOCKS == T2M_THREADS, "one max|t| partial per thread of the tod2map block");
    if ((reinterpret_cast<uintptr_t>(tod) & 15) == 0)
        tod_absmax_kernel<true><<<ABSMAX_BLOCKS, 256, 0, stream>>>(tod, a.nsamp, plan.absmax);
    else
        tod_absmax_kernel<false><<<ABSMAX_BLOCKS, 256, 0, stream>>>(tod, a.nsamp, plan.absmax);
    GMM_HIP(hipGetLastError());

    dim3 grid(unsigned(plan.nchunks)), block(T2M_THREADS);
    auto k = lp.periodic_xcoord ? planned_tod2map_kernel<true> : planned_tod2map_kernel<false>;
    k<<<grid, block, 0, stream>>>(lmap, tod, xp, lp.offsets_gpu, plan.plan_mt, plan.chunks, plan.absmax, err_word,
                                  a.nsamp, int(lp.nypix_global), int(lp.nxpix_global), int(lp.nycells),
                                  int(lp.nxcells), lp.ystride, lp.polstride, a.partial);
    GMM_HIP(hipGetLastError());
}

}  // namespace gmm
