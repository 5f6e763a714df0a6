// Plan construction on gfx950: preplan count, deterministic emit, device scan, stable LSD
// radix sort in LDS, and the operator work list (chunks + multi-cell line list).
//
// All kernels assume 64-lane waves. A 32-sample TOD line maps onto one half-wave, so each wave
// handles two lines at a time; the set of cells touched by a line is found by repeated
// ballot/readlane extraction instead of warp32 match-any.

#include "gmm_internal.hpp"
#include "gmm_locate.hpp"

namespace gmm {

namespace {

constexpr int WAVE = 64;

__device__ inline int lane_id() { return threadIdx.x & 63; }

__device__ inline uint64_t lanemask_lt() { return (1ull << lane_id()) - 1ull; }

// Inclusive scan across a wave.
__device__ inline uint32_t wave_inclusive_scan(uint32_t v)
{
    const int lane = lane_id();
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        uint32_t t = __shfl_up(v, off, WAVE);
        v += (lane >= off) ? t : 0u;
    }
    return v;
}

__device__ inline uint32_t wave_reduce_or(uint32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v |= __shfl_xor(v, off, WAVE);
    return v;
}

__device__ inline uint32_t wave_reduce_add(uint32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, WAVE);
    return v;
}

// Exclusive scan over a block of NT threads (NT/64 waves), returns the block total too.
template<int NT>
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t *s_wave, uint32_t &total)
{
    constexpr int NW = NT / WAVE;
    const int lane = lane_id(), wave = threadIdx.x / WAVE;
    uint32_t inc = wave_inclusive_scan(v);
    if (lane == WAVE - 1)
        s_wave[wave] = inc;
    __syncthreads();
    uint32_t before = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NW; w++) {
        uint32_t t = s_wave[w];
        before += (w < wave) ? t : 0u;
        tot += t;
    }
    __syncthreads();
    total = tot;
    return before + inc - v;
}

// One extraction round over the remaining candidate cells of the two half-waves.
// 'k' holds up to four distinct candidate keys per lane (NO_CELL = none). Returns false when
// both halves are exhausted; otherwise removes one cell per non-empty half from every lane of
// that half and returns it in 'c' (NO_CELL for an exhausted half).
__device__ inline bool extract_one(uint32_t k[4], uint32_t &c, bool &lo_hit, bool &hi_hit)
{
    uint32_t first = (k[0] != NO_CELL) ? k[0] : ((k[1] != NO_CELL) ? k[1] : ((k[2] != NO_CELL) ? k[2] : k[3]));
    uint64_t m = __ballot(first != NO_CELL);
    if (m == 0)
        return false;
    uint32_t mlo = uint32_t(m), mhi = uint32_t(m >> 32);
    uint32_t clo = NO_CELL, chi = NO_CELL;
    if (mlo)
        clo = __builtin_amdgcn_readlane(first, __builtin_ctz(mlo));
    if (mhi)
        chi = __builtin_amdgcn_readlane(first, 32 + __builtin_ctz(mhi));
    lo_hit = (mlo != 0);
    hi_hit = (mhi != 0);
    c = (lane_id() < 32) ? clo : chi;
#pragma unroll
    for (int j = 0; j < 4; j++)
        k[j] = (k[j] == c) ? NO_CELL : k[j];
    return true;
}

// Fast path of the extraction (the usual case, a line inside one cell): true (wave-uniform) when
// in both halves every sample's footprint is the one cell of the half's first sample, k[0] of
// lane 0 or 32. Then that cell, in 'c', is the half's only cell, the one extract_one() would
// return first, and nlo / nhi count the halves whose line exists (first sample active).
__device__ inline bool single_cell_halves(const uint32_t k[4], uint32_t &c, uint32_t &nlo, uint32_t &nhi)
{
    const uint32_t lo = __builtin_amdgcn_readlane(k[0], 0), hi = __builtin_amdgcn_readlane(k[0], 32);
    c = (lane_id() < 32) ? lo : hi;
    nlo = (lo != NO_CELL);
    nhi = (hi != NO_CELL);
    // Inactive lanes (past the end of a partial line, or of a line that does not exist) hold
    // NO_CELL in every slot.
    const bool one = ((k[1] & k[2] & k[3]) == NO_CELL) && (k[0] == c || k[0] == NO_CELL);
    return __ballot(!one) == 0;
}

// The four corner indices of a sample's bilinear footprint (locate_pixel's,
// include/gpu_mm_internals.hpp:124-164), located by locate_op with its range checks deferred to
// 'chk'.
template<bool Periodic>
__device__ inline void sample_corners(float y, float x, const Geom &g, Checks &chk, Pixel<float> &p)
{
    float dy, dx;
    locate_op<Periodic>(y, x, g, chk, p.iy0, p.ix0, dy, dx);
    p.iy1 = p.iy0 + 1;
    p.ix1 = Periodic ? ((p.ix0 + 1 < g.nx) ? p.ix0 + 1 : 0) : p.ix0 + 1;
}

// Footprint cells of one sample. An inactive lane holds NO_CELL everywhere; its y = x = 0 lie in
// range and leave the checks unchanged.
template<bool Periodic>
__device__ inline void sample_cells(float y, float x, const Geom &g, Checks &chk, bool active, uint32_t k[4])
{
    Pixel<float> p;
    sample_corners<Periodic>(y, x, g, chk, p);
    footprint_cells(p, k);
    if (!active)
        k[0] = k[1] = k[2] = k[3] = NO_CELL;
}

// ------------------------------------------------------------------------------------------
// Preplan: number of distinct (line, cell) pairs per planner block
// (src_lib/PointingPrePlan.cu:37-108 computes the same per-block counts).

constexpr int PP_THREADS = 512;
constexpr int PP_UNROLL = 4;

template<bool Periodic>
__global__ void __launch_bounds__(PP_THREADS)
preplan_count_kernel(const float *__restrict__ xp, int64_t nsamp, int64_t ncl, int64_t ncl_per_block,
                     int ny, int nx, uint32_t *__restrict__ nmt_out, uint32_t *__restrict__ err_out)
{
    __shared__ uint32_t s_cnt[PP_THREADS / WAVE], s_err[PP_THREADS / WAVE];
    const int lane = lane_id(), wave = threadIdx.x / WAVE;
    constexpr int NW = PP_THREADS / WAVE;
    const int half = lane >> 5, sub = lane & 31;

    const int64_t l0 = int64_t(blockIdx.x) * ncl_per_block;
    const int64_t l1 = min(ncl, l0 + ncl_per_block);
    uint32_t cnt = 0;   // wave-uniform
    const Geom g = make_geom(ny, nx, Periodic);
    Checks chk;

    for (int64_t lb = l0 + 2 * wave; lb < l1; lb += 2 * NW * PP_UNROLL) {
        float y[PP_UNROLL], x[PP_UNROLL];
        bool act[PP_UNROLL];
#pragma unroll
        for (int u = 0; u < PP_UNROLL; u++) {
            int64_t line = lb + 2 * NW * u + half;
            int64_t s = line * LINE + sub;
            act[u] = (line < l1) && (s < nsamp);        // the last line may be partial
            y[u] = act[u] ? xp[s] : 0.0f;
            x[u] = act[u] ? xp[s + nsamp] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < PP_UNROLL; u++) {
            uint32_t k[4];
            sample_cells<Periodic>(y[u], x[u], g, chk, act[u], k);
            uint32_t c, nlo, nhi;
            if (single_cell_halves(k, c, nlo, nhi)) {
                cnt += nlo + nhi;
                continue;
            }
            bool lo, hi;
            while (extract_one(k, c, lo, hi))
                cnt += uint32_t(lo) + uint32_t(hi);
        }
    }

    uint32_t err = wave_reduce_or(chk.bits(g));
    if (lane == 0) {
        s_cnt[wave] = cnt;
        s_err[wave] = err;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t c = 0, e = 0;
        for (int w = 0; w < NW; w++) {
            c += s_cnt[w];
            e |= s_err[w];
        }
        nmt_out[blockIdx.x] = c;
        err_out[blockIdx.x] = e;
    }
}

// Inclusive scan of the per-block counts (padded to PREPLAN_SIZE with the total) and OR of the
// error words into err[0]; the 64-bit total goes to err[1] (low) and err[2] (high).
// One block of PREPLAN_SIZE threads. Replaces the host round trip of PointingPrePlan.cu:195-216.
__global__ void __launch_bounds__(PREPLAN_SIZE)
preplan_finish_kernel(uint32_t *nmt, uint32_t *err, int nblocks)
{
    __shared__ uint32_t s_wave[PREPLAN_SIZE / WAVE];
    __shared__ unsigned long long s_tot64;
    __shared__ uint32_t s_err;
    const int t = threadIdx.x;
    uint32_t v = (t < nblocks) ? nmt[t] : 0u;
    uint32_t e = (t < nblocks) ? err[t] : 0u;
    if (t == 0) {
        s_tot64 = 0;
        s_err = 0;
    }
    __syncthreads();
    atomicAdd(&s_tot64, (unsigned long long)v);
    atomicOr(&s_err, e);
    uint32_t total;
    uint32_t excl = block_exclusive_scan<PREPLAN_SIZE>(v, s_wave, total);
    nmt[t] = (t < nblocks) ? (excl + v) : total;
    __syncthreads();
    if (t == 0) {
        err[0] = s_err;
        err[1] = uint32_t(s_tot64);
        err[2] = uint32_t(s_tot64 >> 32);
    }
}

// ------------------------------------------------------------------------------------------
// Emit: writes the plan entries of each planner block into its window
// [nmt_cumsum[b-1], nmt_cumsum[b]) in line order; inside a line, cells come in extraction order
// (lowest lane first), which makes the output deterministic. The entry layout and the
// consistency check against the preplan count follow src_lib/PointingPlan.cu:75-89, 116, 227.

constexpr int EM_THREADS = 512;
constexpr int EM_NW = EM_THREADS / WAVE;
constexpr int EM_U = 2;                       // line pairs per wave per step
constexpr int EM_LINES = 2 * EM_NW * EM_U;    // lines per block step (32)

// At most 64 VGPRs (8 waves per SIMD, 4 blocks per CU): the planner grid (1024 blocks at most)
// is then resident at once on 256 CUs instead of leaving a tail.
template<bool Periodic>
__global__ void __launch_bounds__(EM_THREADS) __attribute__((amdgpu_waves_per_eu(8)))
plan_emit_kernel(const float *__restrict__ xp, int64_t nsamp, int64_t ncl, int64_t ncl_per_block,
                 int ny, int nx, const uint32_t *__restrict__ nmt_cumsum,
                 uint64_t *__restrict__ out, uint32_t *__restrict__ err_out)
{
    __shared__ uint32_t s_counts[2][EM_LINES];
    __shared__ uint32_t s_err[EM_NW];
    const int lane = lane_id(), wave = threadIdx.x / WAVE;
    const int half = lane >> 5, sub = lane & 31;
    const int b = blockIdx.x;

    const int64_t l0 = int64_t(b) * ncl_per_block;
    const int64_t l1 = min(ncl, l0 + ncl_per_block);
    const uint32_t base = (b > 0) ? nmt_cumsum[b - 1] : 0u;
    const uint32_t limit = nmt_cumsum[b] - base;
    out += base;

    uint32_t err = 0;
    uint32_t running = 0;   // entries written by earlier steps (block-uniform)
    const Geom g = make_geom(ny, nx, Periodic);
    Checks chk;
    int parity = 0;

    for (int64_t ls = l0; ls < l1; ls += EM_LINES) {
        float y[EM_U], x[EM_U];
        bool act[EM_U];
        int64_t line[EM_U];
#pragma unroll
        for (int u = 0; u < EM_U; u++) {
            line[u] = ls + 2 * (wave * EM_U + u) + half;
            int64_t s = line[u] * LINE + sub;
            act[u] = (line[u] < l1) && (s < nsamp);     // the last line may be partial
            y[u] = act[u] ? xp[s] : 0.0f;
            x[u] = act[u] ? xp[s + nsamp] : 0.0f;
        }

        uint32_t mycell[EM_U], nmy[EM_U];
#pragma unroll
        for (int u = 0; u < EM_U; u++) {
            uint32_t k[4];
            sample_cells<Periodic>(y[u], x[u], g, chk, act[u], k);
            uint32_t c, nlo, nhi;
            mycell[u] = NO_CELL;
            if (single_cell_halves(k, c, nlo, nhi)) {
                mycell[u] = (sub == 0) ? c : NO_CELL;
            }
            else {
                nlo = nhi = 0;
                bool lo, hi;
                while (extract_one(k, c, lo, hi)) {
                    uint32_t j = half ? nhi : nlo;
                    mycell[u] = (uint32_t(sub) == j) ? c : mycell[u];
                    nlo += lo;
                    nhi += hi;
                }
            }
            nmy[u] = half ? nhi : nlo;
            if (sub == 0)
                s_counts[parity][2 * (wave * EM_U + u) + half] = nmy[u];
        }
        __syncthreads();

        // Exclusive offsets of this step's lines, in line order.
        uint32_t cv = (lane < EM_LINES) ? s_counts[parity][lane] : 0u;
        uint32_t inc = wave_inclusive_scan(cv);
        uint32_t step_total = __shfl(inc, EM_LINES - 1, WAVE);
        uint32_t excl = inc - cv;

#pragma unroll
        for (int u = 0; u < EM_U; u++) {
            int idx = 2 * (wave * EM_U + u) + half;
            uint32_t off = running + __shfl(excl, idx, WAVE);
            uint32_t n = nmy[u];
            bool mflag = (n > 1);
            if (act[u] && uint32_t(sub) < n) {
                uint32_t pos = off + sub;
                uint64_t mt = uint64_t(mycell[u]) | (uint64_t(line[u]) << ICL_SHIFT)
                            | (mflag ? MFLAG : 0ull) | ((mflag && sub == 0) ? ZFLAG : 0ull);
                if (pos < limit)
                    out[pos] = mt;
                else
                    err |= ERR_INCONSISTENT_NMT;
            }
            // Lines touching more than 32 cells (not reachable by ordinary scans): the extra
            // cells are re-extracted and written one at a time by the half's first lane.
            uint64_t big = __ballot(n > 32);
            if (big) {
                uint32_t k[4];
                sample_cells<Periodic>(y[u], x[u], g, chk, act[u], k);
                uint32_t c, nlo = 0, nhi = 0;
                bool lo, hi;
                while (extract_one(k, c, lo, hi)) {
                    uint32_t j = half ? nhi : nlo;
                    if (sub == 0 && j >= 32 && c != NO_CELL) {
                        uint32_t pos = off + j;
                        uint64_t mt = uint64_t(c) | (uint64_t(line[u]) << ICL_SHIFT) | MFLAG;
                        if (pos < limit)
                            out[pos] = mt;
                        else
                            err |= ERR_INCONSISTENT_NMT;
                    }
                    nlo += lo;
                    nhi += hi;
                }
            }
        }

        running += step_total;
        parity ^= 1;
    }

    if (running != limit)
        err |= ERR_INCONSISTENT_NMT;

    err = wave_reduce_or(err | chk.bits(g));
    if (lane == 0)
        s_err[wave] = err;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t e = 0;
        for (int w = 0; w < EM_NW; w++)
            e |= s_err[w];
        err_out[b] = e;
    }
}

// ------------------------------------------------------------------------------------------
// Device-wide exclusive scan of uint32 (reduce-then-scan, recursive over tile sums).

constexpr int SC_THREADS = 256;
constexpr int SC_ITEMS = 16;
constexpr int SC_TILE = SC_THREADS * SC_ITEMS;

__global__ void __launch_bounds__(SC_THREADS)
scan_tile_sums_kernel(const uint32_t *__restrict__ in, int64_t n, uint32_t *__restrict__ sums)
{
    __shared__ uint32_t s_wave[SC_THREADS / WAVE];
    const int64_t t0 = int64_t(blockIdx.x) * SC_TILE;
    uint32_t v = 0;
#pragma unroll
    for (int i = 0; i < SC_ITEMS; i++) {
        int64_t j = t0 + int64_t(i) * SC_THREADS + threadIdx.x;
        v += (j < n) ? in[j] : 0u;
    }
    v = wave_reduce_add(v);
    if (lane_id() == 0)
        s_wave[threadIdx.x / WAVE] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (int w = 0; w < SC_THREADS / WAVE; w++)
            s += s_wave[w];
        sums[blockIdx.x] = s;
    }
}

__global__ void __launch_bounds__(SC_THREADS)
scan_tiles_kernel(uint32_t *__restrict__ data, int64_t n, const uint32_t *__restrict__ offsets)
{
    __shared__ uint32_t s_wave[SC_THREADS / WAVE];
    const int64_t t0 = int64_t(blockIdx.x) * SC_TILE + int64_t(threadIdx.x) * SC_ITEMS;
    uint32_t v[SC_ITEMS];
    uint32_t sum = 0;
#pragma unroll
    for (int i = 0; i < SC_ITEMS; i++) {
        int64_t j = t0 + i;
        v[i] = (j < n) ? data[j] : 0u;
        sum += v[i];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan<SC_THREADS>(sum, s_wave, total);
    if (offsets)
        run += offsets[blockIdx.x];
#pragma unroll
    for (int i = 0; i < SC_ITEMS; i++) {
        int64_t j = t0 + i;
        if (j < n)
            data[j] = run;
        run += v[i];
    }
}

// ------------------------------------------------------------------------------------------
// Stable LSD radix sort of plan entries by dense cell key (iycell * kx + ixcell).
// Per pass: per-tile digit histograms -> device scan (digit-major) -> stable scatter.
// Stability inside a tile: wave w owns items [w*1024, (w+1)*1024) of the tile, processed in
// rounds of 64; equal digits inside a round are ranked by lane via ballot matching.

constexpr int RS_THREADS = 256;
constexpr int RS_NW = RS_THREADS / WAVE;
constexpr int RS_ROUNDS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ROUNDS;   // 4096
constexpr int RS_MAXBITS = 8;
constexpr int RS_RADIX = 1 << RS_MAXBITS;

__device__ inline uint32_t dense_key(uint64_t mt, uint32_t kx)
{
    uint32_t c = uint32_t(mt) & CELL_MASK;
    return (c >> 10) * kx + (c & 1023u);
}

__global__ void __launch_bounds__(RS_THREADS)
radix_hist_kernel(const uint64_t *__restrict__ in, int64_t n, uint32_t kx, int shift, uint32_t dmask,
                  uint32_t *__restrict__ hist, int64_t ntiles)
{
    __shared__ uint32_t s_hist[RS_RADIX];
    s_hist[threadIdx.x] = 0;
    __syncthreads();
    const int64_t t0 = int64_t(blockIdx.x) * RS_TILE;
#pragma unroll 4
    for (int r = 0; r < RS_ROUNDS; r++) {
        int64_t j = t0 + int64_t(r) * RS_THREADS + threadIdx.x;
        if (j < n) {
            uint32_t d = (dense_key(in[j], kx) >> shift) & dmask;
            atomicAdd(&s_hist[d], 1u);
        }
    }
    __syncthreads();
    hist[int64_t(threadIdx.x) * ntiles + blockIdx.x] = s_hist[threadIdx.x];
}

__global__ void __launch_bounds__(RS_THREADS)
radix_scatter_kernel(const uint64_t *__restrict__ in, uint64_t *__restrict__ out, int64_t n,
                     uint32_t kx, int shift, int nbits, const uint32_t *__restrict__ hist_scanned,
                     int64_t ntiles)
{
    __shared__ uint32_t s_cnt[RS_NW][RS_RADIX];
    const int lane = lane_id(), wave = threadIdx.x / WAVE;
    const uint32_t dmask = (1u << nbits) - 1u;

    for (int i = threadIdx.x; i < RS_NW * RS_RADIX; i += RS_THREADS)
        (&s_cnt[0][0])[i] = 0;
    __syncthreads();

    const int64_t w0 = int64_t(blockIdx.x) * RS_TILE + int64_t(wave) * (RS_TILE / RS_NW);
    uint64_t item[RS_ROUNDS];
    uint32_t dig[RS_ROUNDS], rank[RS_ROUNDS];

#pragma unroll
    for (int r = 0; r < RS_ROUNDS; r++) {
        int64_t j = w0 + int64_t(r) * WAVE + lane;
        bool valid = (j < n);
        item[r] = valid ? in[j] : 0ull;
        uint32_t d = valid ? ((dense_key(item[r], kx) >> shift) & dmask) : 0u;
        dig[r] = d;

        uint64_t peers = __ballot(valid);
        for (int bit = 0; bit < nbits; bit++) {
            uint64_t bb = __ballot((d >> bit) & 1u);
            peers &= ((d >> bit) & 1u) ? bb : ~bb;
        }
        uint32_t below = __popcll(peers & lanemask_lt());
        uint32_t old = valid ? s_cnt[wave][d] : 0u;
        if (valid && below == 0)
            s_cnt[wave][d] = old + uint32_t(__popcll(peers));
        rank[r] = old + below;
    }
    __syncthreads();

    // Digit-major prefix over waves, plus this tile's global offset for the digit.
    if (threadIdx.x < (1 << nbits)) {
        uint32_t d = threadIdx.x;
        uint32_t acc = hist_scanned[int64_t(d) * ntiles + blockIdx.x];
        for (int w = 0; w < RS_NW; w++) {
            uint32_t c = s_cnt[w][d];
            s_cnt[w][d] = acc;
            acc += c;
        }
    }
    __syncthreads();

#pragma unroll
    for (int r = 0; r < RS_ROUNDS; r++) {
        int64_t j = w0 + int64_t(r) * WAVE + lane;
        if (j < n)
            out[s_cnt[wave][dig[r]] + rank[r]] = item[r];
    }
}

// ------------------------------------------------------------------------------------------
// Work list: per-cell [begin, end) in the sorted plan, cut into chunks of <= chunk_cap entries.

__global__ void fill_u32_kernel(uint32_t *p, int64_t n, uint32_t v)
{
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        p[i] = v;
}

// Entries whose cell lies outside the global map (only possible in a plan whose emit step
// already reported an error) are skipped, so the tables are never indexed out of range.
__global__ void cell_bounds_kernel(const uint64_t *__restrict__ mt, int64_t n, uint32_t kx, int64_t nkeys,
                                   uint32_t *__restrict__ begin, uint32_t *__restrict__ end)
{
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
        uint32_t k = dense_key(mt[i], kx);
        if (((uint32_t(mt[i]) & 1023u) >= kx) || int64_t(k) >= nkeys)
            continue;
        if (i == 0 || dense_key(mt[i - 1], kx) != k)
            begin[k] = uint32_t(i);
        if (i == n - 1 || dense_key(mt[i + 1], kx) != k)
            end[k] = uint32_t(i + 1);
    }
}

__device__ inline uint32_t nchunks_for(uint32_t cnt, uint32_t cap) { return (cnt + cap - 1) / cap; }

__global__ void chunk_count_kernel(const uint32_t *__restrict__ begin, const uint32_t *__restrict__ end,
                                   int64_t nkeys, uint32_t cap, uint32_t *__restrict__ nch)
{
    for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k <= nkeys; k += int64_t(gridDim.x) * blockDim.x)
        nch[k] = (k < nkeys && end[k] > begin[k]) ? nchunks_for(end[k] - begin[k], cap) : 0u;
}

__global__ void chunk_fill_kernel(const uint32_t *__restrict__ begin, const uint32_t *__restrict__ end,
                                  int64_t nkeys, uint32_t kx, uint32_t cap,
                                  const uint32_t *__restrict__ choff, Chunk *__restrict__ chunks,
                                  int64_t max_chunks)
{
    for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < nkeys; k += int64_t(gridDim.x) * blockDim.x) {
        if (end[k] <= begin[k])
            continue;
        uint32_t b = begin[k], cnt = end[k] - begin[k];
        uint32_t nc = nchunks_for(cnt, cap);
        uint32_t icell = cell_key(int(k / kx), int(k % kx));
        for (uint32_t j = 0; j < nc; j++) {
            Chunk c;
            c.icell = icell;
            c.begin = b + uint32_t((uint64_t(cnt) * j) / nc);
            c.end = b + uint32_t((uint64_t(cnt) * (j + 1)) / nc);
            c.flags = (nc > 1) ? CHUNK_SHARED : 0u;
            if (int64_t(choff[k]) + j < max_chunks)
                chunks[choff[k] + j] = c;
        }
    }
}

// Owned-quad masks for the operators (ownership rule at HALO_X in gmm_common.hpp). For entry e
// of cell Y = (cy, cx) and line l, with has(X) = "l has a sample whose (iy0, ix0) corner is in
// X", fit(X) = "every such sample has ix0 % 64 <= HALO_X - 2 and ix0 + 1 < nx" and touches(X) =
// "a bilinear corner of some sample of l is in X" (exactly: the plan has an entry (l, X)):
//   eligible(X) = has(X) && fit(X) && touches(X - 1)        (X - 1: the cell to the left)
//   R(l, X)     = has(X + 1) && fit(X + 1) && !eligible(X)  (X's entry evaluates X + 1's samples)
// A sample of cell X is evaluated by X - 1 iff R(l, X - 1), else by X. R(l, X) excludes that
// X's own samples go left, so no block both gives and takes, and each sample has exactly one
// evaluator; every entry of l computes these from the same 32 samples, so they agree. Bit q of
// qmask[e] is set iff quad q (samples 4q..4q+3) holds a sample that Y evaluates; QM_TAKES_RIGHT
// is R(l, Y). A line that lies in one cell gives all quads that hold a sample to it. Also counts
// the multi-cell lines (one zflag entry each).
template<bool Periodic>
__global__ void __launch_bounds__(256)
qmask_kernel(const uint64_t *__restrict__ mt, int64_t n, const float *__restrict__ xp, int64_t nsamp, int ny,
             int nx, uint16_t *__restrict__ qmask, uint32_t *__restrict__ nmlines)
{
    // Each wave takes 64 entries at a time, one per lane. An entry of a single-cell line owns
    // every quad of its line that holds a sample: written directly. The entries of multi-cell
    // lines are listed in LDS and handled eight at a time, eight lanes each (one quad per lane).
    __shared__ uint32_t s_nml;
    __shared__ uint32_t s_list[256 / WAVE][WAVE];
    if (threadIdx.x == 0)
        s_nml = 0;
    __syncthreads();
    const int lane = lane_id(), wave = threadIdx.x / WAVE;
    const int g = lane >> 3, q = lane & 7;
    // The planner's passes have checked the pointing's range already.
    const Geom geo = make_geom(ny, nx, Periodic);
    Checks unchecked;
    uint32_t *list = s_list[wave];
    // Line summary bits, OR-reduced over the entry's 32 samples.
    constexpr uint32_t HAS_L = 1, BAD_L = 2, HAS_Y = 4, BAD_Y = 8, HAS_R = 16, BAD_R = 32, TOUCH_L = 64,
                       TOUCH_LL = 128;
    const int64_t nwaves = int64_t(gridDim.x) * (256 / WAVE);
    for (int64_t base = (int64_t(blockIdx.x) * (256 / WAVE) + wave) * WAVE; base < n; base += nwaves * WAVE) {
        const int64_t e = base + lane;
        const uint64_t m = (e < n) ? mt[e] : 0ull;
        const bool multi = (e < n) && (m & MFLAG);
        const uint64_t z = __ballot((e < n) && (m & ZFLAG));
        if (lane == 0 && z)
            atomicAdd(&s_nml, uint32_t(__popcll(z)));
        if (e < n && !multi) {
            // Samples past the end of a partial last line do not exist.
            const int64_t s0 = int64_t(uint32_t(m >> ICL_SHIFT) & ICL_MASK) * LINE;
            const int64_t nv = max<int64_t>(0, min<int64_t>(LINE, nsamp - s0));
            qmask[e] = uint16_t((1u << uint32_t((nv + 3) / 4)) - 1u);
        }
        const uint64_t mb = __ballot(multi);
        if (mb == 0)
            continue;
        if (multi)
            list[__builtin_amdgcn_mbcnt_hi(uint32_t(mb >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mb), 0u))] = uint32_t(lane);
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        const int nm = __popcll(mb);
        for (int r0 = 0; r0 < nm; r0 += 8) {
            const bool on = r0 + g < nm;
            const int src = on ? int(list[r0 + g]) : 0;
            const uint64_t em = __shfl(m, src, WAVE);
            const int64_t ee = base + src;
            uint32_t sum = 0, own_y = 0, own_r = 0;     // own_*: per sample k, bit k
            const uint32_t icell = uint32_t(em) & CELL_MASK;
            const int cy = int(icell >> 10), cx = int(icell & 1023u);
            const int64_t s0 = int64_t(uint32_t(em >> ICL_SHIFT) & ICL_MASK) * LINE + 4 * q;
            const uint32_t nv = on ? uint32_t(max<int64_t>(0, min<int64_t>(4, nsamp - s0))) : 0u;
            float yv[4], xv[4];
            if (nv > 0) {
                // All loads first (indices clamped to the last sample), so that they overlap.
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int64_t sk = s0 + min<int64_t>(k, nv - 1);
                    yv[k] = xp[sk];
                    xv[k] = xp[nsamp + sk];
                }
            }
#pragma unroll
            for (int k = 0; k < 4; k++) {
                if (uint32_t(k) >= nv)
                    continue;
                Pixel<float> p;
                sample_corners<Periodic>(yv[k], xv[k], geo, unchecked, p);
                // Cells of the footprint relative to Y: the owner (p.iy0, p.ix0) is in Y's cell
                // row iff iy0 >> 6 == cy; its column offset is d, that of the ix1 corner d1.
                const bool inrow = (p.iy0 >> 6) == cy;
                const bool rowhit = inrow || ((p.iy1 >> 6) == cy);
                const int d = (p.ix0 >> 6) - cx, d1 = (p.ix1 >> 6) - cx;
                const bool fit = ((p.ix0 & 63) <= HALO_X - 2) && (p.ix0 + 1 < nx);
                // HAS/BAD bit pairs of the cells Y - 1, Y, Y + 1 sit at bit 2 * (d + 1).
                if (inrow && d >= -1 && d <= 1)
                    sum |= (fit ? 1u : 3u) << (2 * (d + 1));
                own_y |= (inrow && d == 0) ? (1u << k) : 0u;
                own_r |= (inrow && d == 1) ? (1u << k) : 0u;
                if (rowhit) {
                    sum |= (d == -1 || d1 == -1) ? TOUCH_L : 0u;
                    sum |= (d == -2 || d1 == -2) ? TOUCH_LL : 0u;
                }
            }
            // OR over the 8 lanes of the entry (aligned groups of 8 within the wave).
            sum |= __shfl_xor(sum, 1, WAVE);
            sum |= __shfl_xor(sum, 2, WAVE);
            sum |= __shfl_xor(sum, 4, WAVE);
            const bool elig_l = (sum & HAS_L) && !(sum & BAD_L) && (sum & TOUCH_LL);
            const bool elig_y = (sum & HAS_Y) && !(sum & BAD_Y) && (sum & TOUCH_L);
            const bool taken_y = elig_y && !elig_l;                  // R(l, Y - 1)
            const bool takes_r = (sum & HAS_R) && !(sum & BAD_R) && !elig_y;   // R(l, Y)
            const bool bit = (own_y != 0 && !taken_y) || (own_r != 0 && takes_r);
            const uint64_t b = __ballot(bit);
            if (on && q == 0)
                qmask[ee] = uint16_t(((b >> (lane & ~7)) & 0xffu) | (takes_r ? QM_TAKES_RIGHT : 0u));
        }
        // The list is read before the next batch writes it (LDS operations of a wave complete
        // in issue order; the fence keeps the compiler from reordering them).
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    }
    __syncthreads();
    if (threadIdx.x == 0 && s_nml)
        atomicAdd(nmlines, s_nml);
}

__global__ void copy_word_kernel(const uint32_t *src, uint32_t *dst) { *dst = *src; }

__global__ void __launch_bounds__(1024)
reduce_or_kernel(const uint32_t *__restrict__ in, int64_t n, uint32_t *__restrict__ out)
{
    __shared__ uint32_t s;
    if (threadIdx.x == 0)
        s = 0;
    __syncthreads();
    uint32_t v = 0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x)
        v |= in[i];
    v = wave_reduce_or(v);
    if (lane_id() == 0 && v)
        atomicOr(&s, v);
    __syncthreads();
    if (threadIdx.x == 0)
        *out = s;
}

// ------------------------------------------------------------------------------------------
// Test aids.

template<bool Periodic>
__global__ void quantize_kernel(const float *__restrict__ xp, int64_t nsamp, int ny, int nx,
                                int32_t *__restrict__ iypix, int32_t *__restrict__ ixpix,
                                uint32_t *__restrict__ err_word)
{
    uint32_t err = 0;
    for (int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; s < nsamp; s += int64_t(gridDim.x) * blockDim.x) {
        Pixel<float> p;
        err |= locate_pixel(xp[s], xp[s + nsamp], ny, nx, Periodic, p);
        iypix[s] = p.iy0;
        ixpix[s] = p.ix0;
    }
    err = wave_reduce_or(err);
    if (lane_id() == 0 && err)
        atomicOr(err_word, err);
}

// Walks the work list exactly as the planned operators do: each wave takes its wave_slice() of
// the chunk, in batches of 64 entries (one per lane, one load) and steps of 8 entries (lane
// 8*li + q on entry li of the step, with the entry broadcast from the batch by a shuffle), and
// counts the visits of each entry (lane q == 0); a visit to an entry of another cell sets bit 0
// of the error word, an entry reached twice within one wave bit 1.
__global__ void iterator_test_kernel(const uint64_t *__restrict__ plan_mt, const Chunk *__restrict__ chunks,
                                     uint32_t *__restrict__ visits, uint32_t *__restrict__ err_word)
{
    const int lane = lane_id(), wave = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
    const int li = lane >> 3, q = lane & 7;
    const Chunk ch = chunks[blockIdx.x];
    uint32_t wb, we;
    wave_slice(ch, wave, nw, wb, we);
    uint32_t err = 0, last = 0xffffffffu;
    for (uint32_t base = wb; base < we; base += 64) {
        const uint32_t cnt = min(64u, we - base);
        const uint32_t key = (uint32_t(lane) < cnt) ? (uint32_t(plan_mt[base + lane]) & CELL_MASK) : 0u;
        for (uint32_t j = 0; j < cnt; j += 8) {
            const uint32_t idx = j + li;
            const uint32_t k = __shfl(key, int(min(idx, 63u)), WAVE);
            if (idx >= cnt)
                continue;
            const uint32_t e = base + idx;
            if (k != ch.icell)
                err |= 1u;
            if (last != 0xffffffffu && e <= last)
                err |= 2u;
            last = e;
            if (q == 0)
                atomicAdd(visits + e, 1u);
        }
    }
    err = wave_reduce_or(err);
    if (lane == 0 && err)
        atomicOr(err_word, err);
}

// Largest-first order of the operator work list (below): sort items (index << 20 | key) with
// key = 8192 - entries, then gather the chunks in item order.
__global__ void chunk_key_kernel(const Chunk *__restrict__ chunks, int64_t n, uint64_t *__restrict__ items)
{
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
        const uint32_t len = chunks[i].end - chunks[i].begin;
        const uint32_t key = 8192u - min(len, 8192u);
        items[i] = (uint64_t(i) << 20) | key;
    }
}

__global__ void chunk_copy_kernel(const Chunk *__restrict__ src, int64_t n, Chunk *__restrict__ dst)
{
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}

__global__ void chunk_gather_kernel(const uint64_t *__restrict__ items, const Chunk *__restrict__ src, int64_t n,
                                    Chunk *__restrict__ dst)
{
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        dst[i] = src[items[i] >> 20];
}

inline int grid_for(int64_t n, int threads, int64_t cap = 65536)
{
    int64_t g = (n + threads - 1) / threads;
    g = (g < 1) ? 1 : g;
    return int((g < cap) ? g : cap);
}

}  // namespace

// ==========================================================================================
// Host launchers.

void launch_preplan_count(const float *xp, int64_t nsamp, const gmm_preplan &pp,
                          uint32_t *nmt_out, uint32_t *err_out, hipStream_t stream)
{
    const auto &p = pp.pparams;
    dim3 grid(unsigned(p.planner_nblocks)), block(PP_THREADS);
    if (pp.periodic_xcoord)
        preplan_count_kernel<true><<<grid, block, 0, stream>>>(xp, nsamp, p.ncl, p.ncl_per_threadblock,
            int(pp.nypix_global), int(pp.nxpix_global), nmt_out, err_out);
    else
        preplan_count_kernel<false><<<grid, block, 0, stream>>>(xp, nsamp, p.ncl, p.ncl_per_threadblock,
            int(pp.nypix_global), int(pp.nxpix_global), nmt_out, err_out);
    GMM_HIP(hipGetLastError());
}

void launch_preplan_finish(uint32_t *nmt, const uint32_t *err, int64_t nblocks,
                           uint32_t * /*summary_gpu*/, hipStream_t stream)
{
    preplan_finish_kernel<<<1, PREPLAN_SIZE, 0, stream>>>(nmt, const_cast<uint32_t *>(err), int(nblocks));
    GMM_HIP(hipGetLastError());
}

void launch_plan_emit(const float *xp, const gmm_preplan &pp, uint64_t *out_mt,
                      uint32_t *err_out, hipStream_t stream)
{
    const auto &p = pp.pparams;
    dim3 grid(unsigned(p.planner_nblocks)), block(EM_THREADS);
    if (pp.periodic_xcoord)
        plan_emit_kernel<true><<<grid, block, 0, stream>>>(xp, pp.nsamp, p.ncl, p.ncl_per_threadblock,
            int(pp.nypix_global), int(pp.nxpix_global), pp.nmt_cumsum_gpu, out_mt, err_out);
    else
        plan_emit_kernel<false><<<grid, block, 0, stream>>>(xp, pp.nsamp, p.ncl, p.ncl_per_threadblock,
            int(pp.nypix_global), int(pp.nxpix_global), pp.nmt_cumsum_gpu, out_mt, err_out);
    GMM_HIP(hipGetLastError());
}

int64_t scan_scratch_nbytes(int64_t n)
{
    if (n <= SC_TILE)
        return 0;
    int64_t nt = (n + SC_TILE - 1) / SC_TILE;
    return align128(4 * nt) + scan_scratch_nbytes(nt);
}

void exclusive_scan_u32(uint32_t *data, int64_t n, void *scratch, hipStream_t stream)
{
    if (n <= 0)
        return;
    if (n <= SC_TILE) {
        scan_tiles_kernel<<<1, SC_THREADS, 0, stream>>>(data, n, nullptr);
        GMM_HIP(hipGetLastError());
        return;
    }
    int64_t nt = (n + SC_TILE - 1) / SC_TILE;
    uint32_t *sums = reinterpret_cast<uint32_t *>(scratch);
    scan_tile_sums_kernel<<<unsigned(nt), SC_THREADS, 0, stream>>>(data, n, sums);
    GMM_HIP(hipGetLastError());
    exclusive_scan_u32(sums, nt, reinterpret_cast<uint8_t *>(scratch) + align128(4 * nt), stream);
    scan_tiles_kernel<<<unsigned(nt), SC_THREADS, 0, stream>>>(data, n, sums);
    GMM_HIP(hipGetLastError());
}

namespace {

struct SortLayout
{
    int64_t ntiles;
    int64_t hist_off, hist_scan_off;   // histogram, scan scratch
    int64_t begin_off, end_off, nch_off, nch_scan_off;
    int64_t zc_off, zc_scan_off;
    int64_t total;
};

SortLayout sort_layout(int64_t nmt, int64_t nkeys)
{
    SortLayout L;
    L.ntiles = (nmt + RS_TILE - 1) / RS_TILE;
    int64_t o = 0;
    L.hist_off = o;      o += align128(4 * RS_RADIX * L.ntiles);
    L.hist_scan_off = o; o += align128(scan_scratch_nbytes(RS_RADIX * L.ntiles));
    L.begin_off = o;     o += align128(4 * nkeys);
    L.end_off = o;       o += align128(4 * nkeys);
    L.nch_off = o;       o += align128(4 * (nkeys + 1));
    L.nch_scan_off = o;  o += align128(scan_scratch_nbytes(nkeys + 1));
    int64_t zt = (nmt + SC_TILE - 1) / SC_TILE;
    L.zc_off = o;        o += align128(4 * (zt + 1));
    L.zc_scan_off = o;   o += align128(scan_scratch_nbytes(zt + 1));
    L.total = o;
    return L;
}

}  // namespace

// Reorders the work list of nch chunks largest first (stable: equal sizes stay in cell order),
// so that the operators' last blocks are short ones (measured: tod2map 1.4% faster on the
// 2048x262144 raster). Chunk sizes are at most chunk_cap <= 8192 entries: 13-bit keys, sorted
// with the plan's radix sort (kx = 1024 makes its key the item's low 20 bits). Uses a and b
// (>= 8 * nmt bytes each) and the plan scratch; skipped when 16 * nch bytes do not fit in b.
void order_chunks(const gmm_preplan &pp, Chunk *chunks, int64_t nch, uint64_t *a, uint64_t *b, void *scratch,
                  int64_t scratch_nbytes, hipStream_t stream)
{
    if (nch < 2 || 2 * nch > pp.plan_nmt)
        return;
    SortLayout L = sort_layout(nch, 1);
    GMM_REQUIRE(scratch_nbytes >= L.total, "order_chunks: scratch too small");
    uint8_t *sc = reinterpret_cast<uint8_t *>(scratch);
    uint32_t *hist = reinterpret_cast<uint32_t *>(sc + L.hist_off);
    chunk_key_kernel<<<grid_for(nch, 256), 256, 0, stream>>>(chunks, nch, a);
    GMM_HIP(hipGetLastError());
    const int bits = 13, npasses = 2, per = 7;
    uint64_t *src = a, *dst = b;
    for (int pass = 0, shift = 0; pass < npasses; pass++, shift += per) {
        const int nb = (bits - shift < per) ? (bits - shift) : per;
        radix_hist_kernel<<<unsigned(L.ntiles), RS_THREADS, 0, stream>>>(src, nch, 1024u, shift, (1u << nb) - 1u, hist,
                                                                          L.ntiles);
        GMM_HIP(hipGetLastError());
        exclusive_scan_u32(hist, int64_t(1 << nb) * L.ntiles, sc + L.hist_scan_off, stream);
        radix_scatter_kernel<<<unsigned(L.ntiles), RS_THREADS, 0, stream>>>(src, dst, nch, 1024u, shift, nb, hist,
                                                                             L.ntiles);
        GMM_HIP(hipGetLastError());
        std::swap(src, dst);
    }
    // The sorted items are in src (= a after two passes); copy the chunks into b, gather back.
    Chunk *copy = reinterpret_cast<Chunk *>(dst);
    chunk_copy_kernel<<<grid_for(nch, 256), 256, 0, stream>>>(chunks, nch, copy);
    chunk_gather_kernel<<<grid_for(nch, 256), 256, 0, stream>>>(src, copy, nch, chunks);
    GMM_HIP(hipGetLastError());
}

int64_t plan_scratch_nbytes(const gmm_preplan &pp)
{
    return sort_layout(pp.plan_nmt, pp.nycells_g * pp.nxcells_g).total;
}

void plan_sort(const gmm_preplan &pp, uint64_t *a, uint64_t *b, uint64_t *out,
               void *scratch, int64_t scratch_nbytes, hipStream_t stream)
{
    const int64_t n = pp.plan_nmt;
    const uint32_t kx = uint32_t(pp.nxcells_g);
    SortLayout L = sort_layout(n, pp.nycells_g * pp.nxcells_g);
    GMM_REQUIRE(scratch_nbytes >= L.total, "plan_sort: scratch too small");
    uint8_t *sc = reinterpret_cast<uint8_t *>(scratch);
    uint32_t *hist = reinterpret_cast<uint32_t *>(sc + L.hist_off);

    const int bits = pp.key_bits;
    if (bits == 0) {
        GMM_HIP(hipMemcpyAsync(out, a, size_t(n) * 8, hipMemcpyDeviceToDevice, stream));
        return;
    }
    const int npasses = (bits + RS_MAXBITS - 1) / RS_MAXBITS;
    const int per = (bits + npasses - 1) / npasses;

    const uint64_t *src = a;
    for (int pass = 0, shift = 0; pass < npasses; pass++, shift += per) {
        int nb = (bits - shift < per) ? (bits - shift) : per;
        uint64_t *dst = (pass == npasses - 1) ? out : ((src == a) ? b : a);
        radix_hist_kernel<<<unsigned(L.ntiles), RS_THREADS, 0, stream>>>(src, n, kx, shift, (1u << nb) - 1u, hist, L.ntiles);
        GMM_HIP(hipGetLastError());
        exclusive_scan_u32(hist, int64_t(1 << nb) * L.ntiles, sc + L.hist_scan_off, stream);
        radix_scatter_kernel<<<unsigned(L.ntiles), RS_THREADS, 0, stream>>>(src, dst, n, kx, shift, nb, hist, L.ntiles);
        GMM_HIP(hipGetLastError());
        src = dst;
    }
}

namespace {

// Shared by plan_schedule() and the iterator test: chunk list from a cell-sorted plan.
// counts_gpu[0] <- number of chunks.
void build_chunks(const uint64_t *mt, int64_t n, int64_t nkeys, uint32_t kx, uint32_t cap,
                  Chunk *chunks, int64_t max_chunks, uint32_t *count_gpu, uint8_t *sc, const SortLayout &L,
                  hipStream_t stream)
{
    uint32_t *begin = reinterpret_cast<uint32_t *>(sc + L.begin_off);
    uint32_t *end = reinterpret_cast<uint32_t *>(sc + L.end_off);
    uint32_t *nch = reinterpret_cast<uint32_t *>(sc + L.nch_off);

    fill_u32_kernel<<<grid_for(nkeys, 256, 4096), 256, 0, stream>>>(begin, nkeys, 0u);
    fill_u32_kernel<<<grid_for(nkeys, 256, 4096), 256, 0, stream>>>(end, nkeys, 0u);
    cell_bounds_kernel<<<grid_for(n, 256), 256, 0, stream>>>(mt, n, kx, nkeys, begin, end);
    chunk_count_kernel<<<grid_for(nkeys + 1, 256, 4096), 256, 0, stream>>>(begin, end, nkeys, cap, nch);
    GMM_HIP(hipGetLastError());
    exclusive_scan_u32(nch, nkeys + 1, sc + L.nch_scan_off, stream);
    chunk_fill_kernel<<<grid_for(nkeys, 256, 4096), 256, 0, stream>>>(begin, end, nkeys, kx, cap, nch, chunks, max_chunks);
    copy_word_kernel<<<1, 1, 0, stream>>>(nch + nkeys, count_gpu);
    GMM_HIP(hipGetLastError());
}

}  // namespace

void plan_schedule(const gmm_preplan &pp, const uint64_t *sorted_mt, const float *xp, Chunk *chunks,
                   uint16_t *qmask, uint32_t *counts_gpu, void *scratch,
                   int64_t scratch_nbytes, hipStream_t stream)
{
    const int64_t n = pp.plan_nmt;
    const int64_t nkeys = pp.nycells_g * pp.nxcells_g;
    SortLayout L = sort_layout(n, nkeys);
    GMM_REQUIRE(scratch_nbytes >= L.total, "plan_schedule: scratch too small");
    uint8_t *sc = reinterpret_cast<uint8_t *>(scratch);

    build_chunks(sorted_mt, n, nkeys, uint32_t(pp.nxcells_g), uint32_t(pp.chunk_cap), chunks,
                 pp.max_nchunks, counts_gpu, sc, L, stream);

    // Owned-quad masks and the multi-cell line count.
    fill_u32_kernel<<<1, 64, 0, stream>>>(counts_gpu + 1, 1, 0u);
    const int grid = int(std::min<int64_t>((n + 255) / 256, 2048));
    if (pp.periodic_xcoord)
        qmask_kernel<true><<<grid, 256, 0, stream>>>(sorted_mt, n, xp, pp.nsamp, int(pp.nypix_global),
                                                     int(pp.nxpix_global), qmask, counts_gpu + 1);
    else
        qmask_kernel<false><<<grid, 256, 0, stream>>>(sorted_mt, n, xp, pp.nsamp, int(pp.nypix_global),
                                                      int(pp.nxpix_global), qmask, counts_gpu + 1);
    GMM_HIP(hipGetLastError());
}

int64_t build_chunks_generic_scratch(int64_t nmt)
{
    return sort_layout(nmt, int64_t(1) << 20).total;
}

int64_t build_chunks_generic(const uint64_t *plan_mt_gpu, int64_t nmt, int64_t chunk_cap,
                             Chunk *chunks_gpu, int64_t max_chunks, void *scratch,
                             int64_t scratch_nbytes, hipStream_t stream)
{
    const int64_t nkeys = int64_t(1) << 20;
    SortLayout L = sort_layout(nmt, nkeys);
    GMM_REQUIRE(scratch_nbytes >= L.total, "build_chunks_generic: scratch too small");
    uint8_t *sc = reinterpret_cast<uint8_t *>(scratch);
    uint32_t *count = reinterpret_cast<uint32_t *>(sc + L.zc_off);
    build_chunks(plan_mt_gpu, nmt, nkeys, 1024u, uint32_t(chunk_cap), chunks_gpu, max_chunks, count, sc, L, stream);
    uint32_t h = 0;
    GMM_HIP(hipMemcpyAsync(&h, count, 4, hipMemcpyDeviceToHost, stream));
    GMM_HIP(hipStreamSynchronize(stream));
    return int64_t(h);
}

void launch_fill_u32(uint32_t *p, int64_t n, uint32_t v, hipStream_t stream)
{
    if (n <= 0)
        return;
    fill_u32_kernel<<<grid_for(n, 256, 4096), 256, 0, stream>>>(p, n, v);
    GMM_HIP(hipGetLastError());
}

void launch_reduce_or(const uint32_t *in, int64_t n, uint32_t *out, hipStream_t stream)
{
    reduce_or_kernel<<<1, 1024, 0, stream>>>(in, n, out);
    GMM_HIP(hipGetLastError());
}

void launch_quantize(const float *xp, int64_t nsamp, int ny, int nx, bool periodic,
                     int32_t *iypix, int32_t *ixpix, uint32_t *err_word, hipStream_t stream)
{
    fill_u32_kernel<<<1, 64, 0, stream>>>(err_word, 1, 0u);
    int g = grid_for(nsamp, 256, 16384);
    if (periodic)
        quantize_kernel<true><<<g, 256, 0, stream>>>(xp, nsamp, ny, nx, iypix, ixpix, err_word);
    else
        quantize_kernel<false><<<g, 256, 0, stream>>>(xp, nsamp, ny, nx, iypix, ixpix, err_word);
    GMM_HIP(hipGetLastError());
}

void launch_iterator_test(const uint64_t *plan_mt, const Chunk *chunks, int64_t nchunks,
                          int waves_per_block, uint32_t *visit_counts, uint32_t *err_word,
                          hipStream_t stream)
{
    fill_u32_kernel<<<1, 64, 0, stream>>>(err_word, 1, 0u);
    if (nchunks > 0)
        iterator_test_kernel<<<unsigned(nchunks), unsigned(64 * waves_per_block), 0, stream>>>(plan_mt, chunks, visit_counts, err_word);
    GMM_HIP(hipGetLastError());
}

}  // namespace gmm
