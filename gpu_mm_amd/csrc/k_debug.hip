// Debug checks (debug=True) on gfx950: device-side invariant checks of plans and of the planned
// operators, reported as counts of violations that the host turns into an exception. The
// reference asserts the same kind of invariants inside its kernels under a Debug template
// parameter (include/plan_iterator.hpp:64-68, 91-103, 137-141; include/gpu_mm_internals.hpp:
// 224-235; src_lib/tod2map.cu:136; src_lib/map2tod.cu:107); here a violation never traps the GPU:
// it is counted, and every kernel runs to completion.
//
// Plan invariants (check_plan_invariants):
//   - entries sorted by cell, and inside a cell by line, each (cell, line) once;
//   - every line index below ceil(nsamp/32); every line has an entry; mflag set on all entries of
//     a multi-cell line and on none of a single-cell line; exactly one zflag per multi-cell line;
//   - the operator work list covers every entry exactly once, each chunk within one cell.
// Operator invariants (the planned operators' debug instantiations count into DebugBuffers):
//   - each plan entry is walked by exactly one wave;
//   - each sample is evaluated by exactly one block (sample ownership and the halo rule);
//   - no halo-tile access falls outside the tile.

#include "gmm_internal.hpp"

#include <sstream>

namespace gmm {

namespace {

constexpr int DB_THREADS = 256;

inline int db_grid(int64_t n)
{
    return int(std::max<int64_t>(1, std::min<int64_t>((n + DB_THREADS - 1) / DB_THREADS, 8192)));
}

// viol += number of a[i] != expect, i < n.
__global__ void __launch_bounds__(DB_THREADS)
count_ne_kernel(const uint32_t *__restrict__ a, int64_t n, uint32_t expect, uint32_t *__restrict__ viol)
{
    uint32_t c = 0;
    for (int64_t i = int64_t(blockIdx.x) * DB_THREADS + threadIdx.x; i < n; i += int64_t(gridDim.x) * DB_THREADS)
        c += (a[i] != expect) ? 1u : 0u;
    if (c)
        atomicAdd(viol, c);
}

// viol[0]: order violations; viol[1]: line indices out of range. Per line: entries, zflags, mflags.
__global__ void __launch_bounds__(DB_THREADS)
plan_entries_kernel(const uint64_t *__restrict__ mt, int64_t nmt, uint32_t ncl, uint32_t *__restrict__ n_ent,
                    uint32_t *__restrict__ n_z, uint32_t *__restrict__ n_m, uint32_t *__restrict__ viol)
{
    for (int64_t i = int64_t(blockIdx.x) * DB_THREADS + threadIdx.x; i < nmt; i += int64_t(gridDim.x) * DB_THREADS) {
        const uint64_t e = mt[i];
        const uint32_t cell = uint32_t(e) & CELL_MASK, icl = uint32_t(e >> ICL_SHIFT) & ICL_MASK;
        if (i + 1 < nmt) {
            const uint64_t f = mt[i + 1];
            const uint32_t cell2 = uint32_t(f) & CELL_MASK, icl2 = uint32_t(f >> ICL_SHIFT) & ICL_MASK;
            if (cell > cell2 || (cell == cell2 && icl >= icl2))
                atomicAdd(viol, 1u);
        }
        if (icl >= ncl) {
            atomicAdd(viol + 1, 1u);
            continue;
        }
        atomicAdd(n_ent + icl, 1u);
        if (e & ZFLAG)
            atomicAdd(n_z + icl, 1u);
        if (e & MFLAG)
            atomicAdd(n_m + icl, 1u);
    }
}

// viol[2]: lines without an entry; viol[3]: mflag/zflag rule violations.
__global__ void __launch_bounds__(DB_THREADS)
plan_lines_kernel(const uint32_t *__restrict__ n_ent, const uint32_t *__restrict__ n_z, const uint32_t *__restrict__ n_m,
                  uint32_t ncl, uint32_t *__restrict__ viol)
{
    uint32_t c_empty = 0, c_flag = 0;
    for (int64_t l = int64_t(blockIdx.x) * DB_THREADS + threadIdx.x; l < ncl; l += int64_t(gridDim.x) * DB_THREADS) {
        const uint32_t n = n_ent[l];
        c_empty += (n == 0) ? 1u : 0u;
        const bool multi = n > 1;
        c_flag += (n_m[l] != (multi ? n : 0u) || n_z[l] != (multi ? 1u : 0u)) ? 1u : 0u;
    }
    if (c_empty)
        atomicAdd(viol + 2, c_empty);
    if (c_flag)
        atomicAdd(viol + 3, c_flag);
}

// Marks the entries of every chunk in 'cover'; viol[4]: chunks out of range or spanning cells.
__global__ void __launch_bounds__(DB_THREADS)
chunk_cover_kernel(const Chunk *__restrict__ chunks, int64_t nch, const uint64_t *__restrict__ mt, int64_t nmt,
                   uint32_t *__restrict__ cover, uint32_t *__restrict__ viol)
{
    for (int64_t c = blockIdx.x; c < nch; c += gridDim.x) {
        const Chunk ch = chunks[c];
        if (!(ch.begin < ch.end && int64_t(ch.end) <= nmt)) {
            if (threadIdx.x == 0)
                atomicAdd(viol + 4, 1u);
            continue;
        }
        uint32_t bad = 0;
        for (uint32_t i = ch.begin + threadIdx.x; i < ch.end; i += DB_THREADS) {
            atomicAdd(cover + i, 1u);
            bad += ((uint32_t(mt[i]) & CELL_MASK) != ch.icell) ? 1u : 0u;
        }
        if (bad)
            atomicAdd(viol + 4, bad);
    }
}

struct DevBuf
{
    void *p = nullptr;
    explicit DevBuf(int64_t nbytes) { GMM_HIP(hipMalloc(&p, size_t(std::max<int64_t>(nbytes, 4)))); }
    ~DevBuf() { (void)hipFree(p); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    uint32_t *u32(int64_t word_offset = 0) const { return static_cast<uint32_t *>(p) + word_offset; }
};

}  // namespace

void check_plan_invariants(const gmm_plan &plan, const char *where, hipStream_t stream)
{
    const int64_t nmt = plan.pp.plan_nmt;
    const uint32_t ncl = uint32_t((plan.pp.nsamp + 31) / 32);
    DevBuf viol(8 * 4), lines(3 * int64_t(ncl) * 4), cover(nmt * 4);
    GMM_HIP(hipMemsetAsync(viol.p, 0, 8 * 4, stream));
    GMM_HIP(hipMemsetAsync(lines.p, 0, size_t(3 * int64_t(ncl) * 4), stream));
    GMM_HIP(hipMemsetAsync(cover.p, 0, size_t(std::max<int64_t>(nmt * 4, 4)), stream));
    uint32_t *n_ent = lines.u32(), *n_z = lines.u32(ncl), *n_m = lines.u32(2 * int64_t(ncl));
    if (nmt > 0)
        plan_entries_kernel<<<db_grid(nmt), DB_THREADS, 0, stream>>>(plan.plan_mt, nmt, ncl, n_ent, n_z, n_m, viol.u32());
    GMM_HIP(hipGetLastError());
    plan_lines_kernel<<<db_grid(ncl), DB_THREADS, 0, stream>>>(n_ent, n_z, n_m, ncl, viol.u32());
    GMM_HIP(hipGetLastError());
    if (plan.nchunks > 0)
        chunk_cover_kernel<<<int(std::min<int64_t>(plan.nchunks, 8192)), DB_THREADS, 0, stream>>>(
            plan.chunks, plan.nchunks, plan.plan_mt, nmt, cover.u32(), viol.u32());
    GMM_HIP(hipGetLastError());
    if (nmt > 0)
        count_ne_kernel<<<db_grid(nmt), DB_THREADS, 0, stream>>>(cover.u32(), nmt, 1u, viol.u32(5));
    GMM_HIP(hipGetLastError());
    uint32_t v[8];
    GMM_HIP(hipMemcpyAsync(v, viol.p, sizeof(v), hipMemcpyDeviceToHost, stream));
    GMM_HIP(hipStreamSynchronize(stream));
    if (v[0] | v[1] | v[2] | v[3] | v[4] | v[5]) {
        std::ostringstream ss;
        ss << where << ": plan invariant violated (debug checks): " << v[0] << " entries out of (cell, line) order, "
           << v[1] << " line indices out of range, " << v[2] << " lines without an entry, " << v[3]
           << " lines breaking the mflag/zflag rule, " << v[4] << " work-list chunks out of range or spanning cells, "
           << v[5] << " entries not covered exactly once by the work list";
        fail(ss.str());
    }
}

OpDebug::OpDebug(int64_t nmt, int64_t nsamp, hipStream_t stream) : nmt_(nmt), nsamp_(nsamp), stream_(stream)
{
    const int64_t words = nmt + nsamp + 8;
    GMM_HIP(hipMalloc(&mem_, size_t(words * 4)));
    GMM_HIP(hipMemsetAsync(mem_, 0, size_t(words * 4), stream));
    uint32_t *m = static_cast<uint32_t *>(mem_);
    bufs.entry_visits = m + 8;
    bufs.sample_evals = m + 8 + nmt;
    bufs.tile_bad = m;
}

OpDebug::~OpDebug()
{
    (void)hipFree(mem_);
}

void OpDebug::check(const char *where)
{
    uint32_t *viol = static_cast<uint32_t *>(mem_) + 1;
    if (nmt_ > 0)
        count_ne_kernel<<<db_grid(nmt_), DB_THREADS, 0, stream_>>>(bufs.entry_visits, nmt_, 1u, viol);
    GMM_HIP(hipGetLastError());
    count_ne_kernel<<<db_grid(nsamp_), DB_THREADS, 0, stream_>>>(bufs.sample_evals, nsamp_, 1u, viol + 1);
    GMM_HIP(hipGetLastError());
    uint32_t v[3];
    GMM_HIP(hipMemcpyAsync(v, mem_, sizeof(v), hipMemcpyDeviceToHost, stream_));
    GMM_HIP(hipStreamSynchronize(stream_));
    if (v[0] | v[1] | v[2]) {
        std::ostringstream ss;
        ss << where << ": operator invariant violated (debug checks): " << v[1]
           << " plan entries not walked exactly once, " << v[2] << " samples not evaluated exactly once, " << v[0]
           << " halo-tile accesses outside the tile";
        fail(ss.str());
    }
}

}  // namespace gmm
