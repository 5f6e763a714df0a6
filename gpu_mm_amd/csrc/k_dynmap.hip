// Dynamic map expansion on gfx950: adds the cells of a plan to a growing cell-list map.
//
// cell_offsets[iycell][ixcell] holds -1 for a targeted cell not yet in the map, -2 for an
// untargeted one, and 3*64*64*i for the i-th cell present. Given plan entries sorted by cell (the
// low 20 bits), every targeted cell that appears in the plan and is not yet present gets the next
// free index. Reference: src_lib/expand_dynamic_map.cu:18-104, which assigns the indices in
// atomic (nondeterministic) order; here they are assigned in increasing cell-key order
// (flag, device exclusive scan, scatter), so the map layout is reproducible.

#include "gmm_internal.hpp"

namespace gmm {

namespace {

constexpr int DM_THREADS = 256;

__device__ inline bool dyn_candidate(const uint64_t *__restrict__ mt, int64_t i, int64_t n,
                                     const int64_t *__restrict__ cell_offsets, uint32_t nycells, uint32_t nxcells,
                                     int64_t &ioff)
{
    const uint32_t icell = uint32_t(mt[i]) & CELL_MASK;
    const bool last = (i == n - 1) || ((uint32_t(mt[i + 1]) & CELL_MASK) != icell);
    const uint32_t cy = icell >> 10, cx = icell & 1023u;
    ioff = int64_t(cy) * nxcells + cx;
    return last && (cy < nycells) && (cx < nxcells) && (cell_offsets[ioff] == -1);
}

__global__ void __launch_bounds__(DM_THREADS)
dyn_flag_kernel(const uint64_t *__restrict__ mt, int64_t n, const int64_t *__restrict__ cell_offsets,
                uint32_t nycells, uint32_t nxcells, uint32_t *__restrict__ pos)
{
    for (int64_t i = int64_t(blockIdx.x) * DM_THREADS + threadIdx.x; i <= n; i += int64_t(gridDim.x) * DM_THREADS) {
        int64_t ioff;
        pos[i] = (i < n && dyn_candidate(mt, i, n, cell_offsets, nycells, nxcells, ioff)) ? 1u : 0u;
    }
}

// After the exclusive scan, entry i is a new cell iff pos[i+1] != pos[i]; pos[n] is the count.
__global__ void __launch_bounds__(DM_THREADS)
dyn_assign_kernel(const uint64_t *__restrict__ mt, int64_t n, int64_t *__restrict__ cell_offsets,
                  uint32_t nxcells, const uint32_t *__restrict__ pos, uint32_t *__restrict__ global_ncells)
{
    const uint32_t base = *global_ncells;
    for (int64_t i = int64_t(blockIdx.x) * DM_THREADS + threadIdx.x; i < n; i += int64_t(gridDim.x) * DM_THREADS) {
        if (pos[i + 1] != pos[i]) {
            const uint32_t icell = uint32_t(mt[i]) & CELL_MASK;
            const int64_t ioff = int64_t(icell >> 10) * nxcells + (icell & 1023u);
            cell_offsets[ioff] = int64_t(base + pos[i]) * int64_t(3 * CELL_NPIX);
        }
    }
}

// Runs after every block of dyn_assign_kernel has read the old count (stream order).
__global__ void dyn_count_kernel(const uint32_t *__restrict__ pos, int64_t n, uint32_t *__restrict__ global_ncells)
{
    *global_ncells += pos[n];
}

}  // namespace

int64_t expand_dynamic_scratch_nbytes(int64_t nmt)
{
    return align128(4 * (nmt + 1)) + scan_scratch_nbytes(nmt + 1);
}

void launch_expand_dynamic(uint32_t *global_ncells, int64_t *cell_offsets, int64_t nycells, int64_t nxcells,
                           const uint64_t *plan_mt, int64_t nmt, void *scratch, hipStream_t stream)
{
    if (nmt <= 0)
        return;
    uint32_t *pos = static_cast<uint32_t *>(scratch);
    const int grid = int(std::min<int64_t>((nmt + DM_THREADS) / DM_THREADS, 4096));
    dyn_flag_kernel<<<grid, DM_THREADS, 0, stream>>>(plan_mt, nmt, cell_offsets, uint32_t(nycells), uint32_t(nxcells),
                                                     pos);
    GMM_HIP(hipGetLastError());
    exclusive_scan_u32(pos, nmt + 1, static_cast<uint8_t *>(scratch) + align128(4 * (nmt + 1)), stream);
    dyn_assign_kernel<<<grid, DM_THREADS, 0, stream>>>(plan_mt, nmt, cell_offsets, uint32_t(nxcells), pos,
                                                       global_ncells);
    GMM_HIP(hipGetLastError());
    dyn_count_kernel<<<1, 1, 0, stream>>>(pos, nmt, global_ncells);
    GMM_HIP(hipGetLastError());
}

}  // namespace gmm
