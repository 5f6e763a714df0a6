// Dynamic map expansion on gfx950: adds the cells of a plan to a growing cell-list map.
//
// cell_offsets[iycell][ixcell] holds -1 for a targeted cell not yet in the map, -2 for an
// untargeted one, and 3*64*64*i for the i-th cell present. Given plan entries sorted by cell (the
// low 20 bits), every targeted cell that appears in the plan and is not yet present gets the next
// free index. Reference: src_lib/expand_dynamic_map.cu:18-104, which assigns the indices in
// atomic (nondeterministic) order; here they are assigned in increasing cell-key order
// (flag, device exclusive scan, scatter), so the map layout is reproducible.

#include <mutex>

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

// Allocation-free variant: per-cell flags over the (nycells, nxcells) grid instead of per-entry
// positions, so its scratch has a fixed maximum size (nycells, nxcells <= 1024). Scanned in
// row-major cell order, which is cell-key order: the new cells get the same indices as above.

__global__ void __launch_bounds__(DM_THREADS)
dyn_grid_clear_kernel(uint32_t *__restrict__ flags, int64_t n)
{
    for (int64_t i = int64_t(blockIdx.x) * DM_THREADS + threadIdx.x; i < n; i += int64_t(gridDim.x) * DM_THREADS)
        flags[i] = 0u;
}

// One store per run of entries of a cell (the plan is sorted by cell).
__global__ void __launch_bounds__(DM_THREADS)
dyn_grid_mark_kernel(const uint64_t *__restrict__ mt, int64_t n, const int64_t *__restrict__ cell_offsets,
                     uint32_t nycells, uint32_t nxcells, uint32_t *__restrict__ flags)
{
    for (int64_t i = int64_t(blockIdx.x) * DM_THREADS + threadIdx.x; i < n; i += int64_t(gridDim.x) * DM_THREADS) {
        int64_t ioff;
        if (dyn_candidate(mt, i, n, cell_offsets, nycells, nxcells, ioff))
            flags[ioff] = 1u;
    }
}

// After the exclusive scan over ncg + 1 flags, cell ioff is new iff pos[ioff+1] != pos[ioff].
__global__ void __launch_bounds__(DM_THREADS)
dyn_grid_assign_kernel(int64_t *__restrict__ cell_offsets, int64_t ncg, const uint32_t *__restrict__ pos,
                       const uint32_t *__restrict__ global_ncells)
{
    const uint32_t base = *global_ncells;
    for (int64_t i = int64_t(blockIdx.x) * DM_THREADS + threadIdx.x; i < ncg; i += int64_t(gridDim.x) * DM_THREADS) {
        if (pos[i + 1] != pos[i])
            cell_offsets[i] = int64_t(base + pos[i]) * int64_t(3 * CELL_NPIX);
    }
}

// The variant's scratch: flags for the largest grid, 1024 x 1024 cells plus the total, and the
// scan's tile sums. Each GPU holds its own copy for the life of the process; every call on a
// GPU uses that GPU's copy.
constexpr int64_t NOALLOC_MAX_CELLS = int64_t(1024) * 1024;
constexpr int64_t NOALLOC_SCRATCH_NBYTES = int64_t(4) * 1024 * 1024 + 4096;
static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_NBYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex noalloc_mutex;
void *noalloc_ptr[NOALLOC_MAX_DEVICES];

// Device address of the current device's copy of 'scratch', looked up once per device (the
// array's name is not a device address in host code).
void *noalloc_scratch_ptr()
{
    int dev = -1;
    GMM_HIP(hipGetDevice(&dev));
    GMM_REQUIRE(dev >= 0 && dev < NOALLOC_MAX_DEVICES, "expand_dynamic_map_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(noalloc_mutex);
    if (!noalloc_ptr[dev]) {
        void *p = nullptr;
        GMM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        GMM_REQUIRE(p != nullptr, "expand_dynamic_map_noalloc: no device address for the scratch array");
        noalloc_ptr[dev] = p;
    }
    return noalloc_ptr[dev];
}

}  // namespace

int64_t expand_dynamic_noalloc_max_cells()
{
    return NOALLOC_MAX_CELLS;
}

void launch_expand_dynamic_noalloc(uint32_t *global_ncells, int64_t *cell_offsets, int64_t nycells, int64_t nxcells,
                                   const uint64_t *plan_mt, int64_t nmt, hipStream_t stream)
{
    const int64_t ncg = nycells * nxcells;
    GMM_REQUIRE(nycells > 0 && nxcells > 0 && ncg <= NOALLOC_MAX_CELLS,
                "expand_dynamic_map_noalloc: the cell grid exceeds 1024 x 1024 cells");
    const int64_t flags_nbytes = align128(4 * (ncg + 1));
    GMM_REQUIRE(flags_nbytes + scan_scratch_nbytes(ncg + 1) <= NOALLOC_SCRATCH_NBYTES,
                "expand_dynamic_map_noalloc: scratch layout exceeds the static array");
    uint8_t *base = static_cast<uint8_t *>(noalloc_scratch_ptr());
    uint32_t *flags = reinterpret_cast<uint32_t *>(base);
    const int ggrid = int(std::min<int64_t>((ncg + DM_THREADS) / DM_THREADS, 4096));
    // The array keeps the previous call's contents: clear the flags (a kernel, not a memset).
    dyn_grid_clear_kernel<<<ggrid, DM_THREADS, 0, stream>>>(flags, ncg + 1);
    GMM_HIP(hipGetLastError());
    if (nmt > 0) {
        const int mgrid = int(std::min<int64_t>((nmt + DM_THREADS - 1) / DM_THREADS, 4096));
        dyn_grid_mark_kernel<<<mgrid, DM_THREADS, 0, stream>>>(plan_mt, nmt, cell_offsets, uint32_t(nycells),
                                                               uint32_t(nxcells), flags);
        GMM_HIP(hipGetLastError());
    }
    exclusive_scan_u32(flags, ncg + 1, base + flags_nbytes, stream);
    dyn_grid_assign_kernel<<<ggrid, DM_THREADS, 0, stream>>>(cell_offsets, ncg, flags, global_ncells);
    GMM_HIP(hipGetLastError());
    dyn_count_kernel<<<1, 1, 0, stream>>>(flags, ncg, global_ncells);
    GMM_HIP(hipGetLastError());
}

int64_t expand_dynamic_scratch_nbytes(int64_t nmt)
{
    return align128(4 * (nmt + 1)) + scan_scratch_nbytes(nmt + 1);
}

void launch_expand_dynamic(uint32_t *global_ncells, int64_t *cell_offsets, int64_t nycells, int64_t nxcells,
                           const uint64_t *plan_mt, int64_t nmt, void *work, hipStream_t stream)
{
    if (nmt <= 0)
        return;
    uint32_t *pos = static_cast<uint32_t *>(work);
    const int grid = int(std::min<int64_t>((nmt + DM_THREADS) / DM_THREADS, 4096));
    dyn_flag_kernel<<<grid, DM_THREADS, 0, stream>>>(plan_mt, nmt, cell_offsets, uint32_t(nycells), uint32_t(nxcells),
                                                     pos);
    GMM_HIP(hipGetLastError());
    exclusive_scan_u32(pos, nmt + 1, static_cast<uint8_t *>(work) + align128(4 * (nmt + 1)), stream);
    dyn_assign_kernel<<<grid, DM_THREADS, 0, stream>>>(plan_mt, nmt, cell_offsets, uint32_t(nxcells), pos,
                                                       global_ncells);
    GMM_HIP(hipGetLastError());
    dyn_count_kernel<<<1, 1, 0, stream>>>(pos, nmt, global_ncells);
    GMM_HIP(hipGetLastError());
}

}  // namespace gmm
