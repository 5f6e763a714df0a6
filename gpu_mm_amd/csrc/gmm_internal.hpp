// Internal definitions of libgpu_mm_amd.so: handle structs, error plumbing, launchers.

#pragma once

#include "gmm_common.hpp"

#include <hip/hip_runtime.h>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

// Opaque C-ABI handles (declared in include/gpu_mm_amd.h).

struct gmm_lpix
{
    int64_t nypix_global = 0;
    int64_t nxpix_global = 0;
    bool periodic_xcoord = true;
    const int64_t *offsets_cpu = nullptr;  // caller-owned int64[nycells][nxcells]
    const int64_t *offsets_gpu = nullptr;  // caller-owned device copy
    int64_t nycells = 0, nxcells = 0;
    int64_t ystride = 0, polstride = 0;
    int64_t npix = 0;
};

struct gmm_preplan
{
    int64_t nsamp = 0, nypix_global = 0, nxpix_global = 0;
    bool periodic_xcoord = true;
    gmm::PreplanParams pparams{};
    uint32_t *nmt_cumsum_gpu = nullptr;   // caller-owned uint32[PREPLAN_SIZE]

    int64_t plan_nmt = 0;
    double overhead = 0.0;
    int64_t nmt_per_threadblock = 0, pointing_nblocks = 0;

    // Derived sizes (computed once in the constructor).
    int64_t nycells_g = 0, nxcells_g = 0;  // cells of the global map
    int key_bits = 0;                      // bits of the dense cell key iycell*nxcells_g + ixcell
    int64_t chunk_cap = 0;                 // max plan entries per operator work unit
    int64_t max_nchunks = 0;
    int64_t plan_nbytes = 0, tmp_nbytes = 0, sort_nbytes = 0;

    // Offsets inside the plan buffer.
    int64_t off_err = 0, off_chunks = 0, off_qmask = 0, n_err = 0;
};

struct gmm_plan
{
    gmm_preplan pp;                 // copy (small)
    uint8_t *buf = nullptr;         // caller-owned device buffer, >= pp.plan_nbytes
    uint64_t *plan_mt = nullptr;    // sorted plan entries
    uint32_t *err = nullptr;        // per-block error words (n_err of them)
    gmm::Chunk *chunks = nullptr;   // operator work list
    uint16_t *qmask = nullptr;      // per entry: the quads of its line it evaluates (+ QM_TAKES_RIGHT)
    int64_t nchunks = 0;
    int64_t nmlines = 0;
};

namespace gmm {

struct Error : public std::runtime_error
{
    explicit Error(const std::string &s) : std::runtime_error(s) { }
};

[[noreturn]] void fail(const std::string &msg);
void hip_check(hipError_t e, const char *what);

#define GMM_HIP(call) ::gmm::hip_check((call), #call)
#define GMM_REQUIRE(cond, msg) do { if (!(cond)) ::gmm::fail(std::string(msg)); } while (0)

// C-ABI plumbing (capi.cpp): every entry point runs its body through guarded(), which turns an
// exception into status 1 and the thread's last-error message.
void set_last_error(const char *msg);

template<typename F>
int guarded(F &&f)
{
    try {
        f();
        return 0;
    }
    catch (const std::exception &e) {
        set_last_error(e.what());
        return 1;
    }
    catch (...) {
        set_last_error("unknown C++ exception");
        return 1;
    }
}

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }
void require_ptr(const void *p, const char *where, const char *name);
// Synchronous read / clear of one device word on 'stream'.
uint32_t read_word(const uint32_t *p, hipStream_t stream);
void clear_word(uint32_t *p, hipStream_t stream);

// Converts a set of error bits into the exception the reference raises
// (src_lib/check_arguments.cu:159-175).
void throw_errflags(uint32_t err, const char *where, uint32_t ignore = 0);

// Argument checks (src_lib/check_arguments.cu:137-156).
void check_nsamp(int64_t nsamp, const char *where);
void check_nsamp_quads(int64_t nsamp, const char *where);
// Per-detector gains float[2][ndet] of a TOD of nsamp samples (ndet > 0 divides nsamp).
void check_response(const float *response_gpu, int64_t ndet, int64_t nsamp, const char *where);
void check_npix_global(int64_t n, const char *what, const char *where);

// ------------------------------------------------------------------------------------------
// Device scan (exclusive, in place) over uint32. Scratch size from scan_scratch_nbytes().

int64_t scan_scratch_nbytes(int64_t n);
void exclusive_scan_u32(uint32_t *data, int64_t n, void *scratch, hipStream_t stream);

// ------------------------------------------------------------------------------------------
// Planner launchers (k_planner.hip).

void launch_preplan_count(const float *xp, int64_t nsamp, const gmm_preplan &pp,
                          uint32_t *nmt_out, uint32_t *err_out, hipStream_t stream);

// Inclusive scan of nmt[0:nblocks] in place, padding to PREPLAN_SIZE with the total, and
// OR of err[0:nblocks] into err[PREPLAN_SIZE-1]... returned through 'summary_gpu' (2 words:
// total nmt, OR of error bits).
void launch_preplan_finish(uint32_t *nmt, const uint32_t *err, int64_t nblocks,
                           uint32_t *summary_gpu, hipStream_t stream);

void launch_plan_emit(const float *xp, const gmm_preplan &pp, uint64_t *out_mt,
                      uint32_t *err_out, hipStream_t stream);

// Sorts plan entries by cell (stable), ping-ponging between a and b, result in 'out'.
void plan_sort(const gmm_preplan &pp, uint64_t *a, uint64_t *b, uint64_t *out,
               void *scratch, int64_t scratch_nbytes, hipStream_t stream);

// Builds the chunk list and the multi-cell line list from the sorted plan. Writes the
// two counts to counts_gpu[0], counts_gpu[1].
void plan_schedule(const gmm_preplan &pp, const uint64_t *sorted_mt, const float *xp, Chunk *chunks,
                   uint16_t *qmask, uint32_t *counts_gpu, void *scratch,
                   int64_t scratch_nbytes, hipStream_t stream);

// Reorders the work list (nch chunks, count known on the host) largest first; a and b are the
// plan sort's ping-pong buffers, free after plan_sort().
void order_chunks(const gmm_preplan &pp, Chunk *chunks, int64_t nch, uint64_t *a, uint64_t *b, void *scratch,
                  int64_t scratch_nbytes, hipStream_t stream);

// Size of the scratch used by plan_sort() + plan_schedule().
int64_t plan_scratch_nbytes(const gmm_preplan &pp);

// Fills n words with v (a kernel, so it is safe inside captured graphs).
void launch_fill_u32(uint32_t *p, int64_t n, uint32_t v, hipStream_t stream);

// OR-reduce n words into out[0].
void launch_reduce_or(const uint32_t *in, int64_t n, uint32_t *out, hipStream_t stream);

// Quantize kernel for the plan tester (src_lib/PointingPlanTester.cu:23-63).
void launch_quantize(const float *xp, int64_t nsamp, int ny, int nx, bool periodic,
                     int32_t *iypix, int32_t *ixpix, uint32_t *err_word, hipStream_t stream);

// Work-list coverage test (src_lib/test_plan_iterator.cu:52-100 role).
void launch_iterator_test(const uint64_t *plan_mt, const Chunk *chunks, int64_t nchunks,
                          int waves_per_block, uint32_t *visit_counts, uint32_t *err_word,
                          hipStream_t stream);

// Builds chunks from an arbitrary cell-sorted plan (used by the iterator test).
int64_t build_chunks_generic(const uint64_t *plan_mt_gpu, int64_t nmt, int64_t chunk_cap,
                             Chunk *chunks_gpu, int64_t max_chunks, void *scratch,
                             int64_t scratch_nbytes, hipStream_t stream);
int64_t build_chunks_generic_scratch(int64_t nmt);

// ------------------------------------------------------------------------------------------
// Operator launchers (k_planned.hip, k_unplanned.hip).

// Device counters of a debug launch of the planned operators (k_debug.hip): visits per plan
// entry, evaluations per sample, halo-tile accesses outside the tile.
struct DebugBuffers
{
    uint32_t *entry_visits = nullptr;   // [nmt]
    uint32_t *sample_evals = nullptr;   // [nsamp]
    uint32_t *tile_bad = nullptr;       // one word
};

struct OpArgs
{
    const gmm_lpix *lpix;
    int64_t nsamp;
    bool partial;
    const DebugBuffers *debug = nullptr;   // debug=True on a float32 TOD: instrumented kernels
};

// Debug checks (k_debug.hip). check_plan_invariants() raises if the plan breaks its order, flag
// or work-list invariants. OpDebug allocates and clears the operator counters; check() raises
// if an entry was not walked exactly once, a sample not evaluated exactly once, or a tile access
// fell outside the tile. Both synchronize the stream.
void check_plan_invariants(const gmm_plan &plan, const char *where, hipStream_t stream);

class OpDebug
{
public:
    OpDebug(int64_t nmt, int64_t nsamp, hipStream_t stream);
    ~OpDebug();
    OpDebug(const OpDebug &) = delete;
    OpDebug &operator=(const OpDebug &) = delete;
    void check(const char *where);
    DebugBuffers bufs;

private:
    int64_t nmt_, nsamp_;
    hipStream_t stream_;
    void *mem_ = nullptr;
};

// Kernel timer (measurement aid): event pairs around the main planned kernels while enabled.
enum TimedKernel { TK_MAP2TOD = 0, TK_TOD2MAP = 1, TK_COUNT = 2 };
void kernel_timer_mark(int which, bool end, hipStream_t stream);

// TOD element type and per-sample gains of a planned operator call.
enum TodType { TOD_F32 = 0, TOD_F16 = 1, TOD_BF16 = 2 };

struct TodGains
{
    int tod_type = TOD_F32;
    const float *response = nullptr;   // float[2][resp_ndet] on the device, or null
    int64_t resp_ndet = 0;
    const float *scale = nullptr;      // float[scale_nrows] on the device (16-bit TODs), or null
    int64_t scale_nrows = 0;
};

void launch_planned_map2tod(void *tod, const float *lmap, const float *xp, const OpArgs &a,
                            const gmm_plan &plan, bool accum, const TodGains &g, hipStream_t stream);
void launch_planned_tod2map(float *lmap, const void *tod, const float *xp, const OpArgs &a,
                            const gmm_plan &plan, const TodGains &g, hipStream_t stream);
// Batched operators (k_planned.hip): device tables of TODs, segments and work items, the queue
// counters (2 words, zero between launches) and the error word. grid > 0: persistent grid of
// that many blocks taking items from the queue; grid == 0: one block per item. threads: block
// size of map2tod (256 or 512); tod2map always uses 1024.
struct BatchLaunch
{
    const BatchTod *tods = nullptr;
    const BatchSeg *segs = nullptr;
    const BatchItem *items = nullptr;
    int64_t nitems = 0;
    uint32_t *counters = nullptr;
    uint32_t *err_word = nullptr;
    int64_t grid = 0;
    int threads = 512;
    bool has_resp = false;      // every TOD has per-detector gains (FP32 TODs only)
};

void launch_batched_map2tod(const BatchLaunch &bl, const float *lmap, const gmm_lpix &lp, int tod_type, bool partial,
                            bool accum, hipStream_t stream);
void launch_batched_tod2map(const BatchLaunch &bl, float *lmap, const gmm_lpix &lp, int tod_type, bool partial,
                            hipStream_t stream);

// Dynamic map expansion (cells of a cell-sorted plan_mt into cell_offsets, new indices in
// cell-key order); scratch of expand_dynamic_scratch_nbytes(nmt) bytes.
int64_t expand_dynamic_scratch_nbytes(int64_t nmt);
void launch_expand_dynamic(uint32_t *global_ncells, int64_t *cell_offsets, int64_t nycells, int64_t nxcells,
                           const uint64_t *plan_mt, int64_t nmt, void *work, hipStream_t stream);
// The same result without allocating: per-cell flags over the cell grid in a static device
// array of the library (at most expand_dynamic_noalloc_max_cells() cells); calls that overlap
// on one GPU race on that array.
int64_t expand_dynamic_noalloc_max_cells();
void launch_expand_dynamic_noalloc(uint32_t *global_ncells, int64_t *cell_offsets, int64_t nycells, int64_t nxcells,
                                   const uint64_t *plan_mt, int64_t nmt, hipStream_t stream);

// Cell utilities (k_cells.hip). 'bad' (device word) is OR-ed with 1 on an out-of-range index.
void launch_cell_broadcast(float *dst, int64_t ncells, const float *src, int64_t nsrc, const int64_t *index_map,
                           uint32_t *bad, hipStream_t stream);
void launch_cell_reduce(float *dst, int64_t ndst, const float *src, int64_t nsrc, const int64_t *index_map,
                        uint32_t *bad, hipStream_t stream);
void launch_local_to_global(float *dst, const float *src, const gmm_lpix &lp, hipStream_t stream);

// Streaming-read probe (k_probe.hip, a measurement aid): reads n floats (n % 4 == 0, 16-byte
// aligned) with non-temporal 16-byte loads and adds their sum to *sum.
void launch_stream_read(const float *p, int64_t n, float *sum, hipStream_t stream);

// Per-row 16-bit encoding of a float32 TOD [nrows][row_len]: scale[r] = max(max_s |tod[r][s]| /
// (0.5 * largest finite value of the type), FLT_MIN), out[r][s] = tod[r][s] / scale[r] rounded
// to nearest even, with error feedback of order 0 (none), 1 or 2 along each row.
void launch_tod16_encode(void *out, int tod_type, float *scale, const float *tod, int64_t nrows,
                         int64_t row_len, int error_feedback, hipStream_t stream);
// Words of tod2map scratch kept in the plan buffer.

int64_t unplanned_nblocks(int64_t nsamp, int64_t max_blocks);
void launch_unplanned_map2tod(float *tod, const float *lmap, const float *xp, const OpArgs &a,
                              uint32_t *errflags, int64_t nblocks, bool accum, hipStream_t stream);
void launch_unplanned_tod2map(float *lmap, const float *tod, const float *xp, const OpArgs &a,
                              uint32_t *errflags, int64_t nblocks, hipStream_t stream);

// ------------------------------------------------------------------------------------------
// Host code (host_ref.cpp).

void reference_map2tod(float *tod, const float *lmap, const float *xp, const OpArgs &a, bool accum);
void reference_tod2map(float *lmap, const float *tod, const float *xp, const OpArgs &a);
double toy_pointing_fill(float *xp, int64_t nsamp, int64_t ny, int64_t nx, double scan_speed,
                         double total_drift);
void plan_tester_host(const gmm_preplan &pp, const int32_t *iypix, const int32_t *ixpix,
                      uint32_t *nmt_cumsum, uint64_t *sorted_mt);
std::vector<uint64_t> make_random_plan_mt(int64_t ncells, int64_t min_nmt, int64_t max_nmt,
                                          uint64_t seed);

}  // namespace gmm
