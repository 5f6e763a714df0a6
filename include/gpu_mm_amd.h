/*
 * gpu_mm_amd: planned pointing operators (map2tod / tod2map) for AMD MI355X (gfx950).
 *
 * C ABI of libgpu_mm_amd.so. Plain pointers and sizes only; every array is float32 unless
 * stated. "gpu" pointers are device memory, "cpu" pointers are host memory. Streams are
 * hipStream_t passed as void* (NULL = the default stream).
 *
 * Every function that can fail returns 0 on success and a nonzero status otherwise; the
 * message is then available from gmm_last_error() (thread-local). Messages for pointing
 * errors follow the reference's wording (src_lib/check_arguments.cu:159-175).
 *
 * Array conventions (reference gpu_mm/gpu_mm.py:52-84, 117-124):
 *   tod        float[nsamp]            (a (ndet,nt) array flattened); any nsamp >= 1 (the
 *                                      reference asserts nsamp % 32 == 0, check_arguments.cu:24)
 *   xpointing  float[3][nsamp]         rows {y, x, alpha}
 *   local map  float[3*npix]           element (p,iy,ix) of cell (cy,cx) is at
 *              cell_offsets[cy,cx] + p*polstride + iy*ystride + ix
 *
 * Each declaration below names the reference interface it replaces (file:line in
 * kmsmith137/gpu_mm).
 */

#ifndef GPU_MM_AMD_H
#define GPU_MM_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GMM_PREPLAN_SIZE 1024

/* Error bits reported by the operators (include/gpu_mm_internals.hpp:26-29). */
#define GMM_ERR_BAD_YPIX 0x1u
#define GMM_ERR_BAD_XPIX 0x2u
#define GMM_ERR_INCONSISTENT_NMT 0x4u
#define GMM_ERR_NOT_IN_PIXELIZATION 0x8u

typedef struct gmm_lpix gmm_lpix;
typedef struct gmm_preplan gmm_preplan;
typedef struct gmm_plan gmm_plan;

/* Thread-local message for the last failed call on this thread. */
const char *gmm_last_error(void);

/* Library version (major*10000 + minor*100 + patch). */
int gmm_version(void);

/* ---------------------------------------------------------------------------------------
 * LocalPixelization
 * Replaces LocalPixelization (include/gpu_mm.hpp:17-58; src_lib/LocalPixelization.cu:33-114;
 * bound at src_pybind11/gpu_mm_pybind11.cu:77-99). The handle keeps the two caller-owned
 * cell_offsets pointers (int64[nycells][nxcells], one host copy, one device copy); the caller
 * keeps both alive and identical.
 * --------------------------------------------------------------------------------------- */

typedef struct {
    int64_t nypix_global, nxpix_global;
    int32_t periodic_xcoord;
    int64_t ystride, polstride;
    int64_t nycells, nxcells;
    int64_t npix; /* local pixels, without the factor 3 for T/Q/U */
} gmm_lpix_info;

/* LocalPixelization.cu:33-71 (constructor and its stride/offset checks). */
int gmm_lpix_create(gmm_lpix **out, int64_t nypix_global, int64_t nxpix_global,
                    const int64_t *cell_offsets_cpu, const int64_t *cell_offsets_gpu,
                    int64_t nycells, int64_t nxcells, int64_t ystride, int64_t polstride,
                    int periodic_xcoord);
void gmm_lpix_destroy(gmm_lpix *lpix);
int gmm_lpix_get_info(const gmm_lpix *lpix, gmm_lpix_info *out);
/* gpu_mm_pybind11.cu:94 (npix is read/write). */
int gmm_lpix_set_npix(gmm_lpix *lpix, int64_t npix);
/* LocalPixelization.cu:104-114. */
int gmm_lpix_copy_gpu_offsets_to_cpu(gmm_lpix *lpix);
int gmm_lpix_copy_cpu_offsets_to_gpu(gmm_lpix *lpix);

/* ---------------------------------------------------------------------------------------
 * PointingPrePlan
 * Replaces PointingPrePlan (include/gpu_mm.hpp:61-108; src_lib/PointingPrePlan.cu:129-280;
 * bound at gpu_mm_pybind11.cu:102-130). nmt_gpu and err_gpu are caller-owned device arrays of
 * GMM_PREPLAN_SIZE uint32; the preplan keeps nmt_gpu (it holds nmt_cumsum) and uses err_gpu
 * only during construction. The xpointing array is not retained. Blocks on the stream once.
 * --------------------------------------------------------------------------------------- */

typedef struct {
    int64_t nsamp, nypix_global, nxpix_global;
    int32_t periodic_xcoord;
    int64_t plan_nbytes;                 /* minimum size of the plan's 'buf' */
    int64_t plan_constructor_tmp_nbytes; /* minimum size of the plan's 'tmp_buf' */
    double overhead;                     /* 32*nmt/nsamp - 1 */
    int64_t ncl_per_threadblock, planner_nblocks;
    int64_t nmt_per_threadblock, pointing_nblocks; /* reference launch formulas, informational */
    int64_t plan_nmt;
    int64_t cub_nbytes;                  /* bytes of radix-sort scratch inside tmp_buf */
    int64_t max_nchunks;                 /* upper bound on operator work units */
} gmm_preplan_info;

int gmm_preplan_create(gmm_preplan **out, const float *xpointing_gpu, int64_t nsamp,
                       int64_t nypix_global, int64_t nxpix_global,
                       uint32_t *nmt_gpu, uint32_t *err_gpu,
                       int periodic_xcoord, int debug, void *stream);
void gmm_preplan_destroy(gmm_preplan *pp);
int gmm_preplan_get_info(const gmm_preplan *pp, gmm_preplan_info *out);
/* PointingPrePlan.cu:254-259: copies nmt_cumsum[0:planner_nblocks] to host memory. */
int gmm_preplan_get_nmt_cumsum(const gmm_preplan *pp, uint32_t *out_cpu, int64_t n);

/* ---------------------------------------------------------------------------------------
 * PointingPlan
 * Replaces PointingPlan (include/gpu_mm.hpp:111-152; src_lib/PointingPlan.cu:237-362;
 * bound at gpu_mm_pybind11.cu:133-153). 'buf' (>= plan_nbytes, 128-byte aligned) is kept by
 * the plan and must not be shared; 'tmp_buf' (>= plan_constructor_tmp_nbytes) is used only
 * during construction. Layout of buf: uint64 plan_mt[nmt] at offset 0, then uint32 err[...]
 * at align128(8*nmt), then the operator work list. Blocks on the stream once. debug=1 also
 * checks the plan's invariants on the device (gmm_plan_check_invariants) and fails if one is
 * broken (PointingPlan.cu:272-283 launches its Debug planner kernel instead).
 * --------------------------------------------------------------------------------------- */

typedef struct {
    int64_t nsamp, nypix_global, nxpix_global;
    int32_t periodic_xcoord;
    int64_t plan_nmt;
    int64_t nchunks;       /* operator work units (blocks per map2tod/tod2map launch) */
    int64_t nmlines;       /* number of 32-sample lines that touch more than one cell */
} gmm_plan_info;

int gmm_plan_create(gmm_plan **out, const gmm_preplan *pp, const float *xpointing_gpu,
                    int64_t nsamp, void *buf_gpu, int64_t buf_nbytes,
                    void *tmp_buf_gpu, int64_t tmp_nbytes, int debug, void *stream);
void gmm_plan_destroy(gmm_plan *plan);
int gmm_plan_get_info(const gmm_plan *plan, gmm_plan_info *out);
/* PointingPlan.cu:319-327: copies plan_mt (sorted by cell) to 'out' (host or device). */
int gmm_plan_get_plan_mt(const gmm_plan *plan, uint64_t *out, int64_t n, int out_on_gpu,
                         void *stream);
/* PointingPlan.cu:331-335: synchronous check of the plan's error flags. */
int gmm_plan_check_errflags(const gmm_plan *plan, const char *where);
/* The plan invariants that a debug=1 gmm_plan_create checks after construction, on demand:
 * entries in (cell, line) order with each pair once, line indices in range, every line present,
 * the mflag/zflag rule, and a work list covering every entry once within one cell per chunk.
 * Returns nonzero with a message counting each kind of violation. The reference asserts its
 * iterator and enumerator invariants on the device under Debug (plan_iterator.hpp:64-68, 91-103,
 * 137-141; gpu_mm_internals.hpp:224-235); this is the same role as a check that never traps. */
int gmm_plan_check_invariants(const gmm_plan *plan, void *stream);

/* ---------------------------------------------------------------------------------------
 * Operators. map2tod: tod (=, or += when accum) A * map. tod2map: map += A^T * tod.
 * 'debug' makes the call check error flags synchronously (as the reference always does);
 * with check_errors == 0 and debug == 0 the call is asynchronous and errors are only
 * visible through gmm_plan_check_errflags(). With debug = 1 the planned float32 operators also
 * run an instrumented build of their kernel and fail unless every plan entry was walked by
 * exactly one wave, every sample evaluated by exactly one block (the ownership and halo rules)
 * and every halo-tile access stayed inside the tile (the reference's Debug asserts, map2tod.cu:
 * 105-109, tod2map.cu:134-139, plan_iterator.hpp:64-141). 16-bit TODs are not instrumented.
 * --------------------------------------------------------------------------------------- */

/* launch_planned_map2tod, include/gpu_mm.hpp:155-165; src_lib/map2tod.cu:194-272. */
int gmm_planned_map2tod(float *tod_gpu, int64_t tod_nsamp,
                        const float *local_map_gpu, int64_t local_map_size,
                        const float *xpointing_gpu, int64_t xpointing_nsamp,
                        const gmm_lpix *lpix, const gmm_plan *plan,
                        int partial_pixelization, int debug, int accum, int check_errors,
                        void *stream);

/* launch_planned_tod2map, include/gpu_mm.hpp:168-177; src_lib/tod2map.cu:157-226. */
int gmm_planned_tod2map(float *local_map_gpu, int64_t local_map_size,
                        const float *tod_gpu, int64_t tod_nsamp,
                        const float *xpointing_gpu, int64_t xpointing_nsamp,
                        const gmm_lpix *lpix, const gmm_plan *plan,
                        int partial_pixelization, int debug, int check_errors, void *stream);

/* launch_response_map2tod, include/gpu_mm.hpp:179-190; src_lib/map2tod_response.cu:205-280.
 * response_gpu: float[2][ndet] (row 0 scales T, row 1 scales Q and U); sample s belongs to
 * detector s / (nsamp/ndet); ndet must divide nsamp. Otherwise as gmm_planned_map2tod. */
int gmm_response_map2tod(float *tod_gpu, int64_t tod_nsamp,
                         const float *local_map_gpu, int64_t local_map_size,
                         const float *xpointing_gpu, int64_t xpointing_nsamp,
                         const float *response_gpu, int64_t ndet,
                         const gmm_lpix *lpix, const gmm_plan *plan,
                         int partial_pixelization, int debug, int accum, int check_errors,
                         void *stream);

/* launch_response_tod2map, include/gpu_mm.hpp:192-202; src_lib/tod2map_response.cu:165-240.
 * Deposits response[0][det]*d into T and response[1][det]*d*{cos,sin}(2 alpha) into Q, U. */
int gmm_response_tod2map(float *local_map_gpu, int64_t local_map_size,
                         const float *tod_gpu, int64_t tod_nsamp,
                         const float *xpointing_gpu, int64_t xpointing_nsamp,
                         const float *response_gpu, int64_t ndet,
                         const gmm_lpix *lpix, const gmm_plan *plan,
                         int partial_pixelization, int debug, int check_errors, void *stream);

/* ---- Dynamic maps --------------------------------------------------------------------------
 * expand_dynamic_map, include/gpu_mm.hpp:268-276; src_lib/expand_dynamic_map.cu:107-141.
 * cell_offsets_gpu: int64[nycells][nxcells] with -1 = targeted cell not yet present, -2 =
 * untargeted, 3*64*64*i = the i-th cell present. Every targeted absent cell named by plan_mt
 * (device, sorted by its low 20 bits) becomes present, with indices global_ncells, +1, ... in
 * increasing cell-key order (the reference's order is nondeterministic). *global_ncells_gpu is
 * advanced; the new count is also returned in *new_ncells (host). Synchronizes the stream. */
int gmm_expand_dynamic_map(uint32_t *global_ncells_gpu, int64_t *cell_offsets_gpu, int64_t nycells, int64_t nxcells,
                           const uint64_t *plan_mt_gpu, int64_t nmt, uint32_t *new_ncells, void *stream);

/* No counterpart in the reference: gmm_expand_dynamic_map without any allocation, for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same arguments and the same result (cell indices in
 * cell-key order). Its scratch is a static device array of the library (about 4 MiB per GPU,
 * held until the process ends) covering the largest cell grid, 1024 x 1024 cells; a larger grid
 * is rejected. Calls that overlap on one GPU (several streams or threads) race on that array and
 * must be serialized by the caller. With new_ncells NULL the call only enqueues work on the
 * stream (no copy, no synchronization); otherwise it copies the new count to *new_ncells and
 * synchronizes the stream. */
int gmm_expand_dynamic_map_noalloc(uint32_t *global_ncells_gpu, int64_t *cell_offsets_gpu, int64_t nycells,
                                   int64_t nxcells, const uint64_t *plan_mt_gpu, int64_t nmt,
                                   uint32_t *new_ncells, void *stream);

/* expand_dynamic_map2, include/gpu_mm.hpp:278-283; src_lib/expand_dynamic_map.cu:144-160: the
 * same over a plan's entries, into the LocalPixelization's GPU cell offsets (the full
 * (nycells, nxcells) table). The caller updates npix and the CPU copy, as in the reference. */
int gmm_lpix_expand_dynamic(gmm_lpix *lpix, const gmm_plan *plan, uint32_t *global_ncells_gpu, uint32_t *new_ncells,
                            void *stream);

/* ---- Cell utilities --------------------------------------------------------------------------
 * Cells are float[3][64][64] blocks (12288 floats) in cell-list arrays; all pointers are device
 * pointers, 16-byte aligned; index_map_gpu is int64 on the device. Both raise an error for an
 * index outside the source/destination range. */

/* cell_broadcast, include/gpu_mm.hpp:251-257; src_lib/cell_broadcast.cu:19-36 (a host loop
 * there): dst cell i = src cell index_map[i], i < ncells. */
int gmm_cell_broadcast(float *dst_gpu, int64_t ncells, const float *src_gpu, int64_t nsrc_cells,
                       const int64_t *index_map_gpu, void *stream);

/* cell_reduce, include/gpu_mm.hpp:259-265; src_lib/cell_reduce.cu:19-44 (a host loop there):
 * dst cell d = sum over i < nsrc_cells with index_map[i] == d of src cell i (zero if none), summed
 * in increasing i, so the result is deterministic. */
int gmm_cell_reduce(float *dst_gpu, int64_t ndst_cells, const float *src_gpu, int64_t nsrc_cells,
                    const int64_t *index_map_gpu, void *stream);

/* local_map_to_global, include/gpu_mm.hpp:286-291; src_lib/local_map_to_global.cu:11-45 (host
 * there): dst_gpu float[3][nypix_global][nxpix_global] from the local map (3*npix floats) on the
 * device; pixels of absent cells are zero. */
int gmm_local_map_to_global(const gmm_lpix *lpix, float *dst_gpu, const float *src_gpu, int64_t src_size,
                            void *stream);

/* ---- Kernel timer -------------------------------------------------------------------------
 * No counterpart in the reference; a measurement aid. While enabled, every planned operator
 * records a HIP event pair on its launch stream around its main kernel (planned_map2tod_kernel
 * = GMM_TIMED_MAP2TOD, planned_tod2map_kernel = GMM_TIMED_TOD2MAP). enable() resets the counts;
 * read() waits for the recorded events and returns the summed milliseconds and the number of
 * launches since the last enable/read, then resets that kernel's count. */
enum { GMM_TIMED_MAP2TOD = 0, GMM_TIMED_TOD2MAP = 1 };
int gmm_kernel_timer_enable(int enable);
int gmm_kernel_timer_read(int which, double *total_ms, int64_t *count);

/* ---- Streaming-read probe -----------------------------------------------------------------
 * No counterpart in the reference; a measurement aid for the benchmark's bandwidth ceilings.
 * Reads src_gpu[0..n) once with non-temporal 16-byte loads (the access of the planned operators'
 * pointing and TOD streams) and adds the sum of the values to *sum_gpu (a device float; atomic,
 * so the order of the additions varies). n % 4 == 0, src_gpu 16-byte aligned. Asynchronous on
 * 'stream'. */
int gmm_stream_read(const float *src_gpu, int64_t n, float *sum_gpu, void *stream);

/* ---- 16-bit TODs -------------------------------------------------------------------------
 * No counterpart in the reference (which stores TODs as float32 only). A 16-bit TOD of nsamp
 * samples in tod_nrows rows holds value / tod_scale[row] as FP16 or BF16, with one FP32 scale
 * per row; the operators compute in FP32 exactly as the float32 ones do. tod_gpu must be 8-byte
 * aligned and nsamp divisible by 4. response_gpu may be null, or float[2][resp_ndet] as in gmm_response_map2tod. */
enum { GMM_TOD_F16 = 1, GMM_TOD_BF16 = 2 };

/* map2tod into a 16-bit TOD: each output is rounded once to the 16-bit type, except on lines
 * that touch more than one cell, whose per-cell partials are rounded and then added with packed
 * 16-bit atomics. */
int gmm_planned_map2tod_q16(void *tod_gpu, int tod_type, const float *tod_scale_gpu, int64_t tod_nrows,
                            int64_t tod_nsamp, const float *local_map_gpu, int64_t local_map_size,
                            const float *xpointing_gpu, int64_t xpointing_nsamp,
                            const float *response_gpu, int64_t resp_ndet,
                            const gmm_lpix *lpix, const gmm_plan *plan,
                            int partial_pixelization, int debug, int accum, int check_errors, void *stream);

/* tod2map from a 16-bit TOD (samples decoded as value * scale in FP32, accumulated exactly). */
int gmm_planned_tod2map_q16(float *local_map_gpu, int64_t local_map_size,
                            const void *tod_gpu, int tod_type, const float *tod_scale_gpu, int64_t tod_nrows,
                            int64_t tod_nsamp, const float *xpointing_gpu, int64_t xpointing_nsamp,
                            const float *response_gpu, int64_t resp_ndet,
                            const gmm_lpix *lpix, const gmm_plan *plan,
                            int partial_pixelization, int debug, int check_errors, void *stream);

/* Encodes a float32 TOD [nrows][row_len] (device) into out_gpu (16-bit, same shape) and
 * scale_gpu (float[nrows]): scale = max(max|row| / (0.5 * largest finite value), FLT_MIN),
 * out = tod / scale rounded to nearest even. error_feedback 1 or 2 carries each sample's rounding
 * error (first or second order) into the next sample of its row before rounding, which shapes
 * the encoding error towards high frequencies, where tod2map averages it out; 0 rounds each
 * sample independently. No counterpart in the reference (float32 TODs only). */
int gmm_tod16_encode(void *out_gpu, int tod_type, float *scale_gpu, const float *tod_gpu, int64_t nrows,
                     int64_t row_len, int error_feedback, void *stream);

/* ---- Batched operators -------------------------------------------------------------------
 * No counterpart in the reference, which launches one TOD at a time (map2tod.cu:198-272,
 * tod2map.cu:160-226). A batch binds several TODs, each with its own plan and xpointing, to one
 * LocalPixelization; one kernel launch then runs map2tod (or tod2map) for all of them. The work
 * items are runs of about item_entries plan entries of one cell gathered across the TODs, so
 * the cell's halo tile is staged (map2tod) or accumulated and flushed (tod2map) once for all the
 * TODs that touch it. With blocks_per_cu > 0 the launch is a persistent grid of that many blocks
 * per CU taking items from a device queue, largest first; with 0 it has one block per item.
 * Results equal those of the plain per-TOD calls (map2tod bit for bit; tod2map up to the
 * order of the float additions into the map). The plans, xpointing, TOD and scale arrays must
 * outlive the batch; TODs and xpointing must be 16-byte aligned (8 bytes for 16-bit TODs).
 * Calls on one batch must not overlap in time (they share its queue counters). */
typedef struct gmm_batch gmm_batch;
typedef struct gmm_graph gmm_graph;

enum { GMM_TOD_F32 = 0 };
enum { GMM_OP_MAP2TOD = 0, GMM_OP_TOD2MAP = 1 };

typedef struct {
    int32_t m2t_item_entries;    /* plan entries per map2tod work item (64 .. 2^24) */
    int32_t m2t_threads;         /* map2tod threads per block: 256 or 512 */
    int32_t m2t_blocks_per_cu;   /* map2tod persistent grid, blocks per CU (0 = one block per item) */
    int32_t t2m_item_entries;    /* plan entries per tod2map work item */
    int32_t t2m_blocks_per_cu;   /* tod2map persistent grid, blocks per CU (0 = one block per item) */
    int32_t reserved[3];
} gmm_batch_config;

typedef struct {
    int32_t ntod, tod_type;
    gmm_batch_config config;
    int64_t nentries;            /* plan entries over all TODs */
    int64_t ncells;              /* distinct cells */
    int64_t m2t_nitems, m2t_nsegs, m2t_grid;
    int64_t t2m_nitems, t2m_nsegs, t2m_grid;
    int64_t generation;          /* incremented by every gmm_batch_configure() */
} gmm_batch_info;

typedef struct {
    int32_t op;                  /* GMM_OP_MAP2TOD or GMM_OP_TOD2MAP */
    int32_t item_entries, threads, blocks_per_cu;
    double ms;                   /* median kernel time */
} gmm_batch_tune_entry;

void gmm_batch_default_config(gmm_batch_config *cfg);

/* tod_type: GMM_TOD_F32, GMM_TOD_F16 or GMM_TOD_BF16 (then tod_scale_gpu[i] is float[tod_nrows[i]],
 * otherwise both may be null). nsamp[i] is the sample count of TOD i (its plan's nsamp), which
 * must be divisible by 4. */
int gmm_batch_create(gmm_batch **out, const gmm_lpix *lpix, int32_t ntod, const gmm_plan *const *plans,
                     const float *const *xpointing_gpu, void *const *tod_gpu, const int64_t *nsamp,
                     int32_t tod_type, const float *const *tod_scale_gpu, const int64_t *tod_nrows,
                     const gmm_batch_config *config);

/* No counterpart in the reference. Per-detector gains for every TOD of the batch, as in
 * gmm_response_map2tod: response_gpu[i] is float[2][resp_ndet[i]] (row 0 multiplies T, row 1 Q
 * and U) and resp_ndet[i] divides TOD i's nsamp; the batched operators then compute
 * response_map2tod / response_tod2map per TOD. response_gpu == NULL removes the gains. Float32
 * TODs only. The arrays must outlive the batch's use of them; graphs captured before the call
 * must be captured again. */
int gmm_batch_set_response(gmm_batch *b, const float *const *response_gpu, const int64_t *resp_ndet);
void gmm_batch_destroy(gmm_batch *b);
/* Rebuilds the work items for a new configuration (invalidates graphs captured before). */
int gmm_batch_configure(gmm_batch *b, const gmm_batch_config *config);
int gmm_batch_get_info(const gmm_batch *b, gmm_batch_info *out);

/* map2tod into every bound TOD (overwrite, or add with accum) / tod2map from every bound TOD
 * into local_map (always accumulates). check_errors: synchronize and raise on error flags. */
int gmm_batch_map2tod(gmm_batch *b, const float *local_map_gpu, int64_t local_map_size, int partial_pixelization,
                      int accum, int check_errors, void *stream);
int gmm_batch_tod2map(gmm_batch *b, float *local_map_gpu, int64_t local_map_size, int partial_pixelization,
                      int check_errors, void *stream);
/* Synchronous check (and reset) of the error flags accumulated by the batch's launches. */
int gmm_batch_check_errflags(gmm_batch *b, const char *where);

/* Autotuner: times each candidate configuration of each operator (reps launches, median; one
 * untimed launch first) and configures the batch with the fastest. map2tod is timed reading
 * map_in_gpu and writing the bound TODs (which end up holding map2tod(map_in)); tod2map is timed
 * accumulating into map_scratch_gpu. The table receives every measurement (up to capacity). */
int gmm_batch_autotune(gmm_batch *b, const float *map_in_gpu, float *map_scratch_gpu, int64_t local_map_size,
                       int32_t reps, gmm_batch_tune_entry *table, int32_t table_capacity, int32_t *table_size,
                       void *stream);

/* HIP graph of one batched operator launch (op = GMM_OP_MAP2TOD or GMM_OP_TOD2MAP) with fixed
 * arguments, captured on a private stream. The persistent grid rewinds its own queue at the end
 * of every launch, so the graph needs no reset node. The graph must not outlive the batch, and
 * is refused after the batch is reconfigured. */
int gmm_batch_graph_create(gmm_graph **out, gmm_batch *b, int32_t op, float *local_map_gpu, int64_t local_map_size,
                           int partial_pixelization, int accum);
int gmm_graph_launch(gmm_graph *g, void *stream);
void gmm_graph_destroy(gmm_graph *g);

/* launch_unplanned_map2tod, include/gpu_mm.hpp:204-213; src_lib/map2tod_unplanned.cu:74-127.
 * errflags_gpu: caller-supplied uint32[nerrflags]; the launch uses at most nerrflags blocks. */
int gmm_unplanned_map2tod(float *tod_gpu, int64_t tod_nsamp,
                          const float *local_map_gpu, int64_t local_map_size,
                          const float *xpointing_gpu, int64_t xpointing_nsamp,
                          const gmm_lpix *lpix, uint32_t *errflags_gpu, int64_t nerrflags,
                          int partial_pixelization, int accum, int check_errors, void *stream);

/* launch_unplanned_tod2map, include/gpu_mm.hpp:216-224; src_lib/tod2map_unplanned.cu:68-121. */
int gmm_unplanned_tod2map(float *local_map_gpu, int64_t local_map_size,
                          const float *tod_gpu, int64_t tod_nsamp,
                          const float *xpointing_gpu, int64_t xpointing_nsamp,
                          const gmm_lpix *lpix, uint32_t *errflags_gpu, int64_t nerrflags,
                          int partial_pixelization, int check_errors, void *stream);

/* reference_map2tod, include/gpu_mm.hpp:227-236; src_lib/map2tod_reference.cu:16-53.
 * Host arrays; single-threaded. */
int gmm_reference_map2tod(float *tod_cpu, int64_t tod_nsamp,
                          const float *local_map_cpu, int64_t local_map_size,
                          const float *xpointing_cpu, int64_t xpointing_nsamp,
                          const gmm_lpix *lpix, int partial_pixelization, int accum);

/* reference_tod2map, include/gpu_mm.hpp:239-247; src_lib/tod2map_reference.cu:14-46. */
int gmm_reference_tod2map(float *local_map_cpu, int64_t local_map_size,
                          const float *tod_cpu, int64_t tod_nsamp,
                          const float *xpointing_cpu, int64_t xpointing_nsamp,
                          const gmm_lpix *lpix, int partial_pixelization);

/* ---------------------------------------------------------------------------------------
 * Test aids.
 * --------------------------------------------------------------------------------------- */

/* ToyPointing, include/gpu_mm.hpp:314-350; src_lib/ToyPointing.cu:41-109. Fills a host
 * xpointing[3][nsamp] array with the 45-degree bouncing toy scan; returns drift_speed. */
int gmm_toy_pointing_fill(float *xpointing_cpu, int64_t nsamp, int64_t nypix_global,
                          int64_t nxpix_global, double scan_speed, double total_drift,
                          double *drift_speed_out);

/* PointingPlanTester, include/gpu_mm.hpp:388-430; src_lib/PointingPlanTester.cu:69-178.
 * Quantizes the pointing on the GPU (iypix, ixpix = (iy0, ix0) of each sample) and builds the
 * brute-force plan on the host: nmt_cumsum[planner_nblocks] and sorted_mt[plan_nmt]
 * (46-bit entries, no flags, sorted as integers). */
int64_t gmm_plan_tester_tmp_nbytes(const gmm_preplan *pp);
int gmm_plan_tester_run(const gmm_preplan *pp, const float *xpointing_gpu, int64_t nsamp,
                        void *tmp_gpu, int64_t tmp_nbytes,
                        int32_t *iypix_cpu, int32_t *ixpix_cpu,
                        uint32_t *nmt_cumsum_cpu, uint64_t *sorted_mt_cpu, void *stream);

/* test_plan_iterator, include/plan_iterator.hpp:14; src_lib/test_plan_iterator.cu:120-166.
 * Builds the operator work list for an arbitrary cell-sorted plan_mt (chunks of at most
 * nmt_per_block entries), walks it on the GPU with 'waves_per_block' waves per block and
 * checks that every entry is visited exactly once by a block working on its cell. */
int gmm_test_plan_iterator(const uint64_t *plan_mt_cpu, int64_t nmt, int64_t nmt_per_block,
                           int waves_per_block);

/* make_random_plan_mt, include/plan_iterator.hpp:15; src_lib/test_plan_iterator.cu:170-209.
 * Seeded. Call with out_cpu == NULL to get the entry count in *nmt_out first. */
int gmm_make_random_plan_mt(int64_t ncells, int64_t min_nmt_per_cell, int64_t max_nmt_per_cell,
                            uint64_t seed, uint64_t *out_cpu, int64_t capacity, int64_t *nmt_out);

#ifdef __cplusplus
}
#endif

#endif /* GPU_MM_AMD_H */
