/*
 * oracle: CPU restatement of kmsmith137/gpu_mm's pointing operators and planner.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load this library, and only as the checker or as the CPU baseline; the product
 * (gpu_mm_amd/) never links, imports or calls it.
 *
 * Written independently of gpu_mm_amd/csrc (no shared code), in plain C, compiled with
 * -ffp-contract=off so every float operation is a separately rounded IEEE operation.
 *
 * How each function is pinned (tests/test_oracle.py):
 *   orc_locate_*       hand-computed special cases (integer points, clamped edges, periodic
 *                      wrap at -nx, 0, nx, 2nx) and brute-force agreement with a numpy
 *                      restatement on random points.
 *   orc_map2tod_*      closed forms: constant map -> a + b cos2a + c sin2a; a map affine in
 *                      (y, x) is reproduced exactly; integer pointing reads single pixels;
 *                      FP32 agrees with the FP64 evaluation to ~1e-6.
 *   orc_tod2map_*      adjoint identity <A m, d> = <m, A^T d> against orc_map2tod; unit
 *                      deposits at integer pointing; sum rules (t=1, alpha=0 -> sum T = nsamp).
 *   orc_plan_*         brute-force set-based planner written in Python (tests), invariants
 *                      (nmt = number of distinct (line, cell) pairs, mflag <=> line has > 1
 *                      cells, exactly one zflag per multi-cell line, sorted by cell then line).
 *   orc_lpix_check     the reference's acceptance/rejection rules, table-driven.
 * No reference golden vectors exist (the reference's tests are unseeded same-run GPU-vs-CPU
 * comparisons and its ACT pointing files are not in the repository).
 */

#ifndef GMM_ORACLE_H
#define GMM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error bits (include/gpu_mm_internals.hpp:26-29). */
#define ORC_BAD_Y 1u
#define ORC_BAD_X 2u
#define ORC_NOT_IN_PIX 8u

/* Pixel location, include/gpu_mm_internals.hpp:124-164. out[0..3] = iy0, iy1, ix0, ix1;
 * d[0..1] = dy, dx. Returns the error bits. */
uint32_t orc_locate_f32(float y, float x, int ny, int nx, int periodic, int out[4], float d[2]);
uint32_t orc_locate_f64(double y, double x, int ny, int nx, int periodic, int out[4], double d[2]);

/* Local map description (gpu_mm/gpu_mm.py:117-124). */
typedef struct {
    int ny, nx, periodic;
    const int64_t *cell_offsets; /* [nycells][nxcells], negative = absent */
    int nycells, nxcells;
    int64_t ystride, polstride;
} orc_pix;

/* reference_map2tod / reference_tod2map (src_lib/map2tod_reference.cu:16-53,
 * src_lib/tod2map_reference.cu:14-46) in FP32 with the reference's operation order.
 * Return the error bits; ORC_NOT_IN_PIX only when !partial. */
uint32_t orc_map2tod_f32(float *tod, const float *lmap, const float *xp, int64_t nsamp,
                         const orc_pix *p, int partial, int accum);
uint32_t orc_tod2map_f32(float *lmap, const float *tod, const float *xp, int64_t nsamp,
                         const orc_pix *p, int partial);

/* The same operators evaluated in FP64 from the FP32 inputs (map and TOD in double). */
uint32_t orc_map2tod_f64(double *tod, const double *lmap, const float *xp, int64_t nsamp,
                         const orc_pix *p, int partial, int accum);
uint32_t orc_tod2map_f64(double *lmap, const double *tod, const float *xp, int64_t nsamp,
                         const orc_pix *p, int partial);

/* response_map2tod / response_tod2map (src_lib/map2tod_response.cu:81-198,
 * src_lib/tod2map_response.cu:20-160) in FP64: resp[0..ndet) scales T, resp[ndet..2 ndet)
 * scales Q and U, for the samples of detector s / (nsamp/ndet). */
uint32_t orc_response_map2tod_f64(double *tod, const double *lmap, const float *xp, const float *resp,
                                  int64_t ndet, int64_t nsamp, const orc_pix *p, int partial, int accum);
uint32_t orc_response_tod2map_f64(double *lmap, const double *tod, const float *xp, const float *resp,
                                  int64_t ndet, int64_t nsamp, const orc_pix *p, int partial);

/* Multi-threaded FP32 variants (OpenMP; tod2map with per-thread private maps then a
 * reduction), used as the N-core CPU baseline. nthreads <= 0 uses all cores. */
uint32_t orc_map2tod_f32_mt(float *tod, const float *lmap, const float *xp, int64_t nsamp,
                            const orc_pix *p, int partial, int accum, int nthreads);
uint32_t orc_tod2map_f32_mt(float *lmap, int64_t lmap_size, const float *tod, const float *xp,
                            int64_t nsamp, const orc_pix *p, int partial, int nthreads);
int orc_max_threads(void);

/* Planner launch formulas (src_lib/PointingPrePlan.cu:162-168, 236-238). */
void orc_preplan_params(int64_t nsamp, int64_t *ncl_per_block, int64_t *nblocks);
void orc_pointing_params(int64_t nmt, int64_t *nmt_per_block, int64_t *nblocks);

/* Brute-force planner. nmt_cumsum[nblocks] gets the inclusive per-planner-block counts.
 * Returns the error bits; *nmt_out = total entries. */
uint32_t orc_plan_count(const float *xp, int64_t nsamp, int ny, int nx, int periodic,
                        uint32_t *nmt_cumsum, int64_t *nmt_out);

/* Canonical plan: entries icell | line << 20 | mflag << 46 | zflag << 47, sorted by cell then
 * line; mflag on every entry of a line touching > 1 cell; zflag on the one entry of such a line
 * whose cell holds the (iy0, ix0) corner of the line's first sample. mt must hold nmt entries. */
uint32_t orc_plan_build(const float *xp, int64_t nsamp, int ny, int nx, int periodic,
                        uint64_t *mt, int64_t nmt);

/* LocalPixelization checks (src_lib/LocalPixelization.cu:33-101). Returns 0 if valid (and
 * sets *npix), else a nonzero code: 1 bad npix_global, 2 bad cell-array shape, 3 bad ystride,
 * 4 bad polstride, 5 stride relation, 6 offset not a multiple of 64, 7 cells not contiguous. */
int orc_lpix_check(int64_t ny, int64_t nx, const int64_t *offsets, int64_t nycells,
                   int64_t nxcells, int64_t ystride, int64_t polstride, int64_t *npix);

#ifdef __cplusplus
}
#endif

#endif
