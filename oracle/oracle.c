/*
 * oracle.c -- CPU restatement of gpu_mm's pointing operators and planner.
 * TEST INFRASTRUCTURE ONLY (see oracle.h): loaded by tests/, smoke() and bench.py's
 * cpu_baseline leg as the checker; never part of the product.
 */

#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------------------------------
 * Pixel location (include/gpu_mm_internals.hpp:124-164).
 *
 * The reference truncates toward zero and then clamps below at 0; floor() followed by the same
 * clamp gives identical indices (they differ only for negative values, which clamp to 0), so
 * floor is used here. NaN clamps to 0 through fmax/fmin, matching the device's saturating
 * conversion; such samples are always flagged.
 * --------------------------------------------------------------------------------------- */

uint32_t orc_locate_f32(float y, float x, int ny, int nx, int periodic, int out[4], float d[2])
{
    uint32_t err = 0;
    if (!(y >= 0.0f && y <= (float)(ny - 1)))
        err |= ORC_BAD_Y;
    int iy0 = (int)fminf(fmaxf(floorf(y), 0.0f), (float)(ny - 2));
    out[0] = iy0;
    out[1] = iy0 + 1;
    d[0] = y - (float)iy0;

    if (periodic) {
        if (!(x >= (float)(-nx) && x <= (float)(2 * nx)))
            err |= ORC_BAD_X;
        if (!(x >= 0.0f))
            x = x + (float)nx;
        if (!(x <= (float)nx))
            x = x - (float)nx;
        int ix0 = (int)fminf(fmaxf(floorf(x), 0.0f), (float)nx);
        d[1] = x - (float)ix0;
        if (ix0 >= nx)
            ix0 = 0;
        out[2] = ix0;
        out[3] = (ix0 + 1 < nx) ? ix0 + 1 : 0;
    }
    else {
        if (!(x >= 0.0f && x <= (float)(nx - 2)))
            err |= ORC_BAD_X;
        int ix0 = (int)fminf(fmaxf(floorf(x), 0.0f), (float)(nx - 2));
        out[2] = ix0;
        out[3] = ix0 + 1;
        d[1] = x - (float)ix0;
    }
    return err;
}

uint32_t orc_locate_f64(double y, double x, int ny, int nx, int periodic, int out[4], double d[2])
{
    uint32_t err = 0;
    if (!(y >= 0.0 && y <= (double)(ny - 1)))
        err |= ORC_BAD_Y;
    int iy0 = (int)fmin(fmax(floor(y), 0.0), (double)(ny - 2));
    out[0] = iy0;
    out[1] = iy0 + 1;
    d[0] = y - (double)iy0;

    if (periodic) {
        if (!(x >= (double)(-nx) && x <= (double)(2 * nx)))
            err |= ORC_BAD_X;
        if (!(x >= 0.0))
            x = x + (double)nx;
        if (!(x <= (double)nx))
            x = x - (double)nx;
        int ix0 = (int)fmin(fmax(floor(x), 0.0), (double)nx);
        d[1] = x - (double)ix0;
        if (ix0 >= nx)
            ix0 = 0;
        out[2] = ix0;
        out[3] = (ix0 + 1 < nx) ? ix0 + 1 : 0;
    }
    else {
        if (!(x >= 0.0 && x <= (double)(nx - 2)))
            err |= ORC_BAD_X;
        int ix0 = (int)fmin(fmax(floor(x), 0.0), (double)(nx - 2));
        out[2] = ix0;
        out[3] = ix0 + 1;
        d[1] = x - (double)ix0;
    }
    return err;
}

/* Memory offset of pixel (iy, ix) of the local map's T plane, or -1 when absent
 * (gpu_mm/gpu_mm.py:117-124). */
static int64_t pix_offset(const orc_pix *p, int iy, int ix)
{
    int cy = iy / 64, cx = ix / 64;
    if (iy < 0 || ix < 0 || cy >= p->nycells || cx >= p->nxcells)
        return -1;
    int64_t base = p->cell_offsets[(int64_t)cy * p->nxcells + cx];
    if (base < 0)
        return -1;
    return base + (int64_t)(iy - 64 * cy) * p->ystride + (ix - 64 * cx);
}

/* ---------------------------------------------------------------------------------------
 * Operators, FP32, operation order of map_evaluator/map_accumulator
 * (include/gpu_mm_internals.hpp:258-292, 314-354).
 * --------------------------------------------------------------------------------------- */

static float m2t_sample_f32(const float *lmap, const float *xp, int64_t s, int64_t n,
                            const orc_pix *p, int partial, uint32_t *err)
{
    float sn, cs;
    sincosf(2.0f * xp[s + 2 * n], &sn, &cs);
    int c[4];
    float d[2];
    *err |= orc_locate_f32(xp[s], xp[s + n], p->ny, p->nx, p->periodic, c, d);
    float dy = d[0], dx = d[1];
    const int ys[4] = { c[0], c[0], c[1], c[1] };
    const int xs[4] = { c[2], c[3], c[2], c[3] };
    const float wy[4] = { 1.0f - dy, 1.0f - dy, dy, dy };
    const float wx[4] = { 1.0f - dx, dx, 1.0f - dx, dx };
    float r = 0.0f;
    for (int k = 0; k < 4; k++) {
        int64_t o = pix_offset(p, ys[k], xs[k]);
        float v = 0.0f;
        if (o >= 0) {
            v = lmap[o] + cs * lmap[o + p->polstride];
            v = v + sn * lmap[o + 2 * p->polstride];
        }
        else if (!partial)
            *err |= ORC_NOT_IN_PIX;
        float w = wy[k] * wx[k];
        r = (k == 0) ? (w * v) : (r + w * v);
    }
    return r;
}

uint32_t orc_map2tod_f32(float *tod, const float *lmap, const float *xp, int64_t nsamp,
                         const orc_pix *p, int partial, int accum)
{
    uint32_t err = 0;
    for (int64_t s = 0; s < nsamp; s++) {
        float r = m2t_sample_f32(lmap, xp, s, nsamp, p, partial, &err);
        tod[s] = accum ? (tod[s] + r) : r;
    }
    return err;
}

static void t2m_sample_f32(float *lmap, float t, const float *xp, int64_t s, int64_t n,
                           const orc_pix *p, int partial, uint32_t *err)
{
    float sn, cs;
    sincosf(2.0f * xp[s + 2 * n], &sn, &cs);
    float q = t * cs, u = t * sn;
    int c[4];
    float d[2];
    *err |= orc_locate_f32(xp[s], xp[s + n], p->ny, p->nx, p->periodic, c, d);
    float dy = d[0], dx = d[1];
    const int ys[4] = { c[0], c[0], c[1], c[1] };
    const int xs[4] = { c[2], c[3], c[2], c[3] };
    const float wy[4] = { 1.0f - dy, 1.0f - dy, dy, dy };
    const float wx[4] = { 1.0f - dx, dx, 1.0f - dx, dx };
    for (int k = 0; k < 4; k++) {
        int64_t o = pix_offset(p, ys[k], xs[k]);
        float w = wy[k] * wx[k];
        if (o >= 0) {
            lmap[o] += w * t;
            lmap[o + p->polstride] += w * q;
            lmap[o + 2 * p->polstride] += w * u;
        }
        else if (!partial)
            *err |= ORC_NOT_IN_PIX;
    }
}

uint32_t orc_tod2map_f32(float *lmap, const float *tod, const float *xp, int64_t nsamp,
                         const orc_pix *p, int partial)
{
    uint32_t err = 0;
    for (int64_t s = 0; s < nsamp; s++)
        t2m_sample_f32(lmap, tod[s], xp, s, nsamp, p, partial, &err);
    return err;
}

/* ---------------------------------------------------------------------------------------
 * FP64 evaluation.
 * --------------------------------------------------------------------------------------- */

uint32_t orc_map2tod_f64(double *tod, const double *lmap, const float *xp, int64_t nsamp,
                         const orc_pix *p, int partial, int accum)
{
    uint32_t err = 0;
    for (int64_t s = 0; s < nsamp; s++) {
        double sn, cs;
        sincos(2.0 * (double)xp[s + 2 * nsamp], &sn, &cs);
        int c[4];
        double d[2];
        err |= orc_locate_f64(xp[s], xp[s + nsamp], p->ny, p->nx, p->periodic, c, d);
        const int ys[4] = { c[0], c[0], c[1], c[1] };
        const int xs[4] = { c[2], c[3], c[2], c[3] };
        const double w[4] = { (1 - d[0]) * (1 - d[1]), (1 - d[0]) * d[1], d[0] * (1 - d[1]), d[0] * d[1] };
        double r = 0.0;
        for (int k = 0; k < 4; k++) {
            int64_t o = pix_offset(p, ys[k], xs[k]);
            if (o >= 0)
                r += w[k] * (lmap[o] + cs * lmap[o + p->polstride] + sn * lmap[o + 2 * p->polstride]);
            else if (!partial)
                err |= ORC_NOT_IN_PIX;
        }
        tod[s] = accum ? (tod[s] + r) : r;
    }
    return err;
}

uint32_t orc_tod2map_f64(double *lmap, const double *tod, const float *xp, int64_t nsamp,
                         const orc_pix *p, int partial)
{
    uint32_t err = 0;
    for (int64_t s = 0; s < nsamp; s++) {
        double sn, cs, t = tod[s];
        sincos(2.0 * (double)xp[s + 2 * nsamp], &sn, &cs);
        int c[4];
        double d[2];
        err |= orc_locate_f64(xp[s], xp[s + nsamp], p->ny, p->nx, p->periodic, c, d);
        const int ys[4] = { c[0], c[0], c[1], c[1] };
        const int xs[4] = { c[2], c[3], c[2], c[3] };
        const double w[4] = { (1 - d[0]) * (1 - d[1]), (1 - d[0]) * d[1], d[0] * (1 - d[1]), d[0] * d[1] };
        for (int k = 0; k < 4; k++) {
            int64_t o = pix_offset(p, ys[k], xs[k]);
            if (o >= 0) {
                lmap[o] += w[k] * t;
                lmap[o + p->polstride] += w[k] * t * cs;
                lmap[o + 2 * p->polstride] += w[k] * t * sn;
            }
            else if (!partial)
                err |= ORC_NOT_IN_PIX;
        }
    }
    return err;
}

uint32_t orc_response_map2tod_f64(double *tod, const double *lmap, const float *xp, const float *resp,
                                  int64_t ndet, int64_t nsamp, const orc_pix *p, int partial, int accum)
{
    const int64_t nper = nsamp / ndet;
    uint32_t err = 0;
    for (int64_t s = 0; s < nsamp; s++) {
        const double tr = resp[s / nper], pr = resp[ndet + s / nper];
        double sn, cs;
        sincos(2.0 * (double)xp[s + 2 * nsamp], &sn, &cs);
        int c[4];
        double d[2];
        err |= orc_locate_f64(xp[s], xp[s + nsamp], p->ny, p->nx, p->periodic, c, d);
        const int ys[4] = { c[0], c[0], c[1], c[1] };
        const int xs[4] = { c[2], c[3], c[2], c[3] };
        const double w[4] = { (1 - d[0]) * (1 - d[1]), (1 - d[0]) * d[1], d[0] * (1 - d[1]), d[0] * d[1] };
        double r = 0.0;
        for (int k = 0; k < 4; k++) {
            int64_t o = pix_offset(p, ys[k], xs[k]);
            if (o >= 0)
                r += w[k] * (tr * lmap[o] + pr * (cs * lmap[o + p->polstride] + sn * lmap[o + 2 * p->polstride]));
            else if (!partial)
                err |= ORC_NOT_IN_PIX;
        }
        tod[s] = accum ? (tod[s] + r) : r;
    }
    return err;
}

uint32_t orc_response_tod2map_f64(double *lmap, const double *tod, const float *xp, const float *resp,
                                  int64_t ndet, int64_t nsamp, const orc_pix *p, int partial)
{
    const int64_t nper = nsamp / ndet;
    uint32_t err = 0;
    for (int64_t s = 0; s < nsamp; s++) {
        const double tt = tod[s] * resp[s / nper], tp = tod[s] * resp[ndet + s / nper];
        double sn, cs;
        sincos(2.0 * (double)xp[s + 2 * nsamp], &sn, &cs);
        int c[4];
        double d[2];
        err |= orc_locate_f64(xp[s], xp[s + nsamp], p->ny, p->nx, p->periodic, c, d);
        const int ys[4] = { c[0], c[0], c[1], c[1] };
        const int xs[4] = { c[2], c[3], c[2], c[3] };
        const double w[4] = { (1 - d[0]) * (1 - d[1]), (1 - d[0]) * d[1], d[0] * (1 - d[1]), d[0] * d[1] };
        for (int k = 0; k < 4; k++) {
            int64_t o = pix_offset(p, ys[k], xs[k]);
            if (o >= 0) {
                lmap[o] += w[k] * tt;
                lmap[o + p->polstride] += w[k] * tp * cs;
                lmap[o + 2 * p->polstride] += w[k] * tp * sn;
            }
            else if (!partial)
                err |= ORC_NOT_IN_PIX;
        }
    }
    return err;
}

/* ---------------------------------------------------------------------------------------
 * Multi-threaded FP32 (CPU baseline on all cores).
 * --------------------------------------------------------------------------------------- */

int orc_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

uint32_t orc_map2tod_f32_mt(float *tod, const float *lmap, const float *xp, int64_t nsamp,
                            const orc_pix *p, int partial, int accum, int nthreads)
{
    uint32_t err = 0;
    if (nthreads <= 0)
        nthreads = orc_max_threads();
#pragma omp parallel for num_threads(nthreads) reduction(| : err) schedule(static)
    for (int64_t s = 0; s < nsamp; s++) {
        uint32_t e = 0;
        float r = m2t_sample_f32(lmap, xp, s, nsamp, p, partial, &e);
        tod[s] = accum ? (tod[s] + r) : r;
        err |= e;
    }
    return err;
}

uint32_t orc_tod2map_f32_mt(float *lmap, int64_t lmap_size, const float *tod, const float *xp,
                            int64_t nsamp, const orc_pix *p, int partial, int nthreads)
{
    if (nthreads <= 0)
        nthreads = orc_max_threads();
    float *priv = (float *)calloc((size_t)nthreads * (size_t)lmap_size, sizeof(float));
    if (!priv)
        return 0x80000000u;
    uint32_t err = 0;
#pragma omp parallel num_threads(nthreads) reduction(| : err)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        float *mine = priv + (size_t)tid * (size_t)lmap_size;
        uint32_t e = 0;
#pragma omp for schedule(static)
        for (int64_t s = 0; s < nsamp; s++)
            t2m_sample_f32(mine, tod[s], xp, s, nsamp, p, partial, &e);
        err |= e;
#pragma omp barrier
#pragma omp for schedule(static)
        for (int64_t i = 0; i < lmap_size; i++) {
            float acc = 0.0f;
            for (int t = 0; t < nthreads; t++)
                acc += priv[(size_t)t * (size_t)lmap_size + (size_t)i];
            lmap[i] += acc;
        }
    }
    free(priv);
    return err;
}

/* ---------------------------------------------------------------------------------------
 * Planner.
 * --------------------------------------------------------------------------------------- */

void orc_preplan_params(int64_t nsamp, int64_t *ncl_per_block, int64_t *nblocks)
{
    int64_t ncl = (nsamp + 31) / 32;
    int64_t per = (ncl + 1023) / 1024;
    if (per < 16)
        per = 16;
    per = 4 * ((per + 3) / 4);
    *ncl_per_block = per;
    *nblocks = (ncl + per - 1) / per;
}

void orc_pointing_params(int64_t nmt, int64_t *nmt_per_block, int64_t *nblocks)
{
    int64_t r = (int64_t)sqrt((double)nmt);
    r = 32 * ((r + 32) / 32);
    *nmt_per_block = r;
    *nblocks = (nmt + r - 1) / r;
}

/* Distinct cells (iycell << 10 | ixcell) touched by the bilinear corners of the 32 samples of
 * line 'l' (fewer in a partial last line when nsamp % 32 != 0, which the reference does not
 * allow: PointingPrePlan.cu and check_arguments.cu require nsamp % 32 == 0); cells[0] is the
 * cell of the (iy0, ix0) corner of the first sample. */
static int line_cells(const float *xp, int64_t nsamp, int64_t l, int ny, int nx, int periodic,
                      uint32_t cells[128], uint32_t *err)
{
    int n = 0;
    const int64_t s1 = (32 * l + 32 < nsamp) ? 32 * l + 32 : nsamp;
    for (int64_t s = 32 * l; s < s1; s++) {
        int c[4];
        float d[2];
        *err |= orc_locate_f32(xp[s], xp[s + nsamp], ny, nx, periodic, c, d);
        const int ys[4] = { c[0], c[0], c[1], c[1] };
        const int xs[4] = { c[2], c[3], c[2], c[3] };
        for (int k = 0; k < 4; k++) {
            uint32_t key = ((uint32_t)(ys[k] / 64) << 10) | (uint32_t)(xs[k] / 64);
            int seen = 0;
            for (int i = 0; i < n && !seen; i++)
                seen = (cells[i] == key);
            if (!seen)
                cells[n++] = key;
        }
    }
    return n;
}

uint32_t orc_plan_count(const float *xp, int64_t nsamp, int ny, int nx, int periodic,
                        uint32_t *nmt_cumsum, int64_t *nmt_out)
{
    int64_t per, nb;
    orc_preplan_params(nsamp, &per, &nb);
    int64_t ncl = (nsamp + 31) / 32, total = 0;
    uint32_t err = 0, cells[128];
    for (int64_t b = 0; b < nb; b++) {
        int64_t l1 = (b + 1) * per < ncl ? (b + 1) * per : ncl;
        for (int64_t l = b * per; l < l1; l++)
            total += line_cells(xp, nsamp, l, ny, nx, periodic, cells, &err);
        nmt_cumsum[b] = (uint32_t)total;
    }
    *nmt_out = total;
    return err;
}

static int cmp_mt(const void *a, const void *b)
{
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    uint64_t kx = ((x & 0xfffffull) << 26) | ((x >> 20) & 0x3ffffffull);
    uint64_t ky = ((y & 0xfffffull) << 26) | ((y >> 20) & 0x3ffffffull);
    return (kx > ky) - (kx < ky);
}

uint32_t orc_plan_build(const float *xp, int64_t nsamp, int ny, int nx, int periodic,
                        uint64_t *mt, int64_t nmt)
{
    int64_t ncl = (nsamp + 31) / 32, k = 0;
    uint32_t err = 0, cells[128];
    for (int64_t l = 0; l < ncl; l++) {
        int n = line_cells(xp, nsamp, l, ny, nx, periodic, cells, &err);
        for (int i = 0; i < n; i++) {
            if (k >= nmt)
                return err | 0x80000000u;
            uint64_t e = (uint64_t)cells[i] | ((uint64_t)l << 20);
            if (n > 1)
                e |= 1ull << 46;
            if (n > 1 && i == 0)
                e |= 1ull << 47;
            mt[k++] = e;
        }
    }
    if (k != nmt)
        return err | 0x80000000u;
    qsort(mt, (size_t)nmt, sizeof(uint64_t), cmp_mt);
    return err;
}

/* ---------------------------------------------------------------------------------------
 * LocalPixelization checks (src_lib/LocalPixelization.cu:33-101, check_arguments.cu:299-313).
 * --------------------------------------------------------------------------------------- */

int orc_lpix_check(int64_t ny, int64_t nx, const int64_t *offsets, int64_t nycells,
                   int64_t nxcells, int64_t ystride, int64_t polstride, int64_t *npix)
{
    if (ny <= 0 || ny > 65536 || nx <= 0 || nx > 65536)
        return 1;
    if (nycells <= 0 || nxcells <= 0 || nycells > 1024 || nxcells > 1024)
        return 2;
    if (ystride <= 0 || ystride % 64)
        return 3;
    if (polstride <= 0 || polstride % 64)
        return 4;
    if ((ystride % (3 * polstride)) != 0 && (polstride % (64 * ystride)) != 0)
        return 5;
    int64_t n = 0, mx = 0;
    for (int64_t i = 0; i < nycells * nxcells; i++) {
        if (offsets[i] < 0)
            continue;
        if (offsets[i] % 64)
            return 6;
        if (offsets[i] > mx)
            mx = offsets[i];
        n += 4096;
    }
    if (n > 0 && mx + 2 * polstride + 63 * ystride + 63 != 3 * n - 1)
        return 7;
    *npix = n;
    return 0;
}
