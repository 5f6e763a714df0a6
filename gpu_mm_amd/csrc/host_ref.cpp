// Host-side parts of libgpu_mm_amd.so: the single-threaded CPU operators behind
// plan='reference', the ToyPointing scan generator, and the brute-force planner used by
// PointingPlanTester. Compiled with -ffp-contract=off so that every float operation is a
// separately rounded IEEE operation, as in the reference's host build.

#include "gmm_internal.hpp"

#include <algorithm>
#include <cmath>

namespace gmm {

namespace {

// Looks up one pixel of the local map. Returns the flat offset of its T value, or -1 when
// the pixel's cell is not part of the local pixelization
// (map_evaluator / map_accumulator, include/gpu_mm_internals.hpp:258-282, 314-346).
inline int64_t local_offset(const gmm_lpix &lp, int iy, int ix)
{
    int cy = iy >> 6, cx = ix >> 6;
    if (iy < 0 || ix < 0 || cy >= lp.nycells || cx >= lp.nxcells)
        return -1;
    int64_t off = lp.offsets_cpu[int64_t(cy) * lp.nxcells + cx];
    if (off < 0)
        return -1;
    return off + int64_t(iy - (cy << 6)) * lp.ystride + (ix - (cx << 6));
}

}  // namespace


// reference_map2tod, src_lib/map2tod_reference.cu:16-53. Weight order and the T + cQ + sU
// combination follow map_evaluator::eval, include/gpu_mm_internals.hpp:284-292.
void reference_map2tod(float *tod, const float *lmap, const float *xp, const OpArgs &a, bool accum)
{
    const gmm_lpix &lp = *a.lpix;
    const int64_t n = a.nsamp;
    const int ny = int(lp.nypix_global), nx = int(lp.nxpix_global);
    uint32_t err = 0;

    for (int64_t s = 0; s < n; s++) {
        float y = xp[s], x = xp[s + n], alpha = xp[s + 2 * n];
        float sn, cs;
        sincosf(2.0f * alpha, &sn, &cs);

        Pixel<float> p;
        err |= locate_pixel(y, x, ny, nx, lp.periodic_xcoord, p);

        const int cy[4] = { p.iy0, p.iy0, p.iy1, p.iy1 };
        const int cx[4] = { p.ix0, p.ix1, p.ix0, p.ix1 };
        const float w[4] = { (1.0f - p.dy) * (1.0f - p.dx), (1.0f - p.dy) * p.dx,
                             p.dy * (1.0f - p.dx), p.dy * p.dx };
        float acc = 0.0f;
        for (int c = 0; c < 4; c++) {
            int64_t off = local_offset(lp, cy[c], cx[c]);
            float v = 0.0f;
            if (off >= 0)
                v = lmap[off] + cs * lmap[off + lp.polstride] + sn * lmap[off + 2 * lp.polstride];
            else if (!a.partial)
                err |= ERR_NOT_IN_PIXELIZATION;
            float term = w[c] * v;
            acc = (c == 0) ? term : (acc + term);
        }

        tod[s] = accum ? (tod[s] + acc) : acc;
    }

    throw_errflags(err, "reference_map2tod");
}


// reference_tod2map, src_lib/tod2map_reference.cu:14-46; corner order 00, 01, 10, 11
// (map_accumulator::accum, include/gpu_mm_internals.hpp:348-354). Cells missing from the
// pixelization are skipped (and flagged unless partial_pixelization).
void reference_tod2map(float *lmap, const float *tod, const float *xp, const OpArgs &a)
{
    const gmm_lpix &lp = *a.lpix;
    const int64_t n = a.nsamp;
    const int ny = int(lp.nypix_global), nx = int(lp.nxpix_global);
    uint32_t err = 0;

    for (int64_t s = 0; s < n; s++) {
        float y = xp[s], x = xp[s + n], alpha = xp[s + 2 * n];
        float t = tod[s];
        float sn, cs;
        sincosf(2.0f * alpha, &sn, &cs);
        float q = t * cs, u = t * sn;

        Pixel<float> p;
        err |= locate_pixel(y, x, ny, nx, lp.periodic_xcoord, p);

        const int cy[4] = { p.iy0, p.iy0, p.iy1, p.iy1 };
        const int cx[4] = { p.ix0, p.ix1, p.ix0, p.ix1 };
        const float w[4] = { (1.0f - p.dy) * (1.0f - p.dx), (1.0f - p.dy) * p.dx,
                             p.dy * (1.0f - p.dx), p.dy * p.dx };
        for (int c = 0; c < 4; c++) {
            int64_t off = local_offset(lp, cy[c], cx[c]);
            if (off >= 0) {
                lmap[off] += w[c] * t;
                lmap[off + lp.polstride] += w[c] * q;
                lmap[off + 2 * lp.polstride] += w[c] * u;
            }
            else if (!a.partial)
                err |= ERR_NOT_IN_PIXELIZATION;
        }
    }

    throw_errflags(err, "reference_tod2map");
}


// ToyPointing scan, src_lib/ToyPointing.cu:41-109: a 45-degree scan that bounces between
// y0 = 1.1*v and y1 = ny-2-1.1*v, with a slow x drift and periodic wrap in x, generated in
// double precision. drift_speed = total_drift / (3 * xpointing.size), i.e. / (9*nsamp),
// exactly as the reference computes it (the factor is kept so toy instances match).
double toy_pointing_fill(float *xp, int64_t nsamp, int64_t ny, int64_t nx, double scan_speed,
                         double total_drift)
{
    const double drift_speed = total_drift / (3.0 * double(3 * nsamp));
    GMM_REQUIRE(scan_speed > 0.0 && scan_speed <= 1.0, "ToyPointing: expected 0 < scan_speed <= 1");
    GMM_REQUIRE(drift_speed > 0.0 && drift_speed <= 1.0, "ToyPointing: expected 0 < drift_speed <= 1");

    const double ylo = 1.1 * scan_speed;
    const double yhi = double(ny - 2) - 1.1 * scan_speed;
    double y = 1.0, x = 1.0, v = scan_speed;

    for (int64_t s = 0; s < nsamp; s++) {
        xp[s] = float(y);
        xp[s + nsamp] = float(x);
        xp[s + 2 * nsamp] = 1.0f;

        if (v > 0 && y > yhi)
            v = -scan_speed;
        if (v < 0 && y < ylo)
            v = scan_speed;

        y += v;
        x += v + drift_speed;
        if (x < 0.0)
            x += double(nx);
        if (x > double(nx))
            x -= double(nx);
    }
    return drift_speed;
}


// Brute-force planner of PointingPlanTester, src_lib/PointingPlanTester.cu:120-160: for every
// 32-sample line, the distinct cells of the pixels (iy,ix), (iy,ix1), (iy+1,ix), (iy+1,ix1)
// with ix1 = (ix < nx-1) ? ix+1 : 0, counted per planner block and emitted as
// icell | (line << 20); the full list is then sorted as integers. iypix and ixpix hold nsamp
// entries; a partial last line (nsamp % 32 != 0) has only its first nsamp % 32 samples.
void plan_tester_host(const gmm_preplan &pp, const int32_t *iypix, const int32_t *ixpix,
                      uint32_t *nmt_cumsum, uint64_t *sorted_mt)
{
    const int64_t nsamp = pp.nsamp;
    const int ny = int(pp.nypix_global), nx = int(pp.nxpix_global);
    const int64_t spb = 32 * pp.pparams.ncl_per_threadblock;
    int64_t nmt = 0;

    for (int64_t b = 0; b < pp.pparams.planner_nblocks; b++) {
        int64_t s1 = std::min(nsamp, (b + 1) * spb);
        for (int64_t s0 = b * spb; s0 < s1; s0 += 32) {
            uint32_t cells[128];
            int ncells = 0;
            // The last line of a TOD whose length is not a multiple of 32 is partial.
            const int64_t se = std::min(nsamp, s0 + 32);
            for (int64_t s = s0; s < se; s++) {
                int iy = iypix[s], ix = ixpix[s];
                int ix1 = (ix < nx - 1) ? (ix + 1) : 0;
                const int py[4] = { iy, iy, iy + 1, iy + 1 };
                const int px[4] = { ix, ix1, ix, ix1 };
                for (int c = 0; c < 4; c++) {
                    GMM_REQUIRE(py[c] >= 0 && py[c] < ny && px[c] >= 0 && px[c] < nx,
                                "PointingPlanTester: quantized pixel out of range");
                    uint32_t k = cell_key(py[c] >> 6, px[c] >> 6);
                    bool seen = false;
                    for (int i = 0; i < ncells; i++)
                        seen = seen || (cells[i] == k);
                    if (!seen)
                        cells[ncells++] = k;
                }
            }
            std::sort(cells, cells + ncells);
            for (int i = 0; i < ncells; i++) {
                GMM_REQUIRE(nmt < pp.plan_nmt, "PointingPlanTester: more plan entries than the preplan counted");
                sorted_mt[nmt++] = uint64_t(cells[i]) | (uint64_t(s0 >> 5) << ICL_SHIFT);
            }
        }
        nmt_cumsum[b] = uint32_t(nmt);
    }

    GMM_REQUIRE(nmt == pp.plan_nmt, "PointingPlanTester: plan entry count differs from the preplan");
    std::sort(sorted_mt, sorted_mt + nmt);
}


// Random cell-sorted plan for iterator tests (src_lib/test_plan_iterator.cu:170-209):
// 'ncells' distinct random cells in increasing order, each with a random number of entries in
// [min_nmt, max_nmt]; every entry carries a random line index and random junk in bits 46-63.
namespace {
struct SplitMix
{
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    int64_t range(int64_t lo, int64_t hi) { return lo + int64_t(next() % uint64_t(hi - lo + 1)); }
};
}  // namespace

std::vector<uint64_t> make_random_plan_mt(int64_t ncells, int64_t min_nmt, int64_t max_nmt, uint64_t seed)
{
    GMM_REQUIRE(ncells > 0 && ncells <= (1 << 20), "make_random_plan_mt: expected 0 < ncells <= 2^20");
    GMM_REQUIRE(min_nmt >= 1 && min_nmt <= max_nmt && max_nmt <= 16 * 1024,
                "make_random_plan_mt: expected 1 <= min_nmt_per_cell <= max_nmt_per_cell <= 16384");

    SplitMix rng{seed};
    // Partial Fisher-Yates over all 2^20 cell keys.
    std::vector<uint32_t> perm(1 << 20);
    for (uint32_t i = 0; i < perm.size(); i++)
        perm[i] = i;
    for (int64_t i = 0; i < ncells; i++) {
        int64_t j = rng.range(i, (1 << 20) - 1);
        std::swap(perm[i], perm[j]);
    }
    std::vector<uint32_t> cells(perm.begin(), perm.begin() + ncells);
    std::sort(cells.begin(), cells.end());

    std::vector<uint64_t> out;
    for (int64_t c = 0; c < ncells; c++) {
        int64_t n = rng.range(min_nmt, max_nmt);
        for (int64_t j = 0; j < n; j++) {
            uint64_t icl = rng.next() & ICL_MASK;
            uint64_t junk = rng.next() & ((1ull << 18) - 1);
            out.push_back(uint64_t(cells[c]) | (icl << ICL_SHIFT) | (junk << 46));
        }
    }
    GMM_REQUIRE(out.size() <= (1ull << 30), "make_random_plan_mt: too many entries");
    return out;
}

}  // namespace gmm
