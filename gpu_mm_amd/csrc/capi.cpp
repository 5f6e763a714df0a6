// C ABI of libgpu_mm_amd.so (declared in include/gpu_mm_amd.h): argument checks, handle
// lifetime, and the host side of plan construction and operator launches.

#include "gmm_internal.hpp"
#include "gpu_mm_amd.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <mutex>
#include <cstring>
#include <memory>
#include <sstream>

namespace gmm {

namespace {
thread_local std::string g_last_error;
}

void fail(const std::string &msg) { throw Error(msg); }

void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess)
        fail(std::string("HIP error '") + hipGetErrorString(e) + "' in " + what);
}

// Messages as in src_lib/check_arguments.cu:159-175 (the y-range text states the range that
// locate() actually accepts, [0, nypix_global-1]).
void throw_errflags(uint32_t err, const char *where, uint32_t ignore)
{
    err &= ~ignore;
    if (!err)
        return;
    std::string w(where);
    if (err & ERR_BAD_YPIX)
        fail(w + ": xpointing y-value is outside range [0,nypix_global-1].");
    if (err & ERR_BAD_XPIX)
        fail(w + ": xpointing x-value is outside range, perhaps you want the 'periodic_xcoord' flag?");
    if (err & ERR_INCONSISTENT_NMT)
        fail(w + ": inconsistent value of nmt between preplan/plan?! (should never happen)");
    if (err & ERR_NOT_IN_PIXELIZATION)
        fail(w + ": xpointing (y,x) value is not in LocalPixelization (perhaps pixelization is too small, or you want the 'partial_pixelization' flag");
    fail(w + ": bad errflags?! (should never happen)");
}

// ------------------------------------------------------------------------------------------
// Kernel timer: HIP event pairs recorded on the launch stream around each main operator kernel
// while enabled (a measurement aid; off by default, one relaxed load per launch when off).

namespace {
std::atomic<bool> g_timing{false};
std::mutex g_timer_mu;
std::vector<std::pair<hipEvent_t, hipEvent_t>> g_timer_ev[TK_COUNT];
size_t g_timer_used[TK_COUNT] = {};
}  // namespace

void kernel_timer_mark(int which, bool end, hipStream_t stream)
{
    if (!g_timing.load(std::memory_order_relaxed))
        return;
    std::lock_guard<std::mutex> lk(g_timer_mu);
    auto &v = g_timer_ev[which];
    size_t &u = g_timer_used[which];
    if (!end) {
        if (u == v.size()) {
            hipEvent_t a, b;
            GMM_HIP(hipEventCreate(&a));
            GMM_HIP(hipEventCreate(&b));
            v.emplace_back(a, b);
        }
        GMM_HIP(hipEventRecord(v[u].first, stream));
    }
    else if (u < v.size()) {
        GMM_HIP(hipEventRecord(v[u].second, stream));
        u++;
    }
}

// src_lib/check_arguments.cu also requires nsamp % 32 == 0 (README.md:68-70 lists lifting it as
// a to-do); here the last 32-sample line of a plan may be partial.
void check_nsamp(int64_t nsamp, const char *where)
{
    if (nsamp <= 0 || nsamp > (int64_t(1) << 31)) {
        std::ostringstream ss;
        ss << where << ": expected nsamp > 0 and nsamp <= 2^31 (got nsamp=" << nsamp << ")";
        fail(ss.str());
    }
}

// 16-bit TODs and the batched operators access 4 consecutive samples with one vector access.
void check_nsamp_quads(int64_t nsamp, const char *where)
{
    if (nsamp % 4 != 0) {
        std::ostringstream ss;
        ss << where << ": expected nsamp divisible by 4 (got nsamp=" << nsamp << ")";
        fail(ss.str());
    }
}

void check_npix_global(int64_t n, const char *what, const char *where)
{
    if (n <= 0 || n > 64 * 1024) {
        std::ostringstream ss;
        ss << where << ": expected 0 < " << what << " <= 65536 (got " << n << ")";
        fail(ss.str());
    }
}

void set_last_error(const char *msg) { g_last_error = msg; }

void require_ptr(const void *p, const char *where, const char *name)
{
    if (!p)
        fail(std::string(where) + ": '" + name + "' is a null pointer");
}

uint32_t read_word(const uint32_t *p, hipStream_t stream)
{
    uint32_t h = 0;
    GMM_HIP(hipMemcpyAsync(&h, p, 4, hipMemcpyDeviceToHost, stream));
    GMM_HIP(hipStreamSynchronize(stream));
    return h;
}

void clear_word(uint32_t *p, hipStream_t stream)
{
    static const uint32_t zero = 0;
    GMM_HIP(hipMemcpyAsync(p, &zero, 4, hipMemcpyHostToDevice, stream));
    GMM_HIP(hipStreamSynchronize(stream));
}

namespace {

void init_npix(gmm_lpix &lp, const char *where)
{
    // LocalPixelization.cu:74-101: count local cells, and check that the cells tile
    // [0, 3*npix) exactly (largest cell offset + extent == 3*npix - 1).
    int64_t max_offset = 0;
    lp.npix = 0;
    for (int64_t i = 0; i < lp.nycells * lp.nxcells; i++) {
        int64_t o = lp.offsets_cpu[i];
        if (o < 0)
            continue;
        if (o % 64 != 0)
            fail(std::string(where) + ": cell_offsets must be multiples of 64");
        max_offset = std::max(max_offset, o);
        lp.npix += CELL_NPIX;
    }
    max_offset += 2 * lp.polstride + 63 * lp.ystride + 63;
    if (lp.npix > 0 && max_offset != 3 * lp.npix - 1)
        fail(std::string(where) + ": something is wrong with the cell_offsets -- the cells do not fit together contiguously");
}

// Synchronous check of the plan's accumulated error word; clears it when set.
void check_plan_errors(const gmm_plan &plan, const char *where, hipStream_t stream)
{
    uint32_t *w = plan.err + (plan.pp.n_err - 1);
    uint32_t e = read_word(w, stream);
    if (e) {
        clear_word(w, stream);
        throw_errflags(e, where);
    }
}

void check_op_common(const gmm_lpix *lpix, const gmm_plan *plan, int64_t tod_nsamp,
                     int64_t xp_nsamp, int64_t map_size, const char *where)
{
    if (!lpix)
        fail(std::string(where) + ": null LocalPixelization");
    if (tod_nsamp != xp_nsamp) {
        std::ostringstream ss;
        ss << where << ": tod has " << tod_nsamp << " samples but xpointing has " << xp_nsamp;
        fail(ss.str());
    }
    check_nsamp(tod_nsamp, where);
    if (map_size != 3 * lpix->npix) {
        std::ostringstream ss;
        ss << where << ": expected local map of size 3*npix=" << 3 * lpix->npix << ", got " << map_size;
        fail(ss.str());
    }
    if (plan) {
        if (plan->pp.nsamp != tod_nsamp) {
            std::ostringstream ss;
            ss << where << ": plan was built for nsamp=" << plan->pp.nsamp << ", got " << tod_nsamp;
            fail(ss.str());
        }
        // map2tod.cu:211-215: pixelization and plan must agree on the global pixelization.
        if (plan->pp.nypix_global != lpix->nypix_global || plan->pp.nxpix_global != lpix->nxpix_global
            || plan->pp.periodic_xcoord != lpix->periodic_xcoord)
            fail(std::string(where) + ": LocalPixelization and PointingPlan disagree on (nypix_global, nxpix_global, periodic_xcoord)");
    }
}

int bits_for(int64_t n)
{
    int b = 0;
    while ((int64_t(1) << b) < n)
        b++;
    return b;
}

}  // namespace

// check_response(), src_lib/check_arguments.cu:171-180: response is float[2][ndet] (row 0: T
// response, row 1: polarization response) and ndet must divide nsamp.
void check_response(const float *response_gpu, int64_t ndet, int64_t nsamp, const char *where)
{
    require_ptr(response_gpu, where, "response");
    if (ndet <= 0)
        fail(std::string(where) + ": response array must have shape (2, ndet) with ndet > 0");
    if (nsamp % ndet != 0)
        fail(std::string(where) + ": response array has ndet=" + std::to_string(ndet)
             + ", which does not divide nsamp=" + std::to_string(nsamp));
    if (nsamp / ndet > int64_t(UINT32_MAX))
        fail(std::string(where) + ": samples per detector must be < 2^32");
}

}  // namespace gmm

using namespace gmm;

namespace {
// Arguments of the 16-bit TOD entry points.
TodGains q16_gains(int tod_type, const void *tod, const float *scale, int64_t nrows, int64_t nsamp,
                   const float *response, int64_t resp_ndet, const char *where)
{
    if (tod_type != GMM_TOD_F16 && tod_type != GMM_TOD_BF16)
        fail(std::string(where) + ": tod_type must be GMM_TOD_F16 or GMM_TOD_BF16");
    check_nsamp_quads(nsamp, where);
    if (reinterpret_cast<uintptr_t>(tod) & 7)
        fail(std::string(where) + ": a 16-bit TOD must be 8-byte aligned");
    require_ptr(scale, where, "tod_scale");
    if (nrows <= 0 || nsamp % nrows != 0)
        fail(std::string(where) + ": tod_scale has " + std::to_string(nrows)
             + " rows, which does not divide nsamp=" + std::to_string(nsamp));
    TodGains g;
    g.tod_type = (tod_type == GMM_TOD_F16) ? TOD_F16 : TOD_BF16;
    g.scale = scale;
    g.scale_nrows = nrows;
    if (response) {
        check_response(response, resp_ndet, nsamp, where);
        g.response = response;
        g.resp_ndet = resp_ndet;
    }
    return g;
}
}  // namespace

namespace {
void check_cells(const float *p, const char *where, const char *name)
{
    require_ptr(p, where, name);
    if (reinterpret_cast<uintptr_t>(p) & 15)
        fail(std::string(where) + ": '" + name + "' must be 16-byte aligned");
}

// Runs a cell kernel with a device 'bad index' word, then checks it (synchronizes the stream).
template<typename F>
void with_bad_word(const char *where, hipStream_t stream, F f)
{
    uint32_t *bad = nullptr;
    GMM_HIP(hipMalloc(reinterpret_cast<void **>(&bad), 4));
    uint32_t h = 0;
    try {
        launch_fill_u32(bad, 1, 0u, stream);
        f(bad);
        GMM_HIP(hipMemcpyAsync(&h, bad, 4, hipMemcpyDeviceToHost, stream));
        GMM_HIP(hipStreamSynchronize(stream));
    }
    catch (...) {
        (void) hipFree(bad);
        throw;
    }
    GMM_HIP(hipFree(bad));
    if (h)
        fail(std::string(where) + ": index_map entry out of range");
}
}  // namespace

extern "C" {

const char *gmm_last_error(void) { return g_last_error.c_str(); }

int gmm_version(void) { return 100; }

int gmm_kernel_timer_enable(int enable)
{
    return guarded([&] {
        std::lock_guard<std::mutex> lk(g_timer_mu);
        for (int k = 0; k < TK_COUNT; k++)
            g_timer_used[k] = 0;
        g_timing.store(enable != 0);
    });
}

int gmm_kernel_timer_read(int which, double *total_ms, int64_t *count)
{
    return guarded([&] {
        if (which < 0 || which >= TK_COUNT)
            fail("kernel_timer_read: unknown kernel index");
        require_ptr(total_ms, "kernel_timer_read", "total_ms");
        require_ptr(count, "kernel_timer_read", "count");
        std::lock_guard<std::mutex> lk(g_timer_mu);
        double t = 0.0;
        for (size_t i = 0; i < g_timer_used[which]; i++) {
            GMM_HIP(hipEventSynchronize(g_timer_ev[which][i].second));
            float ms = 0.0f;
            GMM_HIP(hipEventElapsedTime(&ms, g_timer_ev[which][i].first, g_timer_ev[which][i].second));
            t += ms;
        }
        *total_ms = t;
        *count = int64_t(g_timer_used[which]);
        g_timer_used[which] = 0;
    });
}

int gmm_stream_read(const float *src_gpu, int64_t n, float *sum_gpu, void *stream_)
{
    return guarded([&] {
        const char *where = "stream_read";
        require_ptr(src_gpu, where, "src");
        require_ptr(sum_gpu, where, "sum");
        if (n < 0 || n % 4 != 0)
            fail(std::string(where) + ": n must be a non-negative multiple of 4");
        if (reinterpret_cast<uintptr_t>(src_gpu) & 15)
            fail(std::string(where) + ": 'src' must be 16-byte aligned");
        if (n == 0)
            return;
        launch_stream_read(src_gpu, n, sum_gpu, as_stream(stream_));
    });
}

// ------------------------------------------------------------------------------------------

int gmm_lpix_create(gmm_lpix **out, int64_t nypix_global, int64_t nxpix_global,
                    const int64_t *cell_offsets_cpu, const int64_t *cell_offsets_gpu,
                    int64_t nycells, int64_t nxcells, int64_t ystride, int64_t polstride,
                    int periodic_xcoord)
{
    return guarded([&] {
        const char *where = "LocalPixelization constructor";
        require_ptr(out, where, "out");
        require_ptr(cell_offsets_cpu, where, "cell_offsets_cpu");
        check_npix_global(nypix_global, "nypix_global", where);
        check_npix_global(nxpix_global, "nxpix_global", where);
        // check_arguments.cu:299-313: 0 < ncells <= 1024 per axis.
        if (nycells <= 0 || nxcells <= 0 || nycells > 1024 || nxcells > 1024)
            fail(std::string(where) + ": expected cell_offsets shape (nycells, nxcells) with 0 < nycells, nxcells <= 1024");
        // LocalPixelization.cu:58-68.
        if (ystride <= 0 || ystride % 64 != 0)
            fail(std::string(where) + ": expected ystride > 0 and a multiple of 64");
        if (polstride <= 0 || polstride % 64 != 0)
            fail(std::string(where) + ": expected polstride > 0 and a multiple of 64");
        if (!((ystride % (3 * polstride)) == 0 || (polstride % (64 * ystride)) == 0))
            fail(std::string(where) + ": expected either ystride to be a multiple of (3*polstride), or polstride to be a multiple of (64*ystride)");

        auto lp = new gmm_lpix;
        lp->nypix_global = nypix_global;
        lp->nxpix_global = nxpix_global;
        lp->periodic_xcoord = (periodic_xcoord != 0);
        lp->offsets_cpu = cell_offsets_cpu;
        lp->offsets_gpu = cell_offsets_gpu;
        lp->nycells = nycells;
        lp->nxcells = nxcells;
        lp->ystride = ystride;
        lp->polstride = polstride;
        try {
            init_npix(*lp, where);
        }
        catch (...) {
            delete lp;
            throw;
        }
        *out = lp;
    });
}

void gmm_lpix_destroy(gmm_lpix *lpix) { delete lpix; }

int gmm_lpix_get_info(const gmm_lpix *lp, gmm_lpix_info *out)
{
    return guarded([&] {
        require_ptr(lp, "gmm_lpix_get_info", "lpix");
        require_ptr(out, "gmm_lpix_get_info", "out");
        out->nypix_global = lp->nypix_global;
        out->nxpix_global = lp->nxpix_global;
        out->periodic_xcoord = lp->periodic_xcoord ? 1 : 0;
        out->ystride = lp->ystride;
        out->polstride = lp->polstride;
        out->nycells = lp->nycells;
        out->nxcells = lp->nxcells;
        out->npix = lp->npix;
    });
}

int gmm_lpix_set_npix(gmm_lpix *lp, int64_t npix)
{
    return guarded([&] {
        require_ptr(lp, "gmm_lpix_set_npix", "lpix");
        lp->npix = npix;
    });
}

int gmm_lpix_copy_gpu_offsets_to_cpu(gmm_lpix *lp)
{
    return guarded([&] {
        const char *where = "LocalPixelization::copy_gpu_offsets_to_cpu()";
        require_ptr(lp, where, "lpix");
        require_ptr(lp->offsets_gpu, where, "cell_offsets_gpu");
        GMM_HIP(hipMemcpy(const_cast<int64_t *>(lp->offsets_cpu), lp->offsets_gpu,
                          size_t(lp->nycells * lp->nxcells) * 8, hipMemcpyDeviceToHost));
        init_npix(*lp, where);
    });
}

int gmm_lpix_copy_cpu_offsets_to_gpu(gmm_lpix *lp)
{
    return guarded([&] {
        const char *where = "LocalPixelization::copy_cpu_offsets_to_gpu()";
        require_ptr(lp, where, "lpix");
        require_ptr(lp->offsets_gpu, where, "cell_offsets_gpu");
        init_npix(*lp, where);
        GMM_HIP(hipMemcpy(const_cast<int64_t *>(lp->offsets_gpu), lp->offsets_cpu,
                          size_t(lp->nycells * lp->nxcells) * 8, hipMemcpyHostToDevice));
    });
}

// ------------------------------------------------------------------------------------------

int gmm_preplan_create(gmm_preplan **out, const float *xpointing_gpu, int64_t nsamp,
                       int64_t nypix_global, int64_t nxpix_global,
                       uint32_t *nmt_gpu, uint32_t *err_gpu,
                       int periodic_xcoord, int debug, void *stream_)
{
    return guarded([&] {
        const char *where = "PointingPrePlan constructor";
        hipStream_t stream = as_stream(stream_);
        require_ptr(out, where, "out");
        require_ptr(xpointing_gpu, where, "xpointing_gpu");
        require_ptr(nmt_gpu, where, "nmt_gpu");
        require_ptr(err_gpu, where, "err_gpu");
        check_nsamp(nsamp, where);
        check_npix_global(nypix_global, "nypix_global", where);
        check_npix_global(nxpix_global, "nxpix_global", where);
        (void)debug;

        gmm_preplan pp;
        pp.nsamp = nsamp;
        pp.nypix_global = nypix_global;
        pp.nxpix_global = nxpix_global;
        pp.periodic_xcoord = (periodic_xcoord != 0);
        pp.nmt_cumsum_gpu = nmt_gpu;
        pp.pparams = preplan_params(nsamp);
        GMM_REQUIRE(pp.pparams.planner_nblocks > 0 && pp.pparams.planner_nblocks <= PREPLAN_SIZE,
                    "PointingPrePlan constructor: bad planner block count");

        launch_preplan_count(xpointing_gpu, nsamp, pp, nmt_gpu, err_gpu, stream);
        launch_preplan_finish(nmt_gpu, err_gpu, pp.pparams.planner_nblocks, nullptr, stream);

        uint32_t h[3];
        GMM_HIP(hipMemcpyAsync(h, err_gpu, sizeof(h), hipMemcpyDeviceToHost, stream));
        GMM_HIP(hipStreamSynchronize(stream));
        throw_errflags(h[0], where);

        const int64_t nmt = int64_t(h[1]) | (int64_t(h[2]) << 32);
        if (nmt <= 0 || nmt > (int64_t(1) << 30))
            fail(std::string(where) + ": plan entry count out of range (expected 0 < nmt <= 2^30)");
        pp.plan_nmt = nmt;

        // PointingPrePlan.cu:236-250 (reference launch formulas, reported as attributes).
        pp.nmt_per_threadblock = int64_t(std::sqrt(double(nmt)));
        pp.nmt_per_threadblock = (pp.nmt_per_threadblock + 32) & ~int64_t(31);
        pp.pointing_nblocks = (nmt + pp.nmt_per_threadblock - 1) / pp.nmt_per_threadblock;
        pp.overhead = 32.0 * double(nmt) / double(nsamp) - 1.0;

        // The plan layout of this library.
        pp.nycells_g = (nypix_global + 63) / 64;
        pp.nxcells_g = (nxpix_global + 63) / 64;
        const int64_t nkeys = pp.nycells_g * pp.nxcells_g;
        pp.key_bits = bits_for(nkeys);
        int64_t cap = (nmt + 3071) / 3072;
        pp.chunk_cap = std::min<int64_t>(8192, std::max<int64_t>(512, cap));
        pp.max_nchunks = (nmt + pp.chunk_cap - 1) / pp.chunk_cap + std::min(nmt, nkeys);
        pp.n_err = std::max<int64_t>(pp.pparams.planner_nblocks, pp.max_nchunks) + 1;

        pp.off_err = align128(8 * nmt);
        pp.off_chunks = pp.off_err + align128(4 * pp.n_err);
        pp.off_qmask = pp.off_chunks + align128(int64_t(sizeof(Chunk)) * pp.max_nchunks);
        pp.plan_nbytes = pp.off_qmask + align128(2 * nmt);

        pp.sort_nbytes = plan_scratch_nbytes(pp) + 128;
        pp.tmp_nbytes = 2 * align128(8 * nmt) + pp.sort_nbytes;

        *out = new gmm_preplan(pp);
    });
}

void gmm_preplan_destroy(gmm_preplan *pp) { delete pp; }

int gmm_preplan_get_info(const gmm_preplan *pp, gmm_preplan_info *out)
{
    return guarded([&] {
        require_ptr(pp, "gmm_preplan_get_info", "preplan");
        require_ptr(out, "gmm_preplan_get_info", "out");
        out->nsamp = pp->nsamp;
        out->nypix_global = pp->nypix_global;
        out->nxpix_global = pp->nxpix_global;
        out->periodic_xcoord = pp->periodic_xcoord ? 1 : 0;
        out->plan_nbytes = pp->plan_nbytes;
        out->plan_constructor_tmp_nbytes = pp->tmp_nbytes;
        out->overhead = pp->overhead;
        out->ncl_per_threadblock = pp->pparams.ncl_per_threadblock;
        out->planner_nblocks = pp->pparams.planner_nblocks;
        out->nmt_per_threadblock = pp->nmt_per_threadblock;
        out->pointing_nblocks = pp->pointing_nblocks;
        out->plan_nmt = pp->plan_nmt;
        out->cub_nbytes = pp->sort_nbytes;
        out->max_nchunks = pp->max_nchunks;
    });
}

int gmm_preplan_get_nmt_cumsum(const gmm_preplan *pp, uint32_t *out_cpu, int64_t n)
{
    return guarded([&] {
        require_ptr(pp, "get_nmt_cumsum", "preplan");
        require_ptr(out_cpu, "get_nmt_cumsum", "out");
        if (n != pp->pparams.planner_nblocks)
            fail("get_nmt_cumsum: output length must equal planner_nblocks");
        GMM_HIP(hipMemcpy(out_cpu, pp->nmt_cumsum_gpu, size_t(n) * 4, hipMemcpyDeviceToHost));
    });
}

// ------------------------------------------------------------------------------------------

int gmm_plan_create(gmm_plan **out, const gmm_preplan *pp, const float *xpointing_gpu,
                    int64_t nsamp, void *buf_gpu, int64_t buf_nbytes,
                    void *tmp_buf_gpu, int64_t tmp_nbytes, int debug, void *stream_)
{
    return guarded([&] {
        const char *where = "PointingPlan constructor";
        hipStream_t stream = as_stream(stream_);
        require_ptr(out, where, "out");
        require_ptr(pp, where, "preplan");
        require_ptr(xpointing_gpu, where, "xpointing_gpu");
        require_ptr(buf_gpu, where, "buf");
        require_ptr(tmp_buf_gpu, where, "tmp_buf");
        if (nsamp != pp->nsamp) {
            std::ostringstream ss;
            ss << where << ": xpointing has " << nsamp << " samples, preplan was built for " << pp->nsamp;
            fail(ss.str());
        }
        // check_buffer(), check_arguments.cu:326-345.
        if (buf_nbytes < pp->plan_nbytes) {
            std::ostringstream ss;
            ss << where << ": expected 'buf' to be 1-d array of length >= " << pp->plan_nbytes << ", actual length=" << buf_nbytes;
            fail(ss.str());
        }
        if (tmp_nbytes < pp->tmp_nbytes) {
            std::ostringstream ss;
            ss << where << ": expected 'tmp_buf' to be 1-d array of length >= " << pp->tmp_nbytes << ", actual length=" << tmp_nbytes;
            fail(ss.str());
        }
        if ((reinterpret_cast<uintptr_t>(buf_gpu) % 128) || (reinterpret_cast<uintptr_t>(tmp_buf_gpu) % 128))
            fail(std::string(where) + ": 'buf' and 'tmp_buf' must be 128-byte aligned");

        auto plan = new gmm_plan;
        plan->pp = *pp;
        plan->buf = reinterpret_cast<uint8_t *>(buf_gpu);
        plan->plan_mt = reinterpret_cast<uint64_t *>(plan->buf);
        plan->err = reinterpret_cast<uint32_t *>(plan->buf + pp->off_err);
        plan->chunks = reinterpret_cast<Chunk *>(plan->buf + pp->off_chunks);
        plan->qmask = reinterpret_cast<uint16_t *>(plan->buf + pp->off_qmask);

        try {
            uint8_t *tmp = reinterpret_cast<uint8_t *>(tmp_buf_gpu);
            const int64_t mt_bytes = align128(8 * pp->plan_nmt);
            uint64_t *a = reinterpret_cast<uint64_t *>(tmp);
            uint64_t *b = reinterpret_cast<uint64_t *>(tmp + mt_bytes);
            uint8_t *scratch = tmp + 2 * mt_bytes;
            const int64_t scratch_nbytes = pp->sort_nbytes - 128;
            uint32_t *counts = reinterpret_cast<uint32_t *>(scratch + scratch_nbytes);

            launch_fill_u32(plan->err, pp->n_err, 0u, stream);
            launch_plan_emit(xpointing_gpu, *pp, a, plan->err, stream);
            launch_reduce_or(plan->err, pp->pparams.planner_nblocks, plan->err + (pp->n_err - 1), stream);
            plan_sort(*pp, a, b, plan->plan_mt, scratch, scratch_nbytes, stream);
            plan_schedule(*pp, plan->plan_mt, xpointing_gpu, plan->chunks, plan->qmask, counts, scratch, scratch_nbytes,
                          stream);

            uint32_t h[2];
            GMM_HIP(hipMemcpyAsync(h, counts, sizeof(h), hipMemcpyDeviceToHost, stream));
            uint32_t e = read_word(plan->err + (pp->n_err - 1), stream);
            if (e) {
                clear_word(plan->err + (pp->n_err - 1), stream);
                throw_errflags(e, where);
            }
            plan->nchunks = h[0];
            plan->nmlines = h[1];
            GMM_REQUIRE(plan->nchunks <= pp->max_nchunks,
                        "PointingPlan constructor: work list larger than its bound (internal error)");
            order_chunks(*pp, plan->chunks, plan->nchunks, a, b, scratch, scratch_nbytes, stream);
            if (debug)
                check_plan_invariants(*plan, where, stream);
        }
        catch (...) {
            delete plan;
            throw;
        }
        *out = plan;
    });
}

void gmm_plan_destroy(gmm_plan *plan) { delete plan; }

int gmm_plan_get_info(const gmm_plan *plan, gmm_plan_info *out)
{
    return guarded([&] {
        require_ptr(plan, "gmm_plan_get_info", "plan");
        require_ptr(out, "gmm_plan_get_info", "out");
        out->nsamp = plan->pp.nsamp;
        out->nypix_global = plan->pp.nypix_global;
        out->nxpix_global = plan->pp.nxpix_global;
        out->periodic_xcoord = plan->pp.periodic_xcoord ? 1 : 0;
        out->plan_nmt = plan->pp.plan_nmt;
        out->nchunks = plan->nchunks;
        out->nmlines = plan->nmlines;
    });
}

int gmm_plan_get_plan_mt(const gmm_plan *plan, uint64_t *out, int64_t n, int out_on_gpu, void *stream_)
{
    return guarded([&] {
        require_ptr(plan, "get_plan_mt", "plan");
        require_ptr(out, "get_plan_mt", "out");
        if (n != plan->pp.plan_nmt)
            fail("get_plan_mt: output length must equal plan_nmt");
        hipStream_t stream = as_stream(stream_);
        GMM_HIP(hipMemcpyAsync(out, plan->plan_mt, size_t(n) * 8,
                               out_on_gpu ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, stream));
        GMM_HIP(hipStreamSynchronize(stream));
    });
}

int gmm_plan_check_errflags(const gmm_plan *plan, const char *where)
{
    return guarded([&] {
        require_ptr(plan, "check_errflags", "plan");
        check_plan_errors(*plan, where ? where : "PointingPlan", nullptr);
    });
}

int gmm_plan_check_invariants(const gmm_plan *plan, void *stream_)
{
    return guarded([&] {
        require_ptr(plan, "PointingPlan check_invariants", "plan");
        check_plan_invariants(*plan, "PointingPlan check_invariants", as_stream(stream_));
    });
}

// ------------------------------------------------------------------------------------------

int gmm_planned_map2tod(float *tod_gpu, int64_t tod_nsamp,
                        const float *local_map_gpu, int64_t local_map_size,
                        const float *xpointing_gpu, int64_t xpointing_nsamp,
                        const gmm_lpix *lpix, const gmm_plan *plan,
                        int partial_pixelization, int debug, int accum, int check_errors,
                        void *stream_)
{
    return guarded([&] {
        const char *where = "map2tod";
        hipStream_t stream = as_stream(stream_);
        require_ptr(plan, where, "plan");
        require_ptr(tod_gpu, where, "tod");
        require_ptr(local_map_gpu, where, "local_map");
        require_ptr(xpointing_gpu, where, "xpointing");
        check_op_common(lpix, plan, tod_nsamp, xpointing_nsamp, local_map_size, where);
        require_ptr(lpix->offsets_gpu, where, "cell_offsets_gpu");
        OpArgs a{lpix, tod_nsamp, partial_pixelization != 0};
        std::unique_ptr<OpDebug> dbg(debug ? new OpDebug(plan->pp.plan_nmt, tod_nsamp, stream) : nullptr);
        a.debug = dbg ? &dbg->bufs : nullptr;
        launch_planned_map2tod(tod_gpu, local_map_gpu, xpointing_gpu, a, *plan, accum != 0, TodGains{}, stream);
        if (debug || check_errors)
            check_plan_errors(*plan, where, stream);
        if (dbg)
            dbg->check(where);
    });
}

int gmm_planned_tod2map(float *local_map_gpu, int64_t local_map_size,
                        const float *tod_gpu, int64_t tod_nsamp,
                        const float *xpointing_gpu, int64_t xpointing_nsamp,
                        const gmm_lpix *lpix, const gmm_plan *plan,
                        int partial_pixelization, int debug, int check_errors, void *stream_)
{
    return guarded([&] {
        const char *where = "tod2map";
        hipStream_t stream = as_stream(stream_);
        require_ptr(plan, where, "plan");
        require_ptr(tod_gpu, where, "tod");
        require_ptr(local_map_gpu, where, "local_map");
        require_ptr(xpointing_gpu, where, "xpointing");
        check_op_common(lpix, plan, tod_nsamp, xpointing_nsamp, local_map_size, where);
        require_ptr(lpix->offsets_gpu, where, "cell_offsets_gpu");
        OpArgs a{lpix, tod_nsamp, partial_pixelization != 0};
        std::unique_ptr<OpDebug> dbg(debug ? new OpDebug(plan->pp.plan_nmt, tod_nsamp, stream) : nullptr);
        a.debug = dbg ? &dbg->bufs : nullptr;
        launch_planned_tod2map(local_map_gpu, tod_gpu, xpointing_gpu, a, *plan, TodGains{}, stream);
        if (debug || check_errors)
            check_plan_errors(*plan, where, stream);
        if (dbg)
            dbg->check(where);
    });
}


int gmm_response_map2tod(float *tod_gpu, int64_t tod_nsamp,
                         const float *local_map_gpu, int64_t local_map_size,
                         const float *xpointing_gpu, int64_t xpointing_nsamp,
                         const float *response_gpu, int64_t ndet,
                         const gmm_lpix *lpix, const gmm_plan *plan,
                         int partial_pixelization, int debug, int accum, int check_errors,
                         void *stream_)
{
    return guarded([&] {
        const char *where = "response_map2tod";
        hipStream_t stream = as_stream(stream_);
        require_ptr(plan, where, "plan");
        require_ptr(tod_gpu, where, "tod");
        require_ptr(local_map_gpu, where, "local_map");
        require_ptr(xpointing_gpu, where, "xpointing");
        check_op_common(lpix, plan, tod_nsamp, xpointing_nsamp, local_map_size, where);
        check_response(response_gpu, ndet, tod_nsamp, where);
        require_ptr(lpix->offsets_gpu, where, "cell_offsets_gpu");
        OpArgs a{lpix, tod_nsamp, partial_pixelization != 0};
        std::unique_ptr<OpDebug> dbg(debug ? new OpDebug(plan->pp.plan_nmt, tod_nsamp, stream) : nullptr);
        a.debug = dbg ? &dbg->bufs : nullptr;
        TodGains g;
        g.response = response_gpu;
        g.resp_ndet = ndet;
        launch_planned_map2tod(tod_gpu, local_map_gpu, xpointing_gpu, a, *plan, accum != 0, g, stream);
        if (debug || check_errors)
            check_plan_errors(*plan, where, stream);
        if (dbg)
            dbg->check(where);
    });
}

int gmm_response_tod2map(float *local_map_gpu, int64_t local_map_size,
                         const float *tod_gpu, int64_t tod_nsamp,
                         const float *xpointing_gpu, int64_t xpointing_nsamp,
                         const float *response_gpu, int64_t ndet,
                         const gmm_lpix *lpix, const gmm_plan *plan,
                         int partial_pixelization, int debug, int check_errors, void *stream_)
{
    return guarded([&] {
        const char *where = "response_tod2map";
        hipStream_t stream = as_stream(stream_);
        require_ptr(plan, where, "plan");
        require_ptr(tod_gpu, where, "tod");
        require_ptr(local_map_gpu, where, "local_map");
        require_ptr(xpointing_gpu, where, "xpointing");
        check_op_common(lpix, plan, tod_nsamp, xpointing_nsamp, local_map_size, where);
        check_response(response_gpu, ndet, tod_nsamp, where);
        require_ptr(lpix->offsets_gpu, where, "cell_offsets_gpu");
        OpArgs a{lpix, tod_nsamp, partial_pixelization != 0};
        std::unique_ptr<OpDebug> dbg(debug ? new OpDebug(plan->pp.plan_nmt, tod_nsamp, stream) : nullptr);
        a.debug = dbg ? &dbg->bufs : nullptr;
        TodGains g;
        g.response = response_gpu;
        g.resp_ndet = ndet;
        launch_planned_tod2map(local_map_gpu, tod_gpu, xpointing_gpu, a, *plan, g, stream);
        if (debug || check_errors)
            check_plan_errors(*plan, where, stream);
        if (dbg)
            dbg->check(where);
    });
}

namespace {
void expand_dynamic(uint32_t *global_ncells_gpu, int64_t *cell_offsets_gpu, int64_t nycells, int64_t nxcells,
                    const uint64_t *plan_mt_gpu, int64_t nmt, uint32_t *new_ncells, hipStream_t stream)
{
    void *scratch = nullptr;
    const int64_t nb = expand_dynamic_scratch_nbytes(nmt);
    GMM_HIP(hipMalloc(&scratch, size_t(nb)));
    try {
        launch_expand_dynamic(global_ncells_gpu, cell_offsets_gpu, nycells, nxcells, plan_mt_gpu, nmt, scratch, stream);
        uint32_t n = 0;
        GMM_HIP(hipMemcpyAsync(&n, global_ncells_gpu, 4, hipMemcpyDeviceToHost, stream));
        GMM_HIP(hipStreamSynchronize(stream));
        if (new_ncells)
            *new_ncells = n;
    }
    catch (...) {
        (void) hipFree(scratch);
        throw;
    }
    GMM_HIP(hipFree(scratch));
}
}  // namespace

int gmm_expand_dynamic_map(uint32_t *global_ncells_gpu, int64_t *cell_offsets_gpu, int64_t nycells, int64_t nxcells,
                           const uint64_t *plan_mt_gpu, int64_t nmt, uint32_t *new_ncells, void *stream_)
{
    return guarded([&] {
        const char *where = "expand_dynamic_map";
        require_ptr(global_ncells_gpu, where, "global_ncells");
        require_ptr(cell_offsets_gpu, where, "cell_offsets");
        if (nycells <= 0 || nycells > 1024 || nxcells <= 0 || nxcells > 1024)
            fail(std::string(where) + ": expected 0 < nycells, nxcells <= 1024");
        if (nmt < 0 || nmt > int64_t(UINT32_MAX) - 1)
            fail(std::string(where) + ": bad nmt");
        if (nmt > 0)
            require_ptr(plan_mt_gpu, where, "plan_mt");
        expand_dynamic(global_ncells_gpu, cell_offsets_gpu, nycells, nxcells, plan_mt_gpu, nmt, new_ncells,
                       as_stream(stream_));
    });
}

int gmm_expand_dynamic_map_noalloc(uint32_t *global_ncells_gpu, int64_t *cell_offsets_gpu, int64_t nycells,
                                   int64_t nxcells, const uint64_t *plan_mt_gpu, int64_t nmt, uint32_t *new_ncells,
                                   void *stream_)
{
    return guarded([&] {
        const char *where = "expand_dynamic_map_noalloc";
        require_ptr(global_ncells_gpu, where, "global_ncells");
        require_ptr(cell_offsets_gpu, where, "cell_offsets");
        if (nycells <= 0 || nycells > 1024 || nxcells <= 0 || nxcells > 1024)
            fail(std::string(where) + ": expected 0 < nycells, nxcells <= 1024");
        if (nycells * nxcells > expand_dynamic_noalloc_max_cells())
            fail(std::string(where) + ": the cell grid exceeds the scratch array's capacity of "
                 + std::to_string(expand_dynamic_noalloc_max_cells()) + " cells");
        if (nmt < 0 || nmt > int64_t(UINT32_MAX) - 1)
            fail(std::string(where) + ": bad nmt");
        if (nmt > 0)
            require_ptr(plan_mt_gpu, where, "plan_mt");
        hipStream_t stream = as_stream(stream_);
        launch_expand_dynamic_noalloc(global_ncells_gpu, cell_offsets_gpu, nycells, nxcells, plan_mt_gpu, nmt, stream);
        if (new_ncells) {
            uint32_t n = 0;
            GMM_HIP(hipMemcpyAsync(&n, global_ncells_gpu, 4, hipMemcpyDeviceToHost, stream));
            GMM_HIP(hipStreamSynchronize(stream));
            *new_ncells = n;
        }
    });
}

int gmm_lpix_expand_dynamic(gmm_lpix *lpix, const gmm_plan *plan, uint32_t *global_ncells_gpu, uint32_t *new_ncells,
                            void *stream_)
{
    return guarded([&] {
        const char *where = "expand_dynamic_map2";
        require_ptr(lpix, where, "lpix");
        require_ptr(plan, where, "plan");
        require_ptr(global_ncells_gpu, where, "global_ncells");
        require_ptr(lpix->offsets_gpu, where, "cell_offsets_gpu");
        const gmm_preplan &pp = plan->pp;
        if (lpix->nypix_global != pp.nypix_global || lpix->nxpix_global != pp.nxpix_global
            || lpix->periodic_xcoord != pp.periodic_xcoord)
            fail(std::string(where) + ": LocalPixelization and PointingPlan disagree on (nypix_global, nxpix_global, periodic_xcoord)");
        if (lpix->nycells != (pp.nypix_global + 63) / 64 || lpix->nxcells != (pp.nxpix_global + 63) / 64)
            fail(std::string(where) + ": a dynamic map needs the full (nycells, nxcells) cell table");
        expand_dynamic(global_ncells_gpu, const_cast<int64_t *>(lpix->offsets_gpu), lpix->nycells, lpix->nxcells,
                       plan->plan_mt, pp.plan_nmt, new_ncells, as_stream(stream_));
    });
}


int gmm_planned_map2tod_q16(void *tod_gpu, int tod_type, const float *tod_scale_gpu, int64_t tod_nrows,
                            int64_t tod_nsamp, const float *local_map_gpu, int64_t local_map_size,
                            const float *xpointing_gpu, int64_t xpointing_nsamp,
                            const float *response_gpu, int64_t resp_ndet,
                            const gmm_lpix *lpix, const gmm_plan *plan,
                            int partial_pixelization, int debug, int accum, int check_errors, void *stream_)
{
    return guarded([&] {
        const char *where = "map2tod (16-bit TOD)";
        hipStream_t stream = as_stream(stream_);
        require_ptr(plan, where, "plan");
        require_ptr(tod_gpu, where, "tod");
        require_ptr(local_map_gpu, where, "local_map");
        require_ptr(xpointing_gpu, where, "xpointing");
        check_op_common(lpix, plan, tod_nsamp, xpointing_nsamp, local_map_size, where);
        TodGains g = q16_gains(tod_type, tod_gpu, tod_scale_gpu, tod_nrows, tod_nsamp, response_gpu, resp_ndet, where);
        require_ptr(lpix->offsets_gpu, where, "cell_offsets_gpu");
        OpArgs a{lpix, tod_nsamp, partial_pixelization != 0};
        launch_planned_map2tod(tod_gpu, local_map_gpu, xpointing_gpu, a, *plan, accum != 0, g, stream);
        if (debug || check_errors)
            check_plan_errors(*plan, where, stream);
    });
}

int gmm_planned_tod2map_q16(float *local_map_gpu, int64_t local_map_size,
                            const void *tod_gpu, int tod_type, const float *tod_scale_gpu, int64_t tod_nrows,
                            int64_t tod_nsamp, const float *xpointing_gpu, int64_t xpointing_nsamp,
                            const float *response_gpu, int64_t resp_ndet,
                            const gmm_lpix *lpix, const gmm_plan *plan,
                            int partial_pixelization, int debug, int check_errors, void *stream_)
{
    return guarded([&] {
        const char *where = "tod2map (16-bit TOD)";
        hipStream_t stream = as_stream(stream_);
        require_ptr(plan, where, "plan");
        require_ptr(tod_gpu, where, "tod");
        require_ptr(local_map_gpu, where, "local_map");
        require_ptr(xpointing_gpu, where, "xpointing");
        check_op_common(lpix, plan, tod_nsamp, xpointing_nsamp, local_map_size, where);
        TodGains g = q16_gains(tod_type, tod_gpu, tod_scale_gpu, tod_nrows, tod_nsamp, response_gpu, resp_ndet, where);
        require_ptr(lpix->offsets_gpu, where, "cell_offsets_gpu");
        OpArgs a{lpix, tod_nsamp, partial_pixelization != 0};
        launch_planned_tod2map(local_map_gpu, tod_gpu, xpointing_gpu, a, *plan, g, stream);
        if (debug || check_errors)
            check_plan_errors(*plan, where, stream);
    });
}

int gmm_tod16_encode(void *out_gpu, int tod_type, float *scale_gpu, const float *tod_gpu, int64_t nrows,
                     int64_t row_len, int error_feedback, void *stream_)
{
    return guarded([&] {
        const char *where = "tod16_encode";
        require_ptr(out_gpu, where, "out");
        require_ptr(scale_gpu, where, "scale");
        require_ptr(tod_gpu, where, "tod");
        if (tod_type != GMM_TOD_F16 && tod_type != GMM_TOD_BF16)
            fail(std::string(where) + ": tod_type must be GMM_TOD_F16 or GMM_TOD_BF16");
        if (nrows < 0 || row_len < 0 || nrows > int64_t(INT32_MAX))
            fail(std::string(where) + ": bad shape");
        if (error_feedback < 0 || error_feedback > 2)
            fail(std::string(where) + ": error_feedback must be 0, 1 or 2");
        if (nrows == 0 || row_len == 0)
            return;
        launch_tod16_encode(out_gpu, tod_type == GMM_TOD_F16 ? TOD_F16 : TOD_BF16, scale_gpu, tod_gpu, nrows,
                            row_len, error_feedback, as_stream(stream_));
    });
}


int gmm_cell_broadcast(float *dst_gpu, int64_t ncells, const float *src_gpu, int64_t nsrc_cells,
                       const int64_t *index_map_gpu, void *stream_)
{
    return guarded([&] {
        const char *where = "cell_broadcast";
        if (ncells < 0 || nsrc_cells < 0 || ncells > INT32_MAX)
            fail(std::string(where) + ": bad cell counts");
        if (ncells == 0)
            return;
        check_cells(dst_gpu, where, "dst");
        check_cells(src_gpu, where, "src");
        require_ptr(index_map_gpu, where, "index_map");
        hipStream_t stream = as_stream(stream_);
        with_bad_word(where, stream, [&](uint32_t *bad) {
            launch_cell_broadcast(dst_gpu, ncells, src_gpu, nsrc_cells, index_map_gpu, bad, stream);
        });
    });
}

int gmm_cell_reduce(float *dst_gpu, int64_t ndst_cells, const float *src_gpu, int64_t nsrc_cells,
                    const int64_t *index_map_gpu, void *stream_)
{
    return guarded([&] {
        const char *where = "cell_reduce";
        if (ndst_cells < 0 || nsrc_cells < 0 || ndst_cells > INT32_MAX)
            fail(std::string(where) + ": bad cell counts");
        if (ndst_cells == 0)
            return;
        check_cells(dst_gpu, where, "dst");
        if (nsrc_cells > 0) {
            check_cells(src_gpu, where, "src");
            require_ptr(index_map_gpu, where, "index_map");
        }
        hipStream_t stream = as_stream(stream_);
        with_bad_word(where, stream, [&](uint32_t *bad) {
            launch_cell_reduce(dst_gpu, ndst_cells, src_gpu, nsrc_cells, index_map_gpu, bad, stream);
        });
    });
}

int gmm_local_map_to_global(const gmm_lpix *lpix, float *dst_gpu, const float *src_gpu, int64_t src_size,
                            void *stream_)
{
    return guarded([&] {
        const char *where = "local_map_to_global";
        require_ptr(lpix, where, "lpix");
        require_ptr(dst_gpu, where, "dst");
        require_ptr(src_gpu, where, "src");
        require_ptr(lpix->offsets_gpu, where, "cell_offsets_gpu");
        if (src_size != 3 * lpix->npix)
            fail(std::string(where) + ": expected local map of size 3*npix=" + std::to_string(3 * lpix->npix)
                 + ", got " + std::to_string(src_size));
        launch_local_to_global(dst_gpu, src_gpu, *lpix, as_stream(stream_));
    });
}

namespace {
void check_unplanned_errflags(const uint32_t *errflags, int64_t nblocks, const char *where,
                              hipStream_t stream)
{
    std::vector<uint32_t> h(static_cast<size_t>(nblocks), 0u);
    GMM_HIP(hipMemcpyAsync(h.data(), errflags, size_t(nblocks) * 4, hipMemcpyDeviceToHost, stream));
    GMM_HIP(hipStreamSynchronize(stream));
    uint32_t e = 0;
    for (uint32_t v : h)
        e |= v;
    throw_errflags(e, where);
}
}  // namespace

int gmm_unplanned_map2tod(float *tod_gpu, int64_t tod_nsamp,
                          const float *local_map_gpu, int64_t local_map_size,
                          const float *xpointing_gpu, int64_t xpointing_nsamp,
                          const gmm_lpix *lpix, uint32_t *errflags_gpu, int64_t nerrflags,
                          int partial_pixelization, int accum, int check_errors, void *stream_)
{
    return guarded([&] {
        const char *where = "unplanned_map2tod";
        hipStream_t stream = as_stream(stream_);
        require_ptr(tod_gpu, where, "tod");
        require_ptr(local_map_gpu, where, "local_map");
        require_ptr(xpointing_gpu, where, "xpointing");
        require_ptr(errflags_gpu, where, "errflags");
        check_op_common(lpix, nullptr, tod_nsamp, xpointing_nsamp, local_map_size, where);
        require_ptr(lpix->offsets_gpu, where, "cell_offsets_gpu");
        if (nerrflags <= 0)
            fail(std::string(where) + ": errflags array must be non-empty");
        OpArgs a{lpix, tod_nsamp, partial_pixelization != 0};
        launch_unplanned_map2tod(tod_gpu, local_map_gpu, xpointing_gpu, a, errflags_gpu, nerrflags, accum != 0, stream);
        if (check_errors)
            check_unplanned_errflags(errflags_gpu, unplanned_nblocks(tod_nsamp, nerrflags), where, stream);
    });
}

int gmm_unplanned_tod2map(float *local_map_gpu, int64_t local_map_size,
                          const float *tod_gpu, int64_t tod_nsamp,
                          const float *xpointing_gpu, int64_t xpointing_nsamp,
                          const gmm_lpix *lpix, uint32_t *errflags_gpu, int64_t nerrflags,
                          int partial_pixelization, int check_errors, void *stream_)
{
    return guarded([&] {
        const char *where = "unplanned_tod2map";
        hipStream_t stream = as_stream(stream_);
        require_ptr(tod_gpu, where, "tod");
        require_ptr(local_map_gpu, where, "local_map");
        require_ptr(xpointing_gpu, where, "xpointing");
        require_ptr(errflags_gpu, where, "errflags");
        check_op_common(lpix, nullptr, tod_nsamp, xpointing_nsamp, local_map_size, where);
        require_ptr(lpix->offsets_gpu, where, "cell_offsets_gpu");
        if (nerrflags <= 0)
            fail(std::string(where) + ": errflags array must be non-empty");
        OpArgs a{lpix, tod_nsamp, partial_pixelization != 0};
        launch_unplanned_tod2map(local_map_gpu, tod_gpu, xpointing_gpu, a, errflags_gpu, nerrflags, stream);
        if (check_errors)
            check_unplanned_errflags(errflags_gpu, unplanned_nblocks(tod_nsamp, nerrflags), where, stream);
    });
}

int gmm_reference_map2tod(float *tod_cpu, int64_t tod_nsamp,
                          const float *local_map_cpu, int64_t local_map_size,
                          const float *xpointing_cpu, int64_t xpointing_nsamp,
                          const gmm_lpix *lpix, int partial_pixelization, int accum)
{
    return guarded([&] {
        const char *where = "reference_map2tod";
        require_ptr(tod_cpu, where, "tod");
        require_ptr(local_map_cpu, where, "local_map");
        require_ptr(xpointing_cpu, where, "xpointing");
        check_op_common(lpix, nullptr, tod_nsamp, xpointing_nsamp, local_map_size, where);
        OpArgs a{lpix, tod_nsamp, partial_pixelization != 0};
        reference_map2tod(tod_cpu, local_map_cpu, xpointing_cpu, a, accum != 0);
    });
}

int gmm_reference_tod2map(float *local_map_cpu, int64_t local_map_size,
                          const float *tod_cpu, int64_t tod_nsamp,
                          const float *xpointing_cpu, int64_t xpointing_nsamp,
                          const gmm_lpix *lpix, int partial_pixelization)
{
    return guarded([&] {
        const char *where = "reference_tod2map";
        require_ptr(tod_cpu, where, "tod");
        require_ptr(local_map_cpu, where, "local_map");
        require_ptr(xpointing_cpu, where, "xpointing");
        check_op_common(lpix, nullptr, tod_nsamp, xpointing_nsamp, local_map_size, where);
        OpArgs a{lpix, tod_nsamp, partial_pixelization != 0};
        reference_tod2map(local_map_cpu, tod_cpu, xpointing_cpu, a);
    });
}

// ------------------------------------------------------------------------------------------

int gmm_toy_pointing_fill(float *xpointing_cpu, int64_t nsamp, int64_t nypix_global,
                          int64_t nxpix_global, double scan_speed, double total_drift,
                          double *drift_speed_out)
{
    return guarded([&] {
        const char *where = "ToyPointing constructor";
        require_ptr(xpointing_cpu, where, "xpointing_cpu");
        check_nsamp(nsamp, where);
        check_npix_global(nypix_global, "nypix_global", where);
        check_npix_global(nxpix_global, "nxpix_global", where);
        double d = toy_pointing_fill(xpointing_cpu, nsamp, nypix_global, nxpix_global, scan_speed, total_drift);
        if (drift_speed_out)
            *drift_speed_out = d;
    });
}

int64_t gmm_plan_tester_tmp_nbytes(const gmm_preplan *pp)
{
    if (!pp)
        return -1;
    return align128(4 * pp->nsamp) * 2 + 128;
}

int gmm_plan_tester_run(const gmm_preplan *pp, const float *xpointing_gpu, int64_t nsamp,
                        void *tmp_gpu, int64_t tmp_nbytes,
                        int32_t *iypix_cpu, int32_t *ixpix_cpu,
                        uint32_t *nmt_cumsum_cpu, uint64_t *sorted_mt_cpu, void *stream_)
{
    return guarded([&] {
        const char *where = "PointingPlanTester constructor";
        hipStream_t stream = as_stream(stream_);
        require_ptr(pp, where, "preplan");
        require_ptr(xpointing_gpu, where, "xpointing_gpu");
        require_ptr(tmp_gpu, where, "tmp");
        require_ptr(iypix_cpu, where, "iypix");
        require_ptr(ixpix_cpu, where, "ixpix");
        require_ptr(nmt_cumsum_cpu, where, "nmt_cumsum");
        require_ptr(sorted_mt_cpu, where, "sorted_mt");
        if (nsamp != pp->nsamp)
            fail(std::string(where) + ": xpointing and preplan disagree on nsamp");
        if (tmp_nbytes < gmm_plan_tester_tmp_nbytes(pp))
            fail(std::string(where) + ": 'tmp' buffer too small");
        if (reinterpret_cast<uintptr_t>(tmp_gpu) % 128)
            fail(std::string(where) + ": 'tmp' must be 128-byte aligned");

        uint8_t *t = reinterpret_cast<uint8_t *>(tmp_gpu);
        int32_t *iy = reinterpret_cast<int32_t *>(t);
        int32_t *ix = reinterpret_cast<int32_t *>(t + align128(4 * nsamp));
        uint32_t *ew = reinterpret_cast<uint32_t *>(t + 2 * align128(4 * nsamp));
        launch_quantize(xpointing_gpu, nsamp, int(pp->nypix_global), int(pp->nxpix_global),
                        pp->periodic_xcoord, iy, ix, ew, stream);
        uint32_t e = read_word(ew, stream);
        throw_errflags(e, "PointingPlanTester constructor (quantize_kernel)");
        GMM_HIP(hipMemcpyAsync(iypix_cpu, iy, size_t(nsamp) * 4, hipMemcpyDeviceToHost, stream));
        GMM_HIP(hipMemcpyAsync(ixpix_cpu, ix, size_t(nsamp) * 4, hipMemcpyDeviceToHost, stream));
        GMM_HIP(hipStreamSynchronize(stream));
        plan_tester_host(*pp, iypix_cpu, ixpix_cpu, nmt_cumsum_cpu, sorted_mt_cpu);
    });
}

int gmm_test_plan_iterator(const uint64_t *plan_mt_cpu, int64_t nmt, int64_t nmt_per_block,
                           int waves_per_block)
{
    return guarded([&] {
        const char *where = "test_plan_iterator";
        require_ptr(plan_mt_cpu, where, "plan_mt");
        if (nmt <= 0 || nmt > (int64_t(1) << 30))
            fail(std::string(where) + ": expected 0 < nmt <= 2^30");
        if (nmt_per_block <= 0)
            fail(std::string(where) + ": expected nmt_per_block > 0");
        if (waves_per_block < 1 || waves_per_block > 16)
            fail(std::string(where) + ": expected 1 <= waves_per_block <= 16");
        for (int64_t i = 1; i < nmt; i++)
            if ((plan_mt_cpu[i - 1] & CELL_MASK) > (plan_mt_cpu[i] & CELL_MASK))
                fail(std::string(where) + ": plan_mt is not sorted by cell");

        const int64_t max_chunks = (nmt + nmt_per_block - 1) / nmt_per_block + std::min<int64_t>(nmt, int64_t(1) << 20);
        const int64_t scratch = build_chunks_generic_scratch(nmt);
        const int64_t bytes = align128(8 * nmt) + align128(int64_t(sizeof(Chunk)) * max_chunks)
                            + align128(4 * nmt) + 128 + scratch;
        void *mem = nullptr;
        GMM_HIP(hipMalloc(&mem, size_t(bytes)));
        std::string failure;
        try {
            uint8_t *m = reinterpret_cast<uint8_t *>(mem);
            uint64_t *d_mt = reinterpret_cast<uint64_t *>(m);
            Chunk *d_ch = reinterpret_cast<Chunk *>(m + align128(8 * nmt));
            uint32_t *d_vis = reinterpret_cast<uint32_t *>(m + align128(8 * nmt) + align128(int64_t(sizeof(Chunk)) * max_chunks));
            uint32_t *d_err = d_vis + nmt;
            uint8_t *d_scr = m + align128(8 * nmt) + align128(int64_t(sizeof(Chunk)) * max_chunks) + align128(4 * nmt) + 128;
            GMM_HIP(hipMemcpy(d_mt, plan_mt_cpu, size_t(nmt) * 8, hipMemcpyHostToDevice));
            GMM_HIP(hipMemset(d_vis, 0, size_t(nmt) * 4));
            int64_t nch = build_chunks_generic(d_mt, nmt, nmt_per_block, d_ch, max_chunks, d_scr, scratch, nullptr);
            GMM_REQUIRE(nch <= max_chunks, "test_plan_iterator: chunk count exceeds bound");
            launch_iterator_test(d_mt, d_ch, nch, waves_per_block, d_vis, d_err, nullptr);
            std::vector<uint32_t> vis(static_cast<size_t>(nmt), 0u);
            uint32_t e = 0;
            GMM_HIP(hipMemcpy(vis.data(), d_vis, size_t(nmt) * 4, hipMemcpyDeviceToHost));
            GMM_HIP(hipMemcpy(&e, d_err, 4, hipMemcpyDeviceToHost));
            if (e)
                failure = (e & 1u) ? "test_plan_iterator: a block visited an entry of another cell"
                                   : "test_plan_iterator: a wave visited an entry twice";
            for (int64_t i = 0; i < nmt && failure.empty(); i++)
                if (vis[size_t(i)] != 1) {
                    std::ostringstream ss;
                    ss << "test_plan_iterator: entry " << i << " visited " << vis[size_t(i)] << " times";
                    failure = ss.str();
                }
        }
        catch (...) {
            (void)hipFree(mem);
            throw;
        }
        GMM_HIP(hipFree(mem));
        if (!failure.empty())
            fail(failure);
    });
}

int gmm_make_random_plan_mt(int64_t ncells, int64_t min_nmt_per_cell, int64_t max_nmt_per_cell,
                            uint64_t seed, uint64_t *out_cpu, int64_t capacity, int64_t *nmt_out)
{
    return guarded([&] {
        std::vector<uint64_t> v = make_random_plan_mt(ncells, min_nmt_per_cell, max_nmt_per_cell, seed);
        if (nmt_out)
            *nmt_out = int64_t(v.size());
        if (out_cpu) {
            if (capacity < int64_t(v.size()))
                fail("make_random_plan_mt: output capacity too small");
            std::memcpy(out_cpu, v.data(), v.size() * 8);
        }
    });
}

}  // extern "C"
