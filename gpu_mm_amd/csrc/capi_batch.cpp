// C ABI of the batched operators (include/gpu_mm_amd.h, "Batched operators"): binding of
// several TODs with their own plans to one local map, construction of the cross-TOD work items,
// launches, the autotuner and HIP graph capture. No counterpart in the reference, which runs
// one TOD per launch (src_lib/map2tod.cu:198-272, src_lib/tod2map.cu:160-226).

#include "gmm_internal.hpp"
#include "gpu_mm_amd.h"

#include <algorithm>
#include <memory>
#include <sstream>

using namespace gmm;

namespace {

// Plan entries [begin, end) of TOD 'tod' that belong to cell 'icell'.
struct CellRun
{
    uint32_t icell, tod, begin, end;
};

struct DeviceTables
{
    BatchSeg *segs = nullptr;
    BatchItem *items = nullptr;
    int64_t nsegs = 0, nitems = 0, grid = 0;
    int threads = 512;

    void release()
    {
        if (segs)
            (void)hipFree(segs);
        if (items)
            (void)hipFree(items);
        segs = nullptr;
        items = nullptr;
        nsegs = nitems = grid = 0;
    }
};

}  // namespace

struct gmm_batch
{
    const gmm_lpix *lpix = nullptr;
    int tod_type = TOD_F32;
    std::vector<BatchTod> tods;
    std::vector<CellRun> runs;      // sorted by (icell, tod)
    int64_t nentries = 0, ncells = 0;
    gmm_batch_config cfg{};
    int ncu = 256;

    BatchTod *tods_dev = nullptr;
    uint32_t *words = nullptr;      // [0,2) map2tod queue, [2,4) tod2map queue, [4] error word
    DeviceTables m2t, t2m;
    bool has_resp = false;          // every TOD has per-detector gains (gmm_batch_set_response)
    int64_t generation = 0;

    ~gmm_batch()
    {
        m2t.release();
        t2m.release();
        if (tods_dev)
            (void)hipFree(tods_dev);
        if (words)
            (void)hipFree(words);
    }
};

struct gmm_graph
{
    const gmm_batch *batch = nullptr;
    int64_t generation = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    ~gmm_graph()
    {
        if (exec)
            (void)hipGraphExecDestroy(exec);
        if (graph)
            (void)hipGraphDestroy(graph);
    }
};

namespace {

constexpr int64_t MIN_ITEM_ENTRIES = 64;
constexpr int64_t MAX_ITEM_ENTRIES = int64_t(1) << 24;

void check_config(const gmm_batch_config &c, const char *where)
{
    auto bad = [&](const char *what) { fail(std::string(where) + ": " + what); };
    if (c.m2t_item_entries < MIN_ITEM_ENTRIES || c.m2t_item_entries > MAX_ITEM_ENTRIES
        || c.t2m_item_entries < MIN_ITEM_ENTRIES || c.t2m_item_entries > MAX_ITEM_ENTRIES)
        bad("item_entries must be in [64, 2^24]");
    if (c.m2t_threads != 256 && c.m2t_threads != 512)
        bad("m2t_threads must be 256 or 512");
    if (c.m2t_blocks_per_cu < 0 || c.m2t_blocks_per_cu > 16 || c.t2m_blocks_per_cu < 0 || c.t2m_blocks_per_cu > 16)
        bad("blocks_per_cu must be in [0, 16]");
}

// Work items over the cell runs: the entries of each cell (all TODs, in TOD order) are cut
// into ceil(total / target) nearly equal items; items of a cell that is cut are flagged
// CHUNK_SHARED (their tod2map flushes add atomically). Items stay in cell order for a grid of
// one block per item (the kernel spreads runs of neighbouring cells over the XCDs, as the plain
// operators do), and are ordered by decreasing size for a persistent grid, which takes them in
// order and so starts the long ones first.
void build_items(const std::vector<CellRun> &runs, int64_t target, bool by_size, std::vector<BatchSeg> &segs,
                 std::vector<BatchItem> &items)
{
    segs.clear();
    items.clear();
    for (size_t g0 = 0; g0 < runs.size();) {
        size_t g1 = g0;
        int64_t total = 0;
        while (g1 < runs.size() && runs[g1].icell == runs[g0].icell) {
            total += runs[g1].end - runs[g1].begin;
            g1++;
        }
        const int64_t nparts = std::max<int64_t>(1, (total + target - 1) / target);
        for (int64_t p = 0; p < nparts; p++) {
            const int64_t lo = (total * p) / nparts, hi = (total * (p + 1)) / nparts;
            BatchItem it{};
            it.icell = runs[g0].icell;
            it.seg_begin = uint32_t(segs.size());
            it.nent = uint32_t(hi - lo);
            it.flags = (nparts > 1) ? CHUNK_SHARED : 0u;
            int64_t pos = 0;
            for (size_t r = g0; r < g1; r++) {
                const int64_t len = runs[r].end - runs[r].begin;
                const int64_t a = std::max(lo, pos), b = std::min(hi, pos + len);
                if (a < b)
                    segs.push_back(BatchSeg{runs[r].tod, uint32_t(runs[r].begin + (a - pos)),
                                            uint32_t(runs[r].begin + (b - pos)), 0u});
                pos += len;
            }
            it.seg_end = uint32_t(segs.size());
            if (it.nent > 0)
                items.push_back(it);
        }
        g0 = g1;
    }
    if (by_size)
        std::stable_sort(items.begin(), items.end(),
                         [](const BatchItem &a, const BatchItem &b) { return a.nent > b.nent; });
}

void upload_tables(gmm_batch &b, DeviceTables &t, int64_t target, int blocks_per_cu, int threads)
{
    std::vector<BatchSeg> segs;
    std::vector<BatchItem> items;
    build_items(b.runs, target, blocks_per_cu > 0, segs, items);
    if (segs.size() >= (size_t(1) << 32) || items.size() >= (size_t(1) << 31))
        fail("batch: too many work items");
    t.release();
    if (!segs.empty()) {
        GMM_HIP(hipMalloc(&t.segs, segs.size() * sizeof(BatchSeg)));
        GMM_HIP(hipMemcpy(t.segs, segs.data(), segs.size() * sizeof(BatchSeg), hipMemcpyHostToDevice));
    }
    if (!items.empty()) {
        GMM_HIP(hipMalloc(&t.items, items.size() * sizeof(BatchItem)));
        GMM_HIP(hipMemcpy(t.items, items.data(), items.size() * sizeof(BatchItem), hipMemcpyHostToDevice));
    }
    t.nsegs = int64_t(segs.size());
    t.nitems = int64_t(items.size());
    t.grid = (blocks_per_cu > 0) ? std::min<int64_t>(int64_t(blocks_per_cu) * b.ncu, t.nitems) : 0;
    t.threads = threads;
}

void configure(gmm_batch &b, const gmm_batch_config &c)
{
    check_config(c, "batch configure");
    upload_tables(b, b.m2t, c.m2t_item_entries, c.m2t_blocks_per_cu, c.m2t_threads);
    upload_tables(b, b.t2m, c.t2m_item_entries, c.t2m_blocks_per_cu, 1024);
    b.cfg = c;
    b.generation++;
}

BatchLaunch launch_of(const gmm_batch &b, const DeviceTables &t, int op)
{
    BatchLaunch bl;
    bl.tods = b.tods_dev;
    bl.segs = t.segs;
    bl.items = t.items;
    bl.nitems = t.nitems;
    bl.counters = b.words + ((op == GMM_OP_MAP2TOD) ? 0 : 2);
    bl.err_word = b.words + 4;
    bl.grid = t.grid;
    bl.threads = t.threads;
    bl.has_resp = b.has_resp;
    return bl;
}

void check_map_size(const gmm_batch &b, int64_t local_map_size, const char *where)
{
    if (local_map_size != 3 * b.lpix->npix) {
        std::ostringstream ss;
        ss << where << ": expected local map of size 3*npix=" << 3 * b.lpix->npix << ", got " << local_map_size;
        fail(ss.str());
    }
}

void run_op(gmm_batch &b, int op, float *lmap, bool partial, bool accum, hipStream_t stream)
{
    if (op == GMM_OP_MAP2TOD)
        launch_batched_map2tod(launch_of(b, b.m2t, op), lmap, *b.lpix, b.tod_type, partial, accum, stream);
    else
        launch_batched_tod2map(launch_of(b, b.t2m, op), lmap, *b.lpix, b.tod_type, partial, stream);
}

void check_batch_errors(gmm_batch &b, const char *where, hipStream_t stream)
{
    const uint32_t e = read_word(b.words + 4, stream);
    if (e) {
        clear_word(b.words + 4, stream);
        throw_errflags(e, where);
    }
}

// Median time of 'reps' launches of one operator (after one untimed launch), in ms. The launches
// run with partial_pixelization set, which changes only the error flags: a batch meant to be run
// with it must not be left holding "not in LocalPixelization" flags from the timing runs.
double time_op(gmm_batch &b, int op, float *lmap, int reps, hipStream_t stream)
{
    run_op(b, op, lmap, true, false, stream);
    std::vector<hipEvent_t> ev(size_t(2 * reps));
    for (auto &e : ev)
        GMM_HIP(hipEventCreate(&e));
    std::vector<double> ms;
    try {
        for (int r = 0; r < reps; r++) {
            GMM_HIP(hipEventRecord(ev[size_t(2 * r)], stream));
            run_op(b, op, lmap, true, false, stream);
            GMM_HIP(hipEventRecord(ev[size_t(2 * r + 1)], stream));
        }
        GMM_HIP(hipStreamSynchronize(stream));
        for (int r = 0; r < reps; r++) {
            float t = 0.0f;
            GMM_HIP(hipEventElapsedTime(&t, ev[size_t(2 * r)], ev[size_t(2 * r + 1)]));
            ms.push_back(double(t));
        }
    }
    catch (...) {
        for (auto &e : ev)
            (void)hipEventDestroy(e);
        throw;
    }
    for (auto &e : ev)
        GMM_HIP(hipEventDestroy(e));
    std::sort(ms.begin(), ms.end());
    return ms[ms.size() / 2];
}

}  // namespace

extern "C" {

void gmm_batch_default_config(gmm_batch_config *cfg)
{
    if (!cfg)
        return;
    *cfg = gmm_batch_config{};
    cfg->m2t_item_entries = 2048;
    cfg->m2t_threads = 512;
    cfg->m2t_blocks_per_cu = 2;
    cfg->t2m_item_entries = 4096;
    cfg->t2m_blocks_per_cu = 1;
}

int gmm_batch_create(gmm_batch **out, const gmm_lpix *lpix, int32_t ntod, const gmm_plan *const *plans,
                     const float *const *xpointing_gpu, void *const *tod_gpu, const int64_t *nsamp,
                     int32_t tod_type, const float *const *tod_scale_gpu, const int64_t *tod_nrows,
                     const gmm_batch_config *config)
{
    return guarded([&] {
        const char *where = "batch create";
        require_ptr(out, where, "out");
        *out = nullptr;
        require_ptr(lpix, where, "lpix");
        require_ptr(lpix->offsets_gpu, where, "cell_offsets_gpu");
        if (ntod <= 0)
            fail(std::string(where) + ": the batch needs at least one TOD");
        require_ptr(plans, where, "plans");
        require_ptr(xpointing_gpu, where, "xpointing");
        require_ptr(tod_gpu, where, "tod");
        require_ptr(nsamp, where, "nsamp");
        if (tod_type != GMM_TOD_F32 && tod_type != GMM_TOD_F16 && tod_type != GMM_TOD_BF16)
            fail(std::string(where) + ": tod_type must be GMM_TOD_F32, GMM_TOD_F16 or GMM_TOD_BF16");
        const bool q16 = (tod_type != GMM_TOD_F32);
        if (q16) {
            require_ptr(tod_scale_gpu, where, "tod_scale");
            require_ptr(tod_nrows, where, "tod_nrows");
        }

        std::unique_ptr<gmm_batch> b(new gmm_batch);
        b->lpix = lpix;
        b->tod_type = (tod_type == GMM_TOD_F32) ? TOD_F32 : (tod_type == GMM_TOD_F16) ? TOD_F16 : TOD_BF16;
        int dev = 0;
        GMM_HIP(hipGetDevice(&dev));
        GMM_HIP(hipDeviceGetAttribute(&b->ncu, hipDeviceAttributeMultiprocessorCount, dev));

        for (int32_t i = 0; i < ntod; i++) {
            std::ostringstream w;
            w << where << " (TOD " << i << ")";
            const std::string ws = w.str();
            const gmm_plan *pl = plans[i];
            require_ptr(pl, ws.c_str(), "plan");
            require_ptr(xpointing_gpu[i], ws.c_str(), "xpointing");
            require_ptr(tod_gpu[i], ws.c_str(), "tod");
            check_nsamp(nsamp[i], ws.c_str());
            check_nsamp_quads(nsamp[i], ws.c_str());
            if (pl->pp.nsamp != nsamp[i])
                fail(ws + ": plan was built for nsamp=" + std::to_string(pl->pp.nsamp) + ", got "
                     + std::to_string(nsamp[i]));
            if (pl->pp.nypix_global != lpix->nypix_global || pl->pp.nxpix_global != lpix->nxpix_global
                || pl->pp.periodic_xcoord != lpix->periodic_xcoord)
                fail(ws + ": LocalPixelization and PointingPlan disagree on (nypix_global, nxpix_global, periodic_xcoord)");
            if ((reinterpret_cast<uintptr_t>(xpointing_gpu[i]) & 15)
                || (reinterpret_cast<uintptr_t>(tod_gpu[i]) & (q16 ? 7 : 15)))
                fail(ws + ": xpointing must be 16-byte aligned, and the TOD 16-byte (float32) or 8-byte (16-bit) aligned");
            BatchTod t{};
            t.xp = xpointing_gpu[i];
            t.tod = tod_gpu[i];
            t.plan_mt = pl->plan_mt;
            t.qmask = pl->qmask;
            t.nsamp = nsamp[i];
            t.resp_ndet = 1;
            if (q16) {
                require_ptr(tod_scale_gpu[i], ws.c_str(), "tod_scale");
                const int64_t nr = tod_nrows[i];
                if (nr <= 0 || nsamp[i] % nr != 0)
                    fail(ws + ": tod_scale has " + std::to_string(nr) + " rows, which does not divide nsamp="
                         + std::to_string(nsamp[i]));
                t.scale = tod_scale_gpu[i];
                t.scale_div = make_rowdiv(uint32_t(nsamp[i] / nr));
            }
            b->tods.push_back(t);

            // Cell runs of this plan: its chunks (stored largest first) back in plan order, merged
            // per cell.
            std::vector<Chunk> ch(size_t(pl->nchunks));
            if (!ch.empty())
                GMM_HIP(hipMemcpy(ch.data(), pl->chunks, ch.size() * sizeof(Chunk), hipMemcpyDeviceToHost));
            std::sort(ch.begin(), ch.end(), [](const Chunk &x, const Chunk &y) { return x.begin < y.begin; });
            for (const Chunk &c : ch) {
                if (!b->runs.empty() && b->runs.back().tod == uint32_t(i) && b->runs.back().icell == c.icell
                    && b->runs.back().end == c.begin)
                    b->runs.back().end = c.end;
                else
                    b->runs.push_back(CellRun{c.icell, uint32_t(i), c.begin, c.end});
                b->nentries += int64_t(c.end) - int64_t(c.begin);
            }
        }
        std::stable_sort(b->runs.begin(), b->runs.end(), [](const CellRun &x, const CellRun &y) {
            return (x.icell != y.icell) ? (x.icell < y.icell) : (x.tod < y.tod);
        });
        for (size_t r = 0; r < b->runs.size(); r++)
            if (r == 0 || b->runs[r].icell != b->runs[r - 1].icell)
                b->ncells++;

        GMM_HIP(hipMalloc(&b->tods_dev, b->tods.size() * sizeof(BatchTod)));
        GMM_HIP(hipMemcpy(b->tods_dev, b->tods.data(), b->tods.size() * sizeof(BatchTod), hipMemcpyHostToDevice));
        GMM_HIP(hipMalloc(&b->words, 8 * sizeof(uint32_t)));
        GMM_HIP(hipMemset(b->words, 0, 8 * sizeof(uint32_t)));

        gmm_batch_config c;
        gmm_batch_default_config(&c);
        configure(*b, config ? *config : c);
        *out = b.release();
    });
}

void gmm_batch_destroy(gmm_batch *b) { delete b; }

int gmm_batch_set_response(gmm_batch *b, const float *const *response_gpu, const int64_t *resp_ndet)
{
    return guarded([&] {
        const char *where = "batch set_response";
        require_ptr(b, where, "batch");
        std::vector<BatchTod> tods = b->tods;
        if (response_gpu) {
            require_ptr(resp_ndet, where, "resp_ndet");
            if (b->tod_type != TOD_F32)
                fail(std::string(where) + ": per-detector gains are supported for float32 TODs only");
            for (size_t i = 0; i < tods.size(); i++) {
                const std::string ws = std::string(where) + " (TOD " + std::to_string(i) + ")";
                check_response(response_gpu[i], resp_ndet[i], tods[i].nsamp, ws.c_str());
                tods[i].resp = response_gpu[i];
                tods[i].resp_ndet = uint32_t(resp_ndet[i]);
                tods[i].resp_div = make_rowdiv(uint32_t(tods[i].nsamp / resp_ndet[i]));
            }
        }
        else {
            for (BatchTod &t : tods) {
                t.resp = nullptr;
                t.resp_ndet = 1;
                t.resp_div = RowDiv();
            }
        }
        GMM_HIP(hipMemcpy(b->tods_dev, tods.data(), tods.size() * sizeof(BatchTod), hipMemcpyHostToDevice));
        b->tods = tods;
        b->has_resp = (response_gpu != nullptr);
        b->generation++;
    });
}

int gmm_batch_configure(gmm_batch *b, const gmm_batch_config *config)
{
    return guarded([&] {
        require_ptr(b, "batch configure", "batch");
        require_ptr(config, "batch configure", "config");
        GMM_HIP(hipDeviceSynchronize());    // no launch may still read the old tables
        configure(*b, *config);
    });
}

int gmm_batch_get_info(const gmm_batch *b, gmm_batch_info *out)
{
    return guarded([&] {
        require_ptr(b, "batch info", "batch");
        require_ptr(out, "batch info", "out");
        gmm_batch_info i{};
        i.ntod = int32_t(b->tods.size());
        i.tod_type = (b->tod_type == TOD_F32) ? GMM_TOD_F32 : (b->tod_type == TOD_F16) ? GMM_TOD_F16 : GMM_TOD_BF16;
        i.config = b->cfg;
        i.nentries = b->nentries;
        i.ncells = b->ncells;
        i.m2t_nitems = b->m2t.nitems;
        i.m2t_nsegs = b->m2t.nsegs;
        i.m2t_grid = (b->m2t.grid > 0) ? b->m2t.grid : b->m2t.nitems;
        i.t2m_nitems = b->t2m.nitems;
        i.t2m_nsegs = b->t2m.nsegs;
        i.t2m_grid = (b->t2m.grid > 0) ? b->t2m.grid : b->t2m.nitems;
        i.generation = b->generation;
        *out = i;
    });
}

int gmm_batch_map2tod(gmm_batch *b, const float *local_map_gpu, int64_t local_map_size, int partial_pixelization,
                      int accum, int check_errors, void *stream_)
{
    return guarded([&] {
        const char *where = "batch map2tod";
        require_ptr(b, where, "batch");
        require_ptr(local_map_gpu, where, "local_map");
        check_map_size(*b, local_map_size, where);
        hipStream_t stream = as_stream(stream_);
        run_op(*b, GMM_OP_MAP2TOD, const_cast<float *>(local_map_gpu), partial_pixelization != 0, accum != 0,
               stream);
        if (check_errors)
            check_batch_errors(*b, where, stream);
    });
}

int gmm_batch_tod2map(gmm_batch *b, float *local_map_gpu, int64_t local_map_size, int partial_pixelization,
                      int check_errors, void *stream_)
{
    return guarded([&] {
        const char *where = "batch tod2map";
        require_ptr(b, where, "batch");
        require_ptr(local_map_gpu, where, "local_map");
        check_map_size(*b, local_map_size, where);
        hipStream_t stream = as_stream(stream_);
        run_op(*b, GMM_OP_TOD2MAP, local_map_gpu, partial_pixelization != 0, false, stream);
        if (check_errors)
            check_batch_errors(*b, where, stream);
    });
}

int gmm_batch_check_errflags(gmm_batch *b, const char *where)
{
    return guarded([&] {
        require_ptr(b, "batch check_errflags", "batch");
        check_batch_errors(*b, where ? where : "batch", nullptr);
    });
}

int gmm_batch_autotune(gmm_batch *b, const float *map_in_gpu, float *map_scratch_gpu, int64_t local_map_size,
                       int32_t reps, gmm_batch_tune_entry *table, int32_t table_capacity, int32_t *table_size,
                       void *stream_)
{
    return guarded([&] {
        const char *where = "batch autotune";
        require_ptr(b, where, "batch");
        require_ptr(map_in_gpu, where, "map_in");
        require_ptr(map_scratch_gpu, where, "map_scratch");
        check_map_size(*b, local_map_size, where);
        if (reps < 1 || reps > 100)
            fail(std::string(where) + ": reps must be in [1, 100]");
        hipStream_t stream = as_stream(stream_);
        GMM_HIP(hipStreamSynchronize(stream));
        int32_t n = 0;
        auto record = [&](int op, int entries, int threads, int bpc, double ms) {
            if (table && n < table_capacity)
                table[n] = gmm_batch_tune_entry{op, entries, threads, bpc, ms};
            n++;
        };
        static const int ENTRIES[] = {1024, 2048, 4096, 8192, 16384};
        static const int M2T_BPC[] = {0, 1, 2, 3, 4};
        static const int T2M_BPC[] = {0, 1, 2};
        const gmm_batch_config start = b->cfg;
        gmm_batch_config best = start;
        double best_m2t = 1e30, best_t2m = 1e30;
        try {
            // map2tod: work-item size x block size x persistent grid.
            for (int e : ENTRIES)
                for (int th : {256, 512})
                    for (int bpc : M2T_BPC) {
                        DeviceTables &t = b->m2t;
                        upload_tables(*b, t, e, bpc, th);
                        const double ms = time_op(*b, GMM_OP_MAP2TOD, const_cast<float *>(map_in_gpu), reps, stream);
                        record(GMM_OP_MAP2TOD, e, th, bpc, ms);
                        if (ms < best_m2t) {
                            best_m2t = ms;
                            best.m2t_item_entries = e;
                            best.m2t_threads = th;
                            best.m2t_blocks_per_cu = bpc;
                        }
                    }
            // tod2map: work-item size x persistent grid (1024 threads, one block per CU fits).
            for (int e : ENTRIES)
                for (int bpc : T2M_BPC) {
                    upload_tables(*b, b->t2m, e, bpc, 1024);
                    const double ms = time_op(*b, GMM_OP_TOD2MAP, map_scratch_gpu, reps, stream);
                    record(GMM_OP_TOD2MAP, e, 1024, bpc, ms);
                    if (ms < best_t2m) {
                        best_t2m = ms;
                        best.t2m_item_entries = e;
                        best.t2m_blocks_per_cu = bpc;
                    }
                }
        }
        catch (...) {
            (void)hipDeviceSynchronize();
            configure(*b, start);
            throw;
        }
        GMM_HIP(hipStreamSynchronize(stream));
        configure(*b, best);
        if (table_size)
            *table_size = n;
    });
}

int gmm_batch_graph_create(gmm_graph **out, gmm_batch *b, int32_t op, float *local_map_gpu, int64_t local_map_size,
                           int partial_pixelization, int accum)
{
    return guarded([&] {
        const char *where = "batch graph";
        require_ptr(out, where, "out");
        *out = nullptr;
        require_ptr(b, where, "batch");
        require_ptr(local_map_gpu, where, "local_map");
        check_map_size(*b, local_map_size, where);
        if (op != GMM_OP_MAP2TOD && op != GMM_OP_TOD2MAP)
            fail(std::string(where) + ": op must be GMM_OP_MAP2TOD or GMM_OP_TOD2MAP");
        std::unique_ptr<gmm_graph> g(new gmm_graph);
        g->batch = b;
        g->generation = b->generation;
        hipStream_t s = nullptr;
        GMM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        try {
            GMM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            try {
                run_op(*b, op, local_map_gpu, partial_pixelization != 0, accum != 0, s);
            }
            catch (...) {
                hipGraph_t dropped = nullptr;
                (void)hipStreamEndCapture(s, &dropped);
                if (dropped)
                    (void)hipGraphDestroy(dropped);
                throw;
            }
            GMM_HIP(hipStreamEndCapture(s, &g->graph));
            GMM_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
        }
        catch (...) {
            (void)hipStreamDestroy(s);
            throw;
        }
        GMM_HIP(hipStreamDestroy(s));
        *out = g.release();
    });
}

int gmm_graph_launch(gmm_graph *g, void *stream_)
{
    return guarded([&] {
        require_ptr(g, "graph launch", "graph");
        if (g->batch->generation != g->generation)
            fail("graph launch: the batch was reconfigured after this graph was captured");
        GMM_HIP(hipGraphLaunch(g->exec, as_stream(stream_)));
    });
}

void gmm_graph_destroy(gmm_graph *g) { delete g; }

}  // extern "C"
