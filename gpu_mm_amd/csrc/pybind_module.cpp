// pybind11 module _gpu_mm_amd_py: a thin layer over the C ABI of libgpu_mm_amd.so.
// Arrays (torch tensors on the GPU, numpy arrays or torch tensors on the host) arrive through
// the DLPack protocol; this file checks shape, dtype, placement and contiguity, then passes
// plain pointers and sizes down. It mirrors the reference's pybind11 module
// (src_pybind11/gpu_mm_pybind11.cu:29-272) for the pointing-operator subset.

#include <pybind11/pybind11.h>
#include <pybind11/numpy.h>

#include "gpu_mm_amd.h"

#include <algorithm>
#include <cstdint>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

namespace py = pybind11;

namespace {

// ------------------------------------------------------------------------------------------
// DLPack ABI (https://dmlc.github.io/dlpack/latest/c_api.html), legacy unversioned capsule.

struct DLDevice { int32_t device_type; int32_t device_id; };
struct DLDataType { uint8_t code; uint8_t bits; uint16_t lanes; };
struct DLTensor
{
    void *data;
    DLDevice device;
    int32_t ndim;
    DLDataType dtype;
    int64_t *shape;
    int64_t *strides;
    uint64_t byte_offset;
};
struct DLManagedTensor
{
    DLTensor dl_tensor;
    void *manager_ctx;
    void (*deleter)(DLManagedTensor *);
};

enum { kDLCPU = 1, kDLCUDA = 2, kDLCUDAHost = 3, kDLROCM = 10, kDLROCMHost = 11 };
enum { kDLInt = 0, kDLUInt = 1, kDLFloat = 2, kDLBfloat = 4 };

[[noreturn]] void raise(const std::string &s) { throw std::runtime_error(s); }

void check(int status)
{
    if (status != 0)
        raise(gmm_last_error());
}

struct Arr
{
    py::object capsule;   // keeps the producer's tensor alive while the pointer is used
    void *ptr = nullptr;
    std::vector<int64_t> shape;
    int64_t size = 1;
    bool on_gpu = false;
    bool contiguous = true;
    uint8_t code = 0, bits = 0;
};

Arr get_arr(py::handle obj, const char *where, const char *name)
{
    if (obj.is_none())
        raise(std::string(where) + ": '" + name + "' is None");
    if (!py::hasattr(obj, "__dlpack__"))
        raise(std::string(where) + ": '" + name + "' does not support the DLPack protocol (expected a torch tensor or numpy array)");

    Arr a;
    a.capsule = obj.attr("__dlpack__")();
    PyObject *c = a.capsule.ptr();
    if (!PyCapsule_IsValid(c, "dltensor"))
        raise(std::string(where) + ": '" + name + "': unsupported DLPack capsule");
    auto *mt = static_cast<DLManagedTensor *>(PyCapsule_GetPointer(c, "dltensor"));
    const DLTensor &t = mt->dl_tensor;

    a.ptr = static_cast<char *>(t.data) + t.byte_offset;
    a.shape.assign(t.shape, t.shape + t.ndim);
    a.code = t.dtype.code;
    a.bits = t.dtype.bits;
    const int dt = t.device.device_type;
    a.on_gpu = (dt == kDLROCM || dt == kDLCUDA);
    for (int64_t s : a.shape)
        a.size *= s;
    if (t.strides) {
        int64_t expect = 1;
        for (int i = t.ndim - 1; i >= 0; i--) {
            if (a.shape[size_t(i)] != 1 && t.strides[i] != expect)
                a.contiguous = false;
            expect *= a.shape[size_t(i)];
        }
    }
    if (t.dtype.lanes != 1)
        raise(std::string(where) + ": '" + name + "': vector dtypes are not supported");
    return a;
}

std::string shape_str(const Arr &a)
{
    std::ostringstream ss;
    ss << "(";
    for (size_t i = 0; i < a.shape.size(); i++)
        ss << (i ? "," : "") << a.shape[i];
    ss << (a.shape.size() == 1 ? ",)" : ")");
    return ss.str();
}

void require(const Arr &a, const char *where, const char *name, uint8_t code, uint8_t bits,
             bool on_gpu, const char *dtype_name)
{
    if (a.code != code || a.bits != bits)
        raise(std::string(where) + ": expected " + name + " to have dtype " + dtype_name);
    if (a.on_gpu != on_gpu)
        raise(std::string(where) + ": expected " + name + " to be in " + (on_gpu ? "GPU" : "CPU") + " memory");
    if (!a.contiguous)
        raise(std::string(where) + ": expected " + name + " to be contiguous");
}

// TOD: (nsamp,) or (ndet,nt); xpointing: (3,nsamp) or (3,ndet,nt)
// (src_lib/check_arguments.cu:244-286).
int64_t tod_nsamp(const Arr &a, const char *where)
{
    if (a.shape.size() != 1 && a.shape.size() != 2)
        raise(std::string(where) + ": expected tod to have shape (nsamp,) or (ndet,nt), got " + shape_str(a));
    return a.size;
}

int64_t xp_nsamp(const Arr &a, const char *where)
{
    if ((a.shape.size() != 2 && a.shape.size() != 3) || a.shape[0] != 3)
        raise(std::string(where) + ": expected xpointing to have shape (3,nsamp) or (3,ndet,nt), got " + shape_str(a));
    return a.size / 3;
}

Arr tod_arr(py::handle o, const char *where, bool gpu)
{
    Arr a = get_arr(o, where, "tod");
    require(a, where, "tod", kDLFloat, 32, gpu, "float32");
    tod_nsamp(a, where);
    return a;
}

Arr xp_arr(py::handle o, const char *where, bool gpu)
{
    Arr a = get_arr(o, where, "xpointing");
    require(a, where, "xpointing", kDLFloat, 32, gpu, "float32");
    xp_nsamp(a, where);
    return a;
}

Arr map_arr(py::handle o, const char *where, bool gpu)
{
    Arr a = get_arr(o, where, "local_map");
    require(a, where, "local_map", kDLFloat, 32, gpu, "float32");
    return a;
}

void *as_stream(uintptr_t s) { return reinterpret_cast<void *>(s); }

// ------------------------------------------------------------------------------------------

struct LocalPixelization
{
    gmm_lpix *h = nullptr;
    py::object cpu_ref, gpu_ref;

    LocalPixelization(int64_t nypix_global, int64_t nxpix_global, py::object cell_offsets_cpu,
                      py::object cell_offsets_gpu, int64_t ystride, int64_t polstride, bool periodic_xcoord)
    {
        const char *where = "LocalPixelization constructor";
        Arr c = get_arr(cell_offsets_cpu, where, "cell_offsets_cpu");
        require(c, where, "cell_offsets_cpu", kDLInt, 64, false, "int64");
        if (c.shape.size() != 2)
            raise(std::string(where) + ": expected cell_offsets to be a 2-d array, got shape " + shape_str(c));
        const int64_t *gptr = nullptr;
        if (!cell_offsets_gpu.is_none()) {
            Arr g = get_arr(cell_offsets_gpu, where, "cell_offsets_gpu");
            require(g, where, "cell_offsets_gpu", kDLInt, 64, true, "int64");
            if (g.shape != c.shape)
                raise(std::string(where) + ": cell_offsets_cpu and cell_offsets_gpu have different shapes");
            gptr = static_cast<const int64_t *>(g.ptr);
        }
        check(gmm_lpix_create(&h, nypix_global, nxpix_global, static_cast<const int64_t *>(c.ptr), gptr,
                              c.shape[0], c.shape[1], ystride, polstride, periodic_xcoord ? 1 : 0));
        cpu_ref = cell_offsets_cpu;
        gpu_ref = cell_offsets_gpu;
    }
    ~LocalPixelization() { gmm_lpix_destroy(h); }

    gmm_lpix_info info() const
    {
        gmm_lpix_info i;
        check(gmm_lpix_get_info(h, &i));
        return i;
    }
};

struct PointingPrePlan
{
    gmm_preplan *h = nullptr;
    py::object nmt_ref;

    PointingPrePlan(py::object xpointing_gpu, int64_t nypix_global, int64_t nxpix_global,
                    py::object nmt_gpu, py::object err_gpu, bool periodic_xcoord, bool debug, uintptr_t stream)
    {
        const char *where = "PointingPrePlan constructor";
        Arr xp = xp_arr(xpointing_gpu, where, true);
        Arr n = get_arr(nmt_gpu, where, "buf");
        Arr e = get_arr(err_gpu, where, "tmp_buf");
        require(n, where, "buf", kDLUInt, 32, true, "uint32");
        require(e, where, "tmp_buf", kDLUInt, 32, true, "uint32");
        if (n.size != GMM_PREPLAN_SIZE || e.size != GMM_PREPLAN_SIZE)
            raise(std::string(where) + ": expected 'buf' and 'tmp_buf' to have length PointingPrePlan.preplan_size");
        int64_t nsamp = xp.size / 3;
        {
            py::gil_scoped_release nogil;
            check(gmm_preplan_create(&h, static_cast<const float *>(xp.ptr), nsamp, nypix_global, nxpix_global,
                                     static_cast<uint32_t *>(n.ptr), static_cast<uint32_t *>(e.ptr),
                                     periodic_xcoord ? 1 : 0, debug ? 1 : 0, as_stream(stream)));
        }
        nmt_ref = nmt_gpu;
    }
    ~PointingPrePlan() { gmm_preplan_destroy(h); }

    gmm_preplan_info info() const
    {
        gmm_preplan_info i;
        check(gmm_preplan_get_info(h, &i));
        return i;
    }
};

struct PointingPlan
{
    gmm_plan *h = nullptr;
    py::object buf_ref;

    PointingPlan(const PointingPrePlan &pp, py::object xpointing_gpu, py::object buf, py::object tmp_buf,
                 bool debug, uintptr_t stream)
    {
        const char *where = "PointingPlan constructor";
        Arr xp = xp_arr(xpointing_gpu, where, true);
        Arr b = get_arr(buf, where, "buf");
        Arr t = get_arr(tmp_buf, where, "tmp_buf");
        require(b, where, "buf", kDLUInt, 8, true, "uint8");
        require(t, where, "tmp_buf", kDLUInt, 8, true, "uint8");
        if (b.shape.size() != 1 || t.shape.size() != 1)
            raise(std::string(where) + ": expected 'buf' and 'tmp_buf' to be 1-d arrays");
        {
            py::gil_scoped_release nogil;
            check(gmm_plan_create(&h, pp.h, static_cast<const float *>(xp.ptr), xp.size / 3, b.ptr, b.size,
                                  t.ptr, t.size, debug ? 1 : 0, as_stream(stream)));
        }
        buf_ref = buf;
    }
    ~PointingPlan() { gmm_plan_destroy(h); }

    gmm_plan_info info() const
    {
        gmm_plan_info i;
        check(gmm_plan_get_info(h, &i));
        return i;
    }
};

// ------------------------------------------------------------------------------------------

void planned_map2tod(py::object tod, py::object local_map, py::object xpointing, const LocalPixelization &lpix,
                     const PointingPlan &plan, bool partial_pixelization, bool debug, bool accum, bool check_errors,
                     uintptr_t stream)
{
    const char *where = "map2tod";
    Arr t = tod_arr(tod, where, true), m = map_arr(local_map, where, true), x = xp_arr(xpointing, where, true);
    py::gil_scoped_release nogil;
    check(gmm_planned_map2tod(static_cast<float *>(t.ptr), t.size, static_cast<const float *>(m.ptr), m.size,
                              static_cast<const float *>(x.ptr), x.size / 3, lpix.h, plan.h,
                              partial_pixelization, debug, accum, check_errors, as_stream(stream)));
}

void planned_tod2map(py::object local_map, py::object tod, py::object xpointing, const LocalPixelization &lpix,
                     const PointingPlan &plan, bool partial_pixelization, bool debug, bool check_errors, uintptr_t stream)
{
    const char *where = "tod2map";
    Arr m = map_arr(local_map, where, true), t = tod_arr(tod, where, true), x = xp_arr(xpointing, where, true);
    py::gil_scoped_release nogil;
    check(gmm_planned_tod2map(static_cast<float *>(m.ptr), m.size, static_cast<const float *>(t.ptr), t.size,
                              static_cast<const float *>(x.ptr), x.size / 3, lpix.h, plan.h,
                              partial_pixelization, debug, check_errors, as_stream(stream)));
}

// Response array: float32 (2, ndet) on the GPU, C-contiguous (src_lib/check_arguments.cu:171-180).
Arr response_arr(py::handle o, const char *where)
{
    Arr r = get_arr(o, where, "response");
    require(r, where, "response", kDLFloat, 32, true, "float32");
    if (r.shape.size() != 2 || r.shape[0] != 2)
        raise(std::string(where) + ": expected 'response' to have shape (2, ndet)");
    return r;
}

void response_map2tod(py::object tod, py::object local_map, py::object xpointing, py::object response,
                      const LocalPixelization &lpix, const PointingPlan &plan, bool partial_pixelization, bool debug,
                      bool accum, bool check_errors, uintptr_t stream)
{
    const char *where = "response_map2tod";
    Arr t = tod_arr(tod, where, true), m = map_arr(local_map, where, true), x = xp_arr(xpointing, where, true);
    Arr r = response_arr(response, where);
    py::gil_scoped_release nogil;
    check(gmm_response_map2tod(static_cast<float *>(t.ptr), t.size, static_cast<const float *>(m.ptr), m.size,
                               static_cast<const float *>(x.ptr), x.size / 3, static_cast<const float *>(r.ptr),
                               r.shape[1], lpix.h, plan.h, partial_pixelization, debug, accum, check_errors,
                               as_stream(stream)));
}

void response_tod2map(py::object local_map, py::object tod, py::object xpointing, py::object response,
                      const LocalPixelization &lpix, const PointingPlan &plan, bool partial_pixelization, bool debug,
                      bool check_errors, uintptr_t stream)
{
    const char *where = "response_tod2map";
    Arr m = map_arr(local_map, where, true), t = tod_arr(tod, where, true), x = xp_arr(xpointing, where, true);
    Arr r = response_arr(response, where);
    py::gil_scoped_release nogil;
    check(gmm_response_tod2map(static_cast<float *>(m.ptr), m.size, static_cast<const float *>(t.ptr), t.size,
                               static_cast<const float *>(x.ptr), x.size / 3, static_cast<const float *>(r.ptr),
                               r.shape[1], lpix.h, plan.h, partial_pixelization, debug, check_errors,
                               as_stream(stream)));
}

// 16-bit TOD: float16 or bfloat16 on the GPU, (nsamp,) or (ndet,nt), with its float32 per-row
// scale array of shape (nrows,).
struct Tod16
{
    Arr t, sc;
    int type;
};

Tod16 tod16_args(py::handle tod, py::handle scale, const char *where)
{
    Tod16 r;
    r.t = get_arr(tod, where, "tod");
    if (r.t.code == kDLFloat && r.t.bits == 16)
        r.type = GMM_TOD_F16;
    else if (r.t.code == kDLBfloat && r.t.bits == 16)
        r.type = GMM_TOD_BF16;
    else
        raise(std::string(where) + ": expected a 16-bit tod to have dtype float16 or bfloat16");
    if (!r.t.on_gpu)
        raise(std::string(where) + ": expected tod to be in GPU memory");
    if (!r.t.contiguous)
        raise(std::string(where) + ": expected tod to be contiguous");
    tod_nsamp(r.t, where);
    r.sc = get_arr(scale, where, "tod_scale");
    require(r.sc, where, "tod_scale", kDLFloat, 32, true, "float32");
    if (r.sc.shape.size() != 1)
        raise(std::string(where) + ": expected tod_scale to be a 1-d array");
    if (r.t.shape.size() == 2 && r.sc.shape[0] != r.t.shape[0])
        raise(std::string(where) + ": expected tod_scale to have one entry per TOD row");
    return r;
}

void planned_map2tod_q16(py::object tod, py::object tod_scale, py::object local_map, py::object xpointing,
                         const LocalPixelization &lpix, const PointingPlan &plan, bool partial_pixelization,
                         bool debug, bool accum, py::object response, bool check_errors, uintptr_t stream)
{
    const char *where = "map2tod (16-bit TOD)";
    Tod16 t = tod16_args(tod, tod_scale, where);
    Arr m = map_arr(local_map, where, true), x = xp_arr(xpointing, where, true);
    Arr r;
    if (!response.is_none())
        r = response_arr(response, where);
    py::gil_scoped_release nogil;
    check(gmm_planned_map2tod_q16(t.t.ptr, t.type, static_cast<const float *>(t.sc.ptr), t.sc.size, t.t.size,
                                  static_cast<const float *>(m.ptr), m.size, static_cast<const float *>(x.ptr),
                                  x.size / 3, static_cast<const float *>(r.ptr), r.ptr ? r.shape[1] : 0, lpix.h,
                                  plan.h, partial_pixelization, debug, accum, check_errors, as_stream(stream)));
}

void planned_tod2map_q16(py::object local_map, py::object tod, py::object tod_scale, py::object xpointing,
                         const LocalPixelization &lpix, const PointingPlan &plan, bool partial_pixelization,
                         bool debug, py::object response, bool check_errors, uintptr_t stream)
{
    const char *where = "tod2map (16-bit TOD)";
    Tod16 t = tod16_args(tod, tod_scale, where);
    Arr m = map_arr(local_map, where, true), x = xp_arr(xpointing, where, true);
    Arr r;
    if (!response.is_none())
        r = response_arr(response, where);
    py::gil_scoped_release nogil;
    check(gmm_planned_tod2map_q16(static_cast<float *>(m.ptr), m.size, t.t.ptr, t.type,
                                  static_cast<const float *>(t.sc.ptr), t.sc.size, t.t.size,
                                  static_cast<const float *>(x.ptr), x.size / 3, static_cast<const float *>(r.ptr),
                                  r.ptr ? r.shape[1] : 0, lpix.h, plan.h, partial_pixelization, debug, check_errors,
                                  as_stream(stream)));
}

// ------------------------------------------------------------------------------------------
// Batched operators.

struct Batch
{
    gmm_batch *h = nullptr;
    py::list refs;            // the pixelization, plans and arrays the batch points into
    int64_t map_size = 0;

    Batch(const LocalPixelization &lpix, py::object lpix_obj, py::list plans, py::list xpointings, py::list tods,
          py::object scales, py::object config)
    {
        const char *where = "PlanBatch constructor";
        const size_t n = py::len(plans);
        if (n == 0 || py::len(xpointings) != n || py::len(tods) != n)
            raise(std::string(where) + ": plans, xpointings and tods must be non-empty lists of equal length");
        const bool q16 = !scales.is_none();
        if (q16 && py::len(scales) != n)
            raise(std::string(where) + ": scales must be None or a list with one entry per TOD");
        std::vector<const gmm_plan *> ph;
        std::vector<const float *> xph, sch;
        std::vector<void *> th;
        std::vector<int64_t> ns, nrows;
        int tod_type = GMM_TOD_F32;
        refs.append(lpix_obj);
        for (size_t i = 0; i < n; i++) {
            const PointingPlan &pl = plans[i].cast<const PointingPlan &>();
            Arr x = xp_arr(xpointings[i], where, true);
            int64_t tn;
            if (q16) {
                Tod16 t = tod16_args(tods[i], py::list(scales)[i], where);
                if (i > 0 && t.type != tod_type)
                    raise(std::string(where) + ": all TODs of a batch must have the same dtype");
                tod_type = t.type;
                tn = t.t.size;
                th.push_back(t.t.ptr);
                sch.push_back(static_cast<const float *>(t.sc.ptr));
                nrows.push_back(t.sc.size);
            }
            else {
                Arr t = tod_arr(tods[i], where, true);
                tn = t.size;
                th.push_back(t.ptr);
            }
            if (tn != x.size / 3)
                raise(std::string(where) + ": TOD " + std::to_string(i) + " has " + std::to_string(tn)
                      + " samples but its xpointing has " + std::to_string(x.size / 3));
            ph.push_back(pl.h);
            xph.push_back(static_cast<const float *>(x.ptr));
            ns.push_back(tn);
            refs.append(plans[i]);
            refs.append(xpointings[i]);
            refs.append(tods[i]);
            if (q16)
                refs.append(py::list(scales)[i]);
        }
        gmm_batch_config c;
        gmm_batch_default_config(&c);
        if (!config.is_none())
            c = config.cast<gmm_batch_config>();
        {
            py::gil_scoped_release nogil;
            check(gmm_batch_create(&h, lpix.h, int32_t(n), ph.data(), xph.data(), th.data(), ns.data(), tod_type,
                                   q16 ? sch.data() : nullptr, q16 ? nrows.data() : nullptr, &c));
        }
        map_size = 3 * lpix.info().npix;
    }
    ~Batch() { gmm_batch_destroy(h); }

    // Per-detector gains of every TOD (a list of (2, ndet) float32 arrays), or None.
    void set_response(py::object responses)
    {
        const char *where = "PlanBatch.set_response";
        py::list keep;
        std::vector<const float *> rp;
        std::vector<int64_t> nd;
        if (!responses.is_none()) {
            py::list l(responses);
            if (int64_t(py::len(l)) != info().ntod)
                raise(std::string(where) + ": expected one response array per TOD");
            for (size_t i = 0; i < py::len(l); i++) {
                Arr r = response_arr(l[i], where);
                rp.push_back(static_cast<const float *>(r.ptr));
                nd.push_back(r.shape[1]);
                keep.append(l[i]);
            }
        }
        {
            py::gil_scoped_release nogil;
            check(gmm_batch_set_response(h, responses.is_none() ? nullptr : rp.data(),
                                         responses.is_none() ? nullptr : nd.data()));
        }
        resp_refs = keep;
    }
    py::list resp_refs;       // the response arrays the batch points into

    gmm_batch_info info() const
    {
        gmm_batch_info i;
        check(gmm_batch_get_info(h, &i));
        return i;
    }

    Arr map(py::handle m, const char *where) const
    {
        Arr a = map_arr(m, where, true);
        if (a.size != map_size)
            raise(std::string(where) + ": expected local map of size 3*npix=" + std::to_string(map_size) + ", got "
                  + std::to_string(a.size));
        return a;
    }
};

struct Graph
{
    gmm_graph *h = nullptr;
    py::object batch_ref, map_ref;
    ~Graph() { gmm_graph_destroy(h); }
};

void tod16_encode(py::object out, py::object scale, py::object tod, int error_feedback, uintptr_t stream)
{
    const char *where = "tod16_encode";
    Tod16 o = tod16_args(out, scale, where);
    Arr t = tod_arr(tod, where, true);
    if (t.shape != o.t.shape)
        raise(std::string(where) + ": expected out and tod to have the same shape");
    const int64_t nrows = (t.shape.size() == 2) ? t.shape[0] : 1;
    if (o.sc.size != nrows)
        raise(std::string(where) + ": expected scale to have one entry per TOD row");
    py::gil_scoped_release nogil;
    check(gmm_tod16_encode(o.t.ptr, o.type, static_cast<float *>(o.sc.ptr), static_cast<const float *>(t.ptr), nrows,
                           nrows ? t.size / nrows : 0, error_feedback, as_stream(stream)));
}

Arr errflags_arr(py::handle o, const char *where)
{
    Arr e = get_arr(o, where, "errflags");
    require(e, where, "errflags", kDLUInt, 32, true, "uint32");
    if (e.shape.size() != 1)
        raise(std::string(where) + ": expected errflags to be a 1-d array");
    return e;
}

void unplanned_map2tod(py::object tod, py::object local_map, py::object xpointing, const LocalPixelization &lpix,
                       py::object errflags, bool partial_pixelization, bool accum, bool check_errors, uintptr_t stream)
{
    const char *where = "unplanned_map2tod";
    Arr t = tod_arr(tod, where, true), m = map_arr(local_map, where, true), x = xp_arr(xpointing, where, true);
    Arr e = errflags_arr(errflags, where);
    py::gil_scoped_release nogil;
    check(gmm_unplanned_map2tod(static_cast<float *>(t.ptr), t.size, static_cast<const float *>(m.ptr), m.size,
                                static_cast<const float *>(x.ptr), x.size / 3, lpix.h,
                                static_cast<uint32_t *>(e.ptr), e.size, partial_pixelization, accum,
                                check_errors, as_stream(stream)));
}

void unplanned_tod2map(py::object local_map, py::object tod, py::object xpointing, const LocalPixelization &lpix,
                       py::object errflags, bool partial_pixelization, bool check_errors, uintptr_t stream)
{
    const char *where = "unplanned_tod2map";
    Arr m = map_arr(local_map, where, true), t = tod_arr(tod, where, true), x = xp_arr(xpointing, where, true);
    Arr e = errflags_arr(errflags, where);
    py::gil_scoped_release nogil;
    check(gmm_unplanned_tod2map(static_cast<float *>(m.ptr), m.size, static_cast<const float *>(t.ptr), t.size,
                                static_cast<const float *>(x.ptr), x.size / 3, lpix.h,
                                static_cast<uint32_t *>(e.ptr), e.size, partial_pixelization,
                                check_errors, as_stream(stream)));
}

void reference_map2tod(py::object tod, py::object local_map, py::object xpointing, const LocalPixelization &lpix,
                       bool partial_pixelization, bool accum)
{
    const char *where = "reference_map2tod";
    Arr t = tod_arr(tod, where, false), m = map_arr(local_map, where, false), x = xp_arr(xpointing, where, false);
    py::gil_scoped_release nogil;
    check(gmm_reference_map2tod(static_cast<float *>(t.ptr), t.size, static_cast<const float *>(m.ptr), m.size,
                                static_cast<const float *>(x.ptr), x.size / 3, lpix.h, partial_pixelization, accum));
}

void reference_tod2map(py::object local_map, py::object tod, py::object xpointing, const LocalPixelization &lpix,
                       bool partial_pixelization)
{
    const char *where = "reference_tod2map";
    Arr m = map_arr(local_map, where, false), t = tod_arr(tod, where, false), x = xp_arr(xpointing, where, false);
    py::gil_scoped_release nogil;
    check(gmm_reference_tod2map(static_cast<float *>(m.ptr), m.size, static_cast<const float *>(t.ptr), t.size,
                                static_cast<const float *>(x.ptr), x.size / 3, lpix.h, partial_pixelization));
}

std::string preplan_str(const PointingPrePlan &p)
{
    gmm_preplan_info i = p.info();
    std::ostringstream ss;
    ss << "PointingPrePlan(nsamp=" << i.nsamp << ", nypix_global=" << i.nypix_global
       << ", nxpix_global=" << i.nxpix_global << ", plan_nbytes=" << i.plan_nbytes
       << ", tmp_nbytes=" << i.plan_constructor_tmp_nbytes << ", overhead=" << i.overhead
       << ", ncl_per_threadblock=" << i.ncl_per_threadblock << ", planner_blocks=" << i.planner_nblocks
       << ", nmt_per_threadblock=" << i.nmt_per_threadblock << ", pointing_nblocks=" << i.pointing_nblocks << ")";
    return ss.str();
}

std::string plan_str(const PointingPlan &p)
{
    gmm_plan_info i = p.info();
    std::ostringstream ss;
    ss << "PointingPlan(nsamp=" << i.nsamp << ", nypix_global=" << i.nypix_global
       << ", nxpix_global=" << i.nxpix_global << ", plan_nmt=" << i.plan_nmt
       << ", nchunks=" << i.nchunks << ", nmlines=" << i.nmlines << ")";
    return ss.str();
}

}  // namespace


PYBIND11_MODULE(_gpu_mm_amd_py, m)
{
    m.doc() = "gpu_mm_amd: planned pointing operators for MI355X (pybind11 layer over libgpu_mm_amd.so)";
    m.def("_get_tsize", []() { return 4; });
    m.def("version", []() { return gmm_version(); });

    py::class_<LocalPixelization>(m, "LocalPixelization", "LocalPixelization: a set of map cells held on one GPU")
        .def(py::init<int64_t, int64_t, py::object, py::object, int64_t, int64_t, bool>(),
             py::arg("nypix_global"), py::arg("nxpix_global"), py::arg("cell_offsets_cpu"),
             py::arg("cell_offsets_gpu"), py::arg("ystride"), py::arg("polstride"), py::arg("periodic_xcoord"))
        .def_property_readonly("nypix_global", [](const LocalPixelization &p) { return p.info().nypix_global; })
        .def_property_readonly("nxpix_global", [](const LocalPixelization &p) { return p.info().nxpix_global; })
        .def_property_readonly("periodic_xcoord", [](const LocalPixelization &p) { return p.info().periodic_xcoord != 0; })
        .def_property_readonly("ystride", [](const LocalPixelization &p) { return p.info().ystride; })
        .def_property_readonly("polstride", [](const LocalPixelization &p) { return p.info().polstride; })
        .def_property_readonly("nycells", [](const LocalPixelization &p) { return p.info().nycells; })
        .def_property_readonly("nxcells", [](const LocalPixelization &p) { return p.info().nxcells; })
        .def_property("npix", [](const LocalPixelization &p) { return p.info().npix; },
                      [](LocalPixelization &p, int64_t n) { check(gmm_lpix_set_npix(p.h, n)); },
                      "counts only local pixels, does not include factor 3 from TQU.")
        .def("copy_gpu_offsets_to_cpu", [](LocalPixelization &p) { check(gmm_lpix_copy_gpu_offsets_to_cpu(p.h)); })
        .def("copy_cpu_offsets_to_gpu", [](LocalPixelization &p) { check(gmm_lpix_copy_cpu_offsets_to_gpu(p.h)); });

    py::class_<PointingPrePlan>(m, "PointingPrePlan")
        .def(py::init<py::object, int64_t, int64_t, py::object, py::object, bool, bool, uintptr_t>(),
             py::arg("xpointing_gpu"), py::arg("nypix_global"), py::arg("nxpix_global"), py::arg("nmt_gpu"),
             py::arg("err_gpu"), py::arg("periodic_xcoord"), py::arg("debug"), py::arg("stream") = 0)
        .def_static("_get_preplan_size", []() { return GMM_PREPLAN_SIZE; })
        .def_property_readonly("nsamp", [](const PointingPrePlan &p) { return p.info().nsamp; })
        .def_property_readonly("nypix_global", [](const PointingPrePlan &p) { return p.info().nypix_global; })
        .def_property_readonly("nxpix_global", [](const PointingPrePlan &p) { return p.info().nxpix_global; })
        .def_property_readonly("periodic_xcoord", [](const PointingPrePlan &p) { return p.info().periodic_xcoord != 0; })
        .def_property_readonly("plan_nbytes", [](const PointingPrePlan &p) { return p.info().plan_nbytes; })
        .def_property_readonly("plan_constructor_tmp_nbytes", [](const PointingPrePlan &p) { return p.info().plan_constructor_tmp_nbytes; })
        .def_property_readonly("overhead", [](const PointingPrePlan &p) { return p.info().overhead; })
        .def_property_readonly("ncl_per_threadblock", [](const PointingPrePlan &p) { return p.info().ncl_per_threadblock; })
        .def_property_readonly("planner_nblocks", [](const PointingPrePlan &p) { return p.info().planner_nblocks; })
        .def_property_readonly("nmt_per_threadblock", [](const PointingPrePlan &p) { return p.info().nmt_per_threadblock; })
        .def_property_readonly("pointing_nblocks", [](const PointingPrePlan &p) { return p.info().pointing_nblocks; })
        .def_property_readonly("plan_nmt", [](const PointingPrePlan &p) { return p.info().plan_nmt; })
        .def_property_readonly("cub_nbytes", [](const PointingPrePlan &p) { return p.info().cub_nbytes; })
        .def_property_readonly("max_nchunks", [](const PointingPrePlan &p) { return p.info().max_nchunks; })
        .def("get_nmt_cumsum", [](const PointingPrePlan &p) {
            int64_t n = p.info().planner_nblocks;
            py::array_t<uint32_t> out(n);
            check(gmm_preplan_get_nmt_cumsum(p.h, out.mutable_data(), n));
            return out;
        }, "Copies nmt_cumsum array to host, and returns it as a numpy array.")
        .def("__str__", &preplan_str);

    py::class_<PointingPlan>(m, "PointingPlan")
        .def(py::init<const PointingPrePlan &, py::object, py::object, py::object, bool, uintptr_t>(),
             py::arg("preplan"), py::arg("xpointing_gpu"), py::arg("buf"), py::arg("tmp_buf"),
             py::arg("debug") = false, py::arg("stream") = 0)
        .def_property_readonly("nsamp", [](const PointingPlan &p) { return p.info().nsamp; })
        .def_property_readonly("nypix_global", [](const PointingPlan &p) { return p.info().nypix_global; })
        .def_property_readonly("nxpix_global", [](const PointingPlan &p) { return p.info().nxpix_global; })
        .def_property_readonly("periodic_xcoord", [](const PointingPlan &p) { return p.info().periodic_xcoord != 0; })
        .def_property_readonly("plan_nmt", [](const PointingPlan &p) { return p.info().plan_nmt; })
        .def_property_readonly("nchunks", [](const PointingPlan &p) { return p.info().nchunks; })
        .def_property_readonly("nmlines", [](const PointingPlan &p) { return p.info().nmlines; })
        .def("_check_errflags", [](const PointingPlan &p, const std::string &where) {
            check(gmm_plan_check_errflags(p.h, where.c_str()));
        }, py::arg("where"))
        .def("_check_invariants", [](const PointingPlan &p, uintptr_t stream) {
            py::gil_scoped_release nogil;
            check(gmm_plan_check_invariants(p.h, as_stream(stream)));
        }, py::arg("stream") = 0)
        .def("get_plan_mt", [](const PointingPlan &p) {
            int64_t n = p.info().plan_nmt;
            py::array_t<uint64_t> out(n);
            check(gmm_plan_get_plan_mt(p.h, out.mutable_data(), n, 0, nullptr));
            return out;
        }, "Length-nmt array of plan entries, sorted by map cell")
        .def("__str__", &plan_str);

    m.def("planned_map2tod", &planned_map2tod, py::arg("tod"), py::arg("local_map"), py::arg("xpointing"),
          py::arg("local_pixelization"), py::arg("plan"), py::arg("partial_pixelization"), py::arg("debug"),
          py::arg("accum"), py::arg("check_errors") = true, py::arg("stream") = 0);
    m.def("planned_tod2map", &planned_tod2map, py::arg("local_map"), py::arg("tod"), py::arg("xpointing"),
          py::arg("local_pixelization"), py::arg("plan"), py::arg("partial_pixelization"), py::arg("debug"),
          py::arg("check_errors") = true, py::arg("stream") = 0);
    m.def("response_map2tod", &response_map2tod, py::arg("tod"), py::arg("local_map"), py::arg("xpointing"),
          py::arg("response"), py::arg("local_pixelization"), py::arg("plan"), py::arg("partial_pixelization"),
          py::arg("debug"), py::arg("accum"), py::arg("check_errors") = true, py::arg("stream") = 0);
    m.def("response_tod2map", &response_tod2map, py::arg("local_map"), py::arg("tod"), py::arg("xpointing"),
          py::arg("response"), py::arg("local_pixelization"), py::arg("plan"), py::arg("partial_pixelization"),
          py::arg("debug"), py::arg("check_errors") = true, py::arg("stream") = 0);
    m.def("planned_map2tod_q16", &planned_map2tod_q16, py::arg("tod"), py::arg("tod_scale"), py::arg("local_map"),
          py::arg("xpointing"), py::arg("local_pixelization"), py::arg("plan"), py::arg("partial_pixelization"),
          py::arg("debug"), py::arg("accum"), py::arg("response") = py::none(), py::arg("check_errors") = true,
          py::arg("stream") = 0);
    m.def("planned_tod2map_q16", &planned_tod2map_q16, py::arg("local_map"), py::arg("tod"), py::arg("tod_scale"),
          py::arg("xpointing"), py::arg("local_pixelization"), py::arg("plan"), py::arg("partial_pixelization"),
          py::arg("debug"), py::arg("response") = py::none(), py::arg("check_errors") = true, py::arg("stream") = 0);
    m.def("expand_dynamic_map", [](py::object global_ncells, py::object cell_offsets, py::object plan_mt, uintptr_t stream) {
        const char *where = "expand_dynamic_map";
        Arr g = get_arr(global_ncells, where, "global_ncells");
        require(g, where, "global_ncells", kDLUInt, 32, true, "uint32");
        Arr c = get_arr(cell_offsets, where, "cell_offsets");
        require(c, where, "cell_offsets", kDLInt, 64, true, "int64");
        Arr m = get_arr(plan_mt, where, "plan_mt");
        require(m, where, "plan_mt", kDLUInt, 64, true, "uint64");
        if (g.size != 1 || c.shape.size() != 2 || m.shape.size() != 1)
            raise(std::string(where) + ": expected global_ncells of shape (1,), cell_offsets 2-d, plan_mt 1-d");
        uint32_t n = 0;
        {
            py::gil_scoped_release nogil;
            check(gmm_expand_dynamic_map(static_cast<uint32_t *>(g.ptr), static_cast<int64_t *>(c.ptr), c.shape[0],
                                         c.shape[1], static_cast<const uint64_t *>(m.ptr), m.size, &n, as_stream(stream)));
        }
        return n;
    }, py::arg("global_ncells"), py::arg("cell_offsets"), py::arg("plan_mt"), py::arg("stream") = 0,
       "Adds the targeted (-1) cells named by plan_mt to cell_offsets; returns the new cell count");
    m.def("expand_dynamic_map2", [](py::object global_ncells, LocalPixelization &lpix, const PointingPlan &plan,
                                    uintptr_t stream) {
        const char *where = "expand_dynamic_map2";
        Arr g = get_arr(global_ncells, where, "global_ncells");
        require(g, where, "global_ncells", kDLUInt, 32, true, "uint32");
        uint32_t n = 0;
        {
            py::gil_scoped_release nogil;
            check(gmm_lpix_expand_dynamic(lpix.h, plan.h, static_cast<uint32_t *>(g.ptr), &n, as_stream(stream)));
        }
        return n;
    }, py::arg("global_ncells"), py::arg("local_pixelization"), py::arg("plan"), py::arg("stream") = 0);
    m.def("cell_broadcast", [](py::object dst, py::object src, py::object index_map, uintptr_t stream) {
        const char *where = "cell_broadcast";
        Arr d = get_arr(dst, where, "dst"), sr = get_arr(src, where, "src"), im = get_arr(index_map, where, "index_map");
        require(d, where, "dst", kDLFloat, 32, true, "float32");
        require(sr, where, "src", kDLFloat, 32, true, "float32");
        require(im, where, "index_map", kDLInt, 64, true, "int64");
        if (im.shape.size() != 1 || d.size != im.size * 12288 || sr.size % 12288)
            raise(std::string(where) + ": expected dst of 3*64*64*len(index_map) floats and src of whole cells");
        py::gil_scoped_release nogil;
        check(gmm_cell_broadcast(static_cast<float *>(d.ptr), im.size, static_cast<const float *>(sr.ptr),
                                 sr.size / 12288, static_cast<const int64_t *>(im.ptr), as_stream(stream)));
    }, py::arg("dst"), py::arg("src"), py::arg("index_map"), py::arg("stream") = 0);
    m.def("cell_reduce", [](py::object dst, py::object src, py::object index_map, uintptr_t stream) {
        const char *where = "cell_reduce";
        Arr d = get_arr(dst, where, "dst"), sr = get_arr(src, where, "src"), im = get_arr(index_map, where, "index_map");
        require(d, where, "dst", kDLFloat, 32, true, "float32");
        require(sr, where, "src", kDLFloat, 32, true, "float32");
        require(im, where, "index_map", kDLInt, 64, true, "int64");
        if (im.shape.size() != 1 || sr.size != im.size * 12288 || d.size % 12288)
            raise(std::string(where) + ": expected src of 3*64*64*len(index_map) floats and dst of whole cells");
        py::gil_scoped_release nogil;
        check(gmm_cell_reduce(static_cast<float *>(d.ptr), d.size / 12288, static_cast<const float *>(sr.ptr), im.size,
                              static_cast<const int64_t *>(im.ptr), as_stream(stream)));
    }, py::arg("dst"), py::arg("src"), py::arg("index_map"), py::arg("stream") = 0);
    m.def("local_map_to_global", [](const LocalPixelization &lpix, py::object dst, py::object src, uintptr_t stream) {
        const char *where = "local_map_to_global";
        Arr d = get_arr(dst, where, "dst"), sr = get_arr(src, where, "src");
        require(d, where, "dst", kDLFloat, 32, true, "float32");
        require(sr, where, "src", kDLFloat, 32, true, "float32");
        const gmm_lpix_info i = lpix.info();
        if (d.shape.size() != 3 || d.shape[0] != 3 || d.shape[1] != i.nypix_global || d.shape[2] != i.nxpix_global)
            raise(std::string(where) + ": expected dst of shape (3, nypix_global, nxpix_global)");
        py::gil_scoped_release nogil;
        check(gmm_local_map_to_global(lpix.h, static_cast<float *>(d.ptr), static_cast<const float *>(sr.ptr), sr.size,
                                      as_stream(stream)));
    }, py::arg("local_pixelization"), py::arg("dst"), py::arg("src"), py::arg("stream") = 0);
    m.def("kernel_timer_enable", [](bool on) { check(gmm_kernel_timer_enable(on ? 1 : 0)); });
    m.def("kernel_timer_read", [](int which) {
        double ms = 0.0;
        int64_t n = 0;
        check(gmm_kernel_timer_read(which, &ms, &n));
        return py::make_tuple(ms, n);
    }, "Returns (total_ms, launches) of GMM_TIMED_MAP2TOD (0) or GMM_TIMED_TOD2MAP (1) since the last enable/read");
    m.def("stream_read", [](py::object src, py::object sum, uintptr_t stream) {
        const char *where = "stream_read";
        Arr a = get_arr(src, where, "src"), b = get_arr(sum, where, "sum");
        require(a, where, "src", kDLFloat, 32, true, "float32");
        require(b, where, "sum", kDLFloat, 32, true, "float32");
        if (b.size < 1)
            raise(std::string(where) + ": 'sum' must hold at least one float");
        py::gil_scoped_release nogil;
        check(gmm_stream_read(static_cast<const float *>(a.ptr), a.size, static_cast<float *>(b.ptr), as_stream(stream)));
    }, py::arg("src"), py::arg("sum"), py::arg("stream") = 0,
       "Reads src once with non-temporal 16-byte loads and adds its sum to sum[0] (bandwidth probe)");
    py::class_<gmm_batch_config>(m, "BatchConfig")
        .def(py::init([]() { gmm_batch_config c; gmm_batch_default_config(&c); return c; }))
        .def_readwrite("m2t_item_entries", &gmm_batch_config::m2t_item_entries)
        .def_readwrite("m2t_threads", &gmm_batch_config::m2t_threads)
        .def_readwrite("m2t_blocks_per_cu", &gmm_batch_config::m2t_blocks_per_cu)
        .def_readwrite("t2m_item_entries", &gmm_batch_config::t2m_item_entries)
        .def_readwrite("t2m_blocks_per_cu", &gmm_batch_config::t2m_blocks_per_cu)
        .def("__repr__", [](const gmm_batch_config &c) {
            std::ostringstream ss;
            ss << "BatchConfig(m2t_item_entries=" << c.m2t_item_entries << ", m2t_threads=" << c.m2t_threads
               << ", m2t_blocks_per_cu=" << c.m2t_blocks_per_cu << ", t2m_item_entries=" << c.t2m_item_entries
               << ", t2m_blocks_per_cu=" << c.t2m_blocks_per_cu << ")";
            return ss.str();
        });

    py::class_<Batch>(m, "Batch")
        .def(py::init([](py::object lpix, py::list plans, py::list xpointings, py::list tods, py::object scales,
                         py::object config) {
                 return new Batch(lpix.cast<const LocalPixelization &>(), lpix, plans, xpointings, tods, scales,
                                  config);
             }),
             py::arg("local_pixelization"), py::arg("plans"), py::arg("xpointings"), py::arg("tods"),
             py::arg("scales") = py::none(), py::arg("config") = py::none())
        .def_property_readonly("ntod", [](const Batch &b) { return b.info().ntod; })
        .def_property_readonly("config", [](const Batch &b) { return b.info().config; })
        .def_property_readonly("nentries", [](const Batch &b) { return b.info().nentries; })
        .def_property_readonly("ncells", [](const Batch &b) { return b.info().ncells; })
        .def_property_readonly("generation", [](const Batch &b) { return b.info().generation; })
        .def("work_items", [](const Batch &b) {
            const gmm_batch_info i = b.info();
            py::dict d;
            d["m2t_nitems"] = i.m2t_nitems;
            d["m2t_nsegs"] = i.m2t_nsegs;
            d["m2t_grid"] = i.m2t_grid;
            d["t2m_nitems"] = i.t2m_nitems;
            d["t2m_nsegs"] = i.t2m_nsegs;
            d["t2m_grid"] = i.t2m_grid;
            return d;
        })
        .def("configure", [](Batch &b, const gmm_batch_config &c) {
            py::gil_scoped_release nogil;
            check(gmm_batch_configure(b.h, &c));
        })
        .def("set_response", &Batch::set_response, py::arg("responses"))
        .def("map2tod", [](Batch &b, py::object local_map, bool partial, bool accum, bool check_errors,
                           uintptr_t stream) {
            Arr m = b.map(local_map, "PlanBatch.map2tod");
            py::gil_scoped_release nogil;
            check(gmm_batch_map2tod(b.h, static_cast<const float *>(m.ptr), m.size, partial, accum, check_errors,
                                    as_stream(stream)));
        }, py::arg("local_map"), py::arg("partial_pixelization") = false, py::arg("accum") = false,
           py::arg("check_errors") = true, py::arg("stream") = 0)
        .def("tod2map", [](Batch &b, py::object local_map, bool partial, bool check_errors, uintptr_t stream) {
            Arr m = b.map(local_map, "PlanBatch.tod2map");
            py::gil_scoped_release nogil;
            check(gmm_batch_tod2map(b.h, static_cast<float *>(m.ptr), m.size, partial, check_errors,
                                    as_stream(stream)));
        }, py::arg("local_map"), py::arg("partial_pixelization") = false, py::arg("check_errors") = true,
           py::arg("stream") = 0)
        .def("check_errflags", [](Batch &b, const std::string &where) {
            check(gmm_batch_check_errflags(b.h, where.c_str()));
        }, py::arg("where") = "PlanBatch")
        .def("autotune", [](Batch &b, py::object map_in, py::object map_scratch, int reps, uintptr_t stream) {
            Arr mi = b.map(map_in, "PlanBatch.autotune"), ms = b.map(map_scratch, "PlanBatch.autotune");
            std::vector<gmm_batch_tune_entry> tab(256);
            int32_t n = 0;
            {
                py::gil_scoped_release nogil;
                check(gmm_batch_autotune(b.h, static_cast<const float *>(mi.ptr), static_cast<float *>(ms.ptr),
                                         mi.size, reps, tab.data(), int32_t(tab.size()), &n, as_stream(stream)));
            }
            py::list out;
            for (int32_t k = 0; k < std::min<int32_t>(n, int32_t(tab.size())); k++)
                out.append(py::make_tuple(tab[size_t(k)].op == GMM_OP_MAP2TOD ? "map2tod" : "tod2map",
                                          tab[size_t(k)].item_entries, tab[size_t(k)].threads,
                                          tab[size_t(k)].blocks_per_cu, tab[size_t(k)].ms));
            return out;
        }, py::arg("map_in"), py::arg("map_scratch"), py::arg("reps") = 3, py::arg("stream") = 0)
        .def("capture", [](py::object self, const std::string &op, py::object local_map, bool partial, bool accum) {
            Batch &b = self.cast<Batch &>();
            int code;
            if (op == "map2tod")
                code = GMM_OP_MAP2TOD;
            else if (op == "tod2map")
                code = GMM_OP_TOD2MAP;
            else
                raise("PlanBatch.capture: op must be 'map2tod' or 'tod2map'");
            Arr m = b.map(local_map, "PlanBatch.capture");
            auto g = std::make_unique<Graph>();
            {
                py::gil_scoped_release nogil;
                check(gmm_batch_graph_create(&g->h, b.h, code, static_cast<float *>(m.ptr), m.size, partial, accum));
            }
            g->batch_ref = self;
            g->map_ref = local_map;
            return g;
        }, py::arg("op"), py::arg("local_map"), py::arg("partial_pixelization") = false, py::arg("accum") = false);

    py::class_<Graph>(m, "BatchGraph")
        .def("launch", [](Graph &g, uintptr_t stream) {
            py::gil_scoped_release nogil;
            check(gmm_graph_launch(g.h, as_stream(stream)));
        }, py::arg("stream") = 0);

    m.def("tod16_encode", &tod16_encode, py::arg("out"), py::arg("scale"), py::arg("tod"), py::arg("error_feedback") = 0,
          py::arg("stream") = 0);
    m.def("unplanned_map2tod", &unplanned_map2tod, py::arg("tod"), py::arg("local_map"), py::arg("xpointing"),
          py::arg("local_pixelization"), py::arg("errflags"), py::arg("partial_pixelization"), py::arg("accum"),
          py::arg("check_errors") = true, py::arg("stream") = 0);
    m.def("unplanned_tod2map", &unplanned_tod2map, py::arg("local_map"), py::arg("tod"), py::arg("xpointing"),
          py::arg("local_pixelization"), py::arg("errflags"), py::arg("partial_pixelization"),
          py::arg("check_errors") = true, py::arg("stream") = 0);
    m.def("reference_map2tod", &reference_map2tod, py::arg("tod"), py::arg("local_map"), py::arg("xpointing"),
          py::arg("local_pixelization"), py::arg("partial_pixelization"), py::arg("accum"));
    m.def("reference_tod2map", &reference_tod2map, py::arg("local_map"), py::arg("tod"), py::arg("xpointing"),
          py::arg("local_pixelization"), py::arg("partial_pixelization"));

    // ---- test aids ----

    m.def("toy_pointing_fill", [](py::object xpointing_cpu, int64_t nypix_global, int64_t nxpix_global,
                                  double scan_speed, double total_drift) {
        const char *where = "ToyPointing constructor";
        Arr x = xp_arr(xpointing_cpu, where, false);
        double drift = 0;
        check(gmm_toy_pointing_fill(static_cast<float *>(x.ptr), x.size / 3, nypix_global, nxpix_global,
                                    scan_speed, total_drift, &drift));
        return drift;
    }, py::arg("xpointing_cpu"), py::arg("nypix_global"), py::arg("nxpix_global"), py::arg("scan_speed"),
       py::arg("total_drift"));

    m.def("plan_tester_tmp_nbytes", [](const PointingPrePlan &p) { return gmm_plan_tester_tmp_nbytes(p.h); });

    m.def("plan_tester_run", [](const PointingPrePlan &p, py::object xpointing_gpu, py::object tmp, uintptr_t stream) {
        const char *where = "PointingPlanTester constructor";
        Arr x = xp_arr(xpointing_gpu, where, true);
        Arr t = get_arr(tmp, where, "tmp");
        require(t, where, "tmp", kDLUInt, 8, true, "uint8");
        gmm_preplan_info i = p.info();
        int64_t nsamp = x.size / 3;
        py::array_t<int32_t> iy(nsamp), ix(nsamp);
        py::array_t<uint32_t> cum(i.planner_nblocks);
        py::array_t<uint64_t> smt(i.plan_nmt);
        int32_t *iyp = iy.mutable_data(), *ixp = ix.mutable_data();
        uint32_t *cp = cum.mutable_data();
        uint64_t *sp = smt.mutable_data();
        {
            py::gil_scoped_release nogil;
            check(gmm_plan_tester_run(p.h, static_cast<const float *>(x.ptr), nsamp, t.ptr, t.size,
                                      iyp, ixp, cp, sp, as_stream(stream)));
        }
        return py::make_tuple(iy, ix, cum, smt);
    }, py::arg("preplan"), py::arg("xpointing_gpu"), py::arg("tmp"), py::arg("stream") = 0);

    m.def("test_plan_iterator", [](py::object plan_mt, int64_t nmt_per_block, int warps_per_threadblock) {
        const char *where = "test_plan_iterator";
        Arr a = get_arr(plan_mt, where, "plan_mt");
        require(a, where, "plan_mt", kDLUInt, 64, false, "uint64");
        py::gil_scoped_release nogil;
        check(gmm_test_plan_iterator(static_cast<const uint64_t *>(a.ptr), a.size, nmt_per_block, warps_per_threadblock));
    }, py::arg("plan_mt"), py::arg("nmt_per_block"), py::arg("warps_per_threadblock"));

    m.def("make_random_plan_mt", [](int64_t ncells, int64_t min_nmt_per_cell, int64_t max_nmt_per_cell, uint64_t seed) {
        int64_t n = 0;
        check(gmm_make_random_plan_mt(ncells, min_nmt_per_cell, max_nmt_per_cell, seed, nullptr, 0, &n));
        py::array_t<uint64_t> out(n);
        check(gmm_make_random_plan_mt(ncells, min_nmt_per_cell, max_nmt_per_cell, seed, out.mutable_data(), n, &n));
        return out;
    }, py::arg("ncells"), py::arg("min_nmt_per_cell"), py::arg("max_nmt_per_cell"), py::arg("seed") = 0);
}
