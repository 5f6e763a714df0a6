// Geometry, error bits and plan-entry format shared by the host code and the gfx950 kernels.
//
// Everything here is __host__ __device__ so that the CPU "reference" operators, the planner
// test aid and the HIP kernels locate a sample with the same IEEE float operations, and
// therefore agree bit for bit on which pixels and cells a sample touches.

#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace gmm {

// ------------------------------------------------------------------------------------------
// Error bits, reported per block by the kernels and turned into exceptions by the host.
// Values follow include/gpu_mm_internals.hpp:26-29 (same meaning, same messages).

constexpr uint32_t ERR_BAD_YPIX = 0x1;
constexpr uint32_t ERR_BAD_XPIX = 0x2;
constexpr uint32_t ERR_INCONSISTENT_NMT = 0x4;
constexpr uint32_t ERR_NOT_IN_PIXELIZATION = 0x8;

// ------------------------------------------------------------------------------------------
// Plan entry ("mt") bit layout, src_lib/PointingPlan.cu:75-89 and include/gpu_mm.hpp:120-129:
//   bits [0,10)  ixcell        (global x cell index)
//   bits [10,20) iycell        (global y cell index)
//   bits [20,46) icl           (TOD cache line = sample index >> 5)
//   bit  46      mflag         (the cache line touches more than one cell)
//   bit  47      zflag         (exactly one entry per multi-cell line carries it: the entry
//                               whose cell holds the (iy0,ix0) corner of the line's first sample)

constexpr int CELL_BITS = 20;
constexpr uint32_t CELL_MASK = (1u << 20) - 1;
constexpr int ICL_SHIFT = 20;
constexpr uint32_t ICL_MASK = (1u << 26) - 1;
constexpr uint64_t MFLAG = 1ull << 46;
constexpr uint64_t ZFLAG = 1ull << 47;

// A cell is 64x64 pixels; one cell of a T/Q/U map holds 3*64*64 floats (48 KiB).
constexpr int CELL_SIDE = 64;
constexpr int CELL_NPIX = 64 * 64;

// Samples per plan entry (one TOD cache line of float32).
constexpr int LINE = 32;

// Operator sample ownership (per plan entry, in the owned-quad mask qmask[e]). A sample belongs
// to the cell of its (iy0, ix0) corner, except that the block of the cell to its left (same cell
// row) evaluates it when the line's samples of that cell all lie in the first HALO_X - 1 columns
// (ix0 % 64 <= HALO_X - 2, and ix0 + 1 < nx: no x-wraparound) and the line touches the left
// cell: that block's halo tile spans HALO_X columns of its right neighbour, so a line crossing
// the boundary is read once instead of by both blocks. Bits 0..7 of qmask[e]: the 4-sample
// groups of the line holding a sample that entry e evaluates; QM_TAKES_RIGHT: entry e evaluates
// the line's samples of the cell to its right.
constexpr int HALO_X = 12;
constexpr uint16_t QM_TAKES_RIGHT = 0x100;

// Preplan block count (include/gpu_mm.hpp:63).
constexpr int PREPLAN_SIZE = 1024;

__host__ __device__ inline uint32_t cell_key(int iycell, int ixcell)
{
    return (uint32_t(iycell) << 10) | uint32_t(ixcell);
}

__host__ __device__ inline int64_t align128(int64_t n) { return (n + 127) & ~int64_t(127); }

// ------------------------------------------------------------------------------------------
// Pixel location. Semantics follow pixel_locator<T>::locate, include/gpu_mm_internals.hpp:124-164:
//
//   - y: iy0 = clamp(trunc(y), 0, ny-2), iy1 = iy0+1, dy = y - iy0; valid iff 0 <= y <= ny-1.
//   - non-periodic x: the same with nx; valid iff 0 <= x <= nx-2.
//   - periodic x: valid iff -nx <= x <= 2nx. One conditional shift by +nx (x < 0) and one by
//     -nx (x > nx) brings x into [-eps, nx+eps]; ix0 = max(trunc(x),0), dx = x - ix0, then
//     ix0 == nx is folded to 0 and ix1 = ix0+1 wraps to 0 at nx.
//
// The arithmetic is a handful of compares, conversions and exact subtractions, plus the two
// float additions x +- nx, so host and device results agree exactly (no fused operations).
// Out-of-range inputs still produce in-range indices; the returned error bits flag them.

template<typename T>
struct Pixel
{
    int iy0, iy1, ix0, ix1;
    T dy, dx;
};

template<typename T>
__host__ __device__ inline int trunc_to_int(T v)
{
#if defined(__HIP_DEVICE_COMPILE__)
    // v_cvt_i32_f32 truncates, saturates out-of-range values and maps NaN to 0.
    if constexpr (sizeof(T) == 4)
        return __float2int_rz(v);
#endif
    // Host (and double): the same saturation made explicit. Out-of-range or NaN samples are
    // always flagged by the range checks, and clamp to the same pixel on host and device.
    if (!(v > T(-2147483520.0)))
        return (v == v) ? -2147483647 : 0;
    if (!(v < T(2147483520.0)))
        return 2147483647;
    return int(v);
}

template<typename T>
__host__ __device__ inline uint32_t locate_pixel(T y, T x, int ny, int nx, bool periodic, Pixel<T> &p)
{
    uint32_t err = 0;

    int iy = trunc_to_int(y);
    iy = (iy > 0) ? iy : 0;
    iy = (iy < ny - 2) ? iy : (ny - 2);
    p.iy0 = iy;
    p.iy1 = iy + 1;
    p.dy = y - T(iy);
    if (!((y >= T(0)) && (y <= T(ny - 1))))
        err |= ERR_BAD_YPIX;

    if (periodic) {
        if (!((x >= T(-nx)) && (x <= T(2 * nx))))
            err |= ERR_BAD_XPIX;
        x = (x >= T(0)) ? x : (x + T(nx));
        x = (x <= T(nx)) ? x : (x - T(nx));
        int ix = trunc_to_int(x);
        ix = (ix > 0) ? ix : 0;
        p.dx = x - T(ix);
        ix = (ix < nx) ? ix : 0;
        p.ix0 = ix;
        p.ix1 = (ix + 1 < nx) ? (ix + 1) : 0;
    }
    else {
        if (!((x >= T(0)) && (x <= T(nx - 2))))
            err |= ERR_BAD_XPIX;
        int ix = trunc_to_int(x);
        ix = (ix > 0) ? ix : 0;
        ix = (ix < nx - 2) ? ix : (nx - 2);
        p.ix0 = ix;
        p.ix1 = ix + 1;
        p.dx = x - T(ix);
    }

    return err;
}

// ------------------------------------------------------------------------------------------
// sin(x), cos(x) on the device for pointing angles x = 2*alpha.
//
// Cody-Waite reduction x = j*pi/2 + r, then polynomials on |r| <= pi/4 fitted for this
// library by relative least squares on Chebyshev nodes (max error 0.8 ulp for sin, 1.3 ulp for
// cos over the reduced range); the quadrant j mod 4 swaps and negates.
// For |x| <= 1e6, pi/2 is split into three floats and each step is one FMA (exact product, one
// rounding), so the reduction error stays at the rounding of r while j is an exact float.
// Larger |x| (not a realistic angle) reduce in double with pi/2 split into three doubles, behind
// a real branch; beyond |x| ~ 2^52 no float argument carries phase information anyway.

// Polynomials on the reduced argument r (|r| <= pi/4) and the quadrant q.
__device__ inline void sincos_reduced(float r, int q, float &s, float &c)
{
    const float z = r * r;
    const float sp = fmaf(r * z, fmaf(z, fmaf(z, -1.95139321e-04f, 8.33214819e-03f), -1.66666538e-01f), r);
    const float cp = fmaf(z * z, fmaf(z, fmaf(z, 2.44305738e-05f, -1.38872897e-03f), 4.16666456e-02f),
                          fmaf(-0.5f, z, 1.0f));
    const float ss = (q & 1) ? cp : sp;
    const float cc = (q & 1) ? sp : cp;
    s = (q & 2) ? -ss : ss;
    c = ((q + 1) & 2) ? -cc : cc;
}

// |x| <= 1e6 only (the float Cody-Waite reduction).
__device__ inline void sincos_angle_small(float x, float &s, float &c)
{
    const float j = rintf(x * 0.636619747f);                 // 2/pi
    float r = fmaf(j, -1.57079637f, x);                       // pi/2 = 1.57079637
    r = fmaf(j, 4.37113883e-08f, r);                          //      - 4.37113883e-8
    r = fmaf(j, 1.71512451e-15f, r);                          //      - 1.71512451e-15
    sincos_reduced(r, int(j), s, c);
}

__device__ inline void sincos_angle(float x, float &s, float &c)
{
    if (__builtin_expect(fabsf(x) <= 1.0e6f, 1)) {
        sincos_angle_small(x, s, c);
        return;
    }
    const double xd = double(x);
    const double jd = rint(xd * 0.6366197723675814);
    double rd = fma(jd, -1.5707963267948966, xd);
    rd = fma(jd, -6.123233995736766e-17, rd);
    rd = fma(jd, 1.4973849048591698e-33, rd);
    sincos_reduced(float(rd), int(jd - 4.0 * floor(jd * 0.25)), s, c);
}

// The (up to four) distinct cells touched by the bilinear footprint of one sample, as
// cell keys; duplicates within the sample are replaced by NO_CELL.
constexpr uint32_t NO_CELL = 0xffffffffu;

template<typename T>
__host__ __device__ inline void footprint_cells(const Pixel<T> &p, uint32_t k[4])
{
    int cy0 = p.iy0 >> 6, cy1 = p.iy1 >> 6;
    int cx0 = p.ix0 >> 6, cx1 = p.ix1 >> 6;
    bool ysplit = (cy1 != cy0);
    bool xsplit = (cx1 != cx0);
    k[0] = cell_key(cy0, cx0);
    k[1] = xsplit ? cell_key(cy0, cx1) : NO_CELL;
    k[2] = ysplit ? cell_key(cy1, cx0) : NO_CELL;
    k[3] = (xsplit && ysplit) ? cell_key(cy1, cx1) : NO_CELL;
}

// ------------------------------------------------------------------------------------------
// Launch-parameter formulas, src_lib/PointingPrePlan.cu:162-168 and 236-250. They define the
// planner block partition (and therefore the nmt_cumsum array that the tests compare), so
// they are kept identical.

struct PreplanParams
{
    int64_t ncl;                  // number of 32-sample lines
    int64_t ncl_per_threadblock;  // lines per planner block
    int64_t planner_nblocks;
};

__host__ inline PreplanParams preplan_params(int64_t nsamp)
{
    PreplanParams r;
    r.ncl = (nsamp + 31) / 32;
    int64_t a = (r.ncl + PREPLAN_SIZE - 1) / PREPLAN_SIZE;
    r.ncl_per_threadblock = (a > 16) ? a : 16;
    r.ncl_per_threadblock = (r.ncl_per_threadblock + 3) & ~int64_t(3);
    r.planner_nblocks = (r.ncl + r.ncl_per_threadblock - 1) / r.ncl_per_threadblock;
    return r;
}

// ------------------------------------------------------------------------------------------
// Operator work unit: a contiguous run of plan entries that all belong to one cell.
// The sorted plan is cut into these at plan time; each map2tod/tod2map block takes one.

struct Chunk
{
    uint32_t icell;   // cell key (iycell << 10 | ixcell)
    uint32_t begin;   // first plan entry
    uint32_t end;     // one past the last plan entry
    uint32_t flags;   // CHUNK_SHARED: the cell's entries are split over more than one chunk
};

constexpr uint32_t CHUNK_SHARED = 1u;

// Entries [wb, we) of a chunk handled by wave 'wave' of 'nw' in the planned operators: the chunk
// split into nw contiguous, near-equal slices. Each wave then walks its slice in batches of 64
// entries (one plan entry per lane, one vector load) and steps of 8 entries (lane 8*li + q takes
// the 4-sample group q of entry li of the step).
__host__ __device__ inline void wave_slice(const Chunk &ch, int wave, int nw, uint32_t &wb, uint32_t &we)
{
    const uint32_t n = ch.end - ch.begin;
    wb = ch.begin + uint32_t((uint64_t(n) * uint32_t(wave)) / uint32_t(nw));
    we = ch.begin + uint32_t((uint64_t(n) * uint32_t(wave + 1)) / uint32_t(nw));
}

// Row of a sample index for rows of 'nper' samples, by multiplication: row(n) = (n * mult) >>
// shift with mult = ceil(2^(31+L) / nper), shift = 31 + L, L = ceil(log2 nper), exact for every
// n < 2^31. (With e = mult*nper - 2^(31+L) in [0, nper): n*mult / 2^(31+L) = n/nper + n*e /
// (nper*2^(31+L)), and the second term is below 1/nper, too small to reach the next integer.)
// mult < 2^32 because nper > 2^(L-1). Replaces an integer division (about twenty instructions on
// the GPU) by one 32x32->64 multiply and a shift.
struct RowDiv
{
    uint32_t nper = 1, mult = 1u << 31, shift = 31;
};

inline RowDiv make_rowdiv(uint32_t nper)
{
    RowDiv d;
    d.nper = nper ? nper : 1u;
    int L = 0;
    while ((uint64_t(1) << L) < d.nper)
        L++;
    d.shift = uint32_t(31 + L);
    d.mult = uint32_t(((uint64_t(1) << d.shift) + d.nper - 1) / d.nper);
    return d;
}

__host__ __device__ inline uint32_t rowdiv(uint32_t n, const RowDiv &d)
{
    return uint32_t((uint64_t(n) * d.mult) >> d.shift);
}

// Batched operators (several TODs with their own plans, one local map). Device-resident tables.
struct BatchTod
{
    const float *xp;              // xpointing [3][nsamp]
    void *tod;                    // TOD [nsamp] (float, or 16-bit with per-row scales)
    const float *scale;           // 16-bit TODs: float[nsamp / scale_div.nper], else null
    const uint64_t *plan_mt;      // the TOD's sorted plan entries
    const uint16_t *qmask;        // the plan's owned-quad masks
    const float *resp;            // per-detector gains float[2][resp_ndet], or null
    int64_t nsamp;
    RowDiv scale_div;             // samples per scale row
    RowDiv resp_div;              // samples per detector of resp
    uint32_t resp_ndet;           // detectors of resp
    uint32_t pad;
};

struct BatchSeg
{
    uint32_t tod;                 // index into the BatchTod table
    uint32_t begin, end;          // plan entries [begin, end) of that TOD, all of one cell
    uint32_t pad;
};

struct BatchItem
{
    uint32_t icell;               // cell key (iycell << 10 | ixcell)
    uint32_t seg_begin, seg_end;  // segments [seg_begin, seg_end)
    uint32_t nent;                // total plan entries of the segments
    uint32_t flags;               // CHUNK_SHARED: the cell's entries are split over several items
    uint32_t pad[3];
};

}  // namespace gmm
