// Planned map2tod / tod2map on gfx950.
//
// One block per work unit ("chunk": a run of plan entries of one cell, cut at plan time). Each
// wave takes a contiguous slice of the chunk's entries and caches 64 plan entries per vector
// load. A wave step covers 8 entries (8 TOD lines of 32 samples): lane 8*li + q handles samples
// 4q..4q+3 of line li, loaded as float4.
//
// Ownership. A sample belongs to the cell of its (iy0, ix0) corner; that cell is in the plan
// for the sample's line, since the plan lists every corner's cell. The block of a cell works on
// a 65 x (64 + HALO_X) "halo" tile per plane: the cell plus the first row of the cells below and
// the first HALO_X columns of the cell to the right (for periodic x in the last cell column, the
// column that wraparound reads), so it evaluates or deposits all four corners of each sample it
// evaluates at tile offsets o, o+1, o+HS, o+HS+1, and skips the others. It evaluates the samples
// it owns, and also those of the right-hand cell on lines where the plan's owned-quad mask says
// so (QM_TAKES_RIGHT: the line's part in that cell lies within the halo columns), while that
// cell's entry for the line then loads nothing: a line that crosses a vertical cell boundary is
// read once. Lines that touch several cells are written once per sample by plain stores: no
// zeroing pass and no atomics on the TOD (the reference zeroes such lines and accumulates them
// atomically, src_lib/map2tod.cu:15-62, 81-191).
//
// tod2map accumulates in LDS with 64-bit integer atomics on fixed-point values: float LDS
// atomics (ds_add_f32) run an order of magnitude slower than integer ones on this GPU. Deposits
// are w*t*S with S = 2^(27-e) for a bound 2^(e+1) on the bulk of |t|, rounded to the nearest
// integer (int32; int64 for the rare samples above the bound, so a glitch does not coarsen the
// quantum of the cell's other samples), and pixel sums are exact in the int64 tile; the flush
// converts each pixel once, float(sum)/S, which is more accurate than float32 accumulation. The
// bound is estimated in the kernel (4x a robust per-wave bound on |t| over the first lines of
// the chunk: a few glitches there do not raise it, wave_probe_bound); a
// chunk holding a value far beyond it runs again with a bound from its exact maximum, and one
// holding Inf/NaN runs again with float LDS atomics (t2m_accumulate). The halo
// row and columns, and the cell's own first row and first HALO_X columns (other blocks' halo
// targets), are flushed to the map with atomics, the rest with a plain read-modify-write.
// Reference: src_lib/tod2map.cu:33-154.
//
// Both kernels also take per-detector gains, float response[2][ndet] (row 0 scales T, row 1
// scales Q and U; src_lib/map2tod_response.cu, tod2map_response.cu), and 16-bit TODs (FP16 or
// BF16) stored as value/scale[row] with one FP32 scale per TOD row; arithmetic stays FP32.

#include "gmm_internal.hpp"
#include "gmm_locate.hpp"

namespace gmm {

namespace {

constexpr int WAVE = 64;
constexpr int HROWS = 65;           // halo tile rows: the cell plus row 0 of the cells below
constexpr int HS = 64 + HALO_X;     // halo tile row stride: the cell plus HALO_X columns to the right
constexpr int HPLANE = HROWS * HS;  // pixels per halo tile plane

// Device-side invariant counters of a debug launch of the planned operators (debug=True), checked
// by the host after the launch (k_debug.hip): visits of each plan entry (exactly one wave walks
// each entry), evaluations of each sample (exactly one block evaluates each sample: the ownership
// and halo rules), and halo-tile accesses outside the tile (none). Unused outside debug launches.
struct Dbg
{
    uint32_t *entry_visits;   // [nmt]
    uint32_t *sample_evals;   // [nsamp]
    uint32_t *tile_bad;       // one word
};

// Tile offset o of a sample's (iy0, ix0) corner: its four corners o, o+1, o+HS, o+HS+1 lie in
// the tile.
__device__ inline void dbg_tile(const Dbg &d, int o, int extra = HS + 1)
{
    if (!(o >= 0 && o + extra < HPLANE))
        atomicAdd(d.tile_bad, 1u);
}

constexpr int M2T_THREADS = 512;
constexpr int T2M_THREADS = 1024;
constexpr int T2M_LIST = 64 * 8;     // per wave: the (entry, quad) items of a 64-entry batch

constexpr uint32_t FINITE_LIMIT = 0x7f800000u;   // bits of +Inf

__device__ inline int clamp_i(int v, int lo, int hi) { return min(max(v, lo), hi); }

__device__ inline uint32_t wave_or(uint32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v |= __shfl_xor(v, off, WAVE);
    return v;
}

__device__ inline uint32_t wave_max(uint32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v = max(v, uint32_t(__shfl_xor(v, off, WAVE)));
    return v;
}

__device__ inline void report_err(uint32_t err, uint32_t *err_word)
{
    err = wave_or(err);
    if ((threadIdx.x & 63) == 0 && err)
        atomicOr(err_word, err);
}

__device__ inline int64_t chunk_cell_offset(const Chunk &ch, const int64_t *__restrict__ cell_offsets,
                                            int nycells, int nxcells)
{
    const int cy = int(ch.icell >> 10), cx = int(ch.icell & 1023u);
    return (cy < nycells && cx < nxcells) ? cell_offsets[int64_t(cy) * nxcells + cx] : int64_t(-1);
}

// Work unit of this block. Workgroups are dispatched to the GPU's 8 XCDs round robin (block b on
// XCD b mod 8), and each XCD has its own L2. Chunks are in cell order, and a line that crosses a
// cell boundary is read by the blocks of both cells. Runs of XCD_RUN consecutive chunks are
// mapped onto one XCD (blocks b, b + 8, ...), so that neighbouring cells are mostly processed
// on one XCD at about the same time and the second read of such a line hits that XCD's L2,
// while the runs still alternate between XCDs for balance. A bijection on [0, n): whole groups
// of 8 * XCD_RUN blocks are permuted, the tail is left in place.
constexpr uint32_t NUM_XCD = 8;
constexpr uint32_t XCD_RUN = 4;      // measured: runs of 4 beat 1 (plain round robin), 16 and whole ranges

__device__ inline uint32_t xcd_chunk(uint32_t b, uint32_t n)
{
    constexpr uint32_t G = NUM_XCD * XCD_RUN;
    if (b >= (n / G) * G)
        return b;
    const uint32_t w = b % G;
    return (b - w) + (w % NUM_XCD) * XCD_RUN + w / NUM_XCD;
}

// Plan entry packed for broadcast: line index in bits [0,26), mflag in bit 31.
__device__ inline uint32_t pack_entry(uint64_t mt)
{
    return (uint32_t(mt >> ICL_SHIFT) & ICL_MASK) | ((mt & MFLAG) ? 0x80000000u : 0u);
}

// ------------------------------------------------------------------------------------------
// Lane layout of both kernels: a wave step covers 8 plan entries (lines); lane = 8*li + q takes
// samples 4q..4q+3 of line li, loaded as float4 (the Vec variant requires 16-byte aligned arrays
// and nsamp % 4 == 0, which keep every 4-sample group of every row aligned). The last line of a
// TOD may be partial: its groups past the end have no owned-quad bit, and a group cut by the end
// (nsamp % 4 != 0, scalar variant only) reads zeros past it and writes nothing there.

template<bool Vec>
__device__ inline float4 load4(const float *__restrict__ p)
{
    if (Vec)
        return *reinterpret_cast<const float4 *>(p);
    return make_float4(p[0], p[1], p[2], p[3]);
}

// The first nv (1..4) of 4 consecutive floats, zeros after: the last group of a TOD whose length
// is not a multiple of 4 (only in the scalar variant: the vector one requires nsamp % 4 == 0).
template<bool Vec>
__device__ inline float4 load4n(const float *__restrict__ p, uint32_t nv)
{
    if (Vec || nv >= 4)
        return load4<Vec>(p);
    return make_float4(p[0], (nv > 1) ? p[1] : 0.0f, (nv > 2) ? p[2] : 0.0f, 0.0f);
}


// Four consecutive floats (16-byte aligned) with a non-temporal load.
__device__ inline float4 load4_nt(const float *__restrict__ p)
{
    typedef float f4v __attribute__((ext_vector_type(4)));
    const f4v v = __builtin_nontemporal_load(reinterpret_cast<const f4v *>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}

// TOD element access, four consecutive samples at a time. TT is float, _Float16 or __bf16.
// 16-bit TODs are always accessed with 8-byte vector operations (the host requires an 8-byte
// aligned TOD), and multi-cell lines use the packed 16-bit global atomics. Whole-group stores
// are non-temporal: map2tod's output is written once and not read back by the kernel (measured:
// 5% faster map2tod). map2tod's pointing loads are non-temporal too (load4_nt): with the x halo
// a line crossing a vertical cell boundary is read by one block only, so the caches gain little
// from keeping pointing lines (measured: 1.8% faster map2tod; before the halo such loads were
// slower).
template<typename TT, bool Vec>
struct TodIO
{
    typedef TT v4 __attribute__((ext_vector_type(4)));
    typedef short s2 __attribute__((ext_vector_type(2)));
    typedef TT t2 __attribute__((ext_vector_type(2)));
    typedef float f4v __attribute__((ext_vector_type(4)));
    typedef unsigned u2v __attribute__((ext_vector_type(2)));

    __device__ static float4 load(const TT *__restrict__ p)
    {
        if constexpr (std::is_same<TT, float>::value)
            return load4<Vec>(p);
        else {
            const v4 v = *reinterpret_cast<const v4 *>(p);
            return make_float4(float(v.x), float(v.y), float(v.z), float(v.w));
        }
    }

    // The first nv samples of the group (16-bit TODs have a length divisible by 4: nv == 4).
    __device__ static float4 load_n(const TT *__restrict__ p, uint32_t nv)
    {
        if constexpr (std::is_same<TT, float>::value)
            return load4n<Vec>(p, nv);
        else
            return load(p);
    }

    __device__ static void store(TT *__restrict__ p, float4 v)
    {
        if constexpr (std::is_same<TT, float>::value) {
            if (Vec) {
                const f4v w = { v.x, v.y, v.z, v.w };
                __builtin_nontemporal_store(w, reinterpret_cast<f4v *>(p));
            }
            else {
                p[0] = v.x; p[1] = v.y; p[2] = v.z; p[3] = v.w;
            }
        }
        else {
            const v4 h = { TT(v.x), TT(v.y), TT(v.z), TT(v.w) };
            __builtin_nontemporal_store(__builtin_bit_cast(u2v, h), reinterpret_cast<u2v *>(p));
        }
    }

    __device__ static void atomic_add(TT *__restrict__ p, float4 v)
    {
        if constexpr (std::is_same<TT, float>::value) {
            unsafeAtomicAdd(p, v.x);
            unsafeAtomicAdd(p + 1, v.y);
            unsafeAtomicAdd(p + 2, v.z);
            unsafeAtomicAdd(p + 3, v.w);
        }
        else if constexpr (std::is_same<TT, _Float16>::value) {
            const t2 a = { TT(v.x), TT(v.y) }, b = { TT(v.z), TT(v.w) };
            __builtin_amdgcn_flat_atomic_fadd_v2f16(reinterpret_cast<t2 *>(p), a);
            __builtin_amdgcn_flat_atomic_fadd_v2f16(reinterpret_cast<t2 *>(p + 2), b);
        }
        else {
            const t2 a = { TT(v.x), TT(v.y) }, b = { TT(v.z), TT(v.w) };
            __builtin_amdgcn_flat_atomic_fadd_v2bf16(reinterpret_cast<s2 *>(p), __builtin_bit_cast(s2, a));
            __builtin_amdgcn_flat_atomic_fadd_v2bf16(reinterpret_cast<s2 *>(p + 2), __builtin_bit_cast(s2, b));
        }
    }
};

// Per-sample gains. response[2][ndet] (src_lib/map2tod_response.cu, tod2map_response.cu): row 0
// multiplies T, row 1 multiplies Q and U; sample s belongs to detector s / resp_div.nper. A 16-bit
// TOD holds value / scale[s / scale_div.nper].
struct Gains
{
    const float *__restrict__ resp;
    const float *__restrict__ scale;
    uint32_t resp_ndet;
    RowDiv resp_div, scale_div;
};

// Row of each of the 4 samples s0..s0+3 (s0 a multiple of 4) for rows of d.nper samples. When
// nper is a multiple of 4 the quad lies in one row (one multiplication); otherwise it crosses at
// most one row boundary when nper >= 4. Only the first nv samples exist (nv < 4 for the last
// group of a TOD whose length is not a multiple of 4): the others take the row of the last.
__device__ inline void quad_rows(uint32_t s0, const RowDiv &d, uint32_t (&r)[4], uint32_t nv = 4)
{
    const uint32_t r0 = rowdiv(s0, d);
    if ((d.nper & 3u) == 0) {
#pragma unroll
        for (int k = 0; k < 4; k++)
            r[k] = r0;
    }
    else if (d.nper >= 4) {
        const uint32_t b = (r0 + 1) * d.nper - s0;  // samples of the quad in row r0
#pragma unroll
        for (int k = 0; k < 4; k++)
            r[k] = r0 + ((min(uint32_t(k), nv - 1) >= b) ? 1u : 0u);
    }
    else {
#pragma unroll
        for (int k = 0; k < 4; k++)
            r[k] = rowdiv(s0 + min(uint32_t(k), nv - 1), d);
    }
}

// T and P gain of the 4 samples of a quad (1 without a response array).
template<bool HasResp>
__device__ inline void quad_resp(const Gains &g, uint32_t s0, float (&tr)[4], float (&pr)[4], uint32_t nv = 4)
{
    if (HasResp) {
        uint32_t d[4];
        quad_rows(s0, g.resp_div, d, nv);
        if ((g.resp_div.nper & 3u) == 0) {
            const float t = g.resp[d[0]], p = g.resp[g.resp_ndet + d[0]];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                tr[k] = t;
                pr[k] = p;
            }
        }
        else {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                tr[k] = g.resp[d[k]];
                pr[k] = g.resp[g.resp_ndet + d[k]];
            }
        }
    }
    else {
#pragma unroll
        for (int k = 0; k < 4; k++)
            tr[k] = pr[k] = 1.0f;
    }
}

// Scale of the 4 samples of a quad (16-bit TODs only; their length is a multiple of 4).
template<typename TT>
__device__ inline void quad_scale(const Gains &g, uint32_t s0, float (&sc)[4])
{
    if constexpr (std::is_same<TT, float>::value) {
#pragma unroll
        for (int k = 0; k < 4; k++)
            sc[k] = 1.0f;
    }
    else {
        uint32_t r[4];
        quad_rows(s0, g.scale_div, r);
        if ((g.scale_div.nper & 3u) == 0) {
            const float v = g.scale[r[0]];
#pragma unroll
            for (int k = 0; k < 4; k++)
                sc[k] = v;
        }
        else {
#pragma unroll
            for (int k = 0; k < 4; k++)
                sc[k] = g.scale[r[k]];
        }
    }
}

// ------------------------------------------------------------------------------------------
// Sample location for the operators: the same index and offset arithmetic as locate_pixel()
// (so the kernels agree with the planner bit for bit; the x offset x - ix is taken as fract(x),
// the same value for every x in range), with the range checks deferred:
//   - y (valid range [0, ny-1]) and non-periodic x ([0, nx-2]): each lane keeps the largest bit
//     pattern of coordinate + 0.0f (the addition turns -0.0 into +0.0). Read as unsigned
//     integers, the bit patterns of the floats in [0, hi] are exactly those <= bits(hi): any
//     negative value has the sign bit set, and +Inf and every NaN lie above bits(hi).
//   - periodic x ([-nx, 2nx]): the largest max(lo - x, x - hi) (> 0 iff out of range; these
//     exact-sign subtractions have exact signs) and a sum of the coordinates (NaN iff some x is
//     NaN).

// ------------------------------------------------------------------------------------------
// Halo tile (HROWS x HS per plane): the block's cell (64x64) plus row 64 from the cells below and
// columns W .. W+ncol-1 from the cells to the right (and below right). Normally W = 64 and ncol =
// HALO_X: the right-hand cell's first HALO_X columns, so the block can evaluate the samples of
// that cell which the ownership rule gives it (local column < 64 + HALO_X - 1). For the last
// cell column ncol = 1: in periodic mode W = (nx-1) % 64 + 1 holds the column that x-wraparound
// (ix1 = 0) reads (W = 64 when nx is a multiple of 64), otherwise there is no right-hand cell.
// Absent cells read as zeros. A sample at local (ly, lx) has its four corners at tile offsets
// o, o+1, o+HS, o+HS+1 with o = ly*HS + lx.

struct HaloCells
{
    int64_t own, down, right, diag;   // cell offsets (-1 = absent)
    int w;                            // first tile column taken from the right-hand cell
    int ncol;                         // number of such columns
    int cx;                           // the block's cell column
};

__device__ inline int64_t cell_off(const int64_t *__restrict__ cell_offsets, int cy, int cx, int nycells, int nxcells)
{
    return (cy < nycells && cx < nxcells) ? cell_offsets[int64_t(cy) * nxcells + cx] : int64_t(-1);
}

__device__ inline HaloCells halo_cells(uint32_t icell, const int64_t *__restrict__ cell_offsets, int nycells,
                                       int nxcells, int nx, bool periodic)
{
    const int cy = int(icell >> 10), cx = int(icell & 1023u);
    const int last_cx = (nx - 1) >> 6;
    const bool wrap = periodic && (cx == last_cx);
    const int rcx = wrap ? 0 : cx + 1;
    HaloCells h;
    h.own = cell_off(cell_offsets, cy, cx, nycells, nxcells);
    h.down = cell_off(cell_offsets, cy + 1, cx, nycells, nxcells);
    h.right = cell_off(cell_offsets, cy, rcx, nycells, nxcells);
    h.diag = cell_off(cell_offsets, cy + 1, rcx, nycells, nxcells);
    h.w = wrap ? ((nx - 1) & 63) + 1 : 64;
    h.ncol = (cx < last_cx) ? HALO_X : 1;
    h.cx = cx;
    return h;
}

// map2tod's tile holds each pixel's (T, Q, U) together in one float4 (the fourth lane unused),
// so the four corners of a sample are four 16-byte LDS reads from one base address. The
// interior is read as float4 groups of 4 pixels of a row from each plane (cell offsets, ystride
// and polstride are multiples of 64 floats, so a 16-byte aligned map keeps every group aligned),
// all of a thread's groups loaded before any is written, so the loads overlap.
template<int NT>
__device__ inline void stage_halo_tile(float4 *__restrict__ tile, const float *__restrict__ lmap, const HaloCells &h,
                                       int64_t ystride, int64_t polstride, int tid)
{
    constexpr int NG = CELL_NPIX / 4 / NT;            // groups of 4 pixels per thread
    static_assert(NG * 4 * NT == CELL_NPIX, "block size must divide the cell");
    const bool aligned = (reinterpret_cast<uintptr_t>(lmap) & 15) == 0;
    float4 t[NG], q[NG], u[NG];
#pragma unroll
    for (int k = 0; k < NG; k++) {
        const int g = tid + k * NT, r = g >> 4, c0 = (g & 15) << 2;
        t[k] = q[k] = u[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (h.own >= 0) {
            const float *p = lmap + h.own + r * ystride + c0;
            if (aligned) {
                t[k] = *reinterpret_cast<const float4 *>(p);
                q[k] = *reinterpret_cast<const float4 *>(p + polstride);
                u[k] = *reinterpret_cast<const float4 *>(p + 2 * polstride);
            }
            else {
                t[k] = make_float4(p[0], p[1], p[2], p[3]);
                q[k] = make_float4(p[polstride], p[polstride + 1], p[polstride + 2], p[polstride + 3]);
                u[k] = make_float4(p[2 * polstride], p[2 * polstride + 1], p[2 * polstride + 2], p[2 * polstride + 3]);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < NG; k++) {
        const int g = tid + k * NT, r = g >> 4, c0 = (g & 15) << 2;
        float4 *row = tile + r * HS + c0;
        // Column W (when W < 64) holds the wrapped column and is written below.
        if (c0 + 0 != h.w) row[0] = make_float4(t[k].x, q[k].x, u[k].x, 0.0f);
        if (c0 + 1 != h.w) row[1] = make_float4(t[k].y, q[k].y, u[k].y, 0.0f);
        if (c0 + 2 != h.w) row[2] = make_float4(t[k].z, q[k].z, u[k].z, 0.0f);
        if (c0 + 3 != h.w) row[3] = make_float4(t[k].w, q[k].w, u[k].w, 0.0f);
    }
    // Row 64 (columns 0 .. W+ncol-1) and columns W .. W+ncol-1 of rows 0..63.
    const int nrow = h.w + h.ncol, nper = nrow + 64 * h.ncol;
    for (int j = tid; j < nper; j += NT) {
        int64_t src;
        int idx;
        if (j < nrow) {
            idx = 64 * HS + j;
            if (j < h.w)
                src = (h.down >= 0) ? h.down + j : -1;
            else
                src = (h.diag >= 0) ? h.diag + (j - h.w) : -1;
        }
        else {
            const int r = (j - nrow) / h.ncol, c = (j - nrow) - r * h.ncol;
            idx = r * HS + h.w + c;
            src = (h.right >= 0) ? h.right + r * ystride + c : -1;
        }
        float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (src >= 0) {
            const float *p = lmap + src;
            v.x = p[0];
            v.y = p[polstride];
            v.z = p[2 * polstride];
        }
        tile[idx] = v;
    }
}

// T + cos2a Q + sin2a U (times the gains), bilinear over the four corners at tile index 'o'.
template<bool HasResp>
__device__ inline float m2t_interp(const float4 *__restrict__ tile, int o, float dy, float dx, float cs, float sn,
                                   float tr, float pr)
{
    if (HasResp) {
        cs *= pr;
        sn *= pr;
    }
    const float4 a = tile[o], b = tile[o + 1], c = tile[o + HS], d = tile[o + HS + 1];
    const float t00 = HasResp ? a.x * tr : a.x, t01 = HasResp ? b.x * tr : b.x;
    const float t10 = HasResp ? c.x * tr : c.x, t11 = HasResp ? d.x * tr : d.x;
    const float v00 = fmaf(sn, a.z, fmaf(cs, a.y, t00));
    const float v01 = fmaf(sn, b.z, fmaf(cs, b.y, t01));
    const float v10 = fmaf(sn, c.z, fmaf(cs, c.y, t10));
    const float v11 = fmaf(sn, d.z, fmaf(cs, d.y, t11));
    const float top = fmaf(dx, v01 - v00, v00);
    const float bot = fmaf(dx, v11 - v10, v10);
    return fmaf(dy, bot - top, top);
}

// ------------------------------------------------------------------------------------------
// sin and cos of 2*alpha from a 256-entry table in LDS: x = k * (2 pi / 256) + d with
// k = rint(x * 256 / (2 pi)) and d from a two-float Cody-Waite split of 2 pi / 256 (the first
// product is exact inside the FMA; the omitted third part, 2.8e-17, shifts d by at most
// |k| * 2.8e-17 < 2.3e-13 for |x| <= 2e5, far below one ulp of sin or cos), then
// sin x = S_k cos d + C_k sin d, cos x = C_k cos d - S_k sin d with cos d = 1 - d^2/2 and
// sin d = d (1 - d^2/6). For |x| <= 2e5 the product x * 256/(2 pi) is below 2^23, so k is within
// 0.75 of the exact quotient and |d| <= 0.0185, where the dropped terms (d^4/24 < 5e-9,
// d^5/120 < 2e-11) are below float rounding. Larger |x| use sincos_angle(). The table holds the
// float-rounded values of the double-precision sin/cos of 2 pi i / 256.

constexpr int SCT = 256;

__device__ inline void fill_sincos_table(float2 *__restrict__ tab, int nthreads)
{
    for (int i = threadIdx.x; i < SCT; i += nthreads) {
        double sd, cd;
        sincos(double(i) * (6.283185307179586 / SCT), &sd, &cd);
        tab[i] = make_float2(float(sd), float(cd));
    }
}

__device__ inline void sincos_tab(const float2 *__restrict__ tab, float x, float &s, float &c)
{
    const float k = rintf(x * 40.7436638f);              // 256 / (2 pi)
    float d = fmaf(k, -0.024543693f, x);                  // 2 pi / 256 = 0.024543693
    d = fmaf(k, 6.8299044e-10f, d);                       //            - 6.8299044e-10 (+ 2.8e-17)
    const float2 t = tab[int(k) & (SCT - 1)];
    const float z = d * d;
    const float cd = fmaf(z, -0.5f, 1.0f);
    const float sd = d * fmaf(z, -0.16666667f, 1.0f);
    s = fmaf(t.x, cd, t.y * sd);
    c = fmaf(t.y, cd, -(t.x * sd));
}

// Small: every angle of the wave's current quads satisfies |2 alpha| <= 2e5 (checked once per
// quad for the whole wave), so the table path needs no per-sample branch.
template<bool Small>
__device__ inline void sincos2(const float2 *__restrict__ tab, float alpha, float &sn, float &cs)
{
    const float x = 2.0f * alpha;
    if (Small || fabsf(x) <= 2.0e5f)
        sincos_tab(tab, x, sn, cs);
    else
        sincos_angle(x, sn, cs);
}

__device__ inline bool quad_small_angles(const float4 &a4)
{
    const float m = fmaxf(fmaxf(fabsf(a4.x), fabsf(a4.y)), fmaxf(fabsf(a4.z), fabsf(a4.w)));
    return !__any(!(m <= 1.0e5f));
}

template<bool Periodic, bool HasResp, bool Small, bool Debug = false>
__device__ inline float m2t_sample(const float4 *__restrict__ tile, const float2 *__restrict__ tab, float y, float x,
                                   float alpha, const Geom &g, int ybase, int xbase, unsigned lxlim, float tr, float pr,
                                   Checks &chk, bool &owned, const Dbg &dbg = Dbg{})
{
    int iy0, ix0;
    float dy, dx;
    locate_op<Periodic>(y, x, g, chk, iy0, ix0, dy, dx);
    const int ly = iy0 - ybase, lx = ix0 - xbase;
    owned = (unsigned(ly) < 64u) && (unsigned(lx) < lxlim);
    float sn, cs;
    sincos2<Small>(tab, alpha, sn, cs);
    const int o = owned ? (ly * HS + lx) : 0;
    if constexpr (Debug) {
        if (owned)
            dbg_tile(dbg, o);
    }
    return m2t_interp<HasResp>(tile, o, dy, dx, cs, sn, tr, pr);
}

// map2tod of plan entries [wb, we) of one TOD by the calling wave, with the halo tile of the
// block's cell (whose top-left pixel is (ybase, xbase)) in LDS. Every sample is written once, by
// the block that owns it: no atomics and no zeroing.
template<bool Periodic, bool Vec, bool HasResp, typename TT, bool Debug = false>
__device__ inline void m2t_entries(TT *__restrict__ tod, const float *__restrict__ xp,
                                   const uint64_t *__restrict__ plan_mt, const uint16_t *__restrict__ qmask,
                                   int64_t nsamp, uint32_t wb, uint32_t we, const float4 *__restrict__ tile,
                                   const float2 *__restrict__ sct, const Geom &g, int ybase, int xbase, bool accum,
                                   const Gains &gn, Checks &chk, uint32_t &err, const Dbg &dbg = Dbg{})
{
    const int lane = threadIdx.x & 63;
    const int li = lane >> 3, q = lane & 7;
    const uint32_t ncl = uint32_t((nsamp + 31) >> 5);     // the last line may be partial
    using IO = TodIO<TT, Vec>;

    for (uint32_t base = wb; base < we; base += 64) {
        const uint32_t cnt = min(64u, we - base);
        const uint32_t packed = (uint32_t(lane) < cnt) ? pack_entry(plan_mt[base + lane]) : 0u;
        const uint32_t qmw = (uint32_t(lane) < cnt) ? uint32_t(qmask[base + lane]) : 0u;

        for (uint32_t j = 0; j < cnt; j += 8) {
            const uint32_t idx = j + li;
            const uint32_t e = __shfl(packed, int(min(idx, 63u)), WAVE);
            const uint32_t qm = __shfl(qmw, int(min(idx, 63u)), WAVE);
            // Columns of the right-hand cell that this entry evaluates (ownership rule at HALO_X).
            const unsigned lxlim = (qm & QM_TAKES_RIGHT) ? unsigned(64 + HALO_X - 1) : 64u;
            const uint32_t icl = e & ICL_MASK;
            const bool act = (idx < cnt) && (icl < ncl);
            if (idx < cnt && icl >= ncl)
                err |= ERR_INCONSISTENT_NMT;
            if constexpr (Debug) {
                if (q == 0 && idx < cnt)
                    atomicAdd(dbg.entry_visits + (base + idx), 1u);
            }
            // Quads without a sample this entry evaluates are neither loaded nor evaluated.
            if (!act || !((qm >> q) & 1u))
                continue;
            const int64_t s0 = int64_t(icl) * LINE + 4 * q;
            const uint32_t nv = uint32_t(min<int64_t>(4, nsamp - s0));   // < 4 only at the end of a TOD
            const float4 y4 = Vec ? load4_nt(xp + s0) : load4n<Vec>(xp + s0, nv);
            const float4 x4 = Vec ? load4_nt(xp + nsamp + s0) : load4n<Vec>(xp + nsamp + s0, nv);
            const float4 a4 = Vec ? load4_nt(xp + 2 * nsamp + s0) : load4n<Vec>(xp + 2 * nsamp + s0, nv);
            float tr[4], pr[4], sc[4];
            quad_resp<HasResp>(gn, uint32_t(s0), tr, pr, nv);
            quad_scale<TT>(gn, uint32_t(s0), sc);

            bool o0, o1, o2, o3;
            float4 v;
            if (quad_small_angles(a4)) {
                // Every angle of the wave's groups takes the table path: straight-line code, so
                // the LDS reads of the four samples overlap.
                v.x = m2t_sample<Periodic, HasResp, true, Debug>(tile, sct, y4.x, x4.x, a4.x, g, ybase, xbase, lxlim, tr[0], pr[0], chk, o0, dbg);
                v.y = m2t_sample<Periodic, HasResp, true, Debug>(tile, sct, y4.y, x4.y, a4.y, g, ybase, xbase, lxlim, tr[1], pr[1], chk, o1, dbg);
                v.z = m2t_sample<Periodic, HasResp, true, Debug>(tile, sct, y4.z, x4.z, a4.z, g, ybase, xbase, lxlim, tr[2], pr[2], chk, o2, dbg);
                v.w = m2t_sample<Periodic, HasResp, true, Debug>(tile, sct, y4.w, x4.w, a4.w, g, ybase, xbase, lxlim, tr[3], pr[3], chk, o3, dbg);
            }
            else {
                v.x = m2t_sample<Periodic, HasResp, false, Debug>(tile, sct, y4.x, x4.x, a4.x, g, ybase, xbase, lxlim, tr[0], pr[0], chk, o0, dbg);
                __builtin_amdgcn_sched_barrier(0);
                v.y = m2t_sample<Periodic, HasResp, false, Debug>(tile, sct, y4.y, x4.y, a4.y, g, ybase, xbase, lxlim, tr[1], pr[1], chk, o1, dbg);
                __builtin_amdgcn_sched_barrier(0);
                v.z = m2t_sample<Periodic, HasResp, false, Debug>(tile, sct, y4.z, x4.z, a4.z, g, ybase, xbase, lxlim, tr[2], pr[2], chk, o2, dbg);
                __builtin_amdgcn_sched_barrier(0);
                v.w = m2t_sample<Periodic, HasResp, false, Debug>(tile, sct, y4.w, x4.w, a4.w, g, ybase, xbase, lxlim, tr[3], pr[3], chk, o3, dbg);
            }
            // Samples past the end of the TOD (read as zeros) are not written.
            o1 = o1 && nv > 1;
            o2 = o2 && nv > 2;
            o3 = o3 && nv > 3;
            if constexpr (Debug) {
                if (o0) atomicAdd(dbg.sample_evals + s0, 1u);
                if (o1) atomicAdd(dbg.sample_evals + s0 + 1, 1u);
                if (o2) atomicAdd(dbg.sample_evals + s0 + 2, 1u);
                if (o3) atomicAdd(dbg.sample_evals + s0 + 3, 1u);
            }
            if constexpr (!std::is_same<TT, float>::value) {
                // The hardware reciprocal (1 ulp) is exact enough ahead of rounding to 16 bits.
                v.x *= __builtin_amdgcn_rcpf(sc[0]);
                v.y *= __builtin_amdgcn_rcpf(sc[1]);
                v.z *= __builtin_amdgcn_rcpf(sc[2]);
                v.w *= __builtin_amdgcn_rcpf(sc[3]);
            }
            TT *t = tod + s0;
            if (o0 && o1 && o2 && o3) {
                if (accum) {
                    const float4 o = IO::load(t);
                    v.x += o.x; v.y += o.y; v.z += o.z; v.w += o.w;
                }
                IO::store(t, v);
            }
            else if (o0 || o1 || o2 || o3) {
                // A quad shared with another cell: store this block's samples one by one.
                if (o0) t[0] = TT(accum ? float(t[0]) + v.x : v.x);
                if (o1) t[1] = TT(accum ? float(t[1]) + v.y : v.y);
                if (o2) t[2] = TT(accum ? float(t[2]) + v.z : v.z);
                if (o3) t[3] = TT(accum ? float(t[3]) + v.w : v.w);
            }
        }
    }
}

// Chunks of at least this many entries take map2tod's dynamic entry batches.
constexpr uint32_t M2T_DYNAMIC_MIN = 2048;

// Entries per claim from a block's shared entry counter: 64, or fewer for a small work unit so
// that every wave of the block gets a share (a multiple of 8: a wave takes 8 entries per step).
__device__ inline uint32_t claim_batch(uint32_t n, uint32_t nwaves)
{
    const uint32_t per = (n + nwaves - 1) / nwaves;
    return min(64u, max(8u, (per + 7u) & ~7u));
}

template<bool Periodic, bool Vec, bool HasResp, typename TT, bool Debug = false>
__global__ void __launch_bounds__(M2T_THREADS, 4)
planned_map2tod_kernel(TT *__restrict__ tod, const float *__restrict__ lmap, const float *__restrict__ xp,
                       const int64_t *__restrict__ cell_offsets, const uint64_t *__restrict__ plan_mt,
                       const Chunk *__restrict__ chunks, const uint16_t *__restrict__ qmask,
                       uint32_t *__restrict__ err_word, int64_t nsamp, int ny, int nx, int nycells, int nxcells,
                       int64_t ystride, int64_t polstride, bool partial, bool accum, Gains gn, Dbg dbg)
{
    __shared__ float4 tile[HPLANE];                       // 77 KiB
    __shared__ float2 sct[SCT];
    __shared__ uint32_t s_claim;

    const Chunk ch = chunks[xcd_chunk(blockIdx.x, gridDim.x)];
    const HaloCells h = halo_cells(ch.icell, cell_offsets, nycells, nxcells, nx, Periodic);
    uint32_t err = (h.own < 0 && !partial) ? ERR_NOT_IN_PIXELIZATION : 0u;

    if (threadIdx.x == 0)
        s_claim = 0;
    fill_sincos_table(sct, M2T_THREADS);
    stage_halo_tile<M2T_THREADS>(tile, lmap, h, ystride, polstride, threadIdx.x);
    __syncthreads();

    const Geom g = make_geom(ny, nx, Periodic);
    Checks chk;
    const int ybase = int(ch.icell >> 10) << 6, xbase = int(ch.icell & 1023u) << 6;

    // A large chunk's entries go to the waves in batches of 64 from a block counter, so that the
    // waves finish together (0.9% faster on the 2048x262144 raster); a small chunk's are split
    // evenly (with the counter, per-TOD calls on TODs of a few hundred detectors were 2% slower).
    const uint32_t n = ch.end - ch.begin;
    if (n < M2T_DYNAMIC_MIN) {
        uint32_t wb, we;
        wave_slice(ch, int(threadIdx.x >> 6), M2T_THREADS / WAVE, wb, we);
        m2t_entries<Periodic, Vec, HasResp, TT, Debug>(tod, xp, plan_mt, qmask, nsamp, wb, we, tile, sct, g, ybase,
                                                       xbase, accum, gn, chk, err, dbg);
    }
    else {
        for (;;) {
            uint32_t b = 0;
            if ((threadIdx.x & 63) == 0)
                b = atomicAdd(&s_claim, 64u);
            b = __builtin_amdgcn_readfirstlane(b);
            if (b >= n)
                break;
            m2t_entries<Periodic, Vec, HasResp, TT, Debug>(tod, xp, plan_mt, qmask, nsamp, ch.begin + b,
                                                           ch.begin + min(b + 64u, n), tile, sct, g, ybase, xbase,
                                                           accum, gn, chk, err, dbg);
        }
    }
    report_err(err | chk.bits(g), err_word);
}

// ------------------------------------------------------------------------------------------

// Nearest integer of a float (ties away from -inf: floor(v + 0.5)) in one instruction; |v| < 2^31.
// Rounding to nearest keeps the fixed-point deposits unbiased (truncation would pull every
// deposit toward zero, which adds up on a TOD with an offset).
__device__ inline int cvt_rpi(float v)
{
    int r;
    asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(r) : "v"(v));
    return r;
}

// Fixed-point deposit of w*v (already scaled, |v| < 2^30) into a 64-bit LDS accumulator.
__device__ inline void dep_fixed(unsigned long long *__restrict__ a, float v)
{
    atomicAdd(a, (unsigned long long)(long long)cvt_rpi(v));
}

// The same for a scaled value beyond the 32-bit range (an outlier sample, |v| < 2^52): rounded
// to an integer in float (exact: a float of magnitude >= 2^23 is an integer already), then
// converted to int64 exactly.
__device__ inline void dep_wide(unsigned long long *__restrict__ a, float v)
{
    atomicAdd(a, (unsigned long long)(long long)rintf(v));
}

// Deposit forms: float LDS atomics (non-finite fallback), 32-bit fixed point, 64-bit fixed point.
enum { DEP_FLOAT = 0, DEP_FIXED = 1, DEP_WIDE = 2 };

// One sample's contributions to the four corners at halo-tile offset o: T, Q, U planes.
template<int Dep>
__device__ inline void deposit4(unsigned long long *__restrict__ acc, float *__restrict__ facc, int o, float dy,
                                float dx, float t, float tq, float tu)
{
    const float wy0 = 1.0f - dy, wx0 = 1.0f - dx;
    const float w[4] = { wy0 * wx0, wy0 * dx, dy * wx0, dy * dx };
    const int ofs[4] = { 0, 1, HS, HS + 1 };
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const int k = o + ofs[c];
        if (Dep == DEP_FIXED) {
            dep_fixed(acc + k, w[c] * t);
            dep_fixed(acc + HPLANE + k, w[c] * tq);
            dep_fixed(acc + 2 * HPLANE + k, w[c] * tu);
        }
        else if (Dep == DEP_WIDE) {
            dep_wide(acc + k, w[c] * t);
            dep_wide(acc + HPLANE + k, w[c] * tq);
            dep_wide(acc + 2 * HPLANE + k, w[c] * tu);
        }
        else {
            atomicAdd(facc + k, w[c] * t);
            atomicAdd(facc + HPLANE + k, w[c] * tq);
            atomicAdd(facc + 2 * HPLANE + k, w[c] * tu);
        }
    }
}

// One sample of tod2map: locate; if the block evaluates it (its (iy0, ix0) corner lies in the
// block's cell, or at local column < lxlim of the cell to the right), deposit all four corners
// into the halo tile.
template<bool Periodic, bool HasResp, int Dep, bool Small, bool Debug = false>
__device__ inline bool t2m_sample(unsigned long long *__restrict__ acc, float *__restrict__ facc,
                                  const float2 *__restrict__ tab, float y, float x, float alpha, float t, const Geom &g,
                                  int ybase, int xbase, unsigned lxlim, float tr, float pr, Checks &chk,
                                  const Dbg &dbg = Dbg{})
{
    int iy0, ix0;
    float dy, dx;
    locate_op<Periodic>(y, x, g, chk, iy0, ix0, dy, dx);
    const int ly = iy0 - ybase, lx = ix0 - xbase;
    const bool owned = (unsigned(ly) < 64u) && (unsigned(lx) < lxlim);
    if (owned) {
        float sn, cs;
        sincos2<Small>(tab, alpha, sn, cs);
        const float ts = HasResp ? t * tr : t;
        const float ps = HasResp ? t * pr : t;
        if constexpr (Debug)
            dbg_tile(dbg, ly * HS + lx);
        deposit4<Dep>(acc, facc, ly * HS + lx, dy, dx, ts, ps * cs, ps * sn);
    }
    return owned;
}

// sin and cos of x0 + d from s0 = sin x0, c0 = cos x0 for |d| <= 1e-4, where cos d = 1 and
// sin d = d are exact to float rounding (the dropped d^2/2 <= 5e-9 and d^3/6 are below half an
// ulp of values near 1, and below the float rounding of the products otherwise): two FMAs.
__device__ inline void angle_step_tiny(float s0, float c0, float d, float &sn, float &cs)
{
    sn = fmaf(d, c0, s0);
    cs = fmaf(-d, s0, c0);
}

// sin and cos of x0 + d from s0 = sin x0, c0 = cos x0 for |d| <= 0.0185, where cos d = 1 - d^2/2
// and sin d = d (1 - d^2/6) are exact to float rounding (the bound of the table's reduction).
__device__ inline void angle_step(float s0, float c0, float d, float &sn, float &cs)
{
    const float z = d * d;
    const float cd = fmaf(z, -0.5f, 1.0f);
    const float sd = d * fmaf(z, -0.16666667f, 1.0f);
    sn = fmaf(s0, cd, c0 * sd);
    cs = fmaf(c0, cd, -(s0 * sd));
}

// Per-sample state of a 4-sample group in fixed-point tod2map: tile-local corner (ly, lx), offsets
// (dy, dx), whether the block evaluates the sample (own), its T value and its polarized value
// before the angle factor (ts, ps: zero when not evaluated), and its angle.
struct QuadSamples
{
    int ly[4], lx[4];
    float dy[4], dx[4];
    bool own[4];
    float ts[4], ps[4], av[4];
};

// Deposits of samples K0..K1-1 whose corners span at most two rows and two columns (rows by..ey,
// columns bx..ex): their bilinear footprint lies in a 3x3 pixel window, summed per pixel in
// registers and added with one LDS atomic per touched pixel and plane (18 when the samples stay
// in one pixel row, at most 27) instead of 12 per sample. A pixel's sum of at most four
// contributions, each below 2^28 in magnitude, rounds to an int32. (s0, c0) are sin and cos of
// x0 = 2 alpha_0 (sample 0 of the group); 'near' and 'tiny' tell that every angle step d_k =
// 2 alpha_k - x0 of the wave's groups is at most 0.0185 or 1e-4. Returns a value that is NaN
// iff a value the block deposits is NaN.
template<bool HasResp, bool Small, bool Debug, int K0, int K1>
__device__ inline float t2m_window(unsigned long long *__restrict__ acc, const float2 *__restrict__ tab,
                                   const QuadSamples &q, int by, int bx, int ey, int ex, float s0, float c0, float x0,
                                   bool near, bool tiny, const Dbg &dbg)
{
    // The window's rows 0..1 and columns 0..1 are always touched; row 2 when the evaluated
    // samples span two rows (ey > by), column 2 likewise. Touched rows and columns lie inside
    // the tile: ey + 1 <= 64 and ex + 1 <= lxlim <= HS - 1.
    const bool row2 = ey > by, col2 = ex > bx;
    const int o = by * HS + bx;
    if constexpr (Debug)
        dbg_tile(dbg, o, (row2 ? 2 : 1) * HS + (col2 ? 2 : 1));   // the window's touched extent
    // Window weights of sample k: rows r[0..2] (times the sample's value) and columns c[0..2].
    // The sample's corner rows are ly and ly + 1, i.e. window rows 0, 1 or 1, 2; likewise
    // columns.
    auto weights = [&](int k, float v, float (&r)[3], float (&c)[3]) {
        const bool ry = q.ly[k] != by, rx = q.lx[k] != bx;
        const float v0 = v * (1.0f - q.dy[k]), v1 = v * q.dy[k];
        r[0] = ry ? 0.0f : v0;
        r[1] = ry ? v0 : v1;
        r[2] = ry ? v1 : 0.0f;
        const float wx0 = 1.0f - q.dx[k], wx1 = q.dx[k];
        c[0] = rx ? 0.0f : wx0;
        c[1] = rx ? wx0 : wx1;
        c[2] = rx ? wx1 : 0.0f;
    };
    if (!HasResp && tiny) {
        // First order in the angle steps (|d_k| <= 1e-4, where angle_step_tiny() is exact to
        // float rounding): cos 2 alpha_k = c0 - d_k s0 and sin 2 alpha_k = s0 + d_k c0, so each
        // pixel's Q = c0 A - s0 B and U = s0 A + c0 B with A the pixel's T sum (sum of w t) and
        // B the sum of w t d. Each pixel is summed where it is deposited, so untouched row 2 /
        // column 2 pixels cost nothing.
        float r[4][3], c[4][3], rd[4][3];
#pragma unroll
        for (int k = K0; k < K1; k++) {
            weights(k, q.ts[k], r[k], c[k]);
            const float d = 2.0f * q.av[k] - x0;              // d_0 = 0: not used
#pragma unroll
            for (int i = 0; i < 3; i++)
                rd[k][i] = r[k][i] * d;
        }
        auto pix = [&](int i, int j) {
            float a = 0.0f, b = 0.0f;
#pragma unroll
            for (int k = K0; k < K1; k++)
                a = fmaf(r[k][i], c[k][j], a);
#pragma unroll
            for (int k = (K0 == 0 ? 1 : K0); k < K1; k++)
                b = fmaf(rd[k][i], c[k][j], b);
            const int w = o + i * HS + j;
            dep_fixed(acc + w, a);
            dep_fixed(acc + HPLANE + w, fmaf(c0, a, -(s0 * b)));
            dep_fixed(acc + 2 * HPLANE + w, fmaf(s0, a, c0 * b));
            return a;
        };
        pix(0, 0);
        pix(0, 1);
        // Every evaluated sample's value enters window row 1 with a factor (times any column
        // weight, zero included), so this sum is NaN iff such a value is NaN.
        const float probe = pix(1, 0);
        pix(1, 1);
        if (col2) {
            pix(0, 2);
            pix(1, 2);
        }
        if (row2) {
            pix(2, 0);
            pix(2, 1);
            if (col2)
                pix(2, 2);
        }
        return probe;
    }
    float sum[3][3][3];
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++)
                sum[p][i][j] = 0.0f;
#pragma unroll
    for (int k = K0; k < K1; k++) {
        float sn = s0, cs = c0;
        if (k > 0) {
            if (tiny)
                angle_step_tiny(s0, c0, 2.0f * q.av[k] - x0, sn, cs);
            else if (near)
                angle_step(s0, c0, 2.0f * q.av[k] - x0, sn, cs);
            else
                sincos2<Small>(tab, q.av[k], sn, cs);
        }
        float r[3], c[3], rp[3];
        weights(k, q.ts[k], r, c);
        if (HasResp) {
            float cp[3];
            weights(k, q.ps[k], rp, cp);
        }
        else {
#pragma unroll
            for (int i = 0; i < 3; i++)
                rp[i] = r[i];
        }
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const float qv = rp[i] * cs, uv = rp[i] * sn;
#pragma unroll
            for (int j = 0; j < 3; j++) {
                sum[0][i][j] = fmaf(r[i], c[j], sum[0][i][j]);
                sum[1][i][j] = fmaf(qv, c[j], sum[1][i][j]);
                sum[2][i][j] = fmaf(uv, c[j], sum[2][i][j]);
            }
        }
    }
    // As above: NaN iff a deposited value is NaN (the T and, with gains, the Q sum).
    const float probe = HasResp ? sum[0][1][0] + sum[1][1][0] : sum[0][1][0];
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
                dep_fixed(acc + p * HPLANE + o + i * HS + j, sum[p][i][j]);
    if (col2) {
#pragma unroll
        for (int p = 0; p < 3; p++)
#pragma unroll
            for (int i = 0; i < 2; i++)
                dep_fixed(acc + p * HPLANE + o + i * HS + 2, sum[p][i][2]);
    }
    if (row2) {
#pragma unroll
        for (int p = 0; p < 3; p++)
#pragma unroll
            for (int j = 0; j < 2; j++)
                dep_fixed(acc + p * HPLANE + o + 2 * HS + j, sum[p][2][j]);
        if (col2) {
#pragma unroll
            for (int p = 0; p < 3; p++)
                dep_fixed(acc + p * HPLANE + o + 2 * HS + 2, sum[p][2][2]);
        }
    }
    return probe;
}

// Fixed-point tod2map of one 4-sample group (t already scaled): one 3x3 window (t2m_window())
// when, in every lane of the wave, the evaluated samples' corners span at most two rows and two
// columns (scan speeds below ~0.33 pixel per sample); otherwise sample by sample (four corners
// each). The choice is per wave, so a wave runs one of the forms. (A window per sample pair for
// the faster scans costs more: its 3x3 sums take 27 FMAs per sample when some lane of the wave
// spans three rows or columns, against 12 products, for about as many LDS atomics.) Returns a
// value that is NaN iff a value the block deposits for the group is NaN (an evaluated sample
// with t, or t times a gain, NaN): the caller's check for the float fallback.
template<bool Periodic, bool HasResp, bool Small, bool Debug = false>
__device__ inline float t2m_quad_fixed(unsigned long long *__restrict__ acc, const float2 *__restrict__ tab,
                                       const float4 &y4, const float4 &x4, const float4 &a4, const float4 &t4,
                                       const Geom &g, int ybase, int xbase, unsigned lxlim, const float (&tr)[4],
                                       const float (&pr)[4], Checks &chk, const Dbg &dbg = Dbg{}, int64_t samp0 = 0,
                                       uint32_t nv = 4, bool count = false)
{
    const float yv[4] = { y4.x, y4.y, y4.z, y4.w }, xv[4] = { x4.x, x4.y, x4.z, x4.w };
    const float tv[4] = { t4.x, t4.y, t4.z, t4.w };
    QuadSamples q;
    q.av[0] = a4.x;
    q.av[1] = a4.y;
    q.av[2] = a4.z;
    q.av[3] = a4.w;
    int by = 64, bx = 1 << 20, ey = -1, ex = -1;          // row and column range of evaluated samples
#pragma unroll
    for (int k = 0; k < 4; k++) {
        int iy0, ix0;
        locate_op<Periodic>(yv[k], xv[k], g, chk, iy0, ix0, q.dy[k], q.dx[k]);
        q.ly[k] = iy0 - ybase;
        q.lx[k] = ix0 - xbase;
        q.own[k] = (unsigned(q.ly[k]) < 64u) && (unsigned(q.lx[k]) < lxlim);
        if (q.own[k]) {
            by = min(by, q.ly[k]);
            ey = max(ey, q.ly[k]);
            bx = min(bx, q.lx[k]);
            ex = max(ex, q.lx[k]);
        }
        if constexpr (Debug) {
            // Samples past the end of the TOD read as zeros: they deposit nothing.
            if (count && q.own[k] && uint32_t(k) < nv)
                atomicAdd(dbg.sample_evals + samp0 + k, 1u);
        }
    }
    if (ey < 0)
        return 0.0f;
    // The block's share of each sample: its T value and its polarized (Q/U) value before the
    // angle factor; zero for a sample the block does not evaluate.
#pragma unroll
    for (int k = 0; k < 4; k++) {
        q.ts[k] = q.own[k] ? (HasResp ? tv[k] * tr[k] : tv[k]) : 0.0f;
        q.ps[k] = HasResp ? (q.own[k] ? tv[k] * pr[k] : 0.0f) : q.ts[k];
    }
    // sin and cos of 2 alpha: from the table for sample 0 and, when every lane's angles stay
    // within |2 alpha_k - 2 alpha_0| <= 0.0185, for the other samples by angle addition from
    // sample 0 (consecutive samples of one detector); otherwise from the table.
    float s0, c0;
    sincos2<Small>(tab, q.av[0], s0, c0);
    const float x0 = 2.0f * q.av[0];
    const float dmax = fmaxf(fmaxf(fabsf(2.0f * q.av[1] - x0), fabsf(2.0f * q.av[2] - x0)), fabsf(2.0f * q.av[3] - x0));
    const bool near = Small && !__any(!(dmax <= 0.0185f));
    const bool tiny = near && !__any(!(dmax <= 1.0e-4f));
    if (!__any(!(ey - by <= 1 && ex - bx <= 1)))
        return t2m_window<HasResp, Small, Debug, 0, 4>(acc, tab, q, by, bx, ey, ex, s0, c0, x0, near, tiny, dbg);
    float probe = 0.0f;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (!q.own[k])
            continue;
        float sn = s0, cs = c0;
        if (k > 0) {
            if (near)
                angle_step(s0, c0, 2.0f * q.av[k] - x0, sn, cs);
            else
                sincos2<Small>(tab, q.av[k], sn, cs);
        }
        probe += HasResp ? q.ts[k] + q.ps[k] : q.ts[k];
        if constexpr (Debug)
            dbg_tile(dbg, q.ly[k] * HS + q.lx[k]);
        deposit4<DEP_FIXED>(acc, nullptr, q.ly[k] * HS + q.lx[k], q.dy[k], q.dx[k], q.ts[k], q.ps[k] * cs,
                            q.ps[k] * sn);
    }
    return probe;
}

// Exclusive prefix sum over the wave of n (0 <= n <= 15) and its total, from four ballots of the
// bits of n (no LDS traffic).
__device__ inline void wave_compact_offsets(uint32_t n, uint32_t &excl, uint32_t &total)
{
    excl = 0;
    total = 0;
#pragma unroll
    for (int b = 0; b < 4; b++) {
        const uint64_t m = __ballot((n >> b) & 1u);
        excl += uint32_t(__builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u))) << b;
        total += uint32_t(__popcll(m)) << b;
    }
}

// Largest |t| (times the gains) over the first 8 plan entries of [wb, we), for the calling
// lane (lane 8*li + q reads quad q of entry wb + li); 'act' tells whether the lane read samples.
template<bool Vec, bool HasResp, typename TT>
__device__ inline float t2m_probe(const TT *__restrict__ tod, const uint64_t *__restrict__ plan_mt, int64_t nsamp,
                                  uint32_t wb, uint32_t we, const Gains &gn, bool &act)
{
    constexpr bool Q16 = !std::is_same<TT, float>::value;
    using IO = TodIO<TT, Vec>;
    const int lane = threadIdx.x & 63;
    const int li = lane >> 3, q = lane & 7;
    const uint32_t ncl = uint32_t((nsamp + 31) >> 5);     // the last line may be partial
    float m = 0.0f;
    act = false;
    const uint32_t idx = wb + uint32_t(li);
    if (idx < we) {
        const uint32_t icl = uint32_t(plan_mt[idx] >> ICL_SHIFT) & ICL_MASK;
        const int64_t s0 = int64_t(icl) * LINE + 4 * q;
        // Groups past the end of a partial last line hold no sample (nothing to read).
        if (icl < ncl && s0 < nsamp) {
            act = true;
            const uint32_t nv = uint32_t(min<int64_t>(4, nsamp - s0));
            const float4 t4 = IO::load_n(tod + s0, nv);
            float tr[4], pr[4], sc[4];
            quad_resp<HasResp>(gn, uint32_t(s0), tr, pr, nv);
            quad_scale<TT>(gn, uint32_t(s0), sc);
            const float tv[4] = { t4.x, t4.y, t4.z, t4.w };
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const float t = Q16 ? tv[k] * sc[k] : tv[k];
                m = fmaxf(m, fabsf(HasResp ? t * tr[k] : t));
                m = fmaxf(m, fabsf(HasResp ? t * pr[k] : t));
            }
        }
    }
    return m;
}

// Robust bound of a wave's probe: over the lanes that read samples (act), the smallest power of
// two that exceeds the largest |t| of at least 3/4 of them, as float bits (0 when no lane read
// anything, +Inf when that bound is not a finite float). A few glitches among the probed samples
// (up to a quarter of the lanes) leave the bound at the level of the bulk, so the cell keeps a
// fine fixed-point quantum; the glitches themselves take the 64-bit path of t2m_entries().
// Binary search over the 8-bit exponent field with wave ballots (8 steps, wave-uniform).
__device__ inline uint32_t wave_probe_bound(float m, bool act)
{
    const int nact = __popcll(__ballot(act));
    if (nact == 0)
        return 0u;
    const int need = (3 * nact + 3) / 4;
    const int ex = int(__float_as_uint(m) >> 23);          // m >= 0: the biased exponent
    int lo = 0, hi = 255;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (__popcll(__ballot(act && ex <= mid)) >= need)
            hi = mid;
        else
            lo = mid + 1;
    }
    // Every float with biased exponent lo is below 2^(lo - 126), whose bits are (lo + 1) << 23.
    return (lo >= 254) ? FINITE_LIMIT : uint32_t(lo + 1) << 23;
}

// tod2map of plan entries [wb, we) of one TOD by the calling wave into the block's halo tile:
// fixed point with scale 'scale' when 'finite', else float atomics. In fixed point a group whose
// largest |t| (times the gains) is below 'limit' takes the 32-bit path, any other group (an
// outlier) deposits sample by sample in 64-bit fixed point and counts in nwide. Tracks per lane
// the largest |t| of its outlier groups (mrun) and a sum that is NaN whenever a value it deposits
// is NaN (tsum; an outlier group adds all of its samples).
template<bool Periodic, bool Vec, bool HasResp, typename TT, bool Debug = false>
__device__ inline void t2m_entries(unsigned long long *__restrict__ acc, float *__restrict__ facc,
                                   const TT *__restrict__ tod, const float *__restrict__ xp,
                                   const uint64_t *__restrict__ plan_mt, const uint16_t *__restrict__ qmask,
                                   int64_t nsamp, uint32_t wb, uint32_t we, const float2 *__restrict__ sct,
                                   const Geom &g, int ybase, int xbase, const Gains &gn, bool finite, float scale,
                                   float limit, float &mrun, float &tsum, uint32_t &nwide, Checks &chk, uint32_t &err,
                                   uint16_t *__restrict__ wlist, const Dbg &dbg = Dbg{}, bool count = false)
{
    constexpr bool Q16 = !std::is_same<TT, float>::value;
    using IO = TodIO<TT, Vec>;
    const int lane = threadIdx.x & 63;
    const uint32_t ncl = uint32_t((nsamp + 31) >> 5);     // the last line may be partial
    // When the rows of a gain array (16-bit scales, response) are a multiple of 32 samples long, a
    // line's samples share one gain: it is loaded once per entry and broadcast with the entry
    // (FP16 tod2map 1% faster than a lookup per group; map2tod, bound by memory latency, was
    // slower this way and keeps the lookup per group).
    const bool line_sc = Q16 && (gn.scale_div.nper & 31u) == 0;
    const bool line_resp = HasResp && (gn.resp_div.nper & 31u) == 0;

    for (uint32_t base = wb; base < we; base += 64) {
        const uint32_t cnt = min(64u, we - base);
        const bool valid = uint32_t(lane) < cnt;
        const uint32_t packed = valid ? pack_entry(plan_mt[base + lane]) : 0u;
        const uint32_t qmw = valid ? uint32_t(qmask[base + lane]) : 0u;
        const uint32_t icl_l = packed & ICL_MASK;
        if (valid && icl_l >= ncl)
            err |= ERR_INCONSISTENT_NMT;
        float scl = 1.0f, trl = 1.0f, prl = 1.0f;
        if ((line_sc || line_resp) && valid && icl_l < ncl) {
            const uint32_t sl = icl_l * LINE;
            if (line_sc)
                scl = gn.scale[rowdiv(sl, gn.scale_div)];
            if (line_resp) {
                const uint32_t d = rowdiv(sl, gn.resp_div);
                trl = gn.resp[d];
                prl = gn.resp[gn.resp_ndet + d];
            }
        }
        if constexpr (Debug) {
            if (count && valid)
                atomicAdd(dbg.entry_visits + (base + lane), 1u);
        }
        // The batch's evaluated 4-sample groups (the quad bits of the entries' masks), compacted
        // into the wave's LDS list as items (entry, quad) in entry order, so that every lane of a
        // step has a group: with one lane per (entry, quad) of 8 entries, the quads that an entry
        // does not evaluate left lanes idle (13% of them on a raster, 20% on fast rotated scans).
        const uint32_t qb = (valid && icl_l < ncl) ? (qmw & 0xFFu) : 0u;
        uint32_t pos, total;
        wave_compact_offsets(__popc(qb), pos, total);
        for (uint32_t bits = qb; bits; bits &= bits - 1)
            wlist[pos++] = uint16_t((uint32_t(lane) << 3) | uint32_t(__builtin_ctz(bits)));
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

        for (uint32_t r = 0; r < total; r += 64) {
            const uint32_t item = r + uint32_t(lane);
            const uint32_t code = (item < total) ? uint32_t(wlist[item]) : 0u;
            const int li = int(code >> 3), q = int(code & 7u);
            const uint32_t e2 = __shfl(packed, li, WAVE);
            const uint32_t qm = __shfl(qmw, li, WAVE);
            const float sce = line_sc ? __shfl(scl, li, WAVE) : 1.0f;
            const float tre = line_resp ? __shfl(trl, li, WAVE) : 1.0f;
            const float pre = line_resp ? __shfl(prl, li, WAVE) : 1.0f;
            // Columns of the right-hand cell that this entry evaluates (ownership rule at HALO_X).
            const unsigned lxlim = (qm & QM_TAKES_RIGHT) ? unsigned(64 + HALO_X - 1) : 64u;
            const uint32_t icl = e2 & ICL_MASK;
            if (item >= total)
                continue;
            const int64_t s0 = int64_t(icl) * LINE + 4 * q;
            // Past the end of the TOD (nv < 4) the samples read as zeros: they deposit nothing.
            const uint32_t nv = uint32_t(min<int64_t>(4, nsamp - s0));
            // Non-temporal: the pointing and TOD stream through once, while the cell's map lines
            // (touched at block start by t2m_warm_cell) should stay cached for the flush.
            const float4 y4 = Vec ? load4_nt(xp + s0) : load4n<Vec>(xp + s0, nv);
            const float4 x4 = Vec ? load4_nt(xp + nsamp + s0) : load4n<Vec>(xp + nsamp + s0, nv);
            const float4 a4 = Vec ? load4_nt(xp + 2 * nsamp + s0) : load4n<Vec>(xp + 2 * nsamp + s0, nv);
            float4 t4;
            if constexpr (std::is_same<TT, float>::value)
                t4 = Vec ? load4_nt(tod + s0) : IO::load_n(tod + s0, nv);
            else
                t4 = IO::load_n(tod + s0, nv);
            float tr[4], pr[4], sc[4];
            if (line_resp) {
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    tr[k] = tre;
                    pr[k] = pre;
                }
            }
            else
                quad_resp<HasResp>(gn, uint32_t(s0), tr, pr, nv);
            if (line_sc) {
#pragma unroll
                for (int k = 0; k < 4; k++)
                    sc[k] = sce;
            }
            else
                quad_scale<TT>(gn, uint32_t(s0), sc);
            if (Q16) {
                t4.x *= sc[0];
                t4.y *= sc[1];
                t4.z *= sc[2];
                t4.w *= sc[3];
            }
            if (finite) {
                // NaN makes fmaxf skip the value: tsum catches it instead (t2m_quad_fixed()
                // returns a value that is NaN iff a value it deposits is; an outlier group adds
                // its samples).
                float tmax;
                if (HasResp) {
                    tmax = fmaxf(fmaxf(fabsf(t4.x * tr[0]), fabsf(t4.x * pr[0])),
                                 fmaxf(fabsf(t4.y * tr[1]), fabsf(t4.y * pr[1])));
                    tmax = fmaxf(tmax, fmaxf(fmaxf(fabsf(t4.z * tr[2]), fabsf(t4.z * pr[2])),
                                             fmaxf(fabsf(t4.w * tr[3]), fabsf(t4.w * pr[3]))));
                }
                else
                    tmax = fmaxf(fmaxf(fabsf(t4.x), fabsf(t4.y)), fmaxf(fabsf(t4.z), fabsf(t4.w)));
                t4.x *= scale;
                t4.y *= scale;
                t4.z *= scale;
                t4.w *= scale;
                if (tmax < limit) {
                    if (quad_small_angles(a4))
                        tsum += t2m_quad_fixed<Periodic, HasResp, true, Debug>(acc, sct, y4, x4, a4, t4, g, ybase, xbase, lxlim, tr, pr, chk, dbg, s0, nv, count);
                    else
                        tsum += t2m_quad_fixed<Periodic, HasResp, false, Debug>(acc, sct, y4, x4, a4, t4, g, ybase, xbase, lxlim, tr, pr, chk, dbg, s0, nv, count);
                }
                else {
                    // Only an outlier group can exceed the 64-bit range: mrun is tracked here.
                    tsum += (t4.x + t4.y) + (t4.z + t4.w);
                    mrun = fmaxf(mrun, tmax);
                    nwide++;
                    bool ob[4];
                    ob[0] = t2m_sample<Periodic, HasResp, DEP_WIDE, false, Debug>(acc, facc, sct, y4.x, x4.x, a4.x, t4.x, g, ybase, xbase, lxlim, tr[0], pr[0], chk, dbg);
                    ob[1] = t2m_sample<Periodic, HasResp, DEP_WIDE, false, Debug>(acc, facc, sct, y4.y, x4.y, a4.y, t4.y, g, ybase, xbase, lxlim, tr[1], pr[1], chk, dbg);
                    ob[2] = t2m_sample<Periodic, HasResp, DEP_WIDE, false, Debug>(acc, facc, sct, y4.z, x4.z, a4.z, t4.z, g, ybase, xbase, lxlim, tr[2], pr[2], chk, dbg);
                    ob[3] = t2m_sample<Periodic, HasResp, DEP_WIDE, false, Debug>(acc, facc, sct, y4.w, x4.w, a4.w, t4.w, g, ybase, xbase, lxlim, tr[3], pr[3], chk, dbg);
                    if constexpr (Debug) {
                        for (int k = 0; k < 4; k++)
                            if (count && ob[k] && uint32_t(k) < nv)
                                atomicAdd(dbg.sample_evals + s0 + k, 1u);
                    }
                }
            }
            else {
                bool ob[4];
                ob[0] = t2m_sample<Periodic, HasResp, DEP_FLOAT, false, Debug>(acc, facc, sct, y4.x, x4.x, a4.x, t4.x, g, ybase, xbase, lxlim, tr[0], pr[0], chk, dbg);
                ob[1] = t2m_sample<Periodic, HasResp, DEP_FLOAT, false, Debug>(acc, facc, sct, y4.y, x4.y, a4.y, t4.y, g, ybase, xbase, lxlim, tr[1], pr[1], chk, dbg);
                ob[2] = t2m_sample<Periodic, HasResp, DEP_FLOAT, false, Debug>(acc, facc, sct, y4.z, x4.z, a4.z, t4.z, g, ybase, xbase, lxlim, tr[2], pr[2], chk, dbg);
                ob[3] = t2m_sample<Periodic, HasResp, DEP_FLOAT, false, Debug>(acc, facc, sct, y4.w, x4.w, a4.w, t4.w, g, ybase, xbase, lxlim, tr[3], pr[3], chk, dbg);
                if constexpr (Debug) {
                    for (int k = 0; k < 4; k++)
                        if (count && ob[k] && uint32_t(k) < nv)
                            atomicAdd(dbg.sample_evals + s0 + k, 1u);
                }
            }
        }
        // Every lane has read its items before the next batch rewrites the list.
        __builtin_amdgcn_wave_barrier();
    }
}

// Fixed-point scale exponent for a bound with float bits 'mbits': S = 2^(27 - e).
__device__ inline int t2m_scale_exp(uint32_t mbits)
{
    return clamp_i(27 - (int(mbits >> 23) - 127), -100, 126);
}

// Clears the halo tile (16 bytes per store): all of it for the fixed-point form, the float
// view's part (the first half) otherwise. The tile is 16-byte aligned.
template<int NT>
__device__ inline void t2m_clear(unsigned long long *__restrict__ acc, bool finite, int tid)
{
    static_assert((3 * HPLANE) % 4 == 0, "the float view's half of the tile is a whole number of 16-byte stores");
    ulonglong2 *a2 = reinterpret_cast<ulonglong2 *>(acc);
    const int n = finite ? (3 * HPLANE) / 2 : (3 * HPLANE) / 4;
    for (int i = tid; i < n; i += NT)
        a2[i] = make_ulonglong2(0ull, 0ull);
}

// Map value of tile word k: in fixed point the int64 sum times the inverse scale (a power of
// two), rounded once to float (the sum is exact in double for |sum| < 2^53, from its two 32-bit
// halves; a larger sum, which only outlier deposits reach, is rounded twice); otherwise the
// float tile's value.
template<bool Finite>
__device__ inline float t2m_value(const unsigned long long *__restrict__ acc, const float *__restrict__ facc, int k,
                                  double inv_scale)
{
    if constexpr (Finite) {
        const unsigned long long v = acc[k];
        const double d = fma(double(int(uint32_t(v >> 32))), 4294967296.0, double(uint32_t(v)));
        return float(d * inv_scale);
    }
    else
        return facc[k];
}

// Touches the cell's map lines (one float per 32-byte sector) at block start, so that the
// flush's read-modify-writes find them in the caches: their loads overlap the probe, and the
// accumulation's pointing and TOD loads are non-temporal (measured: tod2map 2-3% faster). The
// loaded values only feed an empty asm statement that keeps the loads.
template<int NT>
__device__ inline void t2m_warm_cell(const float *__restrict__ lmap, const HaloCells &h, int64_t ystride,
                                     int64_t polstride, int tid)
{
    if (h.own < 0)
        return;
    uint32_t touched = 0;
    for (int g = tid; g < 3 * 64 * 8; g += NT) {
        const int pl = g >> 9, r = (g >> 3) & 63, c0 = (g & 7) << 3;
        touched |= __float_as_uint(lmap[h.own + pl * polstride + r * ystride + c0]);
    }
    asm volatile("" : : "v"(touched));
}

// Flush of the halo tile into the map. The cell's row 0 and its first HALO_X columns also
// receive other blocks' halo deposits (column 0 of cell column 0 the wrapped ones), so they and
// every split cell are added atomically; the rest of the cell with a plain read-modify-write. A
// split cell is flushed one pixel per lane (each atomic instruction then touches few cache
// lines); an unsplit one as float4 groups of 4 pixels per thread where the group allows it.
// Halo row 64 and columns W .. W+ncol-1 go to the neighbouring cells, atomically; deposits into
// absent cells are dropped. Index decoding uses shifts only (16 slots per tile row for the
// column ranges of at most 16 pixels; row 0 and row 64 take 16-pixel strides).
template<int NT, bool Finite>
__device__ inline void t2m_flush_t(float *__restrict__ lmap, const unsigned long long *__restrict__ acc,
                                   const float *__restrict__ facc, const HaloCells &h, bool shared, double inv_scale,
                                   int64_t ystride, int64_t polstride, int tid)
{
    auto val = [&](int k) { return t2m_value<Finite>(acc, facc, k, inv_scale); };
    const bool vec = !shared && ((reinterpret_cast<uintptr_t>(lmap) & 15) == 0);
    // Columns [0, catom) may receive the left-hand cell's halo deposits.
    const int catom = (h.cx >= 1) ? HALO_X : 1;
    static_assert(HALO_X <= 16, "halo columns fit the 16 slots of a tile row");
    if (h.own >= 0 && !vec) {
        // Whole cell, one pixel per lane; atomic where another block may add too.
        for (int i = tid; i < 3 * CELL_NPIX; i += NT) {
            const int pl = i >> 12, r = (i >> 6) & 63, c = i & 63;
            if (c == h.w)
                continue;
            const float v = val(pl * HPLANE + r * HS + c);
            float *gp = lmap + h.own + pl * polstride + r * ystride + c;
            if (shared || r == 0 || c < catom)
                unsafeAtomicAdd(gp, v);
            else
                *gp += v;
        }
    }
    else if (h.own >= 0) {
        // Row 0 and columns [0, catom) of rows 1..63, which other blocks add to as well: one
        // pixel per lane, atomically, so that the lanes of an atomic instruction share few cache
        // lines (the 12 columns of a row are one 48-byte run). Slot c of row 0 takes columns c,
        // c + 16, c + 32, c + 48.
        for (int i = tid; i < 3 * 64 * 16; i += NT) {
            const int pl = i >> 10, r = (i >> 4) & 63, c = i & 15;
            float *gp = lmap + h.own + pl * polstride + r * ystride;
            const int kt = pl * HPLANE + r * HS;
            if (r == 0) {
#pragma unroll
                for (int j = c; j < 64; j += 16)
                    if (j != h.w)
                        unsafeAtomicAdd(gp + j, val(kt + j));
            }
            else if (c < catom && c != h.w)
                unsafeAtomicAdd(gp + c, val(kt + c));
        }
        // Rows 1..63 from column catom on: groups of 4 pixels of a row, all planes, as float4
        // read-modify-writes. Every group's load is issued (all inside the cell, so the loads go
        // out together); a group holding row 0, a column below catom or column W is done pixel
        // by pixel.
        constexpr int NGRP = 3 * CELL_NPIX / 4;
        constexpr int NG = NGRP / NT;
        static_assert(NG * NT == NGRP, "block size must divide the cell's float4 groups");
        float4 old[NG];
#pragma unroll
        for (int k = 0; k < NG; k++) {
            const int g = tid + k * NT;
            const int pl = g >> 10, r = (g >> 4) & 63, c0 = (g & 15) << 2;
            old[k] = *reinterpret_cast<const float4 *>(lmap + h.own + pl * polstride + r * ystride + c0);
        }
#pragma unroll
        for (int k = 0; k < NG; k++) {
            const int g = tid + k * NT;
            const int pl = g >> 10, r = (g >> 4) & 63, c0 = (g & 15) << 2;
            if (r == 0 || c0 + 3 < catom)
                continue;
            const int kt = pl * HPLANE + r * HS + c0;
            float *gp = lmap + h.own + pl * polstride + r * ystride + c0;
            if (c0 >= catom && (h.w < c0 || h.w > c0 + 3)) {
                float4 v = old[k];
                v.x += val(kt);
                v.y += val(kt + 1);
                v.z += val(kt + 2);
                v.w += val(kt + 3);
                *reinterpret_cast<float4 *>(gp) = v;
            }
            else {
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int c = c0 + j;
                    if (c < catom || c == h.w)
                        continue;
                    gp[j] += val(kt + j);
                }
            }
        }
    }
    // Halo row 64 (columns 0 .. W+ncol-1: slot c takes c, c + 16, ...) and columns W .. W+ncol-1
    // of rows 0..63 (slot c < ncol), to the neighbouring cells.
    const int nrow = h.w + h.ncol;
    constexpr int PSLOTS = 65 * 16;
    for (int i = tid; i < 3 * PSLOTS; i += NT) {
        const int pl = i / PSLOTS;
        const int rem = i - pl * PSLOTS;
        const int r = rem >> 4, c = rem & 15;
        const int kt = pl * HPLANE + r * HS;
        if (r == 64) {
            for (int j = c; j < nrow; j += 16) {
                const int64_t dst = (j < h.w) ? ((h.down >= 0) ? h.down + j : int64_t(-1))
                                              : ((h.diag >= 0) ? h.diag + (j - h.w) : int64_t(-1));
                if (dst >= 0)
                    unsafeAtomicAdd(lmap + dst + pl * polstride, val(kt + j));
            }
        }
        else if (c < h.ncol && h.right >= 0)
            unsafeAtomicAdd(lmap + h.right + pl * polstride + r * ystride + c, val(kt + h.w + c));
    }
}

template<int NT>
__device__ inline void t2m_flush(float *__restrict__ lmap, const unsigned long long *__restrict__ acc,
                                 const float *__restrict__ facc, const HaloCells &h, bool shared, bool finite,
                                 float inv_scale, int64_t ystride, int64_t polstride, int tid)
{
    if (finite)
        t2m_flush_t<NT, true>(lmap, acc, facc, h, shared, double(inv_scale), ystride, polstride, tid);
    else
        t2m_flush_t<NT, false>(lmap, acc, facc, h, shared, double(inv_scale), ystride, polstride, tid);
}

// Block-shared words of t2m_accumulate(): the probe bound (float bits), and per round a redo
// word (largest offending |t|, as float bits) and a count of outlier groups.
// W_CLAIM: per round, the next entry batch of the block's waves.
enum { W_MAX = 0, W_REDO = 1, W_WIDE = 4, W_CLAIM = 7, T2M_WORDS = 10 };

// Outliers: the 64-bit path takes scaled values up to 2^WIDE_BITS times the 32-bit range (|v| <
// 2^52), so a pixel's int64 sum stays exact for the first 2^11 deposits at the top of that range.
constexpr int WIDE_BITS = 24;

__device__ inline uint32_t wave_sum(uint32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v += uint32_t(__shfl_xor(v, off, WAVE));
    return v;
}

// Deposits of one block's work unit (nentries plan entries) into its halo tile, in fixed point.
// Before the call the tile is clear (fixed-point form), words[W_MAX] holds the float bits of the
// block's probe bound (the largest wave_probe_bound() of its waves), the other words are zero,
// and the block has synchronized.
// run(finite, scale, limit, mrun, tsum, nwide, chk, first) deposits the calling wave's samples
// (t2m_entries()). With a bound 2^(e+1) on the bulk of |t|, deposits are w * t * S with
// S = 2^(27 - e): a group below the bound converts each (3x3-window) pixel sum to int32 (|.| <
// 2^30); a group above it (an outlier such as a glitch) deposits each contribution as int64,
// exactly, without coarsening the quantum of the other samples of the cell. The tile sums in
// int64; every conversion rounds to nearest. The bound starts from an estimate, 4 x the probe
// bound. A value beyond the 64-bit range (2^WIDE_BITS times the bound) or a non-finite one
// makes the block discard its tile and run again: with float atomics when the exact maximum M
// is not finite; else with the bound at M when at least one group in 8 was an outlier (the
// estimate was far too low, e.g. zeros at the probed lines), or at 2^-WIDE_BITS M otherwise
// (rare huge outliers: the bulk keeps a fine quantum). Either bound puts M inside the 64-bit
// range, so the second round is the last. Each round has its own words. Returns whether the
// tile is fixed point and its inverse scale; ends synchronized.
template<int NT, typename Run>
__device__ inline bool t2m_accumulate(unsigned long long *__restrict__ acc, uint32_t *words, uint32_t nentries,
                                      const Geom &g, Run &&run, float &inv_scale, uint32_t &err, int tid)
{
    const int lane = threadIdx.x & 63;
    uint32_t mbits = words[W_MAX];
    mbits = (mbits < FINITE_LIMIT) ? __float_as_uint(4.0f * __uint_as_float(mbits)) : mbits;
    bool finite = (mbits < FINITE_LIMIT);

    for (int pass = 0;; pass++) {
        const int e = int(mbits >> 23) - 127;            // binary exponent of the bound
        const float scale = __uint_as_float(uint32_t(t2m_scale_exp(mbits) + 127) << 23);
        const float limit = __uint_as_float(uint32_t(e + 128) << 23);   // 2^(e+1)
        const uint32_t wide_limit = uint32_t(min(e + 128 + WIDE_BITS, 255)) << 23;
        float mrun = 0.0f;                               // largest |t| of this lane's outlier groups
        float tsum = 0.0f;                               // NaN if some deposited value is NaN
        uint32_t nwide = 0;                              // outlier groups of this lane
        Checks chk;
        run(finite, scale, limit, mrun, tsum, nwide, chk, pass == 0);
        err |= chk.bits(g);
        if (!finite)
            break;
        const bool lane_nan = !(tsum == tsum);
        const int r = min(pass, 2);
        const uint32_t mw = wave_max(lane_nan ? FINITE_LIMIT : __float_as_uint(mrun));
        const uint32_t nw = wave_sum(nwide);
        if (lane == 0) {
            if (mw >= wide_limit)
                atomicMax(words + W_REDO + r, mw);
            if (nw)
                atomicAdd(words + W_WIDE + r, nw);
        }
        __syncthreads();
        const uint32_t redo = words[W_REDO + r];
        if (redo == 0)
            break;
        finite = (redo < FINITE_LIMIT);
        if (!finite)
            mbits = redo;
        else if (words[W_WIDE + r] >= nentries)
            mbits = redo;
        else
            mbits = uint32_t(max(int(redo >> 23) - 127 - WIDE_BITS, -126) + 127) << 23;
        t2m_clear<NT>(acc, finite, tid);
        __syncthreads();
    }
    __syncthreads();
    inv_scale = __uint_as_float(uint32_t(127 - t2m_scale_exp(mbits)) << 23);
    return finite;
}

template<bool Periodic, bool Vec, bool HasResp, typename TT, bool Debug = false>
__global__ void __launch_bounds__(T2M_THREADS)
planned_tod2map_kernel(float *__restrict__ lmap, const TT *__restrict__ tod, const float *__restrict__ xp,
                       const int64_t *__restrict__ cell_offsets, const uint64_t *__restrict__ plan_mt,
                       const Chunk *__restrict__ chunks, const uint16_t *__restrict__ qmask,
                       uint32_t *__restrict__ err_word, int64_t nsamp, int ny, int nx, int nycells, int nxcells,
                       int64_t ystride, int64_t polstride, bool partial, Gains gn, Dbg dbg)
{
    __shared__ __align__(16) unsigned long long acc[3 * HPLANE];       // 116 KiB
    __shared__ float2 sct[SCT];
    __shared__ uint32_t s_words[T2M_WORDS];
    __shared__ uint16_t s_list[T2M_THREADS / WAVE][T2M_LIST];            // 16 KiB
    float *facc = reinterpret_cast<float *>(acc);         // float tile for the non-finite fallback

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const Chunk ch = chunks[xcd_chunk(blockIdx.x, gridDim.x)];
    const HaloCells h = halo_cells(ch.icell, cell_offsets, nycells, nxcells, nx, Periodic);
    uint32_t err = (h.own < 0 && !partial) ? ERR_NOT_IN_PIXELIZATION : 0u;

    // Samples owned by an absent cell still deposit into present neighbours through the halo,
    // so every block runs; cells that are absent receive nothing (tod2map_reference.cu:14-46
    // skips them too).

    const Geom g = make_geom(ny, nx, Periodic);
    const int ybase = int(ch.icell >> 10) << 6, xbase = int(ch.icell & 1023u) << 6;
    uint32_t wb, we;
    wave_slice(ch, wave, T2M_THREADS / WAVE, wb, we);

    // The fixed-point bound is estimated in the kernel from the first lines of every wave's
    // slice (t2m_accumulate() describes the scheme): no pass over the TOD is needed first.
    if (threadIdx.x < T2M_WORDS)
        s_words[threadIdx.x] = 0;
    fill_sincos_table(sct, T2M_THREADS);
    __syncthreads();
    {
        bool act;
        const float m = t2m_probe<Vec, HasResp, TT>(tod, plan_mt, nsamp, wb, we, gn, act);
        const uint32_t mw = wave_probe_bound(m, act);
        if (lane == 0)
            atomicMax(&s_words[W_MAX], mw);
    }
    t2m_clear<T2M_THREADS>(acc, true, threadIdx.x);
    t2m_warm_cell<T2M_THREADS>(lmap, h, ystride, polstride, threadIdx.x);
    __syncthreads();

    // The waves take the chunk's entries in batches (claim_batch()) from a block counter (one per
    // round), so that they finish together: entries differ in cost (owned groups, window forms),
    // and an even static split left the block waiting on its slowest wave (3.4% of the kernel).
    float inv_scale;
    int round = 0;
    const bool finite = t2m_accumulate<T2M_THREADS>(
        acc, s_words, ch.end - ch.begin, g,
        [&](bool fin, float scale, float limit, float &mrun, float &tsum, uint32_t &nwide, Checks &chk, bool first) {
            uint32_t *claim = s_words + W_CLAIM + min(round, 2);
            round++;
            const uint32_t n = ch.end - ch.begin, bs = claim_batch(n, T2M_THREADS / WAVE);
            for (;;) {
                uint32_t b = 0;
                if (lane == 0)
                    b = atomicAdd(claim, bs);
                b = __builtin_amdgcn_readfirstlane(b);
                if (b >= n)
                    break;
                t2m_entries<Periodic, Vec, HasResp, TT, Debug>(acc, facc, tod, xp, plan_mt, qmask, nsamp, ch.begin + b,
                                                               ch.begin + min(b + bs, n), sct, g, ybase, xbase, gn,
                                                               fin, scale, limit, mrun, tsum, nwide, chk, err,
                                                               s_list[wave], dbg, first);
            }
        },
        inv_scale, err, threadIdx.x);
    t2m_flush<T2M_THREADS>(lmap, acc, facc, h, (ch.flags & CHUNK_SHARED) != 0, finite, inv_scale, ystride,
                           polstride, threadIdx.x);
    report_err(err, err_word);
}

// ------------------------------------------------------------------------------------------
// Batched operators: several TODs, each with its own plan, against one local map, in one launch.
// A work item is a run of plan entries of one cell, gathered from the plans of several TODs as
// a list of segments (BatchSeg: entries [begin, end) of TOD 'tod'). The block of an item stages
// (map2tod) or accumulates (tod2map) the cell's halo tile once for all of its segments; its
// waves split the item's entries evenly, across segment boundaries. With 'persistent', a grid
// of a few blocks per CU takes items from a queue counter, in the host's order (largest first);
// the last block to finish resets the counters, so a launch needs no host-side reset and can be
// captured in a graph. Otherwise there is one block per item.

// The tables themselves are separate __restrict__ kernel parameters, which lets the compiler
// read them through the scalar unit (wave-uniform values in scalar registers).
struct BatchArgs
{
    uint32_t nitems;
    uint32_t persistent;
    uint32_t *counters;            // 2 words: next item, finished blocks
    uint32_t *err_word;
    int ny, nx, nycells, nxcells;
    int64_t ystride, polstride;
    bool partial, accum;
};

// Item of the block's next work unit, the same in every lane: the next one from the queue
// (persistent grid) or the block's own, with runs of neighbouring items on one XCD as for the
// plain operators. Ends synchronized. (Other threads may reset block-shared words before the
// call: the synchronization covers them.)
__device__ inline uint32_t batch_next_item(const BatchArgs &ba, uint32_t *s_item)
{
    if (!ba.persistent)
        return xcd_chunk(blockIdx.x, gridDim.x);
    if (threadIdx.x == 0)
        *s_item = atomicAdd(ba.counters, 1u);
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(*s_item);
}

// threadIdx.x, opaque to the compiler, so that the thread-dependent index arithmetic of the
// per-item staging and flush loops is not hoisted out of the item loop, where it would hold
// registers across the whole item.
__device__ inline int opaque_tid()
{
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
}

__device__ inline void batch_finish(const BatchArgs &ba)
{
    if (ba.persistent && threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(ba.counters + 1, 1u) == gridDim.x - 1) {
            // Every block has taken its last item: the queue can be rewound for the next launch.
            atomicExch(ba.counters, 0u);
            atomicExch(ba.counters + 1, 0u);
        }
    }
}

// Calls f(seg, begin, end) for the parts of the item's segments inside the wave's entry slice
// [wb, we) (item-relative entry indices).
template<typename F>
__device__ inline void batch_wave_segments(const BatchSeg *__restrict__ segs, const BatchItem &it, uint32_t wb,
                                           uint32_t we, F &&f)
{
    uint32_t pos = 0;
    for (uint32_t s = it.seg_begin; s < it.seg_end && pos < we; s++) {
        const BatchSeg sg = segs[s];
        const uint32_t len = sg.end - sg.begin;
        const uint32_t lo = max(wb, pos), hi = min(we, pos + len);
        if (lo < hi)
            f(sg, sg.begin + (lo - pos), sg.begin + (hi - pos));
        pos += len;
    }
}

__device__ inline Gains batch_gains(const BatchTod &td)
{
    return Gains{ td.resp, td.scale, td.resp_ndet, td.resp_div, td.scale_div };
}

template<bool Periodic, typename TT, int NT, bool HasResp = false>
__global__ void __launch_bounds__(NT, 4)
batched_map2tod_kernel(const float *__restrict__ lmap, const BatchTod *__restrict__ tods,
                       const BatchSeg *__restrict__ segs, const BatchItem *__restrict__ items,
                       const int64_t *__restrict__ cell_offsets, BatchArgs ba)
{
    __shared__ float4 tile[HPLANE];                       // 77 KiB
    __shared__ float2 sct[SCT];
    __shared__ uint32_t s_item;

    fill_sincos_table(sct, NT);
    const Geom g = make_geom(ba.ny, ba.nx, Periodic);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    Checks chk;
    uint32_t err = 0;

    for (;;) {
        const uint32_t item = batch_next_item(ba, &s_item);
        if (item >= ba.nitems)
            break;
        const BatchItem it = items[item];
        const HaloCells h = halo_cells(it.icell, cell_offsets, ba.nycells, ba.nxcells, ba.nx, Periodic);
        if (h.own < 0 && !ba.partial)
            err |= ERR_NOT_IN_PIXELIZATION;
        stage_halo_tile<NT>(tile, lmap, h, ba.ystride, ba.polstride, opaque_tid());
        __syncthreads();

        const int ybase = int(it.icell >> 10) << 6, xbase = int(it.icell & 1023u) << 6;
        const uint32_t wb = uint32_t((uint64_t(it.nent) * wave) / (NT / WAVE));
        const uint32_t we = uint32_t((uint64_t(it.nent) * (wave + 1)) / (NT / WAVE));
        batch_wave_segments(segs, it, wb, we, [&](const BatchSeg &sg, uint32_t b, uint32_t e) {
            const BatchTod td = tods[sg.tod];
            m2t_entries<Periodic, true, HasResp, TT>(static_cast<TT *>(td.tod), td.xp, td.plan_mt, td.qmask, td.nsamp,
                                                   b, e, tile, sct, g, ybase, xbase, ba.accum, batch_gains(td), chk,
                                                   err);
        });
        __syncthreads();     // the tile and s_item are rewritten for the next item
        if (!ba.persistent)
            break;
    }
    batch_finish(ba);
    report_err(err | chk.bits(g), ba.err_word);
}

template<bool Periodic, typename TT, bool HasResp = false>
__global__ void __launch_bounds__(T2M_THREADS)
batched_tod2map_kernel(float *__restrict__ lmap, const BatchTod *__restrict__ tods,
                       const BatchSeg *__restrict__ segs, const BatchItem *__restrict__ items,
                       const int64_t *__restrict__ cell_offsets, BatchArgs ba)
{
    __shared__ __align__(16) unsigned long long acc[3 * HPLANE];       // 116 KiB
    __shared__ float2 sct[SCT];
    __shared__ uint32_t s_words[T2M_WORDS], s_item;
    __shared__ uint16_t s_list[T2M_THREADS / WAVE][T2M_LIST];            // 16 KiB
    float *facc = reinterpret_cast<float *>(acc);

    fill_sincos_table(sct, T2M_THREADS);
    const Geom g = make_geom(ba.ny, ba.nx, Periodic);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    uint32_t err = 0;

    for (;;) {
        if (threadIdx.x >= 1 && threadIdx.x <= T2M_WORDS)   // thread 0 may be writing s_item
            s_words[threadIdx.x - 1] = 0;
        if (!ba.persistent)
            __syncthreads();
        const uint32_t item = batch_next_item(ba, &s_item);
        if (item >= ba.nitems)
            break;
        const BatchItem it = items[item];
        const HaloCells h = halo_cells(it.icell, cell_offsets, ba.nycells, ba.nxcells, ba.nx, Periodic);
        if (h.own < 0 && !ba.partial)
            err |= ERR_NOT_IN_PIXELIZATION;
        const int ybase = int(it.icell >> 10) << 6, xbase = int(it.icell & 1023u) << 6;
        const uint32_t wb = uint32_t((uint64_t(it.nent) * wave) / (T2M_THREADS / WAVE));
        const uint32_t we = uint32_t((uint64_t(it.nent) * (wave + 1)) / (T2M_THREADS / WAVE));

        // Probe: the first lines of the wave's first segment.
        {
            float m = 0.0f;
            bool first = true, act = false;
            batch_wave_segments(segs, it, wb, we, [&](const BatchSeg &sg, uint32_t b, uint32_t e) {
                if (first) {
                    const BatchTod td = tods[sg.tod];
                    m = t2m_probe<true, HasResp, TT>(static_cast<const TT *>(td.tod), td.plan_mt, td.nsamp, b, e,
                                                   batch_gains(td), act);
                    first = false;
                }
            });
            const uint32_t mw = wave_probe_bound(m, act);
            if (lane == 0)
                atomicMax(&s_words[W_MAX], mw);
        }
        t2m_clear<T2M_THREADS>(acc, true, opaque_tid());
        t2m_warm_cell<T2M_THREADS>(lmap, h, ba.ystride, ba.polstride, opaque_tid());
        __syncthreads();

        // Entries in batches from a block counter per round, as in the plain kernel.
        const uint32_t bs = claim_batch(it.nent, T2M_THREADS / WAVE);
        float inv_scale;
        int round = 0;
        const bool finite = t2m_accumulate<T2M_THREADS>(
            acc, s_words, it.nent, g,
            [&](bool fin, float scale, float limit, float &mrun, float &tsum, uint32_t &nwide, Checks &chk, bool) {
                uint32_t *claim = s_words + W_CLAIM + min(round, 2);
                round++;
                for (;;) {
                    uint32_t b0 = 0;
                    if (lane == 0)
                        b0 = atomicAdd(claim, bs);
                    b0 = __builtin_amdgcn_readfirstlane(b0);
                    if (b0 >= it.nent)
                        break;
                    batch_wave_segments(segs, it, b0, min(b0 + bs, it.nent),
                                        [&](const BatchSeg &sg, uint32_t b, uint32_t e) {
                        const BatchTod td = tods[sg.tod];
                        t2m_entries<Periodic, true, HasResp, TT>(acc, facc, static_cast<const TT *>(td.tod), td.xp,
                                                               td.plan_mt, td.qmask, td.nsamp, b, e, sct, g, ybase,
                                                               xbase, batch_gains(td), fin, scale, limit, mrun, tsum,
                                                               nwide, chk, err, s_list[wave]);
                    });
                }
            },
            inv_scale, err, opaque_tid());
        t2m_flush<T2M_THREADS>(lmap, acc, facc, h, (it.flags & CHUNK_SHARED) != 0, finite, inv_scale, ba.ystride,
                               ba.polstride, opaque_tid());
        __syncthreads();     // the tile and the block-shared words are reset for the next item
        if (!ba.persistent)
            break;
    }
    batch_finish(ba);
    report_err(err, ba.err_word);
}

template<bool HasResp, typename TT>
void launch_m2t(TT *tod, const float *lmap, const float *xp, const OpArgs &a, const gmm_plan &plan,
                bool accum, const Gains &gn, hipStream_t stream)
{
    const gmm_lpix &lp = *a.lpix;
    uint32_t *err_word = plan.err + (plan.pp.n_err - 1);
    dim3 grid(unsigned(plan.nchunks)), block(M2T_THREADS);
    const bool vec = (((reinterpret_cast<uintptr_t>(tod) | reinterpret_cast<uintptr_t>(xp)) & 15) == 0) && (a.nsamp % 4 == 0);
    auto k = lp.periodic_xcoord ? (vec ? planned_map2tod_kernel<true, true, HasResp, TT> : planned_map2tod_kernel<true, false, HasResp, TT>)
                                : (vec ? planned_map2tod_kernel<false, true, HasResp, TT> : planned_map2tod_kernel<false, false, HasResp, TT>);
    Dbg dbg{ nullptr, nullptr, nullptr };
    if constexpr (std::is_same<TT, float>::value) {
        // debug=True: the instrumented instantiation (float32 TODs).
        if (a.debug) {
            dbg = Dbg{ a.debug->entry_visits, a.debug->sample_evals, a.debug->tile_bad };
            k = lp.periodic_xcoord ? (vec ? planned_map2tod_kernel<true, true, HasResp, TT, true> : planned_map2tod_kernel<true, false, HasResp, TT, true>)
                                   : (vec ? planned_map2tod_kernel<false, true, HasResp, TT, true> : planned_map2tod_kernel<false, false, HasResp, TT, true>);
        }
    }
    kernel_timer_mark(TK_MAP2TOD, false, stream);
    k<<<grid, block, 0, stream>>>(tod, lmap, xp, lp.offsets_gpu, plan.plan_mt, plan.chunks, plan.qmask, err_word,
                                  a.nsamp, int(lp.nypix_global), int(lp.nxpix_global), int(lp.nycells),
                                  int(lp.nxcells), lp.ystride, lp.polstride, a.partial, accum, gn, dbg);
    GMM_HIP(hipGetLastError());
    kernel_timer_mark(TK_MAP2TOD, true, stream);
}

template<bool HasResp, typename TT>
void launch_t2m(float *lmap, const TT *tod, const float *xp, const OpArgs &a, const gmm_plan &plan,
                const Gains &gn, hipStream_t stream)
{
    const gmm_lpix &lp = *a.lpix;
    uint32_t *err_word = plan.err + (plan.pp.n_err - 1);
    dim3 grid(unsigned(plan.nchunks)), block(T2M_THREADS);
    const bool vec = (((reinterpret_cast<uintptr_t>(tod) | reinterpret_cast<uintptr_t>(xp)) & 15) == 0) && (a.nsamp % 4 == 0);
    auto k = lp.periodic_xcoord ? (vec ? planned_tod2map_kernel<true, true, HasResp, TT> : planned_tod2map_kernel<true, false, HasResp, TT>)
                                : (vec ? planned_tod2map_kernel<false, true, HasResp, TT> : planned_tod2map_kernel<false, false, HasResp, TT>);
    Dbg dbg{ nullptr, nullptr, nullptr };
    if constexpr (std::is_same<TT, float>::value) {
        if (a.debug) {
            dbg = Dbg{ a.debug->entry_visits, a.debug->sample_evals, a.debug->tile_bad };
            k = lp.periodic_xcoord ? (vec ? planned_tod2map_kernel<true, true, HasResp, TT, true> : planned_tod2map_kernel<true, false, HasResp, TT, true>)
                                   : (vec ? planned_tod2map_kernel<false, true, HasResp, TT, true> : planned_tod2map_kernel<false, false, HasResp, TT, true>);
        }
    }
    kernel_timer_mark(TK_TOD2MAP, false, stream);
    k<<<grid, block, 0, stream>>>(lmap, tod, xp, lp.offsets_gpu, plan.plan_mt, plan.chunks, plan.qmask, err_word,
                                  a.nsamp, int(lp.nypix_global), int(lp.nxpix_global), int(lp.nycells),
                                  int(lp.nxcells), lp.ystride, lp.polstride, a.partial, gn, dbg);
    GMM_HIP(hipGetLastError());
    kernel_timer_mark(TK_TOD2MAP, true, stream);
}

Gains make_gains(const OpArgs &a, const TodGains &g)
{
    Gains gn{g.response, g.scale, 1u, RowDiv(), RowDiv()};
    if (g.response) {
        gn.resp_ndet = uint32_t(g.resp_ndet);
        gn.resp_div = make_rowdiv(uint32_t(a.nsamp / g.resp_ndet));
    }
    if (g.scale)
        gn.scale_div = make_rowdiv(uint32_t(a.nsamp / g.scale_nrows));
    return gn;
}

template<typename TT>
void map2tod_typed(TT *tod, const float *lmap, const float *xp, const OpArgs &a, const gmm_plan &plan,
                   bool accum, const TodGains &g, hipStream_t stream)
{
    if (plan.nchunks == 0)
        return;
    const Gains gn = make_gains(a, g);
    if (g.response)
        launch_m2t<true, TT>(tod, lmap, xp, a, plan, accum, gn, stream);
    else
        launch_m2t<false, TT>(tod, lmap, xp, a, plan, accum, gn, stream);
}

template<typename TT>
void tod2map_typed(float *lmap, const TT *tod, const float *xp, const OpArgs &a, const gmm_plan &plan,
                   const TodGains &g, hipStream_t stream)
{
    if (plan.nchunks == 0)
        return;
    const Gains gn = make_gains(a, g);
    if (g.response)
        launch_t2m<true, TT>(lmap, tod, xp, a, plan, gn, stream);
    else
        launch_t2m<false, TT>(lmap, tod, xp, a, plan, gn, stream);
}

BatchArgs batch_args(const BatchLaunch &bl, const gmm_lpix &lp, bool partial, bool accum)
{
    BatchArgs ba;
    ba.nitems = uint32_t(bl.nitems);
    ba.persistent = (bl.grid > 0) ? 1u : 0u;
    ba.counters = bl.counters;
    ba.err_word = bl.err_word;
    ba.ny = int(lp.nypix_global);
    ba.nx = int(lp.nxpix_global);
    ba.nycells = int(lp.nycells);
    ba.nxcells = int(lp.nxcells);
    ba.ystride = lp.ystride;
    ba.polstride = lp.polstride;
    ba.partial = partial;
    ba.accum = accum;
    return ba;
}

unsigned batch_grid(const BatchLaunch &bl)
{
    return unsigned((bl.grid > 0) ? std::min<int64_t>(bl.grid, bl.nitems) : bl.nitems);
}

template<typename TT>
void batched_m2t_typed(const BatchLaunch &bl, const float *lmap, const int64_t *cell_offsets, const BatchArgs &ba,
                       bool periodic, hipStream_t stream)
{
    const unsigned grid = batch_grid(bl);
    if (bl.threads == 256) {
        auto k = periodic ? batched_map2tod_kernel<true, TT, 256> : batched_map2tod_kernel<false, TT, 256>;
        k<<<grid, 256, 0, stream>>>(lmap, bl.tods, bl.segs, bl.items, cell_offsets, ba);
    }
    else {
        auto k = periodic ? batched_map2tod_kernel<true, TT, 512> : batched_map2tod_kernel<false, TT, 512>;
        k<<<grid, 512, 0, stream>>>(lmap, bl.tods, bl.segs, bl.items, cell_offsets, ba);
    }
    GMM_HIP(hipGetLastError());
}

template<typename TT>
void batched_t2m_typed(const BatchLaunch &bl, float *lmap, const int64_t *cell_offsets, const BatchArgs &ba,
                       bool periodic, hipStream_t stream)
{
    auto k = periodic ? batched_tod2map_kernel<true, TT> : batched_tod2map_kernel<false, TT>;
    k<<<batch_grid(bl), T2M_THREADS, 0, stream>>>(lmap, bl.tods, bl.segs, bl.items, cell_offsets, ba);
    GMM_HIP(hipGetLastError());
}

// Batched operators with per-detector gains (float32 TODs).
void batched_m2t_resp(const BatchLaunch &bl, const float *lmap, const int64_t *cell_offsets, const BatchArgs &ba,
                      bool periodic, hipStream_t stream)
{
    const unsigned grid = batch_grid(bl);
    if (bl.threads == 256) {
        auto k = periodic ? batched_map2tod_kernel<true, float, 256, true> : batched_map2tod_kernel<false, float, 256, true>;
        k<<<grid, 256, 0, stream>>>(lmap, bl.tods, bl.segs, bl.items, cell_offsets, ba);
    }
    else {
        auto k = periodic ? batched_map2tod_kernel<true, float, 512, true> : batched_map2tod_kernel<false, float, 512, true>;
        k<<<grid, 512, 0, stream>>>(lmap, bl.tods, bl.segs, bl.items, cell_offsets, ba);
    }
    GMM_HIP(hipGetLastError());
}

void batched_t2m_resp(const BatchLaunch &bl, float *lmap, const int64_t *cell_offsets, const BatchArgs &ba,
                      bool periodic, hipStream_t stream)
{
    auto k = periodic ? batched_tod2map_kernel<true, float, true> : batched_tod2map_kernel<false, float, true>;
    k<<<batch_grid(bl), T2M_THREADS, 0, stream>>>(lmap, bl.tods, bl.segs, bl.items, cell_offsets, ba);
    GMM_HIP(hipGetLastError());
}

}  // namespace

// ==========================================================================================

void launch_planned_map2tod(void *tod, const float *lmap, const float *xp, const OpArgs &a,
                            const gmm_plan &plan, bool accum, const TodGains &g, hipStream_t stream)
{
    switch (g.tod_type) {
    case TOD_F32: map2tod_typed(static_cast<float *>(tod), lmap, xp, a, plan, accum, g, stream); break;
    case TOD_F16: map2tod_typed(static_cast<_Float16 *>(tod), lmap, xp, a, plan, accum, g, stream); break;
    case TOD_BF16: map2tod_typed(static_cast<__bf16 *>(tod), lmap, xp, a, plan, accum, g, stream); break;
    default: fail("map2tod: unknown TOD element type");
    }
}

void launch_planned_tod2map(float *lmap, const void *tod, const float *xp, const OpArgs &a,
                            const gmm_plan &plan, const TodGains &g, hipStream_t stream)
{
    switch (g.tod_type) {
    case TOD_F32: tod2map_typed(lmap, static_cast<const float *>(tod), xp, a, plan, g, stream); break;
    case TOD_F16: tod2map_typed(lmap, static_cast<const _Float16 *>(tod), xp, a, plan, g, stream); break;
    case TOD_BF16: tod2map_typed(lmap, static_cast<const __bf16 *>(tod), xp, a, plan, g, stream); break;
    default: fail("tod2map: unknown TOD element type");
    }
}

void launch_batched_map2tod(const BatchLaunch &bl, const float *lmap, const gmm_lpix &lp, int tod_type, bool partial,
                            bool accum, hipStream_t stream)
{
    if (bl.nitems <= 0)
        return;
    GMM_REQUIRE(bl.threads == 256 || bl.threads == 512, "batched map2tod: threads per block must be 256 or 512");
    GMM_REQUIRE(!bl.has_resp || tod_type == TOD_F32, "batched map2tod: per-detector gains need float32 TODs");
    const BatchArgs ba = batch_args(bl, lp, partial, accum);
    switch (tod_type) {
    case TOD_F32:
        if (bl.has_resp)
            batched_m2t_resp(bl, lmap, lp.offsets_gpu, ba, lp.periodic_xcoord, stream);
        else
            batched_m2t_typed<float>(bl, lmap, lp.offsets_gpu, ba, lp.periodic_xcoord, stream);
        break;
    case TOD_F16: batched_m2t_typed<_Float16>(bl, lmap, lp.offsets_gpu, ba, lp.periodic_xcoord, stream); break;
    case TOD_BF16: batched_m2t_typed<__bf16>(bl, lmap, lp.offsets_gpu, ba, lp.periodic_xcoord, stream); break;
    default: fail("batched map2tod: unknown TOD element type");
    }
}

void launch_batched_tod2map(const BatchLaunch &bl, float *lmap, const gmm_lpix &lp, int tod_type, bool partial,
                            hipStream_t stream)
{
    if (bl.nitems <= 0)
        return;
    const BatchArgs ba = batch_args(bl, lp, partial, false);
    GMM_REQUIRE(!bl.has_resp || tod_type == TOD_F32, "batched tod2map: per-detector gains need float32 TODs");
    switch (tod_type) {
    case TOD_F32:
        if (bl.has_resp)
            batched_t2m_resp(bl, lmap, lp.offsets_gpu, ba, lp.periodic_xcoord, stream);
        else
            batched_t2m_typed<float>(bl, lmap, lp.offsets_gpu, ba, lp.periodic_xcoord, stream);
        break;
    case TOD_F16: batched_t2m_typed<_Float16>(bl, lmap, lp.offsets_gpu, ba, lp.periodic_xcoord, stream); break;
    case TOD_BF16: batched_t2m_typed<__bf16>(bl, lmap, lp.offsets_gpu, ba, lp.periodic_xcoord, stream); break;
    default: fail("batched tod2map: unknown TOD element type");
    }
}

}  // namespace gmm
