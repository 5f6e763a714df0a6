// Planned map2tod / tod2map on gfx950.
//
// One block per work unit ("chunk": a run of plan entries of one cell, cut at plan time). The
// block stages the cell's 3x64x64 T/Q/U values in LDS (map2tod) or accumulates them there
// (tod2map). Each wave takes a contiguous slice of the chunk's entries and caches 64 plan
// entries per vector load. A wave step covers 8 entries (8 TOD lines of 32 samples): lane
// 8*li + q owns samples 4q..4q+3 of line li, loaded as float4, and two such steps are loaded
// before either is used.
//
// map2tod keeps a zero-padded 66x66 tile per plane so that corners outside the cell read 0
// without a branch; lines touching several cells are zeroed first and then accumulated with
// float atomics (reference: src_lib/map2tod.cu:15-62, 81-191).
//
// tod2map accumulates in LDS with 64-bit integer atomics on fixed-point values: float LDS
// atomics (ds_add_f32) run an order of magnitude slower than integer ones on this GPU. A pre-pass
// finds max|tod|; with e = its binary exponent, contributions are scaled by S = 2^(27-e). A lane
// sums its consecutive samples that share the same (iy0, ix0) in float registers (at most 4, so
// |sum| < 2^30), converts the run's 12 partials to int32 and adds them into an int64 tile; the
// flush converts each pixel once, float(sum)/S. Pixel sums are exact in fixed point, so the
// result is more accurate than float32 accumulation. Registers take the runs of equal pixels
// that would otherwise make lanes of one atomic instruction collide on the same LDS address.
// If the TOD holds Inf/NaN the kernel falls back to float LDS atomics.
// Reference: src_lib/tod2map.cu:33-154.
//
// Both kernels also take per-detector gains, float response[2][ndet] (row 0 scales T, row 1
// scales Q and U; src_lib/map2tod_response.cu, tod2map_response.cu), and 16-bit TODs (FP16 or
// BF16) stored as value/scale[row] with one FP32 scale per TOD row; arithmetic stays FP32.

#include "gmm_internal.hpp"

namespace gmm {

namespace {

constexpr int WAVE = 64;
constexpr int TP = 66;              // padded tile side (map2tod)
constexpr int TPLANE = TP * TP;     // floats per padded plane

constexpr int M2T_THREADS = 512;
constexpr int M2T_U = 2;            // 8-line groups in flight per wave
constexpr int T2M_THREADS = 1024;
constexpr int T2M_U = 2;            // 8-line groups in flight per wave

constexpr int ABSMAX_BLOCKS = 1024;
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

// Slice [wb, we) of the chunk's entries handled by wave 'wave' of 'nw'.
__device__ inline void wave_slice(const Chunk &ch, int wave, int nw, uint32_t &wb, uint32_t &we)
{
    const uint32_t n = ch.end - ch.begin;
    wb = ch.begin + uint32_t((uint64_t(n) * wave) / nw);
    we = ch.begin + uint32_t((uint64_t(n) * (wave + 1)) / nw);
}

// Plan entry packed for broadcast: line index in bits [0,26), mflag in bit 31.
__device__ inline uint32_t pack_entry(uint64_t mt)
{
    return (uint32_t(mt >> ICL_SHIFT) & ICL_MASK) | ((mt & MFLAG) ? 0x80000000u : 0u);
}

// ------------------------------------------------------------------------------------------
// Lane layout of both kernels: a wave step covers 8 plan entries (lines); lane = 8*li + q takes
// samples 4q..4q+3 of line li, loaded as float4 (the arrays are 16-byte aligned in the Vec
// variant; 32-sample lines make every 4-sample group aligned).

template<bool Vec>
__device__ inline float4 load4(const float *__restrict__ p)
{
    if (Vec)
        return *reinterpret_cast<const float4 *>(p);
    return make_float4(p[0], p[1], p[2], p[3]);
}

__device__ inline float comp(const float4 &v, int k) { return (k == 0) ? v.x : (k == 1) ? v.y : (k == 2) ? v.z : v.w; }

// TOD element access, four consecutive samples at a time. TT is float, _Float16 or __bf16.
// 16-bit TODs are always accessed with 8-byte vector operations (the host requires an 8-byte
// aligned TOD), and multi-cell lines use the packed 16-bit global atomics.
template<typename TT, bool Vec>
struct TodIO
{
    typedef TT v4 __attribute__((ext_vector_type(4)));
    typedef short s2 __attribute__((ext_vector_type(2)));
    typedef TT t2 __attribute__((ext_vector_type(2)));

    __device__ static float4 load(const TT *__restrict__ p)
    {
        if constexpr (std::is_same<TT, float>::value)
            return load4<Vec>(p);
        else {
            const v4 v = *reinterpret_cast<const v4 *>(p);
            return make_float4(float(v.x), float(v.y), float(v.z), float(v.w));
        }
    }

    __device__ static void store(TT *__restrict__ p, float4 v)
    {
        if constexpr (std::is_same<TT, float>::value) {
            if (Vec)
                *reinterpret_cast<float4 *>(p) = v;
            else {
                p[0] = v.x; p[1] = v.y; p[2] = v.z; p[3] = v.w;
            }
        }
        else {
            const v4 h = { TT(v.x), TT(v.y), TT(v.z), TT(v.w) };
            *reinterpret_cast<v4 *>(p) = h;
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
// multiplies T, row 1 multiplies Q and U; sample s belongs to detector s / resp_nper. A 16-bit
// TOD holds value / scale[s / scale_nper].
struct Gains
{
    const float *__restrict__ resp;
    const float *__restrict__ scale;
    uint32_t resp_ndet, resp_nper, scale_nper;
};

// Row of each of the 4 samples s0..s0+3 for rows of 'nper' samples (one division per quad when
// a quad can cross at most one row boundary).
__device__ inline void quad_rows(uint32_t s0, uint32_t nper, uint32_t (&r)[4])
{
    if (nper >= 4) {
        const uint32_t r0 = s0 / nper;
        const uint32_t b = (r0 + 1) * nper - s0;    // samples of the quad in row r0
#pragma unroll
        for (int k = 0; k < 4; k++)
            r[k] = r0 + ((uint32_t(k) >= b) ? 1u : 0u);
    }
    else {
#pragma unroll
        for (int k = 0; k < 4; k++)
            r[k] = (s0 + uint32_t(k)) / nper;
    }
}

// T and P gain of the 4 samples of a quad (1 without a response array).
template<bool HasResp>
__device__ inline void quad_resp(const Gains &g, uint32_t s0, float (&tr)[4], float (&pr)[4])
{
    if (HasResp) {
        uint32_t d[4];
        quad_rows(s0, g.resp_nper, d);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            tr[k] = g.resp[d[k]];
            pr[k] = g.resp[g.resp_ndet + d[k]];
        }
    }
    else {
#pragma unroll
        for (int k = 0; k < 4; k++)
            tr[k] = pr[k] = 1.0f;
    }
}

// Scale of the 4 samples of a quad (16-bit TODs only).
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
        quad_rows(s0, g.scale_nper, r);
#pragma unroll
        for (int k = 0; k < 4; k++)
            sc[k] = g.scale[r[k]];
    }
}

// Bilinear T + cos2a Q + sin2a U of one sample from the padded tile (corners outside the cell
// read the zero border).
template<bool Periodic, bool HasResp>
__device__ inline float m2t_sample(const float *__restrict__ tile, float y, float x, float alpha,
                                   int ny, int nx, int ybase, int xbase, float tr, float pr, uint32_t &err)
{
    float sn, cs;
    sincos_angle(2.0f * alpha, sn, cs);
    if (HasResp) {
        cs *= pr;
        sn *= pr;
    }
    Pixel<float> p;
    err |= locate_pixel(y, x, ny, nx, Periodic, p);

    const int r0 = clamp_i(p.iy0 - ybase + 1, 0, TP - 1) * TP;
    const int r1 = clamp_i(p.iy1 - ybase + 1, 0, TP - 1) * TP;
    const int c0 = clamp_i(p.ix0 - xbase + 1, 0, TP - 1);
    const int c1 = clamp_i(p.ix1 - xbase + 1, 0, TP - 1);
    const float *tq = tile + TPLANE, *tu = tile + 2 * TPLANE;

    const float t00 = HasResp ? tr * tile[r0 + c0] : tile[r0 + c0];
    const float t01 = HasResp ? tr * tile[r0 + c1] : tile[r0 + c1];
    const float t10 = HasResp ? tr * tile[r1 + c0] : tile[r1 + c0];
    const float t11 = HasResp ? tr * tile[r1 + c1] : tile[r1 + c1];
    const float v00 = t00 + cs * tq[r0 + c0] + sn * tu[r0 + c0];
    const float v01 = t01 + cs * tq[r0 + c1] + sn * tu[r0 + c1];
    const float v10 = t10 + cs * tq[r1 + c0] + sn * tu[r1 + c0];
    const float v11 = t11 + cs * tq[r1 + c1] + sn * tu[r1 + c1];

    const float wy0 = 1.0f - p.dy, wx0 = 1.0f - p.dx;
    return (wy0 * wx0) * v00 + (wy0 * p.dx) * v01 + (p.dy * wx0) * v10 + (p.dy * p.dx) * v11;
}

template<bool Periodic, bool Vec, bool HasResp, typename TT>
__global__ void __launch_bounds__(M2T_THREADS, 6)
planned_map2tod_kernel(TT *__restrict__ tod, const float *__restrict__ lmap, const float *__restrict__ xp,
                       const int64_t *__restrict__ cell_offsets, const uint64_t *__restrict__ plan_mt,
                       const Chunk *__restrict__ chunks, uint32_t *__restrict__ err_word,
                       int64_t nsamp, int ny, int nx, int nycells, int nxcells,
                       int64_t ystride, int64_t polstride, bool partial, bool accum, Gains gn)
{
    __shared__ float tile[3 * TPLANE];

    const Chunk ch = chunks[blockIdx.x];
    const int64_t off = chunk_cell_offset(ch, cell_offsets, nycells, nxcells);
    const uint32_t ncl = uint32_t(nsamp >> 5);
    uint32_t err = (off < 0 && !partial) ? ERR_NOT_IN_PIXELIZATION : 0u;

    if (off >= 0) {
        // Interior: 3 planes x 64 rows x 64 columns, one row of 64 floats per wave load.
        for (int i = threadIdx.x; i < 3 * CELL_NPIX; i += M2T_THREADS) {
            int p = i >> 12, r = (i >> 6) & 63, c = i & 63;
            tile[p * TPLANE + (r + 1) * TP + (c + 1)] = lmap[off + p * polstride + r * ystride + c];
        }
        // Zero border: rows 0 and 65 (66 entries each), columns 0 and 65 of rows 1..64.
        for (int i = threadIdx.x; i < 3 * 260; i += M2T_THREADS) {
            int p = i / 260, j = i % 260, idx;
            if (j < 66)
                idx = j;
            else if (j < 132)
                idx = 65 * TP + (j - 66);
            else if (j < 196)
                idx = (j - 132 + 1) * TP;
            else
                idx = (j - 196 + 1) * TP + 65;
            tile[p * TPLANE + idx] = 0.0f;
        }
    }
    __syncthreads();

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane >> 3, q = lane & 7;
    const int ybase = int(ch.icell >> 10) << 6, xbase = int(ch.icell & 1023u) << 6;

    uint32_t wb, we;
    wave_slice(ch, wave, M2T_THREADS / WAVE, wb, we);

    for (uint32_t base = wb; base < we; base += 64) {
        const uint32_t cnt = min(64u, we - base);
        const uint32_t packed = (uint32_t(lane) < cnt) ? pack_entry(plan_mt[base + lane]) : 0u;

        // Variants with gains keep one line group in flight (their extra loads would spill).
        constexpr int UU = (HasResp || !std::is_same<TT, float>::value) ? 1 : M2T_U;
        using IO = TodIO<TT, Vec>;
        for (uint32_t j = 0; j < cnt; j += 8 * UU) {
            float4 y4[UU], x4[UU], a4[UU];
            int64_t s0[UU];
            bool act[UU], mfl[UU];
#pragma unroll
            for (int u = 0; u < UU; u++) {
                const uint32_t idx = j + 8 * u + li;
                const uint32_t e = __shfl(packed, int(min(idx, 63u)), WAVE);
                const uint32_t icl = e & ICL_MASK;
                act[u] = (idx < cnt) && (icl < ncl);
                if (idx < cnt && icl >= ncl)
                    err |= ERR_INCONSISTENT_NMT;
                mfl[u] = (e >> 31) != 0;
                s0[u] = int64_t(act[u] ? icl : 0u) * LINE + 4 * q;
                y4[u] = load4<Vec>(xp + s0[u]);
                x4[u] = load4<Vec>(xp + nsamp + s0[u]);
                a4[u] = load4<Vec>(xp + 2 * nsamp + s0[u]);
            }
#pragma unroll
            for (int u = 0; u < UU; u++) {
                if (!act[u])
                    continue;
                if (off < 0) {
                    if (!accum && !mfl[u])
                        IO::store(tod + s0[u], make_float4(0.f, 0.f, 0.f, 0.f));
                    continue;
                }
                float tr[4], pr[4], sc[4], vv[4];
                quad_resp<HasResp>(gn, uint32_t(s0[u]), tr, pr);
                quad_scale<TT>(gn, uint32_t(s0[u]), sc);
#pragma unroll 2
                for (int k = 0; k < 4; k++) {
                    vv[k] = m2t_sample<Periodic, HasResp>(tile, comp(y4[u], k), comp(x4[u], k), comp(a4[u], k),
                                                          ny, nx, ybase, xbase, tr[k], pr[k], err);
                    if constexpr (!std::is_same<TT, float>::value)
                        vv[k] = vv[k] / sc[k];
                }
                const float4 v = make_float4(vv[0], vv[1], vv[2], vv[3]);
                TT *t = tod + s0[u];
                if (mfl[u])
                    IO::atomic_add(t, v);
                else if (accum) {
                    const float4 o = IO::load(t);
                    IO::store(t, make_float4(o.x + v.x, o.y + v.y, o.z + v.z, o.w + v.w));
                }
                else
                    IO::store(t, v);
            }
        }
    }

    report_err(err, err_word);
}

// Zeroes the TOD lines that touch more than one cell (they are accumulated atomically).
template<typename TT>
__global__ void zero_lines_kernel(TT *__restrict__ tod, const uint32_t *__restrict__ mlines, int64_t nml)
{
    const int64_t n = nml * LINE;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        tod[int64_t(mlines[i >> 5]) * LINE + (i & 31)] = TT(0.0f);
}

// ------------------------------------------------------------------------------------------

// Per-block maxima, as float bits (NaN > Inf > finite in this order), over ABSMAX_BLOCKS
// blocks: partial[b] = max|tod| over block b's share of the samples, partial[B + b] = max|response|
// and partial[2B + b] = max|scale| over its share of those arrays (0 where absent).
template<typename TT, bool Vec>
__global__ void __launch_bounds__(256)
tod_absmax_kernel(const TT *__restrict__ tod, int64_t n, uint32_t *__restrict__ partial,
                  const float *__restrict__ resp, int64_t nresp, const float *__restrict__ scale, int64_t nscale)
{
    __shared__ uint32_t s_w[3][4];
    const int64_t stride = int64_t(gridDim.x) * 256, i0 = int64_t(blockIdx.x) * 256 + threadIdx.x;
    uint32_t m = 0, mr = 0, ms = 0;
    if (Vec || !std::is_same<TT, float>::value) {
        for (int64_t i = i0; i < (n >> 2); i += stride) {
            const float4 v = TodIO<TT, true>::load(tod + 4 * i);
            m = max(m, __float_as_uint(v.x) & 0x7fffffffu);
            m = max(m, __float_as_uint(v.y) & 0x7fffffffu);
            m = max(m, __float_as_uint(v.z) & 0x7fffffffu);
            m = max(m, __float_as_uint(v.w) & 0x7fffffffu);
        }
    }
    else {
        for (int64_t i = i0; i < n; i += stride)
            m = max(m, __float_as_uint(float(tod[i])) & 0x7fffffffu);
    }
    for (int64_t i = i0; i < nresp; i += stride)
        mr = max(mr, __float_as_uint(resp[i]) & 0x7fffffffu);
    for (int64_t i = i0; i < nscale; i += stride)
        ms = max(ms, __float_as_uint(scale[i]) & 0x7fffffffu);
    m = wave_max(m);
    mr = wave_max(mr);
    ms = wave_max(ms);
    if ((threadIdx.x & 63) == 0) {
        s_w[0][threadIdx.x >> 6] = m;
        s_w[1][threadIdx.x >> 6] = mr;
        s_w[2][threadIdx.x >> 6] = ms;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const uint32_t *w = s_w[threadIdx.x];
        partial[threadIdx.x * gridDim.x + blockIdx.x] = max(max(w[0], w[1]), max(w[2], w[3]));
    }
}

__device__ inline void acc_add3(unsigned long long *__restrict__ acc, int k, float a, float b, float c)
{
    atomicAdd(acc + k, (unsigned long long)(long long)int(a));
    atomicAdd(acc + CELL_NPIX + k, (unsigned long long)(long long)int(b));
    atomicAdd(acc + 2 * CELL_NPIX + k, (unsigned long long)(long long)int(c));
}

// Adds one run's register partials (12 values: corners 00, 01, 10, 11 x T, Q, U; already
// scaled) for the corners that lie inside the cell.
__device__ inline void flush_run(unsigned long long *__restrict__ acc, int ry0, int ry1, int rx0, int rx1,
                                 const float (&a)[12])
{
    if (unsigned(ry0) < 64u) {
        if (unsigned(rx0) < 64u)
            acc_add3(acc, (ry0 << 6) | rx0, a[0], a[1], a[2]);
        if (unsigned(rx1) < 64u)
            acc_add3(acc, (ry0 << 6) | rx1, a[3], a[4], a[5]);
    }
    if (unsigned(ry1) < 64u) {
        if (unsigned(rx0) < 64u)
            acc_add3(acc, (ry1 << 6) | rx0, a[6], a[7], a[8]);
        if (unsigned(rx1) < 64u)
            acc_add3(acc, (ry1 << 6) | rx1, a[9], a[10], a[11]);
    }
}

template<bool Periodic, bool Vec, bool HasResp, typename TT>
__global__ void __launch_bounds__(T2M_THREADS)
planned_tod2map_kernel(float *__restrict__ lmap, const TT *__restrict__ tod, const float *__restrict__ xp,
                       const int64_t *__restrict__ cell_offsets, const uint64_t *__restrict__ plan_mt,
                       const Chunk *__restrict__ chunks, const uint32_t *__restrict__ absmax_partial,
                       uint32_t *__restrict__ err_word, int64_t nsamp, int ny, int nx, int nycells, int nxcells,
                       int64_t ystride, int64_t polstride, bool partial, Gains gn)
{
    constexpr bool Q16 = !std::is_same<TT, float>::value;
    using IO = TodIO<TT, Vec>;
    __shared__ unsigned long long acc[3 * CELL_NPIX];    // 96 KiB
    __shared__ uint32_t s_red[3][T2M_THREADS / WAVE];
    __shared__ uint32_t s_nonfinite;
    float *facc = reinterpret_cast<float *>(acc);         // float tile for the non-finite fallback

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const Chunk ch = chunks[blockIdx.x];
    const int64_t off = chunk_cell_offset(ch, cell_offsets, nycells, nxcells);
    const uint32_t ncl = uint32_t(nsamp >> 5);
    uint32_t err = (off < 0 && !partial) ? ERR_NOT_IN_PIXELIZATION : 0u;

    if (off < 0) {
        // Samples whose cells are not in the local map are dropped (tod2map_reference.cu:14-46
        // skips them too); only the error bit is reported.
        report_err(err, err_word);
        return;
    }

    // Scale from the bound B = (max|t| * max|scale|) * max|response|, evaluated in the order the
    // samples are (t * scale, then * response), so that by monotonic rounding every deposited
    // |t| <= B. ABSMAX_BLOCKS == T2M_THREADS partials of each kind, one per thread. A lane sums
    // up to 4 samples in registers before converting, so |w*t*S| < 2^30 keeps the int32 safe.
    // A 16-bit TOD needs no pass over the samples: every finite value is at most the type's
    // largest finite value, which stands in for max|t|; non-finite values are detected while
    // accumulating and send the block through a second, floating-point pass.
    if (threadIdx.x == 0)
        s_nonfinite = 0;
    {
        const uint32_t mt = wave_max(absmax_partial[threadIdx.x]);
        const uint32_t mr = HasResp ? wave_max(absmax_partial[ABSMAX_BLOCKS + threadIdx.x]) : 0u;
        const uint32_t ms = Q16 ? wave_max(absmax_partial[2 * ABSMAX_BLOCKS + threadIdx.x]) : 0u;
        if (lane == 0) {
            s_red[0][wave] = mt;
            s_red[1][wave] = mr;
            s_red[2][wave] = ms;
        }
    }
    for (int i = threadIdx.x; i < 3 * CELL_NPIX; i += T2M_THREADS)
        acc[i] = 0ull;
    __syncthreads();
    uint32_t m = 0, mr = 0, ms = 0;
#pragma unroll
    for (int w = 0; w < T2M_THREADS / WAVE; w++) {
        m = max(m, s_red[0][w]);
        mr = max(mr, s_red[1][w]);
        ms = max(ms, s_red[2][w]);
    }
    if (Q16)
        m = std::is_same<TT, _Float16>::value ? 0x477fe000u : 0x7f7f0000u;   // 65504, 0x1.fep127
    bool finite = (m < FINITE_LIMIT);
    if (Q16) {
        finite = finite && (ms < FINITE_LIMIT);
        m = __float_as_uint(__uint_as_float(m) * __uint_as_float(ms));
    }
    if (HasResp) {
        finite = finite && (mr < FINITE_LIMIT);
        m = __float_as_uint(__uint_as_float(m) * __uint_as_float(mr));
    }
    finite = finite && (m < FINITE_LIMIT);
    const int e = int(m >> 23) - 127;                    // binary exponent of max|t|
    const int sexp = clamp_i(27 - e, -100, 126);
    const float scale = __uint_as_float(uint32_t(sexp + 127) << 23);
    const float inv_scale = __uint_as_float(uint32_t(127 - sexp) << 23);

    const int li = lane >> 3, q = lane & 7;
    const int ybase = int(ch.icell >> 10) << 6, xbase = int(ch.icell & 1023u) << 6;

    uint32_t wb, we;
    wave_slice(ch, wave, T2M_THREADS / WAVE, wb, we);

    for (;;) {
        bool nonfinite = false;
        for (uint32_t base = wb; base < we; base += 64) {
            const uint32_t cnt = min(64u, we - base);
            const uint32_t packed = (uint32_t(lane) < cnt) ? pack_entry(plan_mt[base + lane]) : 0u;

            for (uint32_t j = 0; j < cnt; j += 8 * T2M_U) {
                float4 y4[T2M_U], x4[T2M_U], a4[T2M_U], t4[T2M_U];
                int64_t s0[T2M_U];
                bool act[T2M_U];
    #pragma unroll
                for (int u = 0; u < T2M_U; u++) {
                    const uint32_t idx = j + 8 * u + li;
                    const uint32_t e2 = __shfl(packed, int(min(idx, 63u)), WAVE);
                    const uint32_t icl = e2 & ICL_MASK;
                    act[u] = (idx < cnt) && (icl < ncl);
                    if (idx < cnt && icl >= ncl)
                        err |= ERR_INCONSISTENT_NMT;
                    s0[u] = int64_t(act[u] ? icl : 0u) * LINE + 4 * q;
                    y4[u] = load4<Vec>(xp + s0[u]);
                    x4[u] = load4<Vec>(xp + nsamp + s0[u]);
                    a4[u] = load4<Vec>(xp + 2 * nsamp + s0[u]);
                    t4[u] = IO::load(tod + s0[u]);
                }
    #pragma unroll
                for (int u = 0; u < T2M_U; u++) {
                    if (!act[u])
                        continue;
                    float tr[4], pr[4], sc[4];
                    quad_resp<HasResp>(gn, uint32_t(s0[u]), tr, pr);
                    quad_scale<TT>(gn, uint32_t(s0[u]), sc);
                    if (Q16) {
                        t4[u].x *= sc[0];
                        t4[u].y *= sc[1];
                        t4[u].z *= sc[2];
                        t4[u].w *= sc[3];
                        nonfinite |= !(isfinite(t4[u].x) && isfinite(t4[u].y) && isfinite(t4[u].z) &&
                                       isfinite(t4[u].w));
                    }
                    if (finite) {
                        // Runs of samples with the same (iy0, ix0) are summed in registers; the LDS
                        // atomics are issued once per run.
                        float a[12];
    #pragma unroll
                        for (int i = 0; i < 12; i++)
                            a[i] = 0.0f;
                        uint32_t cur = 0xffffffffu;
                        int ry0 = 0, ry1 = 0, rx0 = 0, rx1 = 0;
    #pragma unroll
                        for (int k = 0; k < 4; k++) {
                            float sn, cs;
                            sincos_angle(2.0f * comp(a4[u], k), sn, cs);
                            Pixel<float> p;
                            err |= locate_pixel(comp(y4[u], k), comp(x4[u], k), ny, nx, Periodic, p);
                            const uint32_t key = (uint32_t(p.iy0) << 16) | uint32_t(p.ix0);
                            if (key != cur) {
                                if (k > 0) {
                                    flush_run(acc, ry0, ry1, rx0, rx1, a);
    #pragma unroll
                                    for (int i = 0; i < 12; i++)
                                        a[i] = 0.0f;
                                }
                                cur = key;
                                ry0 = p.iy0 - ybase;
                                ry1 = p.iy1 - ybase;
                                rx0 = p.ix0 - xbase;
                                rx1 = p.ix1 - xbase;
                            }
                            const float t0 = comp(t4[u], k) * scale;
                            const float ts = HasResp ? t0 * tr[k] : t0;
                            const float ps = HasResp ? t0 * pr[k] : t0;
                            const float qs = ps * cs, us = ps * sn;
                            const float wy0 = 1.0f - p.dy, wx0 = 1.0f - p.dx;
                            const float w[4] = { wy0 * wx0, wy0 * p.dx, p.dy * wx0, p.dy * p.dx };
    #pragma unroll
                            for (int c = 0; c < 4; c++) {
                                a[3 * c] += w[c] * ts;
                                a[3 * c + 1] += w[c] * qs;
                                a[3 * c + 2] += w[c] * us;
                            }
                        }
                        flush_run(acc, ry0, ry1, rx0, rx1, a);
                    }
                    else {
    #pragma unroll
                        for (int k = 0; k < 4; k++) {
                            float sn, cs;
                            sincos_angle(2.0f * comp(a4[u], k), sn, cs);
                            Pixel<float> p;
                            err |= locate_pixel(comp(y4[u], k), comp(x4[u], k), ny, nx, Periodic, p);
                            const float t = comp(t4[u], k) * tr[k], qv = comp(t4[u], k) * pr[k] * cs,
                                        uv = comp(t4[u], k) * pr[k] * sn;
                            const int ry[2] = { p.iy0 - ybase, p.iy1 - ybase };
                            const int rx[2] = { p.ix0 - xbase, p.ix1 - xbase };
                            const float wy[2] = { 1.0f - p.dy, p.dy };
                            const float wx[2] = { 1.0f - p.dx, p.dx };
                            for (int aa = 0; aa < 2; aa++)
                                for (int bb = 0; bb < 2; bb++)
                                    if ((unsigned(ry[aa]) < 64u) && (unsigned(rx[bb]) < 64u)) {
                                        const float w = wy[aa] * wx[bb];
                                        const int kk = (ry[aa] << 6) | rx[bb];
                                        atomicAdd(facc + kk, w * t);
                                        atomicAdd(facc + CELL_NPIX + kk, w * qv);
                                        atomicAdd(facc + 2 * CELL_NPIX + kk, w * uv);
                                    }
                        }
                    }
                }
            }
        }
        if (!Q16 || !finite)
            break;
        // 16-bit TOD: restart the chunk in floating point if a non-finite value was seen.
        if (nonfinite)
            atomicOr(&s_nonfinite, 1u);
        __syncthreads();
        if (s_nonfinite == 0)
            break;
        for (int i = threadIdx.x; i < 3 * CELL_NPIX; i += T2M_THREADS)
            facc[i] = 0.0f;
        finite = false;
        __syncthreads();
    }
    __syncthreads();

    const bool shared = (ch.flags & CHUNK_SHARED) != 0;
    for (int i = threadIdx.x; i < 3 * CELL_NPIX; i += T2M_THREADS) {
        const int pl = i >> 12, r = (i >> 6) & 63, c = i & 63;
        float *g = lmap + off + pl * polstride + r * ystride + c;
        const float v = finite ? float((long long)acc[i]) * inv_scale : facc[i];
        if (shared)
            unsafeAtomicAdd(g, v);
        else
            *g += v;
    }

    report_err(err, err_word);
}

template<bool HasResp, typename TT>
void launch_m2t(TT *tod, const float *lmap, const float *xp, const OpArgs &a, const gmm_plan &plan,
                bool accum, const Gains &gn, hipStream_t stream)
{
    const gmm_lpix &lp = *a.lpix;
    uint32_t *err_word = plan.err + (plan.pp.n_err - 1);
    dim3 grid(unsigned(plan.nchunks)), block(M2T_THREADS);
    const bool vec = ((reinterpret_cast<uintptr_t>(tod) | reinterpret_cast<uintptr_t>(xp)) & 15) == 0;
    auto k = lp.periodic_xcoord ? (vec ? planned_map2tod_kernel<true, true, HasResp, TT> : planned_map2tod_kernel<true, false, HasResp, TT>)
                                : (vec ? planned_map2tod_kernel<false, true, HasResp, TT> : planned_map2tod_kernel<false, false, HasResp, TT>);
    kernel_timer_mark(TK_MAP2TOD, false, stream);
    k<<<grid, block, 0, stream>>>(tod, lmap, xp, lp.offsets_gpu, plan.plan_mt, plan.chunks, err_word,
                                  a.nsamp, int(lp.nypix_global), int(lp.nxpix_global), int(lp.nycells),
                                  int(lp.nxcells), lp.ystride, lp.polstride, a.partial, accum, gn);
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
    const bool vec = ((reinterpret_cast<uintptr_t>(tod) | reinterpret_cast<uintptr_t>(xp)) & 15) == 0;
    auto k = lp.periodic_xcoord ? (vec ? planned_tod2map_kernel<true, true, HasResp, TT> : planned_tod2map_kernel<true, false, HasResp, TT>)
                                : (vec ? planned_tod2map_kernel<false, true, HasResp, TT> : planned_tod2map_kernel<false, false, HasResp, TT>);
    kernel_timer_mark(TK_TOD2MAP, false, stream);
    k<<<grid, block, 0, stream>>>(lmap, tod, xp, lp.offsets_gpu, plan.plan_mt, plan.chunks, plan.absmax, err_word,
                                  a.nsamp, int(lp.nypix_global), int(lp.nxpix_global), int(lp.nycells),
                                  int(lp.nxcells), lp.ystride, lp.polstride, a.partial, gn);
    GMM_HIP(hipGetLastError());
    kernel_timer_mark(TK_TOD2MAP, true, stream);
}

Gains make_gains(const OpArgs &a, const TodGains &g)
{
    Gains gn{g.response, g.scale, 1u, 1u, 1u};
    if (g.response) {
        gn.resp_ndet = uint32_t(g.resp_ndet);
        gn.resp_nper = uint32_t(a.nsamp / g.resp_ndet);
    }
    if (g.scale)
        gn.scale_nper = uint32_t(a.nsamp / g.scale_nrows);
    return gn;
}

template<typename TT>
void map2tod_typed(TT *tod, const float *lmap, const float *xp, const OpArgs &a, const gmm_plan &plan,
                   bool accum, const TodGains &g, hipStream_t stream)
{
    if (!accum && plan.nmlines > 0) {
        int64_t n = plan.nmlines * LINE;
        int grid = int(std::min<int64_t>((n + 255) / 256, 8192));
        zero_lines_kernel<TT><<<grid, 256, 0, stream>>>(tod, plan.mlines, plan.nmlines);
        GMM_HIP(hipGetLastError());
    }
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
    static_assert(ABSMAX_BLOCKS == T2M_THREADS, "one max|t| partial per thread of the tod2map block");
    const int64_t nresp = g.response ? 2 * g.resp_ndet : 0;
    const int64_t nscale = g.scale ? g.scale_nrows : 0;
    // A 16-bit TOD is bounded by its type (the kernel substitutes the largest finite value),
    // so only the gain arrays are scanned.
    const int64_t nt = std::is_same<TT, float>::value ? a.nsamp : 0;
    if (std::is_same<TT, float>::value && (reinterpret_cast<uintptr_t>(tod) & 15) != 0)
        tod_absmax_kernel<TT, false><<<ABSMAX_BLOCKS, 256, 0, stream>>>(tod, nt, plan.absmax, g.response, nresp,
                                                                       g.scale, nscale);
    else
        tod_absmax_kernel<TT, true><<<ABSMAX_BLOCKS, 256, 0, stream>>>(tod, nt, plan.absmax, g.response, nresp,
                                                                      g.scale, nscale);
    GMM_HIP(hipGetLastError());

    const Gains gn = make_gains(a, g);
    if (g.response)
        launch_t2m<true, TT>(lmap, tod, xp, a, plan, gn, stream);
    else
        launch_t2m<false, TT>(lmap, tod, xp, a, plan, gn, stream);
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

int64_t tod2map_scratch_words() { return 3 * ABSMAX_BLOCKS; }

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

}  // namespace gmm
