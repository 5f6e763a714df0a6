// 16-bit TOD encoding on gfx950: one FP32 scale per TOD row, values stored as FP16 or BF16.
//
// scale[r] = max(max_s |tod[r][s]| / (0.5 * M), FLT_MIN), M = largest finite value of the type
// (65504 for FP16, 0x1.fep127 for BF16), so every encoded value lies in [-M/2, M/2]; the FLT_MIN
// floor keeps the scale a normal float (an all-zero row gets scale FLT_MIN and zeros).
// One block per row: a max-reduction pass, then a pass that divides and rounds to nearest even.
//
// Optional error feedback (order 1 or 2): each thread encodes a contiguous run of its row in
// sample order and adds the previous samples' rounding errors e to the next value before rounding
// (v_k = u_k + e_{k-1}, order 1; v_k = u_k + 2 e_{k-1} - e_{k-2}, order 2; e_k = v_k - round(v_k)).
// The decoded TOD then carries the first (second) difference of the rounding errors instead of
// the errors themselves; tod2map sums samples with slowly varying bilinear weights, so in the map
// the differences largely cancel (FP16, 2048x262144 raster: map error 2.1e-4 -> 9.9e-5 / 8.3e-5
// of the map RMS, while the per-sample TOD error grows by ~1.5x / 2.7x). The chain restarts at
// each thread's run (row_len / 256 samples); a non-finite error is dropped.
// This has no counterpart in the reference, which stores TODs as float32 only.

#include "gmm_internal.hpp"

#include <cfloat>
#include <type_traits>

namespace gmm {

namespace {

constexpr int ENC_THREADS = 256;

template<typename TT, int Order>
__global__ void __launch_bounds__(ENC_THREADS)
tod16_encode_kernel(TT *__restrict__ out, float *__restrict__ scale, const float *__restrict__ tod, int64_t row_len)
{
    __shared__ float s_max[ENC_THREADS / 64];
    const float *src = tod + int64_t(blockIdx.x) * row_len;
    TT *dst = out + int64_t(blockIdx.x) * row_len;

    // fmaxf ignores NaN, so NaN values leave the scale alone and are encoded as NaN. A row that
    // holds an infinity gets an infinite scale (its finite values encode as 0, infinities as NaN).
    float m = 0.0f;
    for (int64_t i = threadIdx.x; i < row_len; i += ENC_THREADS)
        m = fmaxf(m, fabsf(src[i]));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        m = fmaxf(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0)
        s_max[threadIdx.x >> 6] = m;
    __syncthreads();
    m = fmaxf(fmaxf(s_max[0], s_max[1]), fmaxf(s_max[2], s_max[3]));

    constexpr float half_max = std::is_same<TT, _Float16>::value ? 32752.0f : 0x1.fep126f;
    const float sc = fmaxf(m / half_max, FLT_MIN);
    if (threadIdx.x == 0)
        scale[blockIdx.x] = sc;
    if (Order == 0) {
        for (int64_t i = threadIdx.x; i < row_len; i += ENC_THREADS)
            dst[i] = TT(src[i] / sc);
    }
    else {
        const int64_t run = (row_len + ENC_THREADS - 1) / ENC_THREADS;
        const int64_t i0 = int64_t(threadIdx.x) * run, i1 = min(row_len, i0 + run);
        float e1 = 0.0f, e2 = 0.0f;
        for (int64_t i = i0; i < i1; i++) {
            const float v = src[i] / sc + ((Order == 1) ? e1 : 2.0f * e1 - e2);
            const TT q = TT(v);
            float e = v - float(q);
            e = (fabsf(e) <= FLT_MAX) ? e : 0.0f;
            e2 = e1;
            e1 = e;
            dst[i] = q;
        }
    }
}

template<typename TT>
void encode_typed(TT *out, float *scale, const float *tod, int64_t nrows, int64_t row_len, int order, hipStream_t stream)
{
    auto k = (order == 2) ? tod16_encode_kernel<TT, 2> : (order == 1) ? tod16_encode_kernel<TT, 1> : tod16_encode_kernel<TT, 0>;
    k<<<unsigned(nrows), ENC_THREADS, 0, stream>>>(out, scale, tod, row_len);
}

}  // namespace

void launch_tod16_encode(void *out, int tod_type, float *scale, const float *tod, int64_t nrows,
                         int64_t row_len, int error_feedback, hipStream_t stream)
{
    if (tod_type == TOD_F16)
        encode_typed(static_cast<_Float16 *>(out), scale, tod, nrows, row_len, error_feedback, stream);
    else
        encode_typed(static_cast<__bf16 *>(out), scale, tod, nrows, row_len, error_feedback, stream);
    GMM_HIP(hipGetLastError());
}

}  // namespace gmm
