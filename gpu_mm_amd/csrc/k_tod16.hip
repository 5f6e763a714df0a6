// 16-bit TOD encoding on gfx950: one FP32 scale per TOD row, values stored as FP16 or BF16.
//
// scale[r] = max(max_s |tod[r][s]| / (0.5 * M), FLT_MIN), M = largest finite value of the type
// (65504 for FP16, 0x1.fep127 for BF16), so every encoded value lies in [-M/2, M/2]; the FLT_MIN
// floor keeps the scale a normal float (an all-zero row gets scale FLT_MIN and zeros).
// One block per row: a max-reduction pass, then a pass that divides and rounds to nearest even.
// This has no counterpart in the reference, which stores TODs as float32 only.

#include "gmm_internal.hpp"

#include <cfloat>
#include <type_traits>

namespace gmm {

namespace {

constexpr int ENC_THREADS = 256;

template<typename TT>
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
    for (int64_t i = threadIdx.x; i < row_len; i += ENC_THREADS)
        dst[i] = TT(src[i] / sc);
}

}  // namespace

void launch_tod16_encode(void *out, int tod_type, float *scale, const float *tod, int64_t nrows,
                         int64_t row_len, hipStream_t stream)
{
    if (tod_type == TOD_F16)
        tod16_encode_kernel<_Float16><<<unsigned(nrows), ENC_THREADS, 0, stream>>>(static_cast<_Float16 *>(out), scale,
                                                                                   tod, row_len);
    else
        tod16_encode_kernel<__bf16><<<unsigned(nrows), ENC_THREADS, 0, stream>>>(static_cast<__bf16 *>(out), scale,
                                                                                 tod, row_len);
    GMM_HIP(hipGetLastError());
}

}  // namespace gmm
