// Streaming-read bandwidth probe on gfx950: a measurement aid for the benchmark (no counterpart
// in the reference). It reads a float array once with non-temporal 16-byte loads, the access
// the planned operators use for their pointing and TOD streams, so that its rate is the read
// ceiling those kernels are compared with. It also sums the array (one atomic per block), so a
// test can check that every element was read.

#include "gmm_internal.hpp"

#include <algorithm>

namespace gmm {

namespace {

constexpr int PR_THREADS = 512;

typedef float f4v __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(PR_THREADS)
stream_read_kernel(const f4v *__restrict__ p, int64_t n4, float *__restrict__ sum)
{
    float acc = 0.0f;
    const int64_t stride = int64_t(gridDim.x) * PR_THREADS;
    for (int64_t i = int64_t(blockIdx.x) * PR_THREADS + threadIdx.x; i < n4; i += stride) {
        const f4v v = __builtin_nontemporal_load(p + i);
        acc += (v.x + v.y) + (v.z + v.w);
    }
    // One atomic per block: same-address atomics serialize in L2 (one per wave, 8192 of them,
    // added ~0.08 ms to every launch).
    __shared__ float s_part[PR_THREADS / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        acc += __shfl_xor(acc, off, 64);
    if ((threadIdx.x & 63) == 0)
        s_part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.0f;
#pragma unroll
        for (int w = 0; w < PR_THREADS / 64; w++)
            t += s_part[w];
        atomicAdd(sum, t);
    }
}

}  // namespace

void launch_stream_read(const float *p, int64_t n, float *sum, hipStream_t stream)
{
    // 4 blocks of 512 threads per CU: enough loads in flight to saturate HBM (grids of 1024 to
    // 8192 blocks measure the same).
    int dev = 0, ncu = 0;
    GMM_HIP(hipGetDevice(&dev));
    GMM_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int64_t n4 = n / 4;
    const int64_t need = (n4 + PR_THREADS - 1) / PR_THREADS;
    const int grid = int(std::max<int64_t>(1, std::min<int64_t>(int64_t(4) * std::max(ncu, 1), need)));
    stream_read_kernel<<<grid, PR_THREADS, 0, stream>>>(reinterpret_cast<const f4v *>(p), n4, sum);
    GMM_HIP(hipGetLastError());
}

}  // namespace gmm
