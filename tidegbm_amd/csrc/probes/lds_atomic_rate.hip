// Probe: LDS atomic update rate on this GPU (lanes per clock per CU), the resource that bounds
// the histogram builds.  Each of `blocks` 256-thread blocks keeps an 8-feature x 256-bin LDS
// histogram and issues iters x 16 atomic adds per thread to pseudo-random bins (the access
// pattern of a histogram over quantile bins), as 64-bit adds (one per row and feature, the
// packed count + fixed-point sum word) and, for comparison, 32-bit adds.
//
// Usage: probe_lds_atomic_rate  (prints one line per variant and a JSON summary)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CK(x)                                                                                   \
    do {                                                                                        \
        hipError_t e_ = (x);                                                                    \
        if (e_ != hipSuccess) {                                                                 \
            std::fprintf(stderr, "HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
            std::exit(1);                                                                       \
        }                                                                                       \
    } while (0)

__device__ inline unsigned mix(unsigned x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

template <bool WIDE>
__global__ __launch_bounds__(256) void k_rate(int iters, int* sink) {
    __shared__ unsigned long long h64[8 * 256];
    unsigned* h32 = reinterpret_cast<unsigned*>(h64);
    for (int j = threadIdx.x; j < 8 * 256; j += blockDim.x) h64[j] = 0;
    __syncthreads();
    unsigned bins[16];
#pragma unroll
    for (int t = 0; t < 16; ++t) bins[t] = mix(threadIdx.x * 977u + blockIdx.x * 131u + t * 7919u) & 255u;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            const unsigned idx = ((t & 7u) << 8) + ((bins[t] + (unsigned)i) & 255u);
            if (WIDE) atomicAdd(&h64[idx], 1ull);
            else atomicAdd(&h32[idx], 1u);
        }
    }
    __syncthreads();
    if (h64[threadIdx.x] == 0x123456789ull) sink[0] = 1;   // keeps the adds live
}

template <class K>
static float time_us(K launch, int reps) {
    hipEvent_t a, b;
    CK(hipEventCreate(&a));
    CK(hipEventCreate(&b));
    launch();
    CK(hipDeviceSynchronize());
    CK(hipEventRecord(a));
    for (int r = 0; r < reps; ++r) launch();
    CK(hipEventRecord(b));
    CK(hipEventSynchronize(b));
    float ms = 0;
    CK(hipEventElapsedTime(&ms, a, b));
    CK(hipEventDestroy(a));
    CK(hipEventDestroy(b));
    return ms * 1000.0f / reps;
}

int main() {
    int dev = 0;
    hipDeviceProp_t pr;
    CK(hipGetDevice(&dev));
    CK(hipGetDeviceProperties(&pr, dev));
    int* sink = nullptr;
    CK(hipMalloc(&sink, sizeof(int)));
    const int blocks = 4096, iters = 64;
    const double adds = (double)blocks * 256 * iters * 16;
    const double cu_clk = pr.multiProcessorCount * (pr.clockRate * 1e3);   // clockRate is in kHz
    const float us64 = time_us([&] { hipLaunchKernelGGL(k_rate<true>, dim3(blocks), dim3(256), 0, 0, iters, sink); }, 10);
    const float us32 = time_us([&] { hipLaunchKernelGGL(k_rate<false>, dim3(blocks), dim3(256), 0, 0, iters, sink); }, 10);
    CK(hipGetLastError());
    const double r64 = adds / (us64 * 1e-6) / cu_clk, r32 = adds / (us32 * 1e-6) / cu_clk;
    std::printf("%s: %d CUs at %.0f MHz\n", pr.gcnArchName, pr.multiProcessorCount, pr.clockRate / 1e3);
    std::printf("64-bit LDS atomic adds: %.1f us for %.0f adds -> %.2f lanes/clk/CU\n", us64, adds, r64);
    std::printf("32-bit LDS atomic adds: %.1f us for %.0f adds -> %.2f lanes/clk/CU\n", us32, adds, r32);
    std::printf("{\"lds_u64_lanes_per_clk_cu\": %.3f, \"lds_u32_lanes_per_clk_cu\": %.3f, \"cus\": %d, \"mhz\": %.0f}\n", r64,
                r32, pr.multiProcessorCount, pr.clockRate / 1e3);
    CK(hipFree(sink));
    return 0;
}
