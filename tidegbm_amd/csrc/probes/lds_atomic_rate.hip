// Probe: LDS atomic update rate on this GPU (lanes per clock per CU), the resource that bounds
// the histogram builds.  Each of `blocks` 256-thread blocks keeps an 8-feature x 256-bin LDS
// histogram and issues iters x 16 atomic adds per thread to pseudo-random bins (the access
// pattern of a histogram over quantile bins), as 64-bit adds (one per row and feature, the
// packed count + fixed-point sum word) and, for comparison, 32-bit adds.
//
// Also measured: the same adds into histograms with K copies per bin (lane l updates copy l % K),
// which spreads the lanes of a wave over the LDS banks.
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

// Replicated histograms: K copies of each bin, lane l updating copy l % K, laid out so that the
// K copies of one bin sit in consecutive words.  For 64-bit words and K = 32 every lane of a
// wave touches its own bank pair whatever its bin; smaller K trade LDS size for conflicts.
// NF features of K * 256 words fill 64 KB.
template <int K>
__global__ __launch_bounds__(256) void k_rate_rep(int iters, int* sink) {
    constexpr int NF = 8192 / (256 * K) > 0 ? 8192 / (256 * K) : 1;
    __shared__ unsigned long long h64[NF * 256 * K];
    for (int j = threadIdx.x; j < NF * 256 * K; j += blockDim.x) h64[j] = 0;
    __syncthreads();
    unsigned bins[16];
#pragma unroll
    for (int t = 0; t < 16; ++t) bins[t] = mix(threadIdx.x * 977u + blockIdx.x * 131u + t * 7919u) & 255u;
    const unsigned rep = threadIdx.x % K;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            const unsigned idx = ((t % NF) * 256u + ((bins[t] + (unsigned)i) & 255u)) * K + rep;
            atomicAdd(&h64[idx], 1ull);
        }
    }
    __syncthreads();
    if (h64[threadIdx.x] == 0x123456789ull) sink[0] = 1;
}

// 32-bit words, 64 copies: every lane owns one bank.
__global__ __launch_bounds__(256) void k_rate_rep32(int iters, int* sink) {
    __shared__ unsigned h32[256 * 64];
    for (int j = threadIdx.x; j < 256 * 64; j += blockDim.x) h32[j] = 0;
    __syncthreads();
    unsigned bins[16];
#pragma unroll
    for (int t = 0; t < 16; ++t) bins[t] = mix(threadIdx.x * 977u + blockIdx.x * 131u + t * 7919u) & 255u;
    const unsigned rep = threadIdx.x & 63u;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int t = 0; t < 16; ++t) atomicAdd(&h32[((bins[t] + (unsigned)i) & 255u) * 64u + rep], 1u);
    }
    __syncthreads();
    if (h32[threadIdx.x] == 0x12345678u) sink[0] = 1;
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
    double rrep[6] = {0};
    const int ks[6] = {2, 4, 8, 16, 32, 64};
    for (int v = 0; v < 6; ++v) {
        float us = 0;
        switch (v) {
            case 0: us = time_us([&] { hipLaunchKernelGGL(k_rate_rep<2>, dim3(blocks), dim3(256), 0, 0, iters, sink); }, 10); break;
            case 1: us = time_us([&] { hipLaunchKernelGGL(k_rate_rep<4>, dim3(blocks), dim3(256), 0, 0, iters, sink); }, 10); break;
            case 2: us = time_us([&] { hipLaunchKernelGGL(k_rate_rep<8>, dim3(blocks), dim3(256), 0, 0, iters, sink); }, 10); break;
            case 3: us = time_us([&] { hipLaunchKernelGGL(k_rate_rep<16>, dim3(blocks), dim3(256), 0, 0, iters, sink); }, 10); break;
            case 4: us = time_us([&] { hipLaunchKernelGGL(k_rate_rep<32>, dim3(blocks), dim3(256), 0, 0, iters, sink); }, 10); break;
            default: us = time_us([&] { hipLaunchKernelGGL(k_rate_rep32, dim3(blocks), dim3(256), 0, 0, iters, sink); }, 10); break;
        }
        CK(hipGetLastError());
        rrep[v] = adds / (us * 1e-6) / cu_clk;
        std::printf("%s-bit LDS atomic adds, %d copies per bin: %.1f us -> %.2f lanes/clk/CU\n", v == 5 ? "32" : "64",
                    ks[v], us, rrep[v]);
    }
    std::printf("%s: %d CUs at %.0f MHz\n", pr.gcnArchName, pr.multiProcessorCount, pr.clockRate / 1e3);
    std::printf("64-bit LDS atomic adds: %.1f us for %.0f adds -> %.2f lanes/clk/CU\n", us64, adds, r64);
    std::printf("32-bit LDS atomic adds: %.1f us for %.0f adds -> %.2f lanes/clk/CU\n", us32, adds, r32);
    std::printf("{\"lds_u64_lanes_per_clk_cu\": %.3f, \"lds_u32_lanes_per_clk_cu\": %.3f, \"lds_u64_copies_2_4_8_16_32\": "
                "[%.3f, %.3f, %.3f, %.3f, %.3f], \"lds_u32_copies_64\": %.3f, \"cus\": %d, \"mhz\": %.0f}\n",
                r64, r32, rrep[0], rrep[1], rrep[2], rrep[3], rrep[4], rrep[5], pr.multiProcessorCount,
                pr.clockRate / 1e3);
    CK(hipFree(sink));
    return 0;
}
