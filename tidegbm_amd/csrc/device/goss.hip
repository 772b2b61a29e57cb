// GOSS row selection on the device (reference booster.cpp:554-585): rows ordered by |g|
// descending, the first ceil(top * S) kept, then ceil(other * S) rows drawn from the rest in the
// order std::shuffle leaves them, with g and h of the drawn rows scaled.
//
// The reference orders with std::sort, which leaves the order of equal |g| unspecified, and
// draws with std::shuffle on its mt19937.  Here the rows are ordered by a radix sort of the |g|
// bit patterns (non-negative doubles order like their bits) and the draw is the permutation that
// the same std::shuffle call applies to the rest (computed on the host from the same generator
// state: it depends on the rest's size only).  When no two rows share a |g| value, the sorted
// order is unique and the selection is the reference's exactly; the launch reports ties so the
// caller can fall back to the host path, which uses the same std::sort.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "dev_common.h"
#include "launch.h"

namespace tgbm {

__global__ __launch_bounds__(256) void k_goss_keys(const double* __restrict__ g, int n,
                                                   unsigned long long* __restrict__ keys, int* __restrict__ vals,
                                                   unsigned char* __restrict__ keep,
                                                   unsigned char* __restrict__ scaled) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        keys[i] = (unsigned long long)__double_as_longlong(fabs(g[i]));
        vals[i] = i;
        keep[i] = 0;
        scaled[i] = 0;
    }
}

// Keep the top rows, mark the drawn rest rows (rest[perm[k]], k < n_other), flag equal keys
// whose order decides the sample.
__global__ __launch_bounds__(256) void k_goss_mark(const unsigned long long* __restrict__ keys,
                                                   const int* __restrict__ order, int n, int n_top,
                                                   const int* __restrict__ perm, int n_other,
                                                   unsigned char* __restrict__ keep,
                                                   unsigned char* __restrict__ scaled, int* tie) {
    const int stride = gridDim.x * blockDim.x;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
        // equal keys both among the kept top rows leave the sample unchanged
        if (j > 0 && j >= n_top && keys[j] == keys[j - 1]) *tie = 1;
        if (j < n_top) keep[order[j]] = 1;
    }
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n_other; k += stride) {
        const int row = order[n_top + perm[k]];
        keep[row] = 1;
        scaled[row] = 1;
    }
}

size_t goss_temp_bytes(int n) {
    size_t a = 0, b = 0;
    (void)hipcub::DeviceRadixSort::SortPairsDescending(nullptr, a, (const unsigned long long*)nullptr,
                                                        (unsigned long long*)nullptr, (const int*)nullptr,
                                                        (int*)nullptr, n);
    (void)hipcub::DeviceSelect::Flagged(nullptr, b, hipcub::CountingInputIterator<int>(0),
                                        (const unsigned char*)nullptr, (int*)nullptr, (int*)nullptr, n);
    return a > b ? a : b;
}

hipError_t launch_goss(hipStream_t st, const double* g, int n, int n_top, int n_other, const int* perm,
                       unsigned long long* k0, unsigned long long* k1, int* v0, int* v1, unsigned char* keep,
                       unsigned char* scaled, int* sample_out, int* n_sample_out, int* tie, void* tmp,
                       size_t tmp_bytes) {
    const int blocks = std::max(1, std::min(4096, (n + 255) / 256));
    hipLaunchKernelGGL(k_goss_keys, dim3(blocks), dim3(256), 0, st, g, n, k0, v0, keep, scaled);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    size_t bytes = tmp_bytes;
    e = hipcub::DeviceRadixSort::SortPairsDescending(tmp, bytes, k0, k1, v0, v1, n, 0, 64, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_goss_mark, dim3(blocks), dim3(256), 0, st, k1, v1, n, n_top, perm, n_other, keep, scaled,
                       tie);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    bytes = tmp_bytes;
    return hipcub::DeviceSelect::Flagged(tmp, bytes, hipcub::CountingInputIterator<int>(0), keep, sample_out,
                                         n_sample_out, n, st);
}

}  // namespace tgbm
