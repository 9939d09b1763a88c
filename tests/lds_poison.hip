// lds_poison.hip - test support: leaves NaN in the LDS of every CU, and
// measures whether what one kernel leaves in LDS is what the next kernel's
// workgroups find there. A solver kernel that reads LDS it has not written
// first then sees NaN instead of whatever an earlier workgroup left, which
// makes such a read show in its results.
#include <hip/hip_runtime.h>

namespace {

constexpr unsigned kFill = 0xFFFFFFFFu;  // a NaN as float and as double
constexpr int kThreads = 1024;

__global__ __launch_bounds__(kThreads) void k_fill(unsigned words) {
    extern __shared__ unsigned lds[];
    for (unsigned i = threadIdx.x; i < words; i += kThreads) lds[i] = kFill;
}

// Per wave: the number of LDS words (of `words`) holding the fill pattern,
// read before anything in this kernel writes LDS.
__global__ __launch_bounds__(kThreads) void k_probe(unsigned words, unsigned* __restrict__ out) {
    extern __shared__ unsigned lds[];
    unsigned n = 0;
    for (unsigned i = threadIdx.x; i < words; i += kThreads) n += lds[i] == kFill ? 1u : 0u;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off, 64);
    if ((threadIdx.x & 63) == 0) out[blockIdx.x * (kThreads / 64) + threadIdx.x / 64] = n;
}

}  // namespace

extern "C" {

// Fills `bytes` of LDS (at most 160 KB) in `blocks` workgroups of 1024
// threads, then waits for them.
__attribute__((visibility("default"))) int lds_poison_fill(int blocks, int bytes) {
    if (blocks < 1 || bytes < 4 || bytes > 160 * 1024) return 1;
    if (hipFuncSetAttribute((const void*)k_fill, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) != hipSuccess)
        return 2;
    k_fill<<<blocks, kThreads, bytes>>>((unsigned)bytes / 4);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) return 2;
    return 0;
}

// Fraction of the LDS words that `blocks` probe workgroups (same size)
// find holding the fill pattern.
__attribute__((visibility("default"))) int lds_poison_probe(int blocks, int bytes, double* frac) {
    if (blocks < 1 || bytes < 4 || bytes > 160 * 1024 || !frac) return 1;
    if (hipFuncSetAttribute((const void*)k_probe, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) != hipSuccess)
        return 2;
    const size_t nw = (size_t)blocks * (kThreads / 64);
    unsigned* d = nullptr;
    if (hipMalloc(&d, nw * sizeof(unsigned)) != hipSuccess) return 2;
    k_probe<<<blocks, kThreads, bytes>>>((unsigned)bytes / 4, d);
    unsigned* h = new unsigned[nw];
    const bool ok = hipGetLastError() == hipSuccess &&
                    hipMemcpy(h, d, nw * sizeof(unsigned), hipMemcpyDeviceToHost) == hipSuccess;
    double sum = 0.0;
    for (size_t k = 0; k < nw; ++k) sum += h[k];
    delete[] h;
    (void)hipFree(d);
    if (!ok) return 2;
    *frac = sum / ((double)blocks * (bytes / 4));
    return 0;
}
}
