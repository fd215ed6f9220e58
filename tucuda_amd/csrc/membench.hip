// membench.hip -- reference points for the benchmark's roofline: how fast a
// plain write stream and a plain copy run on this GPU with the store form
// the filters use (16-byte non-temporal stores).  Built into its own
// library (libtucuda_membench.so), loaded only by bench.py.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

typedef float v4f32 __attribute__((ext_vector_type(4)));
template <class T> using gptr = T __attribute__((address_space(1)))*;

constexpr int TPB = 256;

template <bool NT> __device__ __forceinline__ void store16(v4f32* p, v4f32 x)
{
    const gptr<v4f32> q = reinterpret_cast<gptr<v4f32>>(reinterpret_cast<uintptr_t>(p));
    if constexpr (NT) __builtin_nontemporal_store(x, q);
    else              *q = x;
}

// Every 16-byte element of [p, p + n16) written with `v`; grid-stride, four
// stores per thread and iteration; streaming (non-temporal) stores, the
// filters' form, or ordinary ones.
template <bool NT>
__global__ void __launch_bounds__(TPB) fill_kernel(v4f32* __restrict__ p, int64_t n16, float v)
{
    const v4f32 x = {v, v, v, v};
    const int64_t step = int64_t(gridDim.x) * TPB;
    int64_t i = int64_t(blockIdx.x) * TPB + threadIdx.x;
    for (; i + 3 * step < n16; i += 4 * step) {
#pragma unroll
        for (int k = 0; k < 4; ++k) store16<NT>(p + i + k * step, x);
    }
    for (; i < n16; i += step) store16<NT>(p + i, x);
}

// dst = src, 16 bytes per thread and step, streaming loads and stores.
__global__ void __launch_bounds__(TPB) copy_kernel(v4f32* __restrict__ dst, const v4f32* __restrict__ src,
                                                   int64_t n16)
{
    const int64_t step = int64_t(gridDim.x) * TPB;
    for (int64_t i = int64_t(blockIdx.x) * TPB + threadIdx.x; i < n16; i += step) {
        const v4f32 x = __builtin_nontemporal_load(reinterpret_cast<gptr<const v4f32>>(reinterpret_cast<uintptr_t>(src + i)));
        __builtin_nontemporal_store(x, reinterpret_cast<gptr<v4f32>>(reinterpret_cast<uintptr_t>(dst + i)));
    }
}

int grid_for(int64_t n16, int per_cu = 8)
{
    int dev = 0, ncu = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    const int64_t full = (n16 + TPB - 1) / TPB;
    const int64_t cap = int64_t(ncu) * per_cu;      // 256-thread workgroups per CU
    return int(full < cap ? full : cap);
}

}   // namespace

extern "C" {

// bytes must be a multiple of 16 and p 16-byte aligned.  nontemporal: the
// filters' store form (1) or ordinary stores (0); per_cu: workgroups of
// 256 threads per CU (0: 8).
__attribute__((visibility("default"))) int tucu_membench_fill(void* p, size_t bytes, float v, int nontemporal,
                                                              int per_cu, void* stream)
{
    if (!p || bytes % 16 || reinterpret_cast<uintptr_t>(p) % 16 || per_cu < 0 || per_cu > 64) return 1;
    const int64_t n16 = int64_t(bytes / 16);
    if (n16 == 0) return 0;
    const dim3 grid(grid_for(n16, per_cu ? per_cu : 8));
    if (nontemporal)
        hipLaunchKernelGGL(fill_kernel<true>, grid, dim3(TPB), 0, static_cast<hipStream_t>(stream),
                           static_cast<v4f32*>(p), n16, v);
    else
        hipLaunchKernelGGL(fill_kernel<false>, grid, dim3(TPB), 0, static_cast<hipStream_t>(stream),
                           static_cast<v4f32*>(p), n16, v);
    return hipGetLastError() == hipSuccess ? 0 : 5;
}

__attribute__((visibility("default"))) int tucu_membench_copy(void* dst, const void* src, size_t bytes, void* stream)
{
    if (!dst || !src || bytes % 16 || reinterpret_cast<uintptr_t>(dst) % 16 || reinterpret_cast<uintptr_t>(src) % 16)
        return 1;
    const int64_t n16 = int64_t(bytes / 16);
    if (n16 == 0) return 0;
    hipLaunchKernelGGL(copy_kernel, dim3(grid_for(n16)), dim3(TPB), 0, static_cast<hipStream_t>(stream),
                       static_cast<v4f32*>(dst), static_cast<const v4f32*>(src), n16);
    return hipGetLastError() == hipSuccess ? 0 : 5;
}

}   // extern "C"
