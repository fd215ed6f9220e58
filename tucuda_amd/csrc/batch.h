// batch.h -- persistent kernel over the tiles of many images (batched mode).
//
// The host lays out one SepArgs per image (job) and one list of tiles
// (job, tile number) in device memory.  A fixed grid of workgroups, sized to
// fill the GPU once, pulls tiles from device counters until the list is
// exhausted.  Each tile runs the same sepfir_tile code as the single-image launch, instantiated for the job's
// radius class, so a pixel's arithmetic does not depend on the batch it is
// in.  The last workgroup to leave resets the counters, which lets a
// captured graph replay the launch without clearing memory in between.
#pragma once

#include "sepfir.h"

#include <type_traits>

namespace tucu {

struct BatchTile {
    uint32_t job;
    uint32_t tile;
};

// Tile scheduling.  Workgroup w starts with tile w (no atomic), then takes
// tile grid + counter++ until the list is exhausted, asking for its next
// tile while it computes the current one; workgroups that drew cheap tiles
// take more of them, which balances batches of mixed radii and sizes.  The
// counter is cleared by reset_counter, launched just before the batch
// kernel on the same stream (a kernel rather than a memset, so the pair can
// be captured and replayed in a HIP graph).  Clearing at the start avoids a
// "last workgroup out" atomic at the end, which hundreds of workgroups
// finishing together would serialise on.
template <int TH, class IN, class OUT>
__global__ void __launch_bounds__(NT)
batch_kernel(const SepArgs* __restrict__ jobs_, const BatchTile* __restrict__ tiles, int ntiles,
             int* __restrict__ counter)
{
    __shared__ float4 lds4[tile_lds_floats<TH, MODE_GAUSS>() / 4];
    __shared__ int next;
    float* const lds = reinterpret_cast<float*>(lds4);
    ConstArgs* const jobs = (ConstArgs*)(uintptr_t)jobs_;

    int ti = int(blockIdx.x);
    while (ti < ntiles) {
        if (threadIdx.x == 0)          // the answer is read after this tile
            next = int(gridDim.x) + __hip_atomic_fetch_add(counter, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const BatchTile e = tiles[ti];
        ConstArgs& a = jobs[e.job];
        switch (a.rt) {
          case 2:  sepfir_tile<2, 2, TH, MODE_GAUSS, IN, OUT>(a, int(e.tile), lds); break;
          case 4:  sepfir_tile<4, 4, TH, MODE_GAUSS, IN, OUT>(a, int(e.tile), lds); break;
          case 8:  sepfir_tile<8, 8, TH, MODE_GAUSS, IN, OUT>(a, int(e.tile), lds); break;
          case 12: sepfir_tile<12, 12, TH, MODE_GAUSS, IN, OUT>(a, int(e.tile), lds); break;
          default: sepfir_tile<16, 16, TH, MODE_GAUSS, IN, OUT>(a, int(e.tile), lds); break;
        }
        __syncthreads();               // LDS reuse; `next` was written before this barrier
        ti = __builtin_amdgcn_readfirstlane(next);
        __syncthreads();               // all have read `next` before thread 0 writes it again
    }
}

template <int N>               // a template, so every unit may instantiate it
__global__ void reset_counter(int* counter)
{
    if (threadIdx.x < N) counter[threadIdx.x] = 0;
}

using BatchLaunchFn = hipError_t (*)(const SepArgs*, const BatchTile*, int ntiles, int* counter,
                                     int grid, hipStream_t);

struct BatchEntry {
    BatchLaunchFn fn = nullptr;
    const void*   kernel = nullptr;   // for occupancy queries
    int           th = 0;             // tile rows
    int           tw = 0;             // tile columns (0: NT - 2 x radius class)
    int           nthreads = NT;      // workgroup size
};

template <int TH, class IN, class OUT>
hipError_t launch_batch(const SepArgs* jobs, const BatchTile* tiles, int ntiles, int* counter,
                        int grid, hipStream_t s)
{
    hipLaunchKernelGGL(reset_counter<1>, dim3(1), dim3(64), 0, s, counter);
    hipLaunchKernelGGL((batch_kernel<TH, IN, OUT>), dim3(grid), dim3(NT), 0, s, jobs, tiles, ntiles, counter);
    return hipGetLastError();
}

constexpr int NBTH = 3;
constexpr int BATCH_THS[NBTH] = {4, 8, 16};
constexpr int BATCH_TH_DEFAULT = 1;

// [tile-height index][in dtype][out dtype].  Every height for the same-type
// f32 and bf16 pairs, the default height for the other pairs.
extern BatchEntry g_batch_table[NBTH][4][4];

// Matrix-core batched kernels, [workgroup-shape index][in dtype][out dtype]
// (tile rows 16, 32, 64), filled by the mfma_<type> units.
extern BatchEntry g_batch_mfma_table[NBTH][4][4];

template <class T> constexpr int dtype_index()
{
    return std::is_same_v<T, uint8_t> ? 0 : std::is_same_v<T, float> ? 1 : std::is_same_v<T, f16_t> ? 2 : 3;
}


}   // namespace tucu
