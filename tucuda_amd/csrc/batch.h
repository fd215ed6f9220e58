// batch.h -- persistent kernel over the tiles of many images (batched mode).
//
// The host lays out one SepArgs per image (job) and one list of tiles
// (job, tile number) in device memory.  A fixed grid of workgroups, sized to
// fill the GPU once, pulls tiles from a queue in device memory until the
// list is exhausted.  Each tile runs the same code as the single-image
// launch, instantiated for the job's radius class, so a pixel's arithmetic
// does not depend on the batch it is in.
#pragma once

#include "sepfir.h"

#include <type_traits>

namespace tucu {

struct BatchTile {
    uint32_t job;
    uint32_t tile;
};

// Tile queue.  The tile list is ordered by radius class, then job, then
// tile number (strip-major within a job), so neighbours in the list are
// vertically adjacent tiles that read the same halo rows.  It is cut into QN
// contiguous ranges, one per XCD, each with its own counter.  Workgroups are
// dispatched round-robin over the 8 XCDs, so workgroup b draws from range
// b % QN first, which keeps neighbouring tiles on one L2; once that range is
// empty it moves on to the following ranges in turn, which balances ranges
// of cheaper tiles and the tail.  Workgroup b starts with tile b / QN of its
// range (no atomic) and asks for its next tile while it computes the
// current one.  Each counter starts at the number of workgroups that took a
// first tile from its range; reset_queue sets them, launched just before the
// batch kernel on the same stream (a kernel rather than a memset, so the
// pair can be captured and replayed in a HIP graph).  Every tile is handed
// out exactly once whatever the placement of the workgroups, and a
// workgroup gives up after at most QN grabs that find a range empty.
constexpr int QN = 8;
constexpr int QSTRIDE = 32;                          // ints between counters: one 128-byte line each
constexpr size_t QUEUE_BYTES = QN * QSTRIDE * sizeof(int);

__device__ __forceinline__ int queue_lo(int ntiles, int r) { return int(int64_t(ntiles) * r / QN); }

template <int N = QN>          // a template, so every unit may instantiate it
__global__ void reset_queue(int* q, int grid)
{
    const int r = int(threadIdx.x);
    if (r < N) q[r * QSTRIDE] = (grid - r + N - 1) / N;    // workgroups b < grid with b % N == r
}

struct TileQueue {
    int r = 0, left = QN;      // current range, ranges not yet found empty

    // This workgroup's first tile, or -1 when its range has fewer tiles than
    // workgroups.
    __device__ int first(int ntiles)
    {
        r = int(blockIdx.x) % QN;
        const int i = queue_lo(ntiles, r) + int(blockIdx.x) / QN;
        return i < queue_lo(ntiles, r + 1) ? i : -1;
    }

    // The next tile, or ntiles when every range is empty.
    __device__ int grab(int* q, int ntiles)
    {
        for (; left > 0; --left, r = r + 1 == QN ? 0 : r + 1) {
            const int i = queue_lo(ntiles, r) +
                          __hip_atomic_fetch_add(q + r * QSTRIDE, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (i < queue_lo(ntiles, r + 1)) return i;
        }
        return ntiles;
    }
};

template <int TH, class IN, class OUT>
__global__ void __launch_bounds__(NT)
batch_kernel(const SepArgs* __restrict__ jobs_, const BatchTile* __restrict__ tiles, int ntiles,
             int* __restrict__ counter)
{
    __shared__ float4 lds4[tile_lds_floats<TH, MODE_GAUSS>() / 4];
    __shared__ int next;
    float* const lds = reinterpret_cast<float*>(lds4);
    ConstArgs* const jobs = (ConstArgs*)(uintptr_t)jobs_;

    TileQueue q;
    int ti = q.first(ntiles);
    for (;;) {
        if (threadIdx.x == 0)          // the answer is read after this tile
            next = q.grab(counter, ntiles);
        if (ti >= 0) {
            const BatchTile e = tiles[ti];
            ConstArgs& a = jobs[e.job];
            switch (a.rt) {
              case 2:  sepfir_tile<2, 2, TH, MODE_GAUSS, IN, OUT>(a, int(e.tile), lds); break;
              case 4:  sepfir_tile<4, 4, TH, MODE_GAUSS, IN, OUT>(a, int(e.tile), lds); break;
              case 8:  sepfir_tile<8, 8, TH, MODE_GAUSS, IN, OUT>(a, int(e.tile), lds); break;
              case 12: sepfir_tile<12, 12, TH, MODE_GAUSS, IN, OUT>(a, int(e.tile), lds); break;
              default: sepfir_tile<16, 16, TH, MODE_GAUSS, IN, OUT>(a, int(e.tile), lds); break;
            }
        }
        __syncthreads();               // LDS reuse; `next` was written before this barrier
        ti = __builtin_amdgcn_readfirstlane(next);
        __syncthreads();               // all have read `next` before thread 0 writes it again
        if (ti >= ntiles) break;
    }
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
    hipLaunchKernelGGL(reset_queue<>, dim3(1), dim3(64), 0, s, counter, grid);
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
