// icia.h -- ICIA's error moment on the matrix-core strip code.
//
// TU/cu/ICIA.h:344-369 computes diffH and diffV of the image into two
// images, then sums ICIAErrorMoment (:61-112) over every pixel:
// image_derivative0(s u, s v, eH, eV).ext(), the upper triangle (row-major,
// TU/array.h:319-325) of a a^T with a the D-vector of the affine (D = 6,
// TU/cu/vec.h:1418-1422) or projective (D = 8, :1285-1290) map, s = 1 /
// max(nrow, ncol), u, v the pixel's column and row.  Here one pass computes
// both gradients with the strip code (MODE_ICIA: the accumulators go to a
// sink instead of the outputs) and reduces them on chip; a gradient is 0
// outside its valid region, as the reference's freshly allocated gradient
// images are there.
//
// Per lane (one row of 16 pixels per strip, so v is fixed): the components
// of a are sums of terms sign * e_X * u^i * v^j (X = H, V), so each moment
// is a sum of sign * A[XY][i + i'] * v^(j + j') over term pairs, with
// A[XY][n] = sum over the lane's pixels of e_X e_Y u^n.  A lane accumulates
// the 3 x (3 or 5) sums A (three products and 9 or 15 FMAs per pixel,
// instead of D (D + 1) / 2 FMAs and the components), then adds the moments
// they give into its FP32 moment sums.  A workgroup filters up to
// ICIA_MAX_TILES tiles; its lanes' sums are added within each row of 16
// lanes by DPP, and across rows and waves in FP64, into one partial per
// workgroup; icia_reduce_kernel adds the partials in FP64 in a fixed order
// (one wave per moment), so the result does not depend on timing.
#pragma once

#include "mfma.h"

namespace tucu {

template <int D> constexpr int icia_nm() { return D * (D + 1) / 2; }

// Tiles per workgroup at most (bounds the FP32 sums of a lane).
constexpr int ICIA_MAX_TILES = 8;

// A component of image_derivative0 as up to two terms sign * e_X u^i v^j.
struct IciaTerm { int x, i, j, sgn; };
struct IciaComp { int n; IciaTerm t[2]; };
// affine: (eH u, eH v, eH, eV u, eV v, eV); projective adds
// -(eH u + eV v) u = -eH u^2 - eV u v and -(eH u + eV v) v = -eH u v - eV v^2
constexpr IciaComp ICIA_COMP[8] = {
    {1, {{0, 1, 0, 1}, {}}}, {1, {{0, 0, 1, 1}, {}}}, {1, {{0, 0, 0, 1}, {}}},
    {1, {{1, 1, 0, 1}, {}}}, {1, {{1, 0, 1, 1}, {}}}, {1, {{1, 0, 0, 1}, {}}},
    {2, {{0, 2, 0, -1}, {1, 1, 1, -1}}}, {2, {{0, 1, 1, -1}, {1, 0, 2, -1}}},
};

// Sum over each row of 16 lanes, in every lane of the row.
__device__ __forceinline__ float row_sum(float v)
{
    v += dpp_f32<0xb1>(v);                      // quad_perm [1, 0, 3, 2]
    v += dpp_f32<0x4e>(v);                      // quad_perm [2, 3, 0, 1]
    v += dpp_f32<0x141>(v);                     // row_half_mirror
    v += dpp_f32<0x140>(v);                     // row_mirror
    return v;
}

// mom[k * ms] += the moments of this lane's 16 pixels (row y, columns x0 +
// 16 o + 4 g + i of the accumulators gH, gV), zero outside each gradient's
// valid region.
template <int D>
__device__ __forceinline__ void icia_lane_moments(const SepArgs& a, int y0, int x0, const v4f32* gH,
                                                  const v4f32* gV, float* mom, int ms)
{
    constexpr int NU = D == 6 ? 3 : 5;          // powers of u: 0 .. 2 (affine), 0 .. 4
    const int lane = threadIdx.x & 63, g = lane >> 4, m = lane & 15;
    const float sc = 1.f / float(max(a.nrow, a.ncol));
    const int y = y0 + m, r0 = a.r[0], r1 = a.r[1];
    const bool rowH = y >= r0 && y < a.nrow - r0;          // diffH: V lobe 0, H lobe 1
    const bool rowV = y >= r1 && y < a.nrow - r1;          // diffV: V lobe 1, H lobe 0
    float A[3][NU];
#pragma unroll
    for (int xy = 0; xy < 3; ++xy)
#pragma unroll
        for (int n = 0; n < NU; ++n) A[xy][n] = 0.f;
#pragma unroll
    for (int o = 0; o < MNB; ++o)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int x = x0 + 16 * o + 4 * g + i;
            const float eH = rowH && x >= r1 && x < a.ncol - r1 ? gH[o][i] : 0.f;
            const float eV = rowV && x >= r0 && x < a.ncol - r0 ? gV[o][i] : 0.f;
            const float ee[3] = {eH * eH, eH * eV, eV * eV};
            float up[NU];
            up[0] = 1.f;
            up[1] = sc * float(x);
#pragma unroll
            for (int n = 2; n < NU; ++n) up[n] = up[n - 1] * up[1];
#pragma unroll
            for (int xy = 0; xy < 3; ++xy) {
                A[xy][0] += ee[xy];
#pragma unroll
                for (int n = 1; n < NU; ++n) A[xy][n] = __builtin_fmaf(ee[xy], up[n], A[xy][n]);
            }
        }
    float vp[5];
    vp[0] = 1.f;
    vp[1] = sc * float(y);
#pragma unroll
    for (int n = 2; n < 5; ++n) vp[n] = vp[n - 1] * vp[1];
    int k = 0;
#pragma unroll
    for (int p = 0; p < D; ++p)
#pragma unroll
        for (int q = p; q < D; ++q, ++k) {
            float s = 0.f;
#pragma unroll
            for (int tp = 0; tp < ICIA_COMP[p].n; ++tp)
#pragma unroll
                for (int tq = 0; tq < ICIA_COMP[q].n; ++tq) {
                    const IciaTerm P = ICIA_COMP[p].t[tp], Q = ICIA_COMP[q].t[tq];
                    const float t = A[P.x + Q.x][P.i + Q.i] * vp[P.j + Q.j];
                    s += P.sgn * Q.sgn > 0 ? t : -t;
                }
            mom[k * ms] += s;
        }
}

// Workgroup b filters tiles t, t + grid, t + 2 grid, ... with t =
// xcd_remap(b) (strip-major tile numbers: the workgroups of one XCD take
// contiguous ranges, whose tiles share halo rows in its L2) and writes its
// FP64 partial moments to partial[k][b] (k = 0 .. NM - 1).
template <int RC, class IN, int D>
__global__ void __launch_bounds__(64 * MWAVES) __attribute__((amdgpu_waves_per_eu(2)))   // 3 spill
icia_mfma_kernel(const SepArgs a, int ntiles, double* __restrict__ partial)
{
    constexpr int WY = MFMA_WYS[MFMA_WY_ICIA];
    constexpr int NM = icia_nm<D>();
    __shared__ MfmaTabs T;
    __shared__ v4f32 stage4[MWAVES * 16 * MOPITCH / 4];
    __shared__ float red[MWAVES][4][NM];
    // Each lane's FP32 moment sums: in registers, or, for FP32 input at
    // radius class 16 with the projective map, in LDS (moment k of thread t
    // at momsh[k][t]), where in registers they stayed live through the strip
    // code and it spilled 80 bytes per lane (1080p: 14.2 against 16.7 us per
    // call; the LDS round trips made the other instances 4-14 % slower).
    constexpr bool MLDS = RC == 16 && D == 8 && std::is_same_v<IN, float>;
    constexpr int MS = MLDS ? 64 * MWAVES : 1;
    __shared__ float momsh[MLDS ? NM * MS : 1];
    float momr[MLDS ? 1 : NM];
    float* const mom = MLDS ? momsh + threadIdx.x : momr;
    mfma_args_early(a);
    float* const stage = reinterpret_cast<float*>(stage4) + (threadIdx.x >> 6) * 16 * MOPITCH;
    const int grid = int(gridDim.x);
    int t = xcd_remap(int(blockIdx.x), grid);
    int y0, x0;
    bool live = mfma_strip_pos<WY>(a, t, y0, x0);
    MfmaRaw<IN> raw;
    if (live) mfma_load<RC, IN>(a, y0, x0, raw);    // in flight while the tables are built
    mfma_get_tabs<MODE_GAUSS>(a, T);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NM; ++k) mom[k * MS] = 0.f;
    // Radius class 16 takes the lane index afresh in each tile, as the
    // six-output tile loop does (3840 x 2160 u8 projective 3.6 %, 1080p f32
    // affine 2 % faster; radius class 8 within +-1.5 %, so it does not).
    constexpr bool FRESH = RC == 16;
    while (t < ntiles) {                            // uniform over the workgroup
        if (live) {
            v4f32 gH[MNB], gV[MNB];
            mfma_compute<RC, IN, float, MODE_ICIA, CPOL_NT, FRESH>(
                a, y0, x0, raw, T, stage, [] {},
                [&](int k, const v4f32* acc, float u) {
#pragma unroll
                    for (int o = 0; o < MNB; ++o) {
                        if (k == 1) gH[o] = acc[o] * u;
                        else        gV[o] = acc[o] * u;
                    }
                }, a.mask);
            icia_lane_moments<D>(a, y0, x0, gH, gV, mom, MS);
        }
        t += grid;
        live = t < ntiles && mfma_strip_pos<WY>(a, t, y0, x0);
        if (live) mfma_load<RC, IN, FRESH>(a, y0, x0, raw);
    }
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < NM; ++k) {
        const float s = row_sum(mom[k * MS]);
        if ((lane & 15) == 0) red[w][lane >> 4][k] = s;
    }
    __syncthreads();
    if (threadIdx.x < NM) {
        double s = 0.0;
#pragma unroll
        for (int v = 0; v < MWAVES; ++v)
#pragma unroll
            for (int r = 0; r < 4; ++r) s += double(red[v][r][threadIdx.x]);
        partial[int64_t(threadIdx.x) * gridDim.x + blockIdx.x] = s;   // moment-major
    }
}

// moment[k] = the sum of partial[k][b] over the nblocks workgroups, in FP64
// in a fixed order; 0 when nblocks is 0.  One wave per moment (workgroup k
// sums moment k): lane l adds b = l, l + 64, l + 128, ... in order, eight
// loads in flight at a time (the partials are stored moment-major, so each
// load reads 64 consecutive doubles), then the 64 lane sums are added by a
// butterfly (xor 32, 16, ..., 1), whose two partners form the same sum, so
// every lane ends with the same value and the result does not depend on
// timing.
template <int NM>
__global__ void __launch_bounds__(64) icia_reduce_kernel(const double* __restrict__ partial, int nblocks,
                                                         float* __restrict__ moment)
{
    constexpr int UNROLL = 8;
    const int lane = int(threadIdx.x);
    const double* p = partial + int64_t(blockIdx.x) * nblocks;
    double s = 0.0;
    for (int b0 = 0; b0 < nblocks; b0 += 64 * UNROLL) {
        double x[UNROLL];
#pragma unroll
        for (int j = 0; j < UNROLL; ++j) {
            const int b = b0 + 64 * j + lane;
            x[j] = b < nblocks ? p[b] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < UNROLL; ++j)
            if (b0 + 64 * j + lane < nblocks) s += x[j];
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m);
    if (lane == 0) moment[blockIdx.x] = float(s);
}

// Workgroups of one ICIA launch over ntiles tiles: the resident workgroups,
// or more so that none takes more than ICIA_MAX_TILES tiles.
template <class K>
int icia_grid(K kernel, int ntiles)
{
    static const int resident = mfma_resident(kernel);
    const int cap = int((int64_t(ntiles) + ICIA_MAX_TILES - 1) / ICIA_MAX_TILES);
    return std::min(ntiles, std::max(resident, cap));
}

// partial holds at least ntiles * icia_nm<D>() doubles.
template <int RC, class IN, int D>
hipError_t launch_icia(const SepArgs& a, int ntiles, double* partial, float* moment, hipStream_t s)
{
    int grid = 0;
    if (ntiles > 0) {
        grid = icia_grid(icia_mfma_kernel<RC, IN, D>, ntiles);
        hipLaunchKernelGGL((icia_mfma_kernel<RC, IN, D>), dim3(grid), dim3(64 * MWAVES), 0, s, a, ntiles, partial);
    }
    hipLaunchKernelGGL(icia_reduce_kernel<icia_nm<D>()>, dim3(icia_nm<D>()), dim3(64), 0, s, partial, grid, moment);
    return hipGetLastError();
}

using IciaLaunchFn = hipError_t (*)(const SepArgs&, int, double*, float*, hipStream_t);

extern IciaLaunchFn g_icia_table[2][4][2];          // [radius class][in][affine, projective]

}   // namespace tucu
