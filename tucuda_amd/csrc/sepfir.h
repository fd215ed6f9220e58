// sepfir.h -- fused separable FIR kernel for gfx950 (MI355X).
//
// One launch computes up to six outputs of a separable FIR from one read of
// the input image.  The reference (TU/cu/FIRFilter.h:181-231) runs a
// horizontal kernel into an FP32 intermediate image in global memory and a
// vertical kernel out of it, once per output; here the intermediate never
// leaves the workgroup.
//
// Workgroup = 256 threads = 4 wave64s, owning a tile of TH centre rows by
// TWO = 256 - 2*RH centre columns.
//   1. Vertical pass.  Thread t owns input column cx - RH + t.  It loads the
//      TH + 2*RV rows its column needs straight into registers (each load
//      instruction of a wave reads 64 consecutive pixels), and computes the
//      NF vertical fields of TH rows with the paired tap sums of
//      FIRFilter.h:66-176 (pair sums/differences shared between lobes).
//      Fields go to LDS, one 256-float row per tile row.
//   2. Horizontal pass.  A task is 4 consecutive centre columns of one tile
//      row.  The thread reads the 2*RH + 4 floats its task needs with aligned
//      ds_read_b128, slides the window in registers and writes every
//      requested output with one 4-element store per output.
// Only the vertical halo is re-read (from L2), and only the horizontal
// halo (2*RH of 256 columns) is computed twice.
//
// Coordinates are centre coordinates: output k of a pixel whose kernel
// window is centred on input (y, x) lands at (y + dy_k, x + dx_k) and is
// written only when (y, x) lies in output k's valid rectangle.  Lobes are
// passed centre-indexed (c[0] = centre tap, c[k] = taps at distance k) and
// zero-padded to the template radius, so one instantiation serves every
// radius up to RV / RH.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tucu {

enum { MODE_FIR = 0, MODE_GAUSS = 1 };
constexpr int NT = 256;        // threads per workgroup = input columns per tile
constexpr int KMAX = 17;       // centre-indexed coefficients, radius <= 16

struct bf16_t { uint16_t bits; };
using f16_t = _Float16;

struct OutDesc {
    void*   ptr;
    int64_t stride;            // elements between rows
    int     dy, dx;            // output position = centre position + (dy, dx)
    int     y0, y1, x0, x1;    // valid centre rows [y0, y1), columns [x0, x1)
};

struct SepArgs {
    const void* in;
    int64_t in_stride;
    int     nrow, ncol;
    int     cy0, cy1, cx0, cx1;    // centre region covered by the grid
    int     tiles_y;               // tiles per column strip
    unsigned mask;                 // FIR: bit 0.  GAUSS: smooth, diffH, diffV, diffHH, diffHV, diffVV
    float   sgnv, sgnh;            // FIR: +1 symmetric lobe, -1 antisymmetric lobe
    float   cv[3][KMAX];           // vertical lobes (centre-indexed, zero padded)
    float   ch[3][KMAX];           // horizontal lobes
    OutDesc out[6];
};

// ----------------------------------------------------------------- pixels
template <class T> struct Px;

template <> struct Px<float> {
    static __device__ __forceinline__ float ld(const float* p) { return *p; }
    static __device__ __forceinline__ void st(float* p, float v) { *p = v; }
    static __device__ __forceinline__ void st4(float* p, float a, float b, float c, float d)
    {
        *reinterpret_cast<float4*>(p) = make_float4(a, b, c, d);
    }
};

template <> struct Px<uint8_t> {
    static __device__ __forceinline__ float ld(const uint8_t* p) { return float(*p); }
    // float -> u_char: saturate to [0, 255], truncate toward zero, NaN -> 0.
    static __device__ __forceinline__ uint32_t cvt(float v)
    {
        return uint32_t(fminf(fmaxf(v, 0.f), 255.f));
    }
    static __device__ __forceinline__ void st(uint8_t* p, float v) { *p = uint8_t(cvt(v)); }
    static __device__ __forceinline__ void st4(uint8_t* p, float a, float b, float c, float d)
    {
        *reinterpret_cast<uint32_t*>(p) = cvt(a) | cvt(b) << 8 | cvt(c) << 16 | cvt(d) << 24;
    }
};

template <> struct Px<f16_t> {
    static __device__ __forceinline__ float ld(const f16_t* p) { return float(*p); }
    static __device__ __forceinline__ uint32_t bits(float v)
    {
        const f16_t h = f16_t(v);      // round to nearest even
        return uint32_t(__builtin_bit_cast(uint16_t, h));
    }
    static __device__ __forceinline__ void st(f16_t* p, float v) { *p = f16_t(v); }
    static __device__ __forceinline__ void st4(f16_t* p, float a, float b, float c, float d)
    {
        *reinterpret_cast<uint2*>(p) = make_uint2(bits(a) | bits(b) << 16, bits(c) | bits(d) << 16);
    }
};

template <> struct Px<bf16_t> {
    static __device__ __forceinline__ float ld(const bf16_t* p) { return __uint_as_float(uint32_t(p->bits) << 16); }
    // round to nearest even; NaN stays a quiet NaN
    static __device__ __forceinline__ uint32_t bits(float v)
    {
        uint32_t x = __float_as_uint(v);
        if ((x & 0x7fffffffu) > 0x7f800000u) return (x >> 16) | 0x40u;
        x += 0x7fffu + ((x >> 16) & 1u);
        return x >> 16;
    }
    static __device__ __forceinline__ void st(bf16_t* p, float v) { p->bits = uint16_t(bits(v)); }
    static __device__ __forceinline__ void st4(bf16_t* p, float a, float b, float c, float d)
    {
        *reinterpret_cast<uint2*>(p) = make_uint2(bits(a) | bits(b) << 16, bits(c) | bits(d) << 16);
    }
};

// Writes the 4 results of centre columns xc .. xc+3 of centre row y into
// output o, each only inside o's valid rectangle.
template <class OUT> __device__ __forceinline__ void
store4(const OutDesc& o, int y, int xc, float v0, float v1, float v2, float v3)
{
    if (y < o.y0 || y >= o.y1) return;
    OUT* p = static_cast<OUT*>(o.ptr) + (int64_t(y + o.dy) * o.stride + int64_t(xc + o.dx));
    if (xc >= o.x0 && xc + 3 < o.x1 && (reinterpret_cast<uintptr_t>(p) & (4 * sizeof(OUT) - 1)) == 0) {
        Px<OUT>::st4(p, v0, v1, v2, v3);
        return;
    }
    if (xc     >= o.x0 && xc     < o.x1) Px<OUT>::st(p,     v0);
    if (xc + 1 >= o.x0 && xc + 1 < o.x1) Px<OUT>::st(p + 1, v1);
    if (xc + 2 >= o.x0 && xc + 2 < o.x1) Px<OUT>::st(p + 2, v2);
    if (xc + 3 >= o.x0 && xc + 3 < o.x1) Px<OUT>::st(p + 3, v3);
}

// Symmetric lobe: sum_k c[k] * s[k] (k = R .. 1, outermost first, as the
// reference orders its taps), then the centre tap.
template <int R> __device__ __forceinline__ float
sym_sum(const float* c, const float* s, float centre)
{
    float acc = c[R] * s[R];
#pragma unroll
    for (int k = R - 1; k >= 1; --k) acc = __builtin_fmaf(c[k], s[k], acc);
    return __builtin_fmaf(c[0], centre, acc);
}

// Antisymmetric lobe: sum_k c[k] * d[k] with d[k] = (left or top) - (right or bottom).
template <int R> __device__ __forceinline__ float
anti_sum(const float* c, const float* d)
{
    float acc = c[R] * d[R];
#pragma unroll
    for (int k = R - 1; k >= 1; --k) acc = __builtin_fmaf(c[k], d[k], acc);
    return acc;
}

// Workgroups are dispatched round-robin over the 8 XCDs.  Give each XCD a
// contiguous range of tile numbers so that vertically adjacent tiles, which
// read the same halo rows, share one L2.  A bijection of [0, n).
__device__ __forceinline__ int xcd_remap(int b, int n)
{
    const int q = n >> 3, rem = n & 7;
    const int xcd = b & 7, slot = b >> 3;
    return xcd < rem ? xcd * (q + 1) + slot : rem * (q + 1) + (xcd - rem) * q + slot;
}

template <int RV, int RH, int TH, int MODE, class IN, class OUT>
__global__ void __launch_bounds__(NT)
sepfir_kernel(const SepArgs a)
{
    constexpr int NF    = MODE == MODE_GAUSS ? 3 : 1;
    constexpr int TWO   = NT - 2 * RH;          // centre columns per tile
    constexpr int PITCH = NT + 4;               // LDS row pitch (floats), keeps 16-byte rows
    constexpr int NWIN  = TH + 2 * RV;          // input rows per column
    constexpr int NRUN  = TWO / 4;              // 4-column tasks per tile row
    constexpr int HW    = 2 * RH + 4;           // window of one task
    static_assert(TWO % 4 == 0 && HW % 4 == 0, "tile geometry");
    static_assert(RV <= 16 && RH <= 16, "radius");

    __shared__ float4 lds4[NF * TH * PITCH / 4];
    float* const lds = reinterpret_cast<float*>(lds4);

    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const int tx = tile / a.tiles_y;
    const int ty = tile - tx * a.tiles_y;
    const int cx = a.cx0 + tx * TWO;
    const int cy = a.cy0 + ty * TH;
    const int t  = threadIdx.x;

    const bool needF0 = MODE == MODE_FIR || (a.mask & 0x0bu);
    const bool needF1 = MODE == MODE_GAUSS && (a.mask & 0x14u);
    const bool needF2 = MODE == MODE_GAUSS && (a.mask & 0x20u);

    // ------------------------------------------------------ vertical pass
    {
        const int  xi    = cx - RH + t;
        const bool colok = unsigned(xi) < unsigned(a.ncol);
        const IN*  src   = static_cast<const IN*>(a.in) + (colok ? xi : 0);
        float w[NWIN];
#pragma unroll
        for (int i = 0; i < NWIN; ++i) {
            const int yi = cy - RV + i;
            w[i] = (colok && unsigned(yi) < unsigned(a.nrow)) ? Px<IN>::ld(src + int64_t(yi) * a.in_stride) : 0.f;
        }
#pragma unroll
        for (int i = 0; i < TH; ++i) {
            const float* win = w + i;          // centre at win[RV]
            if constexpr (MODE == MODE_GAUSS) {
                float s[RV + 1], d[RV + 1];
#pragma unroll
                for (int k = 1; k <= RV; ++k) {
                    s[k] = win[RV - k] + win[RV + k];
                    d[k] = win[RV - k] - win[RV + k];
                }
                if (needF0) lds[(0 * TH + i) * PITCH + t] = sym_sum<RV>(a.cv[0], s, win[RV]);
                if (needF1) lds[(1 * TH + i) * PITCH + t] = anti_sum<RV>(a.cv[1], d);
                if (needF2) lds[(2 * TH + i) * PITCH + t] = sym_sum<RV>(a.cv[2], s, win[RV]);
            } else {
                float p[RV + 1];
#pragma unroll
                for (int k = 1; k <= RV; ++k) p[k] = __builtin_fmaf(a.sgnv, win[RV + k], win[RV - k]);
                lds[i * PITCH + t] = sym_sum<RV>(a.cv[0], p, win[RV]);
            }
        }
    }
    __syncthreads();

    // ---------------------------------------------------- horizontal pass
    for (int task = t; task < TH * NRUN; task += NT) {
        const int i  = task / NRUN;
        const int j  = task - i * NRUN;
        const int y  = cy + i;
        const int xc = cx + 4 * j;
        if (y >= a.cy1 || xc >= a.cx1) continue;

#pragma unroll
        for (int f = 0; f < NF; ++f) {
            if (f == 0 && !needF0) continue;
            if (f == 1 && !needF1) continue;
            if (f == 2 && !needF2) continue;
            float v[HW];
            const float4* row = reinterpret_cast<const float4*>(lds + (f * TH + i) * PITCH + 4 * j);
#pragma unroll
            for (int m = 0; m < HW / 4; ++m) {
                const float4 q = row[m];
                v[4 * m] = q.x; v[4 * m + 1] = q.y; v[4 * m + 2] = q.z; v[4 * m + 3] = q.w;
            }
            if constexpr (MODE == MODE_GAUSS) {
                // field 0 -> smooth (h0), diffH (h1), diffHH (h2)
                // field 1 -> diffV (h0), diffHV (h1)
                // field 2 -> diffVV (h0)
                const unsigned m0 = f == 0 ? 1u << 0 : f == 1 ? 1u << 2 : 1u << 5;
                const unsigned m1 = f == 0 ? 1u << 1 : f == 1 ? 1u << 4 : 0u;
                const unsigned m2 = f == 0 ? 1u << 3 : 0u;
                const int o0 = f == 0 ? 0 : f == 1 ? 2 : 5;
                const int o1 = f == 0 ? 1 : 4;
                const bool want0 = a.mask & m0, want1 = a.mask & m1, want2 = a.mask & m2;
                float r0[4], r1[4], r2[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float* win = v + q;   // centre at win[RH]
                    float s[RH + 1], d[RH + 1];
#pragma unroll
                    for (int k = 1; k <= RH; ++k) {
                        s[k] = win[RH - k] + win[RH + k];
                        d[k] = win[RH - k] - win[RH + k];
                    }
                    r0[q] = want0 ? sym_sum<RH>(a.ch[0], s, win[RH]) : 0.f;
                    r1[q] = want1 ? anti_sum<RH>(a.ch[1], d) : 0.f;
                    r2[q] = want2 ? sym_sum<RH>(a.ch[2], s, win[RH]) : 0.f;
                }
                if (want0) store4<OUT>(a.out[o0], y, xc, r0[0], r0[1], r0[2], r0[3]);
                if (want1) store4<OUT>(a.out[o1], y, xc, r1[0], r1[1], r1[2], r1[3]);
                if (want2) store4<OUT>(a.out[3], y, xc, r2[0], r2[1], r2[2], r2[3]);
            } else {
                float r[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float* win = v + q;
                    float p[RH + 1];
#pragma unroll
                    for (int k = 1; k <= RH; ++k) p[k] = __builtin_fmaf(a.sgnh, win[RH + k], win[RH - k]);
                    r[q] = sym_sum<RH>(a.ch[0], p, win[RH]);
                }
                store4<OUT>(a.out[0], y, xc, r[0], r[1], r[2], r[3]);
            }
        }
    }
}

// Host-side launch entry of one instantiation.
using LaunchFn = hipError_t (*)(const SepArgs&, int nblocks, hipStream_t);

struct KernelEntry {
    LaunchFn fn;
    int      th;     // tile rows
    int      rv, rh; // template radii
};

template <int RV, int RH, int TH, int MODE, class IN, class OUT>
hipError_t launch_sepfir(const SepArgs& a, int nblocks, hipStream_t s)
{
    hipLaunchKernelGGL((sepfir_kernel<RV, RH, TH, MODE, IN, OUT>), dim3(nblocks), dim3(NT), 0, s, a);
    return hipGetLastError();
}

// Radius classes of the instantiated kernels.
constexpr int FIR_RADII[4]   = {2, 4, 8, 16};
constexpr int GAUSS_RADII[5] = {2, 4, 8, 12, 16};
constexpr int FIR_TH   = 32;
constexpr int GAUSS_TH = 16;

// Tables filled by the instantiation units: [rv class][rh class][in dtype][out dtype]
// and [radius class][in dtype][out dtype]; dtype order u8, f32, f16, bf16.
extern KernelEntry g_fir_table[4][4][4][4];
extern KernelEntry g_gauss_table[5][4][4];

}   // namespace tucu
