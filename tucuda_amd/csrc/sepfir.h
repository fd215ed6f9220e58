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
//      TH + 2*RV rows its column needs straight into registers with buffer
//      loads whose row offsets live in SGPRs (each load instruction of a wave
//      reads 64 consecutive pixels; no per-load address arithmetic), and
//      computes the NF vertical fields of TH rows with the paired tap sums of
//      FIRFilter.h:66-176, pair sums / differences shared between lobes.
//      Fields go to LDS, one 256-float row per tile row.
//   2. Horizontal pass.  A task is 4 consecutive centre columns of one tile
//      row.  The thread reads the 2*RH + 4 floats its task needs with aligned
//      ds_read_b128, slides the window in registers and writes every
//      requested output with one 4-element store per output.
// Only the vertical halo is re-read (from L2), and only the horizontal halo
// (2*RH of 256 columns) is computed twice.
//
// Coordinates are centre coordinates: output k of a pixel whose kernel
// window is centred on input (y, x) lands at (y, x) + (dy_k, dx_k), with
// (dy_k, dx_k) = shift ? (0, 0) : (-rv_k, -rh_k), and is written only when
// rv_k <= y < nrow - rv_k and rh_k <= x < ncol - rh_k (FIRFilter.h:37-40,
// 286-287).  Lobes are passed centre-indexed (c[0] = centre tap, c[k] =
// taps at distance k) and zero-padded to the template radius, so one
// instantiation serves every radius up to RV / RH.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tucu {

// MODE_ICIA: the fused diffH / diffV pass of ICIA's error moment; the two
// gradients go to the caller's sink instead of memory (mfma.h).
enum { MODE_FIR = 0, MODE_GAUSS = 1, MODE_TAPS = 2, MODE_ICIA = 3 };
constexpr int TMAX = 33;      // full taps, radius <= 16
constexpr int NT = 256;        // threads per workgroup = input columns per tile
constexpr int KMAX = 17;       // centre-indexed coefficients, radius <= 16
constexpr int BUF_DWORD3 = 0x00020000;   // raw buffer resource word 3 on gfx9-family parts

struct bf16_t { uint16_t bits; };
using f16_t = _Float16;
template <class T> using gptr = T __attribute__((address_space(1)))*;   // global memory
typedef float    v4f32 __attribute__((ext_vector_type(4)));
typedef float    v2f32 __attribute__((ext_vector_type(2)));
typedef uint32_t v2u32 __attribute__((ext_vector_type(2)));
typedef _Float16 v2f16 __attribute__((ext_vector_type(2)));
typedef __bf16   v2bf16 __attribute__((ext_vector_type(2)));

// Two floats to two packed 16-bit values, round to nearest even
// (v_cvt_pk_f16_f32 / v_cvt_pk_bf16_f32 on gfx950).
__device__ __forceinline__ uint32_t pack_f16(float a, float b)
{
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((v2f32){a, b}, v2f16));
}
__device__ __forceinline__ uint32_t pack_bf16(float a, float b)
{
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((v2f32){a, b}, v2bf16));
}

struct SepArgs {
    const void* in;
    int32_t  in_stride_b;          // bytes between input rows
    uint32_t in_bytes;             // bytes from row 0 to the end of the last row
    int      nrow, ncol;
    int      cy0, cy1, cx0, cx1;   // centre region covered by the grid
    int      fy0, fy1, fx0, fx1;   // centre rectangle on which every live output is valid and
                                   // takes aligned 4-element stores (empty if stores are unaligned)
    int      tiles_y;              // tiles per column strip
    unsigned mask;                 // FIR: bit 0.  GAUSS: smooth, diffH, diffV, diffHH, diffHV, diffVV
    int      r[3];                 // GAUSS: radius of lobe 0, 1, 2.  FIR: r[0] vertical, r[1] horizontal
    float    sgnv, sgnh;           // FIR: +1 symmetric lobe, -1 antisymmetric lobe
    float    cv[3][KMAX];          // vertical lobes (centre-indexed, zero padded); GAUSS: both passes
    float    ch[KMAX];             // FIR: horizontal lobe
    float    tv[TMAX], th[TMAX];   // TAPS: full vertical / horizontal taps, tap k at offset k - R
    int      rt;                   // template radius class the tile grid was laid out for (batches)
    float    mp[3], mq;            // matrix-core kernels: power-of-two scales of each lobe's taps
    float    mu[6];                // and of the input, and the factor that undoes them per output
    uintptr_t obase[6];            // output k: address of the element for centre (0, 0), i.e.
                                   // data + (dy_k * stride + dx_k) * elem (may lie before data)
    int32_t  ostride[6];           // output row stride, elements
    const void* tabs;              // matrix-core kernels: the operand tables (MfmaTabs) prebuilt in
                                   // device memory, or null: each workgroup builds them
};

// Job descriptors read by the batched kernel live in device memory the
// kernel never writes; reading them through the constant address space lets
// the compiler treat every field as invariant (scalar loads, hoisted out of
// the tile loops), as it does for kernel arguments.
typedef const __attribute__((address_space(4))) SepArgs ConstArgs;

// Cache policies of the matrix-core kernels' output stores (buffer
// instruction bits: nt = 2, sc1 = 16).  CPOL_NT: streaming.  CPOL_NT_WT:
// streaming and at agent scope, so written through the XCD's L2: the
// kernel-end release then finds no dirty output lines to write back, which
// takes 4 % off a one-round launch (1080p six outputs) but costs 3-7 % in
// the tile loops, where the L2 absorbs the stores of many rounds.
constexpr int CPOL_NT = 2;
constexpr int CPOL_NT_WT = 2 | 16;

typedef uint32_t v4u32 __attribute__((ext_vector_type(4)));

// ----------------------------------------------------------------- pixels
template <class T> struct Px;

template <> struct Px<float> {
    static __device__ __forceinline__ float ld(__amdgpu_buffer_rsrc_t r, int voff, int soff)
    {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
    }
    static __device__ __forceinline__ void st(gptr<float> p, float v) { *p = v; }
    static __device__ __forceinline__ void st4(gptr<float> p, float a, float b, float c, float d)
    {
        *reinterpret_cast<gptr<v4f32>>(p) = v4f32{a, b, c, d};
    }
    static __device__ __forceinline__ void st4nt(gptr<float> p, float a, float b, float c, float d)
    {
        __builtin_nontemporal_store(v4f32{a, b, c, d}, reinterpret_cast<gptr<v4f32>>(p));
    }
    template <int CP>
    static __device__ __forceinline__ void st4buf(__amdgpu_buffer_rsrc_t r, uint32_t off, float a, float b, float c, float d)
    {
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(v4u32, v4f32{a, b, c, d}), r, int(off), 0, CP);
    }
};

template <> struct Px<uint8_t> {
    static __device__ __forceinline__ float ld(__amdgpu_buffer_rsrc_t r, int voff, int soff)
    {
        return float(__builtin_amdgcn_raw_buffer_load_b8(r, voff, soff, 0));
    }
    // float -> u_char: saturate to [0, 255], truncate toward zero, NaN -> 0.
    static __device__ __forceinline__ uint32_t cvt(float v)
    {
        return uint32_t(fminf(fmaxf(v, 0.f), 255.f));
    }
    static __device__ __forceinline__ void st(gptr<uint8_t> p, float v) { *p = uint8_t(cvt(v)); }
    static __device__ __forceinline__ void st4(gptr<uint8_t> p, float a, float b, float c, float d)
    {
        *reinterpret_cast<gptr<uint32_t>>(p) = cvt(a) | cvt(b) << 8 | cvt(c) << 16 | cvt(d) << 24;
    }
    static __device__ __forceinline__ void st4nt(gptr<uint8_t> p, float a, float b, float c, float d)
    {
        __builtin_nontemporal_store(cvt(a) | cvt(b) << 8 | cvt(c) << 16 | cvt(d) << 24, reinterpret_cast<gptr<uint32_t>>(p));
    }
    template <int CP>
    static __device__ __forceinline__ void st4buf(__amdgpu_buffer_rsrc_t r, uint32_t off, float a, float b, float c, float d)
    {
        __builtin_amdgcn_raw_buffer_store_b32(cvt(a) | cvt(b) << 8 | cvt(c) << 16 | cvt(d) << 24, r, int(off), 0, CP);
    }
};

template <> struct Px<f16_t> {
    static __device__ __forceinline__ float ld(__amdgpu_buffer_rsrc_t r, int voff, int soff)
    {
        return float(__builtin_bit_cast(f16_t, __builtin_amdgcn_raw_buffer_load_b16(r, voff, soff, 0)));
    }
    static __device__ __forceinline__ void st(gptr<f16_t> p, float v) { *p = f16_t(v); }
    static __device__ __forceinline__ void st4(gptr<f16_t> p, float a, float b, float c, float d)
    {
        *reinterpret_cast<gptr<v2u32>>(p) = v2u32{pack_f16(a, b), pack_f16(c, d)};
    }
    static __device__ __forceinline__ void st4nt(gptr<f16_t> p, float a, float b, float c, float d)
    {
        __builtin_nontemporal_store(v2u32{pack_f16(a, b), pack_f16(c, d)}, reinterpret_cast<gptr<v2u32>>(p));
    }
    template <int CP>
    static __device__ __forceinline__ void st4buf(__amdgpu_buffer_rsrc_t r, uint32_t off, float a, float b, float c, float d)
    {
        __builtin_amdgcn_raw_buffer_store_b64(v2u32{pack_f16(a, b), pack_f16(c, d)}, r, int(off), 0, CP);
    }
};

template <> struct Px<bf16_t> {
    static __device__ __forceinline__ float ld(__amdgpu_buffer_rsrc_t r, int voff, int soff)
    {
        return __uint_as_float(uint32_t(__builtin_amdgcn_raw_buffer_load_b16(r, voff, soff, 0)) << 16);
    }
    static __device__ __forceinline__ void st(gptr<bf16_t> p, float v)
    {
        p->bits = __builtin_bit_cast(uint16_t, __bf16(v));
    }
    static __device__ __forceinline__ void st4(gptr<bf16_t> p, float a, float b, float c, float d)
    {
        *reinterpret_cast<gptr<v2u32>>(p) = v2u32{pack_bf16(a, b), pack_bf16(c, d)};
    }
    static __device__ __forceinline__ void st4nt(gptr<bf16_t> p, float a, float b, float c, float d)
    {
        __builtin_nontemporal_store(v2u32{pack_bf16(a, b), pack_bf16(c, d)}, reinterpret_cast<gptr<v2u32>>(p));
    }
    template <int CP>
    static __device__ __forceinline__ void st4buf(__amdgpu_buffer_rsrc_t r, uint32_t off, float a, float b, float c, float d)
    {
        __builtin_amdgcn_raw_buffer_store_b64(v2u32{pack_bf16(a, b), pack_bf16(c, d)}, r, int(off), 0, CP);
    }
};

// Results of centre columns xc .. xc+3 of centre row y for output k, whose
// valid centre rectangle is rows [vy0, vy1) x columns [vx0, vx1).
// fast: the caller has established that all four are valid and aligned.
template <class OUT, class A> __device__ __forceinline__ void
store4(const A& a, int k, bool fast, int vy0, int vy1, int vx0, int vx1,
       int y, int xc, float v0, float v1, float v2, float v3)
{
    const gptr<OUT> q = reinterpret_cast<gptr<OUT>>(a.obase[k] + uintptr_t(int64_t(y) * a.ostride[k] + xc) * sizeof(OUT));
    if (fast) {                                 // outputs are written once: streaming stores
        Px<OUT>::st4nt(q, v0, v1, v2, v3);
        return;
    }
    if (y < vy0 || y >= vy1) return;
    if (xc >= vx0 && xc + 3 < vx1 && (reinterpret_cast<uintptr_t>(q) & (4 * sizeof(OUT) - 1)) == 0) {
        Px<OUT>::st4(q, v0, v1, v2, v3);
        return;
    }
    if (xc     >= vx0 && xc     < vx1) Px<OUT>::st(q,     v0);
    if (xc + 1 >= vx0 && xc + 1 < vx1) Px<OUT>::st(q + 1, v1);
    if (xc + 2 >= vx0 && xc + 2 < vx1) Px<OUT>::st(q + 2, v2);
    if (xc + 3 >= vx0 && xc + 3 < vx1) Px<OUT>::st(q + 3, v3);
}

// Symmetric lobe: sum_k c[k] * s[k] (k = R .. 1, outermost first, as the
// reference orders its taps), then the centre tap.
template <int R, class P> __device__ __forceinline__ float
sym_sum(P c, const float* s, float centre)
{
    float acc = c[R] * s[R];
#pragma unroll
    for (int k = R - 1; k >= 1; --k) acc = __builtin_fmaf(c[k], s[k], acc);
    return __builtin_fmaf(c[0], centre, acc);
}

// Antisymmetric lobe: sum_k c[k] * d[k] with d[k] = (left or top) - (right or bottom).
template <int R, class P> __device__ __forceinline__ float
anti_sum(P c, const float* d)
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

// Vertical pass of one tile: NWIN input rows of column xi into w[].
template <int NWIN, class IN, class A> __device__ __forceinline__ void
load_column(const A& a, __amdgpu_buffer_rsrc_t rsrc, int xi, int y_first, float* w)
{
    const bool colok = unsigned(xi) < unsigned(a.ncol);
    const int voff = (colok ? xi : 0) * int(sizeof(IN));
    const bool rows_ok = y_first >= 0 && y_first + NWIN <= a.nrow;       // uniform
    if (rows_ok) {
#pragma unroll
        for (int i = 0; i < NWIN; ++i) w[i] = Px<IN>::ld(rsrc, voff, (y_first + i) * a.in_stride_b);
    } else {
#pragma unroll
        for (int i = 0; i < NWIN; ++i) {
            const int yi = y_first + i;
            w[i] = unsigned(yi) < unsigned(a.nrow) ? Px<IN>::ld(rsrc, voff, yi * a.in_stride_b) : 0.f;
        }
    }
    if (!colok) {
#pragma unroll
        for (int i = 0; i < NWIN; ++i) w[i] = 0.f;
    }
}

// ---------------------------------------------------------------- passes
// Vertical pass, symmetric fields: field 0 (lobe cv[0]) and / or field 2
// (lobe cv[2]) from the pair sums s[k] = w[y-k] + w[y+k].  FIR mode uses
// field 0 with the runtime parity sgnv (s[k] = w[y-k] + sgnv * w[y+k]).
// Tap-outer loop: every row keeps its own accumulator, summed over
// k = RV .. 1 then the centre, the order sym_sum uses.
template <int RV, int TH, int PITCH, bool F0, bool F2, bool FIR, class A>
__device__ __forceinline__ void
vpass_sym(const A& a, const float* w, float* lds, int t)
{
    float acc0[TH], acc2[TH];
#pragma unroll
    for (int k = RV; k >= 1; --k) {
#pragma unroll
        for (int i = 0; i < TH; ++i) {
            const float s = FIR ? __builtin_fmaf(a.sgnv, w[i + RV + k], w[i + RV - k])
                                : w[i + RV - k] + w[i + RV + k];
            if (F0) acc0[i] = k == RV ? a.cv[0][k] * s : __builtin_fmaf(a.cv[0][k], s, acc0[i]);
            if (F2) acc2[i] = k == RV ? a.cv[2][k] * s : __builtin_fmaf(a.cv[2][k], s, acc2[i]);
        }
    }
    // An antisymmetric FIR lobe (sgnv < 0) has no centre tap: its centre
    // pixel is not multiplied at all (a zero weight would turn an Inf there
    // into NaN), as FIRFilter.h:66-176 leaves it out.
    const bool centre = !FIR || a.sgnv > 0.f;
#pragma unroll
    for (int i = 0; i < TH; ++i) {
        if (F0) lds[(0 * TH + i) * PITCH + t] = centre ? __builtin_fmaf(a.cv[0][0], w[i + RV], acc0[i]) : acc0[i];
        if (F2) lds[(2 * TH + i) * PITCH + t] = __builtin_fmaf(a.cv[2][0], w[i + RV], acc2[i]);
    }
}

// Vertical pass, antisymmetric field 1 (lobe cv[1]) from d[k] = w[y-k] - w[y+k].
template <int RV, int TH, int PITCH, class A>
__device__ __forceinline__ void
vpass_anti(const A& a, const float* w, float* lds, int t)
{
    float acc[TH];
#pragma unroll
    for (int k = RV; k >= 1; --k) {
#pragma unroll
        for (int i = 0; i < TH; ++i) {
            const float d = w[i + RV - k] - w[i + RV + k];
            acc[i] = k == RV ? a.cv[1][k] * d : __builtin_fmaf(a.cv[1][k], d, acc[i]);
        }
    }
#pragma unroll
    for (int i = 0; i < TH; ++i) lds[(1 * TH + i) * PITCH + t] = acc[i];
}

// Vertical pass with general taps (no symmetry): field 0 = sum_k tv[k] *
// w[y - RV + k], k = 0 .. 2 RV, top tap first.
template <int RV, int TH, int PITCH, class A>
__device__ __forceinline__ void
vpass_taps(const A& a, const float* w, float* lds, int t)
{
    float acc[TH];
#pragma unroll
    for (int k = 0; k <= 2 * RV; ++k) {
#pragma unroll
        for (int i = 0; i < TH; ++i)
            acc[i] = k == 0 ? a.tv[0] * w[i] : __builtin_fmaf(a.tv[k], w[i + k], acc[i]);
    }
#pragma unroll
    for (int i = 0; i < TH; ++i) lds[i * PITCH + t] = acc[i];
}

// Horizontal pass of field F over the tile: every output the field feeds.
//   GAUSS field 0 -> smooth (lobe 0), diffH (lobe 1), diffHH (lobe 2)
//         field 1 -> diffV (lobe 0), diffHV (lobe 1)
//         field 2 -> diffVV (lobe 0)
//   FIR   field 0 -> output 0 (lobe ch, runtime parity sgnh)
// FAST: every store of the tile is valid and aligned (no checks).
template <int RH, int TH, int PITCH, int TWO, int F, bool FIR, bool FAST, class OUT, bool TAPS, class A>
__device__ __forceinline__ void
hpass_field(const A& a, const float* lds, int t, int cy, int cx)
{
    constexpr int NRUN = TWO / 4;
    constexpr int HW   = 2 * RH + 4;
    constexpr bool H1  = !FIR && F <= 1;        // field feeds an antisymmetric horizontal lobe
    constexpr bool H2  = !FIR && F == 0;        // field feeds horizontal lobe 2
    constexpr int  O0  = FIR ? 0 : F == 0 ? 0 : F == 1 ? 2 : 5;
    constexpr int  O1  = F == 0 ? 1 : 4;
    constexpr int  O2  = 3;
    const unsigned mask = a.mask;
    const bool st0 = FIR || (mask >> O0 & 1u);
    const bool st1 = H1 && (mask >> O1 & 1u);
    const bool st2 = H2 && (mask >> O2 & 1u);
    // Output (h, v) is valid on centre rows [r_v, nrow - r_v), columns [r_h, ncol - r_h).
    const int rvf = FIR ? a.r[0] : a.r[F];
    const int vy0 = rvf, vy1 = a.nrow - rvf;
    const int rh0 = FIR ? a.r[1] : a.r[0];

    for (int task = t; task < TH * NRUN; task += NT) {
        const int i  = task / NRUN;
        const int j  = task - i * NRUN;
        const int y  = cy + i;
        const int xc = cx + 4 * j;
        if (!FAST && (y >= a.cy1 || xc >= a.cx1)) continue;
        float v[HW];
        const float4* row = reinterpret_cast<const float4*>(lds + (F * TH + i) * PITCH + 4 * j);
#pragma unroll
        for (int m = 0; m < HW / 4; ++m) {
            const float4 q = row[m];
            v[4 * m] = q.x; v[4 * m + 1] = q.y; v[4 * m + 2] = q.z; v[4 * m + 3] = q.w;
        }
        float r0[4], r1[4], r2[4];
        if constexpr (TAPS) {
            // general taps, left tap first
#pragma unroll
            for (int k = 0; k <= 2 * RH; ++k)
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    r0[q] = k == 0 ? a.th[0] * v[q] : __builtin_fmaf(a.th[k], v[q + k], r0[q]);
        } else {
        // Tap-outer: 4 columns x up to 3 lobes of running sums, each summed
        // over k = RH .. 1 then the centre (the order sym_sum uses).
        const auto c0 = FIR ? a.ch : a.cv[0];
#pragma unroll
        for (int k = RH; k >= 1; --k) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float lo = v[q + RH - k], hi = v[q + RH + k];
                const float sk = FIR ? __builtin_fmaf(a.sgnh, hi, lo) : lo + hi;
                r0[q] = k == RH ? c0[k] * sk : __builtin_fmaf(c0[k], sk, r0[q]);
                if (H1) {
                    const float dk = lo - hi;
                    r1[q] = k == RH ? a.cv[1][k] * dk : __builtin_fmaf(a.cv[1][k], dk, r1[q]);
                }
                if (H2) r2[q] = k == RH ? a.cv[2][k] * sk : __builtin_fmaf(a.cv[2][k], sk, r2[q]);
            }
            // Keep each tap's pair sums next to their use: hoisting all of
            // them costs 64 registers and occupancy.
            __builtin_amdgcn_sched_barrier(0);
        }
        const bool centre = !FIR || a.sgnh > 0.f;       // no centre tap for an antisymmetric FIR lobe
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            r0[q] = centre ? __builtin_fmaf(c0[0], v[q + RH], r0[q]) : r0[q];
            if (H2) r2[q] = __builtin_fmaf(a.cv[2][0], v[q + RH], r2[q]);
        }
        }
        if (st0) store4<OUT>(a, O0, FAST, vy0, vy1, rh0, a.ncol - rh0, y, xc, r0[0], r0[1], r0[2], r0[3]);
        if (H1 && st1) store4<OUT>(a, O1, FAST, vy0, vy1, a.r[1], a.ncol - a.r[1], y, xc, r1[0], r1[1], r1[2], r1[3]);
        if (H2 && st2) store4<OUT>(a, O2, FAST, vy0, vy1, a.r[2], a.ncol - a.r[2], y, xc, r2[0], r2[1], r2[2], r2[3]);
    }
}

template <int RH, int TH, int PITCH, int TWO, int F, bool FIR, class OUT, bool TAPS = false, class A>
__device__ __forceinline__ void
hpass(const A& a, const float* lds, int t, int cy, int cx, bool fast)
{
    if (fast) hpass_field<RH, TH, PITCH, TWO, F, FIR, true, OUT, TAPS>(a, lds, t, cy, cx);
    else      hpass_field<RH, TH, PITCH, TWO, F, FIR, false, OUT, TAPS>(a, lds, t, cy, cx);
}

// Shared-memory floats one tile needs.
template <int TH, int MODE>
constexpr int tile_lds_floats() { return (MODE == MODE_GAUSS ? 3 : 1) * TH * (NT + 4); }

// One tile (number `tile` of the image described by `a`, strip-major) of the
// fused separable filter.  `lds` holds tile_lds_floats<TH, MODE>() floats,
// 16-byte aligned.  Every thread of the workgroup calls it.
template <int RV, int RH, int TH, int MODE, class IN, class OUT, class A>
__device__ __forceinline__ void
sepfir_tile(const A& a, int tile, float* lds)
{
    constexpr bool TAPS = MODE == MODE_TAPS;
    constexpr bool FIR  = MODE == MODE_FIR || TAPS;     // one field, one output
    constexpr int TWO   = NT - 2 * RH;          // centre columns per tile
    constexpr int PITCH = NT + 4;               // LDS row pitch (floats), keeps 16-byte rows
    constexpr int NWIN  = TH + 2 * RV;          // input rows per column
    static_assert(TWO % 4 == 0 && (2 * RH + 4) % 4 == 0, "tile geometry");
    static_assert(RV <= 16 && RH <= 16, "radius");

    const int tx = tile / a.tiles_y;
    const int ty = tile - tx * a.tiles_y;
    const int cx = a.cx0 + tx * TWO;
    const int cy = a.cy0 + ty * TH;
    const int t  = threadIdx.x;

    // Fields the live outputs need (uniform).
    const unsigned mask = a.mask;
    const bool needF0 = FIR || (mask & 0x0bu);
    const bool needF1 = !FIR && (mask & 0x14u);
    const bool needF2 = !FIR && (mask & 0x20u);

    // ------------------------------------------------------ vertical pass
    {
        const __amdgpu_buffer_rsrc_t rsrc =
            __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(a.in), short(0), int(a.in_bytes), BUF_DWORD3);
        float w[NWIN];
        load_column<NWIN, IN>(a, rsrc, cx - RH + t, cy - RV, w);
        if constexpr (TAPS) {
            vpass_taps<RV, TH, PITCH>(a, w, lds, t);
        } else if constexpr (FIR) {
            vpass_sym<RV, TH, PITCH, true, false, true>(a, w, lds, t);
        } else {
            if (needF0 && needF2)  vpass_sym<RV, TH, PITCH, true, true, false>(a, w, lds, t);
            else if (needF0)       vpass_sym<RV, TH, PITCH, true, false, false>(a, w, lds, t);
            else if (needF2)       vpass_sym<RV, TH, PITCH, false, true, false>(a, w, lds, t);
            if (needF1)            vpass_anti<RV, TH, PITCH>(a, w, lds, t);
        }
    }
    __syncthreads();

    // ---------------------------------------------------- horizontal pass
    // Uniform: is every store of this tile valid and vector-aligned?
    const bool fast = cy >= a.fy0 && cy + TH <= a.fy1 && cx >= a.fx0 && cx + TWO <= a.fx1;
    if constexpr (FIR) {
        hpass<RH, TH, PITCH, TWO, 0, true, OUT, TAPS>(a, lds, t, cy, cx, fast);
    } else {
        if (needF0) hpass<RH, TH, PITCH, TWO, 0, false, OUT>(a, lds, t, cy, cx, fast);
        if (needF1) hpass<RH, TH, PITCH, TWO, 1, false, OUT>(a, lds, t, cy, cx, fast);
        if (needF2) hpass<RH, TH, PITCH, TWO, 2, false, OUT>(a, lds, t, cy, cx, fast);
    }
}

// One image: one workgroup per tile.
template <int RV, int RH, int TH, int MODE, class IN, class OUT>
__global__ void __launch_bounds__(NT)
sepfir_kernel(const SepArgs a)
{
    __shared__ float4 lds4[tile_lds_floats<TH, MODE>() / 4];
    sepfir_tile<RV, RH, TH, MODE, IN, OUT>(a, xcd_remap(blockIdx.x, gridDim.x), reinterpret_cast<float*>(lds4));
}

// Host-side launch entry of one instantiation.
using LaunchFn = hipError_t (*)(const SepArgs&, int nblocks, hipStream_t);

// Every member has an initializer, so the tables below are constant-
// initialized before any registrar (a static constructor in another unit)
// fills them.
struct KernelEntry {
    LaunchFn fn = nullptr;
    int      th = 0;          // tile rows
    int      rv = 0, rh = 0;  // template radii
    int      tw = 0;          // tile columns (0: NT - 2 rh)
};

template <int RV, int RH, int TH, int MODE, class IN, class OUT>
hipError_t launch_sepfir(const SepArgs& a, int nblocks, hipStream_t s)
{
    hipLaunchKernelGGL((sepfir_kernel<RV, RH, TH, MODE, IN, OUT>), dim3(nblocks), dim3(NT), 0, s, a);
    return hipGetLastError();
}

// Radius classes and tile heights of the instantiated kernels.
constexpr int FIR_RADII[4]   = {2, 4, 8, 16};
constexpr int GAUSS_RADII[5] = {2, 4, 8, 12, 16};
constexpr int FIR_TH         = 16;
constexpr int NGTH           = 6;
constexpr int GAUSS_THS[NGTH] = {4, 6, 8, 10, 12, 16};
constexpr int GAUSS_TH_DEFAULT = 2;       // index into GAUSS_THS (8 rows)

// Tables filled by the instantiation units: [rv class][rh class][in dtype][out dtype]
// and [radius class][tile-height index][in dtype][out dtype]; dtype order u8, f32, f16, bf16.
// Every tile height is instantiated for the same-type f32 and bf16 pairs
// (the tuned paths); other pairs carry the default height only, and the
// host falls back to it.
extern KernelEntry g_fir_table[4][4][4][4];
extern KernelEntry g_taps_table[4][4][4][4];     // general taps, same layout as g_fir_table
extern KernelEntry g_gauss_table[5][NGTH][4][4];

}   // namespace tucu
