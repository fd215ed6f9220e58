// subsample.h -- cu::subsample(in, ie, out, op) for gfx950: halves an image
// in both directions, evaluating a small stencil operator at every second
// pixel of every second row (TU/cu/algorithm.h:183-242, operators of
// TU/cu/functional.h:58-256).
//
//   out[y][x] = op(in) at input pixel (2y, 2x),  0 <= y < nrow/2, 0 <= x < ncol/2,
//
// with each operator's own border rule: the guarded operator() of
// functional.h returns 0 wherever its stencil would leave the image.  The
// arithmetic is the operators' source expressions in FP32, evaluated in
// source order with every product and sum rounded on its own (no
// contraction), so it reproduces the CPU restatement bit for bit.
//
// The reference stages a block's input tile in shared memory; here every
// output row of a wave reads its input rows straight into registers:
//   * lane l of a wave owns output columns 2g, 2g+1 (g = global lane
//     column) and loads the four input pixels 4g .. 4g+3 of each row it
//     needs with one 4-element load (one wave load = 256 consecutive
//     pixels); the one or two pixels left of them and the one right of them
//     that 3x3 / 5x5 operators need come from the neighbouring lanes by DPP
//     wave shifts;
//   * a wave owns R output rows; their 2R-1+2RY input rows are loaded once,
//     all loads of the wave in flight together (no load in a divergent
//     branch: the load form is chosen per wave).
// A workgroup is 4 waves stacked vertically: 128 output columns x 4R output
// rows.  Loads go through a buffer resource bounded by the image, so reads
// beyond it return 0 and never fault; their values only reach outputs that
// the border rule zeroes.
#pragma once

#include "sepfir.h"

namespace tucu {

// Operator codes, in the order of functional.h (the numbering of TUCU_OP_*).
enum {
    SUB_IDENTITY = 0,   // identity1x1   functional.h:58-65
    SUB_DIFFH,          // diffH1x3      :68-81
    SUB_DIFFV,          // diffV3x1      :84-98
    SUB_DIFFHH,         // diffHH1x3     :101-114
    SUB_DIFFVV,         // diffVV3x1     :117-130
    SUB_DIFFHV,         // diffHV3x3     :133-148
    SUB_SOBELH,         // sobelH3x3     :151-167
    SUB_SOBELV,         // sobelV3x3     :170-186
    SUB_SOBELABS,       // sobelAbs3x3   :189-198
    SUB_LAPLACIAN,      // laplacian3x3  :201-210
    SUB_DET,            // det3x3        :213-226
    SUB_GAUSSIAN5,      // gaussian5x5   :229-256
    NSUBOPS
};

// Stencil half-extents (OperatorSizeY/X - 1) / 2, which are also the
// border-guard widths of each operator.
constexpr int sub_ry(int op)
{
    return op == SUB_IDENTITY || op == SUB_DIFFH || op == SUB_DIFFHH ? 0 : op == SUB_GAUSSIAN5 ? 2 : 1;
}
constexpr int sub_rx(int op)
{
    return op == SUB_IDENTITY || op == SUB_DIFFV || op == SUB_DIFFVV ? 0 : op == SUB_GAUSSIAN5 ? 2 : 1;
}

constexpr int SUB_R = 2;        // output rows per wave
constexpr int SUB_WAVES = 4;    // waves per workgroup, stacked vertically

struct SubArgs {
    const void* in;
    uint32_t  in_bytes;         // bytes from row 0 to the end of the last row
    int32_t   in_stride_b;      // bytes between input rows
    int32_t   nrow, ncol;       // input extent
    int32_t   onrow, oncol;     // output extent written: nrow/2, ncol/2
    int32_t   vec;              // 4-pixel input loads are aligned (data and stride)
    int32_t   ostride;          // output row stride, elements
    int32_t   tiles_y;          // workgroups per column strip
    uintptr_t out;              // output row 0
};

template <class T> struct SubPx;

template <> struct SubPx<float> {
    static __device__ __forceinline__ void ld4(__amdgpu_buffer_rsrc_t r, int voff, float v[4])
    {
        const auto q = __builtin_amdgcn_raw_buffer_load_b128(r, voff, 0, 0);
        v[0] = __uint_as_float(q[0]); v[1] = __uint_as_float(q[1]);
        v[2] = __uint_as_float(q[2]); v[3] = __uint_as_float(q[3]);
    }
    static __device__ __forceinline__ void st2(gptr<float> p, float a, float b)
    {
        *reinterpret_cast<gptr<v2f32>>(p) = v2f32{a, b};
    }
};

template <> struct SubPx<uint8_t> {
    static __device__ __forceinline__ void ld4(__amdgpu_buffer_rsrc_t r, int voff, float v[4])
    {
        const uint32_t q = __builtin_amdgcn_raw_buffer_load_b32(r, voff, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] = float((q >> (8 * i)) & 0xffu);
    }
    static __device__ __forceinline__ void st2(gptr<uint8_t> p, float a, float b)
    {
        *reinterpret_cast<gptr<uint16_t>>(p) = uint16_t(Px<uint8_t>::cvt(a) | Px<uint8_t>::cvt(b) << 8);
    }
};

template <> struct SubPx<f16_t> {
    static __device__ __forceinline__ void ld4(__amdgpu_buffer_rsrc_t r, int voff, float v[4])
    {
        const auto q = __builtin_amdgcn_raw_buffer_load_b64(r, voff, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; ++i)
            v[i] = float(__builtin_bit_cast(f16_t, uint16_t(q[i / 2] >> (16 * (i % 2)))));
    }
    static __device__ __forceinline__ void st2(gptr<f16_t> p, float a, float b)
    {
        *reinterpret_cast<gptr<uint32_t>>(p) = pack_f16(a, b);
    }
};

template <> struct SubPx<bf16_t> {
    static __device__ __forceinline__ void ld4(__amdgpu_buffer_rsrc_t r, int voff, float v[4])
    {
        const auto q = __builtin_amdgcn_raw_buffer_load_b64(r, voff, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] = __uint_as_float((q[i / 2] >> (16 * (i % 2))) << 16);
    }
    static __device__ __forceinline__ void st2(gptr<bf16_t> p, float a, float b)
    {
        *reinterpret_cast<gptr<uint32_t>>(p) = pack_bf16(a, b);
    }
};

// One input row as a lane needs it: e[k + 2] = in[row][4g + k], k = -2 .. 4.
// VEC: the four pixels 4g .. 4g+3 come in one aligned load, and the
// neighbours a 3x3 / 5x5 stencil needs come from the adjacent lanes by DPP
// wave shifts; only lane 0 (left) and lane 63 (right) load theirs, the other
// lanes' offsets of those loads being out of the buffer's range (no memory
// access, result 0).  Otherwise every pixel is loaded on its own.  No load
// sits in a divergent branch, so a wave's loads of all its rows are in
// flight together.
// DPP move across the wave: lane i takes v from lane i - 1 (WAVE_SHR1) or
// i + 1 (WAVE_SHL1); the lane with no source keeps own_edge.
constexpr int WAVE_SHR1 = 0x138, WAVE_SHL1 = 0x130;
template <int CTRL> __device__ __forceinline__ float sub_shift(float own_edge, float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(own_edge), __float_as_int(v),
                                                      CTRL, 0xf, 0xf, false));
}

template <int RX, bool VEC, class IN> __device__ __forceinline__ void
sub_load_row(__amdgpu_buffer_rsrc_t rsrc, int rowoff, int x0, float e[7])
{
    constexpr int es = int(sizeof(IN));
    constexpr int OOR = int(0x80000000u);          // beyond any image (images are below 2 GiB)
    const int voff = rowoff + x0 * es;
    const int lane = int(threadIdx.x);
    float c[4];
    if constexpr (VEC) {
        SubPx<IN>::ld4(rsrc, voff, c);
    } else {
#pragma unroll
        for (int i = 0; i < 4; ++i)
            c[i] = (RX > 0 || i % 2 == 0) ? Px<IN>::ld(rsrc, voff + i * es, 0) : 0.f;
    }
    e[2] = c[0]; e[3] = c[1]; e[4] = c[2]; e[5] = c[3];
    e[0] = e[1] = e[6] = 0.f;
    if constexpr (VEC) {
        if constexpr (RX >= 1)
            e[1] = sub_shift<WAVE_SHR1>(Px<IN>::ld(rsrc, lane == 0 ? voff - es : OOR, 0), c[3]);
        if constexpr (RX >= 2) {
            e[0] = sub_shift<WAVE_SHR1>(Px<IN>::ld(rsrc, lane == 0 ? voff - 2 * es : OOR, 0), c[2]);
            e[6] = sub_shift<WAVE_SHL1>(Px<IN>::ld(rsrc, lane == 63 ? voff + 4 * es : OOR, 0), c[0]);
        }
    } else {
        if constexpr (RX >= 1) e[1] = Px<IN>::ld(rsrc, voff - es, 0);
        if constexpr (RX >= 2) {
            e[0] = Px<IN>::ld(rsrc, voff - 2 * es, 0);
            e[6] = Px<IN>::ld(rsrc, voff + 4 * es, 0);
        }
    }
}

// gaussian5x5's 1-D pass (functional.h:249-254): weights 0.375, 0.25, 0.0625.
__device__ __forceinline__ float sub_g5(float m2, float m1, float c, float p1, float p2)
{
#pragma clang fp contract(off)
    return 0.375f * c + 0.25f * (m1 + p1) + 0.0625f * (m2 + p2);
}

// The operator at input (Y, X); p(dy, dx) returns in[Y + dy][X + dx].  The
// expressions are those of functional.h, in source order.
template <int OP, class P> __device__ __forceinline__ float sub_op(const P& p)
{
#pragma clang fp contract(off)
    if constexpr (OP == SUB_IDENTITY) return p(0, 0);
    if constexpr (OP == SUB_DIFFH)    return 0.5f * (p(0, 1) - p(0, -1));
    if constexpr (OP == SUB_DIFFV)    return 0.5f * (p(1, 0) - p(-1, 0));
    if constexpr (OP == SUB_DIFFHH)   return p(0, -1) - 2.f * p(0, 0) + p(0, 1);
    if constexpr (OP == SUB_DIFFVV)   return p(-1, 0) - 2.f * p(0, 0) + p(1, 0);
    if constexpr (OP == SUB_DIFFHV)
        return 0.25f * (p(-1, -1) - p(-1, 1) - p(1, -1) + p(1, 1));
    if constexpr (OP == SUB_SOBELH)
        return 0.125f * (p(-1, 1) - p(-1, -1) + p(1, 1) - p(1, -1)) + 0.25f * (p(0, 1) - p(0, -1));
    if constexpr (OP == SUB_SOBELV)
        return 0.125f * (p(1, -1) - p(-1, -1) + p(1, 1) - p(-1, 1)) + 0.25f * (p(1, 0) - p(-1, 0));
    if constexpr (OP == SUB_SOBELABS)
        return __builtin_fabsf(sub_op<SUB_SOBELH>(p)) + __builtin_fabsf(sub_op<SUB_SOBELV>(p));
    if constexpr (OP == SUB_LAPLACIAN)
        return p(0, -1) + p(0, 1) + p(-1, 0) + p(1, 0) - 4.f * p(0, 0);
    if constexpr (OP == SUB_DET) {
        const float dxy = sub_op<SUB_DIFFHV>(p);
        return sub_op<SUB_DIFFHH>(p) * sub_op<SUB_DIFFVV>(p) - dxy * dxy;
    }
    if constexpr (OP == SUB_GAUSSIAN5) {
        float b[5];
#pragma unroll
        for (int i = 0; i < 5; ++i)
            b[i] = sub_g5(p(i - 2, -2), p(i - 2, -1), p(i - 2, 0), p(i - 2, 1), p(i - 2, 2));
        return sub_g5(b[0], b[1], b[2], b[3], b[4]);
    }
    return 0.f;
}

template <int OP, bool VEC, class IN, class OUT> __device__ __forceinline__ void
subsample_wave(const SubArgs& a, int g, int yo0)
{
    constexpr int RY = sub_ry(OP), RX = sub_rx(OP);
    constexpr int NR = RY == 0 ? SUB_R : 2 * SUB_R - 1 + 2 * RY;   // input rows per wave
    const int x0 = 4 * g;
    const __amdgpu_buffer_rsrc_t rsrc =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(a.in), short(0), int(a.in_bytes), BUF_DWORD3);
    float w[NR][7];
#pragma unroll
    for (int i = 0; i < NR; ++i) {
        int y = RY == 0 ? 2 * yo0 + 2 * i : 2 * yo0 - RY + i;
        y = min(max(y, 0), a.nrow - 1);                           // rows outside feed only zeroed outputs
        sub_load_row<RX, VEC, IN>(rsrc, y * a.in_stride_b, x0, w[i]);
    }

#pragma unroll
    for (int t = 0; t < SUB_R; ++t) {
        const int yo = yo0 + t;
        if (yo >= a.onrow) break;
        const int Y = 2 * yo;
        const int wc = RY == 0 ? t : 2 * t + RY;                  // window row of Y
        const bool yin = Y >= RY && Y + RY < a.nrow;
        float v[2];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int X = x0 + 2 * j;
            const bool inside = yin && X >= RX && X + RX < a.ncol;
            const auto p = [&](int dy, int dx) { return w[wc + dy][2 + 2 * j + dx]; };
            v[j] = inside ? sub_op<OP>(p) : 0.f;
        }
        const int xo = 2 * g;
        const gptr<OUT> q = reinterpret_cast<gptr<OUT>>(a.out + uintptr_t(int64_t(yo) * a.ostride + xo) * sizeof(OUT));
        if (xo + 1 < a.oncol && (reinterpret_cast<uintptr_t>(q) & (2 * sizeof(OUT) - 1)) == 0) {
            SubPx<OUT>::st2(q, v[0], v[1]);
        } else {
            if (xo < a.oncol) Px<OUT>::st(q, v[0]);
            if (xo + 1 < a.oncol) Px<OUT>::st(q + 1, v[1]);
        }
    }
}

template <int OP, class IN, class OUT>
__global__ __launch_bounds__(64 * SUB_WAVES) void subsample_kernel(SubArgs a)
{
    // Workgroups in column-strip order, one contiguous range per XCD, so
    // that vertically adjacent ones (sharing stencil halo rows) meet in one L2.
    const int wg = xcd_remap(blockIdx.x, gridDim.x);
    const int bx = wg / a.tiles_y, by = wg % a.tiles_y;
    const int g = bx * 64 + threadIdx.x;                          // owns output columns 2g, 2g+1
    const int yo0 = (by * SUB_WAVES + threadIdx.y) * SUB_R;       // first output row of the wave
    if (yo0 >= a.onrow) return;                                   // wave-uniform
    // Wave-uniform load form: aligned 4-pixel loads when the image allows
    // them, unless a group of the last input row would reach past the end of
    // the image (the one place where that is not padding or the next row).
    // Pixels past a row's end feed only outputs that are zeroed or not
    // stored.
    constexpr int RY = sub_ry(OP);
    const int last_row = 2 * yo0 + (RY == 0 ? 2 * (SUB_R - 1) : 2 * (SUB_R - 1) + RY);
    if (a.vec && (4 * (bx * 64 + 63) + 3 < a.ncol || last_row < a.nrow - 1))
        subsample_wave<OP, true, IN, OUT>(a, g, yo0);
    else
        subsample_wave<OP, false, IN, OUT>(a, g, yo0);
}

template <int OP, class IN, class OUT>
hipError_t launch_subsample(const SubArgs& a, hipStream_t s)
{
    const int nwg = (a.oncol + 127) / 128 * a.tiles_y;
    hipLaunchKernelGGL((subsample_kernel<OP, IN, OUT>), dim3(nwg), dim3(64, SUB_WAVES), 0, s, a);
    return hipGetLastError();
}

using SubLaunchFn = hipError_t (*)(const SubArgs&, hipStream_t);
extern SubLaunchFn g_sub_table[NSUBOPS][4][4];     // [operator][input dtype][output dtype]

// ------------------------------------------------------------------------
// The pyramid step fused: FIRGaussianConvolver2::smooth into a zero-filled
// image followed by subsample(identity1x1) (FIRGaussianConvolver.h:178-182,
// algorithm.h:183-242), in one pass that computes the smoothed image only
// where the subsample reads it:
//
//   half[y][x] = sum_{i,j} k[i] k[j] in[Yc + i][Xc + j],  |i|, |j| <= r,
//   (Yc, Xc) = (2y, 2x) + (shift ? 0 : (r, r)),
//
// where the window lies inside the image, else 0 (the zero border of the
// intermediate), for y < nrow/2, x < ncol/2.  k is the symmetric lobe
// (centre-indexed, zero beyond r).  No counterpart in the reference, which
// writes the full-size smoothed image and reads it back.
//
// A workgroup owns PYR_TH x PYR_TW output pixels.  Vertical pass: thread t
// owns input column Xc(x0) - R + t and loads the 2*PYR_TH - 1 + 2R rows its
// column needs straight into registers (each wave load reads 64 consecutive
// pixels), then forms the vertical sums centred on the PYR_TH even rows into
// LDS.  Horizontal pass: a thread takes 4 adjacent outputs, reading the
// 2R + 7 sums their windows span as 16-byte LDS vectors (one 4-byte read
// per tap and output, at a 2-float stride, made the LDS the bottleneck).  Sums run outer tap pairs first, as FIRFilter.h:66-176
// does (vertical pass first here, horizontal first there: the results
// differ by FP32 rounding only).
constexpr int PYR_TH = 8;       // output rows per workgroup (16 and 32 measured slower: fewer workgroups)
constexpr int PYR_TW = 112;     // output columns per workgroup: 2*111 + 2*16 + 1 <= 256 input columns
constexpr int PYR_P  = 264;     // LDS row pitch, floats

struct PyrArgs {
    const void* in;
    uint32_t  in_bytes;         // bytes from row 0 to the end of the last row
    int32_t   in_stride_b;      // bytes between input rows
    int32_t   nrow, ncol;       // input extent
    int32_t   onrow, oncol;     // output extent written: nrow/2, ncol/2
    int32_t   r;                // lobe radius
    int32_t   o;                // centre offset: shift ? 0 : r
    int32_t   ostride;          // output row stride, elements
    int32_t   tiles_y;          // tiles per column strip
    uintptr_t out;
    float     c[KMAX];          // centre-indexed lobe, zero beyond r
};

template <int R> __device__ __forceinline__ float pyr_sum(const PyrArgs& a, const float* s)
{
    float acc = a.c[R] * (s[0] + s[2 * R]);
#pragma unroll
    for (int j = R - 1; j >= 1; --j) acc = __builtin_fmaf(a.c[j], s[R - j] + s[R + j], acc);
    return __builtin_fmaf(a.c[0], s[R], acc);
}

template <int R, class IN, class OUT>
__global__ __launch_bounds__(256) void pyramid_kernel(PyrArgs a)
{
    constexpr int NW = 2 * (PYR_TH - 1) + 2 * R + 1;   // input rows per column
    constexpr int NC = 2 * (PYR_TW - 1) + 2 * R + 1;   // input columns per tile
    static_assert(NC <= 256, "tile wider than the workgroup");
    __shared__ float V[PYR_TH][PYR_P];
    const int t = threadIdx.x;
    // Tiles in column-strip order, cut into one contiguous range per XCD:
    // vertically adjacent tiles, which read the same 2R halo rows, run at
    // the same time on one L2 (dispatch order alone spreads them over the
    // eight XCDs and their L2s, and the halo rows come from HBM twice).
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const int y0 = (tile % a.tiles_y) * PYR_TH, x0 = (tile / a.tiles_y) * PYR_TW;
    const int yc0 = 2 * y0 + a.o, xc0 = 2 * x0 + a.o;     // centre of output (y0, x0)

    if (t < NC) {
        const __amdgpu_buffer_rsrc_t rsrc =
            __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(a.in), short(0), int(a.in_bytes), BUF_DWORD3);
        // Row and column both go into the vector offset, the part of the
        // address the buffer range check covers (the scalar offset is not
        // checked): a column left of the image reads the end of the row
        // above or, on row 0, returns 0; a column right of it reads the
        // next row or, on the last row, returns 0.  Either feeds only
        // outputs that are zeroed.
        const int coff = (xc0 - R + t) * int(sizeof(IN));
        float w[NW];
#pragma unroll
        for (int i = 0; i < NW; ++i) {
            const int row = min(max(yc0 - R + i, 0), a.nrow - 1);
            w[i] = Px<IN>::ld(rsrc, row * a.in_stride_b + coff, 0);
        }
#pragma unroll
        for (int k = 0; k < PYR_TH; ++k) V[k][t] = pyr_sum<R>(a, &w[2 * k]);
    }
    __syncthreads();

    // Horizontal pass: a thread forms 4 adjacent outputs from the 2R + 7
    // vertical sums their windows span, read as 16-byte LDS vectors.
    constexpr int NQ = PYR_TW / 4;                        // 4-output groups per row
    constexpr int NS = 2 * R + 8;
    for (int gi = t; gi < PYR_TH * NQ; gi += 256) {
        const int k = gi / NQ, xl = 4 * (gi % NQ);
        const int y = y0 + k, x = x0 + xl;
        if (y >= a.onrow || x >= a.oncol) continue;
        const int yc = yc0 + 2 * k;
        const bool yin = yc >= a.r && yc + a.r < a.nrow;
        float s[NS];
#pragma unroll
        for (int m = 0; m < NS / 4; ++m) {
            const v4f32 q = *reinterpret_cast<const v4f32*>(&V[k][2 * xl + 4 * m]);
            s[4 * m] = q[0]; s[4 * m + 1] = q[1]; s[4 * m + 2] = q[2]; s[4 * m + 3] = q[3];
        }
        float v[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int xc = xc0 + 2 * (xl + i);
            v[i] = yin && xc >= a.r && xc + a.r < a.ncol ? pyr_sum<R>(a, &s[2 * i]) : 0.f;
        }
        const gptr<OUT> q = reinterpret_cast<gptr<OUT>>(a.out + uintptr_t(int64_t(y) * a.ostride + x) * sizeof(OUT));
        if (x + 3 < a.oncol && (reinterpret_cast<uintptr_t>(q) & (4 * sizeof(OUT) - 1)) == 0) {
            Px<OUT>::st4(q, v[0], v[1], v[2], v[3]);
        } else {
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (x + i < a.oncol) Px<OUT>::st(q + i, v[i]);
        }
    }
}

template <int R, class IN, class OUT>
hipError_t launch_pyramid(const PyrArgs& a, hipStream_t s)
{
    const int tiles = (a.oncol + PYR_TW - 1) / PYR_TW * a.tiles_y;
    hipLaunchKernelGGL((pyramid_kernel<R, IN, OUT>), dim3(tiles), dim3(256), 0, s, a);
    return hipGetLastError();
}

constexpr int PYR_RADII[3] = {4, 8, 16};
using PyrLaunchFn = hipError_t (*)(const PyrArgs&, hipStream_t);
extern PyrLaunchFn g_pyr_table[3][4][4];           // [radius class][input dtype][output dtype]

}   // namespace tucu
