// mfma.h -- matrix-core (MFMA) variant of the fused six-output Gaussian filter.
//
// The direct FP32 filter needs about 16 r + 6 vector operations per output
// pixel for the six outputs (262 at radius 16), which makes it VALU-bound at
// a quarter of the HBM roofline.  Here both passes run on the matrix cores as
// banded matrix products, and the vector units only split, scale and store.
//
// Strip: one wave64 filters 16 centre rows x 64 centre columns (4 blocks of
// 16 columns); a workgroup puts 4 waves side by side (16 x 256), so that the
// rows it writes are long runs (the write stream dominates the traffic).  With RC the radius
// class (8 or 16; lobes zero-padded to it) and K = 16 + 2 RC:
//   vertical pass, per field block b (16 input columns starting at
//   x0 - RC + 16 b) and lobe l:
//       F^T (16 cols x 16 rows) = X^T (16 cols x K input rows) * T_l^T (K x 16)
//   horizontal pass, per output block o (16 columns starting at x0 + 16 o):
//       O^T (16 cols x 16 rows) = T_h^T (16 x K field cols) * F^T (K x 16 rows)
//   with T the band matrices of the lobes.  The vertical product's
//   accumulator layout (lane l: row l % 16, columns 4 (l / 16) .. +3) is
//   exactly the layout the horizontal product wants for its B operand when
//   the K index of the band matrix is permuted to match, so fields never
//   leave registers.
// Precision: matrix operands are f16.  A float v is scaled by a power of two
// and split as hi = f16(v), lo = f16(v - hi); the products hi*hi + hi*lo +
// lo*hi carry about 22 significant bits (the dropped lo*lo term and the
// split residual are each below 2^-22 relative), summed in FP32.  Scales
// (all powers of two): taps of lobe l by mp[l] (largest tap in [2^5, 2^6),
// MFMA_TAP_EXP); the input of each strip by mq times a factor putting its
// largest magnitude in [2^12, 2^13) (u8 input: the factor for 255, 2^5,
// whatever the strip holds), mq <= 1 / (L1 norm * mp) of every lobe,
// so that the vertical fields stay below 2^13 and no operand leaves the f16
// range; each output is multiplied back by mu[k] / (the strip's factor).
// Where lo falls below f16's normal range (2^-14) it keeps an absolute
// 2^-25 instead of 11 more bits, so the error of an output is bounded by
// about 2^-25 (L1 / X + n tapmax / T) times the strip's largest input, with
// X and T the scaled input and tap peaks; the tap peak of 2^6 balances the
// two terms (a peak of 2^12 left the inputs at a quarter and their lo
// halves subnormal below half the strip's largest value: 20x the error on
// an image whose strips mix 0..255 data with values of 10^6).
// A strip whose input holds a NaN or an Inf, or whose largest magnitude the
// scale cannot bring into range, is filtered by the FP32 form
// (mfma_strip_fp32), which reaches exactly the outputs whose window holds
// a non-finite input, as the reference's tap sums do.
#pragma once

#include "sepfir.h"
#include "batch.h"

#include <algorithm>
#include <type_traits>

namespace tucu {

typedef _Float16 v4f16 __attribute__((ext_vector_type(4)));
typedef _Float16 v8f16 __attribute__((ext_vector_type(8)));

constexpr int MSTRIP_W = 64;              // centre columns per strip
constexpr int MNB      = MSTRIP_W / 16;   // output blocks per strip
constexpr int MWAVES   = 4;               // strips (waves) per workgroup, side by side
constexpr int MTILE_H  = 16;              // centre rows per workgroup tile
constexpr int MTILE_W  = MSTRIP_W * MWAVES;   // centre columns per workgroup tile
constexpr int MOPITCH  = MSTRIP_W + 4;    // row pitch (floats) of a wave's output staging area
constexpr int MT4      = 112;             // tap-run table entries per lobe
constexpr int MBASE    = 64;              // run t starts with tap t - MBASE
constexpr int MHP      = MT4 - 16;        // entries of the horizontal pair table
constexpr int MINVALID = int(0x80000000u);   // buffer offset past any image: loads return 0
constexpr int MFMA_TAP_EXP = 6;           // largest scaled tap of each lobe in [2^5, 2^6)

template <int RC> struct MGeo {
    static constexpr bool K48 = RC == 16;                 // K = 48 (two 32-entry slices), else 32
    static constexpr int  KV  = 16 + 2 * RC;              // input rows per strip
    static constexpr int  NVB = MNB + (K48 ? 2 : 1);      // field blocks per strip
    static constexpr int  NLD = K48 ? 12 : 8;             // input values per lane per block
};

#ifdef TUCU_TIMELINE
// Phase timeline of the one-round six-output launch (radius class 16, float
// in and out), in a build made for it (make timeline; read by
// tucuda_amd/timeline.py): lane 0 of each wave writes the 100 MHz real-time
// counter into slot i of its row of tucu_timeline at each phase -- 0 start,
// 10 strip loads issued, 11 operand tables copied, 1 barrier passed, 2
// input scaled and split, 3 + k output k stored, 9 end.  The default build
// has none of it.
constexpr int TL_SLOTS = 12;
constexpr int TL_WAVES = 2048;
static __device__ uint64_t tucu_timeline[TL_WAVES * TL_SLOTS];
__device__ __forceinline__ void tl_stamp(int i)
{
    const uint64_t t = __builtin_amdgcn_s_memrealtime();
    const unsigned w = blockIdx.x * MWAVES + (threadIdx.x >> 6);
    if ((threadIdx.x & 63) == 0 && w < unsigned(TL_WAVES)) tucu_timeline[w * TL_SLOTS + i] = t;
}
#endif

// Per-lane operand runs, in LDS: run[(part * 3 + l) * MT4 + t] holds taps
// t - MBASE .. t - MBASE + 3 of lobe l, times mp[l], as four f16 (part 0 the high
// half, part 1 the low half).  Tap o is the weight of the input at offset o
// (row or column) from the centre, zero beyond the lobe: symmetric lobes (0,
// 2) weigh both sides c[|o|]; the antisymmetric lobe 1 weighs the upper /
// left side +c[k] and the lower / right side -c[k] (FIRFilter.h:66-176 pairs
// them the same way).  A lane's operand is one or two runs, so it loads
// with 8-byte LDS reads and needs no conversion.
struct alignas(16) MfmaTabs {
    v4f16 run[2 * 3 * MT4];
    v8f16 hpair[2 * 3 * MHP];  // horizontal operands: runs t and t + 16 side by side
    float lobe[2 * 33];        // staging: cv[l][k] at 17 l + k (GAUSS), tv / th at 33 l + o + 16 (TAPS)
    float tap[3][MT4 + 4];     // staging: tap o of lobe l times mp[l] at [l][o + MBASE]
};

// f(std::integral_constant<int, I>) for I = 0 .. N-1, unrolled.
template <int I, int N, class F>
__device__ __forceinline__ void static_for(F&& f)
{
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}

// Tap o of lobe l from the staged lobe words (MfmaTabs::lobe).  MODE_GAUSS:
// the three lobes of the Gaussian (centre-indexed cv[l]).  MODE_TAPS (one
// output of FIRFilter2 or of general taps): table 0 holds the vertical taps,
// table 1 the horizontal ones, both as full taps with tap o at index o + 16
// of tv / th.
template <int MODE>
__device__ __forceinline__ float mfma_tap_lds(const MfmaTabs& T, int l, int o)
{
    const int k = o < 0 ? -o : o;
    if (k > 16) return 0.f;
    if constexpr (MODE == MODE_TAPS) {
        return l < 2 ? T.lobe[33 * l + o + 16] : 0.f;
    } else {
        const float c = T.lobe[17 * l + k];
        return l != 1 ? c : o == 0 ? 0.f : o < 0 ? c : -c;
    }
}

template <int MODE> constexpr int mfma_nlobes() { return MODE == MODE_TAPS ? 2 : 3; }

// Builds the operand tables of one job in LDS.  Every thread of the
// workgroup calls it (it holds a barrier); the caller puts another barrier
// between it and the first use of T.  The lobe words are read with scalar
// loads (their own counter, so vector loads issued before the call, such as
// a strip's input, stay in flight) and written to LDS by one lane; every
// entry is then built from LDS.
template <int MODE = MODE_GAUSS, class A>
__device__ __forceinline__ void mfma_build_tabs(const A& a, MfmaTabs& T)
{
    constexpr int NL = mfma_nlobes<MODE>();
    constexpr int NW = MODE == MODE_TAPS ? 66 : 51;
    if (threadIdx.x == 0) {
        static_for<0, NW>([&](auto ic) {
            constexpr int i = ic;
            if constexpr (MODE == MODE_TAPS) T.lobe[i] = i < 33 ? a.tv[i] : a.th[i - 33];
            else                             T.lobe[i] = a.cv[i / 17][i % 17];
        });
    }
    const float mp0 = a.mp[0], mp1 = a.mp[1], mp2 = a.mp[2];
    __syncthreads();
    constexpr int NT4 = MT4 + 4;
    for (int e = threadIdx.x; e < NL * NT4; e += blockDim.x) {
        const int l = e / NT4, t = e - l * NT4;
        T.tap[l][t] = mfma_tap_lds<MODE>(T, l, t - MBASE) * (l == 0 ? mp0 : l == 1 ? mp1 : mp2);
    }
    __syncthreads();
    // runs (vertical operands: taps t - MBASE .. +3) and pairs of runs 16
    // apart (horizontal operands)
    for (int e = threadIdx.x; e < NL * (MT4 + MHP); e += blockDim.x) {
        const bool pair = e >= NL * MT4;
        const int ee = pair ? e - NL * MT4 : e;
        const int n = pair ? MHP : MT4;
        const int l = ee / n, t = ee - l * n;
        v8f16 hi, lo;
#pragma unroll
        for (int i = 0; i < 8; ++i) {               // taps t .. t + 3, then (pairs) t + 16 .. t + 19
            const float c = i < 4 || pair ? T.tap[l][t + (i < 4 ? i : 12 + i)] : 0.f;
            const _Float16 h = _Float16(c);
            hi[i] = h;
            lo[i] = _Float16(c - float(h));
        }
        if (pair) {
            T.hpair[(0 * 3 + l) * MHP + t] = hi;
            T.hpair[(1 * 3 + l) * MHP + t] = lo;
        } else {
            T.run[(0 * 3 + l) * MT4 + t] = __builtin_shufflevector(hi, hi, 0, 1, 2, 3);
            T.run[(1 * 3 + l) * MT4 + t] = __builtin_shufflevector(lo, lo, 0, 1, 2, 3);
        }
    }
}

// The operand tables of a launch in LDS: copied from a.tabs, where the
// host has built them once for these lobes (the same values, bit for bit:
// mfma_host_tabs restates mfma_build_tabs), else built here.  Every thread
// of the workgroup calls it; the caller puts a barrier between it and the
// first use of T.  (Building them took 2-3 us of a 1080p launch.)
template <int MODE = MODE_GAUSS, class A>
__device__ __forceinline__ void mfma_get_tabs(const A& a, MfmaTabs& T)
{
    if (a.tabs) {                               // uniform
        const gptr<const v4f32> src = reinterpret_cast<gptr<const v4f32>>(reinterpret_cast<uintptr_t>(a.tabs));
        v4f32* const dst = reinterpret_cast<v4f32*>(&T);
        for (int i = threadIdx.x; i < int(sizeof(MfmaTabs) / 16); i += blockDim.x) dst[i] = src[i];
    } else {
        mfma_build_tabs<MODE>(a, T);
    }
}

__device__ __forceinline__ v8f16 cat8(v4f16 a, v4f16 b)
{
    return __builtin_shufflevector(a, b, 0, 1, 2, 3, 4, 5, 6, 7);
}

// v of the lane that DPP control CTRL selects (within rows of 16 lanes).
template <int CTRL> __device__ __forceinline__ float dpp_f32(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xf, 0xf, true));
}

// Largest v over the wave (every lane active), in every lane: DPP within
// each row of 16 lanes (quad swaps, half-row and row mirrors), then the four
// row results read as scalars.  (Six ds_bpermute rounds, each waiting for
// the LDS, did the same with a longer dependent chain.)
__device__ __forceinline__ float wave_max(float v)
{
    v = fmaxf(v, dpp_f32<0xb1>(v));             // quad_perm [1, 0, 3, 2]
    v = fmaxf(v, dpp_f32<0x4e>(v));             // quad_perm [2, 3, 0, 1]
    v = fmaxf(v, dpp_f32<0x141>(v));            // row_half_mirror
    v = fmaxf(v, dpp_f32<0x140>(v));            // row_mirror
    const float a = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 0));
    const float b = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 16));
    const float c = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 32));
    const float d = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 48));
    return fmaxf(fmaxf(a, b), fmaxf(c, d));
}

// Two u8 values as f16 times s, exactly: with M = 1024 s a power of two in
// f16's normal range, M's bits with x (< 256) in the low mantissa bits are
// the f16 M + x s, and subtracting M leaves x s.  mb = M's bits in both
// halves, nm = -M in both halves.
__device__ __forceinline__ v2f16 u8pair_f16(uint32_t lo, uint32_t hi, uint32_t mb, v2f16 nm)
{
    const uint32_t w = (hi << 16) | lo | mb;
    return __builtin_bit_cast(v2f16, w) + nm;
}

// One input value as loaded, unconverted: float, the 16 bits of an f16 /
// bf16, or the byte.  mfma_compute converts them where it uses them, so that
// no conversion sits right after the loads and makes the wave wait for them
// there (the loads of a strip stay in flight through the table build, or
// through the previous strip's work in the tile loop).
// Narrow inputs are held zero-extended in 32 bits.
template <class IN> using RawElem = std::conditional_t<sizeof(IN) == 4, float, uint32_t>;

// Range-checked load of a float input (a pixel outside the image reads 0).
template <class IN>
__device__ __forceinline__ RawElem<IN> mfma_ld(__amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    static_assert(sizeof(IN) == 4, "narrow inputs load with mfma_ld_in");
    return Px<IN>::ld(r, voff, soff);
}

// Unchecked load at row base + lane offset, for windows inside the image:
// a narrow value arrives zero-extended by the load itself (with the buffer
// form the compiler masks it again, and that mask, placed right after the
// loads, makes the wave wait for them there).
template <class IN>
__device__ __forceinline__ RawElem<IN> mfma_ld_in(const char* row, uint32_t off)
{
    const gptr<const uint8_t> base = reinterpret_cast<gptr<const uint8_t>>(reinterpret_cast<uintptr_t>(row));
    if constexpr (sizeof(IN) == 4)
        return *reinterpret_cast<const gptr<const float>>(base + off);
    else if constexpr (sizeof(IN) == 2)
        return *reinterpret_cast<const gptr<const uint16_t>>(base + off);
    else
        return base[off];
}

// The value of a loaded element as a float.
template <class IN>
__device__ __forceinline__ float raw_value(RawElem<IN> v)
{
    if constexpr (std::is_same_v<IN, f16_t>)
        return float(__builtin_bit_cast(_Float16, uint16_t(v)));
    else if constexpr (std::is_same_v<IN, bf16_t>)
        return __uint_as_float(uint32_t(v) << 16);
    else
        return float(v);
}

// Input values of field block b for lane (g, m): column x0 - RC + 16 b + m,
// rows y0 - RC + 8 g + j (j < 8) and, for K = 48, y0 - RC + 32 + 4 g + j
// (j < 4).  INSIDE: every pixel of the strip's window lies in the image
// (row offsets as scalar offsets); otherwise pixels outside the image read
// as 0 (buffer range check for float input; for narrow input a load at the
// clamped address and a select).
template <int RC, bool INSIDE, class IN, class A>
__device__ __forceinline__ void
mfma_load_block(const A& a, __amdgpu_buffer_rsrc_t rsrc, int y0, int x0, int b, int g, int m, RawElem<IN>* v)
{
    constexpr int NLD = MGeo<RC>::NLD;
    const int es = int(sizeof(IN));
    const int col = x0 - RC + 16 * b + m;
    const int sb = a.in_stride_b;
    const int r0 = y0 - RC + 8 * g;            // first row of entries 0 .. 7
    const int r1 = y0 - RC + 32 + 4 * g;       // first row of entries 8 .. 11
    if constexpr (INSIDE && sizeof(IN) < 4) {
        // row bases uniform (y0 - RC + j), lane offsets 32-bit
        const char* const in = static_cast<const char*>(a.in);
        const uint32_t o0 = uint32_t(8 * g * sb + col * es), o1 = uint32_t((32 + 4 * g) * sb + col * es);
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = mfma_ld_in<IN>(in + int64_t(y0 - RC + j) * sb, o0);
        if constexpr (MGeo<RC>::K48) {
#pragma unroll
            for (int j = 0; j < 4; ++j) v[8 + j] = mfma_ld_in<IN>(in + int64_t(y0 - RC + j) * sb, o1);
        }
    } else if constexpr (INSIDE) {
        const int v0 = r0 * sb + col * es, v1 = r1 * sb + col * es;
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = mfma_ld<IN>(rsrc, v0, j * sb);
        if constexpr (MGeo<RC>::K48) {
#pragma unroll
            for (int j = 0; j < 4; ++j) v[8 + j] = mfma_ld<IN>(rsrc, v1, j * sb);
        }
    } else if constexpr (sizeof(IN) < 4) {
        // the same extending loads at addresses clamped into the image, and
        // 0 where the pixel lies outside (so that the values of both paths
        // meet without a mask after them)
        const char* const in = static_cast<const char*>(a.in);
        const bool colok = unsigned(col) < unsigned(a.ncol);
        const int colc = min(max(col, 0), a.ncol - 1);
#pragma unroll
        for (int j = 0; j < NLD; ++j) {
            const int row = j < 8 ? r0 + j : r1 + (j - 8);
            const bool ok = colok && unsigned(row) < unsigned(a.nrow);
            const int rowc = min(max(row, 0), a.nrow - 1);
            const RawElem<IN> x = mfma_ld_in<IN>(in + int64_t(rowc) * sb, uint32_t(colc * es));
            v[j] = ok ? x : 0u;
        }
    } else {
        const bool colok = unsigned(col) < unsigned(a.ncol);
#pragma unroll
        for (int j = 0; j < NLD; ++j) {
            const int row = j < 8 ? r0 + j : r1 + (j - 8);
            const bool ok = colok && unsigned(row) < unsigned(a.nrow);
            v[j] = mfma_ld<IN>(rsrc, ok ? row * sb + col * es : MINVALID, 0);
        }
    }
}

// f16 hi / lo split of 4 floats times s.
__device__ __forceinline__ void split4(float v0, float v1, float v2, float v3, float s, v4f16& hi, v4f16& lo)
{
    const v4f32 x = {v0 * s, v1 * s, v2 * s, v3 * s};
    hi = __builtin_convertvector(x, v4f16);
    lo = __builtin_convertvector(x - __builtin_convertvector(hi, v4f32), v4f16);
}

// Entries 4 O .. 4 O + 7 of a field vector: field blocks O and O + 1.
template <int O, class V>
__device__ __forceinline__ v8f16 pair8(const V& f)
{
    return __builtin_shufflevector(f, f, 4 * O, 4 * O + 1, 4 * O + 2, 4 * O + 3,
                                   4 * O + 4, 4 * O + 5, 4 * O + 6, 4 * O + 7);
}

// Field blocks 2 P and 2 P + 1 (pairs that do not overlap).
template <int P, class V>
__device__ __forceinline__ v8f16 pairP(const V& f)
{
    return pair8<2 * P>(f);
}

// ------------------------------------------------------------------------
// FP32 form of one strip, for input the split-f16 products cannot carry:
// non-finite values (a band product multiplies every input of a block by
// every band entry, zeros included, so one NaN or Inf would reach all 16 x
// 16 outputs of its block window, and Inf - Inf in the hi / lo split turns
// Inf into NaN), or a strip whose largest magnitude lies outside the range
// its power-of-two scale covers.  Each output is summed over exactly the
// taps of its lobes -- the true radius, zero taps inside it included, none
// beyond it, no centre tap for an antisymmetric lobe, paired as
// FIRFilter.h:66-176 pairs them -- so a NaN or an Inf reaches exactly the
// outputs whose window holds it, with the sign the reference's sums give.
// Vertical pass first, one half strip (32 output columns, 32 + 2 RC field
// columns) at a time through the wave's staging area.  Kept out of line so
// that the matrix-core path's register allocation and schedule are its own.

// Tap sum over w(-R) .. w(R) around w(0) of a lobe of radius r <= R whose
// tap at distance k is c(k) on the upper / left side: symmetric: sum_k c(k)
// (w(-k) + w(k)) + c(0) w(0); antisymmetric: sum_k c(k) (w(-k) - w(k));
// outermost tap first.  Taps beyond r are skipped, never multiplied.
template <int R, class C, class W>
__device__ __forceinline__ float fp32_paired(const C& c, int r, bool sym, const W& w)
{
    float acc = 0.f;
#pragma unroll
    for (int k = R; k >= 1; --k) {
        const float lo = w(-k), hi = w(k);
        const float s = sym ? lo + hi : lo - hi;
        acc = k <= r ? (k == r ? c(k) * s : __builtin_fmaf(c(k), s, acc)) : acc;
    }
    return sym ? __builtin_fmaf(c(0), w(0), acc) : acc;
}

// General taps t(o), o = -r .. r, top / left tap first.
template <int R, class C, class W>
__device__ __forceinline__ float fp32_taps(const C& t, int r, const W& w)
{
    float acc = 0.f;
#pragma unroll
    for (int o = -R; o <= R; ++o) {
        const int k = o < 0 ? -o : o;
        acc = k <= r ? (o == -r ? t(o) * w(o) : __builtin_fmaf(t(o), w(o), acc)) : acc;
    }
    return acc;
}

// The taps come from the lobe words staged in LDS (MfmaTabs::lobe): MODE_GAUSS
// c_l(k) at 17 l + k; MODE_TAPS the full vertical taps at 16 + o and the
// horizontal ones at 33 + 16 + o, which for FIRFilter2 lobes (sgnv, sgnh
// nonzero) are c(k) on the upper / left side.
template <int RC, class IN, class OUT, int MODE, class A>
__device__ __forceinline__ void
mfma_strip_fp32(const A& a, int y0, int x0, const MfmaTabs& T, float* stage, unsigned mask)
{
    constexpr bool ONE = MODE == MODE_TAPS;
    constexpr int NFC = 32 + 2 * RC;            // field columns per half strip
    static_assert(NFC <= MOPITCH, "staging area too narrow");
    // Values the compiler cannot see through: nothing of this rarely taken
    // path is computed ahead of it (in a tile loop, such hoisted values would
    // hold registers across the matrix-core path).
    int lane = threadIdx.x & 63;
    asm volatile("" : "+v"(lane), "+s"(y0), "+s"(x0), "+s"(mask));
    const int es = int(sizeof(IN));
    const int sb = a.in_stride_b;
    const __amdgpu_buffer_rsrc_t rsrc =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(a.in), short(0), int(a.in_bytes), BUF_DWORD3);
    constexpr int HL[6] = {ONE ? 1 : 0, 1, 0, 2, 1, 0};   // horizontal lobe of output k (TAPS: table 1)
    constexpr int VL[6] = {0, 0, 1, 0, 1, 2};             // vertical lobe of output k
    const unsigned fneed[3] = {ONE ? 1u : mask & 0x0bu, ONE ? 0u : mask & 0x14u, ONE ? 0u : mask & 0x20u};
    const float* const lw = T.lobe;

#pragma unroll 1
    for (int l = 0; l < 3; ++l) {
        if (!fneed[l]) continue;
#pragma unroll 1
        for (int q = 0; q < 2; ++q) {
            // vertical pass: lane = field column x0 + 32 q - RC + lane, rows y0 .. y0 + 15;
            // one row at a time, its window loaded again (L2-resident), so that
            // this rarely taken path stays small in registers
            const int col = x0 + 32 * q - RC + lane;
            if (lane < NFC) {
                const bool colok = unsigned(col) < unsigned(a.ncol);
#pragma unroll 1
                for (int i = 0; i < 16; ++i) {
                    const auto win = [&](int d) {
                        const int row = y0 + i + d;
                        const bool ok = colok && unsigned(row) < unsigned(a.nrow);
                        return Px<IN>::ld(rsrc, ok ? row * sb + col * es : MINVALID, 0);
                    };
                    float f;
                    if constexpr (ONE) {
                        f = a.sgnv != 0.f ? fp32_paired<RC>([&](int k) { return lw[16 - k]; }, a.r[0], a.sgnv > 0.f, win)
                                          : fp32_taps<RC>([&](int o) { return lw[16 + o]; }, a.r[0], win);
                    } else {
                        f = fp32_paired<RC>([&](int k) { return lw[17 * l + k]; }, a.r[l], l != 1, win);
                    }
                    stage[i * MOPITCH + lane] = f;
                }
            }
            __builtin_amdgcn_wave_barrier();
            // horizontal pass: lane = output column x0 + 32 q + (lane & 31), rows (lane >> 5) * 8 + 0 .. 7
            const int j = lane & 31, i0 = (lane >> 5) * 8;
            const int x = x0 + 32 * q + j;
#pragma unroll 1
            for (int k = 0; k < (ONE ? 1 : 6); ++k) {
                if (VL[k] != l || !(mask >> k & 1u)) continue;
                const int h = HL[k];
                const int rv = ONE ? a.r[0] : a.r[VL[k]], rh = ONE ? a.r[1] : a.r[h];
#pragma unroll 1
                for (int i = 0; i < 8; ++i) {
                    const float* srow = stage + (i0 + i) * MOPITCH + RC + j;
                    const auto win = [&](int d) { return srow[d]; };
                    float v;
                    if constexpr (ONE) {
                        v = a.sgnh != 0.f ? fp32_paired<RC>([&](int kk) { return lw[49 - kk]; }, a.r[1], a.sgnh > 0.f, win)
                                          : fp32_taps<RC>([&](int o) { return lw[49 + o]; }, a.r[1], win);
                    } else {
                        v = fp32_paired<RC>([&](int kk) { return lw[17 * h + kk]; }, a.r[h], h != 1, win);
                    }
                    const int y = y0 + i0 + i;
                    if (y >= rv && y < a.nrow - rv && x >= rh && x < a.ncol - rh) {
                        const gptr<OUT> p = reinterpret_cast<gptr<OUT>>(
                            a.obase[k] + uintptr_t(int64_t(y) * a.ostride[k] + x) * sizeof(OUT));
                        Px<OUT>::st(p, v);
                    }
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
}

#define TUCU_MFMA(A_, B_, C_) __builtin_amdgcn_mfma_f32_16x16x32_f16((A_), (B_), (C_), 0, 0, 0)

// One strip: centre rows y0 .. y0+15, centre columns x0 .. x0+63.  T holds
// the operand runs (mfma_build_tabs).  Every lane of the wave calls it.
//
// Operand layouts (lane = 16 g + m; m is also the row / output column n):
//   vertical, block b, K32 slice: A = input rows y0 - RC + 8 g + j (j < 8)
//     of column x0 - RC + 16 b + m; B = taps of offsets 8 g + j - RC - n.
//   vertical, K48 second slice: A = rows 32 + 4 g .. +3 of the lane's
//     input, high halves then low halves; B = taps of offsets
//     32 + 4 g + j - RC - n twice (high taps: both halves of the input in
//     one product), or once and then zeros (low taps x high input).
//   horizontal, K32 (radius class 8), output block o: B = field blocks o,
//     o + 1 (lane holds columns 4 g .. 4 g + 3 of each, window columns
//     4 g + j and 16 + 4 g + j); A = taps of offsets (window column) - RC - m.
//   horizontal, K48: two K32 slices over the fixed pairs P_(o/2) and
//     P_(o/2 + 1) of field blocks (2 p, 2 p + 1); A = taps as above, shifted
//     by d = 32 p - 16 o (zero for the block outside the window).
// A strip's input, as loaded: field block b, entry j (mfma_load_block's
// order); radius class 8 uses blocks 0 .. 4 and entries 0 .. 7.
template <class IN> using MfmaRaw = RawElem<IN>[6][12];

// Register budget of the kernels: two waves per SIMD for float input (the
// low halves of the input and of the tail rows stay live), three for 16-bit
// and u8 input (exact in one f16; measured 9 % faster in the batched mode).
// The single-image tile loop keeps two: inlined in a loop the strip code
// needs more registers.
template <class IN> constexpr int MFMA_WAVES_PER_SIMD = sizeof(IN) == 4 ? 2 : 3;

// The single-image tile loop (large images): two waves per SIMD for the
// six-output Gaussian (the strip code inlined in a loop needs more
// registers); the one-output mode (FIRFilter2 lobes, general taps) needs
// fewer, but at three or four waves per SIMD it spills (176-236 bytes per
// lane at four): two by default, TUCU_LOOP_WAVES_TAPS at build time to try.
#ifndef TUCU_LOOP_WAVES_TAPS
#define TUCU_LOOP_WAVES_TAPS 2
#endif
template <int MODE, class IN> constexpr int MFMA_LOOP_WAVES = MODE == MODE_TAPS ? TUCU_LOOP_WAVES_TAPS : 2;

// Issues the loads of the strip at (y0, x0); the values arrive in raw.
// FRESH: the lane index is opaque to the compiler (a caller's loop then
// recomputes what derives from it in each pass instead of holding it).
template <int RC, class IN, bool FRESH = false, class A>
__device__ __forceinline__ void mfma_load(const A& a, int y0, int x0, MfmaRaw<IN>& raw)
{
    using G = MGeo<RC>;
    int lane = threadIdx.x & 63;
    if constexpr (FRESH) asm volatile("" : "+v"(lane));
    const int g = lane >> 4, m = lane & 15;
    const __amdgpu_buffer_rsrc_t rsrc =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(a.in), short(0), int(a.in_bytes), BUF_DWORD3);
    // One test for the whole window (y0, x0 are wave-uniform): with a test
    // per block, the compiler makes each block's edge path wait for the
    // loads of the blocks before it.
    const bool inside = y0 - RC >= 0 && y0 - RC + G::KV <= a.nrow &&
                        x0 - RC >= 0 && x0 - RC + 16 * G::NVB <= a.ncol;
    if (inside) {
#pragma unroll
        for (int b = 0; b < G::NVB; ++b) mfma_load_block<RC, true, IN>(a, rsrc, y0, x0, b, g, m, raw[b]);
    } else {
#pragma unroll
        for (int b = 0; b < G::NVB; ++b) mfma_load_block<RC, false, IN>(a, rsrc, y0, x0, b, g, m, raw[b]);
    }
    if constexpr (!G::K48) {
#pragma unroll
        for (int b = 0; b < G::NVB; ++b) raw[b][8] = raw[b][9] = raw[b][10] = raw[b][11] = 0;
    }
}

// One strip: centre rows y0 .. y0+15, centre columns x0 .. x0+63, input in
// raw (mfma_load).  T holds the operand runs (mfma_build_tabs), stage is the
// wave's output staging area.  after_vpass() runs once raw has been
// consumed (it may reuse raw, e.g. to prefetch the next strip).  Every lane
// of the wave calls it.
//
// Operand layouts (lane = 16 g + m; m is also the row / output column n):
//   vertical, block b, K32 slice: A = input rows y0 - RC + 8 g + j (j < 8)
//     of column x0 - RC + 16 b + m; B = taps of offsets 8 g + j - RC - n.
//   vertical, K48 second slice: A = rows 32 + 4 g .. +3 of the lane's
//     input, high halves then low halves; B = taps of offsets
//     32 + 4 g + j - RC - n twice (high taps: both halves of the input in
//     one product), or once and then zeros (low taps x high input).
//   horizontal, K32 (radius class 8), output block o: B = field blocks o,
//     o + 1 (lane holds columns 4 g .. 4 g + 3 of each, window columns
//     4 g + j and 16 + 4 g + j); A = taps of offsets (window column) - RC - m.
//   horizontal, K48: two K32 slices over the fixed pairs P_(o/2) and
//     P_(o/2 + 1) of field blocks (2 p, 2 p + 1); A = taps as above, shifted
//     by d = 32 p - 16 o (zero for the block outside the window).
template <int RC, class IN, class OUT, int MODE = MODE_GAUSS, int CPOL = CPOL_NT, bool FRESH = false, class A, class F,
          class S>
__device__ __forceinline__ void
mfma_compute(const A& a, int y0, int x0, MfmaRaw<IN>& raw, const MfmaTabs& T, float* stage, F&& after_vpass,
             S&& sink, unsigned mask)
{
    constexpr bool ONE = MODE == MODE_TAPS;     // one field (table 0), one output (table 1 across)
    using G = MGeo<RC>;
    constexpr int NVB = G::NVB, NLD = G::NLD;
    constexpr bool XLO = sizeof(IN) == 4;       // f16 / bf16 / u8 inputs are exact in one f16
    constexpr int NF = 4 * NVB;                 // field entries per lane
    typedef _Float16 vfield __attribute__((ext_vector_type(NF)));
    int lane = threadIdx.x & 63;
    if constexpr (FRESH) asm volatile("" : "+v"(lane));    // as for mfma_load
    const int g = lane >> 4, m = lane & 15;
    const v4f16 z4 = {0, 0, 0, 0};

    constexpr bool U8 = std::is_same_v<IN, uint8_t>;
    // Scale exponent: the largest scaled input in [2^12, 2^13).  u8 input
    // takes the one for 255 (2^5) whatever the strip holds: every byte
    // times 2^5 mq is exact in f16, and the error bound, relative to 255,
    // stays that of the brightest strip.
    int ex = 5;
    if constexpr (!U8) {
        float mx = 0.f;
#pragma unroll
        for (int b = 0; b < NVB; ++b)
#pragma unroll
            for (int j = 0; j < NLD; ++j) mx = fmaxf(mx, fabsf(raw_value<IN>(raw[b][j])));
        mx = wave_max(mx);
        (void)frexpf(mx, &ex);                  // mx < 2^ex
        ex = 13 - ex;
    }
    // Strips whose largest magnitude is below 2^-87 or above 2^113 are out
    // of the scale's reach (their scaled values would lose precision or
    // overflow f16): the FP32 form filters them.  u8 input never is.
    constexpr bool CHECK = !U8 && MODE != MODE_ICIA;
    bool redo = CHECK && (ex < -100 || ex > 100);   // wave-uniform
    const float s = ldexpf(a.mq, ex), inv_s = ldexpf(1.f, -ex);

    // xh / xl: entries 0 .. 7 (rows 8 g + j), high and low halves; xt (K = 48):
    // entries 8 .. 11 (rows 32 + 4 g + j), high then low halves, so that one
    // product with the high taps twice covers both halves of those rows.
    // Inputs exact in f16 have no low half: xt then holds the high halves
    // twice, and one product with the high and the low taps side by side
    // covers both tap halves (one product per block instead of two)
    v8f16 xh[NVB], xl[NVB], xt[NVB];
    if constexpr (U8) {
        // s = 2^5 mq lies in [2^-7, 2^5] (mq in [2^-12, 1]), so M = 1024 s
        // is a normal f16
        const _Float16 M = _Float16(1024.f * s);
        const uint32_t mb = uint32_t(__builtin_bit_cast(uint16_t, M)) * 0x10001u;
        const v2f16 nm = {-M, -M};
#pragma unroll
        for (int b = 0; b < NVB; ++b) {
            const v2f16 p0 = u8pair_f16(raw[b][0], raw[b][1], mb, nm), p1 = u8pair_f16(raw[b][2], raw[b][3], mb, nm);
            const v2f16 p2 = u8pair_f16(raw[b][4], raw[b][5], mb, nm), p3 = u8pair_f16(raw[b][6], raw[b][7], mb, nm);
            xh[b] = v8f16{p0[0], p0[1], p1[0], p1[1], p2[0], p2[1], p3[0], p3[1]};
            if constexpr (G::K48) {
                const v2f16 p4 = u8pair_f16(raw[b][8], raw[b][9], mb, nm);
                const v2f16 p5 = u8pair_f16(raw[b][10], raw[b][11], mb, nm);
                xt[b] = v8f16{p4[0], p4[1], p5[0], p5[1], p4[0], p4[1], p5[0], p5[1]};
            }
        }
    } else {
#pragma unroll
        for (int b = 0; b < NVB; ++b) {
            v4f16 h0, l0, h1, l1;
            const auto r = [&](int j) { return raw_value<IN>(raw[b][j]); };
            split4(r(0), r(1), r(2), r(3), s, h0, l0);
            split4(r(4), r(5), r(6), r(7), s, h1, l1);
            xh[b] = cat8(h0, h1);
            xl[b] = cat8(l0, l1);
            if constexpr (G::K48) {
                v4f16 h2, l2;
                split4(r(8), r(9), r(10), r(11), s, h2, l2);
                xt[b] = cat8(h2, XLO ? l2 : h2);        // 16-bit inputs: lo is 0
            }
        }
    }
    after_vpass();

    // --------------------------------------------------- vertical pass
    // fh[l] / fl[l]: entries 4 b .. 4 b + 3 = field l at row y0 + n, columns
    // x0 - RC + 16 b + 4 g .. +3 (the accumulator layout), split in f16.
    // One lobe at a time, the chains of the field blocks interleaved.
    // 1.0 in a register whose value the compiler does not see: v * one - hi
    // then stays one v_fma_mix (v - hi alone would be a convert, a subtract
    // and another convert)
    float one = 1.f;
    asm("" : "+v"(one));
    // vpass(l, ...) also returns, when `sum` is set, the sum of the field's
    // values: a NaN or an Inf among the strip's inputs makes it non-finite
    // (every input of a block meets every band entry in the products).
    auto vpass = [&](int l, vfield& fh, vfield& fl, bool sum) {
        const int t1 = MBASE + 8 * g - RC - m;
        const v8f16 ch1 = cat8(T.run[(0 * 3 + l) * MT4 + t1], T.run[(0 * 3 + l) * MT4 + t1 + 4]);
        const v8f16 cl1 = cat8(T.run[(1 * 3 + l) * MT4 + t1], T.run[(1 * 3 + l) * MT4 + t1 + 4]);
        v8f16 cht, clt;
        if constexpr (G::K48) {                 // rows 32 + 4 g + j: high taps for both halves
            const int t2 = MBASE + 32 + 4 * g - RC - m;
            const v4f16 th = T.run[(0 * 3 + l) * MT4 + t2];
            const v4f16 tl = T.run[(1 * 3 + l) * MT4 + t2];
            cht = XLO ? cat8(th, th) : cat8(th, tl);
            clt = cat8(tl, z4);
        }
        v4f32 acc[NVB];
#pragma unroll
        for (int b = 0; b < NVB; ++b) acc[b] = v4f32{0.f, 0.f, 0.f, 0.f};
        // smaller products first
        if constexpr (XLO) {
#pragma unroll
            for (int b = 0; b < NVB; ++b) acc[b] = TUCU_MFMA(xl[b], ch1, acc[b]);
        }
#pragma unroll
        for (int b = 0; b < NVB; ++b) acc[b] = TUCU_MFMA(xh[b], cl1, acc[b]);
        if constexpr (G::K48) {
            if constexpr (XLO) {
#pragma unroll
                for (int b = 0; b < NVB; ++b) acc[b] = TUCU_MFMA(xt[b], clt, acc[b]);   // hi x lo taps
            }
            // XLO: (hi + lo) x hi taps; otherwise hi x (hi + lo taps)
#pragma unroll
            for (int b = 0; b < NVB; ++b) acc[b] = TUCU_MFMA(xt[b], cht, acc[b]);
        }
#pragma unroll
        for (int b = 0; b < NVB; ++b) acc[b] = TUCU_MFMA(xh[b], ch1, acc[b]);
        float total = 0.f;
        if (sum) {
#pragma unroll
            for (int b = 0; b < NVB; ++b) total += (acc[b][0] + acc[b][1]) + (acc[b][2] + acc[b][3]);
        }
#pragma unroll
        for (int b = 0; b < NVB; ++b) {
            v4f16 h, lo;
            // hi = f16(v), lo = f16(v - hi): one mixed-precision fma each
            h = __builtin_convertvector(acc[b], v4f16);
#pragma unroll
            for (int i = 0; i < 4; ++i) lo[i] = _Float16(__builtin_fmaf(acc[b][i], one, -float(h[i])));
#pragma unroll
            for (int i = 0; i < 4; ++i) { fh[4 * b + i] = h[i]; fl[4 * b + i] = lo[i]; }
        }
        return total;
    };

    // ------------------------------------------------- horizontal pass
    const bool fast = y0 >= a.fy0 && y0 + 16 <= a.fy1 && x0 >= a.fx0 && x0 + MSTRIP_W <= a.fx1;
    const int y = y0 + m;
    constexpr int NOUT  = ONE ? 1 : 6;
    constexpr int HL[6] = {ONE ? 1 : 0, 1, 0, 2, 1, 0};   // horizontal table of output k
    constexpr int VL[6] = {0, 0, 1, 0, 1, 2};             // vertical field of output k
    auto hpass = [&](int k, const vfield& fh, const vfield& fl) {
        const int h = HL[k], v = VL[k];
        v4f32 acc[MNB];
#pragma unroll
        for (int o = 0; o < MNB; ++o) acc[o] = v4f32{0.f, 0.f, 0.f, 0.f};
        if constexpr (G::K48) {
            // Fixed pairs of field blocks P_p = (2 p, 2 p + 1), so that no
            // two B operands overlap: output block o takes P_(o/2) and
            // P_(o/2 + 1), whose columns lie d = 32 p - 16 o from its own
            // (d in {-16, 0, 16, 32}; taps beyond the lobe are zero).
            v8f16 ah[4], al[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int t = MBASE + 16 * i - 16 + 4 * g - RC - m;
                ah[i] = T.hpair[(0 * 3 + h) * MHP + t];
                al[i] = T.hpair[(1 * 3 + h) * MHP + t];
            }
            // smaller products first, as in the vertical pass
            static_for<0, MNB>([&](auto oc) { constexpr int o = oc, p = o / 2, i = (o & 1) ? 0 : 1;
                acc[o] = TUCU_MFMA(al[i], pairP<p>(fh), acc[o]); });
            static_for<0, MNB>([&](auto oc) { constexpr int o = oc, p = o / 2, i = (o & 1) ? 0 : 1;
                acc[o] = TUCU_MFMA(ah[i], pairP<p>(fl), acc[o]); });
            static_for<0, MNB>([&](auto oc) { constexpr int o = oc, p = o / 2 + 1, i = (o & 1) ? 2 : 3;
                acc[o] = TUCU_MFMA(al[i], pairP<p>(fh), acc[o]); });
            static_for<0, MNB>([&](auto oc) { constexpr int o = oc, p = o / 2 + 1, i = (o & 1) ? 2 : 3;
                acc[o] = TUCU_MFMA(ah[i], pairP<p>(fl), acc[o]); });
            static_for<0, MNB>([&](auto oc) { constexpr int o = oc, p = o / 2 + 1, i = (o & 1) ? 2 : 3;
                acc[o] = TUCU_MFMA(ah[i], pairP<p>(fh), acc[o]); });
            static_for<0, MNB>([&](auto oc) { constexpr int o = oc, p = o / 2, i = (o & 1) ? 0 : 1;
                acc[o] = TUCU_MFMA(ah[i], pairP<p>(fh), acc[o]); });
        } else {
            const int t1 = MBASE + 4 * g - RC - m;
            const v8f16 ch1 = T.hpair[(0 * 3 + h) * MHP + t1];
            const v8f16 cl1 = T.hpair[(1 * 3 + h) * MHP + t1];
            static_for<0, MNB>([&](auto oc) { constexpr int o = oc; acc[o] = TUCU_MFMA(cl1, pair8<o>(fh), acc[o]); });
            static_for<0, MNB>([&](auto oc) { constexpr int o = oc; acc[o] = TUCU_MFMA(ch1, pair8<o>(fl), acc[o]); });
            static_for<0, MNB>([&](auto oc) { constexpr int o = oc; acc[o] = TUCU_MFMA(ch1, pair8<o>(fh), acc[o]); });
        }
        const float u = a.mu[k] * inv_s;
        if constexpr (MODE == MODE_ICIA) {      // the gradient stays on chip
            sink(k, acc, u);
            return;
        }
        if (fast) {
            // Through the wave's LDS staging area, so that each store writes
            // 4 rows x 64 contiguous pixels instead of 16 rows x 16.
#pragma unroll
            for (int o = 0; o < MNB; ++o)
                *reinterpret_cast<v4f32*>(stage + m * MOPITCH + 16 * o + 4 * g) = acc[o] * u;
            __builtin_amdgcn_wave_barrier();
            if constexpr (CPOL == CPOL_NT) {
#pragma unroll
                for (int o = 0; o < MNB; ++o) {
                    const int row = 4 * o + g;
                    const v4f32 q = *reinterpret_cast<const v4f32*>(stage + row * MOPITCH + 4 * m);
                    const gptr<OUT> p = reinterpret_cast<gptr<OUT>>(
                        a.obase[k] + uintptr_t(int64_t(y0 + row) * a.ostride[k] + x0 + 4 * m) * sizeof(OUT));
                    Px<OUT>::st4nt(p, q[0], q[1], q[2], q[3]);    // written once: streaming stores
                }
            } else {
                // buffer stores carry the cache policy: rows from the strip's
                // first row (a uniform base), lane offsets in 32 bits (the host
                // keeps 16 rows of every output within 2^31 bytes, or no strip
                // takes this path)
                const __amdgpu_buffer_rsrc_t orsrc = __builtin_amdgcn_make_buffer_rsrc(
                    reinterpret_cast<void*>(a.obase[k] + uintptr_t(int64_t(y0) * a.ostride[k]) * sizeof(OUT)),
                    short(0), int(-1), BUF_DWORD3);
#pragma unroll
                for (int o = 0; o < MNB; ++o) {
                    const int row = 4 * o + g;
                    const v4f32 q = *reinterpret_cast<const v4f32*>(stage + row * MOPITCH + 4 * m);
                    const uint32_t off = uint32_t((row * a.ostride[k] + x0 + 4 * m) * int(sizeof(OUT)));
                    Px<OUT>::template st4buf<CPOL>(orsrc, off, q[0], q[1], q[2], q[3]);
                }
            }
            __builtin_amdgcn_wave_barrier();
        } else {
            const int rv = a.r[v], rh = a.r[h];
#pragma unroll
            for (int o = 0; o < MNB; ++o)
                store4<OUT>(a, k, false, rv, a.nrow - rv, rh, a.ncol - rh, y, x0 + 16 * o + 4 * g,
                            acc[o][0] * u, acc[o][1] * u, acc[o][2] * u, acc[o][3] * u);
        }
#ifdef TUCU_TIMELINE
        if constexpr (CPOL == CPOL_NT_WT && MODE == MODE_GAUSS && RC == 16 && std::is_same_v<IN, float> &&
                      std::is_same_v<OUT, float>)
            tl_stamp(3 + k);
#endif
    };

    // One field at a time, each followed by the outputs it feeds, so that
    // the first stores leave after a third of the vertical work.
    const unsigned fneed[3] = {ONE ? 1u : mask & 0x0bu, ONE ? 0u : mask & 0x14u, ONE ? 0u : mask & 0x20u};
    bool check = CHECK;                         // the first field computed vets the strip's input
#pragma unroll
    for (int l = 0; l < 3; ++l) {
        if (!fneed[l]) continue;
        vfield fh, fl;
        const float total = vpass(l, fh, fl, check);
        if (check) {
            check = false;
            redo |= __builtin_amdgcn_ballot_w64(!(__builtin_fabsf(total) <= 3.0e38f)) != 0;
        }
#pragma unroll
        for (int k = 0; k < NOUT; ++k)
            if (VL[k] == l && (mask >> k & 1u)) hpass(k, fh, fl);
    }
    // Rare: the products could not carry this strip's input.  Its outputs are
    // written again by the FP32 form, after the stores above have completed.
    if constexpr (CHECK) {
        if (redo) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            mfma_strip_fp32<RC, IN, OUT, MODE>(a, y0, x0, T, stage, mask);
        }
    }
}

// The same without a sink, for the outputs of a.mask (every mode that
// writes its outputs).
template <int RC, class IN, class OUT, int MODE = MODE_GAUSS, int CPOL = CPOL_NT, bool FRESH = false, class A, class F>
__device__ __forceinline__ void
mfma_compute(const A& a, int y0, int x0, MfmaRaw<IN>& raw, const MfmaTabs& T, float* stage, F&& after_vpass)
{
    mfma_compute<RC, IN, OUT, MODE, CPOL, FRESH>(a, y0, x0, raw, T, stage, after_vpass, 0, a.mask);
}
#undef TUCU_MFMA

// Workgroup tiles (strip-major: tile = tx * tiles_y + ty) hold WY strips
// down by 4 / WY across, i.e. 16 WY rows x 256 / WY columns.  Every shape
// filters the same 16 x 64 strips (at rows cy0 + 16 i, columns cx0 + 64 j),
// so the shape changes no result.  Position of this wave's strip in `tile`;
// false if the strip lies outside the region.
template <int WY, class A>
__device__ __forceinline__ bool mfma_strip_pos(const A& a, int tile, int& y0, int& x0)
{
    constexpr int WX = MWAVES / WY;
    const int tx = tile / a.tiles_y;
    const int ty = tile - tx * a.tiles_y;
    const int w  = int(threadIdx.x) >> 6;
    // wave-uniform, and held in scalar registers so that every test on the
    // strip's position is a scalar branch: a divergent if / else around the
    // loads makes the compiler wait for the loads already in flight before
    // the other side's address arithmetic
    y0 = __builtin_amdgcn_readfirstlane(a.cy0 + ty * (16 * WY) + 16 * (w / WX));
    x0 = __builtin_amdgcn_readfirstlane(a.cx0 + tx * (MSTRIP_W * WX) + MSTRIP_W * (w % WX));
    return y0 < a.cy1 && x0 < a.cx1;
}

// First tile of XCD x's range when ntiles tiles are shared by grid
// workgroups: ranges proportional to the workgroups each XCD receives, so
// that with as many workgroups as tiles every workgroup takes exactly one.
__device__ __forceinline__ int mfma_share_lo(int ntiles, int grid, int x)
{
    const int q = grid / QN, r = grid - q * QN;
    return int(int64_t(ntiles) * (q * x + min(x, r)) / grid);
}

// Reads, in one round of scalar loads, every kernel argument a strip's
// loads depend on.  Left to itself the compiler reads each where it is first
// used, behind branches that need the one before: six dependent round trips
// to the argument block before a wave could issue its first input load
// (1080p six outputs: 2 % of the launch).
template <class A>
__device__ __forceinline__ void mfma_args_early(const A& a)
{
    const int grid = int(gridDim.x);
    const int ty = a.tiles_y, cy0 = a.cy0, cy1 = a.cy1, cx0 = a.cx0, cx1 = a.cx1;
    const int nr = a.nrow, nc = a.ncol, sb = a.in_stride_b;
    const uint32_t nb = a.in_bytes;
    const void* const in = a.in;
    const void* const tabs = a.tabs;
    asm volatile("" ::"s"(grid), "s"(ty), "s"(cy0), "s"(cy1), "s"(cx0), "s"(cx1), "s"(nr), "s"(nc), "s"(sb), "s"(nb),
                 "s"(in), "s"(tabs));
}

// One image, one workgroup per tile (images with no more tiles than
// resident workgroups, such as 1080p).
template <int RC, int WY, class IN, class OUT, int MODE = MODE_GAUSS>
__global__ void __launch_bounds__(64 * MWAVES) __attribute__((amdgpu_waves_per_eu(MFMA_WAVES_PER_SIMD<IN>)))
sepfir_mfma_kernel(const SepArgs a)
{
    __shared__ MfmaTabs T;
    __shared__ v4f32 stage4[MWAVES * 16 * MOPITCH / 4];
#ifdef TUCU_TIMELINE
    constexpr bool TL = MODE == MODE_GAUSS && RC == 16 && std::is_same_v<IN, float> && std::is_same_v<OUT, float>;
    const auto stamp = [](int i) {
        if constexpr (TL) {
            __builtin_amdgcn_sched_barrier(0);
            tl_stamp(i);
            __builtin_amdgcn_sched_barrier(0);
        }
    };
#else
    const auto stamp = [](int) {};
#endif
    stamp(0);
    mfma_args_early(a);
    int y0, x0;
    const bool live = mfma_strip_pos<WY>(a, xcd_remap(blockIdx.x, gridDim.x), y0, x0);
    MfmaRaw<IN> raw;
    if (live) mfma_load<RC, IN>(a, y0, x0, raw);    // in flight while the tables are built
    stamp(10);
    mfma_get_tabs<MODE>(a, T);
    stamp(11);
    __syncthreads();
    stamp(1);
    if (!live) return;
    // one round of workgroups: outputs written through L2 (CPOL_NT_WT)
    mfma_compute<RC, IN, OUT, MODE, CPOL_NT_WT>(a, y0, x0, raw, T,
                                                reinterpret_cast<float*>(stage4) + (threadIdx.x >> 6) * 16 * MOPITCH,
                                                [&] { stamp(2); });
    stamp(9);
}

// One large image: a grid of the resident workgroups; each builds the tap
// tables once and filters a static share of the tiles: the tile list
// (strip-major, so consecutive tiles are vertically adjacent) is cut into one
// contiguous range per XCD, and the workgroups of XCD x take the tiles of
// range x in turn, so neighbouring tiles, which read the same halo rows, run
// at the same time on one L2.  The tables are read-only after the build and
// each wave has its own staging area, so the tiles need no barrier between
// them.  (A 4096 x 4096 FIRFilter2 launch: 21.5 us against 25.5 us with one
// workgroup per tile, which rebuilds the tables 4096 times.)
template <int RC, int WY, class IN, class OUT, int MODE = MODE_GAUSS>
__global__ void __launch_bounds__(64 * MWAVES) __attribute__((amdgpu_waves_per_eu(MFMA_LOOP_WAVES<MODE, IN>)))
sepfir_mfma_loop_kernel(const SepArgs a, int ntiles)
{
    __shared__ MfmaTabs T;
    __shared__ v4f32 stage4[MWAVES * 16 * MOPITCH / 4];
    mfma_args_early(a);
    const int grid = int(gridDim.x), b = int(blockIdx.x);
    const int x = b % QN, step = (grid - x + QN - 1) / QN;           // workgroups on this XCD
    const int hi = mfma_share_lo(ntiles, grid, x + 1);
    int t = mfma_share_lo(ntiles, grid, x) + b / QN;
    float* const stage = reinterpret_cast<float*>(stage4) + (threadIdx.x >> 6) * 16 * MOPITCH;
    int y0, x0;
    bool live = t < hi && mfma_strip_pos<WY>(a, t, y0, x0);
    MfmaRaw<IN> raw;
    if (live) mfma_load<RC, IN>(a, y0, x0, raw);    // in flight while the tables are built
    mfma_get_tabs<MODE>(a, T);
    __syncthreads();
    // The six-output strip code for float input takes its lane index afresh
    // in each tile: left to itself the compiler computes the lane-derived
    // offsets once, before the loop, and holds them through it (3840 x 2160
    // six outputs 1 % faster).  The one-output loop ran 2.5-6 % slower so,
    // and u8 input at radius class 8 3.5 % slower (its registers then fit
    // three waves per SIMD, and the grid, the resident workgroups, grew by
    // half): they keep the hoisted offsets.
    constexpr bool FRESH = MODE == MODE_GAUSS && sizeof(IN) == 4;
    while (t < hi) {                                // uniform over the workgroup
        if (live) mfma_compute<RC, IN, OUT, MODE, CPOL_NT, FRESH>(a, y0, x0, raw, T, stage, [] {});
        t += step;
        live = t < hi && mfma_strip_pos<WY>(a, t, y0, x0);
        if (live) mfma_load<RC, IN, FRESH>(a, y0, x0, raw);
    }
}

// Workgroups resident at once on the whole GPU for this instantiation.
template <class K>
int mfma_resident(K kernel)
{
    int dev = 0, ncu = 0, per_cu = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kernel), 64 * MWAVES, 0)
            != hipSuccess || per_cu < 1)
        per_cu = 1;
    return std::max(1, ncu) * per_cu;
}

template <int RC, int WY, class IN, class OUT, int MODE = MODE_GAUSS>
hipError_t launch_sepfir_mfma(const SepArgs& a, int nblocks, hipStream_t s)
{
    static const int resident = mfma_resident(sepfir_mfma_loop_kernel<RC, WY, IN, OUT, MODE>);
    if (nblocks <= resident)
        hipLaunchKernelGGL((sepfir_mfma_kernel<RC, WY, IN, OUT, MODE>), dim3(nblocks), dim3(64 * MWAVES), 0, s, a);
    else
        hipLaunchKernelGGL((sepfir_mfma_loop_kernel<RC, WY, IN, OUT, MODE>), dim3(resident), dim3(64 * MWAVES), 0, s,
                           a, nblocks);
    return hipGetLastError();
}

// One-output launches (MODE_TAPS), given the grid of the tile loop's shape
// WYL: an image whose 16 WYR-row tiles fit the resident workgroups of the
// one-round kernel gets one workgroup per such tile (1080p 17 x 17 lobes,
// 1 x 4 strips: 7.0 against 7.3 us with 2 x 2); a larger one the tile loop
// with WYL (the 4096^2 u8 launches 3-7 % slower with 1 x 4 than with 2 x 2).
template <int RC, int WYR, int WYL, class IN, class OUT>
hipError_t launch_fir_mfma(const SepArgs& a, int nblocks, hipStream_t s)
{
    static const int resident = mfma_resident(sepfir_mfma_kernel<RC, WYR, IN, OUT, MODE_TAPS>);
    const int ty = (a.cy1 - a.cy0 + 16 * WYR - 1) / (16 * WYR);
    const int tx = (a.cx1 - a.cx0 + MSTRIP_W * (MWAVES / WYR) - 1) / (MSTRIP_W * (MWAVES / WYR));
    const int64_t n = int64_t(ty) * tx;
    if (n > 0 && n <= resident) {
        SepArgs r = a;
        r.tiles_y = ty;
        hipLaunchKernelGGL((sepfir_mfma_kernel<RC, WYR, IN, OUT, MODE_TAPS>), dim3(int(n)), dim3(64 * MWAVES), 0, s,
                           r);
        return hipGetLastError();
    }
    return launch_sepfir_mfma<RC, WYL, IN, OUT, MODE_TAPS>(a, nblocks, s);
}

// Matrix-core variant of batch_kernel: the same tile queue; the tap
// tables in LDS are rebuilt when the job changes (tiles of a job are
// contiguous in the list).
template <int WY, class IN, class OUT>
__global__ void __launch_bounds__(64 * MWAVES) __attribute__((amdgpu_waves_per_eu(MFMA_WAVES_PER_SIMD<IN>)))
batch_mfma_kernel(const SepArgs* __restrict__ jobs_, const BatchTile* __restrict__ tiles, int ntiles,
                  int* __restrict__ counter)
{
    __shared__ MfmaTabs T;
    __shared__ v4f32 stage4[MWAVES * 16 * MOPITCH / 4];
    __shared__ int next;
    ConstArgs* const jobs = (ConstArgs*)(uintptr_t)jobs_;
    float* const stage = reinterpret_cast<float*>(stage4) + (threadIdx.x >> 6) * 16 * MOPITCH;
    int cur = -1;
    TileQueue q;
    int ti = q.first(ntiles);
    for (;;) {
        if (threadIdx.x == 0)          // the answer is read after this tile
            next = q.grab(counter, ntiles);
        if (ti >= 0) {
            const BatchTile e = tiles[ti];
            ConstArgs& a = jobs[e.job];
            if (int(e.job) != cur) {   // uniform over the workgroup
                mfma_build_tabs(a, T);
                cur = int(e.job);
                __syncthreads();
            }
            int y0, x0;
            if (mfma_strip_pos<WY>(a, int(e.tile), y0, x0)) {
                MfmaRaw<IN> raw;
                if (a.rt <= 8) {
                    mfma_load<8, IN>(a, y0, x0, raw);
                    mfma_compute<8, IN, OUT>(a, y0, x0, raw, T, stage, [] {});
                } else {
                    mfma_load<16, IN>(a, y0, x0, raw);
                    mfma_compute<16, IN, OUT>(a, y0, x0, raw, T, stage, [] {});
                }
            }
        }
        __syncthreads();               // T reuse; `next` was written before this barrier
        ti = __builtin_amdgcn_readfirstlane(next);
        __syncthreads();               // all have read `next` before thread 0 writes it again
        if (ti >= ntiles) break;
    }
}

// Batched, one wave per workgroup: a tile is one strip (16 x 64; tile =
// tx * tiles_y + ty), every wave draws its own strips from the counter, so
// the waves of a SIMD drift apart and their loads, matrix work and stores
// overlap.  Each wave keeps its own tap tables.
template <class IN, class OUT>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2)))   // LDS allows no more
batch_mfma_wave_kernel(const SepArgs* __restrict__ jobs_, const BatchTile* __restrict__ tiles, int ntiles,
                       int* __restrict__ counter)
{
    __shared__ MfmaTabs T;
    __shared__ v4f32 stage4[16 * MOPITCH / 4];
    ConstArgs* const jobs = (ConstArgs*)(uintptr_t)jobs_;
    int cur = -1;
    TileQueue q;
    int ti = q.first(ntiles);
    for (;;) {
        int nxt = 0;
        if (threadIdx.x == 0)          // the answer is read after this strip
            nxt = q.grab(counter, ntiles);
        if (ti >= 0) {
            const BatchTile e = tiles[ti];
            ConstArgs& a = jobs[e.job];
            if (int(e.job) != cur) {   // one wave: its barrier orders its LDS accesses
                mfma_build_tabs(a, T);
                cur = int(e.job);
                __syncthreads();
            }
            const int tx = int(e.tile) / a.tiles_y;
            const int y0 = __builtin_amdgcn_readfirstlane(a.cy0 + 16 * (int(e.tile) - tx * a.tiles_y));
            const int x0 = __builtin_amdgcn_readfirstlane(a.cx0 + MSTRIP_W * tx);
            if (y0 < a.cy1 && x0 < a.cx1) {
                MfmaRaw<IN> raw;
                float* const st = reinterpret_cast<float*>(stage4);
                if (a.rt <= 8) {
                    mfma_load<8, IN>(a, y0, x0, raw);
                    mfma_compute<8, IN, OUT>(a, y0, x0, raw, T, st, [] {});
                } else {
                    mfma_load<16, IN>(a, y0, x0, raw);
                    mfma_compute<16, IN, OUT>(a, y0, x0, raw, T, st, [] {});
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
        ti = __builtin_amdgcn_readfirstlane(nxt);
        if (ti >= ntiles) break;
    }
}

template <class IN, class OUT>
hipError_t launch_batch_mfma_wave(const SepArgs* jobs, const BatchTile* tiles, int ntiles, int* counter,
                                  int grid, hipStream_t s)
{
    hipLaunchKernelGGL(reset_queue<>, dim3(1), dim3(64), 0, s, counter, grid);
    hipLaunchKernelGGL((batch_mfma_wave_kernel<IN, OUT>), dim3(grid), dim3(64), 0, s, jobs, tiles, ntiles, counter);
    return hipGetLastError();
}

template <int WY, class IN, class OUT>
hipError_t launch_batch_mfma(const SepArgs* jobs, const BatchTile* tiles, int ntiles, int* counter,
                             int grid, hipStream_t s)
{
    hipLaunchKernelGGL(reset_queue<>, dim3(1), dim3(64), 0, s, counter, grid);
    hipLaunchKernelGGL((batch_mfma_kernel<WY, IN, OUT>), dim3(grid), dim3(64 * MWAVES), 0, s, jobs, tiles, ntiles, counter);
    return hipGetLastError();
}

// Workgroup shapes: WY = 1, 2, 4 strips down (tile rows 16, 32, 64).
constexpr int MFMA_RADII[2] = {8, 16};
constexpr int NMWY = 3;
constexpr int MFMA_WYS[NMWY] = {1, 2, 4};
// Default shapes (indices into MFMA_WYS), measured: the six-output kernels
// 1 x 4 strips, 64 x 64 pixels (1080p f32 / bf16 / f16 2.5-4 % faster than
// 2 x 2, 3840 x 2160 3 %: the strips of a workgroup share their vertical
// halo rows in L1); the one-output FIR / general-taps tile loop 2 x 2, 32 x
// 128 (the 4096^2 u8 launches 2.5-7 % slower with 1 x 4), its one-round
// launches 1 x 4 (launch_fir_mfma); the ICIA pass 1 x 4 (1080p f32 sigma 3
// projective 8 %, 3840 x 2160 u8 1-3 % faster, 1080p sigma 2 2 % slower).
constexpr int MFMA_WY_DEFAULT = 2;
constexpr int MFMA_WY_ONE = 1;
constexpr int MFMA_WY_ICIA = 2;

extern KernelEntry g_gauss_mfma_table[2][NMWY][4][4];
extern KernelEntry g_fir_mfma_table[2][4][4];       // MODE_TAPS, default shape: [radius class][in][out]

}   // namespace tucu
