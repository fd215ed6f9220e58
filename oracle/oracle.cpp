// oracle.cpp -- CPU restatement of TUCuda's separable FIR / Gaussian-derivative
// path.  TEST INFRASTRUCTURE ONLY: the parity tests, the smoke check and the
// CPU-baseline leg of the benchmark load this library as the checker.  The
// product (libtucuda_amd.so) never links, loads or calls it.
//
// Two flavours are restated:
//   * "f32": faithful to the reference arithmetic -- float coefficients made by
//     the reference rule, an FP32 intermediate image, horizontal pass first,
//     tap sums evaluated in source order with every product and sum rounded
//     separately (this file is compiled with -ffp-contract=off).
//   * "f64": the same algorithm in double precision (coefficients, conversion
//     and accumulation), used as the precision reference.
//
// Reference files followed (paths relative to the TUCuda tree):
//   TU/cu/FIRFilter.h:37-40      offsets / output sizes
//   TU/cu/FIRFilter.h:66-176     paired tap sums, lobe[0] = outermost tap
//   TU/cu/FIRFilter.h:181-231    buf[y][x] = window starting at x; valid-only writes
//   TU/cu/FIRFilter.h:245-261    lobe length limit (17)
//   TU/cu/FIRFilter.h:286-287    shifted output slice
//   TU/cu/FIRFilter.h:298-377    early return, intermediate type, size dispatch
//   TU/cu/FIRGaussianConvolver.h:123-170  Gaussian lobes and their normalisation
//                                         (also applied to the ceil(4 sigma) radius rule)
//   TU/cu/FIRGaussianConvolver.h:178-242  (H, V) lobe pairs of the six methods
//   FIRGaussianConvolver.cu:52-98         detail::lobeSize (1 % threshold, size classes)
//   TU/cu/algorithm.h:119-144             element conversion on tile load
//   TU/cu/fp16.h:55-80                    round-to-nearest-even half conversion
//   TU/cu/algorithm.h:183-242             subsample: out[y][x] = op at (2y, 2x), early return
//   TU/cu/functional.h:58-256             the stencil operators and their border guards
//   TU/cu/ICIA.h:61-112, 352-369          ICIA's error moment over the gradient images
//   TU/cu/vec.h:1285-1290, 1418-1422      image_derivative0 of the projective / affine maps
//   TU/cu/array.h:101-118, 319-325        array::ext, the upper triangle of an outer product
//
// Status codes: 0 ok, 2 lobe longer than 17, 3 unsupported lobe length,
// 1 invalid argument.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstddef>
#include <thread>
#include <vector>
#include <algorithm>

namespace {

enum { DT_U8 = 0, DT_F32 = 1, DT_F16 = 2, DT_BF16 = 3 };
constexpr size_t LobeSizeMax = 17;          // FIRFilter.h:26

// ---------------------------------------------------------------- conversions
inline float half_to_float(uint16_t h)
{
    const uint32_t sign = uint32_t(h & 0x8000u) << 16;
    uint32_t exp  = (h >> 10) & 0x1f;
    uint32_t mant = h & 0x3ffu;
    uint32_t bits;
    if (exp == 0) {
        if (mant == 0) bits = sign;
        else {                                  // subnormal: normalise
            int e = -1;
            do { ++e; mant <<= 1; } while (!(mant & 0x400u));
            bits = sign | uint32_t(127 - 15 - e) << 23 | (mant & 0x3ffu) << 13;
        }
    } else if (exp == 0x1f) {
        bits = sign | 0x7f800000u | mant << 13;
    } else {
        bits = sign | (exp + 127 - 15) << 23 | mant << 13;
    }
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}

// Round-to-nearest-even float -> binary16 (the rounding fp16.h:60-61 asks the
// F16C unit for).  Written from the IEEE-754 definition.
inline uint16_t float_to_half(float f)
{
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t ax = x & 0x7fffffffu;
    if (ax > 0x7f800000u) return uint16_t(sign | 0x7e00u);          // NaN
    if (ax >= 0x477ff000u) return uint16_t(sign | 0x7c00u);         // overflow -> inf
    if (ax < 0x38800000u) {                                          // result subnormal or zero
        // value = m * 2^(e-150); in units of the half subnormal step 2^-24
        // that is m * 2^(e-126), so drop (126 - e) bits and round.
        if (ax < 0x33000000u) return uint16_t(sign);                 // < 2^-25 rounds to 0
        const uint32_t e = ax >> 23;
        const uint32_t m = (ax & 0x7fffffu) | 0x800000u;
        const int shift = 126 - int(e);                              // 14 .. 24
        const uint32_t q = m >> shift;
        const uint32_t rem = m & ((1u << shift) - 1);
        const uint32_t halfway = 1u << (shift - 1);
        uint32_t r = q;
        if (rem > halfway || (rem == halfway && (q & 1))) ++r;
        return uint16_t(sign | r);
    }
    uint32_t h = ((ax >> 23) - 127 + 15) << 10 | ((ax >> 13) & 0x3ffu);
    const uint32_t rem = ax & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1))) ++h;           // may carry into exponent
    return uint16_t(sign | h);
}

inline float bf16_to_float(uint16_t b)
{
    const uint32_t bits = uint32_t(b) << 16;
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}

inline uint16_t float_to_bf16(float f)
{
    uint32_t x;
    std::memcpy(&x, &f, 4);
    if ((x & 0x7fffffffu) > 0x7f800000u) return uint16_t((x >> 16) | 0x40u);
    x += 0x7fffu + ((x >> 16) & 1u);
    return uint16_t(x >> 16);
}

// float -> u_char store (FIRFilter.h:203, 229 store a float into a u_char
// element).  Out-of-range values are undefined in C++; this restatement
// defines the conversion as saturate to [0,255], then truncate toward zero,
// with NaN -> 0.
template <class A> inline uint8_t to_u8(A v)
{
    if (!(v > A(0))) return 0;
    if (v >= A(255)) return 255;
    return uint8_t(int(v));
}

template <class A> inline A load_px(const void* base, int dt, int64_t idx)
{
    switch (dt) {
      case DT_U8:   return A(static_cast<const uint8_t*>(base)[idx]);
      case DT_F32:  return A(static_cast<const float*>(base)[idx]);
      case DT_F16:  return A(half_to_float(static_cast<const uint16_t*>(base)[idx]));
      default:      return A(bf16_to_float(static_cast<const uint16_t*>(base)[idx]));
    }
}

template <class A> inline void store_px(void* base, int dt, int64_t idx, A v)
{
    switch (dt) {
      case DT_U8:   static_cast<uint8_t*>(base)[idx] = to_u8(v); break;
      case DT_F32:  static_cast<float*>(base)[idx] = float(v); break;
      case DT_F16:  static_cast<uint16_t*>(base)[idx] = float_to_half(float(v)); break;
      default:      static_cast<uint16_t*>(base)[idx] = float_to_bf16(float(v)); break;
    }
}

inline bool supported_lobe(int n)
{
    return n == 17 || n == 16 || n == 9 || n == 8 || n == 5 || n == 4 || n == 3 || n == 2;
}

// Paired tap sum of FIRFilter.h:66-176 over the window in(0) .. in(2r),
// r = L & ~1.  Odd L: symmetric kernel, lobe[r] weights the centre.  Even L:
// antisymmetric kernel (left minus right), no centre tap.  Evaluated left to
// right as the source expression is.
template <class A, class GET> inline A paired_sum(const GET& in, const A* lobe, int L)
{
    const int r = L & ~1;
    A acc;
    if (L & 1) {
        acc = lobe[0] * (in(0) + in(2 * r));
        for (int i = 1; i < r; ++i) acc = acc + lobe[i] * (in(i) + in(2 * r - i));
        acc = acc + lobe[r] * in(r);
    } else {
        acc = lobe[0] * (in(0) - in(2 * r));
        for (int i = 1; i < r; ++i) acc = acc + lobe[i] * (in(i) - in(2 * r - i));
    }
    return acc;
}

template <class F> void parallel_rows(int n, int nthreads, const F& f)
{
    if (nthreads <= 1 || n < 2 * nthreads) { f(0, n); return; }
    std::vector<std::thread> th;
    const int chunk = (n + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; ++t) {
        const int b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        th.emplace_back([&f, b, e] { f(b, e); });
    }
    for (auto& x : th) x.join();
}

// FIRFilter2::convolve (FIRFilter.h:298-377) in accumulate type A.
template <class A>
int fir2(const A* lobeH, int nH, const A* lobeV, int nV,
         const void* in, int in_dt, int64_t in_stride, int nrow, int ncol,
         void* out, int out_dt, int64_t out_stride, int shift, int nthreads)
{
    if (nH < 0 || nV < 0 || nrow < 0 || ncol < 0) return 1;
    if (size_t(nH) > LobeSizeMax || size_t(nV) > LobeSizeMax) return 2;   // :250-251
    const int rV = nV & ~1, rH = nH & ~1;
    if (nrow <= 2 * rV) return 0;                                          // :304-307
    if (ncol <= 2 * rH) return 0;                                          // :309-312
    if (!supported_lobe(nH) || !supported_lobe(nV)) return 3;              // :343-344, 374-375
    const int ncb = ncol - 2 * rH;
    std::vector<A> buf(size_t(nrow) * ncb);                                // _buf, T = float
    // Horizontal pass (fir_filterH, :181-205): buf[y][x] over in[y][x .. x+2rH].
    parallel_rows(nrow, nthreads, [&](int yb, int ye) {
        std::vector<A> row(ncol);
        for (int y = yb; y < ye; ++y) {
            for (int x = 0; x < ncol; ++x) row[x] = load_px<A>(in, in_dt, int64_t(y) * in_stride + x);
            A* b = &buf[size_t(y) * ncb];
            for (int x = 0; x < ncb; ++x) {
                const A* w = &row[x];
                b[x] = paired_sum<A>([w](int i) { return w[i]; }, lobeH, nH);
            }
        }
    });
    // Vertical pass (fir_filterV, :207-231) into the shifted slice (:286-287).
    const int oy = shift ? rV : 0, ox = shift ? rH : 0;
    parallel_rows(nrow - 2 * rV, nthreads, [&](int yb, int ye) {
        for (int y = yb; y < ye; ++y)
            for (int x = 0; x < ncb; ++x) {
                const A* c = &buf[size_t(y) * ncb + x];
                const A v = paired_sum<A>([c, ncb](int i) { return c[size_t(i) * ncb]; }, lobeV, nV);
                store_px<A>(out, out_dt, int64_t(y + oy) * out_stride + (x + ox), v);
            }
    });
    return 0;
}

// General taps (an extension with no counterpart in the reference): the
// FIRFilter2 window convention with an arbitrary odd number of taps,
// tap i of the window weighting in[x + i], summed left to right.
template <class A>
int fir2_taps(const A* tapsH, int nH, const A* tapsV, int nV,
              const void* in, int in_dt, int64_t in_stride, int nrow, int ncol,
              void* out, int out_dt, int64_t out_stride, int shift, int nthreads)
{
    if (nH < 1 || nV < 1 || !(nH & 1) || !(nV & 1) || nH > 33 || nV > 33) return 3;
    const int rV = nV / 2, rH = nH / 2;
    if (nrow <= 2 * rV || ncol <= 2 * rH) return 0;
    const int ncb = ncol - 2 * rH;
    std::vector<A> buf(size_t(nrow) * ncb);
    parallel_rows(nrow, nthreads, [&](int yb, int ye) {
        std::vector<A> row(ncol);
        for (int y = yb; y < ye; ++y) {
            for (int x = 0; x < ncol; ++x) row[x] = load_px<A>(in, in_dt, int64_t(y) * in_stride + x);
            for (int x = 0; x < ncb; ++x) {
                A acc = tapsH[0] * row[x];
                for (int j = 1; j < nH; ++j) acc = acc + tapsH[j] * row[x + j];
                buf[size_t(y) * ncb + x] = acc;
            }
        }
    });
    const int oy = shift ? rV : 0, ox = shift ? rH : 0;
    parallel_rows(nrow - 2 * rV, nthreads, [&](int yb, int ye) {
        for (int y = yb; y < ye; ++y)
            for (int x = 0; x < ncb; ++x) {
                A acc = tapsV[0] * buf[size_t(y) * ncb + x];
                for (int i = 1; i < nV; ++i) acc = acc + tapsV[i] * buf[size_t(y + i) * ncb + x];
                store_px<A>(out, out_dt, int64_t(y + oy) * out_stride + (x + ox), acc);
            }
    });
    return 0;
}

// cu::subsample's operators (functional.h:58-256) at input pixel (y, x) of
// an FP32 image: the guarded operator() of each, i.e. 0 unless the stencil
// fits inside the image, then the source expression in source order.
float subsample_op(int op, const std::vector<float>& im, int nrow, int ncol, int y, int x)
{
    static const int RY[12] = {0, 0, 1, 0, 1, 1, 1, 1, 1, 1, 1, 2};
    static const int RX[12] = {0, 1, 0, 1, 0, 1, 1, 1, 1, 1, 1, 2};
    const int ry = RY[op], rx = RX[op];
    if (y < ry || y + ry >= nrow || x < rx || x + rx >= ncol) return 0.f;
    auto in = [&](int yy, int xx) { return im[size_t(yy) * ncol + xx]; };
    auto dHH = [&] { return in(y, x - 1) - 2.f * in(y, x) + in(y, x + 1); };        // :104-108
    auto dVV = [&] { return in(y - 1, x) - 2.f * in(y, x) + in(y + 1, x); };        // :120-124
    auto dHV = [&] {                                                               // :136-141
        return 0.25f * (in(y - 1, x - 1) - in(y - 1, x + 1) - in(y + 1, x - 1) + in(y + 1, x + 1));
    };
    auto sH = [&] {                                                                // :154-159
        return 0.125f * (in(y - 1, x + 1) - in(y - 1, x - 1) + in(y + 1, x + 1) - in(y + 1, x - 1))
             + 0.250f * (in(y, x + 1) - in(y, x - 1));
    };
    auto sV = [&] {                                                                // :173-178
        return 0.125f * (in(y + 1, x - 1) - in(y - 1, x - 1) + in(y + 1, x + 1) - in(y - 1, x + 1))
             + 0.250f * (in(y + 1, x) - in(y - 1, x));
    };
    // gaussian5x5::smooth (:246-253), weights {0.375, 0.25, 0.0625}
    auto g5 = [](const float* a, int i) {
        return 0.375f * a[i] + 0.25f * (a[i - 1] + a[i + 1]) + 0.0625f * (a[i - 2] + a[i + 2]);
    };
    switch (op) {
      case 0:  return in(y, x);                                                    // :61-64
      case 1:  return 0.5f * (in(y, x + 1) - in(y, x - 1));                         // :71-74
      case 2:  return 0.5f * (in(y + 1, x) - in(y - 1, x));                         // :87-90
      case 3:  return dHH();
      case 4:  return dVV();
      case 5:  return dHV();
      case 6:  return sH();
      case 7:  return sV();
      case 8:  return std::fabs(sH()) + std::fabs(sV());                           // :192-197
      case 9:  return in(y, x - 1) + in(y, x + 1) + in(y - 1, x) + in(y + 1, x) - 4.f * in(y, x);   // :204-209
      case 10: { const float dxy = dHV(); return dHH() * dVV() - dxy * dxy; }      // :216-225
      default: {                                                                   // :233-244
        float b[5];
        for (int i = 0; i < 5; ++i) b[i] = g5(&im[size_t(y - 2 + i) * ncol], x);
        return g5(b, 2);
      }
    }
}

// cu::subsample(in, ie, out, op) (algorithm.h:183-242): out[y][x] = op at
// (2y, 2x) for y < nrow/2, x < ncol/2; fewer than 2 rows or columns: no-op.
int subsample(const void* in, int in_dt, int64_t in_stride, int nrow, int ncol,
              void* out, int out_dt, int64_t out_stride, int op)
{
    if (nrow < 0 || ncol < 0 || op < 0 || op > 11) return 1;
    if (nrow < 2 || ncol < 2) return 0;                                    // :227-233
    std::vector<float> im(size_t(nrow) * ncol);
    for (int y = 0; y < nrow; ++y)
        for (int x = 0; x < ncol; ++x) im[size_t(y) * ncol + x] = load_px<float>(in, in_dt, int64_t(y) * in_stride + x);
    for (int y = 0; y < nrow / 2; ++y)
        for (int x = 0; x < ncol / 2; ++x)
            store_px<float>(out, out_dt, int64_t(y) * out_stride + x, subsample_op(op, im, nrow, ncol, 2 * y, 2 * x));
    return 0;
}

// detail::lobeSize (FIRGaussianConvolver.cu:52-98), generic in the element type.
template <class A> size_t lobe_size(const A lobe[], bool even)
{
    const A epsilon = A(0.01);                  // .cu:58 (a float constant in the reference)
    A th = 0;
    for (size_t i = LobeSizeMax; i-- > 0; )
        if (std::abs(lobe[i]) >= th) th = std::abs(lobe[i]);
    th *= epsilon;
    size_t siz;
    for (siz = LobeSizeMax; siz-- > 0; )
        if (std::abs(lobe[siz]) > th) { ++siz; break; }
    if (even) {
        if (siz <= 2) return 3;
        else if (siz <= 5) return 5;
        else if (siz <= 9) return 9;
        else return 17;
    } else {
        if (siz <= 1) return 2;
        else if (siz <= 4) return 4;
        else if (siz <= 8) return 8;
        else return 16;
    }
}

// FIRGaussianConvolver2::initialize(sigma, lobe0, lobe1, lobe2)
// (FIRGaussianConvolver.h:123-170), generic in the element type.
template <class A>
void gauss_lobes(A sigma, A* lobe0, int* n0, A* lobe1, int* n1, A* lobe2, int* n2)
{
    A lb0[LobeSizeMax], lb1[LobeSizeMax], lb2[LobeSizeMax];
    for (size_t i = 0; i < LobeSizeMax; ++i) {
        const A dx = A(i) / sigma, dxdx = dx * dx;
        lb0[i] = std::exp(A(-0.5) * dxdx);
        lb1[i] = -dx * lb0[i];
        lb2[i] = (dxdx - A(1)) * lb0[i];
    }
    const size_t s0 = lobe_size(lb0, true);
    A sum = lb0[0];
    for (size_t i = 1; i < s0; ++i) sum += (A(2) * lb0[i]);
    for (size_t i = 0; i < s0; ++i) lobe0[i] = lb0[s0 - 1 - i] / std::abs(sum);

    const size_t s1 = lobe_size(lb1, false);
    sum = A(0);
    for (size_t i = 0; i < s1; ++i) sum += (A(2) * A(i) * lb1[i]);
    for (size_t i = 0; i < s1; ++i) lobe1[i] = lb1[s1 - i] / std::abs(sum);

    const size_t s2 = lobe_size(lb2, true);
    sum = A(0);
    for (size_t i = 1; i < s2; ++i) sum += (A(i * i) * lb2[i]);
    for (size_t i = 0; i < s2; ++i) lobe2[i] = lb2[s2 - 1 - i] / std::abs(sum);

    *n0 = int(s0); *n1 = int(s1); *n2 = int(s2);
}

// Lobes for a radius fixed at r = ceil(4 sigma), clamped to [1, 16] (the
// batched and served workloads ask for it; the reference has no such rule).
// Symmetric lobes 0 and 2 get r + 1 entries, the antisymmetric lobe 1 gets r,
// and each is normalised by the formula FIRGaussianConvolver.h:141-163
// applies to the lengths lobeSize picks (lobe 1 keeps its off-by-one: taps
// at offsets r .. 1, normalised over offsets 0 .. r-1).
template <class A>
void gauss_lobes_ceil4(A sigma, A* lobe0, int* n0, A* lobe1, int* n1, A* lobe2, int* n2)
{
    A lb0[LobeSizeMax], lb1[LobeSizeMax], lb2[LobeSizeMax];
    for (size_t i = 0; i < LobeSizeMax; ++i) {
        const A dx = A(i) / sigma, dxdx = dx * dx;
        lb0[i] = std::exp(A(-0.5) * dxdx);
        lb1[i] = -dx * lb0[i];
        lb2[i] = (dxdx - A(1)) * lb0[i];
    }
    int r = int(std::ceil(A(4) * sigma));
    r = r < 1 ? 1 : r > 16 ? 16 : r;
    const size_t s0 = size_t(r) + 1, s1 = size_t(r), s2 = size_t(r) + 1;
    A sum = lb0[0];
    for (size_t i = 1; i < s0; ++i) sum += (A(2) * lb0[i]);
    for (size_t i = 0; i < s0; ++i) lobe0[i] = lb0[s0 - 1 - i] / std::abs(sum);
    sum = A(0);
    for (size_t i = 0; i < s1; ++i) sum += (A(2) * A(i) * lb1[i]);
    for (size_t i = 0; i < s1; ++i) lobe1[i] = lb1[s1 - i] / std::abs(sum);
    sum = A(0);
    for (size_t i = 1; i < s2; ++i) sum += (A(i * i) * lb2[i]);
    for (size_t i = 0; i < s2; ++i) lobe2[i] = lb2[s2 - 1 - i] / std::abs(sum);
    *n0 = int(s0); *n1 = int(s1); *n2 = int(s2);
}

// ---------------------------------------------------------------------------
// ICIA's error moment (TU/cu/ICIA.h:61-112 ICIAErrorMoment, summed by
// ICIA.h:352-369 with cub::DeviceReduce::Sum over every pixel).  Pixel
// (v, u) -- row v, column u of the full-size gradient images, whose border
// the reference leaves at the allocator's zeros -- contributes the upper
// triangle, row by row, of the outer product a a^T (array::ext,
// TU/cu/array.h:101-118, 319-325) of
//   affine      (TU/cu/vec.h:1418-1422): a = (eH u', eH v', eH, eV u', eV v', eV)
//   projective  (TU/cu/vec.h:1285-1290): a = (eH u', eH v', eH, eV u', eV v', eV,
//                                           -(eH u' + eV v') u', -(eH u' + eV v') v')
// with u' = s u, v' = s v, s = 1 / max(nrow, ncol) (ICIA.h:84-89).  Summed
// in double; `abs_out` receives the sum of the terms' magnitudes (the scale
// of a summation's rounding error).
int icia_moment(const float* eH, const float* eV, int64_t stride, int nrow, int ncol, int model,
                double* out, double* abs_out)
{
    if (nrow < 0 || ncol < 0 || (model != 6 && model != 8)) return 1;
    const int D = model, NM = D * (D + 1) / 2;
    std::vector<double> acc(NM, 0.0), aacc(NM, 0.0);
    const double s = double(1.0f / float(std::max(nrow, ncol)));
    double a[8];
    for (int v = 0; v < nrow; ++v)
        for (int u = 0; u < ncol; ++u) {
            const double h = eH[int64_t(v) * stride + u], w = eV[int64_t(v) * stride + u];
            const double uf = s * u, vf = s * v;
            a[0] = h * uf; a[1] = h * vf; a[2] = h;
            a[3] = w * uf; a[4] = w * vf; a[5] = w;
            if (D == 8) {
                const double t = -(h * uf + w * vf);
                a[6] = t * uf; a[7] = t * vf;
            }
            int k = 0;
            for (int i = 0; i < D; ++i)
                for (int j = i; j < D; ++j, ++k) {
                    const double p = a[i] * a[j];
                    acc[k] += p;
                    aacc[k] += std::fabs(p);
                }
        }
    for (int k = 0; k < NM; ++k) { out[k] = acc[k]; if (abs_out) abs_out[k] = aacc[k]; }
    return 0;
}

}   // namespace

extern "C" {

int orc_lobe_size_f32(const float* lobe17, int even)
{
    return int(lobe_size<float>(lobe17, even != 0));
}

int orc_lobe_size_f64(const double* lobe17, int even)
{
    return int(lobe_size<double>(lobe17, even != 0));
}

void orc_gauss_lobes_f32(float sigma, float* l0, int* n0, float* l1, int* n1, float* l2, int* n2)
{
    gauss_lobes<float>(sigma, l0, n0, l1, n1, l2, n2);
}

void orc_gauss_lobes_f64(double sigma, double* l0, int* n0, double* l1, int* n1, double* l2, int* n2)
{
    gauss_lobes<double>(sigma, l0, n0, l1, n1, l2, n2);
}

void orc_gauss_lobes_ceil4_f32(float sigma, float* l0, int* n0, float* l1, int* n1, float* l2, int* n2)
{
    gauss_lobes_ceil4<float>(sigma, l0, n0, l1, n1, l2, n2);
}

void orc_gauss_lobes_ceil4_f64(double sigma, double* l0, int* n0, double* l1, int* n1, double* l2, int* n2)
{
    gauss_lobes_ceil4<double>(sigma, l0, n0, l1, n1, l2, n2);
}

int orc_fir2_f32(const float* lobeH, int nH, const float* lobeV, int nV,
                 const void* in, int in_dt, int64_t in_stride, int nrow, int ncol,
                 void* out, int out_dt, int64_t out_stride, int shift, int nthreads)
{
    return fir2<float>(lobeH, nH, lobeV, nV, in, in_dt, in_stride, nrow, ncol,
                       out, out_dt, out_stride, shift, nthreads);
}

int orc_fir2_f64(const double* lobeH, int nH, const double* lobeV, int nV,
                 const void* in, int in_dt, int64_t in_stride, int nrow, int ncol,
                 void* out, int out_dt, int64_t out_stride, int shift, int nthreads)
{
    return fir2<double>(lobeH, nH, lobeV, nV, in, in_dt, in_stride, nrow, ncol,
                        out, out_dt, out_stride, shift, nthreads);
}

int orc_fir2_taps_f32(const float* tapsH, int nH, const float* tapsV, int nV,
                      const void* in, int in_dt, int64_t in_stride, int nrow, int ncol,
                      void* out, int out_dt, int64_t out_stride, int shift, int nthreads)
{
    return fir2_taps<float>(tapsH, nH, tapsV, nV, in, in_dt, in_stride, nrow, ncol,
                            out, out_dt, out_stride, shift, nthreads);
}

int orc_fir2_taps_f64(const double* tapsH, int nH, const double* tapsV, int nV,
                      const void* in, int in_dt, int64_t in_stride, int nrow, int ncol,
                      void* out, int out_dt, int64_t out_stride, int shift, int nthreads)
{
    return fir2_taps<double>(tapsH, nH, tapsV, nV, in, in_dt, in_stride, nrow, ncol,
                             out, out_dt, out_stride, shift, nthreads);
}

int orc_subsample_f32(const void* in, int in_dt, int64_t in_stride, int nrow, int ncol,
                      void* out, int out_dt, int64_t out_stride, int op)
{
    return subsample(in, in_dt, in_stride, nrow, ncol, out, out_dt, out_stride, op);
}

int orc_icia_moment(const float* eH, const float* eV, int64_t stride, int nrow, int ncol, int model,
                    double* out, double* abs_out)
{
    return icia_moment(eH, eV, stride, nrow, ncol, model, out, abs_out);
}

// Element conversions, exposed so the tests can pin them against numpy.
float    orc_half_to_float(uint16_t h)   { return half_to_float(h); }
uint16_t orc_float_to_half(float f)      { return float_to_half(f); }
float    orc_bf16_to_float(uint16_t b)   { return bf16_to_float(b); }
uint16_t orc_float_to_bf16(float f)      { return float_to_bf16(f); }
uint8_t  orc_float_to_u8(float f)        { return to_u8(f); }

}   // extern "C"

