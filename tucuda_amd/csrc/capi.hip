// capi.hip -- the C ABI of libtucuda_amd.so (declared in include/tucu.h).
//
// Validation and error behaviour follow FIRFilter2 / FIRGaussianConvolver2:
// lobe-length limit at initialize (FIRFilter.h:250-251), silent no-op for
// images too small for the kernel (FIRFilter.h:304-312), unsupported lobe
// length at convolve (FIRFilter.h:343-344, 374-375).  Where the reference
// prints and exits on a device error (algorithm.h:85-94) this layer returns
// TUCU_ERR_DEVICE and leaves the decision to the caller.
#include "tucu.h"
#include "sepfir.h"
#include "mfma.h"
#include "icia.h"
#include "subsample.h"

#include <hip/hip_runtime.h>
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>

namespace tucu {

KernelEntry g_fir_table[4][4][4][4];
KernelEntry g_taps_table[4][4][4][4];
KernelEntry g_gauss_table[5][NGTH][4][4];
KernelEntry g_gauss_mfma_table[2][NMWY][4][4];
KernelEntry g_fir_mfma_table[2][4][4];
IciaLaunchFn g_icia_table[2][4][2];
SubLaunchFn g_sub_table[NSUBOPS][4][4];
PyrLaunchFn g_pyr_table[3][4][4];

size_t lobe_size(const float lobe[], bool even);
void gauss_lobes_reference(float, float*, int*, float*, int*, float*, int*);
void gauss_lobes_ceil4(float, float*, int*, float*, int*, float*, int*);

namespace host {

thread_local std::string t_error;

int fail(int code, const std::string& msg)
{
    t_error = msg;
    return code;
}

int device_fail(hipError_t e, const char* what)
{
    return fail(TUCU_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

size_t dtype_size(int dt)
{
    switch (dt) {
      case TUCU_U8:   return 1;
      case TUCU_F32:  return 4;
      case TUCU_F16:
      case TUCU_BF16: return 2;
      default:        return 0;
    }
}

bool supported_lobe(int n)
{
    return n == 17 || n == 16 || n == 9 || n == 8 || n == 5 || n == 4 || n == 3 || n == 2;
}

int fir_class(int r)
{
    return r == 2 ? 0 : r == 4 ? 1 : r == 8 ? 2 : r == 16 ? 3 : -1;
}

int gauss_class(int r)
{
    for (int i = 0; i < 5; ++i)
        if (r <= GAUSS_RADII[i]) return i;
    return -1;
}

// Tile-height choice of the fused Gaussian kernel: TUCU_GAUSS_TH (one of
// GAUSS_THS) overrides the default.
// Kernel families: matrix cores (default) or the direct FP32 vector kernel,
// TUCU_GAUSS_PATH=valu for the six-output Gaussian (single, batched,
// served), TUCU_FIR_PATH=valu for FIRFilter2 and general taps.  Read at
// every call.
bool gauss_use_mfma()
{
    const char* e = std::getenv("TUCU_GAUSS_PATH");
    return !(e && std::strcmp(e, "valu") == 0);
}

bool fir_use_mfma()
{
    const char* e = std::getenv("TUCU_FIR_PATH");
    return !(e && std::strcmp(e, "valu") == 0);
}

int launch(const KernelEntry& k, SepArgs& a, hipStream_t s);
const void* mfma_tabs_for(const SepArgs& a, int mode, hipStream_t s);

// One-output matrix-core launch: full taps vt / ht (tap o at index o + 16)
// go to a.tv / a.th with their power-of-two scales (mfma.h, MODE_TAPS).
int launch_fir_mfma(SepArgs& a, const float* vt, const float* ht, int rV, int rH, int in_dt, int out_dt,
                    hipStream_t s)
{
    std::memcpy(a.tv, vt, sizeof(a.tv));
    std::memcpy(a.th, ht, sizeof(a.th));
    int e[2] = {0, 0}, eq = 0;
    for (int l = 0; l < 2; ++l) {
        const float* t = l == 0 ? vt : ht;
        float cmax = 0.f, l1 = 0.f;
        for (int i = 0; i < TMAX; ++i) { cmax = std::max(cmax, std::fabs(t[i])); l1 += std::fabs(t[i]); }
        int x = 0;
        (void)std::frexp(cmax, &x);             // cmax < 2^x
        e[l] = MFMA_TAP_EXP - x;
        a.mp[l] = std::ldexp(1.f, e[l]);
        if (l == 0) (void)std::frexp(l1 * a.mp[0], &eq);   // l1 * mp < 2^eq
    }
    eq = std::max(eq, 0);
    a.mp[2] = 1.f;
    a.mq = std::ldexp(1.f, -eq);
    a.mu[0] = std::ldexp(1.f, -(e[0] + e[1] - eq));
    a.tabs = mfma_tabs_for(a, MODE_TAPS, s);
    return launch(g_fir_mfma_table[std::max(rV, rH) <= 8 ? 0 : 1][in_dt][out_dt], a, s);
}

// Workgroup shape of the single-image six-output matrix-core kernel:
// TUCU_MFMA_ROWS (16, 32 or 64 tile rows) overrides the default (64).
int mfma_wy_index()
{
    static const int idx = [] {
        const char* e = std::getenv("TUCU_MFMA_ROWS");
        if (e)
            for (int i = 0; i < NMWY; ++i)
                if (std::atoi(e) == 16 * MFMA_WYS[i]) return i;
        return MFMA_WY_DEFAULT;
    }();
    return idx;
}

int gauss_th_index()
{
    static const int idx = [] {
        const char* e = std::getenv("TUCU_GAUSS_TH");
        if (e)
            for (int i = 0; i < NGTH; ++i)
                if (std::atoi(e) == GAUSS_THS[i]) return i;
        return GAUSS_TH_DEFAULT;
    }();
    return idx;
}

int check_image(const tucu_image* im, const char* name, bool need_data)
{
    if (!im) return fail(TUCU_ERR_INVALID, std::string(name) + " descriptor is null");
    if (dtype_size(im->dtype) == 0) return fail(TUCU_ERR_INVALID, std::string(name) + ": unknown dtype");
    if (im->nrow < 0 || im->ncol < 0) return fail(TUCU_ERR_INVALID, std::string(name) + ": negative size");
    if (need_data && !im->data) return fail(TUCU_ERR_INVALID, std::string(name) + ": null data");
    if (reinterpret_cast<uintptr_t>(im->data) % dtype_size(im->dtype) != 0)
        return fail(TUCU_ERR_INVALID, std::string(name) + ": data not aligned to its element size");
    if (im->stride < im->ncol) return fail(TUCU_ERR_INVALID, std::string(name) + ": stride shorter than a row");
    return TUCU_OK;
}

// A lobe in the reference layout (lobe[0] outermost, centre last for
// symmetric lobes) turned into centre-indexed coefficients c[0..16]:
// c[k] weights the taps at distance k, c[0] the centre (0 if antisymmetric).
void centre_index(const float* lobe, int r, bool symmetric, float c[KMAX])
{
    std::fill(c, c + KMAX, 0.f);
    for (int k = 1; k <= r; ++k) c[k] = lobe[r - k];
    if (symmetric) c[0] = lobe[r];
}

// Power-of-two operand scales of the matrix-core kernels (mfma.h): taps of
// lobe l times mp[l] peak in [2^5, 2^6); mq makes mq * mp[l] * (L1 norm of
// lobe l) <= 1 for every lobe, so that fields of an input scaled below 2^13
// stay below 2^13; mu[k] undoes the tap and input scales of output k = (H
// lobe, V lobe).
void mfma_scales(SepArgs& a)
{
    static const int HV[6][2] = {{0, 0}, {1, 0}, {0, 1}, {2, 0}, {1, 1}, {0, 2}};
    int ep[3], emax = 0;
    for (int l = 0; l < 3; ++l) {
        float cmax = 0.f, l1 = 0.f;
        for (int k = 0; k < KMAX; ++k) {
            const float c = std::fabs(a.cv[l][k]);
            cmax = std::max(cmax, c);
            l1 += k == 0 ? c : 2.f * c;
        }
        int e = 0;
        (void)std::frexp(cmax, &e);             // cmax < 2^e
        ep[l] = MFMA_TAP_EXP - e;
        a.mp[l] = std::ldexp(1.f, ep[l]);
        (void)std::frexp(l1 * a.mp[l], &e);     // l1 * mp < 2^e
        emax = std::max(emax, e);
    }
    a.mq = std::ldexp(1.f, -emax);
    for (int k = 0; k < 6; ++k) {
        const int h = HV[k][0], v = HV[k][1];
        a.mu[k] = std::ldexp(1.f, -(ep[h] + ep[v] - emax));
    }
}

// Host image of the operand tables the matrix-core kernels build in LDS
// (mfma_build_tabs in mfma.h, restated step by step: the same float
// products and the same round-to-nearest-even f16 conversions, so the
// bytes are identical).
void mfma_host_tabs(const SepArgs& a, int mode, MfmaTabs& T)
{
    const bool taps = mode == MODE_TAPS;
    const int nl = taps ? 2 : 3, nw = taps ? 66 : 51;
    std::memset(&T, 0, sizeof(T));
    for (int i = 0; i < nw; ++i) T.lobe[i] = taps ? (i < 33 ? a.tv[i] : a.th[i - 33]) : a.cv[i / 17][i % 17];
    const auto tap = [&](int l, int o) -> float {
        const int k = o < 0 ? -o : o;
        if (k > 16) return 0.f;
        if (taps) return l < 2 ? T.lobe[33 * l + o + 16] : 0.f;
        const float c = T.lobe[17 * l + k];
        return l != 1 ? c : o == 0 ? 0.f : o < 0 ? c : -c;
    };
    constexpr int NT4 = MT4 + 4;
    for (int l = 0; l < nl; ++l)
        for (int t = 0; t < NT4; ++t) T.tap[l][t] = tap(l, t - MBASE) * a.mp[l];
    for (int e = 0; e < nl * (MT4 + MHP); ++e) {
        const bool pair = e >= nl * MT4;
        const int ee = pair ? e - nl * MT4 : e;
        const int n = pair ? MHP : MT4;
        const int l = ee / n, t = ee - l * n;
        v8f16 hi, lo;
        for (int i = 0; i < 8; ++i) {
            const float c = i < 4 || pair ? T.tap[l][t + (i < 4 ? i : 12 + i)] : 0.f;
            const _Float16 h = _Float16(c);
            hi[i] = h;
            lo[i] = _Float16(c - float(h));
        }
        if (pair) {
            T.hpair[(0 * 3 + l) * MHP + t] = hi;
            T.hpair[(1 * 3 + l) * MHP + t] = lo;
        } else {
            T.run[(0 * 3 + l) * MT4 + t] = v4f16{hi[0], hi[1], hi[2], hi[3]};
            T.run[(1 * 3 + l) * MT4 + t] = v4f16{lo[0], lo[1], lo[2], lo[3]};
        }
    }
}

// Prebuilt operand tables in device memory, one per distinct (mode, lobe
// words, tap scales), at most TAB_SLOTS per device; the slots are never
// reused or freed (they live until the process ends), so a table a launch
// points at stays valid for every later replay of a captured graph.  A new
// table is uploaded on the caller's stream, which is then synchronised, so
// that launches on any stream may use it; under stream capture no table
// is made (a launch captured before its lobes were seen builds its tables
// itself), and neither when the slots are used up or memory is short.
constexpr int TAB_SLOTS = 64;

struct TabCache {
    std::mutex mu;
    struct Dev {
        char* arena = nullptr;
        MfmaTabs* host = nullptr;
        bool failed = false;
        std::unordered_map<std::string, int> slot;
    };
    std::unordered_map<int, Dev> dev;
};

TabCache& tab_cache()
{
    static TabCache* c = new TabCache;          // never destroyed: launches may run at exit
    return *c;
}

const void* mfma_tabs_for(const SepArgs& a, int mode, hipStream_t s)
{
    const char* env = std::getenv("TUCU_PREBUILT_TABLES");   // "0": every workgroup builds them (read per call)
    if (env && std::strcmp(env, "0") == 0) return nullptr;
    const bool taps = mode == MODE_TAPS;
    std::string key(1, char(mode));
    const int nw = taps ? 33 : 51;
    if (taps) {
        key.append(reinterpret_cast<const char*>(a.tv), nw * sizeof(float));
        key.append(reinterpret_cast<const char*>(a.th), nw * sizeof(float));
    } else {
        key.append(reinterpret_cast<const char*>(a.cv), nw * sizeof(float));
    }
    key.append(reinterpret_cast<const char*>(a.mp), sizeof(a.mp));
    int d = 0;
    if (hipGetDevice(&d) != hipSuccess) return nullptr;
    TabCache& c = tab_cache();
    std::lock_guard<std::mutex> lock(c.mu);
    TabCache::Dev& D = c.dev[d];
    const auto it = D.slot.find(key);
    if (it != D.slot.end()) return D.arena + size_t(it->second) * sizeof(MfmaTabs);
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return nullptr;
    if (D.failed || int(D.slot.size()) >= TAB_SLOTS) return nullptr;
    if (!D.arena) {
        void* p = nullptr;
        if (hipMalloc(&p, TAB_SLOTS * sizeof(MfmaTabs)) != hipSuccess) {
            (void)hipGetLastError();
            D.failed = true;
            return nullptr;
        }
        D.arena = static_cast<char*>(p);
        D.host = new MfmaTabs[TAB_SLOTS];
    }
    const int k = int(D.slot.size());
    mfma_host_tabs(a, mode, D.host[k]);
    char* const dst = D.arena + size_t(k) * sizeof(MfmaTabs);
    if (hipMemcpyAsync(dst, &D.host[k], sizeof(MfmaTabs), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    D.slot.emplace(std::move(key), k);
    return dst;
}

// Output k of the kernel: base address of centre (0, 0) and row stride.
int set_out(SepArgs& a, int k, const tucu_image& im, int rv, int rh, bool shift)
{
    if (im.stride > INT32_MAX) return fail(TUCU_ERR_INVALID, "output row stride too large");
    const int64_t es = int64_t(dtype_size(im.dtype));
    const int64_t dy = shift ? 0 : -rv, dx = shift ? 0 : -rh;
    a.obase[k] = reinterpret_cast<uintptr_t>(im.data) + uintptr_t((dy * im.stride + dx) * es);
    a.ostride[k] = int32_t(im.stride);
    return TUCU_OK;
}

// Input buffer and the fast-store rectangle of a launch whose live outputs
// are `live`, output k valid on centre rows [rv[k], nrow - rv[k]) and
// columns [rh[k], ncol - rh[k]).
int finish_args(SepArgs& a, const tucu_image& in, unsigned live, const tucu_image* outs,
                const int* rv, const int* rh)
{
    const int64_t es = int64_t(dtype_size(in.dtype));
    const int64_t stride_b = in.stride * es;
    if (int64_t(in.nrow) * stride_b > INT32_MAX)
        return fail(TUCU_ERR_INVALID, "input images are limited to 2 GiB");
    a.in = in.data;
    a.in_stride_b = int32_t(stride_b);
    a.in_bytes = uint32_t(int64_t(in.nrow - 1) * stride_b + in.ncol * es);
    a.nrow = in.nrow; a.ncol = in.ncol;
    // Start the column grid on a 128-byte boundary of the outputs at or before
    // the written region (possibly left of the image; tiles mask the columns
    // outside, and their loads read zeros there): tiles then write whole
    // cache lines and load aligned segments, whatever the radius.
    for (int k = 0; k < 6; ++k) {
        if (!(live & (1u << k))) continue;
        const int oes = int(dtype_size(outs[k].dtype));
        for (int d = 0; d < 128 / oes; ++d)
            if ((a.obase[k] + uintptr_t(intptr_t(a.cx0 - d) * oes)) % 128 == 0) { a.cx0 -= d; break; }
        break;                                  // the first live output sets the phase
    }
    a.fy0 = a.fx0 = 0; a.fy1 = in.nrow; a.fx1 = in.ncol;
    bool aligned = true;
    for (int k = 0; k < 6; ++k) {
        if (!(live & (1u << k))) continue;
        a.fy0 = std::max(a.fy0, rv[k]); a.fy1 = std::min(a.fy1, in.nrow - rv[k]);
        a.fx0 = std::max(a.fx0, rh[k]); a.fx1 = std::min(a.fx1, in.ncol - rh[k]);
        const uintptr_t oes = dtype_size(outs[k].dtype);
        if (a.ostride[k] % 4 != 0 || (a.obase[k] + uintptr_t(a.cx0) * oes) % (4 * oes) != 0) aligned = false;
        // the fast stores address 16 rows from a strip's first row in 32 bits
        if ((int64_t(16) * a.ostride[k] + in.ncol + 256) * int64_t(oes) >= (int64_t(1) << 31)) aligned = false;
    }
    if (!aligned) a.fy0 = a.fy1 = a.fx0 = a.fx1 = 0;
    return TUCU_OK;
}

// The written rectangle must fit the output in each dimension whose extent
// is known (nonzero); the reference itself checks neither.
int check_extent(const tucu_image& o, int rv, int rh, int nrow, int ncol, bool shift, const char* name)
{
    const int64_t rows = int64_t(shift ? rv : 0) + nrow - 2 * rv;
    const int64_t cols = int64_t(shift ? rh : 0) + ncol - 2 * rh;
    if ((o.nrow > 0 && rows > o.nrow) || (o.ncol > 0 && cols > o.ncol))
        return fail(TUCU_ERR_INVALID, std::string(name) + ": output smaller than the region written");
    return TUCU_OK;
}

// ---------------------------------------------------------------- aliasing
// The reference's FIRFilter2::convolve runs its horizontal pass into _buf
// and its vertical pass from _buf into the output (FIRFilter.h:298-377), so
// a call whose output is its own input is well defined there.  The fused
// launches here read the input's halo while other workgroups write outputs,
// so an output that overlaps the input goes through a copy of the input.

// Bytes [lo, hi) an image spans, rows r0 .. r0 + nr - 1, columns c0 .. c0 + nc - 1.
struct Span { uintptr_t lo = 0, hi = 0; };

Span span_of(const tucu_image& im, int64_t r0, int64_t nr, int64_t c0, int64_t nc)
{
    Span s;
    if (nr <= 0 || nc <= 0 || !im.data) return s;
    const int64_t es = int64_t(dtype_size(im.dtype));
    s.lo = reinterpret_cast<uintptr_t>(im.data) + uintptr_t((r0 * im.stride + c0) * es);
    s.hi = reinterpret_cast<uintptr_t>(im.data) + uintptr_t(((r0 + nr - 1) * im.stride + c0 + nc) * es);
    return s;
}

bool overlap(const Span& a, const Span& b) { return a.lo < b.hi && b.lo < a.hi; }

Span input_span(const tucu_image& in) { return span_of(in, 0, in.nrow, 0, in.ncol); }

// Region a FIRFilter2-style call writes: (nrow - 2 rv) x (ncol - 2 rh) at
// (rv, rh) when shifted, at (0, 0) otherwise (FIRFilter.h:286-287).
Span written_span(const tucu_image& o, int rv, int rh, int nrow, int ncol, bool shift)
{
    return span_of(o, shift ? rv : 0, nrow - 2 * rv, shift ? rh : 0, ncol - 2 * rh);
}

// A device copy of an input image, made on the call's stream when an output
// overlaps it, freed once the call's work is done (the call then waits for
// its stream: in-place calls are synchronous).
struct InputCopy {
    void*      tmp = nullptr;
    tucu_image img{};

    int make(const tucu_image& in, hipStream_t s)
    {
        const size_t es = dtype_size(in.dtype);
        const size_t stride = tucu_row_stride(size_t(in.ncol), es);
        if (hipMalloc(&tmp, std::max<size_t>(1, size_t(in.nrow) * stride * es)) != hipSuccess) {
            (void)hipGetLastError();
            tmp = nullptr;
            return fail(TUCU_ERR_BAD_ALLOC, "device allocation for the copy of an aliased input failed");
        }
        const hipError_t e = hipMemcpy2DAsync(tmp, stride * es, in.data, size_t(in.stride) * es,
                                              size_t(in.ncol) * es, size_t(in.nrow), hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return device_fail(e, "hipMemcpy2DAsync (aliased input)");
        img = in;
        img.data = tmp;
        img.stride = int64_t(stride);
        return TUCU_OK;
    }

    // Waits for the stream's work, then frees the copy; passes rc through
    // unless waiting fails.
    int finish(int rc, hipStream_t s)
    {
        if (!tmp) return rc;
        const hipError_t e = hipStreamSynchronize(s);
        (void)hipFree(tmp);
        tmp = nullptr;
        if (rc == TUCU_OK && e != hipSuccess) return device_fail(e, "hipStreamSynchronize (aliased input)");
        return rc;
    }

    ~InputCopy()
    {
        if (tmp) {
            (void)hipDeviceSynchronize();
            (void)hipFree(tmp);
        }
    }
};

int launch(const KernelEntry& k, SepArgs& a, hipStream_t s)
{
    if (!k.fn) return fail(TUCU_ERR_INVALID, "no kernel for this type / radius combination");
    const int two = k.tw ? k.tw : NT - 2 * k.rh;
    const int tiles_x = (a.cx1 - a.cx0 + two - 1) / two;
    const int tiles_y = (a.cy1 - a.cy0 + k.th - 1) / k.th;
    a.tiles_y = tiles_y;
    const int64_t nblocks = int64_t(tiles_x) * tiles_y;
    if (nblocks <= 0) return TUCU_OK;
    if (nblocks > (int64_t(1) << 31) - 1) return fail(TUCU_ERR_INVALID, "image too large");
    const hipError_t e = k.fn(a, int(nblocks), s);
    if (e != hipSuccess) return device_fail(e, "kernel launch");
    return TUCU_OK;
}

// Validated kernel arguments of a fused Gaussian call: lobes (roles fix the
// parities: l0, l2 symmetric with r+1 entries, l1 antisymmetric with r),
// outputs selected by `mask`.  On success a.mask holds the outputs that will
// be written (0: nothing to do), rmax the largest radius in use and odt the
// common output dtype.
int gauss_args(const float* const lobes[3], const int n[3], const tucu_image* in,
               const tucu_image* outs, unsigned mask, int shift, SepArgs& a, int& rmax, int& odt)
{
    int rc;
    std::memset(&a, 0, sizeof(a));
    if ((rc = check_image(in, "input", false))) return rc;
    if (mask & ~unsigned(TUCU_ALL6)) return fail(TUCU_ERR_INVALID, "unknown output bits in mask");
    if (mask == 0) return TUCU_OK;
    if (!outs) return fail(TUCU_ERR_INVALID, "null output array");
    const int r[3] = {n[0] - 1, n[1], n[2] - 1};
    static const int HV[6][2] = {{0, 0}, {1, 0}, {0, 1}, {2, 0}, {1, 1}, {0, 2}};   // (H lobe, V lobe)
    const int nrow = in->nrow, ncol = in->ncol;

    unsigned live = 0;
    int cy0 = nrow, cy1 = 0, cx0 = ncol, cx1 = 0;
    rmax = 0;
    for (int k = 0; k < 6; ++k) {
        if (!(mask & (1u << k))) continue;
        const int h = HV[k][0], v = HV[k][1];
        for (int l : {h, v}) {
            if (r[l] < 1 || r[l] > 16)
                return fail(TUCU_ERR_UNSUPPORTED_LOBE, "lobe " + std::to_string(l) + " length outside the supported range");
            if (!lobes[l]) return fail(TUCU_ERR_INVALID, "null lobe");
        }
        const int rv = r[v], rh = r[h];
        if (nrow <= 2 * rv || ncol <= 2 * rh) continue;               // FIRFilter.h:304-312
        const std::string name = "output " + std::to_string(k);
        if ((rc = check_image(&outs[k], name.c_str(), true))) return rc;
        if ((rc = check_extent(outs[k], rv, rh, nrow, ncol, shift != 0, name.c_str()))) return rc;
        live |= 1u << k;
        rmax = std::max(rmax, std::max(rv, rh));
        cy0 = std::min(cy0, rv); cy1 = std::max(cy1, nrow - rv);
        cx0 = std::min(cx0, rh); cx1 = std::max(cx1, ncol - rh);
    }
    if (!live) return TUCU_OK;
    if (!in->data) return fail(TUCU_ERR_INVALID, "input: null data");
    odt = -1;
    for (int k = 0; k < 6; ++k) {
        if (!(live & (1u << k))) continue;
        if (odt < 0) odt = outs[k].dtype;
        else if (outs[k].dtype != odt)
            return fail(TUCU_ERR_INVALID, "all outputs of one call must share an element type");
    }

    a.cy0 = cy0; a.cy1 = cy1; a.cx0 = cx0; a.cx1 = cx1;
    a.sgnv = a.sgnh = 1.f;
    int rvk[6] = {0}, rhk[6] = {0};
    for (int l = 0; l < 3; ++l) {
        if (r[l] < 1 || r[l] > 16 || !lobes[l]) continue;
        a.r[l] = r[l];
        centre_index(lobes[l], r[l], l != 1, a.cv[l]);
    }
    for (int k = 0; k < 6; ++k) {
        if (!(live & (1u << k))) continue;
        rvk[k] = r[HV[k][1]]; rhk[k] = r[HV[k][0]];
        if ((rc = set_out(a, k, outs[k], rvk[k], rhk[k], shift != 0))) return rc;
    }
    if ((rc = finish_args(a, *in, live, outs, rvk, rhk))) return rc;
    mfma_scales(a);
    a.mask = live;
    a.rt = GAUSS_RADII[gauss_class(rmax)];
    return TUCU_OK;
}

// Aliasing of a fused Gaussian call whose live outputs are `live` (radii
// r[] of lobes 0, 1, 2): TUCU_ERR_INVALID if two outputs overlap (one
// launch writes both, so neither result would hold), -1 if an output
// overlaps the input, TUCU_OK otherwise.
int gauss_aliasing(const tucu_image* in, const tucu_image* outs, unsigned live, const int* r, bool shift)
{
    static const int HV[6][2] = {{0, 0}, {1, 0}, {0, 1}, {2, 0}, {1, 1}, {0, 2}};
    Span w[6];
    for (int k = 0; k < 6; ++k)
        if (live & (1u << k))
            w[k] = written_span(outs[k], r[HV[k][1]], r[HV[k][0]], in->nrow, in->ncol, shift);
    for (int k = 0; k < 6; ++k)
        for (int j = 0; j < k; ++j)
            if ((live >> k & 1u) && (live >> j & 1u) && overlap(w[k], w[j]))
                return fail(TUCU_ERR_INVALID, "outputs " + std::to_string(j) + " and " + std::to_string(k) +
                                              " of one call overlap");
    const Span src = input_span(*in);
    for (int k = 0; k < 6; ++k)
        if ((live >> k & 1u) && overlap(w[k], src)) return -1;
    return TUCU_OK;
}

}   // namespace host
}   // namespace tucu

using namespace tucu;
using namespace tucu::host;

extern "C" {

const char* tucu_last_error(void) { return t_error.c_str(); }

int tucu_version(void) { return 100; }

int tucu_malloc(void** ptr, size_t bytes)
{
    if (!ptr) return fail(TUCU_ERR_INVALID, "null out pointer");
    *ptr = nullptr;
    if (bytes == 0) return TUCU_OK;                 // allocator.h:74-75
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess || !p) {
        (void)hipGetLastError();
        return fail(TUCU_ERR_BAD_ALLOC, "device allocation of " + std::to_string(bytes) + " bytes failed");
    }
    const hipError_t e = hipMemset(p, 0, bytes);    // allocator.h:85
    if (e != hipSuccess) {
        (void)hipFree(p);
        return device_fail(e, "hipMemset");
    }
    *ptr = p;
    return TUCU_OK;
}

int tucu_free(void* ptr)
{
    if (!ptr) return TUCU_OK;
    const hipError_t e = hipFree(ptr);
    return e == hipSuccess ? TUCU_OK : device_fail(e, "hipFree");
}

size_t tucu_row_stride(size_t ncol, size_t elem_size)
{
    if (elem_size == 0) return ncol;
    const size_t bytes = (ncol * elem_size + 255) / 256 * 256;
    return bytes / elem_size;
}

int tucu_memcpy2d(void* dst, size_t dpitch, const void* src, size_t spitch,
                  size_t width_bytes, size_t height, int kind)
{
    if (kind < 0 || kind > 2) return fail(TUCU_ERR_INVALID, "copy kind must be 0, 1 or 2");
    if (width_bytes == 0 || height == 0) return TUCU_OK;
    if (!dst || !src) return fail(TUCU_ERR_INVALID, "null pointer");
    if (dpitch < width_bytes || spitch < width_bytes) return fail(TUCU_ERR_INVALID, "pitch shorter than a row");
    const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice
                          : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    const hipError_t e = hipMemcpy2D(dst, dpitch, src, spitch, width_bytes, height, k);
    return e == hipSuccess ? TUCU_OK : device_fail(e, "hipMemcpy2D");
}

int tucu_synchronize(void)
{
    const hipError_t e = hipDeviceSynchronize();
    return e == hipSuccess ? TUCU_OK : device_fail(e, "hipDeviceSynchronize");
}

int tucu_fir2_check(int nH, int nV)
{
    if (nH < 0 || nV < 0) return fail(TUCU_ERR_INVALID, "negative lobe size");
    if (nH > 17 || nV > 17)
        return fail(TUCU_ERR_LOBE_TOO_LARGE, "FIRFilter2<T, BLOCK_TRAITS>::initialize: too large lobe size!");
    return TUCU_OK;
}

int tucu_fir2_convolve(const float* lobeH, int nH, const float* lobeV, int nV,
                       const tucu_image* in, const tucu_image* out, int shift, void* stream)
{
    int rc = tucu_fir2_check(nH, nV);
    if (rc) return rc;
    if ((rc = check_image(in, "input", false))) return rc;
    if ((rc = check_image(out, "output", false))) return rc;
    const int nrow = in->nrow, ncol = in->ncol;
    const int rV = nV & ~1, rH = nH & ~1;
    if (nrow <= 2 * rV || ncol <= 2 * rH) return TUCU_OK;           // FIRFilter.h:304-312
    if (!supported_lobe(nH))
        return fail(TUCU_ERR_UNSUPPORTED_LOBE,
                    "FIRFilter2::convolve: unsupported horizontal lobe size[" + std::to_string(nH) + "]!");
    if (!supported_lobe(nV))
        return fail(TUCU_ERR_UNSUPPORTED_LOBE,
                    "FIRFilter2::convolve: unsupported vertical lobe size[" + std::to_string(nV) + "!");
    if (!lobeH || !lobeV) return fail(TUCU_ERR_INVALID, "null lobe");
    if (!in->data || !out->data) return fail(TUCU_ERR_INVALID, "null image data");
    if ((rc = check_extent(*out, rV, rH, nrow, ncol, shift != 0, "output"))) return rc;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    InputCopy copy;
    if (overlap(written_span(*out, rV, rH, nrow, ncol, shift != 0), input_span(*in))) {   // in place
        if ((rc = copy.make(*in, st))) return copy.finish(rc, st);
        in = &copy.img;
    }

    SepArgs a;
    std::memset(&a, 0, sizeof(a));
    a.cy0 = rV; a.cy1 = nrow - rV;
    a.cx0 = rH; a.cx1 = ncol - rH;
    a.mask = 1;
    a.sgnv = (nV & 1) ? 1.f : -1.f;
    a.sgnh = (nH & 1) ? 1.f : -1.f;
    a.r[0] = rV; a.r[1] = rH;
    centre_index(lobeV, rV, nV & 1, a.cv[0]);
    centre_index(lobeH, rH, nH & 1, a.ch);
    if ((rc = set_out(a, 0, *out, rV, rH, shift != 0))) return copy.finish(rc, st);
    if ((rc = finish_args(a, *in, 1u, out, &rV, &rH))) return copy.finish(rc, st);
    if (fir_use_mfma()) {
        // full taps: c[k] on the upper / left side, sgn * c[k] on the other (FIRFilter.h:66-176)
        float vt[TMAX] = {0.f}, ht[TMAX] = {0.f};
        for (int k = 0; k <= 16; ++k) {
            vt[16 - k] = a.cv[0][k]; vt[16 + k] = k ? a.sgnv * a.cv[0][k] : a.cv[0][0];
            ht[16 - k] = a.ch[k];    ht[16 + k] = k ? a.sgnh * a.ch[k] : a.ch[0];
        }
        return copy.finish(launch_fir_mfma(a, vt, ht, rV, rH, in->dtype, out->dtype, st), st);
    }
    const KernelEntry& k = g_fir_table[fir_class(rV)][fir_class(rH)][in->dtype][out->dtype];
    return copy.finish(launch(k, a, st), st);
}

int tucu_fir2_convolve_taps(const float* tapsH, int nH, const float* tapsV, int nV,
                            const tucu_image* in, const tucu_image* out, int shift, void* stream)
{
    int rc;
    if (nH < 1 || nV < 1 || nH > TMAX || nV > TMAX || !(nH & 1) || !(nV & 1))
        return fail(TUCU_ERR_UNSUPPORTED_LOBE, "general taps: counts must be odd, 1 to 33");
    if ((rc = check_image(in, "input", false))) return rc;
    if ((rc = check_image(out, "output", false))) return rc;
    const int nrow = in->nrow, ncol = in->ncol;
    const int rV = nV / 2, rH = nH / 2;
    if (nrow <= 2 * rV || ncol <= 2 * rH) return TUCU_OK;
    if (!tapsH || !tapsV) return fail(TUCU_ERR_INVALID, "null taps");
    if (!in->data || !out->data) return fail(TUCU_ERR_INVALID, "null image data");
    if ((rc = check_extent(*out, rV, rH, nrow, ncol, shift != 0, "output"))) return rc;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    InputCopy copy;
    if (overlap(written_span(*out, rV, rH, nrow, ncol, shift != 0), input_span(*in))) {   // in place
        if ((rc = copy.make(*in, st))) return copy.finish(rc, st);
        in = &copy.img;
    }
    // template radius: the smallest class that holds the taps (a centred
    // 1-tap filter still uses the radius-2 kernel)
    auto cls = [](int r) { return r <= 2 ? 0 : r <= 4 ? 1 : r <= 8 ? 2 : 3; };
    const int RV = FIR_RADII[cls(rV)], RH = FIR_RADII[cls(rH)];

    SepArgs a;
    std::memset(&a, 0, sizeof(a));
    a.cy0 = rV; a.cy1 = nrow - rV;
    a.cx0 = rH; a.cx1 = ncol - rH;
    a.mask = 1;
    a.r[0] = rV; a.r[1] = rH;
    for (int i = 0; i < nV; ++i) a.tv[RV - rV + i] = tapsV[i];
    for (int i = 0; i < nH; ++i) a.th[RH - rH + i] = tapsH[i];
    if ((rc = set_out(a, 0, *out, rV, rH, shift != 0))) return copy.finish(rc, st);
    if ((rc = finish_args(a, *in, 1u, out, &rV, &rH))) return copy.finish(rc, st);
    if (fir_use_mfma()) {
        float vt[TMAX] = {0.f}, ht[TMAX] = {0.f};
        for (int i = 0; i < nV; ++i) vt[16 - rV + i] = tapsV[i];
        for (int i = 0; i < nH; ++i) ht[16 - rH + i] = tapsH[i];
        return copy.finish(launch_fir_mfma(a, vt, ht, rV, rH, in->dtype, out->dtype, st), st);
    }
    const KernelEntry& k = g_taps_table[cls(rV)][cls(rH)][in->dtype][out->dtype];
    return copy.finish(launch(k, a, st), st);
}

int tucu_lobe_size(const float* lobe17, int even)
{
    if (!lobe17) { fail(TUCU_ERR_INVALID, "null lobe"); return -TUCU_ERR_INVALID; }
    return int(lobe_size(lobe17, even != 0));
}

int tucu_gauss2_lobes(float sigma, int rule, float* lobe0, int* n0, float* lobe1, int* n1,
                      float* lobe2, int* n2)
{
    if (!lobe0 || !lobe1 || !lobe2 || !n0 || !n1 || !n2) return fail(TUCU_ERR_INVALID, "null pointer");
    if (rule == TUCU_RULE_REFERENCE) {
        gauss_lobes_reference(sigma, lobe0, n0, lobe1, n1, lobe2, n2);
        return TUCU_OK;
    }
    if (rule == TUCU_RULE_CEIL4SIGMA) {
        if (!(sigma > 0.f) || !(sigma < 1e6f)) return fail(TUCU_ERR_INVALID, "sigma must be positive and finite");
        gauss_lobes_ceil4(sigma, lobe0, n0, lobe1, n1, lobe2, n2);
        return TUCU_OK;
    }
    return fail(TUCU_ERR_INVALID, "unknown lobe rule");
}

int tucu_gauss2_convolve(const float* lobe0, int n0, const float* lobe1, int n1,
                         const float* lobe2, int n2,
                         const tucu_image* in, const tucu_image* outs,
                         unsigned mask, int shift, void* stream)
{
    const float* lobes[3] = {lobe0, lobe1, lobe2};
    const int n[3] = {n0, n1, n2};
    SepArgs a;
    int rmax = 0, odt = 0;
    int rc = gauss_args(lobes, n, in, outs, mask, shift, a, rmax, odt);
    if (rc || !a.mask) return rc;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    InputCopy copy;
    if ((rc = gauss_aliasing(in, outs, a.mask, a.r, shift != 0))) {
        if (rc != -1) return rc;
        if ((rc = copy.make(*in, st))) return copy.finish(rc, st);     // an output overlaps the input
        in = &copy.img;
        if ((rc = gauss_args(lobes, n, in, outs, mask, shift, a, rmax, odt))) return copy.finish(rc, st);
    }
    if (gauss_use_mfma()) {
        const int rc = rmax <= 8 ? 0 : 1;
        const KernelEntry* m = &g_gauss_mfma_table[rc][mfma_wy_index()][in->dtype][odt];
        if (!m->fn) m = &g_gauss_mfma_table[rc][MFMA_WY_DEFAULT][in->dtype][odt];
        a.tabs = mfma_tabs_for(a, MODE_GAUSS, st);
        return copy.finish(launch(*m, a, st), st);
    }
    const KernelEntry* e = &g_gauss_table[gauss_class(rmax)][gauss_th_index()][in->dtype][odt];
    if (!e->fn) e = &g_gauss_table[gauss_class(rmax)][GAUSS_TH_DEFAULT][in->dtype][odt];
    return copy.finish(launch(*e, a, st), st);
}

int tucu_subsample(const tucu_image* in, const tucu_image* out, int op, void* stream)
{
    int rc;
    if (op < 0 || op >= NSUBOPS) return fail(TUCU_ERR_INVALID, "unknown subsample operator");
    if ((rc = check_image(in, "input", false))) return rc;
    if ((rc = check_image(out, "output", false))) return rc;
    const int nrow = in->nrow, ncol = in->ncol;
    if (nrow < 2 || ncol < 2) return TUCU_OK;                       // algorithm.h:227-233
    if (op != SUB_IDENTITY && in->dtype == TUCU_U8)
        return fail(TUCU_ERR_INVALID, "subsample: this operator needs a floating-point input");
    if (!in->data || !out->data) return fail(TUCU_ERR_INVALID, "null image data");
    if ((out->nrow > 0 && out->nrow < nrow / 2) || (out->ncol > 0 && out->ncol < ncol / 2))
        return fail(TUCU_ERR_INVALID, "output: smaller than nrow/2 x ncol/2");
    if (out->stride > INT32_MAX) return fail(TUCU_ERR_INVALID, "output row stride too large");
    const int64_t es = int64_t(dtype_size(in->dtype));
    if (int64_t(nrow) * in->stride * es > INT32_MAX) return fail(TUCU_ERR_INVALID, "input images are limited to 2 GiB");
    const hipStream_t st = static_cast<hipStream_t>(stream);
    InputCopy copy;
    if (overlap(span_of(*out, 0, nrow / 2, 0, ncol / 2), input_span(*in))) {       // in place
        if ((rc = copy.make(*in, st))) return copy.finish(rc, st);
        in = &copy.img;
    }
    const int64_t stride_b = in->stride * es;

    SubArgs a;
    std::memset(&a, 0, sizeof(a));
    a.in = in->data;
    a.in_stride_b = int32_t(stride_b);
    a.in_bytes = uint32_t(int64_t(nrow - 1) * stride_b + ncol * es);
    a.nrow = nrow; a.ncol = ncol;
    a.onrow = nrow / 2; a.oncol = ncol / 2;
    a.vec = reinterpret_cast<uintptr_t>(in->data) % uintptr_t(4 * es) == 0 && stride_b % (4 * es) == 0;
    a.tiles_y = (a.onrow + SUB_WAVES * SUB_R - 1) / (SUB_WAVES * SUB_R);
    a.ostride = int32_t(out->stride);
    a.out = reinterpret_cast<uintptr_t>(out->data);
    const SubLaunchFn fn = g_sub_table[op][in->dtype][out->dtype];
    if (!fn) return copy.finish(fail(TUCU_ERR_INVALID, "no subsample kernel for this operator / type combination"), st);
    const hipError_t e = fn(a, st);
    return copy.finish(e == hipSuccess ? TUCU_OK : device_fail(e, "kernel launch"), st);
}

int tucu_gauss2_pyramid(const float* lobe0, int n0, const tucu_image* in, const tucu_image* out,
                        int shift, void* stream)
{
    int rc;
    const int r = n0 - 1;
    if (r < 1 || r > 16) return fail(TUCU_ERR_UNSUPPORTED_LOBE, "lobe 0 length outside the supported range");
    if (!lobe0) return fail(TUCU_ERR_INVALID, "null lobe");
    if ((rc = check_image(in, "input", false))) return rc;
    if ((rc = check_image(out, "output", false))) return rc;
    const int nrow = in->nrow, ncol = in->ncol;
    if (nrow < 2 || ncol < 2) return TUCU_OK;                       // algorithm.h:227-233
    if (!in->data || !out->data) return fail(TUCU_ERR_INVALID, "null image data");
    if ((out->nrow > 0 && out->nrow < nrow / 2) || (out->ncol > 0 && out->ncol < ncol / 2))
        return fail(TUCU_ERR_INVALID, "output: smaller than nrow/2 x ncol/2");
    if (out->stride > INT32_MAX) return fail(TUCU_ERR_INVALID, "output row stride too large");
    const int64_t es = int64_t(dtype_size(in->dtype));
    if (int64_t(nrow) * in->stride * es > INT32_MAX) return fail(TUCU_ERR_INVALID, "input images are limited to 2 GiB");
    const hipStream_t st = static_cast<hipStream_t>(stream);
    InputCopy copy;
    if (overlap(span_of(*out, 0, nrow / 2, 0, ncol / 2), input_span(*in))) {       // in place
        if ((rc = copy.make(*in, st))) return copy.finish(rc, st);
        in = &copy.img;
    }
    const int64_t stride_b = in->stride * es;

    PyrArgs a;
    std::memset(&a, 0, sizeof(a));
    a.in = in->data;
    a.in_stride_b = int32_t(stride_b);
    a.in_bytes = uint32_t(int64_t(nrow - 1) * stride_b + ncol * es);
    a.nrow = nrow; a.ncol = ncol;
    a.onrow = nrow / 2; a.oncol = ncol / 2;
    a.r = r;
    a.o = shift ? 0 : r;
    a.ostride = int32_t(out->stride);
    a.tiles_y = (a.onrow + PYR_TH - 1) / PYR_TH;
    a.out = reinterpret_cast<uintptr_t>(out->data);
    centre_index(lobe0, r, true, a.c);
    const int ri = r <= 4 ? 0 : r <= 8 ? 1 : 2;
    const PyrLaunchFn fn = g_pyr_table[ri][in->dtype][out->dtype];
    if (!fn) return copy.finish(fail(TUCU_ERR_INVALID, "no kernel for this type combination"), st);
    const hipError_t e = fn(a, st);
    return copy.finish(e == hipSuccess ? TUCU_OK : device_fail(e, "kernel launch"), st);
}

int tucu_icia_moment(const float* lobe0, int n0, const float* lobe1, int n1, const tucu_image* in,
                     int model, float* moment, void* workspace, size_t* workspace_bytes, void* stream)
{
    int rc;
    if (model != TUCU_ICIA_AFFINE && model != TUCU_ICIA_PROJECTIVE) return fail(TUCU_ERR_INVALID, "unknown ICIA model");
    if (!workspace_bytes) return fail(TUCU_ERR_INVALID, "null workspace size");
    const int r0 = n0 - 1, r1 = n1;                  // lobe 0 symmetric (r + 1 entries), lobe 1 antisymmetric (r)
    if (r0 < 1 || r0 > 16 || r1 < 1 || r1 > 16)
        return fail(TUCU_ERR_UNSUPPORTED_LOBE, "lobe length outside the supported range");
    if (!lobe0 || !lobe1) return fail(TUCU_ERR_INVALID, "null lobe");
    if ((rc = check_image(in, "input", false))) return rc;
    const int nrow = in->nrow, ncol = in->ncol;
    SepArgs a;
    std::memset(&a, 0, sizeof(a));
    // diffH = (H lobe 1, V lobe 0) is valid on rows [r0, nrow - r0), columns
    // [r1, ncol - r1); diffV = (H lobe 0, V lobe 1) on rows [r1, nrow - r1),
    // columns [r0, ncol - r0) (FIRFilter.h:37-40, 286-287); an image too small
    // for one of them leaves it all zero (FIRFilter.h:304-312).
    unsigned live = 0;
    int cy0 = nrow, cy1 = 0, cx0 = ncol, cx1 = 0;
    if (nrow > 2 * r0 && ncol > 2 * r1) {
        live |= TUCU_DIFFH;
        cy0 = std::min(cy0, r0); cy1 = std::max(cy1, nrow - r0); cx0 = std::min(cx0, r1); cx1 = std::max(cx1, ncol - r1);
    }
    if (nrow > 2 * r1 && ncol > 2 * r0) {
        live |= TUCU_DIFFV;
        cy0 = std::min(cy0, r1); cy1 = std::max(cy1, nrow - r1); cx0 = std::min(cx0, r0); cx1 = std::max(cx1, ncol - r0);
    }
    const int D = model, NM = D * (D + 1) / 2;
    constexpr int WY = MFMA_WYS[MFMA_WY_ICIA];
    int64_t nblocks = 0;
    if (live) {
        a.tiles_y = (cy1 - cy0 + 16 * WY - 1) / (16 * WY);
        nblocks = int64_t((cx1 - cx0 + MSTRIP_W * (MWAVES / WY) - 1) / (MSTRIP_W * (MWAVES / WY))) * a.tiles_y;
        if (nblocks > (int64_t(1) << 31) - 1) return fail(TUCU_ERR_INVALID, "image too large");
    }
    const size_t need = (size_t(nblocks) * size_t(NM) * sizeof(double) + 255) / 256 * 256 + 256;
    if (!workspace) { *workspace_bytes = need; return TUCU_OK; }  // the size query (cub's two-call form, ICIA.h:355-366)
    if (*workspace_bytes < need) return fail(TUCU_ERR_INVALID, "workspace smaller than tucu_icia_moment asked for");
    if (!moment) return fail(TUCU_ERR_INVALID, "null moment");
    if (live && !in->data) return fail(TUCU_ERR_INVALID, "input: null data");
    {
        // The moment pass writes its partials while other workgroups still
        // read the input, and the reduction writes the moments while others
        // still read partials: the workspace may overlap neither.
        const uintptr_t w0 = reinterpret_cast<uintptr_t>(workspace), m0 = reinterpret_cast<uintptr_t>(moment);
        const Span ws{w0, w0 + need}, ms{m0, m0 + size_t(NM) * sizeof(float)};
        if (overlap(ws, ms)) return fail(TUCU_ERR_INVALID, "the moment overlaps the workspace");
        if (live && overlap(ws, input_span(*in))) return fail(TUCU_ERR_INVALID, "the workspace overlaps the input");
    }
    if (live) {
        const int64_t es = int64_t(dtype_size(in->dtype));
        const int64_t stride_b = in->stride * es;
        if (int64_t(nrow) * stride_b > INT32_MAX) return fail(TUCU_ERR_INVALID, "input images are limited to 2 GiB");
        a.in = in->data;
        a.in_stride_b = int32_t(stride_b);
        a.in_bytes = uint32_t(int64_t(nrow - 1) * stride_b + ncol * es);
        a.nrow = nrow; a.ncol = ncol;
        a.cy0 = cy0; a.cy1 = cy1; a.cx0 = cx0; a.cx1 = cx1;
        a.r[0] = r0; a.r[1] = r1;
        centre_index(lobe0, r0, true, a.cv[0]);
        centre_index(lobe1, r1, false, a.cv[1]);
        mfma_scales(a);
        a.mask = live;
        a.rt = std::max(r0, r1) <= 8 ? 8 : 16;
        a.tabs = mfma_tabs_for(a, MODE_GAUSS, static_cast<hipStream_t>(stream));
    }
    double* partial = reinterpret_cast<double*>((reinterpret_cast<uintptr_t>(workspace) + 255) / 256 * 256);
    const IciaLaunchFn fn = g_icia_table[std::max(r0, r1) <= 8 ? 0 : 1][in->dtype][model == TUCU_ICIA_AFFINE ? 0 : 1];
    if (!fn) return fail(TUCU_ERR_INVALID, "no ICIA kernel for this input type");
    const hipError_t e = fn(a, int(nblocks), partial, moment, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? TUCU_OK : device_fail(e, "kernel launch");
}

}   // extern "C"
