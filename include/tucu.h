/*
 * tucu.h -- C ABI of the MI355X (gfx950) separable FIR / Gaussian-derivative
 * library libtucuda_amd.so.
 *
 * Plain C: pointers, sizes and POD descriptors only.  Every entry point
 * returns a tucu_status; on failure tucu_last_error() returns a message for
 * the calling thread.  Work is enqueued on the HIP stream passed as `stream`
 * (a hipStream_t, NULL = the default stream) and is asynchronous to the
 * host, as the reference's default-stream launches are.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the TUCuda tree).
 */
#ifndef TUCU_H
#define TUCU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TUCU_API __attribute__((visibility("default")))

typedef enum {
    TUCU_OK = 0,
    TUCU_ERR_INVALID = 1,          /* bad argument (null pointer, negative size, output too small) */
    TUCU_ERR_LOBE_TOO_LARGE = 2,   /* FIRFilter.h:250-251 runtime_error("... too large lobe size!") */
    TUCU_ERR_UNSUPPORTED_LOBE = 3, /* FIRFilter.h:343-344, 374-375 runtime_error("... unsupported ... lobe size") */
    TUCU_ERR_BAD_ALLOC = 4,        /* allocator.h:83-84 std::bad_alloc */
    TUCU_ERR_DEVICE = 5            /* a HIP runtime error (algorithm.h:85-94 prints and exits instead) */
} tucu_status;

typedef enum {
    TUCU_U8 = 0,    /* u_char */
    TUCU_F32 = 1,   /* float */
    TUCU_F16 = 2,   /* IEEE binary16 storage, converted to float on load (extension) */
    TUCU_BF16 = 3   /* bfloat16 storage, converted to float on load (extension) */
} tucu_dtype;

/* A 2-D image in device memory: the kernel-visible form of
 * range<range_iterator<device_ptr<T>>> (iterator.h:385-434, 439-527):
 * row-0 pointer, row stride in elements, rows, columns, element type. */
typedef struct {
    void*   data;
    int64_t stride;
    int32_t nrow;
    int32_t ncol;
    int32_t dtype;
    int32_t reserved;
} tucu_image;

/* Output bits of tucu_gauss2_convolve, in FIRGaussianConvolver2's method order. */
enum {
    TUCU_SMOOTH = 1u << 0,  /* FIRGaussianConvolver.h:178-182 */
    TUCU_DIFFH  = 1u << 1,  /* :190-194 */
    TUCU_DIFFV  = 1u << 2,  /* :202-206 */
    TUCU_DIFFHH = 1u << 3,  /* :214-218 */
    TUCU_DIFFHV = 1u << 4,  /* :226-230 */
    TUCU_DIFFVV = 1u << 5,  /* :238-242 */
    TUCU_ALL6   = 0x3fu
};

/* Lobe rule for tucu_gauss2_lobes. */
enum {
    TUCU_RULE_REFERENCE = 0,   /* FIRGaussianConvolver.h:123-170 + FIRGaussianConvolver.cu:52-98 */
    TUCU_RULE_CEIL4SIGMA = 1   /* extension: radius ceil(4 sigma) (at most 16), same normalisation */
};

/* Message describing the last failure on this thread ("" if none). */
TUCU_API const char* tucu_last_error(void);

/* Library version, major*10000 + minor*100 + patch. */
TUCU_API int tucu_version(void);

/* ---- device memory (cu::allocator / cu::Array2 storage) ---------------- */

/* cu::allocator<T>::allocate (allocator.h:71-87): device memory, zero-filled,
 * 256-byte aligned; bytes == 0 yields NULL.  Failure: TUCU_ERR_BAD_ALLOC. */
TUCU_API int tucu_malloc(void** ptr, size_t bytes);

/* cu::allocator<T>::deallocate (allocator.h:88-94); NULL is a no-op. */
TUCU_API int tucu_free(void* ptr);

/* Row stride (in elements) of a cu::Array2<T> row of `ncol` elements of
 * `elem_size` bytes: rows padded to the allocator's 256-byte alignment
 * (allocator.h:68, Array++.h:142). */
TUCU_API size_t tucu_row_stride(size_t ncol, size_t elem_size);

/* 2-D copy between host and device memory: the transfers behind
 * cu::Array2<T>(host array) and Image<T>(cu::Array2<T>) (Testprog/gaussTest/
 * main.cu:91, 108).  kind: 0 host->device, 1 device->host, 2 device->device.
 * Pitches and width in bytes.  Synchronous, like the reference's copies. */
TUCU_API int tucu_memcpy2d(void* dst, size_t dpitch, const void* src, size_t spitch,
                           size_t width_bytes, size_t height, int kind);

/* hipDeviceSynchronize (the caller's cudaDeviceSynchronize(),
 * Testprog/gaussTest/main.cu:95). */
TUCU_API int tucu_synchronize(void);

/* ---- FIRFilter2 --------------------------------------------------------- */

/* FIRFilter2::initialize(lobeH, lobeV) size check (FIRFilter.h:245-261):
 * TUCU_ERR_LOBE_TOO_LARGE when either length exceeds 17. */
TUCU_API int tucu_fir2_check(int nH, int nV);

/* FIRFilter2::initialize(lobeH, lobeV).convolve(in, ie, out, shift)
 * (FIRFilter.h:245-261, 298-377; instantiated for u_char/float in and out,
 * FIRFilter.cu:49-68).  Lobe lengths must be in {2,3,4,5,8,9,16,17}: odd
 * lengths are symmetric half-kernels with lobe[len-1] at the centre, even
 * lengths antisymmetric ones; lobe[0] is the outermost tap.  With
 * rH = nH & ~1 and rV = nV & ~1, the output is written only on the
 * (nrow-2rV) x (ncol-2rH) valid region, at offset (rV, rH) when shift != 0,
 * else at (0, 0); nothing else of `out` is touched.  Images with
 * nrow <= 2rV or ncol <= 2rH are a silent no-op (FIRFilter.h:304-312).
 * in->dtype and out->dtype may be any tucu_dtype.  If out->nrow and
 * out->ncol are nonzero they are checked against the region written.
 * Coefficients travel as kernel arguments: no per-call upload. */
TUCU_API int tucu_fir2_convolve(const float* lobeH, int nH, const float* lobeV, int nV,
                                const tucu_image* in, const tucu_image* out,
                                int shift, void* stream);

/* Extension (no counterpart in the reference, whose FIRFilter2 lobes are
 * always symmetric or antisymmetric): a separable filter with general taps.
 * tapsH / tapsV hold nH / nV (odd, 1..33) taps; tap i of a window starting
 * at (y, x) weights in[y][x+i] (horizontal) or in[y+i][.] (vertical), the
 * window convention of FIRFilter2.  With r = (n-1)/2 per direction the
 * output placement, valid region, no-op rule and dtypes are those of
 * tucu_fir2_convolve; symmetric taps reproduce FIRFilter2 up to rounding. */
TUCU_API int tucu_fir2_convolve_taps(const float* tapsH, int nH, const float* tapsV, int nV,
                                     const tucu_image* in, const tucu_image* out,
                                     int shift, void* stream);

/* ---- FIRGaussianConvolver2 ---------------------------------------------- */

/* detail::lobeSize(lobe, even) (FIRGaussianConvolver.cu:52-98) over a
 * 17-element lobe: returns the size (2..17), or -TUCU_ERR_INVALID for a
 * null lobe (a size, not a tucu_status, so errors are negative). */
TUCU_API int tucu_lobe_size(const float* lobe17, int even);

/* FIRGaussianConvolver2::initialize(sigma, lobe0, lobe1, lobe2)
 * (FIRGaussianConvolver.h:123-170).  Each lobe buffer holds 17 floats; the
 * lengths are returned in n0, n1, n2.  rule: TUCU_RULE_*. */
TUCU_API int tucu_gauss2_lobes(float sigma, int rule,
                               float* lobe0, int* n0, float* lobe1, int* n1,
                               float* lobe2, int* n2);

/* The six FIRGaussianConvolver2 methods (FIRGaussianConvolver.h:178-242)
 * fused: one read of `in`, every output selected by `mask` (TUCU_SMOOTH ..
 * TUCU_DIFFVV) written by the same launch.  outs[k] is the destination of
 * method k; entries whose mask bit is clear are ignored.  Each output obeys
 * the single-method contract of tucu_fir2_convolve with its own (H, V) lobe
 * pair: smooth (l0,l0), diffH (l1,l0), diffV (l0,l1), diffHH (l2,l0),
 * diffHV (l1,l1), diffVV (l0,l2).  The roles fix the parities: l0 and l2
 * are symmetric lobes of length r+1 (centre last), l1 an antisymmetric lobe
 * of length r, each radius r in 1..16.  Lobes of the reference rule
 * ({3,5,9,17} and {2,4,8,16}) mean exactly what they mean to FIRFilter2;
 * other radii serve TUCU_RULE_CEIL4SIGMA lobes. */
TUCU_API int tucu_gauss2_convolve(const float* lobe0, int n0, const float* lobe1, int n1,
                                  const float* lobe2, int n2,
                                  const tucu_image* in, const tucu_image* outs,
                                  unsigned mask, int shift, void* stream);

/* ---- subsample: the pyramid step ----------------------------------------- */

/* Stencil operators of cu::subsample (TU/cu/functional.h). */
enum {
    TUCU_OP_IDENTITY1X1  = 0,   /* identity1x1   functional.h:58-65 */
    TUCU_OP_DIFFH1X3     = 1,   /* diffH1x3      :68-81 */
    TUCU_OP_DIFFV3X1     = 2,   /* diffV3x1      :84-98 */
    TUCU_OP_DIFFHH1X3    = 3,   /* diffHH1x3     :101-114 */
    TUCU_OP_DIFFVV3X1    = 4,   /* diffVV3x1     :117-130 */
    TUCU_OP_DIFFHV3X3    = 5,   /* diffHV3x3     :133-148 */
    TUCU_OP_SOBELH3X3    = 6,   /* sobelH3x3     :151-167 */
    TUCU_OP_SOBELV3X3    = 7,   /* sobelV3x3     :170-186 */
    TUCU_OP_SOBELABS3X3  = 8,   /* sobelAbs3x3   :189-198 */
    TUCU_OP_LAPLACIAN3X3 = 9,   /* laplacian3x3  :201-210 */
    TUCU_OP_DET3X3       = 10,  /* det3x3        :213-226 */
    TUCU_OP_GAUSSIAN5X5  = 11   /* gaussian5x5   :229-256 */
};

/* cu::subsample(in, ie, out, op) (TU/cu/algorithm.h:183-242): halves the
 * image in both directions, out[y][x] = op evaluated at input pixel
 * (2y, 2x), for y < nrow/2 and x < ncol/2.  The operator's border rule is
 * that of its guarded form in functional.h: 0 wherever its stencil would
 * leave the image.  Images with fewer than 2 rows or columns are a silent
 * no-op (algorithm.h:227-233).  identity1x1 takes any tucu_dtype input; the
 * other operators compute in FP32 and need a floating-point input (f32, or
 * f16 / bf16 storage as an extension).  out may be any tucu_dtype (float ->
 * u_char saturates, then truncates); if out->nrow / out->ncol are nonzero
 * they must hold nrow/2 x ncol/2.  Exactly nrow/2 x ncol/2 pixels are
 * written: for an odd extent the reference's grid can also write row nrow/2
 * or column ncol/2, past the output it is sized for. */
TUCU_API int tucu_subsample(const tucu_image* in, const tucu_image* out, int op, void* stream);

/* Extension (no counterpart in the reference): the pyramid step fused.
 * Equals FIRGaussianConvolver2::smooth(in -> tmp, shift) into a zero-filled
 * nrow x ncol image tmp followed by subsample(tmp -> out, identity1x1)
 * (FIRGaussianConvolver.h:178-182, algorithm.h:183-242), without writing or
 * reading tmp: out[y][x] = the smoothed pixel at (2y, 2x), or 0 where that
 * pixel lies outside smooth's valid region.  lobe0 is a symmetric lobe of
 * n0 = r + 1 entries (centre last), r in 1..16, as for tucu_gauss2_convolve.
 * Every one of the nrow/2 x ncol/2 output pixels is written; images with
 * fewer than 2 rows or columns are a no-op.  Any tucu_dtype in and out. */
TUCU_API int tucu_gauss2_pyramid(const float* lobe0, int n0, const tucu_image* in,
                                 const tucu_image* out, int shift, void* stream);

/* ---- ICIA's error moment ---------------------------------------------
 * Replaces ICIA::operator()'s gradient and moment steps (TU/cu/ICIA.h:346-369:
 * FIRGaussianConvolver2::diffH and diffV into two images, then
 * cub::DeviceReduce::Sum of ICIAErrorMoment, ICIA.h:61-112): diffH and diffV
 * of `in` with the Gaussian lobes lobe0 (n0 = r0 + 1 entries) and lobe1 (n1 =
 * r1) from one fused pass that never writes them, and the sum over every
 * pixel (row v, column u) of the upper triangle, row by row, of a a^T with
 * a = image_derivative0(s u, s v, eH, eV) (TU/cu/vec.h:1418-1422 affine, 6
 * entries, 21 moments; :1285-1290 projective, 8 entries, 36 moments), s = 1 /
 * max(nrow, ncol), a gradient taken as 0 outside its valid region (the
 * reference's freshly allocated gradient images hold zeros there).
 * `moment`: device memory for 21 or 36 floats.  Call first with workspace ==
 * NULL to get the workspace size in *workspace_bytes (as cub's first call,
 * ICIA.h:355-360), then with a device buffer at least that large; it must
 * overlap neither the image nor `moment` (TUCU_ERR_INVALID).  A NaN or an
 * Inf in the image makes the moment non-finite, as the reference's sum. */
enum { TUCU_ICIA_AFFINE = 6, TUCU_ICIA_PROJECTIVE = 8 };
TUCU_API int tucu_icia_moment(const float* lobe0, int n0, const float* lobe1, int n1, const tucu_image* in,
                              int model, float* moment, void* workspace, size_t* workspace_bytes, void* stream);

/* ---- batched mode (extension) ------------------------------------------
 * Many FIRGaussianConvolver2 jobs of uneven sizes and lobe radii in one
 * persistent launch.  No counterpart in the reference, which filters one
 * image per call (FIRGaussianConvolver.h:178-242); each job obeys the
 * contract of tucu_gauss2_convolve and produces exactly what that call
 * produces for it.  The launch allocates nothing and can be captured in a
 * HIP graph; a batch object must not be launched concurrently with itself
 * (it owns the launch's tile counters). */

typedef struct tucu_batch_s* tucu_batch;

typedef struct {
    tucu_image   in;
    tucu_image   out[6];      /* destination of method k (TUCU_SMOOTH .. TUCU_DIFFVV order) */
    const float* lobe0;       /* lobes as for tucu_gauss2_convolve */
    const float* lobe1;
    const float* lobe2;
    int32_t      n0, n1, n2;
    uint32_t     mask;        /* outputs wanted */
    int32_t      shift;
} tucu_gauss_job;

/* Validates every job, copies its arguments and lays out the device tile
 * tables (allocates device memory).  All jobs share input and output
 * element types. */
TUCU_API int tucu_batch_create(tucu_batch* batch, const tucu_gauss_job* jobs, int njobs);
TUCU_API int tucu_batch_destroy(tucu_batch batch);
/* Enqueue the whole batch on `stream` (one kernel launch). */
TUCU_API int tucu_batch_launch(tucu_batch batch, void* stream);
/* Tile height in rows (4, 8 or 16 where instantiated; 0 = the tuned value
 * for batches of this shape if the tuning cache has one, else 8). */
TUCU_API int tucu_batch_set_tile_rows(tucu_batch batch, int rows);
/* Times `reps` launches per available tile height on `stream` (the outputs
 * are written), keeps the fastest for this batch and records it in the
 * process-wide tuning cache under (dtypes, largest radius, log2 pixels). */
TUCU_API int tucu_batch_autotune(tucu_batch batch, void* stream, int reps, int* best_rows, float* best_us);
TUCU_API int tucu_batch_info(tucu_batch batch, int* njobs, int* ntiles, int* grid, int* tile_rows);

/* One batched launch of `jobs` that allocates nothing.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * stream capture, tight serving loops, harnesses that forbid hipMalloc).
 * Validates the jobs as tucu_batch_create does, builds the job table and
 * tile list in host memory, copies them (hipMemcpyAsync from pageable
 * memory, which returns once the host tables are consumed) into a static
 * device scratch array of TUCU_STATIC_SCRATCH_BYTES that each GPU holds
 * until the process ends, and launches the batched kernel there, with the
 * default tile shape.  Results equal tucu_batch_launch's bit for bit.  A
 * batch whose tables exceed the scratch is refused with TUCU_ERR_INVALID.
 * Every call on one GPU shares that GPU's scratch: calls ordered on one
 * stream are safe, calls that overlap on one GPU (other streams, other
 * threads) race. */
#define TUCU_STATIC_SCRATCH_BYTES (4u << 20)
TUCU_API int tucu_batch_run_static(const tucu_gauss_job* jobs, int njobs, void* stream);
/* Tuning cache as JSON lines. */
TUCU_API int tucu_tune_cache_save(const char* path);
TUCU_API int tucu_tune_cache_load(const char* path);

/* ---- serving (extension) -------------------------------------------------
 * A continuous-batching server for Gaussian requests, no counterpart in the
 * reference.  Requests are submitted from any thread; a scheduler thread
 * admits them in arrival order while the output pool has room (pool_bytes
 * reserved at create; when a request does not fit, the pool grows by
 * tucu_server_admit_bytes(...) bytes if the device's free HBM less
 * reserve_bytes holds it within max_pool_bytes, and the request waits for
 * released outputs otherwise), packs all admitted requests into one batched launch
 * on a free in-flight slot (a stream each, max_inflight of them), so requests
 * that arrive while the GPU works join the next launch; completions are
 * drained with tucu_server_poll and their outputs, which live in the pool,
 * are handed back with tucu_server_release.  Each output equals what
 * tucu_gauss2_convolve writes for the request alone. */

typedef struct tucu_server_s* tucu_server;

typedef struct {
    size_t  pool_bytes;       /* HBM budget for request outputs (admission control) */
    int32_t in_dtype;         /* element type of every request's input */
    int32_t out_dtype;        /* element type of the outputs */
    int32_t rule;             /* TUCU_RULE_* for the lobes of each sigma */
    int32_t tile_rows;        /* batched-kernel tile height (0: the library default; as tucu_batch_set_tile_rows) */
    int32_t max_batch_jobs;   /* requests per launch (0: 1024) */
    int32_t max_inflight;     /* launches in flight at once (0: 2) */
    int32_t max_tiles;        /* tiles per launch (0: 2^21) */
    int32_t max_batch_mpix;   /* input megapixels per launch (0: 64); a launch holds at least one request */
    int32_t reserved0;
    size_t  max_pool_bytes;   /* the pool grows up to this (0: while free HBM allows; pool_bytes: fixed) */
    size_t  reserve_bytes;    /* device memory the pool leaves free when it grows (0: 1 GiB) */
    size_t  grow_bytes;       /* smallest growth step (0: 256 MiB) */
} tucu_server_config;

typedef struct {
    tucu_image in;            /* input in device memory, valid until the request completes */
    float      sigma;
    uint32_t   mask;          /* outputs wanted (TUCU_SMOOTH ...) */
    int32_t    shift;
    int32_t    reserved;
    double     arrival;       /* arrival time on the server clock, reported back */
} tucu_request;

typedef struct {
    uint64_t   id;
    int32_t    status;        /* tucu_status of the request */
    int32_t    reserved;
    double     arrival, admitted, launched, completed;   /* server clock, seconds */
    tucu_image out[6];        /* outputs (pool memory), valid until tucu_server_release(id) */
} tucu_completion;

typedef struct {
    uint64_t batches, requests_launched, requests_done, max_batch;
    uint64_t pool_bytes, pool_used, pool_peak;
} tucu_server_counters;

/* The server's admission rule for a request needing `need` bytes of outputs
 * that its pool (now pool_bytes) cannot place: the bytes to grow the pool by,
 * max(need, grow_step) clipped to device_free - reserve and to
 * max_pool_bytes - pool_bytes (max_pool_bytes 0: no cap), or 0 -- wait for
 * released memory -- when need exceeds either limit.  Host-only. */
TUCU_API size_t tucu_server_admit_bytes(size_t need, size_t device_free, size_t reserve, size_t pool_bytes,
                                        size_t max_pool_bytes, size_t grow_step);
TUCU_API int    tucu_server_create(tucu_server* server, const tucu_server_config* cfg);
/* Fails requests not yet admitted, waits for in-flight launches, frees everything. */
TUCU_API int    tucu_server_destroy(tucu_server server);
/* Seconds since the server was created (the clock of all timestamps). */
TUCU_API double tucu_server_now(tucu_server server);
TUCU_API int    tucu_server_submit(tucu_server server, const tucu_request* req, uint64_t* id);
/* Submits n requests at once, in order (one lock, one wake-up of the
 * scheduler): all are checked first and none is queued if one is invalid;
 * ids (n entries, may be null) receives their ids. */
TUCU_API int    tucu_server_submit_many(tucu_server server, const tucu_request* reqs, int n, uint64_t* ids);
/* Moves up to max completions into out; *n = number moved. */
TUCU_API int    tucu_server_poll(tucu_server server, tucu_completion* out, int max, int* n);
/* Hands the outputs of a completed request back to the pool; a request not
 * yet completed (queued or in flight) is refused with TUCU_ERR_INVALID. */
TUCU_API int    tucu_server_release(tucu_server server, uint64_t id);
TUCU_API int    tucu_server_stats(tucu_server server, tucu_server_counters* stats);

#ifdef __cplusplus
}
#endif

#endif /* TUCU_H */
