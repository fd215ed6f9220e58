// batch_host.hip -- batched mode of the C ABI: many Gaussian jobs of uneven
// sizes and radii in one persistent launch, with a tile-height autotuner.
#include "hostutil.h"
#include "batch.h"

#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

namespace tucu {

BatchEntry g_batch_table[NBTH][4][4];
BatchEntry g_batch_mfma_table[NBTH][4][4];

namespace host {

int batch_family()
{
    return gauss_use_mfma() ? BATCH_MFMA : BATCH_VALU;
}

const BatchEntry& batch_entry(int fam, int thi, int in_dt, int out_dt)
{
    return fam == BATCH_MFMA ? g_batch_mfma_table[thi][in_dt][out_dt] : g_batch_table[thi][in_dt][out_dt];
}

// Tile list of `jobs` for the tile shape of e (sets every job's tiles_y):
// tiles of the larger radius classes first (they cost most; handing them out
// early shortens the tail), each job's tiles in strip-major order.
int make_tiles(std::vector<SepArgs>& jobs, const BatchEntry& e, std::vector<BatchTile>& tiles)
{
    const int th = e.th;
    for (SepArgs& a : jobs) a.tiles_y = (a.cy1 - a.cy0 + th - 1) / th;
    std::vector<int> order(jobs.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = int(i);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return jobs[x].rt > jobs[y].rt; });
    tiles.clear();
    for (int j : order) {
        const SepArgs& a = jobs[j];
        const int two = e.tw ? e.tw : NT - 2 * a.rt;
        const int64_t n = int64_t((a.cx1 - a.cx0 + two - 1) / two) * a.tiles_y;
        for (int64_t t = 0; t < n; ++t) tiles.push_back({uint32_t(j), uint32_t(t)});
    }
    if (tiles.size() >= size_t(INT32_MAX)) return fail(TUCU_ERR_INVALID, "batch too large");
    return TUCU_OK;
}

// Workgroups of a batched launch: enough to fill the GPU once, at most one
// per tile.
int batch_grid(const BatchEntry& e, int ntiles)
{
    int dev = 0, ncu = 0, per_cu = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, e.kernel, e.nthreads, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    return std::max(1, std::min(ntiles, std::max(1, ncu) * per_cu));
}

}   // namespace host

namespace {

using namespace host;

// Device copy of a batch laid out for one tile height.
struct Layout {
    SepArgs*   d_jobs  = nullptr;
    BatchTile* d_tiles = nullptr;
    int        ntiles  = 0;
    int        grid    = 0;
    bool       built   = false;
};

// Tuning cache: (kernel family, in dtype, out dtype, radius class of the
// largest job, log2 of the tile count at the default height) -> best
// tile-height index.
using TuneKey = std::tuple<int, int, int, int, int>;
struct TuneVal { int th_index; float us; };
std::mutex                  g_tune_mutex;
std::map<TuneKey, TuneVal>  g_tune;

}   // namespace
}   // namespace tucu

using namespace tucu;
using namespace tucu::host;

struct tucu_batch_s {
    int                  fam = 0;                // kernel family, fixed at create
    int                  in_dt = 0, out_dt = 0;
    std::vector<SepArgs> jobs;
    Layout               layout[NBTH];
    int*                 d_counter = nullptr;    // tile counter of the launches (cleared by each)
    int                  th_index = -1;          // -1: tuned value if cached, else default
    int                  active = BATCH_TH_DEFAULT;   // layout the next launch uses (always built)
    int                  rclass = 0;
    int                  bucket = 0;
};

namespace {

int free_layout(Layout& l)
{
    if (l.d_jobs) (void)hipFree(l.d_jobs);
    if (l.d_tiles) (void)hipFree(l.d_tiles);
    l = Layout();
    return TUCU_OK;
}

int build_layout(tucu_batch_s& b, int thi)
{
    Layout& l = b.layout[thi];
    if (l.built) return TUCU_OK;
    const BatchEntry& e = batch_entry(b.fam, thi, b.in_dt, b.out_dt);
    if (!e.fn) return fail(TUCU_ERR_INVALID, "no batched kernel for this tile height and type pair");
    std::vector<SepArgs> jobs = b.jobs;
    std::vector<BatchTile> all;
    int rc = make_tiles(jobs, e, all);
    if (rc) return rc;
    hipError_t err;
    if ((err = hipMalloc(&l.d_jobs, jobs.size() * sizeof(SepArgs))) != hipSuccess ||
        (err = hipMalloc(&l.d_tiles, std::max<size_t>(1, all.size()) * sizeof(BatchTile))) != hipSuccess) {
        free_layout(l);
        (void)hipGetLastError();
        return fail(TUCU_ERR_BAD_ALLOC, "device allocation for the batch tables failed");
    }
    if ((err = hipMemcpy(l.d_jobs, jobs.data(), jobs.size() * sizeof(SepArgs), hipMemcpyHostToDevice)) != hipSuccess ||
        (all.size() && (err = hipMemcpy(l.d_tiles, all.data(), all.size() * sizeof(BatchTile),
                                        hipMemcpyHostToDevice)) != hipSuccess)) {
        free_layout(l);
        return device_fail(err, "hipMemcpy");
    }
    l.ntiles = int(all.size());
    l.grid = batch_grid(e, l.ntiles);
    l.built = true;
    return TUCU_OK;
}

int pick_th(const tucu_batch_s& b)
{
    if (b.th_index >= 0) return b.th_index;
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    const auto it = g_tune.find(TuneKey(b.fam, b.in_dt, b.out_dt, b.rclass, b.bucket));
    return it != g_tune.end() ? it->second.th_index : BATCH_TH_DEFAULT;
}

int launch_layout(tucu_batch_s& b, int thi, hipStream_t s)
{
    int rc;
    if ((rc = build_layout(b, thi))) return rc;
    const Layout& l = b.layout[thi];
    if (l.ntiles == 0) return TUCU_OK;
    const hipError_t e = batch_entry(b.fam, thi, b.in_dt, b.out_dt).fn(l.d_jobs, l.d_tiles, l.ntiles,
                                                                 b.d_counter, l.grid, s);
    return e == hipSuccess ? TUCU_OK : device_fail(e, "batched kernel launch");
}

// Kernel arguments of every job that writes something, validated as
// tucu_gauss2_convolve validates one call; all jobs share one input and one
// output element type, and none filters in place.  rclass: the largest
// radius class; work: centre pixels covered.
int batch_jobs(const tucu_gauss_job* jobs, int njobs, std::vector<SepArgs>& out, int& in_dt, int& out_dt,
               int& rclass, int64_t& work)
{
    if (njobs < 0 || (njobs > 0 && !jobs)) return fail(TUCU_ERR_INVALID, "bad job array");
    out.clear();
    rclass = 0;
    work = 0;
    for (int j = 0; j < njobs; ++j) {
        const tucu_gauss_job& g = jobs[j];
        const float* lobes[3] = {g.lobe0, g.lobe1, g.lobe2};
        const int n[3] = {g.n0, g.n1, g.n2};
        SepArgs a;
        int rmax = 0, odt = 0;
        const int rc = gauss_args(lobes, n, &g.in, g.out, g.mask, g.shift, a, rmax, odt);
        if (rc) return rc;
        if (!a.mask) continue;                                  // nothing to write for this job
        if (const int al = gauss_aliasing(&g.in, g.out, a.mask, a.r, g.shift != 0))
            return al == -1 ? fail(TUCU_ERR_INVALID, "job " + std::to_string(j) +
                                                     ": an output overlaps the input (batched jobs cannot filter in place)")
                            : al;
        if (out.empty()) { in_dt = g.in.dtype; out_dt = odt; }
        else if (g.in.dtype != in_dt || odt != out_dt)
            return fail(TUCU_ERR_INVALID, "all jobs of a batch must share input and output element types");
        out.push_back(a);
        rclass = std::max(rclass, a.rt);
        work += int64_t(a.cy1 - a.cy0) * (a.cx1 - a.cx0);
    }
    return TUCU_OK;
}

// ------------------------------------------------------ no-allocation launch
// Scratch of tucu_batch_run_static: the tile-queue counters, then the job
// table, then the tile list of one launch.  A file-scope device array: each
// GPU the library runs on holds its own copy until the process ends.  Host
// code never passes the array's name to a kernel or a copy (on the host it
// is only a stand-in for the symbol); it uses the device address
// hipGetSymbolAddress returns, looked up once per device.
constexpr size_t SCRATCH_BYTES = TUCU_STATIC_SCRATCH_BYTES;
constexpr size_t SCRATCH_JOBS_OFF = 1024;                   // after the counters (QUEUE_BYTES)
static_assert(QUEUE_BYTES <= SCRATCH_JOBS_OFF, "counters overlap the job table");
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void*      g_scratch_ptr[MAX_DEVICES] = {nullptr};

// Device address of the current device's scratch.
int scratch_ptr(unsigned char** p)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return device_fail(e, "hipGetDevice");
    if (dev < 0 || dev >= MAX_DEVICES) return fail(TUCU_ERR_INVALID, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) return device_fail(e, "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[dev] = ptr;
    }
    *p = static_cast<unsigned char*>(g_scratch_ptr[dev]);
    return TUCU_OK;
}

size_t round256(size_t n) { return (n + 255) / 256 * 256; }

}   // namespace

extern "C" {

int tucu_batch_run_static(const tucu_gauss_job* jobs, int njobs, void* stream)
{
    std::vector<SepArgs> args;
    int in_dt = 0, out_dt = 0, rclass = 0;
    int64_t work = 0;
    int rc = batch_jobs(jobs, njobs, args, in_dt, out_dt, rclass, work);
    if (rc || args.empty()) return rc;
    const BatchEntry& e = batch_entry(batch_family(), BATCH_TH_DEFAULT, in_dt, out_dt);
    if (!e.fn) return fail(TUCU_ERR_INVALID, "no batched kernel for this type pair");
    std::vector<BatchTile> tiles;
    if ((rc = make_tiles(args, e, tiles))) return rc;
    if (tiles.empty()) return TUCU_OK;
    const size_t jobs_bytes = round256(args.size() * sizeof(SepArgs));
    const size_t tiles_off = SCRATCH_JOBS_OFF + jobs_bytes;
    const size_t need = tiles_off + tiles.size() * sizeof(BatchTile);
    if (need > SCRATCH_BYTES)
        return fail(TUCU_ERR_INVALID, "batch tables need " + std::to_string(need) + " bytes; the static scratch holds " +
                                      std::to_string(SCRATCH_BYTES));
    unsigned char* base = nullptr;
    if ((rc = scratch_ptr(&base))) return rc;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    SepArgs* d_jobs = reinterpret_cast<SepArgs*>(base + SCRATCH_JOBS_OFF);
    BatchTile* d_tiles = reinterpret_cast<BatchTile*>(base + tiles_off);
    // Pageable host memory: each copy returns once its source is consumed,
    // in stream order behind the work already on `s`.
    hipError_t err = hipMemcpyAsync(d_jobs, args.data(), args.size() * sizeof(SepArgs), hipMemcpyHostToDevice, s);
    if (err == hipSuccess)
        err = hipMemcpyAsync(d_tiles, tiles.data(), tiles.size() * sizeof(BatchTile), hipMemcpyHostToDevice, s);
    if (err != hipSuccess) return device_fail(err, "hipMemcpyAsync (static scratch)");
    // The counters are set by the launch's own reset kernel.
    err = e.fn(d_jobs, d_tiles, int(tiles.size()), reinterpret_cast<int*>(base), batch_grid(e, int(tiles.size())), s);
    return err == hipSuccess ? TUCU_OK : device_fail(err, "batched kernel launch");
}

int tucu_batch_create(tucu_batch* batch, const tucu_gauss_job* jobs, int njobs)
{
    if (!batch) return fail(TUCU_ERR_INVALID, "null batch handle");
    *batch = nullptr;
    auto b = new tucu_batch_s();
    b->fam = batch_family();
    int64_t work = 0;
    const int rc = batch_jobs(jobs, njobs, b->jobs, b->in_dt, b->out_dt, b->rclass, work);
    if (rc) { delete b; return rc; }
    b->bucket = work > 0 ? int(std::log2(double(work))) : 0;
    if (hipMalloc(&b->d_counter, QUEUE_BYTES) != hipSuccess) {
        (void)hipGetLastError();
        delete b;
        return fail(TUCU_ERR_BAD_ALLOC, "device allocation for the batch counter failed");
    }
    if (!b->jobs.empty()) {
        b->active = pick_th(*b);
        const int rc = build_layout(*b, b->active);
        if (rc) { tucu_batch_destroy(b); return rc; }
    }
    *batch = b;
    return TUCU_OK;
}

int tucu_batch_destroy(tucu_batch batch)
{
    if (!batch) return TUCU_OK;
    for (Layout& l : batch->layout) free_layout(l);
    if (batch->d_counter) (void)hipFree(batch->d_counter);
    delete batch;
    return TUCU_OK;
}

int tucu_batch_set_tile_rows(tucu_batch batch, int rows)
{
    if (!batch) return fail(TUCU_ERR_INVALID, "null batch");
    int thi = -1;
    if (rows == 0) {
        batch->th_index = -1;
        thi = pick_th(*batch);
    } else {
        for (int i = 0; i < NBTH; ++i)
            if (batch_entry(batch->fam, i, batch->in_dt, batch->out_dt).th == rows &&
                batch_entry(batch->fam, i, batch->in_dt, batch->out_dt).fn) thi = i;
        batch->th_index = thi;
    }
    if (thi >= 0) {
        if (batch->jobs.empty()) return TUCU_OK;
        const int rc = build_layout(*batch, thi);
        if (rc == TUCU_OK) batch->active = thi;
        return rc;
    }
    return fail(TUCU_ERR_INVALID, "tile height not available for this type pair");
}

int tucu_batch_launch(tucu_batch batch, void* stream)
{
    if (!batch) return fail(TUCU_ERR_INVALID, "null batch");
    if (batch->jobs.empty()) return TUCU_OK;
    return launch_layout(*batch, batch->active, static_cast<hipStream_t>(stream));
}

int tucu_batch_info(tucu_batch batch, int* njobs, int* ntiles, int* grid, int* tile_rows)
{
    if (!batch) return fail(TUCU_ERR_INVALID, "null batch");
    const int thi = batch->active;
    const Layout& l = batch->layout[thi];
    if (njobs) *njobs = int(batch->jobs.size());
    if (ntiles) *ntiles = l.ntiles;
    if (grid) *grid = l.grid;
    if (tile_rows) *tile_rows = batch_entry(batch->fam, thi, batch->in_dt, batch->out_dt).th;
    return TUCU_OK;
}

int tucu_batch_autotune(tucu_batch batch, void* stream, int reps, int* best_rows, float* best_us)
{
    if (!batch) return fail(TUCU_ERR_INVALID, "null batch");
    if (batch->jobs.empty()) return TUCU_OK;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    reps = std::max(1, reps);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        const hipError_t err = hipGetLastError();
        if (e0) (void)hipEventDestroy(e0);
        return device_fail(err, "hipEventCreate");
    }
    int best = -1;
    float best_t = 0.f;
    int rc = TUCU_OK;
    for (int i = 0; i < NBTH && rc == TUCU_OK; ++i) {
        if (!batch_entry(batch->fam, i, batch->in_dt, batch->out_dt).fn) continue;
        if ((rc = launch_layout(*batch, i, s))) break;                  // builds and warms up
        std::vector<float> t(reps);
        for (int r = 0; r < reps && rc == TUCU_OK; ++r) {
            (void)hipEventRecord(e0, s);
            rc = launch_layout(*batch, i, s);
            (void)hipEventRecord(e1, s);
            if (hipEventSynchronize(e1) != hipSuccess) rc = device_fail(hipGetLastError(), "hipEventSynchronize");
            (void)hipEventElapsedTime(&t[r], e0, e1);
        }
        if (rc) break;
        std::nth_element(t.begin(), t.begin() + reps / 2, t.end());
        const float med = t[reps / 2] * 1000.f;
        if (best < 0 || med < best_t) { best = i; best_t = med; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (best < 0) return fail(TUCU_ERR_INVALID, "no batched kernel for this type pair");
    batch->th_index = -1;
    batch->active = best;
    {
        std::lock_guard<std::mutex> lock(g_tune_mutex);
        g_tune[TuneKey(batch->fam, batch->in_dt, batch->out_dt, batch->rclass, batch->bucket)] = {best, best_t};
    }
    if (best_rows) *best_rows = batch_entry(batch->fam, best, batch->in_dt, batch->out_dt).th;
    if (best_us) *best_us = best_t;
    return TUCU_OK;
}

int tucu_tune_cache_save(const char* path)
{
    if (!path) return fail(TUCU_ERR_INVALID, "null path");
    std::ofstream f(path);
    if (!f) return fail(TUCU_ERR_INVALID, std::string("cannot write ") + path);
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    for (const auto& kv : g_tune) {
        const int fam = std::get<0>(kv.first), in = std::get<1>(kv.first), out = std::get<2>(kv.first);
        f << "{\"family\": " << fam << ", \"in\": " << in << ", \"out\": " << out
          << ", \"radius\": " << std::get<3>(kv.first) << ", \"log2_pixels\": " << std::get<4>(kv.first)
          << ", \"tile_rows\": " << batch_entry(fam, kv.second.th_index, in, out).th
          << ", \"us\": " << kv.second.us << "}\n";
    }
    return TUCU_OK;
}

int tucu_tune_cache_load(const char* path)
{
    if (!path) return fail(TUCU_ERR_INVALID, "null path");
    std::ifstream f(path);
    if (!f) return fail(TUCU_ERR_INVALID, std::string("cannot read ") + path);
    std::string line;
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    while (std::getline(f, line)) {
        int fam = -1, in = -1, out = -1, r = -1, bucket = -1, rows = -1;
        float us = 0.f;
        if (std::sscanf(line.c_str(), "{\"family\": %d, \"in\": %d, \"out\": %d, \"radius\": %d, "
                                      "\"log2_pixels\": %d, \"tile_rows\": %d, \"us\": %f}",
                        &fam, &in, &out, &r, &bucket, &rows, &us) != 7)
            continue;
        if (fam < 0 || fam > 1 || in < 0 || in > 3 || out < 0 || out > 3) continue;
        for (int i = 0; i < NBTH; ++i)
            if (batch_entry(fam, i, in, out).fn && batch_entry(fam, i, in, out).th == rows)
                g_tune[TuneKey(fam, in, out, r, bucket)] = {i, us};
    }
    return TUCU_OK;
}

}   // extern "C"
