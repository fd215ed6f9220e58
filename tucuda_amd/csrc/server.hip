// server.hip -- continuous-batching server for Gaussian filtering requests.
//
// Requests (an input image already in device memory, sigma, the outputs
// wanted) are submitted from any thread.  A scheduler thread admits them in
// arrival order while the output pool has room -- the pool grows in chunks
// while the device's free HBM (hipMemGetInfo) less a reserve allows, up to
// an optional cap (tucu_server_admit_bytes is the rule) -- packs
// every admitted request into the batched kernel's job and tile tables, and
// launches it on one of a few in-flight slots (a stream each), while earlier
// batches are still running; a request that arrives while the GPU is busy
// joins the next launch as soon as a slot and memory are free.  Completed
// requests are drained with tucu_server_poll and give their output memory
// back with tucu_server_release.  Each output equals what
// tucu_gauss2_convolve writes for the request alone: the batched kernel runs
// the same per-tile code whatever else shares the launch.
#include "hostutil.h"
#include "batch.h"

#include <hip/hip_runtime.h>
#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace tucu {
void gauss_lobes_reference(float, float*, int*, float*, int*, float*, int*);
void gauss_lobes_ceil4(float, float*, int*, float*, int*, float*, int*);
}

using namespace tucu;
using namespace tucu::host;

namespace {

constexpr size_t ALIGN = 256;

size_t round_up(size_t n, size_t a) { return (n + a - 1) / a * a; }

// First-fit allocator over device chunks (one arena at create, more as the
// admission rule lets the pool grow), with coalescing inside each chunk.
class Pool {
  public:
    struct Chunk {
        char*                    base = nullptr;
        size_t                   cap = 0;
        std::map<size_t, size_t> free_;           // offset -> length
    };

    void add(char* base, size_t cap)
    {
        Chunk c;
        c.base = base;
        c.cap = cap;
        c.free_[0] = cap;
        chunks_.push_back(std::move(c));
        cap_ += cap;
    }
    bool alloc(size_t n, int& chunk, size_t& off)
    {
        n = round_up(n, ALIGN);
        for (size_t ci = 0; ci < chunks_.size(); ++ci) {
            auto& fr = chunks_[ci].free_;
            for (auto it = fr.begin(); it != fr.end(); ++it) {
                if (it->second < n) continue;
                off = it->first;
                const size_t rest = it->second - n;
                fr.erase(it);
                if (rest) fr[off + n] = rest;
                used_ += n;
                peak_ = std::max(peak_, used_);
                chunk = int(ci);
                return true;
            }
        }
        return false;
    }
    void release(int chunk, size_t off, size_t n)
    {
        n = round_up(n, ALIGN);
        used_ -= n;
        auto& fr = chunks_[size_t(chunk)].free_;
        auto it = fr.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != fr.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            fr.erase(nx);
        }
        if (it != fr.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                fr.erase(it);
            }
        }
    }
    char*  base(int chunk) const { return chunks_[size_t(chunk)].base; }
    size_t cap()  const { return cap_; }
    size_t used() const { return used_; }
    size_t peak() const { return peak_; }
    size_t nchunks() const { return chunks_.size(); }
    const std::vector<Chunk>& chunks() const { return chunks_; }

  private:
    std::vector<Chunk> chunks_;
    size_t cap_ = 0, used_ = 0, peak_ = 0;
};

struct Request {
    uint64_t     id = 0;
    tucu_request req;
    size_t       off[6] = {0}, bytes[6] = {0};
    int          chunk[6] = {0};
    tucu_image   out[6];
    double       admitted = 0, launched = 0;
    bool         done = false;     // completed (guarded by the server's mu); until then the
                                   // scheduler owns the request and may not see it released
};

struct Slot {
    hipStream_t            stream = nullptr;
    hipEvent_t             done = nullptr;
    SepArgs*               d_jobs = nullptr;
    BatchTile*             d_tiles = nullptr;
    int*                   d_counter = nullptr;
    SepArgs*               h_jobs = nullptr;      // pinned staging
    BatchTile*             h_tiles = nullptr;
    std::vector<Request*>  reqs;
    bool                   busy = false;
};

using Lobes = std::array<float, 3 * 17 + 3>;      // three lobes of 17 + their lengths

}   // namespace

struct tucu_server_s {
    tucu_server_config                     cfg;
    int                                    fam = 0;      // kernel family of the batched kernel
    int                                    thi = 0;      // tile-height index of the batched kernel
    int                                    device = 0;
    std::chrono::steady_clock::time_point  t0;
    std::mutex                             pool_mu;
    Pool                                   pool;
    std::vector<Slot>                      slots;
    std::mutex                             mu;
    std::condition_variable                cv;
    std::deque<Request*>                   submitted;    // guarded by mu
    std::vector<tucu_completion>           completed;    // guarded by mu
    std::unordered_map<uint64_t, Request*> live;         // guarded by mu
    uint64_t                               next_id = 1;  // guarded by mu
    std::deque<Request*>                   waiting;      // scheduler thread only
    std::unordered_map<uint32_t, Lobes>    lobe_cache;   // scheduler thread only
    std::thread                            thread;
    std::atomic<bool>                      stop{false};
    std::atomic<uint64_t>                  batches{0}, launched_reqs{0}, done_reqs{0}, max_batch{0};
    std::string                            error;        // guarded by mu

    double now() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

const Lobes& lobes_for(tucu_server_s& s, float sigma)
{
    uint32_t key;
    std::memcpy(&key, &sigma, 4);
    auto it = s.lobe_cache.find(key);
    if (it != s.lobe_cache.end()) return it->second;
    if (s.lobe_cache.size() > 65536) s.lobe_cache.clear();
    Lobes l{};
    int n[3];
    if (s.cfg.rule == TUCU_RULE_CEIL4SIGMA)
        gauss_lobes_ceil4(sigma, &l[0], &n[0], &l[17], &n[1], &l[34], &n[2]);
    else
        gauss_lobes_reference(sigma, &l[0], &n[0], &l[17], &n[1], &l[34], &n[2]);
    l[51] = float(n[0]); l[52] = float(n[1]); l[53] = float(n[2]);
    return s.lobe_cache.emplace(key, l).first->second;
}

void release_outputs(tucu_server_s& s, Request* r)
{
    std::lock_guard<std::mutex> lock(s.pool_mu);
    for (int k = 0; k < 6; ++k)
        if (r->bytes[k]) { s.pool.release(r->chunk[k], r->off[k], r->bytes[k]); r->bytes[k] = 0; }
}

void complete(tucu_server_s& s, Request* r, int status, double t)
{
    tucu_completion c;
    std::memset(&c, 0, sizeof(c));
    c.id = r->id;
    c.status = status;
    c.arrival = r->req.arrival;
    c.admitted = r->admitted;
    c.launched = r->launched;
    c.completed = t;
    if (status == TUCU_OK) {
        std::memcpy(c.out, r->out, sizeof(c.out));   // pool memory stays the request's until release
    } else {
        release_outputs(s, r);
        std::memset(r->out, 0, sizeof(r->out));
    }
    std::lock_guard<std::mutex> lock(s.mu);
    s.completed.push_back(c);
    r->done = true;                    // the last touch of r by the scheduler
    s.done_reqs++;
}

// Bytes of pool memory the outputs of request r take.
size_t output_bytes(const tucu_server_s& s, const tucu_request& q)
{
    const size_t es = dtype_size(s.cfg.out_dtype);
    const size_t row = round_up(size_t(q.in.ncol) * es, 128);
    return size_t(__builtin_popcount(q.mask)) * round_up(row * size_t(q.in.nrow), ALIGN);
}

// Output memory of request r from the pool; false if the pool is full.
bool try_allocate(tucu_server_s& s, Request* r)
{
    const int es = int(dtype_size(s.cfg.out_dtype));
    const int32_t h = r->req.in.nrow, w = r->req.in.ncol;
    const int64_t stride = int64_t(round_up(size_t(w) * es, 128) / es);
    for (int k = 0; k < 6; ++k) {
        if (!(r->req.mask >> k & 1u)) continue;
        const size_t bytes = size_t(h) * stride * es;
        size_t off;
        int ch;
        if (!s.pool.alloc(bytes, ch, off)) {
            for (int j = 0; j < k; ++j)
                if (r->bytes[j]) { s.pool.release(r->chunk[j], r->off[j], r->bytes[j]); r->bytes[j] = 0; }
            return false;
        }
        r->off[k] = off;
        r->chunk[k] = ch;
        r->bytes[k] = bytes;
        r->out[k] = {s.pool.base(ch) + off, stride, h, w, s.cfg.out_dtype, 0};
    }
    return true;
}

// Output memory of request r: from the pool, or from a new chunk when the
// admission rule (tucu_server_admit_bytes: the device's free memory less the
// reserve, within the pool cap) allows one.  1: allocated; 0: wait for
// memory to be released; -1: the request can never fit (nothing is in use,
// and the device cannot hold it).
int allocate_outputs(tucu_server_s& s, Request* r)
{
    std::lock_guard<std::mutex> lock(s.pool_mu);
    if (try_allocate(s, r)) return 1;
    size_t dev_free = 0, dev_total = 0;
    if (hipMemGetInfo(&dev_free, &dev_total) != hipSuccess) {
        (void)hipGetLastError();
        dev_free = 0;
    }
    const size_t need = output_bytes(s, r->req) + ALIGN;
    const size_t grow = tucu_server_admit_bytes(need, dev_free, s.cfg.reserve_bytes, s.pool.cap(),
                                                s.cfg.max_pool_bytes, s.cfg.grow_bytes);
    if (grow > 0) {
        void* p = nullptr;
        if (hipMalloc(&p, grow) == hipSuccess && p) {
            s.pool.add(static_cast<char*>(p), grow);
            if (try_allocate(s, r)) return 1;
        } else {
            (void)hipGetLastError();
        }
    }
    return s.pool.used() == 0 ? -1 : 0;
}

// Reaps finished slots; returns true if any finished.
bool reap(tucu_server_s& s)
{
    bool any = false;
    for (Slot& sl : s.slots) {
        if (!sl.busy) continue;
        const hipError_t q = hipEventQuery(sl.done);
        if (q == hipErrorNotReady) continue;
        const double t = s.now();
        const int status = q == hipSuccess ? TUCU_OK : TUCU_ERR_DEVICE;
        if (status != TUCU_OK) {
            std::lock_guard<std::mutex> lock(s.mu);
            s.error = std::string("batch failed: ") + hipGetErrorString(q);
        }
        for (Request* r : sl.reqs) complete(s, r, status, t);
        sl.reqs.clear();
        sl.busy = false;
        any = true;
    }
    return any;
}

// Admits waiting requests into free slot `sl` and launches them; returns
// true if a batch was launched.
bool admit_and_launch(tucu_server_s& s, Slot& sl)
{
    const BatchEntry& e = batch_entry(s.fam, s.thi, s.cfg.in_dtype, s.cfg.out_dtype);
    const int tile_w = e.tw ? e.tw : NT - 32;
    std::vector<SepArgs> jobs;
    std::vector<Request*> reqs;
    size_t tiles_est = 0;
    int64_t pixels = 0;
    const int64_t max_pixels = int64_t(s.cfg.max_batch_mpix) << 20;
    while (!s.waiting.empty() && int(jobs.size()) < s.cfg.max_batch_jobs) {
        Request* r = s.waiting.front();
        // tiles at the largest radius (16) and one extra row and column, an upper bound
        const size_t est = size_t((r->req.in.nrow + e.th - 1) / e.th + 1) * size_t(r->req.in.ncol / tile_w + 2);
        if (!jobs.empty() && tiles_est + est > size_t(s.cfg.max_tiles)) break;
        // launches of bounded work, so that several are in flight and
        // completions (and the pool memory they free) come steadily
        const int64_t px = int64_t(r->req.in.nrow) * r->req.in.ncol;
        if (!jobs.empty() && pixels + px > max_pixels) break;
        const int got = allocate_outputs(s, r);              // admission by free HBM: FIFO, head blocks
        if (got == 0) break;
        if (got < 0) {                                        // can never fit: fail it, go on
            s.waiting.pop_front();
            r->admitted = r->launched = s.now();
            complete(s, r, TUCU_ERR_BAD_ALLOC, s.now());
            continue;
        }
        pixels += px;
        const Lobes& l = lobes_for(s, r->req.sigma);
        const float* lob[3] = {&l[0], &l[17], &l[34]};
        const int n[3] = {int(l[51]), int(l[52]), int(l[53])};
        SepArgs a;
        int rmax = 0, odt = 0;
        const int rc = gauss_args(lob, n, &r->req.in, r->out, r->req.mask, r->req.shift, a, rmax, odt);
        s.waiting.pop_front();
        r->admitted = s.now();
        if (rc != TUCU_OK || !a.mask) {                       // invalid, or an image too small to write
            r->launched = r->admitted;
            complete(s, r, rc, s.now());
            continue;
        }
        tiles_est += est;
        jobs.push_back(a);
        reqs.push_back(r);
    }
    if (jobs.empty()) return false;
    std::vector<BatchTile> tiles;
    int rc = make_tiles(jobs, e, tiles);
    if (rc == TUCU_OK && tiles.size() > size_t(s.cfg.max_tiles)) rc = TUCU_ERR_INVALID;
    hipError_t err = hipSuccess;
    if (rc == TUCU_OK) {
        std::memcpy(sl.h_jobs, jobs.data(), jobs.size() * sizeof(SepArgs));
        std::memcpy(sl.h_tiles, tiles.data(), tiles.size() * sizeof(BatchTile));
        err = hipMemcpyAsync(sl.d_jobs, sl.h_jobs, jobs.size() * sizeof(SepArgs), hipMemcpyHostToDevice, sl.stream);
        if (err == hipSuccess)
            err = hipMemcpyAsync(sl.d_tiles, sl.h_tiles, tiles.size() * sizeof(BatchTile), hipMemcpyHostToDevice, sl.stream);
        if (err == hipSuccess)
            err = e.fn(sl.d_jobs, sl.d_tiles, int(tiles.size()), sl.d_counter, batch_grid(e, int(tiles.size())), sl.stream);
        if (err == hipSuccess) err = hipEventRecord(sl.done, sl.stream);
        if (err != hipSuccess) rc = TUCU_ERR_DEVICE;
    }
    const double t = s.now();
    if (rc != TUCU_OK) {
        for (Request* r : reqs) complete(s, r, rc, t);
        return false;
    }
    for (Request* r : reqs) r->launched = t;
    sl.reqs = std::move(reqs);
    sl.busy = true;
    s.batches++;
    s.launched_reqs += sl.reqs.size();
    uint64_t m = s.max_batch.load();
    while (sl.reqs.size() > m && !s.max_batch.compare_exchange_weak(m, sl.reqs.size())) {}
    return true;
}

void scheduler(tucu_server_s* s)
{
    (void)hipSetDevice(s->device);
    for (;;) {
        {
            std::lock_guard<std::mutex> lock(s->mu);
            while (!s->submitted.empty()) { s->waiting.push_back(s->submitted.front()); s->submitted.pop_front(); }
        }
        bool progress = reap(*s);
        for (Slot& sl : s->slots) {
            if (s->waiting.empty()) break;
            if (!sl.busy && admit_and_launch(*s, sl)) progress = true;
        }
        bool busy = false;
        for (const Slot& sl : s->slots) busy |= sl.busy;
        if (s->stop) {
            // Shutting down: requests not yet admitted are failed, in-flight batches finish.
            const double t = s->now();
            while (!s->waiting.empty()) {
                complete(*s, s->waiting.front(), TUCU_ERR_INVALID, t);
                s->waiting.pop_front();
            }
            if (!busy) break;
        }
        if (!progress) {
            std::unique_lock<std::mutex> lock(s->mu);
            // Poll in-flight batches every few tens of microseconds; wake at once on a submission.
            s->cv.wait_for(lock, std::chrono::microseconds(busy ? 20 : 200),
                           [s] { return !s->submitted.empty() || s->stop.load(); });
        }
    }
}

void free_pool(tucu_server_s* s)
{
    for (const auto& c : s->pool.chunks()) (void)hipFree(c.base);
}

void destroy_slots(tucu_server_s* s)
{
    for (Slot& sl : s->slots) {
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        if (sl.d_jobs) (void)hipFree(sl.d_jobs);
        if (sl.d_tiles) (void)hipFree(sl.d_tiles);
        if (sl.d_counter) (void)hipFree(sl.d_counter);
        if (sl.h_jobs) (void)hipHostFree(sl.h_jobs);
        if (sl.h_tiles) (void)hipHostFree(sl.h_tiles);
    }
    s->slots.clear();
}

}   // namespace

extern "C" {

size_t tucu_server_admit_bytes(size_t need, size_t device_free, size_t reserve, size_t pool_bytes,
                               size_t max_pool_bytes, size_t grow_step)
{
    const size_t avail = device_free > reserve ? device_free - reserve : 0;
    const size_t room = max_pool_bytes == 0 ? SIZE_MAX : max_pool_bytes > pool_bytes ? max_pool_bytes - pool_bytes : 0;
    const size_t limit = std::min(avail, room);
    if (need == 0 || need > limit) return 0;
    return std::min(limit, std::max(need, grow_step));
}

int tucu_server_create(tucu_server* server, const tucu_server_config* cfg)
{
    if (!server || !cfg) return fail(TUCU_ERR_INVALID, "null argument");
    *server = nullptr;
    if (dtype_size(cfg->in_dtype) == 0 || dtype_size(cfg->out_dtype) == 0)
        return fail(TUCU_ERR_INVALID, "unknown dtype");
    int thi = -1;
    const int fam = batch_family();
    const int rows = cfg->tile_rows > 0 ? cfg->tile_rows
                                        : batch_entry(fam, BATCH_TH_DEFAULT, cfg->in_dtype, cfg->out_dtype).th;
    for (int i = 0; i < NBTH; ++i)
        if (batch_entry(fam, i, cfg->in_dtype, cfg->out_dtype).th == rows &&
            batch_entry(fam, i, cfg->in_dtype, cfg->out_dtype).fn)
            thi = i;
    if (thi < 0) return fail(TUCU_ERR_INVALID, "no batched kernel for this tile height and type pair");
    if (cfg->pool_bytes == 0 && cfg->max_pool_bytes != 0 && cfg->max_pool_bytes < ALIGN)
        return fail(TUCU_ERR_INVALID, "empty output pool");
    auto s = new tucu_server_s();
    s->cfg = *cfg;
    if (s->cfg.max_pool_bytes != 0) s->cfg.max_pool_bytes = std::max(s->cfg.max_pool_bytes, s->cfg.pool_bytes);
    if (s->cfg.reserve_bytes == 0) s->cfg.reserve_bytes = size_t(1) << 30;
    if (s->cfg.grow_bytes == 0) s->cfg.grow_bytes = size_t(256) << 20;
    s->fam = fam;
    s->thi = thi;
    if (s->cfg.max_batch_jobs <= 0) s->cfg.max_batch_jobs = 1024;
    if (s->cfg.max_inflight <= 0) s->cfg.max_inflight = 2;
    if (s->cfg.max_tiles <= 0) s->cfg.max_tiles = 1 << 21;
    if (s->cfg.max_batch_mpix <= 0) s->cfg.max_batch_mpix = 64;
    s->t0 = std::chrono::steady_clock::now();
    (void)hipGetDevice(&s->device);
    if (s->cfg.pool_bytes > 0) {
        void* arena = nullptr;
        if (hipMalloc(&arena, s->cfg.pool_bytes) != hipSuccess || !arena) {
            (void)hipGetLastError();
            delete s;
            return fail(TUCU_ERR_BAD_ALLOC, "server output pool allocation failed");
        }
        s->pool.add(static_cast<char*>(arena), s->cfg.pool_bytes);
    }
    s->slots.resize(size_t(s->cfg.max_inflight));
    for (Slot& sl : s->slots) {
        hipError_t e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);   // default priority
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
        if (e == hipSuccess) e = hipMalloc(&sl.d_jobs, size_t(s->cfg.max_batch_jobs) * sizeof(SepArgs));
        if (e == hipSuccess) e = hipMalloc(&sl.d_tiles, size_t(s->cfg.max_tiles) * sizeof(BatchTile));
        if (e == hipSuccess) e = hipMalloc(&sl.d_counter, QUEUE_BYTES);
        if (e == hipSuccess) e = hipHostMalloc(&sl.h_jobs, size_t(s->cfg.max_batch_jobs) * sizeof(SepArgs));
        if (e == hipSuccess) e = hipHostMalloc(&sl.h_tiles, size_t(s->cfg.max_tiles) * sizeof(BatchTile));
        if (e != hipSuccess) {
            destroy_slots(s);
            free_pool(s);
            delete s;
            return device_fail(e, "server slot setup");
        }
    }
    s->thread = std::thread(scheduler, s);
    *server = s;
    return TUCU_OK;
}

int tucu_server_destroy(tucu_server server)
{
    if (!server) return TUCU_OK;
    server->stop = true;
    server->cv.notify_all();
    if (server->thread.joinable()) server->thread.join();
    destroy_slots(server);
    for (auto& kv : server->live) delete kv.second;
    free_pool(server);
    delete server;
    return TUCU_OK;
}

double tucu_server_now(tucu_server server)
{
    return server ? server->now() : 0.0;
}

namespace {

int check_request(const tucu_server_s& s, const tucu_request& req)
{
    if (req.in.dtype != s.cfg.in_dtype) return fail(TUCU_ERR_INVALID, "input dtype differs from the server's");
    if (req.mask == 0 || (req.mask & ~unsigned(TUCU_ALL6))) return fail(TUCU_ERR_INVALID, "bad output mask");
    if (!(req.sigma > 0.f) || !(req.sigma < 1e6f))                 // the rule tucu_gauss2_lobes applies
        return fail(TUCU_ERR_INVALID, "sigma must be positive and finite");
    if (!req.in.data || req.in.nrow <= 0 || req.in.ncol <= 0) return fail(TUCU_ERR_INVALID, "bad input image");
    const size_t cap = s.cfg.max_pool_bytes != 0 ? s.cfg.max_pool_bytes : SIZE_MAX;
    if (output_bytes(s, req) > cap)
        return fail(TUCU_ERR_BAD_ALLOC, "request larger than the server's output pool can grow");
    return TUCU_OK;
}

}   // namespace

int tucu_server_submit(tucu_server server, const tucu_request* req, uint64_t* id)
{
    return tucu_server_submit_many(server, req, 1, id);
}

int tucu_server_submit_many(tucu_server server, const tucu_request* reqs, int n, uint64_t* ids)
{
    if (!server || !reqs || n < 0) return fail(TUCU_ERR_INVALID, "null argument");
    int rc;
    for (int i = 0; i < n; ++i)
        if ((rc = check_request(*server, reqs[i]))) return rc;
    std::vector<Request*> rs(size_t(n), nullptr);
    for (int i = 0; i < n; ++i) {
        rs[size_t(i)] = new Request();
        rs[size_t(i)]->req = reqs[i];
        std::memset(rs[size_t(i)]->out, 0, sizeof(rs[size_t(i)]->out));
    }
    {
        std::lock_guard<std::mutex> lock(server->mu);
        if (!server->error.empty()) {
            for (Request* r : rs) delete r;
            return fail(TUCU_ERR_DEVICE, server->error);
        }
        for (Request* r : rs) {
            r->id = server->next_id++;
            server->live[r->id] = r;
            server->submitted.push_back(r);
        }
    }
    server->cv.notify_one();
    if (ids)
        for (int i = 0; i < n; ++i) ids[i] = rs[size_t(i)]->id;
    return TUCU_OK;
}

int tucu_server_poll(tucu_server server, tucu_completion* out, int max, int* n)
{
    if (!server || !n || (max > 0 && !out)) return fail(TUCU_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lock(server->mu);
    const int k = std::min<int>(max, int(server->completed.size()));
    for (int i = 0; i < k; ++i) out[i] = server->completed[size_t(i)];
    server->completed.erase(server->completed.begin(), server->completed.begin() + k);
    *n = k;
    return TUCU_OK;
}

int tucu_server_release(tucu_server server, uint64_t id)
{
    if (!server) return fail(TUCU_ERR_INVALID, "null server");
    Request* r = nullptr;
    {
        std::lock_guard<std::mutex> lock(server->mu);
        auto it = server->live.find(id);
        if (it == server->live.end()) return fail(TUCU_ERR_INVALID, "unknown or already released request");
        // Queued or in flight: the scheduler still holds the request and a
        // launch may still write its outputs.
        if (!it->second->done) return fail(TUCU_ERR_INVALID, "request not completed yet");
        r = it->second;
        server->live.erase(it);
    }
    release_outputs(*server, r);
    delete r;
    server->cv.notify_one();            // freed memory may admit waiting requests
    return TUCU_OK;
}

int tucu_server_stats(tucu_server server, tucu_server_counters* st)
{
    if (!server || !st) return fail(TUCU_ERR_INVALID, "null argument");
    std::memset(st, 0, sizeof(*st));
    st->batches = server->batches;
    st->requests_launched = server->launched_reqs;
    st->requests_done = server->done_reqs;
    st->max_batch = server->max_batch;
    std::lock_guard<std::mutex> lock(server->pool_mu);
    st->pool_bytes = server->pool.cap();
    st->pool_used = server->pool.used();
    st->pool_peak = server->pool.peak();
    return TUCU_OK;
}

}   // extern "C"
