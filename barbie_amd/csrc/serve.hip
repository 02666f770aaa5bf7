// Request server: many estimates of fields of any geometry on one GPU.
//
// The reference runs one field per process (barbie.c:58-397; the scripts
// run:1, many_fields.sh start one process per field).  Here one scheduler
// thread owns the device and serves submitted fields:
//  - plans are cached per geometry (host only);
//  - a lane is an estimator (own stream, max_batch fields) plus its input and
//    output buffers; a lane is created when its device bytes fit the HBM
//    budget, after evicting idle lanes of other geometries (least recently
//    used first);
//  - continuous batching: whenever a lane is idle, the next request (shortest
//    expected time first, with a waiting bound) takes it together with the
//    queued requests of the same geometry, up to max_batch, as one launch
//    sequence;
//  - up to max_inflight lanes run at once (one stream each), so small and
//    large estimates overlap on the device.
// A field's bins do not depend on what else runs: its lane's buffers are its
// own and every kernel sums in a fixed order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "barbie.h"
#include "device.hpp"
#include "plan.hpp"

using namespace bb;

namespace {

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// Geometry key: every field of bb_geometry, ring edges included.
std::string geometry_key(const bb_geometry &g)
{
    std::string k;
    auto put = [&](double v) { k.append(reinterpret_cast<const char *>(&v), sizeof v); };
    for (double v : {(double)g.NN, (double)g.Nmax, (double)g.Nr, (double)g.ring_rule,
                     g.ring_width, (double)g.last_ring_to_nmax, (double)g.n_mtbin,
                     (double)g.UL_k1, (double)g.LL_k1, (double)g.BW_k1, g.LL, g.ff, g.beta})
        put(v);
    if (g.ring_rule == BB_RINGS_EXPLICIT && g.ring_low && g.ring_high)
        for (int r = 0; r < g.Nr; ++r) {
            put(g.ring_low[r]);
            put(g.ring_high[r]);
        }
    return k;
}

struct Model {
    bb_plan *plan = nullptr;
    size_t fsz = 0;   // doubles per field: NN * NN * (NN + 2)
    size_t nb = 0;    // doubles per output array: nbins * 16
    bool tuned = false;
    bb_options eo{};  // estimator options of this geometry's lanes
    double svc_ms = 0.0;   // running estimate of one field's time on a lane (0: none yet)
    std::vector<double *> pinned;   // free pinned host field buffers
    ~Model()
    {
        for (double *p : pinned) (void)hipHostFree(p);
        bb_plan_destroy(plan);
    }
};

struct Req {
    long long id = 0;
    Model *m = nullptr;
    const double *src = nullptr;   // device field, or pinned copy of a host field
    double *pinned = nullptr;      // owned pinned copy (host fields)
    int state = BB_REQ_QUEUED;
    int err = BB_OK;
    std::string msg;
    std::vector<double> out;       // bs | norm | biana
    double t_submit = 0, t_start = 0, t_done = 0;
};

struct Lane {
    Model *m = nullptr;
    bb_estimator *est = nullptr;
    hipStream_t stream = nullptr;   // the server stream lent while busy
    int slot = -1;
    double *d_in = nullptr;     // batch fields
    double *d_out = nullptr;    // bs | norm | biana, each batch * nb
    double *h_out = nullptr;    // pinned copy of d_out (host-coherent)
    double *h_out_dev = nullptr;   // h_out's device address
    size_t bytes = 0;
    std::vector<Req *> reqs;    // in flight
    bool busy = false;
    double last_used = 0;
};

}  // namespace

struct bb_server {
    bb_server_options opt{};
    int device = 0;
    size_t budget = 0, used = 0;
    std::mutex mu;
    std::condition_variable wake, done;
    std::deque<Req *> queue;
    std::map<long long, Req *> reqs;
    std::map<std::string, std::unique_ptr<Model>> models;
    std::vector<std::unique_ptr<Lane>> lanes;
    long long next_id = 1;
    // max_inflight streams, lent to busy lanes.  With reserved CUs, streams
    // [0, nshort) run on the reserved CUs and take short requests only; the
    // others run on the remaining CUs.
    std::vector<hipStream_t> streams;
    std::vector<char> stream_busy;
    int reserved = 0;   // reserved CUs (0: no reservation)
    int nshort = 0;     // streams on the reserved CUs
    long long stats[8] = {0};   // submitted, completed, failed, batches, fields in batches,
                                // lanes created, lanes evicted, peak bytes
    bool stop = false;
    bool paused = false;   // no new batches start while set
    std::thread th;

    void loop();
    bool complete_lanes();
    bool dispatch();
    int make_lane(Model *m, Lane **out);
    void free_lane(Lane *L);
    void fail(Req *r, int code, const std::string &msg);
};

namespace {

// Copies a batch's outputs (three arrays of n_nb values, `stride` apart) into
// the lane's host-coherent pinned buffer with stores from the lane's own
// stream.  A copy-engine request (hipMemcpyAsync to the host) would wait in
// the engine's queue for this lane's kernels, and the requests of other lanes
// queued behind it on the same engine would wait too: a short request's
// output would come back only when a long one's does.
__global__ void k_store_out(const double *__restrict__ d, double *__restrict__ h, size_t n_nb,
                            size_t stride)
{
    // System-scope stores write through to the host memory; the kernel's
    // completion, which the scheduler sees before it reads h, orders them.
    // (A system-scope release fence instead would write back every dirty L2
    // line, those of the kernels running beside this one included.)
    const size_t step = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < 3 * n_nb; i += step) {
        const size_t a = i / n_nb, r = i - a * n_nb;
        __hip_atomic_store(h + a * stride + r, d[a * stride + r], __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

hipError_t launch_store_out(const double *d, double *h, size_t n_nb, size_t stride,
                            hipStream_t s)
{
    const size_t blocks = std::min<size_t>((3 * n_nb + 255) / 256, 1024);
    hipLaunchKernelGGL(k_store_out, dim3((unsigned)std::max<size_t>(blocks, 1)), dim3(256), 0, s,
                       d, h, n_nb, stride);
    return hipGetLastError();
}

int lane_bytes(Model *m, const bb_server_options &o, size_t *bytes)
{
    size_t b = 0;
    int rc = bb_estimator_footprint_opts(m->plan, &m->eo, &b);
    if (rc) return rc;
    b += (size_t)o.max_batch * m->fsz * 8 + 3 * (size_t)o.max_batch * m->nb * 8;
    *bytes = b;
    return BB_OK;
}

}  // namespace

void bb_server::fail(Req *r, int code, const std::string &msg)
{
    r->state = BB_REQ_FAILED;
    r->err = code;
    r->msg = msg;
    r->t_done = now_ms();
    stats[2]++;
}

void bb_server::free_lane(Lane *L)
{
    if (L->est) bb_estimator_destroy(L->est);
    if (L->d_in) (void)hipFree(L->d_in);
    if (L->d_out) (void)hipFree(L->d_out);
    if (L->h_out) (void)hipHostFree(L->h_out);
    used -= L->bytes;
}

int bb_server::make_lane(Model *m, Lane **out)
{
    *out = nullptr;
    size_t need = 0;
    int rc = lane_bytes(m, opt, &need);
    if (rc) return rc;
    // Evict idle lanes of other geometries, least recently used first.
    while (used + need > budget) {
        Lane *victim = nullptr;
        size_t vi = 0;
        for (size_t i = 0; i < lanes.size(); ++i) {
            Lane *L = lanes[i].get();
            if (!L->busy && L->m != m && (!victim || L->last_used < victim->last_used)) {
                victim = L;
                vi = i;
            }
        }
        if (!victim) break;
        free_lane(victim);
        lanes.erase(lanes.begin() + vi);
        stats[6]++;
    }
    if (used + need > budget) return BB_E_DEVMEM;   // caller waits or fails the request
    auto L = std::make_unique<Lane>();
    L->m = m;
    rc = bb_estimator_create_opts(m->plan, &m->eo, &L->est);
    const size_t B = (size_t)opt.max_batch;
    if (!rc && hipMalloc(&L->d_in, B * m->fsz * 8) != hipSuccess)
        rc = set_error(BB_E_DEVMEM, "lane input buffer");
    if (!rc && hipMalloc(&L->d_out, 3 * B * m->nb * 8) != hipSuccess)
        rc = set_error(BB_E_DEVMEM, "lane output buffer");
    if (!rc && hipHostMalloc(&L->h_out, 3 * B * m->nb * 8, hipHostMallocCoherent) != hipSuccess)
        rc = set_error(BB_E_NOMEM, "lane pinned output buffer");
    if (!rc && hipHostGetDevicePointer((void **)&L->h_out_dev, L->h_out, 0) != hipSuccess)
        rc = set_error(BB_E_DEVICE, "lane pinned output buffer: no device address");
    L->bytes = need;
    used += need;
    if (rc) {
        (void)hipGetLastError();
        free_lane(L.get());
        return rc;
    }
    stats[5]++;
    stats[7] = std::max<long long>(stats[7], (long long)used);
    *out = L.get();
    lanes.push_back(std::move(L));
    return BB_OK;
}

// Completed lanes hand their outputs to their requests.
bool bb_server::complete_lanes()
{
    bool any = false;
    for (auto &up : lanes) {
        Lane *L = up.get();
        if (!L->busy) continue;
        int d = 0;
        int rc = bb_estimator_query(L->est, &d);
        if (rc == BB_OK && !d) continue;
        const double t = now_ms();
        const size_t nb = L->m->nb, B = (size_t)opt.max_batch;
        for (size_t k = 0; k < L->reqs.size(); ++k) {
            Req *r = L->reqs[k];
            if (rc) {
                fail(r, rc, bb_last_error());
            } else {
                r->out.resize(3 * nb);
                for (int a = 0; a < 3; ++a)
                    std::memcpy(r->out.data() + a * nb, L->h_out + (a * B + k) * nb, nb * 8);
                r->state = BB_REQ_DONE;
                r->t_done = t;
                stats[1]++;
            }
            if (r->pinned) {
                L->m->pinned.push_back(r->pinned);
                r->pinned = nullptr;
            }
        }
        if (!L->reqs.empty() && L->reqs[0]->t_start > 0) {
            const double run = (t - L->reqs[0]->t_start) / (double)L->reqs.size();
            Model *m = L->m;
            m->svc_ms = m->svc_ms > 0 ? 0.8 * m->svc_ms + 0.2 * run : run;
        }
        L->reqs.clear();
        L->busy = false;
        L->last_used = t;
        // The server stream goes back to the pool: the idle lane returns to
        // its own stream, so that destroying it later (eviction) waits for
        // nothing another lane runs.
        (void)bb_estimator_set_stream(L->est, nullptr);
        L->stream = nullptr;
        if (L->slot >= 0) stream_busy[L->slot] = 0;
        L->slot = -1;
        any = true;
    }
    return any;
}

// Scheduling constants: a field expected to take longer than kLongMs on a
// lane is "long"; long fields may hold at most max_inflight - 1 lanes, so a
// short request always finds a lane soon.  A batch holds as many queued
// fields of its geometry as fit in about kBatchMs.
constexpr double kLongMs = 2.0;
constexpr double kBatchMs = 8.0;

// Picks the request that starts next: the one of the geometry with the
// shortest expected time (so short requests do not queue behind long ones),
// except that a request waiting longer than max(50 ms, 8 times its expected
// time) goes first, oldest first (no request waits forever).  Geometries not
// yet timed count as shortest; ties keep arrival order.  Long requests are
// skipped while `long_ok` is false.
Req *pick(std::deque<Req *> &queue, double now, bool long_ok)
{
    Req *best = nullptr;
    for (Req *r : queue) {
        const double svc = r->m->svc_ms;
        if (!long_ok && svc > kLongMs) continue;
        if (now - r->t_submit > std::max(50.0, 8.0 * svc)) return r;   // oldest overdue
        if (!best || svc < best->m->svc_ms) best = r;
    }
    return best;
}

// Starts batches while requests wait and lanes are free; later requests of
// the picked request's geometry join its batch.
bool bb_server::dispatch()
{
    bool any = false;
    while (!queue.empty() && !paused) {
        int busy = 0, busy_long = 0;
        for (auto &L : lanes) {
            busy += L->busy;
            busy_long += L->busy && L->m->svc_ms > kLongMs;
        }
        if (busy >= opt.max_inflight) break;
        // Long requests may not take the last free lane (with reserved CUs:
        // never stream 0), so a short request always finds one soon.
        bool long_ok = opt.max_inflight == 1 || busy_long < opt.max_inflight - 1;
        if (reserved > 0) {
            long_ok = false;
            for (size_t k = nshort; k < streams.size(); ++k) long_ok = long_ok || !stream_busy[k];
        }
        Req *r = pick(queue, now_ms(), long_ok);
        if (!r) break;   // only long requests wait, and the last lane is kept for short ones
        Model *m = r->m;
        const bool is_long = m->svc_ms > kLongMs;
        Lane *lane = nullptr;
        for (auto &L : lanes)
            if (!L->busy && L->m == m) {
                lane = L.get();
                break;
            }
        if (!lane) {
            const int rc = make_lane(m, &lane);
            if (rc == BB_E_DEVMEM && busy > 0) break;   // wait for a lane to finish
            if (rc) {
                fail(r, rc, rc == BB_E_DEVMEM ? "the request does not fit the HBM budget"
                                              : bb_last_error());
                queue.erase(std::find(queue.begin(), queue.end(), r));
                any = true;
                continue;
            }
        }
        std::vector<Req *> batch{r};
        queue.erase(std::find(queue.begin(), queue.end(), r));
        const int cap = m->svc_ms > 0
                            ? std::max(1, std::min(opt.max_batch, (int)(kBatchMs / m->svc_ms)))
                            : opt.max_batch;
        for (auto it = queue.begin(); it != queue.end() && (int)batch.size() < cap;) {
            if ((*it)->m == m) {
                batch.push_back(*it);
                it = queue.erase(it);
            } else {
                ++it;
            }
        }
        const size_t fsz = m->fsz, nb = m->nb, B = (size_t)opt.max_batch, n = batch.size();
        int rc = BB_OK;
        // Lend a free server stream (fewer busy lanes than streams, so one is
        // free): short requests prefer the reserved streams, long ones never
        // take them.
        int slot = (reserved > 0 && is_long) ? nshort : 0;
        while (slot < (int)streams.size() && stream_busy[slot]) ++slot;
        if (slot == (int)streams.size()) {
            rc = set_error(BB_E_DEVICE, "internal: no free stream");
        } else {
            lane->slot = slot;
            stream_busy[slot] = 1;
            lane->stream = streams[slot];
            rc = bb_estimator_set_stream(lane->est, lane->stream);
        }
        const double t = now_ms();
        for (size_t k = 0; k < n && !rc; ++k) {
            batch[k]->t_start = t;
            batch[k]->state = BB_REQ_RUNNING;
            const hipMemcpyKind kind = batch[k]->pinned ? hipMemcpyHostToDevice
                                                        : hipMemcpyDeviceToDevice;
            if (hipMemcpyAsync(lane->d_in + k * fsz, batch[k]->src, fsz * 8, kind,
                               lane->stream) != hipSuccess)
                rc = set_error(BB_E_DEVICE, "field copy failed");
        }
        if (!rc)
            rc = bb_estimate_batch_async(lane->est, (int)n, lane->d_in, lane->d_out,
                                         lane->d_out + B * nb, lane->d_out + 2 * B * nb);
        if (!rc && launch_store_out(lane->d_out, lane->h_out_dev, n * nb, B * nb,
                                    lane->stream) != hipSuccess)
            rc = set_error(BB_E_DEVICE, "output copy failed");
        if (rc) {
            if (lane->stream) (void)hipStreamSynchronize(lane->stream);
            (void)hipGetLastError();
            if (lane->slot >= 0) stream_busy[lane->slot] = 0;
            lane->slot = -1;
            for (Req *b : batch) {
                fail(b, rc, bb_last_error());
                if (b->pinned) {
                    m->pinned.push_back(b->pinned);
                    b->pinned = nullptr;
                }
            }
        } else {
            lane->reqs = batch;
            lane->busy = true;
            stats[3]++;
            stats[4] += (long long)n;
        }
        any = true;
    }
    return any;
}

void bb_server::loop()
{
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    while (!stop) {
        const bool c = complete_lanes();
        const bool d = dispatch();
        if (c || d) done.notify_all();
        if (c || d) continue;
        bool inflight = false;
        for (auto &L : lanes) inflight = inflight || L->busy;
        if (inflight)
            wake.wait_for(lk, std::chrono::microseconds(50));
        else
            wake.wait(lk, [&] { return stop || (!queue.empty() && !paused); });
    }
}

extern "C" {

int bb_server_create(const bb_server_options *o, bb_server **out)
{
    if (!o || !out) return set_error(BB_E_INVALID, "null argument");
    *out = nullptr;
    if (o->max_inflight < 0 || o->max_inflight > 16 || o->max_batch < 0 || o->max_batch > 256)
        return set_error(BB_E_INVALID, "max_inflight must be in [0, 16], max_batch in [0, 256]");
    if (o->precision != BB_FP64 && o->precision != BB_FP32 && o->precision != BB_BF16X3)
        return set_error(BB_E_INVALID, "unknown precision");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        (void)hipGetLastError();
        return set_error(BB_E_DEVICE, "no HIP device available");
    }
    size_t freeb = 0, total = 0;
    if (hipMemGetInfo(&freeb, &total) != hipSuccess) {
        (void)hipGetLastError();
        return set_error(BB_E_DEVICE, "hipMemGetInfo failed");
    }
    bb_server *S = new (std::nothrow) bb_server;
    if (!S) return set_error(BB_E_NOMEM, "out of host memory");
    S->opt = *o;
    if (S->opt.max_inflight == 0) S->opt.max_inflight = 4;
    if (S->opt.max_batch == 0) S->opt.max_batch = 4;
    S->device = dev;
    S->budget = o->hbm_budget ? o->hbm_budget : freeb / 10 * 8;
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        ncu = 0;
    const int want = o->reserved_cus < 0 ? 0 : o->reserved_cus > 0 ? o->reserved_cus : ncu / 8;
    if (S->opt.max_inflight >= 2 && ncu >= 16 && want > 0 && want < ncu) {
        S->reserved = want;
        S->nshort = std::max(1, S->opt.max_inflight / 2);
    }
    for (int k = 0; k < S->opt.max_inflight; ++k) {
        hipStream_t st = nullptr;
        hipError_t e;
        if (S->reserved > 0) {
            std::vector<uint32_t> mask((ncu + 31) / 32, 0u);
            for (int cu = 0; cu < ncu; ++cu)
                if ((cu < S->reserved) == (k < S->nshort)) mask[cu / 32] |= 1u << (cu % 32);
            e = hipExtStreamCreateWithCUMask(&st, (uint32_t)mask.size(), mask.data());
        } else {
            e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        }
        if (e != hipSuccess) {
            (void)hipGetLastError();
            for (hipStream_t x : S->streams) (void)hipStreamDestroy(x);
            delete S;
            return set_error(BB_E_DEVICE, "stream creation failed");
        }
        S->streams.push_back(st);
        S->stream_busy.push_back(0);
    }
    S->th = std::thread([S] { S->loop(); });
    *out = S;
    return BB_OK;
}

void bb_server_destroy(bb_server *S)
{
    if (!S) return;
    {
        std::lock_guard<std::mutex> g(S->mu);
        S->stop = true;
    }
    S->wake.notify_all();
    if (S->th.joinable()) S->th.join();
    for (auto &L : S->lanes) {
        if (L->busy) (void)bb_estimator_wait(L->est);
        S->free_lane(L.get());
    }
    for (hipStream_t x : S->streams) (void)hipStreamDestroy(x);
    for (auto &kv : S->reqs) {
        if (kv.second->pinned) (void)hipHostFree(kv.second->pinned);
        delete kv.second;
    }
    delete S;
}

namespace {

// The model (plan and lane options) of a geometry, created on first use.
int get_model(bb_server *S, const bb_geometry *g, Model **out)
{
    const std::string key = geometry_key(*g);
    {
        std::lock_guard<std::mutex> lk(S->mu);
        auto it = S->models.find(key);
        if (it != S->models.end()) {
            *out = it->second.get();
            return BB_OK;
        }
    }
    auto nm = std::make_unique<Model>();
    int rc = bb_plan_create(g, &nm->plan);
    if (rc) return rc;
    bb_plan_stats st;
    if ((rc = bb_plan_stats_get(nm->plan, &st))) return rc;
    nm->fsz = (size_t)g->NN * g->NN * (g->NN + 2);
    nm->nb = (size_t)st.nbins * BB_NPOLES;
    std::memset(&nm->eo, 0, sizeof nm->eo);
    nm->eo.precision = S->opt.precision;
    nm->eo.max_batch = S->opt.max_batch;
    if (S->opt.tune) {
        // Time the candidate layouts once per geometry, here in the caller's
        // thread (on the server's device) and outside the server lock, so
        // that submit, poll and wait of other requests never block on it
        // (lanes that run meanwhile can slow the candidates;
        // bb_server_prepare before traffic avoids that).  With reserved CUs
        // every lane runs on a CU-masked server stream, and the mixed BF16x3
        // layout, which puts its FP32 share on a stream of its own across
        // all CUs, is not a candidate.
        int prev = 0;
        const bool swap = hipGetDevice(&prev) == hipSuccess && prev != S->device;
        if (swap) (void)hipSetDevice(S->device);
        bb_options best;
        if (bb::tune_layout(nm->plan, &nm->eo, nullptr, 2, S->reserved <= 0, &best, nullptr) ==
            BB_OK)
            nm->eo = best;
        if (swap) (void)hipSetDevice(prev);
        (void)hipGetLastError();
        nm->tuned = true;
    }
    if (S->reserved > 0) nm->eo.mix_fp32 = 0;
    std::lock_guard<std::mutex> lk(S->mu);
    auto ins = S->models.emplace(key, std::move(nm));
    *out = ins.first->second.get();
    return BB_OK;
}

}  // namespace

int bb_server_prepare(bb_server *S, const bb_geometry *g, int nlanes)
{
    if (!S || !g) return set_error(BB_E_INVALID, "null argument");
    Model *m = nullptr;
    int rc = get_model(S, g, &m);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(S->mu);
    int have = 0;
    for (auto &L : S->lanes) have += L->m == m;
    for (; have < std::min(nlanes, S->opt.max_inflight); ++have) {
        Lane *L = nullptr;
        if ((rc = S->make_lane(m, &L))) return rc == BB_E_DEVMEM ? set_error(rc, "lanes do not fit the HBM budget") : rc;
    }
    return BB_OK;
}

int bb_submit(bb_server *S, const bb_geometry *g, const double *field_r2c, int on_device,
              long long *ticket)
{
    if (!S || !g || !field_r2c || !ticket) return set_error(BB_E_INVALID, "null argument");
    Model *m = nullptr;
    int rc = get_model(S, g, &m);
    if (rc) return rc;
    auto *r = new (std::nothrow) Req;
    if (!r) return set_error(BB_E_NOMEM, "out of host memory");
    r->m = m;
    r->t_submit = now_ms();
    if (on_device) {
        r->src = field_r2c;
    } else {
        double *p = nullptr;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            if (!m->pinned.empty()) {
                p = m->pinned.back();
                m->pinned.pop_back();
            }
        }
        if (!p && hipHostMalloc(&p, m->fsz * 8, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            delete r;
            return set_error(BB_E_NOMEM, "pinned host buffer for the field");
        }
        std::memcpy(p, field_r2c, m->fsz * 8);
        r->pinned = p;
        r->src = p;
    }
    {
        std::lock_guard<std::mutex> lk(S->mu);
        r->id = S->next_id++;
        S->reqs[r->id] = r;
        S->queue.push_back(r);
        S->stats[0]++;
    }
    S->wake.notify_all();
    *ticket = r->id;
    return BB_OK;
}

int bb_server_pause(bb_server *S, int paused)
{
    if (!S) return set_error(BB_E_INVALID, "null argument");
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->paused = paused != 0;
    }
    S->wake.notify_all();
    return BB_OK;
}

int bb_poll(bb_server *S, long long ticket, int *state)
{
    if (!S || !state) return set_error(BB_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(ticket);
    if (it == S->reqs.end()) return set_error(BB_E_INVALID, "unknown ticket");
    *state = it->second->state;
    return BB_OK;
}

int bb_wait(bb_server *S, long long ticket, double *bin_bs, double *bin_norm, double *bin_biana,
            double *latency_ms)
{
    if (!S) return set_error(BB_E_INVALID, "null argument");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(ticket);
    if (it == S->reqs.end()) return set_error(BB_E_INVALID, "unknown ticket");
    Req *r = it->second;
    S->done.wait(lk, [&] { return r->state == BB_REQ_DONE || r->state == BB_REQ_FAILED; });
    S->reqs.erase(it);
    lk.unlock();
    std::unique_ptr<Req> own(r);
    if (latency_ms) {
        latency_ms[0] = r->t_start > 0 ? r->t_start - r->t_submit : r->t_done - r->t_submit;
        latency_ms[1] = r->t_start > 0 ? r->t_done - r->t_start : 0.0;
        latency_ms[2] = r->t_done - r->t_submit;
    }
    if (r->state == BB_REQ_FAILED) return set_error(r->err, r->msg);
    const size_t nb = r->m->nb;
    if (bin_bs) std::memcpy(bin_bs, r->out.data(), nb * 8);
    if (bin_norm) std::memcpy(bin_norm, r->out.data() + nb, nb * 8);
    if (bin_biana) std::memcpy(bin_biana, r->out.data() + 2 * nb, nb * 8);
    return BB_OK;
}

int bb_server_stats(bb_server *S, long long *out)
{
    if (!S || !out) return set_error(BB_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    for (int k = 0; k < 8; ++k) out[k] = S->stats[k];
    long long busy = 0;
    for (auto &L : S->lanes) busy += L->busy;
    out[8] = (long long)S->queue.size();
    out[9] = busy;
    out[10] = (long long)S->lanes.size();
    out[11] = (long long)S->used;
    out[12] = (long long)S->budget;
    return BB_OK;
}

}  // extern "C"
