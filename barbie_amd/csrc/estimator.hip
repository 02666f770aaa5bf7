// Estimator: device memory, work lists and the per-field pipeline
// (plane gather -> plane-ring cache -> contraction -> binning).
// Replaces barbie.c:150-373 (cache build, ring triples, hot loop) for one
// field held in HBM.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <array>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "device.hpp"
#include "plan.hpp"

using namespace bb;

#define HIPCHK(expr)                                                                        \
    do {                                                                                    \
        hipError_t _e = (expr);                                                             \
        if (_e != hipSuccess)                                                               \
            return set_error(BB_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

namespace {

__global__ void k_fill_int(int *p, int n, int v)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

bool make_fft_desc(int N, FftDesc &fd)
{
    fd.N = N;
    fd.nst = 0;
    int n = N;
    while (n % 4 == 0 && fd.nst < kMaxStages) { fd.radix[fd.nst++] = 4; n /= 4; }
    while (n % 2 == 0 && fd.nst < kMaxStages) { fd.radix[fd.nst++] = 2; n /= 2; }
    while (n % 3 == 0 && fd.nst < kMaxStages) { fd.radix[fd.nst++] = 3; n /= 3; }
    while (n % 5 == 0 && fd.nst < kMaxStages) { fd.radix[fd.nst++] = 5; n /= 5; }
    return n == 1;
}

bool smooth235(int n)
{
    if (n < 1) return false;
    for (int f : {2, 3, 5})
        while (n % f == 0) n /= f;
    return n == 1;
}

int isqrt_floor(int v)
{
    if (v < 0) return -1;
    int r = (int)std::sqrt((double)v);
    while ((long long)r * r > v) --r;
    while ((long long)(r + 1) * (r + 1) <= v) ++r;
    return r;
}

template <typename T>
int dev_alloc(T **out, size_t count, std::vector<void *> &owned)
{
    *out = nullptr;
    if (count == 0) count = 1;
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return set_error(BB_E_DEVMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) +
                                          " bytes failed: " + hipGetErrorString(e));
    }
    owned.push_back(p);
    *out = (T *)p;
    return BB_OK;
}

// ---------------------------------------------------------------------------
// Work lists.  One contraction of a plane group: rings and its index.
struct CItem { int r1, r2, r3, cid; };

// A pair block: up to 16 (r1, r2) rows of one plane group whose r3 rings
// span at most kWinRings, covered by nwin windows starting at lo.
constexpr int kWinRings = 16 * 3;
struct PairBlock {
    int pa, pb, pc;           // planes of D1, D2, D3 (p1, |p2|, |p3|)
    int flags;                // TW_CONJ2 | TW_CONJ3
    int r1[16], r2[16];
    int nrow;
    int lo, hi, nwin;
    std::vector<std::array<int, 3>> cs;   // (row, r3, cid)
};

void pair_blocks(std::vector<CItem> &cs, int pa, int pb, int pc, int flags,
                 std::vector<PairBlock> &out)
{
    std::sort(cs.begin(), cs.end(), [](const CItem &a, const CItem &b) {
        if (a.r1 != b.r1) return a.r1 < b.r1;
        if (a.r2 != b.r2) return a.r2 < b.r2;
        return a.r3 < b.r3;
    });
    struct Pair { int r1, r2, lo, hi; size_t b, e; };
    std::vector<Pair> pairs;
    for (size_t k = 0; k < cs.size();) {
        size_t e = k;
        while (e < cs.size() && cs[e].r1 == cs[k].r1 && cs[e].r2 == cs[k].r2 &&
               cs[e].r3 - cs[k].r3 < kWinRings)
            ++e;
        pairs.push_back({cs[k].r1, cs[k].r2, cs[k].r3, cs[e - 1].r3, k, e});
        k = e;
    }
    std::sort(pairs.begin(), pairs.end(), [](const Pair &a, const Pair &b) {
        if (a.lo != b.lo) return a.lo < b.lo;
        if (a.hi != b.hi) return a.hi < b.hi;
        if (a.r1 != b.r1) return a.r1 < b.r1;
        return a.r2 < b.r2;
    });
    for (size_t q = 0; q < pairs.size();) {
        size_t e = q + 1;
        int lo = pairs[q].lo, hi = pairs[q].hi;
        while (e < pairs.size() && e - q < 16) {
            const int l2 = std::min(lo, pairs[e].lo), h2 = std::max(hi, pairs[e].hi);
            if (h2 - l2 + 1 > kWinRings) break;
            lo = l2;
            hi = h2;
            ++e;
        }
        PairBlock B;
        B.pa = pa; B.pb = pb; B.pc = pc; B.flags = flags;
        B.nrow = (int)(e - q);
        B.lo = lo; B.hi = hi; B.nwin = (hi - lo) / 16 + 1;
        for (int m = 0; m < 16; ++m) { B.r1[m] = -1; B.r2[m] = -1; }
        for (size_t u = q; u < e; ++u) {
            const int m = (int)(u - q);
            B.r1[m] = pairs[u].r1;
            B.r2[m] = pairs[u].r2;
            for (size_t k = pairs[u].b; k < pairs[u].e; ++k) B.cs.push_back({m, cs[k].r3, cs[k].cid});
        }
        out.push_back(std::move(B));
        q = e;
    }
}

struct WorkLists {
    std::vector<WorkItem> items;
    std::vector<TWin> tws;
    int ntw = 0;               // output tiles
    int max_rows = 0;          // largest staged ring count of an item
    long long tile_slots = 0;  // ntw * 256
    long long tw_used = 0, tw_slots = 0;   // wave balance: sum cnt, sum waves * max cnt
};

// Packs pair blocks (already grouped so that neighbours share planes) into
// work items.  slot(plane, ring) gives the cache slot; split_b: stage the r3
// rings as bf16 hi/lo operands (BF16x3) instead of raw values.
template <typename Slot>
int pack_items(std::vector<PairBlock> &blocks, bool split_b, int tww, int waves, Slot slot,
               WorkLists &W, std::vector<int32_t> &pos)
{
    const int rowcap = rowcap_for(waves);
    struct Key { int plane, split; bool operator<(const Key &o) const { return plane != o.plane ? plane < o.plane : split < o.split; } };
    struct Range { int lo, hi; };
    size_t a = 0;
    std::vector<std::pair<long long, size_t>> order;   // (work, item index) for LPT ordering
    while (a < blocks.size()) {
        std::map<Key, Range> keys;
        int load[kWaves] = {0};
        std::vector<std::vector<size_t>> wave_blocks(kWaves);
        size_t b = a;
        for (; b < blocks.size(); ++b) {
            const PairBlock &B = blocks[b];
            std::map<Key, Range> k2 = keys;
            auto add = [&](Key k, int lo, int hi) {
                auto it = k2.find(k);
                if (it == k2.end()) k2[k] = {lo, hi};
                else { it->second.lo = std::min(it->second.lo, lo); it->second.hi = std::max(it->second.hi, hi); }
            };
            for (int m = 0; m < B.nrow; ++m) {
                add({B.pa, 0}, B.r1[m], B.r1[m]);
                add({B.pb, 0}, B.r2[m], B.r2[m]);
            }
            add({B.pc, split_b ? 1 : 0}, B.lo, B.hi);
            int rows = 0;
            for (auto &kv : k2) rows += kv.second.hi - kv.second.lo + 1;
            int w = 0;
            for (int v = 1; v < waves; ++v)
                if (load[v] < load[w]) w = v;
            const bool fits = rows <= rowcap && (int)k2.size() <= kMaxBlk && load[w] + B.nwin <= tww;
            if (!fits && b == a)
                return set_error(BB_E_INVALID, "a pair block needs " + std::to_string(rows) +
                                                   " staged rings; the contraction kernel holds " +
                                                   std::to_string(rowcap) + " (too many rings)");
            if (!fits) break;
            keys.swap(k2);
            load[w] += B.nwin;
            wave_blocks[w].push_back(b);
        }
        // Lay out the item: blocks in key order.
        WorkItem it;
        std::memset(&it, 0, sizeof it);
        std::map<Key, int> row0;
        int nrows = 0;
        for (auto &kv : keys) {
            const int bi = it.nblk++;
            it.blk_slot[bi] = slot(kv.first.plane, kv.second.lo);
            it.blk_rows[bi] = (int16_t)(kv.second.hi - kv.second.lo + 1);
            it.blk_split[bi] = (int16_t)kv.first.split;
            row0[kv.first] = nrows - kv.second.lo;   // row of ring r = row0 + r
            nrows += it.blk_rows[bi];
        }
        it.nrows = (int16_t)nrows;
        it.tw0 = (int32_t)W.tws.size();
        W.max_rows = std::max(W.max_rows, nrows);
        {
            TWin dummy;   // computes into an accumulator that is never written out
            std::memset(&dummy, 0, sizeof dummy);
            for (int m = 0; m < 16; ++m) dummy.a1[m] = -1;
            dummy.flags = TW_NEWA;
            dummy.out = -1;
            W.tws.resize(W.tws.size() + kWaves * kTWW, dummy);
        }
        long long work = 0;
        for (int w = 0; w < waves; ++w) {
            int j = 0;
            for (size_t bi : wave_blocks[w]) {
                const PairBlock &B = blocks[bi];
                const int ra = row0[{B.pa, 0}], rb = row0[{B.pb, 0}];
                const int rc = row0[{B.pc, split_b ? 1 : 0}];
                const int tile0 = W.ntw;
                for (int v = 0; v < B.nwin; ++v) {
                    TWin &T = W.tws[it.tw0 + w * kTWW + j];
                    std::memset(&T, 0, sizeof T);
                    for (int m = 0; m < 16; ++m) {
                        T.a1[m] = (int16_t)(B.r1[m] >= 0 ? ra + B.r1[m] : -1);
                        T.a2[m] = (int16_t)(B.r2[m] >= 0 ? rb + B.r2[m] : 0);
                    }
                    const int ws = B.lo + 16 * v;
                    T.b0 = (int16_t)(rc + ws);
                    T.nb = (int16_t)std::min(16, B.hi - ws + 1);
                    T.flags = (int16_t)(B.flags | (v == 0 ? TW_NEWA : 0));
                    T.out = W.ntw++;
                    ++j;
                }
                for (const auto &c : B.cs) {
                    const int d = c[1] - B.lo;
                    pos[c[2]] = (tile0 + d / 16) * 256 + c[0] * 16 + d % 16;
                }
            }
            it.cnt[w] = (int8_t)j;
            work = std::max<long long>(work, j);
            W.tw_used += j;
        }
        W.tw_slots += (long long)waves * work;
        order.push_back({work * 1000000LL + nrows, W.items.size()});
        W.items.push_back(it);
        a = b;
    }
    // Largest items first (longest-processing-time order).
    std::stable_sort(order.begin(), order.end(),
                     [](const std::pair<long long, size_t> &x, const std::pair<long long, size_t> &y) {
                         return x.first > y.first;
                     });
    std::vector<WorkItem> sorted;
    sorted.reserve(W.items.size());
    for (auto &o : order) sorted.push_back(W.items[o.second]);
    W.items.swap(sorted);
    W.tile_slots = (long long)W.ntw * 256;
    return BB_OK;
}

// Checks the invariants the contraction kernels rely on (every LDS row index
// inside the item's staged rows, every output slot used once).
int check_lists(const WorkLists &W, const std::vector<int32_t> &pos, size_t ncontr)
{
    std::vector<uint8_t> seen((size_t)std::max<long long>(W.tile_slots, 1), 0);
    for (size_t c = 0; c < ncontr; ++c) {
        if (pos[c] < 0 || pos[c] >= W.tile_slots || seen[pos[c]])
            return set_error(BB_E_INVALID, "internal: contraction output slots are not unique");
        seen[pos[c]] = 1;
    }
    for (const WorkItem &it : W.items) {
        int rows = 0;
        for (int b = 0; b < it.nblk; ++b) rows += it.blk_rows[b];
        if (rows != it.nrows || it.nrows > kRowCap || it.nblk > kMaxBlk)
            return set_error(BB_E_INVALID, "internal: work item row blocks are inconsistent");
        for (int w = 0; w < kWaves; ++w) {
            if (it.cnt[w] < 0 || it.cnt[w] > kTWW)
                return set_error(BB_E_INVALID, "internal: too many tile windows in a wave");
            for (int j = it.cnt[w]; j < kTWW; ++j)
                if (W.tws[(size_t)it.tw0 + w * kTWW + j].out != -1)
                    return set_error(BB_E_INVALID, "internal: unused tile window slot is live");
            for (int j = 0; j < it.cnt[w]; ++j) {
                const TWin &T = W.tws[(size_t)it.tw0 + w * kTWW + j];
                bool ok = T.b0 >= 0 && T.b0 < it.nrows && T.nb >= 1 && T.nb <= 16 &&
                          T.out >= 0 && T.out < W.ntw && (j > 0 || (T.flags & TW_NEWA));
                for (int m = 0; m < 16; ++m)
                    ok = ok && T.a1[m] >= -1 && T.a1[m] < it.nrows && T.a2[m] >= 0 &&
                         T.a2[m] < it.nrows;
                if (!ok) return set_error(BB_E_INVALID, "internal: tile window out of range");
            }
        }
    }
    return BB_OK;
}

int auto_splits(int nitems)
{
    int s = (1024 + std::max(nitems, 1) - 1) / std::max(nitems, 1);
    return std::max(1, std::min(s, 64));
}

}  // namespace

struct bb_estimator {
    const bb_plan *plan = nullptr;
    bb_options opt{};
    CPrec cprec = CPrec::F64;
    hipStream_t stream = nullptr;
    FftDesc fd{};
    int NN = 0, NG = 0, kmax = 0;
    double scale = 1.0;            // (NN/NG)^2
    long long npts = 0;            // NG^2
    std::vector<int> rlo, rhi, base;  // per plane ring range and first slot
    int nfields = 0;
    std::vector<void *> owned;
    double2 *d_tw = nullptr;
    double2 *d_field = nullptr;    // staging for bb_estimate (host input)
    double2 *d_planes = nullptr;
    void *d_cache = nullptr;
    void *d_split = nullptr;       // BF16x3: bf16 hi/lo operands of every field
    double2 *d_rowbuf = nullptr;   // FP32 caches: FP64 row-pass scratch
    int rowbuf_fields = 0;
    FieldTask *d_tasks = nullptr;
    int ntasks = 0;
    int max_rows = 0;
    WorkItem *d_items = nullptr;
    TWin *d_tws = nullptr;
    int nitems = 0, nsplit = 1, chunk = 32, max_fields = 0, waves = kWaves, tww = kTWW;
    double wave_balance = 1.0;
    long long staged_rows = 0;
    long long tile_slots = 0;
    double *d_part = nullptr;      // [nsplit][tile_slots]
    std::vector<int32_t> pos;      // contraction -> tile position
    BinRec *d_recs = nullptr;
    int32_t *d_seg = nullptr;
    double *d_num = nullptr;
    std::vector<double> num_host;
    int *d_nz = nullptr;
    double *d_avg = nullptr;
    double *d_bins = nullptr;
    BinParams bp{};
    hipEvent_t ev[5] = {};
    float ms[4] = {0, 0, 0, 0};

    int fid(int plane, int ring) const
    {
        if (plane < 0 || plane >= (int)rlo.size() || rlo[plane] < 0) return -1;
        if (ring < rlo[plane] || ring > rhi[plane]) return -1;
        return base[plane] + ring - rlo[plane];
    }
    ~bb_estimator()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
        for (void *p : owned) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

int upload(void *dst, const void *src, size_t bytes, hipStream_t s)
{
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    return BB_OK;
}

// FFT pipeline for a list of tasks into a cache (FP64 in place, or FP32 via
// the FP64 row-pass scratch).
int run_fields(bb_estimator *E, const double2 *planes, const FieldTask *d_tasks, int ntasks,
               int max_rows, void *cache, bool f32, int *nz)
{
    const long long NG2 = E->npts;
    if (!f32) {
        HIPCHK(launch_rowpass(planes, (double2 *)cache, NG2, d_tasks, ntasks, max_rows, E->fd,
                              E->d_tw, nz, E->stream));
        HIPCHK(launch_colpass_f64((double2 *)cache, NG2, (double2 *)cache, d_tasks, ntasks, E->fd,
                                  E->d_tw, E->stream));
        return BB_OK;
    }
    for (int t0 = 0; t0 < ntasks; t0 += E->rowbuf_fields) {
        int nb = std::min(E->rowbuf_fields, ntasks - t0);
        HIPCHK(launch_rowpass(planes, E->d_rowbuf, -NG2, d_tasks + t0, nb, max_rows, E->fd,
                              E->d_tw, nz, E->stream));
        HIPCHK(launch_colpass_f32(E->d_rowbuf, -NG2, (float2 *)cache, E->d_split, d_tasks + t0,
                                  nb, E->fd, E->d_tw, E->stream));
    }
    return BB_OK;
}

int choose_grid(bb_estimator *E, const PlanData &d)
{
    const int NN = d.g.NN;
    int kmax = 0;
    for (int r = 0; r < d.g.Nr; ++r)
        if (d.nhi[r] > 0) kmax = std::max(kmax, isqrt_floor(d.nhi[r] - 1));
    E->kmax = kmax;
    const int need = 3 * kmax + 1;
    E->NG = NN;
    if (!E->opt.full_grid && NN >= need) {
        for (int n = need; n <= NN; ++n)
            if ((n % 2 == 0) && smooth235(n)) {
                E->NG = n;
                break;
            }
    }
    E->scale = ((double)NN * NN) / ((double)E->NG * E->NG);
    E->npts = (long long)E->NG * E->NG;
    return BB_OK;
}

int estimator_init(bb_estimator *E)
{
    const PlanData &d = E->plan->d;
    const int NN = d.g.NN, Nmax = d.g.Nmax, Nr = d.g.Nr;
    E->NN = NN;
    choose_grid(E, d);
    const int NG = E->NG;
    if (!make_fft_desc(NG, E->fd)) return set_error(BB_E_INVALID, "unsupported FFT length");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        (void)hipGetLastError();
        return set_error(BB_E_DEVICE, "no HIP device available");
    }
    HIPCHK(hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking));
    for (auto &e : E->ev) HIPCHK(hipEventCreate(&e));
    int rc;
    const bool f32 = (E->cprec != CPrec::F64);

    // Twiddles exp(+2 pi i t / NG), FP64.
    std::vector<double2> tw(NG);
    for (int t = 0; t < NG; ++t) {
        long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)t / NG;
        tw[t] = make_double2((double)cosl(a), (double)sinl(a));
    }
    if ((rc = dev_alloc(&E->d_tw, NG, E->owned))) return rc;
    if ((rc = upload(E->d_tw, tw.data(), sizeof(double2) * NG, E->stream))) return rc;

    std::vector<int> rowband(Nr);
    for (int r = 0; r < Nr; ++r)
        rowband[r] = d.nhi[r] > 0 ? std::min(NG / 2, isqrt_floor(d.nhi[r] - 1)) : -1;

    // --- Triangle counts Num on the GPU (calCR, funcs.c:328-424) -----------
    // Unit fields I_r = IFFT2(mask_r) in FP64, then Num = sum_x I1 I2 I3 for
    // every ring triple with the FP64 contraction kernel.
    {
        const int rcount = (int)(d.tri.size() / 3);
        std::vector<FieldTask> ut(Nr);
        int mr = 0;
        for (int r = 0; r < Nr; ++r) {
            ut[r] = FieldTask{-1, r, r, rowband[r], d.nlo[r], d.nhi[r]};
            mr = std::max(mr, std::min(NG, 2 * rowband[r] + 1));
        }
        std::vector<CItem> cs;
        for (int rp = 0; rp < rcount; ++rp)
            cs.push_back({d.tri[3 * rp], d.tri[3 * rp + 1], d.tri[3 * rp + 2], rp});
        WorkLists W;
        std::vector<int32_t> upos(std::max(rcount, 1), 0);
        std::vector<PairBlock> blocks;
        if (rcount > 0) pair_blocks(cs, 0, 0, 0, 0, blocks);
        rc = pack_items(blocks, false, contract_tww(CPrec::F64), kWaves,
                        [](int, int r) { return r; }, W, upos);
        if (!rc) rc = check_lists(W, upos, rcount);
        if (rc) return rc;
        const int usplit = auto_splits((int)W.items.size());
        const long long uslots = std::max<long long>(W.tile_slots, 1);
        std::vector<void *> tmp;
        FieldTask *d_ut;
        double2 *d_unit;
        WorkItem *d_uitems;
        TWin *d_utw;
        double *d_upart;
        rc = dev_alloc(&d_ut, Nr, tmp);
        if (!rc) rc = dev_alloc(&d_unit, (size_t)Nr * E->npts, tmp);
        if (!rc) rc = dev_alloc(&d_uitems, W.items.size(), tmp);
        if (!rc) rc = dev_alloc(&d_utw, W.tws.size(), tmp);
        if (!rc) rc = dev_alloc(&d_upart, (size_t)usplit * uslots, tmp);
        if (!rc) rc = upload(d_ut, ut.data(), sizeof(FieldTask) * Nr, E->stream);
        if (!rc && !W.items.empty())
            rc = upload(d_uitems, W.items.data(), sizeof(WorkItem) * W.items.size(), E->stream);
        if (!rc && !W.tws.empty())
            rc = upload(d_utw, W.tws.data(), sizeof(TWin) * W.tws.size(), E->stream);
        if (!rc) rc = run_fields(E, nullptr, d_ut, Nr, mr, d_unit, false, nullptr);
        if (!rc && rcount > 0) {
            hipError_t e = launch_contract(CPrec::F64, kWaves, contract_tww(CPrec::F64), d_unit,
                                           nullptr, E->npts, d_uitems,
                                           (int)W.items.size(), d_utw, usplit, W.max_rows,
                                           d_upart, uslots, E->scale, E->stream);
            if (e != hipSuccess) rc = set_error(BB_E_DEVICE, hipGetErrorString(e));
        }
        std::vector<double> part((size_t)usplit * uslots);
        if (!rc && hipMemcpyAsync(part.data(), d_upart, sizeof(double) * part.size(),
                                  hipMemcpyDeviceToHost, E->stream) != hipSuccess)
            rc = set_error(BB_E_DEVICE, "triangle-count copy failed");
        if (!rc && hipStreamSynchronize(E->stream) != hipSuccess)
            rc = set_error(BB_E_DEVICE, "triangle-count kernels failed");
        for (void *p : tmp) (void)hipFree(p);
        if (rc) return rc;
        E->num_host.assign(std::max(rcount, 1), 0.0);
        for (int rp = 0; rp < rcount; ++rp) {
            double v = 0.0;
            for (int s = 0; s < usplit; ++s) v += part[(size_t)s * uslots + upos[rp]];
            E->num_host[rp] = v;
        }
        if ((rc = dev_alloc(&E->d_num, E->num_host.size(), E->owned))) return rc;
        if ((rc = upload(E->d_num, E->num_host.data(), sizeof(double) * E->num_host.size(),
                         E->stream)))
            return rc;
    }

    // --- Field layout: per plane the contiguous range of rings in use ------
    E->rlo.assign(Nmax, -1);
    E->rhi.assign(Nmax, -2);
    E->base.assign(Nmax, 0);
    for (int p = 0; p < Nmax; ++p)
        for (int r = 0; r < Nr; ++r)
            if (d.field_used[(size_t)p * Nr + r]) {
                if (E->rlo[p] < 0) E->rlo[p] = r;
                E->rhi[p] = r;
            }
    E->nfields = 0;
    std::vector<FieldTask> tasks;
    for (int p = 0; p < Nmax; ++p) {
        E->base[p] = E->nfields;
        if (E->rlo[p] < 0) continue;
        for (int r = E->rlo[p]; r <= E->rhi[p]; ++r) {
            tasks.push_back(FieldTask{p, r, E->nfields, rowband[r], d.nlo[r], d.nhi[r]});
            E->max_rows = std::max(E->max_rows, std::min(NG, 2 * rowband[r] + 1));
            E->nfields++;
        }
    }
    E->ntasks = (int)tasks.size();

    // --- Contraction work lists ------------------------------------------------
    // Pair blocks per plane group; groups with the same plane set
    // {p1, |p2|, |p3|} are adjacent so that work items share staged rings.
    std::map<std::pair<int, int>, std::vector<CItem>> groups;
    for (size_t c = 0; c < d.contr.size(); ++c) {
        const Contraction &k = d.contr[c];
        groups[{k.p1, k.p2}].push_back({k.r1, k.r2, k.r3, (int)c});
    }
    std::map<std::array<int, 3>, std::vector<std::pair<int, int>>> by_set;
    for (auto &kv : groups) {
        const int p1 = kv.first.first, p2 = kv.first.second, p3 = -p1 - p2;
        std::array<int, 3> key{p1, std::abs(p2), std::abs(p3)};
        std::sort(key.begin(), key.end());
        by_set[key].push_back(kv.first);
    }
    std::vector<PairBlock> blocks;
    for (auto &kv : by_set)
        for (auto &gk : kv.second) {
            const int p1 = gk.first, p2 = gk.second, p3 = -p1 - p2;
            const int flags = (p2 < 0 ? TW_CONJ2 : 0) | (p3 < 0 ? TW_CONJ3 : 0);
            pair_blocks(groups[gk], p1, std::abs(p2), std::abs(p3), flags, blocks);
        }
    WorkLists W;
    E->pos.assign(std::max<size_t>(d.contr.size(), 1), 0);
    E->waves = kWaves;
    if (const char *ev = std::getenv("BARBIE_CONTRACT_WAVES")) {
        const int w = std::atoi(ev);
        if (w == 4 || w == 8) E->waves = w;
    }
    E->tww = contract_tww(E->cprec);
    if (const char *ev = std::getenv("BARBIE_CONTRACT_TWW")) {
        const int t = std::atoi(ev);
        if (E->cprec == CPrec::F64 && (t == kTWW || t == kTWWF64Narrow)) E->tww = t;
    }
    rc = pack_items(blocks, E->cprec == CPrec::BF16X3, E->tww, E->waves,
                    [&](int plane, int ring) { return E->fid(plane, ring); }, W, E->pos);
    E->wave_balance = W.tw_slots ? (double)W.tw_used / (double)W.tw_slots : 1.0;
    if (!rc) rc = check_lists(W, E->pos, d.contr.size());
    if (rc) return rc;
    E->nitems = (int)W.items.size();
    E->max_fields = std::max(W.max_rows, 1);
    for (const WorkItem &it : W.items) E->staged_rows += it.nrows;
    E->tile_slots = std::max<long long>(W.tile_slots, 1);
    E->nsplit = E->opt.x_splits > 0 ? E->opt.x_splits : auto_splits(E->nitems);
    E->chunk = contract_chunk(E->cprec);

    // --- Binning records, sorted by bin --------------------------------------
    std::vector<BinRec> recs(d.contr.size());
    std::vector<int32_t> seg(d.nbins + 1, 0);
    for (const Contraction &k : d.contr) seg[k.bin + 1]++;
    for (int b = 0; b < d.nbins; ++b) seg[b + 1] += seg[b];
    {
        std::vector<int32_t> at(seg.begin(), seg.end() - 1);
        for (size_t c = 0; c < d.contr.size(); ++c) {
            const Contraction &k = d.contr[c];
            BinRec R;
            R.p1 = k.p1; R.p2 = k.p2; R.p3 = k.p3;
            R.r1 = k.r1; R.r2 = k.r2; R.r3 = k.r3;
            R.rp = k.rp;
            R.pos = E->pos[c];
            R.f1 = E->fid(k.p1, k.r1);
            R.f2 = E->fid(std::abs(k.p2), k.r2);
            R.f3 = E->fid(std::abs(k.p3), k.r3);
            recs[at[k.bin]++] = R;
        }
    }
    E->bp.nbins = d.nbins;
    E->bp.n_mtbin = d.g.n_mtbin;
    E->bp.NN = NN;
    E->bp.nsplit = E->nsplit;
    E->bp.split_stride = E->tile_slots;
    E->bp.Cunits = 2. * M_PI / (NN * d.g.LL);
    E->bp.ff = d.g.ff;
    E->bp.beta = d.g.beta;

    // --- Device buffers --------------------------------------------------------
    if ((rc = dev_alloc(&E->d_planes, (size_t)Nmax * E->npts, E->owned))) return rc;
    const size_t nf = (size_t)std::max(E->nfields, 1);
    if (f32) {
        float2 *c;
        if ((rc = dev_alloc(&c, nf * E->npts, E->owned))) return rc;
        E->d_cache = c;
        E->rowbuf_fields = std::max(1, std::min(E->ntasks, 128));
        if (E->cprec == CPrec::BF16X3) {
            float2 *sp;   // 8 bytes per point, like the FP32 values
            if ((rc = dev_alloc(&sp, nf * E->npts, E->owned))) return rc;
            E->d_split = sp;
        }
        if ((rc = dev_alloc(&E->d_rowbuf, (size_t)E->rowbuf_fields * E->npts, E->owned))) return rc;
    } else {
        double2 *c;
        if ((rc = dev_alloc(&c, nf * E->npts, E->owned))) return rc;
        E->d_cache = c;
    }
    if ((rc = dev_alloc(&E->d_tasks, tasks.size(), E->owned))) return rc;
    if ((rc = dev_alloc(&E->d_items, W.items.size(), E->owned))) return rc;
    if ((rc = dev_alloc(&E->d_tws, W.tws.size(), E->owned))) return rc;
    if ((rc = dev_alloc(&E->d_part, (size_t)E->nsplit * E->tile_slots, E->owned))) return rc;
    if ((rc = dev_alloc(&E->d_recs, recs.size(), E->owned))) return rc;
    if ((rc = dev_alloc(&E->d_seg, seg.size(), E->owned))) return rc;
    if ((rc = dev_alloc(&E->d_nz, nf, E->owned))) return rc;
    if ((rc = dev_alloc(&E->d_avg, Nr, E->owned))) return rc;
    if ((rc = dev_alloc(&E->d_bins, (size_t)3 * d.nbins * BB_NPOLES, E->owned))) return rc;
    if (!tasks.empty() && (rc = upload(E->d_tasks, tasks.data(), sizeof(FieldTask) * tasks.size(), E->stream))) return rc;
    if (!W.items.empty() && (rc = upload(E->d_items, W.items.data(), sizeof(WorkItem) * W.items.size(), E->stream))) return rc;
    if (!W.tws.empty() && (rc = upload(E->d_tws, W.tws.data(), sizeof(TWin) * W.tws.size(), E->stream))) return rc;
    if (!recs.empty() && (rc = upload(E->d_recs, recs.data(), sizeof(BinRec) * recs.size(), E->stream))) return rc;
    if ((rc = upload(E->d_seg, seg.data(), sizeof(int32_t) * seg.size(), E->stream))) return rc;
    if ((rc = upload(E->d_avg, d.avg.data(), sizeof(double) * Nr, E->stream))) return rc;
    HIPCHK(hipMemsetAsync(E->d_part, 0, sizeof(double) * E->nsplit * E->tile_slots, E->stream));
    HIPCHK(hipStreamSynchronize(E->stream));
    return BB_OK;
}

int estimate_device(bb_estimator *E, const double *d_field_in, double *d_bs, double *d_norm,
                    double *d_biana)
{
    const PlanData &d = E->plan->d;
    const bool f32 = (E->cprec != CPrec::F64);
    hipStream_t s = E->stream;
    HIPCHK(hipEventRecord(E->ev[0], s));
    const int nf = std::max(E->nfields, 1);
    hipLaunchKernelGGL(k_fill_int, dim3((nf + 255) / 256), dim3(256), 0, s, E->d_nz, nf, 0);
    HIPCHK(hipGetLastError());
    HIPCHK(launch_gather((const double2 *)d_field_in, E->d_planes, E->NN, E->NG, E->kmax,
                         d.g.Nmax, s));
    HIPCHK(hipEventRecord(E->ev[1], s));
    int rc = run_fields(E, E->d_planes, E->d_tasks, E->ntasks, E->max_rows, E->d_cache, f32, E->d_nz);
    if (rc) return rc;
    HIPCHK(hipEventRecord(E->ev[2], s));
    HIPCHK(launch_contract(E->cprec, E->waves, E->tww, E->d_cache, E->d_split, E->npts, E->d_items, E->nitems, E->d_tws,
                           E->nsplit, E->max_fields, E->d_part, E->tile_slots, E->scale, s));
    HIPCHK(hipEventRecord(E->ev[3], s));
    HIPCHK(launch_binning(E->d_seg, E->d_recs, E->d_part, E->d_num, E->d_nz, E->d_avg, E->bp,
                          d_bs, d_norm, d_biana, s));
    HIPCHK(hipEventRecord(E->ev[4], s));
    HIPCHK(hipStreamSynchronize(s));
    for (int k = 0; k < 4; ++k) HIPCHK(hipEventElapsedTime(&E->ms[k], E->ev[k], E->ev[k + 1]));
    return BB_OK;
}

size_t cache_fields(const PlanData &d)
{
    size_t nf = 0;
    for (int pl = 0; pl < d.g.Nmax; ++pl) {
        int lo = -1, hi = -2;
        for (int r = 0; r < d.g.Nr; ++r)
            if (d.field_used[(size_t)pl * d.g.Nr + r]) {
                if (lo < 0) lo = r;
                hi = r;
            }
        if (lo >= 0) nf += (size_t)(hi - lo + 1);
    }
    return nf;
}

}  // namespace

extern "C" {

int bb_estimator_footprint(const bb_plan *p, bb_precision prec, size_t *bytes)
{
    if (!p || !bytes) return set_error(BB_E_INVALID, "null argument");
    const PlanData &d = p->d;
    // Upper bound: full-size grid.
    const size_t NN = d.g.NN, NN2 = NN * NN, Nr = d.g.Nr, Nmax = d.g.Nmax;
    const size_t nf = cache_fields(d);
    size_t b = NN * NN * (NN / 2 + 1) * 16 + Nmax * NN2 * 16;
    b += nf * NN2 * (prec == BB_FP64 ? 16 : 8);
    if (prec == BB_BF16X3) b += nf * NN2 * 8;   // bf16 hi/lo operand copy
    if (prec != BB_FP64) b += std::min<size_t>(nf, 128) * NN2 * 16;
    b += d.contr.size() * (sizeof(BinRec) + 8 * 3 * 2);
    b += Nr * NN2 * 16;  // transient unit fields (triangle counts)
    *bytes = b;
    return BB_OK;
}

int bb_estimator_create_opts(const bb_plan *p, const bb_options *opt, bb_estimator **out)
{
    if (!p || !opt || !out) return set_error(BB_E_INVALID, "null argument");
    *out = nullptr;
    CPrec cp;
    switch (opt->precision) {
    case BB_FP64: cp = CPrec::F64; break;
    case BB_FP32: cp = CPrec::F32; break;
    case BB_BF16X3: cp = CPrec::BF16X3; break;
    default: return set_error(BB_E_INVALID, "unknown precision");
    }
    if (opt->x_splits < 0 || opt->x_splits > 1024)
        return set_error(BB_E_INVALID, "x_splits must be in [0, 1024]");
    bb_estimator *E = new (std::nothrow) bb_estimator;
    if (!E) return set_error(BB_E_NOMEM, "out of host memory");
    E->plan = p;
    E->opt = *opt;
    E->cprec = cp;
    int rc;
    try {
        rc = estimator_init(E);
    } catch (const std::bad_alloc &) {
        rc = set_error(BB_E_NOMEM, "out of host memory");
    }
    if (rc) {
        delete E;
        return rc;
    }
    *out = E;
    return BB_OK;
}

int bb_estimator_create(const bb_plan *p, bb_precision prec, bb_estimator **out)
{
    bb_options o;
    std::memset(&o, 0, sizeof o);
    o.precision = prec;
    return bb_estimator_create_opts(p, &o, out);
}

void bb_estimator_destroy(bb_estimator *e) { delete e; }

int bb_estimator_grid(const bb_estimator *e, int *NG)
{
    if (!e || !NG) return set_error(BB_E_INVALID, "null argument");
    *NG = e->NG;
    return BB_OK;
}

int bb_estimate_device(bb_estimator *e, const double *d_field, double *d_bs, double *d_norm,
                       double *d_biana)
{
    if (!e || !d_field) return set_error(BB_E_INVALID, "null argument");
    return estimate_device(e, d_field, d_bs, d_norm, d_biana);
}

int bb_estimate(bb_estimator *e, const double *field, double *bs, double *norm, double *biana)
{
    if (!e || !field) return set_error(BB_E_INVALID, "null argument");
    const PlanData &d = e->plan->d;
    const size_t NN = d.g.NN;
    const size_t bytes = NN * NN * (NN + 2) * sizeof(double);
    if (!e->d_field) {
        int rc = dev_alloc(&e->d_field, NN * NN * (NN / 2 + 1), e->owned);
        if (rc) return rc;
    }
    HIPCHK(hipMemcpyAsync(e->d_field, field, bytes, hipMemcpyHostToDevice, e->stream));
    const size_t nb = (size_t)d.nbins * BB_NPOLES;
    double *B = e->d_bins;
    int rc = estimate_device(e, (const double *)e->d_field, B, B + nb, B + 2 * nb);
    if (rc) return rc;
    if (bs) HIPCHK(hipMemcpy(bs, B, nb * sizeof(double), hipMemcpyDeviceToHost));
    if (norm) HIPCHK(hipMemcpy(norm, B + nb, nb * sizeof(double), hipMemcpyDeviceToHost));
    if (biana) HIPCHK(hipMemcpy(biana, B + 2 * nb, nb * sizeof(double), hipMemcpyDeviceToHost));
    return BB_OK;
}

int bb_estimator_contraction_sums(bb_estimator *e, double *S)
{
    if (!e || !S) return set_error(BB_E_INVALID, "null argument");
    std::vector<double> part((size_t)e->nsplit * e->tile_slots);
    HIPCHK(hipMemcpy(part.data(), e->d_part, sizeof(double) * part.size(), hipMemcpyDeviceToHost));
    const size_t n = e->plan->d.contr.size();
    for (size_t c = 0; c < n; ++c) {
        double v = 0.0;
        for (int s = 0; s < e->nsplit; ++s) v += part[(size_t)s * e->tile_slots + e->pos[c]];
        S[c] = v;
    }
    return BB_OK;
}

int bb_estimator_triangle_counts(bb_estimator *e, double *num)
{
    if (!e || !num) return set_error(BB_E_INVALID, "null argument");
    const size_t n = e->plan->d.tri.size() / 3;
    for (size_t k = 0; k < n; ++k) num[k] = e->num_host[k];
    return BB_OK;
}

int bb_estimator_field(bb_estimator *e, int plane, int ring, double *out)
{
    if (!e || !out) return set_error(BB_E_INVALID, "null argument");
    const int f = e->fid(plane, ring);
    if (f < 0) return set_error(BB_E_INVALID, "field not held in the cache (unused by the plan)");
    const size_t n = (size_t)e->npts;
    if (e->cprec == CPrec::F64) {
        HIPCHK(hipMemcpy(out, (const double2 *)e->d_cache + (size_t)f * n, n * sizeof(double2),
                         hipMemcpyDeviceToHost));
    } else {
        std::vector<float2> tmp(n);
        HIPCHK(hipMemcpy(tmp.data(), (const float2 *)e->d_cache + (size_t)f * n,
                         n * sizeof(float2), hipMemcpyDeviceToHost));
        for (size_t k = 0; k < n; ++k) {
            out[2 * k] = tmp[k].x;
            out[2 * k + 1] = tmp[k].y;
        }
    }
    return BB_OK;
}

int bb_estimator_worklist(const bb_estimator *e, long long *o)
{
    if (!e || !o) return set_error(BB_E_INVALID, "null argument");
    o[0] = e->nitems;
    o[1] = e->nsplit;
    o[2] = e->tile_slots / 256;
    o[3] = e->max_fields;
    o[4] = e->staged_rows;
    o[5] = e->staged_rows * e->npts * (e->cprec == CPrec::F64 ? 16 : 8);
    o[6] = e->NG;
    o[7] = e->chunk;
    o[8] = e->waves;
    o[9] = (long long)std::llround(e->wave_balance * 1000.0);
    return BB_OK;
}

int bb_estimator_timings(bb_estimator *e, double *ms4)
{
    if (!e || !ms4) return set_error(BB_E_INVALID, "null argument");
    for (int k = 0; k < 4; ++k) ms4[k] = e->ms[k];
    return BB_OK;
}

}  // extern "C"
