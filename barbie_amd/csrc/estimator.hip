// Estimator: device memory, work lists and the per-field pipeline
// (plane gather -> plane-ring cache -> contraction -> binning).
// Replaces barbie.c:150-373 (cache build, ring triples, hot loop) for one
// field held in HBM.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
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

// Builds contraction tiles for the contractions of one plane group.
// items: (r1, r2, r3, cidx).  fid1(r1), fid2(r2), fid3(r3) map rings to cache
// slots; consecutive r3 of the group's third plane occupy consecutive slots.
struct Item { int r1, r2, r3, cidx; };

template <typename F1, typename F2, typename F3>
void build_tiles(std::vector<Item> &items, int conj2, int conj3, F1 fid1, F2 fid2, F3 fid3,
                 std::vector<Tile> &tiles, std::vector<int32_t> &outmap)
{
    // group by pair
    std::sort(items.begin(), items.end(), [](const Item &a, const Item &b) {
        if (a.r1 != b.r1) return a.r1 < b.r1;
        if (a.r2 != b.r2) return a.r2 < b.r2;
        return a.r3 < b.r3;
    });
    struct Pair { int r1, r2, lo, hi; size_t b, e; };
    std::vector<Pair> pairs;
    for (size_t k = 0; k < items.size();) {
        size_t e = k;
        while (e < items.size() && items[e].r1 == items[k].r1 && items[e].r2 == items[k].r2) ++e;
        pairs.push_back({items[k].r1, items[k].r2, items[k].r3, items[e - 1].r3, k, e});
        k = e;
    }
    std::sort(pairs.begin(), pairs.end(), [](const Pair &a, const Pair &b) {
        if (a.lo != b.lo) return a.lo < b.lo;
        if (a.hi != b.hi) return a.hi < b.hi;
        if (a.r1 != b.r1) return a.r1 < b.r1;
        return a.r2 < b.r2;
    });
    for (size_t pb = 0; pb < pairs.size(); pb += kTileMP) {
        size_t pe = std::min(pairs.size(), pb + kTileMP);
        int lo = pairs[pb].lo, hi = pairs[pb].hi;
        for (size_t q = pb; q < pe; ++q) { lo = std::min(lo, pairs[q].lo); hi = std::max(hi, pairs[q].hi); }
        for (int w = lo; w <= hi; w += kTileNT) {
            Tile t;
            std::memset(&t, 0, sizeof t);
            int nt = std::min(kTileNT, hi - w + 1);
            bool any = false;
            t.out = (int32_t)outmap.size();
            outmap.resize(outmap.size() + kTileMP * kTileNT, -1);
            for (int i = 0; i < kTileMP; ++i) {
                t.f1[i] = -1;
                t.f2[i] = -1;
            }
            for (size_t q = pb; q < pe; ++q) {
                int i = (int)(q - pb);
                t.f1[i] = fid1(pairs[q].r1);
                t.f2[i] = fid2(pairs[q].r2);
                for (size_t k = pairs[q].b; k < pairs[q].e; ++k) {
                    int j = items[k].r3 - w;
                    if (j >= 0 && j < nt) {
                        outmap[t.out + i * kTileNT + j] = items[k].cidx;
                        any = true;
                    }
                }
            }
            if (!any) {
                outmap.resize(t.out);
                continue;
            }
            t.f3 = fid3(w);
            t.nt = nt;
            t.conj2 = conj2;
            t.conj3 = conj3;
            tiles.push_back(t);
        }
    }
}

}  // namespace

struct bb_estimator {
    const bb_plan *plan = nullptr;
    bb_precision prec = BB_FP64;
    hipStream_t stream = nullptr;
    FftDesc fd{};
    int NN = 0;
    long long npts = 0;
    std::vector<int> rlo, rhi, base;  // per plane ring range and first slot
    int nfields = 0;
    std::vector<void *> owned;
    double2 *d_tw = nullptr;
    double2 *d_field = nullptr;
    double2 *d_planes = nullptr;
    void *d_cache = nullptr;
    double2 *d_rowbuf = nullptr;  // FP32 path scratch
    int rowbuf_fields = 0;
    FieldTask *d_tasks = nullptr;
    int ntasks = 0;
    int max_rows = 0;
    Tile *d_tiles = nullptr;
    int ntiles = 0;
    int32_t *d_outmap = nullptr;
    double *d_S = nullptr;
    BinRec *d_recs = nullptr;
    int32_t *d_seg = nullptr;
    double *d_num = nullptr;
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

// FFT pipeline for a list of tasks into a cache of the given element type.
int run_fields(bb_estimator *E, const double2 *planes, const FieldTask *d_tasks, int ntasks,
               int max_rows, void *cache, bool f32, int *nz)
{
    const long long NN2 = E->npts;
    if (!f32) {
        HIPCHK(launch_rowpass(planes, (double2 *)cache, NN2, d_tasks, ntasks, max_rows, E->fd,
                              E->d_tw, nz, E->stream));
        HIPCHK(launch_colpass_f64((double2 *)cache, NN2, (double2 *)cache, d_tasks, ntasks, E->fd,
                                  E->d_tw, E->stream));
        return BB_OK;
    }
    for (int t0 = 0; t0 < ntasks; t0 += E->rowbuf_fields) {
        int nb = std::min(E->rowbuf_fields, ntasks - t0);
        HIPCHK(launch_rowpass(planes, E->d_rowbuf, -NN2, d_tasks + t0, nb, max_rows, E->fd,
                              E->d_tw, nz, E->stream));
        HIPCHK(launch_colpass_f32(E->d_rowbuf, -NN2, (float2 *)cache, d_tasks + t0, nb, E->fd,
                                  E->d_tw, E->stream));
    }
    return BB_OK;
}

int estimator_init(bb_estimator *E)
{
    const PlanData &d = E->plan->d;
    const int NN = d.g.NN, Nmax = d.g.Nmax, Nr = d.g.Nr;
    E->NN = NN;
    E->npts = (long long)NN * NN;
    if (!make_fft_desc(NN, E->fd)) return set_error(BB_E_INVALID, "unsupported FFT length");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        (void)hipGetLastError();
        return set_error(BB_E_DEVICE, "no HIP device available");
    }
    HIPCHK(hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking));
    for (auto &e : E->ev) HIPCHK(hipEventCreate(&e));
    int rc;
    const bool f32 = (E->prec != BB_FP64);

    // Twiddles exp(+2 pi i t / N), FP64.
    std::vector<double2> tw(NN);
    for (int t = 0; t < NN; ++t) {
        long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)t / NN;
        tw[t] = make_double2((double)cosl(a), (double)sinl(a));
    }
    if ((rc = dev_alloc(&E->d_tw, NN, E->owned))) return rc;
    if ((rc = upload(E->d_tw, tw.data(), sizeof(double2) * NN, E->stream))) return rc;

    // Ring bounds and row bands.
    std::vector<int> rowband(Nr);
    for (int r = 0; r < Nr; ++r)
        rowband[r] = d.nhi[r] > 0 ? std::min(NN / 2, isqrt_floor(d.nhi[r] - 1)) : -1;

    // --- Triangle counts Num on the GPU (calCR, funcs.c:328-424) -----------
    {
        const int rcount = (int)(d.tri.size() / 3);
        std::vector<FieldTask> ut(Nr);
        int mr = 0;
        for (int r = 0; r < Nr; ++r) {
            ut[r] = FieldTask{-1, r, r, rowband[r], d.nlo[r], d.nhi[r]};
            mr = std::max(mr, std::min(NN, 2 * rowband[r] + 1));
        }
        std::vector<void *> tmp_owned;
        FieldTask *d_ut;
        double2 *d_unit;
        Tile *d_ut_tiles;
        int32_t *d_ut_out;
        std::vector<Item> items;
        for (int rp = 0; rp < rcount; ++rp)
            items.push_back({d.tri[3 * rp], d.tri[3 * rp + 1], d.tri[3 * rp + 2], rp});
        std::vector<Tile> tiles;
        std::vector<int32_t> outmap;
        auto id = [](int r) { return r; };
        build_tiles(items, 0, 0, id, id, id, tiles, outmap);
        if ((rc = dev_alloc(&E->d_num, std::max(rcount, 1), E->owned))) return rc;
        rc = dev_alloc(&d_ut, Nr, tmp_owned);
        if (!rc) rc = dev_alloc(&d_unit, (size_t)Nr * E->npts, tmp_owned);
        if (!rc) rc = dev_alloc(&d_ut_tiles, tiles.size(), tmp_owned);
        if (!rc) rc = dev_alloc(&d_ut_out, outmap.size(), tmp_owned);
        if (!rc) rc = upload(d_ut, ut.data(), sizeof(FieldTask) * Nr, E->stream);
        if (!rc) rc = upload(d_ut_tiles, tiles.data(), sizeof(Tile) * tiles.size(), E->stream);
        if (!rc) rc = upload(d_ut_out, outmap.data(), sizeof(int32_t) * outmap.size(), E->stream);
        if (!rc) rc = run_fields(E, nullptr, d_ut, Nr, mr, d_unit, false, nullptr);
        if (!rc) {
            hipError_t e = launch_contract_f64(d_unit, E->npts, d_ut_tiles, (int)tiles.size(),
                                               d_ut_out, E->d_num, E->stream);
            if (e != hipSuccess) rc = set_error(BB_E_DEVICE, hipGetErrorString(e));
        }
        if (!rc && hipStreamSynchronize(E->stream) != hipSuccess)
            rc = set_error(BB_E_DEVICE, "triangle-count kernels failed");
        for (void *p : tmp_owned) (void)hipFree(p);
        if (rc) return rc;
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
            E->max_rows = std::max(E->max_rows, std::min(NN, 2 * rowband[r] + 1));
            E->nfields++;
        }
    }
    E->ntasks = (int)tasks.size();

    // --- Contraction tiles per plane group ----------------------------------
    std::map<std::pair<int, int>, std::vector<Item>> groups;
    for (size_t c = 0; c < d.contr.size(); ++c) {
        const Contraction &k = d.contr[c];
        groups[{k.p1, k.p2}].push_back({k.r1, k.r2, k.r3, (int)c});
    }
    std::vector<Tile> tiles;
    std::vector<int32_t> outmap;
    for (auto &kv : groups) {
        const int p1 = kv.first.first, p2 = kv.first.second, p3 = -p1 - p2;
        const int a2 = std::abs(p2), a3 = std::abs(p3);
        auto f1 = [&](int r) { return E->fid(p1, r); };
        auto f2 = [&](int r) { return E->fid(a2, r); };
        auto f3 = [&](int r) { return E->fid(a3, r); };
        build_tiles(kv.second, p2 < 0, p3 < 0, f1, f2, f3, tiles, outmap);
    }
    E->ntiles = (int)tiles.size();

    // --- Binning records, sorted by bin --------------------------------------
    std::vector<BinRec> recs(d.contr.size());
    std::vector<int32_t> seg(d.nbins + 1, 0);
    for (const Contraction &k : d.contr) seg[k.bin + 1]++;
    for (int b = 0; b < d.nbins; ++b) seg[b + 1] += seg[b];
    {
        std::vector<int32_t> pos(seg.begin(), seg.end() - 1);
        for (size_t c = 0; c < d.contr.size(); ++c) {
            const Contraction &k = d.contr[c];
            BinRec R;
            R.p1 = k.p1; R.p2 = k.p2; R.p3 = k.p3;
            R.r1 = k.r1; R.r2 = k.r2; R.r3 = k.r3;
            R.rp = k.rp;
            R.cidx = (int32_t)c;
            R.f1 = E->fid(k.p1, k.r1);
            R.f2 = E->fid(std::abs(k.p2), k.r2);
            R.f3 = E->fid(std::abs(k.p3), k.r3);
            recs[pos[k.bin]++] = R;
        }
    }
    E->bp.nbins = d.nbins;
    E->bp.n_mtbin = d.g.n_mtbin;
    E->bp.NN = NN;
    E->bp.LL_k1 = d.g.LL_k1;
    E->bp.BW_k1 = d.g.BW_k1;
    E->bp.Cunits = 2. * M_PI / (NN * d.g.LL);
    E->bp.ff = d.g.ff;
    E->bp.beta = d.g.beta;

    // --- Device buffers --------------------------------------------------------
    const size_t H = NN / 2 + 1;
    if ((rc = dev_alloc(&E->d_field, (size_t)NN * NN * H, E->owned))) return rc;
    if ((rc = dev_alloc(&E->d_planes, (size_t)Nmax * E->npts, E->owned))) return rc;
    if (f32) {
        float2 *c;
        if ((rc = dev_alloc(&c, (size_t)std::max(E->nfields, 1) * E->npts, E->owned))) return rc;
        E->d_cache = c;
        E->rowbuf_fields = std::max(1, std::min(E->ntasks, 128));
        if ((rc = dev_alloc(&E->d_rowbuf, (size_t)E->rowbuf_fields * E->npts, E->owned))) return rc;
    } else {
        double2 *c;
        if ((rc = dev_alloc(&c, (size_t)std::max(E->nfields, 1) * E->npts, E->owned))) return rc;
        E->d_cache = c;
    }
    if ((rc = dev_alloc(&E->d_tasks, tasks.size(), E->owned))) return rc;
    if ((rc = dev_alloc(&E->d_tiles, tiles.size(), E->owned))) return rc;
    if ((rc = dev_alloc(&E->d_outmap, outmap.size(), E->owned))) return rc;
    if ((rc = dev_alloc(&E->d_S, d.contr.size(), E->owned))) return rc;
    if ((rc = dev_alloc(&E->d_recs, recs.size(), E->owned))) return rc;
    if ((rc = dev_alloc(&E->d_seg, seg.size(), E->owned))) return rc;
    if ((rc = dev_alloc(&E->d_nz, std::max(E->nfields, 1), E->owned))) return rc;
    if ((rc = dev_alloc(&E->d_avg, Nr, E->owned))) return rc;
    if ((rc = dev_alloc(&E->d_bins, (size_t)3 * d.nbins * BB_NPOLES, E->owned))) return rc;
    if ((rc = upload(E->d_tasks, tasks.data(), sizeof(FieldTask) * tasks.size(), E->stream))) return rc;
    if ((rc = upload(E->d_tiles, tiles.data(), sizeof(Tile) * tiles.size(), E->stream))) return rc;
    if ((rc = upload(E->d_outmap, outmap.data(), sizeof(int32_t) * outmap.size(), E->stream))) return rc;
    if ((rc = upload(E->d_recs, recs.data(), sizeof(BinRec) * recs.size(), E->stream))) return rc;
    if ((rc = upload(E->d_seg, seg.data(), sizeof(int32_t) * seg.size(), E->stream))) return rc;
    if ((rc = upload(E->d_avg, d.avg.data(), sizeof(double) * Nr, E->stream))) return rc;
    HIPCHK(hipMemsetAsync(E->d_S, 0, sizeof(double) * std::max<size_t>(d.contr.size(), 1), E->stream));
    HIPCHK(hipStreamSynchronize(E->stream));
    return BB_OK;
}

int estimate_device(bb_estimator *E, const double *d_field_in, double *d_bs, double *d_norm,
                    double *d_biana)
{
    const PlanData &d = E->plan->d;
    const bool f32 = (E->prec != BB_FP64);
    hipStream_t s = E->stream;
    HIPCHK(hipEventRecord(E->ev[0], s));
    const int nf = std::max(E->nfields, 1);
    hipLaunchKernelGGL(k_fill_int, dim3((nf + 255) / 256), dim3(256), 0, s, E->d_nz, nf, 0);
    HIPCHK(hipGetLastError());
    HIPCHK(launch_gather((const double2 *)d_field_in, E->d_planes, E->NN, d.g.Nmax, s));
    HIPCHK(hipEventRecord(E->ev[1], s));
    int rc = run_fields(E, E->d_planes, E->d_tasks, E->ntasks, E->max_rows, E->d_cache, f32, E->d_nz);
    if (rc) return rc;
    HIPCHK(hipEventRecord(E->ev[2], s));
    if (f32)
        HIPCHK(launch_contract_f32((const float2 *)E->d_cache, E->npts, E->d_tiles, E->ntiles,
                                   E->d_outmap, E->d_S, s));
    else
        HIPCHK(launch_contract_f64((const double2 *)E->d_cache, E->npts, E->d_tiles, E->ntiles,
                                   E->d_outmap, E->d_S, s));
    HIPCHK(hipEventRecord(E->ev[3], s));
    HIPCHK(launch_binning(E->d_seg, E->d_recs, E->d_S, E->d_num, E->d_nz, E->d_avg, E->bp, d_bs,
                          d_norm, d_biana, s));
    HIPCHK(hipEventRecord(E->ev[4], s));
    HIPCHK(hipStreamSynchronize(s));
    for (int k = 0; k < 4; ++k) HIPCHK(hipEventElapsedTime(&E->ms[k], E->ev[k], E->ev[k + 1]));
    return BB_OK;
}

}  // namespace

extern "C" {

int bb_estimator_footprint(const bb_plan *p, bb_precision prec, size_t *bytes)
{
    if (!p || !bytes) return set_error(BB_E_INVALID, "null argument");
    const PlanData &d = p->d;
    const size_t NN = d.g.NN, NN2 = NN * NN, Nr = d.g.Nr, Nmax = d.g.Nmax;
    size_t nf = 0;
    for (size_t pl = 0; pl < Nmax; ++pl) {
        int lo = -1, hi = -2;
        for (size_t r = 0; r < Nr; ++r)
            if (d.field_used[pl * Nr + r]) { if (lo < 0) lo = (int)r; hi = (int)r; }
        if (lo >= 0) nf += (size_t)(hi - lo + 1);
    }
    size_t b = NN * NN * (NN / 2 + 1) * 16 + Nmax * NN2 * 16;
    b += nf * NN2 * (prec == BB_FP64 ? 16 : 8);
    if (prec != BB_FP64) b += std::min<size_t>(nf, 128) * NN2 * 16;
    b += d.contr.size() * (8 + sizeof(BinRec) + 4 * kTileNT);
    b += Nr * NN2 * 16;  // transient unit fields (triangle counts)
    *bytes = b;
    return BB_OK;
}

int bb_estimator_create(const bb_plan *p, bb_precision prec, bb_estimator **out)
{
    if (!p || !out) return set_error(BB_E_INVALID, "null argument");
    *out = nullptr;
    if (prec != BB_FP64 && prec != BB_FP32)
        return set_error(BB_E_INVALID, prec == BB_BF16X3 ? "BF16x3 contraction is not built yet"
                                                         : "unknown precision");
    bb_estimator *E = new (std::nothrow) bb_estimator;
    if (!E) return set_error(BB_E_NOMEM, "out of host memory");
    E->plan = p;
    E->prec = prec;
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

void bb_estimator_destroy(bb_estimator *e) { delete e; }

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
    HIPCHK(hipMemcpy(S, e->d_S, e->plan->d.contr.size() * sizeof(double), hipMemcpyDeviceToHost));
    return BB_OK;
}

int bb_estimator_triangle_counts(bb_estimator *e, double *num)
{
    if (!e || !num) return set_error(BB_E_INVALID, "null argument");
    HIPCHK(hipMemcpy(num, e->d_num, (e->plan->d.tri.size() / 3) * sizeof(double),
                     hipMemcpyDeviceToHost));
    return BB_OK;
}

int bb_estimator_field(bb_estimator *e, int plane, int ring, double *out)
{
    if (!e || !out) return set_error(BB_E_INVALID, "null argument");
    const int f = e->fid(plane, ring);
    if (f < 0) return set_error(BB_E_INVALID, "field not held in the cache (unused by the plan)");
    const size_t n = (size_t)e->npts;
    if (e->prec == BB_FP64) {
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

int bb_estimator_timings(bb_estimator *e, double *ms4)
{
    if (!e || !ms4) return set_error(BB_E_INVALID, "null argument");
    for (int k = 0; k < 4; ++k) ms4[k] = e->ms[k];
    return BB_OK;
}

}  // extern "C"
