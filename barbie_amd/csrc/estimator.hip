// Estimator: grid levels, device memory, work lists and the per-field
// pipeline (plane gather -> plane-ring caches per level -> contraction ->
// binning).  Replaces barbie.c:150-373 (cache build, ring triples, hot loop)
// for one field held in HBM.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <array>
#include <chrono>
#include <map>
#include <new>
#include <string>
#include <thread>
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

bool make_fft_desc(int N, FftDesc &fd)
{
    fd.N = N;
    fd.nst = 0;
    int n = N;
    while (n % 4 == 0 && fd.nst < kMaxStages) { fd.radix[fd.nst++] = 4; n /= 4; }
    while (n % 2 == 0 && fd.nst < kMaxStages) { fd.radix[fd.nst++] = 2; n /= 2; }
    while (n % 3 == 0 && fd.nst < kMaxStages) { fd.radix[fd.nst++] = 3; n /= 3; }
    while (n % 5 == 0 && fd.nst < kMaxStages) { fd.radix[fd.nst++] = 5; n /= 5; }
    if (n != 1 || N >= 4096) return false;
    fd.mN = udiv_magic((unsigned)N);
    int p = 1;
    for (int s = 0; s < fd.nst; ++s) {
        fd.mNR[s] = udiv_magic((unsigned)(N / fd.radix[s]));
        fd.mP[s] = udiv_magic((unsigned)p);
        p *= fd.radix[s];
    }
    return true;
}

// The ring pass's FFT plan: the radix-2..5 stages, adjacent pairs merged
// into one composite stage in registers where the product is 6, 8, 9, 10,
// 12, 15 or 16, in the order largest odd factor first, then ascending
// (160 = 10 x 16, 108 = 9 x 12, 80 = 5 x 16, 54 = 9 x 6, 384 = 3 x 8 x 16).
// A Stockham stage with p = 1 writes its R outputs at a stride of R
// elements across the lanes; a power of two R puts many lanes on one LDS
// bank there, so the powers of two go last, where p is large and the
// writes are contiguous runs.
bool make_ring_fft_desc(int N, FftDesc &fd)
{
    FftDesc base;
    if (!make_fft_desc(N, base)) return false;
    fd = base;
    fd.nst = 0;
    for (int s = 0; s < base.nst;) {
        const int m = s + 1 < base.nst ? base.radix[s] * base.radix[s + 1] : 0;
        const bool merge = m == 6 || m == 8 || m == 9 || m == 10 || m == 12 || m == 15 || m == 16;
        fd.radix[fd.nst++] = merge ? m : base.radix[s];
        s += merge ? 2 : 1;
    }
    auto odd = [](int r) {
        while (r % 2 == 0) r /= 2;
        return r;
    };
    std::stable_sort(fd.radix, fd.radix + fd.nst, [&](int x, int y) {
        return odd(x) != odd(y) ? odd(x) > odd(y) : x < y;
    });
    int p = 1;
    for (int s = 0; s < fd.nst; ++s) {
        fd.mNR[s] = udiv_magic((unsigned)(N / fd.radix[s]));
        fd.mP[s] = udiv_magic((unsigned)p);
        p *= fd.radix[s];
    }
    return true;
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
struct CItem { int r1, r2, r3, cid, pa, pb; };

// A pair block: up to 16 (r1, r2) rows of one plane group whose r3 rings
// span at most kWinRings, covered by nwin windows starting at lo.
constexpr int kWinRings = 16 * 3;
constexpr bool kRepack = true;   // pack the partly filled blocks of a plane set together
struct PairBlock {
    int pc;                   // plane of D3 (|p3|)
    int flags;                // TW_CONJ2 | TW_CONJ3
    int pa[16], pb[16];       // per row: planes of D1, D2 (p1, |p2|)
    int r1[16], r2[16];
    int nrow;
    int lo, hi, nwin;
    std::vector<std::array<int, 3>> cs;   // (row, r3, cid)
};

// cs: contractions of one third plane pc and one conjugation pattern; each
// carries the planes (pa, pb) of its pair, so that the rows of one block may
// come from different plane pairs (they share the B operand, plane pc).
void pair_blocks(std::vector<CItem> &cs, int pc, int flags, std::vector<PairBlock> &out)
{
    std::sort(cs.begin(), cs.end(), [](const CItem &a, const CItem &b) {
        if (a.pa != b.pa) return a.pa < b.pa;
        if (a.pb != b.pb) return a.pb < b.pb;
        if (a.r1 != b.r1) return a.r1 < b.r1;
        if (a.r2 != b.r2) return a.r2 < b.r2;
        return a.r3 < b.r3;
    });
    struct Pair { int pa, pb, r1, r2, lo, hi; size_t b, e; };
    std::vector<Pair> pairs;
    for (size_t k = 0; k < cs.size();) {
        size_t e = k;
        while (e < cs.size() && cs[e].pa == cs[k].pa && cs[e].pb == cs[k].pb &&
               cs[e].r1 == cs[k].r1 && cs[e].r2 == cs[k].r2 && cs[e].r3 - cs[k].r3 < kWinRings)
            ++e;
        pairs.push_back({cs[k].pa, cs[k].pb, cs[k].r1, cs[k].r2, cs[k].r3, cs[e - 1].r3, k, e});
        k = e;
    }
    std::sort(pairs.begin(), pairs.end(), [](const Pair &a, const Pair &b) {
        if (a.lo != b.lo) return a.lo < b.lo;
        if (a.hi != b.hi) return a.hi < b.hi;
        if (a.pa != b.pa) return a.pa < b.pa;
        if (a.pb != b.pb) return a.pb < b.pb;
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
        B.pc = pc; B.flags = flags;
        B.nrow = (int)(e - q);
        B.lo = lo; B.hi = hi; B.nwin = (hi - lo) / 16 + 1;
        for (int m = 0; m < 16; ++m) { B.r1[m] = -1; B.r2[m] = -1; B.pa[m] = 0; B.pb[m] = 0; }
        for (size_t u = q; u < e; ++u) {
            const int m = (int)(u - q);
            B.pa[m] = pairs[u].pa;
            B.pb[m] = pairs[u].pb;
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
    std::vector<uint8_t> kind;    // per item: 0 the layout's precision, 1 FP32 MFMA (mixed layouts)
    std::vector<TWin> tws;
    std::vector<int> item_work;   // tile windows of the busiest wave of each item
    int ntw = 0;                  // output tiles
    int max_rows = 0;             // largest staged ring count of an item
    long long tw_used = 0, tw_slots = 0;   // wave balance: sum cnt, sum waves * max cnt
};

}  // namespace

namespace bb {

// One grid level: the NG x NG grid, the fields it holds (per plane the
// contiguous range of rings in use) and their FFT tasks.
struct Level {
    int NG = 0;
    long long npts = 0;
    long long stride = 0;              // points between fields: npts rounded up to 32
    double scale = 1.0;                // (NN / NG)^2
    long long off = 0;                 // first point of the level in the cache
    std::vector<int> rlo, rhi, base;   // per plane
    int nfields = 0;
    int max_rows = 0;                  // rows of the widest ring band
    std::vector<FieldTask> tasks;

    int fid(int plane, int ring) const
    {
        if (plane < 0 || plane >= (int)rlo.size() || rlo[plane] < 0) return -1;
        if (ring < rlo[plane] || ring > rhi[plane]) return -1;
        return base[plane] + ring - rlo[plane];
    }
};

struct Layout {
    CPrec prec = CPrec::F64;
    int waves = kWaves, tww = kTWW, chunk = 32;
    int NN = 0, NGtop = 0, kmax = 0;
    std::vector<Level> lv;
    std::vector<uint8_t> lvl;          // contraction -> level
    long long cache_pts = 0;           // points of all levels' fields
    std::vector<WorkItem> items;
    std::vector<TWin> tws;
    std::vector<WorkUnit> units;       // the layout's precision
    std::vector<WorkUnit> units_f;     // mixed BF16x3 layouts: items run by the FP32 kernel
    int mix = 0;                       // percent of the cost given to the FP32 kernel
    long long ntiles = 0, part_tiles = 0;
    std::vector<int32_t> pos;          // contraction -> first partial sum
    std::vector<int16_t> nsp;          // contraction -> partial sums
    std::vector<BinRec> recs;          // bin-sorted
    std::vector<int32_t> seg;
    std::vector<int32_t> sg, bin_sg;   // binning segments (record ranges) and each bin's first
    BinParams bp{};
    int max_rows = 0;
    long long staged_rows = 0, staged_bytes = 0, useful_pts = 0;
    long long total_cost = 0;          // tile windows of the busiest waves x chunks, all items
    double wave_balance = 1.0;
};

}  // namespace bb

namespace {

// LDS rows of a work item's staged rings.  Rows are read by the
// contraction as 16 lanes at a time, one ring per lane, and the BF16x3
// layout puts bytes 16 (r mod 16) .. +15 of every 256 of row r into bank
// group r mod 16 (dma_off in contract.hip): lanes that read distinct rings of
// the same group wait for each other.  The B operand reads 16 consecutive
// rows (one per group) and is conflict-free; the A rows (the D1 and the D2
// rings of a pair block) are not consecutive.  With split_b (BF16x3), the B
// rings take the first rows as contiguous ranges, and the pair rings the
// remaining rows, assigned greedily so that the distinct D1 rings of each
// block, and its distinct D2 rings, fall into distinct groups where
// possible (empty A rows read zeros in bank group nrows mod 16, which
// counts as a member of their block).  Otherwise every range is contiguous, in key order.
template <class KeyMap, class Key>
void assign_rows(const std::vector<PairBlock> &blocks, const std::vector<size_t> &item_blocks,
                 const KeyMap &keys, bool split_b, int nrows, std::map<Key, int> &row0,
                 std::map<std::pair<int, int>, int> &prow)
{
    int next = 0;
    if (!split_b) {
        for (auto &kv : keys) {
            row0[kv.first] = next - kv.second.lo;
            for (int r = kv.second.lo; r <= kv.second.hi; ++r) prow[{kv.first.plane, r}] = next + r - kv.second.lo;
            next += kv.second.hi - kv.second.lo + 1;
        }
        return;
    }
    std::vector<std::pair<int, int>> rings;   // pair rings in key order
    for (auto &kv : keys) {
        if (kv.first.split) {
            row0[kv.first] = next - kv.second.lo;
            next += kv.second.hi - kv.second.lo + 1;
        } else {
            for (int r = kv.second.lo; r <= kv.second.hi; ++r) rings.push_back({kv.first.plane, r});
        }
    }
    const int first = next, zero_group = nrows & 15;
    std::map<std::pair<int, int>, int> id;
    for (size_t k = 0; k < rings.size(); ++k) id[rings[k]] = (int)k;
    // Cliques: per block, the distinct D1 rings and the distinct D2 rings.
    std::vector<std::vector<int>> member(rings.size());   // ring -> cliques
    std::vector<std::array<int, 16>> count;               // clique -> rings per group
    for (size_t bi : item_blocks) {
        const PairBlock &B = blocks[bi];
        for (int side = 0; side < 2; ++side) {
            std::vector<int> ids;
            bool empty = false;
            for (int m = 0; m < 16; ++m) {
                if (B.r1[m] < 0) {
                    empty = true;
                    continue;
                }
                ids.push_back(side ? id.at({B.pb[m], B.r2[m]}) : id.at({B.pa[m], B.r1[m]}));
            }
            std::sort(ids.begin(), ids.end());
            ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
            const int q = (int)count.size();
            count.push_back({});
            count.back().fill(0);
            if (empty) count.back()[zero_group] = 1;
            for (int x : ids) member[x].push_back(q);
        }
    }
    std::array<std::vector<int>, 16> free_rows;
    for (int r = first + (int)rings.size() - 1; r >= first; --r) free_rows[r & 15].push_back(r);
    std::vector<int> order(rings.size());
    for (size_t k = 0; k < order.size(); ++k) order[k] = (int)k;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return member[a].size() > member[b].size(); });
    for (int x : order) {
        int best = -1;
        long long best_cost = 0;
        for (int g = 0; g < 16; ++g) {
            if (free_rows[g].empty()) continue;
            long long c = 0;
            for (int q : member[x]) c += count[q][g];
            if (best < 0 || c < best_cost ||
                (c == best_cost && free_rows[g].size() > free_rows[best].size())) {
                best = g;
                best_cost = c;
            }
        }
        prow[rings[x]] = free_rows[best].back();
        free_rows[best].pop_back();
        for (int q : member[x]) count[q][best]++;
    }
}

// Tile rows of a pair block's pairs in the BF16x3 kernel.  ds_read_b128
// serves the 16 rows of an A tile in phases that hold rows {0-3, 12-15} of
// one quarter of the chunk and rows {4-11} of the next (contract.hip), so a
// ring read by rows of both halves is two addresses in one bank group.
// Rows that share a D1 ring or a D2 ring (transitively) are placed in one
// half where they fit: such groups largest first, each into the half with
// more room (the other if only that one holds it; a group larger than both
// is split); empty rows take the slots left.  Returns row m -> tile row.
std::array<int, 16> tile_rows(const PairBlock &B)
{
    static const int half[2][8] = {{0, 1, 2, 3, 12, 13, 14, 15}, {4, 5, 6, 7, 8, 9, 10, 11}};
    int parent[16];
    for (int m = 0; m < 16; ++m) parent[m] = m;
    auto root = [&](int x) {
        while (parent[x] != x) x = parent[x] = parent[parent[x]];
        return x;
    };
    for (int a = 0; a < 16; ++a)
        for (int b = a + 1; b < 16; ++b)
            if (B.r1[a] >= 0 && B.r1[b] >= 0 &&
                ((B.pa[a] == B.pa[b] && B.r1[a] == B.r1[b]) ||
                 (B.pb[a] == B.pb[b] && B.r2[a] == B.r2[b])))
                parent[root(a)] = root(b);
    std::vector<std::vector<int>> groups;
    int gid[16];
    for (int m = 0; m < 16; ++m) gid[m] = -1;
    for (int m = 0; m < 16; ++m) {
        if (B.r1[m] < 0) continue;
        const int r = root(m);
        if (gid[r] < 0) {
            gid[r] = (int)groups.size();
            groups.emplace_back();
        }
        groups[gid[r]].push_back(m);
    }
    std::stable_sort(groups.begin(), groups.end(),
                     [](const std::vector<int> &x, const std::vector<int> &y) { return x.size() > y.size(); });
    std::array<int, 16> to;
    to.fill(-1);
    int used[2] = {0, 0};
    for (const std::vector<int> &g : groups) {
        const int n = (int)g.size();
        int h = used[0] <= used[1] ? 0 : 1;
        if (used[h] + n > 8 && used[1 - h] + n <= 8) h = 1 - h;
        for (int m : g) {
            if (used[h] == 8) h = 1 - h;
            to[m] = half[h][used[h]++];
        }
    }
    for (int m = 0; m < 16; ++m)
        if (to[m] < 0) {
            const int h = used[0] < 8 ? 0 : 1;
            to[m] = half[h][used[h]++];
        }
    return to;
}

// Packs pair blocks of one level (already grouped so that neighbours share
// planes) into work items.  split_b: stage the r3 rings as bf16 hi/lo
// operands (BF16x3) instead of raw values.  pos[cid] receives tile * 256 +
// row * 16 + column of each contraction.
int pack_items(std::vector<PairBlock> &blocks, const Level &L, bool split_b, int tww, int waves,
               int rowcap, WorkLists &W, std::vector<int32_t> &pos)
{
    struct Key { int plane, split; bool operator<(const Key &o) const { return plane != o.plane ? plane < o.plane : split < o.split; } };
    struct Range { int lo, hi; };
    size_t a = 0;
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
                add({B.pa[m], 0}, B.r1[m], B.r1[m]);
                add({B.pb[m], 0}, B.r2[m], B.r2[m]);
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
        // Lay out the item: the staged field ranges in key order (blk_*), and
        // the LDS row of every staged ring.
        WorkItem it;
        std::memset(&it, 0, sizeof it);
        it.off = L.off;
        it.npts = L.npts;
        it.fstride = L.stride;
        it.scale = L.scale;
        int nrows = 0;
        for (auto &kv : keys) {
            const int bi = it.nblk++;
            const int f0 = L.fid(kv.first.plane, kv.second.lo);
            if (f0 < 0 || L.fid(kv.first.plane, kv.second.hi) != f0 + kv.second.hi - kv.second.lo)
                return set_error(BB_E_INVALID, "internal: staged rings missing from the level");
            it.blk_slot[bi] = f0;
            it.blk_rows[bi] = (int16_t)(kv.second.hi - kv.second.lo + 1);
            it.blk_split[bi] = (int16_t)kv.first.split;
            nrows += it.blk_rows[bi];
        }
        it.nrows = (int16_t)nrows;
        std::vector<size_t> item_blocks;
        for (int w = 0; w < waves; ++w)
            item_blocks.insert(item_blocks.end(), wave_blocks[w].begin(), wave_blocks[w].end());
        std::map<Key, int> row0;             // B operand rings: row of ring r = row0 + r
        std::map<std::pair<int, int>, int> prow;   // pair rings (plane, ring) -> row
        assign_rows(blocks, item_blocks, keys, split_b, nrows, row0, prow);
        for (auto &kv : keys) {
            for (int r = kv.second.lo; r <= kv.second.hi; ++r) {
                const bool pair_ring = !(split_b && kv.first.split);
                const int row = pair_ring ? prow.at({kv.first.plane, r}) : row0.at(kv.first) + r;
                it.rslot[row] = L.fid(kv.first.plane, r) | (kv.first.split ? kRowSplit : 0);
            }
        }
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
        int work = 0;
        for (int w = 0; w < waves; ++w) {
            int j = 0;
            for (size_t bi : wave_blocks[w]) {
                const PairBlock &B = blocks[bi];
                const int rc = row0.at({B.pc, split_b ? 1 : 0});
                const int tile0 = W.ntw;
                std::array<int, 16> to;
                for (int m = 0; m < 16; ++m) to[m] = m;
                if (split_b) to = tile_rows(B);
                for (int v = 0; v < B.nwin; ++v) {
                    TWin &T = W.tws[it.tw0 + w * kTWW + j];
                    std::memset(&T, 0, sizeof T);
                    for (int m = 0; m < 16; ++m) {
                        T.a1[to[m]] = (int16_t)(B.r1[m] >= 0 ? prow.at({B.pa[m], B.r1[m]}) : -1);
                        T.a2[to[m]] = (int16_t)(B.r2[m] >= 0 ? prow.at({B.pb[m], B.r2[m]}) : 0);
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
                    pos[c[2]] = (tile0 + d / 16) * 256 + to[c[0]] * 16 + d % 16;
                }
            }
            it.cnt[w] = (int8_t)j;
            work = std::max(work, j);
            W.tw_used += j;
        }
        W.tw_slots += (long long)waves * work;
        W.item_work.push_back(work);
        W.items.push_back(it);
        W.kind.push_back(0);
        a = b;
    }
    return BB_OK;
}

// Workgroups of the launch: each item's chunks are split into x-ranges so
// that units carry similar work (or into x_splits ranges when given), and
// every tile gets one partial tile per x-range.  Rewrites the tile indices
// of the windows and of pos into partial-tile positions; units are ordered
// largest first.  The units of one launch: about kTargetUnits workgroups per
// batch of fields, so an estimator of F fields per batch splits each field's
// work into kTargetUnits / F units (a batch fills the chip by its fields, and
// every x-split costs a partial tile written by the contraction and read by
// the binning).  The split depends on the estimator's batch size, never on
// the size of the batch it runs: a field gets the same bits in every batch.
constexpr long long kTargetUnits = 2048;
long long target_units(int max_batch)
{
    return std::max<long long>(1, kTargetUnits / std::max(1, max_batch));
}
// Largest cost of one unit (tile windows of the busiest wave x chunks).  Long
// workgroups delay the kernels of other streams that wait for a free CU
// (the request server runs several estimates at once), so large estimates get
// more units rather than longer ones.
constexpr long long kMaxUnitCost = 1024;

int make_units(WorkLists &W, int chunk, int x_splits, std::vector<WorkUnit> &units,
               std::vector<int32_t> &pos, std::vector<int16_t> &nsp, long long &part_tiles,
               long long *total_cost, long long target = kTargetUnits)
{
    const size_t n = W.items.size();
    std::vector<int> nch(n), sp(n);
    long long total = 0;
    for (size_t i = 0; i < n; ++i) {
        nch[i] = (int)((W.items[i].npts + chunk - 1) / chunk);
        total += (long long)W.item_work[i] * nch[i];
    }
    if (total_cost) *total_cost = total;
    const long long unit_cost =
        std::max<long long>(1, std::min(total / std::max<long long>(1, target), kMaxUnitCost));
    for (size_t i = 0; i < n; ++i) {
        const long long cost = (long long)W.item_work[i] * nch[i];
        int s = x_splits > 0 ? x_splits : (int)std::min<long long>(256, (cost + unit_cost - 1) / unit_cost);
        sp[i] = std::max(1, std::min(s, nch[i]));
    }
    std::vector<int> tile_item(W.ntw, -1);
    for (size_t i = 0; i < n; ++i) {
        const WorkItem &it = W.items[i];
        for (int w = 0; w < kWaves; ++w)
            for (int j = 0; j < it.cnt[w]; ++j) {
                const int t = W.tws[(size_t)it.tw0 + w * kTWW + j].out;
                if (t < 0 || t >= W.ntw) return set_error(BB_E_INVALID, "internal: tile index out of range");
                tile_item[t] = (int)i;
            }
    }
    std::vector<long long> pbase(W.ntw);
    long long run = 0;
    for (int t = 0; t < W.ntw; ++t) {
        if (tile_item[t] < 0) return set_error(BB_E_INVALID, "internal: tile without a window");
        pbase[t] = run;
        run += sp[tile_item[t]];
    }
    if (run * 256 > 0x7fffffffLL)
        return set_error(BB_E_INVALID, "too many partial sums (" + std::to_string(run) +
                                           " partial tiles); lower x_splits");
    part_tiles = run;
    for (TWin &T : W.tws)
        if (T.out >= 0) T.out = (int32_t)pbase[T.out];
    nsp.assign(pos.size(), 1);
    for (size_t c = 0; c < pos.size(); ++c) {
        const int t = pos[c] / 256, r = pos[c] % 256;
        pos[c] = (int32_t)(pbase[t] * 256 + r);
        nsp[c] = (int16_t)sp[tile_item[t]];
    }
    units.clear();
    std::vector<std::pair<long long, size_t>> order;
    for (size_t i = 0; i < n; ++i)
        for (int s = 0; s < sp[i]; ++s) {
            WorkUnit u;
            u.item = (int32_t)i;
            u.split = s;
            u.cb = (int32_t)((long long)nch[i] * s / sp[i]);
            u.ce = (int32_t)((long long)nch[i] * (s + 1) / sp[i]);
            order.push_back({(long long)(u.ce - u.cb) * W.item_work[i], units.size()});
            units.push_back(u);
        }
    std::stable_sort(order.begin(), order.end(),
                     [](const std::pair<long long, size_t> &x, const std::pair<long long, size_t> &y) {
                         return x.first > y.first;
                     });
    std::vector<WorkUnit> sorted;
    sorted.reserve(units.size());
    for (auto &o : order) sorted.push_back(units[o.second]);
    units.swap(sorted);
    return BB_OK;
}

// Checks the invariants the contraction kernels rely on: every staged ring
// inside its level, every LDS row index inside the item's staged rows, every
// unit's chunk range inside its item, every partial sum written once.
int check_lists(const Layout &L, size_t ncontr)
{
    const long long nslots = L.part_tiles * 256;
    std::vector<uint8_t> seen((size_t)std::max<long long>(nslots, 1), 0);
    for (size_t c = 0; c < ncontr; ++c)
        for (int s = 0; s < L.nsp[c]; ++s) {
            const long long q = (long long)L.pos[c] + 256LL * s;
            if (L.pos[c] < 0 || L.nsp[c] < 1 || q >= nslots || seen[q])
                return set_error(BB_E_INVALID, "internal: contraction output slots are not unique");
            seen[q] = 1;
        }
    std::vector<int> item_sp(L.items.size(), 0);
    // Tile windows per wave the kernel that runs an item executes: the
    // layout's for its own units, the FP32 kernel's for units_f.
    std::vector<int> item_tww(L.items.size(), L.tww);
    for (const WorkUnit &u : L.units_f)
        if (u.item >= 0 && (size_t)u.item < L.items.size())
            item_tww[u.item] = std::min(item_tww[u.item], contract_tww(CPrec::F32));
    std::vector<WorkUnit> all(L.units);
    all.insert(all.end(), L.units_f.begin(), L.units_f.end());
    for (const WorkUnit &u : all) {
        if (u.item < 0 || (size_t)u.item >= L.items.size())
            return set_error(BB_E_INVALID, "internal: unit without an item");
        const WorkItem &it = L.items[u.item];
        const long long nch = (it.npts + L.chunk - 1) / L.chunk;
        if (u.cb < 0 || u.cb >= u.ce || u.ce > nch)
            return set_error(BB_E_INVALID, "internal: unit chunk range out of range");
        item_sp[u.item]++;
    }
    for (size_t i = 0; i < L.items.size(); ++i) {
        const WorkItem &it = L.items[i];
        int rows = 0;
        for (int b = 0; b < it.nblk; ++b) {
            rows += it.blk_rows[b];
            if (it.blk_slot[b] < 0 || it.fstride < it.npts || it.fstride % 32 != 0 ||
                it.off + (long long)(it.blk_slot[b] + it.blk_rows[b]) * it.fstride > L.cache_pts)
                return set_error(BB_E_INVALID, "internal: staged ring outside the cache");
            // the contraction addresses a ring as a 31-bit count of 8-byte words
            const long long words = (long long)(it.blk_slot[b] + it.blk_rows[b]) * it.fstride *
                                    (L.prec == CPrec::F64 ? 2 : 1);
            if (words >= (1LL << 31))
                return set_error(BB_E_RANGE, "a grid level's cache exceeds 16 GiB");
        }
        if (rows != it.nrows || it.nrows > kRowCap || it.nblk > kMaxBlk || item_sp[i] < 1)
            return set_error(BB_E_INVALID, "internal: work item row blocks are inconsistent");
        {
            // every staged field appears on exactly one row
            std::vector<long long> want, got;
            for (int b = 0; b < it.nblk; ++b)
                for (int r = 0; r < it.blk_rows[b]; ++r)
                    want.push_back(((long long)it.blk_slot[b] + r) * 2 + (it.blk_split[b] ? 1 : 0));
            for (int r = 0; r < it.nrows; ++r)
                got.push_back((long long)(it.rslot[r] & (kRowSplit - 1)) * 2 +
                              ((it.rslot[r] & kRowSplit) ? 1 : 0));
            std::sort(want.begin(), want.end());
            std::sort(got.begin(), got.end());
            if (want != got || std::adjacent_find(got.begin(), got.end()) != got.end())
                return set_error(BB_E_INVALID, "internal: staged rows do not match the fields");
        }
        for (int w = 0; w < kWaves; ++w) {
            const int limit = w < L.waves ? std::min(item_tww[i], (int)kTWW) : 0;
            if (it.cnt[w] < 0 || it.cnt[w] > limit)
                return set_error(BB_E_INVALID, "internal: too many tile windows in a wave");
            for (int j = it.cnt[w]; j < kTWW; ++j)
                if (L.tws[(size_t)it.tw0 + w * kTWW + j].out != -1)
                    return set_error(BB_E_INVALID, "internal: unused tile window slot is live");
            for (int j = 0; j < it.cnt[w]; ++j) {
                const TWin &T = L.tws[(size_t)it.tw0 + w * kTWW + j];
                bool ok = T.b0 >= 0 && T.b0 < it.nrows && T.nb >= 1 && T.nb <= 16 &&
                          T.out >= 0 && (long long)T.out + item_sp[i] <= L.part_tiles &&
                          (j > 0 || (T.flags & TW_NEWA));
                for (int m = 0; m < 16; ++m)
                    ok = ok && T.a1[m] >= -1 && T.a1[m] < it.nrows && T.a2[m] >= 0 &&
                         T.a2[m] < it.nrows;
                if (!ok) return set_error(BB_E_INVALID, "internal: tile window out of range");
            }
        }
    }
    return BB_OK;
}

// Largest |k_x| of any mode of ring r (0 for a ring without modes): the
// largest ia for which some ja puts ia^2 + ja^2 in [nlo, nhi).
std::vector<int> ring_kx(const PlanData &d)
{
    std::vector<int> K(d.g.Nr, 0);
    for (int r = 0; r < d.g.Nr; ++r) {
        const int lo = d.nlo[r], hi = d.nhi[r];
        for (int ia = hi > 0 ? isqrt_floor(hi - 1) : -1; ia >= 0; --ia) {
            const int top = hi - 1 - ia * ia;            // ja^2 <= top
            const int j = isqrt_floor(top);
            if (j >= 0 && (long long)j * j >= (long long)lo - (long long)ia * ia) {
                K[r] = ia;
                break;
            }
        }
    }
    return K;
}

// Grid levels.  A contraction of rings (r1, r2, r3) sums exactly on any grid
// of NG > K1 + K2 + K3 points per side (K_r the largest |k_x| of ring r: no
// sum of three modes wraps around), and the FFT of a ring on NG points holds
// all its modes when NG >= 2 K_r + 2; need(c) is the larger bound.  Both
// hold on the NN grid when NN >= 3 kmax + 1 (kmax over all rings); when NN
// is smaller the NN-grid sums alias, and only the NN grid reproduces them.
// Otherwise up to max_levels grids are chosen among the even 2,3,5-smooth
// sizes to minimise the sum over contractions of NG(c)^2 (dynamic
// programming over the histogram of needs); lvl[c] is the level of
// contraction c and NGtop the largest grid (the plane grid).  kmax is the
// largest |k_x| of any ring a contraction uses.
void choose_grids(const PlanData &d, bool full_grid, int max_levels, int &kmax, int &NGtop,
                  std::vector<int> &grids, std::vector<uint8_t> &lvl)
{
    const int NN = d.g.NN;
    const std::vector<int> K = ring_kx(d);
    int kall = 0;
    for (int k : K) kall = std::max(kall, k);
    lvl.assign(d.contr.size(), 0);
    kmax = kall;
    NGtop = NN;
    grids.assign(1, NN);
    if (full_grid || NN < 3 * kall + 1 || d.contr.empty()) return;

    std::vector<int> need(d.contr.size());
    int maxneed = 1;
    kmax = 0;
    for (size_t c = 0; c < d.contr.size(); ++c) {
        const Contraction &k = d.contr[c];
        const int a = K[k.r1], b = K[k.r2], e = K[k.r3];
        need[c] = std::max(a + b + e + 1, 2 * std::max(a, std::max(b, e)) + 2);
        maxneed = std::max(maxneed, need[c]);
        kmax = std::max(kmax, std::max(a, std::max(b, e)));
    }
    std::vector<int> cand;
    for (int n = 8; n <= NN; n += 2)
        if (smooth235(n)) {
            cand.push_back(n);
            if (n >= maxneed) break;
        }
    const int top = cand.back();   // NN itself qualifies (even, smooth, >= maxneed)
    std::vector<long long> cum(top + 2, 0);   // cum[x + 1] = contractions needing <= x
    for (int n : need) cum[n + 1]++;
    for (int x = 1; x <= top + 1; ++x) cum[x] += cum[x - 1];
    auto below = [&](int x) { return cum[std::min(x, top) + 1]; };
    const int C = (int)cand.size(), T = C - 1;
    const int Lmax = std::max(1, std::min(max_levels, C));
    const double inf = 1e300;
    std::vector<std::vector<double>> F(Lmax + 1, std::vector<double>(C, inf));
    std::vector<std::vector<int>> from(Lmax + 1, std::vector<int>(C, -1));
    for (int j = 0; j < C; ++j) F[1][j] = (double)below(cand[j]) * cand[j] * cand[j];
    for (int l = 2; l <= Lmax; ++l)
        for (int j = 0; j < C; ++j)
            for (int i = 0; i < j; ++i) {
                if (F[l - 1][i] >= inf) continue;
                const double v = F[l - 1][i] +
                                 (double)(below(cand[j]) - below(cand[i])) * cand[j] * cand[j];
                if (v < F[l][j]) {
                    F[l][j] = v;
                    from[l][j] = i;
                }
            }
    int best = 1;
    for (int l = 2; l <= Lmax; ++l)
        if (F[l][T] < F[best][T]) best = l;
    std::vector<int> g;
    for (int l = best, j = T; l >= 1 && j >= 0; j = from[l][j], --l) g.push_back(cand[j]);
    std::sort(g.begin(), g.end());
    grids.clear();
    int prev = 0;
    for (int n : g) {
        if (below(n) > below(prev) || n == g.back()) grids.push_back(n);   // skip empty levels
        prev = n;
    }
    NGtop = grids.back();
    for (size_t c = 0; c < d.contr.size(); ++c) {
        int l = 0;
        while (grids[l] < need[c]) ++l;
        lvl[c] = (uint8_t)l;
    }
}

int default_levels(const bb_options &opt)
{
    if (opt.grid_levels > 0) return opt.grid_levels;
    if (const char *ev = std::getenv("BARBIE_GRID_LEVELS")) {
        const int v = std::atoi(ev);
        if (v >= 1 && v <= BB_MAX_LEVELS) return v;
    }
    return 6;
}

int layout_waves(const bb_options &opt)
{
    if (opt.waves == 4 || opt.waves == 8) return opt.waves;
    if (const char *ev = std::getenv("BARBIE_CONTRACT_WAVES")) {
        const int w = std::atoi(ev);
        if (w == 4 || w == 8) return w;
    }
    return kWaves;
}

// BF16x3 at 3 workgroups per CU: fewer tile windows and staged rings.
bool layout_occ3(CPrec p, const bb_options &opt)
{
    return p == CPrec::BF16X3 && opt.occupancy == 3 && layout_waves(opt) == 4;
}
int layout_tww(CPrec p, const bb_options &opt)
{
    return layout_occ3(p, opt) ? kTWWOcc3 : contract_tww(p);
}
int layout_rowcap(CPrec p, const bb_options &opt)
{
    return layout_occ3(p, opt) ? rowcap_occ3(4) : rowcap_for(layout_waves(opt));
}

int build_layout(const PlanData &d, CPrec prec, const bb_options &opt, int waves, int tww,
                 int max_levels, Layout &L)
{
    const int NN = d.g.NN, Nmax = d.g.Nmax, Nr = d.g.Nr;
    L.prec = prec;
    L.waves = waves;
    L.tww = tww;
    L.chunk = contract_chunk(prec);
    L.NN = NN;
    std::vector<int> grids;
    choose_grids(d, opt.full_grid != 0, max_levels, L.kmax, L.NGtop, grids, L.lvl);
    const std::vector<uint8_t> &lvl = L.lvl;
    const std::vector<int> K = ring_kx(d);
    const size_t elem = prec == CPrec::F64 ? 16 : 8;

    // Fields of every level: per plane the contiguous range of rings in use.
    const int nl = (int)grids.size();
    L.lv.assign(nl, Level());
    {
        std::vector<std::vector<uint8_t>> used(nl, std::vector<uint8_t>((size_t)Nmax * Nr, 0));
        for (size_t c = 0; c < d.contr.size(); ++c) {
            const Contraction &k = d.contr[c];
            std::vector<uint8_t> &u = used[lvl[c]];
            u[(size_t)k.p1 * Nr + k.r1] = 1;
            u[(size_t)std::abs(k.p2) * Nr + k.r2] = 1;
            u[(size_t)std::abs(k.p3) * Nr + k.r3] = 1;
        }
        for (int l = 0; l < nl; ++l) {
            Level &V = L.lv[l];
            V.NG = grids[l];
            V.npts = (long long)V.NG * V.NG;
            V.stride = (V.npts + 31) / 32 * 32;
            V.scale = ((double)NN * NN) / ((double)V.NG * V.NG);
            V.off = L.cache_pts;
            V.rlo.assign(Nmax, -1);
            V.rhi.assign(Nmax, -2);
            V.base.assign(Nmax, 0);
            for (int p = 0; p < Nmax; ++p)
                for (int r = 0; r < Nr; ++r)
                    if (used[l][(size_t)p * Nr + r]) {
                        if (V.rlo[p] < 0) V.rlo[p] = r;
                        V.rhi[p] = r;
                    }
            for (int p = 0; p < Nmax; ++p) {
                V.base[p] = V.nfields;
                if (V.rlo[p] < 0) continue;
                for (int r = V.rlo[p]; r <= V.rhi[p]; ++r) {
                    const int band = d.nhi[r] > 0 ? std::min(V.NG / 2, K[r]) : -1;
                    V.tasks.push_back(FieldTask{p, r, V.nfields, band, d.nlo[r], d.nhi[r], p * Nr + r, l, 3});
                    V.max_rows = std::max(V.max_rows, std::min(V.NG, 2 * band + 1));
                    V.nfields++;
                }
            }
            L.cache_pts += (long long)V.nfields * V.stride;
        }
    }

    // Pair blocks per (level, plane group); groups of one level with the same
    // plane set {p1, |p2|, |p3|} are adjacent so that work items share staged
    // rings.
    WorkLists W;
    L.pos.assign(std::max<size_t>(d.contr.size(), 1), 0);
    L.mix = prec == CPrec::BF16X3 ? std::max(0, std::min(opt.mix_fp32, 90)) : 0;
    long long cost_all = 0, cost_f = 0;
    for (int l = 0; l < nl; ++l) {
        std::map<std::pair<int, int>, std::vector<CItem>> groups;
        for (size_t c = 0; c < d.contr.size(); ++c) {
            if (lvl[c] != l) continue;
            const Contraction &k = d.contr[c];
            groups[{k.p1, k.p2}].push_back({k.r1, k.r2, k.r3, (int)c, k.p1, std::abs(k.p2)});
        }
        std::map<std::array<int, 3>, std::vector<std::pair<int, int>>> by_set;
        for (auto &kv : groups) {
            const int p1 = kv.first.first, p2 = kv.first.second, p3 = -p1 - p2;
            std::array<int, 3> key{p1, std::abs(p2), std::abs(p3)};
            std::sort(key.begin(), key.end());
            by_set[key].push_back(kv.first);
        }
        // Mixed layouts: whole plane sets go to the FP32 kernel while its share
        // of the cost (tile windows x points) stays within `mix` percent.
        // Blocks per plane pair; the rows of a plane pair's last, partly
        // filled blocks are packed again with those of the other plane pairs
        // of the set that share the third plane and conjugation pattern (the
        // B operand), so that fewer A tile rows are empty.
        std::vector<PairBlock> blocks, blocks_f;
        for (auto &kv : by_set) {
            std::vector<PairBlock> sb;
            std::map<std::pair<int, int>, std::vector<CItem>> rest;
            for (auto &gk : kv.second) {
                const int p1 = gk.first, p2 = gk.second, p3 = -p1 - p2;
                const int flags = (p2 < 0 ? TW_CONJ2 : 0) | (p3 < 0 ? TW_CONJ3 : 0);
                std::vector<PairBlock> gb;
                pair_blocks(groups[gk], std::abs(p3), flags, gb);
                for (PairBlock &B : gb) {
                    if (B.nrow == 16 || !kRepack) {
                        sb.push_back(std::move(B));
                        continue;
                    }
                    std::vector<CItem> &dst = rest[{std::abs(p3), flags}];
                    for (const auto &c : B.cs)
                        dst.push_back({B.r1[c[0]], B.r2[c[0]], c[1], c[2], B.pa[c[0]], B.pb[c[0]]});
                }
            }
            for (auto &rk : rest) pair_blocks(rk.second, rk.first.first, rk.first.second, sb);
            long long c = 0;
            for (const PairBlock &B : sb) c += B.nwin;
            c *= L.lv[l].npts;
            const bool to_f = L.mix > 0 && (cost_f + c) * 100 <= (long long)L.mix * (cost_all + c);
            cost_all += c;
            if (to_f) cost_f += c;
            std::vector<PairBlock> &dst = to_f ? blocks_f : blocks;
            for (PairBlock &B : sb) dst.push_back(std::move(B));
        }
        // What each field of the level is read as: the BF16x3 kernel reads
        // pair fields as FP32 values and third fields as bf16 operands; the
        // FP32 kernel of mixed layouts reads all three as FP32 values.  The
        // ring pass writes only those forms.
        if (prec == CPrec::BF16X3) {
            std::vector<uint8_t> role((size_t)Nmax * Nr, 0);
            for (int f = 0; f < 2; ++f)
                for (const PairBlock &B : f ? blocks_f : blocks) {
                    for (int m = 0; m < B.nrow; ++m) {
                        role[(size_t)B.pa[m] * Nr + B.r1[m]] |= 1;
                        role[(size_t)B.pb[m] * Nr + B.r2[m]] |= 1;
                    }
                    for (const auto &c : B.cs) role[(size_t)B.pc * Nr + c[1]] |= f ? 1 : 2;
                }
            for (FieldTask &t : L.lv[l].tasks) t.forms = role[(size_t)t.plane * Nr + t.ring];
        }
        int rc = pack_items(blocks, L.lv[l], prec == CPrec::BF16X3, tww, waves,
                            layout_rowcap(prec, opt), W, L.pos);
        if (rc) return rc;
        const size_t first_f = W.items.size();
        rc = pack_items(blocks_f, L.lv[l], false, contract_tww(CPrec::F32), waves,
                        rowcap_for(waves), W, L.pos);
        if (rc) return rc;
        for (size_t i = first_f; i < W.items.size(); ++i) W.kind[i] = 1;
    }
    L.ntiles = W.ntw;
    L.max_rows = W.max_rows;
    L.wave_balance = W.tw_slots ? (double)W.tw_used / (double)W.tw_slots : 1.0;
    for (const WorkItem &it : W.items) {
        L.staged_rows += it.nrows;
        L.staged_bytes += (long long)it.nrows * it.npts * (long long)elem;   // padding excluded
    }
    for (size_t c = 0; c < d.contr.size(); ++c) L.useful_pts += L.lv[lvl[c]].npts;
    std::vector<WorkUnit> units;
    int rc = make_units(W, L.chunk, opt.x_splits, units, L.pos, L.nsp, L.part_tiles,
                        &L.total_cost, target_units(opt.max_batch));
    if (rc) return rc;
    for (const WorkUnit &u : units) (W.kind[u.item] ? L.units_f : L.units).push_back(u);
    L.items.swap(W.items);
    L.tws.swap(W.tws);
    if ((rc = check_lists(L, d.contr.size()))) return rc;

    // Binning records, sorted by bin.
    if (Nr > 255 || Nmax > 32767)
        return set_error(BB_E_INVALID, "the binning records hold up to 255 rings and 32767 planes");
    L.recs.resize(d.contr.size());
    L.seg.assign(d.nbins + 1, 0);
    for (const Contraction &k : d.contr) L.seg[k.bin + 1]++;
    for (int b = 0; b < d.nbins; ++b) L.seg[b + 1] += L.seg[b];
    {
        std::vector<int32_t> at(L.seg.begin(), L.seg.end() - 1);
        for (size_t c = 0; c < d.contr.size(); ++c) {
            const Contraction &k = d.contr[c];
            BinRec R;
            std::memset(&R, 0, sizeof R);
            R.p1 = (int16_t)k.p1; R.p2 = (int16_t)k.p2; R.p3 = (int16_t)k.p3;
            R.r1 = (uint8_t)k.r1; R.r2 = (uint8_t)k.r2; R.r3 = (uint8_t)k.r3;
            R.nsp = L.nsp[c];
            R.rp = k.rp;
            R.pos = L.pos[c];
            L.recs[at[k.bin]++] = R;
        }
    }
    // Within a bin, records in partial-sum order: neighbouring threads then
    // read neighbouring partial sums.  (The bin's summation order is fixed by
    // this order.)
    for (int b = 0; b < d.nbins; ++b)
        std::sort(L.recs.begin() + L.seg[b], L.recs.begin() + L.seg[b + 1],
                  [](const BinRec &x, const BinRec &y) { return x.pos < y.pos; });
    // Binning segments: each bin's records in pieces of at most kBinSeg.
    L.sg.assign(1, 0);
    L.bin_sg.assign(d.nbins + 1, 0);
    for (int b = 0; b < d.nbins; ++b) {
        L.bin_sg[b] = (int32_t)L.sg.size() - 1;
        for (int32_t r = L.seg[b]; r < L.seg[b + 1]; r += kBinSeg)
            L.sg.push_back(std::min(r + kBinSeg, L.seg[b + 1]));
    }
    L.bin_sg[d.nbins] = (int32_t)L.sg.size() - 1;
    L.bp.nbins = d.nbins;
    L.bp.n_mtbin = d.g.n_mtbin;
    L.bp.NN = NN;
    L.bp.Nr = Nr;
    L.bp.Cunits = 2. * M_PI / (NN * d.g.LL);
    L.bp.ff = d.g.ff;
    L.bp.beta = d.g.beta;
    return BB_OK;
}

// The layout of (plan, precision, options), built once and cached in the plan.
int get_layout(const bb_plan *p, CPrec prec, const bb_options &opt,
               std::shared_ptr<const Layout> &out)
{
    const int waves = layout_waves(opt), tww = layout_tww(prec, opt), lev = default_levels(opt);
    const std::string key = std::to_string((int)prec) + "/" + std::to_string(opt.full_grid != 0) +
                            "/" + std::to_string(opt.x_splits) + "/" + std::to_string(lev) + "/" +
                            std::to_string(waves) + "/" + std::to_string(tww) + "/" +
                            std::to_string(prec == CPrec::BF16X3 ? opt.mix_fp32 : 0) + "/" +
                            std::to_string(layout_occ3(prec, opt)) + "/" +
                            std::to_string(target_units(opt.max_batch));
    bb_plan *mp = const_cast<bb_plan *>(p);
    {
        std::lock_guard<std::mutex> g(mp->mu);
        auto it = mp->layouts.find(key);
        if (it != mp->layouts.end()) {
            out = it->second;
            return BB_OK;
        }
    }
    auto L = std::make_shared<Layout>();
    int rc = build_layout(p->d, prec, opt, waves, tww, lev, *L);
    if (rc) return rc;
    std::lock_guard<std::mutex> g(mp->mu);
    auto ins = mp->layouts.emplace(key, std::shared_ptr<const Layout>(L));
    out = ins.first->second;
    return BB_OK;
}

}  // namespace

struct bb_estimator {
    const bb_plan *plan = nullptr;
    bb_options opt{};
    CPrec cprec = CPrec::F64;
    hipStream_t stream = nullptr;
    std::shared_ptr<const Layout> L;
    std::vector<LevelDesc> lvd;       // FFT geometry per level (host copy)
    LevelDesc *d_lvd = nullptr;
    double2 *d_tw = nullptr;          // twiddles of every level
    FftBlocks fb;                     // work units of the ring pass
    std::vector<void *> owned;
    int batch = 1;                    // fields per launch held in device memory
    BatchStrides bst;                 // per-field buffer strides
    bool pending = false;             // an enqueued batch not yet waited for
    double2 *d_field = nullptr;       // staging for bb_estimate (host input)
    int field_cap = 0;                // fields the staging holds
    MixUnit *d_mix = nullptr;         // bb_estimate_mixed (this estimator first): unit list
    size_t mix_cap = 0;
    int *d_mix_counter = nullptr;
    std::vector<unsigned long long> mix_key;   // what the uploaded unit list was built for
    int mix_n = 0;
    unsigned long long serial = 0;    // creation number: unique per estimator in the process
    double2 *h_stage = nullptr;       // pinned host staging of the modes the estimate reads
    double2 *d_stage = nullptr;       // its device copy
    size_t stage_cap = 0;             // complex values both hold
    double2 *d_planes = nullptr;      // NGtop x NGtop per plane
    void *d_cache = nullptr;
    void *d_split = nullptr;          // BF16x3: bf16 hi/lo operands of every field
    FieldTask *d_tasks = nullptr;
    WorkItem *d_items = nullptr;
    TWin *d_tws = nullptr;
    WorkUnit *d_units = nullptr;
    WorkUnit *d_units_f = nullptr;    // mixed layouts: units of the FP32 kernel
    hipStream_t aux = nullptr;        // mixed layouts: the FP32 kernel's stream
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    double *d_part = nullptr;         // partial tiles
    BinRec *d_recs = nullptr;
    int32_t *d_sg = nullptr, *d_bin_sg = nullptr;
    double *d_binpart = nullptr;      // [segment][48] per field
    double *d_bincoef = nullptr;      // [33][records] binning coefficients (batches), or null
    BinStatic bstat;                  // field-independent norm and analytic sums
    double *d_num = nullptr;
    std::vector<double> num_host;
    int *d_nz = nullptr;              // [Nmax * Nr]
    double *d_avg = nullptr;
    double *d_bins = nullptr;
    hipEvent_t ev[5] = {};
    float ms[5] = {0, 0, 0, 0, 0};
    // Captured launch sequence (use_graph) and the arguments it was captured for.
    hipGraphExec_t gexec = nullptr;
    int g_nf = 0;
    const double *g_in = nullptr;
    double *g_out[3] = {nullptr, nullptr, nullptr};
    bool graph_run = false;   // the pending batch ran as a graph
    hipStream_t own_stream = nullptr;   // created with the estimator; `stream` may be lent

    ~bb_estimator()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        if (h_stage) (void)hipHostFree(h_stage);
        if (gexec) (void)hipGraphExecDestroy(gexec);
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
        if (aux) (void)hipStreamSynchronize(aux);
        for (void *p : owned) (void)hipFree(p);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        if (aux) (void)hipStreamDestroy(aux);
        if (own_stream) (void)hipStreamDestroy(own_stream);
    }
};

namespace {

int upload(void *dst, const void *src, size_t bytes, hipStream_t s)
{
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    return BB_OK;
}

std::vector<double2> twiddles(int NG)
{
    // exp(+2 pi i t / NG), FP64
    std::vector<double2> tw(NG);
    for (int t = 0; t < NG; ++t) {
        long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)t / NG;
        tw[t] = make_double2((double)cosl(a), (double)sinl(a));
    }
    return tw;
}

// FFT geometry of a level: radices, twiddle offset, batch sizes, cache
// placement.
int level_desc(const Level &V, int tw_off, int elem, LevelDesc &D)
{
    std::memset(&D, 0, sizeof D);
    if (!make_ring_fft_desc(V.NG, D.fd)) return set_error(BB_E_INVALID, "unsupported FFT length");
    D.tw_off = tw_off;
    D.cpb = fft_cols_per_block(V.NG, elem);
    while ((1 << D.cpb_log2) < D.cpb) ++D.cpb_log2;
    if ((1 << D.cpb_log2) != D.cpb || (D.cpb * V.NG) % 8 != 0)
        return set_error(BB_E_INVALID, "internal: FFT batch sizes");
    D.max_rows = V.max_rows;
    D.off = V.off;
    D.stride = V.stride;
    return BB_OK;
}

// Plane-ring fields of every level into the cache: FP64, or FP32 plus the
// bf16 operands.
int run_fields(bb_estimator *E, const double2 *planes, int NP, int *nz, int nf)
{
    const int nl = (int)E->lvd.size();
    if (E->cprec == CPrec::F64) {
        HIPCHK(launch_fft_f64(planes, NP, (double2 *)E->d_cache, E->d_tasks, E->fb, E->d_lvd,
                              E->lvd.data(), nl, E->d_tw, nz, E->bst, nf, E->stream));
    } else {
        HIPCHK(launch_fft_f32(planes, NP, (float2 *)E->d_cache, E->d_split, E->d_tasks, E->fb,
                              E->d_lvd, E->lvd.data(), nl, E->d_tw, nz, E->bst, nf,
                              E->stream));
    }
    return BB_OK;
}

// Triangle counts Num on the GPU (calCR, funcs.c:328-424): unit fields
// I_r = IFFT2(mask_r) in FP64 on the NN grid, then Num = sum_x I1 I2 I3 for
// every ring triple with the FP64 contraction kernel.  Cached in the plan.
// The NN grid, as in the reference, and not the estimator's grid levels:
// calCR counts every ring triple, also those of rings beyond UL_k1 that no
// contraction reads, and for those a smaller grid would alias (their
// K1 + K2 + K3 exceeds the levels' sizes).
int compute_num(bb_estimator *E)
{
    bb_plan *mp = const_cast<bb_plan *>(E->plan);
    {
        std::lock_guard<std::mutex> g(mp->mu);
        if (mp->have_num) {
            E->num_host = mp->num;
            return BB_OK;
        }
    }
    const PlanData &d = E->plan->d;
    const int Nr = d.g.Nr, NN = d.g.NN, NG = NN;
    const int rcount = (int)(d.tri.size() / 3);
    const std::vector<int> K = ring_kx(d);
    Level U;
    U.NG = NG;
    U.npts = (long long)NG * NG;
    U.stride = (U.npts + 31) / 32 * 32;
    U.scale = ((double)NN * NN) / ((double)NG * NG);
    U.rlo.assign(1, 0);
    U.rhi.assign(1, Nr - 1);
    U.base.assign(1, 0);
    U.nfields = Nr;
    for (int r = 0; r < Nr; ++r) {
        const int band = d.nhi[r] > 0 ? std::min(NG / 2, K[r]) : -1;
        U.tasks.push_back(FieldTask{-1, r, r, band, d.nlo[r], d.nhi[r], 0, 0, 3});
        U.max_rows = std::max(U.max_rows, std::min(NG, 2 * band + 1));
    }
    Layout L;
    L.chunk = contract_chunk(CPrec::F64);
    L.tww = contract_tww(CPrec::F64);
    L.cache_pts = (long long)Nr * U.stride;
    std::vector<CItem> cs;
    for (int rp = 0; rp < rcount; ++rp)
        cs.push_back({d.tri[3 * rp], d.tri[3 * rp + 1], d.tri[3 * rp + 2], rp, 0, 0});
    WorkLists W;
    L.pos.assign(std::max(rcount, 1), 0);
    std::vector<PairBlock> blocks;
    if (rcount > 0) pair_blocks(cs, 0, 0, blocks);
    int rc = pack_items(blocks, U, false, contract_tww(CPrec::F64), kWaves, rowcap_for(kWaves), W,
                        L.pos);
    if (!rc) rc = make_units(W, L.chunk, 0, L.units, L.pos, L.nsp, L.part_tiles, nullptr);
    if (rc) return rc;
    L.items.swap(W.items);
    L.tws.swap(W.tws);
    if ((rc = check_lists(L, rcount))) return rc;
    LevelDesc ud;
    if ((rc = level_desc(U, 0, 8, ud))) return rc;
    const std::vector<double2> tw = twiddles(NG);
    const long long uslots = std::max<long long>(L.part_tiles, 1) * 256;
    std::vector<void *> tmp;
    FieldTask *d_ut;
    LevelDesc *d_ud;
    int2 *d_ublk;
    std::vector<int2> ublk;
    fft_blocks(U.tasks.data(), Nr, &ud, ublk);
    double2 *d_unit, *d_utw;
    WorkItem *d_uitems;
    WorkUnit *d_uunits;
    TWin *d_utws;
    double *d_upart;
    rc = dev_alloc(&d_ut, Nr, tmp);
    if (!rc) rc = dev_alloc(&d_ud, 1, tmp);
    if (!rc) rc = dev_alloc(&d_ublk, std::max<size_t>(ublk.size(), 1), tmp);
    if (!rc && !ublk.empty()) rc = upload(d_ublk, ublk.data(), sizeof(int2) * ublk.size(), E->stream);
    if (!rc) rc = dev_alloc(&d_utw, NG, tmp);
    if (!rc) rc = dev_alloc(&d_unit, (size_t)L.cache_pts, tmp);
    if (!rc && hipMemsetAsync(d_unit, 0, sizeof(double2) * L.cache_pts, E->stream) != hipSuccess)
        rc = set_error(BB_E_DEVICE, "unit-field clear failed");
    if (!rc) rc = dev_alloc(&d_uitems, L.items.size(), tmp);
    if (!rc) rc = dev_alloc(&d_uunits, L.units.size(), tmp);
    if (!rc) rc = dev_alloc(&d_utws, L.tws.size(), tmp);
    if (!rc) rc = dev_alloc(&d_upart, (size_t)uslots, tmp);
    if (!rc) rc = upload(d_ut, U.tasks.data(), sizeof(FieldTask) * Nr, E->stream);
    if (!rc) rc = upload(d_ud, &ud, sizeof(LevelDesc), E->stream);
    if (!rc) rc = upload(d_utw, tw.data(), sizeof(double2) * NG, E->stream);
    if (!rc && !L.items.empty())
        rc = upload(d_uitems, L.items.data(), sizeof(WorkItem) * L.items.size(), E->stream);
    if (!rc && !L.units.empty())
        rc = upload(d_uunits, L.units.data(), sizeof(WorkUnit) * L.units.size(), E->stream);
    if (!rc && !L.tws.empty())
        rc = upload(d_utws, L.tws.data(), sizeof(TWin) * L.tws.size(), E->stream);
    if (!rc) {
        FftBlocks ub;
        ub.units = d_ublk;
        ub.n = (int)ublk.size();
        const BatchStrides one;
        hipError_t e = launch_fft_f64(nullptr, NG, d_unit, d_ut, ub, d_ud, &ud, 1, d_utw, nullptr,
                                      one, 1, E->stream);
        if (e == hipSuccess && rcount > 0)
            e = launch_contract(CPrec::F64, kWaves, contract_tww(CPrec::F64), d_unit, nullptr,
                                d_uitems, d_uunits, (int)L.units.size(), d_utws, d_upart, one, 1,
                                E->stream);
        if (e != hipSuccess) rc = set_error(BB_E_DEVICE, hipGetErrorString(e));
    }
    std::vector<double> part((size_t)uslots);
    if (!rc && hipMemcpyAsync(part.data(), d_upart, sizeof(double) * part.size(),
                              hipMemcpyDeviceToHost, E->stream) != hipSuccess)
        rc = set_error(BB_E_DEVICE, "triangle-count copy failed");
    if (!rc && hipStreamSynchronize(E->stream) != hipSuccess)
        rc = set_error(BB_E_DEVICE, "triangle-count kernels failed");
    for (void *q : tmp) (void)hipFree(q);
    if (rc) return rc;
    E->num_host.assign(std::max(rcount, 1), 0.0);
    for (int rp = 0; rp < rcount; ++rp) {
        double v = 0.0;
        for (int s = 0; s < L.nsp[rp]; ++s) v += part[(size_t)L.pos[rp] + 256 * (size_t)s];
        E->num_host[rp] = v;
    }
    std::lock_guard<std::mutex> g(mp->mu);
    mp->num = E->num_host;
    mp->have_num = true;
    return BB_OK;
}

// The reference's on-disk caches (funcs.c:328-452, 537-566) from a field:
// closed_rings_<NN> (rcount x {r1, r2, r3, Num} doubles, calCR order),
// rcountFor_<NN> (the count as text), and for every plane p < Nmax and ring
// r < Nr dataPlaneRing_<NN>_<p>_<r>_R<NRi>: D_{p,r} on the NN grid, 2 NN^2
// doubles, row-major interleaved complex (wrapid_fast's layout).  The fields
// are built on the device in FP64 by the estimator's transforms, nb at a
// time, and written as they arrive.  A text file cache_config_<NN>_R<NRi>.txt
// beside them records the geometry and its 64-bit FNV-1a hash, the key the
// reference's file names lack.
// 64-bit FNV-1a key of a geometry: NN, Nmax, Nr, n_mtbin and the ring edges.
// The reference's cache file names hold only NN, p, r and NRi, so caches of
// two geometries with the same NN cannot be told apart by name.
unsigned long long geometry_hash(const PlanData &d)
{
    unsigned long long h = 1469598103934665603ULL;
    auto mix = [&](const void *p, size_t n) {
        const unsigned char *b = (const unsigned char *)p;
        for (size_t k = 0; k < n; ++k) h = (h ^ b[k]) * 1099511628211ULL;
    };
    const int ints[4] = {d.g.NN, d.g.Nmax, d.g.Nr, d.g.n_mtbin};
    mix(ints, sizeof ints);
    mix(d.low.data(), d.low.size() * sizeof(double));
    mix(d.high.data(), d.high.size() * sizeof(double));
    return h;
}

std::string cache_name(const PlanData &d, const std::string &dir, int p, int r, int NRi)
{
    return dir + "/dataPlaneRing_" + std::to_string(d.g.NN) + "_" + std::to_string(p) + "_" +
           std::to_string(r) + "_R" + std::to_string(NRi);
}

// The plane-ring fields D_{p,r} on the NN grid (FP64, the estimator's gather
// and transforms, in batches of about 256 MiB) of the fields (p, r) =
// (f / Nr, f % Nr) with f a multiple of `stride`; fn(p, r, data) receives
// each in wrapid_fast's layout (funcs.c:554-563: 2 NN^2 doubles, row-major
// interleaved complex) and returns 0 or an error.
template <class Fn>
int visit_nn_fields(bb_estimator *E, const double *field_r2c, int stride, Fn fn)
{
    const PlanData &d = E->plan->d;
    const int NN = d.g.NN, Nmax = d.g.Nmax, Nr = d.g.Nr;
    const std::vector<int> K = ring_kx(d);
    Level V;
    V.NG = NN;
    V.npts = (long long)NN * NN;
    V.stride = (V.npts + 31) / 32 * 32;
    LevelDesc D;
    if (int rc = level_desc(V, 0, 8, D)) return rc;
    std::vector<int> sel;
    for (int f = 0; f < Nmax * Nr; f += std::max(1, stride)) sel.push_back(f);
    const int nsel = (int)sel.size();
    const int nb = std::max(1, std::min(nsel, (int)((256LL << 20) / (V.stride * 16))));
    const std::vector<double2> tw = twiddles(NN);
    const size_t fsz = (size_t)NN * NN * (NN / 2 + 1);
    std::vector<void *> tmp;
    double2 *d_field, *d_planes, *d_cache, *d_tw;
    FieldTask *d_tasks;
    LevelDesc *d_ld;
    int2 *d_blk;
    const size_t maxblk = (size_t)nb * (NN / D.cpb + 2);
    int rc = dev_alloc(&d_field, fsz, tmp);
    if (!rc) rc = dev_alloc(&d_planes, (size_t)Nmax * NN * NN, tmp);
    if (!rc) rc = dev_alloc(&d_cache, (size_t)nb * V.stride, tmp);
    if (!rc) rc = dev_alloc(&d_tw, (size_t)NN, tmp);
    if (!rc) rc = dev_alloc(&d_tasks, (size_t)nb, tmp);
    if (!rc) rc = dev_alloc(&d_ld, 1, tmp);
    if (!rc) rc = dev_alloc(&d_blk, maxblk, tmp);
    hipStream_t s = E->stream;
    if (!rc) rc = upload(d_field, field_r2c, fsz * sizeof(double2), s);
    if (!rc) rc = upload(d_tw, tw.data(), sizeof(double2) * NN, s);
    if (!rc) rc = upload(d_ld, &D, sizeof D, s);
    if (!rc) {
        hipError_t e = launch_gather(d_field, d_planes, NN, NN, NN / 2, Nmax, 1, s);
        if (e != hipSuccess) rc = set_error(BB_E_DEVICE, hipGetErrorString(e));
    }
    std::vector<double2> host((size_t)nb * V.stride);
    std::vector<double> row((size_t)2 * NN * NN);
    for (int s0 = 0; s0 < nsel && !rc; s0 += nb) {
        const int n = std::min(nb, nsel - s0);
        std::vector<FieldTask> tasks;
        for (int k = 0; k < n; ++k) {
            const int p = sel[s0 + k] / Nr, r = sel[s0 + k] % Nr;
            const int band = d.nhi[r] > 0 ? std::min(NN / 2, K[r]) : -1;
            tasks.push_back(FieldTask{p, r, k, band, d.nlo[r], d.nhi[r], 0, 0, 3});
        }
        std::vector<int2> blk;
        fft_blocks(tasks.data(), n, &D, blk);
        if (blk.size() > maxblk) {
            rc = set_error(BB_E_INVALID, "internal: field block list");
            break;
        }
        rc = upload(d_tasks, tasks.data(), sizeof(FieldTask) * n, s);
        if (!rc && !blk.empty()) rc = upload(d_blk, blk.data(), sizeof(int2) * blk.size(), s);
        if (!rc) {
            FftBlocks fb;
            fb.units = d_blk;
            fb.n = (int)blk.size();
            const BatchStrides one;
            hipError_t e = launch_fft_f64(d_planes, NN, d_cache, d_tasks, fb, d_ld, &D, 1, d_tw,
                                          nullptr, one, 1, s);
            if (e == hipSuccess)
                e = hipMemcpyAsync(host.data(), d_cache, sizeof(double2) * n * V.stride,
                                   hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            if (e != hipSuccess) rc = set_error(BB_E_DEVICE, hipGetErrorString(e));
        }
        for (int k = 0; k < n && !rc; ++k) {
            const int p = sel[s0 + k] / Nr, r = sel[s0 + k] % Nr;
            const double2 *src = host.data() + (size_t)k * V.stride;   // column-major
            for (int i = 0; i < NN; ++i)
                for (int j = 0; j < NN; ++j) {
                    row[2 * ((size_t)i * NN + j)] = src[(size_t)j * NN + i].x;
                    row[2 * ((size_t)i * NN + j) + 1] = src[(size_t)j * NN + i].y;
                }
            rc = fn(p, r, row.data());
        }
    }
    (void)hipStreamSynchronize(s);
    for (void *q : tmp) (void)hipFree(q);
    return rc;
}

// The reference's on-disk caches (funcs.c:328-452, 537-566) from a field:
// closed_rings_<NN> (rcount x {r1, r2, r3, Num} doubles, calCR order),
// rcountFor_<NN> (the count as text), and for every plane p < Nmax and ring
// r < Nr dataPlaneRing_<NN>_<p>_<r>_R<NRi>.  A text file
// cache_config_<NN>_R<NRi>.txt beside them records the geometry and its key.
int export_cache(bb_estimator *E, const double *field_r2c, const std::string &dir, int NRi)
{
    const PlanData &d = E->plan->d;
    const int NN = d.g.NN, Nmax = d.g.Nmax, Nr = d.g.Nr;
    const int rcount = (int)(d.tri.size() / 3);
    const std::string nn = std::to_string(NN);
    auto path = [&](const std::string &f) { return dir + "/" + f; };
    // Ring triples and triangle counts.
    {
        FILE *f = std::fopen(path("closed_rings_" + nn).c_str(), "wb");
        if (!f) return set_error(BB_E_IO, "cannot create " + path("closed_rings_" + nn));
        bool ok = true;
        for (int rp = 0; rp < rcount && ok; ++rp) {
            const double v[4] = {(double)d.tri[3 * rp], (double)d.tri[3 * rp + 1],
                                 (double)d.tri[3 * rp + 2], E->num_host[rp]};
            ok = std::fwrite(v, sizeof(double), 4, f) == 4;
        }
        ok = (std::fclose(f) == 0) && ok;
        FILE *g = std::fopen(path("rcountFor_" + nn).c_str(), "w");
        if (!g) return set_error(BB_E_IO, "cannot create " + path("rcountFor_" + nn));
        ok = std::fprintf(g, "%d", rcount) > 0 && ok;
        ok = (std::fclose(g) == 0) && ok;
        if (!ok) return set_error(BB_E_IO, "write failed in " + dir);
    }
    // Geometry key.
    {
        const std::string cf = path("cache_config_" + nn + "_R" + std::to_string(NRi) + ".txt");
        FILE *f = std::fopen(cf.c_str(), "w");
        if (!f) return set_error(BB_E_IO, "cannot create " + cf);
        std::fprintf(f, "NN %d\nNmax %d\nNr %d\nn_mtbin %d\nNRi %d\nkey %016llx\n", NN, Nmax, Nr,
                     d.g.n_mtbin, NRi, geometry_hash(d));
        for (int r = 0; r < Nr; ++r)
            std::fprintf(f, "ring %d %.17g %.17g\n", r, d.low[r], d.high[r]);
        if (std::fclose(f) != 0) return set_error(BB_E_IO, "write failed: " + cf);
    }
    // Plane-ring fields on the NN grid, FP64.
    return visit_nn_fields(E, field_r2c, 1, [&](int p, int r, const double *data) {
        const std::string fn = cache_name(d, dir, p, r, NRi);
        FILE *f = std::fopen(fn.c_str(), "wb");
        if (!f) return set_error(BB_E_IO, "cannot create " + fn);
        const size_t n = (size_t)2 * NN * NN;
        const bool ok = std::fwrite(data, sizeof(double), n, f) == n;
        if ((std::fclose(f) != 0) || !ok) return set_error(BB_E_IO, "write failed: " + fn);
        return (int)BB_OK;
    });
}

// The key recorded in cache_config_<NN>_R<NRi>.txt, if the file exists:
// 1 and *key set, 0 when there is no such file, an error when it is
// unreadable.
int read_cache_key(const std::string &file, unsigned long long *key)
{
    FILE *f = std::fopen(file.c_str(), "r");
    if (!f) return 0;
    char line[256];
    int found = 0;
    while (std::fgets(line, sizeof line, f))
        if (std::sscanf(line, "key %llx", key) == 1) found = 1;
    std::fclose(f);
    if (!found) return set_error(BB_E_IO, file + " holds no key");
    return 1;
}

// read_rings (funcs.c:429-452) with the checks the reference lacks.
int import_rings(bb_estimator *E, const std::string &dir, int NRi, double rel_tol, int use_num,
                 long long *nmismatch)
{
    const PlanData &d = E->plan->d;
    const std::string nn = std::to_string(d.g.NN);
    unsigned long long key = 0;
    const std::string cf = dir + "/cache_config_" + nn + "_R" + std::to_string(NRi) + ".txt";
    const int have = read_cache_key(cf, &key);
    if (have < 0) return have;
    if (have == 1 && key != geometry_hash(d))
        return set_error(BB_E_INVALID, cf + ": the cache was written for another geometry");
    int rcount = -1;
    {
        const std::string fn = dir + "/rcountFor_" + nn;
        FILE *f = std::fopen(fn.c_str(), "r");
        if (!f) return set_error(BB_E_IO, "cannot open " + fn);
        const int ok = std::fscanf(f, "%d", &rcount);
        std::fclose(f);
        if (ok != 1 || rcount < 0) return set_error(BB_E_IO, fn + ": no ring-triple count");
    }
    const int want = (int)(d.tri.size() / 3);
    if (rcount != want)
        return set_error(BB_E_INVALID, "rcountFor_" + nn + " holds " + std::to_string(rcount) +
                                           " ring triples; this geometry has " + std::to_string(want));
    std::vector<double> data((size_t)4 * rcount);
    {
        const std::string fn = dir + "/closed_rings_" + nn;
        FILE *f = std::fopen(fn.c_str(), "rb");
        if (!f) return set_error(BB_E_IO, "cannot open " + fn);
        const size_t got = std::fread(data.data(), sizeof(double), data.size(), f);
        std::fclose(f);
        if (got != data.size()) return set_error(BB_E_IO, fn + " is short");
    }
    double scale = 0.0;
    for (int rp = 0; rp < rcount; ++rp) scale = std::max(scale, std::fabs(E->num_host[rp]));
    long long bad = 0;
    for (int rp = 0; rp < rcount; ++rp) {
        for (int q = 0; q < 3; ++q)
            if (data[4 * rp + q] != (double)d.tri[3 * rp + q])
                return set_error(BB_E_INVALID, "closed_rings_" + nn + ": ring triple " +
                                                   std::to_string(rp) + " differs from this geometry's");
        if (!(std::fabs(data[4 * rp + 3] - E->num_host[rp]) <= rel_tol * scale)) ++bad;
    }
    if (nmismatch) *nmismatch = bad;
    if (bad)
        return set_error(BB_E_INVALID, "closed_rings_" + nn + ": " + std::to_string(bad) +
                                           " triangle counts differ from this geometry's");
    if (use_num) {
        for (int rp = 0; rp < rcount; ++rp) E->num_host[rp] = data[4 * rp + 3];
        if (int rc = upload(E->d_num, E->num_host.data(), sizeof(double) * rcount, E->stream))
            return rc;
        if (hipStreamSynchronize(E->stream) != hipSuccess)
            return set_error(BB_E_DEVICE, "triangle-count upload failed");
    }
    return BB_OK;
}

// Estimators of more than one field per batch keep the records' binning
// coefficients in a table when it is at most kBinTableBytes: every field of
// a batch then reads them instead of computing them again.
size_t bin_table_bytes(const Layout &L, size_t F)
{
    const size_t b = L.recs.size() * 33 * sizeof(double);
    return (F > 1 && b <= kBinTableBytes && L.recs.size() <= 0x7fffffff) ? b : 0;
}

// Device bytes held by an estimator of layout L with F fields per batch:
// per-field buffers times F, shared work lists once, and the staging of one
// host field (bb_estimate; bb_estimate_batch grows it to the batch).
size_t layout_bytes(const PlanData &d, const Layout &L, size_t F)
{
    const size_t NN = d.g.NN, Nr = d.g.Nr, Nmax = d.g.Nmax;
    const size_t top2 = (size_t)L.NGtop * L.NGtop;
    size_t b = NN * NN * (NN / 2 + 1) * 16;                  // host-field staging
    {
        const size_t km = std::min(NN / 2, (size_t)std::max(0, L.kmax));
        const size_t nr = std::min(NN, 2 * km + 1);
        b += nr * nr * Nmax * 16;                            // its compact upload
    }
    size_t per = Nmax * top2 * 16;                           // gathered planes
    per += (size_t)L.cache_pts * (L.prec == CPrec::F64 ? 16 : 8);
    if (L.prec == CPrec::BF16X3) per += (size_t)L.cache_pts * 8;
    per += (size_t)std::max<long long>(L.part_tiles, 1) * 256 * 8;
    per += Nmax * Nr * 4 + (size_t)3 * d.nbins * BB_NPOLES * 8;
    b += per * F;
    size_t tw = 0, nt = 0, nblk = 0;
    for (const Level &V : L.lv) {
        tw += V.NG;
        nt += V.tasks.size();
        const int cpb = fft_cols_per_block(V.NG, L.prec == CPrec::F64 ? 8 : 4);
        nblk += V.tasks.size() * (size_t)((V.NG + cpb - 1) / cpb);   // at most one unit per block
    }
    b += nblk * sizeof(int2);
    b += tw * 16 + nt * sizeof(FieldTask) + L.lv.size() * sizeof(LevelDesc);
    b += L.items.size() * sizeof(WorkItem) + L.tws.size() * sizeof(TWin) +
         L.units.size() * sizeof(WorkUnit);
    b += L.recs.size() * sizeof(BinRec) + (L.sg.size() + L.bin_sg.size()) * 4;
    b += (L.sg.size() - 1) * 48 * 8 * (size_t)F;
    b += bin_table_bytes(L, F);
    b += (L.sg.size() - 1) * 16 * 8 * (size_t)F + (size_t)3 * d.nbins * BB_NPOLES * 8 +
         (size_t)Nmax * Nr * 4 + (size_t)F * 4;   // field-independent bins (static_bins)
    b += d.tri.size() / 3 * 8 + Nr * 8;
    return b;
}

// The field-independent bins (BinStatic): the reference nonzero flags are
// those of any field whose ring modes are not all zero, i.e. every field a
// task builds except the rings with no mode (plane 0: no mode but the DC
// mode, which the reference zeroes).  One full binning pass with those flags
// (and zero contraction sums) gives the norm and analytic sums.
int static_bins(bb_estimator *E)
{
    const PlanData &d = E->plan->d;
    const Layout &L = *E->L;
    const int Nmax = d.g.Nmax, Nr = d.g.Nr, nseg = (int)L.sg.size() - 1;
    std::vector<int> ref((size_t)Nmax * Nr, 0);
    for (const Level &V : L.lv)
        for (const FieldTask &T : V.tasks) {
            bool any = false;
            for (int ia = 0; ia * ia < T.nhi && !any; ++ia)
                for (int ja = 0; ia * ia + ja * ja < T.nhi && !any; ++ja) {
                    const int n = ia * ia + ja * ja;
                    any = n >= T.nlo && (T.plane != 0 || n > 0);
                }
            if (any) ref[T.zid] = 1;
        }
    int rc;
    int *d_ref, *d_match;
    double *d_sums, *d_bps;
    if ((rc = dev_alloc(&d_ref, ref.size(), E->owned))) return rc;
    if ((rc = dev_alloc(&d_match, (size_t)E->batch, E->owned))) return rc;
    if ((rc = dev_alloc(&d_sums, (size_t)3 * d.nbins * BB_NPOLES, E->owned))) return rc;
    if ((rc = dev_alloc(&d_bps, (size_t)nseg * 16 * E->batch, E->owned))) return rc;
    if ((rc = upload(d_ref, ref.data(), sizeof(int) * ref.size(), E->stream))) return rc;
    const size_t nb = (size_t)d.nbins * BB_NPOLES;
    HIPCHK(launch_binning(E->d_sg, nseg, E->d_bin_sg, E->d_recs, E->d_part, E->d_num, d_ref,
                          E->d_avg, L.bp, E->d_binpart, d_sums, d_sums + nb, d_sums + 2 * nb,
                          E->bst, 1, E->stream));
    E->bst.binpart_s = (long long)nseg * 16;
    E->bstat.ref = d_ref;
    E->bstat.nflags = Nmax * Nr;
    E->bstat.sums = d_sums;
    E->bstat.match = d_match;
    E->bstat.binpart_s = d_bps;
    return BB_OK;
}

int estimator_init(bb_estimator *E)
{
    const PlanData &d = E->plan->d;
    const int Nmax = d.g.Nmax, Nr = d.g.Nr;
    int rc = get_layout(E->plan, E->cprec, E->opt, E->L);
    if (rc) return rc;
    const Layout &L = *E->L;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        (void)hipGetLastError();
        return set_error(BB_E_DEVICE, "no HIP device available");
    }
    HIPCHK(hipStreamCreateWithFlags(&E->own_stream, hipStreamNonBlocking));
    E->stream = E->own_stream;
    for (auto &e : E->ev) HIPCHK(hipEventCreate(&e));
    if ((rc = compute_num(E))) return rc;
    if ((rc = dev_alloc(&E->d_num, E->num_host.size(), E->owned))) return rc;
    if ((rc = upload(E->d_num, E->num_host.data(), sizeof(double) * E->num_host.size(), E->stream)))
        return rc;

    // Per level: FFT geometry and twiddles; the tasks of every level.
    const int nl = (int)L.lv.size();
    E->lvd.resize(nl);
    std::vector<FieldTask> tasks;
    std::vector<double2> tw_all;
    for (int l = 0; l < nl; ++l) {
        const Level &V = L.lv[l];
        if ((rc = level_desc(V, (int)tw_all.size(), E->cprec == CPrec::F64 ? 8 : 4, E->lvd[l])))
            return rc;
        const std::vector<double2> tw = twiddles(V.NG);
        tw_all.insert(tw_all.end(), tw.begin(), tw.end());
        tasks.insert(tasks.end(), V.tasks.begin(), V.tasks.end());
    }
    {
        std::vector<int2> blk;
        fft_blocks(tasks.data(), (int)tasks.size(), E->lvd.data(), blk);
        int2 *d;
        if ((rc = dev_alloc(&d, std::max<size_t>(blk.size(), 1), E->owned))) return rc;
        if (!blk.empty() && (rc = upload(d, blk.data(), sizeof(int2) * blk.size(), E->stream)))
            return rc;
        E->fb.units = d;
        E->fb.n = (int)blk.size();
    }
    if ((rc = dev_alloc(&E->d_tw, tw_all.size(), E->owned))) return rc;
    if ((rc = upload(E->d_tw, tw_all.data(), sizeof(double2) * tw_all.size(), E->stream))) return rc;
    if ((rc = dev_alloc(&E->d_lvd, nl, E->owned))) return rc;
    if ((rc = upload(E->d_lvd, E->lvd.data(), sizeof(LevelDesc) * nl, E->stream))) return rc;

    // --- Device buffers (one copy per field of a batch) ---------------------------
    const size_t F = (size_t)E->batch;
    const size_t top2 = (size_t)L.NGtop * L.NGtop;
    const size_t cpts1 = (size_t)std::max<long long>(L.cache_pts, 1);
    const size_t cpts = cpts1 * F;
    const size_t nslots1 = (size_t)std::max<long long>(L.part_tiles, 1) * 256;
    E->bst.planes = (long long)(Nmax * top2);
    E->bst.cache = (long long)cpts1;
    E->bst.nz = (long long)Nmax * Nr;
    E->bst.part = (long long)nslots1;
    E->bst.bins = (long long)d.nbins * BB_NPOLES;
    E->bst.binpart = (long long)(L.sg.size() - 1) * 48;
    if ((rc = dev_alloc(&E->d_planes, (size_t)Nmax * top2 * F, E->owned))) return rc;
    if (E->cprec != CPrec::F64) {
        float2 *c;
        if ((rc = dev_alloc(&c, cpts, E->owned))) return rc;
        E->d_cache = c;
        if (E->cprec == CPrec::BF16X3) {
            float2 *sp;   // 8 bytes per point, like the FP32 values
            if ((rc = dev_alloc(&sp, cpts, E->owned))) return rc;
            E->d_split = sp;
        }
    } else {
        double2 *c;
        if ((rc = dev_alloc(&c, cpts, E->owned))) return rc;
        E->d_cache = c;
    }
    const size_t nslots = nslots1 * F;
    if ((rc = dev_alloc(&E->d_tasks, tasks.size(), E->owned))) return rc;
    if ((rc = dev_alloc(&E->d_items, L.items.size(), E->owned))) return rc;
    if ((rc = dev_alloc(&E->d_tws, L.tws.size(), E->owned))) return rc;
    if ((rc = dev_alloc(&E->d_units, L.units.size(), E->owned))) return rc;
    if (!L.units_f.empty()) {
        if ((rc = dev_alloc(&E->d_units_f, L.units_f.size(), E->owned))) return rc;
        if ((rc = upload(E->d_units_f, L.units_f.data(), sizeof(WorkUnit) * L.units_f.size(),
                         E->stream)))
            return rc;
        HIPCHK(hipStreamCreateWithFlags(&E->aux, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&E->ev_fork, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&E->ev_join, hipEventDisableTiming));
    }
    if ((rc = dev_alloc(&E->d_part, nslots, E->owned))) return rc;
    if ((rc = dev_alloc(&E->d_recs, L.recs.size(), E->owned))) return rc;
    if ((rc = dev_alloc(&E->d_sg, L.sg.size(), E->owned))) return rc;
    if ((rc = dev_alloc(&E->d_bin_sg, L.bin_sg.size(), E->owned))) return rc;
    if ((rc = dev_alloc(&E->d_binpart, std::max<size_t>(L.sg.size() - 1, 1) * 48 * F, E->owned)))
        return rc;
    if ((rc = dev_alloc(&E->d_nz, (size_t)Nmax * Nr * F, E->owned))) return rc;
    if ((rc = dev_alloc(&E->d_avg, Nr, E->owned))) return rc;
    if ((rc = dev_alloc(&E->d_bins, (size_t)3 * d.nbins * BB_NPOLES * F, E->owned))) return rc;
    if (!tasks.empty() && (rc = upload(E->d_tasks, tasks.data(), sizeof(FieldTask) * tasks.size(), E->stream))) return rc;
    if (!L.items.empty() && (rc = upload(E->d_items, L.items.data(), sizeof(WorkItem) * L.items.size(), E->stream))) return rc;
    if (!L.tws.empty() && (rc = upload(E->d_tws, L.tws.data(), sizeof(TWin) * L.tws.size(), E->stream))) return rc;
    if (!L.units.empty() && (rc = upload(E->d_units, L.units.data(), sizeof(WorkUnit) * L.units.size(), E->stream))) return rc;
    if (!L.recs.empty() && (rc = upload(E->d_recs, L.recs.data(), sizeof(BinRec) * L.recs.size(), E->stream))) return rc;
    if ((rc = upload(E->d_sg, L.sg.data(), sizeof(int32_t) * L.sg.size(), E->stream))) return rc;
    if ((rc = upload(E->d_bin_sg, L.bin_sg.data(), sizeof(int32_t) * L.bin_sg.size(), E->stream)))
        return rc;
    if ((rc = upload(E->d_avg, d.avg.data(), sizeof(double) * Nr, E->stream))) return rc;
    if (bin_table_bytes(L, F) > 0) {
        if ((rc = dev_alloc(&E->d_bincoef, L.recs.size() * 33, E->owned))) return rc;
        HIPCHK(launch_bin_coeffs(E->d_recs, (int)L.recs.size(), E->d_avg, L.bp, E->d_bincoef,
                                 E->stream));
    }
    HIPCHK(hipMemsetAsync(E->d_part, 0, sizeof(double) * nslots, E->stream));
    // Padding points between fields are read by the contraction and never
    // written by the transforms: zero them (and everything else) once.
    HIPCHK(hipMemsetAsync(E->d_cache, 0, cpts * (E->cprec == CPrec::F64 ? 16 : 8), E->stream));
    if (E->d_split) HIPCHK(hipMemsetAsync(E->d_split, 0, cpts * 8, E->stream));
    if (L.sg.size() > 1 && (rc = static_bins(E))) return rc;
    HIPCHK(hipStreamSynchronize(E->stream));
    return BB_OK;
}

// The launch sequence of one batch: nonzero flags cleared by a fill kernel,
// plane gather, transforms, contraction, binning.  `timed` records the stage
// events between the launches (not inside a captured graph).
int enqueue_kernels(bb_estimator *E, int nf, const double *d_fields, double *d_bs,
                    double *d_norm, double *d_biana, bool timed)
{
    const PlanData &d = E->plan->d;
    const Layout &L = *E->L;
    hipStream_t s = E->stream;
    const int nz = d.g.Nmax * d.g.Nr * nf;
    // The gather also clears the nonzero flags that the ring pass raises and
    // sets the flag-match results that the binning clears.
    HIPCHK(launch_gather((const double2 *)d_fields, E->d_planes, L.NN, L.NGtop, L.kmax,
                         d.g.Nmax, nf, s, E->d_nz, nz, E->bstat.match));
    if (timed) HIPCHK(hipEventRecord(E->ev[1], s));
    int rc = run_fields(E, E->d_planes, L.NGtop, E->d_nz, nf);
    if (rc) return rc;
    if (timed) HIPCHK(hipEventRecord(E->ev[2], s));
    const bool mixed = !L.units_f.empty();
    if (mixed) {   // the FP32 items run concurrently on the auxiliary stream
        HIPCHK(hipEventRecord(E->ev_fork, s));
        HIPCHK(hipStreamWaitEvent(E->aux, E->ev_fork, 0));
        HIPCHK(launch_contract(CPrec::F32, L.waves, contract_tww(CPrec::F32), E->d_cache, nullptr,
                               E->d_items, E->d_units_f, (int)L.units_f.size(), E->d_tws,
                               E->d_part, E->bst, nf, E->aux));
        HIPCHK(hipEventRecord(E->ev_join, E->aux));
    }
    HIPCHK(launch_contract(E->cprec, L.waves, L.tww, E->d_cache, E->d_split, E->d_items,
                           E->d_units, (int)L.units.size(), E->d_tws, E->d_part, E->bst, nf, s));
    if (mixed) HIPCHK(hipStreamWaitEvent(s, E->ev_join, 0));
    if (timed) HIPCHK(hipEventRecord(E->ev[3], s));
    HIPCHK(launch_binning(E->d_sg, (int)L.sg.size() - 1, E->d_bin_sg, E->d_recs, E->d_part,
                          E->d_num, E->d_nz, E->d_avg, L.bp, E->d_binpart, d_bs, d_norm, d_biana,
                          E->bst, nf, s, E->d_bincoef, (int)L.recs.size(), E->bstat));
    return BB_OK;
}

// The stages before and after the contraction, on the estimator's stream
// (bb_estimate_mixed runs the contractions of several estimators in one
// launch between them).
int enqueue_pre(bb_estimator *E, int nf, const double *d_fields)
{
    const PlanData &d = E->plan->d;
    const Layout &L = *E->L;
    const int nz = d.g.Nmax * d.g.Nr * nf;
    HIPCHK(launch_gather((const double2 *)d_fields, E->d_planes, L.NN, L.NGtop, L.kmax,
                         d.g.Nmax, nf, E->stream, E->d_nz, nz, E->bstat.match));
    return run_fields(E, E->d_planes, L.NGtop, E->d_nz, nf);
}
int enqueue_post(bb_estimator *E, int nf, double *d_bs, double *d_norm, double *d_biana)
{
    const Layout &L = *E->L;
    HIPCHK(launch_binning(E->d_sg, (int)L.sg.size() - 1, E->d_bin_sg, E->d_recs, E->d_part,
                          E->d_num, E->d_nz, E->d_avg, L.bp, E->d_binpart, d_bs, d_norm, d_biana,
                          E->bst, nf, E->stream, E->d_bincoef, (int)L.recs.size(), E->bstat));
    return BB_OK;
}

int estimate_finish(bb_estimator *E);

// Several estimates of different geometries with their contractions in one
// persistent launch (bb_estimate_mixed).
int estimate_mixed(int n, bb_estimator *const *ests, const int *nfs, const double *const *fields,
                   double *const *bs, double *const *nm, double *const *bi)
{
    bb_estimator *E0 = ests[0];
    const Layout &L0 = *E0->L;
    for (int k = 0; k < n; ++k) {
        bb_estimator *E = ests[k];
        const Layout &L = *E->L;
        if (E->cprec != CPrec::BF16X3 || L.waves != 4 || L.tww != L0.tww || !L.units_f.empty())
            return set_error(BB_E_INVALID, "mixed estimates need BF16x3 estimators of one "
                                           "4-wave layout shape (same occupancy, no mix_fp32)");
        if (nfs[k] < 1 || nfs[k] > E->batch)
            return set_error(BB_E_RANGE, "estimator " + std::to_string(k) + ": batch of " +
                                             std::to_string(nfs[k]) + " fields");
        if (!fields[k]) return set_error(BB_E_INVALID, "null field pointer");
        for (int q = 0; q < k; ++q)
            if (ests[q] == E) return set_error(BB_E_INVALID, "an estimator appears twice");
        if (int rc = estimate_finish(E)) return rc;
    }
    // The unit list: every unit of every field of every estimate, largest
    // first (chunks x tile windows of the item's busiest wave).  It holds the
    // estimators' device addresses, so it is keyed by their creation numbers
    // (an estimator created at a destroyed one's address gets a new number).
    std::vector<unsigned long long> key;
    for (int k = 0; k < n; ++k) {
        key.push_back(ests[k]->serial);
        key.push_back((unsigned long long)nfs[k]);
    }
    if (key != E0->mix_key) {
        // Estimators in order of decreasing mean unit cost, each one's fields
        // in turn, each field's units in the layout's largest-first order:
        // the large units of every size come first (a short tail), and the
        // units of one field run together (its cache stays in the caches).
        auto unit_cost = [](const Layout &L, const WorkUnit &u) {
            const WorkItem &it = L.items[u.item];
            int w = 0;
            for (int q = 0; q < kWaves; ++q) w = std::max(w, (int)it.cnt[q]);
            return (double)(u.ce - u.cb) * w;
        };
        std::vector<std::pair<double, int>> order;
        for (int k = 0; k < n; ++k) {
            const Layout &L = *ests[k]->L;
            double c = 0;
            for (const WorkUnit &u : L.units) c += unit_cost(L, u);
            order.push_back({L.units.empty() ? 0.0 : c / (double)L.units.size(), k});
        }
        std::stable_sort(order.begin(), order.end(),
                         [](const auto &a, const auto &b) { return a.first > b.first; });
        std::vector<MixUnit> host;
        for (const auto &o : order) {
            bb_estimator *E = ests[o.second];
            const Layout &L = *E->L;
            for (int f = 0; f < nfs[o.second]; ++f)
                for (const WorkUnit &u : L.units) {
                    MixUnit M;
                    M.cache = reinterpret_cast<const float2 *>(E->d_cache) + (size_t)f * E->bst.cache;
                    M.split = reinterpret_cast<const __bf16 *>(E->d_split) + (size_t)f * E->bst.cache * 4;
                    M.item = E->d_items + u.item;
                    M.tws = E->d_tws;
                    M.out = E->d_part + (size_t)f * E->bst.part;
                    M.unit = u;
                    host.push_back(M);
                }
        }
        if (E0->mix_cap < host.size()) {
            MixUnit *p = nullptr;
            if (int rc = dev_alloc(&p, host.size(), E0->owned)) return rc;
            E0->d_mix = p;
            E0->mix_cap = host.size();
        }
        if (!E0->d_mix_counter)
            if (int rc = dev_alloc(&E0->d_mix_counter, 1, E0->owned)) return rc;
        if (int rc = upload(E0->d_mix, host.data(), sizeof(MixUnit) * host.size(), E0->stream))
            return rc;
        HIPCHK(hipStreamSynchronize(E0->stream));
        E0->mix_key = key;
        E0->mix_n = (int)host.size();
    }
    // Stages before the contraction on each estimator's stream; the fused
    // contraction on the first's once all are done; binning after it.
    std::vector<hipEvent_t> ev(n + 1, nullptr);
    for (auto &e : ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    int rc = BB_OK;
    for (int k = 0; k < n && !rc; ++k) {
        rc = enqueue_pre(ests[k], nfs[k], fields[k]);
        if (!rc && hipEventRecord(ev[k], ests[k]->stream) != hipSuccess)
            rc = set_error(BB_E_DEVICE, "event record failed");
    }
    for (int k = 0; k < n && !rc; ++k)
        if (hipStreamWaitEvent(E0->stream, ev[k], 0) != hipSuccess)
            rc = set_error(BB_E_DEVICE, "stream wait failed");
    if (!rc) {
        const hipError_t e = launch_contract_mixed(L0.tww, E0->d_mix, E0->mix_n, E0->d_mix_counter,
                                                   E0->stream);
        if (e != hipSuccess) rc = set_error(BB_E_DEVICE, hipGetErrorString(e));
    }
    if (!rc && hipEventRecord(ev[n], E0->stream) != hipSuccess)
        rc = set_error(BB_E_DEVICE, "event record failed");
    for (int k = 0; k < n && !rc; ++k) {
        if (hipStreamWaitEvent(ests[k]->stream, ev[n], 0) != hipSuccess)
            rc = set_error(BB_E_DEVICE, "stream wait failed");
        if (!rc) rc = enqueue_post(ests[k], nfs[k], bs[k], nm[k], bi[k]);
    }
    for (int k = 0; k < n; ++k)
        if (hipStreamSynchronize(ests[k]->stream) != hipSuccess && !rc)
            rc = set_error(BB_E_DEVICE, "mixed estimate failed");
    for (auto &e : ev) (void)hipEventDestroy(e);
    (void)hipGetLastError();
    return rc;
}

// Enqueues the estimates of nf fields (back to back from d_fields) on the
// estimator's stream; outputs back to back per field.  Nothing waits.  With
// use_graph, the first call for a set of arguments runs eagerly and captures
// the sequence into a graph; later calls with the same arguments replay it.
int estimate_enqueue(bb_estimator *E, int nf, const double *d_fields, double *d_bs,
                     double *d_norm, double *d_biana)
{
    if (nf < 1 || nf > E->batch)
        return set_error(BB_E_RANGE, "batch of " + std::to_string(nf) + " fields; the estimator "
                                         "holds " + std::to_string(E->batch));
    hipStream_t s = E->stream;
    const bool same = E->gexec && E->g_nf == nf && E->g_in == d_fields && E->g_out[0] == d_bs &&
                      E->g_out[1] == d_norm && E->g_out[2] == d_biana;
    HIPCHK(hipEventRecord(E->ev[0], s));
    if (E->opt.use_graph && same) {
        HIPCHK(hipGraphLaunch(E->gexec, s));
        E->graph_run = true;
    } else {
        int rc = enqueue_kernels(E, nf, d_fields, d_bs, d_norm, d_biana, true);
        if (rc) return rc;
        E->graph_run = false;
    }
    HIPCHK(hipEventRecord(E->ev[4], s));
    E->pending = true;
    if (E->opt.use_graph && !same) {
        if (E->gexec) {
            HIPCHK(hipGraphExecDestroy(E->gexec));
            E->gexec = nullptr;
        }
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
        const int rc = enqueue_kernels(E, nf, d_fields, d_bs, d_norm, d_biana, false);
        const hipError_t e = hipStreamEndCapture(s, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return set_error(BB_E_DEVICE, "graph capture failed");
        const hipError_t ei = hipGraphInstantiate(&E->gexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            E->gexec = nullptr;
            return set_error(BB_E_DEVICE, "graph instantiation failed");
        }
        E->g_nf = nf;
        E->g_in = d_fields;
        E->g_out[0] = d_bs;
        E->g_out[1] = d_norm;
        E->g_out[2] = d_biana;
    }
    return BB_OK;
}

// Waits for the enqueued batch and records its stage times ([4]: total; a
// replayed graph has no stage boundaries, its stages read 0).
int estimate_finish(bb_estimator *E)
{
    if (!E->pending) return BB_OK;
    E->pending = false;
    HIPCHK(hipStreamSynchronize(E->stream));
    HIPCHK(hipEventElapsedTime(&E->ms[4], E->ev[0], E->ev[4]));
    if (E->graph_run) {
        for (int k = 0; k < 4; ++k) E->ms[k] = 0.0f;
    } else {
        for (int k = 0; k < 4; ++k)
            HIPCHK(hipEventElapsedTime(&E->ms[k], E->ev[k], k == 3 ? E->ev[4] : E->ev[k + 1]));
    }
    return BB_OK;
}

int estimate_device(bb_estimator *E, int nf, const double *d_fields, double *d_bs,
                    double *d_norm, double *d_biana)
{
    int rc = estimate_enqueue(E, nf, d_fields, d_bs, d_norm, d_biana);
    if (rc) {
        (void)hipStreamSynchronize(E->stream);
        E->pending = false;
        return rc;
    }
    return estimate_finish(E);
}

// Scatters the compact modes (stage: per field [nr][nr][Nmax], rows and
// columns in the order of `idx`) into their places of the r2c buffers.
__global__ __launch_bounds__(256) void k_expand_modes(const double2 *__restrict__ stage,
                                                      double2 *__restrict__ fields,
                                                      const int *__restrict__ idx, int nr,
                                                      int Nmax, int NN, long long nper)
{
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t >= nper) return;
    const int f = blockIdx.y;
    const int p = (int)(t % Nmax);
    const long long ij = t / Nmax;
    const int cj = (int)(ij % nr), ci = (int)(ij / nr);
    const long long H = NN / 2 + 1;
    fields[(long long)f * NN * NN * H + ((long long)idx[ci] * NN + idx[cj]) * H + p] =
        stage[(long long)f * nper + t];
}

// Host fields: only the modes the estimate reads cross PCIe -- kz < Nmax and
// |kx|, |ky| <= kmax (the largest |k_x| of a ring mode the estimate uses;
// the plane gather zeroes or masks every other mode).  Host threads gather
// them into pinned memory, one copy moves them, a kernel puts them in their
// places of a device r2c buffer whose other entries stay zero.  Outputs are
// copied back.
int estimate_host(bb_estimator *E, int nf, const double *fields, double *bs, double *norm,
                  double *biana)
{
    const PlanData &d = E->plan->d;
    const size_t NN = d.g.NN, H = NN / 2 + 1, Nmax = d.g.Nmax;
    const size_t fsz = NN * H;                   // complex values per field ...
    const size_t fszf = fsz * NN;                // ... all of them
    if (nf < 1 || nf > E->batch)
        return set_error(BB_E_RANGE, "batch of " + std::to_string(nf) + " fields; the estimator "
                                         "holds " + std::to_string(E->batch));
    if (E->field_cap < nf) {
        double2 *p = nullptr;
        int rc = dev_alloc(&p, fszf * (size_t)nf, E->owned);
        if (rc) return rc;
        HIPCHK(hipMemsetAsync(p, 0, fszf * 16 * (size_t)nf, E->stream));
        E->d_field = p;   // an older, smaller staging stays owned until destroy
        E->field_cap = nf;
    }
    // rows (and columns) ia with |ia| <= kmax, in buffer order
    const int km = std::min((int)NN / 2, std::max(0, E->L->kmax));
    std::vector<int> idx;
    for (int i = 0; i < (int)NN; ++i) {
        const int ia = i <= (int)NN / 2 ? i : i - (int)NN;
        if (ia <= km && -ia <= km) idx.push_back(i);
    }
    const int nr = (int)idx.size();
    const size_t nper = (size_t)nr * nr * Nmax;
    if (E->stage_cap < nper * (size_t)nf) {
        double2 *h = nullptr, *dv = nullptr;
        if (hipHostMalloc(&h, nper * nf * sizeof(double2), hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            return set_error(BB_E_NOMEM, "pinned staging for host fields");
        }
        if (int rc = dev_alloc(&dv, nper * (size_t)nf + (size_t)nr / 2 + 1, E->owned)) {
            (void)hipHostFree(h);
            return rc;
        }
        HIPCHK(hipStreamSynchronize(E->stream));   // the old staging may still be read
        if (E->h_stage) (void)hipHostFree(E->h_stage);
        E->h_stage = h;
        E->d_stage = dv;
        E->stage_cap = nper * (size_t)nf;
    }
    HIPCHK(hipStreamSynchronize(E->stream));   // earlier uploads from the staging are done
    {
        const double2 *src = reinterpret_cast<const double2 *>(fields);
        double2 *dst = E->h_stage;
        const int rows = nf * nr;
        const int nth = std::max(1, std::min(8, rows / 16));
        auto work = [&](int t) {
            for (int q = t; q < rows; q += nth) {
                const int f = q / nr, ci = q % nr;
                for (int cj = 0; cj < nr; ++cj)
                    std::memcpy(dst + (size_t)f * nper + ((size_t)ci * nr + cj) * Nmax,
                                src + (size_t)f * fszf + ((size_t)idx[ci] * NN + idx[cj]) * H,
                                Nmax * sizeof(double2));
            }
        };
        std::vector<std::thread> pool;
        for (int t = 1; t < nth; ++t) pool.emplace_back(work, t);
        work(0);
        for (auto &th : pool) th.join();
    }
    int *d_idx = reinterpret_cast<int *>(E->d_stage + nper * (size_t)nf);
    HIPCHK(hipMemcpyAsync(E->d_stage, E->h_stage, nper * nf * sizeof(double2),
                          hipMemcpyHostToDevice, E->stream));
    HIPCHK(hipMemcpyAsync(d_idx, idx.data(), sizeof(int) * nr, hipMemcpyHostToDevice, E->stream));
    {
        dim3 grid((unsigned)((nper + 255) / 256), (unsigned)nf);
        hipLaunchKernelGGL(k_expand_modes, grid, dim3(256), 0, E->stream, E->d_stage, E->d_field,
                           d_idx, nr, (int)Nmax, (int)NN, (long long)nper);
        HIPCHK(hipGetLastError());
    }
    const size_t nb = (size_t)d.nbins * BB_NPOLES;
    double *B = E->d_bins;   // [bs | norm | biana], each nf * nb
    const size_t span = nb * (size_t)E->batch;
    int rc = estimate_device(E, nf, (const double *)E->d_field, B, B + span, B + 2 * span);
    if (rc) return rc;
    if (bs) HIPCHK(hipMemcpy(bs, B, nb * nf * sizeof(double), hipMemcpyDeviceToHost));
    if (norm) HIPCHK(hipMemcpy(norm, B + span, nb * nf * sizeof(double), hipMemcpyDeviceToHost));
    if (biana)
        HIPCHK(hipMemcpy(biana, B + 2 * span, nb * nf * sizeof(double), hipMemcpyDeviceToHost));
    return BB_OK;
}

int precision_of(bb_precision p, CPrec &cp)
{
    switch (p) {
    case BB_FP64: cp = CPrec::F64; return BB_OK;
    case BB_FP32: cp = CPrec::F32; return BB_OK;
    case BB_BF16X3: cp = CPrec::BF16X3; return BB_OK;
    default: return set_error(BB_E_INVALID, "unknown precision");
    }
}

int check_options(const bb_options &o)
{
    if (o.x_splits < 0 || o.x_splits > 1024)
        return set_error(BB_E_INVALID, "x_splits must be in [0, 1024]");
    if (o.grid_levels < 0 || o.grid_levels > BB_MAX_LEVELS)
        return set_error(BB_E_INVALID, "grid_levels must be in [0, BB_MAX_LEVELS]");
    if (o.max_batch < 0 || o.max_batch > 1024)
        return set_error(BB_E_INVALID, "max_batch must be in [0, 1024]");
    if (o.waves != 0 && o.waves != 4 && o.waves != 8)
        return set_error(BB_E_INVALID, "waves must be 0, 4 or 8");
    if (o.mix_fp32 < 0 || o.mix_fp32 > 90)
        return set_error(BB_E_INVALID, "mix_fp32 must be in [0, 90]");
    if (o.occupancy != 0 && o.occupancy != 2 && o.occupancy != 3)
        return set_error(BB_E_INVALID, "occupancy must be 0, 2 or 3");
    if (o.occupancy == 3 && (o.precision != BB_BF16X3 || layout_waves(o) != 4))
        return set_error(BB_E_INVALID, "occupancy 3 needs precision BB_BF16X3 and 4 waves");
    return BB_OK;
}

}  // namespace

namespace bb {
bool fft_desc(int N, FftDesc &fd) { return make_fft_desc(N, fd); }
std::vector<double2> fft_twiddles(int N) { return twiddles(N); }
}  // namespace bb

extern "C" {

int bb_estimator_footprint_opts(const bb_plan *p, const bb_options *opt, size_t *bytes)
{
    if (!p || !opt || !bytes) return set_error(BB_E_INVALID, "null argument");
    CPrec cp;
    int rc = precision_of(opt->precision, cp);
    if (!rc) rc = check_options(*opt);
    if (rc) return rc;
    std::shared_ptr<const Layout> L;
    try {
        rc = get_layout(p, cp, *opt, L);
    } catch (const std::bad_alloc &) {
        rc = set_error(BB_E_NOMEM, "out of host memory");
    }
    if (rc) return rc;
    *bytes = layout_bytes(p->d, *L, (size_t)std::max(1, opt->max_batch));
    return BB_OK;
}

int bb_plan_grid_levels(const bb_plan *p, const bb_options *opt, int *nlevels, int *grids,
                        int *clevel)
{
    if (!p || !opt || !nlevels) return set_error(BB_E_INVALID, "null argument");
    CPrec cp;
    int rc = precision_of(opt->precision, cp);
    if (!rc) rc = check_options(*opt);
    if (rc) return rc;
    std::shared_ptr<const Layout> L;
    try {
        rc = get_layout(p, cp, *opt, L);
    } catch (const std::bad_alloc &) {
        rc = set_error(BB_E_NOMEM, "out of host memory");
    }
    if (rc) return rc;
    *nlevels = (int)L->lv.size();
    if (grids)
        for (size_t l = 0; l < L->lv.size(); ++l) grids[l] = L->lv[l].NG;
    if (clevel)
        for (size_t c = 0; c < L->lvl.size(); ++c) clevel[c] = L->lvl[c];
    return BB_OK;
}

int bb_tune(const bb_plan *p, const bb_options *base, const double *d_field, int repeats,
            bb_options *best, double *best_ms)
{
    return bb::tune_layout(p, base, d_field, repeats, true, best, best_ms);
}

}  // extern "C"

// bb_tune; allow_mix = false keeps BF16x3 candidates to the BF16x3 kernel
// alone (mix_fp32 = 0), for callers whose streams are CU-masked.
int bb::tune_layout(const bb_plan *p, const bb_options *base, const double *d_field, int repeats,
                    bool allow_mix, bb_options *best, double *best_ms)
{
    if (!p || !base || !best) return set_error(BB_E_INVALID, "null argument");
    CPrec cp;
    int rc = precision_of(base->precision, cp);
    if (!rc) rc = check_options(*base);
    if (rc) return rc;
    repeats = std::max(1, std::min(repeats, 16));
    const PlanData &d = p->d;
    const size_t fsz = (size_t)d.g.NN * d.g.NN * (d.g.NN / 2 + 1);
    const size_t nb = (size_t)d.nbins * BB_NPOLES;
    // Timing does not depend on the field's values: without one, zeros.
    double2 *zf = nullptr;
    double *bins = nullptr;
    if (!d_field) {
        if (hipMalloc(&zf, fsz * sizeof(double2)) != hipSuccess || hipMemset(zf, 0, fsz * 16) != hipSuccess) {
            (void)hipGetLastError();
            if (zf) (void)hipFree(zf);
            return set_error(BB_E_DEVMEM, "tuning field");
        }
        d_field = (const double *)zf;
    }
    if (hipMalloc(&bins, 3 * nb * sizeof(double)) != hipSuccess) {
        (void)hipGetLastError();
        if (zf) (void)hipFree(zf);
        return set_error(BB_E_DEVMEM, "tuning bins");
    }
    double best_t = 1e300;
    bb_options win = *base;
    std::vector<int> levs, mixes{base->mix_fp32};
    for (int lev : (cp == CPrec::BF16X3 ? std::vector<int>{4, 6} : std::vector<int>{4, 5, 6, 8}))
        if (base->grid_levels <= 0 || lev <= base->grid_levels) levs.push_back(lev);
    if (levs.empty()) levs.push_back(base->grid_levels);
    if (cp == CPrec::BF16X3) mixes = allow_mix ? std::vector<int>{0, 30, 45} : std::vector<int>{0};
    // (waves, occupancy): occupancy 3 exists for BF16x3 with 4 waves only.
    std::vector<std::pair<int, int>> shapes{{8, 0}, {4, 0}};
    if (cp == CPrec::BF16X3) shapes.push_back({4, 3});
    for (auto shape : shapes)
        for (int lev : levs)
            for (int mix : mixes) {
            bb_options o = *base;
            o.waves = shape.first;
            o.occupancy = shape.second;
            o.grid_levels = lev;
            o.mix_fp32 = mix;
            o.max_batch = 1;
            bb_estimator *E = nullptr;
            if (bb_estimator_create_opts(p, &o, &E)) continue;   // e.g. out of memory: skip
            double t = 1e300;
            int ok = estimate_device(E, 1, d_field, bins, bins + nb, bins + 2 * nb) == BB_OK;
            for (int r = 0; ok && r < repeats; ++r) {
                const auto t0 = std::chrono::steady_clock::now();
                ok = estimate_device(E, 1, d_field, bins, bins + nb, bins + 2 * nb) == BB_OK;
                t = std::min(t, std::chrono::duration<double, std::milli>(
                                    std::chrono::steady_clock::now() - t0).count());
            }
            bb_estimator_destroy(E);
            if (ok && t < best_t) {
                best_t = t;
                win = o;
            }
        }
    (void)hipFree(bins);
    if (zf) (void)hipFree(zf);
    if (best_t >= 1e300) return set_error(BB_E_DEVICE, "no candidate configuration ran");
    win.max_batch = base->max_batch;
    *best = win;
    if (best_ms) *best_ms = best_t;
    return BB_OK;
}

extern "C" {

int bb_estimator_footprint(const bb_plan *p, bb_precision prec, size_t *bytes)
{
    bb_options o;
    std::memset(&o, 0, sizeof o);
    o.precision = prec;
    return bb_estimator_footprint_opts(p, &o, bytes);
}

int bb_estimator_create_opts(const bb_plan *p, const bb_options *opt, bb_estimator **out)
{
    if (!p || !opt || !out) return set_error(BB_E_INVALID, "null argument");
    *out = nullptr;
    CPrec cp;
    int rc = precision_of(opt->precision, cp);
    if (!rc) rc = check_options(*opt);
    if (rc) return rc;
    bb_estimator *E = new (std::nothrow) bb_estimator;
    if (!E) return set_error(BB_E_NOMEM, "out of host memory");
    static std::atomic<unsigned long long> serials{0};
    E->serial = ++serials;
    E->plan = p;
    E->opt = *opt;
    E->cprec = cp;
    E->batch = std::max(1, opt->max_batch);
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
    *NG = e->L->NGtop;
    return BB_OK;
}

int bb_estimator_levels(const bb_estimator *e, int *nlevels, int *grids)
{
    if (!e || !nlevels) return set_error(BB_E_INVALID, "null argument");
    *nlevels = (int)e->L->lv.size();
    if (grids)
        for (size_t l = 0; l < e->L->lv.size() && l < BB_MAX_LEVELS; ++l) grids[l] = e->L->lv[l].NG;
    return BB_OK;
}

int bb_estimate_device(bb_estimator *e, const double *d_field, double *d_bs, double *d_norm,
                       double *d_biana)
{
    if (!e || !d_field) return set_error(BB_E_INVALID, "null argument");
    int rc = estimate_finish(e);
    if (rc) return rc;
    return estimate_device(e, 1, d_field, d_bs, d_norm, d_biana);
}

int bb_estimate(bb_estimator *e, const double *field, double *bs, double *norm, double *biana)
{
    if (!e || !field) return set_error(BB_E_INVALID, "null argument");
    int rc = estimate_finish(e);
    if (rc) return rc;
    return estimate_host(e, 1, field, bs, norm, biana);
}

int bb_estimate_batch(bb_estimator *e, int nfields, const double *fields, double *bs,
                      double *norm, double *biana)
{
    if (!e || !fields) return set_error(BB_E_INVALID, "null argument");
    int rc = estimate_finish(e);
    if (rc) return rc;
    return estimate_host(e, nfields, fields, bs, norm, biana);
}

int bb_estimate_batch_device(bb_estimator *e, int nfields, const double *d_fields, double *d_bs,
                             double *d_norm, double *d_biana)
{
    if (!e || !d_fields) return set_error(BB_E_INVALID, "null argument");
    int rc = estimate_finish(e);
    if (rc) return rc;
    return estimate_device(e, nfields, d_fields, d_bs, d_norm, d_biana);
}

int bb_estimate_batch_async(bb_estimator *e, int nfields, const double *d_fields, double *d_bs,
                            double *d_norm, double *d_biana)
{
    if (!e || !d_fields) return set_error(BB_E_INVALID, "null argument");
    // A pending batch is not waited for: the new one runs after it on the
    // estimator's stream (the stage events then describe the newest).
    const int rc = estimate_enqueue(e, nfields, d_fields, d_bs, d_norm, d_biana);
    if (rc) {
        (void)hipStreamSynchronize(e->stream);
        e->pending = false;
    }
    return rc;
}

int bb_estimator_query(bb_estimator *e, int *done)
{
    if (!e || !done) return set_error(BB_E_INVALID, "null argument");
    if (!e->pending) {
        *done = 1;
        return BB_OK;
    }
    hipError_t q = hipStreamQuery(e->stream);
    if (q == hipErrorNotReady) {
        *done = 0;
        return BB_OK;
    }
    if (q != hipSuccess) return set_error(BB_E_DEVICE, hipGetErrorString(q));
    *done = 1;
    return estimate_finish(e);
}

int bb_estimator_wait(bb_estimator *e)
{
    if (!e) return set_error(BB_E_INVALID, "null argument");
    return estimate_finish(e);
}

int bb_estimator_stream(const bb_estimator *e, void **stream)
{
    if (!e || !stream) return set_error(BB_E_INVALID, "null argument");
    *stream = (void *)e->stream;
    return BB_OK;
}

int bb_estimator_set_stream(bb_estimator *e, void *stream)
{
    if (!e) return set_error(BB_E_INVALID, "null argument");
    if (int rc = estimate_finish(e)) return rc;
    e->stream = stream ? (hipStream_t)stream : e->own_stream;
    return BB_OK;
}

int bb_estimator_contraction_sums(bb_estimator *e, double *S)
{
    if (!e || !S) return set_error(BB_E_INVALID, "null argument");
    if (int rc = estimate_finish(e)) return rc;
    const Layout &L = *e->L;
    std::vector<double> part((size_t)std::max<long long>(L.part_tiles, 1) * 256);
    HIPCHK(hipMemcpy(part.data(), e->d_part, sizeof(double) * part.size(), hipMemcpyDeviceToHost));
    const size_t n = e->plan->d.contr.size();
    for (size_t c = 0; c < n; ++c) {
        double v = 0.0;
        for (int s = 0; s < L.nsp[c]; ++s) v += part[(size_t)L.pos[c] + 256 * (size_t)s];
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

int bb_estimator_field(bb_estimator *e, int level, int plane, int ring, double *out)
{
    if (!e || !out) return set_error(BB_E_INVALID, "null argument");
    if (int rc = estimate_finish(e)) return rc;
    const Layout &L = *e->L;
    if (level < 0 || level >= (int)L.lv.size()) return set_error(BB_E_INVALID, "no such level");
    const Level &V = L.lv[level];
    const int f = V.fid(plane, ring);
    if (f < 0) return set_error(BB_E_INVALID, "field not held at this level (unused by the plan)");
    const size_t n = (size_t)V.npts, N = (size_t)V.NG;
    const size_t at = (size_t)V.off + (size_t)f * (size_t)V.stride;
    // the cache holds point (i, j) at j * N + i; out is row-major
    if (e->cprec == CPrec::F64) {
        std::vector<double2> tmp(n);
        HIPCHK(hipMemcpy(tmp.data(), (const double2 *)e->d_cache + at, n * sizeof(double2),
                         hipMemcpyDeviceToHost));
        for (size_t i = 0; i < N; ++i)
            for (size_t j = 0; j < N; ++j) {
                out[2 * (i * N + j)] = tmp[j * N + i].x;
                out[2 * (i * N + j) + 1] = tmp[j * N + i].y;
            }
    } else {
        std::vector<float2> tmp(n);
        HIPCHK(hipMemcpy(tmp.data(), (const float2 *)e->d_cache + at, n * sizeof(float2),
                         hipMemcpyDeviceToHost));
        for (size_t i = 0; i < N; ++i)
            for (size_t j = 0; j < N; ++j) {
                out[2 * (i * N + j)] = tmp[j * N + i].x;
                out[2 * (i * N + j) + 1] = tmp[j * N + i].y;
            }
    }
    return BB_OK;
}

int bb_export_cache(bb_estimator *e, const double *field_r2c, const char *dir, int NRi)
{
    if (!e || !field_r2c || !dir) return set_error(BB_E_INVALID, "null argument");
    if (int rc = estimate_finish(e)) return rc;
    return export_cache(e, field_r2c, dir, NRi);
}

int bb_estimate_mixed(int n, bb_estimator *const *ests, const int *nfields,
                      const double *const *d_fields, double *const *d_bs, double *const *d_norm,
                      double *const *d_biana)
{
    if (n < 1 || !ests || !nfields || !d_fields || !d_bs || !d_norm || !d_biana)
        return set_error(BB_E_INVALID, "null argument");
    for (int k = 0; k < n; ++k)
        if (!ests[k] || !d_bs[k] || !d_norm[k] || !d_biana[k])
            return set_error(BB_E_INVALID, "null argument");
    return estimate_mixed(n, ests, nfields, d_fields, d_bs, d_norm, d_biana);
}

int bb_import_rings(bb_estimator *e, const char *dir, int NRi, double rel_tol, int use_num,
                    long long *nmismatch)
{
    if (!e || !dir) return set_error(BB_E_INVALID, "null argument");
    if (!(rel_tol >= 0)) return set_error(BB_E_INVALID, "rel_tol must be >= 0");
    if (nmismatch) *nmismatch = 0;
    if (int rc = estimate_finish(e)) return rc;
    return import_rings(e, dir, NRi, rel_tol, use_num, nmismatch);
}

int bb_check_plane_cache(bb_estimator *e, const double *field_r2c, const char *dir, int NRi,
                         int stride, double *max_rel)
{
    if (!e || !field_r2c || !dir || !max_rel) return set_error(BB_E_INVALID, "null argument");
    if (stride < 1) return set_error(BB_E_INVALID, "stride must be >= 1");
    if (int rc = estimate_finish(e)) return rc;
    const PlanData &d = e->plan->d;
    const size_t n = (size_t)2 * d.g.NN * d.g.NN;
    std::vector<double> disk(n);
    double worst = 0.0;
    const int rc = visit_nn_fields(e, field_r2c, stride, [&](int p, int r, const double *data) {
        const std::string fn = cache_name(d, dir, p, r, NRi);
        FILE *f = std::fopen(fn.c_str(), "rb");
        if (!f) return set_error(BB_E_IO, "cannot open " + fn);
        const size_t got = std::fread(disk.data(), sizeof(double), n, f);
        std::fclose(f);
        if (got != n) return set_error(BB_E_IO, fn + " is short");
        double scale = 0.0, diff = 0.0;
        for (size_t k = 0; k < n; ++k) {
            scale = std::max(scale, std::fabs(data[k]));
            diff = std::max(diff, std::fabs(disk[k] - data[k]));
        }
        const double rel = scale > 0 ? diff / scale : (diff > 0 ? INFINITY : 0.0);
        worst = std::max(worst, rel);
        return (int)BB_OK;
    });
    *max_rel = worst;
    return rc;
}

static void worklist_of(const Layout &L, long long *o)
{
    o[0] = (long long)L.items.size();
    o[1] = (long long)(L.units.size() + L.units_f.size());
    o[2] = L.ntiles;
    o[3] = L.max_rows;
    o[4] = L.staged_rows;
    o[5] = L.staged_bytes;
    o[6] = L.NGtop;
    o[7] = L.chunk;
    o[8] = L.waves;
    o[9] = (long long)std::llround(L.wave_balance * 1000.0);
    o[10] = (long long)L.lv.size();
    o[11] = L.part_tiles;
    o[12] = L.useful_pts;
    o[13] = (long long)L.items.size() * L.waves * L.tww;
    o[14] = L.total_cost;
    o[15] = L.cache_pts;
    long long band_rows = 0;
    for (const Level &V : L.lv)
        for (const FieldTask &T : V.tasks) band_rows += std::min(V.NG, 2 * T.rowband + 1);
    o[16] = band_rows;
    // Pair products formed (16 rows per window that starts a block), the
    // rows of them that hold a pair, and tile windows, each times the points
    // of the item's grid.
    long long formed = 0, used = 0, wpts = 0;
    for (const WorkItem &it : L.items)
        for (int w = 0; w < L.waves; ++w)
            for (int j = 0; j < it.cnt[w]; ++j) {
                const TWin &T = L.tws[(size_t)it.tw0 + w * kTWW + j];
                wpts += it.npts;
                if (!(T.flags & TW_NEWA)) continue;
                formed += 16 * it.npts;
                for (int m = 0; m < 16; ++m) used += T.a1[m] >= 0 ? it.npts : 0;
            }
    o[17] = formed;
    o[18] = used;
    o[19] = wpts;
}

int bb_estimator_worklist(const bb_estimator *e, long long *o)
{
    if (!e || !o) return set_error(BB_E_INVALID, "null argument");
    worklist_of(*e->L, o);
    return BB_OK;
}

int bb_plan_worklist(const bb_plan *p, const bb_options *opt, long long *o)
{
    if (!p || !opt || !o) return set_error(BB_E_INVALID, "null argument");
    CPrec cp;
    int rc = precision_of(opt->precision, cp);
    if (!rc) rc = check_options(*opt);
    if (rc) return rc;
    std::shared_ptr<const Layout> L;
    try {
        rc = get_layout(p, cp, *opt, L);
    } catch (const std::bad_alloc &) {
        rc = set_error(BB_E_NOMEM, "out of host memory");
    }
    if (rc) return rc;
    worklist_of(*L, o);
    return BB_OK;
}

int bb_estimator_timings(bb_estimator *e, double *ms5)
{
    if (!e || !ms5) return set_error(BB_E_INVALID, "null argument");
    if (int rc = estimate_finish(e)) return rc;
    for (int k = 0; k < 5; ++k) ms5[k] = e->ms[k];
    return BB_OK;
}

}  // extern "C"
