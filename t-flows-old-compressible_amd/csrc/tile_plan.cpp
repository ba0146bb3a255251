// tile_plan.cpp -- host construction of the tiles of the tiled one-launch
// DIC sweeps (see tile_plan.h).
#include "tile_plan.h"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <thread>
#include <utility>

namespace tflows {

int sorted_pattern_levels(int n, const int* rowptr, const int* col, const int* dia, bool lower,
                          std::vector<int>& lev) {
    lev.assign(n, 0);
    int nlev = 0;
    if (lower) {
        for (int r = 0; r < n; ++r) {
            int l = 0;
            for (int j = rowptr[r]; j < dia[r]; ++j) {
                const int c = col[j];
                if (c < 0 || c >= r) return -1;
                l = std::max(l, lev[c] + 1);
            }
            lev[r] = l;
            nlev = std::max(nlev, l + 1);
        }
    } else {
        for (int r = n - 1; r >= 0; --r) {
            int l = 0;
            for (int j = dia[r] + 1; j < rowptr[r + 1]; ++j) {
                const int c = col[j];
                if (c <= r || c >= n) return -1;
                l = std::max(l, lev[c] + 1);
            }
            lev[r] = l;
            nlev = std::max(nlev, l + 1);
        }
    }
    return nlev;
}

namespace {

// f(lo, hi) over [0, n) in contiguous chunks on up to 16 host threads
template <class F>
void parallel_for(int n, F&& f) {
    const int hw = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const int nt = n < (1 << 16) ? 1 : hw;
    if (nt == 1) {
        f(0, n);
        return;
    }
    std::vector<std::thread> th;
    const int chunk = (n + nt - 1) / nt;
    for (int t = 0; t < nt; ++t) {
        const int lo = t * chunk, hi = std::min(n, lo + chunk);
        if (lo < hi) th.emplace_back([&f, lo, hi] { f(lo, hi); });
    }
    for (auto& x : th) x.join();
}

// median of v (v is reordered); 0 when empty
int64_t median(std::vector<int64_t>& v) {
    if (v.empty()) return 0;
    auto mid = v.begin() + v.size() / 2;
    std::nth_element(v.begin(), mid, v.end());
    return *mid;
}

// Stable LSD radix sort of idx by key[idx] (16-bit digits over the bits in use).
void radix_sort(std::vector<int>& idx, const std::vector<uint64_t>& key, uint64_t kmax) {
    int bits = 0;
    while (bits < 64 && (kmax >> bits) != 0) ++bits;
    std::vector<int> tmp(idx.size());
    std::vector<size_t> cnt(1 << 16);
    for (int shift = 0; shift < bits; shift += 16) {
        std::fill(cnt.begin(), cnt.end(), 0);
        for (int i : idx) cnt[(key[i] >> shift) & 0xFFFF]++;
        size_t s = 0;
        for (auto& c : cnt) {
            const size_t t = c;
            c = s;
            s += t;
        }
        for (int i : idx) tmp[cnt[(key[i] >> shift) & 0xFFFF]++] = i;
        idx.swap(tmp);
    }
}

// One attempt with the key (c1, c2, c3); false (why set) when a dependency
// breaks the order.
bool attempt(int n, const int* rowptr, const int* dia, const int* flev, const int* iperm, const int* rp,
             const int* cl, int64_t R1, int64_t S, int64_t R2, TilePlan& out) {
    const int W = kTilePhases;
    const uint64_t n2 = (uint64_t)((n - 1) / R1 + 1), n3 = (uint64_t)((S - 1) / R2 + 1);
    std::vector<uint64_t> key(n);
    parallel_for(n, [&](int lo, int hi) {
        for (int r = lo; r < hi; ++r) {
            const uint64_t c1 = (uint64_t)(flev[r] / W), c2 = (uint64_t)(r / R1), c3 = (uint64_t)((r % S) / R2);
            const uint64_t ph = (uint64_t)(flev[r] % W);
            key[r] = ((c1 * n2 + c2) * n3 + c3) * W + ph;
        }
    });
    uint64_t kmax = 0;
    for (int r = 0; r < n; ++r) kmax = std::max(kmax, key[r]);
    std::vector<int> rows(n);
    for (int r = 0; r < n; ++r) rows[r] = r;
    radix_sort(rows, key, kmax);

    // tiles: a new one where the (c1, c2, c3) group changes or the current
    // phase already holds kTileWave rows
    out.order.assign(n, 0);
    std::vector<int> tstart, phstart;   // per tile: first position; all phase starts
    std::vector<int> tfirst_ph;         // per tile: index of its first phase in phstart
    uint64_t gcur = ~0ull;
    int phcur = -1, phcount = 0;
    for (int k = 0; k < n; ++k) {
        const int r = rows[k];
        const uint64_t g = key[r] / W;
        const int ph = (int)(key[r] % W);
        bool new_tile = g != gcur;
        if (!new_tile && ph == phcur && phcount == kTileWave) new_tile = true;
        if (new_tile) {
            tstart.push_back(k);
            tfirst_ph.push_back((int)phstart.size());
            phstart.push_back(k);
            gcur = g;
            phcur = ph;
            phcount = 0;
        } else if (ph != phcur) {
            phstart.push_back(k);
            phcur = ph;
            phcount = 0;
        }
        ++phcount;
        out.order[k] = iperm ? iperm[r] : r;
    }
    const int nt = (int)tstart.size();
    tfirst_ph.push_back((int)phstart.size());
    tstart.push_back(n);
    out.ntiles = nt;
    // columns: the group key without c1, numbered densely in order of first use
    {
        out.tcol.assign(nt, 0);
        std::vector<std::pair<uint64_t, int>> cols(nt);
        for (int t = 0; t < nt; ++t) cols[t] = {(key[rows[tstart[t]]] / W) % (n2 * n3), t};
        std::sort(cols.begin(), cols.end());
        int id = -1;
        for (int i = 0; i < nt; ++i) {
            if (i == 0 || cols[i].first != cols[i - 1].first) ++id;
            out.tcol[cols[i].second] = id;
        }
        out.ncols = id + 1;
    }
    out.tph.assign((size_t)nt * (W + 1), 0);
    // tile / phase bounds of every position
    std::vector<int> pos(n), ktile(n), kph_lo(n), kph_hi(n);
    for (int t = 0; t < nt; ++t) {
        const int q0 = tfirst_ph[t], q1 = tfirst_ph[t + 1];
        if (q1 - q0 > W) {
            out.why = "a tile with more phases than kTilePhases";
            return false;
        }
        for (int q = 0; q <= W; ++q)
            out.tph[(size_t)t * (W + 1) + q] = q0 + q < q1 ? phstart[q0 + q] : tstart[t + 1];
        for (int q = q0; q < q1; ++q) {
            const int a = phstart[q], b = q + 1 < q1 ? phstart[q + 1] : tstart[t + 1];
            if (b - a > kTileWave) {
                out.why = "a phase with more than kTileWave rows";
                return false;
            }
            for (int k = a; k < b; ++k) {
                ktile[k] = t;
                kph_lo[k] = a;
                kph_hi[k] = b;
            }
        }
    }
    for (int k = 0; k < n; ++k) pos[out.order[k]] = k;

    const int E = out.slots;
    out.dmap.assign(n, 0);
    out.lcol.assign((size_t)E * n, kTileNone);
    out.lmap.assign((size_t)E * n, -1);
    out.ucol.assign((size_t)E * n, kTileNone);
    out.umap.assign((size_t)E * n, -1);
    // codes of every entry; a dependency out of order fails the attempt
    static const char* const kWhy[] = {nullptr, "a lower dependency inside the tile but not in an earlier phase",
                                       "a lower dependency in a later tile",
                                       "an upper dependency inside the tile but not in a later phase",
                                       "an upper dependency in an earlier tile"};
    std::atomic<int> fail{0};
    parallel_for(n, [&](int lo, int hi) {
        for (int k = lo; k < hi && fail.load(std::memory_order_relaxed) == 0; ++k) {
            const int r = rows[k];
            const int i = out.order[k];
            const int ts = tstart[ktile[k]], te = tstart[ktile[k] + 1];
            const int nl = dia[r] - rowptr[r], nu = rowptr[r + 1] - dia[r] - 1;
            const int j0 = rp[i];               // the internal row keeps the caller's entry order
            out.dmap[k] = j0 + nl;
            for (int s = 0; s < nl; ++s) {
                const int j = j0 + s;
                const int c = pos[cl[j]];
                int code;
                if (c >= ts && c < te) {
                    if (c >= kph_lo[k]) {
                        fail = 1;
                        return;
                    }
                    code = -(c - ts) - 1;
                } else {
                    if (c >= ts) {
                        fail = 2;
                        return;
                    }
                    code = c;
                }
                out.lcol[(size_t)s * n + k] = code;
                out.lmap[(size_t)s * n + k] = j;
            }
            for (int s = 0; s < nu; ++s) {
                const int j = j0 + nl + 1 + s;
                const int c = pos[cl[j]];
                int code;
                if (c >= ts && c < te) {
                    if (c < kph_hi[k]) {
                        fail = 3;
                        return;
                    }
                    code = -(c - ts) - 1;
                } else {
                    if (c < te) {
                        fail = 4;
                        return;
                    }
                    code = c;
                }
                out.ucol[(size_t)s * n + k] = code;
                out.umap[(size_t)s * n + k] = j;
            }
        }
    });
    if (fail.load()) {
        out.why = kWhy[fail.load()];
        return false;
    }
    return true;
}

}  // namespace

bool build_tile_plan(int n, const int* rowptr, const int* col, const int* dia, const int* flev, const int* iperm,
                     const int* rp, const int* cl, TilePlan& out) {
    out = TilePlan{};
    out.n = n;
    if (n <= 0) {
        out.why = "no rows";
        return false;
    }
    int el = 0, eu = 0;
    std::vector<int64_t> hi, lo;
    hi.reserve(n);
    lo.reserve(n);
    for (int r = 0; r < n; ++r) {
        const int a = rowptr[r], d = dia[r], e = rowptr[r + 1];
        if (d < a || d >= e) {
            out.why = "a row without its diagonal";
            return false;
        }
        el = std::max(el, d - a);
        eu = std::max(eu, e - d - 1);
        for (int j = a; j < d; ++j)
            if (col[j] >= r || col[j] < 0) {
                out.why = "a lower column not below its row";
                return false;
            }
        for (int j = d + 1; j < e; ++j)
            if (col[j] <= r || col[j] >= n) {
                out.why = "an upper column not above its row";
                return false;
            }
        if (d - a >= 1) hi.push_back((int64_t)r - col[a]);
        if (d - a >= 3) lo.push_back((int64_t)r - col[d - 2]);
    }
    if (el > kTileEntries || eu > kTileEntries) {
        out.why = "a row with more than kTileEntries entries in a triangle";
        return false;
    }
    out.slots = std::max(el, eu) <= 3 ? 3 : 4;
    const int64_t o3 = std::max<int64_t>(1, median(hi)), o2 = median(lo);
    out.scale_hi = o3;
    out.scale_lo = o2;
    const int64_t T = 8;   // planes and lines per tile of a box
    const bool three = o2 > 1 && o2 * T <= o3;
    const int64_t R1 = three ? T * o3 : (int64_t)kTileWave * o3;
    const int64_t R2 = three ? T * o2 : R1;
    // the box-shaped key: c3 from the row's place within its plane
    if (three && attempt(n, rowptr, dia, flev, iperm, rp, cl, R1, o3, R2, out)) {
        out.box_key = true;
        return true;
    }
    // the key whose order holds for every sorted pattern
    const std::string first = out.why;
    if (attempt(n, rowptr, dia, flev, iperm, rp, cl, R1, R1, R2, out)) return true;
    if (!first.empty()) out.why = first + "; " + out.why;
    out.order.clear();
    out.tph.clear();
    out.tcol.clear();
    out.dmap.clear();
    out.lcol.clear();
    out.lmap.clear();
    out.ucol.clear();
    out.umap.clear();
    return false;
}

void tile_schedule(const TilePlan& plan, int waves, std::vector<int>& wstart, std::vector<int>& wtiles) {
    waves = std::max(1, waves);
    wstart.assign(waves + 1, 0);
    for (int t = 0; t < plan.ntiles; ++t) wstart[plan.tcol[t] % waves + 1]++;
    for (int w = 0; w < waves; ++w) wstart[w + 1] += wstart[w];
    wtiles.assign(plan.ntiles, 0);
    std::vector<int> fill(wstart.begin(), wstart.end() - 1);
    for (int t = 0; t < plan.ntiles; ++t) wtiles[fill[plan.tcol[t] % waves]++] = t;
}

}  // namespace tflows
