// Host check of the DIC tile plan (t-flows-old-compressible_amd/csrc/tile_plan.cpp).
//
// For a set of generated patterns it builds the plan and then runs, on the
// CPU and in the kernels' order, what the tiled sweeps do on the device:
// tiles in order (backward: in reverse), phases in order, a tile's own
// values from its "LDS", other tiles' values from their slots.  It checks
// that every code and map is in range, that every dependency read is
// already written when read (an earlier tile, or an earlier phase of the
// own tile), that each wavefront schedule covers every tile once in
// increasing order, and that the pivots (Prec1.f90:51-59) and both sweeps
// (CG.f90:135-156) equal the sequential row-order sweeps bit for bit.
//
// Prints one line per case: "case <name> plan=<0|1> box=<0|1> tiles=<n>
// splits=<n> ok" or "FAIL ..."; exits non-zero on a failure.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "tile_plan.h"

using namespace tflows;

namespace {

struct Pattern {
    int n = 0;
    std::vector<int> rowptr, col, dia;
};

// each row's columns ascending, the diagonal among them
Pattern from_rows(int n, std::vector<std::vector<int>>& nb) {
    Pattern p;
    p.n = n;
    p.rowptr.assign(n + 1, 0);
    p.dia.assign(n, 0);
    for (int r = 0; r < n; ++r) {
        nb[r].push_back(r);
        std::sort(nb[r].begin(), nb[r].end());
        p.rowptr[r + 1] = p.rowptr[r] + (int)nb[r].size();
        for (int c : nb[r]) {
            if (c == r) p.dia[r] = (int)p.col.size();
            p.col.push_back(c);
        }
    }
    return p;
}

Pattern box(int nx, int ny, int nz, const std::vector<int>* number = nullptr) {
    const int n = nx * ny * nz;
    std::vector<std::vector<int>> nb(n);
    auto id = [&](int i, int j, int k) {
        const int r = i + nx * (j + ny * k);
        return number ? (*number)[r] : r;
    };
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const int r = id(i, j, k);
                if (i > 0) nb[r].push_back(id(i - 1, j, k));
                if (i < nx - 1) nb[r].push_back(id(i + 1, j, k));
                if (j > 0) nb[r].push_back(id(i, j - 1, k));
                if (j < ny - 1) nb[r].push_back(id(i, j + 1, k));
                if (k > 0) nb[r].push_back(id(i, j, k - 1));
                if (k < nz - 1) nb[r].push_back(id(i, j, k + 1));
            }
    return from_rows(n, nb);
}

// structurally symmetric, at most 3 lower and 3 upper entries per row, each
// lower neighbour within `reach` rows: many rows of one level fall into one
// tile key, so phases of more than kTileWave rows are split
Pattern random_sorted(int n, int reach, unsigned seed) {
    std::mt19937 g(seed);
    std::vector<std::vector<int>> nb(n);
    std::vector<int> up(n, 0);
    for (int r = 1; r < n; ++r) {
        int low = 0;
        for (int t = 0; t < 8 && low < 3; ++t) {
            const int c = std::max(0, r - 1 - (int)(g() % (unsigned)reach));
            if (c >= r || up[c] >= 3 || std::find(nb[r].begin(), nb[r].end(), c) != nb[r].end()) continue;
            nb[r].push_back(c);
            nb[c].push_back(r);
            ++up[c];
            ++low;
        }
    }
    return from_rows(n, nb);
}

int fails = 0;
#define CHECK(c, ...)                              \
    do {                                           \
        if (!(c)) {                                \
            std::printf("FAIL %s: ", name.c_str()); \
            std::printf(__VA_ARGS__);              \
            std::printf("\n");                     \
            ++fails;                               \
            return;                                \
        }                                          \
    } while (0)

void run(const std::string& name, const Pattern& pc, bool permute, int expect_plan, int expect_box,
         bool expect_splits) {
    const int n = pc.n;
    const int nnz = (int)pc.col.size();
    std::mt19937_64 g(7);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    std::vector<double> val(nnz);
    for (int r = 0; r < n; ++r)
        for (int j = pc.rowptr[r]; j < pc.rowptr[r + 1]; ++j) val[j] = j == pc.dia[r] ? 8.0 + U(g) : -0.5 + 0.3 * U(g);
    // internal numbering: rows permuted, each row's entries in the caller's order
    std::vector<int> iperm(n), perm(n);
    for (int i = 0; i < n; ++i) iperm[i] = i;
    if (permute) std::shuffle(iperm.begin(), iperm.end(), g);
    for (int r = 0; r < n; ++r) perm[iperm[r]] = r;
    std::vector<int> rp(n + 1, 0), cl(nnz), diai(n);
    std::vector<double> vi(nnz);
    for (int i = 0; i < n; ++i) rp[i + 1] = rp[i] + pc.rowptr[perm[i] + 1] - pc.rowptr[perm[i]];
    for (int i = 0; i < n; ++i) {
        const int r = perm[i];
        for (int q = 0; q < pc.rowptr[r + 1] - pc.rowptr[r]; ++q) {
            cl[rp[i] + q] = iperm[pc.col[pc.rowptr[r] + q]];
            vi[rp[i] + q] = val[pc.rowptr[r] + q];
        }
        diai[i] = rp[i] + (pc.dia[r] - pc.rowptr[r]);
    }
    std::vector<int> flev;
    const int nl = sorted_pattern_levels(n, pc.rowptr.data(), pc.col.data(), pc.dia.data(), true, flev);
    TilePlan P;
    const bool ok = nl >= 0 && build_tile_plan(n, pc.rowptr.data(), pc.col.data(), pc.dia.data(), flev.data(),
                                               permute ? iperm.data() : nullptr, rp.data(), cl.data(), P);
    if (!ok) {
        CHECK(expect_plan == 0, "no plan: %s", P.why.c_str());
        std::printf("case %s plan=0 why=\"%s\" ok\n", name.c_str(), nl < 0 ? "unsorted pattern" : P.why.c_str());
        return;
    }
    CHECK(expect_plan == 1, "a plan was built where none was expected");
    CHECK(expect_box < 0 || (int)P.box_key == expect_box, "box key %d", (int)P.box_key);
    const int W = kTilePhases, E = P.slots;
    CHECK((int)P.order.size() == n && (int)P.tph.size() == P.ntiles * (W + 1) && (int)P.tcol.size() == P.ntiles,
          "array sizes");
    CHECK((int)P.lcol.size() == E * n && (int)P.ucol.size() == E * n && (int)P.dmap.size() == n, "entry sizes");
    std::vector<int> seen(n, 0);
    for (int k = 0; k < n; ++k) {
        CHECK(P.order[k] >= 0 && P.order[k] < n, "order out of range");
        seen[P.order[k]]++;
    }
    for (int i = 0; i < n; ++i) CHECK(seen[i] == 1, "order is not a permutation");
    int prev = 0, splits = 0;
    for (int t = 0; t < P.ntiles; ++t) {
        const int* ph = &P.tph[(size_t)t * (W + 1)];
        CHECK(ph[0] == prev && ph[W] > ph[0] && ph[W] - ph[0] <= kTileMaxRows, "tile %d bounds", t);
        for (int q = 0; q < W; ++q) CHECK(ph[q + 1] >= ph[q] && ph[q + 1] - ph[q] <= kTileWave, "tile %d phase", t);
        CHECK(P.tcol[t] >= 0 && P.tcol[t] < P.ncols, "column id");
        if (t > 0 && P.tcol[t] == P.tcol[t - 1]) {
            const int last = perm[P.order[ph[0] - 1]], first = perm[P.order[ph[0]]];
            if (flev[first] <= flev[last]) ++splits;   // one tile key cut into two tiles
        }
        prev = ph[W];
    }
    CHECK(prev == n, "tiles do not cover the positions");
    CHECK(!expect_splits || splits > 0, "the split path was not exercised");
    for (int waves : {1, 3, 64, 1000}) {
        std::vector<int> ws, wt;
        tile_schedule(P, waves, ws, wt);
        std::vector<int> cnt(P.ntiles, 0);
        for (int w = 0; w < waves; ++w)
            for (int q = ws[w]; q < ws[w + 1]; ++q) {
                cnt[wt[q]]++;
                CHECK(q == ws[w] || wt[q] > wt[q - 1], "schedule out of order");
            }
        for (int t = 0; t < P.ntiles; ++t) CHECK(cnt[t] == 1, "schedule misses a tile");
    }
    // sequential references in the caller's numbering
    const double tiny = 1.0e-64;   // TINY of Library/allp_mod.f90:31
    std::vector<double> u(n), Dr(n, 0.0), zr(n, 0.0);
    for (int r = 0; r < n; ++r) u[r] = U(g);
    for (int r = 0; r < n; ++r) {
        double s = val[pc.dia[r]];
        for (int j = pc.rowptr[r]; j < pc.dia[r]; ++j) s = s - Dr[pc.col[j]] * val[j] * val[j];
        Dr[r] = 1.0 / s;
    }
    for (int r = 0; r < n; ++r) {
        double s = u[r];
        for (int j = pc.rowptr[r]; j < pc.dia[r]; ++j) s = s - val[j] * zr[pc.col[j]];
        zr[r] = s * Dr[r];
    }
    for (int r = n - 1; r >= 0; --r) {
        double s = zr[r] / (Dr[r] + tiny);
        for (int j = pc.dia[r] + 1; j < pc.rowptr[r + 1]; ++j) s = s - val[j] * zr[pc.col[j]];
        zr[r] = s * Dr[r];
    }
    // the tiled kernels' data flow
    std::vector<double> ui(n), Di(n, 0.0), zi(n, 0.0), slot(n), dp(n), zf(n);
    std::vector<char> ready(n, 0);
    for (int r = 0; r < n; ++r) ui[iperm[r]] = u[r];
    double lds[kTileMaxRows];
    auto sweep = [&](bool backward, int kind) {   // kind 0 pivots, 1 forward, 2 backward
        std::fill(ready.begin(), ready.end(), 0);
        for (int tt = 0; tt < P.ntiles; ++tt) {
            const int t = backward ? P.ntiles - 1 - tt : tt;
            const int* ph = &P.tph[(size_t)t * (W + 1)];
            const int base = ph[0], end = ph[W];
            for (int qq = 0; qq < W; ++qq) {
                const int q = backward ? W - 1 - qq : qq;
                for (int k = ph[q]; k < ph[q + 1]; ++k) {
                    const int i = P.order[k];
                    const std::vector<int>& codes = backward ? P.ucol : P.lcol;
                    const std::vector<int>& maps = backward ? P.umap : P.lmap;
                    double s = kind == 0 ? vi[P.dmap[k]] : kind == 1 ? ui[i] : zf[k] / (dp[k] + tiny);
                    if (kind == 0 && P.dmap[k] != diai[i]) {
                        ++fails;
                        std::printf("FAIL %s: dmap\n", name.c_str());
                        return;
                    }
                    for (int j = 0; j < E; ++j) {
                        const int c = codes[(size_t)j * n + k];
                        if (c == kTileNone) break;
                        const int e = maps[(size_t)j * n + k];
                        if (e < 0 || e >= nnz) {
                            ++fails;
                            std::printf("FAIL %s: entry map\n", name.c_str());
                            return;
                        }
                        double v;
                        if (c < 0) {
                            const int l = -c - 1;
                            const bool earlier = backward ? (base + l >= ph[q + 1] && base + l < end)
                                                          : (base + l < ph[q]);
                            if (!earlier || !ready[base + l]) {
                                ++fails;
                                std::printf("FAIL %s: own-tile value not yet made\n", name.c_str());
                                return;
                            }
                            v = lds[l];
                        } else {
                            const bool other = backward ? c >= end : c < base;
                            if (c >= n || !other || !ready[c]) {
                                ++fails;
                                std::printf("FAIL %s: other-tile value not yet published\n", name.c_str());
                                return;
                            }
                            v = slot[c];
                        }
                        s = kind == 0 ? s - v * vi[e] * vi[e] : s - vi[e] * v;
                    }
                    const double out = kind == 0 ? 1.0 / s : kind == 1 ? s * Di[i] : s * dp[k];
                    lds[k - base] = out;
                    slot[k] = out;
                    ready[k] = 1;
                    if (kind == 0) Di[i] = out;
                    if (kind == 1) {
                        zf[k] = out;
                        dp[k] = Di[i];
                    }
                    if (kind == 2) zi[i] = out;
                }
            }
        }
    };
    sweep(false, 0);
    if (fails) return;
    for (int r = 0; r < n; ++r) CHECK(std::memcmp(&Di[iperm[r]], &Dr[r], 8) == 0, "pivot of row %d", r);
    sweep(false, 1);
    if (fails) return;
    sweep(true, 2);
    if (fails) return;
    for (int r = 0; r < n; ++r) CHECK(std::memcmp(&zi[iperm[r]], &zr[r], 8) == 0, "z of row %d", r);
    std::printf("case %s plan=1 box=%d tiles=%d cols=%d splits=%d ok\n", name.c_str(), (int)P.box_key, P.ntiles,
                P.ncols, splits);
}

}  // namespace

int main() {
    run("box_20x17x13", box(20, 17, 13), false, 1, 1, false);
    run("box_20x17x13_internal_permutation", box(20, 17, 13), true, 1, 1, false);
    run("box_64x64x16", box(64, 64, 16), false, 1, 1, false);
    run("plane_70x50", box(70, 50, 1), false, 1, 0, false);
    run("random_sorted_symmetric", random_sorted(60000, 400, 11), false, 1, -1, true);
    run("random_sorted_symmetric_internal_permutation", random_sorted(20000, 40, 12), true, 1, -1, false);
    {
        std::vector<int> number(12 * 11 * 10);
        for (size_t i = 0; i < number.size(); ++i) number[i] = (int)i;
        std::shuffle(number.begin(), number.end(), std::mt19937(5));
        run("box_randomly_renumbered", box(12, 11, 10, &number), false, 0, -1, false);
    }
    std::printf(fails ? "FAILED\n" : "ALL OK\n");
    return fails ? 1 : 0;
}
