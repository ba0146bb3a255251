// Host check of the shared-offset sliced layouts (t-flows-old-compressible_amd/csrc/slot_layout.cpp).
//
// For a set of generated patterns it builds the full shared-offset layout
// and its symmetric form with the library's own builders, then runs on the
// CPU what the SpMV kernels of kernels.h do with them -- the FP64 row loop
// (sell_rows: slots in ascending offset order, absent slots skipped) and the
// FP32 difference-form loop (sell_tile: absent slots add a (p_i - p_i) with
// a = 0), each lower value of the symmetric form read at the position the
// device computes -- and checks:
//   * every index read lies inside its array;
//   * the three forms (CSR loop in CSR order, full layout, symmetric layout)
//     give the same bits, FP64 and FP32, on random symmetric values;
//   * every lower entry appears once in the pairs list, with its mirror, and
//     one perturbed value makes some pair differ;
//   * patterns that are not structurally symmetric, or whose slices need
//     too many offsets, are refused.
// Values follow the reference's SpMV (Process/CG.f90:193-199): each row's
// products added in its CSR order.
//
// Prints one line per case: "case <name> shared=<0|1> sym=<0|1>
// stored=<n> full=<n> ok" or "FAIL ..."; "ALL OK" at the end; exits
// non-zero on a failure.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <set>
#include <string>
#include <vector>

#include "slot_layout.h"

using namespace tflows;

namespace {

struct Pattern {
    int n = 0;
    std::vector<int> rp, cl;     // 0-based CSR, columns ascending per row
};

Pattern from_sets(int n, const std::vector<std::set<int>>& rows) {
    Pattern p;
    p.n = n;
    p.rp.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) {
        for (int c : rows[i]) p.cl.push_back(c);
        p.rp[i + 1] = (int)p.cl.size();
    }
    return p;
}

// 7-point box (2D when nz == 1), natural order; periodic in x when asked
std::vector<std::set<int>> box_rows(int nx, int ny, int nz, bool diag, bool periodic_x) {
    const int n = nx * ny * nz;
    std::vector<std::set<int>> rows(n);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const int r = i + nx * (j + ny * k);
                if (diag) rows[r].insert(r);
                auto add = [&](int ii, int jj, int kk) {
                    if (periodic_x) ii = (ii + nx) % nx;
                    if (ii < 0 || ii >= nx || jj < 0 || jj >= ny || kk < 0 || kk >= nz) return;
                    const int c = ii + nx * (jj + ny * kk);
                    if (c != r) rows[r].insert(c);
                };
                add(i - 1, j, k), add(i + 1, j, k), add(i, j - 1, k), add(i, j + 1, k);
                add(i, j, k - 1), add(i, j, k + 1);
            }
    return rows;
}

// symmetric random values on a symmetric pattern (a_ij == a_ji bit for bit)
std::vector<double> sym_values(const Pattern& p, std::mt19937_64& rng) {
    std::uniform_real_distribution<double> u(-2.0, 2.0);
    std::vector<double> v(p.cl.size());
    for (int i = 0; i < p.n; ++i)
        for (int e = p.rp[i]; e < p.rp[i + 1]; ++e) {
            const int c = p.cl[e];
            if (c > i) continue;
            v[e] = u(rng);
            if (c == i) continue;
            for (int f = p.rp[c]; f < p.rp[c + 1]; ++f)
                if (p.cl[f] == i) v[f] = v[e];
        }
    return v;
}

bool fail(const std::string& name, const std::string& what) {
    std::printf("FAIL %s: %s\n", name.c_str(), what.c_str());
    return false;
}

// the device's mirror position (kernels.h: sym_mirror_pos)
int mirror_pos(const SymSlots& S, int sb, int j, int sl, int row, int d) {
    const SlotPair lb = S.lbase[sb + j];
    const int ip = row + d;
    const int sa = (sl * 64 + d) >> 6;
    return ((ip >> 6) == sa ? lb.x : lb.y) + ip;
}

template <typename T>
bool in_range(const std::vector<T>& a, int64_t i) {
    return i >= 0 && i < (int64_t)a.size();
}

bool run_case(const std::string& name, const Pattern& p, bool expect_shared, bool expect_sym) {
    std::mt19937_64 rng(12345 + p.n);
    std::vector<int> off, soff, map;
    std::vector<unsigned char> mask;
    const bool shared = shared_slot_layout(p.n, p.rp, p.cl, off, soff, mask, map);
    SymSlots S;
    const bool sym = shared && symmetric_slot_layout(p.n, p.rp, p.cl, off, soff, mask, map, S);
    if (shared != expect_shared) return fail(name, "shared layout " + std::string(shared ? "made" : "refused"));
    if (sym != expect_sym) return fail(name, "symmetric layout " + std::string(sym ? "made" : "refused"));
    if (!shared) {
        std::printf("case %s shared=0 sym=0 stored=0 full=0 ok\n", name.c_str());
        return true;
    }
    const int nsl = (p.n + 63) / 64;
    const std::vector<double> a = sym_values(p, rng);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    std::vector<double> x(p.n);
    for (double& v : x) v = u(rng);
    std::vector<float> x32(p.n);
    for (int i = 0; i < p.n; ++i) x32[i] = (float)x[i];
    // the stored values, gathered through the maps as k_slot_values does
    std::vector<double> vfull(map.size()), vsym(sym ? S.map.size() : 0);
    for (size_t q = 0; q < map.size(); ++q) vfull[q] = map[q] < 0 ? 0.0 : a[map[q]];
    for (size_t q = 0; q < vsym.size(); ++q) {
        if (S.map[q] >= (int)a.size()) return fail(name, "map entry out of range");
        vsym[q] = S.map[q] < 0 ? 0.0 : a[S.map[q]];
    }
    std::vector<float> vfull32(vfull.size()), vsym32(vsym.size());
    for (size_t q = 0; q < vfull.size(); ++q) vfull32[q] = (float)vfull[q];
    for (size_t q = 0; q < vsym.size(); ++q) vsym32[q] = (float)vsym[q];
    // FP32 row sums of the difference form (any fixed value: the forms must agree)
    std::vector<float> rs(p.n);
    for (int i = 0; i < p.n; ++i) {
        double s = 0.0;
        for (int e = p.rp[i]; e < p.rp[i + 1]; ++e) s += a[e];
        rs[i] = (float)s;
    }
    for (int row = 0; row < p.n; ++row) {
        const int sl = row / 64, l = row % 64;
        const int w = (off[sl + 1] - off[sl]) / 64;
        const int* so = soff.data() + off[sl] / 64;
        const unsigned m = mask[row];
        // CSR loop
        double ref = 0.0;
        for (int e = p.rp[row]; e < p.rp[row + 1]; ++e) ref += a[e] * x[p.cl[e]];
        // full layout (FP64 row loop and FP32 difference form)
        double yf = 0.0;
        float yf32 = 0.0f;
        const float own = x32[row];
        for (int j = 0; j < w; ++j) {
            const int64_t q = (int64_t)off[sl] + 64 * j + l;
            if (!in_range(vfull, q)) return fail(name, "full slot out of range");
            const bool has = (m >> j) & 1u;
            const int c = has ? row + so[j] : row;
            if (c < 0 || c >= p.n) return fail(name, "full column out of range");
            if (has) yf += vfull[q] * x[c];
            yf32 = yf32 + vfull32[q] * (x32[c] - own);
        }
        yf32 = yf32 + rs[row] * own;
        if (std::memcmp(&yf, &ref, sizeof(double)) != 0) return fail(name, "full layout differs from CSR at row " + std::to_string(row));
        if (!sym) continue;
        // symmetric layout: lower slots from mirrors, then the stored ones
        const int sb = S.sidx[sl];
        const int ws = S.sidx[sl + 1] - sb;
        const int wst = (S.off[sl + 1] - S.off[sl]) / 64;
        const int nl = ws - wst;
        if (ws != w || nl < 0) return fail(name, "slot counts disagree");
        double ys = 0.0;
        float ys32 = 0.0f;
        for (int j = 0; j < nl; ++j) {
            const bool has = (m >> j) & 1u;
            const int d = soff[sb + j];
            if (d >= 0) return fail(name, "a lower slot with a non-negative offset");
            const int pos = mirror_pos(S, sb, j, sl, row, d);
            const int64_t idx = has ? pos : 0;
            if (!in_range(vsym, idx)) return fail(name, "mirror position out of range at row " + std::to_string(row));
            if (has) {
                const int e = map[off[sl] + 64 * j + l];
                if (e < 0 || S.map[pos] < 0 || a[S.map[pos]] != a[e] || p.cl[S.map[pos]] != row)
                    return fail(name, "mirror is not the transposed entry at row " + std::to_string(row));
                ys += vsym[pos] * x[row + d];
            }
            const float av = vsym32[idx];
            ys32 = ys32 + (has ? av : 0.0f) * (x32[has ? row + d : row] - own);
        }
        for (int j = nl; j < ws; ++j) {
            const int64_t q = (int64_t)S.off[sl] + 64 * (j - nl) + l;
            if (!in_range(vsym, q)) return fail(name, "stored slot out of range");
            const bool has = (m >> j) & 1u;
            const int c = has ? row + soff[sb + j] : row;
            if (c < 0 || c >= p.n) return fail(name, "stored column out of range");
            if (has) ys += vsym[q] * x[c];
            ys32 = ys32 + vsym32[q] * (x32[c] - own);
        }
        ys32 = ys32 + rs[row] * own;
        if (std::memcmp(&ys, &ref, sizeof(double)) != 0)
            return fail(name, "symmetric layout differs from CSR at row " + std::to_string(row));
        if (std::memcmp(&ys32, &yf32, sizeof(float)) != 0)
            return fail(name, "FP32 symmetric layout differs from the full one at row " + std::to_string(row));
    }
    if (sym) {
        // pairs: every lower entry once, with its mirror; a perturbed value shows
        std::vector<int> seen(p.cl.size(), 0);
        for (size_t q = 0; q + 1 < S.pairs.size(); q += 2) {
            const int e = S.pairs[q], f = S.pairs[q + 1];
            if (e < 0 || e >= (int)p.cl.size() || f < 0 || f >= (int)p.cl.size()) return fail(name, "pair out of range");
            ++seen[e];
            int re = 0;
            while (p.rp[re + 1] <= e) ++re;
            int rf = 0;
            while (p.rp[rf + 1] <= f) ++rf;
            if (p.cl[e] >= re || p.cl[e] != rf || p.cl[f] != re) return fail(name, "pair is not (lower entry, mirror)");
        }
        int lower = 0;
        for (int i = 0; i < p.n; ++i)
            for (int e = p.rp[i]; e < p.rp[i + 1]; ++e)
                if (p.cl[e] < i) {
                    ++lower;
                    if (seen[e] != 1) return fail(name, "a lower entry is not paired exactly once");
                }
        if ((int)S.pairs.size() != 2 * lower) return fail(name, "pairs count");
        std::vector<double> b = a;
        b[S.pairs[0]] *= 1.0 + 1e-15;
        bool differs = false;
        for (size_t q = 0; q + 1 < S.pairs.size(); q += 2) differs |= b[S.pairs[q]] != b[S.pairs[q + 1]];
        if (!differs) return fail(name, "a perturbed value went unnoticed");
        if (nsl + 1 != (int)S.sidx.size()) return fail(name, "sidx size");
    }
    std::printf("case %s shared=1 sym=%d stored=%zu full=%zu ok\n", name.c_str(), sym ? 1 : 0,
                sym ? S.map.size() : map.size(), map.size());
    return true;
}

}  // namespace

int main() {
    bool ok = true;
    auto box = [](int nx, int ny, int nz, bool diag, bool per) {
        auto r = box_rows(nx, ny, nz, diag, per);
        return from_sets(nx * ny * nz, r);
    };
    ok &= run_case("box_20x17x13", box(20, 17, 13, true, false), true, true);
    ok &= run_case("box_20x17x13_offdiag", box(20, 17, 13, false, false), true, true);
    // mostly wall rows: the padding exceeds the 1.15 x bound, so no shared layout
    ok &= run_case("box_5x5x5", box(5, 5, 5, true, false), false, false);
    ok &= run_case("plane_40x37", box(40, 37, 1, true, false), true, true);
    ok &= run_case("box_96x7x3", box(96, 7, 3, true, false), true, true);
    // periodic in x: the wrap-around offsets exist on 2 rows of 24, too much padding
    ok &= run_case("box_periodic_x_offdiag", box(24, 9, 5, false, true), false, false);
    {
        // symmetric holes: a random set of face pairs removed on both sides
        auto r = box_rows(29, 21, 11, true, false);
        std::mt19937 g(7);
        for (int i = 0; i < 29 * 21 * 11; ++i)
            for (int c : std::vector<int>(r[i].begin(), r[i].end()))
                if (c > i && g() % 40 == 0) {
                    r[i].erase(c);
                    r[c].erase(i);
                }
        ok &= run_case("box_with_holes", from_sets(29 * 21 * 11, r), true, true);
    }
    {
        // one mirror missing: not structurally symmetric, the symmetric form is refused
        auto r = box_rows(16, 8, 4, true, false);
        r[300].erase(301);
        ok &= run_case("box_missing_mirror", from_sets(16 * 8 * 4, r), true, false);
    }
    {
        // random symmetric graph with scattered offsets: too many per slice
        const int n = 500;
        std::vector<std::set<int>> r(n);
        std::mt19937 g(11);
        for (int i = 0; i < n; ++i) {
            r[i].insert(i);
            for (int t = 0; t < 3; ++t) {
                const int c = (int)(g() % n);
                if (c != i) r[i].insert(c), r[c].insert(i);
            }
        }
        ok &= run_case("random_symmetric", from_sets(n, r), false, false);
    }
    if (ok) std::printf("ALL OK\n");
    return ok ? 0 : 1;
}
