// slot_layout.cpp -- the shared-offset sliced layouts of the SpMV operators
// (host side, plain C++; the kernels that read them are in kernels.h).
#include "slot_layout.h"

#include <algorithm>
#include <climits>
#include <cstdint>

namespace tflows {

bool shared_slot_layout(int n, const std::vector<int>& rp, const std::vector<int>& cl, std::vector<int>& off,
                        std::vector<int>& soff, std::vector<unsigned char>& mask, std::vector<int>& map) {
    const int nsl = (n + kSliceRows - 1) / kSliceRows;
    const int64_t nnz = rp[n];
    off.assign(nsl + 1, 0);
    soff.clear();
    std::vector<int> set;
    for (int sl = 0; sl < nsl; ++sl) {
        set.clear();
        for (int i = sl * kSliceRows; i < std::min(n, (sl + 1) * kSliceRows); ++i)
            for (int e = rp[i]; e < rp[i + 1]; ++e) {
                if (e > rp[i] && cl[e] <= cl[e - 1]) return false;   // the slots must run in CSR order
                set.push_back(cl[e] - i);
            }
        std::sort(set.begin(), set.end());
        set.erase(std::unique(set.begin(), set.end()), set.end());
        if ((int)set.size() > kMaxSliceSlots) return false;
        const int64_t next = (int64_t)off[sl] + (int64_t)set.size() * kSliceRows;
        if (next > INT32_MAX) return false;
        off[sl + 1] = (int)next;
        soff.insert(soff.end(), set.begin(), set.end());
    }
    if ((double)off[nsl] > 1.15 * (double)std::max<int64_t>(nnz, 1)) return false;
    mask.assign((size_t)nsl * kSliceRows, 0);
    map.assign(off[nsl], -1);
    for (int sl = 0; sl < nsl; ++sl) {
        const int* so = soff.data() + off[sl] / kSliceRows;
        for (int i = sl * kSliceRows; i < std::min(n, (sl + 1) * kSliceRows); ++i) {
            int j = 0;
            for (int e = rp[i]; e < rp[i + 1]; ++e) {
                while (so[j] != cl[e] - i) ++j;                      // both ascend; every offset is in the set
                mask[i] |= (unsigned char)(1u << j);
                map[off[sl] + j * kSliceRows + (i - sl * kSliceRows)] = e;
            }
        }
    }
    return true;
}

static int floor_div64(int64_t v) { return (int)(v >= 0 ? v / 64 : -((-v + 63) / 64)); }

bool symmetric_slot_layout(int n, const std::vector<int>& rp, const std::vector<int>& cl, const std::vector<int>& off,
                           const std::vector<int>& soff, const std::vector<unsigned char>& mask,
                           const std::vector<int>& map, SymSlots& out) {
    const int nsl = (int)off.size() - 1;
    if (nsl < 1) return false;
    std::vector<int> nl(nsl), w(nsl);
    out.sidx.assign(nsl + 1, 0);
    out.off.assign(nsl + 1, 0);
    for (int sl = 0; sl < nsl; ++sl) {
        w[sl] = (off[sl + 1] - off[sl]) / kSliceRows;
        out.sidx[sl] = off[sl] / kSliceRows;
        const int* so = soff.data() + out.sidx[sl];
        int c = 0;
        while (c < w[sl] && so[c] < 0) ++c;   // the offsets ascend: the negative ones first
        nl[sl] = c;
        const int64_t next = (int64_t)out.off[sl] + (int64_t)(w[sl] - c) * kSliceRows;
        if (next > INT32_MAX) return false;
        out.off[sl + 1] = (int)next;
    }
    out.sidx[nsl] = off[nsl] / kSliceRows;
    if (out.off[nsl] == 0) return false;
    out.map.assign(out.off[nsl], -1);
    for (int sl = 0; sl < nsl; ++sl)
        for (int j = nl[sl]; j < w[sl]; ++j)
            for (int l = 0; l < kSliceRows; ++l)
                out.map[out.off[sl] + kSliceRows * (j - nl[sl]) + l] = map[off[sl] + kSliceRows * j + l];
    // the slot of column offset v in slice t, -1 when the slice has none
    auto find = [&](int t, int v) {
        const int* so = soff.data() + out.sidx[t];
        for (int j = 0; j < w[t]; ++j)
            if (so[j] == v) return j;
        return -1;
    };
    constexpr int kNone = INT32_MIN / 2;
    out.lbase.assign(out.sidx[nsl], SlotPair{0, 0});
    for (int sl = 0; sl < nsl; ++sl) {
        const int* so = soff.data() + out.sidx[sl];
        for (int j = 0; j < nl[sl]; ++j) {
            const int d = so[j];
            int base[2];
            for (int t = 0; t < 2; ++t) {
                const int sp = floor_div64((int64_t)kSliceRows * sl + (t ? kSliceRows - 1 : 0) + d);
                base[t] = kNone;
                if (sp >= 0 && sp < nsl) {
                    const int jp = find(sp, -d);
                    if (jp >= nl[sp]) base[t] = out.off[sp] + kSliceRows * (jp - nl[sp]) - kSliceRows * sp;
                }
            }
            out.lbase[out.sidx[sl] + j] = SlotPair{base[0], base[1]};
        }
    }
    // every lower entry, read where the kernels will read it: the stored
    // slot there must be its mirror (j, i); the pairs let the device check
    // that the values agree
    out.pairs.clear();
    for (int sl = 0; sl < nsl; ++sl) {
        const int* so = soff.data() + out.sidx[sl];
        for (int l = 0; l < kSliceRows; ++l) {
            const int i = sl * kSliceRows + l;
            if (i >= n) break;
            for (int j = 0; j < nl[sl]; ++j) {
                if (!((mask[i] >> j) & 1u)) continue;
                const int d = so[j];
                const int e = map[off[sl] + kSliceRows * j + l];
                const int ip = i + d;
                if (e < 0 || ip < 0 || ip >= n) return false;
                const int sa = floor_div64((int64_t)kSliceRows * sl + d);
                const SlotPair lb = out.lbase[out.sidx[sl] + j];
                const int b = (ip / kSliceRows == sa) ? lb.x : lb.y;
                if (b == kNone) return false;
                const int64_t pos = (int64_t)b + ip;
                if (pos < 0 || pos >= out.off[nsl]) return false;
                const int em = out.map[pos];
                if (em < 0 || em < rp[ip] || em >= rp[ip + 1] || cl[em] != i) return false;
                out.pairs.push_back(e);
                out.pairs.push_back(em);
            }
        }
    }
    return true;
}

}  // namespace tflows
