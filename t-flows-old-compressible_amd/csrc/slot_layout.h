// slot_layout.h -- the shared-offset sliced layouts of the SpMV operators
// (host side, plain C++).
//
// The sliced ELL operators (kernels.h: Sell, DiffSell32) hold rows in slices
// of kSliceRows; slot j of row 64 s + l sits at off[s] + 64 j + l.  In the
// shared-offset form every row of slice s uses the slice's ascending list of
// column offsets col - row (slot j holding offset soff[off[s] / 64 + j]), a
// mask byte per row marks the slots the row has, and no column is stored
// per entry.
#pragma once

#include <cstdint>
#include <vector>

namespace tflows {

constexpr int kSliceRows = 64;       // rows of a slice (one wavefront)
constexpr int kMaxSliceSlots = 8;    // most offsets a slice may share (bits of the row mask)

// The shared-offset layout of a CSR pattern with row pointer rp and columns
// cl (0-based): slice value offsets, each slice's ascending column offsets,
// the per-row slot masks and the CSR entry of every slot (-1 where the row
// lacks the slot: value 0, never read).  Made when every row's columns
// ascend without repeats, no slice needs more than kMaxSliceSlots offsets
// and the slots stay within 1.15 x the entries; false otherwise.
bool shared_slot_layout(int n, const std::vector<int>& rp, const std::vector<int>& cl, std::vector<int>& off,
                        std::vector<int>& soff, std::vector<unsigned char>& mask, std::vector<int>& map);

// Two bases of one lower slot's mirror positions (the device reads them as
// int2).
struct SlotPair {
    int x, y;
};
static_assert(sizeof(SlotPair) == 8, "two ints, as the device's int2");

// The symmetric form of a shared-offset layout, for a structurally
// symmetric pattern whose values equal their mirrors: only the slots of
// non-negative offsets keep values (off, map as for the full form); sidx =
// each slice's first offset in soff (soff itself is unchanged); lbase = per
// slice and lower slot (indexed like soff), where the mirror values lie:
// lower slot j of slice s (offset d < 0) of row i reads the stored value of
// row i' = i + d at lbase.x + i' when i' lies in slice floor((64 s + d) /
// 64), else at lbase.y + i' (the slice after); pairs = (lower entry, its
// mirror entry) for every lower entry, flattened, so that a set of values
// can be checked for symmetry.  Every lower entry's mirror position is
// checked here against the pattern: false when one fails (or nothing would
// be stored).
struct SymSlots {
    std::vector<int> off, sidx, map, pairs;
    std::vector<SlotPair> lbase;
};
bool symmetric_slot_layout(int n, const std::vector<int>& rp, const std::vector<int>& cl, const std::vector<int>& off,
                           const std::vector<int>& soff, const std::vector<unsigned char>& mask,
                           const std::vector<int>& map, SymSlots& out);

}  // namespace tflows
