// tile_plan.h -- the tiles of the tiled one-launch DIC sweeps (host side,
// plain C++).
//
// The DIC sweeps of Process/CG.f90:135-156 (and the pivots of
// Process/Prec1.f90:51-59) solve with the lower and upper triangles of the
// T-FlowS pattern.  Row r of the forward sweep waits on the rows of its
// lower entries, all of which have a smaller index in the caller's
// numbering (TopolM.f90:72-77 sorts every row); its forward level is the
// length of the longest such chain.
//
// A tile is a set of rows that one wavefront sweeps on its own: at most
// kTilePhases consecutive forward levels ("phases") of at most kTileWave
// rows each.  Tiles are keyed by
//   c1 = level / kTilePhases,  c2 = r / R1,  c3 = (r mod S) / R2
// and numbered in lexicographic key order, the rows of a tile by (level,
// r).  R1, R2 and S come from the two largest scales of the lower offsets
// r - c (for a naturally numbered box of nx * ny * nz cells: nx * ny and
// nx), so that a tile of a box is a parallelepiped of 8 planes by 8 lines
// by 8 levels, 512 rows.  Every lower dependency of a row then lies in an
// earlier tile or in an earlier phase of its own tile; the builder checks
// it (and, for the backward sweep, that every upper dependency lies in a
// later tile or a later phase of the row's tile).  When the box-shaped key
// fails the check it falls back to S = R1, a key under which the order
// holds for every pattern whose rows are sorted.
//
// Output: positions (one per row, tile by tile, phase by phase), the phase
// starts of every tile, and the entries of each row in slot-major ELL form
// with their columns coded as
//   code >= 0:          the position of a dependency in another tile,
//   -1 >= code > kNone: -(1 + its index inside the row's own tile),
//   kTileNone:          a padding slot (the row has fewer entries).
#pragma once

#include <climits>
#include <cstdint>
#include <string>
#include <vector>

namespace tflows {

constexpr int kTilePhases = 8;              // forward levels per tile
constexpr int kTileWave = 64;               // rows per phase (one wavefront)
constexpr int kTileMaxRows = kTilePhases * kTileWave;
constexpr int kTileEntries = 4;             // most entries per triangle row
constexpr int kTileNone = INT_MIN;          // padding slot

struct TilePlan {
    int n = 0;
    int ntiles = 0;
    int slots = 0;                  // entry slots per row (3 or 4)
    bool box_key = false;           // the box-shaped key was used
    int64_t scale_hi = 0, scale_lo = 0;   // the two offset scales found
    std::vector<int> order;         // position -> internal row
    std::vector<int> tph;           // ntiles * (kTilePhases + 1) phase starts
    std::vector<int> tcol;          // the column of every tile: its key without c1, numbered densely
    int ncols = 0;
    std::vector<int> dmap;          // n: the internal diagonal entry of every position
    std::vector<int> lcol, lmap;    // slots * n: lower codes, internal entry (-1 padding)
    std::vector<int> ucol, umap;    // slots * n: upper codes, internal entry (-1 padding)
    std::string why;                // why no plan was built
};

// n rows; the caller's pattern rowptr / col / dia (0-based: row r holds
// entries [rowptr[r], rowptr[r+1]), its diagonal at dia[r]); flev the
// forward level of every caller row; iperm the internal row of each caller
// row (nullptr: the same); rp / cl the internal pattern (row iperm[r]
// holds the entries of caller row r in their order, columns renamed).
// Returns false, with the reason in out.why, when the pattern does not
// allow the tiled sweeps.
bool build_tile_plan(int n, const int* rowptr, const int* col, const int* dia, const int* flev, const int* iperm,
                     const int* rp, const int* cl, TilePlan& out);

// The tiles each of `waves` wavefronts sweeps: wavefront w takes every tile
// of the columns c with c mod waves == w, in increasing tile index, so that
// it climbs its columns level chunk by level chunk.  wstart (waves + 1)
// indexes the lists in wtiles.  Any assignment that keeps each list in
// increasing index is free of deadlock when all wavefronts are resident.
void tile_schedule(const TilePlan& plan, int waves, std::vector<int>& wstart, std::vector<int>& wtiles);

// The forward (lower = true) or backward levels of the caller's pattern when
// every lower column is below its row and every upper column above it (the
// levels then follow in one pass in index order); returns the number of
// levels, or -1 when the pattern is not so ordered.
int sorted_pattern_levels(int n, const int* rowptr, const int* col, const int* dia, bool lower,
                          std::vector<int>& lev);

}  // namespace tflows
