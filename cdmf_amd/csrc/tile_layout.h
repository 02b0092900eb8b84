// Host-side tables of the tiled (CSR5-style) copy of one compressed layout.
// Plain C++ (no HIP): the library builds them before uploading, and a CPU
// checker built with the address and undefined-behaviour sanitizers decodes
// them back to the layout.
//
// A tile holds 64 * sigma consecutive elements of the layout; lane l of a
// tile owns elements l*sigma .. l*sigma + sigma - 1 of it.  Per tile and lane
// a 32-bit word flags the elements that start a row.  In a split layout the
// rows are virtual: virtual row r < rows is the part of row r whose
// elements are gathered from below the split (below[r] leading elements),
// virtual rows rows, rows + 1, ... the rest of each row that has one, in row
// order.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

namespace ccd {

// One row spanning tiles t0..t1 (t1 > t0) of a tiled layout: its row, first
// and last tile and length.
struct SpanRow {
    uint32_t row, t0, t1, len;
};

// Row r of a split layout: its first part (virtual row r) over tiles a0..a1
// (a0 = 0xffffffff: empty), its second part (virtual row b, 0xffffffff:
// none) over tiles b0..b1, and its length.
struct SplitRow {
    uint32_t a0, a1, b, b0, b1, len;
};

struct TileTables {
    int sigma = 0;
    bool split = false;
    uint64_t ntiles = 0;
    // virtual rows of a split layout: row pointer, the layout position of
    // every element in tiled order, the real row of every second part, and
    // each real row's parts with their tile ranges
    std::vector<uint32_t> vptr, vpos, cold_row;
    std::vector<SplitRow> srows;
    // per tile and lane: row-start bits of the lane's elements
    std::vector<uint32_t> flags;
    // per tile: first (virtual) row, segment-table offset (0xffffffff: the
    // tile's rows are consecutive), open flags (1: a row enters from the
    // previous tile, 2: a row leaves into the next one); the open array is
    // padded to whole 32-bit words
    std::vector<uint32_t> row0, segoff, segrows;
    std::vector<uint8_t> open;
    // rows spanning tiles, two-tile spans first
    std::vector<SpanRow> spans;
    uint32_t nshort = 0;
    // split layouts: real rows, and the first tile holding second parts only
    uint32_t nreal = 0;
    uint32_t cold_tile0 = 0xffffffffu;
};

// Tables for tiles of 64 * sigma elements of the layout with row pointer ptr
// (rows + 1 entries, ptr[rows] = nnz).  split: keep each row's two parts
// (below[r] leading elements, then the rest) as two virtual rows, all first
// parts before all second parts.  Returns an empty string or the reason it
// cannot be built.
std::string build_tile_tables(const uint32_t* ptr, int64_t rows, uint64_t nnz, const uint32_t* below, int sigma,
                              bool split, TileTables& T);

}  // namespace ccd
