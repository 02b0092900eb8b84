// CPU check of the tiled layout tables (cdmf_amd/csrc/tile_layout.cpp) --
// test infrastructure, built by tests/test_tile_layout.py with the address
// and undefined-behaviour sanitizers.
//
// usage: tile_check <layout.bin> <sigma> <split_at>
//   layout.bin: int64 rows, int64 gdim, uint64 nnz, uint32 ptr[rows + 1],
//   uint32 idx[nnz] (one compressed layout).  split_at > 0: order each row's
//   elements as the library does for a split layout (indices below split_at
//   first, in order, then the rest sorted) and build the split tables.
//
// Checks, for tiles of 64 * sigma elements:
//   1. every index, written to the tiled index planes by the position rules
//      of the retile kernels (tiled_kernels.hip) for its format (u16 below
//      2^16 gathered entries, 3 bytes below 2^24, else u32) and read back by
//      the lane-load rules of load_tile (tile_body.h), is the layout's index
//      at the position vpos maps to; values likewise;
//   2. a restatement of a plain sweep over the tiles (tile_core: lane loop,
//      row starts, lane_seg0, the segmented scan over lanes, head / tail
//      partials, row lengths from the start positions) followed by the span
//      finish (plain layouts) or the split finish (split layouts) gives, for
//      every row, exactly the sums g = sum x v and h = sum x^2 of a direct
//      pass over the layout (small integer data, so every sum is exact), the
//      row's length where a row is finished, and finishes every row once;
//   3. split layouts: part tile ranges and lengths of every row, and tiles
//      from cold_tile0 on hold only elements gathered from beyond the split.
// Prints "OK ..." and exits 0, or prints the first mismatch and exits 1.
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <vector>

#include "tile_layout.h"

namespace {

constexpr int WAVE = 64;
constexpr uint32_t NONE = 0xffffffffu;
int failures = 0;

#define CHECK(cond, ...)                                  \
    do {                                                  \
        if (!(cond)) {                                    \
            if (failures++ < 10) {                        \
                fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
                fprintf(stderr, __VA_ARGS__);             \
                fprintf(stderr, "\n");                    \
            }                                             \
        }                                                 \
    } while (0)

struct GH {
    double g = 0, h = 0;
};

bool read_all(FILE* f, void* p, size_t bytes) { return fread(p, 1, bytes, f) == bytes; }

}  // namespace

int main(int argc, char** argv) {
    if (argc != 4) {
        fprintf(stderr, "usage: tile_check <layout.bin> <sigma> <split_at>\n");
        return 2;
    }
    const int sigma = atoi(argv[2]);
    const uint32_t split_at = (uint32_t)strtoul(argv[3], nullptr, 10);
    FILE* f = fopen(argv[1], "rb");
    if (!f) return 2;
    int64_t rows = 0, gdim = 0;
    uint64_t nnz = 0;
    if (!read_all(f, &rows, 8) || !read_all(f, &gdim, 8) || !read_all(f, &nnz, 8)) return 2;
    std::vector<uint32_t> ptr((size_t)rows + 1), idx0(nnz);
    if (!read_all(f, ptr.data(), 4 * ptr.size()) || !read_all(f, idx0.data(), 4 * nnz)) return 2;
    fclose(f);

    // element order of the library's layouts: split -> below first, rest sorted
    const bool split = split_at > 0;
    std::vector<uint32_t> idx(nnz), below(split ? rows : 0);
    for (int64_t r = 0; r < rows; ++r) {
        std::vector<uint32_t> lo, hi;
        for (uint32_t e = ptr[r]; e < ptr[r + 1]; ++e) (split && idx0[e] >= split_at ? hi : lo).push_back(idx0[e]);
        std::sort(hi.begin(), hi.end());
        if (split) below[r] = (uint32_t)lo.size();
        uint32_t e = ptr[r];
        for (uint32_t v : lo) idx[e++] = v;
        for (uint32_t v : hi) idx[e++] = v;
    }
    // values and the gathered vector: small integers, so every sum is exact
    std::vector<double> val(nnz), x((size_t)gdim);
    for (uint64_t e = 0; e < nnz; ++e) val[e] = (double)(1 + (e * 7 + 3) % 5);
    for (int64_t j = 0; j < gdim; ++j) x[j] = (double)((j * 13 + 5) % 9);

    ccd::TileTables T;
    const std::string why = ccd::build_tile_tables(ptr.data(), rows, nnz, split ? below.data() : nullptr, sigma,
                                                   split, T);
    if (!why.empty()) {
        fprintf(stderr, "build_tile_tables: %s\n", why.c_str());
        return 1;
    }
    const uint64_t TS = (uint64_t)WAVE * sigma, ntiles = T.ntiles, padded = ntiles * TS;
    CHECK(ntiles == (nnz + TS - 1) / TS, "tile count");
    // virtual layout: row pointer and the layout position of each element
    const std::vector<uint32_t>& vptr = split ? T.vptr : ptr;
    const int64_t vrows = (int64_t)vptr.size() - 1;
    auto lpos = [&](uint64_t p) -> uint64_t { return split ? T.vpos[p] : p; };
    std::vector<int64_t> vrow_of(nnz, -1);
    for (int64_t r = 0; r < vrows; ++r)
        for (uint32_t p = vptr[r]; p < vptr[r + 1]; ++p) vrow_of[p] = r;
    auto real_row = [&](int64_t v) -> int64_t { return split && v >= rows ? (int64_t)T.cold_row[v - rows] : v; };
    if (split) {
        CHECK((int64_t)T.srows.size() == rows && T.nreal == (uint32_t)rows, "split row table size");
        std::vector<uint8_t> seen(nnz, 0);
        for (uint64_t p = 0; p < nnz; ++p) {
            CHECK(T.vpos[p] < nnz && !seen[T.vpos[p]], "vpos is not a permutation at %llu", (unsigned long long)p);
            if (T.vpos[p] < nnz) seen[T.vpos[p]] = 1;
        }
    }

    // 1. index and value planes through the retile and lane-load rules
    const int ib = gdim <= 65536 ? 2 : gdim <= (1 << 24) ? 3 : 4;
    std::vector<uint16_t> lo16(ib == 4 ? 0 : padded, 0);
    std::vector<uint8_t> hi8(ib == 3 ? padded : 0, 0);
    std::vector<uint32_t> w32(ib == 4 ? padded : 0, 0);
    std::vector<double> tv(padded, 0.0);
    for (uint64_t e = 0; e < padded; ++e) {  // retile_idx(_24)_kernel, retile_f32_kernel
        const uint64_t tile = e / TS, rem = e % TS, i = rem % sigma, l = rem / sigma;
        const uint32_t v = e < nnz ? idx[lpos(e)] : 0u;
        if (ib == 4) w32[tile * TS + (i / 4) * (4 * WAVE) + 4 * l + (i % 4)] = v;
        else lo16[tile * TS + (i / 8) * (8 * WAVE) + 8 * l + (i % 8)] = (uint16_t)(v & 0xffffu);
        if (ib == 3) hi8[e] = (uint8_t)(v >> 16);
        tv[tile * TS + (i >> 2) * (4 * WAVE) + 4 * l + (i & 3)] = e < nnz ? val[lpos(e)] : 0.0;
    }
    std::vector<uint32_t> tidx(padded);  // read back as load_tile does: lane `lane`, element i
    std::vector<double> tval(padded);
    for (uint64_t t = 0; t < ntiles; ++t)
        for (int lane = 0; lane < WAVE; ++lane) {
            const int G = ib == 4 ? 4 : 8;
            for (int q = 0; q < sigma / G; ++q)
                for (int j = 0; j < G; ++j) {
                    const uint64_t at = t * TS + (uint64_t)q * G * WAVE + (uint64_t)G * lane + j;
                    uint32_t c = ib == 4 ? w32[at] : lo16[at];
                    if (ib == 3) c |= (uint32_t)hi8[t * TS + (uint64_t)lane * sigma + (G * q + j)] << 16;
                    tidx[t * TS + (uint64_t)lane * sigma + G * q + j] = c;
                }
            for (int q = 0; q < sigma / 4; ++q)
                for (int j = 0; j < 4; ++j)
                    tval[t * TS + (uint64_t)lane * sigma + 4 * q + j] = tv[t * TS + (uint64_t)q * 256 + 4 * lane + j];
        }
    for (uint64_t p = 0; p < nnz; ++p) {
        CHECK(tidx[p] == idx[lpos(p)], "index at tiled position %llu", (unsigned long long)p);
        CHECK(tval[p] == val[lpos(p)], "value at tiled position %llu", (unsigned long long)p);
    }

    // 2. a plain sweep over the tiles, restated
    std::vector<GH> head(ntiles), tail(ntiles), part(vrows);
    std::vector<GH> fin((size_t)rows);       // finished real rows (plain layouts)
    std::vector<int> finished(vrows, 0);     // finish_row calls per (virtual) row
    std::vector<int> decoded(nnz, 0);
    auto finish = [&](int64_t row, double g, double h, int64_t d) {
        CHECK(row >= 0 && row < vrows, "finish of row %lld out of range", (long long)row);
        if (row < 0 || row >= vrows) return;
        ++finished[row];
        if (split) {
            part[row] = {g, h};
        } else {
            fin[row] = {g, h};
            CHECK(d == (int64_t)(vptr[row + 1] - vptr[row]), "row %lld length %lld, kernel sees %lld", (long long)row,
                  (long long)(vptr[row + 1] - vptr[row]), (long long)d);
        }
    };
    for (uint64_t t = 0; t < ntiles; ++t) {
        const uint32_t row0 = T.row0[t], segoff = T.segoff[t], open = T.open[t];
        auto seg_row = [&](uint32_t s) -> int64_t {
            if (segoff == NONE) return (int64_t)row0 + s;
            CHECK(segoff + s < T.segrows.size(), "segment %u of tile %llu past the table", s, (unsigned long long)t);
            return segoff + s < T.segrows.size() ? (int64_t)T.segrows[segoff + s] : -1;
        };
        uint32_t fl[WAVE], seg0[WAVE];
        for (int l = 0; l < WAVE; ++l) fl[l] = T.flags[t * WAVE + l];
        {  // lane_seg0
            uint32_t incl = 0;
            const uint32_t start0 = fl[0] & 1u;
            for (int l = 0; l < WAVE; ++l) {
                const uint32_t pc = (uint32_t)__builtin_popcount(fl[l]);
                incl += pc;
                seg0[l] = l == 0 ? 0u : incl - pc - start0 + (fl[l] & 1u);
            }
        }
        double g[WAVE], h[WAVE], gA[WAVE], hA[WAVE];
        bool has[WAVE];
        int st[WAVE], first[WAVE];
        uint32_t segend[WAVE];
        int64_t rowend[WAVE];
        for (int l = 0; l < WAVE; ++l) {
            uint32_t seg = seg0[l];
            int64_t row = seg_row(seg);
            double gg = 0, hh = 0;
            gA[l] = hA[l] = 0;
            has[l] = (fl[l] & 1u) != 0;
            st[l] = 0;
            first[l] = 0;
            for (int i = 0; i < sigma; ++i) {
                if (i > 0 && ((fl[l] >> i) & 1u)) {
                    if (!has[l]) {
                        gA[l] = gg;
                        hA[l] = hh;
                        first[l] = i;
                    } else {
                        finish(row, gg, hh, i - st[l]);
                    }
                    st[l] = i;
                    has[l] = true;
                    gg = hh = 0;
                    row = seg_row(++seg);
                }
                const uint64_t p = t * TS + (uint64_t)l * sigma + i;
                if (p < nnz) {
                    CHECK(row == vrow_of[p], "tile %llu lane %d element %d: row %lld, layout row %lld",
                          (unsigned long long)t, l, i, (long long)row, (long long)vrow_of[p]);
                    ++decoded[p];
                    const double xv = x[tidx[p]];
                    gg += xv * tval[p];
                    hh += xv * xv;
                }
            }
            g[l] = gg;
            h[l] = hh;
            segend[l] = seg;
            rowend[l] = row;
        }
        double S[WAVE], Sh[WAVE];
        int LS[WAVE];
        for (int l = 0; l < WAVE; ++l) {  // segmented inclusive scan, running maximum of row starts
            S[l] = has[l] || l == 0 ? g[l] : S[l - 1] + g[l];
            Sh[l] = has[l] || l == 0 ? h[l] : Sh[l - 1] + h[l];
            const int ls = has[l] ? l * sigma + st[l] : -1;
            LS[l] = l == 0 ? ls : std::max(LS[l - 1], ls);
        }
        for (int l = 0; l < WAVE; ++l) {
            const double cg = l == 0 ? 0 : S[l - 1], ch = l == 0 ? 0 : Sh[l - 1];
            const int ls_in = l == 0 ? -1 : LS[l - 1];
            if (has[l]) {
                const bool bit0 = (fl[l] & 1u) != 0;
                if (!(bit0 && l == 0)) {
                    const uint32_t seg_in = bit0 ? seg0[l] - 1u : seg0[l];
                    const double tg = bit0 ? cg : cg + gA[l], th = bit0 ? ch : ch + hA[l];
                    const int end = l * sigma + (bit0 ? 0 : first[l]);
                    if (seg_in == 0 && (open & 1u)) head[t] = {tg, th};
                    else finish(seg_row(seg_in), tg, th, end - ls_in);
                }
            }
        }
        {
            const uint64_t left = nnz - t * TS;
            const int end = left < TS ? (int)left : (int)TS;
            if (segend[WAVE - 1] == 0 && (open & 1u)) head[t] = {S[WAVE - 1], Sh[WAVE - 1]};
            else if (open & 2u) tail[t] = {S[WAVE - 1], Sh[WAVE - 1]};
            else finish(rowend[WAVE - 1], S[WAVE - 1], Sh[WAVE - 1], end - LS[WAVE - 1]);
        }
    }
    for (uint64_t p = 0; p < nnz; ++p) CHECK(decoded[p] == 1, "element %llu read %d times", (unsigned long long)p, decoded[p]);

    auto span_sum = [&](uint32_t t0, uint32_t t1) {
        GH s = tail[t0];
        for (uint32_t q = t0 + 1; q <= t1; ++q) {
            s.g += head[q].g;
            s.h += head[q].h;
        }
        return s;
    };
    // rows spanning tiles: finished by the span / split finish only
    std::vector<uint8_t> spans_tiles(vrows, 0);
    for (int64_t r = 0; r < vrows; ++r)
        if (vptr[r + 1] > vptr[r]) spans_tiles[r] = vptr[r] / TS != (vptr[r + 1] - 1ull) / TS;
    for (int64_t r = 0; r < vrows; ++r) {
        if (vptr[r + 1] == vptr[r]) CHECK(finished[r] == 0, "empty row %lld finished", (long long)r);
        else if (spans_tiles[r]) CHECK(finished[r] == 0, "spanning row %lld finished in a tile", (long long)r);
        else CHECK(finished[r] == 1, "row %lld finished %d times", (long long)r, finished[r]);
    }
    std::vector<GH> got((size_t)rows);
    std::vector<int> done((size_t)rows, 0);
    if (!split) {
        uint32_t nshort = 0;
        for (size_t q = 0; q < T.spans.size(); ++q) {
            const ccd::SpanRow& sp = T.spans[q];
            CHECK(sp.row < (uint64_t)rows && spans_tiles[sp.row], "span %zu", q);
            if (sp.row >= (uint64_t)rows) continue;
            CHECK(sp.t0 == vptr[sp.row] / TS && sp.t1 == (vptr[sp.row + 1] - 1ull) / TS, "span %zu tiles", q);
            CHECK(sp.len == vptr[sp.row + 1] - vptr[sp.row], "span %zu length", q);
            CHECK((q < T.nshort) == (sp.t1 == sp.t0 + 1), "span %zu order", q);
            nshort += sp.t1 == sp.t0 + 1;
            fin[sp.row] = span_sum(sp.t0, sp.t1);
            ++finished[sp.row];
        }
        CHECK(nshort == T.nshort, "short span count");
        for (int64_t r = 0; r < rows; ++r) {
            got[r] = fin[r];
            done[r] = finished[r];
        }
    } else {
        for (int64_t r = 0; r < rows; ++r) {
            const ccd::SplitRow& sr = T.srows[r];
            const uint32_t len = ptr[r + 1] - ptr[r];
            CHECK(sr.len == len, "split row %lld length", (long long)r);
            if (below[r]) CHECK(sr.a0 == vptr[r] / TS && sr.a1 == (vptr[r + 1] - 1ull) / TS, "row %lld part 1 tiles", (long long)r);
            else CHECK(sr.a0 == NONE, "row %lld empty part 1", (long long)r);
            if (len > below[r]) {
                CHECK(sr.b != NONE && sr.b >= (uint32_t)rows && sr.b < (uint64_t)vrows && real_row(sr.b) == r,
                      "row %lld part 2", (long long)r);
                if (sr.b == NONE || sr.b >= (uint64_t)vrows) continue;
                CHECK(sr.b0 == vptr[sr.b] / TS && sr.b1 == (vptr[sr.b + 1] - 1ull) / TS, "row %lld part 2 tiles", (long long)r);
            } else {
                CHECK(sr.b == NONE, "row %lld has no part 2", (long long)r);
            }
            if (len == 0) continue;
            GH p1, p2;
            if (sr.a0 != NONE) p1 = sr.a0 == sr.a1 ? part[r] : span_sum(sr.a0, sr.a1);
            if (sr.b != NONE) p2 = sr.b0 == sr.b1 ? part[sr.b] : span_sum(sr.b0, sr.b1);
            got[r] = {p1.g + p2.g, p1.h + p2.h};
            done[r] = 1;
        }
        // 3. tiles holding second parts only
        const uint64_t first_cold = vptr[rows];
        CHECK(T.cold_tile0 == (first_cold + TS - 1) / TS, "cold_tile0");
        for (uint64_t p = 0; p < nnz; ++p) {
            if (p < first_cold) CHECK(tidx[p] < split_at, "first-part element %llu beyond the split", (unsigned long long)p);
            else CHECK(tidx[p] >= split_at, "second-part element %llu below the split", (unsigned long long)p);
        }
    }
    // direct pass over the layout
    uint64_t nonempty = 0;
    for (int64_t r = 0; r < rows; ++r) {
        if (ptr[r + 1] == ptr[r]) {
            CHECK(done[r] == 0, "empty row %lld finished", (long long)r);
            continue;
        }
        ++nonempty;
        GH d;
        for (uint32_t e = ptr[r]; e < ptr[r + 1]; ++e) {
            d.g += x[idx[e]] * val[e];
            d.h += x[idx[e]] * x[idx[e]];
        }
        CHECK(done[r] == 1, "row %lld finished %d times", (long long)r, done[r]);
        CHECK(got[r].g == d.g && got[r].h == d.h, "row %lld: tiled (%.17g, %.17g), direct (%.17g, %.17g)", (long long)r,
              got[r].g, got[r].h, d.g, d.h);
    }
    if (failures) {
        fprintf(stderr, "%d failures\n", failures);
        return 1;
    }
    printf("OK sigma=%d split_at=%u rows=%lld nnz=%llu tiles=%llu spans=%zu short=%u segrows=%zu nonempty=%llu ib=%d\n",
           sigma, split_at, (long long)rows, (unsigned long long)nnz, (unsigned long long)ntiles, T.spans.size(),
           T.nshort, T.segrows.size(), (unsigned long long)nonempty, ib);
    return 0;
}
