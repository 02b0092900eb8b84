// Host-side tables of the tiled copy of a compressed layout (tile_layout.h).
// O(rows + tiles) except the position map of a split layout (O(nnz)).
#include "tile_layout.h"

#include <algorithm>

namespace ccd {

std::string build_tile_tables(const uint32_t* hptr, int64_t nrows, uint64_t nnz, const uint32_t* below, int sigma,
                              bool split, TileTables& T) {
    T = TileTables{};
    if (sigma <= 0 || sigma > 32 || sigma % 8) return "sigma must be 8, 16, 24 or 32";
    if (split && !below) return "a split layout needs the per-row split counts";
    if (nrows < 0 || (uint64_t)hptr[nrows] != nnz) return "row pointer does not end at nnz";
    const uint64_t TS = 64ull * (uint64_t)sigma;
    const uint64_t ntiles = (nnz + TS - 1) / TS;
    if (ntiles >= 0xffffffffull / 64) return "too many tiles";
    T.sigma = sigma;
    T.split = split;
    T.ntiles = ntiles;
    if (split) {
        std::vector<uint32_t>& vptr = T.vptr;
        vptr.assign(1, 0u);
        vptr.reserve((size_t)nrows * 2 + 1);
        T.srows.resize((size_t)nrows);
        for (int64_t r = 0; r < nrows; ++r) {
            if (below[r] > hptr[r + 1] - hptr[r]) return "split count exceeds a row's length";
            vptr.push_back(vptr.back() + below[r]);
        }
        for (int64_t r = 0; r < nrows; ++r) {
            const uint32_t rest = hptr[r + 1] - hptr[r] - below[r];
            SplitRow& q = T.srows[r];
            q = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, hptr[r + 1] - hptr[r]};
            if (below[r]) {
                q.a0 = (uint32_t)(vptr[r] / TS);
                q.a1 = (uint32_t)((vptr[r + 1] - 1ull) / TS);
            }
            if (rest) {
                q.b = (uint32_t)(nrows + T.cold_row.size());
                q.b0 = (uint32_t)(vptr.back() / TS);
                q.b1 = (uint32_t)((vptr.back() + rest - 1ull) / TS);
                T.cold_row.push_back((uint32_t)r);
                vptr.push_back(vptr.back() + rest);
            }
        }
        T.vpos.resize(nnz);
        for (int64_t r = 0; r < nrows; ++r)
            for (uint32_t q = 0; q < below[r]; ++q) T.vpos[vptr[r] + q] = hptr[r] + q;
        for (size_t k = 0; k < T.cold_row.size(); ++k) {
            const uint32_t r = T.cold_row[k], b = vptr[nrows + k];
            for (uint32_t q = 0; q < vptr[nrows + k + 1] - b; ++q) T.vpos[b + q] = hptr[r] + below[r] + q;
        }
        T.nreal = (uint32_t)nrows;
        T.cold_tile0 = (uint32_t)((vptr[nrows] + TS - 1) / TS);
    }
    const uint32_t* ptr = split ? T.vptr.data() : hptr;
    const int64_t rows = split ? (int64_t)T.vptr.size() - 1 : nrows;
    T.flags.assign(ntiles * 64, 0u);
    T.row0.assign(ntiles, 0u);
    T.segoff.assign(ntiles, 0u);
    T.open.assign((ntiles + 3) & ~(uint64_t)3, 0u);  // whole words: the kernels read it a word at a time
    std::vector<uint32_t> emp(rows + 1, 0u);
    for (int64_t r = 0; r < rows; ++r) {
        const bool empty = ptr[r + 1] == ptr[r];
        emp[r + 1] = emp[r] + (empty ? 1u : 0u);
        if (empty) continue;
        const uint64_t e = ptr[r], rem = e % TS;
        T.flags[(e / TS) * 64 + rem / sigma] |= 1u << (rem % sigma);
        const uint64_t t0 = e / TS, t1 = (ptr[r + 1] - 1ull) / TS;
        if (t0 != t1) T.spans.push_back({(uint32_t)r, (uint32_t)t0, (uint32_t)t1, ptr[r + 1] - ptr[r]});
    }
    int64_t r = 0;
    for (uint64_t t = 0; t < ntiles; ++t) {
        const uint64_t tbase = t * TS;
        while ((uint64_t)ptr[r + 1] <= tbase) ++r;
        T.row0[t] = (uint32_t)r;
        uint8_t o = (uint64_t)ptr[r] < tbase ? 1 : 0;
        const uint64_t tend = std::min<uint64_t>(tbase + TS, nnz) - 1;
        int64_t r2 = r;
        while ((uint64_t)ptr[r2 + 1] <= tend) ++r2;
        if ((uint64_t)ptr[r2 + 1] > tbase + TS) o |= 2;
        T.open[t] = o;
        if (r2 > r + 1 && emp[r2] - emp[r + 1] > 0) {
            T.segoff[t] = (uint32_t)T.segrows.size();
            for (int64_t q = r; q <= r2; ++q)
                if (ptr[q + 1] > ptr[q]) T.segrows.push_back((uint32_t)q);
        } else {
            T.segoff[t] = 0xffffffffu;
        }
    }
    // two-tile spans first (finished one per lane), the rest after them
    std::stable_partition(T.spans.begin(), T.spans.end(), [](const SpanRow& q) { return q.t1 == q.t0 + 1; });
    T.nshort = (uint32_t)std::count_if(T.spans.begin(), T.spans.end(), [](const SpanRow& q) { return q.t1 == q.t0 + 1; });
    return std::string();
}

}  // namespace ccd
