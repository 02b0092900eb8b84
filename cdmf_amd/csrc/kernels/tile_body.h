// Per-tile body of the tiled sweeps, shared by the single-problem kernels
// (tiled_kernels.hip) and the batched persistent kernel (batch_kernels.hip).
// Layout and arithmetic are described at the top of tiled_kernels.hip.
#pragma once
#include "ccd_kernels.h"

namespace ccd {
namespace tiled {

constexpr int WAVE = 64;

// Row of segment `seg` of a tile: consecutive rows from row0, or (SEGTAB:
// the tile holds empty rows) from the tile's slice of the segment table.
template <bool SEGTAB>
__device__ __forceinline__ uint32_t seg_row(uint32_t seg, uint32_t row0, uint32_t segoff,
                                            const uint32_t* __restrict__ segrows) {
    if constexpr (SEGTAB) return segrows[segoff + seg];
    else return row0 + seg;
}

template <class TA>
struct Pair {
    TA g, h;
};

// y_row = g / (lambda |row| + h); the row's length d comes from the tile's
// row-start flags, not from the row pointer, so no load waits behind the
// prefetched streams of the next tile.
template <typename TF, typename TA>
__device__ __forceinline__ void finish_row(const TiledArgs& a, uint32_t row, TA g, TA h, uint32_t d) {
    stf<TF>(a.y, row, (float)(g / ((TA)a.lambda * (TA)d + h)));
}

// The per-lane stream words of one tile: raw index groups (four indices a
// load), value groups, the lane's row-start flags and first segment.
template <typename IdxT>
struct IdxGroup;
template <>
struct IdxGroup<uint16_t> {
    using type = vu2;
};
template <>
struct IdxGroup<uint32_t> {
    using type = vu4;
};

template <typename IdxT, int SIGMA>
struct Streams {
    typename IdxGroup<IdxT>::type ci[SIGMA / 4];
    vf4 rv[SIGMA / 4];
    uint32_t flags, seg0;
};

template <typename IdxT, int SIGMA>
__device__ __forceinline__ void load_streams(const TiledArgs& a, const uint32_t tile, const int lane,
                                             Streams<IdxT, SIGMA>& s) {
    using IV = typename IdxGroup<IdxT>::type;
    constexpr uint32_t TS = WAVE * SIGMA;
    const IdxT* __restrict__ tidx = static_cast<const IdxT*>(a.tidx) + (size_t)tile * TS;
    const float* __restrict__ tval = a.tval + (size_t)tile * TS;
    s.flags = a.lane_flags[(size_t)tile * WAVE + lane];
    s.seg0 = a.lane_seg0[(size_t)tile * WAVE + lane];
    // indices first: the gathers wait for them, not for the values
#pragma unroll
    for (int q = 0; q < SIGMA / 4; ++q)
        s.ci[q] = __builtin_nontemporal_load(reinterpret_cast<const IV*>(tidx + q * 256 + 4 * lane));
#pragma unroll
    for (int q = 0; q < SIGMA / 4; ++q)
        s.rv[q] = __builtin_nontemporal_load(reinterpret_cast<const vf4*>(tval + q * 256 + 4 * lane));
}

// Residual update, per-row sums and row finishes of one tile (the part of
// tile_body_s after the gathers).
template <typename IdxT, int SIGMA, typename TF, typename TA, bool SUB, bool ADD, bool SWEEP, bool SEGTAB>
__device__ __forceinline__ void tile_rows(const TiledArgs& a, const uint32_t tile, const int lane, const uint32_t flags,
                                          const uint32_t seg0, const uint32_t row0, const uint32_t segoff,
                                          const uint32_t open, const uint32_t (&c)[SIGMA], float (&r)[SIGMA], const float (&xsv)[SIGMA],
                                          const float (&xav)[SIGMA], const float (&xv)[SIGMA]) {
    constexpr uint32_t TS = WAVE * SIGMA;
    float* __restrict__ tval = a.tval + (size_t)tile * TS;
    const uint64_t lbase = (uint64_t)tile * TS + (uint64_t)lane * SIGMA;
    uint32_t seg = seg0;
    uint32_t row = seg_row<SEGTAB>(seg, row0, segoff, a.segrows);
    float ysr = SUB ? ldf<TF>(a.ys, row) : 0.0f;
    float yar = ADD ? ldf<TF>(a.ya, row) : 0.0f;
    TA g = 0, h = 0, gA = 0, hA = 0;
    bool has_start = (flags & 1u) != 0;
    // st: lane index where the current row started (valid once has_start);
    // first: lane index of the first row start after index 0
    int st = 0, first = 0;
#pragma unroll
    for (int i = 0; i < SIGMA; ++i) {
        if (i > 0 && ((flags >> i) & 1u)) {
            if (SWEEP) {
                if (!has_start) {  // the row that entered this lane ends here
                    gA = g;
                    hA = h;
                    first = i;
                } else {  // a row that started inside this lane is complete
                    finish_row<TF, TA>(a, row, g, h, (uint32_t)(i - st));
                }
            }
            st = i;
            has_start = true;
            g = 0;
            h = 0;
            ++seg;
            row = seg_row<SEGTAB>(seg, row0, segoff, a.segrows);
            if (SUB) ysr = ldf<TF>(a.ys, row);
            if (ADD) yar = ldf<TF>(a.ya, row);
        }
        const bool valid = lbase + (uint64_t)i < a.nnz;
        float v = r[i];
        if (SUB) v = v - xsv[i] * ysr;
        if (ADD) v = v + xav[i] * yar;
        r[i] = v;
        if (SWEEP && valid) {
            g += (TA)xv[i] * (TA)v;
            h += (TA)xv[i] * (TA)xv[i];
        }
    }
    if (SUB || ADD) {
        // padding positions of the last tile are rewritten too; they are never read as data
#pragma unroll
        for (int q = 0; q < SIGMA / 4; ++q)
            __builtin_nontemporal_store(vf4{r[4 * q], r[4 * q + 1], r[4 * q + 2], r[4 * q + 3]},
                                        reinterpret_cast<vf4*>(tval + q * 256 + 4 * lane));
    }
    if (!SWEEP) return;

    // Segmented inclusive scan over lanes of the lane's contribution to the
    // row open at its end; a lane with a row start begins a new segment.
    TA sg = g, sh = h;
    uint32_t f = has_start ? 1u : 0u;
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        const TA pg = __shfl_up(sg, d);
        const TA ph = __shfl_up(sh, d);
        const uint32_t pf = __shfl_up(f, d);
        if (lane >= d && !f) {
            sg = pg + sg;
            sh = ph + sh;
            f = pf;
        }
    }
    TA cg = __shfl_up(sg, 1);
    TA ch = __shfl_up(sh, 1);
    if (lane == 0) {
        cg = 0;
        ch = 0;
    }
    // Tile position where the row open at the end of each lane started: a
    // running maximum over lanes of each lane's last row start (-1: the row
    // started before this tile).
    int ls = has_start ? lane * SIGMA + st : -1;
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        const int p = __shfl_up(ls, d);
        if (lane >= d && p > ls) ls = p;
    }
    int ls_in = __shfl_up(ls, 1);  // for the row entering each lane
    if (lane == 0) ls_in = -1;
    if (has_start) {
        const bool bit0 = (flags & 1u) != 0;
        if (!(bit0 && lane == 0)) {
            // the row that entered this lane closes at the lane's first row start
            const uint32_t seg_in = bit0 ? seg0 - 1u : seg0;
            const TA tg = bit0 ? cg : cg + gA;
            const TA th = bit0 ? ch : ch + hA;
            const int end = lane * SIGMA + (bit0 ? 0 : first);
            if (seg_in == 0 && (open & 1u)) static_cast<Pair<TA>*>(a.head)[tile] = Pair<TA>{tg, th};
            else finish_row<TF, TA>(a, seg_row<SEGTAB>(seg_in, row0, segoff, a.segrows), tg, th, (uint32_t)(end - ls_in));
        }
    }
    if (lane == WAVE - 1) {
        // the row open at the end of the tile; it ends at the tile's last
        // element (the padding of the last tile is not part of it)
        const uint64_t left = a.nnz - (uint64_t)tile * TS;
        const int end = left < TS ? (int)left : (int)TS;
        if (seg == 0 && (open & 1u)) static_cast<Pair<TA>*>(a.head)[tile] = Pair<TA>{sg, sh};
        else if (open & 2u) static_cast<Pair<TA>*>(a.tail)[tile] = Pair<TA>{sg, sh};
        else finish_row<TF, TA>(a, row, sg, sh, (uint32_t)(end - ls));
    }
}

// XCUR (fused passes only): the sweep's x is the pair's current value.
// `tile` is wave-uniform, so the per-tile descriptors are scalar loads.
template <typename IdxT, int SIGMA, typename TF, typename TA, bool SUB, bool ADD, bool SWEEP, bool XCUR>
__device__ __forceinline__ void tile_body_s(const TiledArgs& a, const TF* lds, const FPair<TF>* ldsp,
                                            const uint32_t tile, const int lane, const Streams<IdxT, SIGMA>& cur) {
    constexpr bool PAIR = SWEEP && (SUB || ADD);
    const uint32_t flags = cur.flags;
    const uint32_t seg0 = cur.seg0;
    const uint32_t row0 = a.tile_row0[tile];
    const uint32_t segoff = a.tile_segoff[tile];
    // the byte of this tile, read as its 32-bit word (the array is padded to
    // whole words) so that a wave-uniform tile gives a scalar load
    const uint32_t open =
        (reinterpret_cast<const uint32_t*>(a.tile_open)[tile >> 2] >> (8u * (tile & 3u))) & 0xffu;

    uint32_t c[SIGMA];
    float r[SIGMA];
#pragma unroll
    for (int q = 0; q < SIGMA / 4; ++q) {
        if constexpr (sizeof(IdxT) == 2) {
            const vu2 cv = cur.ci[q];
            c[4 * q + 0] = cv.x & 0xffffu;
            c[4 * q + 1] = cv.x >> 16;
            c[4 * q + 2] = cv.y & 0xffffu;
            c[4 * q + 3] = cv.y >> 16;
        } else {
            const vu4 cv = cur.ci[q];
            c[4 * q + 0] = cv.x;
            c[4 * q + 1] = cv.y;
            c[4 * q + 2] = cv.z;
            c[4 * q + 3] = cv.w;
        }
        const vf4 rv = cur.rv[q];
        r[4 * q + 0] = rv.x;
        r[4 * q + 1] = rv.y;
        r[4 * q + 2] = rv.z;
        r[4 * q + 3] = rv.w;
    }

    // Every gather of the lane is issued before the tile's first store, so
    // all of them are in flight together rather than one per element.
    float xsv[SIGMA], xav[SIGMA], xv[SIGMA];
#pragma unroll
    for (int i = 0; i < SIGMA; ++i) {
        if constexpr (PAIR) {
            const FPair<TF> pr = gather_pair<TF>(ldsp, a.hot_p, a.xp, c[i]);
            xsv[i] = SUB ? tof<TF>(pr.prev) : 0.0f;
            xav[i] = tof<TF>(pr.cur);
            xv[i] = XCUR ? xav[i] : gather<TF>(lds, a.hot_x, a.x, c[i]);
        } else {
            xsv[i] = SUB ? ldf<TF>(a.xs, c[i]) : 0.0f;
            xav[i] = ADD ? ldf<TF>(a.xa, c[i]) : 0.0f;
            xv[i] = SWEEP ? gather<TF>(lds, a.hot_x, a.x, c[i]) : 0.0f;
        }
    }
    // The tile's rows are consecutive unless it holds empty rows: two copies
    // of the rest, chosen by a wave-uniform branch, so the common one has no
    // segment-table loads (a load on only one of two merged paths makes the
    // compiler wait for every outstanding load at each later use).
    if (segoff == 0xffffffffu)
        tile_rows<IdxT, SIGMA, TF, TA, SUB, ADD, SWEEP, false>(a, tile, lane, flags, seg0, row0, segoff, open, c, r, xsv,
                                                             xav, xv);
    else
        tile_rows<IdxT, SIGMA, TF, TA, SUB, ADD, SWEEP, true>(a, tile, lane, flags, seg0, row0, segoff, open, c, r, xsv,
                                                            xav, xv);
}

// One tile: its streams, then the body.
template <typename IdxT, int SIGMA, typename TF, typename TA, bool SUB, bool ADD, bool SWEEP, bool XCUR>
__device__ __forceinline__ void tile_body(const TiledArgs& a, const TF* lds, const FPair<TF>* ldsp, const uint32_t tile,
                                          const int lane) {
    Streams<IdxT, SIGMA> s;
    load_streams<IdxT, SIGMA>(a, tile, lane, s);
    tile_body_s<IdxT, SIGMA, TF, TA, SUB, ADD, SWEEP, XCUR>(a, lds, ldsp, tile, lane, s);
}

// Rows spanning tiles t0..t1: tail[t0] + head[t0+1] + ... + head[t1], by one
// wave, lane-strided partial sums reduced in a fixed tree.
template <typename TF, typename TA>
__device__ __forceinline__ void span_row(const SpanArgs& a, uint32_t i, int lane) {
    const uint32_t row = a.span_rows[i];
    const uint32_t s = a.ptr[row], e = a.ptr[row + 1];
    const uint32_t t0 = s / a.tile_elems, t1 = (e - 1) / a.tile_elems;
    const Pair<TA>* __restrict__ head = static_cast<const Pair<TA>*>(a.head);
    TA g = 0, h = 0;
    for (uint32_t t = t0 + 1 + lane; t <= t1; t += WAVE) {
        const Pair<TA> p = head[t];
        g += p.g;
        h += p.h;
    }
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) {
        g += __shfl_xor(g, off);
        h += __shfl_xor(h, off);
    }
    if (lane == 0) {
        const Pair<TA> p = static_cast<const Pair<TA>*>(a.tail)[t0];
        g = p.g + g;
        h = p.h + h;
        stf<TF>(a.y, row, (float)(g / ((TA)a.lambda * (TA)(e - s) + h)));
    }
}

}  // namespace tiled
}  // namespace ccd
