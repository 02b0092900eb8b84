// Per-tile body of the tiled sweeps, shared by the single-problem kernels
// (tiled_kernels.hip) and the batched persistent kernel (batch_kernels.hip).
// Layout and arithmetic are described at the top of tiled_kernels.hip.
#pragma once
#include "ccd_kernels.h"

namespace ccd {
namespace tiled {

constexpr int WAVE = 64;

__device__ __forceinline__ uint32_t seg_row(uint32_t seg, uint32_t row0, uint32_t segoff,
                                            const uint32_t* __restrict__ segrows) {
    return segoff == 0xffffffffu ? row0 + seg : segrows[segoff + seg];
}

template <class TA>
struct Pair {
    TA g, h;
};

template <typename TF, typename TA>
__device__ __forceinline__ void finish_row(const TiledArgs& a, uint32_t row, TA g, TA h) {
    const uint32_t d = a.ptr[row + 1] - a.ptr[row];
    stf<TF>(a.y, row, (float)(g / ((TA)a.lambda * (TA)d + h)));
}

// XCUR (fused passes only): the sweep's x is the pair's current value.
template <typename IdxT, int SIGMA, typename TF, typename TA, bool SUB, bool ADD, bool SWEEP, bool XCUR>
__device__ __forceinline__ void tile_body(const TiledArgs& a, const TF* lds, const FPair<TF>* ldsp, const uint32_t tile,
                                          const int lane) {
    constexpr bool PAIR = SWEEP && (SUB || ADD);
    constexpr uint32_t TS = WAVE * SIGMA;
    const IdxT* __restrict__ tidx = static_cast<const IdxT*>(a.tidx) + (size_t)tile * TS;
    float* __restrict__ tval = a.tval + (size_t)tile * TS;
    const uint32_t flags = a.lane_flags[(size_t)tile * WAVE + lane];
    const uint32_t seg0 = a.lane_seg0[(size_t)tile * WAVE + lane];
    const uint32_t row0 = a.tile_row0[tile];
    const uint32_t segoff = a.tile_segoff[tile];
    const uint64_t lbase = (uint64_t)tile * TS + (uint64_t)lane * SIGMA;

    // indices first: the gathers below wait for them, not for the values
    uint32_t c[SIGMA];
    float r[SIGMA];
#pragma unroll
    for (int q = 0; q < SIGMA / 4; ++q) {
        if constexpr (sizeof(IdxT) == 2) {
            const vu2 cv = __builtin_nontemporal_load(reinterpret_cast<const vu2*>(tidx + q * 256 + 4 * lane));
            c[4 * q + 0] = cv.x & 0xffffu;
            c[4 * q + 1] = cv.x >> 16;
            c[4 * q + 2] = cv.y & 0xffffu;
            c[4 * q + 3] = cv.y >> 16;
        } else {
            const vu4 cv = __builtin_nontemporal_load(reinterpret_cast<const vu4*>(tidx + q * 256 + 4 * lane));
            c[4 * q + 0] = cv.x;
            c[4 * q + 1] = cv.y;
            c[4 * q + 2] = cv.z;
            c[4 * q + 3] = cv.w;
        }
    }
#pragma unroll
    for (int q = 0; q < SIGMA / 4; ++q) {
        const vf4 rv = __builtin_nontemporal_load(reinterpret_cast<const vf4*>(tval + q * 256 + 4 * lane));
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

    uint32_t seg = seg0;
    uint32_t row = seg_row(seg, row0, segoff, a.segrows);
    float ysr = SUB ? ldf<TF>(a.ys, row) : 0.0f;
    float yar = ADD ? ldf<TF>(a.ya, row) : 0.0f;
    TA g = 0, h = 0, gA = 0, hA = 0;
    bool has_start = (flags & 1u) != 0;
#pragma unroll
    for (int i = 0; i < SIGMA; ++i) {
        if (i > 0 && ((flags >> i) & 1u)) {
            if (SWEEP) {
                if (!has_start) {  // the row that entered this lane ends here
                    gA = g;
                    hA = h;
                } else {  // a row that started inside this lane is complete
                    finish_row<TF, TA>(a, row, g, h);
                }
            }
            has_start = true;
            g = 0;
            h = 0;
            ++seg;
            row = seg_row(seg, row0, segoff, a.segrows);
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
    const uint32_t open = a.tile_open[tile];
    if (has_start) {
        const bool bit0 = (flags & 1u) != 0;
        if (!(bit0 && lane == 0)) {
            // the row that entered this lane closes at the lane's first row start
            const uint32_t seg_in = bit0 ? seg0 - 1u : seg0;
            const TA tg = bit0 ? cg : cg + gA;
            const TA th = bit0 ? ch : ch + hA;
            if (seg_in == 0 && (open & 1u)) static_cast<Pair<TA>*>(a.head)[tile] = Pair<TA>{tg, th};
            else finish_row<TF, TA>(a, seg_row(seg_in, row0, segoff, a.segrows), tg, th);
        }
    }
    if (lane == WAVE - 1) {
        // the row open at the end of the tile
        if (seg == 0 && (open & 1u)) static_cast<Pair<TA>*>(a.head)[tile] = Pair<TA>{sg, sh};
        else if (open & 2u) static_cast<Pair<TA>*>(a.tail)[tile] = Pair<TA>{sg, sh};
        else finish_row<TF, TA>(a, row, sg, sh);
    }
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
