// Per-tile body of the tiled sweeps, shared by the single-problem kernels
// (tiled_kernels.hip) and the batched persistent kernel (batch_kernels.hip).
// Layout and arithmetic are described at the top of tiled_kernels.hip.
#pragma once
#include "ccd_kernels.h"

#include <type_traits>

namespace ccd {
namespace tiled {

constexpr int WAVE = 64;

// Row of segment `seg` of a tile: consecutive rows from row0, or, for a
// tile that holds empty rows, from the tile's slice of the segment table.
// SEG: SEG_RUN / SEG_TABLE when the caller knows which, SEG_ANY to decide
// per tile from segoff.
constexpr int SEG_RUN = 0, SEG_TABLE = 1, SEG_ANY = 2;
template <int SEG>
__device__ __forceinline__ uint32_t seg_row(uint32_t seg, uint32_t row0, uint32_t segoff,
                                            const uint32_t* __restrict__ segrows) {
    if constexpr (SEG == SEG_TABLE) return segrows[segoff + seg];
    else if constexpr (SEG == SEG_RUN) return row0 + seg;
    else return segoff == 0xffffffffu ? row0 + seg : segrows[segoff + seg];
}

template <class TA>
struct Pair {
    TA g, h;
};

// y_row = g / (lambda |row| + h); the row's length d comes from the tile's
// row-start flags, not from two row-pointer loads per finished row.  In a
// split layout the virtual row's (g, h) is kept for split_finish_kernel.
template <typename TF, typename TA>
__device__ __forceinline__ void finish_row(const TiledArgs& a, uint32_t row, TA g, TA h, uint32_t d) {
    if (a.part) static_cast<Pair<TA>*>(a.part)[row] = Pair<TA>{g, h};
    else stf<TF>(a.y, row, (float)(g / ((TA)a.lambda * (TA)d + h)));
}

// Row whose per-row vectors a (virtual) row of the tiles uses.
__device__ __forceinline__ uint32_t real_row(const TiledArgs& a, uint32_t v) {
    return a.cold_row && v >= a.nreal ? a.cold_row[v - a.nreal] : v;
}

// Index plane element: u16 (also the low plane of the 3-byte format) or
// u32; either way a lane reads 16 bytes a load (8 or 4 indices).
template <typename IdxT>
struct IdxGroup {
    using elem = uint16_t;
};
template <>
struct IdxGroup<uint32_t> {
    using elem = uint32_t;
};

// High bytes of the 3-byte format: the lane's SIGMA bytes, one or two 16-byte
// loads (8 bytes for SIGMA = 8).
template <typename IdxT, int SIGMA>
struct HiPlane {
    __device__ __forceinline__ void load(const TiledArgs&, uint32_t, int) {}
    __device__ __forceinline__ uint32_t at(int) const { return 0u; }
};
template <int SIGMA>
struct HiPlane<Idx24, SIGMA> {
    uint32_t w[SIGMA / 4];
    __device__ __forceinline__ void load(const TiledArgs& a, uint32_t tile, int lane) {
        constexpr uint32_t TS = WAVE * SIGMA;
        const uint8_t* hi = reinterpret_cast<const uint8_t*>(static_cast<const uint16_t*>(a.tidx) +
                                                             (size_t)a.ntiles * TS) +
                            (size_t)tile * TS + (size_t)lane * SIGMA;
        if constexpr (SIGMA == 8) {
            const vu2 v = __builtin_nontemporal_load(reinterpret_cast<const vu2*>(hi));
            w[0] = v.x;
            w[1] = v.y;
        } else {
#pragma unroll
            for (int q = 0; q < SIGMA / 16; ++q) {
                const vu4 v = __builtin_nontemporal_load(reinterpret_cast<const vu4*>(hi) + q);
                w[4 * q + 0] = v.x;
                w[4 * q + 1] = v.y;
                w[4 * q + 2] = v.z;
                w[4 * q + 3] = v.w;
            }
        }
    }
    // bits 16..23 of element i, in place
    __device__ __forceinline__ uint32_t at(int i) const { return ((w[i >> 2] >> (8 * (i & 3))) & 0xffu) << 16; }
};

// Segment (row of the tile, counted from 0) of a lane's element 0, from the
// lanes' row-start flags: the row starts in earlier lanes (a wave prefix sum
// of popcounts), less one when the tile's first element starts a row, plus
// one when the lane's element 0 starts a row; lane 0 is segment 0.
__device__ __forceinline__ uint32_t lane_seg0(uint32_t flags, int lane) {
    const uint32_t pc = (uint32_t)__popc(flags);
    uint32_t incl = pc;
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        const uint32_t t = __shfl_up(incl, d);
        if (lane >= d) incl += t;
    }
    const uint32_t start0 = (uint32_t)__shfl(flags, 0) & 1u;
    return lane == 0 ? 0u : incl - pc - start0 + (flags & 1u);
}

// The lane's streams of one tile, unpacked: SIGMA indices and values, its
// row-start flags and first segment.  Indices are loaded first: the gathers
// wait for them, not for the values.  HI: the high bytes of the 3-byte
// format are loaded and added; a tile whose indices are all below 2^16
// leaves them out.
template <typename IdxT, int SIGMA, bool ORDERED, bool HI = true>
__device__ __forceinline__ void load_tile(const TiledArgs& a, const uint32_t tile, const int lane, uint32_t (&c)[SIGMA],
                                          float (&r)[SIGMA], uint32_t& flags, uint32_t& seg0) {
    using IE = typename IdxGroup<IdxT>::elem;
    constexpr int G = 16 / (int)sizeof(IE);  // indices per load
    constexpr uint32_t TS = WAVE * SIGMA;
    const IE* __restrict__ tidx = static_cast<const IE*>(a.tidx) + (size_t)tile * TS;
    const float* __restrict__ tval = a.tval + (size_t)tile * TS;
    if constexpr (SIGMA <= 16) flags = static_cast<const uint16_t*>(a.lane_flags)[(size_t)tile * WAVE + lane];
    else flags = static_cast<const uint32_t*>(a.lane_flags)[(size_t)tile * WAVE + lane];
    // ORDERED: keep this load ahead of the streams (the row lookups wait for
    // it, and loads complete in issue order)
    if constexpr (ORDERED) __builtin_amdgcn_sched_barrier(0);
    vu4 ci[SIGMA / G];
#pragma unroll
    for (int q = 0; q < SIGMA / G; ++q)
        ci[q] = __builtin_nontemporal_load(reinterpret_cast<const vu4*>(tidx + q * G * WAVE + G * lane));
    HiPlane<IdxT, SIGMA> hi;
    if constexpr (HI) hi.load(a, tile, lane);
#pragma unroll
    for (int q = 0; q < SIGMA / 4; ++q) {
        const vf4 rv = __builtin_nontemporal_load(reinterpret_cast<const vf4*>(tval + q * 256 + 4 * lane));
        r[4 * q + 0] = rv.x;
        r[4 * q + 1] = rv.y;
        r[4 * q + 2] = rv.z;
        r[4 * q + 3] = rv.w;
    }
#pragma unroll
    for (int q = 0; q < SIGMA / G; ++q) {
        const uint32_t wd[4] = {ci[q].x, ci[q].y, ci[q].z, ci[q].w};
#pragma unroll
        for (int j = 0; j < G; ++j) {
            if constexpr (G == 8)
                c[8 * q + j] = ((wd[j >> 1] >> (16 * (j & 1))) & 0xffffu) | (HI ? hi.at(8 * q + j) : 0u);
            else c[4 * q + j] = wd[j];
        }
    }
    seg0 = lane_seg0(flags, lane);
}

// One tile after its streams are loaded: the lane's per-row vectors and
// gathers, the residual update, per-row sums and row finishes.  SEG: how the
// tile's rows are found (seg_row).  XCUR (fused passes only): the sweep's x
// is the pair's current value.  HOT (sweeps only): every gather of the tile
// is below the staged counts and the tile has no padding (it is not a
// partial last tile), so the gathers are plain LDS reads and no element is
// masked.  FLAT (plain sweeps of the batched kernel): one flat load per
// gather instead of separate LDS and global loads.
template <typename IdxT, int SIGMA, typename TF, typename TA, bool SUB, bool ADD, bool SWEEP, bool XCUR, int SEG,
          bool HOT = false, bool FLAT = false, bool COLD = false>
__device__ __forceinline__ void tile_core(const TiledArgs& a, const TF* lds, const FPair<TF>* ldsp, const uint32_t tile,
                                          const int lane, const uint32_t flags, const uint32_t seg0,
                                          const uint32_t row0, const uint32_t segoff, const uint32_t open,
                                          const uint32_t (&c)[SIGMA], float (&r)[SIGMA]) {
    constexpr bool PAIR = SWEEP && (SUB || ADD);
    static_assert(SWEEP || !HOT, "only sweeps gather from LDS");
    constexpr uint32_t TS = WAVE * SIGMA;
    float* __restrict__ tval = a.tval + (size_t)tile * TS;
    const uint64_t lbase = (uint64_t)tile * TS + (uint64_t)lane * SIGMA;
    uint32_t seg = seg0;
    uint32_t row = seg_row<SEG>(seg, row0, segoff, a.segrows);
    // Per-row vectors of the lane's first two rows (a lane of SIGMA elements
    // rarely holds more), loaded before the gathers so that their latency
    // overlaps; a third row's are loaded where it starts.
    float ysr = SUB ? ldf<TF>(a.ys, real_row(a, row)) : 0.0f;
    float yar = ADD ? ldf<TF>(a.ya, real_row(a, row)) : 0.0f;
    float ysn = 0.0f, yan = 0.0f;
    if constexpr (SUB || ADD) {
        if (flags >> 1) {
            const uint32_t rn = seg_row<SEG>(seg + 1, row0, segoff, a.segrows);
            if (SUB) ysn = ldf<TF>(a.ys, real_row(a, rn));
            if (ADD) yan = ldf<TF>(a.ya, real_row(a, rn));
        }
    }

    // Every gather of the lane is issued before the tile's first store, so
    // all of them are in flight together rather than one per element.  HOT
    // tiles read LDS only.  Other tiles read LDS for every element (entry 0
    // past the staged prefix), then overwrite the elements past it with
    // global loads.  All LDS reads are issued first, so the global loads
    // wait for them once rather than one by one (a global load into a
    // register must wait for the LDS read it overwrites).
    float xsv[SIGMA], xav[SIGMA], xv[SIGMA];
    const lds_ptr<TF> l3 = (lds_ptr<TF>)lds;
    const lds_ptr<PairWord<TF>> p3 = (lds_ptr<PairWord<TF>>)ldsp;
    if constexpr (PAIR) {
        PairWord<TF> pr[SIGMA];
        TF xg[SIGMA];
#pragma unroll
        for (int i = 0; i < SIGMA; ++i) {
            pr[i] = p3[HOT || c[i] < a.hot_p ? c[i] : 0u];
            if (!XCUR) xg[i] = l3[HOT || c[i] < a.hot_x ? c[i] : 0u];
        }
        if constexpr (!HOT) {
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int i = 0; i < SIGMA; ++i) {
                if (c[i] >= a.hot_p) pr[i] = ((glb_ptr<PairWord<TF>>)a.xp)[c[i]];
                if (!XCUR && c[i] >= a.hot_x) xg[i] = ((glb_ptr<TF>)a.x)[c[i]];
            }
        }
#pragma unroll
        for (int i = 0; i < SIGMA; ++i) {
            const FPair<TF> p = to_pair<TF>(pr[i]);
            xsv[i] = SUB ? tof<TF>(p.prev) : 0.0f;
            xav[i] = tof<TF>(p.cur);
            xv[i] = XCUR ? xav[i] : tof<TF>(xg[i]);
        }
    } else if constexpr (SWEEP && FLAT) {
#pragma unroll
        for (int i = 0; i < SIGMA; ++i) {
            xsv[i] = 0.0f;
            xav[i] = 0.0f;
            xv[i] = gather<TF>(lds, a.hot_x, a.x, c[i]);
        }
    } else if constexpr (SWEEP && COLD) {
        // every gather of the tile misses the staged prefix
        TF xg[SIGMA];
#pragma unroll
        for (int i = 0; i < SIGMA; ++i) xg[i] = ((glb_ptr<TF>)a.x)[c[i]];
#pragma unroll
        for (int i = 0; i < SIGMA; ++i) {
            xsv[i] = 0.0f;
            xav[i] = 0.0f;
            xv[i] = tof<TF>(xg[i]);
        }
    } else if constexpr (SWEEP) {
        TF xg[SIGMA];
#pragma unroll
        for (int i = 0; i < SIGMA; ++i) xg[i] = l3[HOT || c[i] < a.hot_x ? c[i] : 0u];
        if constexpr (!HOT) {
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int i = 0; i < SIGMA; ++i)
                if (c[i] >= a.hot_x) xg[i] = ((glb_ptr<TF>)a.x)[c[i]];
        }
#pragma unroll
        for (int i = 0; i < SIGMA; ++i) {
            xsv[i] = 0.0f;
            xav[i] = 0.0f;
            xv[i] = tof<TF>(xg[i]);
        }
    } else {
#pragma unroll
        for (int i = 0; i < SIGMA; ++i) {
            xsv[i] = SUB ? ldf<TF>(a.xs, c[i]) : 0.0f;
            xav[i] = ADD ? ldf<TF>(a.xa, c[i]) : 0.0f;
            xv[i] = 0.0f;
        }
    }

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
            row = seg_row<SEG>(++seg, row0, segoff, a.segrows);
            if (seg == seg0 + 1) {
                ysr = ysn;
                yar = yan;
            } else {
                if (SUB) ysr = ldf<TF>(a.ys, real_row(a, row));
                if (ADD) yar = ldf<TF>(a.ya, real_row(a, row));
            }
        }
        const bool valid = HOT || lbase + (uint64_t)i < a.nnz;
        float v = r[i];
        if (SUB) v = v - xsv[i] * ysr;
        if (ADD) v = v + xav[i] * yar;
        r[i] = v;
        if (SWEEP && valid) {
            g += (TA)xv[i] * (TA)v;
            h += (TA)xv[i] * (TA)xv[i];
        }
        // Each group of four updated values is stored as soon as it is done,
        // so the stores are complete by the time the next tile's loads reuse
        // their registers.  (Padding positions of the last tile are
        // rewritten too; they are never read as data.)
        if ((SUB || ADD) && (i & 3) == 3)
            __builtin_nontemporal_store(vf4{r[i - 3], r[i - 2], r[i - 1], r[i]},
                                        reinterpret_cast<vf4*>(tval + (i >> 2) * 256 + 4 * lane));
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
            else finish_row<TF, TA>(a, seg_row<SEG>(seg_in, row0, segoff, a.segrows), tg, th, (uint32_t)(end - ls_in));
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

// Largest gathered index for which a tile of a sweep is HOT: the staged
// count of everything the sweep gathers (x; for a fused pass the pairs, and
// x unless it is the pair's current value), at most 2^16 for 3-byte indices
// (HOT tiles skip the high bytes).
template <typename IdxT, bool SUB, bool ADD, bool XCUR>
__device__ __forceinline__ uint32_t hot_limit(const TiledArgs& a) {
    uint32_t lim;
    if constexpr (!(SUB || ADD)) lim = a.hot_x;
    else if constexpr (XCUR) lim = a.hot_p;
    else lim = a.hot_p < a.hot_x ? a.hot_p : a.hot_x;
    if constexpr (std::is_same<IdxT, Idx24>::value) lim = lim < 0x10000u ? lim : 0x10000u;
    return lim;
}

// One tile: streams, then tile_core.  `tile` is wave-uniform, so the
// per-tile descriptors are scalar loads.  SPLIT: separate code for sweep
// tiles whose gathers all hit LDS (most tiles once rows are sorted or split
// by gathered index), for the other tiles with consecutive rows and for
// tiles with empty rows (faster; the batched kernel, which inlines every
// pass mode, keeps one copy, with flat gathers, to stay within its
// registers).  gmax: the tile's tile_gmax entry (sweeps), loaded by the
// caller one tile ahead so that the choice is made before the tile's
// streams are issued (each copy issues its own).
template <typename IdxT, int SIGMA, typename TF, typename TA, bool SUB, bool ADD, bool SWEEP, bool XCUR,
          bool SPLIT = true>
__device__ __forceinline__ void tile_body(const TiledArgs& a, const TF* lds, const FPair<TF>* ldsp, const uint32_t tile,
                                          const int lane, const uint32_t gmax = 0xffffffffu) {
    // Passes that update the residual look up per-row vectors first thing,
    // so their tile descriptors and the lane's flags and first segment are
    // loaded ahead of the streams: a wait for them must not also wait for
    // the streams.
    constexpr bool EARLY = SUB || ADD;
    uint32_t row0 = 0, segoff = 0, open = 0;
    auto descriptors = [&]() {
        row0 = ld_const(a.tile_row0, tile);
        segoff = ld_const(a.tile_segoff, tile);
        // the byte of this tile, read as its 32-bit word (the array is padded
        // to whole words) so that a wave-uniform tile can be a scalar load
        open = (ld_const(reinterpret_cast<const uint32_t*>(a.tile_open), tile >> 2) >> (8u * (tile & 3u))) & 0xffu;
    };
    uint32_t c[SIGMA];
    float r[SIGMA];
    uint32_t flags, seg0;

    // A sweep tile with gmax below the staged counts has consecutive rows
    // (tiles with empty rows carry 0xffffffff) and gathers only from LDS.
    // Copies for the common cases are chosen by wave-uniform branches, so
    // the common one has no segment-table loads (a load on only one of two
    // merged paths makes the compiler wait for every outstanding load at
    // each later use), no global gathers and, for 3-byte indices, no high
    // bytes.  Plain sweeps: a tile of second parts only (consecutive
    // non-empty rows, no padding) gathers straight from global memory.
    // Fused passes keep only the general copies (per-tile choice of row
    // lookup, LDS or global gathers; with 3-byte indices, with or without
    // the high bytes): every copy inlined costs scalar registers, and past
    // the 102 a wave has, the compiler parks values in vector lanes and
    // reads them back with a VALU op at each use.
    if constexpr (SPLIT && SWEEP && !SUB && !ADD) {
        if (tile >= a.cold_tile0 && gmax != 0xffffffffu) {
            load_tile<IdxT, SIGMA, false>(a, tile, lane, c, r, flags, seg0);
            descriptors();
            tile_core<IdxT, SIGMA, TF, TA, SUB, ADD, SWEEP, XCUR, SEG_RUN, true, false, true>(
                a, lds, ldsp, tile, lane, flags, seg0, row0, segoff, open, c, r);
            return;
        }
    }
    if (SPLIT && SWEEP && !(SUB || ADD) && gmax < hot_limit<IdxT, SUB, ADD, XCUR>(a)) {
        if constexpr (EARLY) descriptors();
        load_tile<IdxT, SIGMA, EARLY, false>(a, tile, lane, c, r, flags, seg0);
        if constexpr (!EARLY) descriptors();
        tile_core<IdxT, SIGMA, TF, TA, SUB, ADD, SWEEP, XCUR, SEG_RUN, SWEEP>(a, lds, ldsp, tile, lane, flags, seg0,
                                                                            row0, segoff, open, c, r);
        return;
    }
    // 3-byte indices all below 2^16 (with consecutive rows: gmax is below
    // 0xffffffff) but not all staged, as in the first parts of a split
    // layout under a fused pass, whose pairs take twice the LDS: the general
    // gathers without the high-byte plane.
    if constexpr (SPLIT && SWEEP && std::is_same<IdxT, Idx24>::value) {
        if (gmax < 0x10000u) {
            if constexpr (EARLY) descriptors();
            load_tile<IdxT, SIGMA, EARLY, false>(a, tile, lane, c, r, flags, seg0);
            if constexpr (!EARLY) descriptors();
            tile_core<IdxT, SIGMA, TF, TA, SUB, ADD, SWEEP, XCUR, (SUB || ADD) ? SEG_ANY : SEG_RUN>(
                a, lds, ldsp, tile, lane, flags, seg0, row0, segoff, open, c, r);
            return;
        }
    }
    if constexpr (EARLY) descriptors();
    load_tile<IdxT, SIGMA, EARLY>(a, tile, lane, c, r, flags, seg0);
    if constexpr (!EARLY) descriptors();
    if constexpr (!SPLIT)
        tile_core<IdxT, SIGMA, TF, TA, SUB, ADD, SWEEP, XCUR, SEG_ANY, false, true>(a, lds, ldsp, tile, lane, flags,
                                                                                  seg0, row0, segoff, open, c, r);
    else
        tile_core<IdxT, SIGMA, TF, TA, SUB, ADD, SWEEP, XCUR, SEG_ANY>(a, lds, ldsp, tile, lane, flags, seg0, row0,
                                                                     segoff, open, c, r);
}

// A row spanning two tiles, by one lane: tail[t0] + head[t0 + 1], rounded
// exactly as span_row rounds it (the head first passes through a sum with
// zero, as in span_row's tree).
template <typename TF, typename TA>
__device__ __forceinline__ void span_short(const SpanArgs& a, uint32_t i) {
    const SpanRow sp = a.spans[i];
    const Pair<TA> tl = static_cast<const Pair<TA>*>(a.tail)[sp.t0];
    const Pair<TA> hd = static_cast<const Pair<TA>*>(a.head)[sp.t1];
    const TA g = tl.g + ((TA)0 + hd.g);
    const TA h = tl.h + ((TA)0 + hd.h);
    if (a.part) static_cast<Pair<TA>*>(a.part)[sp.row] = Pair<TA>{g, h};
    else stf<TF>(a.y, sp.row, (float)(g / ((TA)a.lambda * (TA)sp.len + h)));
}

// tail[t0] + head[t0+1] + ... + head[t1] by one wave: lane-strided partial
// sums reduced in a fixed tree; the total is valid on lane 0.
template <typename TA>
__device__ __forceinline__ Pair<TA> span_sum(const void* headp, const void* tailp, uint32_t t0, uint32_t t1,
                                             int lane) {
    const Pair<TA>* __restrict__ head = static_cast<const Pair<TA>*>(headp);
    Pair<TA> tl{0, 0};
    if (lane == 0) tl = static_cast<const Pair<TA>*>(tailp)[t0];
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
    return Pair<TA>{tl.g + g, tl.h + h};
}

// Rows spanning tiles t0..t1, by one wave (span_sum).  The span table holds
// the tile range and length, so the partials are the only dependent loads.
template <typename TF, typename TA>
__device__ __forceinline__ void span_row(const SpanArgs& a, uint32_t i, int lane) {
    const SpanRow sp = a.spans[i];
    const Pair<TA> s = span_sum<TA>(a.head, a.tail, sp.t0, sp.t1, lane);
    if (lane == 0) {
        if (a.part) static_cast<Pair<TA>*>(a.part)[sp.row] = s;
        else stf<TF>(a.y, sp.row, (float)(s.g / ((TA)a.lambda * (TA)sp.len + s.h)));
    }
}

// Row r of a split layout, by one wave (SplitArgs).
template <typename TF, typename TA>
__device__ __forceinline__ void split_row(const SplitArgs& a, uint32_t r, int lane) {
    const SplitRow sr = a.rows[r];
    if (sr.len == 0) return;
    const Pair<TA>* P = static_cast<const Pair<TA>*>(a.part);
    Pair<TA> p1{0, 0}, p2{0, 0};
    if (sr.a0 != 0xffffffffu) p1 = sr.a0 == sr.a1 ? P[r] : span_sum<TA>(a.head, a.tail, sr.a0, sr.a1, lane);
    if (sr.b != 0xffffffffu) p2 = sr.b0 == sr.b1 ? P[sr.b] : span_sum<TA>(a.head, a.tail, sr.b0, sr.b1, lane);
    if (lane != 0) return;
    TA g, h;
    if (sr.a0 != 0xffffffffu) {
        g = p1.g;
        h = p1.h;
        if (sr.b != 0xffffffffu) {
            g = g + p2.g;
            h = h + p2.h;
        }
    } else {
        g = p2.g;
        h = p2.h;
    }
    stf<TF>(a.y, r, (float)(g / ((TA)a.lambda * (TA)sr.len + h)));
}

}  // namespace tiled
}  // namespace ccd
