// Batched CCD++: many independent FP32 factorizations in one persistent
// launch (the batching extension; the reference solves one matrix per
// process, src/main.cpp:38-64).
//
// Every problem owns a team of blocks.  A team runs its problem's whole
// outer/rank/inner loop exactly as ccd_run does for one matrix: the same
// tiled passes (tile_body, shared with tiled_kernels.hip), span finishes,
// rank commits and final subtraction, in the same order and with the same
// tile-to-lane assignment, so every problem's W, H and residual are
// bit-identical to a single-problem tiled run with the same sigma.  Between
// dependent phases the team meets at a barrier: a per-problem counter in
// global memory that each block bumps once and waits on (monotonic targets,
// so it is never reset inside the kernel).  Teams never wait for each other.
//
// Termination: the host sizes the grid to what is co-resident on the device
// (occupancy query), and every barrier wait is bounded by wall_clock64(); a
// wait that runs out sets *status and the whole team leaves the kernel, so
// the grid always drains.
//
// Block placement: a block takes virtual id x * B/8 + (its ticket on XCD
// x), so the virtual ids [x B/8, (x+1) B/8) run on XCD x.  The host packs
// every team into one such range: a team shares one L2 (cheap barriers, see
// team_sync) and keeps its problem's gathered vectors there.
#include "tile_body.h"

namespace ccd {
namespace {

constexpr int WAVE = tiled::WAVE;
constexpr int WPB = BATCH_BLOCK / WAVE;

struct Team {
    uint32_t* cnt;
    uint32_t nb;
    uint32_t target;
    int* status;
    uint64_t wait_ticks;
};

// Barrier of the blocks of one team; all of a team's blocks run on one XCD
// and share its L2.  Release: every wave waits until its stores are
// acknowledged (s_waitcnt vmcnt(0)), then one arrival per block is counted
// with an L2 atomic.  Acquire: after the count is complete, an agent-scope
// acquire fence drops the CU's stale cached copies, so the loads that follow
// see the other blocks' data.  (No L2 write-back is needed inside one XCD;
// an agent-scope release fence, which does one, costs 30-40 us a barrier.)
// false: the wait ran out (status set); the caller returns.
__device__ __forceinline__ bool team_sync(Team& t) {
    __shared__ int s_ok;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (t.nb == 1) return true;
    t.target += t.nb;
    if (threadIdx.x == 0) {
        int ok = 1;
        __hip_atomic_fetch_add(t.cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const long long t0 = wall_clock64();
        while (__hip_atomic_load(t.cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < t.target) {
            __builtin_amdgcn_s_sleep(1);
            if ((uint64_t)(wall_clock64() - t0) > t.wait_ticks) {
                __hip_atomic_store(t.status, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

// The gathered and row-side vectors of one pass.
struct PassVecs {
    const float* x;
    float* y;
    const float* xs;
    const float* ys;
    const float* xa;
    const float* ya;
    uint32_t xcur;
    float lambda;
};

__device__ __forceinline__ TiledArgs layout_args(const BatchLayout& L, const PassVecs& v) {
    TiledArgs a;
    a.ptr = L.ptr;
    a.tidx = L.tidx;
    a.tval = L.tval;
    a.x = v.x;
    a.y = v.y;
    a.xs = v.xs;
    a.ys = v.ys;
    a.xa = v.xa;
    a.ya = v.ya;
    a.xp = L.pair;
    a.lane_flags = L.lane_flags;
    a.tile_row0 = L.tile_row0;
    a.tile_segoff = L.tile_segoff;
    a.segrows = L.segrows;
    a.tile_open = L.tile_open;
    a.tile_gmax = nullptr;  // one tile_core copy here: every tile takes the general path
    a.cold_row = nullptr;   // the batched kernel works on plain layouts
    a.nreal = 0;
    a.cold_tile0 = 0xffffffffu;
    a.spread_tile0 = 0xffffffffu;
    a.part = nullptr;
    a.head = L.head;
    a.tail = L.tail;
    a.nnz = L.nnz;
    a.ntiles = L.ntiles;
    a.hot_x = 0;  // set per pass by the caller (LDS staging)
    a.hot_p = 0;
    a.x_is_cur = v.xcur;
    a.lambda = v.lambda;
    return a;
}

// What a block staged in its LDS for one pass: the first hot_x entries of
// the gathered vector x, or the first hot_p (previous, current) pairs.
struct Staged {
    const float* lds;
    const FPair<float>* ldsp;
    uint32_t hot_x, hot_p;
};

// Who takes which tile of a pass.  Wave w of the team's nw takes tiles w,
// w + nw, ...; or, for a sweep over a layout with at least ba.deal_min
// tiles per wave (dyn), block r of the team's nb owns tiles r, r + nb, ... and its
// waves take the next one from the block's counter in LDS (*pos, set to
// WPB by the caller: positions 0 .. WPB-1 go to the waves in order) when
// they finish one, so they finish within about a tile of each other (the
// largest problems bound the batch).  Which wave runs a tile changes
// nothing in its results.
struct Deal {
    uint32_t w, nw;  // static
    uint32_t r, nb;  // dynamic
    uint32_t* pos;   // LDS
    bool dyn;
};

template <typename IdxT, int SIGMA, bool SUB, bool ADD, bool SWEEP, bool XCUR>
__device__ __forceinline__ void tile_loop(const TiledArgs& a, const float* lds, const FPair<float>* ldsp,
                                          const Deal& d) {
    const int lane = threadIdx.x & (WAVE - 1);
    if (!SWEEP || !d.dyn) {
        const uint32_t w = __builtin_amdgcn_readfirstlane(d.w);  // wave-uniform: scalar descriptor loads
        for (uint32_t t = w; t < a.ntiles; t += d.nw)
            tiled::tile_body<IdxT, SIGMA, float, float, SUB, ADD, SWEEP, XCUR, false>(a, lds, ldsp, t, lane);
        return;
    }
    const uint32_t cnt = d.r < a.ntiles ? (a.ntiles - 1 - d.r) / d.nb + 1 : 0u;
    uint32_t p = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    while (p < cnt) {
        // the wave's next position, taken before this tile runs; positions
        // only grow, so every wave leaves the loop
        uint32_t pn = 0;
        if (lane == 0) pn = atomicAdd(d.pos, 1u);
        pn = __builtin_amdgcn_readfirstlane(pn);
        tiled::tile_body<IdxT, SIGMA, float, float, SUB, ADD, SWEEP, XCUR, false>(a, lds, ldsp, d.r + p * d.nb,
                                                                                  lane);
        p = pn;
    }
}

template <int SIGMA, bool SUB, bool ADD, bool SWEEP, bool XCUR>
__device__ __forceinline__ void tiles_t(const BatchLayout* Lp, PassVecs v, const Staged& st, const Deal& d) {
    const BatchLayout& L = *Lp;
    TiledArgs a = layout_args(L, v);
    a.hot_x = st.hot_x;
    a.hot_p = st.hot_p;
    if (L.ib == 2) tile_loop<uint16_t, SIGMA, SUB, ADD, SWEEP, XCUR>(a, st.lds, st.ldsp, d);
    else if (L.ib == 3) tile_loop<Idx24, SIGMA, SUB, ADD, SWEEP, XCUR>(a, st.lds, st.ldsp, d);
    else tile_loop<uint32_t, SIGMA, SUB, ADD, SWEEP, XCUR>(a, st.lds, st.ldsp, d);
}

// Modes used by the loop: plain sweep, fused sweeps (subtract, or subtract
// and add; x = pair's current value or separate), the add-back alone and the
// final subtraction.  (The first rank of a continued run adds back in a pass
// of its own followed by plain sweeps: the same operations in the same order
// as the fused add-only pass of the single-problem path, which is left out
// here because it makes the compiler spill the whole kernel.)
template <int SIGMA>
__device__ __forceinline__ void tiles(const BatchLayout* L, const PassVecs& v, const Staged& st, bool sub, bool add,
                                      bool sweep, const Deal& d) {
    const int code = (sub ? 4 : 0) | (add ? 2 : 0) | (sweep ? 1 : 0) | (sweep && (sub || add) && v.xcur ? 8 : 0);
    switch (code) {
        case 1: tiles_t<SIGMA, false, false, true, false>(L, v, st, d); break;
        case 2: tiles_t<SIGMA, false, true, false, false>(L, v, st, d); break;
        case 4: tiles_t<SIGMA, true, false, false, false>(L, v, st, d); break;
        case 5: tiles_t<SIGMA, true, false, true, false>(L, v, st, d); break;
        case 7: tiles_t<SIGMA, true, true, true, false>(L, v, st, d); break;
        case 8 + 5: tiles_t<SIGMA, true, false, true, true>(L, v, st, d); break;
        case 8 + 7: tiles_t<SIGMA, true, true, true, true>(L, v, st, d); break;
        default: break;  // not produced by the loop below
    }
}

// Stage the hot prefix a sweep pass gathers (the caller has passed a
// barrier or __syncthreads since the LDS was last read): x for plain
// sweeps and for fused passes with a separate x, the pairs for fused passes
// whose x is the pair's current value.  Empty when the layout does not
// stage.
__device__ __forceinline__ Staged stage_pass(const BatchLayout& L, const PassVecs& v, bool fused, float* lds,
                                             uint32_t cap) {
    Staged st{lds, reinterpret_cast<const FPair<float>*>(lds), 0u, 0u};
    if (!L.stage) return st;
    if (fused && v.xcur) {
        st.hot_p = min(L.gdim, cap / 2);
        const FPair<float>* src = static_cast<const FPair<float>*>(L.pair);
        FPair<float>* dst = reinterpret_cast<FPair<float>*>(lds);
        stage_lds<BATCH_BLOCK>(dst, src, st.hot_p);
    } else {
        st.hot_x = min(L.gdim, cap);
        stage_lds<BATCH_BLOCK>(lds, v.x, st.hot_x);
    }
    __syncthreads();
    return st;
}

template <int SIGMA>
__global__ __launch_bounds__(BATCH_BLOCK) void batch_kernel(BatchArgs ba) {
    extern __shared__ __align__(16) float batch_lds[];
    // Virtual block id: the XCD this block runs on (hardware XCC id
    // register) and a ticket among that XCD's blocks.  Workgroups are dealt
    // to the 8 XCDs round-robin, so each XCD receives nblocks / 8 of them
    // (the starting XCD depends on earlier launches, hence the tickets).
    __shared__ uint32_t s_v;
    if (threadIdx.x == 0) {
        const uint32_t C = ba.nblocks / 8u;
        const uint32_t xcc = __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20);  // HW_REG_XCC_ID[3:0]
        uint32_t v = 0xffffffffu;
        if (xcc < 8u) {
            const uint32_t slot = __hip_atomic_fetch_add(ba.xcd_tickets + 16u * xcc, 1u, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT);
            if (slot < C) v = xcc * C + slot;
        }
        if (v == 0xffffffffu) __hip_atomic_store(ba.status, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_v = v;
    }
    __syncthreads();
    const uint32_t v = s_v;
    if (v == 0xffffffffu) return;
    const uint32_t p = ba.block_problem[v];
    if (p >= ba.nprob) return;
    const BatchProblem& P = ba.probs[p];
    const uint32_t rank = v - P.first;
    Team tm{ba.barriers + p, P.nblocks, 0u, ba.status, ba.wait_ticks};
    const int lane = threadIdx.x & (WAVE - 1);
    const uint32_t w = rank * WPB + threadIdx.x / WAVE, nw = P.nblocks * WPB;
    const uint32_t tid = rank * BATCH_BLOCK + threadIdx.x, nth = P.nblocks * BATCH_BLOCK;
    const uint32_t dF = P.F.rows, dS = P.S.rows;
    const float lambda = P.lambda;

    const Staged none{nullptr, nullptr, 0u, 0u};

    __shared__ uint32_t s_pos;  // the block's next tile position in a dealt sweep
    const Deal fixed{w, nw, 0u, 0u, nullptr, false};

    // one sweep (or fused) pass: LDS staging, tiles, barrier, rows spanning
    // tiles, barrier
    auto pass = [&](const BatchLayout& L, const PassVecs& pv, bool sub, bool add) -> bool {
        __syncthreads();  // every wave of the block is done with the LDS of the last pass
        const bool dyn = (uint64_t)L.ntiles >= (uint64_t)ba.deal_min * nw;
        if (dyn && threadIdx.x == 0) s_pos = WPB;
        const Staged st = stage_pass(L, pv, sub || add, batch_lds, ba.lds_floats);
        if (dyn) __syncthreads();  // s_pos seen by every wave
        tiles<SIGMA>(&L, pv, st, sub, add, true, Deal{w, nw, rank, P.nblocks, &s_pos, dyn});
        if (!team_sync(tm)) return false;
        if (L.nspan) {
            const SpanArgs sa{L.spans, L.head, L.tail, pv.y, L.nspan, L.nshort, lambda};
            for (uint32_t i = tid; i < L.nshort; i += nth) tiled::span_short<float, float>(sa, i);
            for (uint32_t i = L.nshort + w; i < L.nspan; i += nw) tiled::span_row<float, float>(sa, i, lane);
            if (!team_sync(tm)) return false;
        }
        return true;
    };
    auto commit = [&](const CommitArgs& ca) -> bool {
        for (uint64_t e = tid; e < ca.nS + ca.nF; e += nth) commit_elem<float>(ca, e);
        return team_sync(tm);
    };
    auto plain = [&](const float* x, float* y) { return PassVecs{x, y, nullptr, nullptr, nullptr, nullptr, 0u, lambda}; };

    if (ba.trial) {  // autotuner: plain sweeps of rank 0; no state changes but vF, vS
        for (uint32_t it = 0; it < P.trial_reps; ++it) {
            if (!pass(P.F, plain(P.XS, P.vF), false, false)) return;
            if (!pass(P.S, plain(P.vF, P.vS), false, false)) return;
        }
        return;
    }

    bool pending = false;  // rank prev's final u v^T still to be subtracted
    uint32_t prev = 0;
    for (uint32_t oiter = 1; oiter <= P.T; ++oiter) {
        for (uint32_t t = 0; t < P.k; ++t) {
            const float* St = P.XS + (size_t)t * dS;
            const float* Ft = P.XF + (size_t)t * dF;
            const float* Sp = P.XS + (size_t)prev * dS;
            const float* Fp = P.XF + (size_t)prev * dF;
            const bool add = oiter > 1 || P.continued;
            if (add && !pending) {  // rank 0 of a continued run: add back, then plain sweeps
                tiles<SIGMA>(&P.F, PassVecs{nullptr, nullptr, nullptr, nullptr, St, Ft, 0u, lambda}, none, false, true,
                             false, fixed);
                tiles<SIGMA>(&P.S, PassVecs{nullptr, nullptr, nullptr, nullptr, Ft, St, 0u, lambda}, none, false, true,
                             false, fixed);
                // a dealt sweep may give a tile to another wave than the
                // add-back did
                if (!team_sync(tm)) return;
                if (!pass(P.F, plain(St, P.vF), false, false)) return;
                if (!pass(P.S, plain(P.vF, P.vS), false, false)) return;
            } else {
                if (!pass(P.F, PassVecs{St, P.vF, Sp, Fp, St, Ft, 1u, lambda}, pending, add)) return;
                if (!pass(P.S, PassVecs{P.vF, P.vS, Fp, Sp, Ft, St, 0u, lambda}, pending, add)) return;
            }
            for (uint32_t it = 2; it <= P.Tin; ++it) {
                if (!pass(P.F, plain(P.vS, P.vF), false, false)) return;
                if (!pass(P.S, plain(P.vF, P.vS), false, false)) return;
            }
            const bool last = oiter == P.T && t + 1 == P.k;
            const uint32_t tn = t + 1 == P.k ? 0u : t + 1;
            const CommitArgs ca{P.vS, P.XS + (size_t)t * dS, nullptr, P.XS + (size_t)tn * dS,
                                last ? nullptr : P.F.pair, dS,
                                P.vF, P.XF + (size_t)t * dF, nullptr, P.XF + (size_t)tn * dF,
                                last ? nullptr : P.S.pair, dF};
            if (!commit(ca)) return;
            pending = true;
            prev = t;
        }
    }
    if (pending) {  // rank prev's final subtraction; the two layouts are independent
        const float* Sp = P.XS + (size_t)prev * dS;
        const float* Fp = P.XF + (size_t)prev * dF;
        tiles<SIGMA>(&P.F, PassVecs{nullptr, nullptr, Sp, Fp, nullptr, nullptr, 0u, lambda}, none, true, false, false,
                     fixed);
        tiles<SIGMA>(&P.S, PassVecs{nullptr, nullptr, Fp, Sp, nullptr, nullptr, 0u, lambda}, none, true, false, false,
                     fixed);
    }
}

}  // namespace

hipError_t launch_batch(const BatchArgs& a, int sigma, hipStream_t s) {
    if (a.nblocks == 0 || a.nblocks % 8u) return hipErrorInvalidValue;
    const size_t lds = (size_t)a.lds_floats * sizeof(float);
    switch (sigma) {
        case 8: hipLaunchKernelGGL(batch_kernel<8>, dim3(a.nblocks), dim3(BATCH_BLOCK), lds, s, a); break;
        case 16: hipLaunchKernelGGL(batch_kernel<16>, dim3(a.nblocks), dim3(BATCH_BLOCK), lds, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

int batch_blocks_per_cu(int sigma, size_t lds_bytes) {
    int nb = 0;
    hipError_t e = hipErrorInvalidValue;
    if (sigma == 8) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_kernel<8>, BATCH_BLOCK, lds_bytes);
    if (sigma == 16) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_kernel<16>, BATCH_BLOCK, lds_bytes);
    return e == hipSuccess && nb > 0 ? nb : 0;
}

}  // namespace ccd
