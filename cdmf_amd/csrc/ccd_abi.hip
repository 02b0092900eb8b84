// C ABI of the gfx950 CCD++ solver (include/cdmf_hip.h).
//
// Device state of one context: both compressed copies of R (pointer, index,
// residual value and a pristine copy of the values), the factors W [k][m]
// and H [k][n] (rank-major, as MatData [t][i] in src/pmf_util.h:7-8), the
// working vectors u, v, the test set and, per layout, the wave work plan of
// the binned sweeps.
//
// The outer loop follows cdmf_ref.cpp:69-137 step for step, with one
// reordering that leaves every rounded value unchanged: the subtraction of
// rank t-1's final u v^T (cdmf_ref.cpp:120-121) is deferred and done in the
// same pass as the add-back of rank t (cdmf_ref.cpp:86-89) and the first
// sweep of rank t.  Per element the arithmetic is the reference's,
// (val - u_{t-1} v_{t-1}) + w_t h_t, each product and sum rounded to FP32.
// After the last rank the pending subtraction is applied on its own, so the
// residual left on the device equals the one the reference leaves in R.
//
// Internal numbering: rows and columns are renumbered by falling degree at
// upload (stable, so ties keep their order).  The highest-degree rows and
// columns -- the entries gathered most often -- become a contiguous prefix
// of every factor vector, which the sweeps stage in LDS, and the sweep work
// units become contiguous row ranges.  When the whole gathered vector of a
// layout fits the LDS prefix of a plain sweep, the elements of each row are
// sorted by internal index as well: the fused passes, which split their LDS
// between two gathered arrays, then find most tiles of the tiled copy
// gathering only from LDS.  (For a larger gathered vector the sort would
// pack the gathers beyond the prefix into the row ends, which measured
// slower, so element order is kept there.)  Sorting reorders the terms of
// each per-row sum, which are trees anyway, not the reference's serial
// order.  W, H, the test set and the residual are mapped at the boundary;
// callers only ever see their own numbering.
#include "ccd_ctx.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <map>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace ccdi {

int set_err(ccd_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

void dfree(void* p) {
    if (p) (void)hipFree(p);
}

void free_tiled(TiledLayout& T) {
    dfree(T.tidx);
    dfree(T.tval);
    dfree(T.flags);
    dfree(T.row0);
    dfree(T.segoff);
    dfree(T.segrows);
    dfree(T.open);
    dfree(T.gmax);
    dfree(T.head);
    dfree(T.tail);
    dfree(T.spans);
    dfree(T.vpos);
    dfree(T.cold_row);
    dfree(T.srows);
    dfree(T.part);
    T = TiledLayout{};
}

void free_layout(DevLayout& L) {
    free_tiled(L.tl);
    dfree(L.ptr);
    dfree(L.idx);
    dfree(L.val);
    dfree(L.val0);
    dfree(L.units);
    dfree(L.split_rows);
    dfree(L.split_beg);
    dfree(L.partials);
    dfree(L.pair);
    L = DevLayout{};
}

// ----------------------------------------------------------- planning ---
struct HostPlan {
    std::vector<uint4> units;
    std::vector<uint32_t> split_rows, split_beg;
};

inline int group_log2(uint32_t d, uint32_t epl) {
    uint32_t want = (d + epl - 1) / (epl ? epl : 1);
    int lg = 0;
    while ((1u << lg) < want && lg < 6) ++lg;
    return lg;
}

HostPlan make_plan(const uint32_t* ptr, int64_t rows, const ccd_plan_params& pp) {
    // rows arrive sorted by falling degree: split rows first, then groups of
    // equal width as contiguous ranges
    HostPlan P;
    auto deg = [&](int64_t r) { return ptr[r + 1] - ptr[r]; };
    uint32_t slot = 0;
    int64_t r = 0;
    for (; r < rows && deg(r) > pp.split_threshold; ++r) {
        P.split_rows.push_back((uint32_t)r);
        P.split_beg.push_back(slot);
        for (uint32_t s = ptr[r]; s < ptr[r + 1]; s += pp.piece) {
            const uint32_t e = std::min<uint64_t>((uint64_t)s + pp.piece, ptr[r + 1]);
            P.units.push_back(make_uint4(ccd::UNIT_SPLIT | (slot << 8), (uint32_t)r, s, e));
            ++slot;
        }
    }
    P.split_beg.push_back(slot);
    const uint32_t max_rows = 4096;
    int64_t i = r;
    while (i < rows) {
        const int lg = group_log2(deg(i), pp.elems_per_lane);
        const uint32_t rpp = 64u >> lg;
        int64_t j = i;
        uint64_t elems = 0;
        while (j < rows && group_log2(deg(j), pp.elems_per_lane) == lg) {
            const uint64_t cnt = (uint64_t)(j - i);
            if (cnt % rpp == 0 && (elems >= pp.unit_elems || cnt >= max_rows)) break;
            elems += deg(j);
            ++j;
        }
        P.units.push_back(make_uint4(ccd::UNIT_GROUP | ((uint32_t)lg << 8), (uint32_t)i, (uint32_t)j, 0u));
        i = j;
    }
    return P;
}

void drop_graph(ccd_ctx* c) {
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    c->graph_exec = nullptr;
    c->graph_key.clear();
}

// ----------------------------------------------------------- profiling --
hipEvent_t take_event(ccd_ctx* c) {
    if (c->pool_used == c->pool.size()) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        c->pool.push_back(e);
    }
    return c->pool[c->pool_used++];
}

// Times the launches made in its scope as one record of `kind` when that
// kind is profiled.  Default: event records before and after.  bound: the
// events are handed to a single launch (ev()) that binds them to its own
// dispatch, so nothing is added between kernels; a bound pair still costs
// about 5 us of device time, so when the caller selects kinds
// (profile_kinds) one bound launch in PROF_SAMPLE of each is timed.
constexpr uint64_t PROF_SAMPLE = 5;
struct Timed {
    ccd_ctx* c;
    int kind;
    bool bound;
    hipEvent_t a = nullptr, b = nullptr;
    Timed(ccd_ctx* ctx, int k, bool bind = false) : c(ctx), kind(k), bound(bind) {
        if (!c->prof || (k >= 0 && c->prof_kinds && !((c->prof_kinds >> k) & 1u))) return;
        if (bound && k >= 0 && c->prof_kinds && c->kind_seen[k]++ % PROF_SAMPLE != 0) return;
        a = take_event(c);
        b = take_event(c);
        if (!a || !b) {
            a = b = nullptr;
            return;
        }
        if (!bound) (void)hipEventRecord(a, c->stream);
    }
    ccd::LaunchEvents ev_ = {};
    const ccd::LaunchEvents* ev() {
        if (!a) return nullptr;
        ev_ = {a, b};
        return &ev_;
    }
    ~Timed() {
        if (!a) return;
        if (!bound) (void)hipEventRecord(b, c->stream);
        c->recs.push_back({kind, a, b});
    }
};

// Algorithmic HBM bytes of one pass over layout L: the index and value
// streams (values read, and written back when the residual is updated), the
// row pointer and the row list, one read of every distinct gathered vector
// and per-row vector, and the output vector.  In a split tiled layout with
// 3-byte indices the first parts' elements (indices below the split, under
// 2^16) sit in tiles that read only the 16-bit plane: 2 bytes each.
int tiled_index_bytes(const DevLayout& L) { return L.gdim <= 65536 ? 2 : L.gdim <= (1 << 24) ? 3 : 4; }

double layout_bytes(const DevLayout& L, bool tiled, bool split, bool residual_write, int gathered, int per_row,
                    bool writes_y) {
    const double ib = tiled ? (double)tiled_index_bytes(L) : L.idx16 ? 2.0 : 4.0;
    double b = (double)L.nnz * (ib + 4.0 + (residual_write ? 4.0 : 0.0));
    if (tiled && split && L.split_at && ib == 3.0) {
        uint64_t nb = 0;
        for (uint32_t v : L.below) nb += v;
        b -= (double)nb;
    }
    b += 4.0 * (double)(L.rows + 1) + 4.0 * (double)L.rows;
    b += 4.0 * (double)L.gdim * gathered + 4.0 * (double)L.rows * per_row;
    if (writes_y) b += 4.0 * (double)L.rows;
    return b;
}

struct RunState {
    ccd_ctx* c;
    const ccd_params* p;
};

size_t factor_bytes(const ccd_ctx* c) { return c->prec == CCD_PREC_BF16 ? 2 : 4; }

// Launch shape of a persistent sweep kernel: as many 1024-thread blocks per
// CU as its register use allows (at most 2, or the plan's blocks_per_cu),
// and the LDS those blocks share split evenly between them, 1 KiB spare.
struct Launch {
    uint32_t blocks;
    size_t lds_budget;
};

// ib: bytes per index (2 or 4 for the binned sweeps, 2, 3 or 4 for the tiled ones)
Launch launch_shape(ccd_ctx* c, bool tiled, int sigma, int ib, bool sub, bool add, bool sweep, bool xcur) {
    const uint32_t key = (uint32_t)ib << 20 | (xcur ? 1u << 17 : 0u) | (tiled ? 1u << 16 : 0u) |
                         (uint32_t)sigma << 8 | (uint32_t)c->prec << 4 | (sub ? 4u : 0u) | (add ? 2u : 0u) |
                         (sweep ? 1u : 0u);
    auto it = c->occupancy.find(key);
    int bpc;
    if (it != c->occupancy.end()) {
        bpc = it->second;
    } else {
        bpc = tiled ? ccd::tiled_blocks_per_cu(sigma, c->prec, ib, sub, add, sweep, xcur)
                    : ccd::sweep_blocks_per_cu(c->prec, ib == 2, sub, add, sweep);
        c->occupancy[key] = bpc;
    }
    bpc = std::max(1, std::min(bpc, 2));
    if (c->plan.blocks_per_cu) bpc = std::min<int>(bpc, (int)c->plan.blocks_per_cu);
    const int lds = std::min(c->lds_block, c->lds_cu / bpc) - 1024;
    return {(uint32_t)(bpc * c->ncu), (size_t)std::max(0, lds) & ~(size_t)255};
}


// Split of cap factor entries of LDS between x (one entry each) and the pair
// array (two) that covers the most gathers: both prefixes cost one gather
// per element of L, so the split maximises gcum(hx) + gcum(hp) subject to
// hx + 2 hp <= cap.
void choose_split(DevLayout& L, uint64_t cap) {
    if (L.split_cap == cap && cap) return;
    const uint64_t g = (uint64_t)L.gdim;
    uint64_t best = 0, bx = std::min(cap, g), bp = 0;
    for (uint64_t hp = 0; 2 * hp <= cap && hp <= g; ++hp) {
        const uint64_t hx = std::min(cap - 2 * hp, g);
        const uint64_t cov = L.gcum[hx] + L.gcum[hp];
        if (cov > best) {
            best = cov;
            bx = hx;
            bp = hp;
        }
    }
    L.split_cap = cap;
    L.split_x = (uint32_t)bx;
    L.split_p = (uint32_t)bp;
}

// Leading (highest-degree) entries of the gathered vectors staged in LDS.
// A plain sweep stages x only; a fused pass whose x is the pair's current
// value stages pairs only; a fused pass with a separate x splits the budget
// by choose_split.
struct Hot {
    uint32_t x, p;
};

Hot hot_prefix(const ccd_ctx* c, DevLayout& L, size_t budget, bool fused, bool sweep, bool x_is_cur) {
    const size_t cap = budget / factor_bytes(c);
    if (!sweep) return {0, 0};
    if (!fused) return {(uint32_t)std::min<size_t>(cap, (size_t)L.gdim), 0};
    if (x_is_cur) return {0, (uint32_t)std::min<size_t>(cap / 2, (size_t)L.gdim)};
    // the 16-byte alignment of the pair prefix may take up to 15 bytes
    choose_split(L, cap > 8 ? cap - 8 : 0);
    return {L.split_x, L.split_p};
}

// Split bound of the tiled sweeps with the plan's split_rows: the LDS prefix
// of a plain FP32 sweep with one block per CU (FP32 entries of the LDS a
// block may use, 1 KiB spare, as launch_shape sizes it); rows of a layout
// whose gathered vector exceeds it are split there.  0: no split.
uint32_t split_bound(const ccd_ctx* c) {
    if (c->plan.split_rows == 0) return 0;
    return (uint32_t)(((size_t)std::max(0, std::min(c->lds_block, c->lds_cu) - 1024) & ~(size_t)255) / 4);
}

int do_pass(ccd_ctx* c, DevLayout& L, const void* x, void* y, const void* xs, const void* ys, const void* xa,
            const void* ya, bool x_is_cur, bool sub, bool add, bool sweep, float lambda, int kind) {
    const bool fused = sweep && (sub || add);
    const Launch sh = launch_shape(c, false, 0, L.idx16 ? 2 : 4, sub, add, sweep, false);
    const Hot hot = hot_prefix(c, L, sh.lds_budget, fused, sweep, x_is_cur);
    ccd::SweepArgs a{L.ptr, L.idx, L.val, sweep ? x : nullptr, y, xs, ys, xa, ya, fused ? L.pair : nullptr,
                     L.units, L.partials, L.nunits, hot.x, hot.p, (uint32_t)(fused && x_is_cur), lambda};
    {
        Timed tm(c, kind, true);
        HIP_TRY(c, ccd::launch_sweep(a, c->prec, L.idx16, sub, add, sweep, sh.blocks, c->stream, tm.ev()));
    }
    if (sweep && L.nsplit) {
        Timed tm(c, CCD_K_FINALIZE);
        ccd::FinalizeArgs f{L.split_rows, L.split_beg, L.partials, L.ptr, y, lambda, L.nsplit};
        HIP_TRY(c, ccd::launch_finalize(f, c->prec, c->stream));
    }
    return CCD_OK;
}

using PassFn = int (*)(ccd_ctx*, DevLayout&, const void*, void*, const void*, const void*, const void*,
                      const void*, bool, bool, bool, bool, float, int);

int upload_layout(ccd_ctx* c, DevLayout& L, int64_t rows, int64_t gdim, uint64_t nnz, const uint32_t* ptr,
                  const uint32_t* idx, const float* val) {
    L.rows = rows;
    L.gdim = gdim;
    L.nnz = nnz;
    L.idx16 = gdim <= 65536;
    L.hptr.assign(ptr, ptr + rows + 1);
    L.split_cap = 0;
    HIP_TRY(c, dalloc(&L.ptr, rows + 1));
    HIP_TRY(c, ccd::copy_sync(c->stream, L.ptr, ptr, sizeof(uint32_t) * (rows + 1), hipMemcpyHostToDevice));
    if (L.idx16) {
        std::vector<uint16_t> tmp(nnz);
        for (uint64_t e = 0; e < nnz; ++e) tmp[e] = (uint16_t)idx[e];
        uint16_t* d = nullptr;
        HIP_TRY(c, dalloc(&d, nnz));
        L.idx = d;
        HIP_TRY(c, ccd::copy_sync(c->stream, d, tmp.data(), sizeof(uint16_t) * nnz, hipMemcpyHostToDevice));
    } else {
        uint32_t* d = nullptr;
        HIP_TRY(c, dalloc(&d, nnz));
        L.idx = d;
        HIP_TRY(c, ccd::copy_sync(c->stream, d, idx, sizeof(uint32_t) * nnz, hipMemcpyHostToDevice));
    }
    HIP_TRY(c, dalloc(&L.val, nnz));
    HIP_TRY(c, dalloc(&L.val0, nnz));
    HIP_TRY(c, ccd::copy_sync(c->stream, L.val0, val, sizeof(float) * nnz, hipMemcpyHostToDevice));
    HIP_TRY(c, ccd::copy_sync(c->stream, L.val, L.val0, sizeof(float) * nnz, hipMemcpyDeviceToDevice));

    HostPlan P = make_plan(ptr, rows, c->plan);
    if (P.split_beg.back() >= (1u << 24) || P.units.size() >= 0xffffffffull)
        return set_err(c, CCD_ERR_INVALID, "work plan too large: raise ccd_plan_params.piece");
    L.nunits = (uint32_t)P.units.size();
    L.nsplit = (uint32_t)P.split_rows.size();
    L.npieces = P.split_beg.back();
    HIP_TRY(c, dalloc(&L.units, P.units.size()));
    if (!P.units.empty())
        HIP_TRY(c, ccd::copy_sync(c->stream, L.units, P.units.data(), sizeof(uint4) * P.units.size(), hipMemcpyHostToDevice));
    HIP_TRY(c, dalloc(&L.split_rows, P.split_rows.size()));
    if (!P.split_rows.empty())
        HIP_TRY(c, ccd::copy_sync(c->stream, L.split_rows, P.split_rows.data(), sizeof(uint32_t) * P.split_rows.size(),
                                  hipMemcpyHostToDevice));
    HIP_TRY(c, dalloc(&L.split_beg, P.split_beg.size()));
    HIP_TRY(c, ccd::copy_sync(c->stream, L.split_beg, P.split_beg.data(), sizeof(uint32_t) * P.split_beg.size(), hipMemcpyHostToDevice));
    HIP_TRY(c, dalloc(&L.partials, L.npieces));
    return CCD_OK;
}

template <class T>
int upload_vec(ccd_ctx* c, T** dst, const std::vector<T>& v) {
    HIP_TRY(c, dalloc(dst, v.size()));
    if (!v.empty()) HIP_TRY(c, ccd::copy_sync(c->stream, *dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    return CCD_OK;
}

// Tile descriptors from the row pointer (host, tile_layout.cpp), index array
// tiled on the device.
// split_at > 0: split each row at internal index split_at (its elements
// ascend by index) into two virtual rows, all first parts (rows in order)
// before all second parts, so that a tile gathers either from below the
// bound only or from above it only; 0, or a bound past the gathered
// vector: one tile range (the batched kernel asks for it).
int build_tiled(ccd_ctx* c, DevLayout& L, int sigma, uint32_t split_at) {
    if ((int64_t)split_at >= L.gdim) split_at = 0;
    const bool split = split_at != 0;
    if (L.tl.sigma == sigma && L.split_at == split_at) return CCD_OK;
    // a captured outer iteration (or batch) holds the old layout's pointers
    drop_graph(c);
    ++c->gen;
    free_tiled(L.tl);
    L.split_at = 0;
    L.below.clear();
    TiledLayout& T = L.tl;
    if (split) {
        uint32_t* d = nullptr;
        HIP_TRY(c, dalloc(&d, (size_t)std::max<int64_t>(L.rows, 1)));
        L.below.resize((size_t)L.rows);
        hipError_t e = ccd::launch_count_below(L.ptr, L.idx, L.idx16, (uint64_t)L.rows, split_at, d, c->stream);
        if (e == hipSuccess && L.rows)
            e = ccd::copy_sync(c->stream, L.below.data(), d, sizeof(uint32_t) * (size_t)L.rows, hipMemcpyDeviceToHost);
        dfree(d);
        HIP_TRY(c, e);
    }
    ccd::TileTables H;
    const std::string why =
        ccd::build_tile_tables(L.hptr.data(), L.rows, L.nnz, split ? L.below.data() : nullptr, sigma, split, H);
    if (!why.empty()) return set_err(c, CCD_ERR_INVALID, "tile layout: " + why);
    const uint64_t TS = 64ull * (uint64_t)sigma, nnz = L.nnz, ntiles = H.ntiles;
    T.ntiles = (uint32_t)ntiles;
    T.nspan = (uint32_t)H.spans.size();
    T.nshort = H.nshort;
    int rc;
    if (sigma <= 16) {  // the row-start bits of a lane fit 16 bits: half the stream
        std::vector<uint16_t> f16(H.flags.begin(), H.flags.end());
        uint16_t* d = nullptr;
        rc = upload_vec(c, &d, f16);
        T.flags = d;
    } else {
        uint32_t* d = nullptr;
        rc = upload_vec(c, &d, H.flags);
        T.flags = d;
    }
    if (rc) return rc;
    if ((rc = upload_vec(c, &T.row0, H.row0)) ||
        (rc = upload_vec(c, &T.segoff, H.segoff)) || (rc = upload_vec(c, &T.segrows, H.segrows)) ||
        (rc = upload_vec(c, &T.open, H.open)) || (rc = upload_vec(c, &T.spans, H.spans)))
        return rc;
    T.cold_tile0 = H.cold_tile0;
    if (split) {
        T.nreal = H.nreal;
        if ((rc = upload_vec(c, &T.vpos, H.vpos)) || (rc = upload_vec(c, &T.cold_row, H.cold_row)) ||
            (rc = upload_vec(c, &T.srows, H.srows)))
            return rc;
        HIP_TRY(c, dalloc(&T.part, H.vptr.size() - 1));
    }
    const uint64_t padded = ntiles * TS;
    HIP_TRY(c, dalloc(&T.tval, padded));
    HIP_TRY(c, dalloc(&T.gmax, ntiles));
    HIP_TRY(c, dalloc(&T.head, ntiles));
    HIP_TRY(c, dalloc(&T.tail, ntiles));
    T.ib = tiled_index_bytes(L);
    {
        uint8_t* d = nullptr;  // the 3-byte format: u16 plane, then u8 plane
        HIP_TRY(c, dalloc(&d, padded * (size_t)T.ib));
        T.tidx = d;
    }
    if (ntiles) {
        HIP_TRY(c, ccd::launch_retile_idx(L.idx, L.idx16, T.tidx, T.ib, nnz, T.ntiles, sigma, T.vpos, c->stream));
        HIP_TRY(c, ccd::launch_tile_gmax(L.idx, L.idx16, T.segoff, T.gmax, nnz, T.ntiles, sigma, T.vpos, c->stream));
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    T.sigma = sigma;  // only a complete layout is reused
    T.split = split;
    L.split_at = split_at;
    return CCD_OK;
}

int do_tiled_pass(ccd_ctx* c, DevLayout& L, const void* x, void* y, const void* xs, const void* ys, const void* xa,
                  const void* ya, bool x_is_cur, bool sub, bool add, bool sweep, float lambda, int kind) {
    TiledLayout& T = L.tl;
    const bool fused = sweep && (sub || add);
    const Launch sh = launch_shape(c, true, T.sigma, T.ib, sub, add, sweep, fused && x_is_cur);
    const Hot hot = hot_prefix(c, L, sh.lds_budget, fused, sweep, x_is_cur);
    // tiles of second parts gather from beyond the LDS prefix only when the
    // split is at or past it (a prefix staged with two blocks per CU, or in
    // bfloat16, differs from the one the layout was split at)
    const uint32_t cold0 = L.split_at >= hot.x ? T.cold_tile0 : 0xffffffffu;
    ccd::TiledArgs a{L.ptr, T.tidx, T.tval, sweep ? x : nullptr, y, xs, ys, xa, ya, fused ? L.pair : nullptr,
                     T.flags, T.row0, T.segoff, T.segrows, T.open, T.gmax, T.head, T.tail, L.nnz, T.ntiles,
                     hot.x, hot.p, (uint32_t)(fused && x_is_cur), lambda, T.cold_row, T.nreal,
                     sweep ? T.part : nullptr, cold0, T.cold_tile0};
    {
        Timed tm(c, kind, true);
        HIP_TRY(c, ccd::launch_tiled(a, T.sigma, c->prec, T.ib, sub, add, sweep, sh.blocks, c->stream, tm.ev()));
    }
    if (sweep && T.split) {
        Timed tm(c, CCD_K_FINALIZE);
        const ccd::SplitArgs fa{T.srows, T.part, T.head, T.tail, y, lambda, T.nreal};
        HIP_TRY(c, ccd::launch_split_finish(fa, c->prec, c->stream));
    } else if (sweep && T.nspan) {
        Timed tm(c, CCD_K_FINALIZE);
        ccd::SpanArgs sa{T.spans, T.head, T.tail, y, T.nspan, T.nshort, lambda, nullptr};
        HIP_TRY(c, ccd::launch_span(sa, c->prec, c->stream));
    }
    return CCD_OK;
}

// Positions sorted by falling degree (stable).
std::vector<uint32_t> degree_order(const uint32_t* ptr, int64_t rows) {
    std::vector<uint32_t> perm((size_t)rows);
    for (int64_t r = 0; r < rows; ++r) perm[r] = (uint32_t)r;
    std::stable_sort(perm.begin(), perm.end(),
                     [&](uint32_t a, uint32_t b) { return ptr[a + 1] - ptr[a] > ptr[b + 1] - ptr[b]; });
    return perm;
}

// Element order inside each row of a renumbered layout: the caller's, or
// ascending internal index (so that a row splits at any bound into the
// elements gathered from below it and the rest, and a lane's consecutive
// elements beyond the LDS prefix share cache lines).
enum class RowOrder { keep, sort };

// One compressed layout in internal numbering: rows in `perm` order, indices
// mapped through `ipos` (caller's index -> internal), elements inside each
// row in `order` (equal sort keys keep their order).  src[o]: the caller's
// position of internal element o.  Rows are processed in parallel, in
// chunks taken from a shared counter.
void renumber(int64_t rows, const uint32_t* ptr, const uint32_t* idx, const float* val,
              const std::vector<uint32_t>& perm, const std::vector<uint32_t>& ipos, std::vector<uint32_t>& optr,
              std::vector<uint32_t>& oidx, std::vector<float>& oval, std::vector<uint32_t>& src, RowOrder order) {
    optr.assign((size_t)rows + 1, 0u);
    for (int64_t i = 0; i < rows; ++i) optr[i + 1] = optr[i] + (ptr[perm[i] + 1] - ptr[perm[i]]);
    oidx.resize(optr[rows]);
    oval.resize(optr[rows]);
    src.resize(optr[rows]);
    constexpr int64_t CHUNK = 512;
    std::atomic<int64_t> next{0};
    auto work = [&]() {
        std::vector<uint64_t> key;
        for (int64_t b; (b = next.fetch_add(CHUNK)) < rows;) {
            for (int64_t i = b; i < std::min(rows, b + CHUNK); ++i) {
                const uint32_t r = perm[i], e0 = ptr[r], len = ptr[r + 1] - ptr[r];
                key.resize(len);
                for (uint32_t q = 0; q < len; ++q)
                    key[q] = (uint64_t)(order == RowOrder::sort ? ipos[idx[e0 + q]] : 0u) << 32 | q;
                if (order != RowOrder::keep) std::sort(key.begin(), key.end());
                const uint32_t o = optr[i];
                for (uint32_t q = 0; q < len; ++q) {
                    const uint32_t e = e0 + (uint32_t)key[q];
                    oidx[o + q] = ipos[idx[e]];
                    oval[o + q] = val[e];
                    src[o + q] = e;
                }
            }
        }
    };
    const uint64_t nnz = optr[rows];
    const unsigned nt = nnz < (1u << 20) ? 1u : std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < nt; ++t) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
}

const char* check_compressed(int64_t rows, int64_t dim, uint64_t nnz, const uint32_t* ptr, const uint32_t* idx) {
    if (ptr[0] != 0) return "pointer array must start at 0";
    for (int64_t r = 0; r < rows; ++r)
        if (ptr[r + 1] < ptr[r]) return "pointer array is not monotone";
    if ((uint64_t)ptr[rows] != nnz) return "last pointer differs from nnz";
    for (uint64_t e = 0; e < nnz; ++e)
        if ((int64_t)idx[e] >= dim) return "index out of range";
    return nullptr;
}

}  // namespace ccdi

using namespace ccdi;

extern "C" {

int32_t ccd_abi_version(void) { return CCD_ABI_VERSION; }

int32_t ccd_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

void ccd_params_default(ccd_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->k = 10;
    p->lambda = 0.1f;
    p->maxiter = 5;
    p->maxinneriter = 5;
    p->variant = CCD_VARIANT_NATIVE;
    p->precision = CCD_PREC_FP32;
    p->order = CCD_ORDER_H_FIRST;
    p->zero_first = 1;
    p->rmse_every = 1;
    p->profile = 0;
    p->use_graph = 0;
}

int32_t ccd_create(int32_t device, ccd_ctx** out) {
    if (!out) return CCD_ERR_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return CCD_ERR_HIP;
    if (device < 0 || device >= count) return CCD_ERR_INVALID;
    ccd_ctx* c = new (std::nothrow) ccd_ctx();
    if (!c) return CCD_ERR_ALLOC;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return CCD_ERR_HIP;
    }
    int ncu = 0, lds_block = 0, lds_cu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || ncu <= 0) ncu = 256;
    if (hipDeviceGetAttribute(&lds_block, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess)
        lds_block = 65536;
    if (hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device) != hipSuccess ||
        lds_cu <= 0)
        lds_cu = lds_block;
    c->ncu = ncu;
    c->lds_block = lds_block;
    c->lds_cu = lds_cu;
    if (dalloc(&c->rmse_bsum, ccd::RMSE_BLOCKS) != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return CCD_ERR_ALLOC;
    }
    *out = c;
    return CCD_OK;
}

int32_t ccd_destroy(ccd_ctx* c) {
    if (!c) return CCD_ERR_INVALID;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_layout(c->csr);
    free_layout(c->csc);
    dfree(c->trow);
    dfree(c->tcol);
    dfree(c->tval);
    dfree(c->W);
    dfree(c->H);
    dfree(c->u);
    dfree(c->v);
    dfree(c->W16);
    dfree(c->H16);
    dfree(c->u16);
    dfree(c->v16);
    dfree(c->rmse_bsum);
    dfree(c->rmse_dev);
    dfree(c->obj_dev);
    drop_graph(c);
    for (hipEvent_t e : c->pool) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return CCD_OK;
}

const char* ccd_last_error(const ccd_ctx* c) { return c ? c->err.c_str() : "null context"; }

int32_t ccd_set_plan(ccd_ctx* c, const ccd_plan_params* pp) {
    if (!c || !pp) return CCD_ERR_INVALID;
    if (pp->elems_per_lane == 0 || pp->piece == 0 || pp->unit_elems == 0 || pp->split_threshold < pp->piece)
        return set_err(c, CCD_ERR_INVALID, "plan: elems_per_lane, piece, unit_elems > 0 and split_threshold >= piece");
    if (pp->sigma != 8 && pp->sigma != 16 && pp->sigma != 32)
        return set_err(c, CCD_ERR_INVALID, "plan: sigma must be 8, 16 or 32");
    if (pp->blocks_per_cu > 2) return set_err(c, CCD_ERR_INVALID, "plan: blocks_per_cu must be 0, 1 or 2");
    if (pp->split_rows > 1) return set_err(c, CCD_ERR_INVALID, "plan: split_rows must be 0 or 1");
    c->plan = *pp;
    ++c->gen;
    return CCD_OK;
}

int32_t ccd_upload(ccd_ctx* c, int64_t m, int64_t n, uint64_t nnz, const uint32_t* row_ptr, const uint32_t* col_idx,
                   const float* csr_val, const uint32_t* col_ptr, const uint32_t* row_idx, const float* csc_val) {
    if (!c) return CCD_ERR_INVALID;
    if (m <= 0 || n <= 0 || m > 0x7fffffff || n > 0x7fffffff || nnz > 0xffffffffull)
        return set_err(c, CCD_ERR_INVALID, "shape out of range (m, n < 2^31, nnz < 2^32)");
    if (!row_ptr || !col_ptr || (nnz && (!col_idx || !csr_val || !row_idx || !csc_val)))
        return set_err(c, CCD_ERR_INVALID, "null array");
    if (const char* why = check_compressed(m, n, nnz, row_ptr, col_idx)) return set_err(c, CCD_ERR_INVALID, std::string("CSR: ") + why);
    if (const char* why = check_compressed(n, m, nnz, col_ptr, row_idx)) return set_err(c, CCD_ERR_INVALID, std::string("CSC: ") + why);
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    drop_graph(c);
    ++c->gen;
    free_layout(c->csr);
    free_layout(c->csc);
    c->uploaded = false;
    c->factors = false;
    c->m = m;
    c->n = n;
    c->nnz = nnz;
    {
        std::vector<uint32_t> rperm = degree_order(row_ptr, m), cperm = degree_order(col_ptr, n);
        std::vector<uint32_t> rpos((size_t)m), cpos((size_t)n);
        for (int64_t i = 0; i < m; ++i) rpos[rperm[i]] = (uint32_t)i;
        for (int64_t j = 0; j < n; ++j) cpos[cperm[j]] = (uint32_t)j;
        std::vector<uint32_t> p2, i2, src;
        std::vector<float> v2;
        // Rows sorted by gathered index: the tiled sweeps can then split
        // each row at any bound (build_tiled) and keep the elements gathered
        // from beyond an LDS prefix in tiles of their own.
        renumber(m, row_ptr, col_idx, csr_val, rperm, cpos, p2, i2, v2, src, RowOrder::sort);
        int rc = upload_layout(c, c->csr, m, n, nnz, p2.data(), i2.data(), v2.data());
        if (rc) return rc;
        c->csr.src = std::move(src);
        renumber(n, col_ptr, row_idx, csc_val, cperm, rpos, p2, i2, v2, src, RowOrder::sort);
        rc = upload_layout(c, c->csc, n, m, nnz, p2.data(), i2.data(), v2.data());
        if (rc) return rc;
        c->csc.src = std::move(src);
        c->csr.perm = std::move(rperm);
        c->csc.perm = std::move(cperm);
    }
    for (DevLayout* L : {&c->csr, &c->csc}) {  // gather counts of each layout's gathered vector
        const DevLayout& O = L == &c->csr ? c->csc : c->csr;
        L->gcum.assign((size_t)L->gdim + 1, 0);
        for (int64_t i = 0; i < L->gdim; ++i) L->gcum[i + 1] = L->gcum[i] + (O.hptr[i + 1] - O.hptr[i]);
    }
    dfree(c->u);
    dfree(c->v);
    HIP_TRY(c, dalloc(&c->u, m));
    HIP_TRY(c, dalloc(&c->v, n));
    c->uploaded = true;
    return CCD_OK;
}

int32_t ccd_upload_test(ccd_ctx* c, uint64_t nnz, const uint32_t* row, const uint32_t* col, const float* val) {
    if (!c) return CCD_ERR_INVALID;
    if (!c->uploaded) return set_err(c, CCD_ERR_STATE, "ccd_upload_test before ccd_upload");
    if (nnz && (!row || !col || !val)) return set_err(c, CCD_ERR_INVALID, "null array");
    for (uint64_t e = 0; e < nnz; ++e)
        if ((int64_t)row[e] >= c->m || (int64_t)col[e] >= c->n) return set_err(c, CCD_ERR_INVALID, "test index out of range");
    HIP_TRY(c, hipSetDevice(c->device));
    dfree(c->trow);
    dfree(c->tcol);
    dfree(c->tval);
    c->trow = c->tcol = nullptr;
    c->tval = nullptr;
    c->nnz_test = nnz;
    HIP_TRY(c, dalloc(&c->trow, nnz));
    HIP_TRY(c, dalloc(&c->tcol, nnz));
    HIP_TRY(c, dalloc(&c->tval, nnz));
    if (nnz) {
        // Internal numbering of the test entries, sorted by (row, column):
        // neighbouring threads of the RMSE kernel then read neighbouring
        // entries of each rank-major factor row, one or two cache lines a
        // load instead of one per lane.  Only the order of the FP64 sum of
        // squared errors changes.
        std::vector<uint32_t> rpos((size_t)c->m), cpos((size_t)c->n), r2(nnz), c2(nnz);
        std::vector<float> v2(nnz);
        std::vector<uint64_t> key(nnz);
        for (int64_t i = 0; i < c->m; ++i) rpos[c->csr.perm[i]] = (uint32_t)i;
        for (int64_t j = 0; j < c->n; ++j) cpos[c->csc.perm[j]] = (uint32_t)j;
        for (uint64_t e = 0; e < nnz; ++e) key[e] = (uint64_t)rpos[row[e]] << 32 | cpos[col[e]];
        std::vector<uint64_t> order(nnz);
        for (uint64_t e = 0; e < nnz; ++e) order[e] = e;
        std::stable_sort(order.begin(), order.end(), [&](uint64_t x, uint64_t y) { return key[x] < key[y]; });
        for (uint64_t q = 0; q < nnz; ++q) {
            const uint64_t e = order[q];
            r2[q] = (uint32_t)(key[e] >> 32);
            c2[q] = (uint32_t)key[e];
            v2[q] = val[e];
        }
        HIP_TRY(c, ccd::copy_sync(c->stream, c->trow, r2.data(), 4 * nnz, hipMemcpyHostToDevice));
        HIP_TRY(c, ccd::copy_sync(c->stream, c->tcol, c2.data(), 4 * nnz, hipMemcpyHostToDevice));
        HIP_TRY(c, ccd::copy_sync(c->stream, c->tval, v2.data(), 4 * nnz, hipMemcpyHostToDevice));
    }
    return CCD_OK;
}

int32_t ccd_set_factors(ccd_ctx* c, uint32_t k, const float* W, const float* H) {
    if (!c) return CCD_ERR_INVALID;
    if (!c->uploaded) return set_err(c, CCD_ERR_STATE, "ccd_set_factors before ccd_upload");
    if (k == 0 || !W || !H) return set_err(c, CCD_ERR_INVALID, "k must be > 0 and W, H non-null");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (k != c->k) {
        drop_graph(c);
        ++c->gen;
        dfree(c->W);
        dfree(c->H);
        c->W = c->H = nullptr;
        HIP_TRY(c, dalloc(&c->W, (size_t)k * c->m));
        HIP_TRY(c, dalloc(&c->H, (size_t)k * c->n));
        c->k = k;
    }
    {
        std::vector<float> t2((size_t)k * std::max(c->m, c->n));
        for (uint32_t t = 0; t < k; ++t)
            for (int64_t i = 0; i < c->m; ++i) t2[(size_t)t * c->m + i] = W[(size_t)t * c->m + c->csr.perm[i]];
        HIP_TRY(c, ccd::copy_sync(c->stream, c->W, t2.data(), sizeof(float) * (size_t)k * c->m, hipMemcpyHostToDevice));
        for (uint32_t t = 0; t < k; ++t)
            for (int64_t j = 0; j < c->n; ++j) t2[(size_t)t * c->n + j] = H[(size_t)t * c->n + c->csc.perm[j]];
        HIP_TRY(c, ccd::copy_sync(c->stream, c->H, t2.data(), sizeof(float) * (size_t)k * c->n, hipMemcpyHostToDevice));
    }
    HIP_TRY(c, ccd::copy_sync(c->stream, c->csr.val, c->csr.val0, 4 * c->nnz, hipMemcpyDeviceToDevice));
    HIP_TRY(c, ccd::copy_sync(c->stream, c->csc.val, c->csc.val0, 4 * c->nnz, hipMemcpyDeviceToDevice));
    c->factors = true;
    c->continued = false;
    return CCD_OK;
}

int32_t ccd_run(ccd_ctx* c, const ccd_params* p, ccd_iter_cb cb, void* user) {
    if (!c || !p) return CCD_ERR_INVALID;
    if (!c->uploaded || !c->factors) return set_err(c, CCD_ERR_STATE, "ccd_run needs ccd_upload and ccd_set_factors");
    if (p->k != c->k) return set_err(c, CCD_ERR_INVALID, "params.k differs from the k of ccd_set_factors");
    if (p->maxiter < 0 || p->maxinneriter < 0 || !std::isfinite(p->lambda) || p->lambda < 0.0f)
        return set_err(c, CCD_ERR_INVALID, "maxiter, maxinneriter >= 0 and lambda >= 0 finite");
    if (p->variant != CCD_VARIANT_NATIVE && p->variant != CCD_VARIANT_BATCHING && p->variant != CCD_VARIANT_TILED)
        return set_err(c, CCD_ERR_INVALID, "variant must be 1, 2 or 3");
    if (p->precision != CCD_PREC_FP32 && p->precision != CCD_PREC_BF16 && p->precision != CCD_PREC_FP64_SUMS)
        return set_err(c, CCD_ERR_INVALID, "precision must be CCD_PREC_FP32, CCD_PREC_BF16 or CCD_PREC_FP64_SUMS");
    if (p->variant == CCD_VARIANT_TILED && p->precision != CCD_PREC_FP32 && c->plan.sigma != 16)
        return set_err(c, CCD_ERR_UNSUPPORTED, "tiled sweeps in BF16 or FP64-sums precision need sigma = 16");
    if (p->order != CCD_ORDER_H_FIRST && p->order != CCD_ORDER_W_FIRST) return set_err(c, CCD_ERR_INVALID, "bad order");
    if (p->zero_first && c->continued)
        return set_err(c, CCD_ERR_INVALID, "zero_first on a run that continues a previous one (call ccd_set_factors first)");
    HIP_TRY(c, hipSetDevice(c->device));

    const bool hfirst = p->order == CCD_ORDER_H_FIRST;
    const bool bf = p->precision == CCD_PREC_BF16;
    c->prec = p->precision;
    DevLayout& F = hfirst ? c->csc : c->csr;
    DevLayout& S = hfirst ? c->csr : c->csc;
    const int64_t dF = F.rows, dS = S.rows;
    const uint32_t k = c->k;
    // FP32 master factors; the zeroing of the first-swept one happens there
    if (p->zero_first)
        HIP_TRY(c, hipMemsetAsync(hfirst ? c->H : c->W, 0, sizeof(float) * (size_t)k * dF, c->stream));
    if (bf) {
        if (c->k16 != k) {
            dfree(c->W16);
            dfree(c->H16);
            c->W16 = c->H16 = nullptr;
            HIP_TRY(c, dalloc(&c->W16, (size_t)k * c->m));
            HIP_TRY(c, dalloc(&c->H16, (size_t)k * c->n));
            c->k16 = k;
        }
        if (!c->u16) HIP_TRY(c, dalloc(&c->u16, c->m));
        if (!c->v16) HIP_TRY(c, dalloc(&c->v16, c->n));
        HIP_TRY(c, ccd::launch_to_bf16(c->W, c->W16, (uint64_t)k * c->m, c->stream));
        HIP_TRY(c, ccd::launch_to_bf16(c->H, c->H16, (uint64_t)k * c->n, c->stream));
    }
    // factor storage of this run: byte pointers, element size es
    const size_t es = bf ? 2 : 4;
    char* const Wr = bf ? reinterpret_cast<char*>(c->W16) : reinterpret_cast<char*>(c->W);
    char* const Hr = bf ? reinterpret_cast<char*>(c->H16) : reinterpret_cast<char*>(c->H);
    char* const ur = bf ? reinterpret_cast<char*>(c->u16) : reinterpret_cast<char*>(c->u);
    char* const vr = bf ? reinterpret_cast<char*>(c->v16) : reinterpret_cast<char*>(c->v);
    char* const XF = hfirst ? Hr : Wr;
    char* const XS = hfirst ? Wr : Hr;
    char* const vF = hfirst ? vr : ur;  // output of the first sweep
    char* const vS = hfirst ? ur : vr;
    const float lambda = p->lambda;
    const int T = p->maxiter, Tin = p->maxinneriter;
    const bool do_rmse = p->rmse_every && c->nnz_test > 0;

    if (T > c->rmse_cap) {
        dfree(c->rmse_dev);
        c->rmse_dev = nullptr;
        HIP_TRY(c, dalloc(&c->rmse_dev, T));
        c->rmse_cap = T;
    }
    c->iter_rmse.assign(T, NAN);
    c->prof = p->profile != 0;
    c->prof_kinds = (uint32_t)p->profile_kinds;
    c->recs.clear();
    c->pool_used = 0;
    for (int q = 0; q < CCD_K_COUNT; ++q) {
        c->kind_ms[q] = 0.0;
        c->kind_n[q] = 0;
        c->kind_seen[q] = 0;
    }
    const bool tiled = p->variant == CCD_VARIANT_TILED;
    if (tiled) {
        int rc0 = build_tiled(c, F, (int)c->plan.sigma, split_bound(c));
        if (!rc0) rc0 = build_tiled(c, S, (int)c->plan.sigma, split_bound(c));
        if (rc0) return rc0;
    }
    const bool tl = tiled, sp = c->plan.split_rows != 0;
    c->kind_bytes[CCD_K_SWEEP_F] = layout_bytes(F, tl, sp, false, 1, 0, true);
    c->kind_bytes[CCD_K_SWEEP_S] = layout_bytes(S, tl, sp, false, 1, 0, true);
    c->kind_bytes[CCD_K_FUSED_F] = layout_bytes(F, tl, sp, true, 2, 2, true);
    c->kind_bytes[CCD_K_FUSED_S] = layout_bytes(S, tl, sp, true, 3, 2, true);
    c->kind_bytes[CCD_K_RESIDUAL] =
        0.5 * (layout_bytes(F, tl, sp, true, 1, 1, false) + layout_bytes(S, tl, sp, true, 1, 1, false));
    c->kind_bytes[CCD_K_FINALIZE] = 8.0 * (F.npieces + S.npieces) / 2.0;
    c->kind_bytes[CCD_K_RMSE] = 12.0 * (double)c->nnz_test + 4.0 * (double)k * (double)(c->m + c->n);
    c->kind_bytes[CCD_K_COPY] = 5.0 * (p->precision == CCD_PREC_BF16 ? 2.0 : 4.0) * (double)(c->m + c->n);

    const PassFn pass = tiled ? do_tiled_pass : do_pass;
    if (tiled) {
        // residual into tile order; rows without entries keep y = 0
        if (F.tl.ntiles)
            HIP_TRY(c, ccd::launch_retile_f32(F.val, F.tl.tval, F.nnz, F.tl.ntiles, F.tl.sigma, true, c->stream, F.tl.vpos));
        if (S.tl.ntiles)
            HIP_TRY(c, ccd::launch_retile_f32(S.val, S.tl.tval, S.nnz, S.tl.ntiles, S.tl.sigma, true, c->stream, S.tl.vpos));
        if (bf) {
            HIP_TRY(c, ccd::launch_fill_bytes16(reinterpret_cast<uint16_t*>(vF), 0, (uint64_t)dF, c->stream));
            HIP_TRY(c, ccd::launch_fill_bytes16(reinterpret_cast<uint16_t*>(vS), 0, (uint64_t)dS, c->stream));
        } else {
            HIP_TRY(c, ccd::launch_fill(reinterpret_cast<float*>(vF), 0.0f, (uint64_t)dF, c->stream));
            HIP_TRY(c, ccd::launch_fill(reinterpret_cast<float*>(vS), 0.0f, (uint64_t)dS, c->stream));
        }
    }
    hipEvent_t loop_a = nullptr, loop_b = nullptr;
    if (c->prof) {
        loop_a = take_event(c);
        loop_b = take_event(c);
    }
    if (loop_a) HIP_TRY(c, hipEventRecord(loop_a, c->stream));

    // Fused passes gather (previous, current) factor pairs: F.pair holds
    // (XS[prev], XS[t]) for the first-swept layout, S.pair (XF[prev], XF[t])
    // for the second.  The commit kernel at the end of each rank stores u, v
    // into W, H and builds the pairs of the next rank; a run that continues a
    // previous one starts with a fused rank 0 whose pairs are built here.
    const bool sweeps = Tin >= 1 && k > 0;
    if (sweeps) {
        for (DevLayout* L : {&F, &S}) {  // sized for FP32 pairs, so every precision fits
            if (L->pair) continue;
            double2* d = nullptr;
            HIP_TRY(c, dalloc(&d, ((size_t)L->gdim + 1) / 2));
            L->pair = d;
        }
        if (c->continued) {
            ccd::CommitArgs ca{nullptr, nullptr, XS, XS, F.pair, (uint64_t)dS, nullptr, nullptr, XF, XF, S.pair, (uint64_t)dF};
            HIP_TRY(c, ccd::launch_commit(ca, bf, c->stream));
        }
    }

    bool pending = false;  // rank `prev`'s final u v^T still to be subtracted
    uint32_t prev = 0;
    int rc;
    size_t rec_mark = 0;
    // One outer iteration's ranks.  build_last_pairs = false: the commit of
    // the run's last rank skips the pairs of a next rank that never comes.
    auto outer = [&](int oiter, bool build_last_pairs) -> int {
        for (uint32_t t = 0; t < k; ++t) {
            const char* St = XS + (size_t)t * dS * es;
            const char* Ft = XF + (size_t)t * dF * es;
            const char* Sp = XS + (size_t)prev * dS * es;
            const char* Fp = XF + (size_t)prev * dF * es;
            const bool add = oiter > 1 || c->continued;
            const bool fused = pending || add;
            if (sweeps) {
                int rc1 = pass(c, F, St, vF, Sp, Fp, St, Ft, true, pending, add, true, lambda, fused ? CCD_K_FUSED_F : CCD_K_SWEEP_F);
                if (rc1) return rc1;
                rc1 = pass(c, S, vF, vS, Fp, Sp, Ft, St, false, pending, add, true, lambda, fused ? CCD_K_FUSED_S : CCD_K_SWEEP_S);
                if (rc1) return rc1;
                for (int it = 2; it <= Tin; ++it) {
                    rc1 = pass(c, F, vS, vF, nullptr, nullptr, nullptr, nullptr, false, false, false, true, lambda, CCD_K_SWEEP_F);
                    if (rc1) return rc1;
                    rc1 = pass(c, S, vF, vS, nullptr, nullptr, nullptr, nullptr, false, false, false, true, lambda, CCD_K_SWEEP_S);
                    if (rc1) return rc1;
                }
                // W, H <- u, v; pairs of the next rank (none after the last)
                const bool last = !build_last_pairs && oiter == T && t + 1 == k;
                const uint32_t tn = t + 1 == k ? 0 : t + 1;
                char* XSt = XS + (size_t)t * dS * es;
                char* XFt = XF + (size_t)t * dF * es;
                ccd::CommitArgs ca{vS, XSt, nullptr, XS + (size_t)tn * dS * es, last ? nullptr : F.pair, (uint64_t)dS,
                                   vF, XFt, nullptr, XF + (size_t)tn * dF * es, last ? nullptr : S.pair, (uint64_t)dF};
                Timed tm(c, CCD_K_COPY);
                HIP_TRY(c, ccd::launch_commit(ca, bf, c->stream));
            } else if (fused) {
                int rc1 = pass(c, F, nullptr, nullptr, Sp, Fp, St, Ft, false, pending, add, false, lambda, CCD_K_RESIDUAL);
                if (rc1) return rc1;
                rc1 = pass(c, S, nullptr, nullptr, Fp, Sp, Ft, St, false, pending, add, false, lambda, CCD_K_RESIDUAL);
                if (rc1) return rc1;
            }
            pending = true;
            prev = t;
        }
        return CCD_OK;
    };
    // From outer iteration 2 on every iteration launches the same kernels
    // with the same arguments (rank 0 subtracts rank k-1 and adds itself
    // back), so with use_graph it is captured once and replayed.
    const bool graph = p->use_graph != 0 && k > 0 && (sweeps || Tin == 0);
    std::vector<uint64_t> gkey;
    if (graph) {
        uint32_t lam_bits;
        std::memcpy(&lam_bits, &lambda, 4);
        gkey = {c->gen, k, (uint64_t)Tin, lam_bits, (uint64_t)p->variant, (uint64_t)p->precision, (uint64_t)p->order,
                c->plan.sigma, c->plan.blocks_per_cu, (uint64_t)(uintptr_t)XF, (uint64_t)(uintptr_t)XS,
                (uint64_t)(uintptr_t)vF, (uint64_t)(uintptr_t)vS, (uint64_t)(uintptr_t)F.pair,
                (uint64_t)(uintptr_t)S.pair, (uint64_t)(uintptr_t)F.tl.tval, (uint64_t)(uintptr_t)S.tl.tval};
    }
    for (int oiter = 1; oiter <= T; ++oiter) {
        if (graph && oiter >= 2) {
            if (!c->graph_exec || c->graph_key != gkey) {
                HIP_TRY(c, hipStreamSynchronize(c->stream));
                drop_graph(c);
                const bool prof = c->prof;
                c->prof = false;  // no event records inside the capture
                HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
                const bool pending0 = pending;
                const uint32_t prev0 = prev;
                rc = outer(oiter, true);
                hipGraph_t g = nullptr;
                const hipError_t e = hipStreamEndCapture(c->stream, &g);
                c->prof = prof;
                pending = pending0;  // nothing ran yet
                prev = prev0;
                if (rc) {
                    if (g) (void)hipGraphDestroy(g);
                    return rc;
                }
                HIP_TRY(c, e);
                const hipError_t ei = hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (ei != hipSuccess) c->graph_exec = nullptr;
                HIP_TRY(c, ei);
                c->graph_key = gkey;
            }
            {
                Timed tm(c, -1);
                HIP_TRY(c, hipGraphLaunch(c->graph_exec, c->stream));
            }
            pending = true;
            prev = k - 1;
        } else {
            rc = outer(oiter, false);
            if (rc) return rc;
        }
        if (do_rmse) {
            Timed tm(c, CCD_K_RMSE);
            HIP_TRY(c, ccd::launch_rmse(c->trow, c->tcol, c->tval, c->nnz_test, Wr, Hr, bf, c->m, c->n, k,
                                        c->rmse_bsum, c->rmse_dev + (oiter - 1), c->stream));
        }
        if (cb) {
            HIP_TRY(c, hipStreamSynchronize(c->stream));
            double rmse = NAN;
            if (do_rmse) HIP_TRY(c, hipMemcpy(&rmse, c->rmse_dev + (oiter - 1), sizeof(double), hipMemcpyDeviceToHost));
            double rank_t = 0.0, upd_t = 0.0, rmse_t = 0.0;
            for (size_t q = rec_mark; q < c->recs.size(); ++q) {
                float ms = 0.0f;
                (void)hipEventElapsedTime(&ms, c->recs[q].a, c->recs[q].b);
                const int kd = c->recs[q].kind;
                if (kd == CCD_K_RMSE) rmse_t += ms * 1e-3;
                else if (kd == CCD_K_RESIDUAL || kd == CCD_K_COPY) upd_t += ms * 1e-3;
                else rank_t += ms * 1e-3;
            }
            rec_mark = c->recs.size();
            cb(user, oiter, rank_t, upd_t, rmse, rmse_t);
        }
    }
    if (pending) {
        const char* Sp = XS + (size_t)prev * dS * es;
        const char* Fp = XF + (size_t)prev * dF * es;
        rc = pass(c, F, nullptr, nullptr, Sp, Fp, nullptr, nullptr, false, true, false, false, lambda, CCD_K_RESIDUAL);
        if (rc) return rc;
        rc = pass(c, S, nullptr, nullptr, Fp, Sp, nullptr, nullptr, false, true, false, false, lambda, CCD_K_RESIDUAL);
        if (rc) return rc;
    }
    if (loop_b) HIP_TRY(c, hipEventRecord(loop_b, c->stream));
    if (bf) {  // back to the FP32 master copies (exact: every bf16 is a float)
        HIP_TRY(c, ccd::launch_from_bf16(c->W16, c->W, (uint64_t)k * c->m, c->stream));
        HIP_TRY(c, ccd::launch_from_bf16(c->H16, c->H, (uint64_t)k * c->n, c->stream));
    }
    if (tiled) {
        if (F.tl.ntiles)
            HIP_TRY(c, ccd::launch_retile_f32(F.tl.tval, F.val, F.nnz, F.tl.ntiles, F.tl.sigma, false, c->stream, F.tl.vpos));
        if (S.tl.ntiles)
            HIP_TRY(c, ccd::launch_retile_f32(S.tl.tval, S.val, S.nnz, S.tl.ntiles, S.tl.sigma, false, c->stream, S.tl.vpos));
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (T > 0) c->continued = true;
    if (do_rmse && T > 0) HIP_TRY(c, hipMemcpy(c->iter_rmse.data(), c->rmse_dev, sizeof(double) * T, hipMemcpyDeviceToHost));
    c->loop_s = 0.0;
    if (c->prof) {
        for (const ProfRec& r : c->recs) {
            if (r.kind < 0) continue;  // a replayed outer iteration
            float ms = 0.0f;
            HIP_TRY(c, hipEventElapsedTime(&ms, r.a, r.b));
            c->kind_ms[r.kind] += ms;
            c->kind_n[r.kind] += 1;
        }
        float ms = 0.0f;
        HIP_TRY(c, hipEventElapsedTime(&ms, loop_a, loop_b));
        c->loop_s = ms * 1e-3;
    }
    return CCD_OK;
}

int32_t ccd_get_factors(ccd_ctx* c, float* W, float* H) {
    if (!c) return CCD_ERR_INVALID;
    if (!c->factors) return set_err(c, CCD_ERR_STATE, "no factors");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const uint32_t k = c->k;
    std::vector<float> t2((size_t)k * std::max(c->m, c->n));
    if (W) {
        HIP_TRY(c, hipMemcpy(t2.data(), c->W, sizeof(float) * (size_t)k * c->m, hipMemcpyDeviceToHost));
        for (uint32_t t = 0; t < k; ++t)
            for (int64_t i = 0; i < c->m; ++i) W[(size_t)t * c->m + c->csr.perm[i]] = t2[(size_t)t * c->m + i];
    }
    if (H) {
        HIP_TRY(c, hipMemcpy(t2.data(), c->H, sizeof(float) * (size_t)k * c->n, hipMemcpyDeviceToHost));
        for (uint32_t t = 0; t < k; ++t)
            for (int64_t j = 0; j < c->n; ++j) H[(size_t)t * c->n + c->csc.perm[j]] = t2[(size_t)t * c->n + j];
    }
    return CCD_OK;
}

int32_t ccd_get_residual(ccd_ctx* c, float* csr_val, float* csc_val) {
    if (!c) return CCD_ERR_INVALID;
    if (!c->uploaded) return set_err(c, CCD_ERR_STATE, "nothing uploaded");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    std::vector<float> tmp(c->nnz);
    for (int q = 0; q < 2; ++q) {
        float* out = q == 0 ? csr_val : csc_val;
        const DevLayout& L = q == 0 ? c->csr : c->csc;
        if (!out) continue;
        HIP_TRY(c, hipMemcpy(tmp.data(), L.val, 4 * c->nnz, hipMemcpyDeviceToHost));
        for (uint64_t e = 0; e < c->nnz; ++e) out[L.src[e]] = tmp[e];
    }
    return CCD_OK;
}

int32_t ccd_rmse(ccd_ctx* c, double* out) {
    if (!c || !out) return CCD_ERR_INVALID;
    if (!c->factors) return set_err(c, CCD_ERR_STATE, "no factors");
    if (c->nnz_test == 0) return set_err(c, CCD_ERR_STATE, "no test set");
    HIP_TRY(c, hipSetDevice(c->device));
    if (c->rmse_cap < 1) {
        HIP_TRY(c, dalloc(&c->rmse_dev, 1));
        c->rmse_cap = 1;
    }
    HIP_TRY(c, ccd::launch_rmse(c->trow, c->tcol, c->tval, c->nnz_test, c->W, c->H, false, c->m, c->n, c->k,
                                c->rmse_bsum, c->rmse_dev, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(out, c->rmse_dev, sizeof(double), hipMemcpyDeviceToHost));
    return CCD_OK;
}

int32_t ccd_objective(ccd_ctx* c, float lambda, double* out) {
    if (!c || !out) return CCD_ERR_INVALID;
    if (!c->factors) return set_err(c, CCD_ERR_STATE, "no factors");
    if (!std::isfinite(lambda) || lambda < 0.0f) return set_err(c, CCD_ERR_INVALID, "lambda >= 0 finite");
    HIP_TRY(c, hipSetDevice(c->device));
    if (!c->obj_dev) HIP_TRY(c, dalloc(&c->obj_dev, 3));
    // the block partials fit the RMSE scratch (2 * OBJ_BLOCKS <= RMSE_BLOCKS doubles)
    static_assert(2 * ccd::OBJ_BLOCKS <= ccd::RMSE_BLOCKS, "objective partials exceed the scratch");
    HIP_TRY(c, ccd::launch_objective(c->csr.val, c->nnz, c->csr.ptr, c->csc.ptr, c->W, c->H, c->m, c->n, c->k,
                                     (double)lambda, c->rmse_bsum, c->obj_dev, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(out, c->obj_dev, 3 * sizeof(double), hipMemcpyDeviceToHost));
    return CCD_OK;
}

int32_t ccd_iter_rmse(ccd_ctx* c, double* out, int32_t max_count) {
    if (!c || !out) return CCD_ERR_INVALID;
    const int32_t n = std::min<int32_t>(max_count, (int32_t)c->iter_rmse.size());
    for (int32_t q = 0; q < n; ++q) out[q] = c->iter_rmse[q];
    return n;
}

int32_t ccd_kernel_stats(ccd_ctx* c, int32_t kind, double* total_ms, uint64_t* launches, double* bytes_per_launch) {
    if (!c || kind < 0 || kind >= CCD_K_COUNT) return CCD_ERR_INVALID;
    if (total_ms) *total_ms = c->kind_ms[kind];
    if (launches) *launches = c->kind_n[kind];
    if (bytes_per_launch) *bytes_per_launch = c->kind_bytes[kind];
    return CCD_OK;
}

int32_t ccd_loop_time(ccd_ctx* c, double* seconds) {
    if (!c || !seconds) return CCD_ERR_INVALID;
    *seconds = c->loop_s;
    return CCD_OK;
}

int32_t ccd_copy_probe(ccd_ctx* c, uint64_t bytes, int32_t reps, double* ms_per_copy) {
    if (!c || !ms_per_copy || bytes == 0 || bytes % 16 || reps <= 0) return CCD_ERR_INVALID;
    HIP_TRY(c, hipSetDevice(c->device));
    float4 *a = nullptr, *b = nullptr;
    HIP_TRY(c, dalloc(&a, bytes / 16));
    hipError_t e = dalloc(&b, bytes / 16);
    if (e != hipSuccess) {
        dfree(a);
        return set_err(c, CCD_ERR_ALLOC, "copy probe allocation");
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = CCD_OK;
    if (hipMemsetAsync(a, 0, bytes, c->stream) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess || ccd::launch_copy(a, b, bytes / 16, c->stream) != hipSuccess ||
        hipEventRecord(e0, c->stream) != hipSuccess) {
        rc = set_err(c, CCD_ERR_HIP, "copy probe launch");
    } else {
        for (int r = 0; r < reps && rc == CCD_OK; ++r)
            if (ccd::launch_copy(a, b, bytes / 16, c->stream) != hipSuccess) rc = set_err(c, CCD_ERR_HIP, "copy probe launch");
        float ms = 0.0f;
        if (rc == CCD_OK && (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                             hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            rc = set_err(c, CCD_ERR_HIP, "copy probe timing");
        *ms_per_copy = ms / reps;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(c->stream);
    dfree(a);
    dfree(b);
    return rc;
}

}  // extern "C"
