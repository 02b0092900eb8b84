// C ABI of the batched solver (include/cdmf_hip.h, ccd_batch_*): many
// independent factorizations, each an uploaded ccd_ctx, solved by one
// persistent launch of batch_kernel (kernels/batch_kernels.hip), optionally
// replayed from a captured HIP graph, with the tile shape and the split of
// the grid among problems chosen by a measuring autotuner.
#include "ccd_ctx.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <numeric>
#include <string>
#include <vector>

struct ccd_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    struct Item {
        ccd_ctx* ctx;
        ccd_params p;
    };
    std::vector<Item> items;
    // problems of the persistent launch (indices into items); the others
    // of a run are solved alone first (ccd_batch_plan.solo)
    std::vector<size_t> active;
    ccd_batch_plan plan{16, 0, 1.0f, 0, 0};
    int ncu = 0;
    int lds_cu = 0, lds_block = 0;  // LDS bytes of a CU and of a block
    uint32_t lds_floats = 0;        // dynamic LDS of a block of the last grid_blocks, in 4-byte entries
    // device copies of the launch description
    ccd::BatchProblem* d_probs = nullptr;
    uint32_t* d_block_problem = nullptr;
    uint32_t* d_barriers = nullptr;
    uint32_t* d_tickets = nullptr;  // 8 x 16 words
    int* d_status = nullptr;
    size_t cap_probs = 0, cap_blocks = 0;
    // captured run
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<uint64_t> graph_key;
    hipEvent_t ev_a = nullptr, ev_b = nullptr;
    double last_seconds = 0.0;
    std::vector<uint32_t> team;  // blocks per problem of the last launch
    // the last tune's whole-run times: solo = 0, 1, ..., then the deal
    // candidates of tune_deals (with the best solo count)
    std::vector<double> tune_solo_ms;
};

namespace {

int berr(ccd_batch* b, int code, const std::string& msg) {
    if (b) b->err = msg;
    return code;
}

#define BATCH_TRY(b, expr)                                                                             \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) return berr((b), CCD_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// A context call's error, carried over with the problem's index.
int ctx_err(ccd_batch* b, size_t i, int rc) {
    return berr(b, rc, "problem " + std::to_string(i) + ": " + b->items[i].ctx->err);
}

void drop_graph(ccd_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_key.clear();
}

ccdi::DevLayout& first_layout(ccd_ctx* c, const ccd_params& p) {
    return p.order == CCD_ORDER_H_FIRST ? c->csc : c->csr;
}
ccdi::DevLayout& second_layout(ccd_ctx* c, const ccd_params& p) {
    return p.order == CCD_ORDER_H_FIRST ? c->csr : c->csc;
}

ccd::BatchLayout batch_layout(const ccdi::DevLayout& L) {
    const ccdi::TiledLayout& T = L.tl;
    ccd::BatchLayout d;
    d.ptr = L.ptr;
    d.tidx = T.tidx;
    d.tval = T.tval;
    d.lane_flags = T.flags;
    d.tile_row0 = T.row0;
    d.tile_segoff = T.segoff;
    d.segrows = T.segrows;
    d.tile_open = T.open;
    d.head = T.head;
    d.tail = T.tail;
    d.spans = T.spans;
    d.pair = L.pair;
    d.nnz = L.nnz;
    d.ntiles = T.ntiles;
    d.nspan = T.nspan;
    d.nshort = T.nshort;
    d.ib = (uint32_t)T.ib;
    d.rows = (uint32_t)L.rows;
    d.gdim = (uint32_t)L.gdim;
    d.stage = 0;
    return d;
}

// Tiled layouts of every problem for sigma, pair buffers, FP32 state.
int prepare(ccd_batch* b, int sigma) {
    for (size_t i : b->active) {
        ccd_ctx* c = b->items[i].ctx;
        c->prec = CCD_PREC_FP32;
        for (ccdi::DevLayout* L : {&c->csr, &c->csc}) {
            if (int rc = ccdi::build_tiled(c, *L, sigma, 0u)) return ctx_err(b, i, rc);
            if (!L->pair) {
                double2* d = nullptr;
                if (ccdi::dalloc(&d, ((size_t)L->gdim + 1) / 2) != hipSuccess)
                    return berr(b, CCD_ERR_ALLOC, "problem " + std::to_string(i) + ": pair buffer");
                L->pair = d;
            }
        }
    }
    return CCD_OK;
}

// Work of one problem in units of layout elements streamed: (2 Tin + 2)
// passes per rank, k ranks, T outer iterations.
double problem_work(const ccd_batch::Item& it) {
    return (double)it.p.maxiter * it.p.k * (2.0 * it.p.maxinneriter + 2.0) * (double)it.ctx->nnz;
}

// Blocks per problem and their placement.  Sizes: one block each, then one
// at a time to the problem with the largest work^e per block, never more
// than its tiles can use nor than one XCD holds (blocks / 8).  Placement:
// largest team first into the XCD with the most free blocks (shrunk if it
// no longer fits), then each XCD's leftover blocks go to its own teams by
// the same rule.  first[i]: the team's first virtual block id.
struct Split {
    std::vector<uint32_t> nb, first;
};

Split split_grid(const ccd_batch* b, uint32_t blocks, float expo, int sigma) {
    const size_t np = b->active.size();
    const uint32_t C = blocks / 8;
    std::vector<uint32_t> nb(np, 1u), cap(np), xcd(np);
    std::vector<double> w(np);
    for (size_t i = 0; i < np; ++i) {
        ccd_ctx* c = b->items[b->active[i]].ctx;
        const uint64_t te = 64ull * (uint64_t)sigma;
        const uint64_t tiles = (c->nnz + te - 1) / te;
        cap[i] = (uint32_t)std::max<uint64_t>(1, (tiles + ccd::BATCH_BLOCK / 64 - 1) / (ccd::BATCH_BLOCK / 64));
        cap[i] = std::min(cap[i], C);
        w[i] = std::pow(std::max(problem_work(b->items[b->active[i]]), 1.0), (double)expo);
    }
    auto grow = [&](uint32_t left, int only_xcd) {
        for (; left > 0; --left) {
            size_t best = np;
            double bw = -1.0;
            for (size_t i = 0; i < np; ++i)
                if ((only_xcd < 0 || (int)xcd[i] == only_xcd) && nb[i] < cap[i] && w[i] / nb[i] > bw) {
                    bw = w[i] / nb[i];
                    best = i;
                }
            if (best == np) return left;
            ++nb[best];
        }
        return 0u;
    };
    grow(blocks - (uint32_t)np, -1);
    std::vector<size_t> order(np);
    for (size_t i = 0; i < np; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t c) { return nb[a] > nb[c]; });
    std::vector<uint32_t> freeb(8, C);
    for (size_t i : order) {
        const uint32_t x = (uint32_t)(std::max_element(freeb.begin(), freeb.end()) - freeb.begin());
        nb[i] = std::max(1u, std::min(nb[i], freeb[x]));  // np <= blocks keeps freeb[x] >= 1 here
        xcd[i] = x;
        freeb[x] -= nb[i];
    }
    for (uint32_t x = 0; x < 8; ++x) freeb[x] = grow(freeb[x], (int)x);
    Split sp{nb, std::vector<uint32_t>(np)};
    std::vector<uint32_t> used(8, 0);
    for (size_t i = 0; i < np; ++i) {
        sp.first[i] = xcd[i] * C + used[xcd[i]];
        used[xcd[i]] += nb[i];
    }
    return sp;
}

// Tiles per wave from which a team's sweep deals its tiles (plan.deal: 0 is
// the default).
constexpr uint32_t DEAL_DEFAULT = 16;
uint32_t deal_min(uint32_t deal) { return deal ? deal : DEAL_DEFAULT; }
// deal values the tuner tries after the default
constexpr uint32_t tune_deals[] = {8u, 32u, 0xffffffffu};

// Upload the launch description (trial: the autotuner's short run).
int describe(ccd_batch* b, int sigma, uint32_t blocks, const Split& sp, bool trial, ccd::BatchArgs* args) {
    const std::vector<uint32_t>& nb = sp.nb;
    const size_t np = b->active.size();
    std::vector<ccd::BatchProblem> probs(np);
    std::vector<uint32_t> owner(blocks, 0xffffffffu);
    for (size_t i = 0; i < np; ++i) {
        const ccd_batch::Item& it = b->items[b->active[i]];
        ccd_ctx* c = it.ctx;
        const bool hfirst = it.p.order == CCD_ORDER_H_FIRST;
        ccd::BatchProblem& P = probs[i];
        P.F = batch_layout(first_layout(c, it.p));
        P.S = batch_layout(second_layout(c, it.p));
        // stage a layout's gathered vector when each block streams at least
        // twice as many elements as it would stage
        for (ccd::BatchLayout* L : {&P.F, &P.S})
            L->stage = L->nnz / std::max(nb[i], 1u) >= 2ull * std::min<uint64_t>(L->gdim, b->lds_floats) ? 1u : 0u;
        P.XF = hfirst ? c->H : c->W;
        P.XS = hfirst ? c->W : c->H;
        P.vF = hfirst ? c->v : c->u;
        P.vS = hfirst ? c->u : c->v;
        P.k = it.p.k;
        P.T = trial ? 0u : (uint32_t)it.p.maxiter;
        P.Tin = (uint32_t)it.p.maxinneriter;
        P.continued = c->continued ? 1u : 0u;
        P.first = sp.first[i];
        P.nblocks = nb[i];
        P.lambda = it.p.lambda;
        // a trial runs about 1/64 of each problem's sweeps, so that the
        // slowest team sets its time as it sets the real run's
        P.trial_reps = (uint32_t)std::max(1.0, std::round((double)it.p.k * it.p.maxiter * it.p.maxinneriter / 64.0));
        for (uint32_t v = 0; v < nb[i]; ++v) {
            if (owner[sp.first[i] + v] != 0xffffffffu) return berr(b, CCD_ERR_INVALID, "internal: overlapping teams");
            owner[sp.first[i] + v] = (uint32_t)i;
        }
    }
    if (np > b->cap_probs) {
        ccdi::dfree(b->d_probs);
        ccdi::dfree(b->d_barriers);
        b->d_probs = nullptr;
        b->d_barriers = nullptr;
        BATCH_TRY(b, ccdi::dalloc(&b->d_probs, np));
        BATCH_TRY(b, ccdi::dalloc(&b->d_barriers, np));
        b->cap_probs = np;
    }
    if (blocks > b->cap_blocks) {
        ccdi::dfree(b->d_block_problem);
        b->d_block_problem = nullptr;
        BATCH_TRY(b, ccdi::dalloc(&b->d_block_problem, blocks));
        b->cap_blocks = blocks;
    }
    BATCH_TRY(b, ccd::copy_sync(b->stream, b->d_probs, probs.data(), sizeof(ccd::BatchProblem) * np, hipMemcpyHostToDevice));
    BATCH_TRY(b, ccd::copy_sync(b->stream, b->d_block_problem, owner.data(), sizeof(uint32_t) * blocks, hipMemcpyHostToDevice));
    args->probs = b->d_probs;
    args->block_problem = b->d_block_problem;
    args->nprob = (uint32_t)np;
    args->nblocks = blocks;
    args->barriers = b->d_barriers;
    args->xcd_tickets = b->d_tickets;
    args->status = b->d_status;
    args->wait_ticks = 100000000ull * 4;  // 4 s of the 100 MHz wall clock
    args->trial = trial ? 1u : 0u;
    args->lds_floats = b->lds_floats;
    args->deal_min = deal_min(b->plan.deal);
    return CCD_OK;
}

// Launches of one run on b->stream: per-problem set-up, the persistent
// kernel, per-problem tear-down.  Graph-safe (fill kernels, no memsets).
int enqueue_run(ccd_batch* b, int sigma, const ccd::BatchArgs& a, bool trial) {
    hipStream_t s = b->stream;
    BATCH_TRY(b, ccd::launch_fill(reinterpret_cast<float*>(b->d_barriers), 0.0f, a.nprob, s));
    BATCH_TRY(b, ccd::launch_fill(reinterpret_cast<float*>(b->d_status), 0.0f, 1, s));
    BATCH_TRY(b, ccd::launch_fill(reinterpret_cast<float*>(b->d_tickets), 0.0f, 8 * 16, s));
    for (size_t i : b->active) {
        const ccd_batch::Item& it = b->items[i];
        ccd_ctx* c = it.ctx;
        ccdi::DevLayout& F = first_layout(c, it.p);
        ccdi::DevLayout& S = second_layout(c, it.p);
        const bool hfirst = it.p.order == CCD_ORDER_H_FIRST;
        if (!trial && it.p.zero_first && !c->continued)
            BATCH_TRY(b, ccd::launch_fill(hfirst ? c->H : c->W, 0.0f, (uint64_t)it.p.k * F.rows, s));
        if (F.tl.ntiles) BATCH_TRY(b, ccd::launch_retile_f32(F.val, F.tl.tval, F.nnz, F.tl.ntiles, sigma, true, s));
        if (S.tl.ntiles) BATCH_TRY(b, ccd::launch_retile_f32(S.val, S.tl.tval, S.nnz, S.tl.ntiles, sigma, true, s));
        BATCH_TRY(b, ccd::launch_fill(hfirst ? c->v : c->u, 0.0f, (uint64_t)F.rows, s));
        BATCH_TRY(b, ccd::launch_fill(hfirst ? c->u : c->v, 0.0f, (uint64_t)S.rows, s));
    }
    BATCH_TRY(b, ccd::launch_batch(a, sigma, s));
    if (!trial) {
        for (size_t i : b->active) {
            const ccd_batch::Item& it = b->items[i];
            ccd_ctx* c = it.ctx;
            ccdi::DevLayout& F = first_layout(c, it.p);
            ccdi::DevLayout& S = second_layout(c, it.p);
            if (F.tl.ntiles)
                BATCH_TRY(b, ccd::launch_retile_f32(F.tl.tval, F.val, F.nnz, F.tl.ntiles, sigma, false, s));
            if (S.tl.ntiles)
                BATCH_TRY(b, ccd::launch_retile_f32(S.tl.tval, S.val, S.nnz, S.tl.ntiles, sigma, false, s));
        }
    }
    return CCD_OK;
}

// Grid of the persistent launch: co-resident blocks, a multiple of 8.
// Each block's dynamic LDS (staged hot prefixes) is its share of the CU's
// LDS at the occupancy its registers allow, so staging never lowers the
// number of co-resident blocks; sets b->lds_floats.
int grid_blocks(ccd_batch* b, int sigma, uint32_t want, uint32_t* out) {
    const int regs = ccd::batch_blocks_per_cu(sigma, 0);
    if (regs <= 0) return berr(b, CCD_ERR_HIP, "occupancy query of the batched kernel failed");
    const int share = std::min(b->lds_block, b->lds_cu / regs) - 1024;  // 1 KiB for the kernel's own LDS
    // at least 256 bytes: fused passes read entry 0 of the staged area for
    // lanes past a (possibly empty) staged prefix
    const size_t lds = share > 256 ? (size_t)share & ~(size_t)255 : 256;
    const int bpc = ccd::batch_blocks_per_cu(sigma, lds);
    if (bpc <= 0) return berr(b, CCD_ERR_HIP, "occupancy query of the batched kernel failed");
    b->lds_floats = (uint32_t)(lds / sizeof(float));
    uint32_t most = (uint32_t)(bpc * b->ncu) & ~7u;
    uint32_t g = want ? std::min(want, most) & ~7u : most;
    if (g < b->active.size())
        return berr(b, CCD_ERR_UNSUPPORTED, "more problems (" + std::to_string(b->active.size()) +
                                                ") than co-resident blocks (" + std::to_string(g) + "): split the batch");
    *out = g;
    return CCD_OK;
}

int check_status(ccd_batch* b) {
    int st = 0;
    BATCH_TRY(b, hipMemcpy(&st, b->d_status, sizeof(int), hipMemcpyDeviceToHost));
    if (st == 2) return berr(b, CCD_ERR_HIP, "blocks were not spread evenly over 8 XCDs: team barriers are unsafe here");
    if (st) return berr(b, CCD_ERR_HIP, "a team barrier wait ran out (blocks not co-resident?)");
    return CCD_OK;
}

// A short trial launch of the problems in b->active (no state change).
int timed_trial(ccd_batch* b, int sigma, uint32_t blocks, float expo, double* ms) {
    if (int rc = prepare(b, sigma)) return rc;
    const Split sp = split_grid(b, blocks, expo, sigma);
    ccd::BatchArgs a{};
    if (int rc = describe(b, sigma, blocks, sp, true, &a)) return rc;
    if (int rc = enqueue_run(b, sigma, a, true)) return rc;  // warm
    BATCH_TRY(b, hipEventRecord(b->ev_a, b->stream));
    if (int rc = enqueue_run(b, sigma, a, true)) return rc;
    BATCH_TRY(b, hipEventRecord(b->ev_b, b->stream));
    BATCH_TRY(b, hipEventSynchronize(b->ev_b));
    if (int rc = check_status(b)) return rc;
    float f = 0.0f;
    BATCH_TRY(b, hipEventElapsedTime(&f, b->ev_a, b->ev_b));
    *ms = f;
    return CCD_OK;
}

}  // namespace

extern "C" {

int32_t ccd_batch_create(int32_t device, ccd_batch** out) {
    if (!out) return CCD_ERR_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return CCD_ERR_HIP;
    if (device < 0 || device >= count) return CCD_ERR_INVALID;
    ccd_batch* b = new (std::nothrow) ccd_batch();
    if (!b) return CCD_ERR_ALLOC;
    b->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&b->ev_a) != hipSuccess || hipEventCreate(&b->ev_b) != hipSuccess ||
        ccdi::dalloc(&b->d_status, 1) != hipSuccess || ccdi::dalloc(&b->d_tickets, 8 * 16) != hipSuccess) {
        delete b;
        return CCD_ERR_HIP;
    }
    if (hipDeviceGetAttribute(&b->ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || b->ncu <= 0)
        b->ncu = 1;
    if (hipDeviceGetAttribute(&b->lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device) != hipSuccess)
        b->lds_cu = 0;
    if (hipDeviceGetAttribute(&b->lds_block, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess)
        b->lds_block = 0;
    *out = b;
    return CCD_OK;
}

int32_t ccd_batch_destroy(ccd_batch* b) {
    if (!b) return CCD_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    ccdi::dfree(b->d_probs);
    ccdi::dfree(b->d_block_problem);
    ccdi::dfree(b->d_barriers);
    ccdi::dfree(b->d_status);
    ccdi::dfree(b->d_tickets);
    if (b->ev_a) (void)hipEventDestroy(b->ev_a);
    if (b->ev_b) (void)hipEventDestroy(b->ev_b);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return CCD_OK;
}

const char* ccd_batch_last_error(ccd_batch* b) { return b ? b->err.c_str() : "null batch"; }

int32_t ccd_batch_add(ccd_batch* b, ccd_ctx* c, const ccd_params* p) {
    if (!b || !c || !p) return CCD_ERR_INVALID;
    if (!c->uploaded || !c->factors) return berr(b, CCD_ERR_STATE, "ccd_batch_add needs ccd_upload and ccd_set_factors");
    if (c->device != b->device) return berr(b, CCD_ERR_INVALID, "context on another device");
    if (p->k != c->k) return berr(b, CCD_ERR_INVALID, "params.k differs from the k of ccd_set_factors");
    if (p->maxiter < 0 || p->maxinneriter < 1 || !std::isfinite(p->lambda) || p->lambda < 0.0f)
        return berr(b, CCD_ERR_INVALID, "batched problems need maxiter >= 0, maxinneriter >= 1, lambda >= 0 finite");
    if (p->precision != CCD_PREC_FP32) return berr(b, CCD_ERR_UNSUPPORTED, "batched problems run in FP32");
    if (p->order != CCD_ORDER_H_FIRST && p->order != CCD_ORDER_W_FIRST) return berr(b, CCD_ERR_INVALID, "bad order");
    for (const ccd_batch::Item& it : b->items)
        if (it.ctx == c) return berr(b, CCD_ERR_INVALID, "context added twice");
    b->items.push_back({c, *p});
    drop_graph(b);
    return (int32_t)(b->items.size() - 1);
}

int32_t ccd_batch_set_plan(ccd_batch* b, const ccd_batch_plan* pp) {
    if (!b || !pp) return CCD_ERR_INVALID;
    if (pp->sigma != 8 && pp->sigma != 16) return berr(b, CCD_ERR_INVALID, "batch plan: sigma must be 8 or 16");
    if (!(pp->work_exponent > 0.0f && pp->work_exponent <= 2.0f))
        return berr(b, CCD_ERR_INVALID, "batch plan: work_exponent in (0, 2]");
    if (pp->blocks % 8u) return berr(b, CCD_ERR_INVALID, "batch plan: blocks must be 0 or a multiple of 8");
    if (pp->solo > 64u) return berr(b, CCD_ERR_INVALID, "batch plan: solo must be at most 64");
    b->plan = *pp;
    drop_graph(b);
    return CCD_OK;
}

int32_t ccd_batch_get_plan(ccd_batch* b, ccd_batch_plan* out) {
    if (!b || !out) return CCD_ERR_INVALID;
    *out = b->plan;
    return CCD_OK;
}

}  // extern "C"

namespace {

// The `solo` problems of largest work (ties: lower index first) and the
// rest, in index order.
void choose_solo(ccd_batch* b, uint32_t solo, std::vector<size_t>& alone) {
    const size_t np = b->items.size();
    std::vector<size_t> order(np);
    std::iota(order.begin(), order.end(), (size_t)0);
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t x, size_t y) { return problem_work(b->items[x]) > problem_work(b->items[y]); });
    const size_t ns = std::min<size_t>(solo, np);
    alone.assign(order.begin(), order.begin() + ns);
    std::sort(alone.begin(), alone.end());
    std::vector<uint8_t> is_alone(np, 0);
    for (size_t i : alone) is_alone[i] = 1;
    b->active.clear();
    for (size_t i = 0; i < np; ++i)
        if (!is_alone[i]) b->active.push_back(i);
}

// One problem on the whole GPU: ccd_run of the tiled sweeps on the plain
// tile layout at the batch's sigma, exactly what the persistent kernel
// computes for it.  The context's own plan is put back afterwards.
int run_alone(ccd_batch* b, size_t i, int sigma, bool graph) {
    ccd_ctx* c = b->items[i].ctx;
    const ccd_plan_params keep = c->plan;
    c->plan.sigma = (uint32_t)sigma;
    c->plan.split_rows = 0;
    ccd_params p = b->items[i].p;
    p.variant = CCD_VARIANT_TILED;
    p.precision = CCD_PREC_FP32;
    p.rmse_every = 0;
    p.profile = 0;
    p.use_graph = graph ? 1 : 0;
    const int rc = ccd_run(c, &p, nullptr, nullptr);
    c->plan = keep;
    return rc ? ctx_err(b, i, rc) : CCD_OK;
}

// Device copies of each problem's factors, to put back the start of a run
// (tuner runs); the residual goes back to R as in ccd_set_factors.
struct Snapshot {
    std::vector<float*> W, H;
};

int take_snapshot(ccd_batch* b, Snapshot& sn) {
    for (const ccd_batch::Item& it : b->items) {
        ccd_ctx* c = it.ctx;
        float *w = nullptr, *h = nullptr;
        BATCH_TRY(b, ccdi::dalloc(&w, (size_t)c->k * c->m));
        sn.W.push_back(w);
        BATCH_TRY(b, ccdi::dalloc(&h, (size_t)c->k * c->n));
        sn.H.push_back(h);
        BATCH_TRY(b, ccd::copy_sync(c->stream, w, c->W, 4 * (size_t)c->k * c->m, hipMemcpyDeviceToDevice));
        BATCH_TRY(b, ccd::copy_sync(c->stream, h, c->H, 4 * (size_t)c->k * c->n, hipMemcpyDeviceToDevice));
    }
    return CCD_OK;
}

int restore_snapshot(ccd_batch* b, const Snapshot& sn) {
    for (size_t i = 0; i < b->items.size(); ++i) {
        ccd_ctx* c = b->items[i].ctx;
        BATCH_TRY(b, ccd::copy_sync(c->stream, c->W, sn.W[i], 4 * (size_t)c->k * c->m, hipMemcpyDeviceToDevice));
        BATCH_TRY(b, ccd::copy_sync(c->stream, c->H, sn.H[i], 4 * (size_t)c->k * c->n, hipMemcpyDeviceToDevice));
        BATCH_TRY(b, ccd::copy_sync(c->stream, c->csr.val, c->csr.val0, 4 * c->nnz, hipMemcpyDeviceToDevice));
        BATCH_TRY(b, ccd::copy_sync(c->stream, c->csc.val, c->csc.val0, 4 * c->nnz, hipMemcpyDeviceToDevice));
        c->continued = false;
    }
    return CCD_OK;
}

void free_snapshot(Snapshot& sn) {
    for (float* p : sn.W) ccdi::dfree(p);
    for (float* p : sn.H) ccdi::dfree(p);
    sn.W.clear();
    sn.H.clear();
}

}  // namespace

extern "C" {

int32_t ccd_batch_tune(ccd_batch* b, double* best_ms) {
    if (!b) return CCD_ERR_INVALID;
    if (b->items.empty()) return berr(b, CCD_ERR_STATE, "empty batch");
    BATCH_TRY(b, hipSetDevice(b->device));
    ccd_batch_plan best = b->plan;
    double bms = INFINITY;
    b->active.resize(b->items.size());
    std::iota(b->active.begin(), b->active.end(), (size_t)0);
    for (uint32_t sigma : {16u, 8u}) {
        uint32_t blocks = 0;
        if (int rc = grid_blocks(b, (int)sigma, 0, &blocks)) return rc;
        for (float expo : {1.0f, 0.85f, 0.7f}) {
            double ms = 0.0;
            if (int rc = timed_trial(b, (int)sigma, blocks, expo, &ms)) return rc;
            if (ms < bms) {
                bms = ms;
                best = ccd_batch_plan{sigma, 0u, expo, 0u, 0u};
            }
        }
    }
    // Solving the largest problems alone, on the whole GPU, before the
    // persistent launch takes the rest: each team of the launch is confined
    // to one XCD, so a problem with more than an eighth of the work bounds
    // the launch.  Whole runs of each candidate, from the problems' start
    // state (put back after every run), twice each (the first warms up);
    // only from a fresh start (no problem continues a run).
    bool fresh = b->items.size() > 1;
    for (const ccd_batch::Item& it : b->items) fresh = fresh && !it.ctx->continued;
    b->tune_solo_ms.clear();
    if (fresh) {
        Snapshot sn;
        int rc = take_snapshot(b, sn);
        double best_s = INFINITY;
        uint32_t best_solo = 0;
        const uint32_t most = (uint32_t)std::min<size_t>(3, b->items.size() - 1);
        for (uint32_t solo = 0; !rc && solo <= most; ++solo) {
            b->plan = best;
            b->plan.solo = solo;
            double secs = 0.0;
            for (int rep = 0; rep < 2 && !rc; ++rep) {
                drop_graph(b);
                rc = ccd_batch_run(b, 0);
                if (!rc) secs = b->last_seconds;
                if (!rc) rc = restore_snapshot(b, sn);
            }
            if (!rc) b->tune_solo_ms.push_back(1e3 * secs);
            if (!rc && secs < best_s) {
                best_s = secs;
                best_solo = solo;
            }
        }
        // Then the tile deal of the sweeps (the default, 16 tiles per wave,
        // was the runs above): more or fewer teams' sweeps dealt.
        best.solo = best_solo;
        uint32_t best_deal = 0;
        for (uint32_t deal : tune_deals) {
            if (rc) break;
            b->plan = best;
            b->plan.deal = deal;
            double secs = 0.0;
            for (int rep = 0; rep < 2 && !rc; ++rep) {
                drop_graph(b);
                rc = ccd_batch_run(b, 0);
                if (!rc) secs = b->last_seconds;
                if (!rc) rc = restore_snapshot(b, sn);
            }
            if (!rc) b->tune_solo_ms.push_back(1e3 * secs);
            if (!rc && secs < best_s) {
                best_s = secs;
                best_deal = deal;
            }
        }
        free_snapshot(sn);
        if (rc) return rc;
        best.deal = best_deal;
        bms = 1e3 * best_s;
    }
    b->plan = best;
    drop_graph(b);
    if (best_ms) *best_ms = bms;
    return CCD_OK;
}

int32_t ccd_batch_run(ccd_batch* b, int32_t use_graph) {
    if (!b) return CCD_ERR_INVALID;
    if (b->items.empty()) return berr(b, CCD_ERR_STATE, "empty batch");
    BATCH_TRY(b, hipSetDevice(b->device));
    const int sigma = (int)b->plan.sigma;
    for (size_t i = 0; i < b->items.size(); ++i)
        if (b->items[i].p.zero_first && b->items[i].ctx->continued)
            return berr(b, CCD_ERR_INVALID, "problem " + std::to_string(i) + ": zero_first on a continued context");
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<size_t> alone;
    choose_solo(b, b->plan.solo, alone);
    for (size_t i : alone)
        if (int rc = run_alone(b, i, sigma, use_graph != 0)) return rc;
    const double alone_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    b->team.assign(b->items.size(), 0u);
    b->last_seconds = alone_s;
    if (b->active.empty()) return CCD_OK;
    if (int rc = prepare(b, sigma)) return rc;
    uint32_t blocks = 0;
    if (int rc = grid_blocks(b, sigma, b->plan.blocks, &blocks)) return rc;
    const Split sp = split_grid(b, blocks, b->plan.work_exponent, sigma);
    ccd::BatchArgs a{};
    if (int rc = describe(b, sigma, blocks, sp, false, &a)) return rc;
    for (size_t q = 0; q < b->active.size(); ++q) b->team[b->active[q]] = sp.nb[q];

    // the captured sequence depends on the plan, the grid, the problems of
    // the launch, each one's continued / zero_first state and the buffers
    // it was captured with: a context's generation moves whenever it
    // reallocates any of them (upload, a new k, a rebuilt tile layout), and
    // the addresses the capture bakes in are keyed as well
    std::vector<uint64_t> key{(uint64_t)sigma, blocks, (uint64_t)b->active.size(), deal_min(b->plan.deal)};
    for (size_t i : b->active) {
        const ccd_batch::Item& it = b->items[i];
        const ccd_ctx* c = it.ctx;
        key.push_back(i);
        key.push_back((it.ctx->continued ? 1u : 0u) | (it.p.zero_first ? 2u : 0u));
        key.push_back(c->gen);
        for (const void* q : {(const void*)c->W, (const void*)c->H, (const void*)c->u, (const void*)c->v,
                              (const void*)c->csr.val, (const void*)c->csc.val, (const void*)c->csr.tl.tval,
                              (const void*)c->csc.tl.tval, (const void*)c->csr.pair, (const void*)c->csc.pair})
            key.push_back((uint64_t)(uintptr_t)q);
    }
    BATCH_TRY(b, hipEventRecord(b->ev_a, b->stream));
    if (use_graph) {
        if (!b->exec || key != b->graph_key) {
            drop_graph(b);
            BATCH_TRY(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            int rc = enqueue_run(b, sigma, a, false);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(b->stream, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            BATCH_TRY(b, e);
            b->graph = g;
            BATCH_TRY(b, hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
            b->graph_key = key;
        }
        BATCH_TRY(b, hipGraphLaunch(b->exec, b->stream));
    } else {
        if (int rc = enqueue_run(b, sigma, a, false)) return rc;
    }
    BATCH_TRY(b, hipEventRecord(b->ev_b, b->stream));
    BATCH_TRY(b, hipEventSynchronize(b->ev_b));
    if (int rc = check_status(b)) return rc;
    float ms = 0.0f;
    BATCH_TRY(b, hipEventElapsedTime(&ms, b->ev_a, b->ev_b));
    b->last_seconds = alone_s + ms * 1e-3;
    for (size_t i : b->active)
        if (b->items[i].p.maxiter > 0) b->items[i].ctx->continued = true;
    return CCD_OK;
}

int32_t ccd_batch_time(ccd_batch* b, double* seconds) {
    if (!b || !seconds) return CCD_ERR_INVALID;
    *seconds = b->last_seconds;
    return CCD_OK;
}

int32_t ccd_batch_tune_times(ccd_batch* b, double* ms, int32_t max_count) {
    if (!b || !ms) return CCD_ERR_INVALID;
    const int32_t n = std::min<int32_t>(max_count, (int32_t)b->tune_solo_ms.size());
    for (int32_t i = 0; i < n; ++i) ms[i] = b->tune_solo_ms[i];
    return (int32_t)b->tune_solo_ms.size();
}

int32_t ccd_batch_teams(ccd_batch* b, uint32_t* blocks, int32_t max_count) {
    if (!b || !blocks) return CCD_ERR_INVALID;
    const int32_t n = std::min<int32_t>(max_count, (int32_t)b->team.size());
    for (int32_t i = 0; i < n; ++i) blocks[i] = b->team[i];
    return (int32_t)b->team.size();
}

}  // extern "C"
