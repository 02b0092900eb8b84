// Internal state of a solver context, shared by the single-problem C ABI
// (ccd_abi.hip) and the batched solver (batch.hip).  Not part of the
// public interface.
#pragma once
#include "cdmf_hip.h"
#include "ccd_kernels.h"

#include <map>
#include <string>
#include <vector>

namespace ccdi {

// Tiled copy of one layout for the tiled sweeps (see ccd_kernels.h).  The
// index array is tiled once; residual values are tiled at the start of each
// tiled run and written back to the layout's canonical order at its end.
struct TiledLayout {
    int sigma = 0;
    int ib = 0;  // bytes per index of tidx: 2, 3 or 4 (ccd_kernels.h)
    uint32_t ntiles = 0;
    void* tidx = nullptr;
    float* tval = nullptr;
    void* flags = nullptr;  // per lane: u16 words for sigma <= 16, u32 above
    uint32_t* row0 = nullptr;
    uint32_t* segoff = nullptr;
    uint32_t* segrows = nullptr;
    uint8_t* open = nullptr;
    uint32_t* gmax = nullptr;  // largest gathered index per tile
    double2* head = nullptr;  // (g, h) pairs; double-sized so every precision fits
    double2* tail = nullptr;
    ccd::SpanRow* spans = nullptr;  // rows spanning tiles: two-tile ones first
    uint32_t nspan = 0;
    uint32_t nshort = 0;
    // Split layout (rows split at DevLayout::split_at): the tiles hold
    // virtual rows.  Virtual row r < nreal is the part of row r gathered
    // from below the split, virtual row nreal + k the rest of row
    // cold_row[k].  Sweeps leave (g, h) per virtual row in part (or as tile
    // partials); split_finish_kernel adds the two parts of each row and
    // finishes it.
    bool split = false;
    uint32_t nreal = 0;
    uint32_t cold_tile0 = 0xffffffffu;  // first tile holding only second parts (split layouts)
    uint32_t* vpos = nullptr;      // position in the layout of each element, in tiled order
    uint32_t* cold_row = nullptr;  // real row of each second part
    ccd::SplitRow* srows = nullptr;  // per row: tile ranges of its two parts
    double2* part = nullptr;       // per virtual row, double-sized so every precision fits
};

struct DevLayout {
    int64_t rows = 0;   // rows of this layout (n for CSC, m for CSR)
    int64_t gdim = 0;   // length of the gathered vector (m for CSC, n for CSR)
    uint64_t nnz = 0;
    uint32_t* ptr = nullptr;
    void* idx = nullptr;
    bool idx16 = false;
    float* val = nullptr;   // residual
    float* val0 = nullptr;  // R itself, to reset the residual
    uint4* units = nullptr;
    uint32_t nunits = 0;
    uint32_t* split_rows = nullptr;
    uint32_t* split_beg = nullptr;
    uint32_t nsplit = 0;
    double2* partials = nullptr;  // (g, h) per split piece, double-sized for every precision
    uint32_t npieces = 0;
    // (previous, current) factor pairs of the gathered side for fused passes:
    // gdim entries of 8 bytes (FP32) or 4 bytes (BF16)
    void* pair = nullptr;
    // gcum[c]: gathers of entries 0 .. c-1 of the gathered vector per pass
    // (entry c is gathered once per element of row c of the other layout)
    std::vector<uint64_t> gcum;
    // LDS split of a fused pass whose sweep gathers a separate x, for a
    // budget of split_cap factor entries: hot entries of x and of the pairs
    uint64_t split_cap = 0;
    uint32_t split_x = 0, split_p = 0;
    std::vector<uint32_t> hptr;      // host copy of the internal ptr (planning)
    std::vector<uint32_t> src;       // caller's position of each internal element (residual download)
    // Split of the current tiled layout (build_tiled): rows, whose elements
    // ascend by internal index, split at split_at (0: not split); below[i]
    // leading elements of row i are gathered from below it.
    uint32_t split_at = 0;
    std::vector<uint32_t> below;
    std::vector<uint32_t> perm;      // internal row -> caller's row
    TiledLayout tl;
};

struct ProfRec {
    int kind;
    hipEvent_t a, b;
};

}  // namespace ccdi

struct ccd_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    bool uploaded = false, factors = false;
    // true once a run has left the residual at R - W H^T: the next run then
    // adds each rank back from its first outer iteration on
    bool continued = false;
    int64_t m = 0, n = 0;
    uint64_t nnz = 0;
    ccdi::DevLayout csr, csc;
    uint64_t nnz_test = 0;
    uint32_t* trow = nullptr;
    uint32_t* tcol = nullptr;
    float* tval = nullptr;
    uint32_t k = 0;
    float* W = nullptr;
    float* H = nullptr;
    float* u = nullptr;
    float* v = nullptr;
    // bfloat16 copies of W, H, u, v for runs with CCD_PREC_BF16
    uint16_t* W16 = nullptr;
    uint16_t* H16 = nullptr;
    uint16_t* u16 = nullptr;
    uint16_t* v16 = nullptr;
    uint32_t k16 = 0;
    int prec = CCD_PREC_FP32;  // precision of the run in progress
    int ncu = 0;               // compute units
    int lds_block = 0;         // LDS bytes a block may use
    int lds_cu = 0;            // LDS bytes of a CU
    std::map<uint32_t, int> occupancy;  // persistent kernel -> blocks per CU by register use
    double* rmse_bsum = nullptr;
    double* rmse_dev = nullptr;  // one slot per outer iteration
    double* obj_dev = nullptr;   // objective, loss, regulariser
    int32_t rmse_cap = 0;
    std::vector<double> iter_rmse;
    ccd_plan_params plan{4, 8192, 4096, 4096, 16, 0, 1};
    // profiling
    bool prof = false;
    uint32_t prof_kinds = 0;  // bitmask of timed kernel kinds (1 << ccd_kernel_kind)
    std::vector<ccdi::ProfRec> recs;
    std::vector<hipEvent_t> pool;
    size_t pool_used = 0;
    double kind_ms[CCD_K_COUNT] = {};
    uint64_t kind_n[CCD_K_COUNT] = {};
    uint64_t kind_seen[CCD_K_COUNT] = {};  // bound launches of each kind this run (timed or not)
    double kind_bytes[CCD_K_COUNT] = {};
    double loop_s = 0.0;
    // outer iteration captured for replay (use_graph), valid for graph_key
    hipGraphExec_t graph_exec = nullptr;
    std::vector<uint64_t> graph_key;
    uint64_t gen = 0;  // bumped whenever buffers or the plan change
};

namespace ccdi {

int set_err(ccd_ctx* c, int code, const std::string& msg);

#define HIP_TRY(ctx, expr)                                                                              \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            return ccdi::set_err((ctx), CCD_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T>
hipError_t dalloc(T** p, size_t count) {
    *p = nullptr;
    return hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * (count ? count : 1));
}

void dfree(void* p);
void drop_graph(ccd_ctx* c);
// Bytes per index of the tiled copy of L: 2 up to 2^16 gathered entries, 3
// up to 2^24, else 4.
int tiled_index_bytes(const DevLayout& L);
// Tile descriptors of layout L for tiles of 64 * sigma elements (cached).
int build_tiled(ccd_ctx* c, DevLayout& L, int sigma, uint32_t split_at);

}  // namespace ccdi
