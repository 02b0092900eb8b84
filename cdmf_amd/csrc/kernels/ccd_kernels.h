// Internal interface between the C ABI (ccd_abi.hip) and the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tile_layout.h"  // SpanRow, SplitRow

#include <type_traits>

namespace ccd {

// (previous rank's final value, current rank's value) of one factor entry,
// interleaved so that a fused residual+sweep pass gathers both in one access.
template <class TF>
struct alignas(2 * sizeof(TF)) FPair {
    TF prev, cur;
};

template <class TF>
__device__ __forceinline__ float tof(TF v) {
    if constexpr (sizeof(TF) == 2) return __uint_as_float((uint32_t)v << 16);
    else return v;
}

template <class TF>
__device__ __forceinline__ FPair<TF> gather_pair(const FPair<TF>* lds, uint32_t hot, const void* g, uint32_t c) {
    return c < hot ? lds[c] : static_cast<const FPair<TF>*>(g)[c];
}

// Gather of a factor value: from the LDS-staged hot prefix when c < hot.
// (The compiler makes this one flat load through the selected pointer.)
template <class TF>
__device__ __forceinline__ float gather(const TF* lds, uint32_t hot, const void* g, uint32_t c) {
    if (c < hot) {
        if constexpr (sizeof(TF) == 2) return __uint_as_float((uint32_t)lds[c] << 16);
        else return lds[c];
    }
    if constexpr (sizeof(TF) == 2) return __uint_as_float((uint32_t) static_cast<const uint16_t*>(g)[c] << 16);
    else return static_cast<const float*>(g)[c];
}

// Pointers of a fixed address space (LDS, global).  The fused tiled passes
// read LDS and global memory through these, so the compiler cannot turn a
// choice between the two into one flat load through a selected pointer
// (which waits on both the LDS and the memory counters).  gfx950's compiler
// also rejects a flat select between an LDS and a global pointer in a kernel
// that reads LDS through lds_ptr ("Operand has incorrect register class"),
// so a kernel uses one form or the other.
template <class T>
using lds_ptr = const __attribute__((address_space(3))) T*;
template <class T>
using glb_ptr = const __attribute__((address_space(1))) T*;

// Read of an array that no kernel writes while the reader runs (the tiled
// layout's per-tile descriptors), through the constant address space: at a
// wave-uniform index it is a scalar load, whatever stores the kernel makes.
template <class T>
__device__ __forceinline__ T ld_const(const T* p, size_t i) {
    return ((const __attribute__((address_space(4))) T*)p)[i];
}

// Copy of n elements into LDS by a block of BLOCK threads.  Each thread
// issues a round of up to U loads before its first LDS write, so a round
// costs one memory round trip rather than one per element (the plain loop
// waits for every load before its write).
template <int BLOCK, class T>
__device__ __forceinline__ void stage_lds(T* __restrict__ dst, const T* __restrict__ src, uint32_t n) {
    constexpr int U = sizeof(T) >= 16 ? 4 : sizeof(T) >= 8 ? 8 : 16;
    for (uint32_t base = threadIdx.x; base < n; base += U * BLOCK) {
        T v[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (base + u * BLOCK < n) v[u] = src[base + u * BLOCK];
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (base + u * BLOCK < n) dst[base + u * BLOCK] = v[u];
    }
}

// Copy made by a set-up call (host to device, device to device): ordered on
// the object's stream s and complete when it returns.  The library's streams
// are non-blocking, so their kernels do not wait for copies on the null
// stream, and a device-to-device hipMemcpy may return before its copy ends.
inline hipError_t copy_sync(hipStream_t s, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    if (bytes == 0) return hipSuccess;
    const hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, s);
    const hipError_t w = hipStreamSynchronize(s);
    return e != hipSuccess ? e : w;
}

// A pair is read as one word (8 bytes for FP32, 4 for BF16) and split.
template <class TF>
using PairWord = typename std::conditional<sizeof(TF) == 4, uint64_t, uint32_t>::type;
template <class TF>
__device__ __forceinline__ FPair<TF> to_pair(PairWord<TF> w) {
    FPair<TF> p;
    __builtin_memcpy(&p, &w, sizeof(p));
    return p;
}

// Native clang vector types: the non-temporal load/store builtins take these
// (HIP's float4/uint4 are structs).
typedef float vf4 __attribute__((ext_vector_type(4)));
typedef uint32_t vu4 __attribute__((ext_vector_type(4)));
typedef uint32_t vu2 __attribute__((ext_vector_type(2)));

// Storage of the factor vectors W, H, u, v: float, or bfloat16 kept as its
// upper 16 bits (bf16 below).  Arithmetic is FP32 either way; per-row sums
// are accumulated in the accumulator type (float, or double for the
// FP64-sums check mode).  Precision codes follow enum ccd_precision.
typedef uint16_t bf16;
constexpr int PREC_F32 = 0, PREC_BF16 = 1, PREC_F64SUM = 2;

template <class TF>
__device__ __forceinline__ float ldf(const void* p, size_t i) {
    if constexpr (sizeof(TF) == 2) return __uint_as_float((uint32_t) static_cast<const bf16*>(p)[i] << 16);
    else return static_cast<const float*>(p)[i];
}

// Round to the nearest bfloat16, ties to even; NaN stays a quiet NaN.
__device__ __forceinline__ bf16 to_bf16(float x) {
    uint32_t b = __float_as_uint(x);
    if ((b & 0x7f800000u) == 0x7f800000u) {
        if (b & 0x007fffffu) b |= 0x00400000u;
        return (bf16)(b >> 16);
    }
    b += 0x7fffu + ((b >> 16) & 1u);
    return (bf16)(b >> 16);
}

template <class TF>
__device__ __forceinline__ void stf(void* p, size_t i, float v) {
    if constexpr (sizeof(TF) == 2) static_cast<bf16*>(p)[i] = to_bf16(v);
    else static_cast<float*>(p)[i] = v;
}

// Wave work-unit kinds (uint4 {x, y, z, w}).
//   GROUP: x = UNIT_GROUP | log2(G) << 8; rows y .. z-1 swept by groups of G
//          lanes, 64/G rows per pass (rows are numbered by falling degree, so
//          a unit is a contiguous row range).
//   SPLIT: x = UNIT_SPLIT | slot << 8; elements [z, w) of row y; the wave's
//          partial (g, h) goes to partials[slot].
constexpr uint32_t UNIT_GROUP = 0;
constexpr uint32_t UNIT_SPLIT = 2;

// One sweep (or residual-only pass) over a compressed layout.
//   residual:  val[e] = (val[e] - xs[idx[e]] * ys[row]) + xa[idx[e]] * ya[row]
//   sweep:     y[row] = sum x[c] r / (lambda |row| + sum x[c]^2), 0 for empty rows
// Factor pointers (x, y, xs, ys, xa, ya) point to float or bf16 storage per
// the precision code; partials hold (g, h) pairs of the accumulator type.
// The sweep and tiled kernels are persistent: a block first copies the hot
// prefixes of its gathered vectors into LDS -- x[0 .. hot_x) and, for fused
// passes, the pair array xp[0 .. hot_p) -- then its waves take work units
// (tiles) in a grid-stride loop.  Indices below the hot count are gathered
// from LDS, the rest from global memory.
//
// Fused passes (residual update + sweep) take the subtracted and added
// factor values of the gathered side from the pair array xp (prev = xs,
// cur = xa); with x_is_cur the sweep's x is the pair's cur as well (the
// first-swept layout, whose x is the current rank's old factor).
constexpr int SWEEP_BLOCK = 1024;

// Byte offset of the staged pair prefix after hot_x factor entries (16-aligned).
template <class TF>
__host__ __device__ inline size_t lds_pair_offset(uint32_t hot_x) {
    return ((size_t)hot_x * sizeof(TF) + 15) & ~(size_t)15;
}

struct SweepArgs {
    const uint32_t* ptr;
    const void* idx;
    float* val;
    const void* x;
    void* y;
    const void* xs;  // residual-only passes: gathered vectors as separate arrays
    const void* ys;
    const void* xa;
    const void* ya;
    const void* xp;  // fused passes: FPair array of the gathered side
    const uint4* units;
    void* partials;
    uint32_t nunits;
    uint32_t hot_x;
    uint32_t hot_p;
    uint32_t x_is_cur;
    float lambda;
};

struct FinalizeArgs {
    const uint32_t* split_rows;
    const uint32_t* split_beg;  // nsplit + 1
    const void* partials;
    const uint32_t* ptr;
    void* y;
    float lambda;
    uint32_t nsplit;
};

// Dynamic LDS of a persistent sweep: hot_x factor entries, then hot_p pairs,
// then 16 bytes that the fused tiled passes read (and drop) for lanes past
// the prefixes.
size_t staged_lds_bytes(int prec, uint32_t hot_x, uint32_t hot_p);

// Optional timing of one launch: events bound to the kernel's own dispatch
// (hipExtLaunchKernel), so timing adds no packets between kernels.  When no
// kernel is launched both events are recorded on the stream instead.
struct LaunchEvents {
    hipEvent_t start = nullptr, stop = nullptr;
};
hipError_t record_unlaunched(const LaunchEvents* ev, hipStream_t s);

// blocks: persistent grid size (a multiple of the CU count)
hipError_t launch_sweep(const SweepArgs& a, int prec, bool idx16, bool sub, bool add, bool sweep, uint32_t blocks,
                        hipStream_t s, const LaunchEvents* ev = nullptr);
hipError_t launch_finalize(const FinalizeArgs& a, int prec, hipStream_t s);
// Blocks of one persistent sweep kernel that fit on a CU by register use
// (the caller sizes the staged LDS prefix to match).
int sweep_blocks_per_cu(int prec, bool idx16, bool sub, bool add, bool sweep);

// Test RMSE: per-block FP64 partial sums of squared errors, then one block
// that reduces them in a fixed order and writes sqrt(sum / nnz) to *out.
constexpr int RMSE_BLOCKS = 1024;
hipError_t launch_rmse(const uint32_t* trow, const uint32_t* tcol, const float* tval, uint64_t nnz,
                       const void* W, const void* H, bool bf16_factors, int64_t m, int64_t n, uint32_t k,
                       double* block_sums, double* out, hipStream_t s);
// Objective: out[0] = loss + lambda * reg, out[1] = loss = sum res^2, out[2]
// = reg = sum_i |row i| ||w_i||^2 + sum_j |col j| ||h_j||^2, all in FP64,
// reduced in a fixed order (block_sums: 2 * OBJ_BLOCKS doubles).  W [k][m],
// H [k][n] FP32; row_ptr / col_ptr give the degrees.
constexpr int OBJ_BLOCKS = 512;
hipError_t launch_objective(const float* res, uint64_t nnz, const uint32_t* row_ptr, const uint32_t* col_ptr,
                            const float* W, const float* H, int64_t m, int64_t n, uint32_t k, double lambda,
                            double* block_sums, double* out, hipStream_t s);
// End of a rank: dstS[i] = srcS[i] (i < nS) and dstF[j] = srcF[j] (j < nF) when
// the sources are given, and the pair arrays of the next rank's fused passes
//   pairS[i] = (prevS[i], nextS[i]),  pairF[j] = (prevF[j], nextF[j])
// where prev is the just-finished rank's value (srcS when given, else prevS);
// a null pair pointer skips that side's pairs.
struct CommitArgs {
    const void* srcS;
    void* dstS;
    const void* prevS;
    const void* nextS;
    void* pairS;
    uint64_t nS;
    const void* srcF;
    void* dstF;
    const void* prevF;
    const void* nextF;
    void* pairF;
    uint64_t nF;
};
hipError_t launch_commit(const CommitArgs& a, bool bf16_factors, hipStream_t s);

// Entry e of a commit: e < nS is entry e of the S side, else entry e - nS
// of the F side.
template <typename TF>
__device__ __forceinline__ void commit_elem(const CommitArgs& a, uint64_t e) {
    const bool sside = e < a.nS;
    const uint64_t i = sside ? e : e - a.nS;
    const TF* src = static_cast<const TF*>(sside ? a.srcS : a.srcF);
    TF* dst = static_cast<TF*>(sside ? a.dstS : a.dstF);
    const TF* prev = static_cast<const TF*>(sside ? a.prevS : a.prevF);
    const TF* next = static_cast<const TF*>(sside ? a.nextS : a.nextF);
    FPair<TF>* pair = static_cast<FPair<TF>*>(sside ? a.pairS : a.pairF);
    const TF nx = pair ? next[i] : TF(0);  // read before dst is written (the same slot when k = 1)
    TF v;
    if (src) {
        v = src[i];
        dst[i] = v;
    } else {
        v = prev[i];
    }
    if (pair) pair[i] = FPair<TF>{v, src && next == dst ? v : nx};
}

// float <-> bfloat16 factor conversion (rounding to nearest even)
hipError_t launch_to_bf16(const float* src, bf16* dst, uint64_t count, hipStream_t s);
hipError_t launch_from_bf16(const bf16* src, float* dst, uint64_t count, hipStream_t s);

hipError_t launch_copy(const float4* src, float4* dst, uint64_t count, hipStream_t s);
// out[r] = elements of row r (of a layout whose rows ascend by index) with
// index below `bound`
hipError_t launch_count_below(const uint32_t* ptr, const void* idx, bool idx16, uint64_t rows, uint32_t bound,
                              uint32_t* out, hipStream_t s);

// ---------------------------------------------------------------- tiled ---
// Tiled layout of one compressed matrix (the CSR5 idea of
// src/detail/kcode/csr5_spmv.cl, re-derived for 64-wide waves): the nnz
// stream is cut into tiles of 64 * sigma elements, one wave per tile; lane l
// owns the sigma consecutive elements l*sigma .. l*sigma+sigma-1 of its tile,
// stored transposed in groups of four (element i of lane l at
// (i/4)*256 + 4*l + i%4) so a 16-byte load of every lane reads 1 KiB of
// consecutive values.
//   lane_flags  bit i: element i of the lane is the first element of a row
//               (16-bit words for sigma <= 16, 32-bit for larger tiles)
//   tile_row0   row of the tile's first element
//   tile_segoff 0xffffffff when the tile's rows are consecutive, else offset
//               into segrows (row of every segment of the tile)
//   tile_open   bit 0: the tile's first row started in an earlier tile;
//               bit 1: its last row continues into the next tile
//   tile_gmax   largest gathered index of the tile (0xffffffff for a last
//               tile with padding and for tiles with empty rows): a sweep
//               whose staged prefix covers it gathers the whole tile from LDS
// Rows that span tiles are finished by tiled_span_kernel from the per-tile
// partial sums: tail[t0] (the row's part in its first tile) plus head[t] for
// every later tile t of the row.
struct TiledArgs {
    const uint32_t* ptr;
    const void* tidx;
    float* tval;
    const void* x;
    void* y;
    const void* xs;
    const void* ys;
    const void* xa;
    const void* ya;
    const void* xp;
    const void* lane_flags;
    const uint32_t* tile_row0;
    const uint32_t* tile_segoff;
    const uint32_t* segrows;
    const uint8_t* tile_open;
    const uint32_t* tile_gmax;
    void* head;  // (g, h) pairs of the accumulator type
    void* tail;
    uint64_t nnz;
    uint32_t ntiles;
    uint32_t hot_x;
    uint32_t hot_p;
    uint32_t x_is_cur;
    float lambda;
    // Split layout (tiles of virtual rows, see build_tiled): a sweep leaves
    // (g, h) of every virtual row in part instead of finishing rows, and the
    // per-row vectors of virtual row v >= nreal are those of row
    // cold_row[v - nreal].  All null / 0 for a plain layout.
    const uint32_t* cold_row;
    uint32_t nreal;
    void* part;  // (g, h) pairs of the accumulator type
    // First tile of a split layout whose elements all belong to second parts
    // (every gather misses the LDS prefix); 0xffffffff: none.
    uint32_t cold_tile0;
    // First tile of the layout's costly tail (second parts of split rows),
    // whose tiles a block spreads over its whole pass; 0xffffffff: none.
    uint32_t spread_tile0;
};

// spans[0, nshort) cover two tiles (t1 = t0 + 1) and are finished one per
// lane; spans[nshort, nspan) cover more and are finished one per wave.
struct SpanArgs {
    const SpanRow* spans;
    const void* head;
    const void* tail;
    void* y;
    uint32_t nspan;
    uint32_t nshort;
    float lambda;
    void* part;  // split layout: (g, h) of each spanning virtual row go here instead of finishing it
};

// Finish of a split layout's sweep, one wave per row: each part's (g, h) is
// the tile kernel's part[] entry when the part lies in one tile, else its
// tail + heads summed as tiled_span_kernel sums them; then
// y[r] = (g1 + g2) / (lambda |row r| + (h1 + h2)), the parts added in a
// fixed order.  Rows without entries keep y.
struct SplitArgs {
    const SplitRow* rows;
    const void* part;
    const void* head;
    const void* tail;
    void* y;
    float lambda;
    uint32_t nreal;
};
hipError_t launch_split_finish(const SplitArgs& a, int prec, hipStream_t s);

// Index formats of the tiled copies, by bytes per index: 2 (u16), 3 (a u16
// plane plus a u8 plane of bits 16..23, for gathered lengths up to 2^24) or
// 4 (u32).  The 3-byte format keeps both planes in one allocation: the u16
// plane of ntiles * 64 * sigma entries, then the u8 plane, where the high
// bytes of lane l of tile t are the sigma bytes at t * 64 * sigma + l * sigma.
struct Idx24 {};

// sigma 8, 16 or 32 for FP32 (32 in 512-thread blocks: at 1024 threads
// its kernels spilled to scratch and the spilled passes gave wrong
// results), 16 for the other precisions; ib: bytes per index
hipError_t launch_tiled(const TiledArgs& a, int sigma, int prec, int ib, bool sub, bool add, bool sweep,
                        uint32_t blocks, hipStream_t s, const LaunchEvents* ev = nullptr);
hipError_t launch_span(const SpanArgs& a, int prec, hipStream_t s);
int tiled_blocks_per_cu(int sigma, int prec, int ib, bool sub, bool add, bool sweep, bool x_is_cur);
// Layout order <-> tiled order (to_tiled = true: src in layout order, dst
// tiled).  map (optional, a split layout's vpos): layout position of each
// element in tiled order; without it the two orders are the same.
hipError_t launch_retile_f32(const float* src, float* dst, uint64_t nnz, uint32_t ntiles, int sigma, bool to_tiled,
                             hipStream_t s, const uint32_t* map = nullptr);
// u16 or u32 indices in, the tiled copy in the format of ib out
hipError_t launch_retile_idx(const void* src, bool src16, void* dst, int ib, uint64_t nnz, uint32_t ntiles, int sigma,
                             const uint32_t* map, hipStream_t s);
// gmax[t] = largest of the indices of tile t (tiled order), 0xffffffff for a
// tile that extends past nnz or holds empty rows (segoff[t] other than
// 0xffffffff)
hipError_t launch_tile_gmax(const void* idx, bool idx16, const uint32_t* segoff, uint32_t* gmax, uint64_t nnz,
                            uint32_t ntiles, int sigma, const uint32_t* map, hipStream_t s);
hipError_t launch_fill(float* dst, float value, uint64_t count, hipStream_t s);
hipError_t launch_fill_bytes16(uint16_t* dst, uint16_t value, uint64_t count, hipStream_t s);

// --------------------------------------------------------------- batched ---
// Many independent FP32 factorizations in one persistent launch.  Each
// problem owns a team of consecutive (XCD-local where possible) blocks that
// runs its whole CCD++ loop -- the same tiled passes, span finishes and
// commits as the single-problem path, in the same order -- separated by
// team barriers (a per-problem counter in global memory).  Problems never
// wait for each other.  The grid never exceeds what is co-resident on the
// device, and every barrier wait is bounded in time: a wait that runs out
// sets *status and the team leaves the kernel.
constexpr int BATCH_BLOCK = 256;

struct BatchLayout {  // one tiled layout of one problem (see TiledArgs)
    const uint32_t* ptr;
    const void* tidx;
    float* tval;
    const void* lane_flags;
    const uint32_t* tile_row0;
    const uint32_t* tile_segoff;
    const uint32_t* segrows;
    const uint8_t* tile_open;
    void* head;
    void* tail;
    const SpanRow* spans;
    void* pair;  // FPair<float>[gathered length]
    uint64_t nnz;
    uint32_t ntiles;
    uint32_t nspan;
    uint32_t nshort;  // spans over two tiles, first in the span table
    uint32_t ib;      // bytes per index of the tiled copy: 2, 3 or 4
    uint32_t rows;
    uint32_t gdim;    // length of the gathered vector
    uint32_t stage;   // 1: each block stages the gathered vector's hot prefix in LDS per pass
};

struct BatchProblem {
    BatchLayout F, S;  // first- and second-swept layouts
    float* XF;         // [k][F.rows] factor produced by the F sweeps
    float* XS;         // [k][S.rows]
    float* vF;         // work vectors
    float* vS;
    uint32_t k, T, Tin;
    uint32_t continued;  // residual already holds R - W H^T: add back from the first rank on
    uint32_t first, nblocks;  // team: virtual block ids [first, first + nblocks)
    uint32_t trial_reps;      // autotuner trial: plain sweep pairs of rank 0 to run
    float lambda;
};

struct BatchArgs {
    const BatchProblem* probs;
    const uint32_t* block_problem;  // virtual block id -> problem (0xffffffff: idle)
    uint32_t nprob;
    uint32_t nblocks;               // grid size (a multiple of 8)
    uint32_t* barriers;             // one counter per problem, zero at launch
    uint32_t* xcd_tickets;          // 8 counters, 64 bytes apart, zero at launch
    int* status;                    // 0; 1 after a barrier wait ran out
    uint64_t wait_ticks;            // bound of one barrier wait in wall_clock64() ticks
    uint32_t trial;                 // autotuner trial: trial_reps plain sweep pairs per problem, nothing else
    uint32_t lds_floats;            // dynamic LDS of a block, in 4-byte entries (staged hot prefixes)
    uint32_t deal_min;              // dealt sweeps: layouts with at least this many tiles per wave of the team
};

// sigma 8 or 16; dynamic LDS of a.lds_floats entries per block
hipError_t launch_batch(const BatchArgs& a, int sigma, hipStream_t s);
// co-resident blocks per CU with lds_bytes of dynamic LDS each (0 on failure)
int batch_blocks_per_cu(int sigma, size_t lds_bytes);

// ---------------------------------------------------------------- fold-in ---
// Rows (CSR, the new users' ratings) folded in against fixed item factors:
// H [k][n] rank-major for the sweeps, HT [n][k] for the features.
constexpr int FOLD_KMAX = 128;
struct FoldArgs {
    const uint32_t* ptr;  // rows + 1
    const uint32_t* idx;
    const float* val;
    const float* H;
    const float* HT;
    uint32_t rows;
    uint32_t k;
    uint32_t n;
    int sweeps;
    float lambda;
};
// Rows longer than FOLD_SHORT ratings are listed in long_rows[0 .. nlong)
// and get a block each.
constexpr uint32_t FOLD_SHORT = 1024;
// F [rows][2k + 2]: rating-weighted mean and variance of the rated items' H
// columns, log(1 + degree), mean rating (FP64 sums)
hipError_t launch_fold_features(const FoldArgs& a, const uint32_t* long_rows, uint32_t nlong, float* F,
                                hipStream_t s);
// W [rows][k] after a.sweeps sweeps from W0 ([rows][k], null = zeros); rbuf:
// nnz floats of scratch (used by rows above 16384 ratings)
hipError_t launch_fold_sweeps(const FoldArgs& a, const uint32_t* long_rows, uint32_t nlong, const float* W0, float* W,
                              float* rbuf, hipStream_t s);
hipError_t launch_transpose(const float* src, float* dst, uint32_t rows, uint32_t cols, hipStream_t s);

// -------------------------------------------------------------------- MLP ---
// C(i, j) = sum_p A(i, p) B(p, j) for i < M, j < N, p < K with
// A(i, p) = A[i sai + p sap] (FP32), B(p, j) = B[p sbp + j sbj] (FP32 or
// bfloat16); epilogues: GEMM_STORE C = acc (+ bias[j]); GEMM_BIAS_GELU
// C = acc + bias[j], C2 = GELU(C); GEMM_GELU_GRAD C = acc * GELU'(aux)
// (aux, C, C2 row-major with leading dimension ldc).
constexpr int GEMM_STORE = 0, GEMM_BIAS_GELU = 1, GEMM_GELU_GRAD = 2;
struct GemmArgs {
    const float* A;
    uint64_t sai, sap, sbp, sbj;
    uint32_t M, N, K, ldc;
    const void* bias;  // same element type as B
    float* C;
    float* C2;
    const float* aux;
    float* part;       // optional workspace of part_cap floats: a plain product (GEMM_STORE, no
    uint64_t part_cap; // bias) with few output tiles splits K over slices, summed in slice order
};
hipError_t launch_gemm(const GemmArgs& g, const void* B, bool bf16_b, int epi, hipStream_t s);

// Whole-network inference on the bfloat16 matrix cores (mlp_mfma.hip):
// Y [rows][out] = out_map(W3 gelu(W2 gelu(W1 in_map(X) + b1) + b2) + b3),
// hidden width MLP_MFMA_HIDDEN, inputs and activations rounded to bfloat16,
// FP32 accumulation.  Weights packed by launch_mlp_pack: W1 [hidden][kin],
// W2 [hidden][hidden], W3 [nout][hidden] bfloat16 with kin = in rounded up
// to 32 and nout = out rounded up to 16 (zero padding), biases FP32 padded
// likewise.  in_map: x -> (x - in_sub) in_mul, out_map: y -> (y - out_sub)
// out_mul, identity where the pointers are null.
constexpr uint32_t MLP_MFMA_HIDDEN = 256;
constexpr uint32_t MLP_MFMA_MAX_KIN = 288;
struct MlpInferArgs {
    const float* X;
    float* Y;
    uint32_t rows, in, out, kin, nout;
    const uint16_t *W1, *W2, *W3;
    const float *b1, *b2, *b3;
    const float *in_sub, *in_mul, *out_sub, *out_mul;
};
hipError_t launch_mlp_infer(const MlpInferArgs& g, hipStream_t s);
hipError_t launch_mlp_pack(const float* P, uint32_t in, uint32_t out, uint32_t kin, uint32_t nout, uint16_t* W1,
                           uint16_t* W2, uint16_t* W3, float* b1, float* b2, float* b3, hipStream_t s);
hipError_t launch_colsum(const float* dZ, uint32_t M, uint32_t N, float* out, hipStream_t s);
// dY = 2 (Y - T) / count; *out = mean (Y - T)^2 (FP64, fixed order); part: MSE_PARTS doubles
constexpr uint32_t MSE_PARTS = 1024;
hipError_t launch_mse(const float* Y, const float* T, uint64_t count, float* dY, double* part, double* out,
                      hipStream_t s);
struct AdamArgs {
    float* param;
    const float* grad;
    float* m;
    float* v;
    uint64_t count;
    float lr, beta1, beta2, eps, bc1, bc2;  // bc = 1 - beta^step
};
hipError_t launch_adam(const AdamArgs& a, hipStream_t s);
// Y[i][j] = (X[r][j] - sub[j]) * mul[j], r = rows ? rows[i] : i (sub null: plain copy / gather)
hipError_t launch_affine_rows(const float* X, const uint32_t* rows, uint32_t M, uint32_t N, const float* sub,
                              const float* mul, float* Y, hipStream_t s);

}  // namespace ccd
