// Internal interface between the C ABI (ccd_abi.hip) and the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ccd {

// Native clang vector types: the non-temporal load/store builtins take these
// (HIP's float4/uint4 are structs).
typedef float vf4 __attribute__((ext_vector_type(4)));
typedef uint32_t vu4 __attribute__((ext_vector_type(4)));
typedef uint32_t vu2 __attribute__((ext_vector_type(2)));

// Wave work-unit kinds (uint4 {x, y, z, w}).
//   GROUP: x = UNIT_GROUP | log2(G) << 8; rows rowlist[y .. z) swept by groups
//          of G lanes, 64/G rows per pass.
//   SPLIT: x = UNIT_SPLIT | slot << 8; elements [z, w) of row y; the wave's
//          partial (g, h) goes to partials[slot].
constexpr uint32_t UNIT_GROUP = 0;
constexpr uint32_t UNIT_SPLIT = 2;

// One sweep (or residual-only pass) over a compressed layout.
//   residual:  val[e] = (val[e] - xs[idx[e]] * ys[row]) + xa[idx[e]] * ya[row]
//   sweep:     y[row] = sum x[c] r / (lambda |row| + sum x[c]^2), 0 for empty rows
struct SweepArgs {
    const uint32_t* ptr;
    const void* idx;
    float* val;
    const float* x;
    float* y;
    const float* xs;
    const float* ys;
    const float* xa;
    const float* ya;
    const uint4* units;
    const uint32_t* rowlist;
    float2* partials;
    uint32_t nunits;
    float lambda;
};

struct FinalizeArgs {
    const uint32_t* split_rows;
    const uint32_t* split_beg;  // nsplit + 1
    const float2* partials;
    const uint32_t* ptr;
    float* y;
    float lambda;
    uint32_t nsplit;
};

hipError_t launch_sweep(const SweepArgs& a, bool idx16, bool sub, bool add, bool sweep, hipStream_t s);
hipError_t launch_finalize(const FinalizeArgs& a, hipStream_t s);

// Test RMSE: per-block FP64 partial sums of squared errors, then one block
// that reduces them in a fixed order and writes sqrt(sum / nnz) to *out.
constexpr int RMSE_BLOCKS = 1024;
hipError_t launch_rmse(const uint32_t* trow, const uint32_t* tcol, const float* tval, uint64_t nnz,
                       const float* W, const float* H, int64_t m, int64_t n, uint32_t k, double* block_sums,
                       double* out, hipStream_t s);

hipError_t launch_copy(const float4* src, float4* dst, uint64_t count, hipStream_t s);

// ---------------------------------------------------------------- tiled ---
// Tiled layout of one compressed matrix (the CSR5 idea of
// src/detail/kcode/csr5_spmv.cl, re-derived for 64-wide waves): the nnz
// stream is cut into tiles of 64 * sigma elements, one wave per tile; lane l
// owns the sigma consecutive elements l*sigma .. l*sigma+sigma-1 of its tile,
// stored transposed in groups of four (element i of lane l at
// (i/4)*256 + 4*l + i%4) so a 16-byte load of every lane reads 1 KiB of
// consecutive values.
//   lane_flags  bit i: element i of the lane is the first element of a row
//   lane_seg0   segment (row of the tile, counted from 0) of the lane's element 0
//   tile_row0   row of the tile's first element
//   tile_segoff 0xffffffff when the tile's rows are consecutive, else offset
//               into segrows (row of every segment of the tile)
//   tile_open   bit 0: the tile's first row started in an earlier tile;
//               bit 1: its last row continues into the next tile
// Rows that span tiles are finished by tiled_span_kernel from the per-tile
// partial sums: tail[t0] (the row's part in its first tile) plus head[t] for
// every later tile t of the row.
struct TiledArgs {
    const uint32_t* ptr;
    const void* tidx;
    float* tval;
    const float* x;
    float* y;
    const float* xs;
    const float* ys;
    const float* xa;
    const float* ya;
    const uint32_t* lane_flags;
    const uint16_t* lane_seg0;
    const uint32_t* tile_row0;
    const uint32_t* tile_segoff;
    const uint32_t* segrows;
    const uint8_t* tile_open;
    float2* head;
    float2* tail;
    uint64_t nnz;
    uint32_t ntiles;
    float lambda;
};

struct SpanArgs {
    const uint32_t* span_rows;
    const uint32_t* ptr;
    const float2* head;
    const float2* tail;
    float* y;
    uint32_t nspan;
    uint32_t tile_elems;
    float lambda;
};

hipError_t launch_tiled(const TiledArgs& a, int sigma, bool idx16, bool sub, bool add, bool sweep, hipStream_t s);
hipError_t launch_span(const SpanArgs& a, hipStream_t s);
// original order <-> tiled order (to_tiled = true: src original, dst tiled)
hipError_t launch_retile_f32(const float* src, float* dst, uint64_t nnz, uint32_t ntiles, int sigma, bool to_tiled,
                             hipStream_t s);
hipError_t launch_retile_idx(const void* src, void* dst, bool idx16, uint64_t nnz, uint32_t ntiles, int sigma,
                             hipStream_t s);
hipError_t launch_fill(float* dst, float value, uint64_t count, hipStream_t s);

}  // namespace ccd
