// Internal interface between the C ABI (ccd_abi.hip) and the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ccd {

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

}  // namespace ccd
