// gfx950 kernels of the CCD++ rank-one loop.
//
// The sweep kernel replaces RankOneUpdate_DUAL_kernel_{u,v} and the four
// UpdateRating_DUAL_kernel_NoLoss_* kernels (kcode/ccd_1.cl:37-189,
// kcode/ccd_2.cl:1-175): one pass over a compressed layout streams the index
// and value arrays once, optionally updates the residual in place
// (val = (val - xs*ys) + xa*ya, the subtract of the previous rank fused with
// the add-back of the current one, cdmf_ref.cpp:86-89 / 120-121) and computes
// the closed-form update y_r = g / (lambda |r| + h) with
// g = sum x_c r_rc, h = sum x_c^2 (cdmf_ref.cpp:7-17).
//
// Work is cut into wave-sized units planned on the host by row degree:
//   * groups of G = 1..64 lanes per row (lane-per-row for the shortest rows,
//     wave-per-row for long ones), 64/G rows per pass;
//   * rows longer than a threshold are split into pieces, each piece one
//     wave; the per-piece (g, h) partials are reduced in piece order by
//     finalize_kernel, so results are run-to-run deterministic.
// Loads of a group are unit-stride across its lanes, so a wave-per-row pass
// reads 256 B per instruction per stream.  The file is built with
// -ffp-contract=off: the residual update rounds the product and the sum
// separately, exactly like the CPU reference.
#include "ccd_kernels.h"

#include <math.h>

namespace ccd {
namespace {

constexpr int WAVE = 64;
constexpr int BLOCK = 256;
constexpr int WAVES_PER_BLOCK = BLOCK / WAVE;

template <typename IdxT, bool SUB, bool ADD, bool SWEEP>
__device__ __forceinline__ void strided_range(const SweepArgs& a, const IdxT* __restrict__ idx, uint32_t p,
                                              const uint32_t end, const uint32_t step, const float ysr,
                                              const float yar, float& g, float& h) {
    float* __restrict__ val = a.val;
    const float* __restrict__ x = a.x;
    const float* __restrict__ xs = a.xs;
    const float* __restrict__ xa = a.xa;
    // four independent elements per lane in flight
    for (; p + 3 * step < end; p += 4 * step) {
        uint32_t c[4];
        float r[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) c[q] = __builtin_nontemporal_load(idx + p + q * step);
#pragma unroll
        for (int q = 0; q < 4; ++q) r[q] = __builtin_nontemporal_load(val + p + q * step);
        if (SUB) {
#pragma unroll
            for (int q = 0; q < 4; ++q) r[q] = r[q] - xs[c[q]] * ysr;
        }
        if (ADD) {
#pragma unroll
            for (int q = 0; q < 4; ++q) r[q] = r[q] + xa[c[q]] * yar;
        }
        if (SUB || ADD) {
#pragma unroll
            for (int q = 0; q < 4; ++q) __builtin_nontemporal_store(r[q], val + p + q * step);
        }
        if (SWEEP) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float xv = x[c[q]];
                g += xv * r[q];
                h += xv * xv;
            }
        }
    }
    for (; p < end; p += step) {
        const uint32_t c = __builtin_nontemporal_load(idx + p);
        float r = __builtin_nontemporal_load(val + p);
        if (SUB) r = r - xs[c] * ysr;
        if (ADD) r = r + xa[c] * yar;
        if (SUB || ADD) __builtin_nontemporal_store(r, val + p);
        if (SWEEP) {
            const float xv = x[c];
            g += xv * r;
            h += xv * xv;
        }
    }
}

// Four consecutive elements starting at 4*q (16-byte aligned): one float4
// value load, one uint4 (or uint2 for 16-bit) index load.
template <typename IdxT, bool SUB, bool ADD, bool SWEEP>
__device__ __forceinline__ void group4(const SweepArgs& a, const IdxT* __restrict__ idx, uint32_t q, float ysr,
                                       float yar, float& g, float& h) {
    vf4* __restrict__ v4 = reinterpret_cast<vf4*>(a.val) + q;
    const vf4 rv = __builtin_nontemporal_load(v4);
    uint32_t c[4];
    if constexpr (sizeof(IdxT) == 2) {
        const vu2 cv = __builtin_nontemporal_load(reinterpret_cast<const vu2*>(idx) + q);
        c[0] = cv.x & 0xffffu;
        c[1] = cv.x >> 16;
        c[2] = cv.y & 0xffffu;
        c[3] = cv.y >> 16;
    } else {
        const vu4 cv = __builtin_nontemporal_load(reinterpret_cast<const vu4*>(idx) + q);
        c[0] = cv.x;
        c[1] = cv.y;
        c[2] = cv.z;
        c[3] = cv.w;
    }
    float r[4] = {rv.x, rv.y, rv.z, rv.w};
    if (SUB) {
#pragma unroll
        for (int j = 0; j < 4; ++j) r[j] = r[j] - a.xs[c[j]] * ysr;
    }
    if (ADD) {
#pragma unroll
        for (int j = 0; j < 4; ++j) r[j] = r[j] + a.xa[c[j]] * yar;
    }
    if (SUB || ADD) __builtin_nontemporal_store(vf4{r[0], r[1], r[2], r[3]}, v4);
    if (SWEEP) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float xv = a.x[c[j]];
            g += xv * r[j];
            h += xv * xv;
        }
    }
}

// Elements [s, e) by a whole wave: an unaligned head and tail of at most
// three elements each, then the aligned body as 16-byte groups, two groups
// (eight elements) per lane in flight.
template <typename IdxT, bool SUB, bool ADD, bool SWEEP>
__device__ __forceinline__ void wave_range(const SweepArgs& a, const IdxT* __restrict__ idx, uint32_t s, uint32_t e,
                                           uint32_t lane, float ysr, float yar, float& g, float& h) {
    uint32_t b0 = (s + 3u) & ~3u;
    if (b0 > e) b0 = e;
    uint32_t b1 = e & ~3u;
    if (b1 < b0) b1 = b0;
    if (lane < b0 - s) strided_range<IdxT, SUB, ADD, SWEEP>(a, idx, s + lane, s + lane + 1, 1, ysr, yar, g, h);
    if (lane < e - b1) strided_range<IdxT, SUB, ADD, SWEEP>(a, idx, b1 + lane, b1 + lane + 1, 1, ysr, yar, g, h);
    const uint32_t q1 = b1 >> 2;
    uint32_t q = (b0 >> 2) + lane;
    for (; q + WAVE < q1; q += 2 * WAVE) {
        group4<IdxT, SUB, ADD, SWEEP>(a, idx, q, ysr, yar, g, h);
        group4<IdxT, SUB, ADD, SWEEP>(a, idx, q + WAVE, ysr, yar, g, h);
    }
    if (q < q1) group4<IdxT, SUB, ADD, SWEEP>(a, idx, q, ysr, yar, g, h);
}

template <typename IdxT, bool SUB, bool ADD, bool SWEEP>
__global__ __launch_bounds__(BLOCK) void sweep_kernel(SweepArgs a) {
    const int lane = threadIdx.x & (WAVE - 1);
    const uint32_t u = blockIdx.x * WAVES_PER_BLOCK + (threadIdx.x >> 6);
    if (u >= a.nunits) return;
    const uint4 unit = a.units[u];
    const IdxT* __restrict__ idx = static_cast<const IdxT*>(a.idx);
    const uint32_t kind = unit.x & 0xffu;

    if (kind == UNIT_SPLIT) {
        const uint32_t row = unit.y;
        const float ysr = SUB ? a.ys[row] : 0.0f;
        const float yar = ADD ? a.ya[row] : 0.0f;
        float g = 0.0f, h = 0.0f;
        wave_range<IdxT, SUB, ADD, SWEEP>(a, idx, unit.z, unit.w, (uint32_t)lane, ysr, yar, g, h);
        if (SWEEP) {
#pragma unroll
            for (int off = WAVE / 2; off > 0; off >>= 1) {
                g += __shfl_xor(g, off);
                h += __shfl_xor(h, off);
            }
            if (lane == 0) a.partials[unit.x >> 8] = make_float2(g, h);
        }
        return;
    }

    const int lg = (int)((unit.x >> 8) & 7u);
    const uint32_t G = 1u << lg;
    const uint32_t sub = (uint32_t)lane >> lg;
    const uint32_t sl = (uint32_t)lane & (G - 1u);
    const uint32_t rows_per_pass = (uint32_t)WAVE >> lg;
    for (uint32_t base = unit.y; base < unit.z; base += rows_per_pass) {
        const uint32_t ri = base + sub;
        const bool valid = ri < unit.z;
        const uint32_t row = valid ? a.rowlist[ri] : 0u;
        const uint32_t s = valid ? a.ptr[row] : 0u;
        const uint32_t e = valid ? a.ptr[row + 1] : 0u;
        const float ysr = (SUB && valid) ? a.ys[row] : 0.0f;
        const float yar = (ADD && valid) ? a.ya[row] : 0.0f;
        float g = 0.0f, h = 0.0f;
        if (lg == 6) wave_range<IdxT, SUB, ADD, SWEEP>(a, idx, s, e, (uint32_t)lane, ysr, yar, g, h);
        else strided_range<IdxT, SUB, ADD, SWEEP>(a, idx, s + sl, e, G, ysr, yar, g, h);
        if (SWEEP) {
            for (uint32_t off = G >> 1; off > 0; off >>= 1) {
                g += __shfl_xor(g, (int)off);
                h += __shfl_xor(h, (int)off);
            }
            if (valid && sl == 0) a.y[row] = (e > s) ? g / (a.lambda * (float)(e - s) + h) : 0.0f;
        }
    }
}

__global__ __launch_bounds__(BLOCK) void finalize_kernel(FinalizeArgs a) {
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= a.nsplit) return;
    const uint32_t row = a.split_rows[i];
    float g = 0.0f, h = 0.0f;
    for (uint32_t q = a.split_beg[i]; q < a.split_beg[i + 1]; ++q) {
        const float2 pq = a.partials[q];
        g += pq.x;
        h += pq.y;
    }
    const uint32_t d = a.ptr[row + 1] - a.ptr[row];
    a.y[row] = g / (a.lambda * (float)d + h);
}

// Per test entry: pred = sum_t W[t][i] * H[t][j] with each product rounded
// to float and accumulated in double (tools.cpp:556-565); NaN squared errors
// are skipped but counted (tools.cpp:566-573).
__global__ __launch_bounds__(BLOCK) void rmse_partial_kernel(const uint32_t* __restrict__ trow,
                                                             const uint32_t* __restrict__ tcol,
                                                             const float* __restrict__ tval, uint64_t nnz,
                                                             const float* __restrict__ W,
                                                             const float* __restrict__ H, int64_t m, int64_t n,
                                                             uint32_t k, double* __restrict__ block_sums) {
    double acc = 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
    for (uint64_t e = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; e < nnz; e += stride) {
        const uint32_t i = trow[e], j = tcol[e];
        double pred = 0.0;
        for (uint32_t t = 0; t < k; ++t) {
            const float p = W[(size_t)t * m + i] * H[(size_t)t * n + j];
            pred += (double)p;
        }
        const double d = pred - (double)tval[e];
        const double sq = d * d;
        if (!isnan(sq)) acc += sq;
    }
    __shared__ double sh[BLOCK];
    sh[threadIdx.x] = acc;
    __syncthreads();
#pragma unroll
    for (int s = BLOCK / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) block_sums[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(BLOCK) void rmse_final_kernel(const double* __restrict__ block_sums, int nb,
                                                           uint64_t nnz, double* __restrict__ out) {
    __shared__ double sh[BLOCK];
    double acc = 0.0;
    for (int b = threadIdx.x; b < nb; b += BLOCK) acc += block_sums[b];
    sh[threadIdx.x] = acc;
    __syncthreads();
#pragma unroll
    for (int s = BLOCK / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = nnz ? sqrt(sh[0] / (double)nnz) : (double)NAN;
}

__global__ __launch_bounds__(BLOCK) void copy_kernel(const float4* __restrict__ src, float4* __restrict__ dst,
                                                     uint64_t count) {
    const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
    for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < count; i += stride) dst[i] = src[i];
}

template <typename IdxT, bool SUB, bool ADD, bool SWEEP>
hipError_t launch_sweep_t(const SweepArgs& a, hipStream_t s) {
    const uint32_t blocks = (a.nunits + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL((sweep_kernel<IdxT, SUB, ADD, SWEEP>), dim3(blocks), dim3(BLOCK), 0, s, a);
    return hipGetLastError();
}

template <typename IdxT>
hipError_t launch_sweep_idx(const SweepArgs& a, bool sub, bool add, bool sweep, hipStream_t s) {
    const int code = (sub ? 4 : 0) | (add ? 2 : 0) | (sweep ? 1 : 0);
    switch (code) {
        case 1: return launch_sweep_t<IdxT, false, false, true>(a, s);
        case 2: return launch_sweep_t<IdxT, false, true, false>(a, s);
        case 3: return launch_sweep_t<IdxT, false, true, true>(a, s);
        case 4: return launch_sweep_t<IdxT, true, false, false>(a, s);
        case 5: return launch_sweep_t<IdxT, true, false, true>(a, s);
        case 6: return launch_sweep_t<IdxT, true, true, false>(a, s);
        case 7: return launch_sweep_t<IdxT, true, true, true>(a, s);
        default: return hipSuccess;  // nothing to do
    }
}

}  // namespace

hipError_t launch_sweep(const SweepArgs& a, bool idx16, bool sub, bool add, bool sweep, hipStream_t s) {
    return idx16 ? launch_sweep_idx<uint16_t>(a, sub, add, sweep, s)
                 : launch_sweep_idx<uint32_t>(a, sub, add, sweep, s);
}

hipError_t launch_finalize(const FinalizeArgs& a, hipStream_t s) {
    if (a.nsplit == 0) return hipSuccess;
    const uint32_t blocks = (a.nsplit + BLOCK - 1) / BLOCK;
    hipLaunchKernelGGL(finalize_kernel, dim3(blocks), dim3(BLOCK), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_rmse(const uint32_t* trow, const uint32_t* tcol, const float* tval, uint64_t nnz,
                       const float* W, const float* H, int64_t m, int64_t n, uint32_t k, double* block_sums,
                       double* out, hipStream_t s) {
    hipLaunchKernelGGL(rmse_partial_kernel, dim3(RMSE_BLOCKS), dim3(BLOCK), 0, s, trow, tcol, tval, nnz, W, H, m,
                       n, k, block_sums);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(rmse_final_kernel, dim3(1), dim3(BLOCK), 0, s, block_sums, RMSE_BLOCKS, nnz, out);
    return hipGetLastError();
}

hipError_t launch_copy(const float4* src, float4* dst, uint64_t count, hipStream_t s) {
    uint64_t blocks = (count + BLOCK - 1) / BLOCK;
    if (blocks > 8192) blocks = 8192;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(copy_kernel, dim3((uint32_t)blocks), dim3(BLOCK), 0, s, src, dst, count);
    return hipGetLastError();
}

}  // namespace ccd
