// Fold-in of new rows against fixed item factors H (the learned warm-start
// extension; the reference has no fold-in).  Folding in a row is CCD++ with
// the item side frozen: per sweep and rank t,
//   a = r^ + w_t h_t,  w_t = sum a h_t / (lambda |row| + sum h_t^2),  r^ = a - w_t h_t
// (the u-side update of src/cdmf_ref.cpp:7-17 on one row).
//
// The sweeps take one launch of 1024-thread blocks for both row classes:
// blocks below nlong take one long row each (more than FOLD_SHORT ratings,
// listed by the host) with all 16 waves; the other blocks take 16 short
// rows, one per wave.  The features take a launch per class.  Residual, indices and gathered h values stay in registers for the
// whole solve (rows above 16384 ratings keep the rest in a global buffer).
// Per-row sums are reduced in a fixed order: deterministic.
#include "ccd_kernels.h"

namespace ccd {
namespace {

constexpr int WAVE = 64;
constexpr int BLOCK = 256;
constexpr int LBLOCK = 1024;
constexpr int LWAVES = LBLOCK / WAVE;
constexpr int FOLD_E = 16;   // elements per lane of a short row (1024 per row)
constexpr int LONG_E = 16;   // elements per thread of a long row (16384 per row)
constexpr int FEAT_U = 32;   // H-row loads in flight per feature step
static_assert(WAVE * FOLD_E == (int)FOLD_SHORT, "short-row bound of the host's long-row list");

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Feature sums of a row's elements s .. e - 1 for rank t, over the chunks
// of 64 elements c0, c0 + cstep, ...: S = sum r, m = sum r h, v = sum r h^2
// (FP64, in element order within each chunk, chunks in order).  A chunk's
// ratings and indices come in with one coalesced, non-temporal load each
// (so that they do not push the item rows of HT, reread for every rating
// of the item, out of L2); each element's index and rating are then read
// from its lane (v_readlane) and the H rows of 32 elements are loaded
// together.  Lanes past k load rank 0's entry and drop it, so there is no
// branch around the loads.
__device__ __forceinline__ void feature_sums(const FoldArgs& a, uint32_t s, uint32_t e, uint32_t c0, uint32_t cstep,
                                             uint32_t t, bool tok, double& S, double& m, double& v) {
    const uint32_t k = a.k;
    const uint32_t tt = tok ? t : 0u;
    const int lane = threadIdx.x & (WAVE - 1);
    for (uint32_t base = s + c0 * WAVE; base < e; base += cstep * WAVE) {
        const uint32_t p = base + (uint32_t)lane;
        const bool okp = p < e;
        const uint32_t cv = okp ? __builtin_nontemporal_load(a.idx + p) : 0u;
        const float rv = okp ? __builtin_nontemporal_load(a.val + p) : 0.0f;
        const uint32_t n = e - base < (uint32_t)WAVE ? e - base : (uint32_t)WAVE;
#pragma unroll
        for (int half = 0; half < WAVE / FEAT_U; ++half) {
            if ((uint32_t)(FEAT_U * half) >= n) break;
            float h[FEAT_U];
#pragma unroll
            for (int u = 0; u < FEAT_U; ++u) {
                const uint32_t c = (uint32_t)__builtin_amdgcn_readlane((int)cv, FEAT_U * half + u);
                h[u] = a.HT[(size_t)c * k + tt];
            }
#pragma unroll
            for (int u = 0; u < FEAT_U; ++u) {
                if ((uint32_t)(FEAT_U * half + u) < n) {
                    const float r = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rv), FEAT_U * half + u));
                    const float hu = tok ? h[u] : 0.0f;
                    S += r;
                    m += (double)r * hu;
                    v += (double)r * hu * hu;
                }
            }
        }
    }
}

// mean = m / S and variance = v / S - mean^2 (FP64, one pass), 0 for an
// empty row or zero rating sum
__device__ __forceinline__ void write_features(float* f, uint32_t k, uint32_t t, bool tok, uint32_t d, double S,
                                               double m, double v) {
    if (!tok) return;
    double mean = 0.0, var = 0.0;
    if (d && S != 0.0) {
        mean = m / S;
        var = v / S - mean * mean;
        if (var < 0.0) var = 0.0;
    }
    f[t] = (float)mean;
    f[k + t] = (float)var;
}

// F [rows][2k + 2] of the long rows (more than FOLD_SHORT ratings), one
// 1024-thread block each: lanes take ranks t0 + lane, the row's elements are
// dealt to the 16 waves and the per-wave sums added in wave order.
__global__ __launch_bounds__(LBLOCK) void fold_features_long_kernel(FoldArgs a, const uint32_t* __restrict__ long_rows,
                                                                    float* __restrict__ F) {
    __shared__ double part[3][LWAVES][WAVE];
    const int lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
    const uint32_t k = a.k;
    const uint32_t row = long_rows[blockIdx.x];
    const uint32_t s = a.ptr[row], e = a.ptr[row + 1], d = e - s;
    float* f = F + (size_t)row * (2 * k + 2);
    double Sr = 0.0;
    for (uint32_t t0 = 0; t0 < k; t0 += WAVE) {
        const uint32_t t = t0 + lane;
        const bool tok = t < k;
        double S = 0.0, m = 0.0, v = 0.0;
        feature_sums(a, s, e, (uint32_t)wv, LWAVES, t, tok, S, m, v);  // wave wv: chunks wv, wv + 16, ...
        part[0][wv][lane] = S;
        part[1][wv][lane] = m;
        part[2][wv][lane] = v;
        __syncthreads();
        if (wv == 0) {
            double St = 0.0, mt = 0.0, vt = 0.0;
            for (int w = 0; w < LWAVES; ++w) {
                St += part[0][w][lane];
                mt += part[1][w][lane];
                vt += part[2][w][lane];
            }
            write_features(f, k, t, tok, d, St, mt, vt);
            Sr = St;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        f[2 * k] = (float)log1p((double)d);
        f[2 * k + 1] = d ? (float)(Sr / d) : 0.0f;
    }
}

// F of the other rows, one single-wave block per row (so a wave that
// finishes a short row frees its slot at once, whatever the lengths of
// its neighbours): lanes take ranks t0 + lane and walk the row's elements
// in order, FEAT_U loads in flight (feature_sums).
__global__ __launch_bounds__(WAVE) void fold_features_short_kernel(FoldArgs a, float* __restrict__ F) {
    const int lane = threadIdx.x;
    const uint32_t k = a.k;
    const uint32_t row = blockIdx.x;
    const uint32_t s = a.ptr[row], e = a.ptr[row + 1], d = e - s;
    if (d > FOLD_SHORT) return;
    float* f = F + (size_t)row * (2 * k + 2);
    double Sr = 0.0;
    for (uint32_t t0 = 0; t0 < k; t0 += WAVE) {
        const uint32_t t = t0 + lane;
        const bool tok = t < k;
        double S = 0.0, m = 0.0, v = 0.0;
        feature_sums(a, s, e, 0, 1, t, tok, S, m, v);
        write_features(f, k, t, tok, d, S, m, v);
        Sr = S;
    }
    if (lane == 0) {
        f[2 * k] = (float)log1p((double)d);
        f[2 * k + 1] = d ? (float)(Sr / d) : 0.0f;
    }
}

// One short row on one wave.  w_t lives in lane t % 64, slot t / 64, and is
// broadcast with a shuffle (t uniform, whole wave active).
__device__ __forceinline__ void sweep_short_row(const FoldArgs& a, uint32_t row, int lane, const float* __restrict__ W0,
                                                float* __restrict__ W) {
    constexpr int NS = FOLD_KMAX / WAVE;
    const uint32_t s = a.ptr[row], e = a.ptr[row + 1], d = e - s;
    const uint32_t k = a.k;
    const float* __restrict__ H = a.H;
    float wr[NS];
#pragma unroll
    for (int j = 0; j < NS; ++j) {
        const uint32_t t = lane + WAVE * j;
        wr[j] = (W0 && t < k) ? W0[(size_t)row * k + t] : 0.0f;
    }
    auto wget = [&](uint32_t t) -> float {
        float v = 0.0f;
#pragma unroll
        for (int j = 0; j < NS; ++j)
            if (t / WAVE == (uint32_t)j) v = __shfl(wr[j], (int)(t % WAVE));
        return v;
    };
    const float lamd = a.lambda * (float)d;
    uint32_t c[FOLD_E];
    float r[FOLD_E], hv[FOLD_E];
#pragma unroll
    for (int q = 0; q < FOLD_E; ++q) {
        const uint32_t p = lane + WAVE * q;
        const bool ok = p < d;
        c[q] = ok ? a.idx[s + p] : 0u;
        r[q] = ok ? a.val[s + p] : 0.0f;
    }
    // residual of the starting factors, ranks in ascending order
    for (uint32_t t = 0; t < k; ++t) {
        const float wt = wget(t);
        const float* ht = H + (size_t)t * a.n;
#pragma unroll
        for (int q = 0; q < FOLD_E; ++q)
            if (lane + WAVE * q < d) r[q] = r[q] - wt * ht[c[q]];
    }
    for (int sw = 0; sw < a.sweeps; ++sw) {
        for (uint32_t t = 0; t < k; ++t) {
            const float wo = wget(t);
            const float* ht = H + (size_t)t * a.n;
            float g = 0.0f, hh = 0.0f;
#pragma unroll
            for (int q = 0; q < FOLD_E; ++q) {
                const bool ok = lane + WAVE * q < d;
                const float h = ok ? ht[c[q]] : 0.0f;
                hv[q] = h;
                const float av = r[q] + wo * h;
                r[q] = av;
                g += av * h;
                hh += h * h;
            }
            g = wave_sum(g);
            hh = wave_sum(hh);
            const float wn = d ? g / (lamd + hh) : 0.0f;
#pragma unroll
            for (int q = 0; q < FOLD_E; ++q) r[q] = r[q] - wn * hv[q];
#pragma unroll
            for (int j = 0; j < NS; ++j)
                if (t / WAVE == (uint32_t)j && (uint32_t)lane == t % WAVE) wr[j] = wn;
        }
    }
#pragma unroll
    for (int j = 0; j < NS; ++j) {
        const uint32_t t = lane + WAVE * j;
        if (t < k) W[(size_t)row * k + t] = wr[j];
    }
}

// One long row on the whole block: thread i holds elements i + 1024 q,
// q < LONG_E; elements beyond 16384 keep their residual in rbuf.  w in LDS;
// per rank one block reduction of (g, h): wave sums, then the 16 wave
// partials in wave order.
__device__ __forceinline__ void sweep_long_row(const FoldArgs& a, uint32_t row, const float* __restrict__ W0,
                                               float* __restrict__ W, float* __restrict__ rbuf, float* wsh,
                                               float (*red)[LWAVES]) {
    const uint32_t s = a.ptr[row], e = a.ptr[row + 1], d = e - s;
    const uint32_t k = a.k;
    const float* __restrict__ H = a.H;
    const int lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
    for (uint32_t t = threadIdx.x; t < k; t += LBLOCK) wsh[t] = W0 ? W0[(size_t)row * k + t] : 0.0f;
    __syncthreads();
    const float lamd = a.lambda * (float)d;
    constexpr uint32_t CAP = LBLOCK * LONG_E;
    uint32_t c[LONG_E];
    float r[LONG_E], hv[LONG_E];
#pragma unroll
    for (int q = 0; q < LONG_E; ++q) {
        const uint32_t p = threadIdx.x + LBLOCK * q;
        const bool ok = p < d;
        c[q] = ok ? a.idx[s + p] : 0u;
        r[q] = ok ? a.val[s + p] : 0.0f;
    }
    for (uint32_t t = 0; t < k; ++t) {
        const float wt = wsh[t];
        const float* ht = H + (size_t)t * a.n;
#pragma unroll
        for (int q = 0; q < LONG_E; ++q)
            if (threadIdx.x + LBLOCK * q < d) r[q] = r[q] - wt * ht[c[q]];
        for (uint32_t p = CAP + threadIdx.x; p < d; p += LBLOCK)
            rbuf[s + p] = (t == 0 ? a.val[s + p] : rbuf[s + p]) - wt * ht[a.idx[s + p]];
    }
    for (int sw = 0; sw < a.sweeps; ++sw) {
        for (uint32_t t = 0; t < k; ++t) {
            const float wo = wsh[t];
            const float* ht = H + (size_t)t * a.n;
            float g = 0.0f, hh = 0.0f;
#pragma unroll
            for (int q = 0; q < LONG_E; ++q) {
                const bool ok = threadIdx.x + LBLOCK * q < d;
                const float h = ok ? ht[c[q]] : 0.0f;
                hv[q] = h;
                const float av = r[q] + wo * h;
                r[q] = av;
                g += av * h;
                hh += h * h;
            }
            for (uint32_t p = CAP + threadIdx.x; p < d; p += LBLOCK) {
                const float h = ht[a.idx[s + p]];
                const float av = rbuf[s + p] + wo * h;
                rbuf[s + p] = av;
                g += av * h;
                hh += h * h;
            }
            g = wave_sum(g);
            hh = wave_sum(hh);
            if (lane == 0) {
                red[0][wv] = g;
                red[1][wv] = hh;
            }
            __syncthreads();
            float gt = 0.0f, ht2 = 0.0f;
            for (int w = 0; w < LWAVES; ++w) {
                gt += red[0][w];
                ht2 += red[1][w];
            }
            const float wn = gt / (lamd + ht2);
#pragma unroll
            for (int q = 0; q < LONG_E; ++q) r[q] = r[q] - wn * hv[q];
            for (uint32_t p = CAP + threadIdx.x; p < d; p += LBLOCK) rbuf[s + p] = rbuf[s + p] - wn * ht[a.idx[s + p]];
            if (threadIdx.x == 0) wsh[t] = wn;
            __syncthreads();  // red is reused by the next rank
        }
    }
    for (uint32_t t = threadIdx.x; t < k; t += LBLOCK) W[(size_t)row * k + t] = wsh[t];
}

__global__ __launch_bounds__(LBLOCK) void fold_sweep_kernel(FoldArgs a, const uint32_t* __restrict__ long_rows,
                                                            uint32_t nlong, const float* __restrict__ W0,
                                                            float* __restrict__ W, float* __restrict__ rbuf) {
    __shared__ float wsh[FOLD_KMAX];
    __shared__ float red[2][LWAVES];
    if (blockIdx.x < nlong) {
        sweep_long_row(a, long_rows[blockIdx.x], W0, W, rbuf, wsh, red);
        return;
    }
    const uint32_t row = (blockIdx.x - nlong) * LWAVES + threadIdx.x / WAVE;
    if (row >= a.rows) return;
    if (a.ptr[row + 1] - a.ptr[row] > FOLD_SHORT) return;
    sweep_short_row(a, row, threadIdx.x & (WAVE - 1), W0, W);
}

__global__ __launch_bounds__(BLOCK) void transpose_kernel(const float* __restrict__ src, float* __restrict__ dst,
                                                          uint32_t rows, uint32_t cols) {
    const uint64_t total = (uint64_t)rows * cols;
    for (uint64_t q = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; q < total; q += (uint64_t)gridDim.x * BLOCK) {
        const uint64_t r = q / cols, c = q % cols;
        dst[c * rows + r] = src[q];
    }
}

uint32_t grid_rows(uint64_t rows, uint32_t nlong) { return nlong + (uint32_t)((rows + LWAVES - 1) / LWAVES); }

}  // namespace

hipError_t launch_fold_features(const FoldArgs& a, const uint32_t* long_rows, uint32_t nlong, float* F,
                                hipStream_t s) {
    if (a.rows == 0) return hipSuccess;
    if (a.k > (uint32_t)FOLD_KMAX) return hipErrorInvalidValue;
    if (nlong) hipLaunchKernelGGL(fold_features_long_kernel, dim3(nlong), dim3(LBLOCK), 0, s, a, long_rows, F);
    hipLaunchKernelGGL(fold_features_short_kernel, dim3(a.rows), dim3(WAVE), 0, s, a, F);
    return hipGetLastError();
}

hipError_t launch_fold_sweeps(const FoldArgs& a, const uint32_t* long_rows, uint32_t nlong, const float* W0, float* W,
                              float* rbuf, hipStream_t s) {
    if (a.rows == 0) return hipSuccess;
    if (a.k > (uint32_t)FOLD_KMAX) return hipErrorInvalidValue;
    hipLaunchKernelGGL(fold_sweep_kernel, dim3(grid_rows(a.rows, nlong)), dim3(LBLOCK), 0, s, a, long_rows, nlong, W0,
                       W, rbuf);
    return hipGetLastError();
}

hipError_t launch_transpose(const float* src, float* dst, uint32_t rows, uint32_t cols, hipStream_t s) {
    const uint64_t total = (uint64_t)rows * cols;
    if (total == 0) return hipSuccess;
    uint64_t g = (total + BLOCK - 1) / BLOCK;
    if (g > 8192) g = 8192;
    hipLaunchKernelGGL(transpose_kernel, dim3((uint32_t)g), dim3(BLOCK), 0, s, src, dst, rows, cols);
    return hipGetLastError();
}

}  // namespace ccd
