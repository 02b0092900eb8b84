// Kernels of the warm-start MLP (in -> hid -> hid -> out, GELU): a tiled
// FP32 GEMM with fused bias/GELU and GELU-gradient epilogues, column sums,
// the MSE gradient, Adam, and row normalisation / gathering.  Weights are
// read as FP32 (training) or bfloat16 (inference); arithmetic is FP32.
// Every output element is produced by one thread summing in a fixed order,
// so results are run-to-run deterministic.
#include "ccd_kernels.h"

#include <algorithm>

namespace ccd {
namespace {

constexpr int BLOCK = 256;
constexpr int TM = 64, TN = 64, TK = 16;

template <class T>
__device__ __forceinline__ float ld(const T* p, size_t i) {
    if constexpr (sizeof(T) == 2) return __uint_as_float((uint32_t)p[i] << 16);
    else return p[i];
}

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * expf(-0.5f * z * z) * 0.39894228040143268f;
}

// C(i, j) = sum_p A(i, p) B(p, j), A(i, p) = A[i sai + p sap], B(p, j) = B[p sbp + j sbj];
// 64 x 64 tile per block, 4 x 4 outputs per thread (rows ty + 16a, columns tx + 16b).
// Tile loads walk the operand's unit-stride direction first (coalesced).
// blockIdx.z: slice of K when the product is split (partials to g.part).
template <class TB, int EPI>
__global__ __launch_bounds__(BLOCK) void gemm_kernel(GemmArgs g, const TB* __restrict__ B) {
    __shared__ float As[TK][TM + 1];
    __shared__ float Bs[TK][TN + 1];
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    const uint32_t i0 = blockIdx.y * TM, j0 = blockIdx.x * TN;
    const uint32_t kc = (g.K + gridDim.z - 1) / gridDim.z;
    const uint32_t kb = blockIdx.z * kc, ke = min(g.K, kb + kc);
    const bool a_i_fast = g.sai == 1, b_p_fast = g.sbp == 1;
    float acc[4][4] = {};
    for (uint32_t p0 = kb; p0 < ke; p0 += TK) {
        for (int q = threadIdx.x; q < TK * TM; q += BLOCK) {
            const int pp = a_i_fast ? q / TM : q % TK, ii = a_i_fast ? q % TM : q / TK;
            const uint32_t i = i0 + ii, p = p0 + pp;
            As[pp][ii] = (i < g.M && p < ke) ? g.A[(size_t)i * g.sai + (size_t)p * g.sap] : 0.0f;
        }
        for (int q = threadIdx.x; q < TK * TN; q += BLOCK) {
            const int pp = b_p_fast ? q % TK : q / TN, jj = b_p_fast ? q / TK : q % TN;
            const uint32_t j = j0 + jj, p = p0 + pp;
            Bs[pp][jj] = (j < g.N && p < ke) ? ld<TB>(B, (size_t)p * g.sbp + (size_t)j * g.sbj) : 0.0f;
        }
        __syncthreads();
#pragma unroll
        for (int pp = 0; pp < TK; ++pp) {
            float a[4], b[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                a[u] = As[pp][ty + 16 * u];
                b[u] = Bs[pp][tx + 16 * u];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[u][v] += a[u] * b[v];
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const uint32_t i = i0 + ty + 16 * u, j = j0 + tx + 16 * v;
            if (i >= g.M || j >= g.N) continue;
            if (gridDim.z > 1) {  // split K: partial of this slice
                g.part[((size_t)blockIdx.z * g.M + i) * g.N + j] = acc[u][v];
                continue;
            }
            const size_t o = (size_t)i * g.ldc + j;
            if constexpr (EPI == GEMM_STORE) {
                g.C[o] = g.bias ? acc[u][v] + ld<TB>(static_cast<const TB*>(g.bias), j) : acc[u][v];
            } else if constexpr (EPI == GEMM_BIAS_GELU) {
                const float z = acc[u][v] + ld<TB>(static_cast<const TB*>(g.bias), j);
                g.C[o] = z;
                g.C2[o] = gelu(z);
            } else {  // GEMM_GELU_GRAD
                g.C[o] = acc[u][v] * gelu_grad(g.aux[o]);
            }
        }
}

// C[i][j] = sum over slices z (in order) of part[z][i][j]
__global__ __launch_bounds__(BLOCK) void splitk_reduce_kernel(GemmArgs g, uint32_t slices) {
    const uint64_t total = (uint64_t)g.M * g.N;
    for (uint64_t q = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; q < total; q += (uint64_t)gridDim.x * BLOCK) {
        float s = 0.0f;
        for (uint32_t z = 0; z < slices; ++z) s += g.part[(size_t)z * total + q];
        g.C[(q / g.N) * g.ldc + q % g.N] = s;
    }
}

// db[j] = sum_i dZ[i][j]: 16 columns x 16 row groups per block; each
// thread sums its rows i = g, g + 16, ... in order, then the 16 group
// partials are added in group order.
__global__ __launch_bounds__(BLOCK) void colsum_kernel(const float* __restrict__ dZ, uint32_t M, uint32_t N,
                                                       float* __restrict__ out) {
    __shared__ float sh[16][17];
    const int c = threadIdx.x % 16, grp = threadIdx.x / 16;
    const uint32_t j = blockIdx.x * 16 + c;
    float s = 0.0f;
    if (j < N)
        for (uint32_t i = grp; i < M; i += 16) s += dZ[(size_t)i * N + j];
    sh[grp][c] = s;
    __syncthreads();
    if (grp == 0 && j < N) {
        float t = 0.0f;
        for (int q = 0; q < 16; ++q) t += sh[q][c];
        out[j] = t;
    }
}

// dY = 2 (Y - T) / count; per-block FP64 sums of (Y - T)^2
__global__ __launch_bounds__(BLOCK) void mse_kernel(const float* __restrict__ Y, const float* __restrict__ T,
                                                    uint64_t count, float* __restrict__ dY, double* __restrict__ part) {
    __shared__ double sh[BLOCK];
    double s = 0.0;
    const float scale = 2.0f / (float)count;
    for (uint64_t q = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; q < count; q += (uint64_t)gridDim.x * BLOCK) {
        const float d = Y[q] - T[q];
        dY[q] = d * scale;
        s += (double)d * (double)d;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = BLOCK / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void sum_parts_kernel(const double* __restrict__ part, uint32_t n, uint64_t count, double* __restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.0;
    for (uint32_t q = 0; q < n; ++q) s += part[q];
    *out = s / (double)count;
}

__global__ __launch_bounds__(BLOCK) void adam_kernel(AdamArgs a) {
    for (uint64_t q = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; q < a.count; q += (uint64_t)gridDim.x * BLOCK) {
        const float g = a.grad[q];
        const float m = a.beta1 * a.m[q] + (1.0f - a.beta1) * g;
        const float v = a.beta2 * a.v[q] + (1.0f - a.beta2) * g * g;
        a.m[q] = m;
        a.v[q] = v;
        a.param[q] -= a.lr * (m / a.bc1) / (sqrtf(v / a.bc2) + a.eps);
    }
}

// Y[i][j] = (X[src_row(i)][j] - sub[j]) * mul[j]   (src_row: rows[i] or i)
__global__ __launch_bounds__(BLOCK) void affine_rows_kernel(const float* __restrict__ X, const uint32_t* __restrict__ rows,
                                                            uint32_t M, uint32_t N, const float* __restrict__ sub,
                                                            const float* __restrict__ mul, float* __restrict__ Y) {
    const uint64_t total = (uint64_t)M * N;
    for (uint64_t q = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; q < total; q += (uint64_t)gridDim.x * BLOCK) {
        const uint64_t i = q / N, j = q % N;
        const uint64_t si = rows ? rows[i] : i;
        const float x = X[si * N + j];
        Y[q] = sub ? (x - sub[j]) * mul[j] : x;
    }
}

uint32_t grid_for(uint64_t count) {
    uint64_t b = (count + BLOCK - 1) / BLOCK;
    if (b > 8192) b = 8192;
    return (uint32_t)(b ? b : 1);
}

}  // namespace

hipError_t launch_gemm(const GemmArgs& g, const void* B, bool bf16_b, int epi, hipStream_t s) {
    if (g.M == 0 || g.N == 0) return hipSuccess;
    // few output tiles and a long K (weight gradients): split K over up to 16 slices
    const uint32_t tiles = ((g.N + TN - 1) / TN) * ((g.M + TM - 1) / TM);
    uint32_t slices = 1;
    if (g.part && epi == GEMM_STORE && !g.bias && tiles < 128 && g.K >= 512)
        slices = std::min<uint32_t>(std::min<uint32_t>(16u, g.K / 256u), g.part_cap / ((uint64_t)g.M * g.N));
    if (slices < 1) slices = 1;
    const dim3 grid((g.N + TN - 1) / TN, (g.M + TM - 1) / TM, slices);
#define CCD_GEMM(TB_, E_) hipLaunchKernelGGL((gemm_kernel<TB_, E_>), grid, dim3(BLOCK), 0, s, g, static_cast<const TB_*>(B))
    if (bf16_b) {
        if (epi == GEMM_STORE) CCD_GEMM(bf16, GEMM_STORE);
        else if (epi == GEMM_BIAS_GELU) CCD_GEMM(bf16, GEMM_BIAS_GELU);
        else return hipErrorInvalidValue;
    } else {
        if (epi == GEMM_STORE) CCD_GEMM(float, GEMM_STORE);
        else if (epi == GEMM_BIAS_GELU) CCD_GEMM(float, GEMM_BIAS_GELU);
        else if (epi == GEMM_GELU_GRAD) CCD_GEMM(float, GEMM_GELU_GRAD);
        else return hipErrorInvalidValue;
    }
#undef CCD_GEMM
    if (slices > 1) {
        const uint64_t total = (uint64_t)g.M * g.N;
        hipLaunchKernelGGL(splitk_reduce_kernel, dim3(grid_for(total)), dim3(BLOCK), 0, s, g, slices);
    }
    return hipGetLastError();
}

hipError_t launch_colsum(const float* dZ, uint32_t M, uint32_t N, float* out, hipStream_t s) {
    if (N == 0) return hipSuccess;
    hipLaunchKernelGGL(colsum_kernel, dim3((N + 15) / 16), dim3(BLOCK), 0, s, dZ, M, N, out);
    return hipGetLastError();
}

hipError_t launch_mse(const float* Y, const float* T, uint64_t count, float* dY, double* part, double* out,
                      hipStream_t s) {
    if (count == 0) return hipSuccess;
    uint32_t g = grid_for(count);
    if (g > MSE_PARTS) g = MSE_PARTS;
    hipLaunchKernelGGL(mse_kernel, dim3(g), dim3(BLOCK), 0, s, Y, T, count, dY, part);
    hipLaunchKernelGGL(sum_parts_kernel, dim3(1), dim3(64), 0, s, part, g, count, out);
    return hipGetLastError();
}

hipError_t launch_adam(const AdamArgs& a, hipStream_t s) {
    if (a.count == 0) return hipSuccess;
    hipLaunchKernelGGL(adam_kernel, dim3(grid_for(a.count)), dim3(BLOCK), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_affine_rows(const float* X, const uint32_t* rows, uint32_t M, uint32_t N, const float* sub,
                              const float* mul, float* Y, hipStream_t s) {
    const uint64_t total = (uint64_t)M * N;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(affine_rows_kernel, dim3(grid_for(total)), dim3(BLOCK), 0, s, X, rows, M, N, sub, mul, Y);
    return hipGetLastError();
}

}  // namespace ccd
