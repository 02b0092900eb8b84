// Inference of the warm-start MLP (in -> 256 -> 256 -> out, GELU) on the
// bfloat16 matrix cores of gfx950 (v_mfma_f32_16x16x32_bf16), all three
// layers in one kernel.
//
// One 256-thread block (4 waves) takes BM = 32 rows: it loads their
// features, standardises them and rounds them to bfloat16 in LDS, then each
// layer is a [32 x K] x [K x N] product on 16 x 16 output tiles, dealt to
// the waves, accumulated in FP32 by the matrix cores over K in steps of 32.
// A hidden layer's epilogue adds the bias, applies GELU and writes the
// activation back to LDS as bfloat16; the last one adds the bias, maps the
// output back (y = (y_n - out_sub) out_mul) and stores FP32 rows.  Nothing
// but the features and the outputs touches HBM; the weights (about 200 KB
// of bfloat16) come from L2.
//
// Weights are a packed bfloat16 copy with padded rows (mlp_pack_kernel):
// W1 [256][kin] with kin = in rounded up to 32 (zero columns), W2 [256][256],
// W3 [nout][256] with nout = out rounded up to 16 (zero rows), and FP32
// biases padded likewise.
//
// Operand layout of v_mfma_f32_16x16x32_bf16 (64 lanes): lane l supplies
// A[l % 16][8 (l / 16) + e] and B[8 (l / 16) + e][l % 16], e = 0..7, as
// eight bfloat16 in four VGPRs, and receives D[4 (l / 16) + r][l % 16],
// r = 0..3.
#include "ccd_kernels.h"

namespace ccd {
namespace {

constexpr int WAVE = 64;
constexpr int MBLOCK = 256;
constexpr int MWAVES = MBLOCK / WAVE;
constexpr int BM = 32;                  // rows per block
constexpr int HID = MLP_MFMA_HIDDEN;    // 256
constexpr int PAD = 8;                  // bfloat16 of padding per LDS row (bank skew)

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

__device__ __forceinline__ float gelu_f(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }

__device__ __forceinline__ bf16x8 load8(const uint16_t* p) {
    const uint4 v = *reinterpret_cast<const uint4*>(p);
    return __builtin_bit_cast(bf16x8, v);
}

// One layer: out[BM][N] = A[BM][K] (LDS, row stride lda) x W^T (W [N][K],
// row stride ldw, global) + bias.  LAST: map back and store to Y; else GELU
// to LDS out (row stride ldo) as bfloat16.
template <bool LAST>
__device__ __forceinline__ void layer(const uint16_t* A, int lda, int K, const uint16_t* __restrict__ W, int ldw, int N,
                                      const float* __restrict__ bias, uint16_t* out, int ldo, const MlpInferArgs& g,
                                      uint32_t row0) {
    const int lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
    const int tiles_n = N / 16, tiles = (BM / 16) * tiles_n;
    const int lr = lane % 16, lk = 8 * (lane / 16);
    for (int tile = wv; tile < tiles; tile += MWAVES) {
        const int rt = tile / tiles_n, ct = tile % tiles_n;
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        const uint16_t* ap = A + (rt * 16 + lr) * lda + lk;
        const uint16_t* wp = W + (size_t)(ct * 16 + lr) * ldw + lk;
        for (int k0 = 0; k0 < K; k0 += 32)
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(load8(ap + k0), load8(wp + k0), acc, 0, 0, 0);
        const int j = ct * 16 + lr;
        const float bj = bias[j];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = rt * 16 + 4 * (lane / 16) + r;
            const float z = acc[r] + bj;
            if constexpr (LAST) {
                const uint32_t row = row0 + (uint32_t)i;
                if (row < g.rows && (uint32_t)j < g.out) {
                    const float y = g.out_sub ? (z - g.out_sub[j]) * g.out_mul[j] : z;
                    g.Y[(size_t)row * g.out + j] = y;
                }
            } else {
                out[i * ldo + j] = to_bf16(gelu_f(z));
            }
        }
    }
}

__global__ __launch_bounds__(MBLOCK) void mlp_infer_kernel(MlpInferArgs g) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int lda0 = (int)g.kin + PAD, lda1 = HID + PAD;
    uint16_t* a0 = reinterpret_cast<uint16_t*>(lds_raw);
    uint16_t* a1 = a0 + BM * lda0;
    uint16_t* a2 = a1 + BM * lda1;
    const uint32_t row0 = blockIdx.x * BM;
    // features -> standardised -> bfloat16; zero past `in` and past the last row
    for (uint32_t q = threadIdx.x; q < (uint32_t)BM * g.kin; q += MBLOCK) {
        const uint32_t i = q / g.kin, c = q % g.kin, row = row0 + i;
        float x = 0.0f;
        if (row < g.rows && c < g.in) {
            x = g.X[(size_t)row * g.in + c];
            if (g.in_sub) x = (x - g.in_sub[c]) * g.in_mul[c];
        }
        a0[i * lda0 + c] = to_bf16(x);
    }
    __syncthreads();
    layer<false>(a0, lda0, (int)g.kin, g.W1, (int)g.kin, HID, g.b1, a1, lda1, g, row0);
    __syncthreads();
    layer<false>(a1, lda1, HID, g.W2, HID, HID, g.b2, a2, lda1, g, row0);
    __syncthreads();
    layer<true>(a2, lda1, HID, g.W3, HID, (int)g.nout, g.b3, nullptr, 0, g, row0);
}

// Packed weights from the flat FP32 parameters (W1 [hid][in], b1, W2, b2,
// W3 [out][hid], b3): bfloat16 rows padded with zeros, FP32 biases padded
// with zeros.  One thread per packed element.
__global__ __launch_bounds__(MBLOCK) void mlp_pack_kernel(const float* __restrict__ P, uint32_t in, uint32_t out,
                                                          uint32_t kin, uint32_t nout, uint16_t* __restrict__ W1,
                                                          uint16_t* __restrict__ W2, uint16_t* __restrict__ W3,
                                                          float* __restrict__ b1, float* __restrict__ b2,
                                                          float* __restrict__ b3) {
    const uint64_t n1 = (uint64_t)HID * kin, n2 = (uint64_t)HID * HID, n3 = (uint64_t)nout * HID;
    const float* pW1 = P;
    const float* pb1 = pW1 + (size_t)HID * in;
    const float* pW2 = pb1 + HID;
    const float* pb2 = pW2 + (size_t)HID * HID;
    const float* pW3 = pb2 + HID;
    const float* pb3 = pW3 + (size_t)out * HID;
    const uint64_t total = n1 + n2 + n3 + 2 * HID + nout;
    for (uint64_t q = (uint64_t)blockIdx.x * MBLOCK + threadIdx.x; q < total; q += (uint64_t)gridDim.x * MBLOCK) {
        uint64_t e = q;
        if (e < n1) {
            const uint32_t j = (uint32_t)(e / kin), c = (uint32_t)(e % kin);
            W1[e] = c < in ? to_bf16(pW1[(size_t)j * in + c]) : (uint16_t)0;
            continue;
        }
        e -= n1;
        if (e < n2) {
            W2[e] = to_bf16(pW2[e]);
            continue;
        }
        e -= n2;
        if (e < n3) {
            const uint32_t j = (uint32_t)(e / HID), c = (uint32_t)(e % HID);
            W3[e] = j < out ? to_bf16(pW3[(size_t)j * HID + c]) : (uint16_t)0;
            continue;
        }
        e -= n3;
        if (e < HID) {
            b1[e] = pb1[e];
            continue;
        }
        e -= HID;
        if (e < HID) {
            b2[e] = pb2[e];
            continue;
        }
        e -= HID;
        b3[e] = e < out ? pb3[e] : 0.0f;
    }
}

}  // namespace

size_t mlp_infer_lds_bytes(uint32_t kin) { return sizeof(uint16_t) * (size_t)BM * ((kin + PAD) + 2 * (HID + PAD)); }

hipError_t launch_mlp_infer(const MlpInferArgs& g, hipStream_t s) {
    if (g.rows == 0) return hipSuccess;
    if (g.kin % 32 || g.nout % 16 || g.kin < g.in || g.nout < g.out || g.kin > MLP_MFMA_MAX_KIN || g.nout > HID)
        return hipErrorInvalidValue;
    const uint32_t blocks = (g.rows + BM - 1) / BM;
    hipLaunchKernelGGL(mlp_infer_kernel, dim3(blocks), dim3(MBLOCK), mlp_infer_lds_bytes(g.kin), s, g);
    return hipGetLastError();
}

hipError_t launch_mlp_pack(const float* P, uint32_t in, uint32_t out, uint32_t kin, uint32_t nout, uint16_t* W1,
                           uint16_t* W2, uint16_t* W3, float* b1, float* b2, float* b3, hipStream_t s) {
    const uint64_t total = (uint64_t)HID * kin + (uint64_t)HID * HID + (uint64_t)nout * HID + 2 * HID + nout;
    uint64_t blocks = (total + MBLOCK - 1) / MBLOCK;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(mlp_pack_kernel, dim3((uint32_t)blocks), dim3(MBLOCK), 0, s, P, in, out, kin, nout, W1, W2, W3,
                       b1, b2, b3);
    return hipGetLastError();
}

}  // namespace ccd
