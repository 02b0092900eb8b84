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
//     wave-per-row for long ones), 64/G rows per pass; rows are numbered by
//     falling degree, so every unit is a contiguous row range;
//   * rows longer than a threshold are split into pieces, each piece one
//     wave; the per-piece (g, h) partials are reduced in piece order by
//     finalize_kernel, so results are run-to-run deterministic.
// Whole-wave ranges (split pieces, wave-per-row groups) read 16-byte groups
// of four elements per lane; every index/value stream is read and written
// with non-temporal hints so it does not evict the gathered factor vector
// from L2.  The kernel is persistent: each 1024-thread block stages the hot
// prefix of the gathered vector (the highest-degree rows or columns) in LDS
// and its waves then walk the work units in a grid-stride loop.
//
// Template parameters: IdxT (u32 or u16 indices), TF (float or bf16 factor
// storage), TA (float or double per-row sums), SUB/ADD (residual update),
// SWEEP (compute the row update).  The file is built with
// -ffp-contract=off: the residual update rounds the product and the sum
// separately, exactly like the CPU reference.
#include "ccd_kernels.h"

#include <math.h>

namespace ccd {
namespace {

constexpr int WAVE = 64;
constexpr int BLOCK = 256;

template <class TA>
struct Pair {
    TA g, h;
};

// Per-block view of the arguments: the LDS copies of the hot prefixes of x
// and of the pair array.
template <typename TF>
struct Ctx {
    const SweepArgs& a;
    const TF* lds;
    const FPair<TF>* ldsp;
};

// One element: residual update (fused passes take both factor values from
// one pair gather), then the sweep's contribution.
template <typename TF, typename TA, bool SUB, bool ADD, bool SWEEP>
__device__ __forceinline__ void one_elem(const Ctx<TF>& k, uint32_t c, float& r, float ysr, float yar, TA& g,
                                         TA& h) {
    constexpr bool PAIR = SWEEP && (SUB || ADD);
    float xv = 0.0f;
    if constexpr (PAIR) {
        const FPair<TF> p = gather_pair<TF>(k.ldsp, k.a.hot_p, k.a.xp, c);
        if (SUB) r = r - tof<TF>(p.prev) * ysr;
        if (ADD) r = r + tof<TF>(p.cur) * yar;
        xv = k.a.x_is_cur ? tof<TF>(p.cur) : gather<TF>(k.lds, k.a.hot_x, k.a.x, c);
    } else {
        if (SUB) r = r - ldf<TF>(k.a.xs, c) * ysr;
        if (ADD) r = r + ldf<TF>(k.a.xa, c) * yar;
        if (SWEEP) xv = gather<TF>(k.lds, k.a.hot_x, k.a.x, c);
    }
    if (SWEEP) {
        g += (TA)xv * (TA)r;
        h += (TA)xv * (TA)xv;
    }
}

template <typename IdxT, typename TF, typename TA, bool SUB, bool ADD, bool SWEEP>
__device__ __forceinline__ void strided_range(const Ctx<TF>& k, const IdxT* __restrict__ idx, uint32_t p,
                                              const uint32_t end, const uint32_t step, const float ysr,
                                              const float yar, TA& g, TA& h) {
    float* __restrict__ val = k.a.val;
    // four independent elements per lane in flight
    for (; p + 3 * step < end; p += 4 * step) {
        uint32_t c[4];
        float r[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) c[q] = __builtin_nontemporal_load(idx + p + q * step);
#pragma unroll
        for (int q = 0; q < 4; ++q) r[q] = __builtin_nontemporal_load(val + p + q * step);
#pragma unroll
        for (int q = 0; q < 4; ++q) one_elem<TF, TA, SUB, ADD, SWEEP>(k, c[q], r[q], ysr, yar, g, h);
        if (SUB || ADD) {
#pragma unroll
            for (int q = 0; q < 4; ++q) __builtin_nontemporal_store(r[q], val + p + q * step);
        }
    }
    for (; p < end; p += step) {
        const uint32_t c = __builtin_nontemporal_load(idx + p);
        float r = __builtin_nontemporal_load(val + p);
        one_elem<TF, TA, SUB, ADD, SWEEP>(k, c, r, ysr, yar, g, h);
        if (SUB || ADD) __builtin_nontemporal_store(r, val + p);
    }
}

// Four consecutive elements starting at 4*q (16-byte aligned): one 16-byte
// value load, one 16-byte (or 8-byte for 16-bit) index load.
template <typename IdxT, typename TF, typename TA, bool SUB, bool ADD, bool SWEEP>
__device__ __forceinline__ void group4(const Ctx<TF>& k, const IdxT* __restrict__ idx, uint32_t q, float ysr,
                                       float yar, TA& g, TA& h) {
    vf4* __restrict__ v4 = reinterpret_cast<vf4*>(k.a.val) + q;
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
#pragma unroll
    for (int j = 0; j < 4; ++j) one_elem<TF, TA, SUB, ADD, SWEEP>(k, c[j], r[j], ysr, yar, g, h);
    if (SUB || ADD) __builtin_nontemporal_store(vf4{r[0], r[1], r[2], r[3]}, v4);
}

// Elements [s, e) by a whole wave: an unaligned head and tail of at most
// three elements each, then the aligned body as 16-byte groups, two groups
// (eight elements) per lane in flight.
template <typename IdxT, typename TF, typename TA, bool SUB, bool ADD, bool SWEEP>
__device__ __forceinline__ void wave_range(const Ctx<TF>& k, const IdxT* __restrict__ idx, uint32_t s, uint32_t e,
                                           uint32_t lane, float ysr, float yar, TA& g, TA& h) {
    uint32_t b0 = (s + 3u) & ~3u;
    if (b0 > e) b0 = e;
    uint32_t b1 = e & ~3u;
    if (b1 < b0) b1 = b0;
    if (lane < b0 - s)
        strided_range<IdxT, TF, TA, SUB, ADD, SWEEP>(k, idx, s + lane, s + lane + 1, 1, ysr, yar, g, h);
    if (lane < e - b1)
        strided_range<IdxT, TF, TA, SUB, ADD, SWEEP>(k, idx, b1 + lane, b1 + lane + 1, 1, ysr, yar, g, h);
    const uint32_t q1 = b1 >> 2;
    uint32_t q = (b0 >> 2) + lane;
    for (; q + WAVE < q1; q += 2 * WAVE) {
        group4<IdxT, TF, TA, SUB, ADD, SWEEP>(k, idx, q, ysr, yar, g, h);
        group4<IdxT, TF, TA, SUB, ADD, SWEEP>(k, idx, q + WAVE, ysr, yar, g, h);
    }
    if (q < q1) group4<IdxT, TF, TA, SUB, ADD, SWEEP>(k, idx, q, ysr, yar, g, h);
}

template <typename IdxT, typename TF, typename TA, bool SUB, bool ADD, bool SWEEP>
__device__ __forceinline__ void sweep_unit(const Ctx<TF>& k, const uint4 unit, const int lane) {
    const SweepArgs& a = k.a;
    const IdxT* __restrict__ idx = static_cast<const IdxT*>(a.idx);
    if ((unit.x & 0xffu) == UNIT_SPLIT) {
        const uint32_t row = unit.y;
        const float ysr = SUB ? ldf<TF>(a.ys, row) : 0.0f;
        const float yar = ADD ? ldf<TF>(a.ya, row) : 0.0f;
        TA g = 0, h = 0;
        wave_range<IdxT, TF, TA, SUB, ADD, SWEEP>(k, idx, unit.z, unit.w, (uint32_t)lane, ysr, yar, g, h);
        if (SWEEP) {
#pragma unroll
            for (int off = WAVE / 2; off > 0; off >>= 1) {
                g += __shfl_xor(g, off);
                h += __shfl_xor(h, off);
            }
            if (lane == 0) static_cast<Pair<TA>*>(a.partials)[unit.x >> 8] = Pair<TA>{g, h};
        }
        return;
    }
    const int lg = (int)((unit.x >> 8) & 7u);
    const uint32_t G = 1u << lg;
    const uint32_t sub = (uint32_t)lane >> lg;
    const uint32_t sl = (uint32_t)lane & (G - 1u);
    const uint32_t rows_per_pass = (uint32_t)WAVE >> lg;
    for (uint32_t base = unit.y; base < unit.z; base += rows_per_pass) {
        const uint32_t row = base + sub;
        const bool valid = row < unit.z;
        const uint32_t s = valid ? a.ptr[row] : 0u;
        const uint32_t e = valid ? a.ptr[row + 1] : 0u;
        const float ysr = (SUB && valid) ? ldf<TF>(a.ys, row) : 0.0f;
        const float yar = (ADD && valid) ? ldf<TF>(a.ya, row) : 0.0f;
        TA g = 0, h = 0;
        if (lg == 6) wave_range<IdxT, TF, TA, SUB, ADD, SWEEP>(k, idx, s, e, (uint32_t)lane, ysr, yar, g, h);
        else strided_range<IdxT, TF, TA, SUB, ADD, SWEEP>(k, idx, s + sl, e, G, ysr, yar, g, h);
        if (SWEEP) {
            for (uint32_t off = G >> 1; off > 0; off >>= 1) {
                g += __shfl_xor(g, (int)off);
                h += __shfl_xor(h, (int)off);
            }
            if (valid && sl == 0)
                stf<TF>(a.y, row, (e > s) ? (float)(g / ((TA)a.lambda * (TA)(e - s) + h)) : 0.0f);
        }
    }
}

template <typename IdxT, typename TF, typename TA, bool SUB, bool ADD, bool SWEEP>
__global__ __launch_bounds__(SWEEP_BLOCK) void sweep_kernel(SweepArgs a) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    TF* lds = reinterpret_cast<TF*>(lds_raw);
    FPair<TF>* ldsp = reinterpret_cast<FPair<TF>*>(lds_raw + lds_pair_offset<TF>(a.hot_x));
    stage_lds<SWEEP_BLOCK>(lds, static_cast<const TF*>(a.x), a.hot_x);
    stage_lds<SWEEP_BLOCK>(ldsp, static_cast<const FPair<TF>*>(a.xp), a.hot_p);
    __syncthreads();
    const Ctx<TF> k{a, lds, ldsp};
    const int lane = threadIdx.x & (WAVE - 1);
    constexpr uint32_t NW = SWEEP_BLOCK / WAVE;
    const uint32_t stride = gridDim.x * NW;
    for (uint32_t u = blockIdx.x * NW + (threadIdx.x >> 6); u < a.nunits; u += stride)
        sweep_unit<IdxT, TF, TA, SUB, ADD, SWEEP>(k, a.units[u], lane);
}

template <typename TF, typename TA>
__global__ __launch_bounds__(BLOCK) void finalize_kernel(FinalizeArgs a) {
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= a.nsplit) return;
    const uint32_t row = a.split_rows[i];
    const Pair<TA>* __restrict__ part = static_cast<const Pair<TA>*>(a.partials);
    TA g = 0, h = 0;
    for (uint32_t q = a.split_beg[i]; q < a.split_beg[i + 1]; ++q) {
        g += part[q].g;
        h += part[q].h;
    }
    const uint32_t d = a.ptr[row + 1] - a.ptr[row];
    stf<TF>(a.y, row, (float)(g / ((TA)a.lambda * (TA)d + h)));
}

// Per test entry: pred = sum_t W[t][i] * H[t][j] with each product rounded
// to float and accumulated in double (tools.cpp:556-565); NaN squared errors
// are skipped but counted (tools.cpp:566-573).
template <typename TF>
__global__ __launch_bounds__(BLOCK) void rmse_partial_kernel(const uint32_t* __restrict__ trow,
                                                             const uint32_t* __restrict__ tcol,
                                                             const float* __restrict__ tval, uint64_t nnz,
                                                             const void* __restrict__ W, const void* __restrict__ H,
                                                             int64_t m, int64_t n, uint32_t k,
                                                             double* __restrict__ block_sums) {
    double acc = 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
    for (uint64_t e = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; e < nnz; e += stride) {
        const uint32_t i = trow[e], j = tcol[e];
        double pred = 0.0;
        for (uint32_t t = 0; t < k; ++t) {
            const float p = ldf<TF>(W, (size_t)t * m + i) * ldf<TF>(H, (size_t)t * n + j);
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

// The objective CCD++ minimises, in FP64: sum over the observed entries of
// the residual squared, plus lambda * (sum_i |row i| ||w_i||^2 + sum_j
// |col j| ||h_j||^2).  The reference computes the first sum in UpdateRating
// and drops it (cdmf_ref.cpp:27,38).  Per-block partial sums in a fixed
// order; term 0 is the loss, term 1 the regulariser.
__global__ __launch_bounds__(BLOCK) void objective_partial_kernel(const float* __restrict__ res, uint64_t nnz,
                                                                  const uint32_t* __restrict__ row_ptr,
                                                                  const uint32_t* __restrict__ col_ptr,
                                                                  const float* __restrict__ W,
                                                                  const float* __restrict__ H, int64_t m, int64_t n,
                                                                  uint32_t k, double* __restrict__ block_sums) {
    double loss = 0.0, reg = 0.0;
    const uint64_t tid = (uint64_t)blockIdx.x * BLOCK + threadIdx.x, stride = (uint64_t)gridDim.x * BLOCK;
    for (uint64_t e = tid; e < nnz; e += stride) loss += (double)res[e] * (double)res[e];
    for (uint64_t i = tid; i < (uint64_t)(m + n); i += stride) {
        const bool w = i < (uint64_t)m;
        const uint64_t r = w ? i : i - (uint64_t)m;
        const uint64_t dim = w ? (uint64_t)m : (uint64_t)n;
        const float* X = w ? W : H;
        const uint32_t* ptr = w ? row_ptr : col_ptr;
        double sq = 0.0;
        for (uint32_t t = 0; t < k; ++t) {
            const double x = (double)X[(size_t)t * dim + r];
            sq += x * x;
        }
        reg += (double)(ptr[r + 1] - ptr[r]) * sq;
    }
    __shared__ double sh[2][BLOCK];
    sh[0][threadIdx.x] = loss;
    sh[1][threadIdx.x] = reg;
    __syncthreads();
#pragma unroll
    for (int s = BLOCK / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        block_sums[2 * blockIdx.x] = sh[0][0];
        block_sums[2 * blockIdx.x + 1] = sh[1][0];
    }
}

__global__ __launch_bounds__(BLOCK) void objective_final_kernel(const double* __restrict__ block_sums, int nb,
                                                                double lambda, double* __restrict__ out) {
    __shared__ double sh[2][BLOCK];
    double loss = 0.0, reg = 0.0;
    for (int b = threadIdx.x; b < nb; b += BLOCK) {
        loss += block_sums[2 * b];
        reg += block_sums[2 * b + 1];
    }
    sh[0][threadIdx.x] = loss;
    sh[1][threadIdx.x] = reg;
    __syncthreads();
#pragma unroll
    for (int s = BLOCK / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = sh[0][0] + lambda * sh[1][0];
        out[1] = sh[0][0];
        out[2] = sh[1][0];
    }
}

__global__ __launch_bounds__(BLOCK) void copy_kernel(const float4* __restrict__ src, float4* __restrict__ dst,
                                                     uint64_t count) {
    const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
    for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < count; i += stride) dst[i] = src[i];
}

__global__ __launch_bounds__(BLOCK) void to_bf16_kernel(const float* __restrict__ src, bf16* __restrict__ dst,
                                                        uint64_t count) {
    const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
    for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < count; i += stride) dst[i] = to_bf16(src[i]);
}

__global__ __launch_bounds__(BLOCK) void from_bf16_kernel(const bf16* __restrict__ src, float* __restrict__ dst,
                                                          uint64_t count) {
    const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
    for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < count; i += stride)
        dst[i] = __uint_as_float((uint32_t)src[i] << 16);
}

// Elements of each row below `bound`, for rows whose indices ascend: a
// binary search per row (one thread each).
template <typename IT>
__global__ __launch_bounds__(BLOCK) void count_below_kernel(const uint32_t* __restrict__ ptr,
                                                            const IT* __restrict__ idx, uint64_t rows,
                                                            uint32_t bound, uint32_t* __restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
    for (uint64_t r = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; r < rows; r += stride) {
        uint32_t lo = ptr[r], hi = ptr[r + 1];
        const uint32_t b = lo;
        while (lo < hi) {  // first position whose index is >= bound
            const uint32_t mid = lo + (hi - lo) / 2;
            if ((uint32_t)idx[mid] < bound) lo = mid + 1;
            else hi = mid;
        }
        out[r] = lo - b;
    }
}

template <typename TF>
__global__ __launch_bounds__(BLOCK) void commit_kernel(CommitArgs a) {
    const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
    for (uint64_t e = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; e < a.nS + a.nF; e += stride) commit_elem<TF>(a, e);
}

uint32_t grid_for(uint64_t count) {
    uint64_t b = (count + BLOCK - 1) / BLOCK;
    if (b > 8192) b = 8192;
    return (uint32_t)(b ? b : 1);
}

using SweepFn = void (*)(SweepArgs);

template <typename IdxT, typename TF, typename TA>
SweepFn sweep_fn_p(int code) {
    switch (code) {
        case 1: return sweep_kernel<IdxT, TF, TA, false, false, true>;
        case 2: return sweep_kernel<IdxT, TF, TA, false, true, false>;
        case 3: return sweep_kernel<IdxT, TF, TA, false, true, true>;
        case 4: return sweep_kernel<IdxT, TF, TA, true, false, false>;
        case 5: return sweep_kernel<IdxT, TF, TA, true, false, true>;
        case 6: return sweep_kernel<IdxT, TF, TA, true, true, false>;
        case 7: return sweep_kernel<IdxT, TF, TA, true, true, true>;
        default: return nullptr;
    }
}

template <typename IdxT>
SweepFn sweep_fn_i(int prec, int code) {
    switch (prec) {
        case PREC_F32: return sweep_fn_p<IdxT, float, float>(code);
        case PREC_BF16: return sweep_fn_p<IdxT, bf16, float>(code);
        case PREC_F64SUM: return sweep_fn_p<IdxT, float, double>(code);
        default: return nullptr;
    }
}

SweepFn sweep_fn(int prec, bool idx16, bool sub, bool add, bool sweep) {
    const int code = (sub ? 4 : 0) | (add ? 2 : 0) | (sweep ? 1 : 0);
    return idx16 ? sweep_fn_i<uint16_t>(prec, code) : sweep_fn_i<uint32_t>(prec, code);
}

}  // namespace

size_t staged_lds_bytes(int prec, uint32_t hot_x, uint32_t hot_p) {
    const size_t es = prec == PREC_BF16 ? 2 : 4;
    return (((size_t)hot_x * es + 15) & ~(size_t)15) + (size_t)hot_p * 2 * es + 16;
}

hipError_t record_unlaunched(const LaunchEvents* ev, hipStream_t s) {
    if (!ev || !ev->start || !ev->stop) return hipSuccess;
    const hipError_t e = hipEventRecord(ev->start, s);
    return e != hipSuccess ? e : hipEventRecord(ev->stop, s);
}

hipError_t launch_sweep(const SweepArgs& a, int prec, bool idx16, bool sub, bool add, bool sweep, uint32_t blocks,
                        hipStream_t s, const LaunchEvents* ev) {
    if (!sub && !add && !sweep) return record_unlaunched(ev, s);
    const SweepFn fn = sweep_fn(prec, idx16, sub, add, sweep);
    if (!fn) return hipErrorInvalidValue;
    const uint32_t need = (a.nunits + SWEEP_BLOCK / WAVE - 1) / (SWEEP_BLOCK / WAVE);
    if (need == 0) return record_unlaunched(ev, s);
    if (blocks > need) blocks = need;
    SweepArgs arg = a;
    void* args[] = {&arg};
    return hipExtLaunchKernel(reinterpret_cast<const void*>(fn), dim3(blocks), dim3(SWEEP_BLOCK), args,
                              staged_lds_bytes(prec, a.hot_x, a.hot_p), s, ev ? ev->start : nullptr,
                              ev ? ev->stop : nullptr, 0);
}

int sweep_blocks_per_cu(int prec, bool idx16, bool sub, bool add, bool sweep) {
    const SweepFn fn = sweep_fn(prec, idx16, sub, add, sweep);
    int nb = 0;
    if (!fn || hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(fn), SWEEP_BLOCK, 0) !=
                   hipSuccess)
        return 1;
    return nb < 1 ? 1 : nb;
}

hipError_t launch_finalize(const FinalizeArgs& a, int prec, hipStream_t s) {
    if (a.nsplit == 0) return hipSuccess;
    const uint32_t blocks = (a.nsplit + BLOCK - 1) / BLOCK;
    switch (prec) {
        case PREC_F32: hipLaunchKernelGGL((finalize_kernel<float, float>), dim3(blocks), dim3(BLOCK), 0, s, a); break;
        case PREC_BF16: hipLaunchKernelGGL((finalize_kernel<bf16, float>), dim3(blocks), dim3(BLOCK), 0, s, a); break;
        case PREC_F64SUM:
            hipLaunchKernelGGL((finalize_kernel<float, double>), dim3(blocks), dim3(BLOCK), 0, s, a);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_rmse(const uint32_t* trow, const uint32_t* tcol, const float* tval, uint64_t nnz, const void* W,
                       const void* H, bool bf16_factors, int64_t m, int64_t n, uint32_t k, double* block_sums,
                       double* out, hipStream_t s) {
    if (bf16_factors)
        hipLaunchKernelGGL(rmse_partial_kernel<bf16>, dim3(RMSE_BLOCKS), dim3(BLOCK), 0, s, trow, tcol, tval, nnz, W,
                           H, m, n, k, block_sums);
    else
        hipLaunchKernelGGL(rmse_partial_kernel<float>, dim3(RMSE_BLOCKS), dim3(BLOCK), 0, s, trow, tcol, tval, nnz, W,
                           H, m, n, k, block_sums);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(rmse_final_kernel, dim3(1), dim3(BLOCK), 0, s, block_sums, RMSE_BLOCKS, nnz, out);
    return hipGetLastError();
}

hipError_t launch_objective(const float* res, uint64_t nnz, const uint32_t* row_ptr, const uint32_t* col_ptr,
                            const float* W, const float* H, int64_t m, int64_t n, uint32_t k, double lambda,
                            double* block_sums, double* out, hipStream_t s) {
    hipLaunchKernelGGL(objective_partial_kernel, dim3(OBJ_BLOCKS), dim3(BLOCK), 0, s, res, nnz, row_ptr, col_ptr, W, H,
                       m, n, k, block_sums);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(objective_final_kernel, dim3(1), dim3(BLOCK), 0, s, block_sums, OBJ_BLOCKS, lambda, out);
    return hipGetLastError();
}

hipError_t launch_copy(const float4* src, float4* dst, uint64_t count, hipStream_t s) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(copy_kernel, dim3(grid_for(count)), dim3(BLOCK), 0, s, src, dst, count);
    return hipGetLastError();
}

hipError_t launch_commit(const CommitArgs& a, bool bf16_factors, hipStream_t s) {
    const uint64_t count = a.nS + a.nF;
    if (count == 0) return hipSuccess;
    if (bf16_factors) hipLaunchKernelGGL(commit_kernel<bf16>, dim3(grid_for(count)), dim3(BLOCK), 0, s, a);
    else hipLaunchKernelGGL(commit_kernel<float>, dim3(grid_for(count)), dim3(BLOCK), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_count_below(const uint32_t* ptr, const void* idx, bool idx16, uint64_t rows, uint32_t bound,
                              uint32_t* out, hipStream_t s) {
    if (rows == 0) return hipSuccess;
    if (idx16)
        hipLaunchKernelGGL(count_below_kernel<uint16_t>, dim3(grid_for(rows)), dim3(BLOCK), 0, s, ptr,
                           static_cast<const uint16_t*>(idx), rows, bound, out);
    else
        hipLaunchKernelGGL(count_below_kernel<uint32_t>, dim3(grid_for(rows)), dim3(BLOCK), 0, s, ptr,
                           static_cast<const uint32_t*>(idx), rows, bound, out);
    return hipGetLastError();
}

hipError_t launch_to_bf16(const float* src, bf16* dst, uint64_t count, hipStream_t s) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(to_bf16_kernel, dim3(grid_for(count)), dim3(BLOCK), 0, s, src, dst, count);
    return hipGetLastError();
}

hipError_t launch_from_bf16(const bf16* src, float* dst, uint64_t count, hipStream_t s) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(from_bf16_kernel, dim3(grid_for(count)), dim3(BLOCK), 0, s, src, dst, count);
    return hipGetLastError();
}

}  // namespace ccd
