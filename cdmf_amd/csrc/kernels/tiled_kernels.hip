// Tiled (CSR5-style) sweeps for gfx950 -- the "Load Balancing" variant.
//
// Replaces the dual-output CSR5 SpMV of src/detail/kcode/csr5_spmv.cl
// (spmv_csr5_compute/calibrate/tail_partition, :134-565) together with the
// CALV/CALU divide of kcode/ccd_3.cl:1-42 and the UpdateRating kernels that
// V3 could only run after transposing the tiles back to CSR
// (src/cdmf_csr5.cpp:245-325).  Here the residual is updated in the tiled
// layout itself, in the same pass as the sweep, so no re-transposition is
// ever needed.
//
// Persistent kernel: each 1024-thread block stages the hot prefixes of the
// gathered vectors in LDS, then its waves take the block's tiles (every B-th
// tile of the layout) one at a time from a counter in LDS.
// One wave per tile of 64*SIGMA elements.  Lane l owns elements
// l*SIGMA .. l*SIGMA+SIGMA-1 of the tile; they are stored in groups of
// G = 16 / (bytes per entry), element i of lane l at (i/G)*64G + G*l + i%G,
// so every 16-byte load of a lane moves 1 KiB of consecutive values or
// indices per wave (G = 4 for values and 32-bit indices, 8 for 16-bit ones;
// 3-byte indices add a plane of their high bytes, see ccd_kernels.h).  The
// streams are read and written with non-temporal hints so that they do not
// evict the gathered factor vector from L2.  Each lane updates its elements
// and keeps a running (g, h) for the row it is in.  A row that starts and
// ends inside a lane is finished by that lane.  Rows crossing lane boundaries
// are combined by a segmented inclusive scan over the 64 lanes (shuffles,
// fixed order); rows crossing tile boundaries leave per-tile partials
// (head/tail) that tiled_span_kernel adds up in tile order.  A row's length
// comes from the row-start flags (its start and end positions in the tile).
// Tiles whose gathers all fall in the staged prefix (per-tile largest index)
// take a copy of the body with plain LDS reads.  A layout whose gathered
// vector exceeds the prefix can hold its rows split in two virtual rows
// (first the parts gathered from the prefix, then the rest), finished by
// split_finish_kernel.  No atomics in the arithmetic (the block's LDS
// counter only decides which wave runs a tile, and a tile's results do not
// depend on that): results are run-to-run deterministic.
// Built with -ffp-contract=off like the binned kernels.
#include "tile_body.h"

#include <algorithm>

namespace ccd {
namespace {

constexpr int WAVE = tiled::WAVE;
constexpr int BLOCK = 256;
constexpr int WPB = BLOCK / WAVE;
// Blocks of a finish launch: 8 per CU of a 256-CU GPU, one resident round
// of waves at this kernel's occupancy.
constexpr uint32_t FINISH_GRID = 2048;
using tiled::tile_body;

// Threads per block of the tiled kernels: 1024 up to sigma 16; 512 for
// sigma 32, whose lanes hold 32 elements each and need more than the 128
// registers a 1024-thread block leaves a lane.
constexpr int tile_block(int sigma) { return sigma > 16 ? 512 : SWEEP_BLOCK; }

template <typename IdxT, int SIGMA, typename TF, typename TA, bool SUB, bool ADD, bool SWEEP, bool XCUR>
__global__ __launch_bounds__(tile_block(SIGMA)) void tiled_kernel(TiledArgs a) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    TF* lds = reinterpret_cast<TF*>(lds_raw);
    FPair<TF>* ldsp = reinterpret_cast<FPair<TF>*>(lds_raw + lds_pair_offset<TF>(a.hot_x));
    // Tiles are dealt to blocks interleaved (block b: tiles b, b + B, b + 2B,
    // ..., B blocks), so every block samples the whole layout and the blocks'
    // shares cost about the same; inside a block each wave takes the block's
    // next tile from a counter in LDS when it finishes one, so the waves of
    // a block finish within about one tile of each other.  A block runs its
    // tiles in ascending order (all blocks then stream one window of the
    // layout), except the costly tail of a split layout (second parts, whose
    // gathers go to global memory): those tiles are spread evenly over the
    // block's positions, so the pass does not end on them.  (Dealt
    // statically, one wave in a block could finish tens of microseconds
    // after another.)  Which wave runs a tile changes nothing in its results.
    __shared__ uint32_t s_next;
    constexpr int BS = tile_block(SIGMA);
    constexpr uint32_t NW = BS / WAVE;
    if (threadIdx.x == 0) s_next = NW;  // positions 0 .. NW-1 go to the waves in order
    stage_lds<BS>(lds, static_cast<const TF*>(a.x), a.hot_x);
    stage_lds<BS>(ldsp, static_cast<const FPair<TF>*>(a.xp), a.hot_p);
    __syncthreads();
    const int lane = threadIdx.x & (WAVE - 1);
    const uint32_t B = gridDim.x, b = blockIdx.x;
    // The block's tiles b + j B, j < cnt; those with j >= jc form the costly
    // tail (nc of them).  Position p < cnt is a tail position when
    // floor((p + 1) nc / cnt) > floor(p nc / cnt), and takes tail tile
    // jc + floor(p nc / cnt); any other position takes head tile
    // p - floor(p nc / cnt).  A bijection onto j < cnt, each part ascending.
    const uint32_t cnt = b < a.ntiles ? (a.ntiles - 1 - b) / B + 1 : 0u;
    const uint32_t c0 = a.spread_tile0;
    const uint32_t jc = c0 >= a.ntiles ? cnt : c0 <= b ? 0u : min(cnt, (uint32_t)(((uint64_t)c0 - b + B - 1) / B));
    const uint32_t nc = cnt - jc;
    // Fused passes take the positions from both ends in turn (first, last,
    // second, ...) before the map above, so the short rows at a layout's
    // end, which cost a fused pass a little more per tile, are spread over
    // the pass as well.
    const bool zigzag = SUB || ADD;
    auto tile_at = [&](uint32_t p) {
        if (zigzag) p = (p & 1u) ? cnt - 1u - (p >> 1) : (p >> 1);
        uint32_t k0, k1;
        if (cnt < 0x10000u) {
            k0 = p * nc / cnt;
            k1 = (p + 1) * nc / cnt;
        } else {
            k0 = (uint32_t)((uint64_t)p * nc / cnt);
            k1 = (uint32_t)((uint64_t)(p + 1) * nc / cnt);
        }
        return b + (k1 > k0 ? jc + k0 : p - k0) * B;
    };
    // wave-uniform position: the per-tile descriptors become scalar loads
    uint32_t pos = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // Sweeps: each tile's tile_gmax entry is loaded while the previous tile
    // runs (unconditionally, at a clamped tile index: a load on only one
    // of two merged paths would make the compiler drain every outstanding
    // load at the end of each tile).
    uint32_t gmax = SWEEP && pos < cnt ? ld_const(a.tile_gmax, tile_at(pos)) : 0xffffffffu;
    while (pos < cnt) {
        const uint32_t tile = tile_at(pos);
        // the wave's next position, taken before this tile runs; positions
        // only grow, so every wave leaves the loop
        uint32_t pn = 0;
        if (lane == 0) pn = atomicAdd(&s_next, 1u);
        pn = __builtin_amdgcn_readfirstlane(pn);
        const uint32_t next = pn < cnt ? tile_at(pn) : tile;
        const uint32_t gnext = SWEEP ? ld_const(a.tile_gmax, next) : 0xffffffffu;
        tile_body<IdxT, SIGMA, TF, TA, SUB, ADD, SWEEP, XCUR>(a, lds, ldsp, tile, lane, gmax);
        gmax = gnext;
        pos = pn;
    }
}

// Blocks [0, ceil(nshort / BLOCK)) finish two-tile spans, one per thread;
// the rest (at most FINISH_GRID blocks, grid-stride) finish longer spans,
// one per wave.
template <typename TF, typename TA>
__global__ __launch_bounds__(BLOCK) void tiled_span_kernel(SpanArgs a) {
    const uint32_t sb = (a.nshort + BLOCK - 1) / BLOCK;
    if (blockIdx.x < sb) {
        const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
        if (i < a.nshort) tiled::span_short<TF, TA>(a, i);
        return;
    }
    const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (uint32_t i = a.nshort + (blockIdx.x - sb) * WPB + w; i < a.nspan; i += (gridDim.x - sb) * WPB)
        tiled::span_row<TF, TA>(a, i, threadIdx.x & (WAVE - 1));
}

// map (optional): layout position of each element in tiled order
__global__ __launch_bounds__(BLOCK) void retile_f32_kernel(const float* __restrict__ src, float* __restrict__ dst,
                                                           const uint32_t* __restrict__ map, uint64_t nnz,
                                                           uint64_t padded, uint32_t sigma, bool to_tiled) {
    const uint64_t ts = (uint64_t)WAVE * sigma;
    const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
    for (uint64_t e = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; e < padded; e += stride) {
        const uint64_t tile = e / ts, rem = e % ts;
        const uint64_t i = rem % sigma, l = rem / sigma;
        const uint64_t pos = tile * ts + (i >> 2) * (4 * WAVE) + 4 * l + (i & 3);
        if (to_tiled) dst[pos] = e < nnz ? src[map ? map[e] : e] : 0.0f;
        else if (e < nnz) dst[map ? map[e] : e] = src[pos];
    }
}

template <typename IdxT>
__global__ __launch_bounds__(BLOCK) void retile_idx_kernel(const IdxT* __restrict__ src, IdxT* __restrict__ dst,
                                                           const uint32_t* __restrict__ map, uint64_t nnz,
                                                           uint64_t padded, uint32_t sigma) {
    const uint64_t ts = (uint64_t)WAVE * sigma;
    const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
    for (uint64_t e = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; e < padded; e += stride) {
        const uint64_t tile = e / ts, rem = e % ts;
        const uint64_t i = rem % sigma, l = rem / sigma;
        constexpr uint64_t G = 16 / sizeof(IdxT);  // indices per 16-byte lane load
        const uint64_t pos = tile * ts + (i / G) * (G * WAVE) + G * l + (i % G);
        dst[pos] = e < nnz ? src[map ? map[e] : e] : (IdxT)0;
    }
}

// 32-bit indices (each < 2^24) to the 3-byte format: the low 16 bits at the
// position of the u16 layout, bits 16..23 in the lane-major u8 plane.
__global__ __launch_bounds__(BLOCK) void retile_idx24_kernel(const uint32_t* __restrict__ src,
                                                             uint16_t* __restrict__ lo, uint8_t* __restrict__ hi,
                                                             const uint32_t* __restrict__ map, uint64_t nnz,
                                                             uint64_t padded, uint32_t sigma) {
    const uint64_t ts = (uint64_t)WAVE * sigma;
    const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
    for (uint64_t e = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; e < padded; e += stride) {
        const uint64_t tile = e / ts, rem = e % ts;
        const uint64_t i = rem % sigma, l = rem / sigma;
        const uint32_t v = e < nnz ? src[map ? map[e] : e] : 0u;
        lo[tile * ts + (i >> 3) * (8 * WAVE) + 8 * l + (i & 7)] = (uint16_t)(v & 0xffffu);
        hi[e] = (uint8_t)(v >> 16);  // e = tile * ts + l * sigma + i
    }
}

// One wave per row of a split layout (tiled::split_row), grid-stride: the
// grid is at most FINISH_GRID blocks, about one resident round of waves.
template <typename TF, typename TA>
__global__ __launch_bounds__(BLOCK) void split_finish_kernel(SplitArgs a) {
    const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (uint32_t r = blockIdx.x * WPB + w; r < a.nreal; r += gridDim.x * WPB)
        tiled::split_row<TF, TA>(a, r, threadIdx.x & (WAVE - 1));
}

// Largest index of each tile, one wave per tile (grid-stride over tiles);
// a tile that extends past nnz (the padded last one) or holds empty rows
// (segoff other than 0xffffffff) gets 0xffffffff.
template <typename IdxT>
__global__ __launch_bounds__(BLOCK) void tile_gmax_kernel(const IdxT* __restrict__ idx,
                                                          const uint32_t* __restrict__ segoff,
                                                          const uint32_t* __restrict__ map,
                                                          uint32_t* __restrict__ gmax, uint64_t nnz, uint32_t ntiles,
                                                          uint32_t sigma) {
    const uint64_t ts = (uint64_t)WAVE * sigma;
    const int lane = threadIdx.x & (WAVE - 1);
    for (uint32_t t = blockIdx.x * WPB + (threadIdx.x >> 6); t < ntiles; t += gridDim.x * WPB) {
        const uint64_t base = (uint64_t)t * ts;
        uint32_t mx = 0;
        for (uint64_t e = base + lane; e < base + ts && e < nnz; e += WAVE) {
            const uint32_t v = idx[map ? map[e] : e];
            mx = v > mx ? v : mx;
        }
#pragma unroll
        for (int off = WAVE / 2; off > 0; off >>= 1) {
            const uint32_t o = __shfl_xor(mx, off);
            mx = o > mx ? o : mx;
        }
        if (lane == 0) gmax[t] = base + ts > nnz || segoff[t] != 0xffffffffu ? 0xffffffffu : mx;
    }
}

__global__ __launch_bounds__(BLOCK) void fill_kernel(float* __restrict__ dst, float value, uint64_t count) {
    const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
    for (uint64_t e = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; e < count; e += stride) dst[e] = value;
}

__global__ __launch_bounds__(BLOCK) void fill16_kernel(uint16_t* __restrict__ dst, uint16_t value, uint64_t count) {
    const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
    for (uint64_t e = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; e < count; e += stride) dst[e] = value;
}

uint32_t grid_for(uint64_t count) {
    uint64_t b = (count + BLOCK - 1) / BLOCK;
    if (b > 16384) b = 16384;
    return (uint32_t)(b ? b : 1);
}

using TiledFn = void (*)(TiledArgs);

// code: SUB << 2 | ADD << 1 | SWEEP, plus 8 for a fused pass whose x is the
// pair's current value
template <typename IdxT, int SIGMA, typename TF, typename TA>
TiledFn tiled_fn_s(int code) {
    switch (code) {
        case 1: return tiled_kernel<IdxT, SIGMA, TF, TA, false, false, true, false>;
        case 2: return tiled_kernel<IdxT, SIGMA, TF, TA, false, true, false, false>;
        case 3: return tiled_kernel<IdxT, SIGMA, TF, TA, false, true, true, false>;
        case 4: return tiled_kernel<IdxT, SIGMA, TF, TA, true, false, false, false>;
        case 5: return tiled_kernel<IdxT, SIGMA, TF, TA, true, false, true, false>;
        case 6: return tiled_kernel<IdxT, SIGMA, TF, TA, true, true, false, false>;
        case 7: return tiled_kernel<IdxT, SIGMA, TF, TA, true, true, true, false>;
        case 8 + 3: return tiled_kernel<IdxT, SIGMA, TF, TA, false, true, true, true>;
        case 8 + 5: return tiled_kernel<IdxT, SIGMA, TF, TA, true, false, true, true>;
        case 8 + 7: return tiled_kernel<IdxT, SIGMA, TF, TA, true, true, true, true>;
        default: return nullptr;
    }
}

template <typename IdxT>
TiledFn tiled_fn_i(int sigma, int prec, int code) {
    if (prec == PREC_F32) {
        switch (sigma) {
            case 8: return tiled_fn_s<IdxT, 8, float, float>(code);
            case 16: return tiled_fn_s<IdxT, 16, float, float>(code);
            case 32: return tiled_fn_s<IdxT, 32, float, float>(code);
            default: return nullptr;
        }
    }
    if (sigma != 16) return nullptr;
    if (prec == PREC_BF16) return tiled_fn_s<IdxT, 16, bf16, float>(code);
    if (prec == PREC_F64SUM) return tiled_fn_s<IdxT, 16, float, double>(code);
    return nullptr;
}

TiledFn tiled_fn(int sigma, int prec, int ib, bool sub, bool add, bool sweep, bool xcur) {
    const bool fused = sweep && (sub || add);
    const int code = (sub ? 4 : 0) | (add ? 2 : 0) | (sweep ? 1 : 0) | (fused && xcur ? 8 : 0);
    switch (ib) {
        case 2: return tiled_fn_i<uint16_t>(sigma, prec, code);
        case 3: return tiled_fn_i<Idx24>(sigma, prec, code);
        case 4: return tiled_fn_i<uint32_t>(sigma, prec, code);
        default: return nullptr;
    }
}

}  // namespace

hipError_t launch_tiled(const TiledArgs& a, int sigma, int prec, int ib, bool sub, bool add, bool sweep,
                        uint32_t blocks, hipStream_t s, const LaunchEvents* ev) {
    if (!sub && !add && !sweep) return record_unlaunched(ev, s);
    const TiledFn fn = tiled_fn(sigma, prec, ib, sub, add, sweep, a.x_is_cur != 0);
    if (!fn) return hipErrorInvalidValue;
    const int bs = tile_block(sigma);
    const uint32_t need = (a.ntiles + bs / WAVE - 1) / (bs / WAVE);
    if (need == 0) return record_unlaunched(ev, s);
    if (blocks > need) blocks = need;
    TiledArgs arg = a;
    void* args[] = {&arg};
    return hipExtLaunchKernel(reinterpret_cast<const void*>(fn), dim3(blocks), dim3(bs), args,
                              staged_lds_bytes(prec, a.hot_x, a.hot_p), s, ev ? ev->start : nullptr,
                              ev ? ev->stop : nullptr, 0);
}

int tiled_blocks_per_cu(int sigma, int prec, int ib, bool sub, bool add, bool sweep, bool xcur) {
    const TiledFn fn = tiled_fn(sigma, prec, ib, sub, add, sweep, xcur);
    int nb = 0;
    if (!fn || hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(fn), tile_block(sigma),
                                                            0) != hipSuccess)
        return 1;
    return nb < 1 ? 1 : nb;
}

hipError_t launch_span(const SpanArgs& a, int prec, hipStream_t s) {
    if (a.nspan == 0) return hipSuccess;
    if (a.nshort > a.nspan) return hipErrorInvalidValue;
    const dim3 grid((a.nshort + BLOCK - 1) / BLOCK + std::min<uint32_t>((a.nspan - a.nshort + WPB - 1) / WPB, FINISH_GRID));
    switch (prec) {
        case PREC_F32: hipLaunchKernelGGL((tiled_span_kernel<float, float>), grid, dim3(BLOCK), 0, s, a); break;
        case PREC_BF16: hipLaunchKernelGGL((tiled_span_kernel<bf16, float>), grid, dim3(BLOCK), 0, s, a); break;
        case PREC_F64SUM: hipLaunchKernelGGL((tiled_span_kernel<float, double>), grid, dim3(BLOCK), 0, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_split_finish(const SplitArgs& a, int prec, hipStream_t s) {
    if (a.nreal == 0) return hipSuccess;
    const dim3 grid(std::min<uint32_t>((a.nreal + WPB - 1) / WPB, FINISH_GRID));
    switch (prec) {
        case PREC_F32: hipLaunchKernelGGL((split_finish_kernel<float, float>), grid, dim3(BLOCK), 0, s, a); break;
        case PREC_BF16: hipLaunchKernelGGL((split_finish_kernel<bf16, float>), grid, dim3(BLOCK), 0, s, a); break;
        case PREC_F64SUM:
            hipLaunchKernelGGL((split_finish_kernel<float, double>), grid, dim3(BLOCK), 0, s, a);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_retile_f32(const float* src, float* dst, uint64_t nnz, uint32_t ntiles, int sigma, bool to_tiled,
                             hipStream_t s, const uint32_t* map) {
    const uint64_t padded = (uint64_t)ntiles * WAVE * (uint64_t)sigma;
    hipLaunchKernelGGL(retile_f32_kernel, dim3(grid_for(padded)), dim3(BLOCK), 0, s, src, dst, map, nnz, padded,
                       (uint32_t)sigma, to_tiled);
    return hipGetLastError();
}

hipError_t launch_retile_idx(const void* src, bool src16, void* dst, int ib, uint64_t nnz, uint32_t ntiles, int sigma,
                             const uint32_t* map, hipStream_t s) {
    const uint64_t padded = (uint64_t)ntiles * WAVE * (uint64_t)sigma;
    const dim3 grid(grid_for(padded));
    if (src16 != (ib == 2)) return hipErrorInvalidValue;
    if (ib == 2) {
        hipLaunchKernelGGL(retile_idx_kernel<uint16_t>, grid, dim3(BLOCK), 0, s, static_cast<const uint16_t*>(src),
                           static_cast<uint16_t*>(dst), map, nnz, padded, (uint32_t)sigma);
    } else if (ib == 3) {
        uint16_t* lo = static_cast<uint16_t*>(dst);
        hipLaunchKernelGGL(retile_idx24_kernel, grid, dim3(BLOCK), 0, s, static_cast<const uint32_t*>(src), lo,
                           reinterpret_cast<uint8_t*>(lo + padded), map, nnz, padded, (uint32_t)sigma);
    } else if (ib == 4) {
        hipLaunchKernelGGL(retile_idx_kernel<uint32_t>, grid, dim3(BLOCK), 0, s, static_cast<const uint32_t*>(src),
                           static_cast<uint32_t*>(dst), map, nnz, padded, (uint32_t)sigma);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_tile_gmax(const void* idx, bool idx16, const uint32_t* segoff, uint32_t* gmax, uint64_t nnz,
                            uint32_t ntiles, int sigma, const uint32_t* map, hipStream_t s) {
    if (ntiles == 0) return hipSuccess;
    uint32_t blocks = (ntiles + WPB - 1) / WPB;
    if (blocks > 16384) blocks = 16384;
    if (idx16)
        hipLaunchKernelGGL(tile_gmax_kernel<uint16_t>, dim3(blocks), dim3(BLOCK), 0, s,
                           static_cast<const uint16_t*>(idx), segoff, map, gmax, nnz, ntiles, (uint32_t)sigma);
    else
        hipLaunchKernelGGL(tile_gmax_kernel<uint32_t>, dim3(blocks), dim3(BLOCK), 0, s,
                           static_cast<const uint32_t*>(idx), segoff, map, gmax, nnz, ntiles, (uint32_t)sigma);
    return hipGetLastError();
}

hipError_t launch_fill_bytes16(uint16_t* dst, uint16_t value, uint64_t count, hipStream_t s) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(fill16_kernel, dim3(grid_for(count)), dim3(BLOCK), 0, s, dst, value, count);
    return hipGetLastError();
}

hipError_t launch_fill(float* dst, float value, uint64_t count, hipStream_t s) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(fill_kernel, dim3(grid_for(count)), dim3(BLOCK), 0, s, dst, value, count);
    return hipGetLastError();
}

}  // namespace ccd
