// device_common.h -- wavefront / block / grid reductions and launch geometry
// shared by the Krylov kernels (gfx950, 64-lane wavefronts).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tflows {

constexpr int kWave = 64;
constexpr int kBlock = 256;              // threads per block of every streaming kernel
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kNumXcd = 8;               // MI355X: 8 XCDs, blocks dispatched round-robin

// ---------------------------------------------------------------------------
// DPP wavefront sum.  Lanes are combined inside each 16-lane row with DPP
// (quad_perm, row_half_mirror, row_mirror), then the four row totals are
// read with readlane and added in a fixed order, so every lane returns the
// same, deterministic value.
// ---------------------------------------------------------------------------
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
    int64_t b = __double_as_longlong(v);
    int lo = (int)(b & 0xffffffff), hi = (int)(b >> 32);
    lo = __builtin_amdgcn_update_dpp(0, lo, CTRL, 0xF, 0xF, false);
    hi = __builtin_amdgcn_update_dpp(0, hi, CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((int64_t)hi << 32) | (uint32_t)lo);
}

// DPP control words (GFX9 encoding)
constexpr int kQuadSwap1 = 0xB1;        // quad_perm [1,0,3,2]
constexpr int kQuadSwap2 = 0x4E;        // quad_perm [2,3,0,1]
constexpr int kRowHalfMirror = 0x141;   // lane i <-> 7-i inside each 8 lanes
constexpr int kRowMirror = 0x140;       // lane i <-> 15-i inside each 16 lanes

__device__ __forceinline__ double readlane_d(double v, int lane) {
    int64_t b = __double_as_longlong(v);
    int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), lane);
    int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((int64_t)hi << 32) | (uint32_t)lo);
}
__device__ __forceinline__ float readlane_f(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_d<kQuadSwap1>(v);
    v += dpp_d<kQuadSwap2>(v);
    v += dpp_d<kRowHalfMirror>(v);
    v += dpp_d<kRowMirror>(v);
    return (readlane_d(v, 0) + readlane_d(v, 16)) + (readlane_d(v, 32) + readlane_d(v, 48));
}
__device__ __forceinline__ float wave_sum(float v) {
    v += dpp_f<kQuadSwap1>(v);
    v += dpp_f<kQuadSwap2>(v);
    v += dpp_f<kRowHalfMirror>(v);
    v += dpp_f<kRowMirror>(v);
    return (readlane_f(v, 0) + readlane_f(v, 16)) + (readlane_f(v, 32) + readlane_f(v, 48));
}

// Block sum of NV values; the result is valid in every thread.
// `sm` needs kWavesPerBlock * NV elements of shared memory.
template <typename T, int NV>
__device__ __forceinline__ void block_sum(T (&v)[NV], T* sm) {
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) sm[k * kWavesPerBlock + w] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        T s = sm[k * kWavesPerBlock];
#pragma unroll
        for (int i = 1; i < kWavesPerBlock; ++i) s += sm[k * kWavesPerBlock + i];
        v[k] = s;
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Grid reductions without fences or atomics.
//
// A producing kernel stores one partial per block and value:
//     partials[k * gridDim.x + block]
// and a later kernel on the same stream (the kernel boundary makes the stores
// visible) sums them with site_total.  Every block of the consumer sums the
// same values in the same order, so every block obtains the same bits and
// can take the same decision (alfa, beta, convergence) independently.
// ---------------------------------------------------------------------------
template <typename T, int NV>
__device__ __forceinline__ void site_write(T (&v)[NV], T* __restrict__ partials) {
    __shared__ T sm[kWavesPerBlock * NV];
    block_sum<T, NV>(v, (T*)sm);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) partials[k * gridDim.x + blockIdx.x] = v[k];
    }
}

// Each thread adds the partials of blocks b0 + t, b0 + t + kBlock, ... of
// [b0, b1) in that order.  The loads are issued B at a time before any of
// them is added (8 floats or 4 doubles per value, 32 bytes): a plain loop
// waits for each load before it issues the next, a chain of memory round
// trips at the start of every consumer.
template <typename T, int NV>
__device__ __forceinline__ void thread_partials(const T* __restrict__ partials, int G, int b0, int b1,
                                                T (&tot)[NV]) {
    constexpr int B = 32 / (int)sizeof(T);
#pragma unroll
    for (int k = 0; k < NV; ++k) tot[k] = T(0);
    int b = b0 + (int)threadIdx.x;
    for (; b + (B - 1) * kBlock < b1; b += B * kBlock) {
        T v[NV][B];
#pragma unroll
        for (int k = 0; k < NV; ++k)
#pragma unroll
            for (int j = 0; j < B; ++j) v[k][j] = partials[k * G + b + j * kBlock];
#pragma unroll
        for (int k = 0; k < NV; ++k)
#pragma unroll
            for (int j = 0; j < B; ++j) tot[k] += v[k][j];
    }
    for (; b < b1; b += kBlock) {
#pragma unroll
        for (int k = 0; k < NV; ++k) tot[k] += partials[k * G + b];
    }
}

template <typename T, int NV>
__device__ __forceinline__ void site_total(const T* __restrict__ partials, T (&tot)[NV]) {
    __shared__ T sm[kWavesPerBlock * NV];
    thread_partials<T, NV>(partials, (int)gridDim.x, 0, (int)gridDim.x, tot);
    block_sum<T, NV>(tot, (T*)sm);
}

// Sum of the partials of blocks [b0, b1) only (one system of a batch).
template <typename T, int NV>
__device__ __forceinline__ void site_total_range(const T* __restrict__ partials, int b0, int b1, T (&tot)[NV]) {
    __shared__ T sm[kWavesPerBlock * NV];
    thread_partials<T, NV>(partials, (int)gridDim.x, b0, b1, tot);
    block_sum<T, NV>(tot, (T*)sm);
}

// ---------------------------------------------------------------------------
// Launch geometry.  The hardware dispatches block b to XCD (b mod 8); the
// logical block index below gives each XCD one contiguous share of the
// tiles, so halo gathers of neighbouring rows hit that XCD's own L2.
// The grid size is a multiple of 8 whenever it is at least 8.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int logical_block() {
    const int G = gridDim.x, b = blockIdx.x;
    if (G < kNumXcd || (G & (kNumXcd - 1))) return b;
    return (b & (kNumXcd - 1)) * (G / kNumXcd) + b / kNumXcd;
}

// Tiles of a block, visited as t = first, first + step, ... < end.
struct TileSpan {
    int first, end, step;
};

// Strided span for the SpMV kernels: XCD x (blocks b with b mod 8 == x)
// owns one contiguous eighth of the tiles, and its G/8 blocks walk that
// chunk with stride G/8, so at any moment an XCD works on a window of G/8
// consecutive tiles.  Rows at +-(one plane) of a structured mesh were then
// gathered by the previous or the next window, still in that XCD's L2.
__device__ __forceinline__ TileSpan block_span_strided(int ntiles) {
    const int G = gridDim.x, b = blockIdx.x;
    if (G < kNumXcd || (G & (kNumXcd - 1))) return TileSpan{b, ntiles, G};
    const int x = b & (kNumXcd - 1), per = G / kNumXcd;
    const int c0 = (int)(((int64_t)x * ntiles) / kNumXcd), c1 = (int)(((int64_t)(x + 1) * ntiles) / kNumXcd);
    return TileSpan{c0 + b / kNumXcd, c1, per};
}

// Tiles [t0, t1) owned by this block when `ntiles` tiles are split evenly.
__device__ __forceinline__ void block_tiles(int ntiles, int& t0, int& t1) {
    const int G = gridDim.x, lb = logical_block();
    t0 = (int)(((int64_t)lb * ntiles) / G);
    t1 = (int)(((int64_t)(lb + 1) * ntiles) / G);
}

}  // namespace tflows
