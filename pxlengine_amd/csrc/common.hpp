// Device helpers shared by the kernel files: wave64 / workgroup reductions and scans.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pxl {

constexpr int WAVE = 64;
constexpr int BLOCK = 256;  // every kernel here runs 256-lane workgroups (4 wave64)

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x + blockDim.x * threadIdx.y) & (WAVE - 1); }
__device__ __forceinline__ int flat_tid() { return (int)(threadIdx.x + blockDim.x * threadIdx.y); }
__device__ __forceinline__ uint32_t flat_block() {
    return blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
}
__device__ __forceinline__ uint32_t num_blocks() { return gridDim.x * gridDim.y * gridDim.z; }

// Butterfly sum across a wave64: every lane ends with the same, order-fixed total.
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = WAVE / 2; off >= 1; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}

// Inclusive prefix sum across a wave64 (Hillis-Steele over lane shuffles).
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
    const int lane = lane_id();
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        uint32_t t = __shfl_up(v, d, WAVE);
        if (lane >= d) v += t;
    }
    return v;
}

// Sum over a workgroup of NW wave64s (default 4); all lanes get the total, added in wave
// order.  `sh` holds NW slots.
template <typename T, int NW = BLOCK / WAVE>
__device__ __forceinline__ T block_sum(T v, T *sh) {
    v = wave_sum(v);
    const int w = flat_tid() / WAVE;
    if (lane_id() == 0) sh[w] = v;
    __syncthreads();
    T t = sh[0];
#pragma unroll
    for (int k = 1; k < NW; k++) t += sh[k];
    __syncthreads();
    return t;
}

// Exclusive scan over a 256-lane workgroup; *total gets the workgroup sum.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *total, uint32_t *sh) {
    uint32_t inc = wave_inclusive_scan(v);
    const int w = flat_tid() / WAVE;
    if (lane_id() == WAVE - 1) sh[w] = inc;
    __syncthreads();
    uint32_t off = 0;
    for (int k = 0; k < w; k++) off += sh[k];
    *total = sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
    return off + inc - v;
}

}  // namespace pxl
