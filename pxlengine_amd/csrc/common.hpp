// Device helpers shared by the kernel files: wave64 / workgroup reductions and scans, and
// the deterministic last-workgroup grid reduction.
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

// Sum over a 256-lane workgroup; all lanes get the total.  `sh` holds 4 slots.
template <typename T>
__device__ __forceinline__ T block_sum(T v, T *sh) {
    v = wave_sum(v);
    const int w = flat_tid() / WAVE;
    if (lane_id() == 0) sh[w] = v;
    __syncthreads();
    T t = ((sh[0] + sh[1]) + sh[2]) + sh[3];
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

template <typename T>
__device__ __forceinline__ void store_agent(T *p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T load_agent(const T *p) {
    return __hip_atomic_load(const_cast<T *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Grid-wide sums of NV values per workgroup.  Each workgroup stores its partials
// (part[k * nb + block]); the last workgroup to arrive (agent-scope acq_rel counter) adds
// all partials in block order -- a fixed tree, so the result does not depend on which
// workgroup finished last -- writes them to total[] in every lane and resets the counter.
// Returns true in the last workgroup only.  Every lane of every workgroup must call it.
template <typename T, int NV>
__device__ bool grid_sum(const T (&v)[NV], T *part, unsigned *ctr, T (&total)[NV]) {
    __shared__ T sh[4 * NV];
    __shared__ int is_last;
    const uint32_t nb = num_blocks(), b = flat_block();
    const int tid = flat_tid();
    T bs[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) bs[k] = block_sum(v[k], sh + 4 * k);
    if (tid == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) store_agent(&part[(size_t)k * nb + b], bs[k]);
        unsigned prev = __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        is_last = prev == nb - 1;
    }
    __syncthreads();
    if (!is_last) return false;
    __atomic_thread_fence(__ATOMIC_ACQUIRE);
#pragma unroll
    for (int k = 0; k < NV; k++) {
        T a = 0;
        for (uint32_t i = tid; i < nb; i += BLOCK) a += load_agent(&part[(size_t)k * nb + i]);
        total[k] = block_sum(a, sh + 4 * k);
    }
    if (tid == 0) store_agent(ctr, 0u);
    return true;
}

}  // namespace pxl
