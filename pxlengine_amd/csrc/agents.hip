// Agent neighbour grid of pxlengine's agent simulations (shaders/sim.slang:27-104, driven per
// step by crates/pxl-render/src/sim.rs:224-291): agents are grouped by the hash of their grid
// cell, giving per-hash counts, their exclusive scan (each hash's first slot) and the agent
// ids in hash order.
//
// The reference counts with one atomic per agent and scatters with an atomic cursor, so the
// ids of one cell come in thread-timing order and a crowded cell serialises its atomics.  Here
// the (hash, id) pairs are radix-sorted (LSD, 8-bit digits, stable), so the ids of a cell are
// ascending -- one deterministic answer -- and no counter is contended: each pass builds
// per-tile digit histograms in LDS, scans them (digit-major, tile-minor) and scatters every
// tile's elements to their stable global positions.  Counts come from the sorted keys (one
// atomic per run of equal keys in a wave) and offsets from their exclusive scan.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>

#include "common.hpp"
#include "kernels.hpp"
#include "pxlflip.h"
#include "sim.hpp"

namespace pxl {

namespace {

constexpr uint32_t ABLK = 256;
constexpr uint32_t ANW = ABLK / WAVE;
constexpr uint32_t RITEMS = 8;
constexpr uint32_t RTILE = ABLK * RITEMS;  // elements per radix tile
constexpr uint32_t RADIX = 256;

inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

__device__ __forceinline__ uint32_t cell_hash(int cx, int cy, int cz, uint32_t table_size) {
    return (((uint32_t)cx * 73856093u) ^ ((uint32_t)cy * 19349663u) ^ ((uint32_t)cz * 83492791u)) & (table_size - 1u);
}

// agent_cell (sim.slang:31): the hash of int3(floor(position / cell)); 64-byte records.
__global__ __launch_bounds__(ABLK) void k_grid_keys(const float *agents, uint32_t n, float cell, uint32_t table_size,
                                                    uint32_t *keys, uint32_t *vals) {
    const uint32_t i = blockIdx.x * ABLK + threadIdx.x;
    if (i >= n) return;
    const float4 p = reinterpret_cast<const float4 *>(agents)[4ull * i];
    keys[i] = cell_hash((int)floorf(p.x / cell), (int)floorf(p.y / cell), (int)floorf(p.z / cell), table_size);
    vals[i] = i;
}

// Digit histogram of one RTILE-element tile (elements base + m * ABLK + t), stored
// digit-major: hist[d * ntiles + tile].
__global__ __launch_bounds__(ABLK) void k_radix_hist(const uint32_t *keys, uint32_t n, uint32_t shift, uint32_t ntiles,
                                                     uint32_t *hist) {
    __shared__ uint32_t h[RADIX];
    for (uint32_t q = threadIdx.x; q < RADIX; q += ABLK) h[q] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * RTILE;
#pragma unroll
    for (uint32_t m = 0; m < RITEMS; m++) {
        const uint32_t i = base + m * ABLK + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & (RADIX - 1u)], 1u);
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < RADIX; d += ABLK) hist[(size_t)d * ntiles + blockIdx.x] = h[d];
}

// Stable scatter of one tile: an element's rank among the tile's elements of its digit that
// come before it (order m, then t), plus the digit's scanned start for this tile.  Within a
// wave, lanes with equal digits are found with 8 ballots; waves then take their places in
// wave order through per-digit LDS counters.
__global__ __launch_bounds__(ABLK) void k_radix_scatter(const uint32_t *keys, const uint32_t *vals, uint32_t n,
                                                        uint32_t shift, uint32_t ntiles, const uint32_t *hist_off,
                                                        uint32_t *keys_out, uint32_t *vals_out) {
    __shared__ uint32_t run[RADIX];           // elements of each digit placed so far in this tile
    __shared__ uint32_t wcnt[ANW][RADIX];     // this item's per-wave digit counts
    __shared__ uint32_t start[RADIX];
    const uint32_t t = threadIdx.x, lane = t & (WAVE - 1u), w = t / WAVE;
    for (uint32_t d = t; d < RADIX; d += ABLK) {
        run[d] = 0;
        start[d] = hist_off[(size_t)d * ntiles + blockIdx.x];
    }
    const uint32_t base = blockIdx.x * RTILE;
    const uint64_t lt = (1ull << lane) - 1ull;
    for (uint32_t m = 0; m < RITEMS; m++) {
        for (uint32_t d = t; d < RADIX; d += ABLK)
#pragma unroll
            for (uint32_t q = 0; q < ANW; q++) wcnt[q][d] = 0;
        __syncthreads();
        const uint32_t i = base + m * ABLK + t;
        const bool valid = i < n;
        const uint32_t key = valid ? keys[i] : 0u, val = valid ? vals[i] : 0u;
        const uint32_t digit = (key >> shift) & (RADIX - 1u);
        uint64_t same = __ballot(valid);
#pragma unroll
        for (uint32_t b = 0; b < 8; b++) {
            const uint64_t ones = __ballot(valid && ((digit >> b) & 1u));
            same &= ((digit >> b) & 1u) ? ones : ~ones;
        }
        const uint32_t before = (uint32_t)__popcll(same & lt);
        if (valid && before == 0) wcnt[w][digit] = (uint32_t)__popcll(same);  // the digit's first lane
        __syncthreads();
        if (valid) {
            uint32_t r = run[digit] + before;
            for (uint32_t q = 0; q < w; q++) r += wcnt[q][digit];
            keys_out[start[digit] + r] = key;
            vals_out[start[digit] + r] = val;
        }
        __syncthreads();
        for (uint32_t d = t; d < RADIX; d += ABLK) {
            uint32_t s = 0;
#pragma unroll
            for (uint32_t q = 0; q < ANW; q++) s += wcnt[q][d];
            run[d] += s;
        }
    }
}

// counts[h] = the number of sorted keys equal to h: one atomic per run of equal keys in a wave.
__global__ __launch_bounds__(ABLK) void k_grid_counts(const uint32_t *keys, uint32_t n, uint32_t *counts) {
    const uint32_t i = blockIdx.x * ABLK + threadIdx.x;
    const uint32_t lane = threadIdx.x & (WAVE - 1u);
    const bool valid = i < n;
    const uint32_t key = valid ? keys[i] : 0xFFFFFFFFu;
    const uint32_t prev = __shfl_up(key, 1, WAVE);
    const bool head = valid && (lane == 0 || prev != key);
    const uint64_t heads = __ballot(head), live = __ballot(valid);
    if (!head) return;
    const uint64_t after = heads & ~((2ull << lane) - 1ull);
    const uint32_t end = after ? (uint32_t)(__ffsll((long long)after) - 1) : (uint32_t)(64 - __clzll(live));
    atomicAdd(&counts[key], end - lane);
}

}  // namespace

// Device scratch of one build: radix ping-pong buffers, tile histograms, scan scratch.
struct GridScratch {
    uint32_t *keys[2] = {nullptr, nullptr}, *vals[2] = {nullptr, nullptr};
    uint32_t *hist = nullptr, *scan_tmp = nullptr;

    // Scan scratch covers the larger of the histograms and the largest table (2^23 entries).
    static uint32_t scan_len(uint32_t max_agents) {
        const uint32_t hist_len = RADIX * cdiv(max_agents ? max_agents : 1u, RTILE);
        return std::max<uint32_t>(hist_len, 1u << 23);
    }
    // Bytes of each array for up to max_agents agents, in layout order, and their 256-byte
    // aligned total.
    static constexpr int PARTS = 6;
    static void sizes(uint32_t max_agents, size_t out[PARTS]) {
        const size_t n = std::max<uint32_t>(max_agents, 1u);
        const size_t hist_len = (size_t)RADIX * cdiv(n, RTILE);
        const size_t tmp = 4ull * (scan_tiles(scan_len(max_agents)) + 1);
        const size_t sz[PARTS] = {4 * n, 4 * n, 4 * n, 4 * n, 4 * hist_len, tmp};
        for (int q = 0; q < PARTS; q++) out[q] = sz[q];
    }
    static size_t total(uint32_t max_agents) {
        size_t sz[PARTS], t = 0;
        sizes(max_agents, sz);
        for (size_t b : sz) t += (b + 255) / 256 * 256;
        return t;
    }
    // Lays the arrays out from `base` (256-byte aligned, total(max_agents) bytes).
    void layout(void *base, uint32_t max_agents) {
        size_t sz[PARTS];
        sizes(max_agents, sz);
        char *cur = static_cast<char *>(base);
        uint32_t **dst[PARTS] = {&keys[0], &vals[0], &keys[1], &vals[1], &hist, &scan_tmp};
        for (int q = 0; q < PARTS; q++) {
            *dst[q] = reinterpret_cast<uint32_t *>(cur);
            cur += (sz[q] + 255) / 256 * 256;
        }
    }
};

// grid_clear .. grid_scatter (sim.slang:31-104) into device arrays, on `stream`, using `w`.
void enqueue_grid_build(const GridScratch &w, const float *agents, uint32_t n, float cell_size, uint32_t table_size,
                        uint32_t *offsets, uint32_t *counts, uint32_t *sorted, hipStream_t stream) {
    const float cell = cell_size > 1e-6f ? cell_size : 1e-6f;  // sim.rs:182
    launch_fill_u32(counts, table_size, 0u, stream);
    if (n) {
        PXL_LAUNCH(k_grid_keys, cdiv(n, ABLK), ABLK, 0, stream, agents, n, cell, table_size, w.keys[0], w.vals[0]);
        uint32_t bits = 0;
        while ((1u << bits) < table_size) bits++;
        const uint32_t passes = std::max<uint32_t>(1u, (bits + 7) / 8), ntiles = cdiv(n, RTILE);
        int cur = 0;
        for (uint32_t p = 0; p < passes; p++) {
            const uint32_t shift = 8 * p;
            PXL_LAUNCH(k_radix_hist, ntiles, ABLK, 0, stream, w.keys[cur], n, shift, ntiles, w.hist);
            launch_exclusive_scan(w.hist, RADIX * ntiles, w.scan_tmp, w.hist, stream);
            uint32_t *vout = p + 1 == passes ? sorted : w.vals[cur ^ 1];
            PXL_LAUNCH(k_radix_scatter, ntiles, ABLK, 0, stream, w.keys[cur], w.vals[cur], n, shift, ntiles, w.hist,
                       w.keys[cur ^ 1], vout);
            cur ^= 1;
        }
        PXL_LAUNCH(k_grid_counts, cdiv(n, ABLK), ABLK, 0, stream, w.keys[cur], n, counts);
    }
    launch_exclusive_scan(counts, table_size, w.scan_tmp, offsets, stream);
}

// One neighbour grid handle: its own stream, events and scratch for up to max_agents agents.
struct AgentGrid {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t max_agents = 0;
    void *arena = nullptr;
    GridScratch w;
    hipEvent_t ev[2] = {nullptr, nullptr};
    float last_ms = 0.0f;
    std::string error;

    void allocate() {
        if (hipMalloc(&arena, GridScratch::total(max_agents)) != hipSuccess) {
            arena = nullptr;
            throw std::bad_alloc();
        }
        w.layout(arena, max_agents);
    }
    ~AgentGrid() {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        if (arena) (void)hipFree(arena);
        if (stream) (void)hipStreamDestroy(stream);
    }

    void build(const float *agents, uint32_t n, float cell_size, uint32_t table_size, uint32_t *offsets,
               uint32_t *counts, uint32_t *sorted) {
        HIP_CHECK(hipEventRecord(ev[0], stream));
        enqueue_grid_build(w, agents, n, cell_size, table_size, offsets, counts, sorted, stream);
        HIP_CHECK(hipEventRecord(ev[1], stream));
        HIP_CHECK(hipStreamSynchronize(stream));
        HIP_CHECK(hipEventElapsedTime(&last_ms, ev[0], ev[1]));
    }
};

// ---- the neighbour grid without allocation (pxlgrid_build_static)
//
// Its scratch is one file-scope device array, sized for PXLGRID_STATIC_MAX_AGENTS agents:
// 4 radix arrays of 4 B per agent, the digit histograms of 256 counters per 2048-agent tile,
// and the scan scratch of the largest table (2^23 entries, 2048 scan tiles).  Every GPU the
// library runs on holds its own copy until the process ends.  Nothing in it needs clearing:
// each build writes every word it reads (keys and ids of its agents, the histograms of its
// tiles, the scan partials) before reading it.
constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t STATIC_SCRATCH_BYTES =
    4 * align256(4ull * PXLGRID_STATIC_MAX_AGENTS) +
    align256(4ull * RADIX * ((PXLGRID_STATIC_MAX_AGENTS + RTILE - 1) / RTILE)) + align256(4ull * (2048 + 1));
constexpr int MAX_STATIC_DEVICES = 64;

}  // namespace pxl

static __device__ __align__(256) unsigned char scratch[pxl::STATIC_SCRATCH_BYTES];

namespace pxl {
namespace {
std::mutex g_static_mu;
void *g_static_scratch[MAX_STATIC_DEVICES] = {};  // the array's address on each device

// The address of `scratch` on `device` (current), looked up once per device.  The array's name
// in host code is only a host-side stand-in; this address is what kernels must be given.
void *static_scratch(int device) {
    std::lock_guard<std::mutex> lk(g_static_mu);
    if (!g_static_scratch[device]) {
        void *p = nullptr;
        HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_static_scratch[device] = p;
    }
    return g_static_scratch[device];
}
}  // namespace
}  // namespace pxl

struct pxlgrid {
    pxl::AgentGrid g;
};

extern "C" {

uint32_t pxlgrid_table_size(uint32_t count) {
    uint64_t want = (uint64_t)count * 2, t = 1;
    while (t < want) t <<= 1;
    if (t < 1024) t = 1024;
    if (t > (1u << 23)) t = 1u << 23;
    return (uint32_t)t;
}

int pxlgrid_create(uint32_t max_agents, int device, pxlgrid **out) {
    if (!out) return PXLFLIP_ERR_INVALID;
    *out = nullptr;
    auto *h = new (std::nothrow) pxlgrid;
    if (!h) return PXLFLIP_ERR_OOM;
    try {
        HIP_CHECK(hipSetDevice(device));
        h->g.device = device;
        h->g.max_agents = max_agents;
        HIP_CHECK(hipStreamCreateWithFlags(&h->g.stream, hipStreamNonBlocking));
        HIP_CHECK(hipEventCreate(&h->g.ev[0]));
        HIP_CHECK(hipEventCreate(&h->g.ev[1]));
        h->g.allocate();
    } catch (const pxl::HipError &e) {
        pxl::g_create_error = std::string("HIP error in ") + e.what + ": " + hipGetErrorString(e.code);
        delete h;
        return PXLFLIP_ERR_DEVICE;
    } catch (const std::bad_alloc &) {
        pxl::g_create_error = "out of device memory";
        delete h;
        return PXLFLIP_ERR_OOM;
    }
    *out = h;
    return PXLFLIP_OK;
}

}  // extern "C"

namespace {
int grid_check(pxlgrid *h, uint32_t count, uint32_t table_size) {
    if (!h) return PXLFLIP_ERR_INVALID;
    if (count > h->g.max_agents) {
        h->g.error = "count " + std::to_string(count) + " exceeds the grid's max_agents " + std::to_string(h->g.max_agents);
        return PXLFLIP_ERR_INVALID;
    }
    if (table_size < 1 || (table_size & (table_size - 1)) != 0 || table_size > (1u << 23)) {
        h->g.error = "table_size must be a power of two <= 2^23";
        return PXLFLIP_ERR_INVALID;
    }
    return PXLFLIP_OK;
}
template <typename F>
int grid_guarded(pxlgrid *h, F &&body) {
    try {
        HIP_CHECK(hipSetDevice(h->g.device));
        body();
        return PXLFLIP_OK;
    } catch (const pxl::HipError &e) {
        h->g.error = std::string("HIP error in ") + e.what + ": " + hipGetErrorString(e.code);
        return PXLFLIP_ERR_DEVICE;
    } catch (const std::bad_alloc &) {
        h->g.error = "out of device memory";
        return PXLFLIP_ERR_OOM;
    }
}
}  // namespace

extern "C" {

int pxlgrid_build(pxlgrid *h, const float *agents, uint32_t count, float cell_size, uint32_t table_size,
                  uint32_t *offsets, uint32_t *counts, uint32_t *sorted) {
    int rc = grid_check(h, count, table_size);
    if (rc != PXLFLIP_OK) return rc;
    if ((count && (!agents || !sorted)) || !offsets || !counts) {
        h->g.error = "agents, offsets, counts and sorted must be device pointers";
        return PXLFLIP_ERR_INVALID;
    }
    if (reinterpret_cast<uintptr_t>(agents) % 16 != 0) {
        h->g.error = "agents must be 16-byte aligned (64-byte records)";
        return PXLFLIP_ERR_INVALID;
    }
    return grid_guarded(h, [&] { h->g.build(agents, count, cell_size, table_size, offsets, counts, sorted); });
}

int pxlgrid_build_host(pxlgrid *h, const float *agents, uint32_t count, float cell_size, uint32_t table_size,
                       uint32_t *offsets, uint32_t *counts, uint32_t *sorted) {
    int rc = grid_check(h, count, table_size);
    if (rc != PXLFLIP_OK) return rc;
    return grid_guarded(h, [&] {
        pxl::DeviceBuffer a(64ull * count), o(4ull * table_size), c(4ull * table_size), s(4ull * count);
        if (count) HIP_CHECK(hipMemcpy(a.p, agents, 64ull * count, hipMemcpyHostToDevice));
        h->g.build(static_cast<const float *>(a.p), count, cell_size, table_size, static_cast<uint32_t *>(o.p),
                   static_cast<uint32_t *>(c.p), static_cast<uint32_t *>(s.p));
        if (offsets) HIP_CHECK(hipMemcpy(offsets, o.p, 4ull * table_size, hipMemcpyDeviceToHost));
        if (counts) HIP_CHECK(hipMemcpy(counts, c.p, 4ull * table_size, hipMemcpyDeviceToHost));
        if (sorted && count) HIP_CHECK(hipMemcpy(sorted, s.p, 4ull * count, hipMemcpyDeviceToHost));
    });
}

int pxlgrid_build_static(int device, void *stream, const float *agents, uint32_t count, float cell_size,
                         uint32_t table_size, uint32_t *offsets, uint32_t *counts, uint32_t *sorted) {
    using namespace pxl;
    g_create_error.clear();
    if (count > PXLGRID_STATIC_MAX_AGENTS) {
        g_create_error = "count " + std::to_string(count) + " exceeds the static scratch's " +
                         std::to_string(PXLGRID_STATIC_MAX_AGENTS) + " agents (use pxlgrid_create / pxlgrid_build)";
        return PXLFLIP_ERR_INVALID;
    }
    if (table_size < 1 || (table_size & (table_size - 1)) != 0 || table_size > (1u << 23)) {
        g_create_error = "table_size must be a power of two <= 2^23";
        return PXLFLIP_ERR_INVALID;
    }
    if ((count && (!agents || !sorted)) || !offsets || !counts) {
        g_create_error = "agents, offsets, counts and sorted must be device pointers";
        return PXLFLIP_ERR_INVALID;
    }
    if (reinterpret_cast<uintptr_t>(agents) % 16 != 0) {
        g_create_error = "agents must be 16-byte aligned (64-byte records)";
        return PXLFLIP_ERR_INVALID;
    }
    if (device < 0 || device >= MAX_STATIC_DEVICES) {
        g_create_error = "device " + std::to_string(device) + " out of range";
        return PXLFLIP_ERR_INVALID;
    }
    static_assert(STATIC_SCRATCH_BYTES >= 4 * align256(4ull * PXLGRID_STATIC_MAX_AGENTS), "scratch layout");
    if (GridScratch::total(PXLGRID_STATIC_MAX_AGENTS) > STATIC_SCRATCH_BYTES) {
        g_create_error = "static scratch smaller than its layout";
        return PXLFLIP_ERR_INTERNAL;
    }
    try {
        HIP_CHECK(hipSetDevice(device));
        GridScratch w;
        w.layout(static_scratch(device), PXLGRID_STATIC_MAX_AGENTS);
        hipStream_t s = static_cast<hipStream_t>(stream);
        enqueue_grid_build(w, agents, count, cell_size, table_size, offsets, counts, sorted, s);
        HIP_CHECK(hipStreamSynchronize(s));
    } catch (const HipError &e) {
        g_create_error = std::string("HIP error in ") + e.what + ": " + hipGetErrorString(e.code);
        return PXLFLIP_ERR_DEVICE;
    }
    return PXLFLIP_OK;
}

double pxlgrid_last_ms(const pxlgrid *h) { return h ? (double)h->g.last_ms : 0.0; }

const char *pxlgrid_last_error(const pxlgrid *h) { return h ? h->g.error.c_str() : pxl::g_create_error.c_str(); }

void pxlgrid_destroy(pxlgrid *h) { delete h; }

}  // extern "C"
