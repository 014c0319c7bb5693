// Particle-side kernels of the FLIP step: counting-sort binning, P2G splat, G2P advection
// and particle IO.  Semantics follow shaders/fluid.slang (cited per kernel); the layout is
// SoA fp32 streams sorted by cell once per step, so that the splat and the gathers of
// neighbouring lanes touch neighbouring faces.
//
// Floating point: this file is compiled with -ffp-contract=off.  Every expression keeps the
// reference's operation order, so the results are bitwise those of the reference formulas
// evaluated in IEEE fp32 (the P2G sums are integers and therefore order-independent).
#include "common.hpp"
#include "kernels.hpp"

namespace pxl {

namespace {

constexpr int SCAN_ITEMS = 16;
constexpr int TILE_BLOCK = 1024;  // lanes of the per-tile P2G / G2P workgroups
constexpr int SCAN_TILE = BLOCK * SCAN_ITEMS;  // 4096 cells per scan workgroup

inline unsigned blocks_for(uint64_t n) { return (unsigned)((n + BLOCK - 1) / BLOCK); }

// Cell of a position: clamp(floor((x - origin) * inv_h), 0, n - 1) (fluid.slang:113).
__device__ __forceinline__ void cell_of_pos(const DevSim &S, float x, float y, float z, int &cx, int &cy, int &cz) {
    cx = (int)floorf((x - S.ox) * S.inv_h);
    cy = (int)floorf((y - S.oy) * S.inv_h);
    cz = (int)floorf((z - S.oz) * S.inv_h);
    cx = cx < 0 ? 0 : (cx > S.nx - 1 ? S.nx - 1 : cx);
    cy = cy < 0 ? 0 : (cy > S.ny - 1 ? S.ny - 1 : cy);
    cz = cz < 0 ? 0 : (cz > S.nz - 1 ? S.nz - 1 : cz);
}

// ---------------------------------------------------------------- clear / binning

__global__ __launch_bounds__(BLOCK) void k_fill_u32(uint32_t *p, uint32_t n, uint32_t value) {
    uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) p[i] = value;
}

// Workgroups of the per-tile kernels, each walking the active tile list with a stride of
// this count.
__host__ __device__ inline uint32_t tile_grid(const DevSim &S) { return S.tile_wgs; }

// clear_grid (fluid.slang:88-91).  The P2G sums live in per-tile boxes that P2G overwrites
// whole, so only the cell counts need clearing, and only in the tiles the previous step's
// tile list holds: every other count is still zero (allocation zeroes them, and a tile gets
// counts only in a step that lists it).  Workgroup b of nb takes listed tiles b, b + nb, ...;
// the list's count is reset by bin_count, after every workgroup here has read it.  Also
// zeroes the fluid list's row counts, which the cell pass accumulates.
__device__ __forceinline__ void body_clear_grid(const DevSim &S, uint32_t b, uint32_t nb) {
    const uint32_t n_prev = *S.tile_count;
    for (uint32_t u = b; u < n_prev; u += nb) {
        uint32_t *c = S.count + (size_t)S.tile_list[u] * TILE_CELLS;
        for (uint32_t q = threadIdx.x; q < TILE_CELLS; q += BLOCK) c[q] = 0;
    }
    for (uint32_t r = b * BLOCK + threadIdx.x; r < fl_row_slots(S); r += nb * BLOCK) S.fl_rows[r] = 0;
}
__global__ __launch_bounds__(BLOCK) void k_clear_grid(DevSim S) { body_clear_grid(S, blockIdx.x, gridDim.x); }
__global__ __launch_bounds__(BLOCK) void kb_clear_grid(WorkList W) {
    for_each_work(W, [&](const DevSim &S, uint32_t b, const uint32_t *) { body_clear_grid(S, b, tile_grid(S)); });
}

// Counting-sort pass 1: tiled cell key of each particle and its arrival slot in that cell
// (the per-cell count of p2g, fluid.slang:113-114).  Particles arrive in last step's cell
// order, so neighbouring lanes mostly share a key: each run of equal keys in a wave64 takes
// its slots with ONE atomic by the run's first lane.
__device__ __forceinline__ void body_bin_count(const DevSim &S, uint32_t b) {
    const uint32_t i = b * BLOCK + threadIdx.x;
    const bool valid = i < S.np;
    if (i == 0) {  // the tile and band lists are rebuilt later in the step (clear_grid has read them)
        *S.tile_count = 0;
        *S.band_count = 0;
        if (S.band_all && S.member == 0) *S.band_all_count = 0;
    }
    uint32_t key = 0xFFFFFFFFu;
    if (valid) {
        int cx, cy, cz;
        cell_of_pos(S, S.ax[i], S.ay[i], S.az[i], cx, cy, cz);
        key = tile_key(S, cx, cy, cz);
    }
    const int lane = threadIdx.x & (WAVE - 1);
    const uint32_t prev = __shfl_up(key, 1, WAVE);
    const bool head = lane == 0 || prev != key;
    const uint64_t heads = __ballot(head);
    const uint64_t le = lane == WAVE - 1 ? ~0ull : ((2ull << lane) - 1ull);
    const int leader = 63 - __clzll(heads & le);
    const uint64_t after = heads & ~le;
    const int next = after ? __ffsll((long long)after) - 1 : WAVE;
    uint32_t base = 0;
    if (valid && lane == leader) base = atomicAdd(&S.count[key], (uint32_t)(next - leader));
    base = __shfl(base, leader, WAVE);
    if (valid) {
        S.key[i] = key;
        S.slot[i] = base + (uint32_t)(lane - leader);
    }
}
__global__ __launch_bounds__(BLOCK) void k_bin_count(DevSim S) { body_bin_count(S, blockIdx.x); }
__global__ __launch_bounds__(BLOCK) void kb_bin_count(WorkList W) {
    for_each_work(W, [&](const DevSim &S, uint32_t b, const uint32_t *) { body_bin_count(S, b); });
}

// Counting-sort pass 2: exclusive scan of the counts (reduce / scan of chunk sums / down-sweep).
__device__ __forceinline__ void body_scan_reduce(const uint32_t *in, uint32_t n, uint32_t *tile_sum, uint32_t b) {
    __shared__ uint32_t sh[4];
    const uint32_t base = b * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    uint32_t s = 0;
    if (base + SCAN_ITEMS <= n) {
        const uint4 *p = reinterpret_cast<const uint4 *>(in + base);
#pragma unroll
        for (int m = 0; m < SCAN_ITEMS / 4; m++) {
            uint4 q = p[m];
            s += q.x + q.y + q.z + q.w;
        }
    } else {
        for (int m = 0; m < SCAN_ITEMS; m++)
            if (base + m < n) s += in[base + m];
    }
    s = block_sum(s, sh);
    if (threadIdx.x == 0) tile_sum[b] = s;
}

// Single workgroup: exclusive scan of the tile sums in place (any count, in chunks).
__device__ __forceinline__ void body_scan_tiles(uint32_t *tile_sum, uint32_t ntiles) {
    __shared__ uint32_t sh[4];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < ntiles; base += BLOCK * 4) {
        uint32_t v[4], t = 0;
        for (int m = 0; m < 4; m++) {
            uint32_t idx = base + threadIdx.x * 4 + m;
            v[m] = idx < ntiles ? tile_sum[idx] : 0;
            t += v[m];
        }
        uint32_t total;
        uint32_t off = block_exclusive_scan(t, &total, sh) + carry;
        for (int m = 0; m < 4; m++) {
            uint32_t idx = base + threadIdx.x * 4 + m;
            if (idx < ntiles) tile_sum[idx] = off;
            off += v[m];
        }
        carry += total;
    }
}

__device__ __forceinline__ void body_scan_down(const uint32_t *in, uint32_t n, const uint32_t *tile_off,
                                               uint32_t *out, uint32_t b) {
    __shared__ uint32_t sh[4];
    const uint32_t base = b * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t t = 0;
    const bool full = base + SCAN_ITEMS <= n;
    if (full) {
        const uint4 *p = reinterpret_cast<const uint4 *>(in + base);
#pragma unroll
        for (int m = 0; m < SCAN_ITEMS / 4; m++) {
            uint4 q = p[m];
            v[4 * m] = q.x; v[4 * m + 1] = q.y; v[4 * m + 2] = q.z; v[4 * m + 3] = q.w;
        }
    } else {
#pragma unroll
        for (int m = 0; m < SCAN_ITEMS; m++) v[m] = base + m < n ? in[base + m] : 0;
    }
#pragma unroll
    for (int m = 0; m < SCAN_ITEMS; m++) t += v[m];
    uint32_t total;
    uint32_t run = block_exclusive_scan(t, &total, sh) + tile_off[b];
    if (full) {
        uint4 *p = reinterpret_cast<uint4 *>(out + base);
#pragma unroll
        for (int m = 0; m < SCAN_ITEMS / 4; m++) {
            uint4 q;
            q.x = run; run += v[4 * m];
            q.y = run; run += v[4 * m + 1];
            q.z = run; run += v[4 * m + 2];
            q.w = run; run += v[4 * m + 3];
            p[m] = q;
        }
    } else {
        for (int m = 0; m < SCAN_ITEMS; m++) {
            if (base + m < n) out[base + m] = run;
            run += v[m];
        }
    }
}

__global__ __launch_bounds__(BLOCK) void k_scan_reduce(const uint32_t *in, uint32_t n, uint32_t *tile_sum) {
    body_scan_reduce(in, n, tile_sum, blockIdx.x);
}
__global__ __launch_bounds__(BLOCK) void k_scan_tiles(uint32_t *tile_sum, uint32_t ntiles) {
    body_scan_tiles(tile_sum, ntiles);
}
__global__ __launch_bounds__(BLOCK) void k_scan_down(const uint32_t *in, uint32_t n, const uint32_t *tile_off,
                                                     uint32_t *out) {
    body_scan_down(in, n, tile_off, out, blockIdx.x);
}

// The two scans of a step: count -> offset (which = 0) and the fluid list's row counts ->
// the rows' first list slots (1).
struct ScanIO {
    const uint32_t *in;
    uint32_t n;
    uint32_t *tmp, *out;
};
__device__ __forceinline__ ScanIO scan_io(const DevSim &S, int which) {
    if (which == 0) return ScanIO{S.count, S.nkeys, S.scan_tmp, S.offset};
    return ScanIO{S.fl_rows, fl_row_slots(S), S.fl_tmp, S.fl_rowoff};
}
__global__ __launch_bounds__(BLOCK) void kb_scan_reduce(WorkList W, int which) {
    for_each_work(W, [&](const DevSim &S, uint32_t b, const uint32_t *) {
        const ScanIO io = scan_io(S, which);
        body_scan_reduce(io.in, io.n, io.tmp, b);
    });
}
__global__ __launch_bounds__(BLOCK) void kb_scan_tiles(WorkList W, int which) {
    for_each_work(W, [&](const DevSim &S, uint32_t, const uint32_t *) {
        const ScanIO io = scan_io(S, which);
        body_scan_tiles(io.tmp, (io.n + SCAN_TILE - 1) / SCAN_TILE);
    });
}
__global__ __launch_bounds__(BLOCK) void kb_scan_down(WorkList W, int which) {
    for_each_work(W, [&](const DevSim &S, uint32_t b, const uint32_t *) {
        const ScanIO io = scan_io(S, which);
        body_scan_down(io.in, io.n, io.tmp, io.out, b);
    });
}

// Counting-sort pass 3: move each particle to offset[cell] + slot (current -> scratch).
__device__ __forceinline__ void body_scatter(const DevSim &S, uint32_t b) {
    uint32_t i = b * BLOCK + threadIdx.x;
    if (i >= S.np) return;
    uint32_t j = S.offset[S.key[i]] + S.slot[i];
    if (j >= S.np) return;  // only if the counts were not cleared (a broken invariant): stay in bounds
    S.bx[j] = S.ax[i];
    S.by[j] = S.ay[i];
    S.bz[j] = S.az[i];
    S.bvx[j] = S.avx[i];
    S.bvy[j] = S.avy[i];
    S.bvz[j] = S.avz[i];
    S.bid[j] = S.aid[i];
}
__global__ __launch_bounds__(BLOCK) void k_scatter(DevSim S) { body_scatter(S, xcd_major(blockIdx.x, gridDim.x)); }
__global__ __launch_bounds__(BLOCK) void kb_scatter(WorkList W) {
    for_each_work<true>(W, [&](const DevSim &S, uint32_t b, const uint32_t *) { body_scatter(S, b); });
}

// ---------------------------------------------------------------- P2G / G2P

__device__ __forceinline__ uint32_t face_base(const DevSim &S, int c) {
    return c == 0 ? 0u : (c == 1 ? S.nu : S.nu + S.nv);
}

__device__ __forceinline__ uint32_t face_index(const DevSim &S, int c, int qx, int qy, int qz) {
    const uint32_t dx = (uint32_t)S.nx + (c == 0), dy = (uint32_t)S.ny + (c == 1);
    return face_base(S, c) + (uint32_t)qx + dx * ((uint32_t)qy + dy * (uint32_t)qz);
}

__device__ __forceinline__ void tile_coords(const DevSim &S, uint32_t t, int &ti, int &tj, int &tk) {
    ti = (int)(t % (uint32_t)S.ntx);
    tj = (int)((t / (uint32_t)S.ntx) % (uint32_t)S.nty);
    tk = (int)(t / ((uint32_t)S.ntx * (uint32_t)S.nty));
}

// Face-box geometry of component c (compile-time after unrolling): extents d0 x d1 x d2 and
// the LDS offset of trilinear corner k (k = ox + 2 oy + 4 oz) from corner 0.
__device__ __forceinline__ int box_d0(int c) { return c == 0 ? TILE + 1 : TILE + 2; }
__device__ __forceinline__ int box_d1(int c) { return c == 1 ? TILE + 1 : TILE + 2; }
__device__ __forceinline__ int box_corner(int c, int k) {
    return (k & 1) + box_d0(c) * (((k >> 1) & 1) + box_d1(c) * (k >> 2));
}

// Active tiles (tiles holding particles) for the per-tile kernels; list order is arbitrary
// (the per-tile results do not depend on it).
__device__ __forceinline__ void body_tile_list(const DevSim &S, uint32_t b) {
    const uint32_t t = b * BLOCK + threadIdx.x;
    if (t >= S.ntiles) return;
    const bool active = S.offset[t * TILE_CELLS] != S.offset[(t + 1) * TILE_CELLS];
    S.tile_act[t] = (uint8_t)(((S.tile_act[t] & 1u) << 1) | (active ? 1u : 0u));
    if (active) S.tile_list[atomicAdd(S.tile_count, 1u)] = t;
}
__global__ __launch_bounds__(BLOCK) void k_tile_list(DevSim S) { body_tile_list(S, blockIdx.x); }

// Band: tiles with an active tile (this step or the last) among their 3x3x3 neighbours.  A
// wave appends its band tiles in tile order with one atomic, so x-neighbouring tiles are
// mostly adjacent in the list (the band passes take contiguous runs of it).
__device__ __forceinline__ void body_band(const DevSim &S, uint32_t b) {
    const uint32_t t = b * BLOCK + threadIdx.x;
    uint32_t any = 0;
    if (t < S.ntiles) {
        const int ti = (int)(t % (uint32_t)S.ntx), tj = (int)((t / (uint32_t)S.ntx) % (uint32_t)S.nty),
                  tk = (int)(t / ((uint32_t)S.ntx * (uint32_t)S.nty));
        for (int dz = -1; dz <= 1; dz++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    const int x = ti + dx, y = tj + dy, z = tk + dz;
                    if (x < 0 || y < 0 || z < 0 || x >= S.ntx || y >= S.nty || z >= S.ntz) continue;
                    any |= S.tile_act[(uint32_t)x + (uint32_t)S.ntx * ((uint32_t)y + (uint32_t)S.nty * (uint32_t)z)];
                }
    }
    const int lane = threadIdx.x & (WAVE - 1);
    const uint64_t mask = __ballot(any != 0);
    if (!mask) return;
    const int leader = __ffsll((long long)mask) - 1;
    const uint32_t rank = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    uint32_t base = 0, all = 0;
    if (lane == leader) {
        base = atomicAdd(S.band_count, (uint32_t)__popcll(mask));
        if (S.band_all) all = atomicAdd(S.band_all_count, (uint32_t)__popcll(mask));
    }
    base = __shfl(base, leader, WAVE);
    all = __shfl(all, leader, WAVE);
    if (any) {
        S.band_list[base + rank] = t;
        if (S.band_all) S.band_all[all + rank] = (S.member << WORK_SIM_SHIFT) | t;
    }
}
__global__ __launch_bounds__(BLOCK) void k_band(DevSim S) { body_band(S, blockIdx.x); }
__global__ __launch_bounds__(BLOCK) void kb_band(WorkList W) {
    for_each_work(W, [&](const DevSim &S, uint32_t b, const uint32_t *) { body_band(S, b); });
}
__global__ __launch_bounds__(BLOCK) void kb_tile_list(WorkList W) {
    for_each_work(W, [&](const DevSim &S, uint32_t b, const uint32_t *) { body_tile_list(S, b); });
}

// p2g (fluid.slang:94-112), one active 8^3 tile at a time per workgroup.  The tile's
// particles are one contiguous run of the cell-sorted scratch copy; their fixed-point splats
// (acc and wacc packed into one int64) are summed with LDS atomics into the tile's face box,
// and the box is stored whole to HBM with plain stores.  The face pass adds the boxes that
// cover each face, so no global atomics are needed.  Integer sums: the order of additions
// does not matter.  Corners outside the face grid are skipped as in the reference; every
// in-grid corner of a particle binned in the tile lies in the box (the box spans the cells
// [8t-1, 8t+8] around the tile; positions beyond the grid are binned to its edge cells, whose
// corners beyond the grid are skipped).
template <int NT>
__device__ __forceinline__ void body_p2g(const DevSim &S, uint32_t b0, uint32_t nb) {
    __shared__ unsigned long long lbox[BOX_ENTRIES];
    const uint32_t n_active = *S.tile_count;
    // lanes of a wave take every 8th particle of a run, so concurrent LDS atomics land in
    // different cells
    const uint32_t local = (threadIdx.x & (NT / 8 - 1)) * 8u + (threadIdx.x / (NT / 8));
    const float o[3] = {S.ox, S.oy, S.oz};
    for (uint32_t u = b0; u < n_active; u += nb) {
        const uint32_t t = S.tile_list[u];
        const uint32_t beg = S.offset[t * TILE_CELLS], end = S.offset[(t + 1) * TILE_CELLS];
        int ti, tj, tk;
        tile_coords(S, t, ti, tj, tk);
        for (int q = threadIdx.x; q < BOX_ENTRIES; q += NT) lbox[q] = 0ull;
        __syncthreads();
        // software-pipelined: the next round's particle is loaded before this round's LDS
        // atomics (a tile that is listed holds particles, so index beg is valid for a lane past
        // the end, whose values are not used)
        uint32_t p = beg + local;
        float xn[3], vn[3];
        {
            const uint32_t pl = p < end ? p : beg;
            xn[0] = S.bx[pl], xn[1] = S.by[pl], xn[2] = S.bz[pl];
            vn[0] = S.bvx[pl], vn[1] = S.bvy[pl], vn[2] = S.bvz[pl];
        }
        for (uint32_t base = beg; base < end; base += NT, p += NT) {
            const float x[3] = {xn[0], xn[1], xn[2]};
            const float v[3] = {vn[0], vn[1], vn[2]};
            {
                const uint32_t pn = p + NT, pl = pn < end ? pn : beg;
                xn[0] = S.bx[pl], xn[1] = S.by[pl], xn[2] = S.bz[pl];
                vn[0] = S.bvx[pl], vn[1] = S.bvy[pl], vn[2] = S.bvz[pl];
            }
            if (p >= end) continue;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                int dim[3] = {S.nx, S.ny, S.nz};
                dim[c] += 1;
                int lo[3], d[3];
                tile_box(c, ti, tj, tk, lo, d);
                float g[3], wl[3], wh[3];
                int b[3];
                bool ok[3][2];
#pragma unroll
                for (int a = 0; a < 3; a++) g[a] = (x[a] - o[a]) * S.inv_h - 0.5f;
                g[c] = g[c] + 0.5f;
                bool in_box = true;
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    b[a] = (int)floorf(g[a]);
                    const float f = g[a] - (float)b[a];
                    wl[a] = 1.0f - f;
                    wh[a] = f;
                    ok[a][0] = b[a] >= 0 && b[a] <= dim[a] - 1;
                    ok[a][1] = b[a] + 1 >= 0 && b[a] + 1 <= dim[a] - 1;
                    in_box = in_box && b[a] >= lo[a] && b[a] + 1 < lo[a] + d[a];
                }
                if (!in_box) continue;  // unreachable for in-grid corners (see above)
                const int li = c * BOX_FACES + (b[0] - lo[0]) + d[0] * ((b[1] - lo[1]) + d[1] * (b[2] - lo[2]));
                const float wxy[4] = {wl[0] * wl[1], wh[0] * wl[1], wl[0] * wh[1], wh[0] * wh[1]};
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    const int ox = k & 1, oy = (k >> 1) & 1, oz = k >> 2;
                    if (!(ok[0][ox] && ok[1][oy] && ok[2][oz])) continue;
                    const float w = wxy[k & 3] * (oz ? wh[2] : wl[2]);
                    const int a = (int)(w * v[c] * 65536.0f), wi = (int)(w * 65536.0f);
                    const unsigned long long packed = ((unsigned long long)(uint32_t)a << 32) + (unsigned long long)(uint32_t)wi;
                    atomicAdd(&lbox[li + box_corner(c, k)], packed);
                }
            }
        }
        __syncthreads();
        // stored in the HBM box order (box_gidx): entry g of the store comes from LDS entry
        // c * BOX_FACES + lx + d0 (ly + d1 lz)
        int64_t *out = S.box + (size_t)t * BOX_ENTRIES;
        for (int g = threadIdx.x; g < BOX_ENTRIES; g += NT) {
            const int c = g / BOX_FACES, r = g - c * BOX_FACES;
            const int d0 = box_d0(c), d1 = box_d1(c), d2 = c == 2 ? TILE + 1 : TILE + 2;
            int lx, yz;
            if (r < TILE * d1 * d2) {
                lx = (c == 0 ? 0 : 1) + r % TILE;
                yz = r / TILE;
            } else {
                const int h = (r - TILE * d1 * d2) / (d1 * d2);
                lx = c == 0 ? TILE : (h ? TILE + 1 : 0);
                yz = (r - TILE * d1 * d2) - h * d1 * d2;
            }
            out[g] = (int64_t)lbox[c * BOX_FACES + lx + d0 * yz];
        }
        __syncthreads();
    }
}
__global__ __launch_bounds__(TILE_BLOCK) void k_p2g_tile(DevSim S) { body_p2g<TILE_BLOCK>(S, blockIdx.x, gridDim.x); }
template <int NT>
__global__ __launch_bounds__(NT) void kb_p2g_tile(WorkList W) {
    for_each_work(W, [&](const DevSim &S, uint32_t b, const uint32_t *) { body_p2g<NT>(S, b, tile_grid(S)); });
}

// Trilinear stencil of sample_face (fluid.slang:67-80) for component c at one position:
// the eight corner weights w[k] = (wx * wy) * wz (k = ox + 2 oy + 4 oz, the reference's
// product order), the clamped corner coordinates, and -- when all corners lie in the tile's
// face box -- the LDS index of corner 0 and the index steps along x, y, z (0 where a
// clamp folds both corners together).
struct Stencil {
    float w[8];
    int q0[3], q1[3];
    int li, sx, sy, sz;
    bool in_box;
    bool fast;  // no clamping and in the box: corner k is at li + box_corner(c, k)
};

__device__ __forceinline__ void make_stencil(const DevSim &S, int c, int ti, int tj, int tk, float x, float y,
                                             float z, Stencil &st) {
    int dim[3] = {S.nx, S.ny, S.nz};
    dim[c] += 1;
    float g[3] = {(x - S.ox) * S.inv_h - 0.5f, (y - S.oy) * S.inv_h - 0.5f, (z - S.oz) * S.inv_h - 0.5f};
    g[c] = g[c] + 0.5f;
    float wl[3], wh[3];
    int lo[3], d[3];
    tile_box(c, ti, tj, tk, lo, d);
    st.in_box = true;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const int b = (int)floorf(g[a]);
        const float f = g[a] - (float)b;
        wl[a] = 1.0f - f;
        wh[a] = f;
        st.q0[a] = b < 0 ? 0 : (b > dim[a] - 1 ? dim[a] - 1 : b);
        st.q1[a] = b + 1 < 0 ? 0 : (b + 1 > dim[a] - 1 ? dim[a] - 1 : b + 1);
        st.in_box = st.in_box && st.q0[a] >= lo[a] && st.q1[a] < lo[a] + d[a];
    }
    const float wxy[4] = {wl[0] * wl[1], wh[0] * wl[1], wl[0] * wh[1], wh[0] * wh[1]};
#pragma unroll
    for (int k = 0; k < 8; k++) st.w[k] = wxy[k & 3] * ((k >> 2) ? wh[2] : wl[2]);
    st.li = (st.q0[0] - lo[0]) + d[0] * ((st.q0[1] - lo[1]) + d[1] * (st.q0[2] - lo[2]));
    st.fast = st.in_box && st.q1[0] == st.q0[0] + 1 && st.q1[1] == st.q0[1] + 1 && st.q1[2] == st.q0[2] + 1;
    st.sx = st.q1[0] - st.q0[0];
    st.sy = (st.q1[1] - st.q0[1]) * d[0];
    st.sz = (st.q1[2] - st.q0[2]) * d[0] * d[1];
}

// s = sum_k w[k] * value[k], corners in order k = 0..7 (the reference's accumulation).
__device__ __forceinline__ float sample_lds_fast(int c, const Stencil &st, const float *__restrict__ lds) {
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; k++) s = s + st.w[k] * lds[st.li + box_corner(c, k)];
    return s;
}

__device__ __forceinline__ float sample_lds(const Stencil &st, const float *__restrict__ lds) {
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; k++)
        s = s + st.w[k] * lds[st.li + (k & 1) * st.sx + ((k >> 1) & 1) * st.sy + (k >> 2) * st.sz];
    return s;
}

__device__ __forceinline__ float sample_global(const DevSim &S, int c, const Stencil &st, const float *__restrict__ buf) {
    const uint32_t dx = (uint32_t)S.nx + (c == 0), dy = (uint32_t)S.ny + (c == 1);
    const uint32_t base = face_base(S, c);
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int qx = (k & 1) ? st.q1[0] : st.q0[0];
        const int qy = ((k >> 1) & 1) ? st.q1[1] : st.q0[1];
        const int qz = (k >> 2) ? st.q1[2] : st.q0[2];
        s = s + st.w[k] * buf[base + (uint32_t)qx + dx * ((uint32_t)qy + dy * (uint32_t)qz)];
    }
    return s;
}

// Obstacle response (an extension), after advection and before the wall clamp: a particle
// whose binning cell is an obstacle cell moves along the ray from that obstacle's centre to
// distance r + h (its new cell's centre is then outside the sphere), and an inward normal
// velocity component is removed: d = x - c, len = sqrt(d.x^2 + d.y^2 + d.z^2), n = d / len,
// x = c + n (r + h), vn = v.n, v -= vn n when vn < 0.  Sums left to right; IEEE sqrt and
// division; len == 0 leaves the particle as it is.
__device__ __forceinline__ void obstacle_push(const DevSim &S, float x[3], float v[3]) {
    int ci, cj, ck;
    cell_of_pos(S, x[0], x[1], x[2], ci, cj, ck);
    const uint32_t id = (uint32_t)ci + (uint32_t)S.nx * ((uint32_t)cj + (uint32_t)S.ny * (uint32_t)ck);
    if (!((S.solid_bits[id >> 5] >> (id & 31u)) & 1u)) return;
    const uint32_t o = S.obst_id[id];
    if (o == 0u || o > S.n_obst) return;
    const float *sp = S.obst + 4u * (o - 1u);
    const float d[3] = {x[0] - sp[0], x[1] - sp[1], x[2] - sp[2]};
    const float len = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (!(len > 0.0f)) return;
    const float nr[3] = {d[0] / len, d[1] / len, d[2] / len};
    const float rad = sp[3] + S.h;
#pragma unroll
    for (int a = 0; a < 3; a++) x[a] = sp[a] + nr[a] * rad;
    const float vn = v[0] * nr[0] + v[1] * nr[1] + v[2] * nr[2];
    if (vn < 0.0f) {
#pragma unroll
        for (int a = 0; a < 3; a++) v[a] = v[a] - vn * nr[a];
    }
}

// g2p_advect (fluid.slang:200-220), one 8^3 tile per workgroup: the tile's vel and vel_old
// face boxes are staged in LDS, then every particle of the tile does the PIC/FLIP blend
// (lerp = a*(1-t) + b*t), RK2 through the projected grid velocity (the midpoint sample
// falls back to HBM when it leaves the box) and the wall clamp with the wall-normal velocity
// clipped (Slang max/min = SPIR-V FMax/FMin).  Reads the sorted scratch copy, writes the
// current particle arrays at the same (sorted) index.
template <int NT>
__device__ __forceinline__ void body_g2p(const DevSim &S, uint32_t b0, uint32_t nb) {
    // padded to whole fill rounds, so the fill below stores without a condition (a condition
    // lets the compiler sink each round's loads into it, one round trip at a time)
    constexpr int FILL_GROUP = 3, FILL_ROUND = FILL_GROUP * NT;
    constexpr int BOX_PAD = (BOX_ENTRIES + FILL_ROUND - 1) / FILL_ROUND * FILL_ROUND;
    __shared__ float lvel[BOX_PAD];
    __shared__ float lold[BOX_PAD];
    const uint32_t n_active = *S.tile_count;
    for (uint32_t u = b0; u < n_active; u += nb) {
    const uint32_t t = S.tile_list[u];
    const uint32_t beg = S.offset[t * TILE_CELLS], end = S.offset[(t + 1) * TILE_CELLS];
    int ti, tj, tk;
    tile_coords(S, t, ti, tj, tk);
    // the box fill, FILL_GROUP faces of a lane at a time: their loads (faces past the grid or
    // the box at index 0, dropped) are all issued before any LDS store, so they are in flight
    // together (all of a lane's faces at once needed 50% more registers than the rest of the
    // kernel)
    for (int q0 = (int)threadIdx.x; q0 < BOX_ENTRIES; q0 += FILL_ROUND) {
        float fa[FILL_GROUP], fo[FILL_GROUP];
#pragma unroll
        for (int m = 0; m < FILL_GROUP; m++) {
            const int q = q0 + m * NT, qq = q < BOX_ENTRIES ? q : 0;
            const int c = qq / BOX_FACES, r = qq - c * BOX_FACES;
            int lo[3], d[3];
            tile_box(c, ti, tj, tk, lo, d);
            const int qx = lo[0] + r % d[0], qy = lo[1] + (r / d[0]) % d[1], qz = lo[2] + r / (d[0] * d[1]);
            const int dx = S.nx + (c == 0), dy = S.ny + (c == 1), dz = S.nz + (c == 2);
            const bool ok = q < BOX_ENTRIES && qx >= 0 && qy >= 0 && qz >= 0 && qx < dx && qy < dy && qz < dz;
            const uint32_t fi = ok ? face_index(S, c, qx, qy, qz) : 0u;
            const float a = S.vel[fi], o = S.vel_old[fi];
            fa[m] = ok ? a : 0.0f;
            fo[m] = ok ? o : 0.0f;
        }
#pragma unroll
        for (int m = 0; m < FILL_GROUP; m++) {
            lvel[q0 + m * NT] = fa[m];  // past the box: the padding
            lold[q0 + m * NT] = fo[m];
        }
    }
    __syncthreads();
    const float tb = S.flip;
    const float half_dt = 0.5f * S.dt;
    const float org[3] = {S.ox, S.oy, S.oz};
    const float nf[3] = {(float)S.nx, (float)S.ny, (float)S.nz};
    for (uint32_t i = beg + threadIdx.x; i < end; i += NT) {
        float x[3] = {S.bx[i], S.by[i], S.bz[i]};
        const float v0[3] = {S.bvx[i], S.bvy[i], S.bvz[i]};
        float pic[3], old[3], v[3], mid[3], adv[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            Stencil st;
            make_stencil(S, c, ti, tj, tk, x[0], x[1], x[2], st);
            // separate paths so the LDS reads compile to ds_read, not generic loads
            if (st.fast) {
                pic[c] = sample_lds_fast(c, st, lvel + c * BOX_FACES);
                old[c] = sample_lds_fast(c, st, lold + c * BOX_FACES);
            } else if (st.in_box) {
                pic[c] = sample_lds(st, lvel + c * BOX_FACES);
                old[c] = sample_lds(st, lold + c * BOX_FACES);
            } else {
                pic[c] = sample_global(S, c, st, S.vel);
                old[c] = sample_global(S, c, st, S.vel_old);
            }
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float flipv = v0[c] + pic[c] - old[c];
            v[c] = pic[c] * (1.0f - tb) + flipv * tb;
            mid[c] = x[c] + half_dt * pic[c];
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            Stencil st;
            make_stencil(S, c, ti, tj, tk, mid[0], mid[1], mid[2], st);
            if (st.fast) adv[c] = sample_lds_fast(c, st, lvel + c * BOX_FACES);
            else if (st.in_box) adv[c] = sample_lds(st, lvel + c * BOX_FACES);
            else adv[c] = sample_global(S, c, st, S.vel);
        }
#pragma unroll
        for (int c = 0; c < 3; c++) x[c] = x[c] + S.dt * adv[c];
        if (S.solid_bits) obstacle_push(S, x, v);
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const float lo = org[a] + S.h * 1.001f;
            const float hi = org[a] + S.h * (nf[a] - 1.001f);
            if (x[a] < lo) { x[a] = lo; v[a] = v[a] < 0.0f ? 0.0f : v[a]; }
            if (x[a] > hi) { x[a] = hi; v[a] = 0.0f < v[a] ? 0.0f : v[a]; }
        }
        S.ax[i] = x[0];
        S.ay[i] = x[1];
        S.az[i] = x[2];
        S.avx[i] = v[0];
        S.avy[i] = v[1];
        S.avz[i] = v[2];
        S.aid[i] = S.bid[i];
    }
    __syncthreads();
    }
}
__global__ __launch_bounds__(TILE_BLOCK) void k_g2p_tile(DevSim S) { body_g2p<TILE_BLOCK>(S, blockIdx.x, gridDim.x); }
template <int NT>
__global__ __launch_bounds__(NT) void kb_g2p_tile(WorkList W) {
    for_each_work(W, [&](const DevSim &S, uint32_t b, const uint32_t *) { body_g2p<NT>(S, b, tile_grid(S)); });
}

// ---------------------------------------------------------------- IO / bins

// Bins in seeding order and reference (x-fastest) cell numbering: cell_of[id] and the
// per-cell id lists at offset_lin[cell] + slot (slots are arrival ranks within a cell).
__global__ __launch_bounds__(BLOCK) void k_bins_out(DevSim S, const uint32_t *off_lin, uint32_t *cell_of,
                                                    uint32_t *sorted) {
    uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= S.np) return;
    int cx, cy, cz;
    cell_of_pos(S, S.ax[i], S.ay[i], S.az[i], cx, cy, cz);
    const uint32_t c = (uint32_t)cx + (uint32_t)S.nx * ((uint32_t)cy + (uint32_t)S.ny * (uint32_t)cz);
    const uint32_t id = S.aid[i];
    if (cell_of) cell_of[id] = c;
    if (sorted) sorted[off_lin[c] + S.slot[i]] = id;
}

// Canonical order inside each cell: ids ascending (insertion sort of the cell's segment).
__global__ __launch_bounds__(BLOCK) void k_sort_segments(uint32_t ncells, const uint32_t *offsets,
                                                         const uint32_t *counts, uint32_t *sorted) {
    uint32_t c = blockIdx.x * BLOCK + threadIdx.x;
    if (c >= ncells) return;
    uint32_t *seg = sorted + offsets[c];
    uint32_t n = counts[c];
    for (uint32_t a = 1; a < n; a++) {
        uint32_t v = seg[a];
        uint32_t b = a;
        while (b > 0 && seg[b - 1] > v) {
            seg[b] = seg[b - 1];
            b--;
        }
        seg[b] = v;
    }
}

// Per-cell counts in reference cell numbering.
__global__ __launch_bounds__(BLOCK) void k_count_linear(DevSim S, uint32_t *out) {
    uint32_t c = blockIdx.x * BLOCK + threadIdx.x;
    if (c >= S.ncells) return;
    const uint32_t i = c % (uint32_t)S.nx, j = (c / (uint32_t)S.nx) % (uint32_t)S.ny,
                   k = c / ((uint32_t)S.nx * (uint32_t)S.ny);
    out[c] = S.count[tile_key(S, (int)i, (int)j, (int)k)];
}

__global__ __launch_bounds__(BLOCK) void k_unpack_aos8(DevSim S, float *out) {
    uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= S.np) return;
    float *r = out + (size_t)S.aid[i] * 8;
    float4 a = make_float4(S.ax[i], S.ay[i], S.az[i], 0.0f);
    float4 b = make_float4(S.avx[i], S.avy[i], S.avz[i], 0.0f);
    reinterpret_cast<float4 *>(r)[0] = a;
    reinterpret_cast<float4 *>(r)[1] = b;
}

__global__ __launch_bounds__(BLOCK) void k_pack_aos8(DevSim S, const float *in) {
    uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= S.np) return;
    const float4 a = reinterpret_cast<const float4 *>(in + (size_t)i * 8)[0];
    const float4 b = reinterpret_cast<const float4 *>(in + (size_t)i * 8)[1];
    S.ax[i] = a.x; S.ay[i] = a.y; S.az[i] = a.z;
    S.avx[i] = b.x; S.avy[i] = b.y; S.avz[i] = b.z;
    S.aid[i] = i;
}

__global__ __launch_bounds__(BLOCK) void k_reset_particles(DevSim S) {
    uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= S.np) return;
    S.ax[i] = S.ix[i]; S.ay[i] = S.iy[i]; S.az[i] = S.iz[i];
    S.avx[i] = S.ivx[i]; S.avy[i] = S.ivy[i]; S.avz[i] = S.ivz[i];
    S.aid[i] = i;
}

}  // namespace

// Largest squared particle speed, for the CFL cap: each workgroup takes a strided share of the
// particles, reduces it (lanes, then waves through LDS) and adds one atomic maximum.  The
// atomics all go to one word, so their number is kept to the workgroups of one pass over the
// chip (one per wave serialised some 7,000 atomics for 460K particles).
constexpr uint32_t VMAX_BLOCKS = 128;
__global__ __launch_bounds__(BLOCK) void k_vmax2(DevSim S, uint32_t *out) {
    __shared__ float sh[BLOCK / WAVE];
    float v2 = 0.0f;
    for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < S.np; i += gridDim.x * BLOCK) {
        const float vx = S.avx[i], vy = S.avy[i], vz = S.avz[i];
        const float q = vx * vx + vy * vy + vz * vz;
        v2 = q >= 0.0f ? fmaxf(v2, q) : v2;  // NaN adds nothing
    }
    for (int off = WAVE / 2; off > 0; off >>= 1) v2 = fmaxf(v2, __shfl_xor(v2, off, WAVE));
    if ((threadIdx.x & (WAVE - 1)) == 0) sh[threadIdx.x / WAVE] = v2;
    __syncthreads();
    if (threadIdx.x == 0) {
        float m = sh[0];
        for (int w = 1; w < BLOCK / WAVE; w++) m = fmaxf(m, sh[w]);
        atomicMax(out, __float_as_uint(m));
    }
}

void launch_vmax2(const DevSim &S, uint32_t *out, hipStream_t s) {
    launch_fill_u32(out, 1, 0u, s);
    if (S.np) PXL_LAUNCH(k_vmax2, std::min(blocks_for(S.np), VMAX_BLOCKS), BLOCK, 0, s, S, out);
}

void launch_fill_u32(uint32_t *p, uint32_t n, uint32_t value, hipStream_t s) {
    if (n) PXL_LAUNCH(k_fill_u32, blocks_for(n), BLOCK, 0, s, p, n, value);
}

void launch_clear_grid(const DevSim &S, hipStream_t s) { PXL_LAUNCH(k_clear_grid, tile_grid(S), BLOCK, 0, s, S); }

void launch_bin_count(const DevSim &S, hipStream_t s) {
    PXL_LAUNCH(k_bin_count, blocks_for(S.np), BLOCK, 0, s, S);
}

uint32_t scan_tiles(uint32_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE; }

void launch_exclusive_scan(const uint32_t *in, uint32_t n, uint32_t *tile_tmp, uint32_t *out, hipStream_t s) {
    const uint32_t ntiles = scan_tiles(n);
    if (ntiles == 0) return;
    PXL_LAUNCH(k_scan_reduce, ntiles, BLOCK, 0, s, in, n, tile_tmp);
    PXL_LAUNCH(k_scan_tiles, 1, BLOCK, 0, s, tile_tmp, ntiles);
    PXL_LAUNCH(k_scan_down, ntiles, BLOCK, 0, s, in, n, tile_tmp, out);
}

void launch_scan_offsets(const DevSim &S, hipStream_t s) { launch_exclusive_scan(S.count, S.nkeys, S.scan_tmp, S.offset, s); }

void launch_scatter(const DevSim &S, hipStream_t s) { PXL_LAUNCH(k_scatter, blocks_for(S.np), BLOCK, 0, s, S); }

void launch_tile_list(const DevSim &S, hipStream_t s) { PXL_LAUNCH(k_tile_list, blocks_for(S.ntiles), BLOCK, 0, s, S); }
void launch_band(const DevSim &S, hipStream_t s) { PXL_LAUNCH(k_band, blocks_for(S.ntiles), BLOCK, 0, s, S); }
void launch_p2g(const DevSim &S, hipStream_t s) { PXL_LAUNCH(k_p2g_tile, tile_grid(S), TILE_BLOCK, 0, s, S); }
void launch_g2p(const DevSim &S, hipStream_t s) { PXL_LAUNCH(k_g2p_tile, tile_grid(S), TILE_BLOCK, 0, s, S); }

void launch_bins_out(const DevSim &S, const uint32_t *off_lin, uint32_t *cell_of, uint32_t *sorted, hipStream_t s) {
    PXL_LAUNCH(k_bins_out, blocks_for(S.np), BLOCK, 0, s, S, off_lin, cell_of, sorted);
}
void launch_sort_segments(uint32_t ncells, const uint32_t *offsets, const uint32_t *counts, uint32_t *sorted,
                          hipStream_t s) {
    PXL_LAUNCH(k_sort_segments, blocks_for(ncells), BLOCK, 0, s, ncells, offsets, counts, sorted);
}
void launch_count_linear(const DevSim &S, uint32_t *out, hipStream_t s) {
    PXL_LAUNCH(k_count_linear, blocks_for(S.ncells), BLOCK, 0, s, S, out);
}
void launch_unpack_aos8(const DevSim &S, float *aos8, hipStream_t s) {
    PXL_LAUNCH(k_unpack_aos8, blocks_for(S.np), BLOCK, 0, s, S, aos8);
}
void launch_pack_aos8(const DevSim &S, const float *aos8, hipStream_t s) {
    PXL_LAUNCH(k_pack_aos8, blocks_for(S.np), BLOCK, 0, s, S, aos8);
}
void launch_reset_particles(const DevSim &S, hipStream_t s) {
    PXL_LAUNCH(k_reset_particles, blocks_for(S.np), BLOCK, 0, s, S);
}

// ---- batched
uint32_t particle_blocks(const DevSim &S) { return blocks_for(S.np); }
uint32_t tile_blocks(const DevSim &S) { return blocks_for(S.ntiles); }
uint32_t tile_grid_blocks(const DevSim &S) { return tile_grid(S); }

void batch_clear_grid(const WorkLists &W, hipStream_t s) {
    PXL_LAUNCH(kb_clear_grid, batch_grid(W[WD_TILEGRID], batch_cap(BF_PART)), BLOCK, 0, s, W[WD_TILEGRID]);
}
void batch_bin_count(const WorkLists &W, hipStream_t s) { PXL_LAUNCH(kb_bin_count, batch_grid(W[WD_PART], batch_cap(BF_PART)), BLOCK, 0, s, W[WD_PART]); }
void batch_scan_reduce(const WorkLists &W, int which, hipStream_t s) {
    PXL_LAUNCH(kb_scan_reduce, batch_grid(W[which == 0 ? WD_SCAN_KEYS : WD_SCAN_ROWS], batch_cap(BF_PART)), BLOCK, 0, s, W[which == 0 ? WD_SCAN_KEYS : WD_SCAN_ROWS], which);
}
void batch_scan_tiles(const WorkLists &W, int which, hipStream_t s) {
    PXL_LAUNCH(kb_scan_tiles, batch_grid(W[WD_SIM], batch_cap(BF_PART)), BLOCK, 0, s, W[WD_SIM], which);
}
void batch_scan_down(const WorkLists &W, int which, hipStream_t s) {
    PXL_LAUNCH(kb_scan_down, batch_grid(W[which == 0 ? WD_SCAN_KEYS : WD_SCAN_ROWS], batch_cap(BF_PART)), BLOCK, 0, s, W[which == 0 ? WD_SCAN_KEYS : WD_SCAN_ROWS], which);
}
void batch_tile_list(const WorkLists &W, hipStream_t s) { PXL_LAUNCH(kb_tile_list, batch_grid(W[WD_TILES], batch_cap(BF_PART)), BLOCK, 0, s, W[WD_TILES]); }
void batch_band(const WorkLists &W, hipStream_t s) { PXL_LAUNCH(kb_band, batch_grid(W[WD_TILES], batch_cap(BF_PART)), BLOCK, 0, s, W[WD_TILES]); }
void batch_scatter(const WorkLists &W, hipStream_t s) { PXL_LAUNCH(kb_scatter, batch_grid(W[WD_PART], batch_cap(BF_PART)), BLOCK, 0, s, W[WD_PART]); }
// The per-tile launches in the workgroup size the tile-lane setting picks (512, the default, or 1024 lanes:
// lanes stride over a tile's particles, so the size changes the speed, not the results).
void batch_p2g(const WorkLists &W, hipStream_t s) {
    const unsigned g = batch_grid(W[WD_TILEGRID], batch_cap(BF_TILE));
    if (batch_shape(BS_TILE) == 512) PXL_LAUNCH(kb_p2g_tile<512>, g, 512, 0, s, W[WD_TILEGRID]);
    else PXL_LAUNCH(kb_p2g_tile<TILE_BLOCK>, g, TILE_BLOCK, 0, s, W[WD_TILEGRID]);
}
void batch_g2p(const WorkLists &W, hipStream_t s) {
    const unsigned g = batch_grid(W[WD_TILEGRID], batch_cap(BF_TILE));
    if (batch_shape(BS_TILE) == 512) PXL_LAUNCH(kb_g2p_tile<512>, g, 512, 0, s, W[WD_TILEGRID]);
    else PXL_LAUNCH(kb_g2p_tile<TILE_BLOCK>, g, TILE_BLOCK, 0, s, W[WD_TILEGRID]);
}

}  // namespace pxl
