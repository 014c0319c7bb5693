// Geometric multigrid preconditioner for the pressure system of the mixed-precision solve
// (the reference has only fixed-count red-black Gauss-Seidel, fluid.slang:167-182; this
// solver is a build extension that converges the same matrix to a relative FP64 residual).
//
// Levels: level 0 is the simulation grid.  Level l+1 halves every axis (ceil); a coarse cell
// is AIR if any of its 8 children is AIR, else FLUID if any child is FLUID, else SOLID
// (children past the grid edge count as SOLID).  Every level uses the unit 7-point stencil L_l
// of its own types (diagonal = non-SOLID neighbours, -1 per FLUID neighbour).  The coarse
// operator stands for 2 L_{l+1} (the Galerkin product of piecewise-constant transfers is
// 4 L_{l+1}; halving it over-corrects the smooth error that aggregation under-resolves), so
// restriction = one half of the sum over the 8 children and prolongation = injection.
//
// One application z = M r is a V(2,2) cycle with red-black Gauss-Seidel smoothing (red:
// i+j+k even).  Pre-smoothing runs red then black, twice, from zero; post-smoothing runs
// black then red, twice, so it is the adjoint of the pre-smoother and M is symmetric (a valid
// CG preconditioner).  A Gauss-Seidel update z_c = (r_c + sum_FLUID nbrs z_n) / diag_c never
// reads z_c itself, which fuses the steps below:
//   pre0:     red z = r / d, black z = (r + sum red r_n / d_n) / d  (the first sweep)
//   sweep:    one colour's update (the second pre sweep is sweep(R), sweep(B))
//   restrict: the residual after a black sweep is zero at black cells, so the coarse
//             right-hand side sums the red children's residuals only
//   post:     black update from red neighbours' z + prolonged correction (the red cells'
//             corrected values are only ever read here), then sweep(R), sweep(B), sweep(R).
// Levels below mg_tail are walked through lists of their FLUID cells, red cells apart from
// black ones (the step's colour-split fluid list on level 0, per-level lists built with the
// coarse types), so the work follows the liquid, not the grid, and a colour sweep keeps every
// lane busy; a launch takes a list segment in level_blocks slices (the solver kernels' count on
// level 0, an eighth of it per coarser level), and a lane has the loads of several cells in
// flight together.  Every level stores its vectors colour-split (red cells in one half, black
// in the other, each in cell order; MGLevel), so a colour sweep reads and writes dense runs.
// Coarse z is kept 0 on non-FLUID cells (the prolongation reads the parent of any FLUID
// cell).  Levels from mg_tail on (a few thousand cells together) run the same cycle in one
// workgroup with everything in LDS.  Arithmetic is fp32 (the FP64 refinement outside fixes
// the rounding).  Single-sim kernels take the levels they touch as MGLevel values: a runtime
// index into DevSim's level array would copy the whole structure into scratch memory.
#include "common.hpp"
#include "kernels.hpp"

namespace pxl {

namespace {

constexpr int MBLOCK = 256;         // grid-walking level kernels (types, stencils, lists)
constexpr int LBLOCK = 1024;        // list-walking level kernels
constexpr int TBLOCK = 1024;        // the tail workgroup
constexpr int COARSEST_SWEEPS = 1;  // sweeps R-B then as many B-R on the coarsest level

__device__ __forceinline__ float rcp_or0(float d) { return d > 0.0f ? __builtin_amdgcn_rcpf(d) : 0.0f; }
__device__ __forceinline__ float st_diag(uint32_t st) { return (float)((st >> 8) & 7u); }

__device__ __forceinline__ uint32_t lidx(const MGLevel &v, int i, int j, int k) {
    return (uint32_t)i + (uint32_t)v.nx * ((uint32_t)j + (uint32_t)v.ny * (uint32_t)k);
}
__device__ __forceinline__ uint32_t ncell(const MGLevel &v) { return (uint32_t)(v.nx * v.ny * v.nz); }
__device__ __forceinline__ void cell_ijk(int nx, int ny, uint32_t c, int &i, int &j, int &k) {
    i = (int)(c % (uint32_t)nx);
    j = (int)((c / (uint32_t)nx) % (uint32_t)ny);
    k = (int)(c / ((uint32_t)nx * (uint32_t)ny));
}
__device__ __forceinline__ uint8_t level_type(const MGLevel &v, int i, int j, int k) {
    if (i < 0 || j < 0 || k < 0 || i >= v.nx || j >= v.ny || k >= v.nz) return CELL_SOLID;
    return v.type[lidx(v, i, j, k)];
}
// Cell-order index of the coarse cell containing fine cell (i, j, k) (the tail's LDS levels).
__device__ __forceinline__ uint32_t parent(int cnx, int cny, int i, int j, int k) {
    return (uint32_t)(i >> 1) + (uint32_t)cnx * ((uint32_t)(j >> 1) + (uint32_t)cny * (uint32_t)(k >> 1));
}
// Colour-split index of cell (i, j, k) of a coarse level (MGLevel::hx).
__device__ __forceinline__ uint32_t lsig(const MGLevel &v, int i, int j, int k) {
    return (((uint32_t)(i + j + k) & 1u) ? v.hsplit : 0u) + (uint32_t)(i >> 1) +
           (uint32_t)v.hx * ((uint32_t)j + (uint32_t)v.ny * (uint32_t)k);
}
// The colour-split index of the coarse cell containing fine cell (i, j, k) (i, j, k >= 0).
__device__ __forceinline__ uint32_t parent_sig(const MGLevel &c, int i, int j, int k) {
    return lsig(c, i >> 1, j >> 1, k >> 1);
}
__device__ __forceinline__ uint32_t sig_last(const MGLevel &v) { return 2u * v.hsplit - 1u; }
// A coarse list entry: sigma | (i & 1) << 31.
__device__ __forceinline__ uint32_t e_sig(uint32_t e) { return e & 0x7fffffffu; }
__device__ __forceinline__ int32_t e_odd(uint32_t e) { return (int32_t)(e >> 31); }
// Offsets from a cell's sigma to its -x, +x, -y, +y, -z, +z neighbours (all of the other
// colour): d - 1 + o, d + o, d -/+ hx, d -/+ hx ny, d = +hsplit from red, -hsplit from black,
// o = i & 1.
__device__ __forceinline__ void sig_offsets(const MGLevel &v, bool black, int32_t o, int32_t off[6]) {
    const int32_t d = black ? -(int32_t)v.hsplit : (int32_t)v.hsplit, hy = v.hx, hz = v.hx * v.ny;
    off[0] = d - 1 + o;
    off[1] = d + o;
    off[2] = d - hy;
    off[3] = d + hy;
    off[4] = d - hz;
    off[5] = d + hz;
}
// Grid coordinates of coarse list entry e.
__device__ __forceinline__ void e_ijk(const MGLevel &v, uint32_t e, int &i, int &j, int &k) {
    const uint32_t sg = e_sig(e), q = sg >= v.hsplit ? sg - v.hsplit : sg;
    const uint32_t row = q / (uint32_t)v.hx;
    i = 2 * (int)(q - row * (uint32_t)v.hx) + e_odd(e);
    j = (int)(row % (uint32_t)v.ny);
    k = (int)(row / (uint32_t)v.ny);
}
// Workgroup b's (of nb) slice [beg, end) of an n-entry list.  Workgroups are dealt to the 8
// XCDs round-robin (workgroup i runs on XCD i % 8) and each XCD has its own L2: with nb a
// multiple of 8, XCD x takes the x-th eighth of the list, so the neighbour rows and planes
// its slices read are mostly in its own L2.
__device__ __forceinline__ void slice(uint32_t n, uint32_t b, uint32_t nb, uint32_t &beg, uint32_t &end) {
    if (nb % 8u == 0) b = (b % 8u) * (nb / 8u) + b / 8u;
    beg = (uint32_t)((uint64_t)n * b / nb);
    end = (uint32_t)((uint64_t)n * (b + 1) / nb);
}

// Fields of a level-0 list entry's word (DevSim::fl_word).
__device__ __forceinline__ float w_diag(uint32_t w) { return (float)((w >> 6) & 7u); }
__device__ __forceinline__ float w_ndiag(uint32_t w, int n) { return (float)((w >> (9 + 3 * n)) & 7u); }
__device__ __forceinline__ uint32_t w_black(uint32_t w) { return (w >> 27) & 1u; }

// ---- level types (c from its children in f)
__device__ __forceinline__ void body_mg_types(const MGLevel &f, const MGLevel &c, uint32_t b) {
    const uint32_t id = b * MBLOCK + threadIdx.x;
    if (id >= ncell(c)) return;
    int i, j, k;
    cell_ijk(c.nx, c.ny, id, i, j, k);
    bool air = false, fluid = false;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const uint8_t t = level_type(f, 2 * i + (q & 1), 2 * j + ((q >> 1) & 1), 2 * k + (q >> 2));
        air = air || t == CELL_AIR;
        fluid = fluid || t == CELL_FLUID;
    }
    c.type[id] = air ? CELL_AIR : (fluid ? CELL_FLUID : CELL_SOLID);
}
__global__ __launch_bounds__(MBLOCK) void k_mg_types(MGLevel f, MGLevel c) { body_mg_types(f, c, blockIdx.x); }

// ---- level stencil words (FLUID bit, non-SOLID-neighbour count in bits 8-10, FLUID-neighbour
// mask in bits 0-5: the layout of the simulation grid's stencil words, plus the colour in
// STEN_BLACK), z = 0, and each 64-cell chunk's red and black FLUID cells (a wave is a chunk)
__device__ __forceinline__ void body_mg_sten(const MGLevel &v, uint32_t b) {
    const uint32_t id = b * MBLOCK + threadIdx.x;
    const bool in = id < ncell(v);
    uint32_t st = 0;
    bool black = false;
    if (in) {
        int i, j, k;
        cell_ijk(v.nx, v.ny, id, i, j, k);
        const uint32_t sg = lsig(v, i, j, k);
        if (v.type[id] == CELL_FLUID) {
            const int nb[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k}, {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
            uint32_t diag = 0;
#pragma unroll
            for (int n = 0; n < 6; n++) {
                const uint8_t t = level_type(v, nb[n][0], nb[n][1], nb[n][2]);
                if (t != CELL_SOLID) diag++;
                if (t == CELL_FLUID) st |= 1u << n;
            }
            black = ((i + j + k) & 1) != 0;
            st |= STEN_FLUID | (diag << 8) | (black ? STEN_BLACK : 0u);
        }
        v.sten[sg] = (uint16_t)st;
        v.z[sg] = 0.0f;
    }
    const bool fl = (st & STEN_FLUID) != 0;
    const uint64_t mr = __ballot(fl && !black), mb = __ballot(fl && black);
    const uint32_t c = id / WAVE, nch = level_chunks(v);
    if ((threadIdx.x & (WAVE - 1)) == 0 && c < nch) {
        v.chunk[c] = (uint32_t)__popcll(mr);
        v.chunk[nch + c] = (uint32_t)__popcll(mb);
    }
}
__global__ __launch_bounds__(MBLOCK) void k_mg_sten(MGLevel v) { body_mg_sten(v, blockIdx.x); }

// ---- the level lists in cell order.  One workgroup of TBLOCK lanes scans a simulation's
// chunk counts, level by level, into exclusive offsets (each lane sums a run of consecutive
// chunks, the workgroup scans the lanes' sums, each lane rewrites its run) and sets the list
// counts; then a wave per chunk writes the chunk's FLUID cells at its offsets.
// Exclusive scan of a[0, n) in place; returns the total.
__device__ __forceinline__ uint32_t block_scan_inplace(uint32_t *a, uint32_t n) {
    __shared__ uint32_t wsum[TBLOCK / WAVE + 1];
    const uint32_t per = (n + TBLOCK - 1) / TBLOCK;
    const uint32_t beg = min(threadIdx.x * per, n), end = min(beg + per, n);
    uint32_t own = 0;
    for (uint32_t i = beg; i < end; i++) own += a[i];
    const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
    uint32_t inc = own;  // inclusive scan of the lanes' sums within the wave
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        const uint32_t y = __shfl_up(inc, o, WAVE);
        if (lane >= o) inc += y;
    }
    if (lane == WAVE - 1) wsum[w] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int q = 0; q < TBLOCK / WAVE; q++) {
            const uint32_t t = wsum[q];
            wsum[q] = acc;
            acc += t;
        }
        wsum[TBLOCK / WAVE] = acc;
    }
    __syncthreads();
    uint32_t run = wsum[w] + inc - own;
    for (uint32_t i = beg; i < end; i++) {
        const uint32_t t = a[i];
        a[i] = run;
        run += t;
    }
    const uint32_t total = wsum[TBLOCK / WAVE];
    __syncthreads();  // wsum is reused by the next scan
    return total;
}
// The level copy with its count pointer taken from a batch member's count slots (`cnt`; null:
// the level's own counts).
__device__ __forceinline__ MGLevel level_counts(const DevSim &S, uint32_t *cnt, int l) {
    MGLevel v = S.mg[l];
    if (cnt) v.count = cnt + 4 * l;
    return v;
}
__device__ __forceinline__ void body_mg_scan(const DevSim &S, uint32_t *cnt) {
#pragma unroll
    for (int q = 1; q < MG_MAX; q++) {
        if (q < S.mg_levels) {
            const MGLevel v = level_counts(S, cnt, q);
            const uint32_t nch = level_chunks(v);
            const uint32_t red = block_scan_inplace(v.chunk, nch);
            const uint32_t black = block_scan_inplace(v.chunk + nch, nch);
            if (threadIdx.x == 0) {
                v.count[0] = red;
                v.count[1] = black;
                v.count[2] = red + black;
            }
        }
    }
}
__device__ __forceinline__ void body_mg_compact(const MGLevel &v, uint32_t b) {
    const uint32_t id = b * MBLOCK + threadIdx.x, n = ncell(v), nch = level_chunks(v);
    const uint32_t c = id / WAVE;  // the same for the whole wave
    if (c >= nch) return;
    int i = 0, j = 0, k = 0;
    if (id < n) cell_ijk(v.nx, v.ny, id, i, j, k);
    const uint32_t sg = lsig(v, i, j, k);
    const uint32_t st = id < n ? (uint32_t)v.sten[sg] : 0u;
    const bool fl = (st & STEN_FLUID) != 0, black = (st & STEN_BLACK) != 0;
    const uint64_t mr = __ballot(fl && !black), mb = __ballot(fl && black);
    if (!fl) return;
    const uint64_t below = (1ull << (threadIdx.x & (WAVE - 1))) - 1ull;
    const uint32_t ro = v.chunk[c], bo = v.chunk[nch + c];
    const uint32_t rank = (uint32_t)__popcll((black ? mb : mr) & below);
    const uint32_t e = sg | ((uint32_t)(i & 1) << 31);
    if (black) v.list[n - 1u - (bo + rank)] = e;
    else v.list[ro + rank] = e;
    v.nlist[ro + bo + (uint32_t)__popcll((mr | mb) & below)] = e;
}
__global__ __launch_bounds__(TBLOCK) void k_mg_scan(DevSim S) { body_mg_scan(S, nullptr); }
__global__ __launch_bounds__(MBLOCK) void k_mg_compact(MGLevel v) { body_mg_compact(v, blockIdx.x); }

// ---- level 0 (the step's colour-split fluid list: red entries [0, n_red), black [n_red, n);
// r, u are the CG vectors, stored colour-split: an entry's vector index is its fls_list /
// fns_list value and its neighbours' are given by split_offsets)
// pre0: red u = r / d, black u = (r + sum FLUID red nbrs r_n / d_n) / d
// The level-0 kernels below give a lane entries t0 + k LBLOCK (k < PRE0_MLP, SWEEP_MLP,
// POST_MLP) and issue the loads of all of them before any store.  An entry's update reads only
// cells of the other colour (or, in pre0, r), so the order of the updates does not matter.
// Measured on the 64-request batch: pre0 and post 9% faster than one entry at a time; the
// sweep is fastest at two.
constexpr int PRE0_MLP = 4, SWEEP_MLP = 2, POST_MLP = 2;
__device__ __forceinline__ void body_l0_pre0(const DevSim &S, const uint32_t *flc, const float *__restrict__ r,
                                             float *__restrict__ u, uint32_t b, uint32_t nb) {
    uint32_t beg, end;
    slice(flc[1], b, nb, beg, end);
    const Buf br = buf_of(r, 8u * S.hsplit);
    for (uint32_t t0 = beg + threadIdx.x; t0 < end; t0 += PRE0_MLP * LBLOCK) {
        uint32_t id[PRE0_MLP], wd[PRE0_MLP];
#pragma unroll
        for (int k = 0; k < PRE0_MLP; k++) {
            const uint32_t t = t0 + k * LBLOCK;
            id[k] = S.fls_list[t < end ? t : t0];  // colour-split order: runs through each half
            wd[k] = S.fl_word[t < end ? t : t0];  // loaded unconditionally, then dropped
            wd[k] = t < end ? wd[k] : 0u;    // past the slice: red, no neighbours, no store
        }
        float rn[PRE0_MLP][6], rc[PRE0_MLP];
#pragma unroll
        for (int k = 0; k < PRE0_MLP; k++) {
            int32_t off[6];
            split_offsets(S, wd[k], off);
            // red entries read no neighbours (one branch per entry, not per load: the list is
            // red then black, so a wave mostly takes one side); a black entry loads all six, a
            // neighbour that is not FLUID at the +x slot (a line its x neighbours share), and
            // drops them afterwards
#pragma unroll
            for (int n = 0; n < 6; n++) rn[k][n] = buf_f32(br, id[k] + off[n], w_black(wd[k]) && ((wd[k] >> n) & 1u));
            rc[k] = S.cg_rt[t0 + k * LBLOCK < end ? t0 + k * LBLOCK : t0];
        }
#pragma unroll
        for (int k = 0; k < PRE0_MLP; k++) {
            if (t0 + k * LBLOCK >= end) break;
            float sum = 0.0f;
            if (w_black(wd[k])) {
#pragma unroll
                for (int n = 0; n < 6; n++)
                    if ((wd[k] >> n) & 1u) sum += rn[k][n] * rcp_or0(w_ndiag(wd[k], n));
            }
            u[id[k]] = (rc[k] + sum) * rcp_or0(w_diag(wd[k]));
        }
    }
}
// One colour's update over its list segment.  The colour's neighbours all lie in the other
// colour's half of the split vector: d +/- 1 or d along x, d +/- hx along y, d +/- hx ny along z.
__device__ __forceinline__ void body_l0_sweep(const DevSim &S, const uint32_t *flc, const float *__restrict__ r,
                                              float *__restrict__ u, uint32_t colour, uint32_t b, uint32_t nb) {
    const uint32_t n_red = flc[0], n = flc[1];
    const uint32_t lo = colour ? n_red : 0u, hi = colour ? n : n_red;
    uint32_t beg, end;
    slice(hi - lo, b, nb, beg, end);
    beg += lo;
    end += lo;
    const int32_t d = colour ? -(int32_t)S.hsplit : (int32_t)S.hsplit;
    const int32_t hy = (int32_t)S.hx, hz = (int32_t)S.hx * S.ny;
    const Buf bu = buf_of(u, 8u * S.hsplit);
    for (uint32_t t0 = beg + threadIdx.x; t0 < end; t0 += SWEEP_MLP * LBLOCK) {
        uint32_t id[SWEEP_MLP], wd[SWEEP_MLP];
#pragma unroll
        for (int k = 0; k < SWEEP_MLP; k++) {
            const uint32_t t = t0 + k * LBLOCK;
            id[k] = S.fls_list[t < end ? t : t0];
            wd[k] = S.fl_word[t < end ? t : t0];  // loaded unconditionally, then dropped
            wd[k] = t < end ? wd[k] : 0u;
        }
        float z[SWEEP_MLP][6], rc[SWEEP_MLP];
#pragma unroll
        for (int k = 0; k < SWEEP_MLP; k++) {
            const int32_t o = (int32_t)((wd[k] >> 28) & 1u);
            const int32_t off[6] = {d - 1 + o, d + o, d - hy, d + hy, d - hz, d + hz};
            // only FLUID neighbours are read (loading the others too, at the +x slot, made the
            // batched sweep 10-15% slower on the 512-request batch)
#pragma unroll
            for (int q = 0; q < 6; q++) z[k][q] = buf_f32(bu, id[k] + off[q], (wd[k] >> q) & 1u);
            rc[k] = S.cg_rt[t0 + k * LBLOCK < end ? t0 + k * LBLOCK : t0];
        }
#pragma unroll
        for (int k = 0; k < SWEEP_MLP; k++) {
            if (t0 + k * LBLOCK >= end) break;
            float sum = 0.0f;
#pragma unroll
            for (int q = 0; q < 6; q++) sum += z[k][q];
            u[id[k]] = (rc[k] + sum) * rcp_or0(w_diag(wd[k]));
        }
    }
}
// post: black update from red neighbours' u + the prolonged correction c.z (the parents come
// from the cell's grid coordinates, fl_list)
__device__ __forceinline__ void body_l0_post(const DevSim &S, const uint32_t *flc, const MGLevel &c,
                                             const float *__restrict__ r, float *__restrict__ u, uint32_t b,
                                             uint32_t nb) {
    const uint32_t n_red = flc[0], n = flc[1];
    uint32_t beg, end;
    slice(n - n_red, b, nb, beg, end);
    beg += n_red;
    end += n_red;
    const Buf bu = buf_of(u, 8u * S.hsplit), bc = buf_of(c.z, 8u * c.hsplit);
    for (uint32_t t0 = beg + threadIdx.x; t0 < end; t0 += POST_MLP * LBLOCK) {
        uint32_t wd[POST_MLP], id[POST_MLP], sid[POST_MLP];
#pragma unroll
        for (int k = 0; k < POST_MLP; k++) {
            const uint32_t t = t0 + k * LBLOCK, tt = t < end ? t : t0;
            wd[k] = S.fl_word[t < end ? t : t0];  // loaded unconditionally, then dropped
            wd[k] = t < end ? wd[k] : 0u;
            id[k] = S.fl_list[tt];
            sid[k] = S.fls_list[tt];
        }
        float zq[POST_MLP][6], pq[POST_MLP][6], rc[POST_MLP];
#pragma unroll
        for (int k = 0; k < POST_MLP; k++) {
            int32_t off[6];
            split_offsets(S, wd[k], off);
            int i, j, kk;
            cell_ijk(S.nx, S.ny, id[k], i, j, kk);
            const int nbc[6][3] = {{i - 1, j, kk}, {i + 1, j, kk}, {i, j - 1, kk}, {i, j + 1, kk}, {i, j, kk - 1}, {i, j, kk + 1}};
#pragma unroll
            for (int q = 0; q < 6; q++) {
                const bool f = (wd[k] >> q) & 1u;
                zq[k][q] = buf_f32(bu, sid[k] + off[q], f);
                pq[k][q] = buf_f32(bc, parent_sig(c, nbc[q][0], nbc[q][1], nbc[q][2]), f);
            }
            rc[k] = S.cg_rt[t0 + k * LBLOCK < end ? t0 + k * LBLOCK : t0];
        }
#pragma unroll
        for (int k = 0; k < POST_MLP; k++) {
            if (t0 + k * LBLOCK >= end) break;
            float sum = 0.0f;
#pragma unroll
            for (int q = 0; q < 6; q++) {
                if ((wd[k] >> q) & 1u) sum += zq[k][q] + pq[k][q];
            }
            u[sid[k]] = (rc[k] + sum) * rcp_or0(w_diag(wd[k]));
        }
    }
}

// ---- coarse levels 1..mg_tail-1 (the level's red / black lists; pre0 walks the cell-order list)
__device__ __forceinline__ void body_lc_pre0(const MGLevel &v, uint32_t b, uint32_t nb) {
    const uint32_t n = v.count[0] + v.count[1];
    uint32_t beg, end;
    slice(n, b, nb, beg, end);
    const uint32_t last = sig_last(v);
    for (uint32_t t = beg + threadIdx.x; t < end; t += LBLOCK) {
        const uint32_t e = v.nlist[t], id = e_sig(e);
        const uint32_t st = v.sten[id];
        float sum = 0.0f;
        if (id >= v.hsplit) {  // black (its colour follows from the index)
            // neighbours are loaded before the cell's stencil word arrives (indices clamped
            // into the level) and the ones that are not FLUID are dropped afterwards
            int32_t off[6];
            sig_offsets(v, true, e_odd(e), off);
            float rq[6];
            uint32_t sq[6];
#pragma unroll
            for (int q = 0; q < 6; q++) {
                const uint32_t nq = min((uint32_t)((int32_t)id + off[q]), last);
                rq[q] = v.r[nq];
                sq[q] = v.sten[nq];
            }
#pragma unroll
            for (int q = 0; q < 6; q++) {
                const float tq = rq[q] * rcp_or0(st_diag(sq[q]));
                sum = (st & (1u << q)) ? sum + tq : sum;
            }
        }
        v.z[id] = (v.r[id] + sum) * rcp_or0(st_diag(st));
    }
}
__device__ __forceinline__ void body_lc_sweep(const MGLevel &v, uint32_t colour, uint32_t b, uint32_t nb) {
    const uint32_t n = v.count[colour];
    const uint32_t *list = v.list;
    const uint32_t last = ncell(v) - 1u;
    uint32_t beg, end;
    slice(n, b, nb, beg, end);
    const int32_t d = colour ? -(int32_t)v.hsplit : (int32_t)v.hsplit, hy = v.hx, hz = v.hx * v.ny;
    const Buf bz = buf_of(v.z, 8u * v.hsplit);
    for (uint32_t t = beg + threadIdx.x; t < end; t += 2 * LBLOCK) {
        const uint32_t t2 = t + LBLOCK;
        const bool two = t2 < end;
        // both entries are loaded at once (the second at the first's index past the slice)
        const uint32_t tb = two ? t2 : t;
        const uint32_t ea = list[colour ? last - t : t];
        const uint32_t eb = list[colour ? last - tb : tb];
        const uint32_t ia = e_sig(ea), ib = e_sig(eb);
        const int32_t oa = e_odd(ea), ob = e_odd(eb);
        const int32_t offa[6] = {d - 1 + oa, d + oa, d - hy, d + hy, d - hz, d + hz};
        const int32_t offb[6] = {d - 1 + ob, d + ob, d - hy, d + hy, d - hz, d + hz};
        // the second stencil word is loaded unconditionally (at the first entry's cell past the
        // slice), so both are in flight together; only FLUID neighbours are loaded (loading
        // every neighbour with the stencil words was slower on large batches)
        const uint32_t sta = v.sten[ia], stb_raw = v.sten[ib];
        const uint32_t stb = two ? stb_raw : 0u;
        float za[6], zb[6];
#pragma unroll
        for (int q = 0; q < 6; q++) {
            za[q] = buf_f32(bz, ia + offa[q], (sta >> q) & 1u);
            zb[q] = buf_f32(bz, ib + offb[q], (stb >> q) & 1u);
        }
        float sa = 0.0f, sb = 0.0f;
#pragma unroll
        for (int q = 0; q < 6; q++) {
            sa += za[q];
            sb += zb[q];
        }
        const float ra = v.r[ia], rb = v.r[ib];
        v.z[ia] = (ra + sa) * rcp_or0(st_diag(sta));
        if (two) v.z[ib] = (rb + sb) * rcp_or0(st_diag(stb));
    }
}
__device__ __forceinline__ void body_lc_post(const MGLevel &v, const MGLevel &c, uint32_t b, uint32_t nb) {
    const uint32_t n = v.count[1];
    const uint32_t last = ncell(v) - 1u;
    uint32_t beg, end;
    slice(n, b, nb, beg, end);
    for (uint32_t t = beg + threadIdx.x; t < end; t += LBLOCK) {
        const uint32_t e = v.list[last - t], id = e_sig(e);
        int i, j, k;
        e_ijk(v, e, i, j, k);
        const int nbc[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k}, {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
        // neighbour and parent values are loaded together with the stencil word (coordinates
        // clamped into the level); the ones that are not FLUID are dropped afterwards
        float zq[6], pq[6];
#pragma unroll
        for (int q = 0; q < 6; q++) {
            const int a = min(max(nbc[q][0], 0), v.nx - 1), bb = min(max(nbc[q][1], 0), v.ny - 1),
                      d = min(max(nbc[q][2], 0), v.nz - 1);
            zq[q] = v.z[lsig(v, a, bb, d)];
            pq[q] = c.z[parent_sig(c, a, bb, d)];
        }
        const uint32_t st = v.sten[id];
        float sum = 0.0f;
#pragma unroll
        for (int q = 0; q < 6; q++) sum = (st & (1u << q)) ? sum + (zq[q] + pq[q]) : sum;
        v.z[id] = (v.r[id] + sum) * rcp_or0(st_diag(st));
    }
}

// ---- restrict level f -> c after f's black sweep: for each FLUID cell of c (its cell-order
// list, so neighbouring lanes read neighbouring children), c.r = 1/2 sum over FLUID red
// children (q = 0, 3, 5, 6) of r - L z.  With f colour-split, red child (fi, fj, fk) of coarse
// cell (ci, cj, ck) sits at (fi >> 1) + hx (fj + ny fk) = ci + hx (fj + ny fk) in f's red half
// and its (black) neighbours at that index + hsplit + the x parity's offsets; every child's
// values and neighbours are loaded before its stencil word arrives (indices clamped into the
// level; children past its edge read index 0 and are skipped, neighbours that are not FLUID
// are dropped afterwards).  f's arrays are given explicitly (level 0: the CG vectors and the
// grid's stencil words, whose red cells sit at the same split index).
template <bool Level0>
__device__ __forceinline__ void restrict_split(int fnx, int fny, int fnz, uint32_t fhx, uint32_t fhsplit,
                                               const uint16_t *__restrict__ fsten, const float *__restrict__ fr,
                                               const float *__restrict__ fz, const MGLevel &c, uint32_t b,
                                               uint32_t nb) {
    const uint32_t n = c.count[0] + c.count[1];
    uint32_t beg, end;
    slice(n, b, nb, beg, end);
    const int32_t hs = (int32_t)fhsplit, hy = (int32_t)fhx, hz = (int32_t)fhx * fny;
    const uint32_t vlast = 2u * fhsplit - 1u;
    for (uint32_t t = beg + threadIdx.x; t < end; t += LBLOCK) {
        const uint32_t ce = c.nlist[t];
        int ci, cj, ck;
        e_ijk(c, ce, ci, cj, ck);
        const int red[4] = {0, 3, 5, 6};
        bool in[4];
        uint32_t st[4];
        float rc[4], zc[4], zn[4][6];
#pragma unroll
        for (int m = 0; m < 4; m++) {
            const int q = red[m];
            const int a = q & 1, fi = 2 * ci + a, fj = 2 * cj + ((q >> 1) & 1), fk = 2 * ck + (q >> 2);
            in[m] = fi < fnx && fj < fny && fk < fnz;
            const uint32_t sid = in[m] ? (uint32_t)ci + fhx * ((uint32_t)fj + (uint32_t)fny * (uint32_t)fk) : 0u;
            // level 0's stencil words are in cell order; a coarse level's are colour-split
            const uint32_t wid =
                Level0 ? (in[m] ? (uint32_t)fi + (uint32_t)fnx * ((uint32_t)fj + (uint32_t)fny * (uint32_t)fk) : 0u)
                       : sid;
            st[m] = fsten[wid];
            rc[m] = fr[sid];
            zc[m] = fz[sid];
            const int32_t off[6] = {hs - 1 + a, hs + a, hs - hy, hs + hy, hs - hz, hs + hz};
#pragma unroll
            for (int e = 0; e < 6; e++) zn[m][e] = fz[min((uint32_t)((int32_t)sid + off[e]), vlast)];
        }
        float acc = 0.0f;
#pragma unroll
        for (int m = 0; m < 4; m++) {
            if (!in[m] || !(st[m] & STEN_FLUID)) continue;
            float sum = 0.0f;
#pragma unroll
            for (int e = 0; e < 6; e++) sum = (st[m] & (1u << e)) ? sum + zn[m][e] : sum;
            acc += rc[m] - st_diag(st[m]) * zc[m] + sum;
        }
        c.r[e_sig(ce)] = 0.5f * acc;
    }
}
__device__ __forceinline__ void body_restrict(const MGLevel &f, const MGLevel &c, uint32_t b, uint32_t nb) {
    restrict_split<false>(f.nx, f.ny, f.nz, (uint32_t)f.hx, f.hsplit, f.sten, f.r, f.z, c, b, nb);
}
__device__ __forceinline__ void body_restrict0(const DevSim &S, const MGLevel &c, uint32_t b, uint32_t nb) {
    restrict_split<true>(S.nx, S.ny, S.nz, (uint32_t)S.hx, S.hsplit, S.sten, S.cg_r0, S.cg_r1, c, b, nb);
}

// ---- the tail: levels mg_tail..L in one workgroup.  LDS holds every tail level's z, r and
// stencil words back to back, and per level the LDS indices of its FLUID cells (red ones from
// the level's start, black ones from its end, listed once at the fill in any order: every
// phase below updates cells independently of each other); the cycle is the one above with
// barriers in place of launches.  A lane issues all LDS loads of a cell (word, r, the six
// neighbours at indices clamped into the level) before using any; neighbours that are not
// FLUID are dropped afterwards.
struct TailLevel {
    int nx, ny, nz, o;      // dims and LDS offset
    int cnt[2];             // FLUID cells of each colour
    const uint16_t *sten;   // the level's stencil words and right-hand side in global memory
    const float *r;         // (colour-split: MGLevel::hx)
    int hx;
    uint32_t hsplit;
};
__device__ __forceinline__ int t_cells(const TailLevel &v) { return v.nx * v.ny * v.nz; }
// LDS index of the level's t-th FLUID cell (red ones first, then black ones)
__device__ __forceinline__ int t_entry(const TailLevel &v, const uint16_t *lcl, int t) {
    return t < v.cnt[0] ? lcl[v.o + t] : lcl[v.o + t_cells(v) - 1 - (t - v.cnt[0])];
}

__device__ __forceinline__ void t_sweep(const TailLevel &v, float *lz, const float *lr, const uint16_t *ls,
                                        const uint16_t *lcl, int colour) {
    const int n = v.cnt[colour], last = v.o + t_cells(v) - 1;
    const int32_t off[6] = {-1, 1, -v.nx, v.nx, -v.nx * v.ny, v.nx * v.ny};
    for (int t = threadIdx.x; t < n; t += TBLOCK) {
        const int c = colour ? lcl[last - t] : lcl[v.o + t];
        const uint32_t st = ls[c];
        const float rc = lr[c];
        float zq[6];
#pragma unroll
        for (int q = 0; q < 6; q++) zq[q] = lz[min(max(c + off[q], v.o), last)];
        float sum = 0.0f;
#pragma unroll
        for (int q = 0; q < 6; q++) sum = (st & (1u << q)) ? sum + zq[q] : sum;
        lz[c] = (rc + sum) * rcp_or0(st_diag(st));
    }
    __syncthreads();
}

// restrict f -> c after f's black sweep, for c's FLUID cells (c's z is still the fill's zero)
__device__ __forceinline__ void t_restrict(const TailLevel &f, const TailLevel &c, float *lz, float *lr,
                                           const uint16_t *ls, const uint16_t *lcl) {
    const int n = c.cnt[0] + c.cnt[1], last = f.o + t_cells(f) - 1;
    const int32_t off[6] = {-1, 1, -f.nx, f.nx, -f.nx * f.ny, f.nx * f.ny};
    const int red[4] = {0, 3, 5, 6};
    for (int t = threadIdx.x; t < n; t += TBLOCK) {
        const int cc = t_entry(c, lcl, t);
        int ci, cj, ck;
        cell_ijk(c.nx, c.ny, (uint32_t)(cc - c.o), ci, cj, ck);
        bool in[4];
        uint32_t st[4];
        float rc[4], zc[4], zn[4][6];
#pragma unroll
        for (int m = 0; m < 4; m++) {
            const int q = red[m];
            const int fi = 2 * ci + (q & 1), fj = 2 * cj + ((q >> 1) & 1), fk = 2 * ck + (q >> 2);
            in[m] = fi < f.nx && fj < f.ny && fk < f.nz;
            const int id = in[m] ? f.o + fi + f.nx * (fj + f.ny * fk) : f.o;
            st[m] = ls[id];
            rc[m] = lr[id];
            zc[m] = lz[id];
#pragma unroll
            for (int e = 0; e < 6; e++) zn[m][e] = lz[min(max(id + off[e], f.o), last)];
        }
        float acc = 0.0f;
#pragma unroll
        for (int m = 0; m < 4; m++) {
            if (!in[m] || !(st[m] & STEN_FLUID)) continue;
            float sum = 0.0f;
#pragma unroll
            for (int e = 0; e < 6; e++) sum = (st[m] & (1u << e)) ? sum + zn[m][e] : sum;
            acc += rc[m] - st_diag(st[m]) * zc[m] + sum;
        }
        lr[cc] = 0.5f * acc;
    }
    __syncthreads();
}

// black update of f from red neighbours' z + prolonged correction from c
__device__ __forceinline__ void t_post(const TailLevel &f, const TailLevel &c, float *lz, const float *lr,
                                       const uint16_t *ls, const uint16_t *lcl) {
    const int n = f.cnt[1], last = f.o + t_cells(f) - 1;
    for (int t = threadIdx.x; t < n; t += TBLOCK) {
        const int id = lcl[last - t];
        int i, j, k;
        cell_ijk(f.nx, f.ny, (uint32_t)(id - f.o), i, j, k);
        const int nb[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k}, {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
        const uint32_t st = ls[id];
        const float rc = lr[id];
        float zq[6], pq[6];
#pragma unroll
        for (int q = 0; q < 6; q++) {
            const int a = min(max(nb[q][0], 0), f.nx - 1), b = min(max(nb[q][1], 0), f.ny - 1),
                      d = min(max(nb[q][2], 0), f.nz - 1);
            zq[q] = lz[f.o + a + f.nx * (b + f.ny * d)];
            pq[q] = lz[c.o + (int)parent(c.nx, c.ny, a, b, d)];
        }
        float sum = 0.0f;
#pragma unroll
        for (int q = 0; q < 6; q++) sum = (st & (1u << q)) ? sum + (zq[q] + pq[q]) : sum;
        lz[id] = (rc + sum) * rcp_or0(st_diag(st));
    }
    __syncthreads();
}

__device__ __forceinline__ void body_mg_tail(const DevSim &S) {
    __shared__ float lz[MG_TAIL_CELLS], lr[MG_TAIL_CELLS];
    __shared__ uint16_t ls[MG_TAIL_CELLS], lcl[MG_TAIL_CELLS];
    __shared__ TailLevel tl[MG_MAX];
    __shared__ int n_tail;
    const int T = S.mg_tail, L = S.mg_levels - 1;
    // level table (constant level indices: no scratch copy of S)
    if (threadIdx.x == 0) {
        int o = 0;
#pragma unroll
        for (int q = 1; q < MG_MAX; q++) {
            if (q >= T && q <= L) {
                const MGLevel &m = S.mg[q];
                tl[q] = TailLevel{m.nx, m.ny, m.nz, o, {0, 0}, m.sten, m.r, m.hx, m.hsplit};
                o += m.nx * m.ny * m.nz;
            }
        }
        n_tail = o;
    }
    __syncthreads();
    // LDS fill and the FLUID lists.  The tail's cells are taken as one range of at most
    // MG_TAIL_CELLS (TAIL_PER_LANE per lane): every lane issues the global loads of all its
    // cells first, then writes them to LDS and lists each FLUID cell with an LDS atomic on its
    // level's colour count (in any order: every phase below updates cells independently).
    constexpr int TAIL_PER_LANE = (MG_TAIL_CELLS + TBLOCK - 1) / TBLOCK;
    const int total = n_tail;
    int lv[TAIL_PER_LANE];
    uint32_t st[TAIL_PER_LANE];
    float rv[TAIL_PER_LANE];
#pragma unroll
    for (int u = 0; u < TAIL_PER_LANE; u++) {
        const int g = u * TBLOCK + (int)threadIdx.x;
        int q = T;
        while (q < L && g >= tl[q + 1].o) q++;
        lv[u] = q;
        st[u] = 0;
        rv[u] = 0.0f;
        if (g < total) {
            const TailLevel &v = tl[q];
            int i, j, k;
            cell_ijk(v.nx, v.ny, (uint32_t)(g - v.o), i, j, k);
            const uint32_t sg = (((uint32_t)(i + j + k) & 1u) ? v.hsplit : 0u) + (uint32_t)(i >> 1) +
                                (uint32_t)v.hx * ((uint32_t)j + (uint32_t)v.ny * (uint32_t)k);
            // the level table sits in LDS, so its pointers are read as global memory explicitly
            // (a generic pointer would make these flat loads); every level has its r, so it is
            // loaded unconditionally and kept only for the first tail level
            st[u] = ld_global(v.sten, sg);
            const float r = ld_global(v.r, sg);
            rv[u] = q == T ? r : 0.0f;
        }
    }
#pragma unroll
    for (int u = 0; u < TAIL_PER_LANE; u++) {
        const int g = u * TBLOCK + (int)threadIdx.x;
        if (g >= total) break;
        const int q = lv[u];
        TailLevel &v = tl[q];
        const int c = g - v.o;
        ls[g] = (uint16_t)st[u];
        lz[g] = 0.0f;
        if (q == T) lr[g] = rv[u];
        if (st[u] & STEN_FLUID) {
            int i, j, k;
            cell_ijk(v.nx, v.ny, (uint32_t)c, i, j, k);
            const int col = (i + j + k) & 1;
            const int slot = atomicAdd(&v.cnt[col], 1);
            lcl[col ? v.o + t_cells(v) - 1 - slot : v.o + slot] = (uint16_t)g;
        }
    }
    __syncthreads();
    // down: pre-smooth (two red-black sweeps from zero) and restrict
    for (int q = T; q < L; q++) {
        const TailLevel f = tl[q], c = tl[q + 1];
        for (int s = 0; s < 2; s++) {
            t_sweep(f, lz, lr, ls, lcl, 0);
            t_sweep(f, lz, lr, ls, lcl, 1);
        }
        t_restrict(f, c, lz, lr, ls, lcl);
    }
    // coarsest: symmetric sweeps
    {
        const TailLevel v = tl[L];
        for (int s = 0; s < COARSEST_SWEEPS; s++) {
            t_sweep(v, lz, lr, ls, lcl, 0);
            t_sweep(v, lz, lr, ls, lcl, 1);
        }
        for (int s = 0; s < COARSEST_SWEEPS; s++) {
            t_sweep(v, lz, lr, ls, lcl, 1);
            t_sweep(v, lz, lr, ls, lcl, 0);
        }
    }
    // up: correct + post-smooth (black then red, twice)
    for (int q = L - 1; q >= T; q--) {
        const TailLevel f = tl[q], c = tl[q + 1];
        t_post(f, c, lz, lr, ls, lcl);
        t_sweep(f, lz, lr, ls, lcl, 0);
        t_sweep(f, lz, lr, ls, lcl, 1);
        t_sweep(f, lz, lr, ls, lcl, 0);
    }
#pragma unroll
    for (int q = 1; q < MG_MAX; q++) {
        if (q == T) {
            const MGLevel &m = S.mg[q];
            const int n = m.nx * m.ny * m.nz;
            for (int c = threadIdx.x; c < n; c += TBLOCK) {
                int i, j, k;
                cell_ijk(m.nx, m.ny, (uint32_t)c, i, j, k);
                m.z[lsig(m, i, j, k)] = lz[c];
            }
        }
    }
    __syncthreads();  // the LDS is reused by the next simulation of a batched launch
}
// The mid levels (mg_mid .. mg_tail-1) and the tail in one workgroup: the mid levels' down
// half (pre-smoothing, restriction) with the list-level bodies over the whole level (b = 0 of
// nb = 1) and a workgroup barrier after each, the LDS tail, then their up half.  Each cell is
// updated by the same arithmetic as in the grid-wide launches, so the result is the same.
// `cnt` is a batch member's count slots (null: the levels' own counts).  Level indices are
// compile-time constants (no scratch copy of S).
__device__ __forceinline__ void body_mg_midtail(const DevSim &S, uint32_t *cnt) {
    const int M = S.mg_mid, T = S.mg_tail;
#pragma unroll
    for (int q = 1; q < MG_MAX - 1; q++) {
        if (q >= M && q < T) {
            const MGLevel v = level_counts(S, cnt, q), c = level_counts(S, cnt, q + 1);
            body_lc_pre0(v, 0, 1);
            __syncthreads();
            for (int r = 1; r < MG_COARSE_ROUNDS; r++) {
                body_lc_sweep(v, 0, 0, 1);
                __syncthreads();
                body_lc_sweep(v, 1, 0, 1);
                __syncthreads();
            }
            body_restrict(v, c, 0, 1);
            __syncthreads();
        }
    }
    body_mg_tail(S);
#pragma unroll
    for (int q = MG_MAX - 2; q >= 1; q--) {
        if (q >= M && q < T) {
            const MGLevel v = level_counts(S, cnt, q), c = level_counts(S, cnt, q + 1);
            body_lc_post(v, c, 0, 1);
            __syncthreads();
            body_lc_sweep(v, 0, 0, 1);
            __syncthreads();
            for (int r = 1; r < MG_COARSE_ROUNDS; r++) {
                body_lc_sweep(v, 1, 0, 1);
                __syncthreads();
                body_lc_sweep(v, 0, 0, 1);
                __syncthreads();
            }
        }
    }
}
__global__ __launch_bounds__(TBLOCK) void k_mg_tail(DevSim S) { body_mg_midtail(S, nullptr); }

// ---- single-sim launches (grid = level_blocks workgroups for the list kernels)
__global__ __launch_bounds__(LBLOCK) void k_l0_pre0(DevSim S) {
    body_l0_pre0(S, S.fl_count, S.cg_r0, S.cg_r1, blockIdx.x, gridDim.x);
}
__global__ __launch_bounds__(LBLOCK) void k_l0_sweep(DevSim S, uint32_t colour) {
    body_l0_sweep(S, S.fl_count, S.cg_r0, S.cg_r1, colour, blockIdx.x, gridDim.x);
}
__global__ __launch_bounds__(LBLOCK) void k_l0_post(DevSim S, MGLevel c) {
    body_l0_post(S, S.fl_count, c, S.cg_r0, S.cg_r1, blockIdx.x, gridDim.x);
}
__global__ __launch_bounds__(LBLOCK) void k_lc_pre0(MGLevel v) { body_lc_pre0(v, blockIdx.x, gridDim.x); }
__global__ __launch_bounds__(LBLOCK) void k_lc_sweep(MGLevel v, uint32_t colour) {
    body_lc_sweep(v, colour, blockIdx.x, gridDim.x);
}
__global__ __launch_bounds__(LBLOCK) void k_lc_post(MGLevel v, MGLevel c) { body_lc_post(v, c, blockIdx.x, gridDim.x); }
__global__ __launch_bounds__(LBLOCK) void k_restrict(MGLevel f, MGLevel c) {
    body_restrict(f, c, blockIdx.x, gridDim.x);
}
__global__ __launch_bounds__(LBLOCK) void k_restrict0(DevSim S, MGLevel c) { body_restrict0(S, c, blockIdx.x, gridDim.x); }

inline unsigned grid_for(uint32_t n) { return (n + MBLOCK - 1) / MBLOCK; }
inline uint32_t cells(const MGLevel &m) { return (uint32_t)(m.nx * m.ny * m.nz); }

// The simulation grid as a level (its types and stencil words).
__host__ __device__ inline MGLevel level0(const DevSim &S) {
    MGLevel v{};
    v.nx = S.nx;
    v.ny = S.ny;
    v.nz = S.nz;
    v.type = S.type;
    v.sten = S.sten;
    return v;
}

// ---- batched: level l of each simulation in the work list
__global__ __launch_bounds__(MBLOCK) void kb_mg_types(WorkList W, int l) {
    for_each_work(W, [&](const DevSim &S, uint32_t b, const uint32_t *) {
        const MGLevel f = l == 1 ? level0(S) : S.mg[l - 1], c = S.mg[l];
        body_mg_types(f, c, b);
    });
}
__global__ __launch_bounds__(MBLOCK) void kb_mg_sten(WorkList W, int l) {
    for_each_work(W, [&](const DevSim &S, uint32_t b, const uint32_t *) {
        const MGLevel v = S.mg[l];
        body_mg_sten(v, b);
    });
}
__global__ __launch_bounds__(TBLOCK) void kb_mg_scan(WorkList W) {
    for_each_work(W, [&](const DevSim &S, uint32_t, uint32_t *cnt) { body_mg_scan(S, cnt); });
}
__global__ __launch_bounds__(MBLOCK) void kb_mg_compact(WorkList W, int l) {
    for_each_work(W, [&](const DevSim &S, uint32_t b, const uint32_t *) {
        const MGLevel v = S.mg[l];
        body_mg_compact(v, b);
    });
}
// A member's level copy with its count pointer taken from the batch's slots (known from the
// work entry alone, so the count loads alongside the level's fields).
__device__ __forceinline__ MGLevel level_in_batch(const DevSim &S, uint32_t *cnt, int l) {
    MGLevel v = S.mg[l];
    v.count = cnt + 4 * l;
    return v;
}
__global__ __launch_bounds__(LBLOCK) void kb_pre0(WorkList W, int l) {
    for_each_work(W, [&](const DevSim &S, uint32_t b, uint32_t *cnt) {
        if (l == 0) {
            body_l0_pre0(S, cnt, S.cg_r0, S.cg_r1, b, S.cg_blocks);
        } else {
            const MGLevel v = level_in_batch(S, cnt, l);
            body_lc_pre0(v, b, level_blocks(S.cg_blocks, l));
        }
    });
}
__global__ __launch_bounds__(LBLOCK) void kb_sweep(WorkList W, int l, uint32_t colour) {
    for_each_work(W, [&](const DevSim &S, uint32_t b, uint32_t *cnt) {
        if (l == 0) {
            body_l0_sweep(S, cnt, S.cg_r0, S.cg_r1, colour, b, S.cg_blocks);
        } else {
            const MGLevel v = level_in_batch(S, cnt, l);
            body_lc_sweep(v, colour, b, level_blocks(S.cg_blocks, l));
        }
    });
}
__global__ __launch_bounds__(LBLOCK) void kb_restrict(WorkList W, int l) {
    for_each_work(W, [&](const DevSim &S, uint32_t b, uint32_t *cnt) {
        const MGLevel c = level_in_batch(S, cnt, l + 1);
        if (l == 0) {
            body_restrict0(S, c, b, S.cg_blocks);
        } else {
            const MGLevel f = S.mg[l];
            body_restrict(f, c, b, level_blocks(S.cg_blocks, l));
        }
    });
}
__global__ __launch_bounds__(LBLOCK) void kb_post(WorkList W, int l) {
    for_each_work(W, [&](const DevSim &S, uint32_t b, uint32_t *cnt) {
        const MGLevel c = S.mg[l + 1];
        if (l == 0) {
            body_l0_post(S, cnt, c, S.cg_r0, S.cg_r1, b, S.cg_blocks);
        } else {
            const MGLevel v = level_in_batch(S, cnt, l);
            body_lc_post(v, c, b, level_blocks(S.cg_blocks, l));
        }
    });
}
__global__ __launch_bounds__(TBLOCK) void kb_mg_tail(WorkList W) {
    for_each_work(W, [&](const DevSim &S, uint32_t, uint32_t *cnt) { body_mg_midtail(S, cnt); });
}

}  // namespace

void launch_mg_types(const DevSim &S, hipStream_t s) {
    for (int l = 1; l < S.mg_levels; l++) {
        const MGLevel f = l == 1 ? level0(S) : S.mg[l - 1];
        PXL_LAUNCH(k_mg_types, grid_for(cells(S.mg[l])), MBLOCK, 0, s, f, S.mg[l]);
        PXL_LAUNCH(k_mg_sten, grid_for(cells(S.mg[l])), MBLOCK, 0, s, S.mg[l]);
    }
    PXL_LAUNCH(k_mg_scan, 1, TBLOCK, 0, s, S);
    for (int l = 1; l < S.mg_levels; l++) PXL_LAUNCH(k_mg_compact, grid_for(cells(S.mg[l])), MBLOCK, 0, s, S.mg[l]);
}

// u = M r with r = cg_r0, u = cg_r1 (the vectors of the MG-CG iteration).
void launch_mg_vcycle(const DevSim &S, const float *r, float *u, hipStream_t s) {
    (void)r;
    (void)u;
    const int T = S.mg_mid;  // levels from mg_mid on run in the tail's workgroup
    for (int l = 0; l < T; l++) {
        const unsigned g = level_blocks(S.cg_blocks, l);
        if (l == 0) {
            PXL_LAUNCH(k_l0_pre0, g, LBLOCK, 0, s, S);
            PXL_LAUNCH(k_l0_sweep, g, LBLOCK, 0, s, S, 0);
            PXL_LAUNCH(k_l0_sweep, g, LBLOCK, 0, s, S, 1);
            PXL_LAUNCH(k_restrict0, g, LBLOCK, 0, s, S, S.mg[1]);
        } else {
            PXL_LAUNCH(k_lc_pre0, g, LBLOCK, 0, s, S.mg[l]);
            for (int q = 1; q < MG_COARSE_ROUNDS; q++) {
                PXL_LAUNCH(k_lc_sweep, g, LBLOCK, 0, s, S.mg[l], 0);
                PXL_LAUNCH(k_lc_sweep, g, LBLOCK, 0, s, S.mg[l], 1);
            }
            PXL_LAUNCH(k_restrict, g, LBLOCK, 0, s, S.mg[l], S.mg[l + 1]);
        }
    }
    PXL_LAUNCH(k_mg_tail, 1, TBLOCK, 0, s, S);
    for (int l = T - 1; l >= 0; l--) {
        const unsigned g = level_blocks(S.cg_blocks, l);
        if (l == 0) {
            PXL_LAUNCH(k_l0_post, g, LBLOCK, 0, s, S, S.mg[1]);
            PXL_LAUNCH(k_l0_sweep, g, LBLOCK, 0, s, S, 0);
            PXL_LAUNCH(k_l0_sweep, g, LBLOCK, 0, s, S, 1);
            PXL_LAUNCH(k_l0_sweep, g, LBLOCK, 0, s, S, 0);
        } else {
            PXL_LAUNCH(k_lc_post, g, LBLOCK, 0, s, S.mg[l], S.mg[l + 1]);
            PXL_LAUNCH(k_lc_sweep, g, LBLOCK, 0, s, S.mg[l], 0);
            for (int q = 1; q < MG_COARSE_ROUNDS; q++) {
                PXL_LAUNCH(k_lc_sweep, g, LBLOCK, 0, s, S.mg[l], 1);
                PXL_LAUNCH(k_lc_sweep, g, LBLOCK, 0, s, S.mg[l], 0);
            }
        }
    }
}

// ---- batched
uint32_t mg_cell_blocks(const DevSim &S, int l) { return grid_for(cells(l == 0 ? level0(S) : S.mg[l])); }

void batch_mg_types(const WorkLists &W, int l, hipStream_t s) { PXL_LAUNCH(kb_mg_types, batch_grid(W[WD_MGTYPES + l], batch_cap(BF_MG_GRID)), MBLOCK, 0, s, W[WD_MGTYPES + l], l); }
void batch_mg_sten(const WorkLists &W, int l, hipStream_t s) { PXL_LAUNCH(kb_mg_sten, batch_grid(W[WD_MGTYPES + l], batch_cap(BF_MG_GRID)), MBLOCK, 0, s, W[WD_MGTYPES + l], l); }
void batch_mg_scan(const WorkLists &W, hipStream_t s) { PXL_LAUNCH(kb_mg_scan, batch_grid(W[WD_SIM], batch_cap(BF_MG_TAIL)), TBLOCK, 0, s, W[WD_SIM]); }
void batch_mg_compact(const WorkLists &W, int l, hipStream_t s) { PXL_LAUNCH(kb_mg_compact, batch_grid(W[WD_MGTYPES + l], batch_cap(BF_MG_GRID)), MBLOCK, 0, s, W[WD_MGTYPES + l], l); }
void batch_mg_pre0(const WorkLists &W, int l, hipStream_t s) { PXL_LAUNCH(kb_pre0, batch_grid(W[WD_A_LVL + l], batch_cap(BF_MG_LIST)), LBLOCK, 0, s, W[WD_A_LVL + l], l); }
void batch_mg_sweep(const WorkLists &W, int l, uint32_t colour, hipStream_t s) {
    PXL_LAUNCH(kb_sweep, batch_grid(W[WD_A_LVL + l], batch_cap(BF_MG_LIST)), LBLOCK, 0, s, W[WD_A_LVL + l], l, colour);
}
void batch_mg_restrict(const WorkLists &W, int l, hipStream_t s) {
    PXL_LAUNCH(kb_restrict, batch_grid(W[WD_A_LVL + l], batch_cap(BF_MG_LIST)), LBLOCK, 0, s, W[WD_A_LVL + l], l);
}
void batch_mg_post(const WorkLists &W, int l, hipStream_t s) { PXL_LAUNCH(kb_post, batch_grid(W[WD_A_LVL + l], batch_cap(BF_MG_LIST)), LBLOCK, 0, s, W[WD_A_LVL + l], l); }
void batch_mg_tail(const WorkLists &W, hipStream_t s) { PXL_LAUNCH(kb_mg_tail, batch_grid(W[WD_A_SIM], batch_cap(BF_MG_TAIL)), TBLOCK, 0, s, W[WD_A_SIM]); }

}  // namespace pxl
