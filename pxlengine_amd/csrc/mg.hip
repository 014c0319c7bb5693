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
// Levels from mg_tail on (a few thousand cells together) run the same cycle in one
// workgroup with everything in LDS.  Arithmetic is fp32 (the FP64 refinement outside fixes
// the rounding).  Kernels take the levels they touch as MGLevel values: a runtime index into
// DevSim's level array would copy the whole structure into scratch memory.
#include "common.hpp"
#include "kernels.hpp"

namespace pxl {

namespace {

constexpr int MBLOCK = 256;
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
// Index of the coarse cell containing fine cell (i, j, k).
__device__ __forceinline__ uint32_t parent(int cnx, int cny, int i, int j, int k) {
    return (uint32_t)(i >> 1) + (uint32_t)cnx * ((uint32_t)(j >> 1) + (uint32_t)cny * (uint32_t)(k >> 1));
}
// Lane t of a colour launch (ceil(nx/2) lanes per row) -> cell (i, j, k) of that colour;
// false past the grid.
__device__ __forceinline__ bool colour_cell(const MGLevel &v, uint32_t t, int colour, int &i, int &j, int &k) {
    const uint32_t nxh = (uint32_t)(v.nx + 1) >> 1;
    const uint32_t row = t / nxh;
    if (row >= (uint32_t)(v.ny * v.nz)) return false;
    j = (int)(row % (uint32_t)v.ny);
    k = (int)(row / (uint32_t)v.ny);
    i = 2 * (int)(t % nxh) + ((j + k + colour) & 1);
    return i < v.nx;
}

// ---- level types: c from its children in f
__global__ __launch_bounds__(MBLOCK) void k_mg_types(MGLevel f, MGLevel c) {
    const uint32_t id = blockIdx.x * MBLOCK + threadIdx.x;
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

// ---- level stencil words: FLUID bit, non-SOLID-neighbour count (bits 8-10), FLUID-neighbour
// mask (bits 0-5), the layout of the simulation grid's stencil words
__global__ __launch_bounds__(MBLOCK) void k_mg_sten(MGLevel v) {
    const uint32_t id = blockIdx.x * MBLOCK + threadIdx.x;
    if (id >= ncell(v)) return;
    int i, j, k;
    cell_ijk(v.nx, v.ny, id, i, j, k);
    uint32_t st = 0;
    if (v.type[id] == CELL_FLUID) {
        const int nb[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k}, {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
        uint32_t diag = 0;
#pragma unroll
        for (int n = 0; n < 6; n++) {
            const uint8_t t = level_type(v, nb[n][0], nb[n][1], nb[n][2]);
            if (t != CELL_SOLID) diag++;
            if (t == CELL_FLUID) st |= 1u << n;
        }
        st |= STEN_FLUID | (diag << 8);
    }
    v.sten[id] = (uint16_t)st;
}

// ---- pre0: the first red-black sweep from zero.  One lane per cell; non-FLUID cells get
// z = 0 when `zero_rest` (coarse levels: the prolongation reads every coarse z).
__global__ __launch_bounds__(MBLOCK) void k_rb_pre0(MGLevel v, const float *__restrict__ r, float *__restrict__ z,
                                                    int zero_rest) {
    const uint32_t c = blockIdx.x * MBLOCK + threadIdx.x;
    if (c >= ncell(v)) return;
    int i, j, k;
    cell_ijk(v.nx, v.ny, c, i, j, k);
    const bool black = (i + j + k) & 1;
    const uint32_t last = ncell(v) - 1u;
    const int32_t off[6] = {-1, 1, -v.nx, v.nx, -v.nx * v.ny, v.nx * v.ny};
    float rn[6];
    uint32_t sn[6];
    if (black) {  // neighbours' r and stencil words, clamped addresses, masked after
#pragma unroll
        for (int n = 0; n < 6; n++) {
            const uint32_t q = min(c + (uint32_t)off[n], last);
            rn[n] = r[q];
            sn[n] = v.sten[q];
        }
    }
    const float rc = r[c];
    const uint32_t st = v.sten[c];
    if (!(st & STEN_FLUID)) {
        if (zero_rest) z[c] = 0.0f;
        return;
    }
    float sum = 0.0f;
    if (black) {
#pragma unroll
        for (int n = 0; n < 6; n++) sum += (st & (1u << n)) ? rn[n] * rcp_or0(st_diag(sn[n])) : 0.0f;
    }
    z[c] = (rc + sum) * rcp_or0(st_diag(st));
}

// ---- one colour's Gauss-Seidel update (one lane per cell of that colour)
__global__ __launch_bounds__(MBLOCK) void k_rb_sweep(MGLevel v, const float *__restrict__ r, float *__restrict__ z,
                                                     int colour) {
    int i, j, k;
    if (!colour_cell(v, blockIdx.x * MBLOCK + threadIdx.x, colour, i, j, k)) return;
    const uint32_t c = lidx(v, i, j, k);
    const uint32_t last = ncell(v) - 1u;
    const int32_t off[6] = {-1, 1, -v.nx, v.nx, -v.nx * v.ny, v.nx * v.ny};
    // every load issued before the stencil word is known (clamped addresses, masked after)
    float zn[6];
#pragma unroll
    for (int n = 0; n < 6; n++) zn[n] = z[min(c + (uint32_t)off[n], last)];
    const float rc = r[c];
    const uint32_t st = v.sten[c];
    if (!(st & STEN_FLUID)) return;
    float sum = 0.0f;
#pragma unroll
    for (int n = 0; n < 6; n++) sum += (st & (1u << n)) ? zn[n] : 0.0f;
    z[c] = (rc + sum) * rcp_or0(st_diag(st));
}

// ---- restrict after a black sweep: c.r = 1/2 sum over FLUID red children of r - L z (one lane
// per coarse cell; the red children are q = 0, 3, 5, 6).  Loads use clamped addresses and
// masks so that the loads of a lane do not wait on each other.
__global__ __launch_bounds__(MBLOCK) void k_rb_restrict(MGLevel f, const float *__restrict__ r,
                                                        const float *__restrict__ z, MGLevel c) {
    const uint32_t cc = blockIdx.x * MBLOCK + threadIdx.x;
    if (cc >= ncell(c)) return;
    int ci, cj, ck;
    cell_ijk(c.nx, c.ny, cc, ci, cj, ck);
    const uint32_t nf = ncell(f);
    const int32_t off[6] = {-1, 1, -f.nx, f.nx, -f.nx * f.ny, f.nx * f.ny};
    const int red[4] = {0, 3, 5, 6};
    float acc = 0.0f;
#pragma unroll
    for (int m = 0; m < 4; m++) {
        const int q = red[m];
        const int fi = 2 * ci + (q & 1), fj = 2 * cj + ((q >> 1) & 1), fk = 2 * ck + (q >> 2);
        const bool in = fi < f.nx && fj < f.ny && fk < f.nz;
        const uint32_t id = in ? lidx(f, fi, fj, fk) : 0u;
        const uint32_t st = in ? (uint32_t)f.sten[id] : 0u;
        float sum = 0.0f;
#pragma unroll
        for (int n = 0; n < 6; n++) {
            const float zn = z[min(id + (uint32_t)off[n], nf - 1u)];
            sum += (st & (1u << n)) ? zn : 0.0f;
        }
        const float res = r[id] - st_diag(st) * z[id] + sum;
        acc += (st & STEN_FLUID) ? res : 0.0f;
    }
    c.r[cc] = 0.5f * acc;
}

// ---- post: black update from red neighbours' z + prolonged coarse correction c.z
__global__ __launch_bounds__(MBLOCK) void k_rb_post(MGLevel f, const float *__restrict__ r, float *__restrict__ z,
                                                    MGLevel c) {
    int i, j, k;
    if (!colour_cell(f, blockIdx.x * MBLOCK + threadIdx.x, 1, i, j, k)) return;
    const uint32_t id = lidx(f, i, j, k);
    // neighbour coordinates clamped into the grid: every load is issued before the stencil
    // word is known and masked after
    const int im = max(i - 1, 0), ip = min(i + 1, f.nx - 1), jm = max(j - 1, 0), jp = min(j + 1, f.ny - 1),
              km = max(k - 1, 0), kp = min(k + 1, f.nz - 1);
    const int nb[6][3] = {{im, j, k}, {ip, j, k}, {i, jm, k}, {i, jp, k}, {i, j, km}, {i, j, kp}};
    float zn[6];
#pragma unroll
    for (int n = 0; n < 6; n++)
        zn[n] = z[lidx(f, nb[n][0], nb[n][1], nb[n][2])] + c.z[parent(c.nx, c.ny, nb[n][0], nb[n][1], nb[n][2])];
    const float rc = r[id];
    const uint32_t st = f.sten[id];
    if (!(st & STEN_FLUID)) return;
    float sum = 0.0f;
#pragma unroll
    for (int n = 0; n < 6; n++) sum += (st & (1u << n)) ? zn[n] : 0.0f;
    z[id] = (rc + sum) * rcp_or0(st_diag(st));
}

// ---- the tail: levels mg_tail..L in one workgroup.  LDS holds every tail level's z, r and
// stencil words back to back; the cycle is the one above with barriers in place of launches.
struct TailLevel {
    int nx, ny, nz, o;  // dims and LDS offset
};
constexpr uint32_t LDS_BLACK = 1u << 12;  // LDS stencil words carry the cell's colour

__device__ __forceinline__ void t_sweep(const TailLevel &v, float *lz, const float *lr, const uint16_t *ls,
                                        int colour) {
    const int n = v.nx * v.ny * v.nz;
    const int32_t off[6] = {-1, 1, -v.nx, v.nx, -v.nx * v.ny, v.nx * v.ny};
    const uint32_t want = STEN_FLUID | (colour ? LDS_BLACK : 0u);
    for (int c = threadIdx.x; c < n; c += TBLOCK) {
        const uint32_t st = ls[v.o + c];
        if ((st & (STEN_FLUID | LDS_BLACK)) != want) continue;
        float sum = 0.0f;
#pragma unroll
        for (int q = 0; q < 6; q++)
            if (st & (1u << q)) sum += lz[v.o + c + off[q]];
        lz[v.o + c] = (lr[v.o + c] + sum) * rcp_or0(st_diag(st));
    }
    __syncthreads();
}

// restrict f -> c after f's black sweep; also zeroes c's z
__device__ __forceinline__ void t_restrict(const TailLevel &f, const TailLevel &c, float *lz, float *lr,
                                           const uint16_t *ls) {
    const int n = c.nx * c.ny * c.nz;
    const int32_t off[6] = {-1, 1, -f.nx, f.nx, -f.nx * f.ny, f.nx * f.ny};
    const int red[4] = {0, 3, 5, 6};
    for (int cc = threadIdx.x; cc < n; cc += TBLOCK) {
        int ci, cj, ck;
        cell_ijk(c.nx, c.ny, (uint32_t)cc, ci, cj, ck);
        float acc = 0.0f;
#pragma unroll
        for (int m = 0; m < 4; m++) {
            const int q = red[m];
            const int fi = 2 * ci + (q & 1), fj = 2 * cj + ((q >> 1) & 1), fk = 2 * ck + (q >> 2);
            if (fi >= f.nx || fj >= f.ny || fk >= f.nz) continue;
            const int id = f.o + fi + f.nx * (fj + f.ny * fk);
            const uint32_t st = ls[id];
            if (!(st & STEN_FLUID)) continue;
            float sum = 0.0f;
#pragma unroll
            for (int e = 0; e < 6; e++)
                if (st & (1u << e)) sum += lz[id + off[e]];
            acc += lr[id] - st_diag(st) * lz[id] + sum;
        }
        lr[c.o + cc] = 0.5f * acc;
        lz[c.o + cc] = 0.0f;
    }
    __syncthreads();
}

// black update of f from red neighbours' z + prolonged correction from c
__device__ __forceinline__ void t_post(const TailLevel &f, const TailLevel &c, float *lz, const float *lr,
                                       const uint16_t *ls) {
    const int n = f.nx * f.ny * f.nz;
    for (int id = threadIdx.x; id < n; id += TBLOCK) {
        const uint32_t st = ls[f.o + id];
        if ((st & (STEN_FLUID | LDS_BLACK)) != (STEN_FLUID | LDS_BLACK)) continue;
        int i, j, k;
        cell_ijk(f.nx, f.ny, (uint32_t)id, i, j, k);
        const int nb[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k}, {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
        float sum = 0.0f;
#pragma unroll
        for (int q = 0; q < 6; q++) {
            if (st & (1u << q)) {
                const int a = nb[q][0], b = nb[q][1], d = nb[q][2];
                sum += lz[f.o + a + f.nx * (b + f.ny * d)] + lz[c.o + (int)parent(c.nx, c.ny, a, b, d)];
            }
        }
        lz[f.o + id] = (lr[f.o + id] + sum) * rcp_or0(st_diag(st));
    }
    __syncthreads();
}

__global__ __launch_bounds__(TBLOCK) void k_mg_tail(DevSim S) {
    __shared__ float lz[MG_TAIL_CELLS], lr[MG_TAIL_CELLS];
    __shared__ uint16_t ls[MG_TAIL_CELLS];
    __shared__ TailLevel tl[MG_MAX];
    const int T = S.mg_tail, L = S.mg_levels - 1;
    int o = 0;
    // level table and LDS fill (constant level indices: no scratch copy of S)
#pragma unroll
    for (int q = 1; q < MG_MAX; q++) {
        if (q >= T && q <= L) {
            const MGLevel &m = S.mg[q];
            const int n = m.nx * m.ny * m.nz;
            if (threadIdx.x == 0) tl[q] = TailLevel{m.nx, m.ny, m.nz, o};
            for (int c = threadIdx.x; c < n; c += TBLOCK) {
                int i, j, k;
                cell_ijk(m.nx, m.ny, (uint32_t)c, i, j, k);
                ls[o + c] = (uint16_t)(m.sten[c] | (((i + j + k) & 1) ? LDS_BLACK : 0u));
                lz[o + c] = 0.0f;
                if (q == T) lr[o + c] = m.r[c];
            }
            o += n;
        }
    }
    __syncthreads();
    // down: pre-smooth (two red-black sweeps from zero) and restrict
    for (int q = T; q < L; q++) {
        const TailLevel f = tl[q], c = tl[q + 1];
        for (int s = 0; s < 2; s++) {
            t_sweep(f, lz, lr, ls, 0);
            t_sweep(f, lz, lr, ls, 1);
        }
        t_restrict(f, c, lz, lr, ls);
    }
    // coarsest: symmetric sweeps
    {
        const TailLevel v = tl[L];
        for (int s = 0; s < COARSEST_SWEEPS; s++) {
            t_sweep(v, lz, lr, ls, 0);
            t_sweep(v, lz, lr, ls, 1);
        }
        for (int s = 0; s < COARSEST_SWEEPS; s++) {
            t_sweep(v, lz, lr, ls, 1);
            t_sweep(v, lz, lr, ls, 0);
        }
    }
    // up: correct + post-smooth (black then red, twice)
    for (int q = L - 1; q >= T; q--) {
        const TailLevel f = tl[q], c = tl[q + 1];
        t_post(f, c, lz, lr, ls);
        t_sweep(f, lz, lr, ls, 0);
        t_sweep(f, lz, lr, ls, 1);
        t_sweep(f, lz, lr, ls, 0);
    }
#pragma unroll
    for (int q = 1; q < MG_MAX; q++) {
        if (q == T) {
            const MGLevel &m = S.mg[q];
            const int n = m.nx * m.ny * m.nz;
            for (int c = threadIdx.x; c < n; c += TBLOCK) m.z[c] = lz[c];
        }
    }
}

inline unsigned grid_for(uint32_t n) { return (n + MBLOCK - 1) / MBLOCK; }
inline uint32_t cells(const MGLevel &m) { return (uint32_t)(m.nx * m.ny * m.nz); }
inline uint32_t colour_cells(const MGLevel &m) { return (uint32_t)((m.nx + 1) / 2) * (uint32_t)(m.ny * m.nz); }

// The simulation grid as a level (its types and stencil words).
inline MGLevel level0(const DevSim &S) {
    MGLevel v{};
    v.nx = S.nx;
    v.ny = S.ny;
    v.nz = S.nz;
    v.type = S.type;
    v.sten = S.sten;
    return v;
}

}  // namespace

void launch_mg_types(const DevSim &S, hipStream_t s) {
    for (int l = 1; l < S.mg_levels; l++) {
        const MGLevel f = l == 1 ? level0(S) : S.mg[l - 1];
        k_mg_types<<<grid_for(cells(S.mg[l])), MBLOCK, 0, s>>>(f, S.mg[l]);
        k_mg_sten<<<grid_for(cells(S.mg[l])), MBLOCK, 0, s>>>(S.mg[l]);
    }
}

void launch_mg_vcycle(const DevSim &S, const float *r, float *u, hipStream_t s) {
    const int T = S.mg_tail;
    auto lev = [&](int l) { return l == 0 ? level0(S) : S.mg[l]; };
    auto rhs = [&](int l) { return l == 0 ? r : (const float *)S.mg[l].r; };
    auto itr = [&](int l) { return l == 0 ? u : S.mg[l].z; };
    for (int l = 0; l < T; l++) {
        const MGLevel f = lev(l);
        const unsigned gc = grid_for(colour_cells(f));
        k_rb_pre0<<<grid_for(cells(f)), MBLOCK, 0, s>>>(f, rhs(l), itr(l), l > 0);
        k_rb_sweep<<<gc, MBLOCK, 0, s>>>(f, rhs(l), itr(l), 0);
        k_rb_sweep<<<gc, MBLOCK, 0, s>>>(f, rhs(l), itr(l), 1);
        k_rb_restrict<<<grid_for(cells(S.mg[l + 1])), MBLOCK, 0, s>>>(f, rhs(l), itr(l), S.mg[l + 1]);
    }
    k_mg_tail<<<1, TBLOCK, 0, s>>>(S);
    for (int l = T - 1; l >= 0; l--) {
        const MGLevel f = lev(l);
        const unsigned gc = grid_for(colour_cells(f));
        k_rb_post<<<gc, MBLOCK, 0, s>>>(f, rhs(l), itr(l), S.mg[l + 1]);
        k_rb_sweep<<<gc, MBLOCK, 0, s>>>(f, rhs(l), itr(l), 0);
        k_rb_sweep<<<gc, MBLOCK, 0, s>>>(f, rhs(l), itr(l), 1);
        k_rb_sweep<<<gc, MBLOCK, 0, s>>>(f, rhs(l), itr(l), 0);
    }
}

}  // namespace pxl
