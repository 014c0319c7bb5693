// Geometric multigrid preconditioner for the pressure system of the mixed-precision solve
// (the reference has only fixed-count red-black Gauss-Seidel, fluid.slang:167-182; this
// solver is a build extension that converges the same matrix to a relative FP64 residual).
//
// Levels: level 0 is the simulation grid (its FLUID cells, stencil and list).  Level l+1
// halves every axis (ceil); a coarse cell is AIR if any of its 8 children is AIR, else FLUID
// if any child is FLUID, else SOLID (children past the grid edge count as SOLID).  The level-l
// operator is A_l = 4^l L_l with L_l the unit 7-point stencil on level-l types (diagonal =
// non-SOLID neighbours, -1 per FLUID neighbour): the Galerkin product of the fine operator
// with piecewise-constant transfer between a cell and its 8 children.
//
// One application z = M r is a V(1,1) cycle with damped Jacobi (omega = 0.8), restriction =
// sum over children, prolongation = injection into the children (restriction^T), and a fixed
// number of damped-Jacobi sweeps on the coarsest level inside one workgroup.  Every step is
// a fixed linear map and the post-smoother is the adjoint of the pre-smoother, so M is
// symmetric positive definite and can precondition CG.  Each level costs two launches:
//   down: pre-smooth from zero (z1 = omega D^-1 r), residual r - A z1 and its restriction,
//         over the coarse cells (each walks its 8 children);
//   up:   z = z1 + prolonged coarse correction, post-smooth z += omega D^-1 (r - A z).
// Fine-level kernels walk the fluid list (up) or the level-1 cells (down); coarse levels are
// small 3-D grids.  Arithmetic is fp32 (the FP64 refinement outside fixes the rounding).
#include "common.hpp"
#include "kernels.hpp"

namespace pxl {

namespace {

constexpr float OMEGA = 0.8f;
constexpr int MBLOCK = 256;
constexpr int CBLOCK = 1024;   // the coarsest-level workgroup
constexpr int COARSE_SWEEPS = 40;

__device__ __forceinline__ float rcp_or0(float d) { return d > 0.0f ? __builtin_amdgcn_rcpf(d) : 0.0f; }

// Type of cell (i, j, k) on a level whose types are in `t` (SOLID outside).
__device__ __forceinline__ uint8_t level_type(const uint8_t *t, int nx, int ny, int nz, int i, int j, int k) {
    if (i < 0 || j < 0 || k < 0 || i >= nx || j >= ny || k >= nz) return CELL_SOLID;
    return t[(uint32_t)i + (uint32_t)nx * ((uint32_t)j + (uint32_t)ny * (uint32_t)k)];
}

struct LevelView {
    int nx, ny, nz;
    const uint8_t *type;
    __device__ uint32_t idx(int i, int j, int k) const {
        return (uint32_t)i + (uint32_t)nx * ((uint32_t)j + (uint32_t)ny * (uint32_t)k);
    }
    __device__ uint8_t at(int i, int j, int k) const { return level_type(type, nx, ny, nz, i, j, k); }
};

__device__ __forceinline__ LevelView view(const DevSim &S, int l) {
    LevelView v;
    if (l == 0) {
        v.nx = S.nx; v.ny = S.ny; v.nz = S.nz; v.type = S.type;
    } else {
        v.nx = S.mg[l].nx; v.ny = S.mg[l].ny; v.nz = S.mg[l].nz; v.type = S.mg[l].type;
    }
    return v;
}

// Number of non-SOLID neighbours (the unit-stencil diagonal) and FLUID-neighbour mask.
__device__ __forceinline__ void stencil_of(const LevelView &v, int i, int j, int k, float &diag, uint32_t &mask) {
    const int nb[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k}, {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
    diag = 0.0f;
    mask = 0;
#pragma unroll
    for (int n = 0; n < 6; n++) {
        const uint8_t t = v.at(nb[n][0], nb[n][1], nb[n][2]);
        if (t != CELL_SOLID) diag += 1.0f;
        if (t == CELL_FLUID) mask |= 1u << n;
    }
}

// ---- level types
__global__ __launch_bounds__(MBLOCK) void k_mg_types(DevSim S, int l) {
    const LevelView f = view(S, l - 1);
    const int nx = S.mg[l].nx, ny = S.mg[l].ny, nz = S.mg[l].nz;
    const uint32_t c = blockIdx.x * MBLOCK + threadIdx.x;
    if (c >= (uint32_t)(nx * ny * nz)) return;
    const int i = (int)(c % (uint32_t)nx), j = (int)((c / (uint32_t)nx) % (uint32_t)ny), k = (int)(c / ((uint32_t)nx * (uint32_t)ny));
    bool air = false, fluid = false;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const uint8_t t = f.at(2 * i + (q & 1), 2 * j + ((q >> 1) & 1), 2 * k + (q >> 2));
        air = air || t == CELL_AIR;
        fluid = fluid || t == CELL_FLUID;
    }
    S.mg[l].type[c] = air ? CELL_AIR : (fluid ? CELL_FLUID : CELL_SOLID);
}

// ---- down: pre-smooth on level l, residual, restriction to level l+1 (one lane per coarse
// cell).  Level l's r is `rin` (the CG residual on level 0); z1 goes to `z1`.
__global__ __launch_bounds__(MBLOCK) void k_mg_down(DevSim S, int l, const float *__restrict__ rin,
                                                    float *__restrict__ z1) {
    const LevelView f = view(S, l);
    const int cx = S.mg[l + 1].nx, cy = S.mg[l + 1].ny, cz = S.mg[l + 1].nz;
    const uint32_t c = blockIdx.x * MBLOCK + threadIdx.x;
    if (c >= (uint32_t)(cx * cy * cz)) return;
    const int i = (int)(c % (uint32_t)cx), j = (int)((c / (uint32_t)cx) % (uint32_t)cy), k = (int)(c / ((uint32_t)cx * (uint32_t)cy));
    const float scale = l == 0 ? 1.0f : S.mg[l].scale;
    float acc = 0.0f;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int fi = 2 * i + (q & 1), fj = 2 * j + ((q >> 1) & 1), fk = 2 * k + (q >> 2);
        if (f.at(fi, fj, fk) != CELL_FLUID) continue;
        const uint32_t id = f.idx(fi, fj, fk);
        float diag;
        uint32_t mask;
        stencil_of(f, fi, fj, fk, diag, mask);
        const float r = rin[id];
        const float zc = OMEGA * r * rcp_or0(scale * diag);
        // A z1 with z1 = omega r / D at the FLUID neighbours
        const int nb[6][3] = {{fi - 1, fj, fk}, {fi + 1, fj, fk}, {fi, fj - 1, fk}, {fi, fj + 1, fk}, {fi, fj, fk - 1}, {fi, fj, fk + 1}};
        float sum = 0.0f;
#pragma unroll
        for (int n = 0; n < 6; n++) {
            if (!(mask & (1u << n))) continue;
            float dn;
            uint32_t mn;
            stencil_of(f, nb[n][0], nb[n][1], nb[n][2], dn, mn);
            sum += OMEGA * rin[f.idx(nb[n][0], nb[n][1], nb[n][2])] * rcp_or0(scale * dn);
        }
        z1[id] = zc;
        acc += r - scale * (diag * zc - sum);
    }
    S.mg[l + 1].r[c] = acc;
}

// ---- coarsest level: COARSE_SWEEPS damped-Jacobi sweeps from zero in one workgroup (LDS).
__global__ __launch_bounds__(CBLOCK) void k_mg_coarse(DevSim S, int l) {
    __shared__ float za[MG_COARSE_MAX], zb[MG_COARSE_MAX];
    const LevelView v = view(S, l);
    const int n = v.nx * v.ny * v.nz;
    const float scale = S.mg[l].scale;
    for (int c = threadIdx.x; c < n; c += CBLOCK) za[c] = 0.0f;
    __syncthreads();
    float *zin = za, *zout = zb;
    for (int it = 0; it < COARSE_SWEEPS; it++) {
        for (int c = threadIdx.x; c < n; c += CBLOCK) {
            const int i = c % v.nx, j = (c / v.nx) % v.ny, k = c / (v.nx * v.ny);
            float out = 0.0f;
            if (v.at(i, j, k) == CELL_FLUID) {
                float diag;
                uint32_t mask;
                stencil_of(v, i, j, k, diag, mask);
                const int nb[6] = {c - 1, c + 1, c - v.nx, c + v.nx, c - v.nx * v.ny, c + v.nx * v.ny};
                float sum = 0.0f;
#pragma unroll
                for (int q = 0; q < 6; q++)
                    if (mask & (1u << q)) sum += zin[nb[q]];
                const float az = scale * (diag * zin[c] - sum);
                out = zin[c] + OMEGA * (S.mg[l].r[c] - az) * rcp_or0(scale * diag);
            }
            zout[c] = out;
        }
        __syncthreads();
        float *tmp = zin;
        zin = zout;
        zout = tmp;
    }
    for (int c = threadIdx.x; c < n; c += CBLOCK) S.mg[l].zf[c] = zin[c];
}

// Corrected value z1 + prolonged coarse correction of level-l cell (i, j, k) (FLUID only).
__device__ __forceinline__ float corrected(const float *__restrict__ z1, const float *__restrict__ zc, int cnx,
                                           int cny, const LevelView &f, int i, int j, int k) {
    return z1[f.idx(i, j, k)] + zc[(uint32_t)(i >> 1) + (uint32_t)cnx * ((uint32_t)(j >> 1) + (uint32_t)cny * (uint32_t)(k >> 1))];
}

// ---- up on a coarse level l >= 1: zf_l = z + omega D^-1 (r - A z), z = z1 + P zf_{l+1}.
__global__ __launch_bounds__(MBLOCK) void k_mg_up(DevSim S, int l) {
    const LevelView f = view(S, l);
    const uint32_t c = blockIdx.x * MBLOCK + threadIdx.x;
    if (c >= (uint32_t)(f.nx * f.ny * f.nz)) return;
    const int i = (int)(c % (uint32_t)f.nx), j = (int)((c / (uint32_t)f.nx) % (uint32_t)f.ny), k = (int)(c / ((uint32_t)f.nx * (uint32_t)f.ny));
    float out = 0.0f;
    if (f.at(i, j, k) == CELL_FLUID) {
        const float scale = S.mg[l].scale;
        const float *z1 = S.mg[l].z1, *zc = S.mg[l + 1].zf;
        const int cnx = S.mg[l + 1].nx, cny = S.mg[l + 1].ny;
        float diag;
        uint32_t mask;
        stencil_of(f, i, j, k, diag, mask);
        const int nb[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k}, {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
        float sum = 0.0f;
#pragma unroll
        for (int n = 0; n < 6; n++)
            if (mask & (1u << n)) sum += corrected(z1, zc, cnx, cny, f, nb[n][0], nb[n][1], nb[n][2]);
        const float z = corrected(z1, zc, cnx, cny, f, i, j, k);
        out = z + OMEGA * (S.mg[l].r[c] - scale * (diag * z - sum)) * rcp_or0(scale * diag);
    }
    S.mg[l].zf[c] = out;
}

// ---- up on level 0 over the fluid list: u = z + omega D^-1 (r - A z), z = z1 + P zf_1.
__global__ __launch_bounds__(MBLOCK) void k_mg_up0(DevSim S, const float *__restrict__ r, float *__restrict__ u) {
    const uint32_t n = S.fl_count[1];
    const LevelView f = view(S, 0);
    const float *z1 = S.mg_z1, *zc = S.mg[1].zf;
    const int cnx = S.mg[1].nx, cny = S.mg[1].ny;
    const int32_t sy = S.nx, sz = S.nx * S.ny;
    const int32_t off[6] = {-1, 1, -sy, sy, -sz, sz};
    for (uint32_t t = blockIdx.x * MBLOCK + threadIdx.x; t < n; t += gridDim.x * MBLOCK) {
        const uint32_t id = S.fl_list[t];
        const uint32_t wd = S.fl_word[t];
        const int i = (int)(id % (uint32_t)S.nx), j = (int)((id / (uint32_t)S.nx) % (uint32_t)S.ny),
                  k = (int)(id / ((uint32_t)S.nx * (uint32_t)S.ny));
        const int nb[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k}, {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
        float sum = 0.0f;
#pragma unroll
        for (int q = 0; q < 6; q++)
            if ((wd >> q) & 1u) sum += corrected(z1, zc, cnx, cny, f, nb[q][0], nb[q][1], nb[q][2]);
        (void)off;
        const float diag = (float)((wd >> 6) & 7u);
        const float z = corrected(z1, zc, cnx, cny, f, i, j, k);
        u[id] = z + OMEGA * (r[id] - (diag * z - sum)) * rcp_or0(diag);
    }
}

inline unsigned grid_for(uint32_t n) { return (n + MBLOCK - 1) / MBLOCK; }
inline uint32_t cells(const MGLevel &m) { return (uint32_t)(m.nx * m.ny * m.nz); }

}  // namespace

void launch_mg_types(const DevSim &S, hipStream_t s) {
    for (int l = 1; l < S.mg_levels; l++) k_mg_types<<<grid_for(cells(S.mg[l])), MBLOCK, 0, s>>>(S, l);
}

void launch_mg_vcycle(const DevSim &S, const float *r, float *u, hipStream_t s) {
    const int L = S.mg_levels - 1;  // coarsest level index
    k_mg_down<<<grid_for(cells(S.mg[1])), MBLOCK, 0, s>>>(S, 0, r, S.mg_z1);
    for (int l = 1; l < L; l++) k_mg_down<<<grid_for(cells(S.mg[l + 1])), MBLOCK, 0, s>>>(S, l, S.mg[l].r, S.mg[l].z1);
    k_mg_coarse<<<1, CBLOCK, 0, s>>>(S, L);
    for (int l = L - 1; l >= 1; l--) k_mg_up<<<grid_for(cells(S.mg[l])), MBLOCK, 0, s>>>(S, l);
    const uint32_t g = grid_for(S.ncells) < 4096u ? grid_for(S.ncells) : 4096u;
    k_mg_up0<<<g, MBLOCK, 0, s>>>(S, r, u);
}

}  // namespace pxl
