// Device data layout of one simulation and the kernel launch interface.
//
// HBM layout (all arrays separately aligned inside one arena per simulation):
//   particles   SoA fp32 x,y,z,vx,vy,vz + u32 id, two copies (current / sorted scratch)
//   keys        u32 cell and u32 slot per particle (counting-sort binning)
//   faces       MAC faces, u block then v then w (reference order): fp32 vel, vel_old,
//               int32 acc, wacc (fixed-point P2G sums, scale 65536)
//   cells       x-fastest: u8 type, u16 stencil, fp32 pressure, div; u32 count and offset
//               in tiled key order (see TILE below)
//   solver      fp32 r, w, s (two copies each), p, e (CG); fp64 x (two copies), r
//               (refinement); workgroup partials, scalars
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pxl {

enum : uint8_t { CELL_AIR = 0, CELL_FLUID = 1, CELL_SOLID = 2 };

// Stencil word of a cell: bit n (n = 0..5, order -x,+x,-y,+y,-z,+z) set when that
// neighbour is FLUID; bits 8..10 = number of non-SOLID neighbours (the matrix diagonal);
// bit 15 set when the cell itself is FLUID.
constexpr uint16_t STEN_FLUID = 0x8000;

// Particles are counting-sorted by a TILED cell key: 8x8x8-cell tiles in x-fastest tile
// order, cells x-fastest inside a tile.  A tile's particles are then one contiguous run,
// which the P2G kernel accumulates in LDS.  count/offset are indexed by this key (the key
// space is ntx*nty*ntz*512 >= ncells; cells past the grid edge stay empty).
constexpr int TILE = 8;
constexpr uint32_t TILE_CELLS = 512;

// Solver scalar slots (fp64 array).
enum ScalarSlot { SC_RZ = 0, SC_ALPHA = 1, SC_BETA = 2, SC_RR = 3, SC_BB = 4, SC_COUNT = 8 };

struct DevSim {
    // grid
    int nx, ny, nz;
    uint32_t np;
    float h, dt, flip, inv_h;
    float ox, oy, oz;
    float gx, gy, gz;
    uint32_t nu, nv, nw;  // faces per component; bases are 0, nu, nu + nv
    uint32_t ncells, nfaces;
    int ntx, nty, ntz;  // tiles per axis
    uint32_t nkeys;     // ntx * nty * ntz * TILE_CELLS
    // particles: current state (a) and sorted scratch (b)
    float *ax, *ay, *az, *avx, *avy, *avz;
    uint32_t *aid;
    float *bx, *by, *bz, *bvx, *bvy, *bvz;
    uint32_t *bid;
    uint32_t *key, *slot;
    // faces
    float *vel, *vel_old;
    int32_t *acc, *wacc;
    // cells (count/offset: tiled key order, offset[nkeys] = np)
    uint32_t *count, *offset;
    uint8_t *type;
    uint16_t *sten;
    float *pressure, *div;
    // solver
    float *cg_r0, *cg_r1, *cg_w0, *cg_w1, *cg_s0, *cg_s1, *cg_p, *e;
    double *x64a, *x64b, *r64;
    float *part32;
    double *part64;
    double *scal;
    unsigned *ctr;
    uint32_t *scan_tmp;
    // initial particles (reset)
    float *ix, *iy, *iz, *ivx, *ivy, *ivz;
};

__device__ __forceinline__ uint32_t tile_key(const DevSim &S, int i, int j, int k) {
    const uint32_t t = (uint32_t)(i >> 3) + (uint32_t)S.ntx * ((uint32_t)(j >> 3) + (uint32_t)S.nty * (uint32_t)(k >> 3));
    return t * TILE_CELLS + (uint32_t)((i & 7) | ((j & 7) << 3) | ((k & 7) << 6));
}

// ---- launch helpers (all asynchronous on `s`) ----
void launch_fill_u32(uint32_t *p, uint32_t n, uint32_t value, hipStream_t s);
void launch_clear_grid(const DevSim &S, hipStream_t s);       // count, acc, wacc = 0
void launch_bin_count(const DevSim &S, hipStream_t s);        // key, slot, count
void launch_scan_offsets(const DevSim &S, hipStream_t s);     // offset = exclusive scan(count)
uint32_t scan_tiles(uint32_t n);                              // scratch entries a scan of n needs
void launch_exclusive_scan(const uint32_t *in, uint32_t n, uint32_t *tile_tmp, uint32_t *out, hipStream_t s);
void launch_scatter(const DevSim &S, hipStream_t s);          // a -> b in cell order
void launch_p2g(const DevSim &S, hipStream_t s);              // b -> acc, wacc
void launch_faces(const DevSim &S, hipStream_t s);            // normalize + forces
void launch_cells(const DevSim &S, hipStream_t s);            // type, stencil, p=0, div
void launch_rbgs(const DevSim &S, uint32_t parity, hipStream_t s);
void launch_project(const DevSim &S, hipStream_t s);
void launch_g2p(const DevSim &S, hipStream_t s);              // b -> a, advected
// Jacobi-preconditioned CG (solver.hip), one fused pass + one finalize per iteration:
//   launch_cg_start(x, inner), launch_cg_finalize(first = true),
//   then for it = 0..iters-1: launch_cg_iter(x, it), launch_cg_finalize(first = false).
// Plain solve: fp64 = false, x = pressure, inner = false.  Inner solve of the refinement:
// fp64 = true (FP64 dot products), x = e, inner = true (b = float(r64), e starts at 0).
dim3 solver_grid(const DevSim &S);
uint32_t solver_blocks(const DevSim &S);  // workgroups of a solver launch (partials per value)
void launch_cg_start(const DevSim &S, bool fp64, float *x, bool inner, hipStream_t s);
void launch_cg_finalize(const DevSim &S, bool fp64, bool first, hipStream_t s);
void launch_cg_iter(const DevSim &S, bool fp64, float *x, uint32_t it, hipStream_t s);
// FP64 refinement (solver.hip); init/update also total the norms into scal[SC_RR / SC_BB].
void launch_mixed_init(const DevSim &S, double *x64, hipStream_t s);  // x64 = p, r64 = b - A x64
void launch_mixed_update(const DevSim &S, const double *xsrc, double *xdst, hipStream_t s);  // xdst = xsrc + e; r64
void launch_mixed_finish(const DevSim &S, const double *x64, hipStream_t s);  // p = float(x64)
// bins / IO
void launch_bins_out(const DevSim &S, const uint32_t *off_lin, uint32_t *cell_of, uint32_t *sorted, hipStream_t s);
void launch_sort_segments(uint32_t ncells, const uint32_t *offsets, const uint32_t *counts, uint32_t *sorted,
                          hipStream_t s);
void launch_count_linear(const DevSim &S, uint32_t *out, hipStream_t s);  // counts, reference cell order
void launch_unpack_aos8(const DevSim &S, float *aos8, hipStream_t s);        // a -> AoS8 by id
void launch_pack_aos8(const DevSim &S, const float *aos8, hipStream_t s);    // AoS8 -> a, id = i
void launch_reset_particles(const DevSim &S, hipStream_t s);                 // initial -> a
void launch_type_u32(const DevSim &S, uint32_t *out, hipStream_t s);

}  // namespace pxl
