// Device data layout of one simulation and the kernel launch interface.
//
// HBM layout (all arrays separately aligned inside one arena per simulation):
//   particles   SoA fp32 x,y,z,vx,vy,vz + u32 id, two copies (current / sorted scratch)
//   keys        u32 cell and u32 slot per particle (counting-sort binning)
//   faces       MAC faces, u block then v then w (reference order): fp32 vel, vel_old,
//               int32 acc, wacc (fixed-point P2G sums, scale 65536)
//   cells       x-fastest: u32 count, u32 offset, u8 type, u16 stencil, fp32 pressure, div
//   solver      fp32 r, d0, d1, q, e (CG); fp64 x (two copies), r (refinement); partials, scalars
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pxl {

enum : uint8_t { CELL_AIR = 0, CELL_FLUID = 1, CELL_SOLID = 2 };

// Stencil word of a cell: bit n (n = 0..5, order -x,+x,-y,+y,-z,+z) set when that
// neighbour is FLUID; bits 8..10 = number of non-SOLID neighbours (the matrix diagonal);
// bit 15 set when the cell itself is FLUID.
constexpr uint16_t STEN_FLUID = 0x8000;

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
    // particles: current state (a) and sorted scratch (b)
    float *ax, *ay, *az, *avx, *avy, *avz;
    uint32_t *aid;
    float *bx, *by, *bz, *bvx, *bvy, *bvz;
    uint32_t *bid;
    uint32_t *key, *slot;
    // faces
    float *vel, *vel_old;
    int32_t *acc, *wacc;
    // cells
    uint32_t *count, *offset;
    uint8_t *type;
    uint16_t *sten;
    float *pressure, *div;
    // solver
    float *r, *d0, *d1, *q, *e;
    double *x64a, *x64b, *r64;
    float *part32;
    double *part64;
    double *scal;
    unsigned *ctr;
    uint32_t *scan_tmp;
    uint32_t nblocks_cells;  // blocks of a cell launch (reduction partial count)
    // initial particles (reset)
    float *ix, *iy, *iz, *ivx, *ivy, *ivz;
};

// ---- launch helpers (all asynchronous on `s`) ----
void launch_fill_u32(uint32_t *p, uint32_t n, uint32_t value, hipStream_t s);
void launch_clear_grid(const DevSim &S, hipStream_t s);       // count, acc, wacc = 0
void launch_bin_count(const DevSim &S, hipStream_t s);        // key, slot, count
void launch_scan_offsets(const DevSim &S, hipStream_t s);     // offset = exclusive scan(count)
void launch_scatter(const DevSim &S, hipStream_t s);          // a -> b in cell order
void launch_p2g(const DevSim &S, hipStream_t s);              // b -> acc, wacc
void launch_faces(const DevSim &S, hipStream_t s);            // normalize + forces
void launch_cells(const DevSim &S, hipStream_t s);            // type, stencil, p=0, div
void launch_rbgs(const DevSim &S, uint32_t parity, hipStream_t s);
void launch_project(const DevSim &S, hipStream_t s);
void launch_g2p(const DevSim &S, hipStream_t s);              // b -> a, advected
// Jacobi-PCG (solver.hip).  Plain FP32 solve of the pressure system:
//   launch_cg_init32, then for it = 0..iters-1: pass_a(it), pass_b(pressure, it).
// Inner solve of the mixed refinement (A e = float(r64), FP64 dot products):
//   launch_inner_init, then pass_a(fp64, it), pass_b(fp64, e, it).
void launch_cg_init32(const DevSim &S, hipStream_t s);
void launch_inner_init(const DevSim &S, hipStream_t s);
void launch_cg_pass_a(const DevSim &S, bool fp64, uint32_t it, hipStream_t s);  // d, q = A d, alpha
void launch_cg_pass_b(const DevSim &S, bool fp64, float *x, uint32_t it, hipStream_t s);  // x, r, beta
// FP64 refinement (solver.hip).
void launch_mixed_init(const DevSim &S, double *x64, hipStream_t s);  // x64 = p, r64 = b - A x64
void launch_mixed_update(const DevSim &S, const double *xsrc, double *xdst, hipStream_t s);  // xdst = xsrc + e; r64
void launch_mixed_finish(const DevSim &S, const double *x64, hipStream_t s);  // p = float(x64)
// bins / IO
void launch_bins_out(const DevSim &S, uint32_t *cell_of, uint32_t *sorted, hipStream_t s);
void launch_sort_segments(const DevSim &S, uint32_t *sorted, hipStream_t s);
void launch_unpack_aos8(const DevSim &S, float *aos8, hipStream_t s);        // a -> AoS8 by id
void launch_pack_aos8(const DevSim &S, const float *aos8, hipStream_t s);    // AoS8 -> a, id = i
void launch_reset_particles(const DevSim &S, hipStream_t s);                 // initial -> a
void launch_type_u32(const DevSim &S, uint32_t *out, hipStream_t s);

}  // namespace pxl
