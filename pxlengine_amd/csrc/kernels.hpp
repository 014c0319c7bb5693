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

#include <hip/hip_ext.h>
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

// Multigrid levels of the mixed-precision solve (level 0 = the simulation grid).  Coarsening
// halves every axis until no axis exceeds 2 cells; the levels from mg_tail on (at most
// MG_TAIL_CELLS cells together) are solved inside one workgroup's LDS.
constexpr int MG_MAX = 12;
constexpr int MG_TAIL_CELLS = 6144;
// Smoothing rounds of the list levels 1..tail-1 (level 0 always runs 2): pre-smoothing runs
// (red, black) that many times, post-smoothing (black, red) as many (the adjoint, so the
// preconditioner stays symmetric).
constexpr int MG_COARSE_ROUNDS = 2;
// List levels of at most MG_MID_CELLS cells just above the tail (the "mid" levels) run inside
// the tail's single workgroup, from global memory, with workgroup barriers between their
// sweeps instead of one grid-wide launch per sweep.
constexpr int MG_MID_CELLS = 32768;

struct MGLevel {
    int nx, ny, nz;
    // Levels >= 1 store sten, r and z colour-split like the MIXED inner vectors: cell (i, j, k)
    // at sigma = ((i + j + k) & 1) hsplit + (i >> 1) + hx (j + ny k), hx = ceil(nx / 2),
    // hsplit = hx ny nz (2 hsplit entries), so a colour sweep reads and writes dense runs.
    int hx;
    uint32_t hsplit;
    uint8_t *type;   // level-l cell types in cell order (levels >= 1)
    uint16_t *sten;  // level-l stencil words (the bits of DevSim::sten plus STEN_BLACK; levels >= 1)
    float *r, *z;    // level-l right-hand side (the scaled restricted residual) and iterate
    // FLUID cells of the level (levels >= 1, rebuilt each step), each list in cell order: red
    // cells from the front (count[0] of them), black cells from the back (count[1]: the t-th
    // at list[cells - 1 - t]); nlist holds all of them (count[2]) for the kernels that treat
    // both colours alike (pre0 and the restriction into the level: walking the colours one
    // after the other there reads the finer level's lines twice, measured 1.9x the bytes).
    // Cell order keeps the neighbour rows and planes that a list slice reads next to the ones
    // its neighbouring slices read.  An entry is the cell's sigma with i & 1 in bit 31 (the x
    // neighbours' offsets depend on it).
    uint32_t *list, *nlist, *count;
    // per 64-cell chunk of the level: its red FLUID cells [0, chunks), then its black ones
    // [chunks, 2 chunks); the level scan turns both halves into exclusive offsets
    uint32_t *chunk;
};
__host__ __device__ inline uint32_t level_chunks(const MGLevel &v) { return (uint32_t)(v.nx * v.ny * v.nz + 63) / 64u; }

// Solver scalar slots (fp64 array): two slots of SC_COUNT by CG-iteration parity for the
// (alpha, beta, gamma) history; the refinement norms live in slot 0.
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
    int cg_blocks;      // workgroups of the fluid-list solver launches (tunable)
    int split_colours;  // fluid list red cells first then black (RBGS, MIXED) or in cell order (CG)
    // particles: current state (a) and sorted scratch (b)
    float *ax, *ay, *az, *avx, *avy, *avz;
    uint32_t *aid;
    float *bx, *by, *bz, *bvx, *bvy, *bvz;
    uint32_t *bid;
    uint32_t *key, *slot;
    // faces
    float *vel, *vel_old;
    // per-tile P2G sums: for each tile, component c and face of the tile's face box (see
    // tile_box), the fixed-point sums packed as (sum_acc << 32) + sum_wacc in one int64
    // (sum_wacc >= 0 and < 2^32, so the halves never carry into each other)
    int64_t *box;
    uint32_t ntiles;
    // workgroups of the per-tile (P2G, G2P) and band (faces, cells, projection) launches, each
    // walking its list of tiles with this stride: sized by the liquid's volume in tiles, not
    // the grid's, so a small liquid in a large grid does not launch idle workgroups
    uint32_t tile_wgs;
    uint32_t *tile_list, *tile_count;  // tiles holding particles this step (any order)
    // per tile: bit 0 = holds particles this step, bit 1 = held particles the step before
    // (3 after allocation, reset or a field write: the next two steps then cover every tile)
    uint8_t *tile_act;
    // the band: tiles within one tile of a tile active this step or the step before.  Faces
    // and cells outside it hold values that do not change from step to step (no particle
    // weight, no FLUID cell near), so the face, cell and projection passes walk only the band.
    uint32_t *band_list, *band_count;
    // in a batch: the batch-wide band list ((member << WORK_SIM_SHIFT) | tile, every member's
    // band tiles) and its count, and this simulation's member index; null alone
    uint32_t *band_all, *band_all_count;
    uint32_t member;
    // cells (count/offset: tiled key order, offset[nkeys] = np)
    uint32_t *count, *offset;
    uint8_t *type;
    uint16_t *sten;
    float *pressure, *div;
    // obstacles (an extension; null without): bit (cell & 31) of solid_bits[cell >> 5] is set
    // for obstacle cells (SOLID like the wall layer), obst_id[cell] = 1 + the index of the
    // first obstacle in list order holding the cell's centre (0 = none), obst = n_obst rows
    // {cx, cy, cz, r}
    uint32_t *solid_bits, *obst_id;
    float *obst;
    uint32_t n_obst;
    // FLUID-cell list of the step, built from per-row counts: fl_rows holds the FLUID cells
    // of each grid row (j, k) (2 ny nz words: rows' red cells, then rows' black cells; with
    // colours not split, all of a row's cells in the first half and zeros in the second),
    // fl_rowoff their exclusive scan (each row's first list slot), fl_list the list (cell
    // order, or red cells first then black, each in cell order) and fl_count
    // {n_red (0 unless split), n_total}
    uint32_t *fl_rows, *fl_rowoff, *fl_list, *fl_count, *fl_tmp;
    // per list entry: bits 0-5 FLUID neighbours (-x,+x,-y,+y,-z,+z), bits 6-8 the cell's
    // non-SOLID neighbour count (matrix diagonal), bits 9+3n the diagonal of FLUID neighbour n,
    // bit 27 the cell's colour (1 = black: i + j + k odd), bit 28 its x parity (i & 1)
    uint32_t *fl_word;
    // reference mode (RBGS, several half sweeps per launch): per core of the grid (rbgs_cores),
    // both colours' lists of the FLUID cells of its region (rbgs_list_slots entries per core:
    // region index, x parity and diagonal in rb_list, the b term in rb_b) and the lists'
    // lengths by distance from the region's faces (8 words per core); null in other modes
    uint32_t *rb_list, *rb_count;
    float *rb_b;
    // the same entries in plain cell order (with colours not split: fl_list / fl_word
    // themselves), for the kernels that treat every FLUID cell alike: their vector accesses
    // then run along x instead of skipping every other cell
    uint32_t *fn_list, *fn_word;
    // MIXED: the FP32 vectors of the inner solve (cg_r0 = r, cg_r1 = u, cg_w0, cg_s0, cg_p, e)
    // are stored colour-split: cell (i, j, k) of colour p = (i + j + k) & 1 at
    // p hsplit + (i >> 1) + hx (j + ny k), hx = ceil(nx / 2), hsplit = hx ny nz, so a colour
    // sweep reads and writes whole lines of one colour.  fls_list / fns_list hold the split index
    // of each fl_list / fn_list entry (fl_word bit 28 = i & 1 gives the x neighbours' offsets).
    uint32_t hx, hsplit;
    uint32_t *fls_list, *fns_list;
    // solver
    float *cg_r0, *cg_r1, *cg_w0, *cg_w1, *cg_s0, *cg_s1, *cg_p, *e;
    // MIXED: the MG-CG iterate in colour-split list order (index t of fl_list, like p, s and w
    // there; storage of cg_s1, which MIXED does not use otherwise), and a list-order copy of
    // r for the kernels that read r only at the entry itself (storage of cg_w1)
    float *cg_x, *cg_rt;
    double *x64a, *x64b, *r64;
    MGLevel mg[MG_MAX];
    int mg_levels;  // levels including level 0 (0 when the multigrid is not allocated)
    int mg_tail;    // first level of the single-workgroup tail (>= 1)
    int mg_mid;     // first level run inside the tail's workgroup (1 <= mg_mid <= mg_tail)
    float *part32;   // two slots of 2 x (workgroups at z-batch 1)
    double *part64;
    double *scal;
    double *norms;  // where the refinement norms {rr, bb} are totalled (scal + SC_RR, or a
                    // batch's per-sim array)
    unsigned *ctr;  // [0]: the CFL cap's largest squared particle speed (launch_vmax2)
    uint32_t *scan_tmp;
    // initial particles (reset)
    float *ix, *iy, *iz, *ivx, *ivy, *ivz;
};

// Colour-split index of cell (i, j, k) (DevSim::hx, hsplit).
__host__ __device__ inline uint32_t split_index(const DevSim &S, uint32_t i, uint32_t j, uint32_t k) {
    return (((i + j + k) & 1u) ? S.hsplit : 0u) + (i >> 1) + S.hx * (j + (uint32_t)S.ny * k);
}
// Split-index offsets of the six neighbours (-x, +x, -y, +y, -z, +z) of a cell whose list word
// has colour bit 27 and x-parity bit 28: the other colour's half, the same half-row position
// along y and z, and along x the position (i -/+ 1) >> 1.
__device__ __forceinline__ void split_offsets(const DevSim &S, uint32_t word, int32_t off[6]) {
    const int32_t d = ((word >> 27) & 1u) ? -(int32_t)S.hsplit : (int32_t)S.hsplit;
    const int32_t o = (int32_t)((word >> 28) & 1u), hy = (int32_t)S.hx, hz = (int32_t)S.hx * S.ny;
    off[0] = d - 1 + o;
    off[1] = d + o;
    off[2] = d - hy;
    off[3] = d + hy;
    off[4] = d - hz;
    off[5] = d + hz;
}

// Face box of a tile for component c: every face a particle binned in the tile can touch.
// Along c the faces [8t, 8t + 8] (9), along the other axes the faces at cells [8t - 1, 8t + 8]
// (10): 900 faces per component, 2700 per tile.
constexpr int BOX_FACES = 900;
constexpr int BOX_ENTRIES = 3 * BOX_FACES;

__device__ __forceinline__ void tile_box(int c, int ti, int tj, int tk, int lo[3], int d[3]) {
    lo[0] = ti * TILE - 1;
    lo[1] = tj * TILE - 1;
    lo[2] = tk * TILE - 1;
    d[0] = d[1] = d[2] = TILE + 2;
    lo[c] += 1;
    d[c] = TILE + 1;
}

// Where face (lx, ly, lz) of component c's box (box coordinates, tile_box dims) is stored in
// HBM (S.box).  The box's own columns along x (lx 0..7 for c = 0, 1..8 otherwise) come first,
// [lz][ly][8 x]; then its x-halo columns (lx 8 for c = 0; lx 0 and 9 otherwise), each
// [lz][ly] contiguous.  The face pass of the x-neighbour tile reads exactly one halo column of
// this box, so it reads it as one contiguous run instead of one 8-byte entry per 128-byte
// line.  (LDS keeps the plain [lz][ly][lx] order; P2G permutes when it stores.)
__host__ __device__ inline int box_gidx(int c, int lx, int ly, int lz) {
    const int d1 = c == 1 ? TILE + 1 : TILE + 2, d2 = c == 2 ? TILE + 1 : TILE + 2;
    const int i0 = c == 0 ? 0 : 1;
    const int yz = ly + d1 * lz;
    if (lx >= i0 && lx < i0 + TILE) return c * BOX_FACES + (lx - i0) + TILE * yz;
    const int h = (c == 0 || lx == 0) ? 0 : 1;
    return c * BOX_FACES + TILE * d1 * d2 + h * d1 * d2 + yz;
}

__device__ __forceinline__ bool tile_active(const DevSim &S, uint32_t t) { return (S.tile_act[t] & 1u) != 0; }

// Packed P2G sum of face q of component c: the sum over the (1 to 8) boxes of active tiles
// that contain it.  Along c a face 8m lies in the boxes of tiles m-1 and m; along another
// axis the face at cell 8m lies in tiles m-1 and m, at cell 8m+7 in tiles m and m+1.
__device__ __forceinline__ int64_t gather_face(const DevSim &S, int c, int qx, int qy, int qz) {
    const int q[3] = {qx, qy, qz};
    const int nt[3] = {S.ntx, S.nty, S.ntz};
    int tlo[3], thi[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const int m = q[a] >> 3, r = q[a] & 7;
        tlo[a] = m - (r == 0 ? 1 : 0);
        thi[a] = m + ((a != c && r == 7) ? 1 : 0);
        tlo[a] = tlo[a] < 0 ? 0 : tlo[a];
        thi[a] = thi[a] > nt[a] - 1 ? nt[a] - 1 : thi[a];
    }
    int64_t sum = 0;
    for (int tz = tlo[2]; tz <= thi[2]; tz++)
        for (int ty = tlo[1]; ty <= thi[1]; ty++)
            for (int tx = tlo[0]; tx <= thi[0]; tx++) {
                const uint32_t t = (uint32_t)tx + (uint32_t)S.ntx * ((uint32_t)ty + (uint32_t)S.nty * (uint32_t)tz);
                if (!tile_active(S, t)) continue;
                int lo[3], d[3];
                tile_box(c, tx, ty, tz, lo, d);
                sum += S.box[(size_t)t * BOX_ENTRIES + (size_t)box_gidx(c, qx - lo[0], qy - lo[1], qz - lo[2])];
            }
    return sum;
}

__device__ __forceinline__ int32_t packed_acc(int64_t v) { return (int32_t)(v >> 32); }
__device__ __forceinline__ int32_t packed_wacc(int64_t v) { return (int32_t)(uint32_t)(uint64_t)v; }

__device__ __forceinline__ uint32_t tile_key(const DevSim &S, int i, int j, int k) {
    const uint32_t t = (uint32_t)(i >> 3) + (uint32_t)S.ntx * ((uint32_t)(j >> 3) + (uint32_t)S.nty * (uint32_t)(k >> 3));
    return t * TILE_CELLS + (uint32_t)((i & 7) | ((j & 7) << 3) | ((k & 7) << 6));
}

// Surface nets of one simulation (fluid.slang:222-355): the surface grid has sx x sy x sz
// cells ((sx+1)(sy+1)(sz+1) nodes, x fastest) of spacing sh from the fluid's origin.
struct DevSurf {
    int sx, sy, sz;
    uint32_t nnodes, ncells;
    uint32_t max_verts, max_tris;
    float sh, radius, iso;
    float ox, oy, oz;
    int32_t *density;     // nodes: splatted density, fixed point (4096)
    float *phi;           // nodes: smoothed density - iso (> 0 inside), -1 on the boundary
    uint32_t *cell_vert;  // cells: vertex number, or SURF_NONE
    uint32_t *qoff;       // nodes: quads at the node's three edges, then their exclusive scan
    uint32_t *counters;   // [0] surface cells, [1] 2 x quads (the reference's counters, unclipped)
    uint32_t *scan_tmp;
};
constexpr uint32_t SURF_NONE = 0xFFFFFFFFu;
constexpr float SURF_SD = 4096.0f;  // fixed-point scale of the surface density
// The surface pass on the current positions, in stages: clear, splat, phi, vertices (numbered
// in surface-cell order by a scan of the surface flags) and triangles (in (node, axis) order,
// by a scan of the per-node quad counts) into verts[8 (vbase + v)] and indices[ibase + 3 t]
// (device memory).
enum SurfStage { SURF_CLEAR, SURF_SPLAT, SURF_PHI, SURF_VFLAG, SURF_VSCAN, SURF_VERTS, SURF_QCOUNT, SURF_QSCAN,
                 SURF_QUADS, SURF_STAGES };
extern const char *const SURF_STAGE_NAMES[SURF_STAGES];
void launch_surface_stage(int stage, const DevSurf &S, const float *px, const float *py, const float *pz, uint32_t np,
                          float *verts, uint32_t vbase, uint32_t *indices, uint32_t ibase, hipStream_t s);

// ---- launches ----
// Every launch goes through PXL_LAUNCH.  While a profiler has set this thread's launch events,
// the first launch records them from its own dispatch (the kernel's start and end, as a
// kernel trace sees them) and counts itself; later launches of the same profiled phase, and
// all launches otherwise, are plain.
struct LaunchEvents {
    hipEvent_t start = nullptr, stop = nullptr;
    int launches = 0;
};
LaunchEvents &launch_events();  // this thread's slot
#define PXL_LAUNCH(kernel, grid, block, shmem, stream, ...)                                                    \
    do {                                                                                                       \
        ::pxl::LaunchEvents &ev_ = ::pxl::launch_events();                                                     \
        if (ev_.start && ev_.launches++ == 0)                                                                  \
            hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(block), (std::uint32_t)(shmem), stream, ev_.start,  \
                                  ev_.stop, 0u, __VA_ARGS__);                                                  \
        else                                                                                                   \
            hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), shmem, stream, __VA_ARGS__);                  \
    } while (0)

// ---- launch helpers (all asynchronous on `s`) ----
void launch_fill_u32(uint32_t *p, uint32_t n, uint32_t value, hipStream_t s);
// *out = the bits of max over the particles of (vx vx + vy vy) + vz vz (f32; non-negative
// floats order as their bit patterns); clears *out first.
void launch_vmax2(const DevSim &S, uint32_t *out, hipStream_t s);
void launch_clear_grid(const DevSim &S, hipStream_t s);       // count = 0 (the last step's tiles), row counts = 0
void launch_bin_count(const DevSim &S, hipStream_t s);        // key, slot, count
void launch_scan_offsets(const DevSim &S, hipStream_t s);     // offset = exclusive scan(count)
uint32_t scan_tiles(uint32_t n);                              // scratch entries a scan of n needs
// Words of the fluid list's per-row counts (two per grid row (j, k)).
__host__ __device__ inline uint32_t fl_row_slots(const DevSim &S) { return 2u * (uint32_t)S.ny * (uint32_t)S.nz; }
void launch_exclusive_scan(const uint32_t *in, uint32_t n, uint32_t *tile_tmp, uint32_t *out, hipStream_t s);
void launch_tile_list(const DevSim &S, hipStream_t s);        // active tiles (after the scan)
void launch_band(const DevSim &S, hipStream_t s);             // band tiles (after the tile list)
void launch_scatter(const DevSim &S, hipStream_t s);          // a -> b in cell order
void launch_p2g(const DevSim &S, hipStream_t s);              // b -> per-tile boxes
void launch_faces(const DevSim &S, hipStream_t s);            // normalize + forces
void launch_cells(const DevSim &S, hipStream_t s);            // type, stencil, p=0, div, flags
void launch_fluid_list(const DevSim &S, hipStream_t s);       // row counts -> fl_list, fl_count
// `iters` red-black iterations, up to two per launch after one launch that lists each core's
// FLUID cells, ping-ponging between the pressure array and cg_r0 (an even number of launches:
// the result ends in the pressure array)
void launch_rbgs_blocked(const DevSim &S, uint32_t iters, hipStream_t s);
uint32_t rbgs_cores(const DevSim &S);  // cores of the grid (rb_count holds 8 words per core)
uint32_t rbgs_list_slots();            // list entries per core (rb_list, rb_b)
void launch_project(const DevSim &S, hipStream_t s);
void launch_g2p(const DevSim &S, hipStream_t s);              // b -> a, advected
// Jacobi-preconditioned CG (solver.hip, over the fluid list), one fused launch per iteration:
//   launch_cg_start(x, inner), then for it = 0..iters-1: launch_cg_iter(x, it).
// Plain solve: fp64 = false, x = pressure, inner = false.  Inner solve of the refinement:
// fp64 = true (FP64 dot products), x = e, inner = true (b = float(r64), e starts at 0).
uint32_t solver_blocks(const DevSim &S);  // workgroups of a solver launch (partials per value)
void launch_cg_start(const DevSim &S, bool fp64, float *x, bool inner, hipStream_t s);
void launch_cg_iter(const DevSim &S, bool fp64, float *x, uint32_t it, hipStream_t s);
// Multigrid-preconditioned CG (mg.hip + solver.hip), the inner solve of the refinement:
//   launch_mg_types once per step; launch_mgcg_start(e); then for it = 0..K-1:
//   launch_mgcg_iter(e, it, it == K-1).  FP64 dot products.
void launch_mg_types(const DevSim &S, hipStream_t s);
void launch_mg_vcycle(const DevSim &S, const float *r, float *u, hipStream_t s);  // u = M r
void launch_mgcg_start(const DevSim &S, float *x, hipStream_t s);
void launch_mgcg_iter(const DevSim &S, float *x, uint32_t it, bool last, hipStream_t s);
// MG-CG iterations of refinement pass `pass`: `inner` in the first pass; 2 in the second, which
// starts from the first pass's FP32 floor and usually only clears the last digits; 2 more in
// each pass after that (at most `inner`), for the hard steps that need them.
inline uint32_t pass_iterations(uint32_t inner, uint32_t pass) {
    return pass == 0 ? inner : (2 * pass < inner ? 2 * pass : inner);
}
// FP64 refinement (solver.hip); init/update also total the norms into scal[SC_RR / SC_BB].
void launch_mixed_init(const DevSim &S, double *x64, hipStream_t s);  // x64 = p, r64 = b - A x64
void launch_mixed_update(const DevSim &S, const double *xsrc, double *xdst, hipStream_t s);  // xdst = xsrc + e; r64
// (launch_mixed_update also stores pressure = float(xdst), so the pressure is current after
// every refinement pass.)
// bins / IO
void launch_bins_out(const DevSim &S, const uint32_t *off_lin, uint32_t *cell_of, uint32_t *sorted, hipStream_t s);
void launch_sort_segments(uint32_t ncells, const uint32_t *offsets, const uint32_t *counts, uint32_t *sorted,
                          hipStream_t s);
void launch_count_linear(const DevSim &S, uint32_t *out, hipStream_t s);  // counts, reference cell order
void launch_unpack_aos8(const DevSim &S, float *aos8, hipStream_t s);        // a -> AoS8 by id
void launch_pack_aos8(const DevSim &S, const float *aos8, hipStream_t s);    // AoS8 -> a, id = i
void launch_reset_particles(const DevSim &S, hipStream_t s);                 // initial -> a
void launch_type_u32(const DevSim &S, uint32_t *out, hipStream_t s);
void launch_acc_fields(const DevSim &S, int32_t *acc, int32_t *wacc, hipStream_t s);  // boxes -> face sums

// ---- batched launches: one launch runs a kernel for many simulations.
// A batched kernel is a persistent grid walking a device-side work list.  Entry
// (sim << WORK_SIM_SHIFT) | block runs the kernel body for workgroup `block` of simulation
// `sim`'s own single-sim grid, so every simulation computes exactly what it computes alone
// (same partition, same reduction order).  `sims` is a device array of DevSim.
constexpr uint32_t WORK_SIM_SHIFT = 20;
constexpr uint32_t WORK_BLOCK_MASK = (1u << WORK_SIM_SHIFT) - 1u;
constexpr uint32_t MAX_BATCH_SIMS = 1u << (32 - WORK_SIM_SHIFT);

// A batch keeps every member's list counts in one array, COUNT_SLOTS words per simulation:
// [0..1] = DevSim::fl_count, [4l..4l+2] = MGLevel::count of level l (the members' DevSim
// copies point there).  A batched kernel derives a member's slots from its work entry alone,
// so the counts load together with the member's DevSim fields instead of after them.
constexpr uint32_t COUNT_SLOTS = 4 * MG_MAX;
// Level stencil words (levels >= 1) carry the cell's colour in this bit (1 = black).
constexpr uint16_t STEN_BLACK = 0x0800;
// Every work list's storage holds at least this many entries.
constexpr uint32_t WORK_LIST_MIN_CAP = 16384;

struct WorkList {
    const DevSim *sims;
    const uint32_t *entries;
    const uint32_t *count;  // device word: number of entries
    uint32_t *slots;        // the batch's list counts (COUNT_SLOTS per simulation)
    const uint32_t *band, *band_count;  // the batch-wide band list (DevSim::band_all) and its count
    uint32_t grid;          // workgroups a batched launch over this list uses (host; 0 = the kernel's maximum)
};
// Batched launch families whose persistent-grid maximum is a run-time setting (the batch
// autotuner's knobs, pxlflip_batch_set_grid; the values are PXLFLIP_GRID_* of pxlflip.h).
// Results do not depend on them: every kernel walks its work list with a grid stride.
enum BatchFamily : int {
    BF_BAND = 0,  // faces, cells, fluid-list compaction, project
    BF_PART,      // particle and scan kernels (BLOCK lanes)
    BF_TILE,      // per-tile P2G / G2P
    BF_MG_GRID,   // multigrid level types and stencils
    BF_MG_LIST,   // multigrid list levels (pre0, sweep, restrict, post)
    BF_MG_TAIL,   // multigrid tail
    BF_SOLVER,    // solver list kernels (spmv, update, residual, refinement)
    BF_COUNT
};
unsigned batch_cap(int family);  // current maximum of a family (a multiple of 8)
bool set_batch_cap(int family, unsigned workgroups);  // false: unknown family or value out of range
// Batched launch families compiled in more than one workgroup size (the values are
// PXLFLIP_SHAPE_* of pxlflip.h): the per-tile P2G / G2P (512 or 1024 lanes) and the band
// passes faces, cells and project (256 or 512 lanes).  Lanes stride over a tile's particles,
// nodes or cells, so the size changes the speed, not the results.
enum BatchShape : int { BS_TILE = 0, BS_BAND, BS_COUNT };
unsigned batch_shape(int family);                 // current lanes per workgroup
bool set_batch_shape(int family, unsigned lanes);  // false: unknown family or a size not compiled

// Persistent grid of a batched launch: the list's grid, at most the kernel's maximum and
// never more than a list's minimum storage (for_each_work reads entry blockIdx.x unchecked).
inline unsigned batch_grid(const WorkList &w, unsigned max_grid) {
    const unsigned g = w.grid == 0 || w.grid > max_grid ? max_grid : w.grid;
    return g < WORK_LIST_MIN_CAP ? g : WORK_LIST_MIN_CAP;
}

// A device array as a raw buffer of `bytes` bytes (a uniform value: the base and size come from
// a member's descriptor).  A raw buffer load at a byte offset past the size returns 0 without
// touching memory, so a stencil kernel gives a neighbour that is not FLUID the offset
// buf_none(...) instead of branching around its load: every load of a lane is then issued
// together, and the masked ones cost no memory traffic.  0x00020000 is the gfx9-family
// descriptor word 3 (32-bit data format, no swizzle, no stride).
struct Buf {
    __amdgpu_buffer_rsrc_t r;
    uint32_t bytes;
};
__device__ __forceinline__ Buf buf_of(const void *p, uint32_t bytes) {
    return Buf{__builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000), bytes};
}
// Element i of a float buffer, or 0 when `use` is false (then no memory access).
__device__ __forceinline__ float buf_f32(const Buf &b, uint32_t i, bool use) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(b.r, (int)(use ? 4u * i : b.bytes), 0, 0));
}

// The 8-byte word at byte offset `off` of a buffer, or 0 for an offset past its size.
__device__ __forceinline__ int64_t buf_i64(const Buf &b, uint32_t off) {
    return __builtin_bit_cast(int64_t, __builtin_amdgcn_raw_buffer_load_b64(b.r, (int)off, 0, 0));
}

__device__ __forceinline__ double buf_f64(const Buf &b, uint32_t i, bool use) {
    return __builtin_bit_cast(double, buf_i64(b, use ? 8u * i : b.bytes));
}

// p[i] read as global memory, for a pointer the compiler cannot tell is global (one kept in
// LDS, say), which it would otherwise access with flat instructions.
template <typename T>
__device__ __forceinline__ T ld_global(const T *p, size_t i) {
    return ((const __attribute__((address_space(1))) T *)p)[i];
}

// A batch member's descriptor, read through the constant address space.  The descriptor
// array is written only by host copies before a launch, so it is constant for a kernel's
// lifetime; the compiler then fetches its fields with scalar loads and takes every pointer
// loaded from it to be a global-memory pointer (AMDGPU's rule for generic pointers read from
// constant memory).  Through a generic reference the batched bodies would address their arrays
// with flat instructions: 64-bit per-lane addresses, and waits that also cover the scalar
// loads, which kept a lane's neighbour loads from being in flight together.
__device__ __forceinline__ const DevSim &batch_member(const DevSim *sims, uint32_t sim) {
    const __attribute__((address_space(4))) DevSim *cs = (const __attribute__((address_space(4))) DevSim *)sims;
    return *(const DevSim *)(cs + sim);
}

// Runs f(S, block, slots) for each work entry of this workgroup.  The first entry is loaded
// together with the entry count (it lies in the list's storage even when the list is
// shorter), so a workgroup's first item waits for one load instead of two.
__device__ __forceinline__ uint32_t xcd_major(uint32_t b, uint32_t nb);
// XCD_MAJOR: workgroup b starts at entry xcd_major(b) instead of b, so the entries one round
// deals to an XCD are consecutive (used by the particle scatter, whose consecutive entries
// write consecutive runs: 15% faster on the 512-request batch; the per-tile and solver walks
// are slower that way).
template <bool XCD_MAJOR = false, typename F>
__device__ __forceinline__ void for_each_work(const WorkList &W, F &&f) {
    const uint32_t n = *W.count;
    uint32_t e = XCD_MAJOR ? xcd_major(blockIdx.x, gridDim.x) : blockIdx.x;
    uint32_t w = W.entries[e];
    while (e < n) {
        const uint32_t sim = w >> WORK_SIM_SHIFT;
        f(batch_member(W.sims, sim), w & WORK_BLOCK_MASK, W.slots + (size_t)sim * COUNT_SLOTS);
        e += gridDim.x;
        if (e < n) w = W.entries[e];
    }
}

// Virtual index of workgroup b of nb, where b mod 8 is the workgroup's XCD (workgroups are
// dealt to the 8 XCDs round-robin, and each XCD has its own L2): with nb a multiple of 8 the
// workgroups of one XCD get consecutive virtual indices, so list entries v and v + 1 taken in
// the same round of a strided loop (x-neighbouring tiles, mostly) share an L2.
__device__ __forceinline__ uint32_t xcd_major(uint32_t b, uint32_t nb) {
    return nb % 8u == 0 ? (b % 8u) * (nb / 8u) + b / 8u : b;
}

// Runs f(S, tile) for the batch's band tiles, one tile per workgroup iteration (the band kernel
// appends every member's band tiles to one batch-wide list, so the band passes balance over
// tiles, not over members' workgroup counts), in the XCD-major order of xcd_major: the rows of
// a tile's faces and cells are 32 bytes wide, and the x-neighbour tile that writes the rest of
// each cache line then writes it through the same L2 (measured on the 64-request batch: faces,
// cells and project 6-10% faster; the same order in the particle tile passes, whose writes are
// contiguous per tile, made them slower).
template <typename F>
__device__ __forceinline__ void for_batch_band(const WorkList &W, F &&f) {
    const uint32_t n = *W.band_count, g = gridDim.x;
    for (uint32_t e = xcd_major(blockIdx.x, g); e < n; e += g) {
        const uint32_t w = W.band[e];
        f(batch_member(W.sims, w >> WORK_SIM_SHIFT), w & WORK_BLOCK_MASK);
    }
}

// Work domains: which workgroups of a simulation a batched launch covers.  The grid-phase
// domains hold every simulation of the batch; the solve domains (WD_A_*) only those still
// iterating their pressure solve, and per multigrid level only those whose level runs as
// separate launches (level < mg_tail).
enum WorkDomain : int {
    WD_SIM = 0,     // one entry per simulation
    WD_PART,        // particles
    WD_SCAN_KEYS,   // scan tiles of the count array
    WD_TILES,       // tiles (active-tile list)
    WD_TILEGRID,    // per-tile workgroups (P2G, G2P, and the band passes: faces, cells, project)
    WD_SCAN_ROWS,   // scan tiles of the fluid list's row counts
    WD_ROWS,        // grid rows, 32 per workgroup (fluid-list compaction)
    WD_MGTYPES,     // + l, l = 1..MG_MAX-1: cells of level l
    WD_A_SIM = WD_MGTYPES + MG_MAX,
    WD_A_SOLVER,     // fluid-list solver workgroups
    WD_A_LVL,        // + l: the solver workgroups again, for simulations with level l < mg_tail
    WD_COUNT = WD_A_LVL + MG_MAX
};
inline bool wd_solve(int d) { return d >= WD_A_SIM; }

// Workgroups of simulation S in domain d (host side; the batched kernels derive the same
// counts from S where they need them).
uint32_t work_blocks(const DevSim &S, int d);
// per-file pieces of work_blocks
uint32_t particle_blocks(const DevSim &S);
uint32_t tile_blocks(const DevSim &S);
uint32_t tile_grid_blocks(const DevSim &S);
uint32_t mg_cell_blocks(const DevSim &S, int l);
uint32_t fluid_row_blocks(const DevSim &S);  // workgroups of the fluid-list compaction (8 rows per wave)
// Workgroups (list slices) of the multigrid list launches on level l: the solver's count on
// level 0, an eighth of it per coarsening after that (a coarse level holds about an eighth of
// the FLUID cells), at least one.
__host__ __device__ inline uint32_t level_blocks(int cg_blocks, int l) {
    if (l == 0) return (uint32_t)cg_blocks;
    const uint32_t g = ((uint32_t)cg_blocks + (1u << (3 * l)) - 1u) >> (3 * l);
    return g ? g : 1u;
}

// Batched launchers: W[d] is the work list of domain d.
using WorkLists = WorkList[WD_COUNT];
// grid phase
void batch_clear_grid(const WorkLists &W, hipStream_t s);
void batch_bin_count(const WorkLists &W, hipStream_t s);
void batch_scan_reduce(const WorkLists &W, int which, hipStream_t s);  // which 0: count -> offset
void batch_scan_tiles(const WorkLists &W, int which, hipStream_t s);   // which 1: fluid row counts
void batch_scan_down(const WorkLists &W, int which, hipStream_t s);
void batch_tile_list(const WorkLists &W, hipStream_t s);
void batch_band(const WorkLists &W, hipStream_t s);
void batch_scatter(const WorkLists &W, hipStream_t s);
void batch_p2g(const WorkLists &W, hipStream_t s);
void batch_faces(const WorkLists &W, hipStream_t s);
void batch_cells(const WorkLists &W, hipStream_t s);
void batch_fluid_compact(const WorkLists &W, hipStream_t s);
void batch_mg_types(const WorkLists &W, int l, hipStream_t s);  // level l >= 1: types
void batch_mg_sten(const WorkLists &W, int l, hipStream_t s);   // stencil words, z = 0, chunk counts
void batch_mg_scan(const WorkLists &W, hipStream_t s);           // every level: chunk offsets, list counts
void batch_mg_compact(const WorkLists &W, int l, hipStream_t s); // level l's FLUID lists in cell order
// solve (WD_A_* lists)
void batch_mixed_init(const WorkLists &W, hipStream_t s);
void batch_norms(const WorkLists &W, int nv, hipStream_t s);
void batch_cg_residual(const WorkLists &W, hipStream_t s);
void batch_mg_pre0(const WorkLists &W, int l, hipStream_t s);
void batch_mg_sweep(const WorkLists &W, int l, uint32_t colour, hipStream_t s);
void batch_mg_restrict(const WorkLists &W, int l, hipStream_t s);  // level l -> l + 1
void batch_mg_tail(const WorkLists &W, hipStream_t s);
void batch_mg_post(const WorkLists &W, int l, hipStream_t s);
void batch_mgcg_spmv(const WorkLists &W, uint32_t slot, int first, hipStream_t s);
void batch_mgcg_update(const WorkLists &W, uint32_t it, int last, hipStream_t s);  // last: x to e
void batch_mixed_update(const WorkLists &W, int cur, hipStream_t s);
// post
void batch_project(const WorkLists &W, hipStream_t s);
void batch_g2p(const WorkLists &W, hipStream_t s);

}  // namespace pxl
