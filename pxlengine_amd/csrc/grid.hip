// Grid-side kernels of the FLIP step: normalize + forces (one pass over faces), classify +
// stencil + divergence (one pass over cells), the fluid list and the pressure-gradient
// projection.  Semantics follow shaders/fluid.slang (cited per kernel).  The face, cell and
// projection passes walk the step's band of 8^3 tiles (DevSim::band_list; outside it every
// value is the one the previous step left), each workgroup taking band tiles with a stride;
// inside a tile lanes run x-fastest.
//
// Compiled with -ffp-contract=off: results are bitwise the reference formulas in fp32.
#include "common.hpp"
#include "kernels.hpp"

namespace pxl {

namespace {

constexpr int TX = 64, TY = 4;        // full-grid read-back launches
constexpr int GBLOCK = 256;           // band-walking workgroups

// Workgroups of the band-walking launches of one simulation (each takes one contiguous run of
// the band list), the same count as the per-tile particle kernels.
__host__ __device__ inline uint32_t band_grid(const DevSim &S) { return S.tile_wgs; }

// Runs f(i, j, k) over tile t's nodes (faces are stored at the node of their low corner: the
// tile's 8^3 nodes, plus the grid's upper boundary nodes on the last tile along an axis) or
// cells (8^3, clipped at the grid edge).  Lanes take the box x-fastest.
template <int NT = GBLOCK, typename F>
__device__ __forceinline__ void for_tile_box(const DevSim &S, uint32_t t, bool nodes, F &&f) {
    const int ti = (int)(t % (uint32_t)S.ntx), tj = (int)((t / (uint32_t)S.ntx) % (uint32_t)S.nty),
              tk = (int)(t / ((uint32_t)S.ntx * (uint32_t)S.nty));
    const int lo[3] = {ti * TILE, tj * TILE, tk * TILE};
    const int n[3] = {S.nx, S.ny, S.nz}, nt[3] = {S.ntx, S.nty, S.ntz}, tt[3] = {ti, tj, tk};
    int ext[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const int end = nodes ? (tt[a] == nt[a] - 1 ? n[a] + 1 : lo[a] + TILE) : (lo[a] + TILE < n[a] ? lo[a] + TILE : n[a]);
        ext[a] = end - lo[a];
    }
    const int total = ext[0] * ext[1] * ext[2];
    for (int q = threadIdx.x; q < total; q += NT) {
        const int i = lo[0] + q % ext[0], j = lo[1] + (q / ext[0]) % ext[1], k = lo[2] + q / (ext[0] * ext[1]);
        f(i, j, k);
    }
}
// Workgroup b's (of nb) contiguous run of simulation S's band list: x-neighbouring tiles are
// mostly adjacent in the list, so one workgroup writes the rest of a cache line row soon after
// the first part, while it is still in L2.  With nb a multiple of 8, XCD x (the workgroups
// b = x mod 8) takes the x-th eighth of the list, so the runs either side of a run share its
// L2 (128^3 dam: faces, cells and project 5-15% faster).
template <typename F>
__device__ __forceinline__ void for_band(const DevSim &S, uint32_t b, uint32_t nb, F &&f) {
    const uint32_t n = *S.band_count;
    if (nb % 8u == 0) b = (b % 8u) * (nb / 8u) + b / 8u;
    const uint32_t beg = (uint32_t)((uint64_t)n * b / nb), end = (uint32_t)((uint64_t)n * (b + 1) / nb);
    for (uint32_t u = beg; u < end; u++) f(S.band_list[u]);
}

__device__ __forceinline__ uint32_t cidx(const DevSim &S, int i, int j, int k) {
    return (uint32_t)i + (uint32_t)S.nx * ((uint32_t)j + (uint32_t)S.ny * (uint32_t)k);
}

// SOLID for cells outside the grid or on its one-cell wall layer (fluid.slang:55-58, 126-128),
// and for obstacle cells (an extension).
__device__ __forceinline__ bool is_solid(const DevSim &S, int i, int j, int k) {
    if (i <= 0 || j <= 0 || k <= 0 || i >= S.nx - 1 || j >= S.ny - 1 || k >= S.nz - 1) return true;
    if (!S.solid_bits) return false;
    const uint32_t id = cidx(S, i, j, k);
    return ((S.solid_bits[id >> 5] >> (id & 31u)) & 1u) != 0;
}
__device__ __forceinline__ uint32_t uidx(const DevSim &S, int i, int j, int k) {
    return (uint32_t)i + (uint32_t)(S.nx + 1) * ((uint32_t)j + (uint32_t)S.ny * (uint32_t)k);
}
__device__ __forceinline__ uint32_t vidx(const DevSim &S, int i, int j, int k) {
    return S.nu + (uint32_t)i + (uint32_t)S.nx * ((uint32_t)j + (uint32_t)(S.ny + 1) * (uint32_t)k);
}
__device__ __forceinline__ uint32_t widx(const DevSim &S, int i, int j, int k) {
    return S.nu + S.nv + (uint32_t)i + (uint32_t)S.nx * ((uint32_t)j + (uint32_t)S.ny * (uint32_t)k);
}

// The 3x3x3 tiles around tile (ti, tj, tk), bit (dx+1) + 3 (dy+1) + 9 (dz+1) (tiles past the
// grid are 0): `act` = tiles holding particles this step (their P2G boxes are current), `core`
// = tiles holding particles this step or the step before.  One load per lane and two ballots.
constexpr uint32_t NB_SELF = 1u << 13;
__device__ __forceinline__ void neighbour_masks(const DevSim &S, int ti, int tj, int tk, uint32_t &act, uint32_t &core) {
    const int lane = threadIdx.x & (WAVE - 1);
    uint32_t v = 0;
    if (lane < 27) {
        const int x = ti + lane % 3 - 1, y = tj + (lane / 3) % 3 - 1, z = tk + lane / 9 - 1;
        if (x >= 0 && y >= 0 && z >= 0 && x < S.ntx && y < S.nty && z < S.ntz)
            v = S.tile_act[(uint32_t)x + (uint32_t)S.ntx * ((uint32_t)y + (uint32_t)S.nty * (uint32_t)z)];
    }
    act = (uint32_t)__ballot((v & 1u) != 0);
    core = (uint32_t)__ballot(v != 0);
}

// A rim tile of the band (no particles this step or the step before; in the band because a
// neighbour has) keeps last step's values except where a core neighbour touches it: its P2G
// faces within one layer of that neighbour (the neighbour's box reaches one cell in; a core
// neighbour that just emptied leaves faces there to reset), and faces between its cells and
// FLUID cells below it.  Everything else in it has no particle weight and no FLUID cell near,
// and was already written with those values when the tile last had a core neighbour there.
// True when node (u0, u1, u2) of the tile (local coordinates, 8 on the grid's upper planes)
// lies within one layer of a core neighbour.
__device__ __forceinline__ bool rim_node_touched(uint32_t core, int u0, int u1, int u2) {
    const int u[3] = {u0, u1, u2};
    uint32_t m[3];  // allowed offsets per axis: bit d+1
#pragma unroll
    for (int a = 0; a < 3; a++) m[a] = 2u | (u[a] == 0 ? 1u : 0u) | (u[a] >= TILE - 1 ? 4u : 0u);
    uint32_t allowed = 0;
#pragma unroll
    for (int dz = 0; dz < 3; dz++)
#pragma unroll
        for (int dy = 0; dy < 3; dy++)
#pragma unroll
            for (int dx = 0; dx < 3; dx++)
                if (((m[0] >> dx) & (m[1] >> dy) & (m[2] >> dz)) & 1u) allowed |= 1u << (dx + 3 * dy + 9 * dz);
    return (allowed & core & ~NB_SELF) != 0u;
}

// Packed P2G sum of face (c, q) for a face pass walking tile t = (t0, t1, t2) (q in t's node
// box): the boxes that contain the face are t's and, per axis, the tile below when q is on t's
// lower face plane, or the tile above when q is on its last cell row along an axis other than
// c (the same tiles gather_face visits).  The active ones come from the mask, and all 8 box
// words are read with raw buffer loads (`boxes`: the simulation's tile boxes), a box that is
// not among them at an offset past the buffer (0, no memory access): no branch, so the loads
// are in flight together.  Integer sums: the order does not matter.
__device__ __forceinline__ int64_t gather_face_in(const DevSim &S, const Buf &boxes, uint32_t act, int c,
                                                  const int q[3], const int t[3]) {
    int alt[3], u[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        u[a] = q[a] - TILE * t[a];
        alt[a] = u[a] == 0 ? -1 : ((a != c && u[a] == TILE - 1) ? 1 : 0);
    }
    const uint32_t t_lin = (uint32_t)t[0] + (uint32_t)S.ntx * ((uint32_t)t[1] + (uint32_t)S.nty * (uint32_t)t[2]);
    int64_t v[8];
#pragma unroll
    for (int e = 0; e < 8; e++) {
        const int dx = (e & 1) ? alt[0] : 0, dy = (e & 2) ? alt[1] : 0, dz = (e & 4) ? alt[2] : 0;
        const bool valid = (!(e & 1) || alt[0]) && (!(e & 2) || alt[1]) && (!(e & 4) || alt[2]);
        const bool use = valid && ((act >> ((dx + 1) + 3 * (dy + 1) + 9 * (dz + 1))) & 1u);
        // the face's position in the box of tile t + d (a box starts one cell below its tile
        // along the axes other than c)
        const int lx = u[0] - TILE * dx + (c != 0), ly = u[1] - TILE * dy + (c != 1), lz = u[2] - TILE * dz + (c != 2);
        const uint32_t tn = t_lin + (uint32_t)(dx + S.ntx * (dy + S.nty * dz));
        v[e] = buf_i64(boxes, use ? 8u * (tn * (uint32_t)BOX_ENTRIES + (uint32_t)box_gidx(c, lx, ly, lz)) : boxes.bytes);
    }
    int64_t sum = 0;
#pragma unroll
    for (int e = 0; e < 8; e++) sum += v[e];
    return sum;
}

// normalize_grid faces + forces (fluid.slang:119-124, 143-151) for one face: the P2G sums
// acc, wacc are gathered from the tile boxes that cover the face; v = acc / wacc (0 without
// weight) -> vel_old; vel = v + g_c * dt, or 0 when either side is SOLID.  SOLID depends only
// on the cell position, so this needs no classified cells.
__device__ __forceinline__ void face_update(const DevSim &S, int64_t packed, uint32_t f, float g, bool keep) {
    const int32_t w = packed_wacc(packed);
    const float v = w > 0 ? (float)packed_acc(packed) / (float)w : 0.0f;
    S.vel_old[f] = v;
    S.vel[f] = keep ? v + g * S.dt : 0.0f;
}

__device__ __forceinline__ void faces_at(const DevSim &S, const Buf &boxes, uint32_t act, const int t[3], int i, int j,
                                         int k) {
    const bool in_c = !is_solid(S, i, j, k);
    const int q[3] = {i, j, k};
    const bool has[3] = {j < S.ny && k < S.nz, i < S.nx && k < S.nz, i < S.nx && j < S.ny};
    int64_t sum[3];
#pragma unroll
    for (int c = 0; c < 3; c++) sum[c] = gather_face_in(S, boxes, has[c] ? act : 0u, c, q, t);
    if (has[0]) face_update(S, sum[0], uidx(S, i, j, k), S.gx, in_c && !is_solid(S, i - 1, j, k));
    if (has[1]) face_update(S, sum[1], vidx(S, i, j, k), S.gy, in_c && !is_solid(S, i, j - 1, k));
    if (has[2]) face_update(S, sum[2], widx(S, i, j, k), S.gz, in_c && !is_solid(S, i, j, k - 1));
}
template <int NT>
__device__ __forceinline__ void faces_tile(const DevSim &S, uint32_t tl) {
    const int t[3] = {(int)(tl % (uint32_t)S.ntx), (int)((tl / (uint32_t)S.ntx) % (uint32_t)S.nty),
                      (int)(tl / ((uint32_t)S.ntx * (uint32_t)S.nty))};
    uint32_t act, core;
    neighbour_masks(S, t[0], t[1], t[2], act, core);
    const Buf boxes = buf_of(S.box, 8u * (uint32_t)BOX_ENTRIES * S.ntiles);
    if (core & NB_SELF) {
        for_tile_box<NT>(S, tl, true, [&](int i, int j, int k) { faces_at(S, boxes, act, t, i, j, k); });
    } else {  // rim tile: only the layers a core neighbour touches
        for_tile_box<NT>(S, tl, true, [&](int i, int j, int k) {
            if (rim_node_touched(core, i - TILE * t[0], j - TILE * t[1], k - TILE * t[2]))
                faces_at(S, boxes, act, t, i, j, k);
        });
    }
}
__device__ __forceinline__ void body_faces(const DevSim &S, uint32_t b, uint32_t nb) {
    for_band(S, b, nb, [&](uint32_t tl) { faces_tile<GBLOCK>(S, tl); });
}
__global__ __launch_bounds__(GBLOCK) void k_faces(DevSim S) { body_faces(S, blockIdx.x, gridDim.x); }
template <int NT>
__global__ __launch_bounds__(NT) void kb_faces(WorkList W) {
    for_batch_band(W, [&](const DevSim &S, uint32_t t) { faces_tile<NT>(S, t); });
}

// The P2G sums as the reference's acc / wacc face buffers (for read-back).
__global__ __launch_bounds__(TX *TY) void k_acc_fields(DevSim S, int32_t *acc, int32_t *wacc) {
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    const int k = blockIdx.z;
    if (i > S.nx || j > S.ny) return;
    int64_t v;
    if (j < S.ny && k < S.nz) { v = gather_face(S, 0, i, j, k); acc[uidx(S, i, j, k)] = packed_acc(v); wacc[uidx(S, i, j, k)] = packed_wacc(v); }
    if (i < S.nx && k < S.nz) { v = gather_face(S, 1, i, j, k); acc[vidx(S, i, j, k)] = packed_acc(v); wacc[vidx(S, i, j, k)] = packed_wacc(v); }
    if (i < S.nx && j < S.ny) { v = gather_face(S, 2, i, j, k); acc[widx(S, i, j, k)] = packed_acc(v); wacc[widx(S, i, j, k)] = packed_wacc(v); }
}

// The types of cell (i, j, k) and its six neighbours (-x, +x, -y, +y, -z, +z): SOLID on or past
// the wall layer (is_solid) and in obstacle cells, else FLUID where the cell holds particles,
// else AIR.  Every load is issued together: a cell on or past the wall layer reads tile key 0
// and solid word 0, and its values are dropped.
__device__ __forceinline__ void types7(const DevSim &S, int i, int j, int k, uint8_t t[7]) {
    const int nb[7][3] = {{i, j, k}, {i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k}, {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
    bool wall[7];
    uint32_t cnt[7], sb[7];
#pragma unroll
    for (int n = 0; n < 7; n++) {
        const int a = nb[n][0], b = nb[n][1], c = nb[n][2];
        wall[n] = a <= 0 || b <= 0 || c <= 0 || a >= S.nx - 1 || b >= S.ny - 1 || c >= S.nz - 1;
        cnt[n] = S.count[wall[n] ? 0u : tile_key(S, a, b, c)];
        sb[n] = 0;
    }
    if (S.solid_bits) {  // obstacles (uniform)
#pragma unroll
        for (int n = 0; n < 7; n++) {
            const uint32_t id = wall[n] ? 0u : cidx(S, nb[n][0], nb[n][1], nb[n][2]);
            sb[n] = (S.solid_bits[id >> 5] >> (id & 31u)) & 1u;
        }
    }
#pragma unroll
    for (int n = 0; n < 7; n++) t[n] = (wall[n] || sb[n]) ? CELL_SOLID : (cnt[n] > 0 ? CELL_FLUID : CELL_AIR);
}

// normalize_grid cells + compute_divergence (fluid.slang:125-131, 154-165), plus the stencil
// word the pressure kernels read instead of re-classifying six neighbours.  Returns the type.
__device__ __forceinline__ uint8_t cells_at(const DevSim &S, int i, int j, int k) {
    const uint32_t id = cidx(S, i, j, k);
    uint8_t tn[7];
    types7(S, i, j, k, tn);
    const uint8_t t = tn[0];
    S.type[id] = t;
    uint16_t st = 0;
    float d = 0.0f;
    if (t == CELL_FLUID) {
        uint32_t diag = 0;
#pragma unroll
        for (int n = 0; n < 6; n++) {
            const uint8_t tq = tn[n + 1];
            if (tq != CELL_SOLID) diag++;
            if (tq == CELL_FLUID) st |= (uint16_t)(1u << n);
        }
        st |= (uint16_t)(STEN_FLUID | (diag << 8));
        d = d + (S.vel[uidx(S, i + 1, j, k)] - S.vel[uidx(S, i, j, k)]);
        d = d + (S.vel[vidx(S, i, j + 1, k)] - S.vel[vidx(S, i, j, k)]);
        d = d + (S.vel[widx(S, i, j, k + 1)] - S.vel[widx(S, i, j, k)]);
    } else {
        S.pressure[id] = 0.0f;
    }
    S.sten[id] = st;
    S.div[id] = d * S.inv_h;
    return t;
}

// Adds a wave's FLUID cells to the fluid list's row counts.  Lanes hold consecutive cells of
// a tile box (x fastest), so lanes of one grid row form a run; the run's first lane adds the
// run's counts with one atomic per colour (integer sums: the result is independent of order).
// `row` is ~0u on lanes without a cell.  Called by every lane of the wave.
__device__ __forceinline__ void count_row_runs(const DevSim &S, uint32_t row, bool first_half, bool second_half) {
    const int lane = threadIdx.x & (WAVE - 1);
    const uint32_t prev = __shfl_up(row, 1, WAVE);
    const bool start = lane == 0 || prev != row;
    const uint64_t starts = __ballot(start), ma = __ballot(first_half), mb = __ballot(second_half);
    if (!start || row == ~0u) return;
    const uint64_t above = starts & ~((2ull << lane) - 1ull);  // run starts after this lane
    const uint64_t upto = above ? (above & (~above + 1ull)) - 1ull : ~0ull;
    const uint64_t run = upto & ~((1ull << lane) - 1ull);
    const uint32_t na = (uint32_t)__popcll(ma & run), nb = (uint32_t)__popcll(mb & run);
    if (na) atomicAdd(&S.fl_rows[row], na);
    if (nb) atomicAdd(&S.fl_rows[(uint32_t)S.ny * (uint32_t)S.nz + row], nb);
}

// The cell pass over band tile t: every lane of the workgroup runs the same number of rounds
// (the row counting needs whole waves).
// (A rim tile is skipped: its cells have no particles this step or the last, so they are AIR
// or SOLID with zero pressure, divergence and stencil, as its last pass as a core tile left them.)
template <int NT>
__device__ __forceinline__ void cells_tile(const DevSim &S, uint32_t t) {
    if (S.tile_act[t] == 0) return;  // the whole workgroup
    const int ti = (int)(t % (uint32_t)S.ntx), tj = (int)((t / (uint32_t)S.ntx) % (uint32_t)S.nty),
              tk = (int)(t / ((uint32_t)S.ntx * (uint32_t)S.nty));
    const int lo[3] = {ti * TILE, tj * TILE, tk * TILE};
    const int ex = min(TILE, S.nx - lo[0]), ey = min(TILE, S.ny - lo[1]), ez = min(TILE, S.nz - lo[2]);
    const int total = ex * ey * ez;
    for (int q0 = 0; q0 < total; q0 += NT) {
        const int q = q0 + (int)threadIdx.x;
        uint32_t row = ~0u;
        bool fluid = false, red = false;
        if (q < total) {
            const int i = lo[0] + q % ex, j = lo[1] + (q / ex) % ey, k = lo[2] + q / (ex * ey);
            fluid = cells_at(S, i, j, k) == CELL_FLUID;
            red = (((uint32_t)(i + j + k)) & 1u) == 0;
            row = (uint32_t)j + (uint32_t)S.ny * (uint32_t)k;
        }
        const bool split = S.split_colours != 0;
        count_row_runs(S, row, fluid && (red || !split), fluid && !red && split);
    }
}
__device__ __forceinline__ void body_cells(const DevSim &S, uint32_t b, uint32_t nb) {
    for_band(S, b, nb, [&](uint32_t t) { cells_tile<GBLOCK>(S, t); });
}
__global__ __launch_bounds__(GBLOCK) void k_cells(DevSim S) { body_cells(S, blockIdx.x, gridDim.x); }
template <int NT>
__global__ __launch_bounds__(NT) void kb_cells(WorkList W) {
    for_batch_band(W, [&](const DevSim &S, uint32_t t) { cells_tile<NT>(S, t); });
}

// Fluid list from the scanned row counts: a wave writes a grid row's FLUID cells in x order
// at the row's slots (red ones at fl_rowoff[row], black ones at fl_rowoff[R + row] when
// colours are split), so the list is in cell order within each colour.  Each entry also gets
// its stencil word (see DevSim::fl_word).  With colours split the row's cells also go, in x
// order, to the cell-order list fn_list at slot rowoff[row] + rowoff[R + row] - n_red (the
// FLUID cells of all earlier rows).  Workgroup 0 writes fl_count = {red cells (0 unless
// split), all fluid cells}.
// A wave takes ROWS_PER_WAVE consecutive rows: its lanes load those rows' counts and offsets
// together, and it walks the non-empty ones, loading each row's stencil words and their six
// neighbours' (ROW_CHUNKS 64-cell chunks at a time, at indices clamped into the grid and used
// only where the cell's FLUID mask says so) before it uses any.  A wave per row spent most of
// its time waiting on each row's offsets, then on its words, then on its neighbours' words
// (measured per launch: 64-request batch 155 -> 147-150 us, 512-request batch 1.20 -> 1.04 ms;
// 16 rows per wave: 173 us and 0.94 ms).
constexpr int ROWS_PER_WAVE = 8;
constexpr int ROWS_PER_BLOCK = (BLOCK / WAVE) * ROWS_PER_WAVE;
constexpr int ROW_CHUNKS = 2;
__device__ __forceinline__ void fluid_row(const DevSim &S, uint32_t row, uint32_t oa, uint32_t ob, uint32_t on,
                                          bool split, int lane) {
    const uint32_t j = row % (uint32_t)S.ny, k = row / (uint32_t)S.ny, base = (uint32_t)S.nx * row;
    const int32_t sy = S.nx, sz = S.nx * S.ny;
    const int32_t off[6] = {-1, 1, -sy, sy, -sz, sz};
    const int32_t last = (int32_t)S.ncells - 1;
    const uint64_t below = (1ull << lane) - 1ull;
    for (int i0 = 0; i0 < S.nx; i0 += ROW_CHUNKS * WAVE) {
        uint32_t st[ROW_CHUNKS], ns[ROW_CHUNKS][6];
#pragma unroll
        for (int c = 0; c < ROW_CHUNKS; c++) {
            // unconditional loads at indices clamped into the grid (a branch per load would make
            // each wait for the one before), values past the row dropped afterwards
            const int i = i0 + c * WAVE + lane;
            const bool in = i < S.nx;
            const int32_t id = min((int32_t)base + i, last);
            const uint32_t s0 = S.sten[id];
            st[c] = in ? s0 : 0u;
#pragma unroll
            for (int n = 0; n < 6; n++) {
                const uint32_t sn = S.sten[min(max(id + off[n], 0), last)];
                ns[c][n] = in ? sn : 0u;
            }
        }
#pragma unroll
        for (int c = 0; c < ROW_CHUNKS; c++) {
            const int i = i0 + c * WAVE + lane;
            const uint32_t id = base + (uint32_t)i;
            const bool fl = (st[c] & STEN_FLUID) != 0;
            const uint32_t black = ((uint32_t)i + j + k) & 1u;
            const bool a = fl && (!black || !split), bk = fl && black && split;
            const uint64_t ma = __ballot(a), mb = __ballot(bk);
            if (a || bk) {
                const uint32_t o = a ? oa + (uint32_t)__popcll(ma & below) : ob + (uint32_t)__popcll(mb & below);
                uint32_t word = (st[c] & 63u) | (((st[c] >> 8) & 7u) << 6);
#pragma unroll
                for (int n = 0; n < 6; n++)
                    if (st[c] & (1u << n)) word |= ((ns[c][n] >> 8) & 7u) << (9 + 3 * n);
                word |= black << 27;
                word |= ((uint32_t)i & 1u) << 28;
                S.fl_list[o] = id;
                S.fl_word[o] = word;
                const uint32_t sid = black * S.hsplit + ((uint32_t)i >> 1) + S.hx * row;
                if (S.fls_list) S.fls_list[o] = sid;
                if (split) {
                    const uint32_t n = on + (uint32_t)__popcll((ma | mb) & below);
                    S.fn_list[n] = id;
                    S.fn_word[n] = word;
                    if (S.fns_list) S.fns_list[n] = sid;
                }
            }
            oa += (uint32_t)__popcll(ma);
            ob += (uint32_t)__popcll(mb);
            on += (uint32_t)__popcll(ma | mb);
        }
    }
}
__device__ __forceinline__ void body_fluid_rows(const DevSim &S, uint32_t b) {
    const uint32_t R = (uint32_t)S.ny * (uint32_t)S.nz;
    const int lane = threadIdx.x & (WAVE - 1);
    const bool split = S.split_colours != 0;
    if (b == 0 && threadIdx.x == 0) {  // row 0's wave: fl_count = {red cells (0 unless split), all}
        S.fl_count[0] = split ? S.fl_rowoff[R] : 0u;
        S.fl_count[1] = S.fl_rowoff[2 * R - 1] + S.fl_rows[2 * R - 1];
    }
    const uint32_t r0 = b * ROWS_PER_BLOCK + (threadIdx.x >> 6) * ROWS_PER_WAVE;
    if (r0 >= R) return;
    const uint32_t myrow = r0 + (uint32_t)lane;
    const bool mine = lane < ROWS_PER_WAVE && myrow < R;
    uint32_t na = 0, nbl = 0, oa = 0, ob = 0;
    if (mine) {
        na = S.fl_rows[myrow];
        nbl = S.fl_rows[R + myrow];
        oa = S.fl_rowoff[myrow];
        ob = S.fl_rowoff[R + myrow];
    }
    const uint32_t red_total = split ? S.fl_rowoff[R] : 0u;
    uint64_t rows = __ballot(mine && na + nbl != 0);
    while (rows) {
        const int l = __ffsll((long long)rows) - 1;
        rows &= rows - 1;
        const uint32_t roa = __shfl(oa, l, WAVE), rob = __shfl(ob, l, WAVE);
        fluid_row(S, r0 + (uint32_t)l, roa, rob, split ? roa + rob - red_total : 0u, split, lane);
    }
}
__global__ __launch_bounds__(BLOCK) void k_fluid_rows(DevSim S) { body_fluid_rows(S, blockIdx.x); }
__global__ __launch_bounds__(BLOCK) void kb_fluid_compact(WorkList W) {
    for_each_work(W, [&](const DevSim &S, uint32_t b, const uint32_t *) { body_fluid_rows(S, b); });
}

__device__ __forceinline__ uint8_t type_at(const DevSim &S, int i, int j, int k) {
    if (i < 0 || j < 0 || k < 0 || i >= S.nx || j >= S.ny || k >= S.nz) return CELL_SOLID;
    return S.type[cidx(S, i, j, k)];
}

// project (fluid.slang:185-197) for one face between cells a (low side) and b (high side).
__device__ __forceinline__ void project_face(const DevSim &S, uint32_t f, int ai, int aj, int ak, int bi,
                                             int bj, int bk) {
    const uint8_t ta = type_at(S, ai, aj, ak), tb = type_at(S, bi, bj, bk);
    if (ta == CELL_SOLID || tb == CELL_SOLID) {
        S.vel[f] = 0.0f;
        return;
    }
    if (ta == CELL_FLUID || tb == CELL_FLUID) {
        const float pa = ta == CELL_FLUID ? S.pressure[cidx(S, ai, aj, ak)] : 0.0f;
        const float pb = tb == CELL_FLUID ? S.pressure[cidx(S, bi, bj, bk)] : 0.0f;
        S.vel[f] = S.vel[f] - S.dt * S.inv_h * (pb - pa);
    }
}

// project_face for the node's three faces (cell b = (i, j, k) on the high side of each) in two
// rounds of loads without a branch: the four cells' types (index 0 for a cell outside the grid,
// whose type then counts as SOLID), then, as masked raw buffer loads (no memory access where
// not needed), the FLUID cells' pressures and the velocities of the faces that change.
__device__ __forceinline__ void project_at(const DevSim &S, const Buf &bp, const Buf &bv, int i, int j, int k) {
    const bool has[3] = {j < S.ny && k < S.nz, i < S.nx && k < S.nz, i < S.nx && j < S.ny};
    const uint32_t f[3] = {has[0] ? uidx(S, i, j, k) : 0u, has[1] ? vidx(S, i, j, k) : 0u, has[2] ? widx(S, i, j, k) : 0u};
    const int ca[3][3] = {{i - 1, j, k}, {i, j - 1, k}, {i, j, k - 1}};
    const bool in_b = i < S.nx && j < S.ny && k < S.nz;
    const uint32_t ib = in_b ? cidx(S, i, j, k) : 0u;
    bool in_a[3];
    uint32_t ia[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        in_a[c] = ca[c][0] >= 0 && ca[c][1] >= 0 && ca[c][2] >= 0 && ca[c][0] < S.nx && ca[c][1] < S.ny && ca[c][2] < S.nz;
        ia[c] = in_a[c] ? cidx(S, ca[c][0], ca[c][1], ca[c][2]) : 0u;
    }
    const uint8_t tb_raw = S.type[ib];
    uint8_t ta_raw[3];
#pragma unroll
    for (int c = 0; c < 3; c++) ta_raw[c] = S.type[ia[c]];
    const uint8_t tb = in_b ? tb_raw : (uint8_t)CELL_SOLID;
    uint8_t ta[3];
    bool solid[3], upd[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        ta[c] = in_a[c] ? ta_raw[c] : (uint8_t)CELL_SOLID;
        solid[c] = ta[c] == CELL_SOLID || tb == CELL_SOLID;
        upd[c] = has[c] && !solid[c] && (ta[c] == CELL_FLUID || tb == CELL_FLUID);
    }
    const float pb = buf_f32(bp, ib, tb == CELL_FLUID);
    float pa[3], v[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        pa[c] = buf_f32(bp, ia[c], upd[c] && ta[c] == CELL_FLUID);
        v[c] = buf_f32(bv, f[c], upd[c]);
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        if (!has[c]) continue;
        if (solid[c]) S.vel[f[c]] = 0.0f;
        else if (upd[c]) S.vel[f[c]] = v[c] - S.dt * S.inv_h * (pb - pa[c]);
    }
}
template <int NT>
__device__ __forceinline__ void project_tile(const DevSim &S, uint32_t t) {
    if (S.tile_act[t] != 0) {
        const Buf bp = buf_of(S.pressure, 4u * S.ncells), bv = buf_of(S.vel, 4u * S.nfaces);
        for_tile_box<NT>(S, t, true, [&](int i, int j, int k) { project_at(S, bp, bv, i, j, k); });
        return;
    }
    // rim tile: its cells are not FLUID, so only faces on its lower planes can border a FLUID
    // cell, in the tile below along that axis, which then holds particles this step
    const int ti = (int)(t % (uint32_t)S.ntx), tj = (int)((t / (uint32_t)S.ntx) % (uint32_t)S.nty),
              tk = (int)(t / ((uint32_t)S.ntx * (uint32_t)S.nty));
    uint32_t act, core;
    neighbour_masks(S, ti, tj, tk, act, core);
    const bool below[3] = {((act >> 12) & 1u) != 0, ((act >> 10) & 1u) != 0, ((act >> 4) & 1u) != 0};
    if (!(below[0] || below[1] || below[2])) return;  // uniform: the masks are the workgroup's
    for_tile_box<NT>(S, t, true, [&](int i, int j, int k) {
        if (below[0] && i == TILE * ti && j < S.ny && k < S.nz) project_face(S, uidx(S, i, j, k), i - 1, j, k, i, j, k);
        if (below[1] && j == TILE * tj && i < S.nx && k < S.nz) project_face(S, vidx(S, i, j, k), i, j - 1, k, i, j, k);
        if (below[2] && k == TILE * tk && i < S.nx && j < S.ny) project_face(S, widx(S, i, j, k), i, j, k - 1, i, j, k);
    });
}
__device__ __forceinline__ void body_project(const DevSim &S, uint32_t b, uint32_t nb) {
    for_band(S, b, nb, [&](uint32_t t) { project_tile<GBLOCK>(S, t); });
}
__global__ __launch_bounds__(GBLOCK) void k_project(DevSim S) { body_project(S, blockIdx.x, gridDim.x); }
template <int NT>
__global__ __launch_bounds__(NT) void kb_project(WorkList W) {
    for_batch_band(W, [&](const DevSim &S, uint32_t t) { project_tile<NT>(S, t); });
}

__global__ __launch_bounds__(TX *TY) void k_type_u32(DevSim S, uint32_t *out) {
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= S.nx || j >= S.ny) return;
    const uint32_t id = cidx(S, i, j, k);
    out[id] = S.type[id];
}

dim3 cell_grid(const DevSim &S) { return dim3((S.nx + TX - 1) / TX, (S.ny + TY - 1) / TY, S.nz); }
dim3 node_grid(const DevSim &S) { return dim3((S.nx + 1 + TX - 1) / TX, (S.ny + 1 + TY - 1) / TY, S.nz + 1); }

}  // namespace

void launch_faces(const DevSim &S, hipStream_t s) { PXL_LAUNCH(k_faces, band_grid(S), GBLOCK, 0, s, S); }
void launch_cells(const DevSim &S, hipStream_t s) { PXL_LAUNCH(k_cells, band_grid(S), GBLOCK, 0, s, S); }
void launch_fluid_list(const DevSim &S, hipStream_t s) {
    launch_exclusive_scan(S.fl_rows, fl_row_slots(S), S.fl_tmp, S.fl_rowoff, s);
    PXL_LAUNCH(k_fluid_rows, fluid_row_blocks(S), BLOCK, 0, s, S);
}
void launch_project(const DevSim &S, hipStream_t s) { PXL_LAUNCH(k_project, band_grid(S), GBLOCK, 0, s, S); }
void launch_acc_fields(const DevSim &S, int32_t *acc, int32_t *wacc, hipStream_t s) {
    PXL_LAUNCH(k_acc_fields, node_grid(S), dim3(TX, TY), 0, s, S, acc, wacc);
}
void launch_type_u32(const DevSim &S, uint32_t *out, hipStream_t s) {
    PXL_LAUNCH(k_type_u32, cell_grid(S), dim3(TX, TY), 0, s, S, out);
}

// ---- batched
uint32_t fluid_row_blocks(const DevSim &S) {
    return ((uint32_t)S.ny * (uint32_t)S.nz + ROWS_PER_BLOCK - 1) / ROWS_PER_BLOCK;
}

// The band launches in the workgroup size the band-lane setting picks (256 or 512 lanes: lanes
// stride over a tile's nodes or cells, so the size changes the speed, not the results).
#define PXL_BAND_LAUNCH(kernel)                                                              \
    do {                                                                                     \
        const unsigned g_ = batch_grid(W[WD_TILEGRID], batch_cap(BF_BAND));                   \
        if (batch_shape(BS_BAND) == 512) PXL_LAUNCH(kernel<512>, g_, 512, 0, s, W[WD_TILEGRID]); \
        else PXL_LAUNCH(kernel<GBLOCK>, g_, GBLOCK, 0, s, W[WD_TILEGRID]);                    \
    } while (0)
void batch_faces(const WorkLists &W, hipStream_t s) { PXL_BAND_LAUNCH(kb_faces); }
void batch_cells(const WorkLists &W, hipStream_t s) { PXL_BAND_LAUNCH(kb_cells); }
void batch_fluid_compact(const WorkLists &W, hipStream_t s) {
    PXL_LAUNCH(kb_fluid_compact, batch_grid(W[WD_ROWS], batch_cap(BF_BAND)), BLOCK, 0, s, W[WD_ROWS]);
}
void batch_project(const WorkLists &W, hipStream_t s) { PXL_BAND_LAUNCH(kb_project); }

}  // namespace pxl
