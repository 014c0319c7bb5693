// Surface nets of the liquid (shaders/fluid.slang:222-355, driven once per frame by
// crates/pxl-render/src/fluid.rs:308-330): particles splat a poly6-style density onto the
// nodes of a grid `surface_refine` times finer than the simulation's, a 7-point smoothing
// minus the iso level gives phi (> 0 inside), every cell whose corners change sign gets one
// vertex at the mean of its edge crossings, and every sign-changing node edge whose four
// cells hold vertices gets a quad (two triangles) wound from inside to outside.
//
// The reference numbers vertices and triangles with atomic counters, so its order follows
// thread timing.  Here both are numbered by exclusive scans -- vertices in surface-cell
// order, quads in (node, axis) order -- so the mesh is deterministic and equals the oracle's
// sequential restatement bit for bit.  The counters keep the reference's meaning (every
// surface cell counts, every quad whose four cells stored a vertex counts twice) and the
// clipping rules are the reference's: a vertex past max_verts is not stored (its cell holds
// none, so quads that need it are dropped) and a quad whose triangles would pass max_tris is
// not written.
//
// Compiled with -ffp-contract=off; division and sqrt are IEEE, as in the oracle.
#include <climits>

#include "common.hpp"
#include "kernels.hpp"

namespace pxl {

namespace {

constexpr uint32_t SBLK = 256;

__device__ __forceinline__ uint32_t node_id(const DevSurf &S, int i, int j, int k) {
    return (uint32_t)i + (uint32_t)(S.sx + 1) * ((uint32_t)j + (uint32_t)(S.sy + 1) * (uint32_t)k);
}
__device__ __forceinline__ uint32_t cell_id(const DevSurf &S, int i, int j, int k) {
    return (uint32_t)i + (uint32_t)S.sx * ((uint32_t)j + (uint32_t)S.sy * (uint32_t)k);
}
__device__ __forceinline__ void node_ijk(const DevSurf &S, uint32_t n, int &i, int &j, int &k) {
    const uint32_t ax = (uint32_t)S.sx + 1u, ay = (uint32_t)S.sy + 1u;
    i = (int)(n % ax);
    j = (int)((n / ax) % ay);
    k = (int)(n / (ax * ay));
}
__device__ __forceinline__ void cell_ijk(const DevSurf &S, uint32_t c, int &i, int &j, int &k) {
    i = (int)(c % (uint32_t)S.sx);
    j = (int)((c / (uint32_t)S.sx) % (uint32_t)S.sy);
    k = (int)(c / ((uint32_t)S.sx * (uint32_t)S.sy));
}

// s_splat (fluid.slang:235-250): g = (x - origin) / sh, r = radius / sh; every node of
// [ceil(g - r), floor(g + r)] (clipped to the grid) with q = 1 - |node - g|^2 / r^2 > 0 gets
// int(q^3 4096).  Integer atomics: the sum does not depend on order.
// Implementation: a workgroup takes SPLAT_CHUNK consecutive particles (the particle
// arrays are in the last step's cell order, so a chunk covers a compact piece of the liquid),
// adds their contributions into an LDS copy of the node box that bounds them, then adds the
// box's nonzero entries to the grid -- one global atomic per node and chunk instead of one per
// particle and node.  A chunk whose box exceeds SPLAT_LDS nodes takes the global path.
constexpr uint32_t SPLAT_CHUNK = 2048;
constexpr uint32_t SPLAT_PER_LANE = SPLAT_CHUNK / SBLK;
constexpr int SPLAT_LDS = 16384;

struct SplatBox {
    float g[3];
    int lo[3], hi[3];
};
__device__ __forceinline__ SplatBox splat_box(const DevSurf &S, float r, float x, float y, float z) {
    SplatBox b;
    const float xs[3] = {x, y, z}, o[3] = {S.ox, S.oy, S.oz};
    const int smax[3] = {S.sx, S.sy, S.sz};
#pragma unroll
    for (int a = 0; a < 3; a++) {
        b.g[a] = (xs[a] - o[a]) / S.sh;
        b.lo[a] = max((int)ceilf(b.g[a] - r), 0);
        b.hi[a] = min((int)floorf(b.g[a] + r), smax[a]);
    }
    return b;
}

__global__ __launch_bounds__(SBLK) void k_s_splat_lds(DevSurf S, const float *px, const float *py, const float *pz,
                                                      uint32_t np) {
    __shared__ int32_t box[SPLAT_LDS];
    __shared__ int red[6][SBLK / WAVE];
    // Workgroups are dispatched round-robin over the 8 XCDs (block b on XCD b % 8); the grid is
    // a multiple of 8 and XCD x takes the x-th eighth of the chunks, so the chunks whose boxes
    // overlap (neighbours in particle order) add into the same L2.
    const uint32_t per_xcd = gridDim.x / 8u;
    const uint32_t chunk = (blockIdx.x % 8u) * per_xcd + blockIdx.x / 8u;
    const uint32_t base = chunk * SPLAT_CHUNK;
    if (base >= np) return;  // the whole workgroup: padding of the grid
    const float r = S.radius / S.sh;
    const float rr = r * r;
    // the chunk's node box (particles whose own box is empty do not count)
    int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
    for (uint32_t m = 0; m < SPLAT_PER_LANE; m++) {
        const uint32_t p = base + m * SBLK + threadIdx.x;
        if (p >= np) break;
        const SplatBox b = splat_box(S, r, px[p], py[p], pz[p]);
        if (b.lo[0] > b.hi[0] || b.lo[1] > b.hi[1] || b.lo[2] > b.hi[2]) continue;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            lo[a] = min(lo[a], b.lo[a]);
            hi[a] = max(hi[a], b.hi[a]);
        }
    }
    const int lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        int l = lo[a], h = hi[a];
        for (int off = WAVE / 2; off > 0; off >>= 1) {
            l = min(l, __shfl_xor(l, off, WAVE));
            h = max(h, __shfl_xor(h, off, WAVE));
        }
        if (lane == 0) {
            red[a][wv] = l;
            red[3 + a][wv] = h;
        }
    }
    __syncthreads();
    int blo[3], ext[3];
    bool empty = false;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        int l = red[a][0], h = red[3 + a][0];
        for (int w = 1; w < (int)(SBLK / WAVE); w++) {
            l = min(l, red[a][w]);
            h = max(h, red[3 + a][w]);
        }
        blo[a] = l;
        ext[a] = h - l + 1;
        if (h < l) empty = true;
    }
    if (empty) return;  // uniform: every lane read the same reduction
    const long vol = (long)ext[0] * ext[1] * ext[2];
    const bool lds = vol <= SPLAT_LDS;
    if (lds) {
        for (int q = threadIdx.x; q < (int)vol; q += SBLK) box[q] = 0;
        __syncthreads();
    }
    for (uint32_t m = 0; m < SPLAT_PER_LANE; m++) {
        const uint32_t p = base + m * SBLK + threadIdx.x;
        if (p >= np) break;
        const SplatBox b = splat_box(S, r, px[p], py[p], pz[p]);
        for (int k = b.lo[2]; k <= b.hi[2]; k++)
            for (int j = b.lo[1]; j <= b.hi[1]; j++)
                for (int i = b.lo[0]; i <= b.hi[0]; i++) {
                    const float d0 = (float)i - b.g[0], d1 = (float)j - b.g[1], d2 = (float)k - b.g[2];
                    const float q = 1.0f - (d0 * d0 + d1 * d1 + d2 * d2) / rr;
                    if (!(q > 0.0f)) continue;
                    const int32_t v = (int32_t)(q * q * q * SURF_SD);
                    if (lds)
                        atomicAdd(&box[(i - blo[0]) + ext[0] * ((j - blo[1]) + ext[1] * (k - blo[2]))], v);
                    else
                        atomicAdd(&S.density[node_id(S, i, j, k)], v);
                }
    }
    if (!lds) return;
    __syncthreads();
    for (int q = threadIdx.x; q < (int)vol; q += SBLK) {
        const int32_t v = box[q];
        if (v == 0) continue;
        const int i = blo[0] + q % ext[0], j = blo[1] + (q / ext[0]) % ext[1], k = blo[2] + q / (ext[0] * ext[1]);
        atomicAdd(&S.density[node_id(S, i, j, k)], v);
    }
}

// s_phi (fluid.slang:253-268).
__global__ __launch_bounds__(SBLK) void k_s_phi(DevSurf S) {
    const uint32_t n = blockIdx.x * SBLK + threadIdx.x;
    if (n >= S.nnodes) return;
    int i, j, k;
    node_ijk(S, n, i, j, k);
    const bool edge = i == 0 || j == 0 || k == 0 || i == S.sx || j == S.sy || k == S.sz;
    if (edge) {  // every neighbour test below would only feed a value that is then discarded
        S.phi[n] = -1.0f;
        return;
    }
    // interior node: all six neighbours exist, summed -x, +x, -y, +y, -z, +z
    const uint32_t ax = (uint32_t)S.sx + 1u, axy = ax * ((uint32_t)S.sy + 1u);
    float s = 2.0f * (float)S.density[n];
    s = s + (float)S.density[n - 1];
    s = s + (float)S.density[n + 1];
    s = s + (float)S.density[n - ax];
    s = s + (float)S.density[n + ax];
    s = s + (float)S.density[n - axy];
    s = s + (float)S.density[n + axy];
    S.phi[n] = s / 8.0f / SURF_SD - S.iso;
}

// Corner phis of surface cell (i, j, k) (bit m = corner (m & 1, m >> 1 & 1, m >> 2)) and the
// inside mask.
__device__ __forceinline__ uint32_t cell_corners(const DevSurf &S, int i, int j, int k, float v[8]) {
    const uint32_t n0 = node_id(S, i, j, k);
    const uint32_t ax = (uint32_t)S.sx + 1u, axy = ax * ((uint32_t)S.sy + 1u);
    uint32_t inside = 0;
#pragma unroll
    for (int m = 0; m < 8; m++) {
        v[m] = S.phi[n0 + (uint32_t)(m & 1) + ((m >> 1) & 1) * ax + (m >> 2) * axy];
        if (v[m] > 0.0f) inside |= 1u << m;
    }
    return inside;
}

__global__ __launch_bounds__(SBLK) void k_s_vflag(DevSurf S) {
    const uint32_t c = blockIdx.x * SBLK + threadIdx.x;
    if (c >= S.ncells) return;
    int i, j, k;
    cell_ijk(S, c, i, j, k);
    float v[8];
    const uint32_t inside = cell_corners(S, i, j, k, v);
    S.cell_vert[c] = (inside != 0u && inside != 255u) ? 1u : 0u;
}

// s_verts (fluid.slang:271-317) for the cells: cell_vert holds the exclusive scan of the
// surface flags on entry (the cell's vertex number) and the stored vertex or SURF_NONE on exit.
__global__ __launch_bounds__(SBLK) void k_s_verts(DevSurf S, float *verts, uint32_t vbase) {
    const uint32_t c = blockIdx.x * SBLK + threadIdx.x;
    if (c >= S.ncells) return;
    int ci, cj, ck;
    cell_ijk(S, c, ci, cj, ck);
    float v[8];
    const uint32_t inside = cell_corners(S, ci, cj, ck, v);
    const bool surf = inside != 0u && inside != 255u;
    const uint32_t idx = S.cell_vert[c];
    if (c == S.ncells - 1u) S.counters[0] = idx + (surf ? 1u : 0u);
    if (!surf || idx >= S.max_verts) {
        S.cell_vert[c] = SURF_NONE;
        return;
    }
    float sum[3] = {0.0f, 0.0f, 0.0f}, cnt = 0.0f;
#pragma unroll
    for (int e = 0; e < 12; e++) {
        const int axis = e >> 2, bits = e & 3, o1 = (axis + 1) % 3, o2 = (axis + 2) % 3;
        const int a0 = ((bits & 1) ? (1 << o1) : 0) | ((bits & 2) ? (1 << o2) : 0);
        const int a1 = a0 | (1 << axis);
        if ((v[a0] > 0.0f) == (v[a1] > 0.0f)) continue;
        const float t = v[a0] / (v[a0] - v[a1]);
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const float p0 = (float)((a0 >> a) & 1), p1 = (float)((a1 >> a) & 1);
            sum[a] = sum[a] + (p0 * (1.0f - t) + p1 * t);  // lerp = FMix
        }
        cnt = cnt + 1.0f;
    }
    float local[3];
#pragma unroll
    for (int a = 0; a < 3; a++) local[a] = sum[a] / cnt;
    float grad[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int m = 0; m < 8; m++) {
        float w[3], sg[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const int o = (m >> a) & 1;
            w[a] = o ? local[a] : 1.0f - local[a];
            sg[a] = o ? 1.0f : -1.0f;
        }
        const float gv[3] = {sg[0] * w[1] * w[2], sg[1] * w[0] * w[2], sg[2] * w[0] * w[1]};
#pragma unroll
        for (int a = 0; a < 3; a++) grad[a] = grad[a] + v[m] * gv[a];
    }
    const float gg[3] = {grad[0], grad[1] + 1e-9f, grad[2]};
    const float len = sqrtf(gg[0] * gg[0] + gg[1] * gg[1] + gg[2] * gg[2]);
    const int cc[3] = {ci, cj, ck};
    const float o[3] = {S.ox, S.oy, S.oz};
    float4 pos, nrm;
    pos.x = o[0] + ((float)cc[0] + local[0]) * S.sh;
    pos.y = o[1] + ((float)cc[1] + local[1]) * S.sh;
    pos.z = o[2] + ((float)cc[2] + local[2]) * S.sh;
    pos.w = 0.0f;
    nrm.x = -(gg[0] / len);
    nrm.y = -(gg[1] / len);
    nrm.z = -(gg[2] / len);
    nrm.w = 0.0f;
    float4 *out = reinterpret_cast<float4 *>(verts) + 2ull * ((uint64_t)vbase + idx);
    out[0] = pos;
    out[1] = nrm;
    S.cell_vert[c] = idx;
}

// Quads at node n's three edges (n, n + e_axis): sign change and all four cells around the
// edge holding a vertex.  Writes the cells' vertices into q[axis][0..3] when asked.
__device__ __forceinline__ uint32_t node_quads(const DevSurf &S, uint32_t n, bool &in0, uint32_t q[3][4]) {
    int i, j, k;
    node_ijk(S, n, i, j, k);
    in0 = S.phi[n] > 0.0f;
    const int p[3] = {i, j, k}, smax[3] = {S.sx, S.sy, S.sz};
    const uint32_t step[3] = {1u, (uint32_t)S.sx + 1u, ((uint32_t)S.sx + 1u) * ((uint32_t)S.sy + 1u)};
    const int offs[4][2] = {{-1, -1}, {0, -1}, {0, 0}, {-1, 0}};
    uint32_t mask = 0;
#pragma unroll
    for (int axis = 0; axis < 3; axis++) {
        if (p[axis] + 1 > smax[axis]) continue;
        if (in0 == (S.phi[n + step[axis]] > 0.0f)) continue;
        const int o1 = (axis + 1) % 3, o2 = (axis + 2) % 3;
        bool ok = true;
#pragma unroll
        for (int e = 0; e < 4; e++) {
            int c[3] = {p[0], p[1], p[2]};
            c[o1] += offs[e][0];
            c[o2] += offs[e][1];
            if (c[0] < 0 || c[1] < 0 || c[2] < 0 || c[0] >= smax[0] || c[1] >= smax[1] || c[2] >= smax[2]) {
                ok = false;
                break;
            }
            q[axis][e] = S.cell_vert[cell_id(S, c[0], c[1], c[2])];
            if (q[axis][e] == SURF_NONE) {
                ok = false;
                break;
            }
        }
        if (ok) mask |= 1u << axis;
    }
    return mask;
}

__global__ __launch_bounds__(SBLK) void k_s_qcount(DevSurf S) {
    const uint32_t n = blockIdx.x * SBLK + threadIdx.x;
    if (n >= S.nnodes) return;
    bool in0;
    uint32_t q[3][4];
    S.qoff[n] = (uint32_t)__popc(node_quads(S, n, in0, q));
}

// s_quads (fluid.slang:319-355): qoff holds the exclusive scan of the per-node quad counts.
__global__ __launch_bounds__(SBLK) void k_s_quads(DevSurf S, uint32_t *indices, uint32_t ibase) {
    const uint32_t n = blockIdx.x * SBLK + threadIdx.x;
    if (n >= S.nnodes) return;
    bool in0;
    uint32_t q[3][4];
    const uint32_t mask = node_quads(S, n, in0, q);
    uint32_t quad = S.qoff[n];
    if (n == S.nnodes - 1u) S.counters[1] = 2u * (quad + (uint32_t)__popc(mask));
#pragma unroll
    for (int axis = 0; axis < 3; axis++) {
        if (!((mask >> axis) & 1u)) continue;
        const uint32_t t = 2u * quad++;
        if ((uint64_t)t + 2u > (uint64_t)S.max_tris) continue;
        uint32_t *b = indices + (uint64_t)ibase + 3ull * t;
        const uint32_t *c = q[axis];
        if (in0) {
            b[0] = c[0]; b[1] = c[1]; b[2] = c[2]; b[3] = c[0]; b[4] = c[2]; b[5] = c[3];
        } else {
            b[0] = c[0]; b[1] = c[2]; b[2] = c[1]; b[3] = c[0]; b[4] = c[3]; b[5] = c[2];
        }
    }
}

inline uint32_t grid_for(uint64_t n) { return (uint32_t)((n + SBLK - 1) / SBLK); }

}  // namespace

const char *const SURF_STAGE_NAMES[SURF_STAGES] = {"s_clear", "s_splat", "s_phi",   "s_vflag", "s_vscan",
                                                   "s_verts", "s_qcount", "s_qscan", "s_quads"};

void launch_surface_stage(int stage, const DevSurf &S, const float *px, const float *py, const float *pz, uint32_t np,
                          float *verts, uint32_t vbase, uint32_t *indices, uint32_t ibase, hipStream_t s) {
    switch (stage) {
        case SURF_CLEAR: launch_fill_u32(reinterpret_cast<uint32_t *>(S.density), S.nnodes, 0u, s); break;
        case SURF_SPLAT:
            if (np) PXL_LAUNCH(k_s_splat_lds, (np + 8u * SPLAT_CHUNK - 1) / (8u * SPLAT_CHUNK) * 8u, SBLK, 0, s, S, px, py, pz,
                               np);
            break;
        case SURF_PHI: PXL_LAUNCH(k_s_phi, grid_for(S.nnodes), SBLK, 0, s, S); break;
        case SURF_VFLAG: PXL_LAUNCH(k_s_vflag, grid_for(S.ncells), SBLK, 0, s, S); break;
        case SURF_VSCAN: launch_exclusive_scan(S.cell_vert, S.ncells, S.scan_tmp, S.cell_vert, s); break;
        case SURF_VERTS: PXL_LAUNCH(k_s_verts, grid_for(S.ncells), SBLK, 0, s, S, verts, vbase); break;
        case SURF_QCOUNT: PXL_LAUNCH(k_s_qcount, grid_for(S.nnodes), SBLK, 0, s, S); break;
        case SURF_QSCAN: launch_exclusive_scan(S.qoff, S.nnodes, S.scan_tmp, S.qoff, s); break;
        case SURF_QUADS: PXL_LAUNCH(k_s_quads, grid_for(S.nnodes), SBLK, 0, s, S, indices, ibase); break;
        default: break;
    }
}

}  // namespace pxl
