/*
 * flip_oracle.c -- TEST INFRASTRUCTURE ONLY (see flip_oracle.h).
 *
 * CPU restatement of pxlengine's FLIP step, written from the semantics of
 * shaders/fluid.slang:9-220 and crates/pxl-render/src/fluid.rs:96-306, and of the scene
 * block derivations in crates/pxl-scene/src/lib.rs:121-129.  Each function cites the
 * reference lines it follows.  Build: oracle/Makefile (gcc -O2 -ffp-contract=off).
 */
#include "flip_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#if defined(__FAST_MATH__)
#error "the oracle must be compiled without fast-math"
#endif

/* ------------------------------------------------------------------ scene block / seeding */

/* Fluid::grid (crates/pxl-scene/src/lib.rs:123-128): glam Vec3 in f32,
 * size = max - min; h = size.max_element() / resolution; n = round(size / h).max(4). */
void orc_fluid_grid(const float lo[3], const float hi[3], uint32_t resolution, uint32_t n_out[3],
                    float *h_out) {
    float size[3];
    for (int a = 0; a < 3; a++) size[a] = hi[a] - lo[a];
    float m = size[0];
    if (size[1] > m) m = size[1];
    if (size[2] > m) m = size[2];
    float h = m / (float)resolution;
    for (int a = 0; a < 3; a++) {
        float n = roundf(size[a] / h); /* f32::round: half away from zero */
        if (n < 4.0f) n = 4.0f;
        n_out[a] = (uint32_t)n;
    }
    *h_out = h;
}

/* lcg (fluid.rs:97-100): 64-bit wrapping LCG, top 24 bits of the state as a float in [0,1). */
float orc_lcg(uint64_t *s) {
    *s = *s * 6364136223846793005ull + 1442695040888963407ull;
    return (float)(*s >> 40) / (float)(1ull << 24);
}

/* Sphere test of the extensions (sphere fills, obstacles): point c lies in the sphere
 * {centre s[0..2], radius s[3]} when (c-s).x^2 + (c-s).y^2 + (c-s).z^2 <= r^2, in f32, summed
 * in that order. */
static int in_sphere(const float c[3], const float *s) {
    const float dx = c[0] - s[0], dy = c[1] - s[1], dz = c[2] - s[2];
    const float d2 = dx * dx + dy * dy + dz * dz;
    return d2 <= s[3] * s[3];
}

/* Obstacle cells (extension): id_out[cell] = 1 + the index of the first obstacle, in list
 * order, whose sphere holds the cell's centre origin + (idx + 0.5) * h, for interior cells
 * (1..n-2 on every axis); 0 elsewhere.  Brute force over every obstacle. */
uint64_t orc_obstacle_cells(const uint32_t n[3], float h, const float origin[3], const float *obst,
                            uint32_t n_obst, uint32_t *id_out) {
    uint64_t count = 0;
    memset(id_out, 0, (size_t)n[0] * n[1] * n[2] * 4);
    for (uint32_t k = 1; k + 1 < n[2]; k++)
        for (uint32_t j = 1; j + 1 < n[1]; j++)
            for (uint32_t i = 1; i + 1 < n[0]; i++) {
                float c[3] = {origin[0] + ((float)i + 0.5f) * h, origin[1] + ((float)j + 0.5f) * h,
                              origin[2] + ((float)k + 0.5f) * h};
                for (uint32_t o = 0; o < n_obst; o++)
                    if (in_sphere(c, obst + 4 * o)) {
                        id_out[i + n[0] * (j + n[1] * k)] = o + 1;
                        count++;
                        break;
                    }
            }
    return count;
}

/* Seeding loop of FluidState::new (fluid.rs:123-148): for each fill block in order, cells
 * k in 1..nz-1, j in 1..ny-1, i in 1..nx-1 whose centre lies inside the block (inclusive)
 * receive ppc particles at origin + (idx + lcg) * h, x then y then z draws.  With obstacles
 * (obst_id non-NULL, an extension) obstacle cells receive none and draw nothing. */
uint64_t orc_seed(const uint32_t n[3], float h, const float origin[3], const float *blocks,
                  uint32_t nb, uint32_t ppc, uint32_t seed, const uint32_t *obst_id, float *x, float *v) {
    uint64_t rng = 0x9e3779b97f4a7c15ull ^ (uint64_t)seed;
    uint64_t count = 0;
    for (uint32_t bi = 0; bi < nb; bi++) {
        const float *b = blocks + ORC_BLOCK_FLOATS * bi;
        for (uint32_t k = 1; k + 1 < n[2]; k++)
            for (uint32_t j = 1; j + 1 < n[1]; j++)
                for (uint32_t i = 1; i + 1 < n[0]; i++) {
                    float c[3] = {origin[0] + ((float)i + 0.5f) * h, origin[1] + ((float)j + 0.5f) * h,
                                  origin[2] + ((float)k + 0.5f) * h};
                    int out = 0;
                    if (b[0] == 0.0f) {  /* box: centre within [min, max] on every axis */
                        for (int a = 0; a < 3; a++)
                            if (c[a] < b[1 + a] || c[a] > b[4 + a]) out = 1;
                    } else {  /* sphere (extension) */
                        out = !in_sphere(c, b + 1);
                    }
                    if (out || (obst_id && obst_id[i + n[0] * (j + n[1] * k)])) continue;
                    for (uint32_t q = 0; q < ppc; q++) {
                        float rx = orc_lcg(&rng);
                        float px = origin[0] + ((float)i + rx) * h;
                        float ry = orc_lcg(&rng);
                        float py = origin[1] + ((float)j + ry) * h;
                        float rz = orc_lcg(&rng);
                        float pz = origin[2] + ((float)k + rz) * h;
                        if (x) {
                            x[3 * count + 0] = px;
                            x[3 * count + 1] = py;
                            x[3 * count + 2] = pz;
                            v[3 * count + 0] = b[7];
                            v[3 * count + 1] = b[8];
                            v[3 * count + 2] = b[9];
                        }
                        count++;
                    }
                }
    }
    return count;
}

/* steps_for_frame (fluid.rs:234-236), all in f64. */
uint64_t orc_steps_for_frame(double warmup, double frame, double fps, uint32_t substeps) {
    double t = round((warmup + frame / fps) * fps * (double)substeps);
    if (t < 0.0) t = 0.0;
    return (uint64_t)t;
}

/* dt push constant (fluid.rs:164): (1.0 / (fps * substeps as f64)) as f32. */
float orc_dt(double fps, uint32_t substeps) { return (float)(1.0 / (fps * (double)substeps)); }

/* ------------------------------------------------------------------ state */

orc_state *orc_create(const orc_params *p) {
    orc_state *s = (orc_state *)calloc(1, sizeof(orc_state));
    if (!s) return NULL;
    s->p = *p;
    uint64_t nx = p->nx, ny = p->ny, nz = p->nz, np = p->n_particles;
    s->n_faces = (nx + 1) * ny * nz + nx * (ny + 1) * nz + nx * ny * (nz + 1);
    s->n_cells = nx * ny * nz;
    size_t P = np ? np : 1, F = s->n_faces, C = s->n_cells;
    s->px = calloc(P, 4); s->py = calloc(P, 4); s->pz = calloc(P, 4);
    s->vx = calloc(P, 4); s->vy = calloc(P, 4); s->vz = calloc(P, 4);
    s->vel = calloc(F, 4); s->vel_old = calloc(F, 4);
    s->acc = calloc(F, 4); s->wacc = calloc(F, 4);
    s->count = calloc(C, 4); s->type = calloc(C, 4);
    s->pressure = calloc(C, 4); /* zeroed: the reference leaves it undefined (fluid.rs:197) */
    s->div = calloc(C, 4);
    return s;
}

void orc_destroy(orc_state *s) {
    if (!s) return;
    free(s->px); free(s->py); free(s->pz); free(s->vx); free(s->vy); free(s->vz);
    free(s->vel); free(s->vel_old); free(s->acc); free(s->wacc);
    free(s->count); free(s->type); free(s->pressure); free(s->div);
    free(s->obst_id); free(s->obst);
    free(s);
}

int orc_set_obstacles(orc_state *s, const float *obst, uint32_t n_obst) {
    free(s->obst_id); free(s->obst);
    s->obst_id = NULL; s->obst = NULL; s->n_obst = 0;
    if (n_obst == 0) return 0;
    s->obst_id = (uint32_t *)calloc(s->n_cells, 4);
    s->obst = (float *)malloc((size_t)n_obst * 16);
    if (!s->obst_id || !s->obst) return -1;
    memcpy(s->obst, obst, (size_t)n_obst * 16);
    s->n_obst = n_obst;
    const uint32_t n[3] = {s->p.nx, s->p.ny, s->p.nz};
    orc_obstacle_cells(n, s->p.h, s->p.origin, obst, n_obst, s->obst_id);
    return 0;
}

/* ------------------------------------------------------------------ MAC indexing (fluid.slang:43-65) */

static inline uint32_t n_u(const orc_params *p) { return (p->nx + 1) * p->ny * p->nz; }
static inline uint32_t n_v(const orc_params *p) { return p->nx * (p->ny + 1) * p->nz; }

static inline void face_dims(const orc_params *p, int c, uint32_t d[3]) {
    d[0] = p->nx; d[1] = p->ny; d[2] = p->nz;
    d[c] += 1;
}
static inline uint32_t face_base(const orc_params *p, int c) {
    return c == 0 ? 0 : (c == 1 ? n_u(p) : n_u(p) + n_v(p));
}
static inline uint32_t face_index(const orc_params *p, int c, const int f[3]) {
    uint32_t d[3];
    face_dims(p, c, d);
    return face_base(p, c) + (uint32_t)f[0] + d[0] * ((uint32_t)f[1] + d[1] * (uint32_t)f[2]);
}
static inline uint32_t cell_index(const orc_params *p, const int c[3]) {
    return (uint32_t)c[0] + p->nx * ((uint32_t)c[1] + p->ny * (uint32_t)c[2]);
}
static inline uint32_t cell_t(const orc_state *s, const int c[3]) {
    const orc_params *p = &s->p;
    if (c[0] < 0 || c[1] < 0 || c[2] < 0 || c[0] >= (int)p->nx || c[1] >= (int)p->ny ||
        c[2] >= (int)p->nz)
        return ORC_SOLID;
    return s->type[cell_index(p, c)];
}
/* face_coord (fluid.slang:61-65): g = (x - origin) * inv_h - 0.5, then g[c] += 0.5. */
static inline void face_coord(const orc_params *p, int c, const float x[3], float g[3]) {
    for (int a = 0; a < 3; a++) g[a] = (x[a] - p->origin[a]) * p->inv_h - 0.5f;
    g[c] = g[c] + 0.5f;
}
/* face_of (fluid.slang:135-140). */
static inline void face_of(const orc_params *p, uint32_t id, int *c, int f[3]) {
    *c = id < n_u(p) ? 0 : (id < n_u(p) + n_v(p) ? 1 : 2);
    uint32_t d[3];
    face_dims(p, *c, d);
    uint32_t r = id - face_base(p, *c);
    f[0] = (int)(r % d[0]);
    f[1] = (int)((r / d[0]) % d[1]);
    f[2] = (int)(r / (d[0] * d[1]));
}
/* SPIR-V FMax / FMin (Slang max / min). */
static inline float fmax_spv(float x, float y) { return x < y ? y : x; }
static inline float fmin_spv(float x, float y) { return y < x ? y : x; }

/* sample_face (fluid.slang:67-80): trilinear gather, indices clamped to [0, dims-1]. */
static float sample_face(const orc_state *s, const float *buf, int c, const float x[3]) {
    const orc_params *p = &s->p;
    float g[3], f[3];
    int b[3], d[3];
    uint32_t fd[3];
    face_coord(p, c, x, g);
    face_dims(p, c, fd);
    for (int a = 0; a < 3; a++) {
        d[a] = (int)fd[a] - 1;
        b[a] = (int)floorf(g[a]);
        f[a] = g[a] - (float)b[a];
    }
    float acc = 0.0f;
    for (int k = 0; k < 8; k++) {
        int o[3] = {k & 1, (k >> 1) & 1, k >> 2};
        int q[3];
        float w3[3];
        for (int a = 0; a < 3; a++) {
            int t = b[a] + o[a];
            q[a] = t < 0 ? 0 : (t > d[a] ? d[a] : t);
            w3[a] = o[a] ? f[a] : 1.0f - f[a]; /* lerp(1 - f, f, o), o constant 0/1 */
        }
        acc = acc + w3[0] * w3[1] * w3[2] * buf[face_index(p, c, q)];
    }
    return acc;
}

/* ------------------------------------------------------------------ step kernels */

/* clear_grid (fluid.slang:88-91). */
void orc_clear_grid(orc_state *s) {
    memset(s->acc, 0, s->n_faces * 4);
    memset(s->wacc, 0, s->n_faces * 4);
    memset(s->count, 0, s->n_cells * 4);
}

/* p2g (fluid.slang:94-115): fixed-point (x65536, truncating) trilinear splat of v_c onto the
 * 8 faces of component c around the particle, out-of-range corners skipped; +1 per cell. */
void orc_p2g(orc_state *s) {
    const orc_params *p = &s->p;
    for (uint32_t i = 0; i < p->n_particles; i++) {
        float x[3] = {s->px[i], s->py[i], s->pz[i]};
        float v[3] = {s->vx[i], s->vy[i], s->vz[i]};
        for (int c = 0; c < 3; c++) {
            float g[3], f[3];
            int b[3], d[3];
            uint32_t fd[3];
            face_coord(p, c, x, g);
            face_dims(p, c, fd);
            for (int a = 0; a < 3; a++) {
                d[a] = (int)fd[a] - 1;
                b[a] = (int)floorf(g[a]);
                f[a] = g[a] - (float)b[a];
            }
            for (int k = 0; k < 8; k++) {
                int o[3] = {k & 1, (k >> 1) & 1, k >> 2};
                int q[3];
                int skip = 0;
                float w3[3];
                for (int a = 0; a < 3; a++) {
                    q[a] = b[a] + o[a];
                    if (q[a] < 0 || q[a] > d[a]) skip = 1;
                    w3[a] = o[a] ? f[a] : 1.0f - f[a];
                }
                if (skip) continue;
                float w = w3[0] * w3[1] * w3[2];
                uint32_t fi = face_index(p, c, q);
                s->acc[fi] += (int32_t)(w * v[c] * 65536.0f);
                s->wacc[fi] += (int32_t)(w * 65536.0f);
            }
        }
        int cell[3];
        int n[3] = {(int)p->nx, (int)p->ny, (int)p->nz};
        for (int a = 0; a < 3; a++) {
            int t = (int)floorf((x[a] - p->origin[a]) * p->inv_h);
            cell[a] = t < 0 ? 0 : (t > n[a] - 1 ? n[a] - 1 : t);
        }
        s->count[cell_index(p, cell)] += 1u;
    }
}

/* normalize_grid (fluid.slang:118-132): face velocity = acc / wacc (0 without weight),
 * copied to vel_old; cell type from the wall layer and the particle count; pressure zeroed
 * outside the fluid. */
void orc_normalize_grid(orc_state *s) {
    const orc_params *p = &s->p;
    for (uint64_t f = 0; f < s->n_faces; f++) {
        int32_t w = s->wacc[f];
        float v = w > 0 ? (float)s->acc[f] / (float)w : 0.0f;
        s->vel[f] = v;
        s->vel_old[f] = v;
    }
    for (uint32_t id = 0; id < (uint32_t)s->n_cells; id++) {
        uint32_t i = id % p->nx, j = (id / p->nx) % p->ny, k = id / (p->nx * p->ny);
        int wall = i == 0 || j == 0 || k == 0 || i == p->nx - 1 || j == p->ny - 1 || k == p->nz - 1;
        if (s->obst_id && s->obst_id[id]) wall = 1;  /* obstacle cell (extension) */
        uint32_t t = wall ? ORC_SOLID : (s->count[id] > 0 ? ORC_FLUID : ORC_AIR);
        s->type[id] = t;
        if (t != ORC_FLUID) s->pressure[id] = 0.0f;
    }
}

/* forces (fluid.slang:143-151): v += g_c * dt, then 0 on faces touching a solid cell. */
void orc_forces(orc_state *s) {
    const orc_params *p = &s->p;
    for (uint32_t id = 0; id < (uint32_t)s->n_faces; id++) {
        int c, f[3], lo[3];
        face_of(p, id, &c, f);
        memcpy(lo, f, sizeof lo);
        lo[c] -= 1;
        float v = s->vel[id] + p->gravity[c] * p->dt;
        if (cell_t(s, lo) == ORC_SOLID || cell_t(s, f) == ORC_SOLID) v = 0.0f;
        s->vel[id] = v;
    }
}

/* compute_divergence (fluid.slang:154-165). */
void orc_divergence(orc_state *s) {
    const orc_params *p = &s->p;
    for (uint32_t id = 0; id < (uint32_t)s->n_cells; id++) {
        int c[3] = {(int)(id % p->nx), (int)((id / p->nx) % p->ny), (int)(id / (p->nx * p->ny))};
        float d = 0.0f;
        if (s->type[id] == ORC_FLUID) {
            for (int a = 0; a < 3; a++) {
                int hi[3] = {c[0], c[1], c[2]};
                hi[a] += 1;
                d = d + (s->vel[face_index(p, a, hi)] - s->vel[face_index(p, a, c)]);
            }
        }
        s->div[id] = d * p->inv_h;
    }
}

/* pressure_rb (fluid.slang:168-182): one red-black Gauss-Seidel half sweep. */
void orc_pressure_rb(orc_state *s, uint32_t parity) {
    const orc_params *p = &s->p;
    for (uint32_t id = 0; id < (uint32_t)s->n_cells; id++) {
        if (s->type[id] != ORC_FLUID) continue;
        int c[3] = {(int)(id % p->nx), (int)((id / p->nx) % p->ny), (int)(id / (p->nx * p->ny))};
        if ((uint32_t)(c[0] + c[1] + c[2]) % 2 != parity) continue;
        float sum = 0.0f, cnt = 0.0f;
        for (int n = 0; n < 6; n++) {
            int q[3] = {c[0], c[1], c[2]};
            q[n >> 1] += (n & 1) ? 1 : -1;
            uint32_t t = cell_t(s, q);
            if (t == ORC_SOLID) continue;
            cnt = cnt + 1.0f;
            if (t == ORC_FLUID) sum = sum + s->pressure[cell_index(p, q)];
        }
        if (cnt > 0.0f) s->pressure[id] = (sum - s->div[id] * p->h * p->h / p->dt) / cnt;
    }
}

/* project (fluid.slang:185-197). */
void orc_project(orc_state *s) {
    const orc_params *p = &s->p;
    for (uint32_t id = 0; id < (uint32_t)s->n_faces; id++) {
        int c, f[3], lo[3];
        face_of(p, id, &c, f);
        memcpy(lo, f, sizeof lo);
        lo[c] -= 1;
        uint32_t ta = cell_t(s, lo), tb = cell_t(s, f);
        if (ta == ORC_SOLID || tb == ORC_SOLID) {
            s->vel[id] = 0.0f;
            continue;
        }
        if (ta == ORC_FLUID || tb == ORC_FLUID) {
            float pa = ta == ORC_FLUID ? s->pressure[cell_index(p, lo)] : 0.0f;
            float pb = tb == ORC_FLUID ? s->pressure[cell_index(p, f)] : 0.0f;
            s->vel[id] = s->vel[id] - p->dt * p->inv_h * (pb - pa);
        }
    }
}

/* Obstacle response (extension), after advection and before the wall clamp: a particle whose
 * cell (the binning cell, clamp(floor((x - origin) * inv_h))) is an obstacle cell moves along
 * the ray from the obstacle's centre to distance r + h (so that its new cell's centre is
 * outside that sphere), and the velocity's inward normal component is removed:
 *   d = x - centre; len = sqrt(d.x^2 + d.y^2 + d.z^2); n = d / len (per component);
 *   x = centre + n * (r + h); vn = v.x n.x + v.y n.y + v.z n.z; if vn < 0: v = v - vn * n.
 * Sums left to right, f32, correctly rounded sqrt and division; len == 0 leaves the
 * particle unchanged. */
static void obstacle_push(const orc_state *s, float x[3], float v[3]) {
    const orc_params *p = &s->p;
    int n[3] = {(int)p->nx, (int)p->ny, (int)p->nz}, cell[3];
    for (int a = 0; a < 3; a++) {
        int t = (int)floorf((x[a] - p->origin[a]) * p->inv_h);
        cell[a] = t < 0 ? 0 : (t > n[a] - 1 ? n[a] - 1 : t);
    }
    const uint32_t o = s->obst_id[cell_index(p, cell)];
    if (!o) return;
    const float *sp = s->obst + 4 * (o - 1);
    const float d[3] = {x[0] - sp[0], x[1] - sp[1], x[2] - sp[2]};
    const float len = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (!(len > 0.0f)) return;
    const float nr[3] = {d[0] / len, d[1] / len, d[2] / len};
    const float rad = sp[3] + p->h;
    for (int a = 0; a < 3; a++) x[a] = sp[a] + nr[a] * rad;
    const float vn = v[0] * nr[0] + v[1] * nr[1] + v[2] * nr[2];
    if (vn < 0.0f)
        for (int a = 0; a < 3; a++) v[a] = v[a] - vn * nr[a];
}

/* g2p_advect (fluid.slang:200-220): PIC/FLIP blend, RK2 through the grid velocity, clamp to
 * the interior with the wall-normal velocity clipped. */
void orc_g2p_advect(orc_state *s) {
    const orc_params *p = &s->p;
    float lo[3], hi[3];
    float nf[3] = {(float)p->nx, (float)p->ny, (float)p->nz};
    for (int a = 0; a < 3; a++) {
        lo[a] = p->origin[a] + p->h * 1.001f;
        hi[a] = p->origin[a] + p->h * (nf[a] - 1.001f);
    }
    const float t = p->flip;
    for (uint32_t i = 0; i < p->n_particles; i++) {
        float x[3] = {s->px[i], s->py[i], s->pz[i]};
        float v0[3] = {s->vx[i], s->vy[i], s->vz[i]};
        float pic[3], old[3], v[3], mid[3], adv[3];
        for (int c = 0; c < 3; c++) pic[c] = sample_face(s, s->vel, c, x);
        for (int c = 0; c < 3; c++) old[c] = sample_face(s, s->vel_old, c, x);
        for (int c = 0; c < 3; c++) {
            float flipv = v0[c] + pic[c] - old[c];
            v[c] = pic[c] * (1.0f - t) + flipv * t; /* lerp = FMix */
            mid[c] = x[c] + 0.5f * p->dt * pic[c];
        }
        for (int c = 0; c < 3; c++) adv[c] = sample_face(s, s->vel, c, mid);
        for (int c = 0; c < 3; c++) x[c] = x[c] + p->dt * adv[c];
        if (s->obst_id) obstacle_push(s, x, v);
        for (int a = 0; a < 3; a++) {
            if (x[a] < lo[a]) { x[a] = lo[a]; v[a] = fmax_spv(v[a], 0.0f); }
            if (x[a] > hi[a]) { x[a] = hi[a]; v[a] = fmin_spv(v[a], 0.0f); }
        }
        s->px[i] = x[0]; s->py[i] = x[1]; s->pz[i] = x[2];
        s->vx[i] = v[0]; s->vy[i] = v[1]; s->vz[i] = v[2];
    }
}

/* One reference step (fluid.rs:289-302). */
void orc_step_rbgs(orc_state *s, uint32_t iters) {
    orc_clear_grid(s);
    orc_p2g(s);
    orc_normalize_grid(s);
    orc_forces(s);
    orc_divergence(s);
    for (uint32_t it = 0; it < iters; it++) {
        orc_pressure_rb(s, 0);
        orc_pressure_rb(s, 1);
    }
    orc_project(s);
    orc_g2p_advect(s);
}

/* ------------------------------------------------------------------ FP32 PCG (build extension) */

/* Diagonal of the pressure matrix: non-solid neighbours of a fluid cell (fluid.slang:172-178). */
static float cell_diag(const orc_state *s, const int c[3]) {
    float cnt = 0.0f;
    for (int n = 0; n < 6; n++) {
        int q[3] = {c[0], c[1], c[2]};
        q[n >> 1] += (n & 1) ? 1 : -1;
        if (cell_t(s, q) != ORC_SOLID) cnt = cnt + 1.0f;
    }
    return cnt;
}

/* (A x)_i = diag_i x_i - sum_{fluid nbrs, order -x,+x,-y,+y,-z,+z} x_j, fp32. */
static float apply32(const orc_state *s, const float *x, const int c[3], uint32_t id) {
    const orc_params *p = &s->p;
    float sum = 0.0f;
    for (int n = 0; n < 6; n++) {
        int q[3] = {c[0], c[1], c[2]};
        q[n >> 1] += (n & 1) ? 1 : -1;
        if (cell_t(s, q) == ORC_FLUID) sum = sum + x[cell_index(p, q)];
    }
    return cell_diag(s, c) * x[id] - sum;
}

static inline float precond32(float r, float dg) { return dg > 0.0f ? r / dg : 0.0f; }

/* Jacobi-preconditioned CG on A p = b, b = -((div*h)*h)/dt, warm-started from pressure,
 * exactly `iters` iterations, in the Chronopoulos-Gear arrangement (one pair of dot
 * products per iteration):
 *   r0 = b - A x, u0 = r0 / diag, w0 = A u0, gamma = (r,u), delta = (w,u),
 *   alpha = gamma / delta, beta = 0; then per iteration
 *   p = u + beta p, s = w + beta s, x += alpha p, r -= alpha s, u = r / diag, w = A u,
 *   beta = gamma' / gamma, alpha = gamma' / (delta' - beta gamma' / alpha).
 * Dot products are fp32 sums in cell order.  Guards: a zero gamma, old gamma or alpha leaves
 * alpha = beta = 0 (no further change). */
void orc_pressure_cg32(orc_state *s, uint32_t iters) {
    const orc_params *p = &s->p;
    uint32_t C = (uint32_t)s->n_cells;
    float *r = calloc(C, 4), *w = calloc(C, 4), *sv = calloc(C, 4), *pv = calloc(C, 4), *dg = calloc(C, 4);
    float *u = calloc(C, 4);
    float *x = s->pressure;
    for (uint32_t id = 0; id < C; id++) {
        if (s->type[id] != ORC_FLUID) { x[id] = 0.0f; continue; }
        int c[3] = {(int)(id % p->nx), (int)((id / p->nx) % p->ny), (int)(id / (p->nx * p->ny))};
        dg[id] = cell_diag(s, c);
    }
    for (uint32_t id = 0; id < C; id++) {
        if (s->type[id] != ORC_FLUID) continue;
        int c[3] = {(int)(id % p->nx), (int)((id / p->nx) % p->ny), (int)(id / (p->nx * p->ny))};
        float b = -(s->div[id] * p->h * p->h / p->dt);
        r[id] = b - apply32(s, x, c, id);
    }
    float gamma = 0.0f, delta = 0.0f;
    for (uint32_t id = 0; id < C; id++) u[id] = s->type[id] == ORC_FLUID ? precond32(r[id], dg[id]) : 0.0f;
    for (uint32_t id = 0; id < C; id++) {
        if (s->type[id] != ORC_FLUID) continue;
        int c[3] = {(int)(id % p->nx), (int)((id / p->nx) % p->ny), (int)(id / (p->nx * p->ny))};
        w[id] = apply32(s, u, c, id);
        gamma = gamma + r[id] * u[id];
        delta = delta + w[id] * u[id];
    }
    float alpha = (gamma != 0.0f && delta != 0.0f) ? gamma / delta : 0.0f, beta = 0.0f;
    for (uint32_t it = 0; it < iters; it++) {
        for (uint32_t id = 0; id < C; id++) {
            if (s->type[id] != ORC_FLUID) continue;
            pv[id] = precond32(r[id], dg[id]) + beta * pv[id];
            sv[id] = w[id] + beta * sv[id];
            x[id] = x[id] + alpha * pv[id];
            r[id] = r[id] - alpha * sv[id];
        }
        for (uint32_t id = 0; id < C; id++) u[id] = s->type[id] == ORC_FLUID ? precond32(r[id], dg[id]) : 0.0f;
        float g = 0.0f, d = 0.0f;
        for (uint32_t id = 0; id < C; id++) {
            if (s->type[id] != ORC_FLUID) continue;
            int c[3] = {(int)(id % p->nx), (int)((id / p->nx) % p->ny), (int)(id / (p->nx * p->ny))};
            w[id] = apply32(s, u, c, id);
            g = g + r[id] * u[id];
            d = d + w[id] * u[id];
        }
        float a_new = 0.0f, b_new = 0.0f;
        if (g != 0.0f && gamma != 0.0f && alpha != 0.0f) {
            b_new = g / gamma;
            float den = d - b_new * g / alpha;
            a_new = den != 0.0f ? g / den : 0.0f;
        }
        alpha = a_new;
        beta = b_new;
        gamma = g;
    }
    free(r); free(w); free(sv); free(pv); free(dg); free(u);
}

void orc_step_cg(orc_state *s, uint32_t iters) {
    orc_clear_grid(s);
    orc_p2g(s);
    orc_normalize_grid(s);
    orc_forces(s);
    orc_divergence(s);
    orc_pressure_cg32(s, iters);
    orc_project(s);
    orc_g2p_advect(s);
}

/* ------------------------------------------------------------------ agent neighbour grid */

uint32_t orc_grid_table_size(uint32_t count) {
    uint64_t want = (uint64_t)count * 2, t = 1;
    while (t < want) t <<= 1;
    if (t < 1024) t = 1024;
    if (t > (1u << 23)) t = 1u << 23;
    return (uint32_t)t;
}

uint32_t orc_cell_hash(int32_t cx, int32_t cy, int32_t cz, uint32_t table_size) {
    return (((uint32_t)cx * 73856093u) ^ ((uint32_t)cy * 19349663u) ^ ((uint32_t)cz * 83492791u)) &
           (table_size - 1u);
}

void orc_grid_build(const float *agents, uint32_t count, float cell_size, uint32_t table_size,
                    uint32_t *offsets, uint32_t *counts, uint32_t *sorted) {
    const float cs = cell_size > 1e-6f ? cell_size : 1e-6f; /* f32::max(1e-6) */
    uint32_t *key = (uint32_t *)malloc((size_t)(count ? count : 1) * 4);
    memset(counts, 0, (size_t)table_size * 4);
    for (uint32_t i = 0; i < count; i++) {
        const float *a = agents + 16 * (uint64_t)i;
        const int32_t cx = (int32_t)floorf(a[0] / cs), cy = (int32_t)floorf(a[1] / cs), cz = (int32_t)floorf(a[2] / cs);
        key[i] = orc_cell_hash(cx, cy, cz, table_size);
        counts[key[i]]++;
    }
    uint32_t run = 0;
    for (uint32_t h = 0; h < table_size; h++) {
        offsets[h] = run;
        run += counts[h];
    }
    /* stable scatter: ids in increasing order fill their cell's slots in order */
    uint32_t *cursor = (uint32_t *)malloc((size_t)table_size * 4);
    memcpy(cursor, offsets, (size_t)table_size * 4);
    for (uint32_t i = 0; i < count; i++) sorted[cursor[key[i]]++] = i;
    free(cursor);
    free(key);
}

/* ------------------------------------------------------------------ CFL cap (extension) */

uint32_t orc_cfl_substeps(const orc_state *s, float dt, float cfl) {
    float vmax2 = 0.0f;
    for (uint32_t i = 0; i < s->p.n_particles; i++) {
        const float v2 = s->vx[i] * s->vx[i] + s->vy[i] * s->vy[i] + s->vz[i] * s->vz[i];
        if (v2 > vmax2) vmax2 = v2;
    }
    const double cells = sqrt((double)vmax2) * (double)dt / ((double)cfl * (double)s->p.h);
    if (!(cells > 1.0)) return 1;
    if (cells >= 64.0) return 64;
    return (uint32_t)ceil(cells);
}

/* ------------------------------------------------------------------ surface nets */

void orc_surface_params(const uint32_t n[3], float h, const float origin[3], uint32_t refine,
                        float particle_radius, uint32_t ppc, uint32_t max_triangles, orc_surface *out) {
    out->sx = n[0] * refine;
    out->sy = n[1] * refine;
    out->sz = n[2] * refine;
    out->sh = h / (float)refine;
    out->radius = particle_radius * out->sh;
    const float rh = out->radius / h;
    const float kernel_volume = 0.638f * (rh * rh * rh);
    out->iso = 0.5f * (float)ppc * kernel_volume;
    out->max_verts = max_triangles;
    out->max_tris = max_triangles;
    for (int a = 0; a < 3; a++) out->origin[a] = origin[a];
}

static inline uint64_t snode(const orc_surface *sp, int i, int j, int k) {
    return (uint64_t)i + (uint64_t)(sp->sx + 1) * ((uint64_t)j + (uint64_t)(sp->sy + 1) * (uint64_t)k);
}
static inline uint64_t scell(const orc_surface *sp, int i, int j, int k) {
    return (uint64_t)i + (uint64_t)sp->sx * ((uint64_t)j + (uint64_t)sp->sy * (uint64_t)k);
}

/* s_clear + s_splat: each particle adds int(q^3 * 4096), q = 1 - |node - g|^2 / r^2 > 0, to
 * the nodes of its box [ceil(g - r), floor(g + r)] clipped to the grid, g = (x - origin) / sh,
 * r = radius / sh. */
void orc_surface_splat(const orc_surface *sp, const float *px, const float *py, const float *pz,
                       uint32_t np, int32_t *density) {
    const uint64_t nn = (uint64_t)(sp->sx + 1) * (sp->sy + 1) * (sp->sz + 1);
    memset(density, 0, nn * 4);
    const float r = sp->radius / sp->sh;
    const int smax[3] = {(int)sp->sx, (int)sp->sy, (int)sp->sz};
    for (uint32_t p = 0; p < np; p++) {
        const float x[3] = {px[p], py[p], pz[p]};
        float g[3];
        int lo[3], hi[3];
        for (int a = 0; a < 3; a++) {
            g[a] = (x[a] - sp->origin[a]) / sp->sh;
            lo[a] = (int)ceilf(g[a] - r);
            if (lo[a] < 0) lo[a] = 0;
            hi[a] = (int)floorf(g[a] + r);
            if (hi[a] > smax[a]) hi[a] = smax[a];
        }
        for (int k = lo[2]; k <= hi[2]; k++)
            for (int j = lo[1]; j <= hi[1]; j++)
                for (int i = lo[0]; i <= hi[0]; i++) {
                    const float d0 = (float)i - g[0], d1 = (float)j - g[1], d2 = (float)k - g[2];
                    const float q = 1.0f - (d0 * d0 + d1 * d1 + d2 * d2) / (r * r);
                    if (q > 0.0f) density[snode(sp, i, j, k)] += (int32_t)(q * q * q * 4096.0f);
                }
    }
}

/* s_phi: s = 2 density(n) + the in-grid neighbours' densities (-x, +x, -y, +y, -z, +z), w = 2 +
 * their number; phi = s / w / 4096 - iso, or -1 on the grid's boundary nodes. */
void orc_surface_phi(const orc_surface *sp, const int32_t *density, float *phi) {
    const int smax[3] = {(int)sp->sx, (int)sp->sy, (int)sp->sz};
    for (int k = 0; k <= smax[2]; k++)
        for (int j = 0; j <= smax[1]; j++)
            for (int i = 0; i <= smax[0]; i++) {
                const int n[3] = {i, j, k};
                float s = 2.0f * (float)density[snode(sp, i, j, k)];
                float w = 2.0f;
                for (int a = 0; a < 6; a++) {
                    int q[3] = {i, j, k};
                    q[a >> 1] += (a & 1) ? 1 : -1;
                    if (q[0] < 0 || q[1] < 0 || q[2] < 0 || q[0] > smax[0] || q[1] > smax[1] || q[2] > smax[2])
                        continue;
                    s = s + (float)density[snode(sp, q[0], q[1], q[2])];
                    w = w + 1.0f;
                }
                const int edge = n[0] == 0 || n[1] == 0 || n[2] == 0 || n[0] == smax[0] || n[1] == smax[1] ||
                                 n[2] == smax[2];
                phi[snode(sp, i, j, k)] = edge ? -1.0f : s / w / 4096.0f - sp->iso;
            }
}

/* s_verts + s_quads.  A cell with both signs among its 8 corner phis (bit k = corner
 * (k & 1, k >> 1 & 1, k >> 2)) gets the vertex at the mean of its edge crossings (edges e:
 * axis e >> 2, the other axes' corner bits e & 3; t = v0 / (v0 - v1); FMix of the corners) and
 * the normal -normalize(grad + (0, 1e-9, 0)) of the trilinear phi there (normalize = g / sqrt(g.g)).
 * Each node edge (n, n + e_axis) whose phis differ in sign, with all 4 cells around it holding a
 * vertex ((-1,-1), (0,-1), (0,0), (-1,0) along the next two axes), emits two triangles wound
 * from inside to outside. */
void orc_surface_mesh(const orc_surface *sp, const float *phi, uint32_t *cell_vert, float *verts,
                      uint32_t *indices, uint32_t counters[2]) {
    const int smax[3] = {(int)sp->sx, (int)sp->sy, (int)sp->sz};
    uint32_t nv = 0, nt = 0;
    for (int ck = 0; ck < smax[2]; ck++)
        for (int cj = 0; cj < smax[1]; cj++)
            for (int ci = 0; ci < smax[0]; ci++) {
                const int c[3] = {ci, cj, ck};
                const uint64_t id = scell(sp, ci, cj, ck);
                float v[8];
                uint32_t inside = 0;
                for (int k = 0; k < 8; k++) {
                    v[k] = phi[snode(sp, ci + (k & 1), cj + ((k >> 1) & 1), ck + (k >> 2))];
                    if (v[k] > 0.0f) inside |= 1u << k;
                }
                cell_vert[id] = ORC_NONE;
                if (inside == 0 || inside == 255) continue;
                float sum[3] = {0.0f, 0.0f, 0.0f}, cnt = 0.0f;
                for (int e = 0; e < 12; e++) {
                    const int axis = e >> 2, bits = e & 3, o1 = (axis + 1) % 3, o2 = (axis + 2) % 3;
                    int a0 = 0;
                    if (bits & 1) a0 |= 1 << o1;
                    if (bits & 2) a0 |= 1 << o2;
                    const int a1 = a0 | (1 << axis);
                    if ((v[a0] > 0.0f) == (v[a1] > 0.0f)) continue;
                    const float t = v[a0] / (v[a0] - v[a1]);
                    for (int a = 0; a < 3; a++) {
                        const float p0 = (float)((a0 >> a) & 1), p1 = (float)((a1 >> a) & 1);
                        sum[a] = sum[a] + (p0 * (1.0f - t) + p1 * t);
                    }
                    cnt = cnt + 1.0f;
                }
                float local[3];
                for (int a = 0; a < 3; a++) local[a] = sum[a] / cnt;
                float grad[3] = {0.0f, 0.0f, 0.0f};
                for (int k = 0; k < 8; k++) {
                    float w[3], sg[3];
                    for (int a = 0; a < 3; a++) {
                        const int o = (k >> a) & 1;
                        w[a] = o ? local[a] : 1.0f - local[a];
                        sg[a] = o ? 1.0f : -1.0f;
                    }
                    const float gv[3] = {sg[0] * w[1] * w[2], sg[1] * w[0] * w[2], sg[2] * w[0] * w[1]};
                    for (int a = 0; a < 3; a++) grad[a] = grad[a] + v[k] * gv[a];
                }
                const uint32_t idx = nv++;
                if (idx >= sp->max_verts) continue;
                const float gg[3] = {grad[0], grad[1] + 1e-9f, grad[2]};
                const float len = sqrtf(gg[0] * gg[0] + gg[1] * gg[1] + gg[2] * gg[2]);
                float *out = verts + 8 * (uint64_t)idx;
                for (int a = 0; a < 3; a++) {
                    out[a] = sp->origin[a] + ((float)c[a] + local[a]) * sp->sh;
                    out[4 + a] = -(gg[a] / len);
                }
                out[3] = 0.0f;
                out[7] = 0.0f;
                cell_vert[id] = idx;
            }
    static const int offs[4][2] = {{-1, -1}, {0, -1}, {0, 0}, {-1, 0}};
    for (int k = 0; k <= smax[2]; k++)
        for (int j = 0; j <= smax[1]; j++)
            for (int i = 0; i <= smax[0]; i++) {
                const int in0 = phi[snode(sp, i, j, k)] > 0.0f;
                for (int axis = 0; axis < 3; axis++) {
                    int m[3] = {i, j, k};
                    m[axis] += 1;
                    if (m[0] > smax[0] || m[1] > smax[1] || m[2] > smax[2]) continue;
                    if (in0 == (phi[snode(sp, m[0], m[1], m[2])] > 0.0f)) continue;
                    const int o1 = (axis + 1) % 3, o2 = (axis + 2) % 3;
                    uint32_t q[4];
                    int ok = 1;
                    for (int e = 0; e < 4 && ok; e++) {
                        int c[3] = {i, j, k};
                        c[o1] += offs[e][0];
                        c[o2] += offs[e][1];
                        if (c[0] < 0 || c[1] < 0 || c[2] < 0 || c[0] >= smax[0] || c[1] >= smax[1] || c[2] >= smax[2]) {
                            ok = 0;
                            break;
                        }
                        q[e] = cell_vert[scell(sp, c[0], c[1], c[2])];
                        if (q[e] == ORC_NONE) ok = 0;
                    }
                    if (!ok) continue;
                    const uint32_t t = nt;
                    nt += 2;
                    if (t + 2 > sp->max_tris) continue;
                    uint32_t *b = indices + 3 * (uint64_t)t;
                    if (in0) {
                        b[0] = q[0]; b[1] = q[1]; b[2] = q[2]; b[3] = q[0]; b[4] = q[2]; b[5] = q[3];
                    } else {
                        b[0] = q[0]; b[1] = q[2]; b[2] = q[1]; b[3] = q[0]; b[4] = q[3]; b[5] = q[2];
                    }
                }
            }
    counters[0] = nv;
    counters[1] = nt;
}

/* ------------------------------------------------------------------ binning */

/* Cell of a particle, as in p2g (fluid.slang:113): clamp(floor((x - origin) * inv_h)). */
static uint32_t particle_cell(const orc_state *s, uint32_t i) {
    const orc_params *p = &s->p;
    float x[3] = {s->px[i], s->py[i], s->pz[i]};
    int n[3] = {(int)p->nx, (int)p->ny, (int)p->nz}, cell[3];
    for (int a = 0; a < 3; a++) {
        int t = (int)floorf((x[a] - p->origin[a]) * p->inv_h);
        cell[a] = t < 0 ? 0 : (t > n[a] - 1 ? n[a] - 1 : t);
    }
    return cell_index(p, cell);
}

/* Stable counting sort of particle ids by cell: counts = histogram, offsets = exclusive
 * scan, ids ascending within each cell. */
void orc_bin(const orc_state *s, uint32_t *cell_of, uint32_t *counts, uint32_t *offsets,
             uint32_t *sorted_ids) {
    uint32_t P = s->p.n_particles, C = (uint32_t)s->n_cells;
    memset(counts, 0, (size_t)C * 4);
    for (uint32_t i = 0; i < P; i++) {
        cell_of[i] = particle_cell(s, i);
        counts[cell_of[i]]++;
    }
    uint32_t run = 0;
    for (uint32_t c = 0; c < C; c++) {
        offsets[c] = run;
        run += counts[c];
    }
    uint32_t *fill = calloc(C ? C : 1, 4);
    for (uint32_t i = 0; i < P; i++) {
        uint32_t c = cell_of[i];
        sorted_ids[offsets[c] + fill[c]++] = i;
    }
    free(fill);
}

/* ------------------------------------------------------------------ FP64 Poisson oracle */

void orc_poisson_apply64(const orc_state *s, const double *x, double *y) {
    const orc_params *p = &s->p;
    for (uint32_t id = 0; id < (uint32_t)s->n_cells; id++) {
        if (s->type[id] != ORC_FLUID) { y[id] = 0.0; continue; }
        int c[3] = {(int)(id % p->nx), (int)((id / p->nx) % p->ny), (int)(id / (p->nx * p->ny))};
        double sum = 0.0, cnt = 0.0;
        for (int n = 0; n < 6; n++) {
            int q[3] = {c[0], c[1], c[2]};
            q[n >> 1] += (n & 1) ? 1 : -1;
            uint32_t t = cell_t(s, q);
            if (t == ORC_SOLID) continue;
            cnt += 1.0;
            if (t == ORC_FLUID) sum += x[cell_index(p, q)];
        }
        y[id] = cnt * x[id] - sum;
    }
}

void orc_poisson_rhs64(const orc_state *s, double *b) {
    const orc_params *p = &s->p;
    double h = p->h, dt = p->dt;
    for (uint32_t id = 0; id < (uint32_t)s->n_cells; id++)
        b[id] = s->type[id] == ORC_FLUID ? -((double)s->div[id] * h * h / dt) : 0.0;
}

double orc_poisson_relres64(const orc_state *s, const float *x32, const double *x64) {
    uint32_t C = (uint32_t)s->n_cells;
    double *x = calloc(C, 8), *y = calloc(C, 8), *b = calloc(C, 8);
    for (uint32_t i = 0; i < C; i++) x[i] = (s->type[i] == ORC_FLUID) ? (x64 ? x64[i] : (double)x32[i]) : 0.0;
    orc_poisson_apply64(s, x, y);
    orc_poisson_rhs64(s, b);
    double rr = 0.0, bb = 0.0;
    for (uint32_t i = 0; i < C; i++) {
        double r = b[i] - y[i];
        rr += r * r;
        bb += b[i] * b[i];
    }
    free(x); free(y); free(b);
    return bb > 0.0 ? sqrt(rr / bb) : sqrt(rr);
}

int orc_poisson_solve64(const orc_state *s, double *x, double tol, int max_iter, double *relres) {
    const orc_params *p = &s->p;
    uint32_t C = (uint32_t)s->n_cells;
    double *r = malloc((size_t)C * 8), *d = calloc(C, 8), *q = malloc((size_t)C * 8);
    double *dg = calloc(C, 8);
    orc_poisson_rhs64(s, r);
    double bb = 0.0, rz = 0.0;
    for (uint32_t id = 0; id < C; id++) {
        x[id] = 0.0;
        bb += r[id] * r[id];
        if (s->type[id] == ORC_FLUID) {
            int c[3] = {(int)(id % p->nx), (int)((id / p->nx) % p->ny), (int)(id / (p->nx * p->ny))};
            dg[id] = (double)cell_diag(s, c);
        }
        if (dg[id] > 0.0) rz += r[id] * r[id] / dg[id];
    }
    double bn = sqrt(bb), rn = bn;
    int it = 0;
    double beta = 0.0;
    while (it < max_iter && rn > tol * bn && bn > 0.0) {
        for (uint32_t id = 0; id < C; id++) d[id] = (dg[id] > 0.0 ? r[id] / dg[id] : 0.0) + beta * d[id];
        orc_poisson_apply64(s, d, q);
        double dq = 0.0;
        for (uint32_t id = 0; id < C; id++) dq += d[id] * q[id];
        if (dq == 0.0) break;
        double alpha = rz / dq, rz_new = 0.0, rr = 0.0;
        for (uint32_t id = 0; id < C; id++) {
            x[id] += alpha * d[id];
            r[id] -= alpha * q[id];
            rr += r[id] * r[id];
            if (dg[id] > 0.0) rz_new += r[id] * r[id] / dg[id];
        }
        beta = rz_new / rz;
        rz = rz_new;
        rn = sqrt(rr);
        it++;
    }
    if (relres) *relres = orc_poisson_relres64(s, NULL, x);
    free(r); free(d); free(q); free(dg);
    return it;
}
