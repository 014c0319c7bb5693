/*
 * pxlflip.h -- C ABI of the MI355X-native FLIP liquid step (HIP, gfx950).
 *
 * Drop-in boundary for pxlengine's fluid path.  In the reference the caller is
 * Renderer::render_sequence (crates/pxl-render/src/lib.rs:204-252), which drives
 * FluidState (crates/pxl-render/src/fluid.rs:74-306) over the Slang kernels of
 * shaders/fluid.slang:86-220.  Those are crate-private Rust today; each entry point below
 * names the reference function whose role it takes over.
 *
 * Conventions
 *  - Every function returning int returns PXLFLIP_OK (0) or a negative PXLFLIP_ERR_*;
 *    the message is available from pxlflip_last_error(handle) (or pxlflip_last_error(NULL)
 *    for failures of the create functions, kept per host thread).
 *  - Plain pointers and sizes only.  The library owns all device memory of a handle (HBM,
 *    released by pxlflip_destroy); the caller owns every host array it passes in.
 *  - A handle is used by one host thread at a time.  Calls return when their GPU work is
 *    complete (the reference's submit + fence wait, crates/pxl-gpu/src/lib.rs:266-293).
 *  - Particle arrays use the reference's particle record: 8 floats per particle,
 *    {x, y, z, 0, vx, vy, vz, 0} (fluid.slang:9, fluid.rs:143), in seeding order.
 */
#ifndef PXLFLIP_H
#define PXLFLIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PXLFLIP_OK 0
#define PXLFLIP_ERR_INVALID (-1)  /* bad scene block / arguments (validation message) */
#define PXLFLIP_ERR_OOM (-2)      /* device or host allocation failed */
#define PXLFLIP_ERR_DEVICE (-3)   /* HIP runtime / device error */
#define PXLFLIP_ERR_INTERNAL (-4) /* library bug */

/* Pressure solver.  RBGS is the reference's red-black Gauss-Seidel with
 * pressure_iterations sweeps (fluid.slang:167-182, fluid.rs:295-300).  CG and MIXED are
 * extensions on the same matrix: Jacobi-preconditioned conjugate gradients in FP32 with a
 * fixed iteration count (CG), or multigrid-preconditioned CG iterations in FP32 with FP64
 * dot products inside an FP64 iterative-refinement loop run to a relative FP64 residual
 * (MIXED). */
#define PXLFLIP_SOLVER_RBGS 0
#define PXLFLIP_SOLVER_CG 1
#define PXLFLIP_SOLVER_MIXED 2

typedef struct pxlflip pxlflip;

/* Options the reference has no field for (it always runs RBGS on device 0). */
typedef struct pxlflip_opts {
    uint32_t struct_size;     /* sizeof(pxlflip_opts) */
    int32_t solver;           /* PXLFLIP_SOLVER_* (default RBGS) */
    uint32_t cg_iterations;   /* CG: iterations per step; 0 = pressure_iterations */
    uint32_t inner_iterations;/* MIXED: FP32 multigrid-CG iterations of the first refinement pass (default 9; pass k >= 1 runs min(2k, first) iterations) */
    uint32_t max_outer;       /* MIXED: refinement passes cap (default 50) */
    double tolerance;         /* MIXED: ||b - A p|| / ||b|| target in FP64 (default 1e-6) */
    int32_t device;           /* HIP device ordinal (default 0) */
    uint32_t scene_index;     /* i in "fluids[i] 'name': ..." validation messages */
    int32_t use_graph;        /* 1 = replay each step as a captured hipGraph (default 1) */
    uint32_t solver_workgroups; /* workgroups of the CG / MIXED solver launches; 0 = default (256), max 4096 (RBGS runs one workgroup per 8x8x16 core) */
    uint32_t max_cells;       /* grid cell limit of the validation: 0 = the reference's 16,000,000
                                 (crates/pxl-scene/src/lib.rs:732); larger admits bigger grids up to the
                                 resolution limit, 256^3 (extension) */
    uint32_t inner_max;       /* MIXED: 0 = the fixed pass schedule (inner_iterations, then 2, 4, 6, ...);
                                 > 0 = adaptive: each pass runs the MG-CG iterations the measured
                                 per-iteration reduction needs to reach the tolerance from the current
                                 residual, between 2 and inner_max (the first pass of a simulation:
                                 inner_iterations) */
} pxlflip_opts;

/* Fills *opts with the defaults above. */
int pxlflip_default_opts(pxlflip_opts *opts);

/* Replaces Fluid deserialisation + Scene::validate's fluid checks
 * (crates/pxl-scene/src/lib.rs:49-119, 710-735) and FluidState::new
 * (crates/pxl-render/src/fluid.rs:109-228): parses one `fluids[]` object (strict: unknown
 * fields are rejected), derives the grid (Fluid::grid, lib.rs:123-128), seeds the particles
 * (fluid.rs:123-149, seed = render.seed) and allocates the simulation in HBM.
 * fps is render.fps (lib.rs:626-627).  opts may be NULL. */
int pxlflip_create_from_json(const char *fluid_json, double fps, uint32_t seed,
                             const pxlflip_opts *opts, pxlflip **out);

/* The host half of create, touching no device: parse + validate the block (as above) and
 * report its grid and seeding count -- what `pxl check` prints for a fluid
 * (crates/pxl-cli/src/check.rs:98-104).  info gets n, h, origin, gravity, flip, dt (from
 * fps), substeps, pressure_iterations, n_particles (for `seed`), n_faces and n_cells;
 * device_bytes is the HBM a handle would allocate with `opts`.  info may be NULL. */
typedef struct pxlflip_info pxlflip_info;
int pxlflip_check_json(const char *fluid_json, double fps, uint32_t seed, const pxlflip_opts *opts,
                       pxlflip_info *info);

/* Whole scene documents.  Replaces Scene::from_json / Scene::validate (crates/pxl-scene/src/lib.rs:
 * 646-745) for what the fluid path reads and the summary of `pxl check` (crates/pxl-cli/src/check.rs:
 * 31-106): the document is parsed strictly at the top level (unknown scene fields are refused, the
 * version must be 1, a camera node must exist), the render block strictly (resolution, fps > 0,
 * seed, motion_blur), every fluids[i] block strictly with Scene::validate's messages, mesh
 * materials' program names checked; the contents of nodes, world, plate, simulations and programs
 * are outside the fluid path and not interpreted.  pxlflip_scene_check touches no device; line gets
 * "ok: N nodes, frames a..b, S simulation(s), F fluid(s)" (check.rs:98-104, without compiling
 * programs).  pxlflip_create_from_scene creates fluids[fluid_index] with the scene's render.fps and
 * render.seed (lib.rs:624-627).  opts as for pxlflip_create_from_json (max_cells applies to every
 * fluid; scene_index is set to fluid_index). */
typedef struct pxlflip_scene_info {
    uint32_t version, n_nodes, n_simulations, n_fluids;
    double fps;
    uint32_t seed;
    int32_t frame_start, frame_end; /* the animation's frame range; 0..0 without animation */
    uint32_t pad;
} pxlflip_scene_info;
int pxlflip_scene_check(const char *scene_json, const pxlflip_opts *opts, pxlflip_scene_info *info, char *line,
                        size_t line_len);
int pxlflip_create_from_scene(const char *scene_json, uint32_t fluid_index, const pxlflip_opts *opts, pxlflip **out);

/* Host seeding only (fluid.rs:123-148): writes the initial particles of `fluid_json`
 * for `seed` as 8-float records into aos8 (capacity max_particles) and their number into
 * *count.  With aos8 NULL only counts. */
int pxlflip_seed_particles(const char *fluid_json, uint32_t seed, float *aos8, uint64_t max_particles,
                           uint64_t *count);

/* Releases the handle and all of its device memory (drop of FluidState, fluid.rs:74-94). */
void pxlflip_destroy(pxlflip *h);

/* Replaces fluid_advance's stepping (fluid.rs:265-306): brings the simulation to
 * steps_for_frame(frame) = round((warmup + frame/fps) * fps * substeps) steps
 * (fluid.rs:234-236), re-seeding from the initial particles first when the target is
 * earlier than the steps already done or the state was never initialised. */
int pxlflip_advance_to_frame(pxlflip *h, double frame);

/* Runs n more steps (the body of the step loop, fluid.rs:289-302). */
int pxlflip_step(pxlflip *h, uint32_t n);

/* Restores the initial particles and zero pressure (fluid.rs:271-275). */
int pxlflip_reset(pxlflip *h);

/* FluidState::particle_count (fluid.rs:230-232). */
uint32_t pxlflip_particle_count(const pxlflip *h);

/* Steps done so far (FluidState::steps_done, fluid.rs:85). */
uint64_t pxlflip_steps_done(const pxlflip *h);

/* FluidState::reservation (fluid.rs:104-106): vertices and indices for the surface mesh. */
int pxlflip_reservation(const pxlflip *h, uint32_t *vertices, uint32_t *indices);

/* Grid of the simulation: cells per axis, cell size, origin, dt and the counts of faces
 * and cells (FluidPush fields, fluid.rs:158-180, and counts(), fluid.rs:238-245). */
struct pxlflip_info {
    uint32_t n[3];
    float h, dt, flip, inv_h;
    float origin[3];
    float gravity[3];
    uint32_t n_particles;
    uint32_t substeps, pressure_iterations;
    uint64_t n_faces, n_cells;
    uint64_t device_bytes; /* HBM held by this handle */
};
int pxlflip_get_info(const pxlflip *h, pxlflip_info *info);

/* Particle state in the reference record layout (8 floats per particle, seeding order).
 * Read = the particle buffer after the last step; write = teacher forcing / checkpoint
 * restore (the initial upload of fluid.rs:188/272). */
int pxlflip_read_particles(pxlflip *h, float *aos8);
int pxlflip_write_particles(pxlflip *h, const float *aos8);

/* Grid fields, in the reference's buffer layout (fluid.slang:22-36): faces are the u block,
 * then v, then w; cells are x-fastest. */
#define PXLFLIP_FIELD_VEL 0        /* float[n_faces] */
#define PXLFLIP_FIELD_VEL_OLD 1    /* float[n_faces] */
#define PXLFLIP_FIELD_ACC 2        /* int32[n_faces] */
#define PXLFLIP_FIELD_WACC 3       /* int32[n_faces] */
#define PXLFLIP_FIELD_COUNT 4      /* uint32[n_cells] */
#define PXLFLIP_FIELD_TYPE 5       /* uint32[n_cells]: 0 air, 1 fluid, 2 solid */
#define PXLFLIP_FIELD_PRESSURE 6   /* float[n_cells] */
#define PXLFLIP_FIELD_DIVERGENCE 7 /* float[n_cells] */
#define PXLFLIP_FIELD_PRESSURE64 8 /* double[n_cells]: FP64 solution of the MIXED solver */
/* CFL cap (extension, the fluid block's `cfl` field; no reference counterpart): the number k of
 * equal substeps (dt / k each, f32) the last step ran; 1 without the cap. */
int pxlflip_last_substeps(const pxlflip *h, uint32_t *k);

/* Surface-net fields of the last pxlflip_surface pass (fluid.slang:39-41; nodes (sx+1)(sy+1)(sz+1),
 * cells sx sy sz, x fastest; sizes in pxlflip_surface_info). */
#define PXLFLIP_FIELD_DENSITY 9    /* int32[nodes]: splatted density, fixed point 4096 */
#define PXLFLIP_FIELD_PHI 10       /* float[nodes]: smoothed density / 4096 - iso, -1 on the boundary */
#define PXLFLIP_FIELD_CELL_VERT 11 /* uint32[cells]: the cell's vertex, 0xffffffff for none */
int pxlflip_read_field(pxlflip *h, int field, void *out);
/* Writable: VEL, VEL_OLD, PRESSURE (teacher forcing of the warm start). */
int pxlflip_write_field(pxlflip *h, int field, const void *in);

/* ---- Surface nets (replaces the per-frame surface pass of fluid_advance,
 * crates/pxl-render/src/fluid.rs:308-330, running shaders/fluid.slang:222-355) ---- */
typedef struct pxlflip_surface_info {
    uint32_t sx, sy, sz;      /* surface cells per axis: n * surface_refine (fluid.rs:151) */
    uint32_t max_triangles;   /* the fluid's max_triangles: capacity in vertices, and in triangles */
    uint32_t vertices;        /* vertices written: min(surface cells, max_triangles) */
    uint32_t triangles;       /* triangles written: min(2 quads, max_triangles), rounded down to even */
    uint32_t vertices_total;  /* the reference's counters[0] and [1] before clipping */
    uint32_t triangles_total;
    uint32_t overflow;        /* 1 when a count was clipped (FluidState::overflow, fluid.rs:328) */
    float sh, radius, iso;    /* node spacing, splat radius, iso level (fluid.rs:152-156) */
} pxlflip_surface_info;

/* Meshes the current particles: splat -> phi -> vertices -> quads.  verts and indices are
 * DEVICE pointers to the renderer's shared buffers (bindings 43 / 44, fluid.rs:257-258):
 * vertex v is written as 8 floats {x, y, z, 0, nx, ny, nz, 0} at verts + 8 (vbase + v), triangle
 * t as 3 vertex numbers local to this fluid at indices + ibase + 3 t, as the reference writes
 * them.  Capacity: max_triangles vertices and 3 max_triangles indices from the bases
 * (FluidState::reservation, fluid.rs:104-106); verts 16-byte aligned.  Vertices are numbered in surface-cell order and
 * triangles in (node, axis) order (the reference's follow its atomic counters).  info may be
 * NULL. */
int pxlflip_surface(pxlflip *h, float *verts, uint32_t vbase, uint32_t *indices, uint32_t ibase,
                    pxlflip_surface_info *info);
/* Vulkan <-> HIP interop for the surface buffers (the renderer's vertex / index buffers live in
 * Vulkan, fluid.rs:257-258): imports a buffer exported as an opaque file descriptor
 * (VK_KHR_external_memory_fd, vkGetMemoryFdKHR) of `bytes` bytes on `device` and maps it as
 * device memory; *dev_ptr can be passed to pxlflip_surface as verts or indices.  On success the
 * descriptor belongs to HIP.  No reference counterpart (the reference meshes inside Vulkan). */
typedef struct pxlflip_external pxlflip_external;
int pxlflip_import_buffer_fd(int fd, uint64_t bytes, int device, void **dev_ptr, pxlflip_external **out);
/* The same for a memory block holding several buffers (the renderer's vertex and index buffers
 * are suballocated by gpu-allocator, crates/pxl-gpu/src/lib.rs:186-213, so each is a byte range of
 * one VkDeviceMemory): imports the whole exported allocation of `bytes` bytes once (flags:
 * PXLFLIP_IMPORT_DEDICATED when it was a dedicated allocation, VkMemoryDedicatedAllocateInfo), then
 * pxlflip_external_map gives device memory for the range [offset, offset + bytes) -- the buffer's
 * memory offset and size (vkGetBufferMemoryRequirements / the allocation's offset) -- as often as
 * needed.  pxlflip_release_buffer unmaps every range and releases the import. */
#define PXLFLIP_IMPORT_DEDICATED 1u
int pxlflip_import_memory_fd(int fd, uint64_t bytes, uint32_t flags, int device, pxlflip_external **out);
int pxlflip_external_map(pxlflip_external *ext, uint64_t offset, uint64_t bytes, void **dev_ptr);
/* Unmaps and releases an imported buffer or memory block. */
void pxlflip_release_buffer(pxlflip_external *ext);

/* The same mesh copied to host arrays (capacity as above; only the written counts are copied).
 * Either array may be NULL. */
int pxlflip_surface_host(pxlflip *h, float *verts, uint32_t *indices, pxlflip_surface_info *info);

/* Counting-sort binning of the current particle positions: the per-cell particle count of
 * p2g (fluid.slang:113-114) plus the sort the build adds.  cell_of[P] = cell of each
 * particle (seeding order), counts[C], offsets[C] = exclusive scan of counts,
 * sorted_ids[P] = particle ids grouped by cell, ascending within a cell.  Any pointer may
 * be NULL. */
int pxlflip_read_bins(pxlflip *h, uint32_t *cell_of, uint32_t *counts, uint32_t *offsets,
                      uint32_t *sorted_ids);

/* Runs only the grid phase of one step on the current particles -- bin, p2g,
 * normalize/classify, forces, divergence (fluid.rs:290-294) -- leaving the fields readable
 * for parity checks.  Does not advance steps_done. */
int pxlflip_grid_phase(pxlflip *h);

/* Statistics of the last pressure solve. */
typedef struct pxlflip_solve_stats {
    double rel_residual;  /* MIXED: final FP64 ||b - A p|| / ||b||; else -1 */
    uint32_t inner_iterations;
    uint32_t outer_iterations;
} pxlflip_solve_stats;
int pxlflip_last_solve(const pxlflip *h, pxlflip_solve_stats *st);

/* Per-kernel timing: runs n more steps without graph replay, every launch bracketed by HIP
 * events on the simulation's stream, and reports per kernel name the launch count and the
 * summed event time, largest first (the reference only times whole frames, FrameStats in
 * crates/pxl-render/src/lib.rs:77-91).  out may hold max_out entries; *n_out gets the
 * number of distinct kernels. */
typedef struct pxlflip_kernel_time {
    char name[32];
    uint32_t launches;
    double total_ms;
    double alg_bytes; /* algorithmic bytes summed over the launches (0 where not priced) */
} pxlflip_kernel_time;
int pxlflip_profile_steps(pxlflip *h, uint32_t n, pxlflip_kernel_time *out, uint32_t max_out, uint32_t *n_out);
/* The same for n surface passes (pxlflip_surface into internal buffers): one entry per stage
 * (s_clear, s_splat, s_phi, s_vflag, s_vscan, s_verts, s_qcount, s_qscan, s_quads). */
int pxlflip_profile_surface(pxlflip *h, uint32_t n, pxlflip_kernel_time *out, uint32_t max_out, uint32_t *n_out);

/* Autotuner of the solver launch shape: for each candidate workgroup count of the CG / MIXED
 * solver kernels (64, 128, 256, 512, 1024) runs `steps` timed steps, keeps the fastest, and
 * restores the particles, ids and pressure so the simulation state and steps_done are
 * unchanged.  Returns the choice and its average event-timed step time.  (RBGS sweeps run one
 * workgroup per 8x8x16 core whatever the choice.) */
int pxlflip_autotune(pxlflip *h, uint32_t steps, uint32_t *chosen_workgroups, double *best_ms);

/* Last error message of a handle; with NULL, the calling thread's last create error. */
const char *pxlflip_last_error(const pxlflip *h);

/* ---- Serving: many independent liquids stepped together.
 * No counterpart in the reference (it steps one fluid per scene, synchronously, inside
 * Renderer::render_sequence, crates/pxl-render/src/lib.rs:204-252).  A server owns a device
 * thread that keeps every admitted simulation resident in HBM and advances all of them one
 * step at a time with batched launches (one launch per kernel for the whole batch, replayed
 * as hipGraphs), using the MIXED solver for each.  Requests are admitted first-in first-out
 * while their HBM footprint fits the budget; a request arriving while others run joins the
 * batch at the next step boundary (continuous batching), and a finished one leaves at the
 * step boundary after its last step.  Each simulation computes exactly what it computes
 * alone: same kernels, same workgroup partition and reduction order, its own convergence.
 * submit / poll / wait / cancel / stats may be called from any host thread. */
typedef struct pxlflip_server pxlflip_server;

typedef struct pxlflip_server_opts {
    uint32_t struct_size;       /* sizeof(pxlflip_server_opts) */
    int32_t device;             /* HIP device ordinal (default 0) */
    uint64_t budget_bytes;      /* HBM admission budget (default 200e9) */
    uint32_t max_sims;          /* most simulations in the batch at once (default 256, max 4096) */
    uint32_t inner_iterations;  /* as pxlflip_opts (default 9) */
    uint32_t max_outer;         /* as pxlflip_opts (default 50) */
    double tolerance;           /* as pxlflip_opts (default 1e-6) */
    uint32_t solver_workgroups; /* per simulation; 0 = chosen by grid size */
    int32_t use_graph;          /* 1 = replay the batch phases as hipGraphs (default 1) */
    uint32_t profile_every;     /* > 0: every profile_every-th batch step runs its launches eagerly, each timed
                                   by HIP events and priced in algorithmic bytes (pxlflip_server_kernel_stats);
                                   0 = never (default) */
} pxlflip_server_opts;

#define PXLFLIP_REQ_QUEUED 0
#define PXLFLIP_REQ_RUNNING 1
#define PXLFLIP_REQ_DONE 2
#define PXLFLIP_REQ_FAILED 3
#define PXLFLIP_REQ_CANCELLED 4

typedef struct pxlflip_request_info {
    int32_t status;          /* PXLFLIP_REQ_* */
    uint32_t steps;          /* requested */
    uint32_t steps_done;
    uint32_t n_particles;
    uint64_t device_bytes;   /* HBM charged against the budget */
    double queue_ms;         /* submit -> admission */
    double latency_ms;       /* submit -> completion (DONE) */
    double rel_residual;     /* FP64 relative residual of the last step's pressure solve */
    uint32_t max_outer;      /* most refinement passes any of its steps took */
    uint32_t pad;
} pxlflip_request_info;

typedef struct pxlflip_server_stats {
    uint64_t submitted, completed, failed, cancelled, queued, running;
    double p50_ms, p99_ms, mean_ms, max_ms; /* submit -> completion over completed requests */
    double particle_steps;   /* sum over completed requests of particles x steps */
    double span_s;           /* first submit -> last completion */
    double throughput;       /* particle_steps / span_s */
    double busy_s;           /* device-thread time spent in batch steps */
    uint64_t batch_steps;    /* batch steps run */
    double mean_batch;       /* mean simulations per batch step */
    uint64_t bytes_in_use, peak_bytes;
    double admit_s;          /* device-thread time spent allocating admitted requests */
    uint64_t rebuilds;       /* work-list rebuilds (batch composition changes) */
    uint64_t graph_captures; /* batch graphs captured */
    double capture_s;        /* device-thread time spent capturing and instantiating them */
    double retire_s;         /* device-thread time spent retiring requests (copy-out, release) */
} pxlflip_server_stats;

/* Persistent-grid maxima of the batched launch families (no counterpart in the reference:
 * the batch engine's tuning knobs, the settings the batch autotuner picks).  A batched launch
 * over a work list uses min(list grid, family maximum) workgroups; results do not depend on
 * them.  Process-wide; read when a launch is recorded, so a change applies to batch graphs
 * captured after the call (batches and servers created later).  Values: multiples of 8 in
 * [8, 16384], else PXLFLIP_ERR_INVALID. */
#define PXLFLIP_GRID_BAND 0      /* faces, cells, fluid-list compaction, project (default 4096) */
#define PXLFLIP_GRID_PARTICLE 1  /* binning, scans, scatter (default 2048) */
#define PXLFLIP_GRID_TILE 2      /* per-tile P2G / G2P (default 2048) */
#define PXLFLIP_GRID_MG_SETUP 3  /* multigrid level types and stencils (default 2048) */
#define PXLFLIP_GRID_MG_LIST 4   /* multigrid list levels (default 512) */
#define PXLFLIP_GRID_MG_TAIL 5   /* multigrid tail (default 512) */
#define PXLFLIP_GRID_SOLVER 6    /* solver list kernels (default 1024) */
#define PXLFLIP_GRID_COUNT 7
int pxlflip_batch_get_grid(int family, uint32_t *workgroups);
int pxlflip_batch_set_grid(int family, uint32_t workgroups);
/* Workgroup sizes of the batched launch families compiled in two sizes (the batch autotuner's
 * shape knobs; no counterpart in the reference): PXLFLIP_SHAPE_TILE, the per-tile P2G / G2P
 * (512 lanes, default, or 1024), and PXLFLIP_SHAPE_BAND, the band passes faces, cells and
 * project (256, default, or 512).  Lanes stride over a tile's particles, nodes or cells, so a size
 * changes the speed and never the results.  Process-wide, read when a launch is recorded (batch
 * graphs captured after the call); another size is PXLFLIP_ERR_INVALID. */
#define PXLFLIP_SHAPE_TILE 0
#define PXLFLIP_SHAPE_BAND 1
#define PXLFLIP_SHAPE_COUNT 2
int pxlflip_batch_get_shape(int family, uint32_t *lanes);
int pxlflip_batch_set_shape(int family, uint32_t lanes);

int pxlflip_server_default_opts(pxlflip_server_opts *opts);
/* opts may be NULL.  Fails (PXLFLIP_ERR_DEVICE) when the device cannot be used. */
int pxlflip_server_create(const pxlflip_server_opts *opts, pxlflip_server **out);
/* Validates the fluid block (as pxlflip_create_from_json) and queues `steps` steps of it.
 * keep_particles = 1 keeps the final particles for pxlflip_server_read_particles. */
int pxlflip_server_submit(pxlflip_server *srv, const char *fluid_json, double fps, uint32_t seed, uint32_t steps,
                          int32_t keep_particles, uint64_t *id);
int pxlflip_server_poll(pxlflip_server *srv, uint64_t id, pxlflip_request_info *info);
/* Blocks until the request is DONE, FAILED or CANCELLED, or timeout_s passes (then returns
 * PXLFLIP_ERR_INVALID with the info of the moment); timeout_s < 0 waits without limit. */
int pxlflip_server_wait(pxlflip_server *srv, uint64_t id, double timeout_s, pxlflip_request_info *info);
/* Blocks until nothing is queued or running, or timeout_s passes. */
int pxlflip_server_drain(pxlflip_server *srv, double timeout_s);
int pxlflip_server_cancel(pxlflip_server *srv, uint64_t id);
/* The final particles of a DONE request submitted with keep_particles (AoS8, seeding order). */
int pxlflip_server_read_particles(pxlflip_server *srv, uint64_t id, float *aos8, uint64_t max_particles);
/* Forgets a finished request (its kept particles). */
int pxlflip_server_release(pxlflip_server *srv, uint64_t id);
int pxlflip_server_get_stats(pxlflip_server *srv, pxlflip_server_stats *stats);
/* Per-kernel totals of the batch steps profiled so far (profile_every): launches, HIP-event time
 * and algorithmic bytes over the members each launch covered, in first-launch order. */
int pxlflip_server_kernel_stats(pxlflip_server *srv, pxlflip_kernel_time *out, uint32_t max_out, uint32_t *n_out);
/* Message of the last failed call on this server (or of a failed request via poll). */
const char *pxlflip_server_last_error(const pxlflip_server *srv);
/* Stops the device thread (requests still queued or running are cancelled) and frees all. */
void pxlflip_server_destroy(pxlflip_server *srv);

/* ---- A fixed batch: the server's batched step without the serving loop.  n simulations
 * (fluid_jsons[i] seeded with seeds[i], or seed 0 when seeds is NULL) are allocated together
 * and advanced in lock step; opts as for the server (budget_bytes and max_sims unused).
 * No counterpart in the reference.  The handle is used by one host thread at a time. */
typedef struct pxlflip_batch pxlflip_batch;
int pxlflip_batch_create(const char *const *fluid_jsons, const uint32_t *seeds, uint32_t n, double fps,
                         const pxlflip_server_opts *opts, pxlflip_batch **out);
int pxlflip_batch_step(pxlflip_batch *b, uint32_t steps);
/* Runs `steps` steps with every launch eager and timed by HIP events on the batch stream;
 * alg_bytes prices each launch over the simulations it covers (their particles, faces, cells
 * and FLUID cells per multigrid level of that step). */
int pxlflip_batch_profile_steps(pxlflip_batch *b, uint32_t steps, pxlflip_kernel_time *out, uint32_t max_out,
                                uint32_t *n_out);
int pxlflip_batch_get_info(pxlflip_batch *b, uint32_t index, pxlflip_info *info);
int pxlflip_batch_read_particles(pxlflip_batch *b, uint32_t index, float *aos8);
int pxlflip_batch_last_solve(pxlflip_batch *b, uint32_t index, pxlflip_solve_stats *st);
/* Member `index`'s fields and binning, as pxlflip_read_field / pxlflip_read_bins (the count field
 * holds the counts of the member's last batched grid phase). */
int pxlflip_batch_read_field(pxlflip_batch *b, uint32_t index, int field, void *out);
int pxlflip_batch_read_bins(pxlflip_batch *b, uint32_t index, uint32_t *cell_of, uint32_t *counts, uint32_t *offsets,
                            uint32_t *sorted_ids);
const char *pxlflip_batch_last_error(const pxlflip_batch *b);
void pxlflip_batch_destroy(pxlflip_batch *b);

/* ---- Agent neighbour grid (shaders/sim.slang:27-104, driven per step by
 * crates/pxl-render/src/sim.rs:224-291): agents grouped by the hash of their grid cell ---- */
typedef struct pxlgrid pxlgrid;
/* The table size of a simulation with `count` agents (sim.rs:129): the next power of two of
 * 2 count, clamped to [1024, 2^23]. */
uint32_t pxlgrid_table_size(uint32_t count);
/* Scratch for grids of up to max_agents agents on `device`. */
int pxlgrid_create(uint32_t max_agents, int device, pxlgrid **out);
/* grid_clear -> grid_count -> scan -> grid_scatter (sim.slang:31-104) on `count` agents in
 * DEVICE memory, 64-byte records with the position in floats 0-2 (struct Agent, sim.slang:8;
 * 16-byte aligned): cell = int3(floor(position / max(cell_size, 1e-6))) (sim.rs:182), hash =
 * (u32(x) 73856093 ^ u32(y) 19349663 ^ u32(z) 83492791) & (table_size - 1); writes
 * counts[table_size], offsets[table_size] (exclusive scan: each hash's first slot) and
 * sorted[count] (agent ids grouped by hash; ascending within a hash, where the reference's
 * order follows its atomic cursor), all DEVICE pointers.  table_size: a power of two <= 2^23
 * (normally pxlgrid_table_size(count)).  Returns when done. */
int pxlgrid_build(pxlgrid *g, const float *agents, uint32_t count, float cell_size, uint32_t table_size,
                  uint32_t *offsets, uint32_t *counts, uint32_t *sorted);
/* The same with host arrays (copied through device scratch); outputs may be NULL. */
int pxlgrid_build_host(pxlgrid *g, const float *agents, uint32_t count, float cell_size, uint32_t table_size,
                       uint32_t *offsets, uint32_t *counts, uint32_t *sorted);
/* The neighbour grid of pxlgrid_build without a handle and without allocating: no counterpart
 * in the reference.  For callers that cannot allocate device memory (code under graph capture,
 * tight serving loops, harnesses that forbid it).  Its scratch is a library-owned static array
 * on each device (17 MB, held until the process ends), so count is at most
 * PXLGRID_STATIC_MAX_AGENTS (else PXLFLIP_ERR_INVALID).  Inputs and outputs as pxlgrid_build
 * (DEVICE pointers); runs on `stream` (a hipStream_t; NULL = the default stream) of `device`
 * and returns when done.  Calls on one device share the array: calls that overlap on one
 * device race.  Errors: pxlgrid_last_error(NULL). */
#define PXLGRID_STATIC_MAX_AGENTS (1u << 20)
int pxlgrid_build_static(int device, void *stream, const float *agents, uint32_t count, float cell_size,
                         uint32_t table_size, uint32_t *offsets, uint32_t *counts, uint32_t *sorted);
/* HIP-event time of the last build's kernels, in ms. */
double pxlgrid_last_ms(const pxlgrid *g);
const char *pxlgrid_last_error(const pxlgrid *g);
void pxlgrid_destroy(pxlgrid *g);

#ifdef __cplusplus
}
#endif
#endif /* PXLFLIP_H */
