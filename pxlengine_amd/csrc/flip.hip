// Host runtime of the FLIP path and its C ABI (include/pxlflip.h).
//
// One Sim = one liquid: its particles, grid and solver state live in one HBM arena for the
// lifetime of the handle.  A step is the reference's dispatch sequence (fluid.rs:289-302)
// re-cut into fused kernels:
//   clear -> bin count -> scan -> scatter (counting sort) -> p2g -> faces (normalize+forces)
//   -> cells (classify+stencil+divergence) -> pressure -> project -> g2p
// and, for the RBGS and CG solvers, captured once into a hipGraph that each step replays.
#include <cstring>
#include <new>

#include "sim.hpp"

namespace pxl {
thread_local std::string g_create_error;
LaunchEvents &launch_events() {
    static thread_local LaunchEvents ev;
    return ev;
}
}  // namespace pxl

using pxl::Sim;

struct pxlflip {
    Sim sim;
};

// Memory of another API imported through an exported file descriptor, and the buffers mapped
// from it (each a byte range of the allocation, as a suballocated Vulkan buffer is).
struct pxlflip_external {
    hipExternalMemory_t mem;
    void *ptr;  // the first mapping (pxlflip_import_buffer_fd)
    int device;
    uint64_t bytes;
    std::vector<void *> maps;
};

extern "C" {

int pxlflip_default_opts(pxlflip_opts *o) {
    if (!o) return PXLFLIP_ERR_INVALID;
    std::memset(o, 0, sizeof *o);
    o->struct_size = sizeof(pxlflip_opts);
    o->solver = PXLFLIP_SOLVER_RBGS;
    o->cg_iterations = 0;
    o->inner_iterations = 9;
    o->max_outer = 50;
    o->tolerance = 1e-6;
    o->device = 0;
    o->scene_index = 0;
    o->use_graph = 1;
    o->solver_workgroups = 0;
    return PXLFLIP_OK;
}

}  // extern "C"

namespace pxl {

int prepare(const char *fluid_json, double fps, uint32_t seed, const pxlflip_opts *opts, Sim &s,
            std::vector<float> *x0, std::vector<float> *v0, const FluidSpec *parsed) {
    g_create_error.clear();
    if (!fluid_json && !parsed) {
        g_create_error = "fluid_json is NULL";
        return PXLFLIP_ERR_INVALID;
    }
    pxlflip_opts o;
    pxlflip_default_opts(&o);
    if (opts) {
        if (opts->struct_size != sizeof(pxlflip_opts)) {
            g_create_error = "pxlflip_opts.struct_size mismatch (use pxlflip_default_opts)";
            return PXLFLIP_ERR_INVALID;
        }
        o = *opts;
    }
    if (o.solver < PXLFLIP_SOLVER_RBGS || o.solver > PXLFLIP_SOLVER_MIXED) {
        g_create_error = "unknown solver " + std::to_string(o.solver);
        return PXLFLIP_ERR_INVALID;
    }
    if (o.solver_workgroups > MAX_SOLVER_BLOCKS) {
        g_create_error = "solver_workgroups must be <= " + std::to_string(MAX_SOLVER_BLOCKS);
        return PXLFLIP_ERR_INVALID;
    }
    if (o.solver == PXLFLIP_SOLVER_MIXED && (o.inner_iterations == 0 || !(o.tolerance > 0.0))) {
        g_create_error = "mixed solver needs inner_iterations > 0 and tolerance > 0";
        return PXLFLIP_ERR_INVALID;
    }
    if (!(fps > 0.0)) {  // Scene::validate (crates/pxl-scene/src/lib.rs:670-672)
        g_create_error = "render.fps must be > 0";
        return PXLFLIP_ERR_INVALID;
    }
    FluidSpec spec;
    std::string err = parsed ? std::string() : parse_fluid(fluid_json, spec);
    if (parsed) spec = *parsed;
    if (err.empty()) err = validate_fluid(spec, o.scene_index, o.max_cells ? (uint64_t)o.max_cells : REFERENCE_MAX_CELLS);
    if (!err.empty()) {
        g_create_error = err;
        return PXLFLIP_ERR_INVALID;
    }
    uint64_t np = seed_particles(spec, seed, x0, v0);
    if (np == 0) {  // fluid.rs:149
        g_create_error = "fluid '" + spec.name + "': fill blocks contain no grid cells";
        return PXLFLIP_ERR_INVALID;
    }
    if (np > 0xFFFFFFFFull) {
        g_create_error = "fluid '" + spec.name + "': too many particles";
        return PXLFLIP_ERR_INVALID;
    }
    s.spec = spec;
    s.opts = o;
    s.fps = fps;
    s.seed = seed;
    s.device = o.device;
    s.iters = (o.solver == PXLFLIP_SOLVER_CG && o.cg_iterations) ? o.cg_iterations : spec.pressure_iterations;
    uint32_t n[3];
    float hcell;
    fluid_grid(spec, n, &hcell);
    DevSim &d = s.d;
    d.nx = (int)n[0];
    d.ny = (int)n[1];
    d.nz = (int)n[2];
    d.np = (uint32_t)np;
    d.h = hcell;
    d.dt = fluid_dt(fps, spec.substeps);
    d.flip = spec.flip;
    d.inv_h = 1.0f / hcell;
    d.ox = spec.bounds[0][0];
    d.oy = spec.bounds[0][1];
    d.oz = spec.bounds[0][2];
    d.gx = spec.gravity[0];
    d.gy = spec.gravity[1];
    d.gz = spec.gravity[2];
    d.nu = (n[0] + 1) * n[1] * n[2];
    d.nv = n[0] * (n[1] + 1) * n[2];
    d.nw = n[0] * n[1] * (n[2] + 1);
    d.nfaces = d.nu + d.nv + d.nw;
    d.ncells = n[0] * n[1] * n[2];
    d.hx = (n[0] + 1) / 2;
    d.hsplit = d.hx * n[1] * n[2];
    d.ntx = (int)((n[0] + TILE - 1) / TILE);
    d.nty = (int)((n[1] + TILE - 1) / TILE);
    d.ntz = (int)((n[2] + TILE - 1) / TILE);
    d.ntiles = (uint32_t)d.ntx * (uint32_t)d.nty * (uint32_t)d.ntz;
    d.nkeys = d.ntiles * TILE_CELLS;
    d.n_obst = (uint32_t)spec.obstacles.size();
    {
        // surface-net parameters (fluid.rs:151-157): sx = n * refine, sh = h / refine,
        // radius = particle_radius * sh, iso = 0.5 * ppc * 0.638 * (radius / h)^3
        DevSurf &sf = s.surf;
        const uint32_t r = spec.surface_refine;
        sf.sx = (int)(n[0] * r);
        sf.sy = (int)(n[1] * r);
        sf.sz = (int)(n[2] * r);
        sf.sh = hcell / (float)r;
        sf.radius = spec.particle_radius * sf.sh;
        const float rh = sf.radius / hcell;
        const float kernel_volume = 0.638f * (rh * rh * rh);
        sf.iso = 0.5f * (float)spec.particles_per_cell * kernel_volume;
        sf.max_verts = spec.max_triangles;
        sf.max_tris = spec.max_triangles;
        sf.ox = spec.bounds[0][0];
        sf.oy = spec.bounds[0][1];
        sf.oz = spec.bounds[0][2];
        sf.nnodes = (uint32_t)(sf.sx + 1) * (uint32_t)(sf.sy + 1) * (uint32_t)(sf.sz + 1);
        sf.ncells = (uint32_t)sf.sx * (uint32_t)sf.sy * (uint32_t)sf.sz;
    }
    {
        // twice the tiles the liquid fills at seeding (room for the band around it and for
        // spreading), at least 64, at most the grid's tile count and 1024 -- 2048 on grids of
        // 16384 tiles or more (measured: on the 64-request batch 2x is 0.7% faster than 4x, 8x
        // 0.3%; the 256^3 two-block scene is 1.4% faster with 2048, the 128^3 dam 0.4% slower)
        const uint64_t liquid_tiles = ((uint64_t)np / std::max<uint32_t>(spec.particles_per_cell, 1u) + TILE_CELLS - 1) / TILE_CELLS;
        const uint64_t cap = d.ntiles >= 16384 ? 2048 : 1024;
        d.tile_wgs = (uint32_t)std::min<uint64_t>(std::min<uint64_t>(d.ntiles, cap), std::max<uint64_t>(64, 2 * liquid_tiles));
    }
    d.cg_blocks = (int)(o.solver_workgroups ? o.solver_workgroups : DEFAULT_SOLVER_BLOCKS);
    d.split_colours = o.solver == PXLFLIP_SOLVER_CG ? 0 : 1;  // RBGS and the multigrid sweep one colour at a time
    d.mg_levels = 0;
    if (o.solver == PXLFLIP_SOLVER_MIXED) {
        int l = 0;
        d.mg[0].nx = d.nx;
        d.mg[0].ny = d.ny;
        d.mg[0].nz = d.nz;
        auto cells_of = [&](int q) { return (size_t)d.mg[q].nx * d.mg[q].ny * d.mg[q].nz; };
        // at least one coarse level, then halve until no axis exceeds 2 cells
        while ((l == 0 || std::max(d.mg[l].nx, std::max(d.mg[l].ny, d.mg[l].nz)) > 2) && l + 1 < MG_MAX) {
            MGLevel &c = d.mg[l + 1];
            c.nx = (d.mg[l].nx + 1) / 2;
            c.ny = (d.mg[l].ny + 1) / 2;
            c.nz = (d.mg[l].nz + 1) / 2;
            c.hx = (c.nx + 1) / 2;
            c.hsplit = (uint32_t)c.hx * (uint32_t)c.ny * (uint32_t)c.nz;
            l++;
        }
        d.mg_levels = l + 1;
        // the multigrid kernels address a colour-split vector (2 hsplit floats) as a raw buffer
        // with 32-bit byte offsets below 2^31
        if (8ull * (uint64_t)d.hsplit >= 0x80000000ull) {
            g_create_error = "fluid '" + spec.name + "': grid too large for the mixed-precision solver";
            return PXLFLIP_ERR_INVALID;
        }
        // the tail: the longest run of coarsest levels (from level 1 on) that fits the LDS
        size_t tail = 0;
        d.mg_tail = l + 1;
        for (int q = l; q >= 1 && tail + cells_of(q) <= (size_t)MG_TAIL_CELLS; q--) {
            tail += cells_of(q);
            d.mg_tail = q;
        }
        if (d.mg_tail > l) {
            g_create_error = "fluid '" + spec.name + "': grid too large for the multigrid hierarchy";
            return PXLFLIP_ERR_INVALID;
        }
        // the mid levels: the run of list levels just above the tail small enough for one workgroup
        d.mg_mid = d.mg_tail;
        for (int q = d.mg_tail - 1; q >= 1 && cells_of(q) <= (size_t)MG_MID_CELLS; q--) d.mg_mid = q;
    }
    return PXLFLIP_OK;
}

void fill_info(const Sim &s, pxlflip_info *info) {
    const DevSim &d = s.d;
    info->n[0] = (uint32_t)d.nx;
    info->n[1] = (uint32_t)d.ny;
    info->n[2] = (uint32_t)d.nz;
    info->h = d.h;
    info->dt = d.dt;
    info->flip = d.flip;
    info->inv_h = d.inv_h;
    info->origin[0] = d.ox;
    info->origin[1] = d.oy;
    info->origin[2] = d.oz;
    info->gravity[0] = d.gx;
    info->gravity[1] = d.gy;
    info->gravity[2] = d.gz;
    info->n_particles = d.np;
    info->substeps = s.spec.substeps;
    info->pressure_iterations = s.iters;
    info->n_faces = d.nfaces;
    info->n_cells = d.ncells;
    info->device_bytes = s.arena_bytes;
}

// Why `field` cannot be read from this simulation ("" when it can).
std::string field_check(const Sim &s, int field) {
    if (field < PXLFLIP_FIELD_VEL || field > PXLFLIP_FIELD_CELL_VERT) return "unknown field " + std::to_string(field);
    if (field == PXLFLIP_FIELD_PRESSURE64 && s.opts.solver != PXLFLIP_SOLVER_MIXED)
        return "PRESSURE64 exists only with the MIXED solver";
    if (field >= PXLFLIP_FIELD_DENSITY && !s.surf_arena) return "no surface pass has run on this handle";
    return "";
}

// Copies one field (pxlflip.h's PXLFLIP_FIELD_*, checked by field_check) in the reference
// layout into host memory, on the simulation's stream; returns when it is there.
void read_field_to_host(Sim &s, int field, void *out) {
    const DevSim &d = s.d;
    const size_t F = d.nfaces, C = d.ncells;
    switch (field) {
        case PXLFLIP_FIELD_ACC:
        case PXLFLIP_FIELD_WACC: {
            DeviceBuffer a(4 * F), w(4 * F);
            launch_acc_fields(d, a.as<int32_t>(), w.as<int32_t>(), s.stream);
            HIP_CHECK(hipMemcpyAsync(out, field == PXLFLIP_FIELD_ACC ? a.p : w.p, 4 * F, hipMemcpyDeviceToHost, s.stream));
            HIP_CHECK(hipStreamSynchronize(s.stream));
            return;
        }
        case PXLFLIP_FIELD_TYPE:
        case PXLFLIP_FIELD_COUNT: {
            DeviceBuffer buf(4 * C);
            if (field == PXLFLIP_FIELD_TYPE) launch_type_u32(d, buf.as<uint32_t>(), s.stream);
            else launch_count_linear(d, buf.as<uint32_t>(), s.stream);
            HIP_CHECK(hipMemcpyAsync(out, buf.p, 4 * C, hipMemcpyDeviceToHost, s.stream));
            HIP_CHECK(hipStreamSynchronize(s.stream));
            return;
        }
        default: break;
    }
    const void *src = nullptr;
    size_t bytes = 0;
    switch (field) {
        case PXLFLIP_FIELD_VEL: src = d.vel; bytes = 4 * F; break;
        case PXLFLIP_FIELD_VEL_OLD: src = d.vel_old; bytes = 4 * F; break;
        case PXLFLIP_FIELD_PRESSURE: src = d.pressure; bytes = 4 * C; break;
        case PXLFLIP_FIELD_DIVERGENCE: src = d.div; bytes = 4 * C; break;
        case PXLFLIP_FIELD_PRESSURE64: src = s.x64(s.x64_cur); bytes = 8 * C; break;
        case PXLFLIP_FIELD_DENSITY: src = s.surf.density; bytes = 4ull * s.surf.nnodes; break;
        case PXLFLIP_FIELD_PHI: src = s.surf.phi; bytes = 4ull * s.surf.nnodes; break;
        default: src = s.surf.cell_vert; bytes = 4ull * s.surf.ncells; break;
    }
    HIP_CHECK(hipMemcpyAsync(out, src, bytes, hipMemcpyDeviceToHost, s.stream));
    HIP_CHECK(hipStreamSynchronize(s.stream));
}

// The counting sort of the current positions in reference cell numbering (pxlflip_read_bins);
// any output may be NULL.  Leaves every count zero, as the next step expects.
void read_bins_to_host(Sim &s, uint32_t *cell_of, uint32_t *counts, uint32_t *offsets, uint32_t *sorted_ids) {
    const DevSim &d = s.d;
    const size_t C = d.ncells;
    DeviceBuffer cbuf(4ull * d.np), sbuf(4ull * d.np), cnt(4 * C), off(4 * C),
        tmp(4ull * (scan_tiles((uint32_t)C) + 1));
    // the step's binning (tiled keys) on the current positions ...
    launch_fill_u32(d.count, d.nkeys, 0, s.stream);
    launch_bin_count(d, s.stream);
    // ... reported in reference cell numbering
    launch_count_linear(d, cnt.as<uint32_t>(), s.stream);
    // the step clears only the counts of the tiles its last tile list holds (bin_count above
    // emptied that list), so every count must be zero again before the next step
    launch_fill_u32(d.count, d.nkeys, 0, s.stream);
    launch_exclusive_scan(cnt.as<uint32_t>(), (uint32_t)C, tmp.as<uint32_t>(), off.as<uint32_t>(), s.stream);
    launch_bins_out(d, off.as<uint32_t>(), cbuf.as<uint32_t>(), sbuf.as<uint32_t>(), s.stream);
    launch_sort_segments((uint32_t)C, off.as<uint32_t>(), cnt.as<uint32_t>(), sbuf.as<uint32_t>(), s.stream);
    if (cell_of) HIP_CHECK(hipMemcpyAsync(cell_of, cbuf.p, 4ull * d.np, hipMemcpyDeviceToHost, s.stream));
    if (counts) HIP_CHECK(hipMemcpyAsync(counts, cnt.p, 4 * C, hipMemcpyDeviceToHost, s.stream));
    if (offsets) HIP_CHECK(hipMemcpyAsync(offsets, off.p, 4 * C, hipMemcpyDeviceToHost, s.stream));
    if (sorted_ids) HIP_CHECK(hipMemcpyAsync(sorted_ids, sbuf.p, 4ull * d.np, hipMemcpyDeviceToHost, s.stream));
    HIP_CHECK(hipStreamSynchronize(s.stream));
}

}  // namespace pxl

extern "C" {

int pxlflip_check_json(const char *fluid_json, double fps, uint32_t seed, const pxlflip_opts *opts,
                       pxlflip_info *info) {
    pxl::Sim s;
    int rc = pxl::prepare(fluid_json, fps, seed, opts, s, nullptr, nullptr);
    if (rc != PXLFLIP_OK) return rc;
    s.arena_bytes = pxl::Sim::plan_total(s.plan());
    if (info) pxl::fill_info(s, info);
    return PXLFLIP_OK;
}

int pxlflip_seed_particles(const char *fluid_json, uint32_t seed, float *aos8, uint64_t max_particles,
                           uint64_t *count) {
    pxl::Sim s;
    std::vector<float> x0, v0;
    int rc = pxl::prepare(fluid_json, 24.0, seed, nullptr, s, aos8 ? &x0 : nullptr, aos8 ? &v0 : nullptr);
    if (rc != PXLFLIP_OK) return rc;
    const uint64_t P = s.d.np;
    if (count) *count = P;
    if (aos8) {
        if (max_particles < P) {
            pxl::g_create_error = "aos8 holds " + std::to_string(max_particles) + " particles, need " + std::to_string(P);
            return PXLFLIP_ERR_INVALID;
        }
        for (uint64_t i = 0; i < P; i++) {
            float *r = aos8 + 8 * i;
            r[0] = x0[i]; r[1] = x0[P + i]; r[2] = x0[2 * P + i]; r[3] = 0.0f;
            r[4] = v0[i]; r[5] = v0[P + i]; r[6] = v0[2 * P + i]; r[7] = 0.0f;
        }
    }
    return PXLFLIP_OK;
}

}  // extern "C"

namespace {
// FluidState::new for a fluid block given as JSON text or already parsed (from a scene).
int create_sim(const char *fluid_json, const pxl::FluidSpec *parsed, double fps, uint32_t seed,
               const pxlflip_opts *opts, pxlflip **out) {
    using namespace pxl;
    if (!out) {
        g_create_error = "out is NULL";
        return PXLFLIP_ERR_INVALID;
    }
    *out = nullptr;
    pxlflip *h = new (std::nothrow) pxlflip();
    if (!h) {
        g_create_error = "out of memory";
        return PXLFLIP_ERR_OOM;
    }
    Sim &s = h->sim;
    int rc = prepare(fluid_json, fps, seed, opts, s, nullptr, nullptr, parsed);
    if (rc != PXLFLIP_OK) {
        delete h;
        return rc;
    }
    try {
        HIP_CHECK(hipSetDevice(s.device));
        HIP_CHECK(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
        s.allocate();
        s.reset();
        s.initialized = false;  // first advance re-seeds, as fluid.rs:271
    } catch (const HipError &e) {
        g_create_error = std::string("HIP error in ") + e.what + ": " + hipGetErrorString(e.code);
        delete h;
        return e.code == hipErrorOutOfMemory ? PXLFLIP_ERR_OOM : PXLFLIP_ERR_DEVICE;
    } catch (const std::bad_alloc &) {
        g_create_error = "fluid '" + s.spec.name + "': out of device memory";
        delete h;
        return PXLFLIP_ERR_OOM;
    } catch (const std::exception &e) {
        g_create_error = "fluid '" + s.spec.name + "': " + e.what();
        delete h;
        return PXLFLIP_ERR_INTERNAL;
    }
    *out = h;
    return PXLFLIP_OK;
}

// Scene::from_json with the opts' cell limit; "" or the message.
std::string load_scene(const char *scene_json, const pxlflip_opts *opts, pxl::SceneSpec &sc) {
    if (!scene_json) return "scene_json is NULL";
    if (opts && opts->struct_size != sizeof(pxlflip_opts)) return "pxlflip_opts.struct_size mismatch (use pxlflip_default_opts)";
    const uint64_t max_cells = opts && opts->max_cells ? (uint64_t)opts->max_cells : pxl::REFERENCE_MAX_CELLS;
    return pxl::parse_scene(scene_json, sc, max_cells);
}
}  // namespace

extern "C" {

int pxlflip_create_from_json(const char *fluid_json, double fps, uint32_t seed, const pxlflip_opts *opts,
                             pxlflip **out) {
    return create_sim(fluid_json, nullptr, fps, seed, opts, out);
}

int pxlflip_scene_check(const char *scene_json, const pxlflip_opts *opts, pxlflip_scene_info *info, char *line,
                        size_t line_len) {
    pxl::g_create_error.clear();
    pxl::SceneSpec sc;
    const std::string err = load_scene(scene_json, opts, sc);
    if (!err.empty()) {
        pxl::g_create_error = err;
        return PXLFLIP_ERR_INVALID;
    }
    // each fluid must also seed particles (fluid.rs:149), as FluidState::new would report
    for (size_t i = 0; i < sc.fluids.size(); i++) {
        if (pxl::seed_particles(sc.fluids[i], sc.seed, nullptr, nullptr) == 0) {
            pxl::g_create_error = "fluid '" + sc.fluids[i].name + "': fill blocks contain no grid cells";
            return PXLFLIP_ERR_INVALID;
        }
    }
    if (info) {
        std::memset(info, 0, sizeof *info);
        info->version = sc.version;
        info->n_nodes = sc.n_nodes;
        info->n_simulations = sc.n_simulations;
        info->n_fluids = (uint32_t)sc.fluids.size();
        info->fps = (double)sc.fps;
        info->seed = sc.seed;
        info->frame_start = sc.animated ? sc.frame_start : 0;
        info->frame_end = sc.animated ? sc.frame_end : 0;
    }
    if (line && line_len) std::snprintf(line, line_len, "%s", pxl::check_line(sc).c_str());
    return PXLFLIP_OK;
}

int pxlflip_create_from_scene(const char *scene_json, uint32_t fluid_index, const pxlflip_opts *opts, pxlflip **out) {
    pxl::g_create_error.clear();
    if (out) *out = nullptr;
    pxl::SceneSpec sc;
    const std::string err = load_scene(scene_json, opts, sc);
    if (!err.empty()) {
        pxl::g_create_error = err;
        return PXLFLIP_ERR_INVALID;
    }
    if (fluid_index >= sc.fluids.size()) {
        pxl::g_create_error = "the scene has " + std::to_string(sc.fluids.size()) + " fluid(s), no fluids[" +
                              std::to_string(fluid_index) + "]";
        return PXLFLIP_ERR_INVALID;
    }
    pxlflip_opts o;
    pxlflip_default_opts(&o);
    if (opts) o = *opts;
    o.scene_index = fluid_index;
    // render.fps and render.seed of the scene (lib.rs:624-627; fluid.rs:120, 164)
    return create_sim(nullptr, &sc.fluids[fluid_index], (double)sc.fps, sc.seed, &o, out);
}

void pxlflip_destroy(pxlflip *h) {
    if (!h) return;
    (void)hipSetDevice(h->sim.device);
    delete h;
}

int pxlflip_advance_to_frame(pxlflip *h, double frame) {
    if (!h) return PXLFLIP_ERR_INVALID;
    Sim *s = &h->sim;
    return pxl::guarded(s, [&] {
        uint64_t target = pxl::steps_for_frame((double)s->spec.warmup, frame, s->fps, s->spec.substeps);
        if (!s->initialized || target < s->steps_done) s->reset();
        while (s->steps_done < target) {
            uint64_t n = std::min<uint64_t>(target - s->steps_done, 1u << 20);
            s->run_steps((uint32_t)n);
        }
    });
}

int pxlflip_step(pxlflip *h, uint32_t n) {
    if (!h) return PXLFLIP_ERR_INVALID;
    Sim *s = &h->sim;
    return pxl::guarded(s, [&] {
        if (!s->initialized) s->reset();
        s->run_steps(n);
    });
}

int pxlflip_reset(pxlflip *h) {
    if (!h) return PXLFLIP_ERR_INVALID;
    Sim *s = &h->sim;
    return pxl::guarded(s, [&] { s->reset(); });
}

uint32_t pxlflip_particle_count(const pxlflip *h) { return h ? h->sim.d.np : 0; }

uint64_t pxlflip_steps_done(const pxlflip *h) { return h ? h->sim.steps_done : 0; }

int pxlflip_reservation(const pxlflip *h, uint32_t *vertices, uint32_t *indices) {
    if (!h) return PXLFLIP_ERR_INVALID;
    if (vertices) *vertices = h->sim.spec.max_triangles;
    if (indices) *indices = h->sim.spec.max_triangles * 3u;
    return PXLFLIP_OK;
}

int pxlflip_get_info(const pxlflip *h, pxlflip_info *info) {
    if (!h || !info) return PXLFLIP_ERR_INVALID;
    pxl::fill_info(h->sim, info);
    return PXLFLIP_OK;
}

int pxlflip_read_particles(pxlflip *h, float *aos8) {
    if (!h || !aos8) return PXLFLIP_ERR_INVALID;
    Sim *s = &h->sim;
    return pxl::guarded(s, [&] {
        if (!s->initialized) s->reset();
        pxl::DeviceBuffer buf(32ull * s->d.np);
        pxl::launch_unpack_aos8(s->d, buf.as<float>(), s->stream);
        HIP_CHECK(hipMemcpyAsync(aos8, buf.p, 32ull * s->d.np, hipMemcpyDeviceToHost, s->stream));
        HIP_CHECK(hipStreamSynchronize(s->stream));
    });
}

int pxlflip_write_particles(pxlflip *h, const float *aos8) {
    if (!h || !aos8) return PXLFLIP_ERR_INVALID;
    Sim *s = &h->sim;
    return pxl::guarded(s, [&] {
        pxl::DeviceBuffer buf(32ull * s->d.np);
        HIP_CHECK(hipMemcpyAsync(buf.p, aos8, 32ull * s->d.np, hipMemcpyHostToDevice, s->stream));
        pxl::launch_pack_aos8(s->d, buf.as<float>(), s->stream);
        s->mark_all_tiles();
        HIP_CHECK(hipStreamSynchronize(s->stream));
        s->initialized = true;
    });
}

int pxlflip_read_field(pxlflip *h, int field, void *out) {
    if (!h || !out) return PXLFLIP_ERR_INVALID;
    Sim *s = &h->sim;
    std::string why = pxl::field_check(*s, field);
    if (!why.empty()) return pxl::fail(s, PXLFLIP_ERR_INVALID, why);
    return pxl::guarded(s, [&] { pxl::read_field_to_host(*s, field, out); });
}

namespace {
void fill_surface_info(const pxl::Sim &s, const uint32_t ctr[2], pxlflip_surface_info *info) {
    if (!info) return;
    std::memset(info, 0, sizeof *info);
    const pxl::DevSurf &sf = s.surf;
    info->sx = (uint32_t)sf.sx;
    info->sy = (uint32_t)sf.sy;
    info->sz = (uint32_t)sf.sz;
    info->max_triangles = sf.max_tris;
    info->sh = sf.sh;
    info->radius = sf.radius;
    info->iso = sf.iso;
    info->vertices_total = ctr[0];
    info->triangles_total = ctr[1];
    // fluid.rs:328-330
    info->overflow = (ctr[0] > sf.max_verts || ctr[1] > sf.max_tris) ? 1u : 0u;
    info->vertices = std::min(ctr[0], sf.max_verts);
    info->triangles = std::min(ctr[1], sf.max_tris) / 2u * 2u;
}
}  // namespace

int pxlflip_import_memory_fd(int fd, uint64_t bytes, uint32_t flags, int device, pxlflip_external **out) {
    pxl::g_create_error.clear();
    if (!out || fd < 0 || bytes == 0 || (flags & ~PXLFLIP_IMPORT_DEDICATED)) {
        pxl::g_create_error = "import_memory_fd needs a file descriptor, a size, known flags and an output pointer";
        return PXLFLIP_ERR_INVALID;
    }
    *out = nullptr;
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) {
        pxl::g_create_error = std::string("hipSetDevice: ") + hipGetErrorString(e);
        return PXLFLIP_ERR_DEVICE;
    }
    hipExternalMemoryHandleDesc desc;
    std::memset(&desc, 0, sizeof desc);
    desc.type = hipExternalMemoryHandleTypeOpaqueFd;
    desc.handle.fd = fd;
    desc.size = bytes;
    desc.flags = (flags & PXLFLIP_IMPORT_DEDICATED) ? hipExternalMemoryDedicated : 0u;
    hipExternalMemory_t mem = nullptr;
    e = hipImportExternalMemory(&mem, &desc);
    if (e != hipSuccess) {
        pxl::g_create_error = std::string("hipImportExternalMemory: ") + hipGetErrorString(e);
        return PXLFLIP_ERR_DEVICE;
    }
    auto *x = new (std::nothrow) pxlflip_external{mem, nullptr, device, bytes, {}};
    if (!x) {
        (void)hipDestroyExternalMemory(mem);
        pxl::g_create_error = "out of memory";
        return PXLFLIP_ERR_OOM;
    }
    *out = x;
    return PXLFLIP_OK;
}

int pxlflip_external_map(pxlflip_external *x, uint64_t offset, uint64_t bytes, void **dev_ptr) {
    pxl::g_create_error.clear();
    if (!x || !dev_ptr || bytes == 0 || offset > x->bytes || bytes > x->bytes - offset) {
        pxl::g_create_error = "external_map: the range must lie inside the imported allocation";
        return PXLFLIP_ERR_INVALID;
    }
    *dev_ptr = nullptr;
    hipError_t e = hipSetDevice(x->device);
    if (e != hipSuccess) {
        pxl::g_create_error = std::string("hipSetDevice: ") + hipGetErrorString(e);
        return PXLFLIP_ERR_DEVICE;
    }
    hipExternalMemoryBufferDesc bd;
    std::memset(&bd, 0, sizeof bd);
    bd.offset = offset;
    bd.size = bytes;
    void *p = nullptr;
    e = hipExternalMemoryGetMappedBuffer(&p, x->mem, &bd);
    if (e != hipSuccess) {
        pxl::g_create_error = std::string("hipExternalMemoryGetMappedBuffer: ") + hipGetErrorString(e);
        return PXLFLIP_ERR_DEVICE;
    }
    x->maps.push_back(p);
    *dev_ptr = p;
    return PXLFLIP_OK;
}

int pxlflip_import_buffer_fd(int fd, uint64_t bytes, int device, void **dev_ptr, pxlflip_external **out) {
    if (!dev_ptr || !out) {
        pxl::g_create_error = "import_buffer_fd needs a file descriptor, a size and output pointers";
        return PXLFLIP_ERR_INVALID;
    }
    *dev_ptr = nullptr;
    pxlflip_external *x = nullptr;
    int rc = pxlflip_import_memory_fd(fd, bytes, 0u, device, &x);
    if (rc != PXLFLIP_OK) return rc;
    rc = pxlflip_external_map(x, 0, bytes, dev_ptr);
    if (rc != PXLFLIP_OK) {
        const std::string msg = pxl::g_create_error;
        pxlflip_release_buffer(x);
        pxl::g_create_error = msg;
        return rc;
    }
    x->ptr = *dev_ptr;
    *out = x;
    return PXLFLIP_OK;
}

void pxlflip_release_buffer(pxlflip_external *x) {
    if (!x) return;
    (void)hipSetDevice(x->device);
    for (void *p : x->maps) (void)hipFree(p);
    (void)hipDestroyExternalMemory(x->mem);
    delete x;
}

int pxlflip_last_substeps(const pxlflip *h, uint32_t *k) {
    if (!h || !k) return PXLFLIP_ERR_INVALID;
    *k = h->sim.last_substeps;
    return PXLFLIP_OK;
}

int pxlflip_surface(pxlflip *h, float *verts, uint32_t vbase, uint32_t *indices, uint32_t ibase,
                    pxlflip_surface_info *info) {
    if (!h) return PXLFLIP_ERR_INVALID;
    Sim *s = &h->sim;
    if (!verts || !indices) return pxl::fail(s, PXLFLIP_ERR_INVALID, "verts and indices must be device pointers");
    if (reinterpret_cast<uintptr_t>(verts) % 16 != 0)
        return pxl::fail(s, PXLFLIP_ERR_INVALID, "verts must be 16-byte aligned (two float4 per vertex)");
    return pxl::guarded(s, [&] {
        if (!s->initialized) s->reset();
        uint32_t ctr[2];
        s->surface(verts, vbase, indices, ibase, ctr);
        fill_surface_info(*s, ctr, info);
    });
}

int pxlflip_surface_host(pxlflip *h, float *verts, uint32_t *indices, pxlflip_surface_info *info) {
    if (!h) return PXLFLIP_ERR_INVALID;
    Sim *s = &h->sim;
    return pxl::guarded(s, [&] {
        if (!s->initialized) s->reset();
        const uint32_t mv = s->surf.max_verts, mt = s->surf.max_tris;
        pxl::DeviceBuffer vb(32ull * mv), ib(12ull * mt);
        uint32_t ctr[2];
        s->surface(vb.as<float>(), 0, ib.as<uint32_t>(), 0, ctr);
        pxlflip_surface_info tmp;
        fill_surface_info(*s, ctr, &tmp);
        if (verts && tmp.vertices)
            HIP_CHECK(hipMemcpyAsync(verts, vb.p, 32ull * tmp.vertices, hipMemcpyDeviceToHost, s->stream));
        if (indices && tmp.triangles)
            HIP_CHECK(hipMemcpyAsync(indices, ib.p, 12ull * tmp.triangles, hipMemcpyDeviceToHost, s->stream));
        HIP_CHECK(hipStreamSynchronize(s->stream));
        if (info) *info = tmp;
    });
}

int pxlflip_write_field(pxlflip *h, int field, const void *in) {
    if (!h || !in) return PXLFLIP_ERR_INVALID;
    Sim *s = &h->sim;
    const pxl::DevSim &d = s->d;
    void *dst = nullptr;
    size_t bytes = 0;
    switch (field) {
        case PXLFLIP_FIELD_VEL: dst = d.vel; bytes = 4ull * d.nfaces; break;
        case PXLFLIP_FIELD_VEL_OLD: dst = d.vel_old; bytes = 4ull * d.nfaces; break;
        case PXLFLIP_FIELD_PRESSURE: dst = d.pressure; bytes = 4ull * d.ncells; break;
        default: return pxl::fail(s, PXLFLIP_ERR_INVALID, "field " + std::to_string(field) + " is not writable");
    }
    return pxl::guarded(s, [&] {
        HIP_CHECK(hipMemcpyAsync(dst, in, bytes, hipMemcpyHostToDevice, s->stream));
        s->mark_all_tiles();
        HIP_CHECK(hipStreamSynchronize(s->stream));
    });
}

int pxlflip_read_bins(pxlflip *h, uint32_t *cell_of, uint32_t *counts, uint32_t *offsets, uint32_t *sorted_ids) {
    if (!h) return PXLFLIP_ERR_INVALID;
    Sim *s = &h->sim;
    return pxl::guarded(s, [&] {
        if (!s->initialized) s->reset();
        pxl::read_bins_to_host(*s, cell_of, counts, offsets, sorted_ids);
    });
}

int pxlflip_grid_phase(pxlflip *h) {
    if (!h) return PXLFLIP_ERR_INVALID;
    Sim *s = &h->sim;
    return pxl::guarded(s, [&] {
        if (!s->initialized) s->reset();
        s->enqueue_grid_phase();
        HIP_CHECK(hipStreamSynchronize(s->stream));
    });
}

}  // extern "C"

namespace {
// Runs work() with every launch event-timed and reports per-kernel totals, longest first.
template <typename F>
int profile_launches(Sim *s, F &&work, pxlflip_kernel_time *out, uint32_t max_out, uint32_t *n_out) {
    std::vector<pxlflip_kernel_time> agg;
    int rc = pxl::guarded(s, [&] {
        if (!s->initialized) s->reset();
        s->timer.clear();
        s->timer.on = true;
        try {
            work();
        } catch (...) {
            s->timer.on = false;
            throw;
        }
        s->timer.on = false;
        for (const auto &r : s->timer.recs) {
            float ms = 0.0f;
            HIP_CHECK(hipEventElapsedTime(&ms, r.a, r.b));
            auto it = std::find_if(agg.begin(), agg.end(),
                                   [&](const pxlflip_kernel_time &k) { return std::strcmp(k.name, r.name) == 0; });
            if (it == agg.end()) {
                pxlflip_kernel_time k{};
                std::snprintf(k.name, sizeof k.name, "%s", r.name);
                agg.push_back(k);
                it = agg.end() - 1;
            }
            it->launches++;
            it->total_ms += ms;
        }
        s->timer.clear();
    });
    if (rc != PXLFLIP_OK) return rc;
    std::sort(agg.begin(), agg.end(),
              [](const pxlflip_kernel_time &a, const pxlflip_kernel_time &b) { return a.total_ms > b.total_ms; });
    uint32_t k = 0;
    for (; k < agg.size() && k < max_out; k++)
        if (out) out[k] = agg[k];
    if (n_out) *n_out = (uint32_t)agg.size();
    return PXLFLIP_OK;
}
}  // namespace

extern "C" {

int pxlflip_profile_steps(pxlflip *h, uint32_t n, pxlflip_kernel_time *out, uint32_t max_out, uint32_t *n_out) {
    if (!h) return PXLFLIP_ERR_INVALID;
    Sim *s = &h->sim;
    return profile_launches(s, [&] { s->run_steps(n); }, out, max_out, n_out);
}

int pxlflip_profile_surface(pxlflip *h, uint32_t n, pxlflip_kernel_time *out, uint32_t max_out, uint32_t *n_out) {
    if (!h) return PXLFLIP_ERR_INVALID;
    Sim *s = &h->sim;
    return profile_launches(
        s,
        [&] {
            pxl::DeviceBuffer vb(32ull * s->surf.max_verts), ib(12ull * s->surf.max_tris);
            uint32_t ctr[2];
            for (uint32_t k = 0; k < n; k++) s->surface(vb.as<float>(), 0, ib.as<uint32_t>(), 0, ctr);
        },
        out, max_out, n_out);
}

int pxlflip_autotune(pxlflip *h, uint32_t steps, uint32_t *chosen_workgroups, double *best_ms) {
    if (!h || steps == 0) return PXLFLIP_ERR_INVALID;
    Sim *s = &h->sim;
    return pxl::guarded(s, [&] {
        if (!s->initialized) s->reset();
        pxl::DevSim &d = s->d;
        // save the state that persists across steps (particles, ids, warm-start pressure)
        const size_t P = d.np, C = d.ncells;
        pxl::DeviceBuffer save(28 * P + 4 * C);
        char *sp = save.as<char>();
        float *src[6] = {d.ax, d.ay, d.az, d.avx, d.avy, d.avz};
        for (int a = 0; a < 6; a++)
            HIP_CHECK(hipMemcpyAsync(sp + 4 * P * a, src[a], 4 * P, hipMemcpyDeviceToDevice, s->stream));
        HIP_CHECK(hipMemcpyAsync(sp + 24 * P, d.aid, 4 * P, hipMemcpyDeviceToDevice, s->stream));
        HIP_CHECK(hipMemcpyAsync(sp + 28 * P, d.pressure, 4 * C, hipMemcpyDeviceToDevice, s->stream));
        const uint64_t steps_before = s->steps_done;
        const int candidates[] = {64, 128, 256, 512, 1024};
        int best = d.cg_blocks;
        double best_t = 1e300;
        for (int g : candidates) {
            d.cg_blocks = g;
            s->timer.clear();
            s->timer.on = true;
            s->run_steps(steps);
            s->timer.on = false;
            double t = 0.0;
            for (const auto &r : s->timer.recs) {
                float ms = 0.0f;
                HIP_CHECK(hipEventElapsedTime(&ms, r.a, r.b));
                t += ms;
            }
            s->timer.clear();
            if (t < best_t) {
                best_t = t;
                best = g;
            }
            for (int a = 0; a < 6; a++)
                HIP_CHECK(hipMemcpyAsync(src[a], sp + 4 * P * a, 4 * P, hipMemcpyDeviceToDevice, s->stream));
            HIP_CHECK(hipMemcpyAsync(d.aid, sp + 24 * P, 4 * P, hipMemcpyDeviceToDevice, s->stream));
            HIP_CHECK(hipMemcpyAsync(d.pressure, sp + 28 * P, 4 * C, hipMemcpyDeviceToDevice, s->stream));
            HIP_CHECK(hipStreamSynchronize(s->stream));
        }
        d.cg_blocks = best;
        s->drop_graphs();
        s->steps_done = steps_before;
        if (chosen_workgroups) *chosen_workgroups = (uint32_t)best;
        if (best_ms) *best_ms = best_t / steps;
    });
}

int pxlflip_last_solve(const pxlflip *h, pxlflip_solve_stats *st) {
    if (!h || !st) return PXLFLIP_ERR_INVALID;
    *st = h->sim.stats;
    return PXLFLIP_OK;
}

const char *pxlflip_last_error(const pxlflip *h) {
    if (!h) return pxl::g_create_error.c_str();
    return h->sim.error.c_str();
}

}  // extern "C"
