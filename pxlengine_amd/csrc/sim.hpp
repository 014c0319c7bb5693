// Host-side state of one simulation (shared by the single-sim C ABI and the batch server).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <array>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "pxlflip.h"
#include "kernels.hpp"
#include "spec.hpp"

namespace pxl {

// Error message of the last failed create/check on this thread.
extern thread_local std::string g_create_error;

struct HipError {
    hipError_t code;
    const char *what;
};

#define HIP_CHECK(expr)                                                   \
    do {                                                                  \
        hipError_t e_ = (expr);                                           \
        if (e_ != hipSuccess) throw ::pxl::HipError{e_, #expr};           \
    } while (0)

constexpr size_t ALIGN = 256;
constexpr uint32_t MAX_SOLVER_BLOCKS = 4096;
constexpr uint32_t DEFAULT_SOLVER_BLOCKS = 256;

// Optional HIP-event timing of every launch (pxlflip_profile_steps): a single-kernel phase gets
// its kernel's own start and end; a phase of several kernels runs from its first kernel's start
// to an event after its last.  Off while a graph is captured or replayed.
struct KernelTimer {
    struct Rec {
        const char *name;
        hipEvent_t a, b;
    };
    bool on = false;
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    size_t used = 0;

    hipEvent_t next() {
        if (used == pool.size()) {
            hipEvent_t e;
            HIP_CHECK(hipEventCreate(&e));
            pool.push_back(e);
        }
        return pool[used++];
    }
    template <typename F>
    void run(const char *name, hipStream_t s, F &&launch) {
        if (!on) {
            launch();
            return;
        }
        hipEvent_t a = next(), b = next();
        LaunchEvents &ev = launch_events();
        ev = LaunchEvents{a, b, 0};
        launch();  // the first kernel records a and b itself
        const int n = ev.launches;
        ev = LaunchEvents{};
        if (n == 0) {  // nothing launched
            HIP_CHECK(hipEventRecord(a, s));
            HIP_CHECK(hipEventRecord(b, s));
        } else if (n > 1) {  // several kernels: stop after the last one
            HIP_CHECK(hipEventRecord(b, s));
        }
        recs.push_back({name, a, b});
    }
    void clear() {
        recs.clear();
        used = 0;
    }
    ~KernelTimer() {
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

#define KRUN(name, call) timer.run(name, stream, [&] { call; })

struct Sim;
// Seeds a simulation's initial particles (ix..ivz) on its stream (seed.hip); flags and rank are
// device scratch of ncells words each.  Throws if the device's count differs from the host's.
void seed_on_device(Sim &s, uint32_t *flags, uint32_t *rank);

struct Sim {
    FluidSpec spec;
    pxlflip_opts opts{};
    double fps = 24.0;
    uint32_t seed = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    DevSim d{};
    void *arena = nullptr;
    size_t arena_bytes = 0;
    uint32_t iters = 80;          // RBGS sweeps or CG iterations per step
    uint64_t steps_done = 0;
    bool initialized = false;
    // false when `stream` is lent by its creator (a serving allocator thread's stream)
    bool owns_stream = true;
    hipGraphExec_t step_graph = nullptr;
    std::map<uint32_t, std::array<hipGraphExec_t, 2>> inner_graph;  // [MG-CG iterations][x64 parity]
    static constexpr uint32_t MAX_SUBSTEP_GRAPHS = 8;
    std::map<uint32_t, hipGraphExec_t> sub_graph;  // [CFL substeps]: the step at dt / substeps
    hipGraphExec_t grid_graph = nullptr, post_graph = nullptr;  // MIXED: phases around the solve
    double *host_scal = nullptr;  // pinned
    int x64_cur = 0;
    pxlflip_solve_stats stats{-1.0, 0, 0};
    std::string error;
    KernelTimer timer;
    // surface nets: parameters set at creation, arrays allocated by the first surface pass
    DevSurf surf{};
    void *surf_arena = nullptr;
    uint32_t *host_ctr = nullptr;  // pinned: the surface counters

    void ensure_surface() {
        if (surf_arena) return;
        const size_t nn = surf.nnodes, nc = surf.ncells;
        const size_t tmp = 4ull * (scan_tiles((uint32_t)std::max(nn, nc)) + 1);
        const size_t sizes[6] = {4 * nn, 4 * nn, 4 * nn, 4 * nc, 64, tmp};
        size_t total = 0;
        for (size_t b : sizes) total += (b + ALIGN - 1) / ALIGN * ALIGN;
        if (hipMalloc(&surf_arena, total) != hipSuccess) {
            surf_arena = nullptr;
            throw std::bad_alloc();
        }
        char *cur = static_cast<char *>(surf_arena);
        void **dst[6] = {reinterpret_cast<void **>(&surf.density), reinterpret_cast<void **>(&surf.phi),
                         reinterpret_cast<void **>(&surf.qoff), reinterpret_cast<void **>(&surf.cell_vert),
                         reinterpret_cast<void **>(&surf.counters), reinterpret_cast<void **>(&surf.scan_tmp)};
        for (int q = 0; q < 6; q++) {
            *dst[q] = cur;
            cur += (sizes[q] + ALIGN - 1) / ALIGN * ALIGN;
        }
        HIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&host_ctr), 16, hipHostMallocDefault));
    }

    // One surface pass on the current particles into device buffers; returns the counters.
    void surface(float *verts, uint32_t vbase, uint32_t *indices, uint32_t ibase, uint32_t ctr[2]) {
        ensure_surface();
        for (int st = 0; st < SURF_STAGES; st++)
            KRUN(SURF_STAGE_NAMES[st], launch_surface_stage(st, surf, d.ax, d.ay, d.az, d.np, verts, vbase, indices, ibase, stream));
        HIP_CHECK(hipMemcpyAsync(host_ctr, surf.counters, 8, hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
        ctr[0] = host_ctr[0];
        ctr[1] = host_ctr[1];
    }

    // Graphs bake kernel arguments (grid shapes, tunables); drop them after a change.
    void drop_graphs() {
        if (step_graph) (void)hipGraphExecDestroy(step_graph);
        step_graph = nullptr;
        for (hipGraphExec_t *g : {&grid_graph, &post_graph}) {
            if (*g) (void)hipGraphExecDestroy(*g);
            *g = nullptr;
        }
        for (auto &e : inner_graph)
            for (hipGraphExec_t g : e.second)
                if (g) (void)hipGraphExecDestroy(g);
        inner_graph.clear();
        for (auto &e : sub_graph)
            if (e.second) (void)hipGraphExecDestroy(e.second);
        sub_graph.clear();
    }

    ~Sim() {
        drop_graphs();
        if (host_scal) (void)hipHostFree(host_scal);
        if (host_ctr) (void)hipHostFree(host_ctr);
        if (surf_arena) (void)hipFree(surf_arena);
        if (arena) (void)hipFree(arena);
        if (stream && owns_stream) (void)hipStreamDestroy(stream);
    }

    // ---- allocation
    using Plan = std::vector<std::pair<void **, size_t>>;
    // Every device array of the simulation, in arena order.
    Plan plan() {
        const size_t P = d.np, F = d.nfaces, C = d.ncells, K = d.nkeys;
        const size_t nb = MAX_SOLVER_BLOCKS;  // partials sized for the largest grid the tuner may pick
        Plan plan;
        auto add = [&](auto **p, size_t bytes) { plan.emplace_back(reinterpret_cast<void **>(p), bytes); };
        for (float **p : {&d.ax, &d.ay, &d.az, &d.avx, &d.avy, &d.avz}) add(p, 4 * P);
        add(&d.aid, 4 * P);
        for (float **p : {&d.bx, &d.by, &d.bz, &d.bvx, &d.bvy, &d.bvz}) add(p, 4 * P);
        add(&d.bid, 4 * P);
        add(&d.key, 4 * P);
        add(&d.slot, 4 * P);
        for (float **p : {&d.ix, &d.iy, &d.iz, &d.ivx, &d.ivy, &d.ivz}) add(p, 4 * P);
        add(&d.vel, 4 * F);
        add(&d.vel_old, 4 * F);
        add(&d.box, 8ull * BOX_ENTRIES * d.ntiles);
        add(&d.tile_list, 4ull * d.ntiles);
        add(&d.tile_count, 64);
        add(&d.tile_act, d.ntiles);
        add(&d.band_list, 4ull * d.ntiles);
        add(&d.band_count, 64);
        add(&d.count, 4 * K);
        add(&d.offset, 4 * (K + 1));
        add(&d.type, C);
        add(&d.sten, 2 * C);
        add(&d.pressure, 4 * C);
        add(&d.div, 4 * C);
        if (d.n_obst) {
            add(&d.solid_bits, 4 * ((C + 31) / 32));
            add(&d.obst_id, 4 * C);
            add(&d.obst, 16ull * d.n_obst);
        }
        add(&d.fl_rows, 4ull * fl_row_slots(d));
        add(&d.fl_rowoff, 4ull * fl_row_slots(d));
        add(&d.fl_list, 4 * C);
        add(&d.fl_word, 4 * C);
        if (d.split_colours) {
            add(&d.fn_list, 4 * C);
            add(&d.fn_word, 4 * C);
        }
        add(&d.fl_count, 64);
        if (opts.solver == PXLFLIP_SOLVER_RBGS) {
            const size_t nl = (size_t)rbgs_cores(d) * rbgs_list_slots();
            add(&d.rb_list, 4 * nl);
            add(&d.rb_b, 4 * nl);
            add(&d.rb_count, 32ull * rbgs_cores(d));
        }
        add(&d.fl_tmp, 4ull * (scan_tiles(fl_row_slots(d)) + 1));
        // the MIXED inner solve stores its vectors colour-split (2 hsplit entries >= C)
        const size_t V = opts.solver == PXLFLIP_SOLVER_MIXED ? std::max<size_t>(C, 2ull * d.hsplit) : C;
        for (float **p : {&d.cg_r0, &d.cg_r1, &d.cg_w0, &d.cg_w1, &d.cg_s0, &d.cg_s1, &d.cg_p}) add(p, 4 * V);
        if (opts.solver == PXLFLIP_SOLVER_MIXED) {
            for (int l = 1; l < d.mg_levels; l++) {
                const size_t cl = (size_t)d.mg[l].nx * d.mg[l].ny * d.mg[l].nz;
                const size_t sl = 2ull * d.mg[l].hsplit;  // colour-split arrays (>= cl)
                add(&d.mg[l].type, cl);
                add(&d.mg[l].sten, 2 * sl);
                add(&d.mg[l].r, 4 * sl);
                add(&d.mg[l].z, 4 * sl);
                add(&d.mg[l].list, 4 * cl);
                add(&d.mg[l].nlist, 4 * cl);
                add(&d.mg[l].count, 64);
                add(&d.mg[l].chunk, 8 * ((cl + 63) / 64));
            }
            add(&d.e, 4 * V);
            add(&d.fls_list, 4 * C);
            add(&d.fns_list, 4 * C);
            add(&d.x64a, 8 * C);
            add(&d.x64b, 8 * C);
            add(&d.r64, 8 * C);
        }
        add(&d.part32, 4 * 4 * nb);
        add(&d.part64, 8 * 4 * nb);
        add(&d.scal, 8 * 2 * SC_COUNT);
        add(&d.ctr, 64);
        add(&d.scan_tmp, 4 * (scan_tiles(K) + 1));
        return plan;
    }
    static size_t plan_total(const Plan &plan) {
        size_t total = 0;
        for (auto &e : plan) total += (e.second + ALIGN - 1) / ALIGN * ALIGN;
        return total;
    }

    // The arena (zeroed), the obstacle arrays and the initial particles, seeded on the device.
    void allocate() {
        Plan plan = this->plan();
        const size_t total = plan_total(plan);
        void *base = nullptr;
        hipError_t err = hipMalloc(&base, total);
        if (err != hipSuccess) throw std::bad_alloc();
        arena = base;
        arena_bytes = total;
        char *cur = static_cast<char *>(base);
        for (auto &e : plan) {
            *e.first = cur;
            cur += (e.second + ALIGN - 1) / ALIGN * ALIGN;
        }
        d.norms = d.scal + SC_RR;
        d.cg_x = d.cg_s1;
        d.cg_rt = d.cg_w1;
        if (!d.split_colours) {  // the list is already in cell order
            d.fn_list = d.fl_list;
            d.fn_word = d.fl_word;
        }
        HIP_CHECK(hipMemsetAsync(arena, 0, arena_bytes, stream));
        // the end of the last tile's particle run (the scan fills offset[0 .. nkeys-1])
        const uint32_t np32 = d.np;
        HIP_CHECK(hipMemcpyAsync(d.offset + d.nkeys, &np32, 4, hipMemcpyHostToDevice, stream));
        // obstacles (extension): cell ids, their bit mask and the spheres
        std::vector<uint32_t> ids, bits;
        std::vector<float> spheres;
        if (d.n_obst) {
            obstacle_cells(spec, ids);
            bits.assign(((size_t)d.ncells + 31) / 32, 0u);
            for (size_t c = 0; c < ids.size(); c++)
                if (ids[c]) bits[c >> 5] |= 1u << (c & 31);
            for (const auto &o : spec.obstacles) spheres.insert(spheres.end(), {o.center[0], o.center[1], o.center[2], o.radius});
            HIP_CHECK(hipMemcpyAsync(d.obst_id, ids.data(), 4 * ids.size(), hipMemcpyHostToDevice, stream));
            HIP_CHECK(hipMemcpyAsync(d.solid_bits, bits.data(), 4 * bits.size(), hipMemcpyHostToDevice, stream));
            HIP_CHECK(hipMemcpyAsync(d.obst, spheres.data(), 4 * spheres.size(), hipMemcpyHostToDevice, stream));
        }
        HIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&host_scal), 8 * (SC_COUNT + 1), hipHostMallocDefault));
        // initial particles (the reset source), with two solver vectors as scratch, cleared after
        seed_on_device(*this, reinterpret_cast<uint32_t *>(d.cg_r0), reinterpret_cast<uint32_t *>(d.cg_r1));
        launch_fill_u32(reinterpret_cast<uint32_t *>(d.cg_r0), (uint32_t)d.ncells, 0u, stream);
        launch_fill_u32(reinterpret_cast<uint32_t *>(d.cg_r1), (uint32_t)d.ncells, 0u, stream);
        HIP_CHECK(hipStreamSynchronize(stream));
    }

    // ---- step pieces
    void enqueue_grid_phase() {
        KRUN("clear_grid", launch_clear_grid(d, stream));
        KRUN("bin_count", launch_bin_count(d, stream));
        KRUN("scan_offsets", launch_scan_offsets(d, stream));
        KRUN("tile_list", launch_tile_list(d, stream));
        KRUN("band", launch_band(d, stream));
        KRUN("scatter", launch_scatter(d, stream));
        KRUN("p2g", launch_p2g(d, stream));
        KRUN("faces", launch_faces(d, stream));
        KRUN("cells", launch_cells(d, stream));
        // the FLUID-cell list feeds the CG and MIXED solvers; the reference mode's blocked sweeps
        // list each core's cells themselves
        if (opts.solver != PXLFLIP_SOLVER_RBGS) KRUN("fluid_list", launch_fluid_list(d, stream));
        if (d.mg_levels) KRUN("mg_types", launch_mg_types(d, stream));
    }
    void enqueue_fixed_pressure() {
        if (opts.solver == PXLFLIP_SOLVER_RBGS) {
            KRUN("rbgs", launch_rbgs_blocked(d, iters, stream));
        } else {
            KRUN("cg_start", launch_cg_start(d, false, d.pressure, false, stream));
            for (uint32_t it = 0; it < iters; it++) KRUN("cg_iter", launch_cg_iter(d, false, d.pressure, it, stream));
        }
    }
    void enqueue_post() {
        KRUN("project", launch_project(d, stream));
        KRUN("g2p", launch_g2p(d, stream));
    }
    void enqueue_fixed_step() {
        enqueue_grid_phase();
        enqueue_fixed_pressure();
        enqueue_post();
    }
    void enqueue_inner(int cur, uint32_t inner) {
        KRUN("mgcg_start", launch_mgcg_start(d, d.e, stream));
        for (uint32_t it = 0; it < inner; it++)
            KRUN("mgcg_iter", launch_mgcg_iter(d, d.e, it, it + 1 == inner, stream));
        KRUN("mixed_update", launch_mixed_update(d, x64(cur), x64(cur ^ 1), stream));
    }
    void enqueue_grid_init() {
        enqueue_grid_phase();
        KRUN("mixed_init", launch_mixed_init(d, x64(0), stream));
    }
    template <typename F>
    hipGraphExec_t capture(F &&fn) {
        hipGraph_t g = nullptr;
        hipGraphExec_t ex = nullptr;
        HIP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        fn();
        HIP_CHECK(hipStreamEndCapture(stream, &g));
        hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIP_CHECK(e);
        return ex;
    }

    double *x64(int k) const { return k == 0 ? d.x64a : d.x64b; }

    void read_scalars() {
        HIP_CHECK(hipMemcpyAsync(host_scal, d.scal, 8 * SC_COUNT, hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
    }

    // Runs `fn` as a graph captured on first use into *g (when graphs are on), else eagerly.
    template <typename F>
    void replay(hipGraphExec_t *g, F &&fn) {
        if (opts.use_graph && !timer.on) {
            if (!*g) *g = capture(fn);
            HIP_CHECK(hipGraphLaunch(*g, stream));
        } else {
            fn();
        }
    }

    // Adaptive pass length (opts.inner_max > 0): the iterations that take the relative
    // residual from `rel0` to a third of the tolerance at the last measured per-iteration
    // reduction rho_est (inner_iterations before any measurement), within [2, inner_max].  A
    // pass is planned for at most a 1e-5 reduction: the FP32 inner solve does not reliably get
    // further, and the next refinement pass takes over from there.
    double rho_est = 0.0;
    uint32_t adaptive_iterations(double rel0) const {
        if (!(rho_est > 0.0)) return std::min(opts.inner_iterations, opts.inner_max);
        const double target = std::max(opts.tolerance / 3.0 / rel0, 1e-5);
        const double need = std::log(target) / std::log(rho_est);
        const double k = std::ceil(need);
        if (!(k >= 2.0)) return std::min<uint32_t>(2u, opts.inner_max);
        return k >= (double)opts.inner_max ? opts.inner_max : (uint32_t)k;
    }

    // Mixed-precision pressure solve after enqueue_grid_init: FP64 refinement driven from
    // the host, one residual read-back per refinement pass.
    void mixed_pressure() {
        x64_cur = 0;
        read_scalars();
        const double bb = host_scal[SC_BB];
        double rr = host_scal[SC_RR];
        auto rel = [&](double r) { return bb > 0.0 ? std::sqrt(r / bb) : std::sqrt(r); };
        uint32_t outer = 0, inner_total = 0;
        while (rel(rr) > opts.tolerance && outer < opts.max_outer) {
            const double before = rel(rr);
            const uint32_t inner = opts.inner_max ? adaptive_iterations(before) : pass_iterations(opts.inner_iterations, outer);
            const int cur = x64_cur;
            replay(&inner_graph[inner][cur], [this, cur, inner] { enqueue_inner(cur, inner); });
            x64_cur ^= 1;
            read_scalars();
            rr = host_scal[SC_RR];
            outer++;
            inner_total += inner;
            if (opts.inner_max) {
                // per-iteration reduction of this pass (a pass that hit the FP32 floor reads slower
                // than the solver converges, which only makes later passes longer)
                const double ratio = before > 0.0 ? rel(rr) / before : 1.0;
                rho_est = ratio > 0.0 && ratio < 1.0 ? std::pow(ratio, 1.0 / inner) : 0.95;
                rho_est = std::min(0.95, std::max(0.02, rho_est));
            }
        }
        stats.rel_residual = rel(rr);
        stats.outer_iterations = outer;
        stats.inner_iterations = inner_total;
    }

    // CFL cap (extension): substeps of the next step from the particles' largest speed.
    uint32_t last_substeps = 1;
    uint32_t cfl_step_count() {
        launch_vmax2(d, d.ctr, stream);
        HIP_CHECK(hipMemcpyAsync(host_scal + SC_COUNT, d.ctr, 4, hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
        uint32_t bits;
        std::memcpy(&bits, host_scal + SC_COUNT, 4);
        float v2;
        std::memcpy(&v2, &bits, 4);
        return cfl_substeps(v2, d.dt, spec.cfl, d.h);
    }
    // One step of the current d.dt (graphs when allowed, else eager).
    // `sg`: the graph slot of the fixed-iteration step at this d.dt (the full step's by default).
    void step_once(hipGraphExec_t *sg = nullptr) {
        if (opts.solver == PXLFLIP_SOLVER_MIXED) {
            replay(&grid_graph, [this] { enqueue_grid_init(); });
            mixed_pressure();
            replay(&post_graph, [this] { enqueue_post(); });
        } else {
            const bool graph = opts.use_graph && !timer.on;
            hipGraphExec_t *gp = sg ? sg : &step_graph;
            if (graph && !*gp) *gp = capture([this] { enqueue_fixed_step(); });
            if (graph) HIP_CHECK(hipGraphLaunch(*gp, stream));
            else enqueue_fixed_step();
            stats.rel_residual = -1.0;
            stats.inner_iterations = iters;
            stats.outer_iterations = 0;
        }
    }
    // With the CFL cap each step of dt runs as k substeps of dt / k (f32).  The fixed-iteration
    // solvers replay one graph per substep count up to MAX_SUBSTEP_GRAPHS (their kernels hold
    // dt / k); larger counts and the MIXED solve (whose graphs hold the full dt) launch eagerly.
    void run_capped_steps(uint32_t n) {
        for (uint32_t s = 0; s < n; s++) {
            const uint32_t k = cfl_step_count();
            // faces away from the liquid hold 0 + g dt from the last pass that covered them:
            // when dt changes, the next two steps cover every tile again
            if (k != last_substeps) mark_all_tiles();
            last_substeps = k;
            if (k == 1) {
                step_once();
            } else {
                const float dt0 = d.dt;
                const int g = opts.use_graph;
                d.dt = dt0 / (float)k;
                hipGraphExec_t *sg = nullptr;
                if (opts.solver != PXLFLIP_SOLVER_MIXED && k <= MAX_SUBSTEP_GRAPHS) sg = &sub_graph[k];
                else opts.use_graph = 0;
                try {
                    for (uint32_t q = 0; q < k; q++) step_once(sg);
                } catch (...) {
                    d.dt = dt0;
                    opts.use_graph = g;
                    throw;
                }
                d.dt = dt0;
                opts.use_graph = g;
            }
            steps_done++;
        }
        HIP_CHECK(hipStreamSynchronize(stream));
    }

    void run_steps(uint32_t n) {
        if (n == 0) return;
        if (spec.cfl > 0.0f) {
            run_capped_steps(n);
            return;
        }
        if (opts.solver == PXLFLIP_SOLVER_MIXED) {
            for (uint32_t s = 0; s < n; s++) {
                replay(&grid_graph, [this] { enqueue_grid_init(); });
                mixed_pressure();
                replay(&post_graph, [this] { enqueue_post(); });
                steps_done++;
            }
        } else {
            const bool graph = opts.use_graph && !timer.on;
            if (graph && !step_graph) step_graph = capture([this] { enqueue_fixed_step(); });
            for (uint32_t s = 0; s < n; s++) {
                if (graph) HIP_CHECK(hipGraphLaunch(step_graph, stream));
                else enqueue_fixed_step();
            }
            steps_done += n;
            stats.rel_residual = -1.0;
            stats.inner_iterations = iters;
            stats.outer_iterations = 0;
        }
        HIP_CHECK(hipStreamSynchronize(stream));
    }

    // The next two steps walk every tile (the band then covers the whole grid), so every face
    // and cell holds this step's values again after a reset or a field write.
    void mark_all_tiles() { HIP_CHECK(hipMemsetAsync(d.tile_act, 3, d.ntiles, stream)); }

    void reset() {
        launch_reset_particles(d, stream);
        HIP_CHECK(hipMemsetAsync(d.pressure, 0, 4 * (size_t)d.ncells, stream));
        mark_all_tiles();
        HIP_CHECK(hipStreamSynchronize(stream));
        steps_done = 0;
        initialized = true;
    }
};

// Host read-back shared by the single-simulation and batch entry points (flip.hip).
std::string field_check(const Sim &s, int field);
void read_field_to_host(Sim &s, int field, void *out);
void read_bins_to_host(Sim &s, uint32_t *cell_of, uint32_t *counts, uint32_t *offsets, uint32_t *sorted_ids);

inline int fail(Sim *s, int code, const std::string &msg) {
    if (s) s->error = msg;
    return code;
}

inline int device_fail(Sim *s, const HipError &e) {
    return fail(s, PXLFLIP_ERR_DEVICE, std::string("HIP error in ") + e.what + ": " + hipGetErrorString(e.code));
}

// Runs body() with the handle's device current; maps exceptions to status codes.
template <typename F>
inline int guarded(Sim *s, F &&body) {
    if (!s) return PXLFLIP_ERR_INVALID;
    try {
        HIP_CHECK(hipSetDevice(s->device));
        body();
        return PXLFLIP_OK;
    } catch (const HipError &e) {
        return device_fail(s, e);
    } catch (const std::bad_alloc &) {
        return fail(s, PXLFLIP_ERR_OOM, "out of memory");
    } catch (const std::exception &e) {
        return fail(s, PXLFLIP_ERR_INTERNAL, e.what());
    }
}

struct DeviceBuffer {
    void *p = nullptr;
    explicit DeviceBuffer(size_t bytes) {
        if (hipMalloc(&p, bytes ? bytes : 4) != hipSuccess) throw std::bad_alloc();
    }
    ~DeviceBuffer() {
        if (p) (void)hipFree(p);
    }
    template <typename T>
    T *as() const { return static_cast<T *>(p); }
};

// Parse + validate + seed + derive the device parameters; no device work.  Returns 0 or an
// error code with g_create_error set.
int prepare(const char *fluid_json, double fps, uint32_t seed, const pxlflip_opts *opts, Sim &s,
            std::vector<float> *x0, std::vector<float> *v0, const FluidSpec *parsed = nullptr);
void fill_info(const Sim &s, pxlflip_info *info);

}  // namespace pxl
