// Batched execution of many simulations and the serving loop around it
// (pxlflip_server_* in include/pxlflip.h).
//
// Batch: the resident simulations are stepped together.  Every kernel of the step is
// launched once for the whole batch as a persistent grid over a device-side work list of
// (simulation, workgroup) entries (kernels.hpp, WorkList); a simulation's entries are
// exactly the workgroups of its own single-sim launch, so it computes bitwise what it
// computes alone.  The step runs as three graph replays and a host-driven refinement loop:
//   grid graph:   clear .. fluid list, multigrid types, FP64 init (x64, r64, norms)
//   inner graph:  one refinement pass (MG-CG + FP64 update) for the simulations still
//                 solving; replayed until every one of them reaches the tolerance (the norms
//                 of all simulations come back in one copy per pass, and the solve work
//                 lists are rewritten to hold only the unconverged ones)
//   post graph:   project, g2p.
// Work lists live in one device buffer laid out by domain; the layout (and with it the
// graphs, which bake the list pointers) only changes when a domain outgrows its capacity
// or the deepest multigrid level changes.
//
// Server: a device thread admits queued requests first-in first-out while their HBM
// footprint fits the budget, allocates them, adds them to the batch, runs batch steps and
// retires simulations at the step boundary after their last step.
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

#include "common.hpp"
#include "sim.hpp"

namespace pxl {

uint32_t work_blocks(const DevSim &S, int d) {
    switch (d) {
        case WD_SIM:
        case WD_A_SIM: return 1;
        case WD_PART: return particle_blocks(S);
        case WD_SCAN_KEYS: return scan_tiles(S.nkeys);
        case WD_TILES: return tile_blocks(S);
        case WD_TILEGRID: return tile_grid_blocks(S);
        case WD_SCAN_ROWS: return scan_tiles(fl_row_slots(S));
        case WD_ROWS: return fluid_row_blocks(S);
        case WD_A_SOLVER: return (uint32_t)S.cg_blocks;
        default: break;
    }
    if (d >= WD_MGTYPES && d < WD_MGTYPES + MG_MAX) {
        const int l = d - WD_MGTYPES;
        return l >= 1 && l < S.mg_levels ? mg_cell_blocks(S, l) : 0;
    }
    if (d >= WD_A_LVL && d < WD_A_LVL + MG_MAX) {
        const int l = d - WD_A_LVL;
        return l < S.mg_mid ? level_blocks(S.cg_blocks, l) : 0;
    }
    return 0;
}

// A work-list entry holds a member's workgroup in its low WORK_SIM_SHIFT bits, so no domain of a
// batched simulation may need more workgroups than that field counts (the particle domain
// reaches it first: 2^20 workgroups of 256 particles).
bool work_fits(const DevSim &S, std::string *why) {
    for (int d = 0; d < WD_COUNT; d++) {
        const uint32_t nb = work_blocks(S, d);
        if (nb > WORK_BLOCK_MASK + 1u) {
            if (why)
                *why = "work domain " + std::to_string(d) + " needs " + std::to_string(nb) +
                       " workgroups, a batch entry addresses " + std::to_string(WORK_BLOCK_MASK + 1u);
            return false;
        }
    }
    return true;
}

namespace {

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

template <typename T>
void host_alloc(T **p, size_t n) {
    if (*p) (void)hipHostFree(*p);
    *p = nullptr;
    HIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(p), sizeof(T) * (n ? n : 1), hipHostMallocDefault));
}
template <typename T>
void dev_alloc(T **p, size_t n) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    if (hipMalloc(reinterpret_cast<void **>(p), sizeof(T) * (n ? n : 1)) != hipSuccess) throw std::bad_alloc();
}

// What a profiled batch step needs to price each launch: per member, the FLUID cells of each
// level (level 0 = the fluid list) and the active P2G tiles of the step.
struct MemberCounts {
    uint32_t fluid[MG_MAX] = {};
    uint32_t tiles = 0, band = 0;
};

// Algorithmic bytes of one simulation's share of a launch: every logical array the kernel
// must touch once, at its element width (fp32/int32 4 B, fp64 8 B, cell type 1 B, stencil
// word 2 B, particle position + velocity 24 B, id 4 B).  P particles, F faces, C cells, K
// count keys, Cf / c_l FLUID cells of level 0 / l, N_l cells of level l, t active tiles (a
// tile's P2G box is 2700 packed 8-byte sums).
double share_bytes(const char *k, const DevSim &d, const MemberCounts &m, int l) {
    const double P = d.np, F = d.nfaces, C = d.ncells, K = d.nkeys, Cf = m.fluid[0], t = m.tiles;
    const double box = 8.0 * BOX_ENTRIES;
    auto cells_of = [&](int q) { return q == 0 ? C : (double)d.mg[q].nx * d.mg[q].ny * d.mg[q].nz; };
    const double cl = l < MG_MAX ? m.fluid[l] : 0, cl1 = l + 1 < MG_MAX ? m.fluid[l + 1] : 0;
    const std::string n(k);
    if (n == "clear_grid") return 4.0 * TILE_CELLS * t + 4.0 * fl_row_slots(d);  // t: about the last step's tiles
    if (n == "bin_count") return 20 * P + 4 * C;
    const double rows = fl_row_slots(d);
    if (n == "scan_reduce") return 4 * (l == 0 ? K : rows);
    if (n == "scan_down") return 8 * (l == 0 ? K : rows);
    if (n == "tile_list") return 8.0 * d.ntiles;
    if (n == "scatter") return 68 * P;
    if (n == "p2g") return 24 * P + box * t;
    // band passes: the fraction of the grid in the step's band of tiles
    const double bf = d.ntiles ? std::min(1.0, (double)m.band / (double)d.ntiles) : 1.0;
    if (n == "band") return 28.0 * d.ntiles;
    if (n == "faces") return 8 * F * bf + box * t;
    if (n == "cells") return bf * (4 * F + 15 * C + 4 * C) - 4 * Cf;
    if (n == "fluid_compact") return 8 * rows + 2 * C * bf + 8 * Cf;
    if (n == "mg_types") return cells_of(l - 1) + cells_of(l);
    if (n == "mg_sten") return 7 * cells_of(l) + 8 * std::ceil(cells_of(l) / 64.0);
    if (n == "mg_scan") {  // every level's chunk counts read twice and written once
        double b = 0;
        for (int q = 1; q < d.mg_levels; q++) b += 12 * std::ceil(cells_of(q) / 64.0);
        return b;
    }
    if (n == "mg_compact") return 2 * cells_of(l) + 8 * std::ceil(cells_of(l) / 64.0) + 8 * cl;
    if (n == "mixed_init") return 32 * Cf;
    if (n == "cg_residual") return 24 * Cf;
    if (n == "mg_pre0") return 16 * cl;
    if (n == "mg_sweep") return (l == 0 ? 14 : 13) * cl;
    if (n == "mg_restrict") return 64 * cl1;
    if (n == "mg_post") return 13 * cl;
    if (n == "mg_tail") {  // the mid levels' sweeps, restrictions and posts, then the LDS tail
        double b = 0;
        for (int q = d.mg_mid; q < d.mg_tail; q++)
            b += (16 + 13 + 13 * (4 * MG_COARSE_ROUNDS - 3)) * (double)m.fluid[q] + 64.0 * m.fluid[q + 1];
        for (int q = d.mg_tail; q < d.mg_levels; q++) b += 2 * cells_of(q);
        return b + 8 * cells_of(d.mg_tail);
    }
    if (n == "mgcg_spmv") return 20 * Cf;
    if (n == "mgcg_update") return 44 * Cf;
    if (n == "mixed_update") return 44 * Cf;
    if (n == "project") return bf * (8 * F + 5 * C);
    if (n == "g2p") return 56 * P + 8 * F;
    return 0;
}

class Batch {
public:
    explicit Batch(const pxlflip_opts &o) : opts_(o) {
        HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        dev_alloc(&d_counts_, WD_COUNT);
        host_alloc(&h_counts_, WD_COUNT);
        dev_alloc(&d_band_count_, 16);
        std::fill(cap_, cap_ + WD_COUNT, 0);
        std::fill(off_, off_ + WD_COUNT, 0);
    }
    ~Batch() {
        if (stream_) (void)hipStreamSynchronize(stream_);
        drop_graphs();
        for (void *p : {(void *)d_sims_, (void *)d_entries_, (void *)d_counts_, (void *)d_norms_, (void *)d_slots_,
                        (void *)d_band_, (void *)d_band_count_})
            if (p) (void)hipFree(p);
        for (void *p : {(void *)h_sims_, (void *)h_entries_, (void *)h_counts_, (void *)h_norms_})
            if (p) (void)hipHostFree(p);
        for (auto &r : recs_) {
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
        if (stream_) (void)hipStreamDestroy(stream_);
    }

    hipStream_t stream() const { return stream_; }
    size_t size() const { return members_.size(); }
    const std::vector<Sim *> &members() const { return members_; }
    uint64_t graph_captures() const { return captures_; }
    double capture_seconds() const { return capture_s_; }
    uint64_t rebuilds() const { return rebuilds_; }

    void add(Sim *s) {
        members_.push_back(s);
        passes_.push_back(0);
        dirty_ = true;
    }
    void remove(Sim *s) {
        HIP_CHECK(hipStreamSynchronize(stream_));
        for (size_t i = 0; i < members_.size(); i++) {
            if (members_[i] == s) {
                members_.erase(members_.begin() + (long)i);
                passes_.erase(passes_.begin() + (long)i);
                break;
            }
        }
        dirty_ = true;
    }

    // One step of every member.
    void step() {
        if (members_.empty()) return;
        if (dirty_) rebuild();
        const size_t n = members_.size();
        std::vector<uint32_t> active(n);
        for (size_t i = 0; i < n; i++) active[i] = (uint32_t)i;
        set_active(active);
        replay(&g_grid_, [&] {
            enqueue_grid_phase();
            if (profiling_) read_counts();
            run("mixed_init", -1, [&] { batch_mixed_init(lists_, stream_); });
            run("norms", -1, [&] { batch_norms(lists_, 2, stream_); });
        });
        read_norms();
        std::vector<double> bb(n), rel(n);
        for (size_t i = 0; i < n; i++) {
            bb[i] = h_norms_[2 * i + 1];
            rel[i] = rel_of(h_norms_[2 * i], bb[i]);
            passes_[i] = 0;
        }
        std::vector<uint32_t> next;
        for (uint32_t i : active)
            if (rel[i] > opts_.tolerance) next.push_back(i);
        uint32_t pass = 0;
        while (!next.empty() && pass < opts_.max_outer) {
            if (next.size() != active.size()) set_active(next);
            active = next;
            const int cur = (int)(pass & 1);
            const uint32_t inner = pass_iterations(opts_.inner_iterations, pass);
            replay(&g_inner_[inner][cur], [&] { enqueue_inner_pass(cur, inner); });
            read_norms();
            next.clear();
            for (uint32_t i : active) {
                passes_[i]++;
                rel[i] = rel_of(h_norms_[2 * i], bb[i]);
                if (rel[i] > opts_.tolerance) next.push_back(i);
            }
            pass++;
        }
        replay(&g_post_, [&] {
            run("project", -1, [&] { batch_project(lists_, stream_); });
            run("g2p", -1, [&] { batch_g2p(lists_, stream_); });
        });
        for (size_t i = 0; i < n; i++) {
            Sim *s = members_[i];
            s->steps_done++;
            s->stats.rel_residual = rel[i];
            s->stats.outer_iterations = passes_[i];
            uint32_t total = 0;
            for (uint32_t q = 0; q < passes_[i]; q++) total += pass_iterations(opts_.inner_iterations, q);
            s->stats.inner_iterations = total;
            s->x64_cur = (int)(passes_[i] & 1);
        }
    }

    void sync() { HIP_CHECK(hipStreamSynchronize(stream_)); }

    // Profiling: every launch eager, timed by HIP events recorded from its own dispatch on the
    // batch stream, and priced with share_bytes over the simulations of its work list.
    void set_profiling(bool on) {
        profiling_ = on;
        recs_used_ = 0;
        agg_.clear();
    }
    struct Agg {
        uint32_t launches = 0;
        double ms = 0, bytes = 0;
    };
    // Per-kernel totals of the profiled steps so far (in first-launch order).
    std::vector<std::pair<std::string, Agg>> profile() {
        HIP_CHECK(hipStreamSynchronize(stream_));
        collect();
        return agg_;
    }

private:
    static double rel_of(double rr, double bb) { return bb > 0.0 ? std::sqrt(rr / bb) : std::sqrt(rr); }

    // ---- launch sequences (the same order as the single-sim step)
    void enqueue_grid_phase() {
        grid_start_ = recs_used_;
        run("clear_grid", -1, [&] { batch_clear_grid(lists_, stream_); });
        run("bin_count", -1, [&] { batch_bin_count(lists_, stream_); });
        run("scan_reduce", 0, [&] { batch_scan_reduce(lists_, 0, stream_); });
        run("scan_tiles", 0, [&] { batch_scan_tiles(lists_, 0, stream_); });
        run("scan_down", 0, [&] { batch_scan_down(lists_, 0, stream_); });
        run("tile_list", -1, [&] { batch_tile_list(lists_, stream_); });
        run("band", -1, [&] { batch_band(lists_, stream_); });
        run("scatter", -1, [&] { batch_scatter(lists_, stream_); });
        run("p2g", -1, [&] { batch_p2g(lists_, stream_); });
        run("faces", -1, [&] { batch_faces(lists_, stream_); });
        run("cells", -1, [&] { batch_cells(lists_, stream_); });
        run("scan_reduce", 1, [&] { batch_scan_reduce(lists_, 1, stream_); });
        run("scan_tiles", 1, [&] { batch_scan_tiles(lists_, 1, stream_); });
        run("scan_down", 1, [&] { batch_scan_down(lists_, 1, stream_); });
        run("fluid_compact", -1, [&] { batch_fluid_compact(lists_, stream_); });
        for (int l = 1; l < max_levels_; l++) {
            run("mg_types", l, [&] { batch_mg_types(lists_, l, stream_); });
            run("mg_sten", l, [&] { batch_mg_sten(lists_, l, stream_); });
        }
        if (max_levels_ > 1) run("mg_scan", -1, [&] { batch_mg_scan(lists_, stream_); });
        for (int l = 1; l < max_levels_; l++) run("mg_compact", l, [&] { batch_mg_compact(lists_, l, stream_); });
    }
    void enqueue_vcycle() {
        // the same sequence as the single-simulation V-cycle (mg.hip): two smoothing rounds on
        // level 0, MG_COARSE_ROUNDS on the coarser list levels
        for (int l = 0; l < max_mid_; l++) {
            const int rounds = l == 0 ? 2 : MG_COARSE_ROUNDS;
            run("mg_pre0", l, [&] { batch_mg_pre0(lists_, l, stream_); });
            for (int q = 1; q < rounds; q++) {
                run("mg_sweep", l, [&] { batch_mg_sweep(lists_, l, 0, stream_); });
                run("mg_sweep", l, [&] { batch_mg_sweep(lists_, l, 1, stream_); });
            }
            run("mg_restrict", l, [&] { batch_mg_restrict(lists_, l, stream_); });
        }
        run("mg_tail", -1, [&] { batch_mg_tail(lists_, stream_); });
        for (int l = max_mid_ - 1; l >= 0; l--) {
            const int rounds = l == 0 ? 2 : MG_COARSE_ROUNDS;
            run("mg_post", l, [&] { batch_mg_post(lists_, l, stream_); });
            run("mg_sweep", l, [&] { batch_mg_sweep(lists_, l, 0, stream_); });
            for (int q = 1; q < rounds; q++) {
                run("mg_sweep", l, [&] { batch_mg_sweep(lists_, l, 1, stream_); });
                run("mg_sweep", l, [&] { batch_mg_sweep(lists_, l, 0, stream_); });
            }
        }
    }
    // One refinement pass for the solving members: MG-CG from e = 0 on float(r64), then the
    // FP64 update x64[cur ^ 1] = x64[cur] + e and the residual norm.
    void enqueue_inner_pass(int cur, uint32_t inner) {
        run("cg_residual", -1, [&] { batch_cg_residual(lists_, stream_); });
        enqueue_vcycle();
        run("mgcg_spmv", -1, [&] { batch_mgcg_spmv(lists_, 0, 1, stream_); });
        for (uint32_t it = 0; it < inner; it++) {
            run("mgcg_update", -1, [&] { batch_mgcg_update(lists_, it, it + 1 == inner ? 1 : 0, stream_); });
            if (it + 1 == inner) break;
            enqueue_vcycle();
            run("mgcg_spmv", -1, [&] { batch_mgcg_spmv(lists_, (it + 1) & 1, 0, stream_); });
        }
        run("mixed_update", -1, [&] { batch_mixed_update(lists_, cur, stream_); });
        run("norms", -1, [&] { batch_norms(lists_, 1, stream_); });
    }

    // ---- profiling
    struct Rec {
        const char *name;
        int level;
        hipEvent_t a, b;
        double bytes;
    };
    template <typename F>
    void run(const char *name, int level, F &&launch) {
        if (!profiling_) {
            launch();
            return;
        }
        if (recs_used_ == recs_.size()) {
            Rec r{};
            HIP_CHECK(hipEventCreate(&r.a));
            HIP_CHECK(hipEventCreate(&r.b));
            recs_.push_back(r);
        }
        Rec &r = recs_[recs_used_++];
        r.name = name;
        r.level = level;
        r.bytes = launch_bytes(name, level);
        LaunchEvents &ev = launch_events();
        ev = LaunchEvents{r.a, r.b, 0};
        launch();  // the kernel records r.a and r.b from its own dispatch
        const int n = ev.launches;
        ev = LaunchEvents{};
        if (n != 1) throw std::logic_error(std::string("profiled batch phase ") + name + " did not launch one kernel");
    }
    // The members a launch covers: the solve kernels take the active ones (per level, those
    // whose level runs as separate launches), the rest take every member.
    double launch_bytes(const char *name, int level) {
        const std::string n(name);
        const bool grid_mg = n == "mg_types" || n == "mg_sten" || n == "mg_scan" || n == "mg_compact";
        const bool solve = n.rfind("mg_", 0) == 0 ? !grid_mg :
                           (n == "mixed_init" || n == "norms" || n == "cg_residual" || n.rfind("mgcg", 0) == 0 ||
                            n == "mixed_update");
        double b = 0;
        for (uint32_t i = 0; i < members_.size(); i++) {
            if (solve && !is_active_[i]) continue;
            const DevSim &d = members_[i]->d;
            if (n == "mg_types" || n == "mg_sten" || n == "mg_compact") {
                if (level >= d.mg_levels) continue;
            } else if (n == "mg_pre0" || n == "mg_sweep" || n == "mg_restrict" || n == "mg_post") {
                if (level >= d.mg_mid) continue;
            }
            b += share_bytes(name, d, counts_[i], level < 0 ? 0 : level);
        }
        return b;
    }
    void read_counts() {
        // device -> host copies of each member's list counts (profiling only; not in graphs):
        // its count slots (fl_count = {red, all}, level l = {red, black}), tiles and band
        const size_t n = members_.size();
        counts_.assign(n, MemberCounts{});
        std::vector<uint32_t> slots(n * COUNT_SLOTS);
        if (n)
            HIP_CHECK(hipMemcpyAsync(slots.data(), d_slots_, 4 * n * COUNT_SLOTS, hipMemcpyDeviceToHost, stream_));
        for (size_t i = 0; i < n; i++) {
            const DevSim &d = members_[i]->d;
            HIP_CHECK(hipMemcpyAsync(&counts_[i].tiles, d.tile_count, 4, hipMemcpyDeviceToHost, stream_));
            HIP_CHECK(hipMemcpyAsync(&counts_[i].band, d.band_count, 4, hipMemcpyDeviceToHost, stream_));
        }
        HIP_CHECK(hipStreamSynchronize(stream_));
        for (size_t i = 0; i < n; i++) {
            const uint32_t *c = slots.data() + i * COUNT_SLOTS;
            counts_[i].fluid[0] = c[1];
            for (int l = 1; l < members_[i]->d.mg_levels && l < MG_MAX; l++)
                counts_[i].fluid[l] = c[4 * l] + c[4 * l + 1];
        }
        // this step's grid-phase launches (recorded before its counts were known) are priced now
        for (size_t k = grid_start_; k < recs_used_; k++) recs_[k].bytes = launch_bytes(recs_[k].name, recs_[k].level);
    }
    void collect() {
        // events of the pending records -> per-kernel totals
        for (size_t k = 0; k < recs_used_; k++) {
            float ms = 0;
            HIP_CHECK(hipEventElapsedTime(&ms, recs_[k].a, recs_[k].b));
            auto it = std::find_if(agg_.begin(), agg_.end(), [&](auto &e) { return e.first == recs_[k].name; });
            if (it == agg_.end()) {
                agg_.emplace_back(recs_[k].name, Agg{});
                it = agg_.end() - 1;
            }
            it->second.launches++;
            it->second.ms += ms;
            it->second.bytes += recs_[k].bytes;
        }
        recs_used_ = 0;
    }

    void drop_graphs() {
        for (hipGraphExec_t *g : {&g_grid_, &g_post_}) {
            if (*g) (void)hipGraphExecDestroy(*g);
            *g = nullptr;
        }
        for (auto &e : g_inner_)
            for (hipGraphExec_t g : e.second)
                if (g) (void)hipGraphExecDestroy(g);
        g_inner_.clear();
    }

    template <typename F>
    void replay(hipGraphExec_t *g, F &&enqueue) {
        if (!opts_.use_graph || profiling_) {
            enqueue();
            return;
        }
        if (!*g) {
            const double t0 = now_s();
            hipGraph_t graph = nullptr;
            HIP_CHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
            enqueue();
            HIP_CHECK(hipStreamEndCapture(stream_, &graph));
            hipError_t e = hipGraphInstantiate(g, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            HIP_CHECK(e);
            captures_++;
            capture_s_ += now_s() - t0;
        }
        HIP_CHECK(hipGraphLaunch(*g, stream_));
    }

    void read_norms() {
        HIP_CHECK(hipMemcpyAsync(h_norms_, d_norms_, sizeof(double) * 2 * members_.size(), hipMemcpyDeviceToHost,
                                 stream_));
        HIP_CHECK(hipStreamSynchronize(stream_));
    }

    // Fills domain d's entries for the members `ids` into the host staging buffer.  The
    // one-workgroup-per-member domains (the multigrid tail, the scans) list the largest grids
    // first, so their long workgroups start in the first round and the small ones fill in
    // behind them (the entry order changes no member's arithmetic).
    uint32_t fill(int d, const std::vector<uint32_t> &ids_in) {
        std::vector<uint32_t> sorted;
        if (d == WD_SIM || d == WD_A_SIM) {
            sorted = ids_in;
            std::stable_sort(sorted.begin(), sorted.end(),
                             [&](uint32_t a, uint32_t b) { return members_[a]->d.ncells > members_[b]->d.ncells; });
        }
        const std::vector<uint32_t> &ids = sorted.empty() ? ids_in : sorted;
        uint32_t *out = h_entries_ + off_[d];
        uint32_t k = 0;
        for (uint32_t i : ids) {
            const uint32_t nb = work_blocks(members_[i]->d, d);
            for (uint32_t b = 0; b < nb; b++) out[k++] = (i << WORK_SIM_SHIFT) | b;
        }
        h_counts_[d] = k;
        return k;
    }

    // Solve domains for the members in `ids` (all of them at the start of a step).
    void set_active(const std::vector<uint32_t> &ids) {
        is_active_.assign(members_.size(), 0);
        for (uint32_t i : ids) is_active_[i] = 1;
        for (int d = WD_A_SIM; d < WD_COUNT; d++) fill(d, ids);
        const size_t a = off_[WD_A_SIM], bytes = 4 * (off_[WD_COUNT - 1] + cap_[WD_COUNT - 1] - a);
        HIP_CHECK(hipMemcpyAsync(d_entries_ + a, h_entries_ + a, bytes, hipMemcpyHostToDevice, stream_));
        HIP_CHECK(hipMemcpyAsync(d_counts_, h_counts_, 4 * WD_COUNT, hipMemcpyHostToDevice, stream_));
    }

    void rebuild() {
        HIP_CHECK(hipStreamSynchronize(stream_));
        const size_t n = members_.size();
        if (n > MAX_BATCH_SIMS) throw std::runtime_error("batch holds more simulations than the work lists address");
        for (Sim *s : members_) {
            std::string why;
            if (!work_fits(s->d, &why)) throw std::runtime_error("batch member too large: " + why);
        }
        bool relayout = false;
        if (n > sims_cap_) {
            sims_cap_ = std::max<size_t>(2 * sims_cap_, std::max<size_t>(n, 64));
            sims_cap_ = std::min<size_t>(sims_cap_, MAX_BATCH_SIMS);
            dev_alloc(&d_sims_, sims_cap_);
            host_alloc(&h_sims_, sims_cap_);
            dev_alloc(&d_norms_, 2 * sims_cap_);
            host_alloc(&h_norms_, 2 * sims_cap_);
            dev_alloc(&d_slots_, COUNT_SLOTS * sims_cap_);
            relayout = true;
        }
        size_t need[WD_COUNT];
        for (int d = 0; d < WD_COUNT; d++) {
            need[d] = 0;
            for (Sim *s : members_) need[d] += work_blocks(s->d, d);
            if (need[d] > cap_[d]) relayout = true;
        }
        // the batch-wide band list holds at most every member's tiles
        size_t tiles = 0;
        for (Sim *s : members_) tiles += s->d.ntiles;
        if (tiles > band_cap_) {
            band_cap_ = std::max<size_t>(2 * tiles, 65536);
            dev_alloc(&d_band_, band_cap_);
            relayout = true;
        }
        if (relayout) {
            size_t total = 0;
            for (int d = 0; d < WD_COUNT; d++) {
                // every domain gets headroom at once, so growth rarely relays out (and
                // re-captures) again
                cap_[d] = std::max<size_t>(cap_[d], std::max<size_t>(2 * need[d], WORK_LIST_MIN_CAP));
                off_[d] = total;
                total += cap_[d];
            }
            dev_alloc(&d_entries_, total);
            host_alloc(&h_entries_, total);
            for (int d = 0; d < WD_COUNT; d++)
                lists_[d] = WorkList{d_sims_, d_entries_ + off_[d], d_counts_ + d, d_slots_, d_band_, d_band_count_,
                                     lists_[d].grid};
        }
        // the graphs launch the level kernels of the deepest hierarchy seen so far (levels a
        // member does not have get empty work lists), so members leaving do not force a
        // re-capture
        int levels = max_levels_, mid = max_mid_;
        for (Sim *s : members_) {
            levels = std::max(levels, s->d.mg_levels);
            mid = std::max(mid, s->d.mg_mid);  // list levels launched grid-wide: below mg_mid
        }
        // A launch's persistent grid is baked into the graphs: twice its list length rounded up
        // to a power of two (at least 64), grown (and the graphs re-captured) only when a list
        // outgrows it, never shrunk: any grid processes every entry, so this only tunes speed,
        // and members coming and going then rarely force a re-capture.
        auto pow2 = [](size_t v) {
            uint32_t g = 64;
            while (g < v) g <<= 1;
            return g;
        };
        bool regrid = false;
        for (int d = 0; d < WD_COUNT; d++)
            if (need[d] > lists_[d].grid) regrid = true;
        if (regrid)
            for (int d = 0; d < WD_COUNT; d++) lists_[d].grid = std::max(lists_[d].grid, pow2(2 * need[d]));
        if (relayout || regrid || levels != max_levels_ || mid != max_mid_) drop_graphs();
        rebuilds_++;
        max_levels_ = levels;
        max_mid_ = mid;
        // each member's copy points its norms and list counts into the batch's arrays (the
        // counts are rebuilt every step before they are read, so slots can move between steps)
        for (size_t i = 0; i < n; i++) {
            h_sims_[i] = members_[i]->d;
            h_sims_[i].norms = d_norms_ + 2 * i;
            uint32_t *slot = d_slots_ + i * COUNT_SLOTS;
            h_sims_[i].band_all = d_band_;
            h_sims_[i].band_all_count = d_band_count_;
            h_sims_[i].member = (uint32_t)i;
            h_sims_[i].fl_count = slot;
            for (int l = 1; l < MG_MAX; l++) h_sims_[i].mg[l].count = slot + 4 * l;
        }
        HIP_CHECK(hipMemcpyAsync(d_sims_, h_sims_, sizeof(DevSim) * n, hipMemcpyHostToDevice, stream_));
        std::vector<uint32_t> all(n);
        for (size_t i = 0; i < n; i++) all[i] = (uint32_t)i;
        for (int d = 0; d < WD_A_SIM; d++) fill(d, all);
        HIP_CHECK(hipMemcpyAsync(d_entries_, h_entries_, 4 * off_[WD_A_SIM], hipMemcpyHostToDevice, stream_));
        // the copies above read pinned staging that the next rebuild rewrites
        HIP_CHECK(hipStreamSynchronize(stream_));
        counts_.assign(n, MemberCounts{});
        dirty_ = false;
    }

    pxlflip_opts opts_;
    hipStream_t stream_ = nullptr;
    std::vector<Sim *> members_;
    std::vector<uint32_t> passes_;
    std::vector<uint8_t> is_active_;
    bool dirty_ = true;
    size_t sims_cap_ = 0;
    DevSim *d_sims_ = nullptr, *h_sims_ = nullptr;
    uint32_t *d_entries_ = nullptr, *h_entries_ = nullptr;
    uint32_t *d_counts_ = nullptr, *h_counts_ = nullptr;
    double *d_norms_ = nullptr, *h_norms_ = nullptr;
    uint32_t *d_slots_ = nullptr;  // list counts, COUNT_SLOTS per member
    uint32_t *d_band_ = nullptr, *d_band_count_ = nullptr;  // the batch-wide band list
    size_t band_cap_ = 0;
    size_t off_[WD_COUNT], cap_[WD_COUNT];
    WorkLists lists_{};
    int max_levels_ = 0, max_mid_ = 0;
    hipGraphExec_t g_grid_ = nullptr, g_post_ = nullptr;
    std::map<uint32_t, std::array<hipGraphExec_t, 2>> g_inner_;  // [MG-CG iterations][x64 parity]
    uint64_t captures_ = 0, rebuilds_ = 0;
    double capture_s_ = 0.0;
    bool profiling_ = false;
    std::vector<Rec> recs_;
    size_t recs_used_ = 0, grid_start_ = 0;
    std::vector<MemberCounts> counts_;
    std::vector<std::pair<std::string, Agg>> agg_;
};

// Persistent-grid maxima of the batched launch families, measured on the 64-request serving
// batch (MI355X; the autotuner can re-pick them): band passes 4096 (0.8% more throughput than
// 2048, 5% more than 1024), particle kernels 2048, per-tile P2G/G2P 2048 (2% more than 1024),
// multigrid level setup 2048, multigrid list kernels 512 (two 1024-lane workgroups per CU,
// so one round on the chip; 1024 made sweeps ~6% slower), tail 512, solver 1024 (2048 no
// faster).  Read when a launch is recorded, so a change applies to graphs captured later.
std::atomic<uint32_t> g_batch_caps[BF_COUNT] = {4096, 2048, 2048, 2048, 512, 512, 1024};
static_assert(BF_BAND == PXLFLIP_GRID_BAND && BF_PART == PXLFLIP_GRID_PARTICLE && BF_TILE == PXLFLIP_GRID_TILE &&
                  BF_MG_GRID == PXLFLIP_GRID_MG_SETUP && BF_MG_LIST == PXLFLIP_GRID_MG_LIST &&
                  BF_MG_TAIL == PXLFLIP_GRID_MG_TAIL && BF_SOLVER == PXLFLIP_GRID_SOLVER && BF_COUNT == PXLFLIP_GRID_COUNT,
              "batch grid families match pxlflip.h");

// Solver workgroups of a batched simulation: one per ~2048 FLUID cells it is expected to hold
// (its particles over the seeding density), so that a workgroup's slice of a list keeps its
// lanes busy.
uint32_t batch_solver_blocks(const Sim &s) {
    // a multiple of 8, so every member's entries start on a multiple of 8 in the solve lists
    // and a persistent workgroup's XCD (its index mod 8) is the slice index mod 8: the XCD
    // slice mapping of the list kernels then holds in the batch as it does alone
    const double fluid = (double)s.d.np / (double)std::max<uint32_t>(1u, s.spec.particles_per_cell);
    const double g = 8.0 * std::ceil(fluid / (8.0 * 2048.0));
    return (uint32_t)std::max(8.0, std::min(256.0, g));
}

struct Request {
    uint64_t id = 0;
    std::string json;
    double fps = 24.0;
    uint32_t seed = 0, steps = 0;
    bool keep = false;
    int status = PXLFLIP_REQ_QUEUED;
    bool cancel = false;
    std::string error;
    double t_submit = 0, t_admit = 0, t_done = 0;
    uint64_t bytes = 0;
    uint32_t np = 0;
    uint32_t steps_done = 0;  // updated by the device thread under the server lock
    std::string alloc_error;  // set when admission failed to allocate
    bool allocated = false;   // set by an allocator thread (under the lock) when r->sim is ready
    bool in_batch = false;    // member of the batch (device thread only)
    std::unique_ptr<Sim> sim;
    std::vector<float> result;
    double rel_residual = -1.0;
    uint32_t max_outer = 0;
};

}  // namespace

unsigned batch_cap(int family) {
    return family >= 0 && family < BF_COUNT ? g_batch_caps[family].load(std::memory_order_relaxed) : 1024u;
}
bool set_batch_cap(int family, unsigned workgroups) {
    // a multiple of 8 (the XCD mappings assume it) and within a work list's minimum storage
    if (family < 0 || family >= BF_COUNT || workgroups < 8 || workgroups > WORK_LIST_MIN_CAP || workgroups % 8)
        return false;
    g_batch_caps[family].store(workgroups, std::memory_order_relaxed);
    return true;
}

// Workgroup sizes of the families compiled in two sizes (the first is the default: on the
// 64-request serving batch it was measured fastest when the sizes were fixed at compile time).
constexpr unsigned kShapeSizes[BS_COUNT][2] = {{512, 1024}, {256, 512}};  // default first
std::atomic<uint32_t> g_batch_shapes[BS_COUNT] = {kShapeSizes[BS_TILE][0], kShapeSizes[BS_BAND][0]};
static_assert(BS_TILE == PXLFLIP_SHAPE_TILE && BS_BAND == PXLFLIP_SHAPE_BAND && BS_COUNT == PXLFLIP_SHAPE_COUNT,
              "batch shape families match pxlflip.h");
unsigned batch_shape(int family) {
    return family >= 0 && family < BS_COUNT ? g_batch_shapes[family].load(std::memory_order_relaxed) : 0u;
}
bool set_batch_shape(int family, unsigned lanes) {
    if (family < 0 || family >= BS_COUNT || (lanes != kShapeSizes[family][0] && lanes != kShapeSizes[family][1]))
        return false;
    g_batch_shapes[family].store(lanes, std::memory_order_relaxed);
    return true;
}
}  // namespace pxl

// Host threads that seed and allocate admitted requests (seeding a large liquid is several
// milliseconds of host work).  The batch fills sooner with more of them: on the 512-request
// capacity run 8 threads give 3% more particle-steps/s than 4 (mean batch 394-404 against
// 366-375), and 16 no more than 8.  Each thread prepares its requests on one stream of its
// own, created with the server: creating a stream per request took ~2 ms of the runtime's
// time each (1.1 s of the 512 requests' admission), one request after another.
constexpr int ALLOCATOR_THREADS = 8;

struct pxlflip_server {
    pxlflip_server_opts opts{};
    pxlflip_opts sim_opts{};
    std::unique_ptr<pxl::Batch> batch;
    std::mutex mu;
    std::condition_variable cv;       // worker wake-up
    std::condition_variable done_cv;  // request state changes
    std::deque<pxl::Request *> queue;        // submitted, waiting for admission
    std::deque<pxl::Request *> to_allocate;  // admitted, waiting for an allocator thread
    std::vector<pxl::Request *> running;     // admitted and not finished (allocating or stepping)
    // simulations of finished requests, freed by the release thread (freeing device memory
    // waits for the device), with the budget each returns
    std::deque<std::pair<std::unique_ptr<pxl::Sim>, uint64_t>> to_free;
    size_t freeing = 0;  // in to_free or being freed
    std::map<uint64_t, std::unique_ptr<pxl::Request>> all;
    uint64_t next_id = 1;
    bool stop = false;
    bool worker_done = false;  // the device thread has left: nothing more is queued to free
    std::thread worker;
    std::vector<std::thread> allocators;  // allocate + seed admitted requests off the device thread
    std::vector<hipStream_t> alloc_streams;  // one per allocator thread
    std::thread releaser;                 // frees finished requests' simulations
    std::string error;
    // accounting
    uint64_t bytes_in_use = 0, peak_bytes = 0;
    uint64_t submitted = 0, completed = 0, failed = 0, cancelled = 0;
    std::vector<double> latencies_ms;
    double particle_steps = 0, first_submit = -1, last_done = -1, busy_s = 0, retire_s = 0;
    uint64_t batch_steps = 0;
    double batch_members = 0;
    std::vector<std::pair<std::string, pxl::Batch::Agg>> kernel_stats;  // profiled batch steps
    double admit_s = 0;  // allocator-thread time spent allocating admitted requests

    ~pxlflip_server() {  // the allocator threads have left (or never started)
        for (hipStream_t st : alloc_streams)
            if (st) (void)hipStreamDestroy(st);
    }
    void run();
    void allocate_loop(int k);
    void free_loop();
    void finish(pxl::Request *r, int status, const std::string &err);
};

namespace pxl {
namespace {

// Allocates a request's simulation on the allocator thread's stream `st` (allocate and reset
// end with that stream drained).  Throws on failure.
void allocate_request(pxlflip_server *srv, Request *r, hipStream_t st) {
    auto sim = std::make_unique<Sim>();
    if (prepare(r->json.c_str(), r->fps, r->seed, &srv->sim_opts, *sim, nullptr, nullptr) != PXLFLIP_OK)
        throw std::runtime_error(g_create_error);
    sim->d.cg_blocks = (int)(srv->opts.solver_workgroups ? srv->opts.solver_workgroups : batch_solver_blocks(*sim));
    sim->stream = st;
    sim->owns_stream = false;
    sim->allocate();
    sim->reset();
    r->sim = std::move(sim);
}

}  // namespace
}  // namespace pxl

void pxlflip_server::finish(pxl::Request *r, int status, const std::string &err) {
    // mu held
    r->status = status;
    r->error = err;
    r->t_done = pxl::now_s();
    if (status == PXLFLIP_REQ_DONE) {
        completed++;
        latencies_ms.push_back(1e3 * (r->t_done - r->t_submit));
        particle_steps += (double)r->np * (double)r->steps;
        last_done = r->t_done;
    } else if (status == PXLFLIP_REQ_FAILED) {
        failed++;
    } else {
        cancelled++;
    }
    done_cv.notify_all();
}

// Frees the simulations of finished requests and returns their budget.  Freeing device
// memory waits for the device, so this runs on its own thread, not the device thread.  It
// leaves only after the device thread has (whose last step can still retire simulations), so
// every retired simulation is freed here.
void pxlflip_server::free_loop() {
    using namespace pxl;
    try {
        HIP_CHECK(hipSetDevice(opts.device));
    } catch (const HipError &) {
    }
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
        cv.wait(lk, [&] { return !to_free.empty() || worker_done; });
        if (to_free.empty()) break;  // the device thread has left and nothing is left to free
        auto item = std::move(to_free.front());
        to_free.pop_front();
        lk.unlock();
        item.first.reset();
        lk.lock();
        bytes_in_use -= item.second;
        freeing--;
        cv.notify_all();       // the returned budget may admit a queued request
        done_cv.notify_all();  // drain waits for the memory too
    }
}

void pxlflip_server::allocate_loop(int k) {
    using namespace pxl;
    try {
        HIP_CHECK(hipSetDevice(opts.device));
    } catch (const HipError &) {
    }
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
        cv.wait(lk, [&] { return stop || !to_allocate.empty(); });
        if (stop) break;
        Request *r = to_allocate.front();
        to_allocate.pop_front();
        lk.unlock();
        const double t0 = now_s();
        std::string err;
        try {
            allocate_request(this, r, alloc_streams[(size_t)k]);
        } catch (const HipError &e) {
            err = std::string("HIP error in ") + e.what + ": " + hipGetErrorString(e.code);
        } catch (const std::bad_alloc &) {
            err = "out of device memory";
        } catch (const std::exception &e) {
            err = e.what();
        }
        if (!err.empty()) r->sim.reset();
        lk.lock();
        admit_s += now_s() - t0;
        r->alloc_error = err;
        r->allocated = true;
        cv.notify_all();
    }
}

void pxlflip_server::run() {
    using namespace pxl;
    try {
        HIP_CHECK(hipSetDevice(opts.device));
    } catch (const HipError &) {
    }
    std::unique_lock<std::mutex> lk(mu);
    while (!stop) {
        // admission: first in, first out, while the budget and the batch size allow.  An
        // admitted request is `running` from here on; an allocator thread prepares it.
        while (!queue.empty()) {
            Request *r = queue.front();
            if (r->cancel) {
                queue.pop_front();
                finish(r, PXLFLIP_REQ_CANCELLED, "");
                continue;
            }
            if (bytes_in_use + r->bytes > opts.budget_bytes || running.size() >= opts.max_sims) break;
            queue.pop_front();
            bytes_in_use += r->bytes;
            peak_bytes = std::max(peak_bytes, bytes_in_use);
            r->status = PXLFLIP_REQ_RUNNING;
            r->t_admit = now_s();
            running.push_back(r);
            to_allocate.push_back(r);
            cv.notify_all();
        }
        // allocated requests join the batch at this step boundary
        std::vector<Request *> joining;
        size_t in_batch = 0;
        for (Request *r : running) {
            if (r->in_batch) in_batch++;
            else if (r->allocated && r->alloc_error.empty() && r->steps > 0) joining.push_back(r);
        }
        bool retire_only = false;
        for (Request *r : running)
            if (r->allocated && (!r->alloc_error.empty() || r->steps == 0 || r->cancel)) retire_only = true;
        if (in_batch + joining.size() == 0 && !retire_only) {
            cv.wait(lk, [&] {
                if (stop) return true;
                for (Request *r : running)
                    if (r->allocated) return true;
                return running.size() < opts.max_sims && !queue.empty() &&
                       bytes_in_use + queue.front()->bytes <= opts.budget_bytes;
            });
            continue;
        }
        lk.unlock();
        std::string step_error;
        double t0 = now_s(), t1 = t0;
        size_t members = 0;
        try {
            for (Request *r : joining) {
                batch->add(r->sim.get());
                r->in_batch = true;
            }
            members = batch->size();
            if (members) {
                const bool prof = opts.profile_every && (batch_steps % opts.profile_every) == opts.profile_every - 1;
                if (prof) batch->set_profiling(true);
                batch->step();
                batch->sync();
                if (prof) {
                    auto agg = batch->profile();
                    batch->set_profiling(false);
                    std::lock_guard<std::mutex> g(mu);
                    for (auto &e : agg) {
                        auto it = std::find_if(kernel_stats.begin(), kernel_stats.end(),
                                               [&](auto &k) { return k.first == e.first; });
                        if (it == kernel_stats.end()) {
                            kernel_stats.push_back(e);
                        } else {
                            it->second.launches += e.second.launches;
                            it->second.ms += e.second.ms;
                            it->second.bytes += e.second.bytes;
                        }
                    }
                }
            }
            t1 = now_s();
        } catch (const HipError &e) {
            step_error = std::string("HIP error in ") + e.what + ": " + hipGetErrorString(e.code);
        } catch (const std::exception &e) {
            step_error = e.what();
        }
        lk.lock();
        if (members) {
            batch_steps++;
            batch_members += (double)members;
            busy_s += t1 - t0;
        }
        // retire finished, cancelled or failed requests at the step boundary
        const double t_retire = now_s();
        std::vector<Request *> keep;
        for (Request *r : running) {
            if (!r->allocated) {  // still with an allocator thread
                keep.push_back(r);
                continue;
            }
            Sim *s = r->sim.get();
            if (!s) {  // allocation failed
                bytes_in_use -= r->bytes;
                finish(r, PXLFLIP_REQ_FAILED, r->alloc_error);
                continue;
            }
            if (r->in_batch && members && step_error.empty()) {
                r->rel_residual = s->stats.rel_residual;
                r->max_outer = std::max(r->max_outer, s->stats.outer_iterations);
                r->steps_done = (uint32_t)s->steps_done;
            }
            const bool done = s->steps_done >= r->steps;
            if (step_error.empty() && !done && !r->cancel) {
                keep.push_back(r);
                continue;
            }
            // leaves the batch now (the batch re-lays out its lists before its next launch);
            // kept particles are copied out here, the release thread frees the simulation
            lk.unlock();
            std::string err = step_error;
            try {
                if (r->in_batch) batch->remove(s);
                r->in_batch = false;
                if (done && err.empty() && r->keep) {
                    r->result.assign((size_t)s->d.np * 8, 0.0f);
                    DeviceBuffer buf((size_t)s->d.np * 32);
                    const hipStream_t bs = batch->stream();  // the device thread's own
                    launch_unpack_aos8(s->d, buf.as<float>(), bs);
                    HIP_CHECK(hipMemcpyAsync(r->result.data(), buf.p, (size_t)s->d.np * 32, hipMemcpyDeviceToHost, bs));
                    HIP_CHECK(hipStreamSynchronize(bs));
                }
            } catch (const HipError &e) {
                err = std::string("HIP error in ") + e.what + ": " + hipGetErrorString(e.code);
            }
            lk.lock();
            to_free.emplace_back(std::move(r->sim), r->bytes);
            freeing++;
            cv.notify_all();
            finish(r, !err.empty() ? PXLFLIP_REQ_FAILED : (done ? PXLFLIP_REQ_DONE : PXLFLIP_REQ_CANCELLED), err);
        }
        running.swap(keep);
        retire_s += now_s() - t_retire;
    }
    // stop: wait for the allocators to leave, then cancel what is left
    lk.unlock();
    for (auto &t : allocators)
        if (t.joinable()) t.join();
    lk.lock();
    for (Request *r : running) {
        lk.unlock();
        try {
            if (r->in_batch) batch->remove(r->sim.get());
        } catch (...) {
        }
        r->sim.reset();
        lk.lock();
        bytes_in_use -= r->bytes;
        finish(r, PXLFLIP_REQ_CANCELLED, "server stopped");
    }
    running.clear();
    to_allocate.clear();
    while (!queue.empty()) {
        finish(queue.front(), PXLFLIP_REQ_CANCELLED, "server stopped");
        queue.pop_front();
    }
    worker_done = true;
    cv.notify_all();
}

namespace {

void fill_request_info(const pxl::Request &r, pxlflip_request_info *info) {
    std::memset(info, 0, sizeof *info);
    info->status = r.status;
    info->steps = r.steps;
    info->steps_done = r.steps_done;
    info->n_particles = r.np;
    info->device_bytes = r.bytes;
    info->queue_ms = r.t_admit > 0 ? 1e3 * (r.t_admit - r.t_submit) : -1.0;
    info->latency_ms = r.status == PXLFLIP_REQ_DONE ? 1e3 * (r.t_done - r.t_submit) : -1.0;
    info->rel_residual = r.rel_residual;
    info->max_outer = r.max_outer;
}

double percentile(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    // nearest-rank percentile
    size_t k = (size_t)std::ceil(q * (double)v.size());
    k = k == 0 ? 0 : k - 1;
    return v[std::min(k, v.size() - 1)];
}

}  // namespace

extern "C" {

int pxlflip_batch_get_grid(int family, uint32_t *workgroups) {
    if (!workgroups || family < 0 || family >= pxl::BF_COUNT) return PXLFLIP_ERR_INVALID;
    *workgroups = pxl::batch_cap(family);
    return PXLFLIP_OK;
}

int pxlflip_batch_set_grid(int family, uint32_t workgroups) {
    return pxl::set_batch_cap(family, workgroups) ? PXLFLIP_OK : PXLFLIP_ERR_INVALID;
}

int pxlflip_batch_get_shape(int family, uint32_t *lanes) {
    if (!lanes || family < 0 || family >= pxl::BS_COUNT) return PXLFLIP_ERR_INVALID;
    *lanes = pxl::batch_shape(family);
    return PXLFLIP_OK;
}

int pxlflip_batch_set_shape(int family, uint32_t lanes) {
    return pxl::set_batch_shape(family, lanes) ? PXLFLIP_OK : PXLFLIP_ERR_INVALID;
}

int pxlflip_server_default_opts(pxlflip_server_opts *o) {
    if (!o) return PXLFLIP_ERR_INVALID;
    std::memset(o, 0, sizeof *o);
    o->struct_size = sizeof(pxlflip_server_opts);
    o->device = 0;
    o->budget_bytes = 200000000000ull;
    o->max_sims = 256;
    o->inner_iterations = 9;
    o->max_outer = 50;
    o->tolerance = 1e-6;
    o->solver_workgroups = 0;
    o->use_graph = 1;
    o->profile_every = 0;
    return PXLFLIP_OK;
}

int pxlflip_server_create(const pxlflip_server_opts *opts, pxlflip_server **out) {
    using namespace pxl;
    g_create_error.clear();
    if (!out) {
        g_create_error = "out is NULL";
        return PXLFLIP_ERR_INVALID;
    }
    *out = nullptr;
    pxlflip_server_opts o;
    pxlflip_server_default_opts(&o);
    if (opts) {
        if (opts->struct_size != sizeof(pxlflip_server_opts)) {
            g_create_error = "pxlflip_server_opts.struct_size mismatch (use pxlflip_server_default_opts)";
            return PXLFLIP_ERR_INVALID;
        }
        o = *opts;
    }
    if (o.max_sims == 0 || o.max_sims > MAX_BATCH_SIMS) {
        g_create_error = "max_sims must be 1.." + std::to_string(MAX_BATCH_SIMS);
        return PXLFLIP_ERR_INVALID;
    }
    if (o.inner_iterations == 0 || o.max_outer == 0 || !(o.tolerance > 0.0)) {
        g_create_error = "inner_iterations, max_outer and tolerance must be positive";
        return PXLFLIP_ERR_INVALID;
    }
    if (o.solver_workgroups > MAX_SOLVER_BLOCKS) {
        g_create_error = "solver_workgroups must be at most " + std::to_string(MAX_SOLVER_BLOCKS);
        return PXLFLIP_ERR_INVALID;
    }
    auto srv = std::make_unique<pxlflip_server>();
    srv->opts = o;
    pxlflip_default_opts(&srv->sim_opts);
    srv->sim_opts.solver = PXLFLIP_SOLVER_MIXED;
    srv->sim_opts.inner_iterations = o.inner_iterations;
    srv->sim_opts.max_outer = o.max_outer;
    srv->sim_opts.tolerance = o.tolerance;
    srv->sim_opts.device = o.device;
    srv->sim_opts.use_graph = o.use_graph;
    try {
        int n = 0;
        HIP_CHECK(hipGetDeviceCount(&n));
        if (o.device < 0 || o.device >= n) {
            g_create_error = "device " + std::to_string(o.device) + " not present";
            return PXLFLIP_ERR_DEVICE;
        }
        HIP_CHECK(hipSetDevice(o.device));
        srv->batch = std::make_unique<Batch>(srv->sim_opts);
        for (int k = 0; k < ALLOCATOR_THREADS; k++) {
            hipStream_t st = nullptr;
            HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
            srv->alloc_streams.push_back(st);
        }
    } catch (const HipError &e) {
        g_create_error = std::string("HIP error in ") + e.what + ": " + hipGetErrorString(e.code);
        return PXLFLIP_ERR_DEVICE;
    } catch (const std::bad_alloc &) {
        g_create_error = "out of memory";
        return PXLFLIP_ERR_OOM;
    }
    pxlflip_server *p = srv.release();
    for (int k = 0; k < ALLOCATOR_THREADS; k++) p->allocators.emplace_back([p, k] { p->allocate_loop(k); });
    p->releaser = std::thread([p] { p->free_loop(); });
    p->worker = std::thread([p] { p->run(); });
    *out = p;
    return PXLFLIP_OK;
}

namespace {
// Sets a prepared simulation's batched solver workgroups and checks that it can join a batch:
// members step in lockstep with one dt (the CFL cap's per-simulation substeps do not fit), and
// every domain's workgroups must be addressable by a work-list entry's block field.
bool batchable(pxl::Sim &s, const pxlflip_server_opts &o) {
    using namespace pxl;
    s.d.cg_blocks = (int)(o.solver_workgroups ? o.solver_workgroups : batch_solver_blocks(s));
    if (s.spec.cfl > 0.0f) {
        g_create_error = "fluid '" + s.spec.name + "': cfl (the CFL cap) is not supported in batched execution";
        return false;
    }
    std::string why;
    if (!work_fits(s.d, &why)) {
        g_create_error = "fluid '" + s.spec.name + "': too large for batched execution (" + why + ")";
        return false;
    }
    return true;
}
}  // namespace

int pxlflip_server_submit(pxlflip_server *srv, const char *fluid_json, double fps, uint32_t seed, uint32_t steps,
                          int32_t keep_particles, uint64_t *id) {
    using namespace pxl;
    if (!srv) return PXLFLIP_ERR_INVALID;
    auto r = std::make_unique<Request>();
    r->t_submit = now_s();
    Sim probe;
    if (prepare(fluid_json, fps, seed, &srv->sim_opts, probe, nullptr, nullptr) != PXLFLIP_OK ||
        !batchable(probe, srv->opts)) {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->error = g_create_error;
        return PXLFLIP_ERR_INVALID;
    }
    r->bytes = Sim::plan_total(probe.plan());
    r->np = probe.d.np;
    r->json = fluid_json;
    r->fps = fps;
    r->seed = seed;
    r->steps = steps;
    r->keep = keep_particles != 0;
    std::lock_guard<std::mutex> lk(srv->mu);
    if (r->bytes > srv->opts.budget_bytes) {
        srv->error = "request needs " + std::to_string(r->bytes) + " bytes of HBM, more than the admission budget (" +
                     std::to_string(srv->opts.budget_bytes) + ")";
        return PXLFLIP_ERR_INVALID;
    }
    r->id = srv->next_id++;
    if (id) *id = r->id;
    if (srv->first_submit < 0) srv->first_submit = r->t_submit;
    srv->submitted++;
    srv->queue.push_back(r.get());
    srv->all.emplace(r->id, std::move(r));
    srv->cv.notify_all();
    return PXLFLIP_OK;
}

int pxlflip_server_poll(pxlflip_server *srv, uint64_t id, pxlflip_request_info *info) {
    if (!srv) return PXLFLIP_ERR_INVALID;
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->all.find(id);
    if (it == srv->all.end()) {
        srv->error = "unknown request " + std::to_string(id);
        return PXLFLIP_ERR_INVALID;
    }
    if (info) fill_request_info(*it->second, info);
    if (it->second->status == PXLFLIP_REQ_FAILED) srv->error = it->second->error;
    return PXLFLIP_OK;
}

int pxlflip_server_wait(pxlflip_server *srv, uint64_t id, double timeout_s, pxlflip_request_info *info) {
    if (!srv) return PXLFLIP_ERR_INVALID;
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->all.find(id);
    if (it == srv->all.end()) {
        srv->error = "unknown request " + std::to_string(id);
        return PXLFLIP_ERR_INVALID;
    }
    pxl::Request *r = it->second.get();
    auto finished = [&] { return r->status >= PXLFLIP_REQ_DONE; };
    bool ok = true;
    if (timeout_s < 0) srv->done_cv.wait(lk, finished);
    else ok = srv->done_cv.wait_for(lk, std::chrono::duration<double>(timeout_s), finished);
    if (info) fill_request_info(*r, info);
    if (!ok) {
        srv->error = "timed out waiting for request " + std::to_string(id);
        return PXLFLIP_ERR_INVALID;
    }
    if (r->status == PXLFLIP_REQ_FAILED) srv->error = r->error;
    return PXLFLIP_OK;
}

int pxlflip_server_drain(pxlflip_server *srv, double timeout_s) {
    if (!srv) return PXLFLIP_ERR_INVALID;
    std::unique_lock<std::mutex> lk(srv->mu);
    auto idle = [&] { return srv->queue.empty() && srv->running.empty() && srv->freeing == 0; };
    bool ok = true;
    if (timeout_s < 0) srv->done_cv.wait(lk, idle);
    else ok = srv->done_cv.wait_for(lk, std::chrono::duration<double>(timeout_s), idle);
    if (!ok) {
        srv->error = "timed out draining the server";
        return PXLFLIP_ERR_INVALID;
    }
    return PXLFLIP_OK;
}

int pxlflip_server_cancel(pxlflip_server *srv, uint64_t id) {
    if (!srv) return PXLFLIP_ERR_INVALID;
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->all.find(id);
    if (it == srv->all.end()) {
        srv->error = "unknown request " + std::to_string(id);
        return PXLFLIP_ERR_INVALID;
    }
    it->second->cancel = true;
    srv->cv.notify_all();
    return PXLFLIP_OK;
}

int pxlflip_server_read_particles(pxlflip_server *srv, uint64_t id, float *aos8, uint64_t max_particles) {
    if (!srv || !aos8) return PXLFLIP_ERR_INVALID;
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->all.find(id);
    if (it == srv->all.end()) {
        srv->error = "unknown request " + std::to_string(id);
        return PXLFLIP_ERR_INVALID;
    }
    const pxl::Request &r = *it->second;
    if (r.status != PXLFLIP_REQ_DONE || !r.keep) {
        srv->error = "request " + std::to_string(id) + " has no kept particles (not done or submitted without keep)";
        return PXLFLIP_ERR_INVALID;
    }
    if (max_particles < r.np) {
        srv->error = "aos8 holds " + std::to_string(max_particles) + " particles, need " + std::to_string(r.np);
        return PXLFLIP_ERR_INVALID;
    }
    std::memcpy(aos8, r.result.data(), r.result.size() * sizeof(float));
    return PXLFLIP_OK;
}

int pxlflip_server_release(pxlflip_server *srv, uint64_t id) {
    if (!srv) return PXLFLIP_ERR_INVALID;
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->all.find(id);
    if (it == srv->all.end()) {
        srv->error = "unknown request " + std::to_string(id);
        return PXLFLIP_ERR_INVALID;
    }
    if (it->second->status < PXLFLIP_REQ_DONE) {
        srv->error = "request " + std::to_string(id) + " has not finished";
        return PXLFLIP_ERR_INVALID;
    }
    srv->all.erase(it);
    return PXLFLIP_OK;
}

int pxlflip_server_get_stats(pxlflip_server *srv, pxlflip_server_stats *st) {
    if (!srv || !st) return PXLFLIP_ERR_INVALID;
    std::lock_guard<std::mutex> lk(srv->mu);
    std::memset(st, 0, sizeof *st);
    st->submitted = srv->submitted;
    st->completed = srv->completed;
    st->failed = srv->failed;
    st->cancelled = srv->cancelled;
    st->queued = srv->queue.size();
    st->running = srv->running.size();
    st->p50_ms = percentile(srv->latencies_ms, 0.50);
    st->p99_ms = percentile(srv->latencies_ms, 0.99);
    double sum = 0, mx = 0;
    for (double v : srv->latencies_ms) {
        sum += v;
        mx = std::max(mx, v);
    }
    st->mean_ms = srv->latencies_ms.empty() ? 0.0 : sum / (double)srv->latencies_ms.size();
    st->max_ms = mx;
    st->particle_steps = srv->particle_steps;
    st->span_s = srv->last_done > srv->first_submit && srv->first_submit >= 0 ? srv->last_done - srv->first_submit : 0.0;
    st->throughput = st->span_s > 0 ? st->particle_steps / st->span_s : 0.0;
    st->busy_s = srv->busy_s;
    st->batch_steps = srv->batch_steps;
    st->mean_batch = srv->batch_steps ? srv->batch_members / (double)srv->batch_steps : 0.0;
    st->bytes_in_use = srv->bytes_in_use;
    st->peak_bytes = srv->peak_bytes;
    st->admit_s = srv->admit_s;
    st->rebuilds = srv->batch ? srv->batch->rebuilds() : 0;
    st->graph_captures = srv->batch ? srv->batch->graph_captures() : 0;
    st->capture_s = srv->batch ? srv->batch->capture_seconds() : 0.0;
    st->retire_s = srv->retire_s;
    return PXLFLIP_OK;
}

int pxlflip_server_kernel_stats(pxlflip_server *srv, pxlflip_kernel_time *out, uint32_t max_out, uint32_t *n_out) {
    if (!srv) return PXLFLIP_ERR_INVALID;
    std::lock_guard<std::mutex> lk(srv->mu);
    if (n_out) *n_out = (uint32_t)srv->kernel_stats.size();
    for (uint32_t k = 0; k < srv->kernel_stats.size() && k < max_out && out; k++) {
        const auto &e = srv->kernel_stats[k];
        std::memset(&out[k], 0, sizeof out[k]);
        std::snprintf(out[k].name, sizeof out[k].name, "%s", e.first.c_str());
        out[k].launches = e.second.launches;
        out[k].total_ms = e.second.ms;
        out[k].alg_bytes = e.second.bytes;
    }
    return PXLFLIP_OK;
}

const char *pxlflip_server_last_error(const pxlflip_server *srv) {
    if (!srv) return pxl::g_create_error.c_str();
    return srv->error.c_str();
}

void pxlflip_server_destroy(pxlflip_server *srv) {
    if (!srv) return;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
        srv->cv.notify_all();
    }
    if (srv->worker.joinable()) srv->worker.join();  // also joins the allocator threads
    if (srv->releaser.joinable()) srv->releaser.join();  // frees every retired simulation, then leaves
    (void)hipSetDevice(srv->opts.device);
    srv->batch.reset();
    delete srv;
}

// ---- fixed batch
}  // extern "C"

struct pxlflip_batch {
    pxlflip_opts sim_opts{};
    std::vector<std::unique_ptr<pxl::Sim>> sims;
    std::unique_ptr<pxl::Batch> batch;
    int device = 0;
    std::string error;
};

namespace {

template <typename F>
int batch_guarded(pxlflip_batch *b, F &&body) {
    using namespace pxl;
    if (!b) return PXLFLIP_ERR_INVALID;
    try {
        HIP_CHECK(hipSetDevice(b->device));
        return body();
    } catch (const HipError &e) {
        b->error = std::string("HIP error in ") + e.what + ": " + hipGetErrorString(e.code);
        return PXLFLIP_ERR_DEVICE;
    } catch (const std::bad_alloc &) {
        b->error = "out of memory";
        return PXLFLIP_ERR_OOM;
    } catch (const std::exception &e) {
        b->error = e.what();
        return PXLFLIP_ERR_INTERNAL;
    }
}

}  // namespace

extern "C" {

int pxlflip_batch_create(const char *const *fluid_jsons, const uint32_t *seeds, uint32_t n, double fps,
                         const pxlflip_server_opts *opts, pxlflip_batch **out) {
    using namespace pxl;
    g_create_error.clear();
    if (!out || (!fluid_jsons && n)) {
        g_create_error = "fluid_jsons / out is NULL";
        return PXLFLIP_ERR_INVALID;
    }
    *out = nullptr;
    if (n == 0 || n > MAX_BATCH_SIMS) {
        g_create_error = "a batch holds 1.." + std::to_string(MAX_BATCH_SIMS) + " simulations";
        return PXLFLIP_ERR_INVALID;
    }
    pxlflip_server_opts o;
    pxlflip_server_default_opts(&o);
    if (opts) {
        if (opts->struct_size != sizeof(pxlflip_server_opts)) {
            g_create_error = "pxlflip_server_opts.struct_size mismatch (use pxlflip_server_default_opts)";
            return PXLFLIP_ERR_INVALID;
        }
        o = *opts;
    }
    auto b = std::make_unique<pxlflip_batch>();
    b->device = o.device;
    pxlflip_default_opts(&b->sim_opts);
    b->sim_opts.solver = PXLFLIP_SOLVER_MIXED;
    b->sim_opts.inner_iterations = o.inner_iterations;
    b->sim_opts.max_outer = o.max_outer;
    b->sim_opts.tolerance = o.tolerance;
    b->sim_opts.device = o.device;
    b->sim_opts.use_graph = o.use_graph;
    // every block is parsed, validated and checked for batching before any device work
    for (uint32_t i = 0; i < n; i++) {
        Sim probe;
        b->sim_opts.scene_index = i;
        const int rc = prepare(fluid_jsons[i], fps, seeds ? seeds[i] : 0u, &b->sim_opts, probe, nullptr, nullptr);
        if (rc != PXLFLIP_OK) return rc;
        if (!batchable(probe, o)) return PXLFLIP_ERR_INVALID;
    }
    try {
        HIP_CHECK(hipSetDevice(o.device));
        for (uint32_t i = 0; i < n; i++) {
            auto sim = std::make_unique<Sim>();
            b->sim_opts.scene_index = i;
            const int rc = prepare(fluid_jsons[i], fps, seeds ? seeds[i] : 0u, &b->sim_opts, *sim, nullptr, nullptr);
            if (rc != PXLFLIP_OK) return rc;
            if (!batchable(*sim, o)) return PXLFLIP_ERR_INVALID;
            HIP_CHECK(hipStreamCreateWithFlags(&sim->stream, hipStreamNonBlocking));
            sim->allocate();
            sim->reset();
            b->sims.push_back(std::move(sim));
        }
        b->batch = std::make_unique<Batch>(b->sim_opts);
        for (auto &sim : b->sims) b->batch->add(sim.get());
    } catch (const HipError &e) {
        g_create_error = std::string("HIP error in ") + e.what + ": " + hipGetErrorString(e.code);
        return e.code == hipErrorOutOfMemory ? PXLFLIP_ERR_OOM : PXLFLIP_ERR_DEVICE;
    } catch (const std::bad_alloc &) {
        g_create_error = "out of device memory";
        return PXLFLIP_ERR_OOM;
    } catch (const std::exception &e) {
        g_create_error = e.what();
        return PXLFLIP_ERR_INTERNAL;
    }
    *out = b.release();
    return PXLFLIP_OK;
}

int pxlflip_batch_step(pxlflip_batch *b, uint32_t steps) {
    return batch_guarded(b, [&] {
        for (uint32_t k = 0; k < steps; k++) b->batch->step();
        b->batch->sync();
        return PXLFLIP_OK;
    });
}

int pxlflip_batch_profile_steps(pxlflip_batch *b, uint32_t steps, pxlflip_kernel_time *out, uint32_t max_out,
                                uint32_t *n_out) {
    return batch_guarded(b, [&] {
        b->batch->set_profiling(true);
        for (uint32_t k = 0; k < steps; k++) b->batch->step();
        auto agg = b->batch->profile();
        b->batch->set_profiling(false);
        if (n_out) *n_out = (uint32_t)agg.size();
        for (uint32_t k = 0; k < agg.size() && k < max_out && out; k++) {
            std::memset(&out[k], 0, sizeof out[k]);
            std::snprintf(out[k].name, sizeof out[k].name, "%s", agg[k].first.c_str());
            out[k].launches = agg[k].second.launches;
            out[k].total_ms = agg[k].second.ms;
            out[k].alg_bytes = agg[k].second.bytes;
        }
        return PXLFLIP_OK;
    });
}

int pxlflip_batch_get_info(pxlflip_batch *b, uint32_t index, pxlflip_info *info) {
    if (!b || !info || index >= b->sims.size()) return PXLFLIP_ERR_INVALID;
    pxl::fill_info(*b->sims[index], info);
    return PXLFLIP_OK;
}

int pxlflip_batch_read_particles(pxlflip_batch *b, uint32_t index, float *aos8) {
    if (!b || !aos8 || index >= b->sims.size()) return PXLFLIP_ERR_INVALID;
    return batch_guarded(b, [&] {
        using namespace pxl;
        Sim &s = *b->sims[index];
        b->batch->sync();
        DeviceBuffer buf((size_t)s.d.np * 32);
        launch_unpack_aos8(s.d, buf.as<float>(), s.stream);
        HIP_CHECK(hipMemcpyAsync(aos8, buf.p, (size_t)s.d.np * 32, hipMemcpyDeviceToHost, s.stream));
        HIP_CHECK(hipStreamSynchronize(s.stream));
        return PXLFLIP_OK;
    });
}

int pxlflip_batch_read_field(pxlflip_batch *b, uint32_t index, int field, void *out) {
    if (!b || !out || index >= b->sims.size()) return PXLFLIP_ERR_INVALID;
    pxl::Sim &s = *b->sims[index];
    const std::string why = pxl::field_check(s, field);
    if (!why.empty()) {
        b->error = why;
        return PXLFLIP_ERR_INVALID;
    }
    return batch_guarded(b, [&] {
        b->batch->sync();
        pxl::read_field_to_host(s, field, out);
        return PXLFLIP_OK;
    });
}

int pxlflip_batch_read_bins(pxlflip_batch *b, uint32_t index, uint32_t *cell_of, uint32_t *counts, uint32_t *offsets,
                            uint32_t *sorted_ids) {
    if (!b || index >= b->sims.size()) return PXLFLIP_ERR_INVALID;
    return batch_guarded(b, [&] {
        b->batch->sync();
        pxl::read_bins_to_host(*b->sims[index], cell_of, counts, offsets, sorted_ids);
        return PXLFLIP_OK;
    });
}

int pxlflip_batch_last_solve(pxlflip_batch *b, uint32_t index, pxlflip_solve_stats *st) {
    if (!b || !st || index >= b->sims.size()) return PXLFLIP_ERR_INVALID;
    *st = b->sims[index]->stats;
    return PXLFLIP_OK;
}

const char *pxlflip_batch_last_error(const pxlflip_batch *b) {
    if (!b) return pxl::g_create_error.c_str();
    return b->error.c_str();
}

void pxlflip_batch_destroy(pxlflip_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    b->batch.reset();
    b->sims.clear();
    delete b;
}

}  // extern "C"
