// The scene's fluid block and the host-side derivations of the FLIP path.
//
// FluidSpec mirrors pxl_scene::Fluid / FluidBlock (crates/pxl-scene/src/lib.rs:49-119),
// parse_fluid() its strict serde deserialisation, validate_fluid() the fluid checks of
// Scene::validate (lib.rs:710-735), fluid_grid() Fluid::grid (lib.rs:123-128), and
// seed_particles() the seeding loop of FluidState::new (crates/pxl-render/src/fluid.rs:96-149).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace pxl {

// A fill region: the reference's box {min, max, velocity}, or a sphere {center, radius,
// velocity} (an extension: the reference's blocks are boxes only).
struct FluidBlockSpec {
    float min[3] = {0, 0, 0};
    float max[3] = {0, 0, 0};
    float velocity[3] = {0, 0, 0};
    bool sphere = false;
    float center[3] = {0, 0, 0};
    float radius = 0.0f;
};

// A solid sphere inside the domain (an extension: the reference has no obstacles).
struct ObstacleSpec {
    float center[3] = {0, 0, 0};
    float radius = 0.0f;
};
constexpr uint32_t MAX_OBSTACLES = 65536;

struct FluidSpec {
    std::string name;
    float bounds[2][3] = {{0, 0, 0}, {0, 0, 0}};
    uint32_t resolution = 64;
    std::vector<FluidBlockSpec> fill;
    uint32_t particles_per_cell = 8;
    uint32_t substeps = 8;
    float flip = 0.95f;
    float gravity[3] = {0.0f, -9.81f, 0.0f};
    uint32_t pressure_iterations = 80;
    uint32_t surface_refine = 2;
    float particle_radius = 2.0f;
    float warmup = 0.0f;
    uint32_t max_triangles = 2000000;
    uint32_t n_materials = 0;  // rendering-only; accepted, not interpreted
    std::vector<ObstacleSpec> obstacles;  // extension
    // extension: CFL cap.  0 = off (the reference's fixed dt).  Otherwise each step of dt is run
    // as k equal substeps, k = ceil(|v|max dt / (cfl h)) (|v|max over the particles at the
    // step's start), at most MAX_CFL_SUBSTEPS.
    float cfl = 0.0f;
};
constexpr uint32_t MAX_CFL_SUBSTEPS = 64;
// Substeps of one step under the CFL cap, from the particles' largest squared speed (f32).
uint32_t cfl_substeps(float vmax2, float dt, float cfl, float h);

// Strict parse of one fluids[] object.  Returns "" on success, else the error message.
std::string parse_fluid(const std::string &json, FluidSpec &out);

// Scene::validate's fluid checks; index is i in "fluids[i] 'name': ...".  "" when valid.
// max_cells: the grid cell limit (the reference's 16,000,000 unless a caller raises it).
constexpr uint64_t REFERENCE_MAX_CELLS = 16000000ull;
std::string validate_fluid(const FluidSpec &f, uint32_t index, uint64_t max_cells = REFERENCE_MAX_CELLS);

// A whole scene document as far as the fluid path reads it (Scene, crates/pxl-scene/src/lib.rs:20-47):
// the render block (fps, seed; RenderSettings, lib.rs:610-638), every fluids[i] block, and what
// Scene::validate and `pxl check` (crates/pxl-cli/src/check.rs:31-106) report.  Nodes,
// simulations, world, plate and programs are checked for their outer type and the node tags
// (a camera must exist; mesh materials' programs must be identifiers with a programs block);
// their contents are outside the fluid path and are not interpreted.
constexpr uint32_t SCENE_VERSION = 1;  // lib.rs:18
struct SceneSpec {
    uint32_t version = 0;
    uint32_t resolution[2] = {640, 480};  // lib.rs:820
    float fps = 24.0f;  // render.fps (lib.rs:813)
    uint32_t seed = 0;  // render.seed
    bool has_motion_blur = false;
    float motion_blur = 0.0f;
    bool has_camera = false;
    uint32_t n_nodes = 0, n_simulations = 0;
    bool animated = false;
    int32_t frame_start = 0, frame_end = 0;
    std::vector<FluidSpec> fluids;
};
// Scene::from_json's deserialisation + validate for the above; "" when valid.
std::string parse_scene(const std::string &json, SceneSpec &out, uint64_t max_cells = 16000000ull);
// The summary line `pxl check` prints (check.rs:98-104, without the compiled-programs part).
std::string check_line(const SceneSpec &sc);

// Fluid::grid: cells per axis and cell size, glam Vec3 f32 arithmetic.
void fluid_grid(const FluidSpec &f, uint32_t n[3], float *h);

// (1.0 / (fps * substeps)) as f32 (fluid.rs:164).
float fluid_dt(double fps, uint32_t substeps);

// round((warmup + frame / fps) * fps * substeps).max(0) (fluid.rs:234-236).
uint64_t steps_for_frame(double warmup, double frame, double fps, uint32_t substeps);

// Obstacle cells (extension): ids[cell] = 1 + the index of the first obstacle (list order)
// whose sphere holds the centre of interior cell `cell` ((c - o).x^2 + (c - o).y^2 +
// (c - o).z^2 <= r^2 in f32, the sphere-fill test), else 0.  Returns the number of obstacle
// cells; ids is left empty without obstacles.
uint64_t obstacle_cells(const FluidSpec &f, std::vector<uint32_t> &ids);

// The interior cells whose centres can pass fill block b's test: index ranges [lo, hi) per axis
// inside [1, n - 1).  Exact for a box (a cell centre's f32 coordinate never decreases with its
// index, so the passing indices of an axis are contiguous); a sphere's bounding box widened by
// a cell (the f32 sphere test decides inside it).  The seeding loop visits nothing outside.
struct SeedRange {
    uint32_t lo[3], hi[3];
    uint64_t cells() const {
        return (uint64_t)(hi[0] - lo[0]) * (uint64_t)(hi[1] - lo[1]) * (uint64_t)(hi[2] - lo[2]);
    }
};
SeedRange seed_range(const FluidBlockSpec &b, const uint32_t n[3], float h, const float origin[3]);

// The reference's cell-centre test of a fill block (fluid.rs:127-133; a sphere's f32 distance
// test for the extension), f32 as the seeding loop evaluates it.
bool block_holds(const FluidBlockSpec &b, const float origin[3], float h, uint32_t i, uint32_t j, uint32_t k);

// Particles each fill block seeds (the RNG stream continues across blocks, so block b's first
// particle is the sum over the earlier blocks).  Obstacle ids as obstacle_cells (empty: none).
std::vector<uint64_t> seed_counts(const FluidSpec &f, const std::vector<uint32_t> &obstacle_ids);

// Reference seeding into SoA arrays (x[3][P], v[3][P]); returns the particle count.
// With out == nullptr only counts.  Obstacle cells (extension) receive no particles.
uint64_t seed_particles(const FluidSpec &f, uint32_t seed, std::vector<float> *x,
                        std::vector<float> *v);

}  // namespace pxl
