// Fluid block parsing, validation, grid derivation and seeding (see spec.hpp).
// Compiled with -ffp-contract=off: every f32 operation below rounds once, in source order,
// as the reference's Rust (which never contracts) does.
#include "spec.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cctype>
#include <set>

#include "json.hpp"

namespace pxl {
namespace {

const char *const kFluidFields[] = {"name", "bounds", "resolution", "fill", "particles_per_cell",
                                    "substeps", "flip", "gravity", "pressure_iterations",
                                    "surface_refine", "particle_radius", "warmup",
                                    "max_triangles", "materials", "obstacles", "cfl"};
const char *const kBlockFields[] = {"min", "max", "velocity"};
const char *const kSphereFields[] = {"center", "radius", "velocity"};  // extension
const char *const kObstacleFields[] = {"center", "radius"};              // extension
// Material and MaterialInputs (crates/pxl-scene/src/lib.rs:365-444, both deny_unknown_fields)
const char *const kMaterialFields[] = {"inputs", "program", "role"};
const char *const kMaterialInputFields[] = {
    "base_color", "base_color_texture", "specular_roughness", "specular_roughness_texture", "base_metalness",
    "base_metalness_texture", "emission_color", "emission_color_texture", "emission_luminance", "normal_texture",
    "normal_scale", "uv_scale", "transmission_weight", "specular_ior", "transmission_color", "transmission_depth"};
const char *const kRoles[] = {"surface", "shadow_catcher", "holdout"};  // Role, rename_all snake_case

template <size_t N>
std::string expected_list(const char *const (&fields)[N]) {
    std::string s;
    for (size_t k = 0; k < N; k++) {
        if (k) s += ", ";
        s += "`";
        s += fields[k];
        s += "`";
    }
    return s;
}

template <size_t N>
bool known(const char *const (&fields)[N], const std::string &key) {
    for (size_t k = 0; k < N; k++)
        if (key == fields[k]) return true;
    return false;
}

std::string get_u32(const JsonValue &v, uint32_t &out) {
    if (v.kind != JsonValue::Number) return "invalid type: " + json_describe(v) + ", expected u32";
    if (!v.is_integer) return "invalid type: " + json_describe(v) + ", expected u32";
    if (v.negative || !v.uint_fits || v.uint_value > 0xFFFFFFFFull)
        return "invalid value: " + json_describe(v) + ", expected u32";
    out = (uint32_t)v.uint_value;
    return "";
}

std::string get_f32(const JsonValue &v, float &out) {
    if (v.kind != JsonValue::Number) return "invalid type: " + json_describe(v) + ", expected f32";
    out = (float)v.num;  // serde_json: decimal -> f64 (correctly rounded) -> `as f32`
    return "";
}

std::string get_vec3(const JsonValue &v, float out[3]) {
    if (v.kind != JsonValue::Array) return "invalid type: " + json_describe(v) + ", expected an array of length 3";
    if (v.items.size() != 3)
        return "invalid length " + std::to_string(v.items.size()) + ", expected an array of length 3";
    for (int a = 0; a < 3; a++) {
        std::string e = get_f32(v.items[a], out[a]);
        if (!e.empty()) return e;
    }
    return "";
}

// A sphere fill (extension): an object with a `center` field.
std::string parse_sphere(const JsonValue &v, FluidBlockSpec &b) {
    std::set<std::string> seen;
    bool has_radius = false;
    b.sphere = true;
    for (const auto &m : v.members) {
        const std::string &k = m.first;
        if (!known(kSphereFields, k)) return "unknown field `" + k + "`, expected one of " + expected_list(kSphereFields);
        if (!seen.insert(k).second) return "duplicate field `" + k + "`";
        std::string e;
        if (k == "center") e = get_vec3(m.second, b.center);
        else if (k == "radius") { e = get_f32(m.second, b.radius); has_radius = true; }
        else e = get_vec3(m.second, b.velocity);
        if (!e.empty()) return e;
    }
    if (!has_radius) return "missing field `radius`";
    return "";
}

std::string parse_obstacle(const JsonValue &v, ObstacleSpec &o) {
    if (v.kind != JsonValue::Object) return "invalid type: " + json_describe(v) + ", expected struct Obstacle";
    std::set<std::string> seen;
    for (const auto &m : v.members) {
        const std::string &k = m.first;
        if (!known(kObstacleFields, k)) return "unknown field `" + k + "`, expected one of " + expected_list(kObstacleFields);
        if (!seen.insert(k).second) return "duplicate field `" + k + "`";
        std::string e = k == "center" ? get_vec3(m.second, o.center) : get_f32(m.second, o.radius);
        if (!e.empty()) return e;
    }
    if (!seen.count("center")) return "missing field `center`";
    if (!seen.count("radius")) return "missing field `radius`";
    return "";
}

std::string parse_block(const JsonValue &v, FluidBlockSpec &b) {
    if (v.kind != JsonValue::Object) return "invalid type: " + json_describe(v) + ", expected struct FluidBlock";
    for (const auto &m : v.members)
        if (m.first == "center") return parse_sphere(v, b);
    std::set<std::string> seen;
    bool has_min = false, has_max = false;
    for (const auto &m : v.members) {
        const std::string &k = m.first;
        if (!known(kBlockFields, k)) return "unknown field `" + k + "`, expected one of " + expected_list(kBlockFields);
        if (!seen.insert(k).second) return "duplicate field `" + k + "`";
        std::string e;
        if (k == "min") { e = get_vec3(m.second, b.min); has_min = true; }
        else if (k == "max") { e = get_vec3(m.second, b.max); has_max = true; }
        else { e = get_vec3(m.second, b.velocity); }
        if (!e.empty()) return e;
    }
    if (!has_min) return "missing field `min`";
    if (!has_max) return "missing field `max`";
    return "";
}

std::string get_floats(const JsonValue &v, float *out, size_t n) {
    const std::string len = "an array of length " + std::to_string(n);
    if (v.kind != JsonValue::Array) return "invalid type: " + json_describe(v) + ", expected " + len;
    if (v.items.size() != n) return "invalid length " + std::to_string(v.items.size()) + ", expected " + len;
    for (size_t a = 0; a < n; a++) {
        std::string e = get_f32(v.items[a], out[a]);
        if (!e.empty()) return e;
    }
    return "";
}

// Option<String>: null or a string.
std::string get_opt_string(const JsonValue &v) {
    if (v.kind == JsonValue::Null || v.kind == JsonValue::String) return "";
    return "invalid type: " + json_describe(v) + ", expected a string";
}

// The rendering material of a fluid (Material, crates/pxl-scene/src/lib.rs:365-377): parsed
// with the reference's field names and types so that a block the reference refuses is refused
// here too; its values do not enter the simulation.
std::string parse_material(const JsonValue &v) {
    if (v.kind != JsonValue::Object) return "invalid type: " + json_describe(v) + ", expected struct Material";
    std::set<std::string> seen;
    for (const auto &m : v.members) {
        const std::string &k = m.first;
        const JsonValue &x = m.second;
        if (!known(kMaterialFields, k)) return "unknown field `" + k + "`, expected one of " + expected_list(kMaterialFields);
        if (!seen.insert(k).second) return "duplicate field `" + k + "`";
        std::string e;
        if (k == "program") {
            e = get_opt_string(x);
        } else if (k == "role") {
            if (x.kind != JsonValue::String) e = "invalid type: " + json_describe(x) + ", expected enum Role";
            else if (!known(kRoles, x.text))
                e = "unknown variant `" + x.text + "`, expected one of " + expected_list(kRoles);
        } else {  // inputs (MaterialInputs, lib.rs:400-444)
            if (x.kind != JsonValue::Object) return "invalid type: " + json_describe(x) + ", expected struct MaterialInputs";
            std::set<std::string> iseen;
            for (const auto &im : x.members) {
                const std::string &ik = im.first;
                const JsonValue &iv = im.second;
                if (!known(kMaterialInputFields, ik))
                    return "unknown field `" + ik + "`, expected one of " + expected_list(kMaterialInputFields);
                if (!iseen.insert(ik).second) return "duplicate field `" + ik + "`";
                float tmp[3];
                const size_t len = ik.size();
                if (len > 8 && ik.compare(len - 8, 8, "_texture") == 0) e = get_opt_string(iv);
                else if (ik == "base_color" || ik == "emission_color" || ik == "transmission_color")
                    e = get_floats(iv, tmp, 3);
                else if (ik == "uv_scale") e = get_floats(iv, tmp, 2);
                else e = get_f32(iv, tmp[0]);
                if (!e.empty()) return e;
            }
        }
        if (!e.empty()) return e;
    }
    return "";
}

bool all_gt(const float a[3], const float b[3]) { return a[0] > b[0] && a[1] > b[1] && a[2] > b[2]; }
bool any_lt(const float a[3], const float b[3]) { return a[0] < b[0] || a[1] < b[1] || a[2] < b[2]; }

}  // namespace

namespace {
// The fields of one Fluid object (error messages without a prefix).
std::string parse_fluid_value(const JsonValue &root, FluidSpec &f) {
    if (root.kind != JsonValue::Object) return "invalid type: " + json_describe(root) + ", expected struct Fluid";
    f = FluidSpec();
    std::set<std::string> seen;
    bool has_name = false, has_bounds = false, has_fill = false;
    for (const auto &m : root.members) {
        const std::string &k = m.first;
        const JsonValue &v = m.second;
        std::string e;
        if (!known(kFluidFields, k)) {
            e = "unknown field `" + k + "`, expected one of " + expected_list(kFluidFields);
        } else if (!seen.insert(k).second) {
            e = "duplicate field `" + k + "`";
        } else if (k == "name") {
            if (v.kind != JsonValue::String) e = "invalid type: " + json_describe(v) + ", expected a string";
            else f.name = v.text;
            has_name = true;
        } else if (k == "bounds") {
            if (v.kind != JsonValue::Array) e = "invalid type: " + json_describe(v) + ", expected an array of length 2";
            else if (v.items.size() != 2) e = "invalid length " + std::to_string(v.items.size()) + ", expected an array of length 2";
            else {
                e = get_vec3(v.items[0], f.bounds[0]);
                if (e.empty()) e = get_vec3(v.items[1], f.bounds[1]);
            }
            has_bounds = true;
        } else if (k == "resolution") {
            e = get_u32(v, f.resolution);
        } else if (k == "fill") {
            if (v.kind != JsonValue::Array) e = "invalid type: " + json_describe(v) + ", expected a sequence";
            else {
                for (const auto &item : v.items) {
                    FluidBlockSpec b;
                    e = parse_block(item, b);
                    if (!e.empty()) break;
                    f.fill.push_back(b);
                }
            }
            has_fill = true;
        } else if (k == "particles_per_cell") {
            e = get_u32(v, f.particles_per_cell);
        } else if (k == "substeps") {
            e = get_u32(v, f.substeps);
        } else if (k == "flip") {
            e = get_f32(v, f.flip);
        } else if (k == "gravity") {
            e = get_vec3(v, f.gravity);
        } else if (k == "pressure_iterations") {
            e = get_u32(v, f.pressure_iterations);
        } else if (k == "surface_refine") {
            e = get_u32(v, f.surface_refine);
        } else if (k == "particle_radius") {
            e = get_f32(v, f.particle_radius);
        } else if (k == "warmup") {
            e = get_f32(v, f.warmup);
        } else if (k == "max_triangles") {
            e = get_u32(v, f.max_triangles);
        } else if (k == "cfl") {
            e = get_f32(v, f.cfl);
        } else if (k == "obstacles") {
            if (v.kind != JsonValue::Array) e = "invalid type: " + json_describe(v) + ", expected a sequence";
            else {
                for (const auto &item : v.items) {
                    ObstacleSpec o;
                    e = parse_obstacle(item, o);
                    if (!e.empty()) break;
                    f.obstacles.push_back(o);
                }
            }
        } else if (k == "materials") {
            // Rendering materials (OpenPBR inputs): checked field by field like the reference's
            // Material; they do not enter the simulation.
            if (v.kind != JsonValue::Array) e = "invalid type: " + json_describe(v) + ", expected a sequence";
            else {
                for (const auto &item : v.items) {
                    e = parse_material(item);
                    if (!e.empty()) break;
                }
                f.n_materials = (uint32_t)v.items.size();
            }
        }
        if (!e.empty()) return e;
    }
    if (!has_name) return "missing field `name`";
    if (!has_bounds) return "missing field `bounds`";
    if (!has_fill) return "missing field `fill`";
    return "";
}
}  // namespace

std::string parse_fluid(const std::string &json, FluidSpec &f) {
    JsonValue root;
    JsonError jerr;
    if (!json_parse(json, root, jerr))
        return "invalid fluid: " + jerr.message + " at line " + std::to_string(jerr.line) + " column " +
               std::to_string(jerr.column);
    std::string e = parse_fluid_value(root, f);
    return e.empty() ? e : "invalid fluid: " + e;
}

void fluid_grid(const FluidSpec &f, uint32_t n[3], float *h_out) {
    float size[3];
    for (int a = 0; a < 3; a++) size[a] = f.bounds[1][a] - f.bounds[0][a];
    float m = size[0];
    if (size[1] > m) m = size[1];
    if (size[2] > m) m = size[2];
    float h = m / (float)f.resolution;
    for (int a = 0; a < 3; a++) {
        float c = std::round(size[a] / h);  // f32::round (half away from zero)
        if (c < 4.0f) c = 4.0f;
        n[a] = (uint32_t)c;
    }
    *h_out = h;
}

std::string validate_fluid(const FluidSpec &f, uint32_t index, uint64_t max_cells) {
    std::string ctx = "fluids[" + std::to_string(index) + "] '" + f.name + "': ";
    const float *lo = f.bounds[0], *hi = f.bounds[1];
    if (!all_gt(hi, lo)) return ctx + "bounds must be [min, max] with max > min on every axis";
    bool flip_ok = f.flip >= 0.0f && f.flip <= 1.0f;  // NaN fails, as RangeInclusive::contains
    if (!(f.resolution >= 16 && f.resolution <= 256) ||
        !(f.particles_per_cell >= 1 && f.particles_per_cell <= 64) ||
        !(f.substeps >= 1 && f.substeps <= 64) || !(f.surface_refine >= 1 && f.surface_refine <= 3) ||
        !flip_ok || f.pressure_iterations == 0)
        return ctx + "resolution 16..256, particles_per_cell 1..64, substeps 1..64, surface_refine 1..3, flip 0..1";
    if (f.fill.empty()) return ctx + "needs at least one `fill` block of liquid";
    for (const auto &b : f.fill) {
        if (b.sphere) {
            float smin[3], smax[3];
            for (int a = 0; a < 3; a++) {
                smin[a] = b.center[a] - b.radius;
                smax[a] = b.center[a] + b.radius;
            }
            if (!(b.radius > 0.0f) || any_lt(smin, lo) || any_lt(hi, smax))
                return ctx + "each fill sphere must have radius > 0 and lie inside `bounds`";
        } else if (!all_gt(b.max, b.min) || any_lt(b.min, lo) || any_lt(hi, b.max)) {
            return ctx + "each fill block must have max > min and lie inside `bounds`";
        }
    }
    if (!(f.cfl >= 0.0f) || !std::isfinite(f.cfl)) return ctx + "cfl must be >= 0 (0 = off)";
    if (f.obstacles.size() > MAX_OBSTACLES) return ctx + "at most 65536 obstacles";
    for (const auto &o : f.obstacles)
        if (!(o.radius > 0.0f) || any_lt(o.center, lo) || any_lt(hi, o.center))
            return ctx + "each obstacle must have radius > 0 and its center inside `bounds`";
    uint32_t n[3];
    float h;
    fluid_grid(f, n, &h);
    if ((uint64_t)n[0] * n[1] * n[2] > max_cells) return ctx + "grid too large (lower resolution)";
    return "";
}

uint32_t cfl_substeps(float vmax2, float dt, float cfl, float h) {
    const double u = std::sqrt((double)vmax2);
    const double cells = u * (double)dt / ((double)cfl * (double)h);
    if (!(cells > 1.0)) return 1;  // also NaN
    if (cells >= (double)MAX_CFL_SUBSTEPS) return MAX_CFL_SUBSTEPS;
    return (uint32_t)std::ceil(cells);
}

float fluid_dt(double fps, uint32_t substeps) { return (float)(1.0 / (fps * (double)substeps)); }

uint64_t steps_for_frame(double warmup, double frame, double fps, uint32_t substeps) {
    double t = std::round((warmup + frame / fps) * fps * (double)substeps);
    return t > 0.0 ? (uint64_t)t : 0;
}

namespace {
inline float lcg_next(uint64_t &s) {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    return (float)(s >> 40) / (float)(1ull << 24);
}
}  // namespace

uint64_t obstacle_cells(const FluidSpec &f, std::vector<uint32_t> &ids) {
    ids.clear();
    if (f.obstacles.empty()) return 0;
    uint32_t n[3];
    float h;
    fluid_grid(f, n, &h);
    const float *org = f.bounds[0];
    ids.assign((size_t)n[0] * n[1] * n[2], 0u);
    uint64_t count = 0;
    // Obstacles in list order, each over the interior cells of its bounding box widened by one
    // cell (the f32 test below decides); a cell keeps the first obstacle that claims it.
    for (size_t q = 0; q < f.obstacles.size(); q++) {
        const ObstacleSpec &ob = f.obstacles[q];
        int lo[3], hi[3];
        for (int a = 0; a < 3; a++) {
            const double c = ((double)ob.center[a] - (double)org[a]) / (double)h - 0.5, r = (double)ob.radius / (double)h;
            lo[a] = std::max(1, (int)std::floor(c - r) - 1);
            hi[a] = std::min((int)n[a] - 2, (int)std::ceil(c + r) + 1);
        }
        for (int k = lo[2]; k <= hi[2]; k++)
            for (int j = lo[1]; j <= hi[1]; j++)
                for (int i = lo[0]; i <= hi[0]; i++) {
                    const size_t id = (size_t)i + (size_t)n[0] * ((size_t)j + (size_t)n[1] * (size_t)k);
                    if (ids[id]) continue;
                    const float cx = org[0] + ((float)i + 0.5f) * h;
                    const float cy = org[1] + ((float)j + 0.5f) * h;
                    const float cz = org[2] + ((float)k + 0.5f) * h;
                    const float dx = cx - ob.center[0], dy = cy - ob.center[1], dz = cz - ob.center[2];
                    const float d2 = dx * dx + dy * dy + dz * dz;
                    if (d2 <= ob.radius * ob.radius) {
                        ids[id] = (uint32_t)q + 1u;
                        count++;
                    }
                }
    }
    return count;
}

namespace {
// Scene (crates/pxl-scene/src/lib.rs:20-47, deny_unknown_fields) and RenderSettings (lib.rs:610-638)
const char *const kSceneFields[] = {"version", "nodes", "world", "render", "animation", "programs",
                                    "simulations", "plate", "fluids"};
const char *const kRenderFields[] = {"resolution", "samples", "bounces", "exposure", "view", "seed",
                                     "fps", "motion_blur", "denoise", "transparent"};
const char *const kViews[] = {"agx", "aces", "standard"};  // View, rename_all lowercase
const char *const kNodeTypes[] = {"mesh", "asset", "light", "camera"};  // Node, tag "type"

// is_ident (lib.rs): a Slang function name
bool is_ident(const std::string &p) {
    if (p.empty() || !(std::isalpha((unsigned char)p[0]) || p[0] == '_')) return false;
    for (char c : p)
        if (!(std::isalnum((unsigned char)c) || c == '_')) return false;
    return true;
}

std::string get_i32(const JsonValue &v, int32_t &out) {
    if (v.kind != JsonValue::Number || !v.is_integer) return "invalid type: " + json_describe(v) + ", expected i32";
    if (v.num < -2147483648.0 || v.num > 2147483647.0) return "invalid value: " + json_describe(v) + ", expected i32";
    out = (int32_t)v.num;
    return "";
}

const JsonValue *member(const JsonValue &o, const char *key) {
    for (const auto &m : o.members)
        if (m.first == key) return &m.second;
    return nullptr;
}

std::string parse_render(const JsonValue &v, SceneSpec &sc) {
    if (v.kind != JsonValue::Object) return "invalid type: " + json_describe(v) + ", expected struct RenderSettings";
    std::set<std::string> seen;
    for (const auto &m : v.members) {
        const std::string &k = m.first;
        const JsonValue &x = m.second;
        if (!known(kRenderFields, k)) return "unknown field `" + k + "`, expected one of " + expected_list(kRenderFields);
        if (!seen.insert(k).second) return "duplicate field `" + k + "`";
        std::string e;
        float f32 = 0.0f;
        uint32_t u = 0;
        if (k == "resolution") {
            if (x.kind != JsonValue::Array) e = "invalid type: " + json_describe(x) + ", expected an array of length 2";
            else if (x.items.size() != 2) e = "invalid length " + std::to_string(x.items.size()) + ", expected an array of length 2";
            else {
                e = get_u32(x.items[0], sc.resolution[0]);
                if (e.empty()) e = get_u32(x.items[1], sc.resolution[1]);
            }
        } else if (k == "samples" || k == "bounces") {
            if (x.kind != JsonValue::Null) e = get_u32(x, u);
        } else if (k == "exposure") {
            e = get_f32(x, f32);
        } else if (k == "view") {
            if (x.kind != JsonValue::String) e = "invalid type: " + json_describe(x) + ", expected enum View";
            else if (!known(kViews, x.text)) e = "unknown variant `" + x.text + "`, expected one of " + expected_list(kViews);
        } else if (k == "seed") {
            e = get_u32(x, sc.seed);
        } else if (k == "fps") {
            e = get_f32(x, sc.fps);
        } else if (k == "motion_blur") {
            if (x.kind != JsonValue::Null) {
                e = get_f32(x, sc.motion_blur);
                sc.has_motion_blur = true;
            }
        } else if (k == "denoise" || k == "transparent") {
            const bool opt = k == "denoise";
            if (!(x.kind == JsonValue::Bool || (opt && x.kind == JsonValue::Null)))
                e = "invalid type: " + json_describe(x) + ", expected a boolean";
        }
        if (!e.empty()) return e;
    }
    return "";
}

// Animation::end (crates/pxl-scene/src/anim.rs:24-35): frame_end, else the last key (or orbit
// end frame) of any track, at least frame_start, rounded up.
std::string parse_animation(const JsonValue &v, SceneSpec &sc) {
    if (v.kind == JsonValue::Null) return "";
    if (v.kind != JsonValue::Object) return "invalid type: " + json_describe(v) + ", expected struct Animation";
    sc.animated = true;
    int32_t start = 0, end = 0;
    bool has_end = false;
    if (const JsonValue *x = member(v, "frame_start")) {
        std::string e = get_i32(*x, start);
        if (!e.empty()) return e;
    }
    if (const JsonValue *x = member(v, "frame_end")) {
        if (x->kind != JsonValue::Null) {
            std::string e = get_i32(*x, end);
            if (!e.empty()) return e;
            has_end = true;
        }
    }
    if (!has_end) {
        double last = (double)start;
        if (const JsonValue *tracks = member(v, "tracks")) {
            if (tracks->kind != JsonValue::Array) return "invalid type: " + json_describe(*tracks) + ", expected a sequence";
            for (const auto &t : tracks->items) {
                if (t.kind != JsonValue::Object) continue;
                const JsonValue *orbit = member(t, "orbit");
                double f = 0.0;
                if (orbit && orbit->kind == JsonValue::Object) {
                    const JsonValue *fr = member(*orbit, "frames");
                    if (fr && fr->kind == JsonValue::Array && fr->items.size() == 2 && fr->items[1].kind == JsonValue::Number)
                        f = fr->items[1].num;
                } else if (const JsonValue *keys = member(t, "keys")) {
                    if (keys->kind == JsonValue::Array && !keys->items.empty()) {
                        const JsonValue &k = keys->items.back();
                        if (k.kind == JsonValue::Array && !k.items.empty() && k.items[0].kind == JsonValue::Number)
                            f = k.items[0].num;
                    }
                }
                last = std::max(last, f);
            }
        }
        end = (int32_t)std::ceil(last);
    }
    sc.frame_start = start;
    sc.frame_end = end;
    return "";
}
}  // namespace

std::string parse_scene(const std::string &json, SceneSpec &sc, uint64_t max_cells) {
    JsonValue root;
    JsonError jerr;
    sc = SceneSpec();
    if (!json_parse(json, root, jerr))
        return "invalid scene: " + jerr.message + " at line " + std::to_string(jerr.line) + " column " +
               std::to_string(jerr.column);
    if (root.kind != JsonValue::Object) return "invalid scene: invalid type: " + json_describe(root) + ", expected struct Scene";
    std::set<std::string> seen;
    bool has_version = false, has_programs = false;
    const JsonValue *nodes = nullptr;
    for (const auto &m : root.members) {
        const std::string &k = m.first;
        const JsonValue &v = m.second;
        std::string e;
        if (!known(kSceneFields, k)) {
            e = "unknown field `" + k + "`, expected one of " + expected_list(kSceneFields);
        } else if (!seen.insert(k).second) {
            e = "duplicate field `" + k + "`";
        } else if (k == "version") {
            e = get_u32(v, sc.version);
            has_version = true;
        } else if (k == "render") {
            e = parse_render(v, sc);
        } else if (k == "fluids") {
            if (v.kind != JsonValue::Array) e = "invalid type: " + json_describe(v) + ", expected a sequence";
            else {
                for (const auto &item : v.items) {
                    FluidSpec f;
                    e = parse_fluid_value(item, f);
                    if (!e.empty()) break;
                    sc.fluids.push_back(f);
                }
            }
        } else if (k == "animation") {
            e = parse_animation(v, sc);
        } else if (k == "nodes" || k == "simulations") {
            if (v.kind != JsonValue::Array) e = "invalid type: " + json_describe(v) + ", expected a sequence";
            else if (k == "nodes") {
                nodes = &v;
                for (const auto &n : v.items) {
                    const JsonValue *t = n.kind == JsonValue::Object ? member(n, "type") : nullptr;
                    if (!t) {
                        e = n.kind == JsonValue::Object ? "missing field `type`"
                                                        : "invalid type: " + json_describe(n) + ", expected internally tagged enum Node";
                        break;
                    }
                    if (t->kind != JsonValue::String || !known(kNodeTypes, t->text)) {
                        e = "unknown variant `" + (t->kind == JsonValue::String ? t->text : json_describe(*t)) +
                            "`, expected one of " + expected_list(kNodeTypes);
                        break;
                    }
                    if (t->text == "camera") sc.has_camera = true;
                }
                sc.n_nodes = (uint32_t)v.items.size();
            } else {
                sc.n_simulations = (uint32_t)v.items.size();
            }
        } else if (k == "programs") {
            has_programs = v.kind != JsonValue::Null;
            if (has_programs && v.kind != JsonValue::Object) e = "invalid type: " + json_describe(v) + ", expected struct Programs";
        } else {  // world, plate: objects the fluid path does not read
            const bool opt = k == "plate";
            if (!(v.kind == JsonValue::Object || (opt && v.kind == JsonValue::Null)))
                e = "invalid type: " + json_describe(v) + ", expected struct " + (k == "world" ? "World" : "Plate");
        }
        if (!e.empty()) return "invalid scene: " + e;
    }
    if (!has_version) return "invalid scene: missing field `version`";
    // Scene::validate (lib.rs:659-745), the checks on what the fluid path reads, in its order
    if (sc.version != SCENE_VERSION)
        return "scene version " + std::to_string(sc.version) + " is not supported (expected " +
               std::to_string(SCENE_VERSION) + ")";
    if (!sc.has_camera)
        return "scene has no camera node; add {\"type\":\"camera\",\"translation\":[x,y,z],\"target\":[0,0,0]}";
    const uint32_t w = sc.resolution[0], h = sc.resolution[1];
    if (w == 0 || h == 0 || w > 16384 || h > 16384)
        return "render.resolution must be within 1..16384, got [" + std::to_string(w) + "," + std::to_string(h) + "]";
    if (!(sc.fps > 0.0f)) return "render.fps must be > 0";
    if (nodes) {  // material programs of mesh nodes (lib.rs:674-686)
        for (const auto &n : nodes->items) {
            const JsonValue *t = member(n, "type");
            const JsonValue *mats = member(n, "materials");
            if (!t || t->text != "mesh" || !mats || mats->kind != JsonValue::Array) continue;
            for (const auto &mat : mats->items) {
                const JsonValue *p = mat.kind == JsonValue::Object ? member(mat, "program") : nullptr;
                if (!p || p->kind != JsonValue::String) continue;
                if (!is_ident(p->text)) return "material program '" + p->text + "' must be the name of a Slang function";
                if (!has_programs)
                    return "a material uses program '" + p->text + "' but the scene has no `programs` block";
            }
        }
    }
    for (size_t i = 0; i < sc.fluids.size(); i++) {
        std::string e = validate_fluid(sc.fluids[i], (uint32_t)i, max_cells);
        if (!e.empty()) return e;
    }
    if (sc.has_motion_blur && !(sc.motion_blur >= 0.0f && sc.motion_blur <= 1.0f)) {
        char buf[64];
        std::snprintf(buf, sizeof buf, "%g", (double)sc.motion_blur);
        return std::string("render.motion_blur is the shutter fraction of a frame, 0..1 (0.5 = 180\u00b0), got ") + buf;
    }
    return "";
}

std::string check_line(const SceneSpec &sc) {
    const int32_t a = sc.animated ? sc.frame_start : 0, b = sc.animated ? sc.frame_end : 0;
    return "ok: " + std::to_string(sc.n_nodes) + " nodes, frames " + std::to_string(a) + ".." + std::to_string(b) + ", " +
           std::to_string(sc.n_simulations) + " simulation(s), " + std::to_string(sc.fluids.size()) + " fluid(s)";
}

SeedRange seed_range(const FluidBlockSpec &b, const uint32_t n[3], float h, const float origin[3]) {
    SeedRange r;
    for (int a = 0; a < 3; a++) {
        r.lo[a] = r.hi[a] = 1;
        bool any = false;
        for (uint32_t idx = 1; idx + 1 < n[a]; idx++) {
            const float c = origin[a] + ((float)idx + 0.5f) * h;
            bool in;
            if (b.sphere) {  // bounding box of the ball, one cell wider on each side
                const double d = (double)c - (double)b.center[a];
                in = std::fabs(d) <= (double)b.radius + (double)h;
            } else {
                in = !(c < b.min[a] || c > b.max[a]);
            }
            if (!in) continue;
            if (!any) r.lo[a] = idx;
            r.hi[a] = idx + 1;
            any = true;
        }
        if (!any) r.lo[a] = r.hi[a] = 1;
    }
    return r;
}

bool block_holds(const FluidBlockSpec &b, const float o[3], float h, uint32_t i, uint32_t j, uint32_t k) {
    const float cx = o[0] + ((float)i + 0.5f) * h;
    const float cy = o[1] + ((float)j + 0.5f) * h;
    const float cz = o[2] + ((float)k + 0.5f) * h;
    if (b.sphere) {  // extension: centre within the sphere (f32, x + y + z)
        const float dx = cx - b.center[0], dy = cy - b.center[1], dz = cz - b.center[2];
        const float d2 = dx * dx + dy * dy + dz * dz;
        return d2 <= b.radius * b.radius;
    }
    return !(cx < b.min[0] || cx > b.max[0] || cy < b.min[1] || cy > b.max[1] || cz < b.min[2] || cz > b.max[2]);
}

std::vector<uint64_t> seed_counts(const FluidSpec &f, const std::vector<uint32_t> &obst) {
    uint32_t n[3];
    float h;
    fluid_grid(f, n, &h);
    const float *o = f.bounds[0];
    std::vector<uint64_t> out;
    for (const auto &b : f.fill) {
        const SeedRange r = seed_range(b, n, h, o);
        uint64_t cells = 0;
        if (!b.sphere && obst.empty()) {
            cells = r.cells();  // every cell of a box's range passes
        } else {
            for (uint32_t k = r.lo[2]; k < r.hi[2]; k++)
                for (uint32_t j = r.lo[1]; j < r.hi[1]; j++)
                    for (uint32_t i = r.lo[0]; i < r.hi[0]; i++) {
                        if (!block_holds(b, o, h, i, j, k)) continue;
                        if (!obst.empty() && obst[(size_t)i + (size_t)n[0] * ((size_t)j + (size_t)n[1] * (size_t)k)])
                            continue;  // extension: no liquid inside obstacles
                        cells++;
                    }
        }
        out.push_back(cells * f.particles_per_cell);
    }
    return out;
}

uint64_t seed_particles(const FluidSpec &f, uint32_t seed, std::vector<float> *x, std::vector<float> *v) {
    std::vector<uint32_t> obst;
    obstacle_cells(f, obst);
    if (!x) {
        uint64_t count = 0;
        for (uint64_t c : seed_counts(f, obst)) count += c;
        return count;
    }
    uint32_t n[3];
    float h;
    fluid_grid(f, n, &h);
    const float *o = f.bounds[0];
    uint64_t rng = 0x9e3779b97f4a7c15ull ^ (uint64_t)seed;
    uint64_t count = 0;
    std::vector<float> xs[3], vs[3];
    // the reference loop (k outer, j, i inner over the interior, fluid.rs:123-148), restricted
    // to each block's seed range: cells outside it fail the block's test and draw nothing
    for (const auto &b : f.fill) {
        const SeedRange r = seed_range(b, n, h, o);
        for (uint32_t k = r.lo[2]; k < r.hi[2]; k++)
            for (uint32_t j = r.lo[1]; j < r.hi[1]; j++)
                for (uint32_t i = r.lo[0]; i < r.hi[0]; i++) {
                    if (!block_holds(b, o, h, i, j, k)) continue;
                    if (!obst.empty() && obst[(size_t)i + (size_t)n[0] * ((size_t)j + (size_t)n[1] * (size_t)k)])
                        continue;  // extension: no liquid inside obstacles
                    for (uint32_t q = 0; q < f.particles_per_cell; q++) {
                        float px = o[0] + ((float)i + lcg_next(rng)) * h;
                        float py = o[1] + ((float)j + lcg_next(rng)) * h;
                        float pz = o[2] + ((float)k + lcg_next(rng)) * h;
                        xs[0].push_back(px); xs[1].push_back(py); xs[2].push_back(pz);
                        for (int a = 0; a < 3; a++) vs[a].push_back(b.velocity[a]);
                        count++;
                    }
                }
    }
    x->clear();
    v->clear();
    x->reserve(3 * count);
    v->reserve(3 * count);
    for (int a = 0; a < 3; a++) {
        x->insert(x->end(), xs[a].begin(), xs[a].end());
        v->insert(v->end(), vs[a].begin(), vs[a].end());
    }
    return count;
}

}  // namespace pxl
