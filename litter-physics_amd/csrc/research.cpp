// Research fixtures over the engine: the wire request, the fixture build, the
// record loop, the observables and the checkpoint codec (core/src/research.rs,
// core/src/mpm/fixtures.rs:94-615), and run_json (core/src/lib.rs:11-49).
//
// Everything here is host code on f64 state.  The engine (the C ABI of
// include/mpm_engine.h) owns the device state; this layer reads it back only
// at record boundaries and at the end of a segment, as the reference's
// research::run reads its Simulation.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mpm_engine.h"
#include "host_math.h"
#include "json.h"

namespace {

using mpm_json::Value;

struct ResearchError : std::runtime_error {
    int code;
    ResearchError(const std::string& m, int c = MPM_ERR_INVALID) : std::runtime_error(m), code(c) {}
};
[[noreturn]] void bail(const std::string& m, int code = MPM_ERR_INVALID) { throw ResearchError(m, code); }

// Wire caps of the research mode (research.rs:21-23).
constexpr size_t kMaxParticles = 100000;
constexpr size_t kMaxNodes = 1000000;
constexpr double kMaxFramePoints = 2000000.0;
constexpr int kParticlesPerAxis = 2;       // fixtures.rs:38
constexpr double kJitterFraction = 0.1;    // fixtures.rs:40
constexpr double kMinCellsPerRadius = 1.5; // fixtures.rs:42
constexpr double kPelletGapCells = 2.0;    // fixtures.rs:44

// f64 as Rust's Display prints it (shortest round-trip digits, positional),
// for error messages.
std::string show(double v) {
    if (std::isnan(v)) return "NaN";
    if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
    std::string s;
    mpm_json::write_f64(s, v);
    size_t e = s.find('e');
    if (e != std::string::npos) {
        // Expand d.ddde±X into positional notation.
        bool neg = s[0] == '-';
        std::string mant = s.substr(neg ? 1 : 0, e - (neg ? 1 : 0));
        int ex = std::stoi(s.substr(e + 1));
        std::string digits;
        int point = 0;
        for (char c : mant) {
            if (c == '.') continue;
            digits += c;
        }
        point = (int)(mant.find('.') == std::string::npos ? mant.size() : mant.find('.')) + ex;
        std::string out = neg ? "-" : "";
        if (point <= 0) {
            out += "0.";
            out.append((size_t)(-point), '0');
            out += digits;
        } else if (point >= (int)digits.size()) {
            out += digits;
            out.append((size_t)(point - (int)digits.size()), '0');
        } else {
            out += digits.substr(0, (size_t)point) + "." + digits.substr((size_t)point);
        }
        return out;
    }
    if (s.size() > 2 && s.compare(s.size() - 2, 2, ".0") == 0) s.resize(s.size() - 2);
    return s;
}

// ------------------------------------------------------------ typed reading
// serde's derive with deny_unknown_fields: every field present (Option fields
// may be missing or null), no other field, numbers of the declared type.

const char* type_name(const Value& v) {
    switch (v.kind) {
        case Value::Null: return "null";
        case Value::Bool: return "boolean";
        case Value::UInt: return "integer";
        case Value::Int: return "integer";
        case Value::Float: return "floating point";
        case Value::String: return "string";
        case Value::Array: return "sequence";
        case Value::Object: return "map";
    }
    return "value";
}

void expect_object(const Value& v, const char* what) {
    if (v.kind != Value::Object) bail(std::string("invalid type: ") + type_name(v) + ", expected struct " + what);
}

void deny_unknown(const Value& obj, std::initializer_list<const char*> fields) {
    for (const auto& kv : obj.o) {
        bool known = false;
        for (const char* f : fields) known = known || kv.first == f;
        if (!known) {
            std::string list;
            for (const char* f : fields) list += std::string(list.empty() ? "" : ", ") + "`" + f + "`";
            bail("unknown field `" + kv.first + "`, expected one of " + list);
        }
    }
}

const Value& field(const Value& obj, const char* name) {
    const Value* v = obj.find(name);
    if (!v) bail(std::string("missing field `") + name + "`");
    return *v;
}

double as_f64(const Value& v, const char* name) {
    switch (v.kind) {
        case Value::UInt: return (double)v.u;
        case Value::Int: return (double)v.i;
        case Value::Float: return v.d;
        default: bail(std::string("invalid type: ") + type_name(v) + ", expected f64 for `" + name + "`");
    }
}
uint64_t as_u64(const Value& v, const char* name) {
    if (v.kind != Value::UInt) bail(std::string("invalid type: ") + type_name(v) + ", expected u64 for `" + name + "`");
    return v.u;
}
uint32_t as_u32(const Value& v, const char* name) {
    uint64_t u = as_u64(v, name);
    if (u > 0xFFFFFFFFull) bail(std::string("invalid value: integer `") + std::to_string(u) + "`, expected u32");
    return (uint32_t)u;
}
std::string as_str(const Value& v, const char* name) {
    if (v.kind != Value::String) bail(std::string("invalid type: ") + type_name(v) + ", expected a string for `" + name + "`");
    return v.s;
}
const std::vector<Value>& as_seq(const Value& v, const char* name) {
    if (v.kind != Value::Array) bail(std::string("invalid type: ") + type_name(v) + ", expected a sequence for `" + name + "`");
    return v.a;
}
void as_fixed(const Value& v, const char* name, double* out, size_t n) {
    const auto& a = as_seq(v, name);
    if (a.size() != n)
        bail("invalid length " + std::to_string(a.size()) + ", expected an array of length " + std::to_string(n) +
             " for `" + name + "`");
    for (size_t k = 0; k < n; ++k) out[k] = as_f64(a[k], name);
}

// ------------------------------------------------------------ the request
// household/request.rs:43-99 (the shared envelope) and research.rs:26-40.

struct BoxSpec {
    std::string id;
    double size_m[3], origin_m[3];
    double slot_width_m, slot_length_m, slot_pitch_m, drawer_depth_m;
    uint64_t pellet_count;
    double pellet_radius_m, pellet_length_m, pellet_density_kg_m3;

    bool operator==(const BoxSpec& o) const {
        return id == o.id && std::equal(size_m, size_m + 3, o.size_m) && std::equal(origin_m, origin_m + 3, o.origin_m) &&
               slot_width_m == o.slot_width_m && slot_length_m == o.slot_length_m && slot_pitch_m == o.slot_pitch_m &&
               drawer_depth_m == o.drawer_depth_m && pellet_count == o.pellet_count &&
               pellet_radius_m == o.pellet_radius_m && pellet_length_m == o.pellet_length_m &&
               pellet_density_kg_m3 == o.pellet_density_kg_m3;
    }
};

struct Materials {
    double friction, restitution, normal_stiffness_n_m, water_capacity_ratio, uptake_rate_s, breakdown_rate_s,
        evaporation_rate_s;
    bool operator==(const Materials& o) const {
        return friction == o.friction && restitution == o.restitution && normal_stiffness_n_m == o.normal_stiffness_n_m &&
               water_capacity_ratio == o.water_capacity_ratio && uptake_rate_s == o.uptake_rate_s &&
               breakdown_rate_s == o.breakdown_rate_s && evaporation_rate_s == o.evaporation_rate_s;
    }
};

struct Event {
    double time_s;
    std::string kind, box_id;
    double position_m[3], direction[3];
    double duration_s, amount_kg, water_fraction, radius_m;
    bool operator==(const Event& o) const {
        return time_s == o.time_s && kind == o.kind && box_id == o.box_id &&
               std::equal(position_m, position_m + 3, o.position_m) && std::equal(direction, direction + 3, o.direction) &&
               duration_s == o.duration_s && amount_kg == o.amount_kg && water_fraction == o.water_fraction &&
               radius_m == o.radius_m;
    }
};

struct Request {
    uint32_t schema_version;
    std::string mode;
    uint64_t seed;
    double duration_s, dt_s, record_interval_s, max_wall_time_s;
    std::vector<BoxSpec> boxes;
    Materials materials;
    std::vector<Event> events;
    bool has_research = false;
    Value research;

    bool operator==(const Request& o) const {
        return schema_version == o.schema_version && mode == o.mode && seed == o.seed && duration_s == o.duration_s &&
               dt_s == o.dt_s && record_interval_s == o.record_interval_s && max_wall_time_s == o.max_wall_time_s &&
               boxes == o.boxes && materials == o.materials && events == o.events && has_research == o.has_research &&
               research == o.research;
    }
};

BoxSpec parse_box(const Value& v) {
    expect_object(v, "BoxSpec");
    deny_unknown(v, {"id", "size_m", "origin_m", "slot_width_m", "slot_length_m", "slot_pitch_m", "drawer_depth_m",
                     "pellet_count", "pellet_radius_m", "pellet_length_m", "pellet_density_kg_m3"});
    BoxSpec b;
    b.id = as_str(field(v, "id"), "id");
    as_fixed(field(v, "size_m"), "size_m", b.size_m, 3);
    as_fixed(field(v, "origin_m"), "origin_m", b.origin_m, 3);
    b.slot_width_m = as_f64(field(v, "slot_width_m"), "slot_width_m");
    b.slot_length_m = as_f64(field(v, "slot_length_m"), "slot_length_m");
    b.slot_pitch_m = as_f64(field(v, "slot_pitch_m"), "slot_pitch_m");
    b.drawer_depth_m = as_f64(field(v, "drawer_depth_m"), "drawer_depth_m");
    b.pellet_count = as_u64(field(v, "pellet_count"), "pellet_count");
    b.pellet_radius_m = as_f64(field(v, "pellet_radius_m"), "pellet_radius_m");
    b.pellet_length_m = as_f64(field(v, "pellet_length_m"), "pellet_length_m");
    b.pellet_density_kg_m3 = as_f64(field(v, "pellet_density_kg_m3"), "pellet_density_kg_m3");
    return b;
}

Materials parse_materials(const Value& v) {
    expect_object(v, "Materials");
    deny_unknown(v, {"friction", "restitution", "normal_stiffness_n_m", "water_capacity_ratio", "uptake_rate_s",
                     "breakdown_rate_s", "evaporation_rate_s"});
    Materials m;
    m.friction = as_f64(field(v, "friction"), "friction");
    m.restitution = as_f64(field(v, "restitution"), "restitution");
    m.normal_stiffness_n_m = as_f64(field(v, "normal_stiffness_n_m"), "normal_stiffness_n_m");
    m.water_capacity_ratio = as_f64(field(v, "water_capacity_ratio"), "water_capacity_ratio");
    m.uptake_rate_s = as_f64(field(v, "uptake_rate_s"), "uptake_rate_s");
    m.breakdown_rate_s = as_f64(field(v, "breakdown_rate_s"), "breakdown_rate_s");
    m.evaporation_rate_s = as_f64(field(v, "evaporation_rate_s"), "evaporation_rate_s");
    return m;
}

Event parse_event(const Value& v) {
    expect_object(v, "Event");
    deny_unknown(v, {"time_s", "kind", "box_id", "position_m", "direction", "duration_s", "amount_kg", "water_fraction",
                     "radius_m"});
    static const std::set<std::string> kinds = {"urinate", "defecate", "dig",         "cover", "stir",
                                                "scoop",   "empty_drawer", "refill", "clean"};
    Event e;
    e.time_s = as_f64(field(v, "time_s"), "time_s");
    e.kind = as_str(field(v, "kind"), "kind");
    if (!kinds.count(e.kind)) bail("unknown variant `" + e.kind + "` for `kind`");
    e.box_id = as_str(field(v, "box_id"), "box_id");
    as_fixed(field(v, "position_m"), "position_m", e.position_m, 3);
    as_fixed(field(v, "direction"), "direction", e.direction, 3);
    e.duration_s = as_f64(field(v, "duration_s"), "duration_s");
    e.amount_kg = as_f64(field(v, "amount_kg"), "amount_kg");
    e.water_fraction = as_f64(field(v, "water_fraction"), "water_fraction");
    e.radius_m = as_f64(field(v, "radius_m"), "radius_m");
    return e;
}

Request parse_request(const Value& v) {
    expect_object(v, "Request");
    deny_unknown(v, {"schema_version", "mode", "seed", "duration_s", "dt_s", "record_interval_s", "max_wall_time_s",
                     "boxes", "materials", "events", "research"});
    Request r;
    r.schema_version = as_u32(field(v, "schema_version"), "schema_version");
    r.mode = as_str(field(v, "mode"), "mode");
    r.seed = as_u64(field(v, "seed"), "seed");
    r.duration_s = as_f64(field(v, "duration_s"), "duration_s");
    r.dt_s = as_f64(field(v, "dt_s"), "dt_s");
    r.record_interval_s = as_f64(field(v, "record_interval_s"), "record_interval_s");
    r.max_wall_time_s = as_f64(field(v, "max_wall_time_s"), "max_wall_time_s");
    for (const Value& b : as_seq(field(v, "boxes"), "boxes")) r.boxes.push_back(parse_box(b));
    r.materials = parse_materials(field(v, "materials"));
    for (const Value& e : as_seq(field(v, "events"), "events")) r.events.push_back(parse_event(e));
    const Value* res = v.find("research");
    r.has_research = res && res->kind != Value::Null;
    if (r.has_research) r.research = *res;
    return r;
}

// household/request.rs:140-202 range checks.
void finite(const std::string& name, double v) {
    if (!std::isfinite(v)) bail("`" + name + "` must be finite, got '" + show(v) + "'");
}
void positive(const std::string& name, double v) {
    finite(name, v);
    if (!(v > 0.0)) bail("`" + name + "` must be positive, got '" + show(v) + "'");
}
void non_negative(const std::string& name, double v) {
    finite(name, v);
    if (!(v >= 0.0)) bail("`" + name + "` must be non-negative, got '" + show(v) + "'");
}
void within(const std::string& name, double v, double lo, double hi) {
    finite(name, v);
    if (!(v >= lo && v <= hi)) bail("`" + name + "` must be in [" + show(lo) + ", " + show(hi) + "], got '" + show(v) + "'");
}

void validate_materials(const Materials& m) {
    non_negative("materials.friction", m.friction);
    within("materials.restitution", m.restitution, 0.0, 1.0);
    positive("materials.normal_stiffness_n_m", m.normal_stiffness_n_m);
    positive("materials.water_capacity_ratio", m.water_capacity_ratio);
    non_negative("materials.uptake_rate_s", m.uptake_rate_s);
    non_negative("materials.breakdown_rate_s", m.breakdown_rate_s);
    non_negative("materials.evaporation_rate_s", m.evaporation_rate_s);
}

// BoxSpec::validate, household/request.rs:217-273 (dem.rs:35 caps the
// multisphere at 16 spheres).
void validate_box(const BoxSpec& b) {
    std::string tag = "boxes['" + b.id + "']";
    if (b.id.empty()) bail("box `id` must not be empty");
    const char* axes[3] = {"x", "y", "z"};
    for (int a = 0; a < 3; ++a) positive(tag + ".size_m." + axes[a], b.size_m[a]);
    for (int a = 0; a < 3; ++a) finite(tag + ".origin_m." + axes[a], b.origin_m[a]);
    positive(tag + ".slot_width_m", b.slot_width_m);
    positive(tag + ".slot_length_m", b.slot_length_m);
    positive(tag + ".slot_pitch_m", b.slot_pitch_m);
    positive(tag + ".drawer_depth_m", b.drawer_depth_m);
    positive(tag + ".pellet_radius_m", b.pellet_radius_m);
    positive(tag + ".pellet_length_m", b.pellet_length_m);
    positive(tag + ".pellet_density_kg_m3", b.pellet_density_kg_m3);
    if (b.slot_width_m >= b.slot_pitch_m) bail(tag + ": `slot_width_m` must be smaller than `slot_pitch_m`");
    if (b.slot_pitch_m > b.size_m[1]) bail(tag + ": `slot_pitch_m` exceeds the box `size_m.y`");
    if (b.slot_length_m + 2.0 * b.slot_width_m > b.size_m[0]) bail(tag + ": one slot plus its bars does not fit `size_m.x`");
    double bounding = std::max(b.pellet_length_m, 2.0 * b.pellet_radius_m);
    if (bounding >= std::min(b.size_m[0], b.size_m[1]) || bounding >= b.size_m[2])
        bail(tag + ": a pellet does not fit inside the box");
    if (b.pellet_length_m / b.pellet_radius_m > 16.0) bail(tag + ": pellet aspect ratio exceeds 16 spheres");
    double pitch = bounding * 1.05;
    double sites = std::floor(b.size_m[0] / pitch) * std::floor(b.size_m[1] / pitch);
    uint64_t per_layer = (sites >= 0.0 && sites < 1e12) ? (uint64_t)sites : 0;
    if (per_layer == 0) bail(tag + ": no pellet lattice site fits the footprint");
    uint64_t layers = (b.pellet_count + per_layer - 1) / per_layer;
    if ((double)layers * bounding > b.size_m[2])
        bail(tag + ": `pellet_count` " + std::to_string(b.pellet_count) + " does not fit the box height as a loose lattice");
}

// validate_envelope, research.rs:135-167.
void validate_envelope(const Request& r) {
    if (r.schema_version != 1 || r.mode != "research" || !r.events.empty())
        bail("research requires schema_version=1, mode=research and no household events");
    for (double v : {r.duration_s, r.dt_s, r.record_interval_s, r.max_wall_time_s})
        if (!std::isfinite(v) || v <= 0.0) bail("research times and budgets must be finite and positive");
    if (r.dt_s > r.duration_s || r.max_wall_time_s > 3600.0)
        bail("research dt exceeds duration or budget exceeds 3600 seconds");
    if (r.duration_s / r.record_interval_s > 10000.0 || r.duration_s / r.dt_s > 10000000.0)
        bail("research recording or nominal-step resource cap exceeded");
    if (r.boxes.size() > 1) bail("research fixtures support at most one box supplying pellet geometry");
    for (const BoxSpec& b : r.boxes) validate_box(b);
    validate_materials(r.materials);
}

// ------------------------------------------------------------ the fixture

struct Spec {
    int fixture = 0;  // 0 slump, 1 hydrostatic, 2 dam_break, 3 coupled_patch
    mpm_material material{};
    double dx = 0.0;
    double domain[3] = {0, 0, 0}, initial_size[3] = {0, 0, 0}, initial_velocity[3] = {0, 0, 0};
    uint64_t seed = 0;
    bool has_pellet = false;
    double pel_radius = 0.0, pel_length = 0.0, pel_density = 0.0;
    mpm_config config{};
    size_t max_particles = kMaxParticles, max_nodes = kMaxNodes;

    size_t axis_count(int a) const {
        double sub = dx / kParticlesPerAxis;
        return (size_t)std::max(std::round(initial_size[a] / sub), 1.0);
    }
    // FixtureSpec::particle_count, fixtures.rs:120-126 (saturating product).
    size_t particle_count() const {
        size_t n = 1;
        for (int a = 0; a < 3; ++a) {
            size_t c = axis_count(a);
            n = (c != 0 && n > SIZE_MAX / c) ? SIZE_MAX : n * c;
        }
        return n;
    }
    double gravity_norm() const {
        const double* g = config.gravity;
        return std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    }
};

// GridLayout::new, grid.rs:37-86.
void grid_layout(const double* domain, double spacing, size_t max_nodes, size_t* cells) {
    if (!(std::isfinite(spacing) && spacing > 0.0))
        bail("`grid_spacing_m` must be positive and finite, got '" + show(spacing) + "'");
    for (int a = 0; a < 3; ++a) {
        double extent = domain[a];
        if (!(std::isfinite(extent) && extent > 0.0))
            bail("`domain_m[" + std::to_string(a) + "]` must be positive and finite, got '" + show(extent) + "'");
        double ratio = extent / spacing, rounded = std::round(ratio);
        double cap = max_nodes >= 3 ? std::min((double)(max_nodes - 3), 4294967295.0) : 0.0;
        if (!std::isfinite(ratio) || rounded > cap) bail("grid axis exceeds the node allocation cap");
        if (std::fabs(ratio - rounded) > 1e-6 * std::max(rounded, 1.0))
            bail("`domain_m[" + std::to_string(a) + "]`='" + show(extent) + "' is not an integer multiple of `grid_spacing_m`='" +
                 show(spacing) + "'");
        if (rounded < 2.0) bail("`domain_m[" + std::to_string(a) + "]` must span at least two cells");
        cells[a] = (size_t)rounded;
    }
    size_t total = (cells[0] + 3) * (cells[1] + 3) * (cells[2] + 3);
    if (total > max_nodes)
        bail("grid would allocate '" + std::to_string(total) + "' nodes, exceeding the cap of '" + std::to_string(max_nodes) +
             "'; coarsen `grid_spacing_m` or shrink `domain_m`");
}

const char* fixture_name(int f) {
    static const char* names[4] = {"slump", "hydrostatic", "dam_break", "coupled_patch"};
    return names[f];
}

// FixtureSpec::validate, fixtures.rs:152-236.
void validate_spec(const Spec& s) {
    for (int a = 0; a < 3; ++a) {
        if (s.initial_size[a] > s.domain[a] + 1e-12)
            bail("`initial_size_m[" + std::to_string(a) + "]`='" + show(s.initial_size[a]) + "' exceeds `domain_m[" +
                 std::to_string(a) + "]`='" + show(s.domain[a]) + "'");
        if (s.initial_size[a] < s.dx)
            bail("`initial_size_m[" + std::to_string(a) + "]`='" + show(s.initial_size[a]) +
                 "' must span at least one cell of '" + show(s.dx) + "'");
    }
    size_t count = s.particle_count();
    if (count > s.max_particles)
        bail("fixture would allocate '" + std::to_string(count) + "' particles, exceeding the cap of '" +
             std::to_string(s.max_particles) + "'");
    if (s.fixture == 1) {
        for (int a = 0; a < 2; ++a)
            if (std::fabs(s.initial_size[a] - s.domain[a]) > 1e-9)
                bail("`hydrostatic` requires `initial_size_m[" + std::to_string(a) + "]` to equal `domain_m[" +
                     std::to_string(a) + "]` so the column spans the domain");
        const double* v = s.initial_velocity;
        if (std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) > 0.0) bail("`hydrostatic` requires zero `initial_velocity_m_s`");
        if (s.initial_size[2] >= s.domain[2]) bail("`hydrostatic` requires free space above the column");
    } else if (s.fixture == 3) {
        if (!s.has_pellet) bail("`coupled_patch` requires exactly one box supplying pellet geometry");
        if (s.material.kind != 0) bail("`coupled_patch` is implemented for `paste` only");
        if (s.pel_radius < kMinCellsPerRadius * s.dx)
            bail("`pellet_radius_m`='" + show(s.pel_radius) + "' is under-resolved: coupling needs at least 1.5 cells per radius at `grid_spacing_m`='" +
                 show(s.dx) + "'");
        double top = s.initial_size[2] + kPelletGapCells * s.dx + 2.0 * s.pel_radius;
        if (top > s.domain[2] || s.pel_length > s.domain[0] || 2.0 * s.pel_radius > s.domain[1])
            bail("`coupled_patch` pellet does not fit above the block inside `domain_m`");
    } else if (s.has_pellet) {
        bail(std::string("`") + fixture_name(s.fixture) + "` does not accept boxes; only `coupled_patch` uses pellet geometry");
    }
}

// Material from the research parameters in shear units, research.rs:272-305.
mpm_material research_material(const std::string& name, double density, double young, double poisson, double yield_stress,
                               double consistency, double flow_index) {
    mpm_material m{};
    if (name == "paste") {
        mpm_material_paste_shear(density, young, poisson, yield_stress, consistency, flow_index, &m);
    } else if (name == "water") {
        if (!(yield_stress == 0.0 && std::fabs(flow_index - 1.0) < 1e-12)) bail("water requires zero yield stress and flow_index=1");
        mpm_material_liquid(density, young, poisson, consistency, &m);
    } else if (name == "fines") {
        bail("porous Drucker-Prager fines are not implemented");
    } else {
        bail("unsupported research material '" + name + "'");
    }
    double modulus = m.kind == 0 ? m.lame_lambda + 2.0 * m.shear_modulus : m.bulk_modulus;
    double wave = std::sqrt(modulus / m.density);
    double diff = m.kind == 0 ? 0.0 : m.consistency / m.density;
    if (!std::isfinite(wave) || wave <= 0.0 || !std::isfinite(m.consistency) || !std::isfinite(m.yield_stress) ||
        !std::isfinite(diff))
        bail("derived material properties are nonfinite");
    return m;
}

// specification, research.rs:170-232, with validate_parameters (235-270).
Spec specification(const Request& r) {
    validate_envelope(r);
    if (!r.has_research) bail("research parameters are required");
    const Value& p = r.research;
    expect_object(p, "Parameters");
    deny_unknown(p, {"fixture", "grid_spacing_m", "domain_m", "material", "density_kg_m3", "young_modulus_pa",
                     "poisson_ratio", "yield_stress_pa", "consistency_pa_s_n", "flow_index", "initial_size_m",
                     "initial_velocity_m_s"});
    std::string fixture = as_str(field(p, "fixture"), "fixture");
    double dx = as_f64(field(p, "grid_spacing_m"), "grid_spacing_m");
    double domain[3], size[3], vel[3];
    as_fixed(field(p, "domain_m"), "domain_m", domain, 3);
    std::string material = as_str(field(p, "material"), "material");
    double density = as_f64(field(p, "density_kg_m3"), "density_kg_m3");
    double young = as_f64(field(p, "young_modulus_pa"), "young_modulus_pa");
    double poisson = as_f64(field(p, "poisson_ratio"), "poisson_ratio");
    double yield_stress = as_f64(field(p, "yield_stress_pa"), "yield_stress_pa");
    double consistency = as_f64(field(p, "consistency_pa_s_n"), "consistency_pa_s_n");
    double flow_index = as_f64(field(p, "flow_index"), "flow_index");
    as_fixed(field(p, "initial_size_m"), "initial_size_m", size, 3);
    as_fixed(field(p, "initial_velocity_m_s"), "initial_velocity_m_s", vel, 3);

    // validate_parameters
    bool bad = false;
    for (double v : {dx, density, young, flow_index}) bad = bad || !std::isfinite(v) || v <= 0.0;
    for (int a = 0; a < 3; ++a) bad = bad || !std::isfinite(domain[a]) || domain[a] <= 0.0;
    for (int a = 0; a < 3; ++a) bad = bad || !std::isfinite(size[a]) || size[a] <= 0.0;
    for (int a = 0; a < 3; ++a) bad = bad || !std::isfinite(vel[a]);
    if (bad) bail("research geometry, density, stiffness and flow index must be finite and positive");
    if (!std::isfinite(poisson) || !(poisson >= 0.0 && poisson < 0.5) || !std::isfinite(yield_stress) || yield_stress < 0.0 ||
        !std::isfinite(consistency) || consistency < 0.0)
        bail("invalid Poisson ratio, yield stress or consistency");
    for (int a = 0; a < 3; ++a) {
        double ratio = size[a] / (0.5 * dx);
        if (size[a] > domain[a] || !std::isfinite(ratio) || std::fabs(ratio - std::round(ratio)) > 1e-6)
            bail("initial block must fit domain and span whole particle subcells");
    }

    Spec s;
    if (fixture == "slump")
        s.fixture = 0;
    else if (fixture == "hydrostatic")
        s.fixture = 1;
    else if (fixture == "dam_break")
        s.fixture = 2;
    else if (fixture == "coupled_patch")
        s.fixture = 3;
    else
        bail("unsupported research fixture '" + fixture + "'");
    if (s.fixture != 3 && !r.boxes.empty()) bail("only coupled_patch accepts a pellet geometry box");
    s.material = research_material(material, density, young, poisson, yield_stress, consistency, flow_index);
    s.dx = dx;
    for (int a = 0; a < 3; ++a) {
        s.domain[a] = domain[a];
        s.initial_size[a] = size[a];
        s.initial_velocity[a] = vel[a];
    }
    s.seed = r.seed;
    if (!r.boxes.empty()) {
        s.has_pellet = true;
        s.pel_radius = r.boxes[0].pellet_radius_m;
        s.pel_length = r.boxes[0].pellet_length_m;
        s.pel_density = r.boxes[0].pellet_density_kg_m3;
    }
    mpm_config& c = s.config;
    c.gravity[0] = 0.0;
    c.gravity[1] = 0.0;
    c.gravity[2] = -9.81;
    c.wall_friction = r.materials.friction;
    c.cfl = 0.3;
    c.max_dt = r.dt_s;
    c.min_dt = 1e-12;
    c.contact_stiffness = r.materials.normal_stiffness_n_m;
    c.contact_restitution = r.materials.restitution;
    c.contact_friction = r.materials.friction;
    for (int f = 0; f < 6; ++f) c.face_friction[f] = r.materials.friction;
    size_t cells[3];
    grid_layout(s.domain, s.dx, s.max_nodes, cells);
    validate_spec(s);
    double frames = std::ceil(r.duration_s / r.record_interval_s) + 1.0;
    size_t count = s.particle_count();
    if (count > 0xFFFFFFFFull) bail("particle count exceeds u32");
    uint32_t c32 = (uint32_t)count;
    uint32_t sat = c32 > 0xFFFFFFFFu - 64 ? 0xFFFFFFFFu : c32 + 64;
    if (frames * (double)sat > kMaxFramePoints) bail("research frame-point resource cap exceeded; reduce recording frequency");
    return s;
}

// Pellet::new, rigid.rs:115-186: a cylinder at rest as a line of spheres.
mpm_pellet make_pellet(double radius, double length, double density, const double* position) {
    if (!(std::isfinite(radius) && radius > 0.0)) bail("pellet radius must be positive, got '" + show(radius) + "'");
    if (!(std::isfinite(length) && length >= 2.0 * radius))
        bail("pellet length must be at least twice the radius, got '" + show(length) + "' for radius '" + show(radius) + "'");
    if (!(std::isfinite(density) && density > 0.0)) bail("pellet density must be positive, got '" + show(density) + "'");
    mpm_pellet p{};
    p.radius = radius;
    p.length = length;
    p.mass = density * M_PI * radius * radius * length;
    p.inertia_body[0] = 0.5 * p.mass * radius * radius;
    p.inertia_body[1] = p.mass * (3.0 * radius * radius + length * length) / 12.0;
    p.inertia_body[2] = p.mass * (3.0 * radius * radius + length * length) / 12.0;
    bool ok = std::isfinite(p.mass) && p.mass > 0.0;
    for (int a = 0; a < 3; ++a) ok = ok && std::isfinite(p.inertia_body[a]) && p.inertia_body[a] > 0.0 && std::isfinite(position[a]);
    if (!ok) bail("pellet mass, inertia or pose is not representable");
    for (int a = 0; a < 3; ++a) p.position[a] = position[a];
    p.orientation[3] = 1.0;  // identity (i, j, k, w)
    double half_span = 0.5 * length - radius;
    int count = 1;
    if (half_span > 0.0) {
        double ratio = std::ceil(2.0 * half_span / radius);
        if (!std::isfinite(ratio) || ratio > 4095.0) bail("pellet proxy exceeds the sphere allocation cap");
        count = (int)ratio + 1;
    }
    if (count > 64) bail("pellet proxy exceeds the engine's 64-sphere cap");
    p.sphere_count = count;
    for (int i = 0; i < count; ++i)
        p.sphere_offsets[i] = count == 1 ? 0.0 : -half_span + 2.0 * half_span * (double)i / (double)(count - 1);
    return p;
}

// ------------------------------------------------------------ state on the host

struct State {
    std::vector<uint32_t> id;
    std::vector<double> pos, vel, aff, def, ratio, tau, eps, mass, vol0;
    std::vector<mpm_pellet> pellets;
    mpm_ledger ledger{};
    mpm_history history{};

    size_t n() const { return id.size(); }
    void resize(size_t n) {
        id.resize(n);
        pos.resize(3 * n);
        vel.resize(3 * n);
        aff.resize(9 * n);
        def.resize(9 * n);
        ratio.resize(n);
        tau.resize(9 * n);
        eps.resize(n);
        mass.resize(n);
        vol0.resize(n);
    }
    mpm_particles view() {
        mpm_particles p{};
        p.n = n();
        p.id = id.data();
        p.position = pos.data();
        p.velocity = vel.data();
        p.affine = aff.data();
        p.deformation = def.data();
        p.volume_ratio = ratio.data();
        p.kirchhoff = tau.data();
        p.plastic_strain = eps.data();
        p.mass = mass.data();
        p.volume0 = vol0.data();
        return p;
    }
    void load(const mpm_particles* ps, const mpm_pellet* pel, size_t npel, const mpm_ledger* led, const mpm_history* hist) {
        resize(ps ? ps->n : 0);
        size_t n = this->n();
        for (size_t p = 0; p < n; ++p) {
            id[p] = ps->id ? ps->id[p] : (uint32_t)p;
            ratio[p] = ps->volume_ratio ? ps->volume_ratio[p] : 1.0;
            eps[p] = ps->plastic_strain ? ps->plastic_strain[p] : 0.0;
            mass[p] = ps->mass[p];
            vol0[p] = ps->volume0[p];
            for (int a = 0; a < 3; ++a) {
                pos[3 * p + a] = ps->position[3 * p + a];
                vel[3 * p + a] = ps->velocity ? ps->velocity[3 * p + a] : 0.0;
            }
            for (int a = 0; a < 9; ++a) {
                aff[9 * p + a] = ps->affine ? ps->affine[9 * p + a] : 0.0;
                def[9 * p + a] = ps->deformation ? ps->deformation[9 * p + a] : (a % 4 == 0 ? 1.0 : 0.0);
                tau[9 * p + a] = ps->kirchhoff ? ps->kirchhoff[9 * p + a] : 0.0;
            }
        }
        pellets.assign(pel, pel + npel);
        ledger = led ? *led : mpm_ledger{};
        if (hist) {
            history = *hist;
        } else {
            history = mpm_history{};
            history.dt_scale = 1.0;
            history.min_dt_used = std::numeric_limits<double>::infinity();
        }
    }
};

// The engine behind one research run (Simulation plus its fixture).
struct Sim {
    mpm_engine* eng = nullptr;
    State st;
    size_t cells[3] = {0, 0, 0};
    double dx = 0.0;
    ~Sim() {
        if (eng) mpm_destroy(eng);
    }
    [[noreturn]] void engine_fail(int rc) const {
        std::string msg = eng ? mpm_last_error(eng) : "engine unavailable";
        bail(msg, rc);
    }
    void pull() {
        size_t n = mpm_particle_count(eng);
        st.resize(n);
        mpm_particles view = st.view();
        st.pellets.resize(mpm_pellet_count(eng));
        int rc = mpm_get_state(eng, &view, st.pellets.data(), st.pellets.size(), &st.ledger, &st.history);
        if (rc) engine_fail(rc);
    }
    void push() {
        mpm_particles view = st.view();
        int rc = mpm_set_state(eng, &view, st.pellets.data(), st.pellets.size(), &st.ledger, &st.history);
        if (rc) engine_fail(rc);
    }
};

// FixtureSpec::build, fixtures.rs:238-270: placement, the pellet, and
// Simulation::new's ledger seed (done by the engine from the f64 state).
void build(const Spec& s, Sim& sim) {
    size_t n = s.particle_count();
    mpm_fixture f{};
    f.fixture = s.fixture;
    f.material = s.material;
    f.grid_spacing = s.dx;
    for (int a = 0; a < 3; ++a) {
        f.domain[a] = s.domain[a];
        f.initial_size[a] = s.initial_size[a];
        f.initial_velocity[a] = s.initial_velocity[a];
    }
    f.seed = s.seed;
    f.gravity_norm = s.gravity_norm();
    f.jitter = kJitterFraction;
    State& st = sim.st;
    st.resize(n);
    mpm_particles view = st.view();
    if (mpm_fixture_place(&f, &view) != MPM_OK || view.n != n) bail("fixture placement failed");
    st.pellets.clear();
    if (s.fixture == 3 && s.has_pellet) {
        double center[3] = {0.5 * s.domain[0], 0.5 * s.domain[1], s.initial_size[2] + kPelletGapCells * s.dx + s.pel_radius};
        st.pellets.push_back(make_pellet(s.pel_radius, s.pel_length, s.pel_density, center));
    }
    mpm_layout layout{};
    for (int a = 0; a < 3; ++a) layout.domain[a] = s.domain[a];
    layout.spacing = s.dx;
    layout.max_nodes = s.max_nodes;
    int rc = mpm_create(&s.material, &s.config, &layout, n, st.pellets.size(), &sim.eng);
    if (rc) bail(std::string("engine creation failed: ") + (sim.eng ? mpm_last_error(sim.eng) : "no device"), rc);
    rc = mpm_set_state(sim.eng, &view, st.pellets.data(), st.pellets.size(), nullptr, nullptr);
    if (rc) sim.engine_fail(rc);
    grid_layout(s.domain, s.dx, s.max_nodes, sim.cells);
    sim.dx = s.dx;
}

// ------------------------------------------------------------ checks

// GridLayout::stencil's validity test, grid.rs:151-161.
bool inside_grid(const double* x, double dx, const size_t* cells) {
    for (int a = 0; a < 3; ++a) {
        double padded = std::floor(x[a] / dx - 0.5) + 1.0;
        if (std::isnan(padded) || padded < 0.0 || padded + 2.0 >= (double)(cells[a] + 3)) return false;
    }
    return true;
}

// ParticleSet::validate, particles.rs:159-211.
void validate_particles(const State& st) {
    size_t n = st.n();
    std::vector<uint32_t> seen(st.id);
    std::sort(seen.begin(), seen.end());
    for (size_t k = 1; k < seen.size(); ++k)
        if (seen[k] == seen[k - 1]) bail("particle ids are not unique");
    for (size_t p = 0; p < n; ++p) {
        bool ok = std::isfinite(st.ratio[p]) && std::isfinite(st.eps[p]) && std::isfinite(st.mass[p]) && std::isfinite(st.vol0[p]);
        for (int a = 0; a < 3; ++a) ok = ok && std::isfinite(st.pos[3 * p + a]) && std::isfinite(st.vel[3 * p + a]);
        for (int a = 0; a < 9; ++a)
            ok = ok && std::isfinite(st.aff[9 * p + a]) && std::isfinite(st.def[9 * p + a]) && std::isfinite(st.tau[9 * p + a]);
        if (!ok) bail("particle id '" + std::to_string(st.id[p]) + "' has non-finite state");
        if (st.mass[p] <= 0.0 || st.vol0[p] <= 0.0)
            bail("particle id '" + std::to_string(st.id[p]) + "' has non-positive mass or volume");
        if (st.ratio[p] <= 0.0 || st.eps[p] < 0.0)
            bail("particle id '" + std::to_string(st.id[p]) + "' has invalid ratio/plastic state");
    }
}

// Ledger::validate, solver.rs:182-243.
void validate_ledger(const mpm_ledger& L) {
    const std::pair<const char*, double> scalars[] = {
        {"initial_mass", L.initial_mass},
        {"outflow_mass", L.outflow_mass},
        {"initial_mechanical_energy", L.initial_mechanical_energy},
        {"wall_normal_projection_energy", L.wall_normal_projection_energy},
        {"wall_friction_dissipation", L.wall_friction_dissipation},
        {"wall_normal_traction_energy", L.wall_normal_traction_energy},
        {"pellet_wall_work", L.pellet_wall_work},
        {"coupling_grid_energy", L.coupling_grid_energy},
        {"coupling_pellet_energy", L.coupling_pellet_energy},
        {"plastic_dissipation", L.plastic_dissipation}};
    for (const auto& kv : scalars)
        if (!std::isfinite(kv.second)) bail(std::string("invalid solver state: ledger `") + kv.first + "` is nonfinite: '" + show(kv.second) + "'");
    const std::pair<const char*, const double*> vectors[] = {
        {"initial_momentum", L.initial_momentum},     {"gravity_impulse", L.gravity_impulse},
        {"wall_impulse", L.wall_impulse},             {"pellet_wall_impulse", L.pellet_wall_impulse},
        {"coupling_impulse", L.coupling_impulse},     {"coupling_angular_impulse", L.coupling_angular_impulse},
        {"outflow_momentum", L.outflow_momentum}};
    for (const auto& kv : vectors)
        for (int a = 0; a < 3; ++a)
            if (!std::isfinite(kv.second[a])) bail(std::string("invalid solver state: ledger `") + kv.first + "` is nonfinite");
    const std::pair<const char*, double> nonneg[] = {{"initial_mass", L.initial_mass},
                                                      {"outflow_mass", L.outflow_mass},
                                                      {"plastic_dissipation", L.plastic_dissipation},
                                                      {"wall_friction_dissipation", L.wall_friction_dissipation}};
    for (const auto& kv : nonneg)
        if (kv.second < 0.0)
            bail(std::string("invalid solver state: ledger `") + kv.first + "` must be non-negative, got '" + show(kv.second) + "'");
    if (L.wall_normal_projection_energy > 0.0)
        bail("invalid solver state: ledger `wall_normal_projection_energy` must be non-positive, got '" +
             show(L.wall_normal_projection_energy) + "'");
}

bool pellet_finite(const mpm_pellet& p) {
    for (int a = 0; a < 3; ++a)
        if (!std::isfinite(p.position[a]) || !std::isfinite(p.velocity[a]) || !std::isfinite(p.angular_momentum[a])) return false;
    for (int a = 0; a < 4; ++a)
        if (!std::isfinite(p.orientation[a])) return false;
    return true;
}

double total_mass(const State& st) {
    double m = 0.0;
    for (double v : st.mass) m += v;
    return m;
}
double mass_residual(const State& st) { return total_mass(st) + st.ledger.outflow_mass - st.ledger.initial_mass; }

// Simulation::validate (solver.rs:1106-1140) and check_closed_state (research.rs:442-457).
void check_closed_state(const State& st, double dx, const size_t* cells) {
    try {
        validate_particles(st);
    } catch (const ResearchError& e) {
        bail(std::string("invalid solver state: ") + e.what());
    }
    validate_ledger(st.ledger);
    const mpm_history& h = st.history;
    if (!std::isfinite(h.time) || h.time < 0.0) bail("invalid solver state: time '" + show(h.time) + "' invalid");
    if (!(std::isfinite(h.dt_scale) && h.dt_scale > 0.0 && h.dt_scale <= 1.0))
        bail("invalid solver state: dt_scale '" + show(h.dt_scale) + "' invalid");
    for (const mpm_pellet& p : st.pellets)
        if (!pellet_finite(p)) bail("invalid solver state: pellet state non-finite");
    for (size_t p = 0; p < st.n(); ++p)
        if (!inside_grid(&st.pos[3 * p], dx, cells))
            bail("invalid solver state: particle id '" + std::to_string(st.id[p]) + "' lies outside the grid");
    const double out = st.ledger.outflow_mass;
    if (!std::isfinite(out) || std::fabs(out) > 0.0)
        bail("closed research fixture leaked '" + show(out) + "' kg beyond the padded grid");
    double r = mass_residual(st);
    if (!std::isfinite(r) || std::fabs(r) > std::max(1e-6 * st.ledger.initial_mass, 1e-12))
        bail("research mass conservation failed: '" + show(r) + "' kg");
}

// ------------------------------------------------------------ observables

// Energies of the state, solver.rs:1027-1099 (fixed-order sums in particle order).
struct Energies {
    double kinetic, potential, elastic;
};
Energies energies(const Spec& s, const State& st) {
    const double* g = s.config.gravity;
    size_t n = st.n();
    double ke = 0.0, pe = 0.0, el = 0.0;
    for (size_t p = 0; p < n; ++p) {
        const double* v = &st.vel[3 * p];
        ke += 0.5 * st.mass[p] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    }
    for (const mpm_pellet& pl : st.pellets) ke += mpm_host::pellet_kinetic(pl);
    for (size_t p = 0; p < n; ++p) {
        const double* x = &st.pos[3 * p];
        pe += -st.mass[p] * (g[0] * x[0] + g[1] * x[1] + g[2] * x[2]);
    }
    for (const mpm_pellet& pl : st.pellets) pe -= pl.mass * (g[0] * pl.position[0] + g[1] * pl.position[1] + g[2] * pl.position[2]);
    const mpm_material& m = s.material;
    for (size_t p = 0; p < n; ++p) {
        const double v0 = st.vol0[p];
        if (m.kind == 0) {
            const double* F = &st.def[9 * p];
            double b[9], vals[3];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) b[3 * i + j] = F[3 * i] * F[3 * j] + F[3 * i + 1] * F[3 * j + 1] + F[3 * i + 2] * F[3 * j + 2];
            if (!mpm_host::eigen3(b, vals)) {
                el += std::numeric_limits<double>::quiet_NaN();
                continue;
            }
            double hh[3];
            for (int a = 0; a < 3; ++a) hh[a] = 0.5 * std::log(std::max(vals[a], 1e-300));
            double tr = hh[0] + hh[1] + hh[2], h2 = hh[0] * hh[0] + hh[1] * hh[1] + hh[2] * hh[2];
            el += v0 * (m.shear_modulus * h2 + 0.5 * m.lame_lambda * tr * tr);
        } else {
            double j = st.ratio[p];
            if (m.tait_gamma == 1.0)
                el += v0 * m.bulk_modulus * (j - 1.0 - std::log(j));
            else
                el += v0 * (m.bulk_modulus / m.tait_gamma) *
                      (j - 1.0 + (std::pow(j, 1.0 - m.tait_gamma) - 1.0) / (m.tait_gamma - 1.0));
        }
    }
    return {ke, pe, el};
}

// Ledger::ledgered_energy, solver.rs:168-175.
double ledgered_energy(const mpm_ledger& L) {
    return L.wall_normal_projection_energy - L.wall_friction_dissipation + L.wall_normal_traction_energy + L.pellet_wall_work +
           L.coupling_grid_energy + L.coupling_pellet_energy - L.plastic_dissipation;
}

// liquid_pressure, constitutive.rs:593-595 (Tait exponent 1 is the reference law).
double liquid_pressure(const mpm_material& m, double j) {
    if (m.tait_gamma == 1.0) return m.bulk_modulus * (1.0 / j - 1.0);
    return m.bulk_modulus / m.tait_gamma * (std::pow(j, -m.tait_gamma) - 1.0);
}

// FixtureSpec::hydrostatic_reference, fixtures.rs:354-366.
double hydrostatic_reference(const Spec& s, double depth) {
    const mpm_material& m = s.material;
    double g = s.gravity_norm();
    if (m.kind == 1) return m.bulk_modulus * (std::exp(m.density * g * depth / m.bulk_modulus) - 1.0);
    return m.density * g * depth;
}

// FixtureSpec::observables, fixtures.rs:369-615.
std::map<std::string, double> observables(const Spec& s, const State& st) {
    std::map<std::string, double> out;
    const mpm_ledger& L = st.ledger;
    const mpm_history& H = st.history;
    size_t n = st.n();
    const double sub = s.dx / kParticlesPerAxis;
    out["mass_residual_kg"] = mass_residual(st);
    double mom[3] = {0.0, 0.0, 0.0};
    for (size_t p = 0; p < n; ++p)
        for (int a = 0; a < 3; ++a) mom[a] += st.vel[3 * p + a] * st.mass[p];
    for (const mpm_pellet& pl : st.pellets)
        for (int a = 0; a < 3; ++a) mom[a] += pl.velocity[a] * pl.mass;
    double res[3];
    for (int a = 0; a < 3; ++a)
        res[a] = mom[a] - L.initial_momentum[a] - L.gravity_impulse[a] + L.outflow_momentum[a] - L.wall_impulse[a] -
                 L.pellet_wall_impulse[a];
    out["momentum_residual_kg_m_s"] = std::sqrt(res[0] * res[0] + res[1] * res[1] + res[2] * res[2]);
    Energies e = energies(s, st);
    double mech = e.kinetic + e.potential + e.elastic;
    out["kinetic_energy_j"] = e.kinetic;
    out["potential_energy_j"] = e.potential;
    out["elastic_energy_j"] = e.elastic;
    out["mechanical_energy_j"] = mech;
    out["plastic_dissipation_j"] = L.plastic_dissipation;
    out["wall_normal_projection_energy_j"] = L.wall_normal_projection_energy;
    out["wall_friction_dissipation_j"] = L.wall_friction_dissipation;
    out["wall_normal_traction_energy_j"] = L.wall_normal_traction_energy;
    out["energy_residual_j"] = mech - L.initial_mechanical_energy - ledgered_energy(L);
    out["outflow_kg"] = L.outflow_mass;
    out["particle_count"] = (double)n;
    out["step_count"] = (double)H.step_count;
    out["limited_steps"] = (double)H.limited_steps;
    out["rejected_steps"] = (double)H.rejected_steps;
    out["min_dt_s"] = std::isfinite(H.min_dt_used) ? H.min_dt_used : 0.0;
    out["max_dt_s"] = H.max_dt_used;
    double max_speed = 0.0, max_plastic = 0.0;
    size_t yielded = 0;
    for (size_t p = 0; p < n; ++p) {
        const double* v = &st.vel[3 * p];
        max_speed = std::max(max_speed, std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
        max_plastic = std::max(max_plastic, st.eps[p]);
        yielded += st.eps[p] > 0.0;
    }
    out["max_speed_m_s"] = max_speed;
    out["max_plastic_strain"] = max_plastic;
    out["yielded_fraction"] = n ? (double)yielded / (double)n : 0.0;
    // shape_observables, fixtures.rs:464-508
    double min_x = INFINITY, max_x = -INFINITY, min_y = INFINITY, max_y = -INFINITY, max_z = -INFINITY;
    for (size_t p = 0; p < n; ++p) {
        const double* x = &st.pos[3 * p];
        min_x = std::min(min_x, x[0]);
        max_x = std::max(max_x, x[0]);
        min_y = std::min(min_y, x[1]);
        max_y = std::max(max_y, x[1]);
        max_z = std::max(max_z, x[2]);
    }
    double height = n ? max_z + 0.5 * sub : 0.0;
    out["height_m"] = height;
    out["spread_x_m"] = n ? max_x - min_x + sub : 0.0;
    out["spread_y_m"] = n ? max_y - min_y + sub : 0.0;
    const double g = s.gravity_norm();
    if (s.fixture == 0) {
        // slump_observables, fixtures.rs:510-527 (Pashias et al. 1996).
        double h0 = s.initial_size[2];
        out["slump_m"] = h0 - height;
        double shear_yield = s.material.yield_stress / std::sqrt(3.0);
        double tp = shear_yield / (s.material.density * g * h0);
        out["dimensionless_yield_stress"] = tp;
        out["pashias_slump_m"] = tp >= 0.5 ? 0.0 : (tp <= 0.0 ? h0 : h0 * (1.0 - 2.0 * tp * (1.0 - std::log(2.0 * tp))));
    } else if (s.fixture == 1) {
        // hydrostatic_observables, fixtures.rs:529-561.
        double h0 = s.initial_size[2];
        double scale = s.material.density * g * h0;
        double sum_sq = 0.0, max_err = 0.0;
        for (size_t p = 0; p < n; ++p) {
            double depth = std::max(h0 - st.pos[3 * p + 2], 0.0);
            double ref = hydrostatic_reference(s, depth);
            double measured = s.material.kind == 1 ? liquid_pressure(s.material, st.ratio[p]) : -st.tau[9 * p + 8];
            double err = (measured - ref) / scale;
            sum_sq += err * err;
            max_err = std::max(max_err, std::fabs(err));
        }
        out["hydrostatic_rms_relative_error"] = n ? std::sqrt(sum_sq / (double)n) : 0.0;
        out["hydrostatic_max_relative_error"] = max_err;
        out["reference_bottom_stress_pa"] = hydrostatic_reference(s, h0);
        double modulus = s.material.kind == 0 ? s.material.lame_lambda + 2.0 * s.material.shear_modulus : s.material.bulk_modulus;
        out["wave_speed_m_s"] = std::sqrt(modulus / s.material.density);
    } else if (s.fixture == 2) {
        out["front_x_m"] = n ? max_x + 0.5 * sub : 0.0;
        double ritter = s.initial_size[0] + 2.0 * std::sqrt(g * s.initial_size[2]) * H.time;
        out["ritter_front_bound_m"] = std::min(ritter, s.domain[0]);
        double column = 0.0;
        for (size_t p = 0; p < n; ++p)
            if (st.pos[3 * p] < s.dx) column = std::max(column, st.pos[3 * p + 2] + 0.5 * sub);
        out["column_height_m"] = column;
    } else if (s.fixture == 3 && !st.pellets.empty()) {
        // coupled_observables, fixtures.rs:563-615.
        const mpm_pellet& pl = st.pellets[0];
        out["pellet_x_m"] = pl.position[0];
        out["pellet_y_m"] = pl.position[1];
        out["pellet_z_m"] = pl.position[2];
        const double* v = pl.velocity;
        out["pellet_speed_m_s"] = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        double om[3];
        mpm_host::pellet_omega(pl, om);
        out["pellet_angular_speed_rad_s"] = std::sqrt(om[0] * om[0] + om[1] * om[1] + om[2] * om[2]);
        out["pellet_kinetic_energy_j"] = mpm_host::pellet_kinetic(pl);
        out["pellet_penetration_m"] = std::max(s.initial_size[2] + pl.radius - pl.position[2], 0.0);
        double r[9];
        mpm_host::qmatrix(pl.orientation, r);
        out["pellet_axis_tilt_rad"] = std::asin(std::min(std::fabs(r[6]), 1.0));
        out["coupling_impulse_x_kg_m_s"] = L.coupling_impulse[0];
        out["coupling_impulse_y_kg_m_s"] = L.coupling_impulse[1];
        out["coupling_impulse_z_kg_m_s"] = L.coupling_impulse[2];
        out["coupling_angular_impulse_x_kg_m2_s"] = L.coupling_angular_impulse[0];
        out["coupling_angular_impulse_y_kg_m2_s"] = L.coupling_angular_impulse[1];
        out["coupling_angular_impulse_z_kg_m2_s"] = L.coupling_angular_impulse[2];
        out["pellet_wall_impulse_z_kg_m_s"] = L.pellet_wall_impulse[2];
        out["pellet_wall_work_j"] = L.pellet_wall_work;
        out["coupling_grid_energy_j"] = L.coupling_grid_energy;
        out["coupling_pellet_energy_j"] = L.coupling_pellet_energy;
    }
    return out;
}

// ------------------------------------------------------------ checkpoint codec
// Snapshot and Checkpoint, research.rs:42-65, serialised as serde_json does:
// Vector3 as [x, y, z], Matrix3 as its nine entries column by column (the
// storage order of nalgebra's ArrayStorage), UnitQuaternion as [i, j, k, w],
// Option as null or the value.

Value mat_json(const double* rm) {
    Value v = Value::array();
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) v.a.push_back(Value::number(rm[3 * r + c]));
    return v;
}

Value particles_json(const State& st) {
    size_t n = st.n();
    Value o = Value::object();
    Value id = Value::array(), pos = Value::array(), vel = Value::array(), aff = Value::array(), def = Value::array(),
          ratio = Value::array(), tau = Value::array(), eps = Value::array(), mass = Value::array(), vol = Value::array();
    for (size_t p = 0; p < n; ++p) {
        id.a.push_back(Value::uint(st.id[p]));
        pos.a.push_back(Value::vec(&st.pos[3 * p], 3));
        vel.a.push_back(Value::vec(&st.vel[3 * p], 3));
        aff.a.push_back(mat_json(&st.aff[9 * p]));
        def.a.push_back(mat_json(&st.def[9 * p]));
        ratio.a.push_back(Value::number(st.ratio[p]));
        tau.a.push_back(mat_json(&st.tau[9 * p]));
        eps.a.push_back(Value::number(st.eps[p]));
        mass.a.push_back(Value::number(st.mass[p]));
        vol.a.push_back(Value::number(st.vol0[p]));
    }
    o["id"] = std::move(id);
    o["position"] = std::move(pos);
    o["velocity"] = std::move(vel);
    o["affine"] = std::move(aff);
    o["deformation"] = std::move(def);
    o["volume_ratio"] = std::move(ratio);
    o["kirchhoff"] = std::move(tau);
    o["plastic_strain"] = std::move(eps);
    o["mass"] = std::move(mass);
    o["volume0"] = std::move(vol);
    return o;
}

Value pellet_json(const mpm_pellet& p) {
    Value o = Value::object();
    o["radius"] = Value::number(p.radius);
    o["length"] = Value::number(p.length);
    o["mass"] = Value::number(p.mass);
    o["inertia_body"] = Value::vec(p.inertia_body, 3);
    o["position"] = Value::vec(p.position, 3);
    o["orientation"] = Value::vec(p.orientation, 4);
    o["velocity"] = Value::vec(p.velocity, 3);
    o["angular_momentum"] = Value::vec(p.angular_momentum, 3);
    o["sphere_offsets"] = Value::vec(p.sphere_offsets, (size_t)p.sphere_count);
    return o;
}

Value ledger_json(const mpm_ledger& L) {
    Value o = Value::object();
    o["initial_mass"] = Value::number(L.initial_mass);
    o["outflow_mass"] = Value::number(L.outflow_mass);
    o["initial_momentum"] = Value::vec(L.initial_momentum, 3);
    o["initial_mechanical_energy"] = Value::number(L.initial_mechanical_energy);
    o["gravity_impulse"] = Value::vec(L.gravity_impulse, 3);
    o["wall_impulse"] = Value::vec(L.wall_impulse, 3);
    o["pellet_wall_impulse"] = Value::vec(L.pellet_wall_impulse, 3);
    o["coupling_impulse"] = Value::vec(L.coupling_impulse, 3);
    o["coupling_angular_impulse"] = Value::vec(L.coupling_angular_impulse, 3);
    o["wall_normal_projection_energy"] = Value::number(L.wall_normal_projection_energy);
    o["wall_friction_dissipation"] = Value::number(L.wall_friction_dissipation);
    o["wall_normal_traction_energy"] = Value::number(L.wall_normal_traction_energy);
    o["pellet_wall_work"] = Value::number(L.pellet_wall_work);
    o["coupling_grid_energy"] = Value::number(L.coupling_grid_energy);
    o["coupling_pellet_energy"] = Value::number(L.coupling_pellet_energy);
    o["plastic_dissipation"] = Value::number(L.plastic_dissipation);
    o["outflow_momentum"] = Value::vec(L.outflow_momentum, 3);
    return o;
}

// snapshot, research.rs:308-321.
Value snapshot_json(const State& st, uint32_t next_record_index) {
    Value o = Value::object();
    o["particles"] = particles_json(st);
    o["pellet"] = st.pellets.empty() ? Value::null() : pellet_json(st.pellets[0]);
    o["ledger"] = ledger_json(st.ledger);
    o["step_count"] = Value::uint(st.history.step_count);
    o["limited_steps"] = Value::uint(st.history.limited_steps);
    o["rejected_steps"] = Value::uint(st.history.rejected_steps);
    o["dt_scale"] = Value::number(st.history.dt_scale);
    o["min_dt_used"] = std::isfinite(st.history.min_dt_used) ? Value::number(st.history.min_dt_used) : Value::null();
    o["max_dt_used"] = Value::number(st.history.max_dt_used);
    o["next_record_index"] = Value::uint(next_record_index);
    return o;
}

Value checkpoint_json(const Value& request, const State& st, uint32_t next_record_index) {
    Value cp = Value::object();
    cp["schema_version"] = Value::uint(1);
    cp["mode"] = Value::string("research");
    cp["request"] = request;
    cp["time_s"] = Value::number(st.history.time);
    cp["state"] = snapshot_json(st, next_record_index);
    return cp;
}

void mat_from(const Value& v, const char* name, double* rm) {
    double col[9];
    as_fixed(v, name, col, 9);
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) rm[3 * r + c] = col[3 * c + r];
}

void parse_particles(const Value& v, State& st) {
    expect_object(v, "ParticleSet");
    deny_unknown(v, {"id", "position", "velocity", "affine", "deformation", "volume_ratio", "kirchhoff", "plastic_strain",
                     "mass", "volume0"});
    const auto& id = as_seq(field(v, "id"), "id");
    const auto& pos = as_seq(field(v, "position"), "position");
    const auto& vel = as_seq(field(v, "velocity"), "velocity");
    const auto& aff = as_seq(field(v, "affine"), "affine");
    const auto& def = as_seq(field(v, "deformation"), "deformation");
    const auto& ratio = as_seq(field(v, "volume_ratio"), "volume_ratio");
    const auto& tau = as_seq(field(v, "kirchhoff"), "kirchhoff");
    const auto& eps = as_seq(field(v, "plastic_strain"), "plastic_strain");
    const auto& mass = as_seq(field(v, "mass"), "mass");
    const auto& vol = as_seq(field(v, "volume0"), "volume0");
    size_t n = id.size();
    const std::pair<const char*, size_t> lengths[] = {{"position", pos.size()},   {"velocity", vel.size()},
                                                       {"affine", aff.size()},     {"deformation", def.size()},
                                                       {"volume_ratio", ratio.size()}, {"kirchhoff", tau.size()},
                                                       {"plastic_strain", eps.size()}, {"mass", mass.size()},
                                                       {"volume0", vol.size()}};
    for (const auto& kv : lengths)
        if (kv.second != n)
            bail(std::string("particle array `") + kv.first + "` has length '" + std::to_string(kv.second) + "', expected '" +
                 std::to_string(n) + "'");
    st.resize(n);
    for (size_t p = 0; p < n; ++p) {
        st.id[p] = as_u32(id[p], "id");
        as_fixed(pos[p], "position", &st.pos[3 * p], 3);
        as_fixed(vel[p], "velocity", &st.vel[3 * p], 3);
        mat_from(aff[p], "affine", &st.aff[9 * p]);
        mat_from(def[p], "deformation", &st.def[9 * p]);
        st.ratio[p] = as_f64(ratio[p], "volume_ratio");
        mat_from(tau[p], "kirchhoff", &st.tau[9 * p]);
        st.eps[p] = as_f64(eps[p], "plastic_strain");
        st.mass[p] = as_f64(mass[p], "mass");
        st.vol0[p] = as_f64(vol[p], "volume0");
    }
}

mpm_pellet parse_pellet(const Value& v) {
    expect_object(v, "Pellet");
    deny_unknown(v, {"radius", "length", "mass", "inertia_body", "position", "orientation", "velocity", "angular_momentum",
                     "sphere_offsets"});
    mpm_pellet p{};
    p.radius = as_f64(field(v, "radius"), "radius");
    p.length = as_f64(field(v, "length"), "length");
    p.mass = as_f64(field(v, "mass"), "mass");
    as_fixed(field(v, "inertia_body"), "inertia_body", p.inertia_body, 3);
    as_fixed(field(v, "position"), "position", p.position, 3);
    as_fixed(field(v, "orientation"), "orientation", p.orientation, 4);
    as_fixed(field(v, "velocity"), "velocity", p.velocity, 3);
    as_fixed(field(v, "angular_momentum"), "angular_momentum", p.angular_momentum, 3);
    const auto& off = as_seq(field(v, "sphere_offsets"), "sphere_offsets");
    if (off.size() > 64) bail("pellet has more than 64 spheres");
    p.sphere_count = (int)off.size();
    for (size_t k = 0; k < off.size(); ++k) p.sphere_offsets[k] = as_f64(off[k], "sphere_offsets");
    return p;
}

mpm_ledger parse_ledger(const Value& v) {
    expect_object(v, "Ledger");
    deny_unknown(v, {"initial_mass", "outflow_mass", "initial_momentum", "initial_mechanical_energy", "gravity_impulse",
                     "wall_impulse", "pellet_wall_impulse", "coupling_impulse", "coupling_angular_impulse",
                     "wall_normal_projection_energy", "wall_friction_dissipation", "wall_normal_traction_energy",
                     "pellet_wall_work", "coupling_grid_energy", "coupling_pellet_energy", "plastic_dissipation",
                     "outflow_momentum"});
    mpm_ledger L{};
    L.initial_mass = as_f64(field(v, "initial_mass"), "initial_mass");
    L.outflow_mass = as_f64(field(v, "outflow_mass"), "outflow_mass");
    as_fixed(field(v, "initial_momentum"), "initial_momentum", L.initial_momentum, 3);
    L.initial_mechanical_energy = as_f64(field(v, "initial_mechanical_energy"), "initial_mechanical_energy");
    as_fixed(field(v, "gravity_impulse"), "gravity_impulse", L.gravity_impulse, 3);
    as_fixed(field(v, "wall_impulse"), "wall_impulse", L.wall_impulse, 3);
    as_fixed(field(v, "pellet_wall_impulse"), "pellet_wall_impulse", L.pellet_wall_impulse, 3);
    as_fixed(field(v, "coupling_impulse"), "coupling_impulse", L.coupling_impulse, 3);
    as_fixed(field(v, "coupling_angular_impulse"), "coupling_angular_impulse", L.coupling_angular_impulse, 3);
    L.wall_normal_projection_energy = as_f64(field(v, "wall_normal_projection_energy"), "wall_normal_projection_energy");
    L.wall_friction_dissipation = as_f64(field(v, "wall_friction_dissipation"), "wall_friction_dissipation");
    L.wall_normal_traction_energy = as_f64(field(v, "wall_normal_traction_energy"), "wall_normal_traction_energy");
    L.pellet_wall_work = as_f64(field(v, "pellet_wall_work"), "pellet_wall_work");
    L.coupling_grid_energy = as_f64(field(v, "coupling_grid_energy"), "coupling_grid_energy");
    L.coupling_pellet_energy = as_f64(field(v, "coupling_pellet_energy"), "coupling_pellet_energy");
    L.plastic_dissipation = as_f64(field(v, "plastic_dissipation"), "plastic_dissipation");
    as_fixed(field(v, "outflow_momentum"), "outflow_momentum", L.outflow_momentum, 3);
    return L;
}

struct Checkpoint {
    uint32_t schema_version;
    std::string mode;
    Value request;
    double time_s;
    State state;
    uint32_t next_record_index;
};

Checkpoint parse_checkpoint(const Value& v) {
    expect_object(v, "Checkpoint");
    deny_unknown(v, {"schema_version", "mode", "request", "time_s", "state"});
    Checkpoint cp;
    cp.schema_version = as_u32(field(v, "schema_version"), "schema_version");
    cp.mode = as_str(field(v, "mode"), "mode");
    cp.request = field(v, "request");
    cp.time_s = as_f64(field(v, "time_s"), "time_s");
    const Value& s = field(v, "state");
    expect_object(s, "Snapshot");
    deny_unknown(s, {"particles", "pellet", "ledger", "step_count", "limited_steps", "rejected_steps", "dt_scale",
                     "min_dt_used", "max_dt_used", "next_record_index"});
    State& st = cp.state;
    parse_particles(field(s, "particles"), st);
    const Value* pel = s.find("pellet");
    st.pellets.clear();
    if (pel && pel->kind != Value::Null) st.pellets.push_back(parse_pellet(*pel));
    st.ledger = parse_ledger(field(s, "ledger"));
    mpm_history& h = st.history;
    h.step_count = as_u64(field(s, "step_count"), "step_count");
    h.limited_steps = as_u64(field(s, "limited_steps"), "limited_steps");
    h.rejected_steps = as_u64(field(s, "rejected_steps"), "rejected_steps");
    h.dt_scale = as_f64(field(s, "dt_scale"), "dt_scale");
    const Value* mn = s.find("min_dt_used");
    h.min_dt_used = (mn && mn->kind != Value::Null) ? as_f64(*mn, "min_dt_used") : std::numeric_limits<double>::infinity();
    h.max_dt_used = as_f64(field(s, "max_dt_used"), "max_dt_used");
    h.time = cp.time_s;
    cp.next_record_index = as_u32(field(s, "next_record_index"), "next_record_index");
    return cp;
}

// validate_snapshot, research.rs:371-427.
void validate_snapshot(const Checkpoint& cp, const Request& r, const State& initial) {
    const State& st = cp.state;
    const mpm_history& h = st.history;
    bool min_bad = std::isfinite(h.min_dt_used) ? (h.min_dt_used <= 0.0 || h.min_dt_used > h.max_dt_used) : false;
    if (!std::isfinite(h.dt_scale) || !(h.dt_scale >= 0.0 && h.dt_scale <= 1.0) || h.dt_scale == 0.0 ||
        !std::isfinite(h.max_dt_used) || h.max_dt_used < 0.0 || min_bad || cp.next_record_index == 0 ||
        cp.next_record_index > 10002 || h.step_count > 1000000000000ull || h.limited_steps > h.step_count ||
        h.rejected_steps > 1000000000000ull)
        bail("invalid research checkpoint integration history");
    double next_time = (double)cp.next_record_index * r.record_interval_s;
    double previous = (double)(cp.next_record_index - 1) * r.record_interval_s;
    if (cp.time_s < r.duration_s && (cp.time_s < previous - 1e-12 || cp.time_s >= next_time))
        bail("research checkpoint recording schedule is inconsistent");
    for (size_t p = 0; p < st.n(); ++p) {
        const double* F = &st.def[9 * p];
        double det = F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) + F[2] * (F[3] * F[7] - F[4] * F[6]);
        if (!std::isfinite(det) || det <= 0.0) bail("research checkpoint contains inverted deformation");
    }
    const mpm_ledger& L = st.ledger;
    const mpm_ledger& I = initial.ledger;
    double dm[3];
    for (int a = 0; a < 3; ++a) dm[a] = L.initial_momentum[a] - I.initial_momentum[a];
    bool ledger_bad = L.outflow_mass < 0.0 || std::fabs(L.initial_mass - I.initial_mass) > 1e-12 ||
                      std::memcmp(&L.initial_mechanical_energy, &I.initial_mechanical_energy, 8) != 0 ||
                      std::sqrt(dm[0] * dm[0] + dm[1] * dm[1] + dm[2] * dm[2]) > 1e-12;
    for (double v : {L.initial_mass, L.outflow_mass, L.initial_mechanical_energy, L.wall_normal_projection_energy,
                     L.wall_friction_dissipation, L.wall_normal_traction_energy, L.pellet_wall_work, L.coupling_grid_energy,
                     L.coupling_pellet_energy, L.plastic_dissipation})
        ledger_bad = ledger_bad || !std::isfinite(v);
    if (ledger_bad) bail("research checkpoint initial inventory or ledger is inconsistent");
    bool pel_ok;
    if (st.pellets.empty() && initial.pellets.empty()) {
        pel_ok = true;
    } else if (st.pellets.size() == 1 && initial.pellets.size() == 1) {
        const mpm_pellet& p = st.pellets[0];
        const mpm_pellet& q = initial.pellets[0];
        const double* o = p.orientation;
        double qn = std::sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2] + o[3] * o[3]);
        pel_ok = pellet_finite(p) && std::memcmp(&p.mass, &q.mass, 8) == 0 && std::memcmp(&p.radius, &q.radius, 8) == 0 &&
                 std::memcmp(&p.length, &q.length, 8) == 0 && std::equal(p.inertia_body, p.inertia_body + 3, q.inertia_body) &&
                 p.sphere_count == q.sphere_count &&
                 std::equal(p.sphere_offsets, p.sphere_offsets + p.sphere_count, q.sphere_offsets) && std::fabs(qn - 1.0) < 1e-10;
    } else {
        pel_ok = false;
    }
    if (!pel_ok) bail("research checkpoint pellet geometry is incompatible");
}

// restore, research.rs:324-368, up to the assignment: the checkpoint checked
// against the request and against the freshly built initial state.
Checkpoint checked_checkpoint(const Spec& spec, const Request& r, const Value& value, const State& initial, double dx,
                              const size_t* cells) {
    Checkpoint cp = parse_checkpoint(value);
    if (cp.schema_version != 1 || cp.mode != "research" || !std::isfinite(cp.time_s) || cp.time_s < 0.0 ||
        cp.time_s > r.duration_s)
        bail("invalid research checkpoint version, mode or time");
    Request original = parse_request(cp.request);
    original.max_wall_time_s = r.max_wall_time_s;
    if (!(original == r)) bail("research checkpoint request differs from resume request");
    validate_particles(cp.state);
    size_t count = spec.particle_count();
    if (cp.state.n() > count) bail("research checkpoint particle identities are incompatible");
    for (uint32_t id : cp.state.id)
        if ((size_t)id >= count) bail("research checkpoint particle identities are incompatible");
    validate_snapshot(cp, r, initial);
    // check_closed_state would refuse a particle outside the grid after the
    // assignment; refusing it before the state reaches the device is the same
    // outcome.
    for (size_t p = 0; p < cp.state.n(); ++p)
        if (!inside_grid(&cp.state.pos[3 * p], dx, cells))
            bail("invalid solver state: particle id '" + std::to_string(cp.state.id[p]) + "' lies outside the grid");
    return cp;
}

uint32_t restore(Sim& sim, const Spec& spec, const Request& r, const Value& value) {
    Checkpoint cp = checked_checkpoint(spec, r, value, sim.st, sim.dx, sim.cells);
    sim.st = cp.state;
    sim.push();
    sim.pull();
    check_closed_state(sim.st, sim.dx, sim.cells);
    return cp.next_record_index;
}

// record, research.rs:460-488.
void record(const Spec& s, const State& st, Value& frames, Value& metrics) {
    const bool water = s.material.kind == 1;
    const char* label = water ? "water" : "waste";
    Value positions = Value::array(), radii = Value::array(), materials = Value::array(), boxes = Value::array();
    for (size_t p = 0; p < st.n(); ++p) {
        positions.a.push_back(Value::vec(&st.pos[3 * p], 3));
        radii.a.push_back(Value::number(s.dx / 4.0));
        materials.a.push_back(Value::string(label));
    }
    for (const mpm_pellet& pl : st.pellets) {
        double r[9];
        mpm_host::qmatrix(pl.orientation, r);
        const double axis[3] = {r[0], r[3], r[6]};
        for (int k = 0; k < pl.sphere_count; ++k) {
            double c[3];
            for (int a = 0; a < 3; ++a) c[a] = pl.position[a] + axis[a] * pl.sphere_offsets[k];
            positions.a.push_back(Value::vec(c, 3));
            radii.a.push_back(Value::number(pl.radius));
            materials.a.push_back(Value::string("wood"));
        }
    }
    for (size_t k = 0; k < positions.a.size(); ++k) boxes.a.push_back(Value::string("research"));
    Value frame = Value::object();
    frame["time_s"] = Value::number(st.history.time);
    frame["positions_m"] = std::move(positions);
    frame["radii_m"] = std::move(radii);
    frame["materials"] = std::move(materials);
    frame["box_ids"] = std::move(boxes);
    frames.a.push_back(std::move(frame));
    double mass = total_mass(st);
    const std::pair<const char*, double> comps[] = {{"domain", mass}, {"outflow", st.ledger.outflow_mass}};
    for (const auto& c : comps) {
        bool domain = std::strcmp(c.first, "domain") == 0;
        Value row = Value::object();
        row["time_s"] = Value::number(st.history.time);
        row["compartment"] = Value::string(c.first);
        row["wood_kg"] = Value::number(domain && !st.pellets.empty() ? st.pellets[0].mass : 0.0);
        row["water_kg"] = Value::number(water ? c.second : 0.0);
        row["waste_kg"] = Value::number(water ? 0.0 : c.second);
        metrics.a.push_back(std::move(row));
    }
}

struct Budget {
    std::chrono::steady_clock::time_point start;
    double limit;
    double elapsed() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count(); }
};
int keep_going(void* ctx) {
    const Budget* b = static_cast<const Budget*>(ctx);
    return b->elapsed() < b->limit ? 1 : 0;
}

Value output_json(const Spec& spec, const Value& request, const State& st, uint32_t next_record, bool completed,
                  Value frames, Value metrics);

// research::run, research.rs:72-132.
Value run(const Value& request, const Value* resume) {
    Budget budget{std::chrono::steady_clock::now(), 0.0};
    Request typed = parse_request(request);
    budget.limit = typed.max_wall_time_s;
    Spec spec = specification(typed);
    Sim sim;
    build(spec, sim);
    sim.pull();
    uint32_t next_record = 1;
    Value frames = Value::array(), metrics = Value::array();
    if (resume) {
        next_record = restore(sim, spec, typed, *resume);
    } else {
        record(spec, sim.st, frames, metrics);
    }
    bool completed = sim.st.history.time >= typed.duration_s;
    while (!completed) {
        double target = std::min((double)next_record * typed.record_interval_s, typed.duration_s);
        int reached = 0;
        int rc = mpm_advance_while(sim.eng, target, keep_going, &budget, &reached);
        if (rc) sim.engine_fail(rc);
        if (!reached) break;
        sim.pull();
        check_closed_state(sim.st, sim.dx, sim.cells);
        record(spec, sim.st, frames, metrics);
        next_record += 1;
        completed = sim.st.history.time >= typed.duration_s;
    }
    sim.pull();
    check_closed_state(sim.st, sim.dx, sim.cells);
    Value out = output_json(spec, request, sim.st, next_record, completed, std::move(frames), std::move(metrics));
    if (budget.elapsed() >= typed.max_wall_time_s) out["status"] = Value::string("budget_exhausted");
    return out;
}

// The output document of a segment, research.rs:102-131.
Value output_json(const Spec& spec, const Value& request, const State& st, uint32_t next_record, bool completed,
                  Value frames, Value metrics) {
    Value checkpoint = checkpoint_json(request, st, next_record);
    std::map<std::string, double> obs = observables(spec, st);
    Value observ = Value::object();
    for (const auto& kv : obs) {
        if (!std::isfinite(kv.second)) bail("research observables became nonfinite");
        observ[kv.first] = Value::number(kv.second);
    }
    Value out = Value::object();
    out["schema_version"] = Value::uint(1);
    out["mode"] = Value::string("research");
    out["status"] = Value::string(completed ? "completed" : "budget_exhausted");
    out["fidelity"] = Value::string("research_unvalidated");
    out["time_s"] = Value::number(st.history.time);
    out["frames"] = std::move(frames);
    out["metrics"] = std::move(metrics);
    out["observables"] = std::move(observ);
    out["checkpoint"] = std::move(checkpoint);
    Value diag = Value::array();
    for (const char* d : {"Unvalidated numerical fixture; no measured surrogate or household calibration",
                          "Single-phase paste mass is booked as waste; no mobile/bound-water split",
                          "No porous fines, wet fragmentation, absorption, adhesion or response-table generation",
                          "Dense padded grid with active-node updates; no sparse-block storage",
                          "Grid-level coupling conserves impulses but can create kinetic energy for light bodies; coupled "
                          "energy stability, leakage and convergence remain unverified",
                          "Wall lattice completion is an approximate contact treatment; its normal transfer is booked as a "
                          "signed grid kinetic-energy diagnostic, not physical wall work; energy closure remains open"})
        diag.a.push_back(Value::string(d));
    out["diagnostics"] = std::move(diag);
    return out;
}

// ------------------------------------------------------------ batched segments
// The reference's sweep and refinement-study runners call run_json once per
// job, one after another (python/litter_physics/sweep.py:206-261,
// verification.py:645-680).  Here every valid job is one instance of a single
// batch engine; all of them advance together, each to its own record
// boundaries, and each job ends with the output run_json would give it.

struct Job {
    Value request;
    Request typed;
    Spec spec;
    int inst = -1;  // batch instance; -1 when the request failed validation
    size_t cells[3] = {0, 0, 0};
    State st;
    uint32_t next_record = 1;
    Value frames = Value::array(), metrics = Value::array();
    bool finished = false;
    Value result;
};

void job_fail(Job& j, const std::string& msg) {
    j.finished = true;
    j.result = Value::object();
    j.result["ok"] = Value::boolean(false);
    j.result["error"] = Value::string(msg);
}

// Final checks and the output document; `elapsed` past the job's budget marks
// it budget_exhausted even when complete, as research::run does.
void job_finish(Job& j, bool completed, double elapsed) {
    try {
        check_closed_state(j.st, j.spec.dx, j.cells);
        Value out = output_json(j.spec, j.request, j.st, j.next_record, completed, std::move(j.frames),
                                std::move(j.metrics));
        if (elapsed >= j.typed.max_wall_time_s) out["status"] = Value::string("budget_exhausted");
        j.finished = true;
        j.result = Value::object();
        j.result["ok"] = Value::boolean(true);
        j.result["output"] = std::move(out);
    } catch (const ResearchError& e) {
        job_fail(j, e.what());
    }
}

void pull_instance(mpm_engine* eng, Job& j) {
    size_t n = mpm_instance_particle_count(eng, (size_t)j.inst);
    j.st.resize(n);
    mpm_particles v = j.st.view();
    j.st.pellets.resize(j.spec.has_pellet && j.spec.fixture == 3 ? 1 : 0);
    int rc = mpm_instance_get_state(eng, (size_t)j.inst, &v, j.st.pellets.data(), j.st.pellets.size(), &j.st.ledger,
                                    &j.st.history);
    if (rc) bail(mpm_last_error(eng), rc);
}

void push_instance(mpm_engine* eng, Job& j, bool seed_ledger) {
    mpm_particles v = j.st.view();
    int rc = mpm_instance_set_state(eng, (size_t)j.inst, &v, j.st.pellets.data(), j.st.pellets.size(),
                                    seed_ledger ? nullptr : &j.st.ledger, seed_ledger ? nullptr : &j.st.history);
    if (rc) bail(mpm_last_error(eng), rc);
}

// Fixture placement and pellet of a job (FixtureSpec::build, fixtures.rs:238-270).
void place(Job& j) {
    const Spec& s = j.spec;
    mpm_fixture f{};
    f.fixture = s.fixture;
    f.material = s.material;
    f.grid_spacing = s.dx;
    for (int a = 0; a < 3; ++a) {
        f.domain[a] = s.domain[a];
        f.initial_size[a] = s.initial_size[a];
        f.initial_velocity[a] = s.initial_velocity[a];
    }
    f.seed = s.seed;
    f.gravity_norm = s.gravity_norm();
    f.jitter = kJitterFraction;
    j.st.resize(s.particle_count());
    mpm_particles view = j.st.view();
    if (mpm_fixture_place(&f, &view) != MPM_OK || view.n != s.particle_count()) bail("fixture placement failed");
    j.st.pellets.clear();
    if (s.fixture == 3 && s.has_pellet) {
        double center[3] = {0.5 * s.domain[0], 0.5 * s.domain[1], s.initial_size[2] + kPelletGapCells * s.dx + s.pel_radius};
        j.st.pellets.push_back(make_pellet(s.pel_radius, s.pel_length, s.pel_density, center));
    }
    grid_layout(s.domain, s.dx, s.max_nodes, j.cells);
}

struct BatchEngine {
    mpm_engine* eng = nullptr;
    ~BatchEngine() {
        if (eng) mpm_destroy(eng);
    }
};

Value run_batch(const Value& requests, const Value* resumes) {
    Budget clock{std::chrono::steady_clock::now(), 0.0};
    if (requests.kind != Value::Array) bail("the batch must be a JSON array of requests");
    if (resumes && (resumes->kind != Value::Array || resumes->a.size() != requests.a.size()))
        bail("resumes must be an array with one checkpoint or null per request");
    std::vector<Job> jobs(requests.a.size());
    std::vector<Job*> live;
    for (size_t k = 0; k < jobs.size(); ++k) {
        Job& j = jobs[k];
        j.request = requests.a[k];
        try {
            j.typed = parse_request(j.request);
            j.spec = specification(j.typed);
            place(j);
            j.inst = (int)live.size();
            live.push_back(&j);
        } catch (const ResearchError& e) {
            job_fail(j, e.what());
        }
    }
    BatchEngine be;
    if (!live.empty()) {
        std::vector<mpm_material> mats;
        std::vector<mpm_config> cfgs;
        std::vector<mpm_layout> lays;
        std::vector<size_t> caps, pels;
        for (Job* j : live) {
            mats.push_back(j->spec.material);
            cfgs.push_back(j->spec.config);
            mpm_layout l{};
            for (int a = 0; a < 3; ++a) l.domain[a] = j->spec.domain[a];
            l.spacing = j->spec.dx;
            l.max_nodes = j->spec.max_nodes;
            lays.push_back(l);
            caps.push_back(j->st.n());
            pels.push_back(j->st.pellets.size());
        }
        int rc = mpm_batch_create(live.size(), mats.data(), cfgs.data(), lays.data(), caps.data(), pels.data(), &be.eng);
        if (rc) bail(std::string("engine creation failed: ") + (be.eng ? mpm_last_error(be.eng) : "no device"), rc);
    }
    // Seed every instance (Simulation::new), then restore or record the first frame.
    for (Job* j : live) {
        try {
            push_instance(be.eng, *j, true);
            pull_instance(be.eng, *j);
            const Value* res = resumes ? &resumes->a[(size_t)(j - jobs.data())] : nullptr;
            if (res && res->kind != Value::Null) {
                Checkpoint cp = checked_checkpoint(j->spec, j->typed, *res, j->st, j->spec.dx, j->cells);
                j->st = cp.state;
                j->next_record = cp.next_record_index;
                push_instance(be.eng, *j, false);
                pull_instance(be.eng, *j);
                check_closed_state(j->st, j->spec.dx, j->cells);
            } else {
                record(j->spec, j->st, j->frames, j->metrics);
            }
            if (j->st.history.time >= j->typed.duration_s) job_finish(*j, true, clock.elapsed());
        } catch (const ResearchError& e) {
            job_fail(*j, e.what());
        }
    }
    // Advance all unfinished jobs towards their next record boundaries, at most
    // 256 trial steps per round so that every job's budget is checked often.
    std::vector<double> targets(live.size());
    std::vector<int> status(live.size());
    for (;;) {
        const double elapsed = clock.elapsed();
        bool any = false;
        for (Job* j : live) {
            if (!j->finished && elapsed >= j->typed.max_wall_time_s) job_finish(*j, false, elapsed);
            any = any || !j->finished;
        }
        if (!any) break;
        for (Job* j : live) {
            mpm_history h{};
            int rc = mpm_instance_get_state(be.eng, (size_t)j->inst, nullptr, nullptr, 0, nullptr, &h);
            if (rc) bail(mpm_last_error(be.eng), rc);
            targets[(size_t)j->inst] =
                j->finished ? h.time
                            : std::max(h.time, std::min((double)j->next_record * j->typed.record_interval_s, j->typed.duration_s));
        }
        int rc = mpm_batch_advance(be.eng, targets.data(), 256, status.data(), nullptr);
        if (rc) bail(mpm_last_error(be.eng), rc);
        const double after = clock.elapsed();
        for (Job* j : live) {
            if (j->finished) continue;
            const int st = status[(size_t)j->inst];
            try {
                if (st < 0) {
                    pull_instance(be.eng, *j);
                    job_fail(*j, "solver aborted at t='" + show(j->st.history.time) +
                                     "' s: step below minimum after '" + std::to_string(j->st.history.rejected_steps) +
                                     "' rejections");
                    continue;
                }
                if (st == 0) {
                    pull_instance(be.eng, *j);
                    check_closed_state(j->st, j->spec.dx, j->cells);
                    record(j->spec, j->st, j->frames, j->metrics);
                    j->next_record += 1;
                    if (j->st.history.time >= j->typed.duration_s) job_finish(*j, true, after);
                }
            } catch (const ResearchError& e) {
                job_fail(*j, e.what());
            }
        }
    }
    Value out = Value::array();
    for (Job& j : jobs) out.a.push_back(std::move(j.result));
    return out;
}

char* dup(const std::string& s) {
    char* p = static_cast<char*>(std::malloc(s.size() + 1));
    if (p) std::memcpy(p, s.c_str(), s.size() + 1);
    return p;
}

int finish(int rc, const std::string& text, char** out) {
    if (out) *out = dup(text);
    return rc;
}

State state_from(const mpm_particles* ps, const mpm_pellet* pel, size_t npel, const mpm_ledger* led, const mpm_history* hist) {
    State st;
    st.load(ps, pel, npel, led, hist);
    return st;
}

}  // namespace

extern "C" {

int mpm_protocol_version(void) { return 1; }

void mpm_string_free(char* s) { std::free(s); }

int mpm_run_json(const char* request_json, const char* resume_json, char** output) {
    if (output) *output = nullptr;
    if (!request_json) return finish(MPM_ERR_INVALID, "request is NULL", output);
    constexpr size_t kMaxJsonBytes = 256u * 1024u * 1024u;
    if (std::strlen(request_json) > kMaxJsonBytes || (resume_json && std::strlen(resume_json) > kMaxJsonBytes))
        return finish(MPM_ERR_INVALID, "native JSON input exceeds 256 MiB", output);
    Value request, resume;
    try {
        request = mpm_json::parse(request_json);
    } catch (const std::exception& e) {
        return finish(MPM_ERR_INVALID, std::string("invalid request JSON: ") + e.what(), output);
    }
    if (resume_json) {
        try {
            resume = mpm_json::parse(resume_json);
        } catch (const std::exception& e) {
            return finish(MPM_ERR_INVALID, std::string("invalid checkpoint JSON: ") + e.what(), output);
        }
    }
    const Value* mode = request.find("mode");
    if (!mode || mode->kind != Value::String || (mode->s != "research" && mode->s != "household"))
        return finish(MPM_ERR_INVALID, "mode must be household or research", output);
    if (mode->s == "household")
        return finish(MPM_ERR_INVALID, "household mode is not part of this engine (research mode only)", output);
    try {
        Value out = run(request, resume_json ? &resume : nullptr);
        return finish(MPM_OK, mpm_json::dump(out), output);
    } catch (const ResearchError& e) {
        return finish(e.code, e.what(), output);
    } catch (const std::exception& e) {
        return finish(MPM_ERR_INVALID, e.what(), output);
    }
}

int mpm_run_json_batch(const char* requests_json, const char* resumes_json, char** output) {
    if (output) *output = nullptr;
    if (!requests_json) return finish(MPM_ERR_INVALID, "requests are NULL", output);
    Value requests, resumes;
    try {
        requests = mpm_json::parse(requests_json);
        if (resumes_json) resumes = mpm_json::parse(resumes_json);
    } catch (const std::exception& e) {
        return finish(MPM_ERR_INVALID, std::string("invalid batch JSON: ") + e.what(), output);
    }
    try {
        return finish(MPM_OK, mpm_json::dump(run_batch(requests, resumes_json ? &resumes : nullptr)), output);
    } catch (const ResearchError& e) {
        return finish(e.code, e.what(), output);
    } catch (const std::exception& e) {
        return finish(MPM_ERR_INVALID, e.what(), output);
    }
}

int mpm_research_observables(const char* request_json, const mpm_particles* particles, const mpm_pellet* pellets,
                             size_t n_pellets, const mpm_ledger* ledger, const mpm_history* history, char** output) {
    if (output) *output = nullptr;
    if (!request_json || !particles || !ledger || !history) return finish(MPM_ERR_INVALID, "NULL argument", output);
    try {
        Request r = parse_request(mpm_json::parse(request_json));
        Spec spec = specification(r);
        State st = state_from(particles, pellets, n_pellets, ledger, history);
        Value o = Value::object();
        for (const auto& kv : observables(spec, st)) o[kv.first] = Value::number(kv.second);
        return finish(MPM_OK, mpm_json::dump(o), output);
    } catch (const std::exception& e) {
        return finish(MPM_ERR_INVALID, e.what(), output);
    }
}

int mpm_checkpoint_encode(const char* request_json, const mpm_particles* particles, const mpm_pellet* pellets,
                          size_t n_pellets, const mpm_ledger* ledger, const mpm_history* history,
                          uint32_t next_record_index, char** output) {
    if (output) *output = nullptr;
    if (!request_json || !particles || !ledger || !history) return finish(MPM_ERR_INVALID, "NULL argument", output);
    try {
        Value request = mpm_json::parse(request_json);
        State st = state_from(particles, pellets, n_pellets, ledger, history);
        return finish(MPM_OK, mpm_json::dump(checkpoint_json(request, st, next_record_index)), output);
    } catch (const std::exception& e) {
        return finish(MPM_ERR_INVALID, e.what(), output);
    }
}

int mpm_checkpoint_decode(const char* checkpoint_json, mpm_particles* particles, mpm_pellet* pellets,
                          size_t pellet_capacity, size_t* n_pellets, mpm_ledger* ledger, mpm_history* history,
                          uint32_t* next_record_index, char** error) {
    if (error) *error = nullptr;
    if (!checkpoint_json || !particles) return finish(MPM_ERR_INVALID, "NULL argument", error);
    try {
        Checkpoint cp = parse_checkpoint(mpm_json::parse(checkpoint_json));
        State& st = cp.state;
        size_t n = st.n();
        if (particles->n < n) {
            particles->n = n;
            return finish(MPM_ERR_CAPACITY, "particle arrays are smaller than the checkpoint", error);
        }
        if (st.pellets.size() > pellet_capacity) return finish(MPM_ERR_CAPACITY, "pellet array is too small", error);
        particles->n = n;
        for (size_t p = 0; p < n; ++p) {
            if (particles->id) particles->id[p] = st.id[p];
            if (particles->volume_ratio) particles->volume_ratio[p] = st.ratio[p];
            if (particles->plastic_strain) particles->plastic_strain[p] = st.eps[p];
            if (particles->mass) particles->mass[p] = st.mass[p];
            if (particles->volume0) particles->volume0[p] = st.vol0[p];
            for (int a = 0; a < 3; ++a) {
                if (particles->position) particles->position[3 * p + a] = st.pos[3 * p + a];
                if (particles->velocity) particles->velocity[3 * p + a] = st.vel[3 * p + a];
            }
            for (int a = 0; a < 9; ++a) {
                if (particles->affine) particles->affine[9 * p + a] = st.aff[9 * p + a];
                if (particles->deformation) particles->deformation[9 * p + a] = st.def[9 * p + a];
                if (particles->kirchhoff) particles->kirchhoff[9 * p + a] = st.tau[9 * p + a];
            }
        }
        for (size_t k = 0; k < st.pellets.size(); ++k) pellets[k] = st.pellets[k];
        if (n_pellets) *n_pellets = st.pellets.size();
        if (ledger) *ledger = st.ledger;
        if (history) *history = st.history;
        if (next_record_index) *next_record_index = cp.next_record_index;
        return MPM_OK;
    } catch (const std::exception& e) {
        return finish(MPM_ERR_INVALID, e.what(), error);
    }
}

int mpm_checkpoint_particle_count(const char* checkpoint_json, size_t* n) {
    if (!checkpoint_json || !n) return MPM_ERR_INVALID;
    try {
        Value v = mpm_json::parse(checkpoint_json);
        const Value* st = v.find("state");
        const Value* ps = st ? st->find("particles") : nullptr;
        const Value* id = ps ? ps->find("id") : nullptr;
        if (!id || id->kind != Value::Array) return MPM_ERR_INVALID;
        *n = id->a.size();
        return MPM_OK;
    } catch (const std::exception&) {
        return MPM_ERR_INVALID;
    }
}

int mpm_research_validate(const char* request_json, char** error) {
    if (error) *error = nullptr;
    if (!request_json) return finish(MPM_ERR_INVALID, "request is NULL", error);
    try {
        Request r = parse_request(mpm_json::parse(request_json));
        Spec spec = specification(r);
        (void)spec;
        return MPM_OK;
    } catch (const std::exception& e) {
        return finish(MPM_ERR_INVALID, e.what(), error);
    }
}

}  // extern "C"
