// Host side of the MI355X MLS-MPM engine: the C ABI of include/mpm_engine.h.
//
// One engine holds a batch of instances (an instance is the reference's
// Simulation, core/src/mpm/solver.rs:324-611).  Every trial step, the adaptive
// dt, accept/reject, the ledgers and the pellets run on the device (kernels.hip);
// the host enqueues a run of trial steps, synchronises once per run and only
// steps in to remove particles that left the grid, to re-sort particles for
// locality and to hand states in and out.  There is no CPU fallback for any
// kernel: without a usable GPU every stepping entry point fails with
// MPM_ERR_DEVICE.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/mpm_engine.h"
#include "engine_internal.h"
#include "host_math.h"

using namespace mpm;

static_assert(sizeof(mpm_ledger) == sizeof(double) * kLedger, "ledger layout");

namespace {

using mpm_host::eigen3;
using mpm_host::qmatrix;

constexpr int kMaxStepsPerSync = 256;
constexpr int kStampsPerStep = 4 * kStampSlots;  // kernel timing stamps of one trial step
// Locality sorts.  Between sorts the particles drift out of their work items'
// cell blocks and the items' tiles grow, so the trial steps slow down, at a
// rate that depends on the flow.  The host re-sorts when the time the rounds
// since the last sort lost to that drift -- each round's time per step above
// the fastest round's since the sort, times its steps -- has reached what a
// sort costs.  For a loss that grows linearly with the steps since the sort
// this is the interval that minimises the total, sqrt(2 C / slope).  The cost
// C is twice the host time of the last periodic sort (the full sort after the
// host set the state is not representative: it sorts by cell and, the first
// time, loads the sort's code; until a periodic sort has been timed,
// kSortCostUs0 + kSortCostUsPerParticle per particle slot).  Twice: sorting
// somewhat later than at the measured break-even was faster on the paste
// column, the dam break and the pellet bed.  Sorts stay kSortMinInterval to
// kSortMaxInterval trial steps apart.
constexpr int kSortMinInterval = 32, kSortMaxInterval = 1024;
constexpr double kSortCostUs0 = 300.0, kSortCostUsPerParticle = 2.5e-4;

// ------------------------------------------------------------ f64 host math

struct V3 {
    double x[3];
    double& operator[](int i) { return x[i]; }
    double operator[](int i) const { return x[i]; }
};
inline V3 v3(double a, double b, double c) { return V3{{a, b, c}}; }
inline double dot(const V3& a, const V3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Pellet::angular_velocity and kinetic energy, rigid.rs:193-210.
V3 pellet_omega(const PelletState& p) {
    double r[9];
    qmatrix(p.q, r);
    V3 b = v3(r[0] * p.L[0] + r[3] * p.L[1] + r[6] * p.L[2], r[1] * p.L[0] + r[4] * p.L[1] + r[7] * p.L[2],
              r[2] * p.L[0] + r[5] * p.L[1] + r[8] * p.L[2]);
    V3 bo = v3(b[0] / p.inertia[0], b[1] / p.inertia[1], b[2] / p.inertia[2]);
    return v3(r[0] * bo[0] + r[1] * bo[1] + r[2] * bo[2], r[3] * bo[0] + r[4] * bo[1] + r[5] * bo[2],
              r[6] * bo[0] + r[7] * bo[1] + r[8] * bo[2]);
}
double pellet_kinetic(const PelletState& p) {
    V3 om = pellet_omega(p);
    return 0.5 * p.mass * (p.vel[0] * p.vel[0] + p.vel[1] * p.vel[1] + p.vel[2] * p.vel[2]) +
           0.5 * (om[0] * p.L[0] + om[1] * p.L[1] + om[2] * p.L[2]);
}

double wave_speed(const mpm_material& m) {
    double modulus = m.kind == 0 ? m.lame_lambda + 2.0 * m.shear_modulus : m.bulk_modulus;
    return std::sqrt(modulus / m.density);
}
double diffusivity(const mpm_material& m) { return m.kind == 0 ? 0.0 : m.consistency / m.density; }

// 2 zeta of the pellet wall contact (rigid.rs:279-290); damping = 2 zeta sqrt(k m).
double damping_ratio2(const mpm_config& c) {
    double zeta;
    if (c.contact_restitution <= 0.0)
        zeta = 1.0;
    else if (c.contact_restitution >= 1.0)
        zeta = 0.0;
    else {
        double lg = std::log(c.contact_restitution);
        zeta = -lg / std::sqrt(M_PI * M_PI + lg * lg);
    }
    return 2.0 * zeta;
}

// Particle mass and reference volume as given, by particle id.  The kernels
// use f32 copies; state exchange, the ledger seed, outflow and the residuals
// use these, so a state read back returns the constants bit for bit and the
// mass ledger closes in f64.  Ids are dense from zero in the reference
// (particles.rs:13); other id sets use a hash map.
struct ExactConstants {
    std::vector<double> dense;  // mass, volume0 per id
    std::unordered_map<uint32_t, std::pair<double, double>> sparse;
    bool use_dense = true;

    void assign(size_t n, const uint32_t* id, const double* mass, const double* volume0) {
        uint32_t top = 0;
        for (size_t p = 0; p < n; ++p) top = std::max(top, id ? id[p] : (uint32_t)p);
        use_dense = (size_t)top < 4 * n + 4096;
        dense.clear();
        sparse.clear();
        if (use_dense) dense.assign(2 * ((size_t)top + 1), 0.0);
        for (size_t p = 0; p < n; ++p) {
            uint32_t k = id ? id[p] : (uint32_t)p;
            if (use_dense) {
                dense[2 * (size_t)k] = mass[p];
                dense[2 * (size_t)k + 1] = volume0[p];
            } else {
                sparse[k] = {mass[p], volume0[p]};
            }
        }
    }
    // Falls back to the stored f32 values for an id it does not hold.
    void get(uint32_t id, float m32, float v32, double& m, double& v0) const {
        m = m32;
        v0 = v32;
        if (use_dense) {
            if (2 * (size_t)id + 1 < dense.size()) m = dense[2 * (size_t)id], v0 = dense[2 * (size_t)id + 1];
        } else {
            auto it = sparse.find(id);
            if (it != sparse.end()) m = it->second.first, v0 = it->second.second;
        }
    }
};

// Host copy of one instance's particles in the f32 slab layout (field-major, n entries).
struct HostParticles {
    size_t n = 0;
    std::vector<float> s, sh;
    std::vector<uint32_t> id;
    std::vector<uint64_t> ref;
};

// Kinetic, potential and elastic energy (solver.rs:1027-1076) from the f32 state.
void energies_of(const mpm_material& mat, const double* g, const HostParticles& h, const ExactConstants& ex,
                 const std::vector<PelletState>& pellets, double* ke_out, double* pe_out, double* el_out) {
    size_t n = h.n;
    double ke = 0.0, pe = 0.0, el = 0.0;
    for (size_t p = 0; p < n; ++p) {
        double m, v0;
        ex.get(h.id[p], h.sh[SMASS * n + p], h.sh[SVOL0 * n + p], m, v0);
        double v2 = 0.0, gx = 0.0;
        for (int a = 0; a < 3; ++a) {
            double v = h.s[(VX + a) * n + p];
            v2 += v * v;
            gx += g[a] * (double)h.s[(PX + a) * n + p];
        }
        ke += 0.5 * m * v2;
        pe += -m * gx;
        if (mat.kind == 0) {
            double F[9], b[9], vals[3];
            for (int a = 0; a < 9; ++a) F[a] = (a % 4 == 0 ? 1.0 : 0.0) + (double)h.s[(H00 + a) * n + p];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    b[3 * i + j] = F[3 * i] * F[3 * j] + F[3 * i + 1] * F[3 * j + 1] + F[3 * i + 2] * F[3 * j + 2];
            if (!eigen3(b, vals)) {
                el = std::numeric_limits<double>::quiet_NaN();
                continue;
            }
            double hh[3], tr = 0.0, h2 = 0.0;
            for (int a = 0; a < 3; ++a) {
                hh[a] = 0.5 * std::log(std::max(vals[a], 1e-300));
                tr += hh[a];
                h2 += hh[a] * hh[a];
            }
            el += v0 * (mat.shear_modulus * h2 + 0.5 * mat.lame_lambda * tr * tr);
        } else {
            double j = 1.0 + (double)h.s[JM1 * n + p];
            double gm = mat.tait_gamma;
            if (gm == 1.0)
                el += v0 * mat.bulk_modulus * (j - 1.0 - std::log(j));
            else
                el += v0 * (mat.bulk_modulus / gm) * (j - 1.0 + (std::pow(j, 1.0 - gm) - 1.0) / (gm - 1.0));
        }
    }
    for (const PelletState& p : pellets) {
        ke += pellet_kinetic(p);
        pe -= p.mass * (g[0] * p.pos[0] + g[1] * p.pos[1] + g[2] * p.pos[2]);
    }
    if (ke_out) *ke_out = ke;
    if (pe_out) *pe_out = pe;
    if (el_out) *el_out = el;
}

// Kinetic, potential and elastic energy (solver.rs:1027-1076) of an f64 state
// in its own order, with the reference's operation order.
void energies_f64(const mpm_material& mat, const double* g, const mpm_particles& ps, const mpm_pellet* pellets,
                  size_t n_pellets, double* ke_out, double* pe_out, double* el_out) {
    double ke = 0.0, pe = 0.0, el = 0.0;
    for (size_t p = 0; p < ps.n; ++p) {
        const double m = ps.mass[p];
        const double* v = ps.velocity ? ps.velocity + 3 * p : nullptr;
        const double v2 = v ? v[0] * v[0] + v[1] * v[1] + v[2] * v[2] : 0.0;
        ke += 0.5 * m * v2;
    }
    for (size_t p = 0; p < ps.n; ++p) {
        const double* x = ps.position + 3 * p;
        pe += -ps.mass[p] * (g[0] * x[0] + g[1] * x[1] + g[2] * x[2]);
    }
    for (size_t p = 0; p < ps.n; ++p) {
        const double v0 = ps.volume0[p];
        if (mat.kind == 0) {
            double F[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, b[9], vals[3];
            if (ps.deformation)
                for (int a = 0; a < 9; ++a) F[a] = ps.deformation[9 * p + a];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    b[3 * i + j] = F[3 * i] * F[3 * j] + F[3 * i + 1] * F[3 * j + 1] + F[3 * i + 2] * F[3 * j + 2];
            if (!eigen3(b, vals)) {
                el = std::numeric_limits<double>::quiet_NaN();
                continue;
            }
            double hh[3];
            for (int a = 0; a < 3; ++a) hh[a] = 0.5 * std::log(std::max(vals[a], 1e-300));
            const double tr = hh[0] + hh[1] + hh[2], h2 = hh[0] * hh[0] + hh[1] * hh[1] + hh[2] * hh[2];
            el += v0 * (mat.shear_modulus * h2 + 0.5 * mat.lame_lambda * tr * tr);
        } else {
            const double j = ps.volume_ratio ? ps.volume_ratio[p] : 1.0;
            const double gm = mat.tait_gamma;
            if (gm == 1.0)
                el += v0 * mat.bulk_modulus * (j - 1.0 - std::log(j));
            else
                el += v0 * (mat.bulk_modulus / gm) * (j - 1.0 + (std::pow(j, 1.0 - gm) - 1.0) / (gm - 1.0));
        }
    }
    for (size_t k = 0; k < n_pellets; ++k) {
        const mpm_pellet& pl = pellets[k];
        ke += mpm_host::pellet_kinetic(pl);
        pe -= pl.mass * (g[0] * pl.position[0] + g[1] * pl.position[1] + g[2] * pl.position[2]);
    }
    if (ke_out) *ke_out = ke;
    if (pe_out) *pe_out = pe;
    if (el_out) *el_out = el;
}

// GridLayout::new's checks, grid.rs:37-86.
int grid_cells(const mpm_layout& l, int* cells) {
    double spacing = l.spacing;
    if (!(std::isfinite(spacing) && spacing > 0.0)) return MPM_ERR_INVALID;
    size_t total = 1;
    for (int a = 0; a < 3; ++a) {
        double ratio = l.domain[a] / spacing;
        double rounded = std::round(ratio);
        if (!std::isfinite(ratio) || std::fabs(ratio - rounded) > 1e-6 * std::max(rounded, 1.0) || rounded < 2.0 ||
            rounded > 4000.0)
            return MPM_ERR_INVALID;
        cells[a] = (int)rounded;
        total *= (size_t)(cells[a] + 3);
    }
    if (total > l.max_nodes) return MPM_ERR_INVALID;
    return MPM_OK;
}

int ceil_log2(uint64_t v) {
    int b = 0;
    while ((1ull << b) < v) ++b;
    return b;
}

}  // namespace

#define HIP_OK(expr)                                      \
    do {                                                  \
        hipError_t e_ = (expr);                           \
        if (e_ != hipSuccess) return fail_device(#expr, e_); \
    } while (0)

struct Instance {
    mpm_material mat{};
    mpm_config cfg{};
    double dx = 0.0;
    int cells[3] = {0, 0, 0};
    int cb[3] = {0, 0, 0};  // 4^3-cell blocks per axis (cells 0..cells inclusive after clamping)
    int pel_cap = 0;
    InstParams ip{};
    bool grid_clear = false;  // both accumulator sets of the instance are zero
    ExactConstants exact;

    void fill_params() {
        Params& P = ip.P;
        P.dx = (float)dx;
        P.inv_dx = (float)(1.0 / dx);
        for (int a = 0; a < 3; ++a) {
            P.cells[a] = cells[a];
            P.nodes[a] = cells[a] + 3;
            P.nb[a] = (cells[a] + 3 + 3) / 4;
            P.extent[a] = (float)(cells[a] * dx);
            P.gravity[a] = (float)cfg.gravity[a];
            cb[a] = (cells[a] + 4) / 4;
        }
        P.kind = mat.kind;
        P.G = (float)mat.shear_modulus;
        P.lam = (float)mat.lame_lambda;
        P.ys = (float)mat.yield_stress;
        P.K = (float)mat.consistency;
        P.inv_n = (float)(1.0 / mat.flow_index);
        P.bulk = (float)mat.bulk_modulus;
        P.gamma = (float)mat.tait_gamma;
        P.visc = (float)mat.consistency;
        for (int i = 0; i < 6; ++i) P.face_mu[i] = (float)cfg.face_friction[i];
        ip.dx = dx;
        ip.cfl = cfg.cfl;
        ip.max_dt = cfg.max_dt;
        ip.min_dt = cfg.min_dt;
        ip.wave = wave_speed(mat);
        ip.diff = diffusivity(mat);
        for (int a = 0; a < 3; ++a) ip.gravity[a] = cfg.gravity[a];
        ip.wall_friction = cfg.wall_friction;
        ip.contact_k = cfg.contact_stiffness;
        ip.contact_damping = damping_ratio2(cfg);
        ip.contact_friction = cfg.contact_friction;
    }
    int node_blocks() const { return ip.P.nb[0] * ip.P.nb[1] * ip.P.nb[2]; }
};

struct mpm_engine {
    int ninst = 0;
    std::vector<Instance> inst;
    size_t cap = 0, na = 0;
    int total_blocks = 0, ngwg = 0, total_pel = 0;
    // particles
    float* slab[2] = {nullptr, nullptr};
    float* shared[2] = {nullptr, nullptr};
    uint32_t* ids[2] = {nullptr, nullptr};
    uint64_t* refkey[2] = {nullptr, nullptr};
    int* nw_list[2] = {nullptr, nullptr};
    int sh = 0;
    // grid
    long long* gmem[2] = {nullptr, nullptr};
    int* gflag[2] = {nullptr, nullptr};
    float* gv = nullptr;
    long long* trS = nullptr;
    long long* trT = nullptr;
    int* blk_inst = nullptr;
    double* gpart = nullptr;
    int* gbad = nullptr;
    double* gchunk = nullptr;
    int* gchunk_bad = nullptr;
    ItemSlot* islot = nullptr;
    // energy audit (allocated on first use)
    double* apart = nullptr;
    double* d_audit = nullptr;
    bool auditing = false;
    // instances
    InstParams* d_ip = nullptr;
    InstState* d_is = nullptr;
    InstState* h_is = nullptr;  // pinned mirror
    int* d_poffs = nullptr;
    int* counters = nullptr;
    // pellets
    PelletState* d_pel = nullptr;
    PelletGeo* d_pgeo = nullptr;
    PelletAccum* d_pacc = nullptr;
    PelletTrial* d_ptrial = nullptr;
    // work items and partials
    int4* d_work = nullptr;
    ItemPre* wpre = nullptr;  // set-up records, 2 x work_cap
    int nwork = 0;
    size_t work_cap = 0;
    float* pmax = nullptr;
    // sort
    uint64_t* keys[2] = {nullptr, nullptr};
    uint32_t* order[2] = {nullptr, nullptr};
    uint64_t* run_unique = nullptr;
    uint32_t* run_count = nullptr;
    int* d_nruns = nullptr;
    // Pinned staging of the run list and the work items (one copy each way per sort).
    int* h_nruns = nullptr;
    uint64_t* h_uq = nullptr;
    uint32_t* h_cnt = nullptr;
    size_t h_run_cap = 0;
    size_t last_nruns = 0;  // run count of the last sort: the size of the next first copy
    int4* h_work = nullptr;
    std::vector<int4> work_build, work_sorted;  // the sort's work-item lists, kept between sorts
    std::vector<int> work_hist;
    // Locality-sort policy (kept across runs): trial steps since the last sort,
    // that sort's host time, the fastest round since it (per trial step; 0: no
    // round yet) and the time the rounds since it lost against that pace.
    int since_sort = 0, sorts_since_full = 0;
    double sort_cost_us = 0.0, best_step_us = 0.0, drift_us = 0.0;  // sort_cost_us 0: not timed yet
    void* sort_temp = nullptr;
    size_t sort_bytes = 0;
    int cell_bits = 0, end_bit = 0;
    bool dirty = true;  // work items, set-up records and near-wall lists must be rebuilt
    // launch
    mpm_launch launch{32, 0, 2, 0, 2, 512};
    hipStream_t stream = nullptr;
    hipGraphExec_t graph = nullptr;
    int graph_steps = 0;
    int graph_twgs = 0;  // traction workgroups the captured graph launches
    bool graph_stale = true;
    // Kernel timing: four device-clock stamps per enqueued step (BatchDev::ktime).
    unsigned long long* d_ktime = nullptr;
    unsigned long long* d_run_clock = nullptr;  // device clock at the start of run()
    unsigned long long* h_ktime = nullptr;  // pinned
    double clock_khz = 0.0;
    bool profiling = false;
    int profile_every = 1;        // time every profile_every-th step
    uint64_t profile_seq = 0;     // steps since profiling was enabled
    bool profiled[kMaxStepsPerSync] = {};  // the enqueued steps whose kernels carry events
    double g2p_ms = 0.0, p2g_ms = 0.0;
    uint64_t profiled_launches = 0;
    // return-map surrogate and the log of exact returns
    const mpm_mlp* mlp = nullptr;
    float mlp_tol = 1e-4f;
    int mlp_polish = 2;
    unsigned long long* mlp_stats = nullptr;
    float* rlog = nullptr;
    int* rlog_count = nullptr;
    int rlog_cap = 0;
    std::string err;

    int fail_device(const char* what, hipError_t e) {
        err = std::string("HIP error in ") + what + ": " + hipGetErrorString(e);
        return MPM_ERR_DEVICE;
    }
    int fail(int code, const std::string& msg) {
        err = msg;
        return code;
    }
    bool any_pellets() const {
        for (const Instance& in : inst)
            if (in.ip.npel > 0) return true;
        return false;
    }

    BatchDev dev() const {
        BatchDev B{};
        B.ninst = ninst;
        B.cap = cap;
        B.na = na;
        B.ip = d_ip;
        B.is = d_is;
        B.poffs = d_poffs;
        for (int b = 0; b < 2; ++b) {
            B.slab[b] = slab[b];
            B.shared[b] = shared[b];
            B.gmem[b] = gmem[b];
            B.gflag[b] = gflag[b];
            B.nw_list[b] = nw_list[b];
        }
        B.sh = sh;
        B.g2p_occupancy = launch.g2p_occupancy;
        B.p2g_block = launch.p2g_block;
        B.gv = gv;
        B.trS = trS;
        B.trT = trT;
        B.ids = ids[sh];
        B.refkey = refkey[sh];
        B.blk_inst = blk_inst;
        B.work = d_work;
        B.wpre[0] = wpre;
        B.wpre[1] = wpre + work_cap;
        B.nwork = nwork;
        B.ngwg = ngwg;
        B.counters = counters;
        B.gpart = gpart;
        B.gbad = gbad;
        B.apart = auditing ? apart : nullptr;
        B.audit = auditing ? d_audit : nullptr;
        B.gchunk = gchunk;
        B.gchunk_bad = gchunk_bad;
        B.islot = islot;
        B.pmax = pmax;
        B.pel = d_pel;
        B.pgeo = d_pgeo;
        B.pacc = d_pacc;
        B.ptrial = d_ptrial;
        B.npel = total_pel;
        B.mlp_fwd = mlp ? mlp_device_fragments(mlp) : nullptr;
        B.mlp_bias = mlp ? mlp_device_biases(mlp) : nullptr;
        B.mlp_tol = mlp_tol;
        B.mlp_polish = mlp_polish;
        B.mlp_stats = mlp_stats;
        B.rlog = rlog;
        B.rlog_count = rlog_count;
        B.rlog_cap = rlog_cap;
        return B;
    }

    // Wait for the stream by polling it.  A blocking synchronisation sleeps, and
    // its wake-up cost tens of microseconds of idle GPU per host round trip.
    int sync() {
        hipError_t e;
        while ((e = hipStreamQuery(stream)) == hipErrorNotReady) std::this_thread::yield();
        HIP_OK(e);
        return MPM_OK;
    }
    int pull() {
        HIP_OK(hipMemcpyAsync(h_is, d_is, sizeof(InstState) * ninst, hipMemcpyDeviceToHost, stream));
        return sync();
    }
    // Upload the mirror.  Without `wait`, the caller must not touch the mirror
    // before a later synchronisation of the stream (a pull() is one).
    int push(bool wait = true) {
        // Any host change may have touched the pellets (state upload, restore):
        // the next trial recomputes their geometry.
        for (int i = 0; i < ninst; ++i) h_is[i].geo_stale = 1;
        HIP_OK(hipMemcpyAsync(d_is, h_is, sizeof(InstState) * ninst, hipMemcpyHostToDevice, stream));
        return wait ? sync() : MPM_OK;
    }
    int upload_params(int i) {
        HIP_OK(hipMemcpyAsync(d_ip + i, &inst[i].ip, sizeof(InstParams), hipMemcpyHostToDevice, stream));
        if (int rc_ = sync()) return rc_;
        return MPM_OK;
    }

    template <class T>
    int dalloc(T** p, size_t n, bool zero = true) {
        HIP_OK(hipMalloc((void**)p, sizeof(T) * std::max<size_t>(n, 1)));
        if (zero) HIP_OK(hipMemset(*p, 0, sizeof(T) * std::max<size_t>(n, 1)));
        return MPM_OK;
    }

    int alloc(const size_t* capacities, const size_t* max_pellets) {
        HIP_OK(hipStreamCreate(&stream));
        std::vector<int> poffs(ninst + 1, 0);
        std::vector<int> binst;
        size_t total = 0, wcap = 0;
        int max_cb = 1;
        for (int i = 0; i < ninst; ++i) {
            Instance& in = inst[i];
            size_t pc = std::max<size_t>(capacities ? capacities[i] : 1, 1);
            if (total + pc > (size_t)INT_MAX / 2) return fail(MPM_ERR_CAPACITY, "total particle capacity too large");
            in.ip.poff = (int)total;
            in.ip.pcap = (int)pc;
            poffs[i] = (int)total;
            total += pc;
            int nb = in.node_blocks();
            int nblk = (nb + kGridWgBlocks - 1) / kGridWgBlocks * kGridWgBlocks;
            in.ip.boff = total_blocks;
            in.ip.nblk = nblk;
            for (int b = 0; b < nblk; ++b) binst.push_back(i);
            total_blocks += nblk;
            in.pel_cap = (int)(max_pellets ? max_pellets[i] : 0);
            in.ip.pel_off = total_pel;
            in.ip.npel = 0;
            total_pel += in.pel_cap;
            int cbs = in.cb[0] * in.cb[1] * in.cb[2];
            max_cb = std::max(max_cb, cbs);
            wcap += (size_t)cbs + pc / (size_t)g2p_chunk() + 16;
        }
        poffs[ninst] = (int)total;
        cap = total;
        na = (size_t)total_blocks * 64;
        ngwg = total_blocks / kGridWgBlocks;
        work_cap = wcap;
        // Device-order key: (instance, cell block, cell within block); the all-ones
        // block value marks the unused entries of an instance's range.
        cell_bits = ceil_log2((uint64_t)max_cb + 1) + 6;
        end_bit = cell_bits + ceil_log2((uint64_t)ninst);
        if (end_bit > 63) return fail(MPM_ERR_CAPACITY, "too many instances or cells for the sort key");
        for (int b = 0; b < 2; ++b) {
            if (int rc = dalloc(&slab[b], NPF * cap)) return rc;
            if (int rc = dalloc(&shared[b], NSF * cap)) return rc;
            if (int rc = dalloc(&ids[b], cap)) return rc;
            if (int rc = dalloc(&refkey[b], cap)) return rc;
            if (int rc = dalloc(&nw_list[b], cap)) return rc;
            if (int rc = dalloc(&keys[b], cap, false)) return rc;
            if (int rc = dalloc(&order[b], cap, false)) return rc;
            if (int rc = dalloc(&gmem[b], (size_t)kAuditFields * na)) return rc;
            if (int rc = dalloc(&gflag[b], (size_t)total_blocks)) return rc;
        }
        if (int rc = dalloc(&gv, 3 * na)) return rc;
        if (int rc = dalloc(&trS, 3 * na)) return rc;
        if (int rc = dalloc(&trT, 3 * na)) return rc;
        if (int rc = dalloc(&blk_inst, (size_t)total_blocks)) return rc;
        HIP_OK(hipMemcpy(blk_inst, binst.data(), sizeof(int) * binst.size(), hipMemcpyHostToDevice));
        if (int rc = dalloc(&gpart, (size_t)kGridVals * ngwg)) return rc;
        if (int rc = dalloc(&gbad, (size_t)ngwg)) return rc;
        if (int rc = dalloc(&gchunk, (size_t)ninst * kGridVals * kGridChunks)) return rc;
        if (int rc = dalloc(&gchunk_bad, (size_t)ninst * kGridChunks)) return rc;
        if (int rc = dalloc(&islot, (size_t)ninst * kItemSlots)) return rc;
        {
            std::vector<ItemSlot> idle((size_t)ninst * kItemSlots);
            for (ItemSlot& sl : idle) {
                std::memset(&sl, 0, sizeof(sl));
                sl.fail = ~0ull;
            }
            HIP_OK(hipMemcpy(islot, idle.data(), sizeof(ItemSlot) * idle.size(), hipMemcpyHostToDevice));
        }
        if (int rc = dalloc(&d_ip, (size_t)ninst)) return rc;
        if (int rc = dalloc(&d_is, (size_t)ninst)) return rc;
        HIP_OK(hipHostMalloc((void**)&h_is, sizeof(InstState) * ninst, hipHostMallocDefault));
        std::memset(h_is, 0, sizeof(InstState) * ninst);
        for (int i = 0; i < ninst; ++i) {
            h_is[i].mode = M_DONE;
            h_is[i].dt_scale = 1.0;
            h_is[i].min_dt_used = std::numeric_limits<double>::infinity();
        }
        if (int rc = dalloc(&d_poffs, (size_t)ninst + 1)) return rc;
        HIP_OK(hipMemcpy(d_poffs, poffs.data(), sizeof(int) * poffs.size(), hipMemcpyHostToDevice));
        if (int rc = dalloc(&counters, 2)) return rc;
        if (int rc = dalloc(&d_pel, (size_t)total_pel)) return rc;
        if (int rc = dalloc(&d_pgeo, (size_t)total_pel)) return rc;
        if (int rc = dalloc(&d_pacc, (size_t)total_pel)) return rc;
        if (int rc = dalloc(&d_ptrial, (size_t)total_pel)) return rc;
        if (total_pel) {
            std::vector<PelletState> empty(total_pel);
            for (PelletState& p : empty) {
                std::memset(&p, 0, sizeof(p));
                p.inst = -1;
            }
            HIP_OK(hipMemcpy(d_pel, empty.data(), sizeof(PelletState) * total_pel, hipMemcpyHostToDevice));
        }
        if (int rc = dalloc(&d_work, work_cap)) return rc;
        if (int rc = dalloc(&wpre, 2 * work_cap)) return rc;
        if (int rc = dalloc(&pmax, 4 * work_cap)) return rc;
        if (int rc = dalloc(&run_unique, cap, false)) return rc;
        if (int rc = dalloc(&run_count, cap, false)) return rc;
        if (int rc = dalloc(&d_nruns, 1)) return rc;
        HIP_OK(hipHostMalloc((void**)&h_nruns, sizeof(int), hipHostMallocDefault));
        HIP_OK(hipHostMalloc((void**)&h_work, sizeof(int4) * std::max<size_t>(work_cap, 1), hipHostMallocDefault));
        if (int rc = dalloc(&mlp_stats, 2)) return rc;
        size_t sort_need = 0, rle_need = 0;
        HIP_OK(sort_temp_bytes((int)cap, end_bit, &sort_need));
        HIP_OK(rle_temp_bytes((int)cap, end_bit, &rle_need));
        sort_bytes = std::max(sort_need, rle_need);
        HIP_OK(hipMalloc(&sort_temp, std::max<size_t>(sort_bytes, 16)));
        // Kernel timing stamps (the kernels write them; no event records, which
        // would leave the GPU idle for microseconds between the kernels they separate).
        if (int rc = dalloc(&d_ktime, (size_t)kStampsPerStep * kMaxStepsPerSync)) return rc;
        if (int rc = dalloc(&d_run_clock, 1)) return rc;
        HIP_OK(hipHostMalloc((void**)&h_ktime, sizeof(unsigned long long) * kStampsPerStep * kMaxStepsPerSync,
                             hipHostMallocDefault));
        {
            int dev = 0, khz = 0;
            HIP_OK(hipGetDevice(&dev));
            HIP_OK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
            clock_khz = (double)khz;
        }
        for (int i = 0; i < ninst; ++i)
            if (int rc = upload_params(i)) return rc;
        return push();
    }

    ~mpm_engine() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (graph) (void)hipGraphExecDestroy(graph);
        for (int b = 0; b < 2; ++b) {
            (void)hipFree(slab[b]);
            (void)hipFree(shared[b]);
            (void)hipFree(ids[b]);
            (void)hipFree(refkey[b]);
            (void)hipFree(nw_list[b]);
            (void)hipFree(keys[b]);
            (void)hipFree(order[b]);
            (void)hipFree(gmem[b]);
            (void)hipFree(gflag[b]);
        }
        (void)hipFree(gv);
        (void)hipFree(trS);
        (void)hipFree(trT);
        (void)hipFree(blk_inst);
        (void)hipFree(gpart);
        (void)hipFree(gbad);
        (void)hipFree(gchunk);
        (void)hipFree(gchunk_bad);
        (void)hipFree(islot);
        (void)hipFree(apart);
        (void)hipFree(d_audit);
        (void)hipFree(d_ip);
        (void)hipFree(d_is);
        if (h_is) (void)hipHostFree(h_is);
        (void)hipFree(d_poffs);
        (void)hipFree(counters);
        (void)hipFree(d_pel);
        (void)hipFree(d_pgeo);
        (void)hipFree(d_pacc);
        (void)hipFree(d_ptrial);
        (void)hipFree(d_work);
        (void)hipFree(wpre);
        (void)hipFree(pmax);
        (void)hipFree(run_unique);
        (void)hipFree(run_count);
        (void)hipFree(d_nruns);
        if (h_nruns) (void)hipHostFree(h_nruns);
        if (h_uq) (void)hipHostFree(h_uq);
        if (h_cnt) (void)hipHostFree(h_cnt);
        if (h_work) (void)hipHostFree(h_work);
        (void)hipFree(mlp_stats);
        (void)hipFree(rlog);
        (void)hipFree(rlog_count);
        (void)hipFree(sort_temp);
        (void)hipFree(d_ktime);
        (void)hipFree(d_run_clock);
        if (h_ktime) (void)hipHostFree(h_ktime);
        if (stream) (void)hipStreamDestroy(stream);
    }

    // Order keys still pending for instances whose last accepted step hit the
    // 32-step cadence and that took no trial step since (the next trial's G2P2G
    // writes them otherwise).  The mirror must be current.
    // With `upload` false the cleared flags reach the device with the caller's
    // next push (nothing launched before it reads them).
    int flush_refkeys(bool upload = true) {
        bool any = false;
        for (int i = 0; i < ninst; ++i) any = any || h_is[i].refsort;
        if (!any) return MPM_OK;
        launch_refkeys(stream, dev());
        HIP_OK(hipGetLastError());
        for (int i = 0; i < ninst; ++i) h_is[i].refsort = 0;
        return upload ? push() : MPM_OK;
    }

    // Pinned staging for at least n runs (contents are not kept).  It grows to
    // twice the request: pinned allocations cost far more than the copies.
    int reserve_runs(size_t n) {
        if (n <= h_run_cap) return MPM_OK;
        n = std::min(cap, 2 * n);
        if (h_uq) (void)hipHostFree(h_uq);
        if (h_cnt) (void)hipHostFree(h_cnt);
        h_uq = nullptr;
        h_cnt = nullptr;
        h_run_cap = 0;
        HIP_OK(hipHostMalloc((void**)&h_uq, sizeof(uint64_t) * n, hipHostMallocDefault));
        HIP_OK(hipHostMalloc((void**)&h_cnt, sizeof(uint32_t) * n, hipHostMallocDefault));
        h_run_cap = n;
        return MPM_OK;
    }

    // Sort every instance's particles by 4^3-cell block (device order, for
    // locality), rebuild the work items and the near-wall lists.  The order keys
    // of the reference order travel with the particles.  The mirror must be
    // current; it is current again on return.
    int sort_all() {
        const auto t_sort = std::chrono::steady_clock::now();
        const bool full = dirty;
        if (int rc = flush_refkeys(false)) return rc;
        BatchDev B = dev();
        launch_sort_keys(stream, B, cell_bits, end_bit, keys[0], order[0]);
        HIP_OK(hipGetLastError());
        // A state set from the host is sorted by cell; the periodic locality
        // sorts order by cell block only (two radix passes at 1M particles in one
        // instance instead of three): within a block the stable order of the
        // previous sort stays.  The drift does disturb it over a long run (the
        // deposit's LDS atomics of particles of one cell then no longer go to
        // lanes side by side), so every fourth periodic sort orders by cell again:
        // 188.5 against 191.5 us per step over 4000 steps of the 1M-particle
        // column, the same over its first 500.
        const int begin_bit = (dirty || ++sorts_since_full >= 4) ? 0 : 6;
        if (begin_bit == 0) sorts_since_full = 0;
        HIP_OK(launch_radix_sort(stream, sort_temp, sort_bytes, keys[0], keys[1], order[0], order[1], (int)cap,
                                 begin_bit, end_bit));
        launch_permute(stream, B, order[1], shared[sh], shared[sh ^ 1], ids[sh], ids[sh ^ 1], refkey[sh],
                       refkey[sh ^ 1]);
        HIP_OK(hipGetLastError());
        HIP_OK(launch_block_runs(stream, sort_temp, sort_bytes, (int)cap, end_bit, keys[1], run_unique, run_count,
                                 d_nruns));
        // The run list comes back with its count in one round trip: the first
        // `guess` runs (the last sort's count and a margin), and the rest in a
        // second copy only if the count outgrew the guess.
        const size_t guess = std::min(cap, last_nruns + last_nruns / 4 + 1024);
        if (int rc = reserve_runs(guess)) return rc;
        HIP_OK(hipMemcpyAsync(h_nruns, d_nruns, sizeof(int), hipMemcpyDeviceToHost, stream));
        HIP_OK(hipMemcpyAsync(h_uq, run_unique, sizeof(uint64_t) * guess, hipMemcpyDeviceToHost, stream));
        HIP_OK(hipMemcpyAsync(h_cnt, run_count, sizeof(uint32_t) * guess, hipMemcpyDeviceToHost, stream));
        if (int rc_ = sync()) return rc_;
        const int nruns = *h_nruns;
        if (nruns < 0 || (size_t)nruns > cap) return fail(MPM_ERR_DEVICE, "run-length encoding count out of range");
        if ((size_t)nruns > guess) {
            if (int rc = reserve_runs((size_t)nruns)) return rc;
            HIP_OK(hipMemcpyAsync(h_uq, run_unique, sizeof(uint64_t) * nruns, hipMemcpyDeviceToHost, stream));
            HIP_OK(hipMemcpyAsync(h_cnt, run_count, sizeof(uint32_t) * nruns, hipMemcpyDeviceToHost, stream));
            if (int rc_ = sync()) return rc_;
        }
        last_nruns = (size_t)nruns;
        const uint64_t* uq = h_uq;
        const uint32_t* cnt = h_cnt;
        sh ^= 1;
        for (int i = 0; i < ninst; ++i) h_is[i].cur ^= 1;
        // Work items: runs of one cell block, merged along z up to launch.merge_blocks
        // blocks of one instance and one chunk of particles; longer runs split.
        const int chunk = g2p_chunk();
        const int shift = cell_bits - 6;
        const uint64_t bmask = (1ull << shift) - 1;
        std::vector<int4>& work = work_build;
        work.clear();
        work.reserve(nruns + cap / chunk + 1);
        int start = 0, ci = -1, cstart = -1, ccount = 0, cblocks = 0, cbz = 1, zpos = 0;
        uint64_t prev = 0;
        auto flush = [&]() {
            if (cstart >= 0) work.push_back(make_int4(ci, cstart, ccount, 0));
            cstart = -1;
            ccount = 0;
            cblocks = 0;
        };
        for (int r = 0; r < nruns; ++r) {
            const int c = (int)cnt[r];
            const int in = (int)(uq[r] >> shift);
            const uint64_t blk = uq[r] & bmask;
            if (blk == bmask || in >= ninst) {
                flush();
                start += c;
                continue;
            }
            // z position of the block in its column: counted along a run of
            // consecutive blocks, a 64-bit modulo only where the run breaks.
            if (in == ci && blk == prev + 1) {
                zpos = zpos + 1 == cbz ? 0 : zpos + 1;
            } else {
                cbz = inst[in].cb[2];
                zpos = (int)(blk % (uint64_t)cbz);
            }
            const bool adjacent =
                cstart >= 0 && ci == in && cblocks < launch.merge_blocks && blk == prev + 1 && zpos != 0 &&
                ccount + c <= chunk;
            if (!adjacent) flush();
            if (c > chunk) {
                for (int off = 0; off < c; off += chunk)
                    work.push_back(make_int4(in, start + off, std::min(chunk, c - off), 0));
            } else if (adjacent) {
                ccount += c;
                cblocks += 1;
            } else {
                cstart = start;
                ccount = c;
                cblocks = 1;
            }
            ci = in;
            prev = blk;
            start += c;
        }
        flush();
        // Largest items first within each instance's range.  A workgroup's time
        // grows with its particle count, and workgroups start in index order, so
        // the small items fill the end of the launch instead of a few large ones
        // running on alone (1M-particle column: 173 us of modelled launch time
        // becomes 151 us).  Each instance's items stay contiguous.  A stable
        // counting sort on the particle count (at most one chunk); the list is
        // already grouped by instance and each group is ordered on its own.
        {
            std::vector<int>& hist = work_hist;
            hist.assign((size_t)chunk + 2, 0);
            std::vector<int4>& out = work_sorted;
            out.resize(work.size());
            for (size_t b = 0; b < work.size();) {
                size_t e = b;
                while (e < work.size() && work[e].x == work[b].x) ++e;
                for (size_t w = b; w < e; ++w) hist[chunk - work[w].z + 1] += 1;
                int run = (int)b;
                for (int k = 0; k <= chunk + 1; ++k) {
                    const int h = hist[k];
                    hist[k] = run;
                    run += h;
                }
                for (size_t w = b; w < e; ++w) out[hist[chunk - work[w].z + 1]++] = work[w];
                std::fill(hist.begin(), hist.end(), 0);
                b = e;
            }
            work.swap(out);
        }
        if (work.size() > work_cap) return fail(MPM_ERR_CAPACITY, "work list overflow");
        nwork = (int)work.size();
        for (int i = 0; i < ninst; ++i) {
            h_is[i].w_off = 0;
            h_is[i].w_cnt = 0;
        }
        for (int w = nwork - 1; w >= 0; --w) {
            InstState& s = h_is[work[w].x];
            s.w_off = w;
            s.w_cnt += 1;
        }
        for (int i = 0; i < ninst; ++i) {
            InstState& s = h_is[i];
            s.nw_count[s.cur] = 0;
        }
        if (nwork) {
            std::memcpy(h_work, work.data(), sizeof(int4) * nwork);
            HIP_OK(hipMemcpyAsync(d_work, h_work, sizeof(int4) * nwork, hipMemcpyHostToDevice, stream));
        }
        // The pull below synchronises the stream before the host touches the
        // mirror or the staged work list again.
        if (int rc = push(false)) return rc;
        launch_setup(stream, dev());
        HIP_OK(hipGetLastError());
        if (int rc = pull()) return rc;
        dirty = false;
        graph_stale = true;
        if (!full)
            sort_cost_us =
                2.0 * std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t_sort).count();
        since_sort = 0;
        best_step_us = drift_us = 0.0;
        return MPM_OK;
    }

    // Zero both accumulator sets of instance i: each trial deposits into the
    // set its predecessor left cleared, and a state set from outside starts the
    // alternation afresh.
    int clear_grid(int i) {
        Instance& in = inst[i];
        size_t noff = (size_t)in.ip.boff * 64, nn = (size_t)in.ip.nblk * 64;
        for (int g = 0; g < 2; ++g) {
            for (int f = 0; f < kAuditFields; ++f)
                HIP_OK(hipMemsetAsync(gmem[g] + f * na + noff, 0, sizeof(long long) * nn, stream));
            HIP_OK(hipMemsetAsync(gflag[g] + in.ip.boff, 0, sizeof(int) * in.ip.nblk, stream));
        }
        if (int rc_ = sync()) return rc_;
        in.grid_clear = true;
        return MPM_OK;
    }

    int prepare() {
        if (dirty)
            if (int rc = sort_all()) return rc;
        for (int i = 0; i < ninst; ++i)
            if (!inst[i].grid_clear)
                if (int rc = clear_grid(i)) return rc;
        return MPM_OK;
    }

    // Copy instance i's committed particles to the host (mirror current).
    int download(int i, HostParticles& h) {
        if (int rc = flush_refkeys()) return rc;
        const Instance& in = inst[i];
        size_t n = (size_t)h_is[i].n, off = (size_t)in.ip.poff;
        int cur = h_is[i].cur;
        h.n = n;
        h.s.assign(NPF * n, 0.0f);
        h.sh.assign(NSF * n, 0.0f);
        h.id.assign(n, 0);
        h.ref.assign(n, 0);
        if (n == 0) return MPM_OK;
        for (int f = 0; f < NPF; ++f)
            HIP_OK(hipMemcpyAsync(h.s.data() + f * n, slab[cur] + f * cap + off, sizeof(float) * n,
                                  hipMemcpyDeviceToHost, stream));
        for (int f = 0; f < NSF; ++f)
            HIP_OK(hipMemcpyAsync(h.sh.data() + f * n, shared[sh] + f * cap + off, sizeof(float) * n,
                                  hipMemcpyDeviceToHost, stream));
        HIP_OK(hipMemcpyAsync(h.id.data(), ids[sh] + off, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, stream));
        HIP_OK(hipMemcpyAsync(h.ref.data(), refkey[sh] + off, sizeof(uint64_t) * n, hipMemcpyDeviceToHost, stream));
        if (int rc_ = sync()) return rc_;
        return MPM_OK;
    }
    // Write instance i's particles into both slabs (and the current shared slab).
    int upload(int i, const HostParticles& h) {
        const Instance& in = inst[i];
        size_t n = h.n, off = (size_t)in.ip.poff;
        if (n) {
            for (int b = 0; b < 2; ++b)
                for (int f = 0; f < NPF; ++f)
                    HIP_OK(hipMemcpyAsync(slab[b] + f * cap + off, h.s.data() + f * n, sizeof(float) * n,
                                          hipMemcpyHostToDevice, stream));
            for (int f = 0; f < NSF; ++f)
                HIP_OK(hipMemcpyAsync(shared[sh] + f * cap + off, h.sh.data() + f * n, sizeof(float) * n,
                                      hipMemcpyHostToDevice, stream));
            HIP_OK(hipMemcpyAsync(ids[sh] + off, h.id.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, stream));
            HIP_OK(hipMemcpyAsync(refkey[sh] + off, h.ref.data(), sizeof(uint64_t) * n, hipMemcpyHostToDevice,
                                  stream));
        }
        if (int rc_ = sync()) return rc_;
        h_is[i].n = (int)n;
        dirty = true;
        return MPM_OK;
    }
    int download_pellets(int i, std::vector<PelletState>& out) {
        const Instance& in = inst[i];
        out.resize(in.ip.npel);
        if (in.ip.npel) {
            HIP_OK(hipMemcpyAsync(out.data(), d_pel + in.ip.pel_off, sizeof(PelletState) * in.ip.npel,
                                  hipMemcpyDeviceToHost, stream));
            if (int rc_ = sync()) return rc_;
        }
        return MPM_OK;
    }

    // Remove instance i's particles that left the padded grid (sweep_outflow,
    // solver.rs:520-532), booking their mass and momentum as outflow.
    int remove_escaped(int i) {
        Instance& in = inst[i];
        HostParticles h, k;
        if (int rc = download(i, h)) return rc;
        size_t n = h.n, keep = 0;
        std::vector<char> inside(n, 1);
        double out_mass = 0.0, out_mom[3] = {0.0, 0.0, 0.0};
        for (size_t p = 0; p < n; ++p) {
            for (int a = 0; a < 3; ++a) {
                double sc = (double)h.s[(PX + a) * n + p] / in.dx;
                double padded = std::floor(sc - 0.5) + 1.0;
                if (std::isnan(padded) || padded < 0.0 || padded + 2.0 >= (double)(in.cells[a] + 3)) inside[p] = 0;
            }
            if (!inside[p]) {
                double m, v0;
                in.exact.get(h.id[p], h.sh[SMASS * n + p], h.sh[SVOL0 * n + p], m, v0);
                out_mass += m;
                for (int a = 0; a < 3; ++a) out_mom[a] += (double)h.s[(VX + a) * n + p] * m;
            } else {
                ++keep;
            }
        }
        k.n = keep;
        k.s.resize(NPF * keep);
        k.sh.resize(NSF * keep);
        k.id.resize(keep);
        k.ref.resize(keep);
        size_t q = 0;
        for (size_t p = 0; p < n; ++p) {
            if (!inside[p]) continue;
            for (int f = 0; f < NPF; ++f) k.s[f * keep + q] = h.s[f * n + p];
            for (int f = 0; f < NSF; ++f) k.sh[f * keep + q] = h.sh[f * n + p];
            k.id[q] = h.id[p];
            k.ref[q] = h.ref[p];
            ++q;
        }
        InstState& s = h_is[i];
        s.ledger[1] += out_mass;
        for (int a = 0; a < 3; ++a) s.ledger[28 + a] += out_mom[a];
        in.ip.particle_mass -= out_mass;
        float mmax = 0.0f;
        for (size_t q2 = 0; q2 < keep; ++q2) mmax = std::max(mmax, k.sh[SMASS * keep + q2]);
        in.ip.mass_max = mmax;
        if (int rc = upload_params(i)) return rc;
        return upload(i, k);
    }

    // Workgroups of the wall-traction kernel: the launch shape's count, or (0)
    // enough threads for one pass over the near-wall lists as of the last
    // synchronisation (nine per particle, a quarter spare), rounded up to a power
    // of two so that a captured graph is rebuilt only when the lists change
    // much.  The kernel's grid-stride loop covers any excess.
    int traction_workgroups() const {
        if (launch.traction_wgs > 0) return launch.traction_wgs;
        long long nw = 0;
        for (int i = 0; i < ninst; ++i)
            if (h_is[i].mode == M_RUN) nw += h_is[i].nw_count[h_is[i].cur];
        const long long want = (9 * nw + 9 * nw / 4 + 255) / 256;
        int w = 64;
        while (w < want && w < 65536) w <<= 1;
        return w;
    }

    // Enqueue `k` trial steps of every running instance.
    int enqueue(int k) {
        BatchDev B = dev();
        bool pel = any_pellets();
        const int twgs = traction_workgroups();
        if (launch.use_graph && !profiling && !auditing) {
            if (graph_stale || graph_steps != k || graph_twgs != twgs || !graph) {
                if (graph) {
                    HIP_OK(hipGraphExecDestroy(graph));
                    graph = nullptr;
                }
                hipGraph_t g = nullptr;
                HIP_OK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
                for (int j = 0; j < k; ++j) {
                    BatchDev Bj = B;
                    Bj.fuse_control = j + 1 < k;
                    launch_step(stream, Bj, twgs, pel, j == 0);
                }
                HIP_OK(hipStreamEndCapture(stream, &g));
                hipError_t e = hipGraphInstantiate(&graph, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                HIP_OK(e);
                graph_steps = k;
                graph_twgs = twgs;
                graph_stale = false;
            }
            HIP_OK(hipGraphLaunch(graph, stream));
        } else {
            if (profiling) {
                // Start stamps all ones (atomic minima), end stamps zero (atomic maxima).
                for (int j = 0; j < k; ++j)
                    for (int q = 0; q < 4; ++q)
                        for (int w = 0; w < kStampSlots; ++w)
                            h_ktime[kStampsPerStep * j + q * kStampSlots + w] = q % 2 == 0 ? ~0ull : 0ull;
                HIP_OK(hipMemcpyAsync(d_ktime, h_ktime, sizeof(unsigned long long) * kStampsPerStep * k,
                                      hipMemcpyHostToDevice, stream));
            }
            for (int j = 0; j < k; ++j) {
                profiled[j] = profiling && profile_seq++ % (uint64_t)profile_every == 0;
                BatchDev Bj = B;
                Bj.ktime = profiled[j] ? d_ktime + kStampsPerStep * j : nullptr;
                Bj.fuse_control = j + 1 < k;
                launch_step(stream, Bj, twgs, pel, j == 0);
            }
        }
        HIP_OK(hipGetLastError());
        return MPM_OK;
    }
    // Enqueue the copy of the last enqueue's stamps (read by collect_profile
    // after the stream has been synchronised).
    int fetch_profile(int k) {
        if (profiling)
            HIP_OK(hipMemcpyAsync(h_ktime, d_ktime, sizeof(unsigned long long) * kStampsPerStep * k,
                                  hipMemcpyDeviceToHost, stream));
        return MPM_OK;
    }
    // Kernel times of the profiled steps of the last enqueue, from the stamps
    // fetch_profile copied (stream synchronised since).  A launch whose
    // workgroups all found their instance idle leaves no end stamp and is not
    // counted.
    int collect_profile(int k) {
        if (!profiling) return MPM_OK;
        const double ms_per_tick = 1.0 / clock_khz;
        for (int j = 0; j < k; ++j) {
            if (!profiled[j]) continue;
            profiled[j] = false;
            // The earliest start and the latest end over the partial slots.
            unsigned long long t[4] = {~0ull, 0ull, ~0ull, 0ull};
            for (int q = 0; q < 4; ++q)
                for (int w = 0; w < kStampSlots; ++w) {
                    const unsigned long long v = h_ktime[kStampsPerStep * j + q * kStampSlots + w];
                    t[q] = q % 2 == 0 ? std::min(t[q], v) : std::max(t[q], v);
                }
            if (t[1] <= t[0] || t[3] <= t[2]) continue;
            p2g_ms += (double)(t[1] - t[0]) * ms_per_tick;
            g2p_ms += (double)(t[3] - t[2]) * ms_per_tick;
            profiled_launches += 1;
        }
        return MPM_OK;
    }

    std::string abort_message(int i) const {
        const InstState& s = h_is[i];
        char buf[320];
        if (s.err == E_MIN_AFTER_REJECT)
            std::snprintf(buf, sizeof(buf),
                          "solver aborted at t='%.17g' s: step '%.17g' s below minimum after '%llu' rejections; "
                          "last failure code %d",
                          s.time, s.dt, (unsigned long long)s.rejected_steps, s.last_failure);
        else
            std::snprintf(buf, sizeof(buf), "solver aborted at t='%.17g' s: step '%.17g' s below minimum", s.time,
                          s.dt);
        std::string msg(buf);
        if (ninst > 1) msg = "instance " + std::to_string(i) + ": " + msg;
        return msg;
    }

    // Run the instances in M_RUN until none is left running, `k_sync` trial steps
    // per host synchronisation, at most `max_rounds` synchronisations (< 0: no
    // limit).  status/done_ms per instance (may be null).
    int run(int k_sync, int* status, double* done_ms, int max_rounds = -1) {
        auto t0 = std::chrono::steady_clock::now();
        // Completion times come from the device clock: the run's start here, each
        // instance's last accepted step in its commit stamp.
        unsigned long long start_clock = 0;
        if (done_ms) {
            launch_clock(stream, d_run_clock);
            HIP_OK(hipGetLastError());
            HIP_OK(hipMemcpyAsync(&start_clock, d_run_clock, sizeof(start_clock), hipMemcpyDeviceToHost, stream));
            if (int rc_ = sync()) return rc_;
        }
        std::vector<char> recorded(ninst, 0);
        int rounds = 0;
        bool first_error = true;
        for (int i = 0; i < ninst; ++i)
            if (h_is[i].mode != M_RUN) recorded[i] = 1;
        while (max_rounds < 0 || rounds < max_rounds) {
            bool running = false;
            int most = 0;  // largest step budget left among the running instances
            for (int i = 0; i < ninst; ++i)
                if (h_is[i].mode == M_RUN) {
                    running = true;
                    most = std::max(most, h_is[i].trial_dt > 0.0 ? 1 : h_is[i].budget);
                }
            if (!running) break;
            // No launches past the last budgeted step (they would find nothing to do).
            const int k = std::max(1, std::min(k_sync, most));
            const auto t_round = std::chrono::steady_clock::now();
            if (int rc = enqueue(k)) return rc;
            // One host round trip per round: the stamps and the mirror together.
            if (int rc = fetch_profile(k)) return rc;
            if (int rc = pull()) return rc;
            if (int rc = collect_profile(k)) return rc;
            const double step_us =
                std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t_round).count() / k;
            if (best_step_us <= 0.0 || step_us < best_step_us) best_step_us = step_us;
            drift_us += (step_us - best_step_us) * k;
            since_sort += k;
            rounds += 1;
            double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            bool changed = false;
            for (int i = 0; i < ninst; ++i) {
                InstState& s = h_is[i];
                if (s.mode == M_PAUSE && (s.pause & P_ESCAPED)) {
                    if (int rc = remove_escaped(i)) return rc;
                    s.pause &= ~P_ESCAPED;
                    s.mode = (s.pause & P_BUDGET) ? M_PAUSE : M_RUN;
                    changed = true;
                }
                // The transitions the control kernel would make at the next trial step.
                if (s.mode == M_RUN && s.trial_dt == 0.0) {
                    if (!(s.time < s.target)) {
                        s.time = s.target;
                        s.mode = M_DONE;
                        changed = true;
                    } else if (s.budget <= 0) {
                        s.mode = M_PAUSE;
                        s.pause |= P_BUDGET;
                        changed = true;
                    }
                }
                if (recorded[i] || s.mode == M_RUN) continue;
                recorded[i] = 1;
                int st = s.mode == M_DONE ? 0 : (s.mode == M_ERROR ? MPM_ERR_MIN_STEP : 1);
                if (st == MPM_ERR_MIN_STEP && first_error) {
                    err = abort_message(i);
                    first_error = false;
                }
                if (status) status[i] = st;
                if (done_ms)
                    done_ms[i] = s.commit_clock > start_clock && clock_khz > 0.0
                                     ? (double)(s.commit_clock - start_clock) / clock_khz
                                     : ms;
            }
            const double sort_cost =
                sort_cost_us > 0.0 ? sort_cost_us : kSortCostUs0 + kSortCostUsPerParticle * (double)cap;
            if (dirty || since_sort >= kSortMaxInterval || (since_sort >= kSortMinInterval && drift_us >= sort_cost)) {
                if (changed)
                    if (int rc = push()) return rc;
                if (int rc = sort_all()) return rc;
            } else if (changed) {
                if (int rc = push()) return rc;
            }
        }
        return MPM_OK;
    }

    // Host copy of Simulation::stability_limit (solver.rs:477-498).
    double stability_limit(int i) {
        const Instance& in = inst[i];
        if (pull()) return std::numeric_limits<double>::quiet_NaN();
        std::vector<PelletState> pel;
        if (download_pellets(i, pel)) return std::numeric_limits<double>::quiet_NaN();
        double limit = in.cfg.cfl * in.dx / (in.ip.wave + h_is[i].max_speed);
        if (in.ip.diff > 0.0) limit = std::min(limit, 0.25 * in.dx * in.dx / in.ip.diff);
        for (const PelletState& p : pel) {
            limit = std::min(limit, 0.2 * std::sqrt(p.mass / in.cfg.contact_stiffness));
            double s = std::sqrt(p.vel[0] * p.vel[0] + p.vel[1] * p.vel[1] + p.vel[2] * p.vel[2]);
            if (s > 0.0) limit = std::min(limit, in.cfg.cfl * in.dx / s);
        }
        return limit;
    }

    // Device copy of every mutable buffer, for the autotuner's trial runs.
    struct Snapshot {
        std::vector<std::pair<void*, size_t>> bufs;  // (backup, bytes), in the order of targets()
        std::vector<InstState> mirror;
        int sh = 0;
    };
    std::vector<std::pair<void*, size_t>> targets() {
        std::vector<std::pair<void*, size_t>> t;
        for (int b = 0; b < 2; ++b) {
            t.push_back({slab[b], sizeof(float) * NPF * cap});
            t.push_back({shared[b], sizeof(float) * NSF * cap});
            t.push_back({ids[b], sizeof(uint32_t) * cap});
            t.push_back({refkey[b], sizeof(uint64_t) * cap});
            t.push_back({gmem[b], sizeof(long long) * kAuditFields * na});
            t.push_back({gflag[b], sizeof(int) * (size_t)total_blocks});
        }
        t.push_back({d_pel, sizeof(PelletState) * (size_t)std::max(total_pel, 1)});
        return t;
    }
    int snapshot(Snapshot& snap) {
        snap.mirror.assign(h_is, h_is + ninst);
        snap.sh = sh;
        for (auto& t : targets()) {
            void* b = nullptr;
            HIP_OK(hipMalloc(&b, t.second));
            snap.bufs.push_back({b, t.second});
            HIP_OK(hipMemcpyAsync(b, t.first, t.second, hipMemcpyDeviceToDevice, stream));
        }
        if (int rc_ = sync()) return rc_;
        return MPM_OK;
    }
    int restore(const Snapshot& snap) {
        auto t = targets();
        for (size_t k = 0; k < t.size(); ++k)
            HIP_OK(hipMemcpyAsync(t[k].first, snap.bufs[k].first, t[k].second, hipMemcpyDeviceToDevice, stream));
        std::copy(snap.mirror.begin(), snap.mirror.end(), h_is);
        sh = snap.sh;
        dirty = true;
        return push();
    }
    static void release(Snapshot& snap) {
        for (auto& b : snap.bufs) (void)hipFree(b.first);
        snap.bufs.clear();
    }

    // Time max(steps, kMinTuneSteps) trial steps of every instance under each
    // candidate launch shape, from the same state, and keep the fastest.  Two
    // stages:
    //  1. the kernel shapes -- cell blocks per work item (tile size and workgroup
    //     count), the G2P register budget (workgroups per CU) and the deposit
    //     workgroup size -- by the device time of the deposit and G2P kernels,
    //     stamped on the device clock in stream mode (the kernels these choices
    //     change; no host noise);
    //  2. the synchronisation interval and graph replay with those kernels, by
    //     wall time, best of kRepeats (the first run of a graph candidate
    //     includes its instantiation).
    // The count spans two of the longest synchronisation intervals, so the
    // candidates differ in how often the host synchronises as they will in use.
    int autotune(int steps, mpm_launch* chosen) {
        if (int rc = pull()) return rc;
        if (int rc = prepare()) return rc;
        if (int rc = flush_refkeys()) return rc;
        Snapshot snap;
        int rc = snapshot(snap);
        // The caller's kernel timing, kept across the tuner's own.
        const bool prof = profiling;
        const int prof_every = profile_every;
        const uint64_t prof_seq = profile_seq, prof_launches = profiled_launches;
        const double prof_g2p = g2p_ms, prof_p2g = p2g_ms;
        constexpr int kMinTuneSteps = 64;
        constexpr int kRepeats = 3;
        const int timed = std::max(steps, kMinTuneSteps);
        // One timed run of `c` from the snapshot; wall ms in *wall, mean kernel
        // ms per step in *kern (profiled runs).
        auto trial = [&](const mpm_launch& c, bool prof_run, double* wall, double* kern) -> int {
            if (int r = restore(snap)) return r;
            launch = c;
            graph_stale = true;
            if (int r = prepare()) return r;
            for (int i = 0; i < ninst; ++i) {
                InstState& s = h_is[i];
                s.target = std::numeric_limits<double>::max();
                s.mode = s.n > 0 ? M_RUN : M_DONE;
                s.pause = 0;
                s.trial_dt = 0.0;
                s.budget = timed;
            }
            if (int r = push()) return r;
            profiling = prof_run;
            profile_every = 1;
            g2p_ms = p2g_ms = 0.0;
            profiled_launches = 0;
            auto t0 = std::chrono::steady_clock::now();
            int r = run(c.steps_per_sync, nullptr, nullptr);
            *wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            *kern = profiled_launches ? (g2p_ms + p2g_ms) / (double)profiled_launches
                                      : std::numeric_limits<double>::infinity();
            profiling = false;
            return r;
        };
        mpm_launch best = launch;
        {
            double best_kern = std::numeric_limits<double>::infinity();
            for (int m : {1, 2})
                for (int o : {2, 3})
                    for (int pb : {256, 512}) {
                        const mpm_launch c{32, 0, m, launch.traction_wgs, o, pb};
                        double wall = 0.0, kern = 0.0;
                        if ((rc = trial(c, true, &wall, &kern))) break;
                        if (kern < best_kern) {
                            best_kern = kern;
                            best = c;
                        }
                    }
        }
        if (!rc) {
            double best_ms = std::numeric_limits<double>::infinity();
            const mpm_launch kernels = best;
            for (int k : {8, 32, 64})
                for (int g : {0, 1}) {
                    mpm_launch c = kernels;
                    c.steps_per_sync = k;
                    c.use_graph = g;
                    double cand_ms = std::numeric_limits<double>::infinity();
                    for (int rep = 0; rep < kRepeats && !rc; ++rep) {
                        double wall = 0.0, kern = 0.0;
                        if ((rc = trial(c, false, &wall, &kern))) break;
                        cand_ms = std::min(cand_ms, wall);
                    }
                    if (rc) break;
                    if (cand_ms < best_ms) {
                        best_ms = cand_ms;
                        best = c;
                    }
                }
        }
        if (!rc) rc = restore(snap);
        release(snap);
        profiling = prof;
        profile_every = prof_every;
        profile_seq = prof_seq;
        profiled_launches = prof_launches;
        g2p_ms = prof_g2p;
        p2g_ms = prof_p2g;
        launch = best;
        graph_stale = true;
        dirty = true;
        if (!rc) rc = prepare();
        if (chosen) *chosen = best;
        return rc;
    }

    int set_state(int i, const mpm_particles* ps, const mpm_pellet* pellets, size_t n_pellets,
                  const mpm_ledger* ledger, const mpm_history* history);
    int get_state(int i, mpm_particles* ps, mpm_pellet* pellets, size_t pellet_capacity, mpm_ledger* ledger,
                  mpm_history* history);
    int energies(int i, double* ke, double* pe, double* el) {
        if (int rc = pull()) return rc;
        HostParticles h;
        if (int rc = download(i, h)) return rc;
        std::vector<PelletState> pel;
        if (int rc = download_pellets(i, pel)) return rc;
        energies_of(inst[i].mat, inst[i].cfg.gravity, h, inst[i].exact, pel, ke, pe, el);
        return MPM_OK;
    }
};

int mpm_engine::set_state(int i, const mpm_particles* ps, const mpm_pellet* pellets, size_t n_pellets,
                          const mpm_ledger* ledger, const mpm_history* history) {
    Instance& in = inst[i];
    if (!ps) return fail(MPM_ERR_INVALID, "particles are required");
    if (ps->n > (size_t)in.ip.pcap) return fail(MPM_ERR_CAPACITY, "more particles than the instance capacity");
    if (n_pellets > (size_t)in.pel_cap) return fail(MPM_ERR_CAPACITY, "more pellets than the instance capacity");
    if (n_pellets && !pellets) return fail(MPM_ERR_INVALID, "pellet array is NULL");
    size_t n = ps->n;
    if (n && (!ps->position || !ps->mass || !ps->volume0))
        return fail(MPM_ERR_INVALID, "position, mass and volume0 are required");
    if (int rc = pull()) return rc;
    HostParticles h;
    h.n = n;
    h.s.assign(NPF * n, 0.0f);
    h.sh.assign(NSF * n, 0.0f);
    h.id.resize(n);
    h.ref.resize(n);
    double max_speed = 0.0, mass = 0.0;
    for (size_t p = 0; p < n; ++p) {
        h.id[p] = ps->id ? ps->id[p] : (uint32_t)p;
        h.ref[p] = p;  // until the first 32-step sort the reference order is the given order
        for (int a = 0; a < 3; ++a) {
            h.s[(PX + a) * n + p] = (float)ps->position[3 * p + a];
            h.s[(VX + a) * n + p] = ps->velocity ? (float)ps->velocity[3 * p + a] : 0.0f;
        }
        double vs = 0.0;
        for (int a = 0; a < 3; ++a) vs += (double)h.s[(VX + a) * n + p] * (double)h.s[(VX + a) * n + p];
        max_speed = std::max(max_speed, std::sqrt(vs));
        for (int a = 0; a < 9; ++a) {
            h.s[(C00 + a) * n + p] = ps->affine ? (float)ps->affine[9 * p + a] : 0.0f;
            double F = ps->deformation ? ps->deformation[9 * p + a] : (a % 4 == 0 ? 1.0 : 0.0);
            h.s[(H00 + a) * n + p] = (float)(F - (a % 4 == 0 ? 1.0 : 0.0));
        }
        h.s[JM1 * n + p] = ps->volume_ratio ? (float)(ps->volume_ratio[p] - 1.0) : 0.0f;
        if (ps->kirchhoff) {
            const double* t = ps->kirchhoff + 9 * p;
            h.s[TXX * n + p] = (float)t[0];
            h.s[TYY * n + p] = (float)t[4];
            h.s[TZZ * n + p] = (float)t[8];
            h.s[TXY * n + p] = (float)(0.5 * (t[1] + t[3]));
            h.s[TXZ * n + p] = (float)(0.5 * (t[2] + t[6]));
            h.s[TYZ * n + p] = (float)(0.5 * (t[5] + t[7]));
        }
        float eps = ps->plastic_strain ? (float)ps->plastic_strain[p] : 0.0f;
        h.s[EPS * n + p] = eps;
        h.sh[SEPS * n + p] = eps;
        h.sh[SMASS * n + p] = (float)ps->mass[p];
        h.sh[SVOL0 * n + p] = (float)ps->volume0[p];
        mass += ps->mass[p];
    }
    in.exact.assign(n, ps->id, ps->mass, ps->volume0);
    // Pellets: the instance's slots, unused ones marked free.
    std::vector<PelletState> pel(in.pel_cap);
    double contact_dt = std::numeric_limits<double>::infinity();
    for (int k = 0; k < in.pel_cap; ++k) {
        PelletState& d = pel[k];
        std::memset(&d, 0, sizeof(d));
        d.inst = -1;
        if ((size_t)k >= n_pellets) continue;
        const mpm_pellet& src = pellets[k];
        d.inst = i;
        d.radius = src.radius;
        d.length = src.length;
        d.mass = src.mass;
        for (int a = 0; a < 3; ++a) {
            d.inertia[a] = src.inertia_body[a];
            d.pos[a] = src.position[a];
            d.vel[a] = src.velocity[a];
            d.L[a] = src.angular_momentum[a];
        }
        for (int a = 0; a < 4; ++a) d.q[a] = src.orientation[a];
        d.spheres = std::min(std::max(src.sphere_count, 0), kMaxSpheres);
        d.half = 0.0;
        for (int s = 0; s < d.spheres; ++s) {
            d.offsets[s] = src.sphere_offsets[s];
            d.half = std::max(d.half, std::fabs(d.offsets[s]));
        }
        contact_dt = std::min(contact_dt, 0.2 * std::sqrt(d.mass / in.cfg.contact_stiffness));
    }
    if (in.pel_cap) {
        HIP_OK(hipMemcpyAsync(d_pel + in.ip.pel_off, pel.data(), sizeof(PelletState) * in.pel_cap,
                              hipMemcpyHostToDevice, stream));
        if (int rc_ = sync()) return rc_;
    }
    pel.resize(n_pellets);
    in.ip.npel = (int)n_pellets;
    in.ip.particle_mass = mass;
    {
        float mmax = 0.0f;
        for (size_t p = 0; p < n; ++p) mmax = std::max(mmax, h.sh[SMASS * n + p]);
        in.ip.mass_max = mmax;
    }
    in.ip.pel_contact_dt = contact_dt;
    if (int rc = upload_params(i)) return rc;
    if (int rc = upload(i, h)) return rc;
    InstState& s = h_is[i];
    std::memset(&s, 0, sizeof(s));
    s.n = (int)n;
    s.mode = M_DONE;
    s.max_speed = max_speed;
    for (const PelletState& p : pel)
        s.pel_vmax = std::max(s.pel_vmax, std::sqrt(p.vel[0] * p.vel[0] + p.vel[1] * p.vel[1] + p.vel[2] * p.vel[2]));
    if (history) {
        s.time = history->time;
        s.step_count = history->step_count;
        s.limited_steps = history->limited_steps;
        s.rejected_steps = history->rejected_steps;
        s.dt_scale = history->dt_scale;
        s.min_dt_used = history->min_dt_used;
        s.max_dt_used = history->max_dt_used;
    } else {
        s.dt_scale = 1.0;
        s.min_dt_used = std::numeric_limits<double>::infinity();
        s.max_dt_used = 0.0;
    }
    s.target = s.time;
    if (ledger) {
        std::memcpy(s.ledger, ledger, sizeof(mpm_ledger));
    } else {
        // Simulation::new seeds the initial inventory from the state as given
        // (f64, in its order), solver.rs:357-390.
        double mom[3] = {0.0, 0.0, 0.0};
        if (ps->velocity)
            for (size_t p = 0; p < n; ++p)
                for (int a = 0; a < 3; ++a) mom[a] += ps->velocity[3 * p + a] * ps->mass[p];
        for (size_t k = 0; k < n_pellets; ++k)
            for (int a = 0; a < 3; ++a) mom[a] += pellets[k].velocity[a] * pellets[k].mass;
        s.ledger[0] = mass;
        for (int a = 0; a < 3; ++a) s.ledger[2 + a] = mom[a];
        double ke = 0, pe = 0, el = 0;
        energies_f64(in.mat, in.cfg.gravity, *ps, pellets, n_pellets, &ke, &pe, &el);
        s.ledger[5] = ke + pe + el;
    }
    in.grid_clear = false;
    dirty = true;
    graph_stale = true;
    return push();
}

int mpm_engine::get_state(int i, mpm_particles* ps, mpm_pellet* pellets, size_t pellet_capacity,
                          mpm_ledger* ledger, mpm_history* history) {
    if (int rc = pull()) return rc;
    const Instance& in = inst[i];
    const InstState& st = h_is[i];
    if (ps) {
        if (ps->n < (size_t)st.n) return fail(MPM_ERR_CAPACITY, "particle arrays are smaller than the state");
        HostParticles h;
        if (int rc = download(i, h)) return rc;
        size_t n = h.n;
        ps->n = n;
        bool fresh = st.committed == 0;
        // Hand the particles back in the reference's order (sorted by the key of the
        // last 32-step sort, solver.rs:501-517), whatever the device order is.
        std::vector<size_t> ord(n);
        for (size_t k = 0; k < n; ++k) ord[k] = k;
        std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return h.ref[a] < h.ref[b]; });
        for (size_t k = 0; k < n; ++k) {
            size_t p = ord[k];
            if (ps->id) ps->id[k] = h.id[p];
            for (int a = 0; a < 3; ++a) {
                if (ps->position) ps->position[3 * k + a] = h.s[(PX + a) * n + p];
                if (ps->velocity) ps->velocity[3 * k + a] = h.s[(VX + a) * n + p];
            }
            for (int a = 0; a < 9; ++a) {
                if (ps->affine) ps->affine[9 * k + a] = h.s[(C00 + a) * n + p];
                if (ps->deformation) {
                    double id_a = a % 4 == 0 ? 1.0 : 0.0;
                    // A committed liquid step resets F to I (solver.rs:965).
                    ps->deformation[9 * k + a] =
                        (in.mat.kind == 1 && !fresh) ? id_a : id_a + (double)h.s[(H00 + a) * n + p];
                }
            }
            if (ps->kirchhoff) {
                double* t = ps->kirchhoff + 9 * k;
                t[0] = h.s[TXX * n + p];
                t[4] = h.s[TYY * n + p];
                t[8] = h.s[TZZ * n + p];
                t[1] = t[3] = h.s[TXY * n + p];
                t[2] = t[6] = h.s[TXZ * n + p];
                t[5] = t[7] = h.s[TYZ * n + p];
            }
            // A committed paste step sets J = 1 (solver.rs:950).
            if (ps->volume_ratio)
                ps->volume_ratio[k] = (in.mat.kind == 0 && !fresh) ? 1.0 : 1.0 + (double)h.s[JM1 * n + p];
            if (ps->plastic_strain) ps->plastic_strain[k] = in.mat.kind == 0 ? h.s[EPS * n + p] : h.sh[SEPS * n + p];
            double m, v0;
            in.exact.get(h.id[p], h.sh[SMASS * n + p], h.sh[SVOL0 * n + p], m, v0);
            if (ps->mass) ps->mass[k] = m;
            if (ps->volume0) ps->volume0[k] = v0;
        }
    }
    if (pellets) {
        std::vector<PelletState> pel;
        if (int rc = download_pellets(i, pel)) return rc;
        for (size_t k = 0; k < pel.size() && k < pellet_capacity; ++k) {
            const PelletState& p = pel[k];
            mpm_pellet& d = pellets[k];
            d.radius = p.radius;
            d.length = p.length;
            d.mass = p.mass;
            for (int a = 0; a < 3; ++a) {
                d.inertia_body[a] = p.inertia[a];
                d.position[a] = p.pos[a];
                d.velocity[a] = p.vel[a];
                d.angular_momentum[a] = p.L[a];
            }
            for (int a = 0; a < 4; ++a) d.orientation[a] = p.q[a];
            d.sphere_count = p.spheres;
            for (int s = 0; s < p.spheres; ++s) d.sphere_offsets[s] = p.offsets[s];
        }
    }
    if (ledger) std::memcpy(ledger, st.ledger, sizeof(mpm_ledger));
    if (history)
        *history = mpm_history{st.time,     st.step_count,  st.limited_steps, st.rejected_steps,
                               st.dt_scale, st.min_dt_used, st.max_dt_used};
    return MPM_OK;
}

namespace {

void material_paste(double density, double young, double poisson, double ys, double k, double n, mpm_material* out) {
    double shear = young / (2.0 * (1.0 + poisson));
    double lame = young * poisson / ((1.0 + poisson) * (1.0 - 2.0 * poisson));
    *out = mpm_material{0, density, shear, lame, lame + 2.0 * shear / 3.0, ys, k, n, 1.0};
}

Params batch_params(const mpm_material* m) {
    Params P{};
    P.dx = 1.0f;
    P.inv_dx = 1.0f;
    P.kind = m->kind;
    P.G = (float)m->shear_modulus;
    P.lam = (float)m->lame_lambda;
    P.ys = (float)m->yield_stress;
    P.K = (float)m->consistency;
    P.inv_n = (float)(1.0 / m->flow_index);
    P.bulk = (float)m->bulk_modulus;
    P.gamma = (float)m->tait_gamma;
    P.visc = (float)m->consistency;
    return P;
}

template <class T>
struct DevBuf {
    T* p = nullptr;
    hipError_t alloc(size_t n) { return hipMalloc(&p, sizeof(T) * std::max<size_t>(n, 1)); }
    ~DevBuf() { (void)hipFree(p); }
};

bool valid_instance(const mpm_engine* e, size_t i) { return e && i < (size_t)e->ninst; }

}  // namespace

extern "C" {

int mpm_abi_version(void) { return MPM_ABI_VERSION; }

void mpm_material_paste(double density, double young, double poisson, double yield_stress, double consistency,
                        double flow_index, mpm_material* out) {
    material_paste(density, young, poisson, yield_stress, consistency, flow_index, out);
}

void mpm_material_paste_shear(double density, double young, double poisson, double shear_yield,
                              double shear_consistency, double flow_index, mpm_material* out) {
    material_paste(density, young, poisson, std::sqrt(3.0) * shear_yield,
                   std::pow(3.0, (1.0 + flow_index) / 2.0) * shear_consistency, flow_index, out);
}

void mpm_material_liquid(double density, double young, double poisson, double viscosity, mpm_material* out) {
    *out = mpm_material{1, density, 0.0, 0.0, young / (3.0 * (1.0 - 2.0 * poisson)), 0.0, viscosity, 1.0, 1.0};
}

// FixtureSpec::particle_count, fixtures.rs:120-133.
size_t mpm_fixture_particle_count(const mpm_fixture* f) {
    double sub = f->grid_spacing / 2.0;
    size_t c = 1;
    for (int a = 0; a < 3; ++a) c *= (size_t)std::max(std::round(f->initial_size[a] / sub), 1.0);
    return c;
}

// FixtureSpec::place_particles + hydrostatic_state, fixtures.rs:272-349, with
// SplitMix64 (rng.rs:35-47) drawing three jitter samples per particle.
int mpm_fixture_place(const mpm_fixture* f, mpm_particles* out) {
    if (!f || !out) return MPM_ERR_INVALID;
    double sub = f->grid_spacing / 2.0;
    size_t counts[3];
    for (int a = 0; a < 3; ++a) counts[a] = (size_t)std::max(std::round(f->initial_size[a] / sub), 1.0);
    size_t total = counts[0] * counts[1] * counts[2];
    if (out->n < total) return MPM_ERR_CAPACITY;
    out->n = total;
    double origin[3] = {0.0, 0.0, 0.0};
    if (f->fixture == 0 || f->fixture == 3) {
        origin[0] = 0.5 * (f->domain[0] - f->initial_size[0]);
        origin[1] = 0.5 * (f->domain[1] - f->initial_size[1]);
    }
    double jitter = f->fixture == 1 ? 0.0 : f->jitter * sub;
    uint64_t state = f->seed;
    auto next_unit = [&]() {
        state += 0x9E3779B97F4A7C15ull;
        uint64_t z = state;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / (double)(1ull << 53));
    };
    const mpm_material& m = f->material;
    size_t p = 0;
    for (size_t i = 0; i < counts[0]; ++i)
        for (size_t j = 0; j < counts[1]; ++j)
            for (size_t k = 0; k < counts[2]; ++k, ++p) {
                double pos[3] = {origin[0] + ((double)i + 0.5) * sub, origin[1] + ((double)j + 0.5) * sub,
                                 origin[2] + ((double)k + 0.5) * sub};
                if (jitter > 0.0)
                    for (int a = 0; a < 3; ++a) pos[a] += (next_unit() - 0.5) * 2.0 * jitter;
                double ratio = 1.0;
                double F[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, tau[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                if (f->fixture == 1) {
                    double depth = f->initial_size[2] - pos[2];
                    if (m.kind == 1) {
                        ratio = std::exp(-m.density * f->gravity_norm * depth / m.bulk_modulus);
                        double pressure = m.tait_gamma == 1.0 ? m.bulk_modulus * (1.0 / ratio - 1.0)
                                                              : m.bulk_modulus / m.tait_gamma *
                                                                    (std::pow(ratio, -m.tait_gamma) - 1.0);
                        tau[0] = tau[4] = tau[8] = -pressure * ratio;
                    } else {
                        double modulus = m.lame_lambda + 2.0 * m.shear_modulus;
                        double strain = -m.density * f->gravity_norm * depth / modulus;
                        double stretch = std::exp(strain);
                        F[8] = stretch;
                        tau[0] = tau[4] = m.lame_lambda * strain;
                        tau[8] = modulus * strain;
                        ratio = stretch;
                    }
                }
                double v0 = sub * sub * sub / ratio;
                if (out->id) out->id[p] = (uint32_t)p;
                for (int a = 0; a < 3; ++a) {
                    if (out->position) out->position[3 * p + a] = pos[a];
                    if (out->velocity) out->velocity[3 * p + a] = f->initial_velocity[a];
                }
                for (int a = 0; a < 9; ++a) {
                    if (out->affine) out->affine[9 * p + a] = 0.0;
                    if (out->deformation) out->deformation[9 * p + a] = F[a];
                    if (out->kirchhoff) out->kirchhoff[9 * p + a] = tau[a];
                }
                if (out->volume_ratio) out->volume_ratio[p] = ratio;
                if (out->plastic_strain) out->plastic_strain[p] = 0.0;
                if (out->mass) out->mass[p] = m.density * v0;
                if (out->volume0) out->volume0[p] = v0;
            }
    return MPM_OK;
}

int mpm_batch_create(size_t n_instances, const mpm_material* materials, const mpm_config* solver_configs,
                     const mpm_layout* layouts, const size_t* capacities, const size_t* max_pellets,
                     mpm_engine** out) {
    if (!out) return MPM_ERR_INVALID;
    *out = nullptr;
    if (n_instances == 0 || n_instances > 65536 || !materials || !solver_configs || !layouts) return MPM_ERR_INVALID;
    std::vector<Instance> inst(n_instances);
    for (size_t i = 0; i < n_instances; ++i) {
        if (max_pellets && max_pellets[i] > 4096) return MPM_ERR_CAPACITY;
        int rc = grid_cells(layouts[i], inst[i].cells);
        if (rc) return rc;
        inst[i].mat = materials[i];
        inst[i].cfg = solver_configs[i];
        inst[i].dx = layouts[i].spacing;
        inst[i].fill_params();
    }
    int devices = 0;
    if (hipGetDeviceCount(&devices) != hipSuccess || devices < 1) return MPM_ERR_DEVICE;
    mpm_engine* e = new mpm_engine();
    e->ninst = (int)n_instances;
    e->inst = std::move(inst);
    int rc = e->alloc(capacities, max_pellets);
    if (rc) {
        delete e;
        return rc;
    }
    *out = e;
    return MPM_OK;
}

int mpm_create(const mpm_material* material, const mpm_config* config, const mpm_layout* layout, size_t capacity,
               size_t max_pellets, mpm_engine** out) {
    if (!material || !config || !layout || !out) return MPM_ERR_INVALID;
    return mpm_batch_create(1, material, config, layout, &capacity, &max_pellets, out);
}

void mpm_destroy(mpm_engine* e) { delete e; }

const char* mpm_last_error(const mpm_engine* e) { return e ? e->err.c_str() : "null engine"; }

size_t mpm_instance_count(const mpm_engine* e) { return e ? (size_t)e->ninst : 0; }

size_t mpm_instance_particle_count(const mpm_engine* e, size_t i) {
    return valid_instance(e, i) ? (size_t)e->h_is[i].n : 0;
}

size_t mpm_particle_count(const mpm_engine* e) { return mpm_instance_particle_count(e, 0); }

size_t mpm_pellet_count(const mpm_engine* e) { return e ? (size_t)e->inst[0].ip.npel : 0; }

int mpm_set_config(mpm_engine* e, const mpm_config* c) {
    if (!e || !c) return MPM_ERR_INVALID;
    Instance& in = e->inst[0];
    in.cfg = *c;
    in.fill_params();
    return e->upload_params(0);
}

int mpm_instance_set_state(mpm_engine* e, size_t i, const mpm_particles* ps, const mpm_pellet* pellets,
                           size_t n_pellets, const mpm_ledger* ledger, const mpm_history* history) {
    if (!valid_instance(e, i)) return MPM_ERR_INVALID;
    return e->set_state((int)i, ps, pellets, n_pellets, ledger, history);
}

int mpm_set_state(mpm_engine* e, const mpm_particles* ps, const mpm_pellet* pellets, size_t n_pellets,
                  const mpm_ledger* ledger, const mpm_history* history) {
    return mpm_instance_set_state(e, 0, ps, pellets, n_pellets, ledger, history);
}

int mpm_instance_get_state(mpm_engine* e, size_t i, mpm_particles* ps, mpm_pellet* pellets, size_t pellet_capacity,
                           mpm_ledger* ledger, mpm_history* history) {
    if (!valid_instance(e, i)) return MPM_ERR_INVALID;
    return e->get_state((int)i, ps, pellets, pellet_capacity, ledger, history);
}

int mpm_get_state(mpm_engine* e, mpm_particles* ps, mpm_pellet* pellets, size_t pellet_capacity, mpm_ledger* ledger,
                  mpm_history* history) {
    return mpm_instance_get_state(e, 0, ps, pellets, pellet_capacity, ledger, history);
}

// Simulation::try_step, solver.rs:535-611: one bare trial step of instance 0.
int mpm_try_step(mpm_engine* e, double dt, mpm_step_report* report) {
    if (!e) return MPM_ERR_INVALID;
    if (report) *report = mpm_step_report{0, 0, 0.0, dt};
    if (!(std::isfinite(dt) && dt > 0.0)) return e->fail(MPM_ERR_INVALID, "dt must be finite and positive");
    if (int rc = e->pull()) return rc;
    if (int rc = e->prepare()) return rc;
    for (int i = 0; i < e->ninst; ++i) {
        InstState& s = e->h_is[i];
        s.mode = i == 0 ? M_RUN : M_DONE;
        s.pause = 0;
        s.trial_dt = i == 0 ? dt : 0.0;
    }
    if (int rc = e->push()) return rc;
    if (int rc = e->enqueue(1)) return rc;
    if (int rc = e->fetch_profile(1)) return rc;
    if (int rc = e->pull()) return rc;
    if (int rc = e->collect_profile(1)) return rc;
    InstState& s = e->h_is[0];
    int code = s.last_failure;
    if (report) {
        report->accepted = code == 0;
        report->failure = code;
        report->max_cells = (double)s.st.max_disp;
    }
    if (s.mode == M_PAUSE && (s.pause & P_ESCAPED)) {
        if (int rc = e->remove_escaped(0)) return rc;
    }
    s.mode = M_DONE;
    s.pause = 0;
    s.trial_dt = 0.0;
    if (int rc = e->push()) return rc;
    if (e->dirty)
        if (int rc = e->sort_all()) return rc;
    return code;
}

}  // extern "C"

// ------------------------------------------------------------ validation and the energy audit

namespace {

// A state read back in the reference's particle order (f64), for host-side checks.
struct HostState {
    std::vector<uint32_t> id;
    std::vector<double> pos, vel, aff, def, ratio, tau, eps, mass, vol0;
    std::vector<mpm_pellet> pellets;
    mpm_ledger ledger{};
    mpm_history history{};
    size_t n() const { return id.size(); }
    mpm_particles view() {
        return mpm_particles{id.size(), id.data(), pos.data(), vel.data(), aff.data(), def.data(), ratio.data(),
                             tau.data(), eps.data(), mass.data(), vol0.data()};
    }
};

int read_state(mpm_engine* e, HostState& h) {
    size_t n = mpm_particle_count(e);
    h.id.resize(n);
    for (auto* v : {&h.pos, &h.vel}) v->resize(3 * n);
    for (auto* v : {&h.aff, &h.def, &h.tau}) v->resize(9 * n);
    for (auto* v : {&h.ratio, &h.eps, &h.mass, &h.vol0}) v->resize(n);
    h.pellets.resize(mpm_pellet_count(e));
    mpm_particles ps = h.view();
    return mpm_get_state(e, &ps, h.pellets.data(), h.pellets.size(), &h.ledger, &h.history);
}

// Simulation::validate, solver.rs:1106-1140, with ParticleSet::validate
// (particles.rs:159-211) and Ledger::validate (solver.rs:182-243).  Returns
// the reason, or an empty string.
std::string validate_state(const HostState& h, double dx, const int* cells) {
    std::vector<uint32_t> ids(h.id);
    std::sort(ids.begin(), ids.end());
    for (size_t k = 1; k < ids.size(); ++k)
        if (ids[k] == ids[k - 1]) return "particle ids are not unique";
    for (size_t p = 0; p < h.n(); ++p) {
        bool ok = std::isfinite(h.ratio[p]) && std::isfinite(h.eps[p]) && std::isfinite(h.mass[p]) &&
                  std::isfinite(h.vol0[p]);
        for (int a = 0; a < 3; ++a) ok = ok && std::isfinite(h.pos[3 * p + a]) && std::isfinite(h.vel[3 * p + a]);
        for (int a = 0; a < 9; ++a)
            ok = ok && std::isfinite(h.aff[9 * p + a]) && std::isfinite(h.def[9 * p + a]) && std::isfinite(h.tau[9 * p + a]);
        std::string who = "particle id '" + std::to_string(h.id[p]) + "'";
        if (!ok) return who + " has non-finite state";
        if (h.mass[p] <= 0.0 || h.vol0[p] <= 0.0) return who + " has non-positive mass or volume";
        if (h.ratio[p] <= 0.0 || h.eps[p] < 0.0) return who + " has invalid ratio/plastic state";
    }
    const mpm_ledger& L = h.ledger;
    const double* flat = reinterpret_cast<const double*>(&L);
    for (int k = 0; k < kLedger; ++k)
        if (!std::isfinite(flat[k])) return "ledger is nonfinite";
    if (L.initial_mass < 0.0 || L.outflow_mass < 0.0 || L.plastic_dissipation < 0.0 || L.wall_friction_dissipation < 0.0)
        return "ledger has a negative mass or dissipation";
    if (L.wall_normal_projection_energy > 0.0) return "ledger `wall_normal_projection_energy` must be non-positive";
    if (!std::isfinite(h.history.time) || h.history.time < 0.0) return "time invalid";
    if (!(std::isfinite(h.history.dt_scale) && h.history.dt_scale > 0.0 && h.history.dt_scale <= 1.0))
        return "dt_scale invalid";
    for (const mpm_pellet& pl : h.pellets) {
        bool ok = true;
        for (int a = 0; a < 3; ++a)
            ok = ok && std::isfinite(pl.position[a]) && std::isfinite(pl.velocity[a]) && std::isfinite(pl.angular_momentum[a]);
        for (int a = 0; a < 4; ++a) ok = ok && std::isfinite(pl.orientation[a]);
        if (!ok) return "pellet state non-finite";
    }
    for (size_t p = 0; p < h.n(); ++p)
        for (int a = 0; a < 3; ++a) {
            double padded = std::floor(h.pos[3 * p + a] / dx - 0.5) + 1.0;
            if (std::isnan(padded) || padded < 0.0 || padded + 2.0 >= (double)(cells[a] + 3))
                return "particle id '" + std::to_string(h.id[p]) + "' lies outside the grid";
        }
    return "";
}

// EnergyState::measure, audit.rs:37-54: translation, affine norm h^2/8 sum m |C|^2,
// elastic, potential (material and pellets), pellet kinetic energy.
void measure_energy(HostState& h, const mpm_material& mat, const double* g, double dx, double* e) {
    double translation = 0.0, affine = 0.0;
    for (size_t p = 0; p < h.n(); ++p) {
        const double* v = &h.vel[3 * p];
        translation += 0.5 * h.mass[p] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    }
    for (size_t p = 0; p < h.n(); ++p) {
        const double* c = &h.aff[9 * p];
        double n2 = 0.0;  // column by column, as nalgebra stores the matrix
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) n2 += c[3 * i + j] * c[3 * i + j];
        affine += h.mass[p] * n2;
    }
    mpm_particles ps = h.view();
    double ke = 0.0, pe = 0.0, el = 0.0;
    energies_f64(mat, g, ps, h.pellets.data(), h.pellets.size(), &ke, &pe, &el);
    double pellet = 0.0;
    for (const mpm_pellet& pl : h.pellets) pellet += mpm_host::pellet_kinetic(pl);
    e[0] = translation;
    e[1] = dx * dx / 8.0 * affine;
    e[2] = el;
    e[3] = pe;
    e[4] = pellet;
}

const char* failure_name(int code) {
    switch (code) {
        case MPM_FAIL_INVERTED: return "constitutive update failed: inverted deformation";
        case MPM_FAIL_EIGEN: return "constitutive update failed: eigen-decomposition did not converge";
        case MPM_FAIL_STRETCH: return "constitutive update failed: degenerate stretch";
        case MPM_FAIL_RETURN: return "constitutive update failed: return map did not converge";
        case MPM_FAIL_NONFINITE: return "constitutive update failed: non-finite quantity";
        case MPM_FAIL_NONFINITE_GRID: return "non-finite grid velocity";
        case MPM_FAIL_DISPLACEMENT: return "particle moved more than one cell in one step";
        case MPM_FAIL_PELLET: return "pellet state became non-finite";
        case MPM_FAIL_TRACTION: return "wall traction kinetic-energy jump became non-finite";
        default: return "step failed";
    }
}

bool all_finite(const double* v, int n) {
    for (int k = 0; k < n; ++k)
        if (!std::isfinite(v[k])) return false;
    return true;
}

}  // namespace

extern "C" {

int mpm_validate(mpm_engine* e) {
    if (!e) return MPM_ERR_INVALID;
    if (e->ninst != 1) return e->fail(MPM_ERR_INVALID, "validation covers a single-instance engine");
    HostState h;
    if (int rc = read_state(e, h)) return rc;
    std::string bad = validate_state(h, e->inst[0].dx, e->inst[0].cells);
    if (!bad.empty()) return e->fail(MPM_ERR_INVALID, "invalid solver state: " + bad);
    return MPM_OK;
}

// Simulation::audit_step, audit.rs:275-345: one trial of the production
// kernels from a device snapshot, with the grid stage energies reduced on the
// device, and the snapshot restored afterwards.  The two coupling stages follow
// from the trial's coupling histories (the kinetic-energy changes of the
// constrained nodes and of the pellets, solver.rs:568-585).
int mpm_audit_step(mpm_engine* e, double dt, mpm_energy_audit* out) {
    if (!e || !out) return MPM_ERR_INVALID;
    if (e->ninst != 1) return e->fail(MPM_ERR_INVALID, "the energy audit covers a single-instance engine");
    HostState before;
    if (int rc = read_state(e, before)) return rc;
    const Instance& in = e->inst[0];
    std::string bad = validate_state(before, in.dx, in.cells);
    if (!bad.empty()) return e->fail(MPM_ERR_INVALID, "invalid solver state: " + bad);
    const mpm_config& c = in.cfg;
    const double stability = e->stability_limit(0);
    if (!all_finite(c.gravity, 3) || !std::isfinite(c.cfl) || c.cfl <= 0.0 || !std::isfinite(c.wall_friction) ||
        c.wall_friction < 0.0 || !std::isfinite(c.min_dt) || c.min_dt <= 0.0 || !std::isfinite(c.max_dt) ||
        c.max_dt < c.min_dt || !std::isfinite(stability) || stability <= 0.0)
        return e->fail(MPM_ERR_INVALID, "invalid solver state: invalid audit step configuration or stability limit");
    const double limit = std::min(before.history.dt_scale * stability, c.max_dt);
    if (!std::isfinite(dt) || dt <= 0.0 || dt < c.min_dt || !std::isfinite(limit) || dt > limit) {
        char buf[160];
        std::snprintf(buf, sizeof(buf), "invalid solver state: audit `dt`='%.17g' outside finite accepted range ending at '%.17g'",
                      dt, limit);
        return e->fail(MPM_ERR_INVALID, buf);
    }
    mpm_energy_audit a{};
    a.time = before.history.time;
    a.dt = dt;
    measure_energy(before, in.mat, c.gravity, in.dx, a.before);
    if (!all_finite(a.before, 5)) return e->fail(MPM_ERR_INVALID, "invalid solver state: nonfinite initial audit energy");
    if (!e->apart) {
        if (int rc = e->dalloc(&e->apart, (size_t)kAuditVals * e->ngwg)) return rc;
        if (int rc = e->dalloc(&e->d_audit, (size_t)kAuditOut * e->ninst)) return rc;
    }
    mpm_engine::Snapshot snap;
    if (int rc = e->snapshot(snap)) {
        mpm_engine::release(snap);
        return rc;
    }
    const InstParams saved = e->inst[0].ip;
    e->auditing = true;
    mpm_step_report report{};
    const int code = mpm_try_step(e, dt, &report);
    e->auditing = false;
    double dev[kAuditOut] = {0, 0, 0, 0, 0, 0};
    HostState after;
    int rc_read = 0;
    if (code == 0) {
        if (hipMemcpy(dev, e->d_audit, sizeof(dev), hipMemcpyDeviceToHost) != hipSuccess) rc_read = MPM_ERR_DEVICE;
        if (!rc_read) rc_read = read_state(e, after);
    }
    e->inst[0].ip = saved;
    int rc_restore = e->upload_params(0);
    if (!rc_restore) rc_restore = e->restore(snap);
    mpm_engine::release(snap);
    if (rc_restore) return rc_restore;
    if (code < 0) return code;
    if (code > 0) return e->fail(MPM_ERR_INVALID, std::string("invalid solver state: audit trial rejected: ") + failure_name(code));
    if (rc_read) return e->fail(rc_read, "device failure reading the audit trial");
    if (after.n() != before.n())
        return e->fail(MPM_ERR_INVALID, "invalid solver state: a particle lies outside the grid after the audit trial");
    bad = validate_state(after, in.dx, in.cells);
    if (!bad.empty()) return e->fail(MPM_ERR_INVALID, "invalid solver state: " + bad);
    measure_energy(after, in.mat, c.gravity, in.dx, a.after);
    const mpm_ledger &B = before.ledger, &A = after.ledger;
    a.increments[0] = A.plastic_dissipation - B.plastic_dissipation;
    a.increments[1] = A.wall_friction_dissipation - B.wall_friction_dissipation;
    a.increments[2] = A.wall_normal_projection_energy - B.wall_normal_projection_energy;
    a.increments[3] = A.wall_normal_traction_energy - B.wall_normal_traction_energy;
    a.increments[4] = A.coupling_grid_energy - B.coupling_grid_energy;
    a.increments[5] = A.coupling_pellet_energy - B.coupling_pellet_energy;
    a.increments[6] = A.pellet_wall_work - B.pellet_wall_work;
    const bool coupled = !before.pellets.empty();
    double* st = a.stages;
    for (int k = 0; k < kAuditVals; ++k) st[k] = dev[k];
    st[5] = st[4] + a.increments[4];
    st[6] = coupled ? a.before[4] + a.increments[5] : 0.0;
    st[7] = dev[kAuditVals];
    // EnergyTerms::measure, audit.rs:193-232.
    const double* be = a.before;
    const double* af = a.after;
    const double p2g = be[0] + be[1] - st[0];
    const double g2p = st[5] - af[0] - af[1];
    const double stress = st[1] - st[0] + af[2] - be[2] + st[7];
    const double gravity = st[3] - st[2] + af[3] - be[3];
    const double integrator = af[4] - st[6];
    const double traction = st[2] - st[1];
    const double walls = st[4] - st[3];
    const double coupling = st[5] - st[4] + st[6] - be[4];
    const double change = (af[0] + af[1] + af[2] + af[3] + af[4]) - (be[0] + be[1] + be[2] + be[3] + be[4]);
    const double sum = -p2g - g2p + stress - st[7] + gravity + integrator + traction + walls + coupling;
    const double terms[11] = {p2g, g2p, stress, gravity, integrator, traction, walls, coupling, change, sum, change - sum};
    std::copy(terms, terms + 11, a.terms);
    a.coupled = coupled ? 1 : 0;
    if (!all_finite(a.after, 5) || !all_finite(a.stages, 8) || !all_finite(a.increments, 7) || !all_finite(a.terms, 11))
        return e->fail(MPM_ERR_INVALID, "invalid solver state: audit energy became nonfinite");
    *out = a;
    return MPM_OK;
}

// Simulation::advance_while, solver.rs:413-472, for instance 0.  With a
// callback, it is asked before every trial step as in the reference; without
// one, trial steps run in device-driven batches.
int mpm_advance_while(mpm_engine* e, double target, int (*keep_going)(void*), void* ctx, int* reached) {
    if (!e) return MPM_ERR_INVALID;
    if (reached) *reached = 0;
    if (int rc = e->pull()) return rc;
    InstState& s0 = e->h_is[0];
    if (!std::isfinite(target) || target < s0.time) return e->fail(MPM_ERR_INVALID, "invalid target time");
    if (int rc = e->prepare()) return rc;
    for (int i = 1; i < e->ninst; ++i) e->h_is[i].mode = M_DONE;
    int status = 0;
    if (!keep_going) {
        s0.target = target;
        s0.mode = M_RUN;
        s0.pause = 0;
        s0.trial_dt = 0.0;
        s0.budget = INT_MAX;
        if (int rc = e->push()) return rc;
        if (int rc = e->run(e->launch.steps_per_sync, &status, nullptr)) return rc;
    } else {
        while (true) {
            InstState& s = e->h_is[0];
            if (!(s.time < target)) {
                s.time = target;
                s.mode = M_DONE;
                status = 0;
                if (int rc = e->push()) return rc;
                break;
            }
            if (!keep_going(ctx)) return MPM_OK;
            s.target = target;
            s.mode = M_RUN;
            s.pause = 0;
            s.trial_dt = 0.0;
            s.budget = 1;
            if (int rc = e->push()) return rc;
            status = 0;
            if (int rc = e->run(1, &status, nullptr, 1)) return rc;
            if (status == MPM_ERR_MIN_STEP) break;
        }
    }
    if (status == MPM_ERR_MIN_STEP) return MPM_ERR_MIN_STEP;
    if (reached) *reached = 1;
    return MPM_OK;
}

int mpm_advance_to(mpm_engine* e, double target) {
    int reached = 0;
    return mpm_advance_while(e, target, nullptr, nullptr, &reached);
}

int mpm_batch_advance(mpm_engine* e, const double* targets, uint64_t max_steps, int* status, double* done_ms) {
    if (!e || !targets) return MPM_ERR_INVALID;
    if (int rc = e->pull()) return rc;
    for (int i = 0; i < e->ninst; ++i)
        if (!std::isfinite(targets[i]) || targets[i] < e->h_is[i].time)
            return e->fail(MPM_ERR_INVALID, "invalid target time for instance " + std::to_string(i));
    if (int rc = e->prepare()) return rc;
    int budget = max_steps == 0 || max_steps > (uint64_t)INT_MAX ? INT_MAX : (int)max_steps;
    for (int i = 0; i < e->ninst; ++i) {
        InstState& s = e->h_is[i];
        s.target = targets[i];
        s.mode = M_RUN;
        s.pause = 0;
        s.trial_dt = 0.0;
        s.budget = budget;
        if (status) status[i] = 0;
        if (done_ms) done_ms[i] = 0.0;
    }
    if (int rc = e->push()) return rc;
    return e->run(e->launch.steps_per_sync, status, done_ms);
}

double mpm_stability_limit(mpm_engine* e) {
    return e ? e->stability_limit(0) : std::numeric_limits<double>::quiet_NaN();
}

double mpm_mass_residual(mpm_engine* e) {
    if (!e || e->pull()) return std::numeric_limits<double>::quiet_NaN();
    HostParticles h;
    if (e->download(0, h)) return std::numeric_limits<double>::quiet_NaN();
    double m = 0.0;
    for (size_t p = 0; p < h.n; ++p) {
        double mp, v0;
        e->inst[0].exact.get(h.id[p], h.sh[SMASS * h.n + p], h.sh[SVOL0 * h.n + p], mp, v0);
        m += mp;
    }
    const double* L = e->h_is[0].ledger;
    return m + L[1] - L[0];
}

int mpm_momentum_residual(mpm_engine* e, double* out) {
    if (!e || !out) return MPM_ERR_INVALID;
    if (int rc = e->pull()) return rc;
    HostParticles h;
    if (int rc = e->download(0, h)) return rc;
    std::vector<PelletState> pel;
    if (int rc = e->download_pellets(0, pel)) return rc;
    size_t n = h.n;
    double p[3] = {0.0, 0.0, 0.0};
    for (size_t i = 0; i < n; ++i) {
        double m, v0;
        e->inst[0].exact.get(h.id[i], h.sh[SMASS * n + i], h.sh[SVOL0 * n + i], m, v0);
        for (int a = 0; a < 3; ++a) p[a] += (double)h.s[(VX + a) * n + i] * m;
    }
    for (const PelletState& pl : pel)
        for (int a = 0; a < 3; ++a) p[a] += pl.vel[a] * pl.mass;
    const double* L = e->h_is[0].ledger;
    for (int a = 0; a < 3; ++a) out[a] = p[a] - L[2 + a] - L[6 + a] + L[28 + a] - L[9 + a] - L[12 + a];
    return MPM_OK;
}

int mpm_energies(mpm_engine* e, double* kinetic, double* potential, double* elastic) {
    if (!e) return MPM_ERR_INVALID;
    return e->energies(0, kinetic, potential, elastic);
}

int mpm_paste_update_batch(const mpm_material* material, double dt, size_t n, const double* trial, double* deformation,
                           double* kirchhoff, double* increment, double* dissipation_density, int* codes) {
    if (!material || !trial) return MPM_ERR_INVALID;
    Params P = batch_params(material);
    std::vector<float> h(9 * n);
    for (size_t p = 0; p < n; ++p)
        for (int a = 0; a < 9; ++a) h[9 * p + a] = (float)(trial[9 * p + a] - (a % 4 == 0 ? 1.0 : 0.0));
    DevBuf<float> din, dout, dtau, dinc, ddd;
    DevBuf<int> dcodes;
    if (din.alloc(9 * n) || dout.alloc(9 * n) || dtau.alloc(6 * n) || dinc.alloc(n) || ddd.alloc(n) ||
        dcodes.alloc(n))
        return MPM_ERR_DEVICE;
    if (hipMemcpy(din.p, h.data(), sizeof(float) * 9 * n, hipMemcpyHostToDevice)) return MPM_ERR_DEVICE;
    launch_paste_batch(nullptr, P, (float)dt, (int)n, din.p, dout.p, dtau.p, dinc.p, ddd.p, dcodes.p);
    if (hipGetLastError() || hipDeviceSynchronize()) return MPM_ERR_DEVICE;
    std::vector<float> ho(9 * n), ht(6 * n), hi(n), hd(n);
    std::vector<int> hc(n);
    if (hipMemcpy(ho.data(), dout.p, sizeof(float) * 9 * n, hipMemcpyDeviceToHost) ||
        hipMemcpy(ht.data(), dtau.p, sizeof(float) * 6 * n, hipMemcpyDeviceToHost) ||
        hipMemcpy(hi.data(), dinc.p, sizeof(float) * n, hipMemcpyDeviceToHost) ||
        hipMemcpy(hd.data(), ddd.p, sizeof(float) * n, hipMemcpyDeviceToHost) ||
        hipMemcpy(hc.data(), dcodes.p, sizeof(int) * n, hipMemcpyDeviceToHost))
        return MPM_ERR_DEVICE;
    for (size_t p = 0; p < n; ++p) {
        for (int a = 0; a < 9; ++a)
            if (deformation) deformation[9 * p + a] = (a % 4 == 0 ? 1.0 : 0.0) + (double)ho[9 * p + a];
        if (kirchhoff) {
            double* t = kirchhoff + 9 * p;
            t[0] = ht[6 * p];
            t[4] = ht[6 * p + 1];
            t[8] = ht[6 * p + 2];
            t[1] = t[3] = ht[6 * p + 3];
            t[2] = t[6] = ht[6 * p + 4];
            t[5] = t[7] = ht[6 * p + 5];
        }
        if (increment) increment[p] = hi[p];
        if (dissipation_density) dissipation_density[p] = hd[p];
        if (codes) codes[p] = hc[p];
    }
    return MPM_OK;
}

int mpm_liquid_update_batch(const mpm_material* material, double dt, size_t n, const double* volume_ratio,
                            const double* gradient, double* new_ratio, double* kirchhoff, int* codes) {
    if (!material || !volume_ratio || !gradient) return MPM_ERR_INVALID;
    Params P = batch_params(material);
    std::vector<float> j(n), c(9 * n);
    for (size_t p = 0; p < n; ++p) {
        j[p] = (float)(volume_ratio[p] - 1.0);
        for (int a = 0; a < 9; ++a) c[9 * p + a] = (float)gradient[9 * p + a];
    }
    DevBuf<float> dj, dc, djo, dt6;
    DevBuf<int> dcodes;
    if (dj.alloc(n) || dc.alloc(9 * n) || djo.alloc(n) || dt6.alloc(6 * n) || dcodes.alloc(n)) return MPM_ERR_DEVICE;
    if (hipMemcpy(dj.p, j.data(), sizeof(float) * n, hipMemcpyHostToDevice) ||
        hipMemcpy(dc.p, c.data(), sizeof(float) * 9 * n, hipMemcpyHostToDevice))
        return MPM_ERR_DEVICE;
    launch_liquid_batch(nullptr, P, (float)dt, (int)n, dj.p, dc.p, djo.p, dt6.p, dcodes.p);
    if (hipGetLastError() || hipDeviceSynchronize()) return MPM_ERR_DEVICE;
    std::vector<float> jo(n), t6(6 * n);
    std::vector<int> hc(n);
    if (hipMemcpy(jo.data(), djo.p, sizeof(float) * n, hipMemcpyDeviceToHost) ||
        hipMemcpy(t6.data(), dt6.p, sizeof(float) * 6 * n, hipMemcpyDeviceToHost) ||
        hipMemcpy(hc.data(), dcodes.p, sizeof(int) * n, hipMemcpyDeviceToHost))
        return MPM_ERR_DEVICE;
    for (size_t p = 0; p < n; ++p) {
        if (new_ratio) new_ratio[p] = 1.0 + (double)jo[p];
        if (kirchhoff) {
            double* t = kirchhoff + 9 * p;
            t[0] = t6[6 * p];
            t[4] = t6[6 * p + 1];
            t[8] = t6[6 * p + 2];
            t[1] = t[3] = t6[6 * p + 3];
            t[2] = t[6] = t6[6 * p + 4];
            t[5] = t[7] = t6[6 * p + 5];
        }
        if (codes) codes[p] = hc[p];
    }
    return MPM_OK;
}

void mpm_enable_profiling(mpm_engine* e, int on) {
    if (!e) return;
    e->profiling = on > 0;
    e->profile_every = on > 0 ? on : 1;
    e->profile_seq = 0;
    e->g2p_ms = 0.0;
    e->p2g_ms = 0.0;
    e->profiled_launches = 0;
}

int mpm_profile_kernels(mpm_engine* e, double* g2p_avg_ms, double* p2g_avg_ms, uint64_t* launches) {
    if (!e) return MPM_ERR_INVALID;
    const double n = (double)e->profiled_launches;
    if (g2p_avg_ms) *g2p_avg_ms = n > 0 ? e->g2p_ms / n : 0.0;
    if (p2g_avg_ms) *p2g_avg_ms = n > 0 ? e->p2g_ms / n : 0.0;
    if (launches) *launches = e->profiled_launches;
    return MPM_OK;
}

int mpm_set_surrogate(mpm_engine* e, const mpm_mlp* mlp, double tolerance, int polish_steps) {
    if (!e) return MPM_ERR_INVALID;
    if (mlp && (!(tolerance > 0.0) || polish_steps < 0 || polish_steps > 8))
        return e->fail(MPM_ERR_INVALID, "surrogate tolerance or correction count out of range");
    e->mlp = mlp;
    e->mlp_tol = (float)tolerance;
    e->mlp_polish = polish_steps;
    e->graph_stale = true;
    if (hipMemset(e->mlp_stats, 0, 2 * sizeof(unsigned long long)) != hipSuccess)
        return e->fail(MPM_ERR_DEVICE, "device failure clearing the surrogate counters");
    return MPM_OK;
}

int mpm_surrogate_stats(mpm_engine* e, uint64_t* considered, uint64_t* accepted) {
    if (!e) return MPM_ERR_INVALID;
    unsigned long long h[2] = {0, 0};
    if (hipMemcpy(h, e->mlp_stats, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess)
        return e->fail(MPM_ERR_DEVICE, "device failure reading the surrogate counters");
    if (considered) *considered = h[0];
    if (accepted) *accepted = h[1];
    return MPM_OK;
}

int mpm_log_returns(mpm_engine* e, size_t capacity) {
    if (!e || capacity > (size_t)INT32_MAX / 2) return MPM_ERR_INVALID;
    (void)hipFree(e->rlog);
    (void)hipFree(e->rlog_count);
    e->rlog = nullptr;
    e->rlog_count = nullptr;
    e->rlog_cap = 0;
    e->graph_stale = true;
    if (capacity == 0) return MPM_OK;
    if (hipMalloc(&e->rlog, sizeof(float) * kLogStride * capacity) != hipSuccess ||
        hipMalloc(&e->rlog_count, sizeof(int)) != hipSuccess || hipMemset(e->rlog_count, 0, sizeof(int)) != hipSuccess)
        return e->fail(MPM_ERR_DEVICE, "cannot allocate the return log");
    e->rlog_cap = (int)capacity;
    return MPM_OK;
}

int mpm_take_returns(mpm_engine* e, float* inputs, float* targets, float* aux, size_t capacity, size_t* n) {
    if (!e || !n) return MPM_ERR_INVALID;
    *n = 0;
    if (!e->rlog) return MPM_OK;
    int count = 0;
    if (hipMemcpy(&count, e->rlog_count, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
        return e->fail(MPM_ERR_DEVICE, "device failure reading the return log");
    size_t k = std::min({(size_t)std::max(count, 0), (size_t)e->rlog_cap, capacity});
    std::vector<float> buf(k * kLogStride);
    if (k && hipMemcpy(buf.data(), e->rlog, sizeof(float) * buf.size(), hipMemcpyDeviceToHost) != hipSuccess)
        return e->fail(MPM_ERR_DEVICE, "device failure reading the return log");
    for (size_t i = 0; i < k; ++i) {
        const float* r = buf.data() + i * kLogStride;
        if (inputs) std::memcpy(inputs + i * 12, r, sizeof(float) * 12);
        if (targets) std::memcpy(targets + i * 7, r + 12, sizeof(float) * 7);
        if (aux) std::memcpy(aux + i * 6, r + 19, sizeof(float) * 6);
    }
    if (hipMemset(e->rlog_count, 0, sizeof(int)) != hipSuccess)
        return e->fail(MPM_ERR_DEVICE, "device failure clearing the return log");
    *n = k;
    return MPM_OK;
}

int mpm_autotune(mpm_engine* e, int steps, mpm_launch* chosen) {
    if (!e || steps < 1) return MPM_ERR_INVALID;
    return e->autotune(steps, chosen);
}

int mpm_get_launch(const mpm_engine* e, mpm_launch* out) {
    if (!e || !out) return MPM_ERR_INVALID;
    *out = e->launch;
    return MPM_OK;
}

int mpm_set_launch(mpm_engine* e, const mpm_launch* l) {
    if (!e || !l) return MPM_ERR_INVALID;
    // More than two cell blocks per work item would not fit the LDS tiles.
    if (l->steps_per_sync < 1 || l->steps_per_sync > kMaxStepsPerSync || l->merge_blocks < 1 ||
        l->merge_blocks > 2 || l->traction_wgs < 0 || l->traction_wgs > 65536 || l->g2p_occupancy < 2 ||
        l->g2p_occupancy > 3 || (l->p2g_block != 256 && l->p2g_block != 512))
        return e->fail(MPM_ERR_INVALID, "launch shape out of range");
    bool resort = l->merge_blocks != e->launch.merge_blocks;
    e->launch = *l;
    e->graph_stale = true;
    if (resort) e->dirty = true;
    return MPM_OK;
}

}  // extern "C"
