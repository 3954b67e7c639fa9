// Host side of the MI355X MLS-MPM engine: the C ABI of include/mpm_engine.h.
//
// The host mirrors the reference's driver (core/src/mpm/solver.rs:357-611) in
// f64: the adaptive-dt loop, accept/reject, ledgers, the 32-step particle sort
// and the rigid pellets (rigid.rs), while every per-particle and per-node phase
// runs in the HIP kernels of kernels.hip.  There is no CPU fallback for any
// kernel: without a usable GPU every stepping entry point fails with
// MPM_ERR_DEVICE.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../../include/mpm_engine.h"
#include "engine_internal.h"

using namespace mpm;

namespace {

constexpr uint64_t kSortInterval = 32;  // solver.rs:39
constexpr double kDtRecovery = 1.25;   // solver.rs:45

struct V3 {
    double x[3];
    double& operator[](int i) { return x[i]; }
    double operator[](int i) const { return x[i]; }
};
inline V3 v3(double a, double b, double c) { return V3{{a, b, c}}; }
inline V3 operator+(const V3& a, const V3& b) { return v3(a[0] + b[0], a[1] + b[1], a[2] + b[2]); }
inline V3 operator-(const V3& a, const V3& b) { return v3(a[0] - b[0], a[1] - b[1], a[2] - b[2]); }
inline V3 operator*(const V3& a, double s) { return v3(a[0] * s, a[1] * s, a[2] * s); }
inline double dot(const V3& a, const V3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline double norm(const V3& a) { return std::sqrt(dot(a, a)); }
inline V3 cross(const V3& a, const V3& b) {
    return v3(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]);
}
inline bool finite(const V3& a) { return std::isfinite(a[0]) && std::isfinite(a[1]) && std::isfinite(a[2]); }

struct Quat {
    double i, j, k, w;
};
Quat qmul(const Quat& a, const Quat& b) {
    return Quat{a.w * b.i + a.i * b.w + a.j * b.k - a.k * b.j, a.w * b.j - a.i * b.k + a.j * b.w + a.k * b.i,
                a.w * b.k + a.i * b.j - a.j * b.i + a.k * b.w, a.w * b.w - a.i * b.i - a.j * b.j - a.k * b.k};
}
V3 qrot(const Quat& q, const V3& v) {
    V3 qv = v3(q.i, q.j, q.k);
    V3 t = cross(qv, v) * 2.0;
    return t * q.w + cross(qv, t) + v;
}
void qmatrix(const Quat& q, double r[9]) {
    double i = q.i, j = q.j, k = q.k, w = q.w;
    double ww = w * w, ii = i * i, jj = j * j, kk = k * k;
    double ij = i * j * 2.0, wk = w * k * 2.0, wj = w * j * 2.0, ik = i * k * 2.0, jk = j * k * 2.0, wi = w * i * 2.0;
    double m[9] = {ww + ii - jj - kk, ij - wk, wj + ik, wk + ij, ww - ii + jj - kk, jk - wi, ik - wj, wi + jk,
                   ww - ii - jj + kk};
    std::memcpy(r, m, sizeof(m));
}
V3 mv(const double r[9], const V3& x) {
    return v3(r[0] * x[0] + r[1] * x[1] + r[2] * x[2], r[3] * x[0] + r[4] * x[1] + r[5] * x[2],
              r[6] * x[0] + r[7] * x[1] + r[8] * x[2]);
}
V3 mtv(const double r[9], const V3& x) {
    return v3(r[0] * x[0] + r[3] * x[1] + r[6] * x[2], r[1] * x[0] + r[4] * x[1] + r[7] * x[2],
              r[2] * x[0] + r[5] * x[1] + r[8] * x[2]);
}
// Exponential of the pure quaternion a/2 (a rotation by |a| about a).
Quat from_scaled_axis(const V3& a) {
    V3 half = a * 0.5;
    double nn = dot(half, half);
    const double eps = std::numeric_limits<double>::epsilon();
    if (nn <= eps * eps) return Quat{0, 0, 0, 1};
    double n = std::sqrt(nn);
    double f = std::sin(n) / n;
    return Quat{half[0] * f, half[1] * f, half[2] * f, std::cos(n)};
}

// Rigid multisphere pellet, rigid.rs:30-344 (f64 on the host).
struct Pellet {
    double radius = 0, length = 0, mass = 0;
    V3 inertia{{0, 0, 0}}, pos{{0, 0, 0}};
    Quat q{0, 0, 0, 1};
    V3 vel{{0, 0, 0}}, L{{0, 0, 0}};
    std::vector<double> offsets;

    V3 axis() const { return qrot(q, v3(1, 0, 0)); }
    V3 omega() const {
        double r[9];
        qmatrix(q, r);
        V3 b = mtv(r, L);
        return mv(r, v3(b[0] / inertia[0], b[1] / inertia[1], b[2] / inertia[2]));
    }
    void bounding_box(V3& lo, V3& hi) const {
        V3 ax = axis();
        double half = 0.0;
        for (double o : offsets) half = std::max(half, std::fabs(o));
        V3 reach = v3(std::fabs(ax[0]) * half + radius, std::fabs(ax[1]) * half + radius,
                      std::fabs(ax[2]) * half + radius);
        lo = pos - reach;
        hi = pos + reach;
    }
    void apply_impulse(const V3& imp, const V3& ang) {
        vel = vel + v3(imp[0] / mass, imp[1] / mass, imp[2] / mass);
        L = L + ang;
    }
    double kinetic() const { return 0.5 * mass * dot(vel, vel) + 0.5 * dot(omega(), L); }
    bool finite_state() const {
        return finite(pos) && finite(vel) && finite(L) && std::isfinite(q.i) && std::isfinite(q.j) &&
               std::isfinite(q.k) && std::isfinite(q.w);
    }
};

struct WallContact {
    V3 impulse{{0, 0, 0}};
    double work = 0.0;
};

double contact_damping(const mpm_config& c, double mass) {
    double zeta;
    if (c.contact_restitution <= 0.0)
        zeta = 1.0;
    else if (c.contact_restitution >= 1.0)
        zeta = 0.0;
    else {
        double lg = std::log(c.contact_restitution);
        zeta = -lg / std::sqrt(M_PI * M_PI + lg * lg);
    }
    return 2.0 * zeta * std::sqrt(c.contact_stiffness * mass);
}

// Pellet::integrate, rigid.rs:279-334.
WallContact integrate(Pellet& p, double dt, const V3& g, const V3& ext, const mpm_config& c) {
    double damping = contact_damping(c, p.mass);
    V3 force{{0, 0, 0}}, torque{{0, 0, 0}};
    WallContact led;
    V3 om = p.omega(), ax = p.axis();
    for (double o : p.offsets) {
        V3 center = p.pos + ax * o;
        for (int a = 0; a < 3; ++a)
            for (int side = 0; side < 2; ++side) {
                double face = side == 0 ? 0.0 : ext[a];
                double sign = side == 0 ? 1.0 : -1.0;
                double pen = p.radius - (center[a] - face) * sign;
                if (pen <= 0.0) continue;
                V3 nrm{{0, 0, 0}};
                nrm[a] = sign;
                V3 lever = center - nrm * p.radius - p.pos;
                V3 pv = p.vel + cross(om, lever);
                double vn = dot(pv, nrm);
                double fn = std::max(c.contact_stiffness * pen - damping * vn, 0.0);
                V3 t = pv - nrm * vn;
                double ts = norm(t);
                V3 fr{{0, 0, 0}};
                if (ts > 1e-12) {
                    double lim = std::min(c.contact_friction * fn, p.mass * ts / dt);
                    fr = v3(-t[0] / ts, -t[1] / ts, -t[2] / ts) * lim;
                }
                V3 total = nrm * fn + fr;
                force = force + total;
                torque = torque + cross(lever, total);
                led.work += dot(total, pv) * dt;
            }
    }
    V3 imp = force * dt;
    led.impulse = imp;
    p.vel = p.vel + (g * dt + v3(imp[0] / p.mass, imp[1] / p.mass, imp[2] / p.mass));
    p.L = p.L + torque * dt;
    p.pos = p.pos + p.vel * dt;
    p.q = qmul(from_scaled_axis(p.omega() * dt), p.q);
    return led;
}

// Symmetric 3x3 cyclic Jacobi in f64 (constitutive.rs:373-425), used for the
// elastic-energy diagnostic.
bool eigen3(const double* m, double* vals) {
    double d[9];
    for (int i = 0; i < 9; ++i) d[i] = 0.5 * (m[i] + m[3 * (i % 3) + i / 3]);
    double scale = 0.0;
    for (int i = 0; i < 9; ++i) scale += d[i] * d[i];
    scale = std::max(scale, std::numeric_limits<double>::min());
    for (int sweep = 0; sweep < 64; ++sweep) {
        double off = d[1] * d[1] + d[2] * d[2] + d[5] * d[5];
        if (off <= 1e-30 * scale) {
            vals[0] = d[0];
            vals[1] = d[4];
            vals[2] = d[8];
            return true;
        }
        const int pq[3][2] = {{0, 1}, {0, 2}, {1, 2}};
        for (auto& r : pq) {
            int p = r[0], q = r[1];
            double apq = d[3 * p + q];
            if (apq == 0.0) continue;
            double th = (d[3 * q + q] - d[3 * p + p]) / (2.0 * apq);
            double t = std::copysign(1.0, th) / (std::fabs(th) + std::sqrt(th * th + 1.0));
            double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
            for (int k = 0; k < 3; ++k) {
                double a = d[3 * k + p], b = d[3 * k + q];
                d[3 * k + p] = c * a - s * b;
                d[3 * k + q] = s * a + c * b;
            }
            for (int k = 0; k < 3; ++k) {
                double a = d[3 * p + k], b = d[3 * q + k];
                d[3 * p + k] = c * a - s * b;
                d[3 * q + k] = s * a + c * b;
            }
        }
    }
    return false;
}

double wave_speed(const mpm_material& m) {
    double modulus = m.kind == 0 ? m.lame_lambda + 2.0 * m.shear_modulus : m.bulk_modulus;
    return std::sqrt(modulus / m.density);
}
double diffusivity(const mpm_material& m) { return m.kind == 0 ? 0.0 : m.consistency / m.density; }

#define HIP_OK(expr)                                                    \
    do {                                                                \
        hipError_t e_ = (expr);                                         \
        if (e_ != hipSuccess) {                                         \
            return fail_device(#expr, e_);                              \
        }                                                               \
    } while (0)

}  // namespace

struct mpm_engine {
    mpm_material mat{};
    mpm_config cfg{};
    Params P{};
    double dx = 0.0;
    int cells[3] = {0, 0, 0};
    size_t cap = 0, n = 0, max_pellets = 0;
    // device particle storage
    float* slab[2] = {nullptr, nullptr};
    float* shared[2] = {nullptr, nullptr};
    uint32_t* ids[2] = {nullptr, nullptr};
    int cur = 0, sh = 0;
    // device grid
    size_t na = 0;
    int nblocks = 0;
    float* gmem[2] = {nullptr, nullptr};
    int* gflag[2] = {nullptr, nullptr};
    GridPtrs grid[2]{};
    int gcur = 0;
    bool grid_valid = false;
    float* gv = nullptr;
    float* trS = nullptr;
    float* trT = nullptr;
    StepStatus* d_st = nullptr;
    StepStatus* h_st = nullptr;
    // pellets
    std::vector<Pellet> pellets;
    PelletDev* d_pel = nullptr;
    PelletAccum* d_acc = nullptr;
    PelletAccum* h_acc = nullptr;
    int* d_bstart = nullptr;
    int* d_blist = nullptr;
    size_t blist_cap = 0;
    // per-workgroup partials of the step reductions
    int ngwg = 0, npwg = 0;
    double* gpart = nullptr;
    int* gbad = nullptr;
    float* pmax = nullptr;
    double* pdiss = nullptr;
    unsigned long long* pfail = nullptr;
    int* pesc = nullptr;
    // particles within half a cell of a face, one list per particle slab
    int* nw_list[2] = {nullptr, nullptr};
    int* nw_count = nullptr;  // two counters
    int nw_host = 0;
    // reference order keys (kept with the shared slab)
    uint64_t* refkey[2] = {nullptr, nullptr};
    // workgroup work items: runs of particles in one or two z-adjacent 4^3-cell blocks
    int2* d_work = nullptr;
    int nwork = 0;
    size_t work_cap = 0;
    uint32_t* run_unique = nullptr;
    uint32_t* run_count = nullptr;
    int* d_nruns = nullptr;
    int cb[3] = {0, 0, 0};  // cell blocks per axis
    // sort
    uint64_t* keys[2] = {nullptr, nullptr};
    uint32_t* order[2] = {nullptr, nullptr};
    void* sort_temp = nullptr;
    size_t sort_bytes = 0;
    int key_bits = 0;
    // host state, solver.rs:324-352
    mpm_ledger ledger{};
    double time = 0.0;
    uint64_t step_count = 0, limited_steps = 0, rejected_steps = 0;
    double dt_scale = 1.0, min_dt_used = std::numeric_limits<double>::infinity(), max_dt_used = 0.0;
    double max_speed = 0.0;
    double particle_mass = 0.0;  // sum of the stored (f32) particle masses, in f64
    uint64_t steps_since_set = 0;
    int last_failure = 0;
    std::string err;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool profiling = false;
    double g2p2g_ms = 0.0;
    uint64_t g2p2g_launches = 0;

    int fail_device(const char* what, hipError_t e) {
        err = std::string("HIP error in ") + what + ": " + hipGetErrorString(e);
        return MPM_ERR_DEVICE;
    }
    int fail(int code, const std::string& msg) {
        err = msg;
        return code;
    }
    GridPtrs make_grid(float* mem, int* flag) const {
        GridPtrs g;
        g.m = mem;
        g.px = mem + na;
        g.py = mem + 2 * na;
        g.pz = mem + 3 * na;
        g.fx = mem + 4 * na;
        g.fy = mem + 5 * na;
        g.fz = mem + 6 * na;
        g.flag = flag;
        return g;
    }
    V3 extents() const { return v3(cells[0] * dx, cells[1] * dx, cells[2] * dx); }
    V3 gravity() const { return v3(cfg.gravity[0], cfg.gravity[1], cfg.gravity[2]); }
    void fill_params() {
        P.dx = (float)dx;
        P.inv_dx = (float)(1.0 / dx);
        for (int a = 0; a < 3; ++a) {
            P.cells[a] = cells[a];
            P.nodes[a] = cells[a] + 3;
            P.nb[a] = (cells[a] + 3 + 3) / 4;
            P.extent[a] = (float)(cells[a] * dx);
            P.gravity[a] = (float)cfg.gravity[a];
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
    }

    int alloc(size_t capacity, size_t pellets_cap) {
        cap = std::max<size_t>(capacity, 1);
        max_pellets = pellets_cap;
        for (int b = 0; b < 2; ++b) {
            HIP_OK(hipMalloc(&slab[b], sizeof(float) * NPF * cap));
            HIP_OK(hipMemset(slab[b], 0, sizeof(float) * NPF * cap));
            HIP_OK(hipMalloc(&shared[b], sizeof(float) * NSF * cap));
            HIP_OK(hipMemset(shared[b], 0, sizeof(float) * NSF * cap));
            HIP_OK(hipMalloc(&ids[b], sizeof(uint32_t) * cap));
            HIP_OK(hipMalloc(&keys[b], sizeof(uint64_t) * cap));
            HIP_OK(hipMalloc(&refkey[b], sizeof(uint64_t) * cap));
            HIP_OK(hipMalloc(&nw_list[b], sizeof(int) * cap));
            HIP_OK(hipMalloc(&order[b], sizeof(uint32_t) * cap));
        }
        nblocks = P.nb[0] * P.nb[1] * P.nb[2];
        na = (size_t)nblocks * 64;
        for (int b = 0; b < 2; ++b) {
            HIP_OK(hipMalloc(&gmem[b], sizeof(float) * 7 * na));
            HIP_OK(hipMemset(gmem[b], 0, sizeof(float) * 7 * na));
            HIP_OK(hipMalloc(&gflag[b], sizeof(int) * nblocks));
            HIP_OK(hipMemset(gflag[b], 0, sizeof(int) * nblocks));
            grid[b] = make_grid(gmem[b], gflag[b]);
        }
        HIP_OK(hipMalloc(&gv, sizeof(float) * 3 * na));
        HIP_OK(hipMemset(gv, 0, sizeof(float) * 3 * na));
        HIP_OK(hipMalloc(&trS, sizeof(float) * 3 * na));
        HIP_OK(hipMemset(trS, 0, sizeof(float) * 3 * na));
        HIP_OK(hipMalloc(&trT, sizeof(float) * 3 * na));
        HIP_OK(hipMemset(trT, 0, sizeof(float) * 3 * na));
        HIP_OK(hipMalloc(&d_st, sizeof(StepStatus)));
        HIP_OK(hipMalloc(&nw_count, sizeof(int) * 2));
        HIP_OK(hipMemset(nw_count, 0, sizeof(int) * 2));
        ngwg = grid_workgroups(nblocks);
        for (int a = 0; a < 3; ++a) cb[a] = (cells[a] + 4) / 4;
        work_cap = (size_t)cb[0] * cb[1] * cb[2] + cap / (size_t)g2p2g_chunk() + 16;
        npwg = (int)work_cap;
        HIP_OK(hipMalloc(&d_work, sizeof(int2) * work_cap));
        HIP_OK(hipMalloc(&run_unique, sizeof(uint32_t) * cap));
        HIP_OK(hipMalloc(&run_count, sizeof(uint32_t) * cap));
        HIP_OK(hipMalloc(&d_nruns, sizeof(int)));
        HIP_OK(hipMalloc(&gpart, sizeof(double) * 12 * ngwg));
        HIP_OK(hipMalloc(&gbad, sizeof(int) * ngwg));
        HIP_OK(hipMalloc(&pmax, sizeof(float) * 2 * npwg));
        HIP_OK(hipMalloc(&pdiss, sizeof(double) * npwg));
        HIP_OK(hipMalloc(&pfail, sizeof(unsigned long long) * npwg));
        HIP_OK(hipMalloc(&pesc, sizeof(int) * npwg));
        HIP_OK(hipHostMalloc(&h_st, sizeof(StepStatus), hipHostMallocDefault));
        if (max_pellets) {
            HIP_OK(hipMalloc(&d_pel, sizeof(PelletDev) * max_pellets));
            HIP_OK(hipMalloc(&d_acc, sizeof(PelletAccum) * max_pellets));
            HIP_OK(hipHostMalloc(&h_acc, sizeof(PelletAccum) * max_pellets, hipHostMallocDefault));
            HIP_OK(hipMalloc(&d_bstart, sizeof(int) * (nblocks + 1)));
        }
        uint64_t kmax = (uint64_t)((cells[0] + 4) / 4) * ((cells[1] + 4) / 4) * ((cells[2] + 4) / 4) * 64;
        key_bits = 32;
        while ((1ull << (key_bits - 32)) < kmax) ++key_bits;
        sort_bytes = std::max(sort_temp_bytes((int)cap, key_bits), rle_temp_bytes((int)cap));
        HIP_OK(hipMalloc(&sort_temp, std::max<size_t>(sort_bytes, 16)));
        HIP_OK(hipStreamCreate(&stream));
        HIP_OK(hipEventCreate(&ev0));
        HIP_OK(hipEventCreate(&ev1));
        return MPM_OK;
    }

    ~mpm_engine() {
        for (int b = 0; b < 2; ++b) {
            (void)hipFree(slab[b]);
            (void)hipFree(shared[b]);
            (void)hipFree(ids[b]);
            (void)hipFree(keys[b]);
            (void)hipFree(refkey[b]);
            (void)hipFree(nw_list[b]);
            (void)hipFree(order[b]);
            (void)hipFree(gmem[b]);
            (void)hipFree(gflag[b]);
        }
        (void)hipFree(gv);
        (void)hipFree(trS);
        (void)hipFree(trT);
        (void)hipFree(d_st);
        (void)hipFree(nw_count);
        (void)hipFree(gpart);
        (void)hipFree(d_work);
        (void)hipFree(run_unique);
        (void)hipFree(run_count);
        (void)hipFree(d_nruns);
        (void)hipFree(gbad);
        (void)hipFree(pmax);
        (void)hipFree(pdiss);
        (void)hipFree(pfail);
        (void)hipFree(pesc);
        if (h_st) (void)hipHostFree(h_st);
        (void)hipFree(d_pel);
        (void)hipFree(d_acc);
        if (h_acc) (void)hipHostFree(h_acc);
        (void)hipFree(d_bstart);
        (void)hipFree(d_blist);
        (void)hipFree(sort_temp);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }

    // Deposit the committed state into the current accumulator set.
    int rebuild_grid() {
        HIP_OK(hipMemsetAsync(gmem[gcur], 0, sizeof(float) * 7 * na, stream));
        HIP_OK(hipMemsetAsync(gflag[gcur], 0, sizeof(int) * nblocks, stream));
        launch_p2g(stream, P, (int)n, slab[cur], cap, shared[sh], grid[gcur]);
        HIP_OK(hipGetLastError());
        int rc = rebuild_nearwall();
        if (rc) return rc;
        grid_valid = true;
        return MPM_OK;
    }

    // Near-wall particle list of the committed slab.
    int rebuild_nearwall() {
        HIP_OK(hipMemsetAsync(nw_count + cur, 0, sizeof(int), stream));
        launch_nearwall(stream, P, (int)n, slab[cur], cap, nw_list[cur], nw_count + cur);
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpyAsync(&nw_host, nw_count + cur, sizeof(int), hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
        return MPM_OK;
    }

    // Pellet geometry and per-block pellet lists for the coupling kernel.
    int upload_pellets() {
        size_t np = pellets.size();
        std::vector<PelletDev> dev(np);
        std::vector<int> count(nblocks + 1, 0);
        std::vector<std::vector<int>> ranges(np);
        for (size_t k = 0; k < np; ++k) {
            const Pellet& p = pellets[k];
            PelletDev& d = dev[k];
            V3 ax = p.axis(), om = p.omega(), lo, hi;
            p.bounding_box(lo, hi);
            double margin = 0.0;
            for (int a = 0; a < 3; ++a) {
                d.pos[a] = p.pos[a];
                d.axis[a] = ax[a];
                d.vel[a] = p.vel[a];
                d.omega[a] = om[a];
                d.lo[a] = lo[a];
                d.hi[a] = hi[a];
            }
            d.radius = p.radius;
            d.spheres = (int)p.offsets.size();
            d.pad = 0;
            for (size_t s = 0; s < p.offsets.size() && s < (size_t)kMaxSpheres; ++s) d.offsets[s] = p.offsets[s];
            int b0[3], b1[3];
            bool empty = false;
            for (int a = 0; a < 3; ++a) {
                // Candidate node coordinates with one node of slack (the kernel re-tests the box exactly).
                double i0 = std::floor((lo[a] - margin) / dx + 1.0) - 1.0,
                       i1 = std::ceil((hi[a] + margin) / dx + 1.0) + 1.0;
                i0 = std::max(i0, 0.0);
                i1 = std::min(i1, (double)(cells[a] + 2));
                if (!(i0 <= i1)) empty = true;
                b0[a] = empty ? 0 : (int)i0 >> 2;
                b1[a] = empty ? -1 : (int)i1 >> 2;
            }
            if (empty) continue;
            for (int bi = b0[0]; bi <= b1[0]; ++bi)
                for (int bj = b0[1]; bj <= b1[1]; ++bj)
                    for (int bk = b0[2]; bk <= b1[2]; ++bk) {
                        int b = (bi * P.nb[1] + bj) * P.nb[2] + bk;
                        ranges[k].push_back(b);
                        count[b + 1] += 1;
                    }
        }
        for (int b = 0; b < nblocks; ++b) count[b + 1] += count[b];
        std::vector<int> list(std::max(count[nblocks], 1));
        std::vector<int> fillp(count.begin(), count.end() - 1);
        for (size_t k = 0; k < np; ++k)
            for (int b : ranges[k]) list[fillp[b]++] = (int)k;
        if ((size_t)list.size() > blist_cap) {
            (void)hipFree(d_blist);
            blist_cap = list.size() * 2;
            HIP_OK(hipMalloc(&d_blist, sizeof(int) * blist_cap));
        }
        HIP_OK(hipMemcpyAsync(d_pel, dev.data(), sizeof(PelletDev) * np, hipMemcpyHostToDevice, stream));
        HIP_OK(hipMemcpyAsync(d_bstart, count.data(), sizeof(int) * (nblocks + 1), hipMemcpyHostToDevice, stream));
        HIP_OK(hipMemcpyAsync(d_blist, list.data(), sizeof(int) * list.size(), hipMemcpyHostToDevice, stream));
        HIP_OK(hipMemsetAsync(d_acc, 0, sizeof(PelletAccum) * np, stream));
        // The host vectors must outlive the asynchronous copies.
        HIP_OK(hipStreamSynchronize(stream));
        return MPM_OK;
    }

    // solver.rs:477-498 (one contact limit per pellet).
    double stability_limit() const {
        double limit = cfg.cfl * dx / (wave_speed(mat) + max_speed);
        double diff = diffusivity(mat);
        if (diff > 0.0) limit = std::min(limit, 0.25 * dx * dx / diff);
        for (const Pellet& p : pellets) {
            limit = std::min(limit, 0.2 * std::sqrt(p.mass / cfg.contact_stiffness));
            double s = norm(p.vel);
            if (s > 0.0) limit = std::min(limit, cfg.cfl * dx / s);
        }
        return limit;
    }

    // Sort by 4^3-cell block (the device order) and rebuild the work items.  At the
    // reference's 32-step cadence (solver.rs:501-517) the reference order key is
    // refreshed too; keep_ref re-sorts for locality only, leaving the reference
    // order of the state unchanged.  The grid sums are order free.
    int sort_particles(int keep_ref = 0) {
        if (n == 0) {
            nwork = 0;
            return MPM_OK;
        }
        launch_sort_keys(stream, P, (int)n, slab[cur], cap, ids[sh], keys[0], order[0], refkey[sh], keep_ref);
        launch_radix_sort(stream, sort_temp, sort_bytes, keys[0], keys[1], order[0], order[1], (int)n, key_bits);
        launch_permute(stream, (int)n, order[1], slab[cur], slab[cur ^ 1], cap, shared[sh], shared[sh ^ 1], ids[sh],
                       ids[sh ^ 1], refkey[sh], refkey[sh ^ 1]);
        launch_block_runs(stream, sort_temp, sort_bytes, (int)n, keys[1], order[0], run_unique, run_count, d_nruns);
        HIP_OK(hipGetLastError());
        cur ^= 1;
        sh ^= 1;
        int nruns = 0;
        HIP_OK(hipMemcpyAsync(&nruns, d_nruns, sizeof(int), hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
        std::vector<uint32_t> uq(nruns), cnt(nruns);
        HIP_OK(hipMemcpyAsync(uq.data(), run_unique, sizeof(uint32_t) * nruns, hipMemcpyDeviceToHost, stream));
        HIP_OK(hipMemcpyAsync(cnt.data(), run_count, sizeof(uint32_t) * nruns, hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
        // Merge at most two z-adjacent blocks per work item; split runs longer than a chunk.
        const int chunk = g2p2g_chunk();
        std::vector<int2> work;
        work.reserve(nruns + n / chunk + 1);
        int start = 0;
        int cur_start = -1, cur_count = 0, cur_blocks = 0;
        uint32_t prev = 0;
        auto flush = [&]() {
            if (cur_start >= 0) work.push_back(make_int2(cur_start, cur_count));
            cur_start = -1;
            cur_count = 0;
            cur_blocks = 0;
        };
        for (int r = 0; r < nruns; ++r) {
            int c = (int)cnt[r];
            uint32_t b = uq[r];
            bool adjacent = cur_blocks == 1 && b == prev + 1 && (b % (uint32_t)cb[2]) != 0 && cur_count + c <= chunk;
            if (!adjacent) flush();
            if (c > chunk) {
                for (int off = 0; off < c; off += chunk) work.push_back(make_int2(start + off, std::min(chunk, c - off)));
            } else if (adjacent) {
                cur_count += c;
                cur_blocks = 2;
            } else {
                cur_start = start;
                cur_count = c;
                cur_blocks = 1;
            }
            prev = b;
            start += c;
        }
        flush();
        if (work.size() > work_cap) return fail(MPM_ERR_CAPACITY, "work list overflow");
        nwork = (int)work.size();
        HIP_OK(hipMemcpyAsync(d_work, work.data(), sizeof(int2) * work.size(), hipMemcpyHostToDevice, stream));
        HIP_OK(hipStreamSynchronize(stream));
        return rebuild_nearwall();
    }

    int try_step(double dt, mpm_step_report* report);
    int remove_escaped();
    std::vector<uint64_t> host_ref;
    int download(std::vector<float>& s, std::vector<float>& shv, std::vector<uint32_t>& id) {
        s.resize(NPF * n);
        shv.resize(NSF * n);
        id.resize(n);
        host_ref.resize(n);
        if (n) HIP_OK(hipMemcpyAsync(host_ref.data(), refkey[sh], sizeof(uint64_t) * n, hipMemcpyDeviceToHost, stream));
        for (int f = 0; f < NPF; ++f)
            HIP_OK(hipMemcpyAsync(s.data() + f * n, slab[cur] + f * cap, sizeof(float) * n, hipMemcpyDeviceToHost,
                                  stream));
        for (int f = 0; f < NSF; ++f)
            HIP_OK(hipMemcpyAsync(shv.data() + f * n, shared[sh] + f * cap, sizeof(float) * n,
                                  hipMemcpyDeviceToHost, stream));
        HIP_OK(hipMemcpyAsync(id.data(), ids[sh], sizeof(uint32_t) * n, hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
        return MPM_OK;
    }
    int upload(const std::vector<float>& s, const std::vector<float>& shv, const std::vector<uint32_t>& id) {
        for (int b = 0; b < 2; ++b) {
            for (int f = 0; f < NPF; ++f)
                HIP_OK(hipMemcpyAsync(slab[b] + f * cap, s.data() + f * n, sizeof(float) * n, hipMemcpyHostToDevice,
                                      stream));
        }
        for (int f = 0; f < NSF; ++f)
            HIP_OK(hipMemcpyAsync(shared[sh] + f * cap, shv.data() + f * n, sizeof(float) * n,
                                  hipMemcpyHostToDevice, stream));
        if (n) HIP_OK(hipMemcpyAsync(ids[sh], id.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, stream));
        if (n) HIP_OK(hipMemcpyAsync(refkey[sh], host_ref.data(), sizeof(uint64_t) * n, hipMemcpyHostToDevice, stream));
        HIP_OK(hipStreamSynchronize(stream));
        return MPM_OK;
    }
};

namespace {

float bits_float(unsigned int b) {
    float f;
    std::memcpy(&f, &b, sizeof(f));
    return f;
}

}  // namespace

// Remove particles that left the padded grid (sweep_outflow, solver.rs:520-532).
int mpm_engine::remove_escaped() {
    std::vector<float> s, shv;
    std::vector<uint32_t> id;
    int rc = download(s, shv, id);
    if (rc) return rc;
    size_t keep = 0;
    std::vector<float> s2(NPF * n), sh2(NSF * n);
    std::vector<uint32_t> id2(n);
    std::vector<uint64_t> ref2(n);
    double out_mass = 0.0;
    V3 out_mom{{0, 0, 0}};
    for (size_t p = 0; p < n; ++p) {
        bool inside = true;
        for (int a = 0; a < 3; ++a) {
            double sc = (double)s[(PX + a) * n + p] / dx;
            double bf = std::floor(sc - 0.5);
            double padded = bf + 1.0;
            if (std::isnan(padded) || padded < 0.0 || padded + 2.0 >= (double)(cells[a] + 3)) inside = false;
        }
        if (!inside) {
            double m = shv[SMASS * n + p];
            out_mass += m;
            for (int a = 0; a < 3; ++a) out_mom[a] += (double)s[(VX + a) * n + p] * m;
            continue;
        }
        for (int f = 0; f < NPF; ++f) s2[f * n + keep] = s[f * n + p];
        for (int f = 0; f < NSF; ++f) sh2[f * n + keep] = shv[f * n + p];
        id2[keep] = id[p];
        ref2[keep] = host_ref[p];
        ++keep;
    }
    size_t old = n;
    n = keep;
    std::vector<float> s3(NPF * n), sh3(NSF * n);
    for (int f = 0; f < NPF; ++f) std::copy(s2.begin() + f * old, s2.begin() + f * old + n, s3.begin() + f * n);
    for (int f = 0; f < NSF; ++f) std::copy(sh2.begin() + f * old, sh2.begin() + f * old + n, sh3.begin() + f * n);
    id2.resize(n);
    ref2.resize(n);
    host_ref = ref2;
    ledger.outflow_mass += out_mass;
    particle_mass -= out_mass;
    for (int a = 0; a < 3; ++a) ledger.outflow_momentum[a] += out_mom[a];
    rc = upload(s3, sh3, id2);
    if (rc) return rc;
    return sort_particles(1);
}

// Simulation::try_step_captured (solver.rs:540-611) on the device.
int mpm_engine::try_step(double dt, mpm_step_report* report) {
    if (report) *report = mpm_step_report{0, 0, 0.0, dt};
    if (!grid_valid) {
        int rc = rebuild_grid();
        if (rc) return rc;
    }
    float fdt = (float)dt;
    launch_reset_status(stream, d_st, nw_count + (cur ^ 1));
    launch_traction(stream, P, fdt, nw_list[cur], nw_count + cur, nw_host, slab[cur], cap, shared[sh], grid[gcur],
                    trS, trT, na);
    launch_grid_update(stream, P, fdt, nblocks, grid[gcur], trS, trT, gv, na, grid[gcur ^ 1], gpart, gbad);
    HIP_OK(hipGetLastError());
    std::vector<Pellet> trial = pellets;
    if (!pellets.empty()) {
        int rc = upload_pellets();
        if (rc) return rc;
        launch_couple(stream, P, dx, nblocks, grid[gcur], gv, na, d_pel, (int)pellets.size(), d_bstart, d_blist,
                      d_acc, cfg.wall_friction);
        HIP_OK(hipGetLastError());
    }
    launch_g2p2g(stream, P, fdt, (int)n, slab[cur], slab[cur ^ 1], cap, shared[sh], gv, na, grid[gcur ^ 1], d_work,
                 nwork, pmax, pdiss, pfail, pesc, nw_list[cur ^ 1], nw_count + (cur ^ 1), profiling ? ev0 : nullptr,
                 profiling ? ev1 : nullptr);
    launch_finalize(stream, ngwg, gpart, gbad, nwork, pmax, pdiss, pfail, pesc, nw_count + (cur ^ 1), d_st);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(h_st, d_st, sizeof(StepStatus), hipMemcpyDeviceToHost, stream));
    if (!pellets.empty())
        HIP_OK(hipMemcpyAsync(h_acc, d_acc, sizeof(PelletAccum) * pellets.size(), hipMemcpyDeviceToHost, stream));
    HIP_OK(hipStreamSynchronize(stream));
    if (profiling) {
        float ms = 0.0f;
        HIP_OK(hipEventElapsedTime(&ms, ev0, ev1));
        g2p2g_ms += ms;
        g2p2g_launches += 1;
    }
    const StepStatus& st = *h_st;
    double max_cells = (double)bits_float(st.max_disp_bits);
    if (report) report->max_cells = max_cells;
    int code = 0;
    if (!std::isfinite(st.traction_energy) || !std::isfinite(ledger.wall_normal_traction_energy + st.traction_energy))
        code = F_TRACTION;
    else if (st.grid_nonfinite)
        code = F_GRID;
    std::vector<double> pellet_energy(pellets.size(), 0.0);
    std::vector<WallContact> walls(pellets.size());
    if (!code) {
        V3 g = gravity(), ext = extents();
        for (size_t k = 0; k < trial.size(); ++k) {
            const PelletAccum& a = h_acc[k];
            double before = trial[k].kinetic();
            trial[k].apply_impulse(v3(a.impulse[0], a.impulse[1], a.impulse[2]),
                                   v3(a.angular[0], a.angular[1], a.angular[2]));
            pellet_energy[k] = trial[k].kinetic() - before;
            walls[k] = integrate(trial[k], dt, g, ext, cfg);
            if (!trial[k].finite_state() || !std::isfinite(pellet_energy[k])) {
                code = F_PELLET;
                break;
            }
        }
    }
    if (!code && st.fail != ~0ull) code = (int)(st.fail & 0xff);
    if (!code && max_cells > 1.0) code = F_DISPLACEMENT;
    if (code) {
        last_failure = code;
        if (report) report->failure = code;
        return code;
    }
    // commit (solver.rs:980-998 and 600-609)
    cur ^= 1;
    gcur ^= 1;
    steps_since_set += 1;
    ledger.plastic_dissipation += st.dissipation;
    for (int a = 0; a < 3; ++a) ledger.wall_impulse[a] += st.traction_impulse[a] + st.wall_impulse[a];
    ledger.wall_normal_projection_energy += st.projection;
    ledger.wall_friction_dissipation += st.friction;
    ledger.wall_normal_traction_energy += st.traction_energy;
    for (int a = 0; a < 3; ++a) ledger.gravity_impulse[a] += cfg.gravity[a] * (particle_mass * dt);
    for (size_t k = 0; k < trial.size(); ++k) {
        const PelletAccum& a = h_acc[k];
        for (int c = 0; c < 3; ++c) {
            ledger.gravity_impulse[c] += cfg.gravity[c] * (trial[k].mass * dt);
            ledger.pellet_wall_impulse[c] += walls[k].impulse[c];
            ledger.coupling_impulse[c] += a.impulse[c];
            ledger.coupling_angular_impulse[c] += a.angular[c];
        }
        ledger.pellet_wall_work += walls[k].work;
        ledger.coupling_grid_energy += a.grid_energy;
        ledger.coupling_pellet_energy += pellet_energy[k];
    }
    pellets = trial;
    max_speed = (double)bits_float(st.max_speed_bits);
    nw_host = st.nearwall;
    if (report) report->accepted = 1;
    if (st.escaped) return remove_escaped();
    return MPM_OK;
}

namespace {

double particle_mass_sum(mpm_engine* e, std::vector<float>& s, std::vector<float>& shv, std::vector<uint32_t>& id) {
    if (e->download(s, shv, id)) return std::numeric_limits<double>::quiet_NaN();
    double m = 0.0;
    for (size_t p = 0; p < e->n; ++p) m += shv[SMASS * e->n + p];
    return m;
}

void material_paste(double density, double young, double poisson, double ys, double k, double n, mpm_material* out) {
    double shear = young / (2.0 * (1.0 + poisson));
    double lame = young * poisson / ((1.0 + poisson) * (1.0 - 2.0 * poisson));
    *out = mpm_material{0, density, shear, lame, lame + 2.0 * shear / 3.0, ys, k, n, 1.0};
}

}  // namespace

namespace {

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

int mpm_create(const mpm_material* material, const mpm_config* config, const mpm_layout* layout, size_t capacity,
               size_t max_pellets, mpm_engine** out) {
    if (!material || !config || !layout || !out) return MPM_ERR_INVALID;
    *out = nullptr;
    if (max_pellets > (size_t)kMaxPellets) return MPM_ERR_CAPACITY;
    double spacing = layout->spacing;
    if (!(std::isfinite(spacing) && spacing > 0.0)) return MPM_ERR_INVALID;
    int cells[3];
    size_t total = 1;
    for (int a = 0; a < 3; ++a) {
        double ratio = layout->domain[a] / spacing;
        double rounded = std::round(ratio);
        // grid.rs:37-86
        if (!std::isfinite(ratio) || std::fabs(ratio - rounded) > 1e-6 * std::max(rounded, 1.0) || rounded < 2.0 ||
            rounded > 4000.0)
            return MPM_ERR_INVALID;
        cells[a] = (int)rounded;
        total *= (size_t)(cells[a] + 3);
    }
    if (total > layout->max_nodes) return MPM_ERR_INVALID;
    int devices = 0;
    if (hipGetDeviceCount(&devices) != hipSuccess || devices < 1) return MPM_ERR_DEVICE;
    mpm_engine* e = new mpm_engine();
    e->mat = *material;
    e->cfg = *config;
    e->dx = spacing;
    for (int a = 0; a < 3; ++a) e->cells[a] = cells[a];
    e->fill_params();
    int rc = e->alloc(capacity, max_pellets);
    if (rc) {
        delete e;
        return rc;
    }
    *out = e;
    return MPM_OK;
}

void mpm_destroy(mpm_engine* e) { delete e; }

const char* mpm_last_error(const mpm_engine* e) { return e ? e->err.c_str() : "null engine"; }

size_t mpm_particle_count(const mpm_engine* e) { return e ? e->n : 0; }
size_t mpm_pellet_count(const mpm_engine* e) { return e ? e->pellets.size() : 0; }

int mpm_set_config(mpm_engine* e, const mpm_config* c) {
    if (!e || !c) return MPM_ERR_INVALID;
    e->cfg = *c;
    e->fill_params();
    return MPM_OK;
}

int mpm_set_state(mpm_engine* e, const mpm_particles* ps, const mpm_pellet* pellets, size_t n_pellets,
                  const mpm_ledger* ledger, const mpm_history* history) {
    if (!e || !ps) return MPM_ERR_INVALID;
    if (ps->n > e->cap) return e->fail(MPM_ERR_CAPACITY, "more particles than the engine capacity");
    if (n_pellets > e->max_pellets) return e->fail(MPM_ERR_CAPACITY, "more pellets than the engine capacity");
    size_t n = ps->n;
    if (n && (!ps->position || !ps->mass || !ps->volume0))
        return e->fail(MPM_ERR_INVALID, "position, mass and volume0 are required");
    e->n = n;
    std::vector<float> s(NPF * n, 0.0f), shv(NSF * n, 0.0f);
    std::vector<uint32_t> id(n);
    double max_speed = 0.0;
    for (size_t p = 0; p < n; ++p) {
        id[p] = ps->id ? ps->id[p] : (uint32_t)p;
        for (int a = 0; a < 3; ++a) {
            s[(PX + a) * n + p] = (float)ps->position[3 * p + a];
            s[(VX + a) * n + p] = ps->velocity ? (float)ps->velocity[3 * p + a] : 0.0f;
        }
        double vs = 0.0;
        for (int a = 0; a < 3; ++a) vs += (double)s[(VX + a) * n + p] * (double)s[(VX + a) * n + p];
        max_speed = std::max(max_speed, std::sqrt(vs));
        for (int a = 0; a < 9; ++a) {
            s[(C00 + a) * n + p] = ps->affine ? (float)ps->affine[9 * p + a] : 0.0f;
            double F = ps->deformation ? ps->deformation[9 * p + a] : (a % 4 == 0 ? 1.0 : 0.0);
            s[(H00 + a) * n + p] = (float)(F - (a % 4 == 0 ? 1.0 : 0.0));
        }
        s[JM1 * n + p] = ps->volume_ratio ? (float)(ps->volume_ratio[p] - 1.0) : 0.0f;
        if (ps->kirchhoff) {
            const double* t = ps->kirchhoff + 9 * p;
            s[TXX * n + p] = (float)t[0];
            s[TYY * n + p] = (float)t[4];
            s[TZZ * n + p] = (float)t[8];
            s[TXY * n + p] = (float)(0.5 * (t[1] + t[3]));
            s[TXZ * n + p] = (float)(0.5 * (t[2] + t[6]));
            s[TYZ * n + p] = (float)(0.5 * (t[5] + t[7]));
        }
        float eps = ps->plastic_strain ? (float)ps->plastic_strain[p] : 0.0f;
        s[EPS * n + p] = eps;
        shv[SEPS * n + p] = eps;
        shv[SMASS * n + p] = (float)ps->mass[p];
        shv[SVOL0 * n + p] = (float)ps->volume0[p];
    }
    e->cur = 0;
    e->sh = 0;
    // Until the first sort the reference order is the given order.
    e->host_ref.resize(n);
    for (size_t p = 0; p < n; ++p) e->host_ref[p] = p;
    int rc = e->upload(s, shv, id);
    if (rc) return rc;
    rc = e->sort_particles(1);
    if (rc) return rc;
    e->pellets.clear();
    for (size_t k = 0; k < n_pellets; ++k) {
        const mpm_pellet& src = pellets[k];
        Pellet p;
        p.radius = src.radius;
        p.length = src.length;
        p.mass = src.mass;
        p.inertia = v3(src.inertia_body[0], src.inertia_body[1], src.inertia_body[2]);
        p.pos = v3(src.position[0], src.position[1], src.position[2]);
        p.q = Quat{src.orientation[0], src.orientation[1], src.orientation[2], src.orientation[3]};
        p.vel = v3(src.velocity[0], src.velocity[1], src.velocity[2]);
        p.L = v3(src.angular_momentum[0], src.angular_momentum[1], src.angular_momentum[2]);
        int count = std::min(std::max(src.sphere_count, 0), kMaxSpheres);
        p.offsets.assign(src.sphere_offsets, src.sphere_offsets + count);
        e->pellets.push_back(p);
    }
    e->max_speed = max_speed;
    e->particle_mass = 0.0;
    for (size_t p = 0; p < n; ++p) e->particle_mass += shv[SMASS * n + p];
    e->steps_since_set = 0;
    e->grid_valid = false;
    if (history) {
        e->time = history->time;
        e->step_count = history->step_count;
        e->limited_steps = history->limited_steps;
        e->rejected_steps = history->rejected_steps;
        e->dt_scale = history->dt_scale;
        e->min_dt_used = history->min_dt_used;
        e->max_dt_used = history->max_dt_used;
    } else {
        e->time = 0.0;
        e->step_count = e->limited_steps = e->rejected_steps = 0;
        e->dt_scale = 1.0;
        e->min_dt_used = std::numeric_limits<double>::infinity();
        e->max_dt_used = 0.0;
    }
    if (ledger) {
        e->ledger = *ledger;
    } else {
        // Simulation::new seeds the initial inventory, solver.rs:357-390.
        std::memset(&e->ledger, 0, sizeof(e->ledger));
        double mass = 0.0;
        V3 mom{{0, 0, 0}};
        for (size_t p = 0; p < n; ++p) {
            double m = shv[SMASS * n + p];
            mass += m;
            for (int a = 0; a < 3; ++a) mom[a] += (double)s[(VX + a) * n + p] * m;
        }
        for (const Pellet& p : e->pellets) mom = mom + p.vel * p.mass;
        e->ledger.initial_mass = mass;
        for (int a = 0; a < 3; ++a) e->ledger.initial_momentum[a] = mom[a];
        double k = 0, pe = 0, el = 0;
        rc = mpm_energies(e, &k, &pe, &el);
        if (rc) return rc;
        e->ledger.initial_mechanical_energy = k + pe + el;
    }
    return MPM_OK;
}

int mpm_get_state(mpm_engine* e, mpm_particles* ps, mpm_pellet* pellets, size_t pellet_capacity, mpm_ledger* ledger,
                  mpm_history* history) {
    if (!e) return MPM_ERR_INVALID;
    if (ps) {
        if (ps->n < e->n) return e->fail(MPM_ERR_CAPACITY, "particle arrays are smaller than the state");
        std::vector<float> s, shv;
        std::vector<uint32_t> id;
        int rc = e->download(s, shv, id);
        if (rc) return rc;
        size_t n = e->n;
        ps->n = n;
        bool fresh = e->steps_since_set == 0;
        // Hand the particles back in the reference's order (sorted by the key of the
        // last 32-step sort, solver.rs:501-517), whatever the device order is.
        std::vector<size_t> order(n);
        for (size_t k = 0; k < n; ++k) order[k] = k;
        std::stable_sort(order.begin(), order.end(),
                         [&](size_t a, size_t b) { return e->host_ref[a] < e->host_ref[b]; });
        std::vector<float> so(NPF * n), sho(NSF * n);
        std::vector<uint32_t> ido(n);
        for (size_t k = 0; k < n; ++k) {
            size_t q = order[k];
            for (int f = 0; f < NPF; ++f) so[f * n + k] = s[f * n + q];
            for (int f = 0; f < NSF; ++f) sho[f * n + k] = shv[f * n + q];
            ido[k] = id[q];
        }
        s.swap(so);
        shv.swap(sho);
        id.swap(ido);
        for (size_t p = 0; p < n; ++p) {
            if (ps->id) ps->id[p] = id[p];
            for (int a = 0; a < 3; ++a) {
                if (ps->position) ps->position[3 * p + a] = s[(PX + a) * n + p];
                if (ps->velocity) ps->velocity[3 * p + a] = s[(VX + a) * n + p];
            }
            for (int a = 0; a < 9; ++a) {
                if (ps->affine) ps->affine[9 * p + a] = s[(C00 + a) * n + p];
                if (ps->deformation) {
                    double id_a = a % 4 == 0 ? 1.0 : 0.0;
                    // A committed liquid step resets F to I (solver.rs:965).
                    ps->deformation[9 * p + a] = (e->mat.kind == 1 && !fresh) ? id_a : id_a + s[(H00 + a) * n + p];
                }
            }
            if (ps->kirchhoff) {
                double* t = ps->kirchhoff + 9 * p;
                t[0] = s[TXX * n + p];
                t[4] = s[TYY * n + p];
                t[8] = s[TZZ * n + p];
                t[1] = t[3] = s[TXY * n + p];
                t[2] = t[6] = s[TXZ * n + p];
                t[5] = t[7] = s[TYZ * n + p];
            }
            // A committed paste step sets J = 1 (solver.rs:950).
            if (ps->volume_ratio)
                ps->volume_ratio[p] = (e->mat.kind == 0 && !fresh) ? 1.0 : 1.0 + (double)s[JM1 * n + p];
            if (ps->plastic_strain) ps->plastic_strain[p] = e->mat.kind == 0 ? s[EPS * n + p] : shv[SEPS * n + p];
            if (ps->mass) ps->mass[p] = shv[SMASS * n + p];
            if (ps->volume0) ps->volume0[p] = shv[SVOL0 * n + p];
        }
    }
    if (pellets) {
        for (size_t k = 0; k < e->pellets.size() && k < pellet_capacity; ++k) {
            const Pellet& p = e->pellets[k];
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
            d.orientation[0] = p.q.i;
            d.orientation[1] = p.q.j;
            d.orientation[2] = p.q.k;
            d.orientation[3] = p.q.w;
            d.sphere_count = (int)p.offsets.size();
            for (size_t s = 0; s < p.offsets.size(); ++s) d.sphere_offsets[s] = p.offsets[s];
        }
    }
    if (ledger) *ledger = e->ledger;
    if (history)
        *history = mpm_history{e->time,     e->step_count,  e->limited_steps, e->rejected_steps,
                               e->dt_scale, e->min_dt_used, e->max_dt_used};
    return MPM_OK;
}

int mpm_try_step(mpm_engine* e, double dt, mpm_step_report* report) {
    if (!e) return MPM_ERR_INVALID;
    if (!(std::isfinite(dt) && dt > 0.0)) return e->fail(MPM_ERR_INVALID, "dt must be finite and positive");
    int rc = e->try_step(dt, report);
    if (rc == MPM_OK) {
        // try_step is the bare trial (solver.rs:535-537); the sort cadence belongs to advance_while.
    }
    return rc;
}

// Simulation::advance_while, solver.rs:413-472.
int mpm_advance_while(mpm_engine* e, double target, int (*keep_going)(void*), void* ctx, int* reached) {
    if (!e) return MPM_ERR_INVALID;
    if (reached) *reached = 0;
    if (!std::isfinite(target) || target < e->time) return e->fail(MPM_ERR_INVALID, "invalid target time");
    while (e->time < target) {
        if (keep_going && !keep_going(ctx)) return MPM_OK;
        double remaining = target - e->time;
        double limit = e->stability_limit();
        double scaled = e->dt_scale * limit;
        bool limited = scaled < e->cfg.max_dt;
        double dt = std::min(std::min(scaled, e->cfg.max_dt), remaining);
        if (remaining - dt < 1e-12 * std::max(target, 1.0)) dt = remaining;
        if (!std::isfinite(dt) || dt <= 0.0 || (dt < e->cfg.min_dt && remaining > e->cfg.min_dt)) {
            char buf[256];
            std::snprintf(buf, sizeof(buf), "solver aborted at t='%.17g' s: step '%.17g' s below minimum", e->time, dt);
            return e->fail(MPM_ERR_MIN_STEP, buf);
        }
        int rc = e->try_step(dt, nullptr);
        if (rc < 0) return rc;
        if (rc == MPM_OK) {
            e->time += dt;
            e->step_count += 1;
            e->limited_steps += limited ? 1 : 0;
            e->min_dt_used = std::min(e->min_dt_used, dt);
            e->max_dt_used = std::max(e->max_dt_used, dt);
            e->dt_scale = std::min(e->dt_scale * kDtRecovery, 1.0);
            if (e->step_count % kSortInterval == 0) {
                rc = e->sort_particles();
                if (rc) return rc;
            }
        } else {
            e->rejected_steps += 1;
            e->dt_scale *= 0.5;
            if (e->dt_scale * limit < e->cfg.min_dt) {
                char buf[256];
                std::snprintf(buf, sizeof(buf),
                              "solver aborted at t='%.17g' s: step '%.17g' s below minimum after '%llu' rejections; "
                              "last failure code %d",
                              e->time, dt, (unsigned long long)e->rejected_steps, rc);
                return e->fail(MPM_ERR_MIN_STEP, buf);
            }
        }
    }
    e->time = target;
    if (reached) *reached = 1;
    return MPM_OK;
}

int mpm_advance_to(mpm_engine* e, double target) {
    int reached = 0;
    return mpm_advance_while(e, target, nullptr, nullptr, &reached);
}

double mpm_stability_limit(mpm_engine* e) { return e ? e->stability_limit() : std::numeric_limits<double>::quiet_NaN(); }

double mpm_mass_residual(mpm_engine* e) {
    if (!e) return std::numeric_limits<double>::quiet_NaN();
    std::vector<float> s, shv;
    std::vector<uint32_t> id;
    double m = particle_mass_sum(e, s, shv, id);
    return m + e->ledger.outflow_mass - e->ledger.initial_mass;
}

int mpm_momentum_residual(mpm_engine* e, double* out) {
    if (!e || !out) return MPM_ERR_INVALID;
    std::vector<float> s, shv;
    std::vector<uint32_t> id;
    int rc = e->download(s, shv, id);
    if (rc) return rc;
    size_t n = e->n;
    V3 p{{0, 0, 0}};
    for (size_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) p[a] += (double)s[(VX + a) * n + i] * (double)shv[SMASS * n + i];
    for (const Pellet& pl : e->pellets) p = p + pl.vel * pl.mass;
    const mpm_ledger& l = e->ledger;
    for (int a = 0; a < 3; ++a)
        out[a] = p[a] - l.initial_momentum[a] - l.gravity_impulse[a] + l.outflow_momentum[a] - l.wall_impulse[a] -
                 l.pellet_wall_impulse[a];
    return MPM_OK;
}

// Kinetic, potential and elastic energy (solver.rs:1027-1076) from the f32 state.
int mpm_energies(mpm_engine* e, double* kinetic, double* potential, double* elastic) {
    if (!e) return MPM_ERR_INVALID;
    std::vector<float> s, shv;
    std::vector<uint32_t> id;
    int rc = e->download(s, shv, id);
    if (rc) return rc;
    size_t n = e->n;
    double ke = 0.0, pe = 0.0, el = 0.0;
    V3 g = e->gravity();
    bool fresh = e->steps_since_set == 0;
    for (size_t p = 0; p < n; ++p) {
        double m = shv[SMASS * n + p], v0 = shv[SVOL0 * n + p];
        double v2 = 0.0, gx = 0.0;
        for (int a = 0; a < 3; ++a) {
            double v = s[(VX + a) * n + p];
            v2 += v * v;
            gx += g[a] * (double)s[(PX + a) * n + p];
        }
        ke += 0.5 * m * v2;
        pe += -m * gx;
        if (e->mat.kind == 0) {
            double F[9], b[9], vals[3];
            for (int a = 0; a < 9; ++a) F[a] = (a % 4 == 0 ? 1.0 : 0.0) + (double)s[(H00 + a) * n + p];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) b[3 * i + j] = F[3 * i] * F[3 * j] + F[3 * i + 1] * F[3 * j + 1] + F[3 * i + 2] * F[3 * j + 2];
            if (!eigen3(b, vals)) {
                el = std::numeric_limits<double>::quiet_NaN();
                continue;
            }
            double h[3], tr = 0.0, h2 = 0.0;
            for (int a = 0; a < 3; ++a) {
                h[a] = 0.5 * std::log(std::max(vals[a], 1e-300));
                tr += h[a];
                h2 += h[a] * h[a];
            }
            el += v0 * (e->mat.shear_modulus * h2 + 0.5 * e->mat.lame_lambda * tr * tr);
        } else {
            double j = 1.0 + (double)s[JM1 * n + p];
            double gm = e->mat.tait_gamma;
            if (gm == 1.0)
                el += v0 * e->mat.bulk_modulus * (j - 1.0 - std::log(j));
            else
                el += v0 * (e->mat.bulk_modulus / gm) * (j - 1.0 + (std::pow(j, 1.0 - gm) - 1.0) / (gm - 1.0));
        }
    }
    (void)fresh;
    for (const Pellet& p : e->pellets) {
        ke += p.kinetic();
        pe -= p.mass * dot(g, p.pos);
    }
    if (kinetic) *kinetic = ke;
    if (potential) *potential = pe;
    if (elastic) *elastic = el;
    return MPM_OK;
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
    e->profiling = on != 0;
    e->g2p2g_ms = 0.0;
    e->g2p2g_launches = 0;
}

int mpm_profile_g2p2g(mpm_engine* e, double* avg_ms, uint64_t* launches) {
    if (!e) return MPM_ERR_INVALID;
    if (avg_ms) *avg_ms = e->g2p2g_launches ? e->g2p2g_ms / (double)e->g2p2g_launches : 0.0;
    if (launches) *launches = e->g2p2g_launches;
    return MPM_OK;
}

}  // extern "C"
