// CPU f64 oracle for the MLS-MPM substep -- TEST INFRASTRUCTURE ONLY.
//
// A restatement of the reference crate's CPU path, kept in its operation order
// (left-to-right sums, nalgebra's column-accumulating products, the serial
// first-touch active list) so that it reproduces the reference's published
// numbers.  Each function names the reference lines it follows.  Only the test
// suite, the smoke check and the benchmark's CPU-baseline leg use it; the HIP
// engine never links it.  Build with -ffp-contract=off: the reference is Rust,
// which never fuses multiply-adds.
#include "mpm_oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <numeric>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using V3 = std::array<double, 3>;
using M3 = std::array<double, 9>;  // row-major

int g_threads = 1;

inline double at(const M3& m, int i, int j) { return m[3 * i + j]; }
inline V3 v3(double a, double b, double c) { return V3{a, b, c}; }
inline V3 add(const V3& a, const V3& b) { return v3(a[0] + b[0], a[1] + b[1], a[2] + b[2]); }
inline V3 sub(const V3& a, const V3& b) { return v3(a[0] - b[0], a[1] - b[1], a[2] - b[2]); }
inline V3 scale(const V3& a, double s) { return v3(a[0] * s, a[1] * s, a[2] * s); }
inline double dot(const V3& a, const V3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline double norm2(const V3& a) { return a[0] * a[0] + a[1] * a[1] + a[2] * a[2]; }
inline double norm(const V3& a) { return std::sqrt(norm2(a)); }
inline V3 cross(const V3& a, const V3& b) {
    return v3(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]);
}
inline bool finite3(const V3& a) {
    return std::isfinite(a[0]) && std::isfinite(a[1]) && std::isfinite(a[2]);
}
inline bool finite9(const M3& a) {
    for (double x : a)
        if (!std::isfinite(x)) return false;
    return true;
}
M3 identity() { return M3{1, 0, 0, 0, 1, 0, 0, 0, 1}; }
M3 zeros9() { return M3{0, 0, 0, 0, 0, 0, 0, 0, 0}; }
M3 madd(const M3& a, const M3& b) {
    M3 r;
    for (int i = 0; i < 9; ++i) r[i] = a[i] + b[i];
    return r;
}
M3 msub(const M3& a, const M3& b) {
    M3 r;
    for (int i = 0; i < 9; ++i) r[i] = a[i] - b[i];
    return r;
}
M3 mscale(const M3& a, double s) {
    M3 r;
    for (int i = 0; i < 9; ++i) r[i] = a[i] * s;
    return r;
}
M3 transpose(const M3& a) {
    return M3{a[0], a[3], a[6], a[1], a[4], a[7], a[2], a[5], a[8]};
}
// nalgebra gemv accumulates columns: r_i = (A_i0 x0 + A_i1 x1) + A_i2 x2.
V3 mv(const M3& a, const V3& x) {
    V3 r;
    for (int i = 0; i < 3; ++i) r[i] = (at(a, i, 0) * x[0] + at(a, i, 1) * x[1]) + at(a, i, 2) * x[2];
    return r;
}
M3 mm(const M3& a, const M3& b) {
    M3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r[3 * i + j] = (at(a, i, 0) * at(b, 0, j) + at(a, i, 1) * at(b, 1, j)) + at(a, i, 2) * at(b, 2, j);
    return r;
}
M3 diag(const V3& d) { return M3{d[0], 0, 0, 0, d[1], 0, 0, 0, d[2]}; }
double trace(const M3& a) { return a[0] + a[4] + a[8]; }
double frob2(const M3& a) {
    // nalgebra norm_squared iterates column-major storage.
    double s = 0.0;
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) s += at(a, i, j) * at(a, i, j);
    return s;
}
// nalgebra Matrix3::determinant (cofactor expansion along the first row).
double det3(const M3& m) {
    double m11 = m[0], m12 = m[1], m13 = m[2];
    double m21 = m[3], m22 = m[4], m23 = m[5];
    double m31 = m[6], m32 = m[7], m33 = m[8];
    double minor_m12_m23 = m22 * m33 - m32 * m23;
    double minor_m11_m23 = m21 * m33 - m31 * m23;
    double minor_m11_m22 = m21 * m32 - m31 * m22;
    return m11 * minor_m12_m23 - m12 * minor_m11_m23 + m13 * minor_m11_m22;
}

// ---------------------------------------------------------------- material

// constitutive.rs:227-243
double wave_speed(const oracle_material& m) {
    double modulus = m.kind == 0 ? m.lame_lambda + 2.0 * m.shear_modulus : m.bulk_modulus;
    return std::sqrt(modulus / m.density);
}
double diffusivity(const oracle_material& m) { return m.kind == 0 ? 0.0 : m.consistency / m.density; }

// constitutive.rs:593-595, with the Tait generalisation (gamma == 1 is the reference).
double liquid_pressure(const oracle_material& m, double j) {
    if (m.tait_gamma == 1.0) return m.bulk_modulus * (1.0 / j - 1.0);
    return m.bulk_modulus / m.tait_gamma * (std::pow(j, -m.tait_gamma) - 1.0);
}

// ------------------------------------------------------------ constitutive

struct ReturnSolution {
    double stress, increment;
    int iterations;
};

// constitutive.rs:280-362
int hb_return(double trial, double ys, double k, double n, double mu, double dt, ReturnSolution& out) {
    double inputs[6] = {trial, ys, k, n, mu, dt};
    for (double v : inputs)
        if (!std::isfinite(v)) return ORACLE_FAIL_NONFINITE_CONST;
    double viscous_scale = 3.0 * mu * dt;
    if (trial <= ys) {
        out = {trial, 0.0, 0};
        return 0;
    }
    if (k <= 0.0) {
        out = {ys, (trial - ys) / (3.0 * mu), 0};
        return 0;
    }
    double exponent = 1.0 / n;
    auto residual = [&](double sigma) {
        double over = std::max((sigma - ys) / k, 0.0);
        return sigma - trial + viscous_scale * std::pow(over, exponent);
    };
    double lower = ys, upper = trial;
    double sigma = (lower + upper) / 2.0;
    double tolerance = 1e-12 * std::max(trial, 1.0);
    double last = residual(sigma);
    for (int iteration = 1; iteration <= 200; ++iteration) {
        if (std::fabs(last) <= tolerance || (upper - lower) <= tolerance) {
            out = {sigma, (trial - sigma) / (3.0 * mu), iteration};
            return 0;
        }
        if (last > 0.0)
            upper = sigma;
        else
            lower = sigma;
        double over = std::max((sigma - ys) / k, 0.0);
        double derivative = over > 0.0
                                ? 1.0 + viscous_scale * exponent * std::pow(over, exponent - 1.0) / k
                                : std::numeric_limits<double>::infinity();
        double newton = sigma - last / derivative;
        sigma = (std::isfinite(newton) && newton > lower && newton < upper) ? newton : (lower + upper) / 2.0;
        last = residual(sigma);
        if (!std::isfinite(last)) return ORACLE_FAIL_NONFINITE_CONST;
    }
    return ORACLE_FAIL_RETURN;
}

// constitutive.rs:373-425 (cyclic Jacobi, pairs (0,1),(0,2),(1,2)).
bool eigen3(const M3& matrix, V3& values, M3& vectors) {
    if (!finite9(matrix)) return false;
    M3 mt = transpose(matrix), d;
    for (int i = 0; i < 9; ++i) d[i] = 0.5 * (matrix[i] + mt[i]);
    vectors = identity();
    double scale = std::max(frob2(d), std::numeric_limits<double>::min());
    for (int sweep = 0; sweep < 64; ++sweep) {
        double a01 = d[1], a02 = d[2], a12 = d[5];
        double off = a01 * a01 + a02 * a02 + a12 * a12;
        if (off <= 1e-30 * scale) {
            values = v3(d[0], d[4], d[8]);
            return true;
        }
        const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
        for (auto& pq : pairs) {
            int p = pq[0], q = pq[1];
            double apq = d[3 * p + q];
            if (apq == 0.0) continue;
            double theta = (d[3 * q + q] - d[3 * p + p]) / (2.0 * apq);
            double tangent = std::copysign(1.0, theta) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
            double cosine = 1.0 / std::sqrt(tangent * tangent + 1.0);
            double sine = tangent * cosine;
            for (int k = 0; k < 3; ++k) {
                double akp = d[3 * k + p], akq = d[3 * k + q];
                d[3 * k + p] = cosine * akp - sine * akq;
                d[3 * k + q] = sine * akp + cosine * akq;
            }
            for (int k = 0; k < 3; ++k) {
                double apk = d[3 * p + k], aqk = d[3 * q + k];
                d[3 * p + k] = cosine * apk - sine * aqk;
                d[3 * q + k] = sine * apk + cosine * aqk;
            }
            for (int k = 0; k < 3; ++k) {
                double vkp = vectors[3 * k + p], vkq = vectors[3 * k + q];
                vectors[3 * k + p] = cosine * vkp - sine * vkq;
                vectors[3 * k + q] = sine * vkp + cosine * vkq;
            }
        }
    }
    return false;
}

struct PasteUpdate {
    M3 deformation, kirchhoff;
    double increment, dissipation_density, equivalent_stress;
};

// constitutive.rs:456-531
int update_paste(const M3& trial, const oracle_material& m, double dt, PasteUpdate& out) {
    double determinant = det3(trial);
    if (!std::isfinite(determinant)) return ORACLE_FAIL_NONFINITE_CONST;
    if (determinant <= 0.0) return ORACLE_FAIL_INVERTED;
    M3 b = mm(trial, transpose(trial));
    V3 eig;
    M3 u;
    if (!eigen3(b, eig, u)) return ORACLE_FAIL_EIGEN;
    V3 stretches, hencky;
    for (int a = 0; a < 3; ++a) {
        double s = std::sqrt(std::max(eig[a], 0.0));
        if (!std::isfinite(s) || s < 1e-6) return ORACLE_FAIL_STRETCH;
        stretches[a] = s;
        hencky[a] = std::log(s);
    }
    double volumetric = hencky[0] + hencky[1] + hencky[2];
    double mean = volumetric / 3.0;
    V3 dev = v3(hencky[0] - mean, hencky[1] - mean, hencky[2] - mean);
    double mu = m.shear_modulus;
    double dev_norm = norm(dev);
    double trial_eq = std::sqrt(1.5) * 2.0 * mu * dev_norm;
    double sc = 1.0, inc = 0.0, eq = trial_eq;
    if (trial_eq > m.yield_stress && dev_norm > 0.0) {
        ReturnSolution sol;
        int rc = hb_return(trial_eq, m.yield_stress, m.consistency, m.flow_index, mu, dt, sol);
        if (rc) return rc;
        sc = sol.stress / trial_eq;
        inc = sol.increment;
        eq = sol.stress;
    }
    V3 dev_new = scale(dev, sc);
    V3 hencky_new = v3(dev_new[0] + mean, dev_new[1] + mean, dev_new[2] + mean);
    V3 ratio = v3(std::exp(hencky_new[0]) / stretches[0], std::exp(hencky_new[1]) / stretches[1],
                  std::exp(hencky_new[2]) / stretches[2]);
    out.deformation = mm(mm(mm(u, diag(ratio)), transpose(u)), trial);
    V3 principal = add(scale(hencky_new, 2.0 * mu), v3(m.lame_lambda * volumetric, m.lame_lambda * volumetric,
                                                       m.lame_lambda * volumetric));
    out.kirchhoff = mm(mm(u, diag(principal)), transpose(u));
    if (!finite9(out.kirchhoff)) return ORACLE_FAIL_NONFINITE_CONST;
    out.increment = inc;
    out.dissipation_density = eq * inc;
    out.equivalent_stress = eq;
    return 0;
}

// constitutive.rs:557-589 (Tait pressure when gamma != 1).
int update_liquid(double j, const M3& c, const oracle_material& m, double dt, double& new_ratio, M3& kirchhoff,
                  double& pressure) {
    new_ratio = j * (1.0 + dt * trace(c));
    if (!std::isfinite(new_ratio)) return ORACLE_FAIL_NONFINITE_CONST;
    if (new_ratio <= 1e-6) return ORACLE_FAIL_INVERTED;
    pressure = liquid_pressure(m, new_ratio);
    M3 ct = transpose(c), rate;
    for (int i = 0; i < 9; ++i) rate[i] = 0.5 * (c[i] + ct[i]);
    M3 dev = msub(rate, mscale(identity(), trace(rate) / 3.0));
    kirchhoff = madd(mscale(identity(), -pressure * new_ratio), mscale(dev, 2.0 * m.consistency * new_ratio));
    if (!finite9(kirchhoff)) return ORACLE_FAIL_NONFINITE_CONST;
    return 0;
}

// ------------------------------------------------------------------ grid

struct Layout {
    double spacing;
    size_t cells[3], nodes[3];
    size_t count() const { return nodes[0] * nodes[1] * nodes[2]; }
    size_t index(size_t i, size_t j, size_t k) const { return (i * nodes[1] + j) * nodes[2] + k; }
    V3 extents() const {
        return v3((double)cells[0] * spacing, (double)cells[1] * spacing, (double)cells[2] * spacing);
    }
    V3 node_position(size_t i, size_t j, size_t k) const {
        return v3(((double)i - 1.0) * spacing, ((double)j - 1.0) * spacing, ((double)k - 1.0) * spacing);
    }
    void coords(size_t index, size_t& i, size_t& j, size_t& k) const {
        k = index % nodes[2];
        size_t rest = index / nodes[2];
        j = rest % nodes[1];
        i = rest / nodes[1];
    }
    // grid.rs:135-143
    int wall_side(int axis, size_t coord) const {
        if (coord <= 1) return 1;
        if (coord > cells[axis]) return -1;
        return 0;
    }
};

// grid.rs:37-86
bool make_layout(const double* domain, double spacing, size_t max_nodes, Layout& out, std::string& err) {
    if (!(std::isfinite(spacing) && spacing > 0.0)) {
        err = "`grid_spacing_m` must be positive and finite";
        return false;
    }
    for (int a = 0; a < 3; ++a) {
        double extent = domain[a];
        if (!(std::isfinite(extent) && extent > 0.0)) {
            err = "`domain_m` must be positive and finite";
            return false;
        }
        double ratio = extent / spacing;
        double rounded = std::round(ratio);
        double cap = (double)std::min<size_t>(max_nodes > 3 ? max_nodes - 3 : 0, 0xffffffffu);
        if (!std::isfinite(ratio) || rounded > cap) {
            err = "grid axis exceeds the node allocation cap";
            return false;
        }
        if (std::fabs(ratio - rounded) > 1e-6 * std::max(rounded, 1.0)) {
            err = "`domain_m` is not an integer multiple of `grid_spacing_m`";
            return false;
        }
        if (rounded < 2.0) {
            err = "`domain_m` must span at least two cells";
            return false;
        }
        out.cells[a] = (size_t)rounded;
    }
    out.spacing = spacing;
    for (int a = 0; a < 3; ++a) out.nodes[a] = out.cells[a] + 3;
    if (out.count() > max_nodes) {
        err = "grid would allocate more nodes than the cap";
        return false;
    }
    return true;
}

struct Stencil {
    size_t base[3];
    double w[3][3];
    double off[3][3];
    double weight(int a, int b, int c) const { return w[0][a] * w[1][b] * w[2][c]; }
    V3 distance(int a, int b, int c) const { return v3(off[0][a], off[1][b], off[2][c]); }
};

// grid.rs:151-183
bool stencil(const Layout& l, const V3& x, Stencil& s) {
    for (int a = 0; a < 3; ++a) {
        double scaled = x[a] / l.spacing;
        double base_f = std::floor(scaled - 0.5);
        double padded = base_f + 1.0;
        if (std::isnan(padded) || padded < 0.0 || padded + 2.0 >= (double)l.nodes[a]) return false;
        double fx = scaled - base_f;
        double t0 = 1.5 - fx, t1 = fx - 1.0, t2 = fx - 0.5;
        s.w[a][0] = 0.5 * (t0 * t0);
        s.w[a][1] = 0.75 - t1 * t1;
        s.w[a][2] = 0.5 * (t2 * t2);
        s.base[a] = (size_t)padded;
        for (int o = 0; o < 3; ++o) s.off[a][o] = (fx - (double)o) * -l.spacing;
    }
    return true;
}

struct Grid {
    Layout layout;
    std::vector<double> mass;
    std::vector<V3> momentum, traction;
    std::vector<size_t> active;
    explicit Grid(const Layout& l) : layout(l), mass(l.count(), 0.0), momentum(l.count(), V3{0, 0, 0}),
                                     traction(l.count(), V3{0, 0, 0}) {}
    // grid.rs:243-250
    void clear() {
        for (size_t i : active) {
            mass[i] = 0.0;
            momentum[i] = V3{0, 0, 0};
            traction[i] = V3{0, 0, 0};
        }
        active.clear();
    }
    // grid.rs:254-263
    void deposit(size_t index, double m, const V3& p) {
        if (m == 0.0) return;
        if (mass[index] == 0.0) active.push_back(index);
        mass[index] += m;
        momentum[index] = add(momentum[index], p);
    }
    // grid.rs:279-295
    bool deposit_reaction(size_t index, int axis, double inward, double inward_impulse, double& jump) {
        double m = mass[index];
        if (m <= 0.0) return false;
        double impulse = inward_impulse * inward;
        jump = impulse * (momentum[index][axis] / m + impulse / (2.0 * m));
        momentum[index][axis] += impulse;
        traction[index][axis] += inward_impulse;
        return true;
    }
};

// ---------------------------------------------------------------- pellet

struct Quat {
    double i, j, k, w;
};
Quat qmul(const Quat& a, const Quat& b) {
    return Quat{a.w * b.i + a.i * b.w + a.j * b.k - a.k * b.j, a.w * b.j - a.i * b.k + a.j * b.w + a.k * b.i,
                a.w * b.k + a.i * b.j - a.j * b.i + a.k * b.w, a.w * b.w - a.i * b.i - a.j * b.j - a.k * b.k};
}
V3 qrot(const Quat& q, const V3& v) {
    V3 qv = v3(q.i, q.j, q.k);
    V3 t = scale(cross(qv, v), 2.0);
    V3 c = cross(qv, t);
    return add(add(scale(t, q.w), c), v);
}
M3 qmatrix(const Quat& q) {
    double i = q.i, j = q.j, k = q.k, w = q.w;
    double ww = w * w, ii = i * i, jj = j * j, kk = k * k;
    double ij = i * j * 2.0, wk = w * k * 2.0, wj = w * j * 2.0, ik = i * k * 2.0, jk = j * k * 2.0,
           wi = w * i * 2.0;
    return M3{ww + ii - jj - kk, ij - wk, wj + ik, wk + ij, ww - ii + jj - kk, jk - wi, ik - wj, wi + jk,
              ww - ii - jj + kk};
}
// nalgebra UnitQuaternion::from_scaled_axis = exp of the pure quaternion a/2.
Quat from_scaled_axis(const V3& a) {
    V3 half = v3(a[0] / 2.0, a[1] / 2.0, a[2] / 2.0);
    double nn = norm2(half);
    const double eps = std::numeric_limits<double>::epsilon();
    if (nn <= eps * eps) return Quat{0, 0, 0, 1};
    double n = std::sqrt(nn);
    double f = 1.0 * std::sin(n) / n;
    return Quat{half[0] * f, half[1] * f, half[2] * f, 1.0 * std::cos(n)};
}

struct Pellet {
    double radius, length, mass;
    V3 inertia_body, position;
    Quat orientation;
    V3 velocity, angular_momentum;
    std::vector<double> offsets;

    V3 axis() const { return qrot(orientation, v3(1, 0, 0)); }
    // rigid.rs:193-198
    V3 angular_velocity() const {
        M3 r = qmatrix(orientation);
        V3 body = mv(transpose(r), angular_momentum);
        V3 bo = v3(body[0] / inertia_body[0], body[1] / inertia_body[1], body[2] / inertia_body[2]);
        return mv(r, bo);
    }
    // rigid.rs:212-230
    double signed_distance(const V3& point, V3& normal) const {
        V3 ax = axis();
        double best = std::numeric_limits<double>::infinity();
        normal = v3(0, 0, 1);
        for (double o : offsets) {
            V3 center = add(position, scale(ax, o));
            V3 delta = sub(point, center);
            double distance = norm(delta);
            double phi = distance - radius;
            if (phi < best) {
                normal = distance > 1e-12 ? v3(delta[0] / distance, delta[1] / distance, delta[2] / distance)
                                          : v3(0, 0, 1);
                best = phi;
            }
        }
        return best;
    }
    V3 point_velocity(const V3& point) const {
        return add(velocity, cross(angular_velocity(), sub(point, position)));
    }
    // rigid.rs:240-248
    void bounding_box(V3& lo, V3& hi) const {
        V3 ax = axis();
        double half = 0.0;
        for (double o : offsets) half = std::max(half, std::fabs(o));
        V3 reach = v3(std::fabs(ax[0]) * half + radius, std::fabs(ax[1]) * half + radius,
                      std::fabs(ax[2]) * half + radius);
        lo = sub(position, reach);
        hi = add(position, reach);
    }
    void apply_impulse(const V3& impulse, const V3& angular) {
        velocity = add(velocity, v3(impulse[0] / mass, impulse[1] / mass, impulse[2] / mass));
        angular_momentum = add(angular_momentum, angular);
    }
    double kinetic_energy() const {
        V3 omega = angular_velocity();
        return 0.5 * mass * norm2(velocity) + 0.5 * dot(omega, angular_momentum);
    }
    bool is_finite() const {
        return finite3(position) && finite3(velocity) && finite3(angular_momentum) &&
               std::isfinite(orientation.i) && std::isfinite(orientation.j) && std::isfinite(orientation.k) &&
               std::isfinite(orientation.w);
    }
};

struct WallContact {
    V3 impulse{0, 0, 0};
    double work = 0.0;
    int contacts = 0;
};

// rigid.rs:81-97
double contact_damping(const oracle_config& c, double mass) {
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
double contact_stable_dt(const oracle_config& c, double mass) { return 0.2 * std::sqrt(mass / c.contact_stiffness); }

// rigid.rs:279-334
WallContact pellet_integrate(Pellet& p, double dt, const V3& gravity, const V3& extents, const oracle_config& c) {
    double damping = contact_damping(c, p.mass);
    V3 force{0, 0, 0}, torque{0, 0, 0};
    WallContact ledger;
    V3 omega = p.angular_velocity();
    V3 ax = p.axis();
    for (double o : p.offsets) {
        V3 center = add(p.position, scale(ax, o));
        for (int axis = 0; axis < 3; ++axis) {
            for (int side = 0; side < 2; ++side) {
                double face = side == 0 ? 0.0 : extents[axis];
                double sign = side == 0 ? 1.0 : -1.0;
                double gap = (center[axis] - face) * sign;
                double penetration = p.radius - gap;
                if (penetration <= 0.0) continue;
                V3 normal{0, 0, 0};
                normal[axis] = sign;
                V3 contact_point = sub(center, scale(normal, p.radius));
                V3 lever = sub(contact_point, p.position);
                V3 pv = add(p.velocity, cross(omega, lever));
                double normal_speed = dot(pv, normal);
                double normal_force = std::max(c.contact_stiffness * penetration - damping * normal_speed, 0.0);
                V3 tangential = sub(pv, scale(normal, normal_speed));
                double ts = norm(tangential);
                V3 friction{0, 0, 0};
                if (ts > 1e-12) {
                    double limit = std::min(c.contact_friction * normal_force, p.mass * ts / dt);
                    friction = scale(v3(-tangential[0] / ts, -tangential[1] / ts, -tangential[2] / ts), limit);
                }
                V3 total = add(scale(normal, normal_force), friction);
                force = add(force, total);
                torque = add(torque, cross(lever, total));
                ledger.work += dot(total, pv) * dt;
                ledger.contacts += 1;
            }
        }
    }
    V3 impulse = scale(force, dt);
    ledger.impulse = impulse;
    p.velocity = add(p.velocity, add(scale(gravity, dt), v3(impulse[0] / p.mass, impulse[1] / p.mass,
                                                            impulse[2] / p.mass)));
    p.angular_momentum = add(p.angular_momentum, scale(torque, dt));
    p.position = add(p.position, scale(p.velocity, dt));
    V3 omega_new = p.angular_velocity();
    Quat inc = from_scaled_axis(scale(omega_new, dt));
    p.orientation = qmul(inc, p.orientation);
    return ledger;
}

struct CouplingResult {
    V3 impulse{0, 0, 0}, angular_impulse{0, 0, 0};
    double grid_energy = 0.0;
    size_t constrained = 0;
};

// rigid.rs:385-438
CouplingResult couple_grid(Grid& grid, const Pellet& pellet, double friction) {
    V3 lo, hi;
    pellet.bounding_box(lo, hi);
    const Layout& l = grid.layout;
    CouplingResult result;
    for (size_t index : grid.active) {
        double m = grid.mass[index];
        if (m <= 0.0) continue;
        size_t i, j, k;
        l.coords(index, i, j, k);
        V3 node = l.node_position(i, j, k);
        bool outside = false;
        for (int a = 0; a < 3; ++a)
            if (node[a] < lo[a] || node[a] > hi[a]) outside = true;
        if (outside) continue;
        V3 normal;
        double phi = pellet.signed_distance(node, normal);
        if (phi > 0.0) continue;
        V3 velocity = grid.momentum[index];
        V3 body = pellet.point_velocity(node);
        V3 relative = sub(velocity, body);
        double ns = dot(relative, normal);
        if (ns >= 0.0) continue;
        V3 tangential = sub(relative, scale(normal, ns));
        double ts = norm(tangential);
        double reduction = friction * (-ns);
        V3 tnew = (ts <= reduction || ts < 1e-14) ? V3{0, 0, 0} : scale(tangential, (ts - reduction) / ts);
        V3 nv = add(body, tnew);
        V3 dp = scale(sub(nv, velocity), m);
        double energy = 0.5 * m * (norm2(nv) - norm2(velocity));
        grid.momentum[index] = nv;
        result.impulse = sub(result.impulse, dp);
        result.angular_impulse = sub(result.angular_impulse, cross(sub(node, pellet.position), dp));
        result.grid_energy += energy;
        result.constrained += 1;
    }
    return result;
}

// ------------------------------------------------------------- particles

struct Particles {
    std::vector<uint32_t> id;
    std::vector<V3> pos, vel;
    std::vector<M3> affine, deformation, kirchhoff;
    std::vector<double> volume_ratio, plastic, mass, volume0;
    size_t size() const { return id.size(); }
    void push(const V3& x, const V3& v, double m, double v0) {
        id.push_back((uint32_t)id.size());
        pos.push_back(x);
        vel.push_back(v);
        affine.push_back(zeros9());
        deformation.push_back(identity());
        volume_ratio.push_back(1.0);
        kirchhoff.push_back(zeros9());
        plastic.push_back(0.0);
        mass.push_back(m);
        volume0.push_back(v0);
    }
    double total_mass() const {
        double s = 0.0;
        for (double m : mass) s += m;
        return s;
    }
    V3 linear_momentum() const {
        V3 acc{0, 0, 0};
        for (size_t p = 0; p < size(); ++p) acc = add(acc, scale(vel[p], mass[p]));
        return acc;
    }
    double kinetic_energy() const {
        double s = 0.0;
        for (size_t p = 0; p < size(); ++p) s += 0.5 * mass[p] * norm2(vel[p]);
        return s;
    }
    template <class T>
    static void apply(std::vector<T>& v, const std::vector<size_t>& order) {
        std::vector<T> r;
        r.reserve(order.size());
        for (size_t i : order) r.push_back(v[i]);
        v.swap(r);
    }
    // particles.rs:123-138
    void permute(const std::vector<size_t>& order) {
        apply(id, order);
        apply(pos, order);
        apply(vel, order);
        apply(affine, order);
        apply(deformation, order);
        apply(volume_ratio, order);
        apply(kirchhoff, order);
        apply(plastic, order);
        apply(mass, order);
        apply(volume0, order);
    }
    // particles.rs:142-152
    double remove_sorted(const std::vector<size_t>& indices) {
        if (indices.empty()) return 0.0;
        double removed = 0.0;
        for (size_t i : indices) removed += mass[i];
        std::vector<size_t> keep;
        size_t next = 0;
        for (size_t i = 0; i < size(); ++i) {
            if (next < indices.size() && indices[next] == i) {
                ++next;
                continue;
            }
            keep.push_back(i);
        }
        permute(keep);
        return removed;
    }
};

struct Trial {
    V3 position, velocity;
    M3 affine, deformation, kirchhoff;
    double volume_ratio, increment, dissipation, displacement;
};

struct NodeTrial {
    V3 velocity{0, 0, 0}, wall_impulse{0, 0, 0};
    double projection = 0.0, friction = 0.0;
};

// solver.rs:1145-1155
V3 reduce_tangential(const V3& t, double reduction) {
    if (reduction == 0.0) return t;
    double speed = norm(t);
    if (speed <= reduction) return V3{0, 0, 0};
    return scale(t, (speed - reduction) / speed);
}

}  // namespace

struct oracle_sim {
    oracle_material material;
    Particles particles;
    Grid grid;
    std::vector<Pellet> pellets;
    oracle_config config;
    oracle_ledger ledger;
    double time = 0.0;
    uint64_t step_count = 0, limited_steps = 0, rejected_steps = 0;
    double dt_scale = 1.0, min_dt_used = std::numeric_limits<double>::infinity(), max_dt_used = 0.0;
    int last_failure = 0;
    explicit oracle_sim(const Layout& l) : grid(l) {}
};

namespace {

double potential_energy(const oracle_sim& s) {
    const V3 g = v3(s.config.gravity[0], s.config.gravity[1], s.config.gravity[2]);
    double e = 0.0;
    for (size_t p = 0; p < s.particles.size(); ++p) e += -s.particles.mass[p] * dot(g, s.particles.pos[p]);
    for (const Pellet& pl : s.pellets) e -= pl.mass * dot(g, pl.position);
    return e;
}

// solver.rs:1046-1070 (Tait stored energy when gamma != 1).
double elastic_energy(const oracle_sim& s) {
    const oracle_material& m = s.material;
    double total = 0.0;
    for (size_t p = 0; p < s.particles.size(); ++p) {
        double v0 = s.particles.volume0[p];
        double e;
        if (m.kind == 0) {
            const M3& f = s.particles.deformation[p];
            M3 b = mm(f, transpose(f));
            V3 eig;
            M3 u;
            if (!eigen3(b, eig, u)) {
                e = std::numeric_limits<double>::quiet_NaN();
            } else {
                V3 h = v3(0.5 * std::log(std::max(eig[0], 1e-300)), 0.5 * std::log(std::max(eig[1], 1e-300)),
                          0.5 * std::log(std::max(eig[2], 1e-300)));
                double tr = h[0] + h[1] + h[2];
                e = v0 * (m.shear_modulus * norm2(h) + 0.5 * m.lame_lambda * tr * tr);
            }
        } else {
            double j = s.particles.volume_ratio[p];
            if (m.tait_gamma == 1.0) {
                e = v0 * m.bulk_modulus * (j - 1.0 - std::log(j));
            } else {
                double g = m.tait_gamma;
                e = v0 * (m.bulk_modulus / g) * (j - 1.0 + (std::pow(j, 1.0 - g) - 1.0) / (g - 1.0));
            }
        }
        total += e;
    }
    return total;
}

double kinetic_energy(const oracle_sim& s) {
    double e = s.particles.kinetic_energy();
    for (const Pellet& p : s.pellets) e += p.kinetic_energy();
    return e;
}

double mechanical_energy(const oracle_sim& s) { return kinetic_energy(s) + potential_energy(s) + elastic_energy(s); }

V3 total_momentum(const oracle_sim& s) {
    V3 m = s.particles.linear_momentum();
    for (const Pellet& p : s.pellets) m = add(m, scale(p.velocity, p.mass));
    return m;
}

double ledgered_energy(const oracle_ledger& l) {
    return l.wall_normal_projection_energy - l.wall_friction_dissipation + l.wall_normal_traction_energy +
           l.pellet_wall_work + l.coupling_grid_energy + l.coupling_pellet_energy - l.plastic_dissipation;
}

void seed_ledger(oracle_sim& s) {
    std::memset(&s.ledger, 0, sizeof(s.ledger));
    V3 momentum = s.particles.linear_momentum();
    for (const Pellet& p : s.pellets) momentum = add(momentum, scale(p.velocity, p.mass));
    s.ledger.initial_mass = s.particles.total_mass();
    for (int a = 0; a < 3; ++a) s.ledger.initial_momentum[a] = momentum[a];
    s.ledger.initial_mechanical_energy = mechanical_energy(s);
}

double face_friction(const oracle_config& c, int axis, int sign) { return c.face_friction[2 * axis + (sign > 0 ? 0 : 1)]; }

// solver.rs:477-498 (one contact limit per pellet).
double stability_limit(const oracle_sim& s) {
    double dx = s.grid.layout.spacing;
    double max_speed = 0.0;
    for (const V3& v : s.particles.vel) max_speed = std::max(max_speed, norm(v));
    double wave = wave_speed(s.material);
    double limit = s.config.cfl * dx / (wave + max_speed);
    double diff = diffusivity(s.material);
    if (diff > 0.0) limit = std::min(limit, 0.25 * dx * dx / diff);
    for (const Pellet& p : s.pellets) {
        limit = std::min(limit, contact_stable_dt(s.config, p.mass));
        double speed = norm(p.velocity);
        if (speed > 0.0) limit = std::min(limit, s.config.cfl * dx / speed);
    }
    return limit;
}

// solver.rs:501-517
void sort_particles(oracle_sim& s) {
    const Layout& l = s.grid.layout;
    size_t n = s.particles.size();
    std::vector<size_t> keys(n), order(n);
    for (size_t p = 0; p < n; ++p) {
        size_t cell = 0;
        for (int a = 0; a < 3; ++a) {
            double scaled = std::floor(s.particles.pos[p][a] / l.spacing);
            scaled = std::isnan(scaled) ? 0.0 : std::max(scaled, 0.0);
            size_t coord = scaled >= 1.8e19 ? (size_t)-1 : (size_t)scaled;
            coord = std::min(coord, l.cells[a]);
            cell = cell * (l.cells[a] + 1) + coord;
        }
        keys[p] = cell;
        order[p] = p;
    }
    std::sort(order.begin(), order.end(), [&](size_t a, size_t b) {
        if (keys[a] != keys[b]) return keys[a] < keys[b];
        return s.particles.id[a] < s.particles.id[b];
    });
    s.particles.permute(order);
}

// solver.rs:520-532
void sweep_outflow(oracle_sim& s) {
    std::vector<size_t> escaped;
    Stencil st;
    for (size_t p = 0; p < s.particles.size(); ++p)
        if (!stencil(s.grid.layout, s.particles.pos[p], st)) escaped.push_back(p);
    if (escaped.empty()) return;
    for (size_t i : escaped)
        for (int a = 0; a < 3; ++a) s.ledger.outflow_momentum[a] += s.particles.vel[i][a] * s.particles.mass[i];
    s.ledger.outflow_mass += s.particles.remove_sorted(escaped);
}

// solver.rs:632-661
void deposit_particles(oracle_sim& s, double dt) {
    s.grid.clear();
    const Layout& l = s.grid.layout;
    double dx2_inv = 4.0 / (l.spacing * l.spacing);
    Stencil st;
    for (size_t p = 0; p < s.particles.size(); ++p) {
        if (!stencil(l, s.particles.pos[p], st)) continue;
        double mass = s.particles.mass[p];
        const V3& velocity = s.particles.vel[p];
        M3 am = msub(mscale(s.particles.affine[p], mass),
                     mscale(s.particles.kirchhoff[p], dt * dx2_inv * s.particles.volume0[p]));
        V3 vm = scale(velocity, mass);
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                for (int c = 0; c < 3; ++c) {
                    double w = st.weight(a, b, c);
                    V3 d = st.distance(a, b, c);
                    size_t index = l.index(st.base[0] + a, st.base[1] + b, st.base[2] + c);
                    V3 momentum = scale(add(vm, mv(am, d)), w);
                    s.grid.deposit(index, w * mass, momentum);
                }
    }
}

struct Transfer {
    V3 impulse{0, 0, 0};
    double energy = 0.0;
};

// solver.rs:725-819
Transfer wall_traction(oracle_sim& s, double dt) {
    const Layout& l = s.grid.layout;
    V3 extents = l.extents();
    double half_width = 0.5 * l.spacing;
    double force_scale = -dt * 4.0 / (l.spacing * l.spacing);
    V3 g = v3(s.config.gravity[0], s.config.gravity[1], s.config.gravity[2]);
    Transfer t;
    Stencil st;
    for (size_t p = 0; p < s.particles.size(); ++p) {
        const V3& position = s.particles.pos[p];
        double mass = s.particles.mass[p];
        double volume0 = s.particles.volume0[p];
        for (int axis = 0; axis < 3; ++axis) {
            double face;
            if (position[axis] < half_width)
                face = 0.0;
            else if (position[axis] > extents[axis] - half_width)
                face = extents[axis];
            else
                continue;
            V3 image = position;
            image[axis] = 2.0 * face - position[axis];
            if (!stencil(l, image, st)) continue;
            size_t plane;
            double inward;
            if (face == 0.0) {
                plane = 1;
                inward = 1.0;
            } else {
                plane = l.cells[axis] + 1;
                inward = -1.0;
            }
            double normal_stress = at(s.particles.kirchhoff[p], axis, axis) * volume0 -
                                   mass * g[axis] * (image[axis] - position[axis]);
            if (normal_stress >= 0.0) continue;
            double weight_momentum = mass * g[axis] * dt;
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b)
                    for (int c = 0; c < 3; ++c) {
                        size_t coords[3] = {st.base[0] + a, st.base[1] + b, st.base[2] + c};
                        if (l.wall_side(axis, coords[axis]) != 0) continue;
                        double w = st.weight(a, b, c);
                        if (w == 0.0) continue;
                        size_t pc[3] = {coords[0], coords[1], coords[2]};
                        pc[axis] = plane;
                        size_t plane_index = l.index(pc[0], pc[1], pc[2]);
                        double plane_mass = s.grid.mass[plane_index];
                        if (plane_mass <= 0.0) continue;
                        double trial = s.grid.momentum[plane_index][axis] / plane_mass + g[axis] * dt;
                        if (trial * inward >= 0.0) continue;
                        int off = axis == 0 ? a : (axis == 1 ? b : c);
                        double distance = st.off[axis][off];
                        double reaction = (normal_stress * distance * force_scale + weight_momentum) * w * inward;
                        if (reaction <= 0.0) continue;
                        size_t index = l.index(coords[0], coords[1], coords[2]);
                        double jump;
                        if (s.grid.deposit_reaction(index, axis, inward, reaction, jump)) {
                            t.impulse[axis] += reaction * inward;
                            t.energy += jump;
                        }
                    }
        }
    }
    return t;
}

// solver.rs:822-892 (per-face friction extension).
int grid_update(const oracle_sim& s, double dt, std::vector<NodeTrial>& trials) {
    const Layout& l = s.grid.layout;
    V3 g = v3(s.config.gravity[0], s.config.gravity[1], s.config.gravity[2]);
    size_t n = s.grid.active.size();
    trials.assign(n, NodeTrial());
    std::vector<char> bad(n, 0);
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (size_t q = 0; q < n; ++q) {
        size_t index = s.grid.active[q];
        double mass = s.grid.mass[index];
        if (mass <= 0.0) continue;
        const V3& mom = s.grid.momentum[index];
        V3 velocity = v3(mom[0] / mass + g[0] * dt, mom[1] / mass + g[1] * dt, mom[2] / mass + g[2] * dt);
        if (!finite3(velocity)) {
            bad[q] = 1;
            continue;
        }
        V3 before = velocity;
        double projection = 0.0, friction = 0.0;
        size_t c[3];
        l.coords(index, c[0], c[1], c[2]);
        for (int axis = 0; axis < 3; ++axis) {
            int sign = l.wall_side(axis, c[axis]);
            if (sign == 0) continue;
            double ns = velocity[axis] * (double)sign;
            if (ns >= 0.0) continue;
            projection -= 0.5 * mass * ns * ns;
            velocity[axis] = 0.0;
            V3 tangential = velocity;
            velocity = reduce_tangential(tangential, face_friction(s.config, axis, sign) * (-ns));
            friction += 0.5 * mass * (norm2(tangential) - norm2(velocity));
        }
        const V3& tr = s.grid.traction[index];
        for (int axis = 0; axis < 3; ++axis) {
            double reaction = tr[axis];
            if (reaction <= 0.0) continue;
            double normal = velocity[axis];
            V3 tangential = velocity;
            tangential[axis] = 0.0;
            int sign = 2 * c[axis] <= l.cells[axis] + 2 ? 1 : -1;
            V3 reduced = reduce_tangential(tangential, face_friction(s.config, axis, sign) * reaction / mass);
            friction += 0.5 * mass * (norm2(tangential) - norm2(reduced));
            velocity = reduced;
            velocity[axis] = normal;
        }
        NodeTrial& t = trials[q];
        t.velocity = velocity;
        t.wall_impulse = scale(sub(velocity, before), mass);
        t.projection = projection;
        t.friction = friction;
    }
    for (size_t q = 0; q < n; ++q)
        if (bad[q]) return ORACLE_FAIL_NONFINITE_GRID;
    return 0;
}

// solver.rs:895-977
int grid_to_particle(const oracle_sim& s, double dt, std::vector<Trial>& trials) {
    const Layout& l = s.grid.layout;
    double dx2_inv = 4.0 / (l.spacing * l.spacing);
    const oracle_material& mat = s.material;
    const Particles& ps = s.particles;
    size_t n = ps.size();
    trials.resize(n);
    std::vector<int> codes(n, 0);
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (size_t p = 0; p < n; ++p) {
        Trial& t = trials[p];
        Stencil st;
        if (!stencil(l, ps.pos[p], st)) {
            t = Trial{ps.pos[p], ps.vel[p], ps.affine[p], ps.deformation[p], ps.kirchhoff[p],
                      ps.volume_ratio[p], 0.0, 0.0, 0.0};
            continue;
        }
        V3 velocity{0, 0, 0};
        M3 affine = zeros9();
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                for (int c = 0; c < 3; ++c) {
                    double w = st.weight(a, b, c);
                    V3 d = st.distance(a, b, c);
                    size_t index = l.index(st.base[0] + a, st.base[1] + b, st.base[2] + c);
                    V3 nv = scale(s.grid.momentum[index], w);
                    velocity = add(velocity, nv);
                    for (int i = 0; i < 3; ++i)
                        for (int j = 0; j < 3; ++j) affine[3 * i + j] += nv[i] * d[j];
                }
        affine = mscale(affine, dx2_inv);
        t.position = add(ps.pos[p], scale(velocity, dt));
        t.velocity = velocity;
        t.affine = affine;
        t.displacement = norm(velocity) * dt / l.spacing;
        if (mat.kind == 0) {
            M3 trial = mm(madd(identity(), mscale(affine, dt)), ps.deformation[p]);
            PasteUpdate up;
            int rc = update_paste(trial, mat, dt, up);
            if (rc) {
                codes[p] = rc;
                continue;
            }
            t.deformation = up.deformation;
            t.volume_ratio = 1.0;
            t.kirchhoff = up.kirchhoff;
            t.increment = up.increment;
            t.dissipation = up.dissipation_density * ps.volume0[p];
        } else {
            double j;
            M3 tau;
            double pressure;
            int rc = update_liquid(ps.volume_ratio[p], affine, mat, dt, j, tau, pressure);
            if (rc) {
                codes[p] = rc;
                continue;
            }
            t.deformation = identity();
            t.volume_ratio = j;
            t.kirchhoff = tau;
            t.increment = 0.0;
            t.dissipation = 0.0;
        }
    }
    for (size_t p = 0; p < n; ++p)
        if (codes[p]) return codes[p];
    return 0;
}

// solver.rs:540-611 (pellets coupled one after another in index order).
// Kinetic snapshots of one trial (audit.rs:70-88), filled when try_step is
// given somewhere to put them (solver.rs:540-611 with `audit: Some`).
struct Stages {
    double transport = 0.0, after_stress = 0.0, after_traction = 0.0, after_gravity = 0.0, after_walls = 0.0,
           after_coupling = 0.0, pellet_after_coupling = 0.0, plastic = 0.0;
};

// grid_kinetic, audit.rs:360-374: active nodes with mass, momentum or velocity storage.
double grid_kinetic(const Grid& g, bool velocities, const V3& kick) {
    double e = 0.0;
    for (size_t i : g.active) {
        double m = g.mass[i];
        if (!(m > 0.0)) continue;
        V3 v = g.momentum[i];
        if (!velocities) v = V3{v[0] / m, v[1] / m, v[2] / m};
        e += 0.5 * m * norm2(add(v, kick));
    }
    return e;
}

int try_step(oracle_sim& s, double dt, Stages* audit = nullptr) {
    sweep_outflow(s);
    deposit_particles(s, dt);
    const V3 kick0{0.0, 0.0, 0.0};
    if (audit) audit->after_stress = grid_kinetic(s.grid, false, kick0);
    Transfer traction = wall_traction(s, dt);
    if (audit) {
        audit->after_traction = grid_kinetic(s.grid, false, kick0);
        audit->after_gravity =
            grid_kinetic(s.grid, false, V3{s.config.gravity[0] * dt, s.config.gravity[1] * dt, s.config.gravity[2] * dt});
    }
    if (!std::isfinite(traction.energy) || !std::isfinite(s.ledger.wall_normal_traction_energy + traction.energy))
        return ORACLE_FAIL_TRACTION;
    std::vector<NodeTrial> nodes;
    int rc = grid_update(s, dt, nodes);
    if (rc) return rc;
    V3 wall_impulse = traction.impulse;
    double projection = 0.0, friction = 0.0;
    for (size_t q = 0; q < s.grid.active.size(); ++q) {
        s.grid.momentum[s.grid.active[q]] = nodes[q].velocity;
        wall_impulse = add(wall_impulse, nodes[q].wall_impulse);
        projection += nodes[q].projection;
        friction += nodes[q].friction;
    }
    if (audit) audit->after_walls = grid_kinetic(s.grid, true, kick0);
    std::vector<Pellet> trial_pellets = s.pellets;
    std::vector<CouplingResult> couplings;
    std::vector<double> pellet_energies;
    std::vector<WallContact> walls;
    V3 g = v3(s.config.gravity[0], s.config.gravity[1], s.config.gravity[2]);
    for (Pellet& pellet : trial_pellets) {
        CouplingResult res = couple_grid(s.grid, pellet, s.config.wall_friction);
        double before = pellet.kinetic_energy();
        pellet.apply_impulse(res.impulse, res.angular_impulse);
        double pe = pellet.kinetic_energy() - before;
        if (audit) audit->pellet_after_coupling += pellet.kinetic_energy();
        WallContact wc = pellet_integrate(pellet, dt, g, s.grid.layout.extents(), s.config);
        if (!pellet.is_finite() || !std::isfinite(pe)) return ORACLE_FAIL_PELLET;
        couplings.push_back(res);
        pellet_energies.push_back(pe);
        walls.push_back(wc);
    }
    if (audit) audit->after_coupling = grid_kinetic(s.grid, true, kick0);
    std::vector<Trial> trials;
    rc = grid_to_particle(s, dt, trials);
    if (rc) return rc;
    if (audit) {
        double sum = 0.0;
        for (const Trial& t : trials) sum += t.dissipation;
        audit->plastic = sum;
    }
    double max_cells = 0.0;
    for (const Trial& t : trials) max_cells = std::max(max_cells, t.displacement);
    if (max_cells > 1.0) return ORACLE_FAIL_DISPLACEMENT;
    // commit, solver.rs:980-998
    double dissipation = 0.0;
    Particles& ps = s.particles;
    for (size_t p = 0; p < trials.size(); ++p) {
        const Trial& t = trials[p];
        ps.pos[p] = t.position;
        ps.vel[p] = t.velocity;
        ps.affine[p] = t.affine;
        ps.deformation[p] = t.deformation;
        ps.volume_ratio[p] = t.volume_ratio;
        ps.kirchhoff[p] = t.kirchhoff;
        ps.plastic[p] += t.increment;
        dissipation += t.dissipation;
    }
    oracle_ledger& L = s.ledger;
    L.plastic_dissipation += dissipation;
    for (int a = 0; a < 3; ++a) L.wall_impulse[a] += wall_impulse[a];
    L.wall_normal_projection_energy += projection;
    L.wall_friction_dissipation += friction;
    L.wall_normal_traction_energy += traction.energy;
    double mdt = ps.total_mass() * dt;
    for (int a = 0; a < 3; ++a) L.gravity_impulse[a] += g[a] * mdt;
    for (size_t k = 0; k < trial_pellets.size(); ++k) {
        double pm = trial_pellets[k].mass * dt;
        for (int a = 0; a < 3; ++a) {
            L.gravity_impulse[a] += g[a] * pm;
            L.pellet_wall_impulse[a] += walls[k].impulse[a];
            L.coupling_impulse[a] += couplings[k].impulse[a];
            L.coupling_angular_impulse[a] += couplings[k].angular_impulse[a];
        }
        L.pellet_wall_work += walls[k].work;
        L.coupling_grid_energy += couplings[k].grid_energy;
        L.coupling_pellet_energy += pellet_energies[k];
    }
    s.pellets = trial_pellets;
    return 0;
}

// ---------------------------------------------------------------- energy audit, audit.rs

// Simulation::validate (solver.rs:1106-1140) with ParticleSet::validate
// (particles.rs:159-211) and Ledger::validate (solver.rs:182-243).
bool validate_sim(const oracle_sim& s) {
    const Particles& ps = s.particles;
    std::vector<uint32_t> ids(ps.id);
    std::sort(ids.begin(), ids.end());
    for (size_t k = 1; k < ids.size(); ++k)
        if (ids[k] == ids[k - 1]) return false;
    for (size_t p = 0; p < ps.size(); ++p) {
        bool ok = finite3(ps.pos[p]) && finite3(ps.vel[p]) && std::isfinite(ps.volume_ratio[p]) &&
                  std::isfinite(ps.plastic[p]) && std::isfinite(ps.mass[p]) && std::isfinite(ps.volume0[p]);
        for (int a = 0; a < 9; ++a)
            ok = ok && std::isfinite(ps.affine[p][a]) && std::isfinite(ps.deformation[p][a]) && std::isfinite(ps.kirchhoff[p][a]);
        if (!ok || ps.mass[p] <= 0.0 || ps.volume0[p] <= 0.0 || ps.volume_ratio[p] <= 0.0 || ps.plastic[p] < 0.0)
            return false;
    }
    const oracle_ledger& L = s.ledger;
    const double scalars[] = {L.initial_mass, L.outflow_mass, L.initial_mechanical_energy, L.wall_normal_projection_energy,
                              L.wall_friction_dissipation, L.wall_normal_traction_energy, L.pellet_wall_work,
                              L.coupling_grid_energy, L.coupling_pellet_energy, L.plastic_dissipation};
    for (double v : scalars)
        if (!std::isfinite(v)) return false;
    const double* vectors[] = {L.initial_momentum, L.gravity_impulse, L.wall_impulse, L.pellet_wall_impulse,
                               L.coupling_impulse, L.coupling_angular_impulse, L.outflow_momentum};
    for (const double* v : vectors)
        for (int a = 0; a < 3; ++a)
            if (!std::isfinite(v[a])) return false;
    if (L.initial_mass < 0.0 || L.outflow_mass < 0.0 || L.plastic_dissipation < 0.0 || L.wall_friction_dissipation < 0.0 ||
        L.wall_normal_projection_energy > 0.0)
        return false;
    if (!std::isfinite(s.time) || s.time < 0.0) return false;
    if (!(std::isfinite(s.dt_scale) && s.dt_scale > 0.0 && s.dt_scale <= 1.0)) return false;
    for (const Pellet& pl : s.pellets)
        if (!pl.is_finite()) return false;
    Stencil st;
    for (size_t p = 0; p < ps.size(); ++p)
        if (!stencil(s.grid.layout, ps.pos[p], st)) return false;
    return true;
}

// EnergyState::measure, audit.rs:37-54.
void measure_energy(const oracle_sim& s, double* e) {
    const Particles& ps = s.particles;
    double sum = 0.0;
    for (size_t p = 0; p < ps.size(); ++p) {
        double n2 = 0.0;  // norm_squared over the column-major storage
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) n2 += at(ps.affine[p], i, j) * at(ps.affine[p], i, j);
        sum += ps.mass[p] * n2;
    }
    double h = s.grid.layout.spacing;
    double pellet = 0.0;
    for (const Pellet& pl : s.pellets) pellet += pl.kinetic_energy();
    e[0] = ps.kinetic_energy();
    e[1] = h * h / 8.0 * sum;
    e[2] = elastic_energy(s);
    e[3] = potential_energy(s);
    e[4] = pellet;
}

// transport_energy, audit.rs:376-398: APIC-only deposit on a fresh grid.
double transport_energy(const oracle_sim& s) {
    const Layout& l = s.grid.layout;
    Grid shadow(l);
    Stencil st;
    for (size_t p = 0; p < s.particles.size(); ++p) {
        if (!stencil(l, s.particles.pos[p], st)) return std::numeric_limits<double>::quiet_NaN();
        double mass = s.particles.mass[p];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                for (int c = 0; c < 3; ++c) {
                    double w = st.weight(a, b, c);
                    V3 velocity = add(s.particles.vel[p], mv(s.particles.affine[p], st.distance(a, b, c)));
                    size_t index = l.index(st.base[0] + a, st.base[1] + b, st.base[2] + c);
                    shadow.deposit(index, w * mass, scale(velocity, w * mass));
                }
    }
    return grid_kinetic(shadow, false, V3{0.0, 0.0, 0.0});
}

// Simulation::audit_step, audit.rs:275-345, and EnergyTerms::measure (193-232).
int audit_step(const oracle_sim& s, double dt, oracle_audit* out) {
    if (!validate_sim(s)) return ORACLE_ERR_INVALID;
    const oracle_config& c = s.config;
    double stability = stability_limit(s);
    if (!std::isfinite(c.gravity[0]) || !std::isfinite(c.gravity[1]) || !std::isfinite(c.gravity[2]) ||
        !std::isfinite(c.cfl) || c.cfl <= 0.0 || !std::isfinite(c.wall_friction) || c.wall_friction < 0.0 ||
        !std::isfinite(c.min_dt) || c.min_dt <= 0.0 || !std::isfinite(c.max_dt) || c.max_dt < c.min_dt ||
        !std::isfinite(stability) || stability <= 0.0)
        return ORACLE_ERR_INVALID;
    double limit = std::min(s.dt_scale * stability, c.max_dt);
    if (!std::isfinite(dt) || dt <= 0.0 || dt < c.min_dt || !std::isfinite(limit) || dt > limit) return ORACLE_ERR_INVALID;
    oracle_audit a{};
    a.time = s.time;
    a.dt = dt;
    measure_energy(s, a.before);
    for (double v : a.before)
        if (!std::isfinite(v)) return ORACLE_ERR_INVALID;
    Stages st;
    st.transport = transport_energy(s);
    oracle_sim trial(s);
    if (try_step(trial, dt, &st) != 0) return ORACLE_ERR_INVALID;
    if (!validate_sim(trial)) return ORACLE_ERR_INVALID;
    measure_energy(trial, a.after);
    const double stages[8] = {st.transport, st.after_stress, st.after_traction, st.after_gravity,
                              st.after_walls, st.after_coupling, st.pellet_after_coupling, st.plastic};
    std::copy(stages, stages + 8, a.stages);
    const oracle_ledger &B = s.ledger, &A = trial.ledger;
    a.increments[0] = A.plastic_dissipation - B.plastic_dissipation;
    a.increments[1] = A.wall_friction_dissipation - B.wall_friction_dissipation;
    a.increments[2] = A.wall_normal_projection_energy - B.wall_normal_projection_energy;
    a.increments[3] = A.wall_normal_traction_energy - B.wall_normal_traction_energy;
    a.increments[4] = A.coupling_grid_energy - B.coupling_grid_energy;
    a.increments[5] = A.coupling_pellet_energy - B.coupling_pellet_energy;
    a.increments[6] = A.pellet_wall_work - B.pellet_wall_work;
    const double* be = a.before;
    const double* af = a.after;
    double p2g = be[0] + be[1] - st.transport;
    double g2p = st.after_coupling - af[0] - af[1];
    double stress = st.after_stress - st.transport + af[2] - be[2] + st.plastic;
    double gravity = st.after_gravity - st.after_traction + af[3] - be[3];
    double integrator = af[4] - st.pellet_after_coupling;
    double traction = st.after_traction - st.after_stress;
    double walls = st.after_walls - st.after_gravity;
    double coupling = st.after_coupling - st.after_walls + st.pellet_after_coupling - be[4];
    double total_after = af[0] + af[1] + af[2] + af[3] + af[4];
    double total_before = be[0] + be[1] + be[2] + be[3] + be[4];
    double change = total_after - total_before;
    double sum = -p2g - g2p + stress - st.plastic + gravity + integrator + traction + walls + coupling;
    const double terms[11] = {p2g, g2p, stress, gravity, integrator, traction, walls, coupling, change, sum, change - sum};
    std::copy(terms, terms + 11, a.terms);
    a.coupled = s.pellets.empty() ? 0 : 1;
    for (double v : a.after)
        if (!std::isfinite(v)) return ORACLE_ERR_INVALID;
    for (double v : a.stages)
        if (!std::isfinite(v)) return ORACLE_ERR_INVALID;
    for (double v : a.increments)
        if (!std::isfinite(v)) return ORACLE_ERR_INVALID;
    for (double v : a.terms)
        if (!std::isfinite(v)) return ORACLE_ERR_INVALID;
    *out = a;
    return 0;
}

// solver.rs:413-472
int advance_while(oracle_sim& s, double target, double* fail_dt) {
    if (!std::isfinite(target) || target < s.time) return ORACLE_ERR_INVALID;
    while (s.time < target) {
        double remaining = target - s.time;
        double limit = stability_limit(s);
        double scaled = s.dt_scale * limit;
        bool limited = scaled < s.config.max_dt;
        double dt = std::min(std::min(scaled, s.config.max_dt), remaining);
        if (remaining - dt < 1e-12 * std::max(target, 1.0)) dt = remaining;
        if (!std::isfinite(dt) || dt <= 0.0 || (dt < s.config.min_dt && remaining > s.config.min_dt)) {
            if (fail_dt) *fail_dt = dt;
            return ORACLE_ERR_MIN_STEP;
        }
        // try_step commits nothing unless every check passes (the grid is scratch).
        int rc = try_step(s, dt);
        if (rc == 0) {
            s.time += dt;
            s.step_count += 1;
            s.limited_steps += limited ? 1 : 0;
            s.min_dt_used = std::min(s.min_dt_used, dt);
            s.max_dt_used = std::max(s.max_dt_used, dt);
            s.dt_scale = std::min(s.dt_scale * 1.25, 1.0);
            if (s.step_count % 32 == 0) sort_particles(s);
        } else {
            s.last_failure = rc;
            s.rejected_steps += 1;
            s.dt_scale *= 0.5;
            if (s.dt_scale * limit < s.config.min_dt) {
                if (fail_dt) *fail_dt = dt;
                return ORACLE_ERR_MIN_STEP;
            }
        }
    }
    s.time = target;
    return 0;
}

// rigid.rs:115-182
int pellet_new(double radius, double length, double density, const double* position, const double* q,
               Pellet& out) {
    if (!(std::isfinite(radius) && radius > 0.0)) return -1;
    if (!(std::isfinite(length) && length >= 2.0 * radius)) return -1;
    if (!(std::isfinite(density) && density > 0.0)) return -1;
    double mass = density * M_PI * radius * radius * length;
    out.radius = radius;
    out.length = length;
    out.mass = mass;
    out.inertia_body = v3(0.5 * mass * radius * radius, mass * (3.0 * radius * radius + length * length) / 12.0,
                          mass * (3.0 * radius * radius + length * length) / 12.0);
    out.position = v3(position[0], position[1], position[2]);
    out.orientation = q ? Quat{q[0], q[1], q[2], q[3]} : Quat{0, 0, 0, 1};
    out.velocity = V3{0, 0, 0};
    out.angular_momentum = V3{0, 0, 0};
    double half_span = 0.5 * length - radius;
    size_t count = 1;
    if (half_span > 0.0) {
        double ratio = std::ceil(2.0 * half_span / radius);
        if (!std::isfinite(ratio) || ratio > 63.0) return -1;
        count = (size_t)ratio + 1;
    }
    out.offsets.clear();
    for (size_t i = 0; i < count; ++i)
        out.offsets.push_back(count == 1 ? 0.0
                                         : -half_span + 2.0 * half_span * (double)i / (double)(count - 1));
    return 0;
}

void to_c(const Pellet& p, oracle_pellet* o) {
    o->radius = p.radius;
    o->length = p.length;
    o->mass = p.mass;
    for (int a = 0; a < 3; ++a) {
        o->inertia_body[a] = p.inertia_body[a];
        o->position[a] = p.position[a];
        o->velocity[a] = p.velocity[a];
        o->angular_momentum[a] = p.angular_momentum[a];
    }
    o->orientation[0] = p.orientation.i;
    o->orientation[1] = p.orientation.j;
    o->orientation[2] = p.orientation.k;
    o->orientation[3] = p.orientation.w;
    o->sphere_count = (int)p.offsets.size();
    for (size_t i = 0; i < p.offsets.size() && i < 64; ++i) o->sphere_offsets[i] = p.offsets[i];
}
void from_c(const oracle_pellet* o, Pellet& p) {
    p.radius = o->radius;
    p.length = o->length;
    p.mass = o->mass;
    for (int a = 0; a < 3; ++a) {
        p.inertia_body[a] = o->inertia_body[a];
        p.position[a] = o->position[a];
        p.velocity[a] = o->velocity[a];
        p.angular_momentum[a] = o->angular_momentum[a];
    }
    p.orientation = Quat{o->orientation[0], o->orientation[1], o->orientation[2], o->orientation[3]};
    p.offsets.assign(o->sphere_offsets, o->sphere_offsets + o->sphere_count);
}

M3 load9(const double* p) {
    M3 m;
    std::memcpy(m.data(), p, sizeof(double) * 9);
    return m;
}

}  // namespace

extern "C" {

void oracle_set_threads(int threads) { g_threads = threads < 1 ? 1 : threads; }

void oracle_material_paste(double density, double young, double poisson, double yield_stress, double consistency,
                           double flow_index, oracle_material* out) {
    double shear = young / (2.0 * (1.0 + poisson));
    double lame = young * poisson / ((1.0 + poisson) * (1.0 - 2.0 * poisson));
    *out = oracle_material{0, density, shear, lame, lame + 2.0 * shear / 3.0, yield_stress, consistency,
                           flow_index, 1.0};
}

void oracle_material_paste_shear(double density, double young, double poisson, double shear_yield,
                                 double shear_consistency, double flow_index, oracle_material* out) {
    double root3 = std::sqrt(3.0);
    oracle_material_paste(density, young, poisson, root3 * shear_yield,
                          std::pow(3.0, (1.0 + flow_index) / 2.0) * shear_consistency, flow_index, out);
}

void oracle_material_liquid(double density, double young, double poisson, double viscosity, oracle_material* out) {
    double bulk = young / (3.0 * (1.0 - 2.0 * poisson));
    *out = oracle_material{1, density, 0.0, 0.0, bulk, 0.0, viscosity, 1.0, 1.0};
}

double oracle_wave_speed(const oracle_material* m) { return wave_speed(*m); }

void oracle_default_config(oracle_config* c) {
    *c = oracle_config{};
    c->gravity[2] = -9.81;
    c->wall_friction = 0.4;
    c->cfl = 0.3;
    c->max_dt = 1e-3;
    c->min_dt = 1e-9;
    c->contact_stiffness = 1000.0;
    c->contact_restitution = 0.2;
    c->contact_friction = 0.4;
    for (int i = 0; i < 6; ++i) c->face_friction[i] = 0.4;
}

int oracle_hb_return(double trial, double ys, double k, double n, double mu, double dt, double* stress,
                     double* increment, int* iterations) {
    ReturnSolution s{};
    int rc = hb_return(trial, ys, k, n, mu, dt, s);
    if (stress) *stress = s.stress;
    if (increment) *increment = s.increment;
    if (iterations) *iterations = s.iterations;
    return rc;
}

int oracle_eigen3(const double* m, double* values, double* vectors) {
    V3 v;
    M3 u;
    if (!eigen3(load9(m), v, u)) return ORACLE_FAIL_EIGEN;
    std::memcpy(values, v.data(), 3 * sizeof(double));
    std::memcpy(vectors, u.data(), 9 * sizeof(double));
    return 0;
}

int oracle_update_paste(const double* trial, const oracle_material* m, double dt, double* deformation,
                        double* kirchhoff, double* increment, double* dissipation_density,
                        double* equivalent_stress) {
    PasteUpdate up{};
    int rc = update_paste(load9(trial), *m, dt, up);
    if (rc) return rc;
    if (deformation) std::memcpy(deformation, up.deformation.data(), 9 * sizeof(double));
    if (kirchhoff) std::memcpy(kirchhoff, up.kirchhoff.data(), 9 * sizeof(double));
    if (increment) *increment = up.increment;
    if (dissipation_density) *dissipation_density = up.dissipation_density;
    if (equivalent_stress) *equivalent_stress = up.equivalent_stress;
    return 0;
}

int oracle_update_liquid(double volume_ratio, const double* gradient, const oracle_material* m, double dt,
                         double* new_ratio, double* kirchhoff, double* pressure) {
    double j = 0.0, p = 0.0;
    M3 tau = zeros9();
    int rc = update_liquid(volume_ratio, load9(gradient), *m, dt, j, tau, p);
    if (new_ratio) *new_ratio = j;
    if (pressure) *pressure = p;
    if (kirchhoff) std::memcpy(kirchhoff, tau.data(), 9 * sizeof(double));
    return rc;
}

double oracle_liquid_pressure(const oracle_material* m, double j) { return liquid_pressure(*m, j); }

uint64_t oracle_splitmix_next(uint64_t* state) {
    *state += 0x9E3779B97F4A7C15ull;
    uint64_t z = *state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double oracle_splitmix_unit(uint64_t* state) {
    return (double)(oracle_splitmix_next(state) >> 11) * (1.0 / (double)(1ull << 53));
}

// fixtures.rs:120-133
size_t oracle_fixture_particle_count(const oracle_fixture* f) {
    double sub = f->grid_spacing / 2.0;
    size_t count = 1;
    for (int a = 0; a < 3; ++a) count *= (size_t)std::max(std::round(f->initial_size[a] / sub), 1.0);
    return count;
}

// fixtures.rs:152-349
int oracle_fixture_build(const oracle_fixture* f, oracle_sim** out, char* err, size_t errlen) {
    auto fail = [&](const std::string& msg) {
        if (err && errlen) std::snprintf(err, errlen, "%s", msg.c_str());
        return ORACLE_ERR_INVALID;
    };
    for (int a = 0; a < 3; ++a) {
        if (f->initial_size[a] > f->domain[a] + 1e-12) return fail("`initial_size_m` exceeds `domain_m`");
        if (f->initial_size[a] < f->grid_spacing) return fail("`initial_size_m` must span at least one cell");
    }
    size_t count = oracle_fixture_particle_count(f);
    if (count > f->max_particles) return fail("fixture would allocate particles exceeding the cap");
    if (f->fixture == 1) {
        for (int a = 0; a < 2; ++a)
            if (std::fabs(f->initial_size[a] - f->domain[a]) > 1e-9)
                return fail("`hydrostatic` requires the column to span the domain");
        if (norm(v3(f->initial_velocity[0], f->initial_velocity[1], f->initial_velocity[2])) > 0.0)
            return fail("`hydrostatic` requires zero `initial_velocity_m_s`");
        if (f->initial_size[2] >= f->domain[2]) return fail("`hydrostatic` requires free space above the column");
    }
    if (f->fixture == 3) {
        if (!f->has_pellet) return fail("`coupled_patch` requires exactly one box supplying pellet geometry");
        if (f->material.kind != 0) return fail("`coupled_patch` is implemented for `paste` only");
        if (f->pellet_radius < 1.5 * f->grid_spacing) return fail("`pellet_radius_m` is under-resolved");
        double top = f->initial_size[2] + 2.0 * f->grid_spacing + 2.0 * f->pellet_radius;
        if (top > f->domain[2] || f->pellet_length > f->domain[0] || 2.0 * f->pellet_radius > f->domain[1])
            return fail("`coupled_patch` pellet does not fit above the block inside `domain_m`");
    } else if (f->has_pellet) {
        return fail("only `coupled_patch` uses pellet geometry");
    }
    Layout layout;
    std::string msg;
    if (!make_layout(f->domain, f->grid_spacing, f->max_nodes, layout, msg)) return fail(msg);
    oracle_sim* s = new oracle_sim(layout);
    s->material = f->material;
    s->config = f->config;
    uint64_t rng = f->seed;
    double sub = f->grid_spacing / 2.0;
    size_t counts[3];
    for (int a = 0; a < 3; ++a) counts[a] = (size_t)std::max(std::round(f->initial_size[a] / sub), 1.0);
    V3 origin{0, 0, 0};
    if (f->fixture == 0 || f->fixture == 3)
        origin = v3(0.5 * (f->domain[0] - f->initial_size[0]), 0.5 * (f->domain[1] - f->initial_size[1]), 0.0);
    double jitter = f->fixture == 1 ? 0.0 : f->jitter * sub;
    double height = f->initial_size[2];
    const oracle_material& m = f->material;
    double gnorm = norm(v3(f->config.gravity[0], f->config.gravity[1], f->config.gravity[2]));
    V3 v0 = v3(f->initial_velocity[0], f->initial_velocity[1], f->initial_velocity[2]);
    for (size_t i = 0; i < counts[0]; ++i)
        for (size_t j = 0; j < counts[1]; ++j)
            for (size_t k = 0; k < counts[2]; ++k) {
                V3 pos = add(origin, v3(((double)i + 0.5) * sub, ((double)j + 0.5) * sub, ((double)k + 0.5) * sub));
                if (jitter > 0.0)
                    for (int a = 0; a < 3; ++a) pos[a] += (oracle_splitmix_unit(&rng) - 0.5) * 2.0 * jitter;
                double current = sub * sub * sub;
                double ratio = 1.0;
                M3 deformation = identity(), tau = zeros9();
                if (f->fixture == 1) {
                    double depth = height - pos[2];
                    if (m.kind == 1) {
                        ratio = std::exp(-m.density * gnorm * depth / m.bulk_modulus);
                        double pressure = liquid_pressure(m, ratio);
                        tau = mscale(identity(), -pressure * ratio);
                    } else {
                        double modulus = m.lame_lambda + 2.0 * m.shear_modulus;
                        double strain = -m.density * gnorm * depth / modulus;
                        double stretch = std::exp(strain);
                        deformation = diag(v3(1.0, 1.0, stretch));
                        tau = diag(v3(m.lame_lambda * strain, m.lame_lambda * strain, modulus * strain));
                        ratio = stretch;
                    }
                }
                double volume0 = current / ratio;
                s->particles.push(pos, v0, m.density * volume0, volume0);
                size_t idx = s->particles.size() - 1;
                s->particles.volume_ratio[idx] = ratio;
                s->particles.deformation[idx] = deformation;
                s->particles.kirchhoff[idx] = tau;
            }
    if (f->fixture == 3) {
        double center[3] = {0.5 * f->domain[0], 0.5 * f->domain[1],
                            f->initial_size[2] + 2.0 * f->grid_spacing + f->pellet_radius};
        Pellet p;
        if (pellet_new(f->pellet_radius, f->pellet_length, f->pellet_density, center, nullptr, p)) {
            delete s;
            return fail("invalid pellet geometry");
        }
        s->pellets.push_back(p);
    }
    seed_ledger(*s);
    *out = s;
    return 0;
}

int oracle_sim_create(const oracle_material* m, const oracle_config* c, const double* domain, double spacing,
                      size_t max_nodes, size_t n, const double* pos, const double* vel, const double* mass,
                      const double* volume0, oracle_sim** out) {
    Layout layout;
    std::string msg;
    if (!make_layout(domain, spacing, max_nodes, layout, msg)) return ORACLE_ERR_INVALID;
    oracle_sim* s = new oracle_sim(layout);
    s->material = *m;
    s->config = *c;
    for (size_t p = 0; p < n; ++p)
        s->particles.push(v3(pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]),
                          v3(vel[3 * p], vel[3 * p + 1], vel[3 * p + 2]), mass[p], volume0[p]);
    seed_ledger(*s);
    *out = s;
    return 0;
}

void oracle_sim_destroy(oracle_sim* s) { delete s; }
oracle_sim* oracle_sim_clone(const oracle_sim* s) { return new oracle_sim(*s); }
size_t oracle_sim_len(const oracle_sim* s) { return s->particles.size(); }

void oracle_sim_get(const oracle_sim* s, uint32_t* id, double* pos, double* vel, double* affine, double* deformation,
                    double* volume_ratio, double* kirchhoff, double* plastic, double* mass, double* volume0) {
    const Particles& p = s->particles;
    for (size_t i = 0; i < p.size(); ++i) {
        if (id) id[i] = p.id[i];
        for (int a = 0; a < 3; ++a) {
            if (pos) pos[3 * i + a] = p.pos[i][a];
            if (vel) vel[3 * i + a] = p.vel[i][a];
        }
        for (int a = 0; a < 9; ++a) {
            if (affine) affine[9 * i + a] = p.affine[i][a];
            if (deformation) deformation[9 * i + a] = p.deformation[i][a];
            if (kirchhoff) kirchhoff[9 * i + a] = p.kirchhoff[i][a];
        }
        if (volume_ratio) volume_ratio[i] = p.volume_ratio[i];
        if (plastic) plastic[i] = p.plastic[i];
        if (mass) mass[i] = p.mass[i];
        if (volume0) volume0[i] = p.volume0[i];
    }
}

int oracle_sim_set(oracle_sim* s, size_t n, const uint32_t* id, const double* pos, const double* vel,
                   const double* affine, const double* deformation, const double* volume_ratio,
                   const double* kirchhoff, const double* plastic, const double* mass, const double* volume0) {
    Particles& p = s->particles;
    if (n != p.size()) {
        Particles fresh;
        for (size_t i = 0; i < n; ++i) fresh.push(V3{0, 0, 0}, V3{0, 0, 0}, 1.0, 1.0);
        p = fresh;
    }
    for (size_t i = 0; i < n; ++i) {
        if (id) p.id[i] = id[i];
        for (int a = 0; a < 3; ++a) {
            if (pos) p.pos[i][a] = pos[3 * i + a];
            if (vel) p.vel[i][a] = vel[3 * i + a];
        }
        for (int a = 0; a < 9; ++a) {
            if (affine) p.affine[i][a] = affine[9 * i + a];
            if (deformation) p.deformation[i][a] = deformation[9 * i + a];
            if (kirchhoff) p.kirchhoff[i][a] = kirchhoff[9 * i + a];
        }
        if (volume_ratio) p.volume_ratio[i] = volume_ratio[i];
        if (plastic) p.plastic[i] = plastic[i];
        if (mass) p.mass[i] = mass[i];
        if (volume0) p.volume0[i] = volume0[i];
    }
    return 0;
}

void oracle_sim_get_ledger(const oracle_sim* s, oracle_ledger* l) { *l = s->ledger; }
void oracle_sim_set_ledger(oracle_sim* s, const oracle_ledger* l) { s->ledger = *l; }
void oracle_sim_get_history(const oracle_sim* s, oracle_history* h) {
    *h = oracle_history{s->time, s->step_count, s->limited_steps, s->rejected_steps, s->dt_scale, s->min_dt_used,
                        s->max_dt_used};
}
void oracle_sim_set_history(oracle_sim* s, const oracle_history* h) {
    s->time = h->time;
    s->step_count = h->step_count;
    s->limited_steps = h->limited_steps;
    s->rejected_steps = h->rejected_steps;
    s->dt_scale = h->dt_scale;
    s->min_dt_used = h->min_dt_used;
    s->max_dt_used = h->max_dt_used;
}
void oracle_sim_get_config(const oracle_sim* s, oracle_config* c) { *c = s->config; }
void oracle_sim_set_config(oracle_sim* s, const oracle_config* c) { s->config = *c; }
void oracle_sim_get_layout(const oracle_sim* s, double* spacing, int* cells) {
    *spacing = s->grid.layout.spacing;
    for (int a = 0; a < 3; ++a) cells[a] = (int)s->grid.layout.cells[a];
}

int oracle_pellet_new(double radius, double length, double density, const double* position, const double* q,
                      oracle_pellet* out) {
    Pellet p;
    if (pellet_new(radius, length, density, position, q, p)) return ORACLE_ERR_INVALID;
    to_c(p, out);
    return 0;
}
int oracle_sim_add_pellet(oracle_sim* s, const oracle_pellet* o) {
    Pellet p;
    from_c(o, p);
    s->pellets.push_back(p);
    return (int)s->pellets.size() - 1;
}
int oracle_sim_pellet_count(const oracle_sim* s) { return (int)s->pellets.size(); }
void oracle_sim_get_pellet(const oracle_sim* s, int index, oracle_pellet* p) { to_c(s->pellets[index], p); }
void oracle_sim_set_pellet(oracle_sim* s, int index, const oracle_pellet* p) { from_c(p, s->pellets[index]); }
double oracle_pellet_kinetic_energy(const oracle_pellet* o) {
    Pellet p;
    from_c(o, p);
    return p.kinetic_energy();
}
void oracle_pellet_signed_distance(const oracle_pellet* o, const double* point, double* phi, double* normal) {
    Pellet p;
    from_c(o, p);
    V3 n;
    *phi = p.signed_distance(v3(point[0], point[1], point[2]), n);
    for (int a = 0; a < 3; ++a) normal[a] = n[a];
}

int oracle_sim_advance_to(oracle_sim* s, double target, double* fail_dt) { return advance_while(*s, target, fail_dt); }

int oracle_sim_try_step(oracle_sim* s, double dt) { return try_step(*s, dt); }

double oracle_sim_stability_limit(const oracle_sim* s) { return stability_limit(*s); }
void oracle_sim_sort(oracle_sim* s) { sort_particles(*s); }

size_t oracle_sim_probe_transfer(const oracle_sim* s, double dt, double* traction_impulse, double* traction_energy,
                                 double* full_scan_jump) {
    oracle_sim copy(*s);
    sweep_outflow(copy);
    deposit_particles(copy, dt);
    auto scan = [&]() {
        double e = 0.0;
        for (size_t i : copy.grid.active) e += 0.5 * norm2(copy.grid.momentum[i]) / copy.grid.mass[i];
        return e;
    };
    double before = scan();
    Transfer t = wall_traction(copy, dt);
    double after = scan();
    if (traction_impulse)
        for (int a = 0; a < 3; ++a) traction_impulse[a] = t.impulse[a];
    if (traction_energy) *traction_energy = t.energy;
    if (full_scan_jump) *full_scan_jump = after - before;
    return copy.grid.active.size();
}

size_t oracle_sim_probe_grid(const oracle_sim* s, double dt, int with_traction, double* mass, double* momentum) {
    oracle_sim copy(*s);
    sweep_outflow(copy);
    deposit_particles(copy, dt);
    if (with_traction) wall_traction(copy, dt);
    size_t n = copy.grid.layout.count();
    for (size_t i = 0; i < n; ++i) {
        if (mass) mass[i] = copy.grid.mass[i];
        if (momentum)
            for (int a = 0; a < 3; ++a) momentum[3 * i + a] = copy.grid.momentum[i][a];
    }
    return copy.grid.active.size();
}

size_t oracle_sim_probe_update(const oracle_sim* s, double dt, double* mass, double* before, double* after,
                               double* budget, int* code) {
    oracle_sim copy(*s);
    sweep_outflow(copy);
    deposit_particles(copy, dt);
    wall_traction(copy, dt);
    std::vector<NodeTrial> trials;
    int rc = grid_update(copy, dt, trials);
    if (code) *code = rc;
    const Grid& g = copy.grid;
    size_t n = g.layout.count();
    for (size_t i = 0; i < n; ++i) {
        if (mass) mass[i] = g.mass[i];
        for (int a = 0; a < 3; ++a) {
            if (before) before[3 * i + a] = 0.0;
            if (after) after[3 * i + a] = 0.0;
            if (budget) budget[3 * i + a] = g.traction[i][a];
        }
    }
    if (rc) return 0;
    for (size_t q = 0; q < g.active.size(); ++q) {
        size_t i = g.active[q];
        double m = g.mass[i];
        if (!(m > 0.0)) continue;
        for (int a = 0; a < 3; ++a) {
            // solver.rs:833: the velocity the walls act on.
            if (before) before[3 * i + a] = g.momentum[i][a] / m + copy.config.gravity[a] * dt;
            if (after) after[3 * i + a] = trials[q].velocity[a];
        }
    }
    return g.active.size();
}

int oracle_couple_nodes(const double* domain, double spacing, size_t n_nodes, const size_t* index, const double* mass,
                        double* velocity, const oracle_pellet* pellets, int n_pellets, double friction,
                        double* results) {
    Layout l;
    std::string err;
    if (!make_layout(domain, spacing, SIZE_MAX, l, err)) return ORACLE_ERR_INVALID;
    Grid g(l);
    for (size_t q = 0; q < n_nodes; ++q) {
        if (index[q] >= l.count()) return ORACLE_ERR_INVALID;
        g.active.push_back(index[q]);
        g.mass[index[q]] = mass[q];
        g.momentum[index[q]] = v3(velocity[3 * q], velocity[3 * q + 1], velocity[3 * q + 2]);
    }
    for (int k = 0; k < n_pellets; ++k) {
        Pellet p;
        from_c(pellets + k, p);
        CouplingResult r = couple_grid(g, p, friction);
        double* o = results + 8 * k;
        for (int a = 0; a < 3; ++a) {
            o[a] = r.impulse[a];
            o[3 + a] = r.angular_impulse[a];
        }
        o[6] = r.grid_energy;
        o[7] = (double)r.constrained;
    }
    for (size_t q = 0; q < n_nodes; ++q)
        for (int a = 0; a < 3; ++a) velocity[3 * q + a] = g.momentum[index[q]][a];
    return 0;
}

double oracle_sim_mass_residual(const oracle_sim* s) {
    return s->particles.total_mass() + s->ledger.outflow_mass - s->ledger.initial_mass;
}
void oracle_sim_momentum_residual(const oracle_sim* s, double* out) {
    V3 p = total_momentum(*s);
    const oracle_ledger& l = s->ledger;
    for (int a = 0; a < 3; ++a)
        out[a] = p[a] - l.initial_momentum[a] - l.gravity_impulse[a] + l.outflow_momentum[a] - l.wall_impulse[a] -
                 l.pellet_wall_impulse[a];
}
double oracle_sim_kinetic_energy(const oracle_sim* s) { return kinetic_energy(*s); }
double oracle_sim_potential_energy(const oracle_sim* s) { return potential_energy(*s); }
double oracle_sim_elastic_energy(const oracle_sim* s) { return elastic_energy(*s); }
double oracle_sim_mechanical_energy(const oracle_sim* s) { return mechanical_energy(*s); }
int oracle_sim_audit_step(const oracle_sim* s, double dt, oracle_audit* out) { return audit_step(*s, dt, out); }

void oracle_sim_get_material(const oracle_sim* s, oracle_material* out) { *out = s->material; }

double oracle_sim_energy_residual(const oracle_sim* s) {
    return mechanical_energy(*s) - s->ledger.initial_mechanical_energy - ledgered_energy(s->ledger);
}

}  // extern "C"
