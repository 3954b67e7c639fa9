// f64 host math shared by the engine's host side and the research layer.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

#include "../../include/mpm_engine.h"

namespace mpm_host {

// Rotation matrix (row-major) of the unit quaternion q = (i, j, k, w).
inline void qmatrix(const double* q, double r[9]) {
    double i = q[0], j = q[1], k = q[2], w = q[3];
    double ww = w * w, ii = i * i, jj = j * j, kk = k * k;
    double ij = i * j * 2.0, wk = w * k * 2.0, wj = w * j * 2.0, ik = i * k * 2.0, jk = j * k * 2.0,
           wi = w * i * 2.0;
    double m[9] = {ww + ii - jj - kk, ij - wk, wj + ik, wk + ij, ww - ii + jj - kk, jk - wi, ik - wj, wi + jk,
                   ww - ii - jj + kk};
    std::memcpy(r, m, sizeof(m));
}

// Symmetric 3x3 cyclic Jacobi in f64 (constitutive.rs:373-425), used for the
// elastic-energy diagnostic.
inline bool eigen3(const double* m, double* vals) {
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

// Pellet::angular_velocity, rigid.rs:193-200: R (I_body^-1 (R^T L)).
inline void pellet_omega(const mpm_pellet& p, double om[3]) {
    double r[9];
    qmatrix(p.orientation, r);
    const double* L = p.angular_momentum;
    double b[3] = {r[0] * L[0] + r[3] * L[1] + r[6] * L[2], r[1] * L[0] + r[4] * L[1] + r[7] * L[2],
                   r[2] * L[0] + r[5] * L[1] + r[8] * L[2]};
    for (int a = 0; a < 3; ++a) b[a] /= p.inertia_body[a];
    for (int a = 0; a < 3; ++a) om[a] = r[3 * a] * b[0] + r[3 * a + 1] * b[1] + r[3 * a + 2] * b[2];
}

// Pellet::kinetic_energy, rigid.rs:258-261.
inline double pellet_kinetic(const mpm_pellet& p) {
    double om[3];
    pellet_omega(p, om);
    const double* v = p.velocity;
    const double* L = p.angular_momentum;
    return 0.5 * p.mass * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) + 0.5 * (om[0] * L[0] + om[1] * L[1] + om[2] * L[2]);
}

}  // namespace mpm_host
