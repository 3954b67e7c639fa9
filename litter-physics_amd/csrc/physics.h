// Device-side physics of the MLS-MPM substep, shared by the kernels:
// stencils (grid.rs:151-183), the constitutive updates (constitutive.rs:280-589),
// the wall velocity law (solver.rs:1145-1155), the deposit and gather helpers
// and exact fixed-point sums.  f32 throughout, f64 where a reduction needs it,
// exact integers for the deposit and for the sums whose order varies.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>
#include <cmath>

#include "engine_internal.h"

namespace mpm {

constexpr int kBlock = 256;

struct Stencil {
    int base[3];
    float w[3][3];
    float d[3][3];  // node minus particle in cells
};

// grid.rs:151-183 in f32: base = floor(x/dx - 1/2) + 1 (padded), quadratic B-spline weights.
__device__ inline bool make_stencil(const Params& P, float x, float y, float z, Stencil& s) {
    float xs[3] = {x * P.inv_dx, y * P.inv_dx, z * P.inv_dx};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float bf = floorf(xs[a] - 0.5f);
        float padded = bf + 1.0f;
        if (!(padded >= 0.0f) || padded + 2.0f >= (float)P.nodes[a]) return false;
        float fx = xs[a] - bf;
        float t0 = 1.5f - fx, t1 = fx - 1.0f, t2 = fx - 0.5f;
        s.w[a][0] = 0.5f * (t0 * t0);
        s.w[a][1] = 0.75f - t1 * t1;
        s.w[a][2] = 0.5f * (t2 * t2);
        s.base[a] = (int)padded;
        s.d[a][0] = -fx;
        s.d[a][1] = 1.0f - fx;
        s.d[a][2] = 2.0f - fx;
    }
    return true;
}

__device__ inline float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ inline double wave_max_d(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ inline int wave_sum_i(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ inline unsigned long long wave_min_u64(unsigned long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        unsigned long long w = __shfl_xor(v, o, 64);
        v = w < v ? w : v;
    }
    return v;
}

__device__ inline bool lane0() { return (threadIdx.x & 63) == 0; }

// ------------------------------------------------------------ constitutive

__device__ inline float det3(const float* f) {
    float m23 = f[4] * f[8] - f[7] * f[5];
    float m13 = f[3] * f[8] - f[6] * f[5];
    float m12 = f[3] * f[7] - f[6] * f[4];
    return f[0] * m23 - f[1] * m13 + f[2] * m12;
}

// Cyclic Jacobi on a symmetric 3x3 (constitutive.rs:373-425), pairs (0,1),(0,2),(1,2).
// Stops when the off-diagonal norm falls below 1e-12 of the Frobenius norm (the
// f64 reference's 1e-30 is below f32 resolution).
__device__ inline void jacobi_rotate(float* a, float* v, int p, int q) {
    float apq = a[3 * p + q];
    if (apq == 0.0f) return;
    float theta = (a[3 * q + q] - a[3 * p + p]) / (2.0f * apq);
    float t = fabsf(theta) > 1e15f ? 0.5f / theta
                                   : copysignf(1.0f, theta) / (fabsf(theta) + sqrtf(theta * theta + 1.0f));
    float c = 1.0f / sqrtf(t * t + 1.0f);
    float s = t * c;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float akp = a[3 * k + p], akq = a[3 * k + q];
        a[3 * k + p] = c * akp - s * akq;
        a[3 * k + q] = s * akp + c * akq;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float apk = a[3 * p + k], aqk = a[3 * q + k];
        a[3 * p + k] = c * apk - s * aqk;
        a[3 * q + k] = s * apk + c * aqk;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float vkp = v[3 * k + p], vkq = v[3 * k + q];
        v[3 * k + p] = c * vkp - s * vkq;
        v[3 * k + q] = s * vkp + c * vkq;
    }
}

__device__ inline int eigen_sym3(float* a, float* v) {
#pragma unroll
    for (int i = 0; i < 9; ++i) v[i] = (i % 4 == 0) ? 1.0f : 0.0f;
    float scale = 0.0f;
#pragma unroll
    for (int i = 0; i < 9; ++i) scale += a[i] * a[i];
    if (!isfinite(scale)) return F_EIGEN;
    scale = fmaxf(scale, FLT_MIN);
    for (int sweep = 0; sweep < 24; ++sweep) {
        float off = a[1] * a[1] + a[2] * a[2] + a[5] * a[5];
        if (off <= 1e-12f * scale) return F_OK;
        jacobi_rotate(a, v, 0, 1);
        jacobi_rotate(a, v, 0, 2);
        jacobi_rotate(a, v, 1, 2);
    }
    float off = a[1] * a[1] + a[2] * a[2] + a[5] * a[5];
    return off <= 1e-8f * scale ? F_OK : F_EIGEN;
}

// Square root by the hardware instruction (within about 1 ulp; the library's
// correctly rounded sqrtf is a dozen instructions more), for norms whose last
// bit does not matter: the stress norm that sets the trial stress, speeds.
__device__ inline float hw_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// (x / K)^(1/n) of the Herschel-Bulkley law for x >= 0, by the hardware base-2
// logarithm and exponential (about 1 ulp each; the library powf is some
// thirty instructions more per Newton step).  0 for x = 0 and for subnormal x,
// whose power is below the return's tolerance.
__device__ inline float hb_pow(float x, float inv_n) {
    return x > 0.0f ? __builtin_amdgcn_exp2f(inv_n * __builtin_amdgcn_logf(x)) : 0.0f;
}

// Herschel-Bulkley implicit return (constitutive.rs:280-362), solved for the
// overstress d = q - sigma_y on [0, q_tr - sigma_y] with the same bracketed Newton;
// tolerance 2.5e-7 of the trial stress (a few f32 ulps).
__device__ inline int hb_return(const Params& P, float dt, float qtr, float& q, float& inc) {
    float G3 = 3.0f * P.G;
    if (!isfinite(qtr)) return F_NONFINITE;
    if (qtr <= P.ys) {
        q = qtr;
        inc = 0.0f;
        return F_OK;
    }
    float avail = qtr - P.ys;
    if (P.K <= 0.0f) {
        q = P.ys;
        inc = avail / G3;
        return F_OK;
    }
    float vs = G3 * dt;
    float inv_k = __builtin_amdgcn_rcpf(P.K);
    float lo = 0.0f, hi = avail, d = 0.5f * avail;
    float tol = 2.5e-7f * fmaxf(qtr, 1.0f);
    // pw = (d/K)^(1/n); its derivative in d is (1/n) pw / d, so one power per iteration.
    float pw = hb_pow(d * inv_k, P.inv_n);
    float r = d - avail + vs * pw;
    for (int it = 0; it < 100; ++it) {
        if (fabsf(r) <= tol || (hi - lo) <= tol) {
            q = P.ys + d;
            inc = (avail - d) * __builtin_amdgcn_rcpf(G3);
            return F_OK;
        }
        if (r > 0.0f)
            hi = d;
        else
            lo = d;
        float deriv = d > 0.0f ? 1.0f + vs * P.inv_n * pw * __builtin_amdgcn_rcpf(d) : INFINITY;
        float nd = d - r * __builtin_amdgcn_rcpf(deriv);
        d = (isfinite(nd) && nd > lo && nd < hi) ? nd : 0.5f * (lo + hi);
        pw = hb_pow(d * inv_k, P.inv_n);
        r = d - avail + vs * pw;
        if (!isfinite(r)) return F_NONFINITE;
    }
    return F_RETURN;
}

// Trial state of update_paste (constitutive.rs:456-500) on H = F - I:
//   X = b - I = H + H^T + H H^T, Hencky e = log(I + X)/2, D = dev e (lab frame,
//   xx yy zz xy xz yz), q_tr = sqrt(3/2) 2G |D|.
// For |X| < 0.1 the logarithm is its power series (the matrix function the
// reference evaluates in the eigenbasis, to f32 round-off, in 3-7 terms); larger
// strains take the cyclic Jacobi eigenbasis as the reference does.
struct PasteTrial {
    float D[6];
    float mean, vol, dn, qtr;
};

// Product of two symmetric 3x3 matrices that commute (powers of one matrix), as xx yy zz xy xz yz.
__device__ inline void sym_mul(const float* a, const float* b, float* c) {
    c[0] = a[0] * b[0] + a[3] * b[3] + a[4] * b[4];
    c[1] = a[3] * b[3] + a[1] * b[1] + a[5] * b[5];
    c[2] = a[4] * b[4] + a[5] * b[5] + a[2] * b[2];
    c[3] = a[0] * b[3] + a[3] * b[1] + a[4] * b[5];
    c[4] = a[0] * b[4] + a[3] * b[5] + a[4] * b[2];
    c[5] = a[3] * b[4] + a[1] * b[5] + a[5] * b[2];
}

__device__ inline float sym_norm2(const float* a) {
    return a[0] * a[0] + a[1] * a[1] + a[2] * a[2] + 2.0f * (a[3] * a[3] + a[4] * a[4] + a[5] * a[5]);
}

__device__ inline int paste_trial(const Params& P, const float* Ht, PasteTrial& s) {
    float F[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) F[i] = Ht[i] + ((i % 4 == 0) ? 1.0f : 0.0f);
    float det = det3(F);
    if (!isfinite(det)) return F_NONFINITE;
    if (det <= 0.0f) return F_INVERTED;
    float B[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            B[3 * i + j] = Ht[3 * i + j] + Ht[3 * j + i] + Ht[3 * i] * Ht[3 * j] + Ht[3 * i + 1] * Ht[3 * j + 1] +
                           Ht[3 * i + 2] * Ht[3 * j + 2];
    const float X[6] = {B[0], B[4], B[8], B[1], B[2], B[5]};
    const float rho2 = sym_norm2(X);
    if (!isfinite(rho2)) return F_NONFINITE;
    float e[6];
    if (rho2 < 0.01f) {
        // log(I + X) = X - X^2/2 + X^3/3 - ...; every eigenvalue of b lies in (0.9, 1.1),
        // so no stretch can be degenerate.
        float pw[6], lg[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) pw[i] = lg[i] = X[i];
        const float tol2 = 1e-15f * rho2;
        // (-1)^(k+1) / k for k = 2 .. 12
        constexpr float kLogC[11] = {-1.0f / 2, 1.0f / 3, -1.0f / 4, 1.0f / 5, -1.0f / 6, 1.0f / 7,
                                     -1.0f / 8, 1.0f / 9, -1.0f / 10, 1.0f / 11, -1.0f / 12};
        for (int k = 2; k <= 12; ++k) {
            float nx[6];
            sym_mul(pw, X, nx);
            const float c = kLogC[k - 2];
#pragma unroll
            for (int i = 0; i < 6; ++i) {
                pw[i] = nx[i];
                lg[i] += c * nx[i];
            }
            if (sym_norm2(pw) * c * c <= tol2) break;
        }
#pragma unroll
        for (int i = 0; i < 6; ++i) e[i] = 0.5f * lg[i];
    } else {
        float U[9];
        int rc = eigen_sym3(B, U);
        if (rc) return rc;
        float beta[3] = {B[0], B[4], B[8]};
        float ev[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float st = sqrtf(fmaxf(1.0f + beta[a], 0.0f));
            if (!isfinite(st) || st < 1e-6f) return F_STRETCH;
            ev[a] = 0.5f * log1pf(beta[a]);
        }
        const int ii[6] = {0, 1, 2, 0, 0, 1}, jj[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
        for (int t = 0; t < 6; ++t) {
            int i = ii[t], j = jj[t];
            e[t] = U[3 * i] * ev[0] * U[3 * j] + U[3 * i + 1] * ev[1] * U[3 * j + 1] + U[3 * i + 2] * ev[2] * U[3 * j + 2];
        }
    }
    s.vol = e[0] + e[1] + e[2];
    s.mean = s.vol * (1.0f / 3.0f);
    s.D[0] = e[0] - s.mean;
    s.D[1] = e[1] - s.mean;
    s.D[2] = e[2] - s.mean;
    s.D[3] = e[3];
    s.D[4] = e[4];
    s.D[5] = e[5];
    s.dn = hw_sqrt(sym_norm2(s.D));
    s.qtr = 1.2247448713915890f * 2.0f * P.G * s.dn;
    return F_OK;
}

// Whether the trial state needs the Herschel-Bulkley return (constitutive.rs:500-508).
__device__ inline bool paste_yields(const Params& P, const PasteTrial& s) { return s.qtr > P.ys && s.dn > 0.0f; }

// Rest of update_paste (constitutive.rs:508-531) for the returned equivalent
// stress q, alpha = q / q_tr:
//   F' = F + M F,  M = expm1((alpha - 1) D)   (the matrix function),
//   tau = 2G alpha D + (2G mean + lambda tr e) I.
// The reference applies exp in the trial eigenbasis; D shares that basis, so M
// is the same matrix.  For |(alpha - 1) D| < 0.1 it is the power series, else
// D is diagonalised again (cyclic Jacobi), which keeps the eigenvectors out of
// the trial state that has to stay live across the return.
__device__ inline int paste_finish(const Params& P, const PasteTrial& s, const float* Ht, float q, bool yielded,
                                   float* Hn, float* tau) {
    const float alpha = yielded ? q * __builtin_amdgcn_rcpf(s.qtr) : 1.0f;
    float M[9];
    const float am1 = alpha - 1.0f;
    float Y[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) Y[i] = am1 * s.D[i];
    const float y2 = sym_norm2(Y);
    if (y2 >= 0.01f) {
        float A[9] = {Y[0], Y[3], Y[4], Y[3], Y[1], Y[5], Y[4], Y[5], Y[2]}, U[9];
        int rc = eigen_sym3(A, U);
        if (rc) return rc;
        float rm1[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) rm1[a] = expm1f(A[4 * a]);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                M[3 * i + j] = U[3 * i] * rm1[0] * U[3 * j] + U[3 * i + 1] * rm1[1] * U[3 * j + 1] +
                               U[3 * i + 2] * rm1[2] * U[3 * j + 2];
    } else {
        // expm1(Y) = Y + Y^2/2! + Y^3/3! + ..., |Y| < 0.1
        float pw[6], ex[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) pw[i] = ex[i] = Y[i];
        const float tol2 = 1e-15f * y2;
        // 1 / k! for k = 2 .. 10
        constexpr float kExpC[9] = {1.0f / 2, 1.0f / 6, 1.0f / 24, 1.0f / 120, 1.0f / 720, 1.0f / 5040,
                                    1.0f / 40320, 1.0f / 362880, 1.0f / 3628800};
        for (int k = 2; k <= 10 && tol2 > 0.0f; ++k) {
            float nx[6];
            sym_mul(pw, Y, nx);
            const float c = kExpC[k - 2];
#pragma unroll
            for (int i = 0; i < 6; ++i) {
                pw[i] = nx[i];
                ex[i] += c * nx[i];
            }
            if (sym_norm2(pw) * c * c <= tol2) break;
        }
        M[0] = ex[0];
        M[4] = ex[1];
        M[8] = ex[2];
        M[1] = M[3] = ex[3];
        M[2] = M[6] = ex[4];
        M[5] = M[7] = ex[5];
    }
    // H' = H + M (I + H)
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            Hn[3 * i + j] = Ht[3 * i + j] + M[3 * i + j] + M[3 * i] * Ht[j] + M[3 * i + 1] * Ht[3 + j] +
                            M[3 * i + 2] * Ht[6 + j];
    const float g2a = 2.0f * P.G * alpha;
    const float iso = 2.0f * P.G * s.mean + P.lam * s.vol;
#pragma unroll
    for (int t = 0; t < 6; ++t) {
        tau[t] = g2a * s.D[t] + (t < 3 ? iso : 0.0f);
        if (!isfinite(tau[t])) return F_NONFINITE;
    }
    return F_OK;
}

// update_paste (constitutive.rs:456-531): trial state, exact return, finish.
__device__ inline int paste_update(const Params& P, float dt, const float* Ht, float* Hn, float* tau, float& inc,
                                   float& dd) {
    PasteTrial s;
    int rc = paste_trial(P, Ht, s);
    if (rc) return rc;
    bool y = paste_yields(P, s);
    float q = s.qtr;
    inc = 0.0f;
    if (y) {
        rc = hb_return(P, dt, s.qtr, q, inc);
        if (rc) return rc;
    }
    rc = paste_finish(P, s, Ht, q, y, Hn, tau);
    if (rc) return rc;
    dd = q * inc;
    return F_OK;
}

// ------------------------------------------------------------ surrogate encoding

// Inputs of the return-map surrogate (12, all O(1)):
//   0-5  trial deviatoric Kirchhoff stress / q_tr (xx, yy, zz, xy, xz, yz, lab frame)
//   6    log10(r) / 4, r = (q_tr - sigma_y) / (3 G dt): the plastic strain rate a full
//        return would need, in units of the elastic rate
//   7    log1p(1e4 dt eps_dot) / 5, eps_dot = sqrt(2/3 dev(D):dev(D)), D = sym(C)
//   8    asinh(p / sigma_y) / 4, p = -tr(tau_tr) / 3
//   9    sigma_y / q_tr
//   10   log10(K / (3 G dt)) / 4
//   11   n
// With r, K / (3 G dt) and n the return is fully determined (the dimensionless
// overstress y solves y + beta y^(1/n) = 1, beta = (1/r) (r / (K / 3G dt))^(1/n)).
__device__ inline void surrogate_features(const Params& P, float dt, const PasteTrial& s, const float* C, float* f) {
    const float sc = 2.0f * P.G / s.qtr;
#pragma unroll
    for (int t = 0; t < 6; ++t) f[t] = sc * s.D[t];
    const float s3 = 3.0f * P.G * dt;
    f[6] = log10f(fmaxf((s.qtr - P.ys) / s3, 1e-30f)) * 0.25f;
    float tr = (C[0] + C[4] + C[8]) / 3.0f;
    float dxx = C[0] - tr, dyy = C[4] - tr, dzz = C[8] - tr;
    float dxy = 0.5f * (C[1] + C[3]), dxz = 0.5f * (C[2] + C[6]), dyz = 0.5f * (C[5] + C[7]);
    float ee = sqrtf((2.0f / 3.0f) * (dxx * dxx + dyy * dyy + dzz * dzz + 2.0f * (dxy * dxy + dxz * dxz + dyz * dyz)));
    f[7] = log1pf(1e4f * dt * ee) * 0.2f;
    float p = -(2.0f * P.G / 3.0f + P.lam) * s.vol;
    f[8] = asinhf(p / P.ys) * 0.25f;
    f[9] = P.ys / s.qtr;
    f[10] = log10f(P.K / s3) * 0.25f;
    f[11] = 1.0f / P.inv_n;
}

// Targets (7): the corrected deviatoric stress / q_tr (alpha times inputs 0-5)
// and y = (q - sigma_y) / (q_tr - sigma_y) = 1 - 3G dlambda / (q_tr - sigma_y),
// the plastic multiplier in units of the full return.
__device__ inline void surrogate_targets(const Params& P, const PasteTrial& s, const float* f, float q, float* t) {
    float alpha = q / s.qtr;
#pragma unroll
    for (int i = 0; i < 6; ++i) t[i] = alpha * f[i];
    t[6] = (q - P.ys) / (s.qtr - P.ys);
}

// `steps` safeguarded Newton steps of the return equation from the predicted
// overstress fraction y, and the residual |f(q)| / max(q_tr, 1) at the result
// (constitutive.rs:301-319, the same equation hb_return solves).  One power per
// step: with pw = (d/K)^(1/n), f'(d) = 1 + 3G dt (1/n) pw / d.  (Stopping as
// soon as the residual passes the acceptance gate left accepted returns up to
// the gate's 1e-4 from the root instead of converged, and put the 256-instance
// sweep's largest stress difference to the exact batch at 9e-3.)
__device__ inline float surrogate_polish(const Params& P, float dt, float qtr, float y, int steps, float& q) {
    const float s3 = 3.0f * P.G * dt, inv_k = 1.0f / P.K;
    const float avail = qtr - P.ys;
    float d = avail * fminf(fmaxf(y, 0.0f), 1.0f);
    float pw = hb_pow(d * inv_k, P.inv_n);
    for (int it = 0; it < steps; ++it) {
        const float r = d - avail + s3 * pw;
        const float deriv = d > 0.0f ? 1.0f + s3 * P.inv_n * pw / d : INFINITY;
        const float nd = d - r / deriv;
        d = (isfinite(nd) && nd > 0.0f && nd < avail) ? nd : d;
        pw = hb_pow(d * inv_k, P.inv_n);
    }
    q = P.ys + d;
    return fabsf(d - avail + s3 * pw) / fmaxf(qtr, 1.0f);
}

// log1p and expm1 of the Tait pressure by their Taylor series (Horner) where the
// argument is small, the common case of a weakly compressible liquid (|J - 1|
// of a percent or less); the library functions beyond.  The first omitted
// terms are below 2^-40 |x| (log1p, |x| <= 1/16, ten terms) and 1e-13 |y|
// (expm1, |y| <= 1/2, twelve terms).
__device__ inline float tait_log1p(float x) {
    if (!(fabsf(x) <= 0.0625f)) return log1pf(x);
    float t = -1.0f / 10.0f;
    t = fmaf(t, x, 1.0f / 9.0f);
    t = fmaf(t, x, -1.0f / 8.0f);
    t = fmaf(t, x, 1.0f / 7.0f);
    t = fmaf(t, x, -1.0f / 6.0f);
    t = fmaf(t, x, 1.0f / 5.0f);
    t = fmaf(t, x, -1.0f / 4.0f);
    t = fmaf(t, x, 1.0f / 3.0f);
    t = fmaf(t, x, -1.0f / 2.0f);
    t = fmaf(t, x, 1.0f);
    return t * x;
}
__device__ inline float tait_expm1(float y) {
    if (!(fabsf(y) <= 0.5f)) return expm1f(y);
    float t = 1.0f / 479001600.0f;  // 1/12!
    t = fmaf(t, y, 1.0f / 39916800.0f);
    t = fmaf(t, y, 1.0f / 3628800.0f);
    t = fmaf(t, y, 1.0f / 362880.0f);
    t = fmaf(t, y, 1.0f / 40320.0f);
    t = fmaf(t, y, 1.0f / 5040.0f);
    t = fmaf(t, y, 1.0f / 720.0f);
    t = fmaf(t, y, 1.0f / 120.0f);
    t = fmaf(t, y, 1.0f / 24.0f);
    t = fmaf(t, y, 1.0f / 6.0f);
    t = fmaf(t, y, 1.0f / 2.0f);
    t = fmaf(t, y, 1.0f);
    return t * y;
}

// update_liquid (constitutive.rs:557-589) on J - 1, with the Tait pressure
// p = (K/g) expm1(-g log1p(J - 1)); g = 1 gives the reference's K (1/J - 1).
__device__ inline int liquid_update(const Params& P, float dt, float jm1, const float* C, float& jn, float* tau) {
    float tr = C[0] + C[4] + C[8];
    float g = dt * tr;
    jn = jm1 + g + g * jm1;
    float J = 1.0f + jn;
    if (!isfinite(jn)) return F_NONFINITE;
    if (J <= 1e-6f) return F_INVERTED;
    float p = P.gamma == 1.0f ? -P.bulk * jn / J : (P.bulk / P.gamma) * tait_expm1(-P.gamma * tait_log1p(jn));
    float c2 = 2.0f * P.visc * J;
    float m3 = tr / 3.0f;
    float pj = -p * J;
    tau[0] = pj + c2 * (C[0] - m3);
    tau[1] = pj + c2 * (C[4] - m3);
    tau[2] = pj + c2 * (C[8] - m3);
    tau[3] = c2 * 0.5f * (C[1] + C[3]);
    tau[4] = c2 * 0.5f * (C[2] + C[6]);
    tau[5] = c2 * 0.5f * (C[5] + C[7]);
#pragma unroll
    for (int t = 0; t < 6; ++t)
        if (!isfinite(tau[t])) return F_NONFINITE;
    return F_OK;
}

// Within half a cell of a face: the only particles the wall traction can touch.
__device__ inline bool near_wall(const Params& P, float x, float y, float z) {
    float half = 0.5f * P.dx;
    return x < half || y < half || z < half || x > P.extent[0] - half || y > P.extent[1] - half ||
           z > P.extent[2] - half;
}

__device__ inline void reduce_tangential(float* v, float reduction) {
    if (reduction == 0.0f) return;
    float speed = hw_sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (speed <= reduction) {
        v[0] = v[1] = v[2] = 0.0f;
    } else {
        float f = (speed - reduction) / speed;
        v[0] *= f;
        v[1] *= f;
        v[2] *= f;
    }
}

// v from the lane CTRL names (a DPP row shift or row broadcast) in the rows
// ROWS enables, 0 elsewhere and where the source lies outside the row.
template <int CTRL, int ROWS>
__device__ inline double dpp_take(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWS, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWS, 0xF, true);
    return __hiloint2double(hi, lo);
}

// Sum over the wavefront in a fixed order, left in lane 63: row prefix sums by
// DPP row shifts, then the row totals by row broadcasts (VALU moves only, no
// LDS permutes).
__device__ inline double wave_sum_last(double v) {
    v += dpp_take<0x111, 0xF>(v);  // row_shr:1
    v += dpp_take<0x112, 0xF>(v);  // row_shr:2
    v += dpp_take<0x114, 0xF>(v);  // row_shr:4
    v += dpp_take<0x118, 0xF>(v);  // row_shr:8: lane 15 of each row holds the row's sum
    v += dpp_take<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
    v += dpp_take<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
    return v;
}

// The same fixed-order sum for 64-bit and 32-bit integers (wrapping), left in
// lane 63.
template <int CTRL, int ROWS>
__device__ inline unsigned long long dpp_take_u64(unsigned long long v) {
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)v, CTRL, ROWS, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(v >> 32), CTRL, ROWS, 0xF, true);
    return ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
}
__device__ inline unsigned long long wave_sum_u64_last(unsigned long long v) {
    v += dpp_take_u64<0x111, 0xF>(v);
    v += dpp_take_u64<0x112, 0xF>(v);
    v += dpp_take_u64<0x114, 0xF>(v);
    v += dpp_take_u64<0x118, 0xF>(v);
    v += dpp_take_u64<0x142, 0xA>(v);
    v += dpp_take_u64<0x143, 0xC>(v);
    return v;
}
__device__ inline int wave_sum_i_last(int v) {
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xF, 0xF, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xF, 0xF, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xA, 0xF, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xC, 0xF, true);
    return v;
}

// Wavefront minimum and maximum by the same DPP pattern, returned as a uniform
// (scalar) value.  Lanes without a DPP source keep their own value.  Every lane
// of the wavefront must be active.
template <int CTRL, int ROWS>
__device__ inline int dpp_self(int v) {
    return __builtin_amdgcn_update_dpp(v, v, CTRL, ROWS, 0xF, false);
}

template <int CTRL, int ROWS>
__device__ inline unsigned long long dpp_self_u64(unsigned long long v) {
    const int lo = (int)(unsigned)v, hi = (int)(unsigned)(v >> 32);
    const int l = __builtin_amdgcn_update_dpp(lo, lo, CTRL, ROWS, 0xF, false);
    const int h = __builtin_amdgcn_update_dpp(hi, hi, CTRL, ROWS, 0xF, false);
    return ((unsigned long long)(unsigned)h << 32) | (unsigned)l;
}

// Wavefront minimum of 64-bit integers, left in lane 63.
__device__ inline unsigned long long wave_min_u64_last(unsigned long long v) {
    unsigned long long w;
    w = dpp_self_u64<0x111, 0xF>(v);
    v = w < v ? w : v;
    w = dpp_self_u64<0x112, 0xF>(v);
    v = w < v ? w : v;
    w = dpp_self_u64<0x114, 0xF>(v);
    v = w < v ? w : v;
    w = dpp_self_u64<0x118, 0xF>(v);
    v = w < v ? w : v;
    w = dpp_self_u64<0x142, 0xA>(v);
    v = w < v ? w : v;
    w = dpp_self_u64<0x143, 0xC>(v);
    v = w < v ? w : v;
    return v;
}

__device__ inline int wave_min_uniform(int v) {
    v = min(v, dpp_self<0x111, 0xF>(v));
    v = min(v, dpp_self<0x112, 0xF>(v));
    v = min(v, dpp_self<0x114, 0xF>(v));
    v = min(v, dpp_self<0x118, 0xF>(v));
    v = min(v, dpp_self<0x142, 0xA>(v));
    v = min(v, dpp_self<0x143, 0xC>(v));
    return __builtin_amdgcn_readlane(v, 63);
}

__device__ inline int wave_max_uniform(int v) {
    v = max(v, dpp_self<0x111, 0xF>(v));
    v = max(v, dpp_self<0x112, 0xF>(v));
    v = max(v, dpp_self<0x114, 0xF>(v));
    v = max(v, dpp_self<0x118, 0xF>(v));
    v = max(v, dpp_self<0x142, 0xA>(v));
    v = max(v, dpp_self<0x143, 0xC>(v));
    return __builtin_amdgcn_readlane(v, 63);
}

// fmaxf over the wavefront (NaN lanes ignored, as fmaxf does).
__device__ inline float wave_fmax_uniform(float v) {
    v = fmaxf(v, __int_as_float(dpp_self<0x111, 0xF>(__float_as_int(v))));
    v = fmaxf(v, __int_as_float(dpp_self<0x112, 0xF>(__float_as_int(v))));
    v = fmaxf(v, __int_as_float(dpp_self<0x114, 0xF>(__float_as_int(v))));
    v = fmaxf(v, __int_as_float(dpp_self<0x118, 0xF>(__float_as_int(v))));
    v = fmaxf(v, __int_as_float(dpp_self<0x142, 0xA>(__float_as_int(v))));
    v = fmaxf(v, __int_as_float(dpp_self<0x143, 0xC>(__float_as_int(v))));
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// q / d for 0 <= q < 2^20 and a divisor d > 0 given as inv = 1/d in f32: the
// quotient plus 1/2d lies at least 1/2d from an integer, far beyond the few-ulp
// error of the product, so the conversion truncates to the exact quotient.
// Three VALU instructions instead of an integer division sequence.
__device__ inline int div_small(int q, float inv) { return (int)(((float)q + 0.5f) * inv); }

// ------------------------------------------------------------ tiles

// LDS tile of one work item: the nodes of the stencils of its particles (the
// box of their stencil bases from the set-up record, plus two nodes).
struct Tile {
    int org[3];  // origin (lowest stencil base)
    int gd[3];   // dims
    int tiled;
};

// grid_to_particle's sums (solver.rs:905-930): v = sum_i w_i v_i and
// B = sum_i w_i v_i d_i^T over the 27 stencil nodes (d in cells).  The weights
// and offsets factor by axis, so the sums run axis by axis: over z for each
// (x, y) column, then y, then x -- about 290 VALU operations instead of 440.
// `node(a, b, c)` returns the velocity of stencil node (a, b, c).
template <class Node>
__device__ inline void gather_sums(const Stencil& s, const Node& node, float* v, float* B) {
    float wdz[3], wdy[3], wdx[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        wdz[c] = s.w[2][c] * s.d[2][c];
        wdy[c] = s.w[1][c] * s.d[1][c];
        wdx[c] = s.w[0][c] * s.d[0][c];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        v[i] = 0.0f;
#pragma unroll
        for (int j = 0; j < 3; ++j) B[3 * i + j] = 0.0f;
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float T[3] = {0.0f, 0.0f, 0.0f}, Ty[3] = {0.0f, 0.0f, 0.0f}, Tz[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            float S[3] = {0.0f, 0.0f, 0.0f}, Sz[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float3 g = node(a, b, c);
                S[0] = fmaf(s.w[2][c], g.x, S[0]);
                S[1] = fmaf(s.w[2][c], g.y, S[1]);
                S[2] = fmaf(s.w[2][c], g.z, S[2]);
                Sz[0] = fmaf(wdz[c], g.x, Sz[0]);
                Sz[1] = fmaf(wdz[c], g.y, Sz[1]);
                Sz[2] = fmaf(wdz[c], g.z, Sz[2]);
            }
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                T[i] = fmaf(s.w[1][b], S[i], T[i]);
                Ty[i] = fmaf(wdy[b], S[i], Ty[i]);
                Tz[i] = fmaf(s.w[1][b], Sz[i], Tz[i]);
            }
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            v[i] = fmaf(s.w[0][a], T[i], v[i]);
            B[3 * i] = fmaf(wdx[a], T[i], B[3 * i]);
            B[3 * i + 1] = fmaf(s.w[0][a], Ty[i], B[3 * i + 1]);
            B[3 * i + 2] = fmaf(s.w[0][a], Tz[i], B[3 * i + 2]);
        }
    }
}

__device__ inline void gather_global(const Params& P, const float* __restrict__ gv, size_t na, const Stencil& s,
                                     float* v, float* B) {
    gather_sums(s, [&](int a, int b, int c) {
        const int ni = node_index(P, s.base[0] + a, s.base[1] + b, s.base[2] + c);
        return make_float3(gv[ni], gv[na + ni], gv[2 * na + ni]);
    }, v, B);
}

__device__ inline void gather_lds(const float4* sg, const Tile& t, const Stencil& s, float* v, float* B) {
    int li = s.base[0] - t.org[0], lj = s.base[1] - t.org[1], lk = s.base[2] - t.org[2];
    // Tile offsets are far below 2^23: 24-bit multiplies run at full rate.
    const int row0 = __mul24(__mul24(li, t.gd[1]) + lj, t.gd[2]) + lk;
    const int s12 = __mul24(t.gd[1], t.gd[2]);
    gather_sums(s, [&](int a, int b, int c) {
        const float4 g = sg[row0 + a * s12 + b * t.gd[2] + c];
        return make_float3(g.x, g.y, g.z);
    }, v, B);
}

// A wavefront-uniform copy of a struct read from memory: every 32-bit word goes
// through readfirstlane, so the compiler knows the fields are uniform and keeps
// them, and every address derived from them, in scalar registers.
template <class T>
__device__ inline T uniform_copy(const T& src) {
    static_assert(sizeof(T) % 4 == 0, "uniform_copy copies 32-bit words");
    int w[sizeof(T) / 4];
    __builtin_memcpy(w, &src, sizeof(T));
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) w[i] = __builtin_amdgcn_readfirstlane(w[i]);
    T out;
    __builtin_memcpy(&out, w, sizeof(T));
    return out;
}

// Select among three register values by a runtime index (no array indexing, which
// would place the stencil in scratch memory).
__device__ inline float sel3(int i, float x0, float x1, float x2) { return i == 0 ? x0 : (i == 1 ? x1 : x2); }

// ------------------------------------------------------------ exact sums

// Exact fixed point (Fix128, units of 2^-kFixBits) of a finite double with
// |x| < 2^27, truncated toward zero below the unit.
__device__ inline Fix128 fix_from(double x) {
    const bool neg = x < 0.0;
    const double a = fabs(x) * 0x1p100;  // kFixBits
    const double hd = floor(a * 0x1p-64);
    const double ld = a - hd * 0x1p64;  // a mod 2^64: a subset of a's mantissa bits, exact
    unsigned long long lo = (unsigned long long)ld, hi = (unsigned long long)hd;
    if (neg) {
        lo = ~lo + 1ull;
        hi = ~hi + (lo == 0ull ? 1ull : 0ull);
    }
    return Fix128{lo, (long long)hi};
}

__device__ inline void fix_add(Fix128& a, const Fix128& b) {
    const unsigned long long lo = a.lo + b.lo;
    a.hi += b.hi + (lo < a.lo ? 1ll : 0ll);
    a.lo = lo;
}

// The Fix128 value of three limb sums l0 + l1 2^32 + h 2^64 (modulo 2^128),
// each limb a sum of fewer than 2^32 terms below 2^32 (l0, l1) or modulo 2^64 (h).
__device__ inline Fix128 fix_from_limbs(unsigned long long l0, unsigned long long l1, unsigned long long h) {
    Fix128 r;
    r.lo = l0 + (l1 << 32);
    const unsigned long long carry = r.lo < l0 ? 1ull : 0ull;
    r.hi = (long long)(h + (l1 >> 32) + carry);
    return r;
}

// The double nearest the fixed-point value (two roundings, a function of the
// integer alone).
__device__ inline double fix_value(const Fix128& a) {
    const bool neg = a.hi < 0;
    unsigned long long lo = a.lo, hi = (unsigned long long)a.hi;
    if (neg) {
        lo = ~lo + 1ull;
        hi = ~hi + (lo == 0ull ? 1ull : 0ull);
    }
    const double v = ((double)hi * 0x1p64 + (double)lo) * 0x1p-100;
    return neg ? -v : v;
}

// Sum over the wavefront (every lane gets it).
__device__ inline Fix128 wave_fix_sum(Fix128 v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        Fix128 w;
        w.lo = __shfl_xor(v.lo, o, 64);
        w.hi = __shfl_xor(v.hi, o, 64);
        fix_add(v, w);
    }
    return v;
}

// ------------------------------------------------------------ deposit

// Largest weight of the quadratic B-spline stencil: 0.75^3.
constexpr float kMaxWeight = 0.421875f;

// Per-particle bounds of the deposit (solver.rs:643-655): every momentum
// contribution w (m v_a + sum_b A_ab d_b), A = m C - dt (4/dx^2) V0 tau and
// |d_b| <= 1.5 dx, is at most kMaxWeight (b1 + dt b2) with
//   b1 = m (max_a |v_a| + 1.5 dx max_a sum_b |C_ab|),
//   b2 = (6/dx) V0 max_a sum_b |tau_ab|,
// and every wall-traction reaction at most kMaxWeight dt (b2 + 7 m |g|).  No
// contraction: the trial step that writes a state and the set-up kernel that
// rereads it get the same bits.
__device__ inline float particle_bound_motion(const Params& P, float m, const float* v, const float* C) {
#pragma clang fp contract(off)
    const float vm = fmaxf(fmaxf(fabsf(v[0]), fabsf(v[1])), fabsf(v[2]));
    float cm = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) cm = fmaxf(cm, (fabsf(C[3 * a]) + fabsf(C[3 * a + 1])) + fabsf(C[3 * a + 2]));
    return m * (vm + (1.5f * P.dx) * cm);
}
__device__ inline float particle_bound_stress(const Params& P, float v0, const float* tau) {
#pragma clang fp contract(off)
    const float r0 = (fabsf(tau[0]) + fabsf(tau[3])) + fabsf(tau[4]),
                r1 = (fabsf(tau[3]) + fabsf(tau[1])) + fabsf(tau[5]),
                r2 = (fabsf(tau[4]) + fabsf(tau[5])) + fabsf(tau[2]);
    return (6.0f * P.inv_dx) * (v0 * fmaxf(fmaxf(r0, r1), r2));
}
__device__ inline void particle_bounds(const Params& P, float m, float v0, const float* v, const float* C,
                                       const float* tau, float& b1, float& b2) {
    b1 = particle_bound_motion(P, m, v, C);
    b2 = particle_bound_stress(P, v0, tau);
}

// |v| as the stability limit takes it (no contraction, as particle_bounds).
__device__ inline float particle_speed(const float* v) {
#pragma clang fp contract(off)
    return hw_sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
}

// 2^k with kMaxWeight-weighted contributions of a particle within `bound`
// below 2^29 units (bound < 2^e): a quarter of the i32 range left for the
// f32 rounding of the contributions.  1 for a zero bound (all contributions 0).
__device__ inline float deposit_scale(double bound) {
    if (!(bound > 0.0) || !(bound < 1e300)) return 1.0f;
    int e;
    frexp((double)kMaxWeight * bound, &e);
    return ldexpf(1.0f, max(min(29 - e, 126), -126));
}

// The trial's deposit scales (InstState::sc_*), read once per workgroup.
struct DepositScales {
    float sc_m, sc_p, sc_mom, sc_f;
};

// One particle's deposit in scaled units.  NF = 4: mass and momentum with the
// stress impulse (the reference's P2G once dt is known); NF = 7 (audit): mass,
// the APIC momentum w (m v + m C d) and the stress sum w V0 tau d.
template <int NF>
struct Deposit {
    float ms;     // m sc_m
    float mv[3];  // m v_a times the momentum scale
    float A[9];   // momentum matrix times dx times its scale (d in cells)
    float T[6];   // audit: V0 tau dx sc_f (xx yy zz xy xz yz)
};

template <int NF>
__device__ inline Deposit<NF> make_deposit(const Params& P, const DepositScales& is, float dt, float m, float v0,
                                           const float* v, const float* C, const float* tau) {
#pragma clang fp contract(off)
    Deposit<NF> q;
    q.ms = m * is.sc_m;
    const float sp = NF == 7 ? is.sc_mom : is.sc_p;
    const float mdx = m * P.dx;
#pragma unroll
    for (int a = 0; a < 3; ++a) q.mv[a] = (m * v[a]) * sp;
    if constexpr (NF == 7) {
#pragma unroll
        for (int i = 0; i < 9; ++i) q.A[i] = (C[i] * mdx) * sp;
        const float vdx = (v0 * P.dx) * is.sc_f;
#pragma unroll
        for (int i = 0; i < 6; ++i) q.T[i] = tau[i] * vdx;
    } else {
        // dt (4/dx^2) V0 tau dx = dt (4/dx) V0 tau
        const float k = (dt * (4.0f * P.inv_dx)) * v0;
        const int sym[9] = {0, 3, 4, 3, 1, 5, 4, 5, 2};
#pragma unroll
        for (int i = 0; i < 9; ++i) q.A[i] = (C[i] * mdx - tau[sym[i]] * k) * sp;
    }
    return q;
}

// Round to the nearest integer (halves up) in one conversion.
__device__ inline int round_int(float x) {
    int q;
    asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(q) : "v"(x));
    return q;
}

// deposit_particles (solver.rs:632-661) of one particle: the integer
// contribution of every field to every stencil node, handed to `sink(field, a,
// b, c, value)`.  Mass rounds up, so every node the particle reaches with a
// nonzero weight is marked occupied, as its f64 mass would be; momentum
// rounds half up.  The tile and the global paths both come through here, so a
// particle gives the same integers whichever path it takes.  Lanes walk the
// nine (a, b) stencil columns from column (ra, rb), both indices advancing
// modulo 3, so that particles of one cell, which share most nodes, hit
// different LDS words at the same time (ra = rot / 3, rb = rot % 3 with
// rot = 5 lane mod 9).  The weights and offsets along x and y are rotated
// once per particle, so the nine unrolled columns index them by constants.
template <int NF, class Sink>
__device__ inline void deposit_particle(const Stencil& s, const Deposit<NF>& q, int ra, int rb, Sink& sink) {
#pragma clang fp contract(off)
    float wx[3], dx[3], wy[3], dy[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int a = ra + i >= 3 ? ra + i - 3 : ra + i, b = rb + i >= 3 ? rb + i - 3 : rb + i;
        wx[i] = sel3(a, s.w[0][0], s.w[0][1], s.w[0][2]);
        dx[i] = sel3(a, s.d[0][0], s.d[0][1], s.d[0][2]);
        wy[i] = sel3(b, s.w[1][0], s.w[1][1], s.w[1][2]);
        dy[i] = sel3(b, s.d[1][0], s.d[1][1], s.d[1][2]);
    }
    sink.rotate(ra, rb);
#pragma unroll
    for (int ka = 0; ka < 3; ++ka) {
        const float px0 = fmaf(q.A[0], dx[ka], q.mv[0]), px1 = fmaf(q.A[3], dx[ka], q.mv[1]),
                    px2 = fmaf(q.A[6], dx[ka], q.mv[2]);
#pragma unroll
        for (int kb = 0; kb < 3; ++kb) {
            const float wab = wx[ka] * wy[kb];
            const float p0 = fmaf(q.A[1], dy[kb], px0), p1 = fmaf(q.A[4], dy[kb], px1),
                        p2 = fmaf(q.A[7], dy[kb], px2);
            float f0 = 0.0f, f1 = 0.0f, f2 = 0.0f;
            if constexpr (NF == 7) {
                // rows of the symmetric stress: (xx xy xz), (xy yy yz), (xz yz zz)
                f0 = fmaf(q.T[3], dy[kb], q.T[0] * dx[ka]);
                f1 = fmaf(q.T[1], dy[kb], q.T[3] * dx[ka]);
                f2 = fmaf(q.T[5], dy[kb], q.T[4] * dx[ka]);
            }
            sink.column(ka, kb);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float w = wab * s.w[2][c];
                const float dz = s.d[2][c];
                // The mass is never negative: its 64-bit word needs no sign extension.
                sink.add_unsigned(0, c, (unsigned)ceilf(w * q.ms));
                sink.add(1, c, round_int(w * fmaf(q.A[2], dz, p0)));
                sink.add(2, c, round_int(w * fmaf(q.A[5], dz, p1)));
                sink.add(3, c, round_int(w * fmaf(q.A[8], dz, p2)));
                if constexpr (NF == 7) {
                    sink.add(4, c, round_int(w * fmaf(q.T[4], dz, f0)));
                    sink.add(5, c, round_int(w * fmaf(q.T[5], dz, f1)));
                    sink.add(6, c, round_int(w * fmaf(q.T[2], dz, f2)));
                }
            }
        }
    }
}

// Deposit into the LDS tile: NF fields of CAP 64-bit words (the exact integer
// sums of one workgroup), node (i, j, k) of the tile at (i sd1 + j) sd2 + k.
template <int NF, int CAP>
struct TileSink {
    unsigned long long* ss;
    int row0, s12, sd2, row;
    int oa[3], ob[3];  // row offsets of the rotated columns
    __device__ inline void rotate(int ra, int rb) {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const int a = ra + i >= 3 ? ra + i - 3 : ra + i, b = rb + i >= 3 ? rb + i - 3 : rb + i;
            oa[i] = row0 + (int)__umul24((unsigned)a, (unsigned)s12);
            ob[i] = (int)__umul24((unsigned)b, (unsigned)sd2);
        }
    }
    __device__ inline void column(int ka, int kb) { row = oa[ka] + ob[kb]; }
    __device__ inline void add(int f, int c, int v) {
        atomicAdd(ss + f * CAP + row + c, (unsigned long long)(long long)v);
    }
    __device__ inline void add_unsigned(int f, int c, unsigned v) {
        atomicAdd(ss + f * CAP + row + c, (unsigned long long)v);
    }
};

// Deposit straight into the global accumulators (64-bit atomics), for particles
// outside their workgroup's tile.
struct GlobalSink {
    long long* v;  // the accumulator set, already offset to the instance's first node
    size_t na;
    const Params* P;
    int base[3];
    int ni[3];
    int ra, rb;
    __device__ inline void rotate(int ra_, int rb_) {
        ra = ra_;
        rb = rb_;
    }
    __device__ inline void column(int ka, int kb) {
        const int a = ra + ka >= 3 ? ra + ka - 3 : ra + ka, b = rb + kb >= 3 ? rb + kb - 3 : rb + kb;
#pragma unroll
        for (int c = 0; c < 3; ++c) ni[c] = node_index(*P, base[0] + a, base[1] + b, base[2] + c);
    }
    __device__ inline void add(int f, int c, int val) {
        if (val) atomicAdd(reinterpret_cast<unsigned long long*>(v + f * na + ni[c]), (unsigned long long)(long long)val);
    }
    __device__ inline void add_unsigned(int f, int c, unsigned val) {
        if (val) atomicAdd(reinterpret_cast<unsigned long long*>(v + f * na + ni[c]), (unsigned long long)val);
    }
};

// Occupancy flags of the (up to 8) node blocks a stencil touches.
__device__ inline void flag_stencil(const Params& P, int* flag, const Stencil& s) {
    int b0[3], b1[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        b0[a] = s.base[a] >> 2;
        b1[a] = (s.base[a] + 2) >> 2;
    }
    for (int bi = b0[0]; bi <= b1[0]; ++bi)
        for (int bj = b0[1]; bj <= b1[1]; ++bj)
            for (int bk = b0[2]; bk <= b1[2]; ++bk) flag[(bi * P.nb[1] + bj) * P.nb[2] + bk] = 1;
}

__device__ inline int wave_min_i(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ inline int wave_max_i(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// Wave-aggregated append of this lane's index when `pred` holds.
__device__ inline void wave_append(bool pred, int value, int* list, int* count) {
    unsigned long long mask = __ballot(pred);
    if (!mask) return;
    int lane = threadIdx.x & 63;
    int leader = __ffsll((long long)mask) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(count, __popcll(mask));
    base = __builtin_amdgcn_readlane(base, leader);  // leader is uniform: no LDS permute
    if (pred) list[base + __popcll(mask & ((1ull << lane) - 1ull))] = value;
}

}  // namespace mpm
