// HIP kernels of the MLS-MPM substep for CDNA4 (gfx950).
//
// One substep (core/src/mpm/solver.rs:540-611) maps to:
//   traction     -- wall_traction_to_grid (solver.rs:725-819) on particles near a face
//   grid_update  -- grid_update (solver.rs:822-892) plus the ledger sums, one wave per
//                   4x4x4-node block; also clears the other accumulator set
//   couple       -- couple_grid (rigid.rs:385-438) for every pellet, in index order
//   g2p2g        -- grid_to_particle (solver.rs:895-977) with update_paste / update_liquid
//                   (constitutive.rs:456-589), fused with the next step's
//                   deposit_particles (solver.rs:632-661) into the other accumulator set.
// The deposit is split into the dt-free APIC momentum and the dt-free stress sum so
// that it can run before the next step's dt is known.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cfloat>
#include <climits>
#include <cmath>

#include "engine_internal.h"

namespace mpm {

namespace {

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

// ------------------------------------------------------------ scatter

// deposit_particles (solver.rs:632-661) for one particle: mass w m, momentum
// w (m v + m C d) and stress w V0 tau d, d = x_node - x_p in metres.
__device__ inline void scatter_particle(const Params& P, GridPtrs g, const Stencil& s, float m, float vx, float vy,
                                        float vz, const float* C, const float* tau, float v0) {
    // Precompute m C dx and V0 tau dx (d in cells times dx is metres).
    float mc[9], vt[6];
    float mdx = m * P.dx, vdx = v0 * P.dx;
#pragma unroll
    for (int i = 0; i < 9; ++i) mc[i] = C[i] * mdx;
#pragma unroll
    for (int i = 0; i < 6; ++i) vt[i] = tau[i] * vdx;
    float mvx = m * vx, mvy = m * vy, mvz = m * vz;
    // Block flags of the (up to 8) node blocks the stencil touches.
    int b0[3], b1[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        b0[a] = s.base[a] >> 2;
        b1[a] = (s.base[a] + 2) >> 2;
    }
    for (int bi = b0[0]; bi <= b1[0]; ++bi)
        for (int bj = b0[1]; bj <= b1[1]; ++bj)
            for (int bk = b0[2]; bk <= b1[2]; ++bk) g.flag[(bi * P.nb[1] + bj) * P.nb[2] + bk] = 1;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            float wab = s.w[0][a] * s.w[1][b];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                float w = wab * s.w[2][c];
                float dx = s.d[0][a], dy = s.d[1][b], dz = s.d[2][c];
                int n = node_index(P, s.base[0] + a, s.base[1] + b, s.base[2] + c);
                float px = w * (mvx + mc[0] * dx + mc[1] * dy + mc[2] * dz);
                float py = w * (mvy + mc[3] * dx + mc[4] * dy + mc[5] * dz);
                float pz = w * (mvz + mc[6] * dx + mc[7] * dy + mc[8] * dz);
                float fx = w * (vt[0] * dx + vt[3] * dy + vt[4] * dz);
                float fy = w * (vt[3] * dx + vt[1] * dy + vt[5] * dz);
                float fz = w * (vt[4] * dx + vt[5] * dy + vt[2] * dz);
                atomicAdd(g.m + n, w * m);
                atomicAdd(g.px + n, px);
                atomicAdd(g.py + n, py);
                atomicAdd(g.pz + n, pz);
                atomicAdd(g.fx + n, fx);
                atomicAdd(g.fy + n, fy);
                atomicAdd(g.fz + n, fz);
            }
        }
    }
}

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
    float inv_k = 1.0f / P.K;
    float lo = 0.0f, hi = avail, d = 0.5f * avail;
    float tol = 2.5e-7f * fmaxf(qtr, 1.0f);
    float r = d - avail + vs * powf(d * inv_k, P.inv_n);
    for (int it = 0; it < 100; ++it) {
        if (fabsf(r) <= tol || (hi - lo) <= tol) {
            q = P.ys + d;
            inc = (avail - d) / G3;
            return F_OK;
        }
        if (r > 0.0f)
            hi = d;
        else
            lo = d;
        float over = d * inv_k;
        float deriv = over > 0.0f ? 1.0f + vs * P.inv_n * powf(over, P.inv_n - 1.0f) * inv_k : INFINITY;
        float nd = d - r / deriv;
        d = (isfinite(nd) && nd > lo && nd < hi) ? nd : 0.5f * (lo + hi);
        r = d - avail + vs * powf(d * inv_k, P.inv_n);
        if (!isfinite(r)) return F_NONFINITE;
    }
    return F_RETURN;
}

// update_paste (constitutive.rs:456-531) on H = F - I:
//   b - I = H + H^T + H H^T, Hencky e = log1p(beta)/2, radial return on dev(e),
//   F' = F + U diag(expm1((alpha - 1) dev e)) U^T F, tau = U diag(2G e' + lambda tr e) U^T.
__device__ inline int paste_update(const Params& P, float dt, const float* Ht, float* Hn, float* tau, float& inc,
                                   float& dd) {
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
    float U[9];
    int rc = eigen_sym3(B, U);
    if (rc) return rc;
    float beta[3] = {B[0], B[4], B[8]};
    float e[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float s = sqrtf(fmaxf(1.0f + beta[a], 0.0f));
        if (!isfinite(s) || s < 1e-6f) return F_STRETCH;
        e[a] = 0.5f * log1pf(beta[a]);
    }
    float vol = e[0] + e[1] + e[2];
    float mean = vol / 3.0f;
    float dev[3] = {e[0] - mean, e[1] - mean, e[2] - mean};
    float dn = sqrtf(dev[0] * dev[0] + dev[1] * dev[1] + dev[2] * dev[2]);
    float qtr = 1.2247448713915890f * 2.0f * P.G * dn;
    float alpha = 1.0f, q = qtr;
    inc = 0.0f;
    if (qtr > P.ys && dn > 0.0f) {
        rc = hb_return(P, dt, qtr, q, inc);
        if (rc) return rc;
        alpha = q / qtr;
    }
    float rm1[3], pr[3];
    float lv = P.lam * vol;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        rm1[a] = expm1f((alpha - 1.0f) * dev[a]);
        pr[a] = 2.0f * P.G * (alpha * dev[a] + mean) + lv;
    }
    // M = U diag(rm1) U^T, H' = H + M (I + H)
    float M[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            M[3 * i + j] = U[3 * i] * rm1[0] * U[3 * j] + U[3 * i + 1] * rm1[1] * U[3 * j + 1] +
                           U[3 * i + 2] * rm1[2] * U[3 * j + 2];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            Hn[3 * i + j] = Ht[3 * i + j] + M[3 * i + j] + M[3 * i] * Ht[j] + M[3 * i + 1] * Ht[3 + j] +
                            M[3 * i + 2] * Ht[6 + j];
    const int ii[6] = {0, 1, 2, 0, 0, 1}, jj[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
    for (int t = 0; t < 6; ++t) {
        int i = ii[t], j = jj[t];
        tau[t] = U[3 * i] * pr[0] * U[3 * j] + U[3 * i + 1] * pr[1] * U[3 * j + 1] + U[3 * i + 2] * pr[2] * U[3 * j + 2];
        if (!isfinite(tau[t])) return F_NONFINITE;
    }
    dd = q * inc;
    return F_OK;
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
    float p = P.gamma == 1.0f ? -P.bulk * jn / J : (P.bulk / P.gamma) * expm1f(-P.gamma * log1pf(jn));
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

// ---------------------------------------------------------------- kernels

__global__ void reset_status_kernel(StepStatus* st, int* nw_count) {
    if (threadIdx.x == 0) {
        StepStatus z = {};
        z.fail = ~0ull;
        *st = z;
        *nw_count = 0;
    }
}

__global__ __launch_bounds__(kBlock) void p2g_kernel(Params P, int n, const float* __restrict__ slab, size_t cap,
                                                     const float* __restrict__ shared, GridPtrs g) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    float x = slab[PX * cap + p], y = slab[PY * cap + p], z = slab[PZ * cap + p];
    Stencil s;
    if (!make_stencil(P, x, y, z, s)) return;
    float C[9], tau[6];
#pragma unroll
    for (int i = 0; i < 9; ++i) C[i] = slab[(C00 + i) * cap + p];
#pragma unroll
    for (int i = 0; i < 6; ++i) tau[i] = slab[(TXX + i) * cap + p];
    scatter_particle(P, g, s, shared[SMASS * cap + p], slab[VX * cap + p], slab[VY * cap + p], slab[VZ * cap + p],
                     C, tau, shared[SVOL0 * cap + p]);
}

__device__ inline bool near_wall(const Params& P, float x, float y, float z) {
    float half = 0.5f * P.dx;
    return x < half || y < half || z < half || x > P.extent[0] - half || y > P.extent[1] - half ||
           z > P.extent[2] - half;
}

// Indices of the particles within half a cell of a face (the only ones the
// wall traction can touch).
__global__ __launch_bounds__(kBlock) void nearwall_kernel(Params P, int n, const float* __restrict__ slab,
                                                          size_t cap, int* list, int* count) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    if (near_wall(P, slab[PX * cap + p], slab[PY * cap + p], slab[PZ * cap + p])) list[atomicAdd(count, 1)] = p;
}

// wall_traction_to_grid (solver.rs:725-819).  Each deposit adds to the signed
// traction impulse (trS) and the Coulomb budget (trT) of its node; the grid
// update books the telescoped kinetic-energy jump per node.
__global__ __launch_bounds__(kBlock) void traction_kernel(Params P, float dt, const int* __restrict__ list,
                                                          const int* __restrict__ count, const float* __restrict__ slab,
                                                          size_t cap, const float* __restrict__ shared, GridPtrs g,
                                                          float* trS, float* trT, size_t na) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= *count) return;
    int p = list[q];
    float pos[3] = {slab[PX * cap + p], slab[PY * cap + p], slab[PZ * cap + p]};
    float half = 0.5f * P.dx;
    float mass = shared[SMASS * cap + p], v0 = shared[SVOL0 * cap + p];
    float k4 = 4.0f * P.inv_dx * P.inv_dx;
    float force_scale = -dt * k4;
    for (int axis = 0; axis < 3; ++axis) {
        float face;
        if (pos[axis] < half)
            face = 0.0f;
        else if (pos[axis] > P.extent[axis] - half)
            face = P.extent[axis];
        else
            continue;
        float img[3] = {pos[0], pos[1], pos[2]};
        img[axis] = 2.0f * face - pos[axis];
        Stencil s;
        if (!make_stencil(P, img[0], img[1], img[2], s)) continue;
        int plane = face == 0.0f ? 1 : P.cells[axis] + 1;
        float inward = face == 0.0f ? 1.0f : -1.0f;
        float g_a = P.gravity[axis];
        float normal_stress = slab[(TXX + axis) * cap + p] * v0 - mass * g_a * (img[axis] - pos[axis]);
        if (normal_stress >= 0.0f) continue;
        float weight_momentum = mass * g_a * dt;
        const float* mom = axis == 0 ? g.px : (axis == 1 ? g.py : g.pz);
        const float* frc = axis == 0 ? g.fx : (axis == 1 ? g.fy : g.fz);
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                for (int c = 0; c < 3; ++c) {
                    int co[3] = {s.base[0] + a, s.base[1] + b, s.base[2] + c};
                    if (co[axis] <= 1 || co[axis] > P.cells[axis]) continue;
                    float w = s.w[0][a] * s.w[1][b] * s.w[2][c];
                    if (w == 0.0f) continue;
                    int pc[3] = {co[0], co[1], co[2]};
                    pc[axis] = plane;
                    int pi = node_index(P, pc[0], pc[1], pc[2]);
                    float pm = g.m[pi];
                    if (pm <= 0.0f) continue;
                    float trial = (mom[pi] - dt * k4 * frc[pi]) / pm + g_a * dt;
                    if (trial * inward >= 0.0f) continue;
                    int off = axis == 0 ? a : (axis == 1 ? b : c);
                    float distance = s.d[axis][off] * P.dx;
                    float reaction = (normal_stress * distance * force_scale + weight_momentum) * w * inward;
                    if (reaction <= 0.0f) continue;
                    int ni = node_index(P, co[0], co[1], co[2]);
                    if (g.m[ni] <= 0.0f) continue;
                    atomicAdd(trS + axis * na + ni, reaction * inward);
                    atomicAdd(trT + axis * na + ni, reaction);
                }
    }
}

__device__ inline void reduce_tangential(float* v, float reduction) {
    if (reduction == 0.0f) return;
    float speed = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (speed <= reduction) {
        v[0] = v[1] = v[2] = 0.0f;
    } else {
        float f = (speed - reduction) / speed;
        v[0] *= f;
        v[1] *= f;
        v[2] *= f;
    }
}

// Block-wide sum of NV doubles held by every thread; thread 0 gets the totals.
template <int NV>
__device__ inline void block_sum(double* vals, double* scratch) {
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; ++i) vals[i] = wave_sum(vals[i]);
    if (lane == 0)
#pragma unroll
        for (int i = 0; i < NV; ++i) scratch[wave * NV + i] = vals[i];
    __syncthreads();
    if (threadIdx.x == 0) {
        int waves = blockDim.x >> 6;
        for (int i = 0; i < NV; ++i) {
            double t = 0.0;
            for (int w = 0; w < waves; ++w) t += scratch[w * NV + i];
            vals[i] = t;
        }
    }
}

// grid_update (solver.rs:822-892).  Four 64-node blocks per workgroup, one per
// wavefront; ledger sums go to per-workgroup partials (no same-address atomics).
constexpr int kGridVals = 12;
__global__ __launch_bounds__(256) void grid_update_kernel(Params P, float dt, GridPtrs A, float* trS, float* trT,
                                                          float* gv, size_t na, GridPtrs B, int nblocks,
                                                          double* partials, int* bad_partials) {
    __shared__ double scratch[4 * kGridVals];
    __shared__ int sbad[4];
    int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    int b = blockIdx.x * 4 + wave;
    double vals[kGridVals] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    int bad = 0;
    if (b < nblocks) {
        int n = (b << 6) | l;
        if (B.flag[b]) {
            B.m[n] = 0.0f;
            B.px[n] = 0.0f;
            B.py[n] = 0.0f;
            B.pz[n] = 0.0f;
            B.fx[n] = 0.0f;
            B.fy[n] = 0.0f;
            B.fz[n] = 0.0f;
            if (l == 0) B.flag[b] = 0;
        }
        if (A.flag[b]) {
            float m = A.m[n];
            float k4 = 4.0f * P.inv_dx * P.inv_dx;
            if (m > 0.0f) {
                float p[3] = {A.px[n] - dt * k4 * A.fx[n], A.py[n] - dt * k4 * A.fy[n], A.pz[n] - dt * k4 * A.fz[n]};
                float v[3], tb[3];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    float D = trS[a * na + n];
                    tb[a] = trT[a * na + n];
                    if (D != 0.0f) {
                        // Telescoped kinetic-energy jump of this node's deposits (grid.rs:265-295).
                        vals[8] += (double)D * ((double)p[a] / (double)m + (double)D / (2.0 * (double)m));
                        vals[3 + a] += D;
                        trS[a * na + n] = 0.0f;
                    }
                    if (tb[a] != 0.0f) trT[a * na + n] = 0.0f;
                    v[a] = (p[a] + D) / m + P.gravity[a] * dt;
                }
                if (!(isfinite(v[0]) && isfinite(v[1]) && isfinite(v[2]))) {
                    bad = 1;
                    v[0] = v[1] = v[2] = 0.0f;
                } else {
                    float before[3] = {v[0], v[1], v[2]};
                    int c[3];
                    node_coords(P, n, c[0], c[1], c[2]);
                    for (int axis = 0; axis < 3; ++axis) {
                        int sign = c[axis] <= 1 ? 1 : (c[axis] > P.cells[axis] ? -1 : 0);
                        if (sign == 0) continue;
                        float ns = v[axis] * (float)sign;
                        if (ns >= 0.0f) continue;
                        vals[6] -= 0.5 * (double)m * (double)ns * (double)ns;
                        v[axis] = 0.0f;
                        double tn = (double)v[0] * v[0] + (double)v[1] * v[1] + (double)v[2] * v[2];
                        reduce_tangential(v, P.face_mu[2 * axis + (sign > 0 ? 0 : 1)] * (-ns));
                        double vn = (double)v[0] * v[0] + (double)v[1] * v[1] + (double)v[2] * v[2];
                        vals[7] += 0.5 * (double)m * (tn - vn);
                    }
                    for (int axis = 0; axis < 3; ++axis) {
                        float reaction = tb[axis];
                        if (reaction <= 0.0f) continue;
                        float normal = v[axis];
                        float t[3] = {v[0], v[1], v[2]};
                        t[axis] = 0.0f;
                        double tn = (double)t[0] * t[0] + (double)t[1] * t[1] + (double)t[2] * t[2];
                        int sign = 2 * c[axis] <= P.cells[axis] + 2 ? 1 : -1;
                        reduce_tangential(t, P.face_mu[2 * axis + (sign > 0 ? 0 : 1)] * reaction / m);
                        double rn = (double)t[0] * t[0] + (double)t[1] * t[1] + (double)t[2] * t[2];
                        vals[7] += 0.5 * (double)m * (tn - rn);
                        v[0] = t[0];
                        v[1] = t[1];
                        v[2] = t[2];
                        v[axis] = normal;
                    }
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        vals[a] = ((double)v[a] - (double)before[a]) * (double)m;
                        vals[9 + a] = (double)v[a] * (double)m;
                    }
                }
                gv[n] = v[0];
                gv[na + n] = v[1];
                gv[2 * na + n] = v[2];
            } else {
                gv[n] = 0.0f;
                gv[na + n] = 0.0f;
                gv[2 * na + n] = 0.0f;
            }
        }
    }
    int nbad = wave_sum_i(bad);
    if (l == 0) sbad[wave] = nbad;
    block_sum<kGridVals>(vals, scratch);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < kGridVals; ++i) partials[blockIdx.x * kGridVals + i] = vals[i];
        bad_partials[blockIdx.x] = sbad[0] + sbad[1] + sbad[2] + sbad[3];
    }
}

// couple_grid (rigid.rs:385-438) for every pellet listed on this node block, in
// pellet index order, so a node sees the pellets one after another as in a
// sequential loop over pellets.  Geometry and the projection run in f64 with the
// reference's operation order; node positions are (i - 1) dx exactly.
__global__ __launch_bounds__(64) void couple_kernel(Params P, double dx, GridPtrs A, float* gv, size_t na,
                                                    const PelletDev* __restrict__ pellets,
                                                    const int* __restrict__ bstart, const int* __restrict__ blist,
                                                    PelletAccum* acc, double friction) {
    int b = blockIdx.x, l = threadIdx.x;
    int n = (b << 6) | l;
    int s0 = bstart[b], s1 = bstart[b + 1];
    if (s0 == s1 || !A.flag[b]) return;
    double m = A.m[n];
    if (!(m > 0.0)) return;
    int ci, cj, ck;
    node_coords(P, n, ci, cj, ck);
    double node[3] = {((double)ci - 1.0) * dx, ((double)cj - 1.0) * dx, ((double)ck - 1.0) * dx};
    double v[3] = {gv[n], gv[na + n], gv[2 * na + n]};
    bool changed = false;
    for (int s = s0; s < s1; ++s) {
        int pi = blist[s];
        const PelletDev& pl = pellets[pi];
        if (node[0] < pl.lo[0] || node[0] > pl.hi[0] || node[1] < pl.lo[1] || node[1] > pl.hi[1] ||
            node[2] < pl.lo[2] || node[2] > pl.hi[2])
            continue;
        double best = INFINITY, nrm[3] = {0.0, 0.0, 1.0};
        for (int k = 0; k < pl.spheres; ++k) {
            double o = pl.offsets[k];
            double dlt[3] = {node[0] - (pl.pos[0] + pl.axis[0] * o), node[1] - (pl.pos[1] + pl.axis[1] * o),
                             node[2] - (pl.pos[2] + pl.axis[2] * o)};
            double dist = sqrt(dlt[0] * dlt[0] + dlt[1] * dlt[1] + dlt[2] * dlt[2]);
            double phi = dist - pl.radius;
            if (phi < best) {
                best = phi;
                if (dist > 1e-12) {
                    nrm[0] = dlt[0] / dist;
                    nrm[1] = dlt[1] / dist;
                    nrm[2] = dlt[2] / dist;
                } else {
                    nrm[0] = 0.0;
                    nrm[1] = 0.0;
                    nrm[2] = 1.0;
                }
            }
        }
        if (best > 0.0) continue;
        double r[3] = {node[0] - pl.pos[0], node[1] - pl.pos[1], node[2] - pl.pos[2]};
        double body[3] = {pl.vel[0] + (pl.omega[1] * r[2] - pl.omega[2] * r[1]),
                          pl.vel[1] + (pl.omega[2] * r[0] - pl.omega[0] * r[2]),
                          pl.vel[2] + (pl.omega[0] * r[1] - pl.omega[1] * r[0])};
        double rel[3] = {v[0] - body[0], v[1] - body[1], v[2] - body[2]};
        double ns = rel[0] * nrm[0] + rel[1] * nrm[1] + rel[2] * nrm[2];
        if (ns >= 0.0) continue;
        double t[3] = {rel[0] - nrm[0] * ns, rel[1] - nrm[1] * ns, rel[2] - nrm[2] * ns};
        double ts = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
        double red = friction * (-ns);
        double f = (ts <= red || ts < 1e-14) ? 0.0 : (ts - red) / ts;
        double nv[3] = {body[0] + t[0] * f, body[1] + t[1] * f, body[2] + t[2] * f};
        double dp[3] = {(nv[0] - v[0]) * m, (nv[1] - v[1]) * m, (nv[2] - v[2]) * m};
        double e = 0.5 * m * ((nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]) - (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
        PelletAccum* a = acc + pi;
        atomicAdd(&a->impulse[0], -dp[0]);
        atomicAdd(&a->impulse[1], -dp[1]);
        atomicAdd(&a->impulse[2], -dp[2]);
        atomicAdd(&a->angular[0], -(r[1] * dp[2] - r[2] * dp[1]));
        atomicAdd(&a->angular[1], -(r[2] * dp[0] - r[0] * dp[2]));
        atomicAdd(&a->angular[2], -(r[0] * dp[1] - r[1] * dp[0]));
        atomicAdd(&a->grid_energy, e);
        atomicAdd(&a->constrained, 1.0);
        v[0] = nv[0];
        v[1] = nv[1];
        v[2] = nv[2];
        changed = true;
    }
    if (changed) {
        gv[n] = (float)v[0];
        gv[na + n] = (float)v[1];
        gv[2 * na + n] = (float)v[2];
    }
}

// ------------------------------------------------------------ fused G2P2G

// Tile geometry of one workgroup's chunk of particles.  Gather tile: the nodes
// of every current stencil; scatter tile: one cell wider on each side, since an
// accepted step moves a particle by at most one cell.
struct Tile {
    int org[3];  // gather origin (lowest stencil base)
    int gd[3];   // gather dims
    int sd[3];   // scatter dims (origin org - 1)
    int tiled;
};

__device__ inline void gather_global(const Params& P, const float* __restrict__ gv, size_t na, const Stencil& s,
                                     float* v, float* B) {
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            float wab = s.w[0][a] * s.w[1][b];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                float w = wab * s.w[2][c];
                int ni = node_index(P, s.base[0] + a, s.base[1] + b, s.base[2] + c);
                float nv[3] = {gv[ni] * w, gv[na + ni] * w, gv[2 * na + ni] * w};
                float d[3] = {s.d[0][a], s.d[1][b], s.d[2][c]};
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    v[i] += nv[i];
#pragma unroll
                    for (int j = 0; j < 3; ++j) B[3 * i + j] += nv[i] * d[j];
                }
            }
        }
}

__device__ inline void gather_lds(const float4* sg, const Tile& t, const Stencil& s, float* v, float* B) {
    int li = s.base[0] - t.org[0], lj = s.base[1] - t.org[1], lk = s.base[2] - t.org[2];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            float wab = s.w[0][a] * s.w[1][b];
            int row = ((li + a) * t.gd[1] + (lj + b)) * t.gd[2] + lk;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                float w = wab * s.w[2][c];
                float4 g = sg[row + c];
                float nv[3] = {g.x * w, g.y * w, g.z * w};
                float d[3] = {s.d[0][a], s.d[1][b], s.d[2][c]};
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    v[i] += nv[i];
#pragma unroll
                    for (int j = 0; j < 3; ++j) B[3 * i + j] += nv[i] * d[j];
                }
            }
        }
}

// Select among three register values by a runtime index (no array indexing, which
// would place the stencil in scratch memory).
__device__ inline float sel3(int i, float x0, float x1, float x2) { return i == 0 ? x0 : (i == 1 ? x1 : x2); }

// deposit_particles for one particle into the f64 LDS accumulator tile (LDS
// f64 atomic adds run at close to full LDS rate on CDNA4; f32 ones do not).
// Lanes walk the nine (a, b) stencil columns starting at a lane-dependent
// column so that particles of one cell, which share most nodes, hit different
// LDS words.
template <int SCAP>
__device__ inline void scatter_lds(double* ss, const int* so, const int* sd, const Params& P, const Stencil& s,
                                   float m, const float* v, const float* C, const float* tau, float v0, int rot) {
    float mc[9], vt[6];
    float mdx = m * P.dx, vdx = v0 * P.dx;
#pragma unroll
    for (int i = 0; i < 9; ++i) mc[i] = C[i] * mdx;
#pragma unroll
    for (int i = 0; i < 6; ++i) vt[i] = tau[i] * vdx;
    float mv0 = m * v[0], mv1 = m * v[1], mv2 = m * v[2];
    int li = s.base[0] - so[0], lj = s.base[1] - so[1], lk = s.base[2] - so[2];
    for (int kk = 0; kk < 9; ++kk) {
        int ab = kk + rot;
        ab = ab >= 9 ? ab - 9 : ab;
        int a = ab >= 6 ? 2 : (ab >= 3 ? 1 : 0);
        int b = ab - 3 * a;
        float wab = sel3(a, s.w[0][0], s.w[0][1], s.w[0][2]) * sel3(b, s.w[1][0], s.w[1][1], s.w[1][2]);
        float dx = sel3(a, s.d[0][0], s.d[0][1], s.d[0][2]), dy = sel3(b, s.d[1][0], s.d[1][1], s.d[1][2]);
        int row = ((li + a) * sd[1] + (lj + b)) * sd[2] + lk;
        float px0 = mv0 + mc[0] * dx + mc[1] * dy, py0 = mv1 + mc[3] * dx + mc[4] * dy,
              pz0 = mv2 + mc[6] * dx + mc[7] * dy;
        float fx0 = vt[0] * dx + vt[3] * dy, fy0 = vt[3] * dx + vt[1] * dy, fz0 = vt[4] * dx + vt[5] * dy;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            float w = wab * s.w[2][c];
            float dz = s.d[2][c];
            int idx = row + c;
            atomicAdd(ss + 0 * SCAP + idx, (double)(w * m));
            atomicAdd(ss + 1 * SCAP + idx, (double)(w * (px0 + mc[2] * dz)));
            atomicAdd(ss + 2 * SCAP + idx, (double)(w * (py0 + mc[5] * dz)));
            atomicAdd(ss + 3 * SCAP + idx, (double)(w * (pz0 + mc[8] * dz)));
            atomicAdd(ss + 4 * SCAP + idx, (double)(w * (fx0 + vt[4] * dz)));
            atomicAdd(ss + 5 * SCAP + idx, (double)(w * (fy0 + vt[5] * dz)));
            atomicAdd(ss + 6 * SCAP + idx, (double)(w * (fz0 + vt[2] * dz)));
        }
    }
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
    base = __shfl(base, leader, 64);
    if (pred) list[base + __popcll(mask & ((1ull << lane) - 1ull))] = value;
}

// grid_to_particle + constitutive update (solver.rs:895-977) fused with the
// next step's particle-to-grid deposit into the other accumulator set.  A
// workgroup owns one or two cell blocks of particles (a work item).  It stages
// the node velocities of their stencils in LDS, updates every particle, then
// deposits the updated particles into an f64 LDS tile covering exactly their
// new stencils and flushes it with one global atomic per touched node and field.
// Work items whose tiles exceed the LDS budget take the direct global path.
template <int CHUNK, int GCAP, int SCAP>
__global__ __launch_bounds__(kBlock) void g2p2g_kernel(Params P, float dt, int n, const float* __restrict__ src,
                                                       float* __restrict__ dst, size_t cap,
                                                       const float* __restrict__ shared,
                                                       const float* __restrict__ gv, size_t na, GridPtrs out,
                                                       const int2* __restrict__ work, float* pmax, double* pdiss,
                                                       unsigned long long* pfail, int* pesc, int* nw_list,
                                                       int* nw_count) {
    constexpr int R = CHUNK / kBlock;
    __shared__ float4 sg[GCAP];
    __shared__ double ss[7 * SCAP];
    __shared__ int sbox[4][6];
    __shared__ Tile tile;
    __shared__ int sorg[3], sdim[3], stiled;
    __shared__ double sdiss[4];
    __shared__ float sred[4][2];
    __shared__ unsigned long long sfail[4];
    __shared__ int sesc[4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int2 item = work[blockIdx.x];
    const int chunk0 = item.x;
    const int chunk_end = min(item.x + item.y, n);

    // Phase A: stencil bases of the current positions and their bounding box.
    float px[R], py[R], pz[R];
    int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
#pragma unroll
    for (int r = 0; r < R; ++r) {
        int p = chunk0 + r * kBlock + tid;
        px[r] = py[r] = pz[r] = 0.0f;
        if (p < chunk_end) {
            px[r] = src[PX * cap + p];
            py[r] = src[PY * cap + p];
            pz[r] = src[PZ * cap + p];
            float xs[3] = {px[r] * P.inv_dx, py[r] * P.inv_dx, pz[r] * P.inv_dx};
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                float bf = floorf(xs[a] - 0.5f) + 1.0f;
                if (bf >= 0.0f && bf + 2.0f < (float)P.nodes[a]) {
                    int bi = (int)bf;
                    lo[a] = min(lo[a], bi);
                    hi[a] = max(hi[a], bi);
                }
            }
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        lo[a] = wave_min_i(lo[a]);
        hi[a] = wave_max_i(hi[a]);
    }
    if (lane == 0)
        for (int a = 0; a < 3; ++a) {
            sbox[wave][a] = lo[a];
            sbox[wave][3 + a] = hi[a];
        }
    __syncthreads();
    if (tid == 0) {
        Tile t;
        int L[3], H[3];
        for (int a = 0; a < 3; ++a) {
            L[a] = min(min(sbox[0][a], sbox[1][a]), min(sbox[2][a], sbox[3][a]));
            H[a] = max(max(sbox[0][3 + a], sbox[1][3 + a]), max(sbox[2][3 + a], sbox[3][3 + a]));
        }
        bool ok = L[0] <= H[0] && L[1] <= H[1] && L[2] <= H[2];
        int gsz = 1;
        for (int a = 0; a < 3; ++a) {
            t.org[a] = ok ? L[a] : 0;
            t.gd[a] = ok ? H[a] - L[a] + 3 : 0;
            t.sd[a] = 0;
            gsz *= t.gd[a];
        }
        t.tiled = ok && gsz <= GCAP;
        tile = t;
    }
    __syncthreads();
    const Tile t = tile;
    if (t.tiled) {
        int gsz = t.gd[0] * t.gd[1] * t.gd[2];
        for (int q = tid; q < gsz; q += kBlock) {
            int k = q % t.gd[2], rest = q / t.gd[2];
            int j = rest % t.gd[1], i = rest / t.gd[1];
            int ni = node_index(P, t.org[0] + i, t.org[1] + j, t.org[2] + k);
            sg[q] = make_float4(gv[ni], gv[na + ni], gv[2 * na + ni], 0.0f);
        }
    }
    __syncthreads();

    // Phase B: G2P, advection, constitutive update; outputs go straight to the
    // trial slab.  The new stencil base of each updated particle is kept packed.
    float speed = 0.0f, disp = 0.0f;
    double diss = 0.0;
    unsigned long long fail = ~0ull;
    int escaped = 0;
    int nbase[R];
    lo[0] = lo[1] = lo[2] = INT_MAX;
    hi[0] = hi[1] = hi[2] = INT_MIN;
#pragma unroll 1
    for (int r = 0; r < R; ++r) {
        int p = chunk0 + r * kBlock + tid;
        bool nw = false;
        nbase[r] = -1;
        if (p < chunk_end) {
            float x = px[r], y = py[r], z = pz[r];
            Stencil s;
            if (!make_stencil(P, x, y, z, s)) {
                // Outside the padded grid: frozen for this trial (solver.rs:905-917).
                for (int f = 0; f < NPF; ++f) dst[f * cap + p] = src[f * cap + p];
                escaped += 1;
            } else {
                float v[3] = {0.0f, 0.0f, 0.0f};
                float B[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                if (t.tiled)
                    gather_lds(sg, t, s, v, B);
                else
                    gather_global(P, gv, na, s, v, B);
                float C[9];
                float k4 = 4.0f * P.inv_dx;
#pragma unroll
                for (int i = 0; i < 9; ++i) C[i] = B[i] * k4;
                float nx = x + v[0] * dt, ny = y + v[1] * dt, nz = z + v[2] * dt;
                float vn = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                speed = fmaxf(speed, vn);
                disp = fmaxf(disp, vn * dt * P.inv_dx);
                float tau[6];
                float inc = 0.0f, dd = 0.0f;
                int rc;
                if (P.kind == 0) {
                    float H[9], Ht[9], Hn[9];
#pragma unroll
                    for (int i = 0; i < 9; ++i) H[i] = src[(H00 + i) * cap + p];
                    // (I + dt C)(I + H) - I = H + dt C + dt C H
#pragma unroll
                    for (int i = 0; i < 3; ++i)
#pragma unroll
                        for (int j = 0; j < 3; ++j)
                            Ht[3 * i + j] = H[3 * i + j] + dt * C[3 * i + j] +
                                            dt * (C[3 * i] * H[j] + C[3 * i + 1] * H[3 + j] + C[3 * i + 2] * H[6 + j]);
                    rc = paste_update(P, dt, Ht, Hn, tau, inc, dd);
                    if (rc == F_OK) {
#pragma unroll
                        for (int i = 0; i < 9; ++i) dst[(H00 + i) * cap + p] = Hn[i];
                        diss += (double)dd * (double)shared[SVOL0 * cap + p];
                        dst[EPS * cap + p] = src[EPS * cap + p] + inc;
                    }
                } else {
                    float jn;
                    rc = liquid_update(P, dt, src[JM1 * cap + p], C, jn, tau);
                    if (rc == F_OK) dst[JM1 * cap + p] = jn;
                }
                if (rc != F_OK) {
                    unsigned long long f = ((unsigned long long)p << 8) | (unsigned long long)rc;
                    fail = f < fail ? f : fail;
                } else {
                    dst[PX * cap + p] = nx;
                    dst[PY * cap + p] = ny;
                    dst[PZ * cap + p] = nz;
                    dst[VX * cap + p] = v[0];
                    dst[VY * cap + p] = v[1];
                    dst[VZ * cap + p] = v[2];
#pragma unroll
                    for (int i = 0; i < 9; ++i) dst[(C00 + i) * cap + p] = C[i];
#pragma unroll
                    for (int i = 0; i < 6; ++i) dst[(TXX + i) * cap + p] = tau[i];
                    float ns[3] = {nx * P.inv_dx, ny * P.inv_dx, nz * P.inv_dx};
                    int nb[3];
                    bool valid = true;
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        float bf = floorf(ns[a] - 0.5f) + 1.0f;
                        valid = valid && bf >= 0.0f && bf + 2.0f < (float)P.nodes[a];
                        nb[a] = valid ? (int)bf : 0;
                    }
                    if (valid) {
                        nbase[r] = (nb[0] << 20) | (nb[1] << 10) | nb[2];
#pragma unroll
                        for (int a = 0; a < 3; ++a) {
                            lo[a] = min(lo[a], nb[a]);
                            hi[a] = max(hi[a], nb[a]);
                        }
                        nw = near_wall(P, nx, ny, nz);
                    } else {
                        escaped += 1;
                    }
                }
            }
        }
        wave_append(nw, p, nw_list, nw_count);
    }

    // Phase C: exact tile of the new stencils.
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        lo[a] = wave_min_i(lo[a]);
        hi[a] = wave_max_i(hi[a]);
    }
    if (lane == 0)
        for (int a = 0; a < 3; ++a) {
            sbox[wave][a] = lo[a];
            sbox[wave][3 + a] = hi[a];
        }
    __syncthreads();
    if (tid == 0) {
        int L[3], H[3];
        for (int a = 0; a < 3; ++a) {
            L[a] = min(min(sbox[0][a], sbox[1][a]), min(sbox[2][a], sbox[3][a]));
            H[a] = max(max(sbox[0][3 + a], sbox[1][3 + a]), max(sbox[2][3 + a], sbox[3][3 + a]));
        }
        bool ok = L[0] <= H[0] && L[1] <= H[1] && L[2] <= H[2];
        int ssz = 1;
        for (int a = 0; a < 3; ++a) {
            sorg[a] = ok ? L[a] : 0;
            sdim[a] = ok ? H[a] - L[a] + 3 : 0;
            ssz *= sdim[a];
        }
        stiled = ok && ssz <= SCAP;
    }
    __syncthreads();
    const int so[3] = {sorg[0], sorg[1], sorg[2]};
    const int sd[3] = {sdim[0], sdim[1], sdim[2]};
    const bool tiled2 = stiled != 0;
    const int ssz = sd[0] * sd[1] * sd[2];
    if (tiled2)
        for (int q = tid; q < 7 * ssz; q += kBlock) ss[(q / ssz) * SCAP + (q % ssz)] = 0.0;
    __syncthreads();

    // Phase D: deposit the updated particles (each thread re-reads its own writes).
    const int rot = lane & 7;
#pragma unroll 1
    for (int r = 0; r < R; ++r) {
        if (nbase[r] < 0) continue;
        int p = chunk0 + r * kBlock + tid;
        float x = dst[PX * cap + p], y = dst[PY * cap + p], z = dst[PZ * cap + p];
        Stencil s2;
        make_stencil(P, x, y, z, s2);
        float v[3] = {dst[VX * cap + p], dst[VY * cap + p], dst[VZ * cap + p]};
        float C[9], tau[6];
#pragma unroll
        for (int i = 0; i < 9; ++i) C[i] = dst[(C00 + i) * cap + p];
#pragma unroll
        for (int i = 0; i < 6; ++i) tau[i] = dst[(TXX + i) * cap + p];
        float m = shared[SMASS * cap + p], v0 = shared[SVOL0 * cap + p];
        if (tiled2)
            scatter_lds<SCAP>(ss, so, sd, P, s2, m, v, C, tau, v0, rot);
        else
            scatter_particle(P, out, s2, m, v[0], v[1], v[2], C, tau, v0);
    }
    __syncthreads();
    if (tiled2) {
        for (int q = tid; q < ssz; q += kBlock) {
            double m = ss[q];
            if (m == 0.0) continue;
            int k = q % sd[2], rest = q / sd[2];
            int j = rest % sd[1], i = rest / sd[1];
            int ni = node_index(P, so[0] + i, so[1] + j, so[2] + k);
            atomicAdd(out.m + ni, (float)m);
            atomicAdd(out.px + ni, (float)ss[1 * SCAP + q]);
            atomicAdd(out.py + ni, (float)ss[2 * SCAP + q]);
            atomicAdd(out.pz + ni, (float)ss[3 * SCAP + q]);
            atomicAdd(out.fx + ni, (float)ss[4 * SCAP + q]);
            atomicAdd(out.fy + ni, (float)ss[5 * SCAP + q]);
            atomicAdd(out.fz + ni, (float)ss[6 * SCAP + q]);
            out.flag[ni >> 6] = 1;
        }
    }
    // Per-workgroup partials of the step reductions.
    speed = wave_max(speed);
    disp = wave_max(disp);
    diss = wave_sum(diss);
    fail = wave_min_u64(fail);
    escaped = wave_sum_i(escaped);
    if (lane == 0) {
        sred[wave][0] = speed;
        sred[wave][1] = disp;
        sdiss[wave] = diss;
        sfail[wave] = fail;
        sesc[wave] = escaped;
    }
    __syncthreads();
    if (tid == 0) {
        float s0 = 0.0f, s1 = 0.0f;
        double d = 0.0;
        unsigned long long f = ~0ull;
        int e = 0;
        for (int w = 0; w < 4; ++w) {
            s0 = fmaxf(s0, sred[w][0]);
            s1 = fmaxf(s1, sred[w][1]);
            d += sdiss[w];
            f = sfail[w] < f ? sfail[w] : f;
            e += sesc[w];
        }
        pmax[2 * blockIdx.x] = s0;
        pmax[2 * blockIdx.x + 1] = s1;
        pdiss[blockIdx.x] = d;
        pfail[blockIdx.x] = f;
        pesc[blockIdx.x] = e;
    }
}

// Fixed-order reduction of the per-workgroup partials into the step status.
__global__ __launch_bounds__(256) void finalize_kernel(int ngrid, const double* gpart, const int* gbad, int npart,
                                                       const float* pmax, const double* pdiss,
                                                       const unsigned long long* pfail, const int* pesc,
                                                       const int* nw_count, StepStatus* st) {
    __shared__ double sd[256 * 4];
    __shared__ float sf[256 * 2];
    __shared__ unsigned long long su[256];
    __shared__ int si[256 * 2];
    int tid = threadIdx.x;
    double acc[kGridVals + 1];
    for (int i = 0; i <= kGridVals; ++i) acc[i] = 0.0;
    int bad = 0;
    for (int q = tid; q < ngrid; q += 256) {
        for (int i = 0; i < kGridVals; ++i) acc[i] += gpart[q * kGridVals + i];
        bad += gbad[q];
    }
    float smax = 0.0f, dmax = 0.0f;
    unsigned long long f = ~0ull;
    int esc = 0;
    for (int q = tid; q < npart; q += 256) {
        smax = fmaxf(smax, pmax[2 * q]);
        dmax = fmaxf(dmax, pmax[2 * q + 1]);
        acc[kGridVals] += pdiss[q];
        f = pfail[q] < f ? pfail[q] : f;
        esc += pesc[q];
    }
    // Tree over threads, one quantity group at a time.
    for (int i = 0; i <= kGridVals; i += 4) {
        for (int k = 0; k < 4; ++k) sd[tid * 4 + k] = (i + k <= kGridVals) ? acc[i + k] : 0.0;
        __syncthreads();
        for (int stride = 128; stride > 0; stride >>= 1) {
            if (tid < stride)
                for (int k = 0; k < 4; ++k) sd[tid * 4 + k] += sd[(tid + stride) * 4 + k];
            __syncthreads();
        }
        if (tid == 0)
            for (int k = 0; k < 4; ++k)
                if (i + k <= kGridVals) acc[i + k] = sd[k];
        __syncthreads();
    }
    sf[tid * 2] = smax;
    sf[tid * 2 + 1] = dmax;
    su[tid] = f;
    si[tid * 2] = bad;
    si[tid * 2 + 1] = esc;
    __syncthreads();
    for (int stride = 128; stride > 0; stride >>= 1) {
        if (tid < stride) {
            sf[tid * 2] = fmaxf(sf[tid * 2], sf[(tid + stride) * 2]);
            sf[tid * 2 + 1] = fmaxf(sf[tid * 2 + 1], sf[(tid + stride) * 2 + 1]);
            su[tid] = su[tid + stride] < su[tid] ? su[tid + stride] : su[tid];
            si[tid * 2] += si[(tid + stride) * 2];
            si[tid * 2 + 1] += si[(tid + stride) * 2 + 1];
        }
        __syncthreads();
    }
    if (tid == 0) {
        for (int a = 0; a < 3; ++a) {
            st->wall_impulse[a] = acc[a];
            st->traction_impulse[a] = acc[3 + a];
            st->grid_momentum[a] = acc[9 + a];
        }
        st->projection = acc[6];
        st->friction = acc[7];
        st->traction_energy = acc[8];
        st->dissipation = acc[kGridVals];
        st->max_speed_bits = __float_as_uint(sf[0]);
        st->max_disp_bits = __float_as_uint(sf[1]);
        st->fail = su[0];
        st->grid_nonfinite = si[0];
        st->escaped = si[1];
        st->nearwall = *nw_count;
    }
}

// Sort keys every 32 accepted steps (solver.rs:501-517).  The device order
// groups cells in 4x4x4-cell blocks (compact workgroup chunks); the reference's
// order key (clamped cell, z fastest, then id) is kept per particle so that the
// state is handed back in the reference's order.
__global__ void sort_keys_kernel(Params P, int n, const float* __restrict__ slab, size_t cap,
                                 const uint32_t* __restrict__ id, uint64_t* keys, uint32_t* order, uint64_t* refkey,
                                 int keep_ref) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    uint64_t c[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float sc = floorf(slab[(PX + a) * cap + p] / P.dx);
        sc = isnan(sc) ? 0.0f : fmaxf(sc, 0.0f);
        c[a] = sc >= (float)P.cells[a] ? (uint64_t)P.cells[a] : (uint64_t)sc;
    }
    uint64_t ref = (c[0] * (uint64_t)(P.cells[1] + 1) + c[1]) * (uint64_t)(P.cells[2] + 1) + c[2];
    uint64_t nb1 = (uint64_t)(P.cells[1] + 4) / 4, nb2 = (uint64_t)(P.cells[2] + 4) / 4;
    uint64_t blk = ((c[0] >> 2) * nb1 + (c[1] >> 2)) * nb2 + (c[2] >> 2);
    uint64_t within = ((c[0] & 3) << 4) | ((c[1] & 3) << 2) | (c[2] & 3);
    keys[p] = ((blk * 64 + within) << 32) | (uint64_t)id[p];
    order[p] = (uint32_t)p;
    if (!keep_ref) refkey[p] = (ref << 32) | (uint64_t)id[p];
}

__global__ void block_ids_kernel(int n, const uint64_t* __restrict__ keys, uint32_t* out) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) out[p] = (uint32_t)(keys[p] >> 38);
}

__global__ void permute_kernel(int n, const uint32_t* __restrict__ order, const float* __restrict__ src,
                               float* __restrict__ dst, size_t cap, const float* __restrict__ ssrc,
                               float* __restrict__ sdst, const uint32_t* __restrict__ isrc,
                               uint32_t* __restrict__ idst, const uint64_t* __restrict__ rsrc,
                               uint64_t* __restrict__ rdst) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    uint32_t o = order[k];
    for (int f = 0; f < NPF; ++f) dst[f * cap + k] = src[f * cap + o];
    for (int f = 0; f < NSF; ++f) sdst[f * cap + k] = ssrc[f * cap + o];
    idst[k] = isrc[o];
    rdst[k] = rsrc[o];
}

__global__ void paste_batch_kernel(Params P, float dt, int n, const float* h_in, float* h_out, float* tau,
                                   float* inc, float* dd, int* codes) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    float Ht[9], Hn[9], t[6];
    for (int i = 0; i < 9; ++i) Ht[i] = h_in[9 * p + i];
    float a = 0.0f, b = 0.0f;
    int rc = paste_update(P, dt, Ht, Hn, t, a, b);
    codes[p] = rc;
    for (int i = 0; i < 9; ++i) h_out[9 * p + i] = rc ? 0.0f : Hn[i];
    for (int i = 0; i < 6; ++i) tau[6 * p + i] = rc ? 0.0f : t[i];
    inc[p] = a;
    dd[p] = b;
}

__global__ void liquid_batch_kernel(Params P, float dt, int n, const float* jm1, const float* c, float* jm1_out,
                                    float* tau, int* codes) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    float C[9], t[6], jn = 0.0f;
    for (int i = 0; i < 9; ++i) C[i] = c[9 * p + i];
    int rc = liquid_update(P, dt, jm1[p], C, jn, t);
    codes[p] = rc;
    jm1_out[p] = jn;
    for (int i = 0; i < 6; ++i) tau[6 * p + i] = rc ? 0.0f : t[i];
}

inline int blocks_for(int n) { return (n + kBlock - 1) / kBlock; }

constexpr int kChunk = 1024;

constexpr int kGatherCap = 768;
constexpr int kScatterCap = 1024;

}  // namespace

int g2p2g_chunk() { return kChunk; }
int grid_workgroups(int nblocks) { return (nblocks + 3) / 4; }

void launch_reset_status(hipStream_t s, StepStatus* st, int* nw_count) {
    reset_status_kernel<<<1, 64, 0, s>>>(st, nw_count);
}

void launch_p2g(hipStream_t s, const Params& P, int n, const float* slab, size_t cap, const float* shared,
                GridPtrs g) {
    if (n > 0) p2g_kernel<<<blocks_for(n), kBlock, 0, s>>>(P, n, slab, cap, shared, g);
}

void launch_nearwall(hipStream_t s, const Params& P, int n, const float* slab, size_t cap, int* list, int* count) {
    if (n > 0) nearwall_kernel<<<blocks_for(n), kBlock, 0, s>>>(P, n, slab, cap, list, count);
}

void launch_traction(hipStream_t s, const Params& P, float dt, const int* list, const int* count, int max_count,
                     const float* slab, size_t cap, const float* shared, GridPtrs g, float* trS, float* trT,
                     size_t na) {
    if (max_count > 0)
        traction_kernel<<<blocks_for(max_count), kBlock, 0, s>>>(P, dt, list, count, slab, cap, shared, g, trS, trT,
                                                                 na);
}

void launch_grid_update(hipStream_t s, const Params& P, float dt, int nblocks, GridPtrs a, float* trS, float* trT,
                        float* gv, size_t na, GridPtrs clear, double* partials, int* bad) {
    grid_update_kernel<<<grid_workgroups(nblocks), 256, 0, s>>>(P, dt, a, trS, trT, gv, na, clear, nblocks, partials,
                                                                bad);
}

void launch_couple(hipStream_t s, const Params& P, double dx, int nblocks, GridPtrs a, float* gv, size_t na,
                   const PelletDev* pellets, int npellets, const int* bstart, const int* blist, PelletAccum* acc,
                   double friction) {
    if (npellets > 0) couple_kernel<<<nblocks, 64, 0, s>>>(P, dx, a, gv, na, pellets, bstart, blist, acc, friction);
}

void launch_g2p2g(hipStream_t s, const Params& P, float dt, int n, const float* src, float* dst, size_t cap,
                  const float* shared, const float* gv, size_t na, GridPtrs out, const int2* work, int nwork,
                  float* pmax, double* pdiss, unsigned long long* pfail, int* pesc, int* nw_list, int* nw_count,
                  hipEvent_t start, hipEvent_t stop) {
    if (start) (void)hipEventRecord(start, s);
    if (n > 0 && nwork > 0)
        g2p2g_kernel<kChunk, kGatherCap, kScatterCap><<<nwork, kBlock, 0, s>>>(
            P, dt, n, src, dst, cap, shared, gv, na, out, work, pmax, pdiss, pfail, pesc, nw_list, nw_count);
    if (stop) (void)hipEventRecord(stop, s);
}

size_t rle_temp_bytes(int n) {
    size_t bytes = 0;
    (void)hipcub::DeviceRunLengthEncode::Encode(nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                                (uint32_t*)nullptr, (int*)nullptr, n);
    return bytes;
}

void launch_block_runs(hipStream_t s, void* temp, size_t temp_bytes, int n, const uint64_t* sorted_keys,
                       uint32_t* ids_tmp, uint32_t* unique, uint32_t* counts, int* nruns) {
    if (n <= 0) return;
    block_ids_kernel<<<blocks_for(n), kBlock, 0, s>>>(n, sorted_keys, ids_tmp);
    (void)hipcub::DeviceRunLengthEncode::Encode(temp, temp_bytes, ids_tmp, unique, counts, nruns, n, s);
}

void launch_finalize(hipStream_t s, int ngrid, const double* gpart, const int* gbad, int npart, const float* pmax,
                     const double* pdiss, const unsigned long long* pfail, const int* pesc, const int* nw_count,
                     StepStatus* st) {
    finalize_kernel<<<1, 256, 0, s>>>(ngrid, gpart, gbad, npart, pmax, pdiss, pfail, pesc, nw_count, st);
}

size_t sort_temp_bytes(int n, int end_bit) {
    size_t bytes = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                             (const uint32_t*)nullptr, (uint32_t*)nullptr, n, 0, end_bit);
    return bytes;
}

void launch_radix_sort(hipStream_t s, void* temp, size_t temp_bytes, const uint64_t* keys_in, uint64_t* keys_out,
                       const uint32_t* vals_in, uint32_t* vals_out, int n, int end_bit) {
    (void)hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, keys_in, keys_out, vals_in, vals_out, n, 0, end_bit,
                                             s);
}

void launch_sort_keys(hipStream_t s, const Params& P, int n, const float* slab, size_t cap, const uint32_t* id,
                      uint64_t* keys, uint32_t* order, uint64_t* refkey, int keep_ref) {
    if (n > 0) sort_keys_kernel<<<blocks_for(n), kBlock, 0, s>>>(P, n, slab, cap, id, keys, order, refkey, keep_ref);
}

void launch_permute(hipStream_t s, int n, const uint32_t* order, const float* src, float* dst, size_t cap,
                    const float* ssrc, float* sdst, const uint32_t* isrc, uint32_t* idst, const uint64_t* rsrc,
                    uint64_t* rdst) {
    if (n > 0)
        permute_kernel<<<blocks_for(n), kBlock, 0, s>>>(n, order, src, dst, cap, ssrc, sdst, isrc, idst, rsrc, rdst);
}

void launch_paste_batch(hipStream_t s, const Params& P, float dt, int n, const float* h_in, float* h_out, float* tau,
                        float* inc, float* dd, int* codes) {
    if (n > 0) paste_batch_kernel<<<blocks_for(n), kBlock, 0, s>>>(P, dt, n, h_in, h_out, tau, inc, dd, codes);
}

void launch_liquid_batch(hipStream_t s, const Params& P, float dt, int n, const float* jm1, const float* c,
                         float* jm1_out, float* tau, int* codes) {
    if (n > 0) liquid_batch_kernel<<<blocks_for(n), kBlock, 0, s>>>(P, dt, n, jm1, c, jm1_out, tau, codes);
}

}  // namespace mpm
