// HIP kernels of the batched MLS-MPM substep for CDNA4 (gfx950).
//
// One trial step of every running instance (core/src/mpm/solver.rs:413-611) is a
// fixed launch sequence with no host round trip:
//   control      -- advance_while's dt choice (solver.rs:421-443) and the stability
//                   limit (477-498) per instance, the deposit scales of the trial;
//                   pellet geometry for the coupling
//   p2g          -- deposit_particles (solver.rs:632-661) as exact integer sums
//   traction     -- wall_traction_to_grid (solver.rs:725-819) over the near-wall lists
//   grid_update  -- grid_update (solver.rs:822-892), one wavefront per 4x4x4-node block
//   couple       -- couple_grid (rigid.rs:385-438) for each instance's pellets in order
//   g2p          -- grid_to_particle with update_paste / update_liquid (solver.rs:895-977,
//                   constitutive.rs:456-589) into the trial slab
//   pellet       -- Pellet::apply_impulse and ::integrate (rigid.rs:251-334) on trial copies
//   finalize     -- reduction of the per-workgroup partials per instance
//   commit       -- accept/reject, commit and ledgers (solver.rs:444-467, 593-611, 980-998)
//   refkeys      -- the 32-step reference order keys (solver.rs:501-517)
// Every sum whose order of additions varies with the scheduling is an exact
// integer (the deposit, the wall reactions, the pellets' coupling sums and the
// plastic dissipation), so a trajectory is a function of the state alone: the
// same on every run and after a restore, whatever the particle order.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_run_length_encode.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

#include "mlp.h"
#include "physics.h"

namespace mpm {

namespace {

constexpr int kChunk = 1024;
// Nodes of a work item's LDS tile: the stencils of two z-adjacent 4^3-cell
// blocks span 7 x 7 x 11 nodes, with room for the particles' drift between
// sorts.  An item whose box outgrows it deposits and gathers through global
// memory at several times the cost, so the cap sets how far the particles may
// drift before that cliff: 1280 nodes (40 KB of deposit tile, four 256-thread
// deposit workgroups per CU) against 768 moved the cliff past where the sort
// policy re-sorts in a fast flow.
constexpr int kTileCap = 1280;
constexpr double kDtRecovery = 1.25;              // solver.rs:45
constexpr unsigned long long kSortInterval = 32;  // solver.rs:39

// ------------------------------------------------------------ pellets (f64)

// nalgebra's UnitQuaternion (i, j, k, w) acting on a vector.
__device__ inline void qrot(const double* q, const double* v, double* out) {
    double t[3] = {2.0 * (q[1] * v[2] - q[2] * v[1]), 2.0 * (q[2] * v[0] - q[0] * v[2]),
                   2.0 * (q[0] * v[1] - q[1] * v[0])};
    double c[3] = {q[1] * t[2] - q[2] * t[1], q[2] * t[0] - q[0] * t[2], q[0] * t[1] - q[1] * t[0]};
    for (int a = 0; a < 3; ++a) out[a] = t[a] * q[3] + c[a] + v[a];
}

__device__ inline void qmatrix(const double* q, double* r) {
    double i = q[0], j = q[1], k = q[2], w = q[3];
    double ww = w * w, ii = i * i, jj = j * j, kk = k * k;
    double ij = i * j * 2.0, wk = w * k * 2.0, wj = w * j * 2.0, ik = i * k * 2.0, jk = j * k * 2.0,
           wi = w * i * 2.0;
    r[0] = ww + ii - jj - kk;
    r[1] = ij - wk;
    r[2] = wj + ik;
    r[3] = wk + ij;
    r[4] = ww - ii + jj - kk;
    r[5] = jk - wi;
    r[6] = ik - wj;
    r[7] = wi + jk;
    r[8] = ww - ii - jj + kk;
}

// Pellet::angular_velocity, rigid.rs:193-198: R I^-1 R^T L.
__device__ inline void pellet_omega(const double* q, const double* L, const double* inertia, double* om) {
    double r[9];
    qmatrix(q, r);
    double b[3] = {r[0] * L[0] + r[3] * L[1] + r[6] * L[2], r[1] * L[0] + r[4] * L[1] + r[7] * L[2],
                   r[2] * L[0] + r[5] * L[1] + r[8] * L[2]};
    double bo[3] = {b[0] / inertia[0], b[1] / inertia[1], b[2] / inertia[2]};
    for (int a = 0; a < 3; ++a) om[a] = r[3 * a] * bo[0] + r[3 * a + 1] * bo[1] + r[3 * a + 2] * bo[2];
}

__device__ inline double pellet_kinetic(double mass, const double* vel, const double* q, const double* L,
                                        const double* inertia) {
    double om[3];
    pellet_omega(q, L, inertia, om);
    return 0.5 * mass * (vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]) +
           0.5 * (om[0] * L[0] + om[1] * L[1] + om[2] * L[2]);
}

// Instance owning slab entry p (binary search over the range starts).
__device__ inline int instance_of(const BatchDev& B, int p) {
    int lo = 0, hi = B.ninst - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (B.poffs[mid] <= p)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Kernel timing on the device's constant clock: the earliest workgroup start
// and the latest workgroup end of a launch.  Workgroups start in index order,
// so the first 64 stamp the start (one address); every workgroup stamps its
// end, spread over kStampSlots partial slots (a single address taking every
// workgroup's atomic cost the step 10 us; a start address that depends on the
// workgroup index made G2P's register allocation 44 VGPRs larger).
// kernel_end is called after the workgroup's last barrier.  q: the timed
// quantity (0..3).
__device__ inline void kernel_start(unsigned long long* stamps, int q) {
    if (threadIdx.x == 0 && blockIdx.x < 64)
        atomicMin(stamps + q * kStampSlots, (unsigned long long)__builtin_amdgcn_s_memrealtime());
}
__device__ inline void kernel_end(unsigned long long* stamps, int q) {
    if (threadIdx.x == 0)
        atomicMax(stamps + q * kStampSlots + (blockIdx.x % kStampSlots),
                  (unsigned long long)__builtin_amdgcn_s_memrealtime());
}

// Reference order key (clamped cell index, z fastest, then id; solver.rs:501-517),
// in f64 as the reference computes it: the f32 position promoted exactly and
// divided by the f64 spacing, so a particle within an ulp of a cell face floors
// to the same cell as in the reference (an f32 quotient by the rounded f32
// spacing need not).  Only evaluated at the 32-step cadence.
__device__ inline uint64_t ref_cell(const InstParams& ip, float x, float y, float z) {
    const double pos[3] = {(double)x, (double)y, (double)z};
    uint64_t c[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double sc = floor(pos[a] / ip.dx);
        sc = isnan(sc) ? 0.0 : fmax(sc, 0.0);
        c[a] = sc >= (double)ip.P.cells[a] ? (uint64_t)ip.P.cells[a] : (uint64_t)sc;
    }
    return (c[0] * (uint64_t)(ip.P.cells[1] + 1) + c[1]) * (uint64_t)(ip.P.cells[2] + 1) + c[2];
}

// ------------------------------------------------------------ control

// The double nearest coupling sum i of a pellet.
__device__ inline double pel_sum(const PelletAccum& a, int i) {
    return fix_value(fix_from_limbs(a.sum[i][0], a.sum[i][1], a.sum[i][2]));
}

// A pellet's start-of-step geometry for the coupling kernel (its axis, angular
// velocity and box) from its committed pose and momenta (pos, vel, L, q) and
// its constant shape in `p`; coupling sums cleared.
__device__ inline PelletGeo pellet_geometry(const BatchDev& B, int k, const PelletState& p, const double* pos,
                                            const double* vel, const double* L, const double* q) {
    PelletGeo g;
    const double xh[3] = {1.0, 0.0, 0.0};
    qrot(q, xh, g.axis);
    pellet_omega(q, L, p.inertia, g.omega);
    const int spheres = p.spheres;
    const double radius = p.radius, half = p.half;  // half: set with the pellet on the host
    for (int a = 0; a < 3; ++a) {
        g.pos[a] = pos[a];
        g.vel[a] = vel[a];
        double reach = fabs(g.axis[a]) * half + radius;
        g.lo[a] = pos[a] - reach;
        g.hi[a] = pos[a] + reach;
    }
    g.radius = radius;
    g.spheres = spheres;
    g.inst = p.inst;
    B.pgeo[k] = g;
    PelletAccum z = {};
    B.pacc[k] = z;
    return g;
}

// advance_while's choice of the next trial (solver.rs:425-443) for one
// instance, and the trial's deposit scales.  Returns the instance's near-wall
// count if it runs a trial, else 0.  Every field is read before the first
// store, so the loads go out together.
__device__ inline int control_instance(InstState& s, const InstParams& ip) {
    const int mode = s.mode, budget = s.budget, cur = s.cur, nwc = s.nw_count[s.cur];
    const double trial_dt = s.trial_dt, time = s.time, target = s.target, dt_scale = s.dt_scale;
    const double max_speed = s.max_speed, pel_vmax = s.pel_vmax;
    const double b1 = s.bound[0], b2 = s.bound[1];
    const double cfl = ip.cfl, dx = ip.dx, wave = ip.wave, diff = ip.diff, pel_contact_dt = ip.pel_contact_dt;
    const double max_dt = ip.max_dt, min_dt = ip.min_dt;
    const int npel = ip.npel;
    const double mass_max = (double)ip.mass_max;
    const double gmax = fmax(fmax(fabs(ip.gravity[0]), fabs(ip.gravity[1])), fabs(ip.gravity[2]));
    int active = 0;
    bool set_dt = false;
    double dt = 0.0;
    if (mode == M_RUN) {
        if (trial_dt > 0.0) {
            // A bare trial step of a given size (solver.rs:535-537).
            dt = trial_dt;
            set_dt = true;
            s.limit = 0.0;
            s.limited = 0;
            active = 1;
        } else if (!(time < target)) {
            s.time = target;  // solver.rs:470
            s.mode = M_DONE;
        } else if (budget <= 0) {
            s.mode = M_PAUSE;
            s.pause |= P_BUDGET;
        } else {
            // solver.rs:425-443, with the stability limit of solver.rs:477-498
            const double remaining = target - time;
            double limit = cfl * dx / (wave + max_speed);
            if (diff > 0.0) limit = fmin(limit, 0.25 * dx * dx / diff);
            if (npel > 0) {
                // min over pellets of cfl dx / |v_k| is cfl dx / max_k |v_k|
                limit = fmin(limit, pel_contact_dt);
                if (pel_vmax > 0.0) limit = fmin(limit, cfl * dx / pel_vmax);
            }
            const double scaled = dt_scale * limit;
            const int limited = scaled < max_dt;
            dt = fmin(fmin(scaled, max_dt), remaining);
            if (remaining - dt < 1e-12 * fmax(target, 1.0)) dt = remaining;
            set_dt = true;
            if (!isfinite(dt) || dt <= 0.0 || (dt < min_dt && remaining > min_dt)) {
                s.mode = M_ERROR;
                s.err = E_MIN_STEP;
            } else {
                s.limit = limit;
                s.limited = limited;
                active = 1;
            }
        }
    }
    s.active = active;
    if (set_dt) s.dt = dt;
    if (!active) return 0;
    StepStatus z = {};
    z.fail = ~0ull;
    s.st = z;
    s.nw_count[cur ^ 1] = 0;
    // Deposit scales of this trial from the committed state's bounds
    // (particle_bounds): no contribution reaches 2^29 units.
    s.sc_m = deposit_scale(mass_max);
    s.sc_p = deposit_scale(b1 + dt * b2);
    s.sc_mom = deposit_scale(b1);
    s.sc_f = deposit_scale(b2 * 0.25 * dx * dx);
    s.sc_t = deposit_scale(dt * (b2 + 7.0 * mass_max * gmax));
    return nwc;
}

// Step control of every instance.  Thread t takes a contiguous run of
// instances; a block scan of their near-wall counts gives each running
// instance its range in the concatenated near-wall list (in instance order),
// which the traction kernel walks.  Pellet geometry is recomputed here only for
// instances whose pellets the host has touched since the last finalize (which
// keeps it current otherwise).
__global__ __launch_bounds__(256) void control_kernel(BatchDev B) {
    __shared__ int sscan[256];
    __shared__ int sstale;
    const int tid = threadIdx.x;
    const int per = (B.ninst + 255) / 256;
    const int t0 = min(tid * per, B.ninst), t1 = min(t0 + per, B.ninst);
    if (tid == 0) sstale = 0;
    int local = 0;
    bool stale = false;
    for (int t = t0; t < t1; ++t) {
        InstState& s = B.is[t];
        const int nwc = control_instance(s, B.ip[t]);
        local += nwc;
        stale = stale || (s.active && s.geo_stale && B.ip[t].npel > 0);
    }
    sscan[tid] = local;
    __syncthreads();
    if (stale) sstale = 1;
    for (int o = 1; o < 256; o <<= 1) {
        const int v = tid >= o ? sscan[tid - o] : 0;
        __syncthreads();
        sscan[tid] += v;
        __syncthreads();
    }
    int acc = sscan[tid] - local;
    for (int t = t0; t < t1; ++t) {
        InstState& s = B.is[t];
        acc += s.active ? s.nw_count[s.cur] : 0;
        s.nw_base = acc;
    }
    if (tid == 255) B.counters[1] = sscan[255];
    if (!sstale) return;
    // Start-of-step pellet geometry and zeroed coupling accumulators of the
    // stale running instances; their coupling kernel skips no node block.
    __syncthreads();
    for (int k = tid; k < B.npel; k += 256) {
        const int inst = B.pel[k].inst;
        if (inst < 0) continue;
        const InstState& s = B.is[inst];
        if (!s.active || !s.geo_stale) continue;
        const PelletState& p = B.pel[k];
        pellet_geometry(B, k, p, p.pos, p.vel, p.L, p.q);
    }
    __syncthreads();
    for (int t = t0; t < t1; ++t) {
        InstState& s = B.is[t];
        if (!s.active || !s.geo_stale) continue;
        for (int a = 0; a < 3; ++a) {
            s.pel_box[a] = -INFINITY;
            s.pel_box[3 + a] = INFINITY;
        }
        s.geo_stale = 0;
    }
}

// ------------------------------------------------------------ deposit

// deposit_particles (solver.rs:632-661) of every running instance's committed
// state, at the start of the trial once its dt is known.  One workgroup per
// work item: the item's particles deposit into an LDS tile of exact 64-bit
// integer sums over the nodes of their stencils (the box of the committed
// slab's set-up record), flushed with one 64-bit global atomic per touched
// node and field.  Particles outside the tile, and items whose tile exceeds
// the LDS budget, deposit with global atomics; a particle gives the same
// integers either way (deposit_particle).  At the 32-step cadence it also
// writes the reference-order keys of the committed positions.
// NF = kDepositFields, or kAuditFields for the energy audit.  BS: threads per
// workgroup, 256 (four rounds of a full item) or 512 (two rounds, the tile's
// clearing and flush over twice the threads); the launch shape picks it.
template <int NF, int BS>
__global__ __launch_bounds__(BS) void p2g_kernel(BatchDev B) {
    constexpr int R = kChunk / BS;
    __shared__ __align__(16) unsigned long long ss[NF * kTileCap];
    if (B.ktime) kernel_start(B.ktime, 0);
    // The item and the instance fields the rounds use, read once into uniform
    // registers (read through references, they would be reloaded after every
    // global store in the loop).
    const int4 item = uniform_copy(B.work[blockIdx.x]);
    const int inst = item.x;
    const InstState& is = B.is[inst];
    if (!is.active) return;
    const InstParams& ip = B.ip[inst];
    const Params P = ip.P;
    const int tid = threadIdx.x, lane = tid & 63;
    const size_t cap = B.cap, na = B.na;
    const float dt = (float)is.dt;
    const int refsort = is.refsort;
    const DepositScales scales{is.sc_m, is.sc_p, is.sc_mom, is.sc_f};
    const float* __restrict__ src = B.slab[is.cur];
    const float* __restrict__ shared = B.shared[B.sh];
    GridPtrs out = grid_set(B, is.gcur);
    out.v += (size_t)ip.boff * 64;
    out.flag += ip.boff;
    const ItemPre pre = B.wpre[is.cur][blockIdx.x];
    Tile t;
    {
        bool ok = true;
        int sz = 1;
#pragma unroll
        for (int a = 0; a < 3; ++a) ok = ok && pre.lo[a] <= pre.hi[a];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            t.org[a] = ok ? pre.lo[a] : 0;
            t.gd[a] = ok ? pre.hi[a] - pre.lo[a] + 3 : 0;
            sz *= t.gd[a];
        }
        t.tiled = ok && sz <= kTileCap;
    }
    // Only the tile's own nodes are cleared (the flush reads no others), two
    // 64-bit words per store.
    if (t.tiled) {
        const int half = (t.gd[0] * t.gd[1] * t.gd[2] + 1) >> 1;
#pragma unroll
        for (int f = 0; f < NF; ++f)
            for (int q = tid; q < half; q += BS)
                reinterpret_cast<ulonglong2*>(ss + f * kTileCap)[q] = make_ulonglong2(0ull, 0ull);
    }
    __syncthreads();
    const int rot = (lane * 5) % 9, ra = rot / 3, rb = rot - 3 * ra;
    const int s12 = __mul24(t.gd[1], t.gd[2]);
    const int end = item.y + item.z;
#pragma unroll 1
    for (int r = 0; r < R; ++r) {
        const int p = item.y + r * BS + tid;
        if (p >= end) continue;
        // All of the particle's reads go out together, before any of them is used.
        const float x = src[PX * cap + p], y = src[PY * cap + p], z = src[PZ * cap + p];
        const float m = shared[SMASS * cap + p], v0 = shared[SVOL0 * cap + p];
        float v[3], C[9], tau[6];
#pragma unroll
        for (int a = 0; a < 3; ++a) v[a] = src[(VX + a) * cap + p];
#pragma unroll
        for (int i = 0; i < 9; ++i) C[i] = src[(C00 + i) * cap + p];
#pragma unroll
        for (int i = 0; i < 6; ++i) tau[i] = src[(TXX + i) * cap + p];
        __builtin_amdgcn_sched_barrier(0);
        if (refsort) B.refkey[p] = (ref_cell(ip, x, y, z) << 32) | (uint64_t)B.ids[p];
        Stencil s;
        if (!make_stencil(P, x, y, z, s)) continue;  // outside the padded grid: no deposit (solver.rs:636)
        const Deposit<NF> q = make_deposit<NF>(P, scales, dt, m, v0, v, C, tau);
        bool in_tile = t.tiled;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int li = s.base[a] - t.org[a];
            in_tile = in_tile && li >= 0 && li + 2 < t.gd[a];
        }
        if (in_tile) {
            TileSink<NF, kTileCap> sink;
            sink.ss = ss;
            sink.s12 = s12;
            sink.sd2 = t.gd[2];
            sink.row0 = __mul24(__mul24(s.base[0] - t.org[0], t.gd[1]) + (s.base[1] - t.org[1]), t.gd[2]) +
                        (s.base[2] - t.org[2]);
            sink.row = 0;
            deposit_particle<NF>(s, q, ra, rb, sink);
        } else {
            GlobalSink sink;
            sink.v = out.v;
            sink.na = na;
            sink.P = &P;
#pragma unroll
            for (int a = 0; a < 3; ++a) sink.base[a] = s.base[a];
            flag_stencil(P, out.flag, s);
            deposit_particle<NF>(s, q, ra, rb, sink);
        }
    }
    __syncthreads();
    if (!t.tiled) {
        if (B.ktime) kernel_end(B.ktime, 1);
        return;
    }
    // Flush: the tile's exact sums into the global ones.  A node with a
    // contribution has mass (mass rounds up), so the mass word marks it.
    const int sz = t.gd[0] * s12;
    const float ig2 = 1.0f / (float)t.gd[2], ig1 = 1.0f / (float)t.gd[1];
    for (int q = tid; q < sz; q += BS) {
        const unsigned long long qm = ss[q];
        if (qm == 0ull) continue;
        const int rest = div_small(q, ig2), k = q - rest * t.gd[2];
        const int i = div_small(rest, ig1), j = rest - i * t.gd[1];
        const int ni = node_index(P, t.org[0] + i, t.org[1] + j, t.org[2] + k);
        atomicAdd(reinterpret_cast<unsigned long long*>(out.v + ni), qm);
#pragma unroll
        for (int f = 1; f < NF; ++f) {
            const unsigned long long qv = ss[f * kTileCap + q];
            if (qv) atomicAdd(reinterpret_cast<unsigned long long*>(out.v + f * na + ni), qv);
        }
        out.flag[ni >> 6] = 1;
    }
    if (B.ktime) {
        __syncthreads();
        kernel_end(B.ktime, 1);
    }
}

// ------------------------------------------------------------ traction

// A workgroup's traction deposits merged by node in LDS before they reach
// memory: open addressing with linear probing on the key axis * na + node.
// Neighbouring near-wall particles share most of their wall nodes, and every
// global 64-bit atomic on one address waits for the previous one, so one
// atomic per node and workgroup instead of one per deposit shortens the
// drain after the kernel's last wave.  A deposit that finds no free entry
// within kTrProbes probes goes to memory directly; integer sums do not
// depend on the order.
constexpr int kTrHash = 1024, kTrProbes = 32;
constexpr unsigned kTrEmpty = 0xffffffffu;

// wall_traction_to_grid (solver.rs:725-819) over the near-wall lists of the
// running instances, concatenated.  Each deposit adds its reaction to the
// Coulomb budget (trT) of its node, which also gives the signed traction
// impulse (trS, kept separately only on grids two cells deep); the grid update
// books the telescoped kinetic-energy jump per node.  Reactions are exact
// integers in units of the instance's traction scale.
__global__ __launch_bounds__(kBlock) void traction_kernel(BatchDev B) {
    const int total = B.counters[1];
    const float* shared = B.shared[B.sh];
    const size_t cap = B.cap, na = B.na;
    if ((long long)blockIdx.x * blockDim.x >= 9LL * total) return;  // no work in this workgroup
    __shared__ unsigned hkey[kTrHash];
    __shared__ unsigned long long hval[kTrHash];
    const bool merge = 3 * na < (size_t)kTrEmpty;  // keys fit 32 bits
    for (int i = threadIdx.x; i < kTrHash; i += kBlock) {
        hkey[i] = kTrEmpty;
        hval[i] = 0ull;
    }
    __syncthreads();
    // One thread per (near-wall particle, stencil column across the wall): a
    // particle's node visits are independent (each reads its own wall-plane node
    // and deposits on its own node), so spreading them over threads overlaps
    // their memory round trips.  A thread takes the three nodes of its column
    // along the wall axis; one of them is past the wall-side nodes.
    const long long items = 9LL * total;
    // A batch of one instance reads the instance's list fields together with
    // the count (both list counts, selected by the slab): the list entry then
    // waits for one memory round trip, not three.
    const InstState& s0 = B.is[0];
    const int cur0 = s0.cur, nwc00 = s0.nw_count[0], nwc01 = s0.nw_count[1], nwb0 = s0.nw_base;
    const int poff0 = B.ip[0].poff;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < items;
         t += (long long)gridDim.x * blockDim.x) {
        const int q = (int)(t / 9), col = (int)(t - 9LL * q);
        const int cu = col / 3, cv = col % 3;
        int lo = 0, hi = B.ninst - 1;
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if (B.is[mid].nw_base > q)
                hi = mid;
            else
                lo = mid + 1;
        }
        const InstState& s = B.is[lo];
        const InstParams& ip = B.ip[lo];
        const Params P = ip.P;
        const bool one = B.ninst == 1;
        const int cur = one ? cur0 : s.cur;
        const int local = q - (one ? nwb0 - (cur0 ? nwc01 : nwc00) : s.nw_base - s.nw_count[cur]);
        const int p = B.nw_list[cur][(one ? poff0 : ip.poff) + local];
        const float* slab = B.slab[cur];
        const float dt = (float)s.dt;
        const long long* gsum = B.gmem[s.gcur];
        const double um = 1.0 / (double)s.sc_m, up = 1.0 / (double)s.sc_p;
        const float tscale = s.sc_t;
        const size_t noff = (size_t)ip.boff * 64;
        float pos[3] = {slab[PX * cap + p], slab[PY * cap + p], slab[PZ * cap + p]};
        const float tnn[3] = {slab[TXX * cap + p], slab[TYY * cap + p], slab[TZZ * cap + p]};
        float half = 0.5f * P.dx;
        float mass = shared[SMASS * cap + p], v0 = shared[SVOL0 * cap + p];
        float k4 = 4.0f * P.inv_dx * P.inv_dx;
        float force_scale = -dt * k4;
#pragma unroll
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
            Stencil st;
            if (!make_stencil(P, img[0], img[1], img[2], st)) continue;
            float g_a = P.gravity[axis];
            float normal_stress = tnn[axis] * v0 - mass * g_a * (img[axis] - pos[axis]);
            if (normal_stress >= 0.0f) continue;
            int plane = face == 0.0f ? 1 : P.cells[axis] + 1;
            float inward = face == 0.0f ? 1.0f : -1.0f;
            float weight_momentum = mass * g_a * dt;
            const long long* mom = gsum + (size_t)(1 + axis) * na;
            // The column's three nodes: all their grid reads first (wall-plane node
            // mass and momentum, deposit node mass), then the deposits, so a
            // thread waits for one memory round trip rather than two per node.
            float w[3], pm[3], pmo[3];
            long long dm[3];
            size_t nis[3];
#pragma unroll
            for (int off = 0; off < 3; ++off) {
                const int sa = axis == 0 ? off : cu, sb = axis == 0 ? cu : (axis == 1 ? off : cv),
                          sc = axis == 2 ? off : cv;
                int co[3] = {st.base[0] + sa, st.base[1] + sb, st.base[2] + sc};
                w[off] = 0.0f;
                pm[off] = pmo[off] = 0.0f;
                dm[off] = 0;
                nis[off] = 0;
                if (co[axis] <= 1 || co[axis] > P.cells[axis]) continue;
                // sel3: the stencil stays in registers (a runtime index would place it in scratch).
                w[off] = sel3(sa, st.w[0][0], st.w[0][1], st.w[0][2]) * sel3(sb, st.w[1][0], st.w[1][1], st.w[1][2]) *
                         sel3(sc, st.w[2][0], st.w[2][1], st.w[2][2]);
                if (w[off] == 0.0f) continue;
                int pc[3] = {co[0], co[1], co[2]};
                pc[axis] = plane;
                const size_t pi = noff + node_index(P, pc[0], pc[1], pc[2]);
                nis[off] = noff + node_index(P, co[0], co[1], co[2]);
                pm[off] = (float)((double)gsum[pi] * um);
                pmo[off] = (float)((double)mom[pi] * up);
                dm[off] = gsum[nis[off]];
            }
#pragma unroll
            for (int off = 0; off < 3; ++off) {
                if (w[off] == 0.0f || pm[off] <= 0.0f) continue;
                float trial = pmo[off] / pm[off] + g_a * dt;
                if (trial * inward >= 0.0f) continue;
                float distance = st.d[axis][off] * P.dx;
                float reaction = (normal_stress * distance * force_scale + weight_momentum) * w[off] * inward;
                if (reaction <= 0.0f) continue;
                if (dm[off] <= 0) continue;
                const int units = round_int(reaction * tscale);
                const size_t key = axis * na + nis[off];
                bool held = false;
                if (merge) {
                    unsigned h = ((unsigned)key * 2654435761u) >> 22;  // 10 bits: kTrHash
                    for (int probe = 0; probe < kTrProbes; ++probe) {
                        const unsigned old = atomicCAS(&hkey[h], kTrEmpty, (unsigned)key);
                        if (old == kTrEmpty || old == (unsigned)key) {
                            atomicAdd(&hval[h], (unsigned long long)(long long)units);
                            held = true;
                            break;
                        }
                        h = (h + 1) & (kTrHash - 1);
                    }
                }
                if (!held)
                    atomicAdd(reinterpret_cast<unsigned long long*>(B.trT + key), (unsigned long long)(long long)units);
                if (P.cells[axis] <= 2)
                    atomicAdd(reinterpret_cast<unsigned long long*>(B.trS + axis * na + nis[off]),
                              (unsigned long long)(long long)(face == 0.0f ? units : -units));
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kTrHash; i += kBlock) {
        const unsigned key = hkey[i];
        if (key != kTrEmpty) atomicAdd(reinterpret_cast<unsigned long long*>(B.trT) + key, hval[i]);
    }
}

// ------------------------------------------------------------ grid update

// grid_update (solver.rs:822-892).  kGridWgBlocks 64-node blocks of one
// instance per workgroup, kGridWgBlocks / 4 per wavefront; ledger sums go to
// per-workgroup partials (no same-address atomics), reduced once per
// workgroup.  The integer sums of the trial's deposit convert exactly to f64
// (they stay below 2^53) and round once to f32.  The other accumulator set,
// which the previous trial deposited into, is cleared on the way for the next
// trial's deposit.
// AUDIT: also the grid kinetic energy sum 1/2 m |v|^2 over nodes with mass at
// the stages where the reference's audit takes it (solver.rs:540-625,
// audit.rs:360-398): the APIC transport momentum alone (the accumulated
// sum w (m v + m C d) is exactly the audit's shadow deposit), after the stress
// impulse, after the wall traction, with gravity's kick, and after the walls.
// One node block per wavefront (kGridWgBlocks = 4): blocks with deposits lie
// together, so smaller workgroups spread them over more wavefronts; at eight
// wavefronts per SIMD the deposit-free ones cost little.
template <bool AUDIT>
__global__ __launch_bounds__(256, AUDIT ? 4 : 8) void grid_update_kernel(BatchDev B) {
    __shared__ double scratch[4 * kGridVals];
    __shared__ double ascratch[AUDIT ? 4 * kAuditVals : 1];
    __shared__ int sbad[4];
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    static_assert(kGridWgBlocks % 4 == 0 && kGridWgBlocks <= 32, "node blocks in whole wavefronts, flags in one ballot");
    constexpr int J = kGridWgBlocks / 4;  // node blocks per wavefront
    // The flags of both accumulator sets go out first, with the instance state
    // (which of the two is the trial's is only known from it), and a batch of
    // one instance needs no block-to-instance lookup: the sums wait for one
    // memory round trip instead of three.  Every wavefront reads both flags of
    // every block of the workgroup (lane i < G = kGridWgBlocks: set 0 of block
    // i, G <= i < 2G: set 1), so each knows whether any block of the workgroup
    // holds a deposit.
    const int fb = blockIdx.x * kGridWgBlocks + l % kGridWgBlocks;
    const int myflag = l < kGridWgBlocks ? B.gflag[0][fb] : (l < 2 * kGridWgBlocks ? B.gflag[1][fb] : 0);
    const unsigned long long fmask = __ballot(myflag != 0);
    int fl0[J], fl1[J];
#pragma unroll
    for (int j = 0; j < J; ++j) {
        fl0[j] = (int)((fmask >> (4 * j + wave)) & 1ull);
        fl1[j] = (int)((fmask >> (kGridWgBlocks + 4 * j + wave)) & 1ull);
    }
    const int inst = B.ninst == 1 ? 0 : B.blk_inst[blockIdx.x * kGridWgBlocks];
    const InstState& s = B.is[inst];
    if (!s.active) return;
    const InstParams& ip = B.ip[inst];
    const Params P = ip.P;  // private copy: no reloads behind the grid stores
    GridPtrs A = grid_set(B, s.gcur), C = grid_set(B, s.gcur ^ 1);
    const float dt = (float)s.dt;
    const size_t na = B.na;
    const float k4 = 4.0f * P.inv_dx * P.inv_dx;
    constexpr int NF = AUDIT ? kAuditFields : kDepositFields;
    const double um = 1.0 / (double)s.sc_m, up = 1.0 / (double)(AUDIT ? s.sc_mom : s.sc_p),
                 uf = 1.0 / (double)s.sc_f, ut = 1.0 / (double)s.sc_t;
    double vals[kGridVals] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double avals[kAuditVals] = {0, 0, 0, 0, 0};
    int bad = 0;
    {
        // All loads of the wavefront's blocks are issued before any result is used:
        // the flags of both sets, then the sums and traction of the touched blocks,
        // so a wavefront waits for two memory round trips instead of two per block.
        int cf[J], af[J];
#pragma unroll
        for (int j = 0; j < J; ++j) {
            cf[j] = s.gcur ? fl0[j] : fl1[j];
            af[j] = s.gcur ? fl1[j] : fl0[j];
        }
        float fm[J], fpx[J], fpy[J], fpz[J], ffx[J], ffy[J], ffz[J], ftb[J][3];
        bool wall[J];
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const int b = blockIdx.x * kGridWgBlocks + 4 * j + wave;
            const size_t n = ((size_t)b << 6) | l;
            fm[j] = fpx[j] = fpy[j] = fpz[j] = ffx[j] = ffy[j] = ffz[j] = 0.0f;
            wall[j] = false;
#pragma unroll
            for (int a = 0; a < 3; ++a) ftb[j][a] = 0.0f;
            if (!af[j]) continue;
            // The traction buffers can be nonzero only where wall_traction_to_grid
            // deposits: a particle's wall image lands its deposits on node 2 or node
            // `cells` of the wall axis, in the first block or the block holding node
            // `cells` along that axis.
            int bc[3];
            node_coords(P, (b - ip.boff) << 6, bc[0], bc[1], bc[2]);
#pragma unroll
            for (int a = 0; a < 3; ++a) wall[j] = wall[j] || bc[a] == 0 || (bc[a] >> 2) == (P.cells[a] >> 2);
            fm[j] = (float)((double)A.v[n] * um);
            fpx[j] = (float)((double)A.v[na + n] * up);
            fpy[j] = (float)((double)A.v[2 * na + n] * up);
            fpz[j] = (float)((double)A.v[3 * na + n] * up);
            if (AUDIT) {
                ffx[j] = (float)((double)A.v[4 * na + n] * uf);
                ffy[j] = (float)((double)A.v[5 * na + n] * uf);
                ffz[j] = (float)((double)A.v[6 * na + n] * uf);
            }
            if (wall[j])
#pragma unroll
                for (int a = 0; a < 3; ++a) ftb[j][a] = (float)((double)B.trT[a * na + n] * ut);
        }
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const int b = blockIdx.x * kGridWgBlocks + 4 * j + wave;
            const size_t n = ((size_t)b << 6) | l;
            if (cf[j]) {
#pragma unroll
                for (int f = 0; f < NF; ++f) C.v[f * na + n] = 0;
                if (l == 0) C.flag[b] = 0;
            }
        }
#pragma unroll
        for (int j = 0; j < J; ++j) {
            if (!af[j]) continue;
            const int b = blockIdx.x * kGridWgBlocks + 4 * j + wave;
            const size_t n = ((size_t)b << 6) | l;
            const bool wall_block = wall[j];
            const float m = fm[j];
            if (!(m > 0.0f)) {
                B.gv[n] = 0.0f;
                B.gv[na + n] = 0.0f;
                B.gv[2 * na + n] = 0.0f;
                continue;
            }
            // The audit deposits the APIC momentum and the stress sum apart; the
            // reference's P2G momentum is their combination (solver.rs:643-655).
            float p[3] = {fpx[j], fpy[j], fpz[j]};
            if (AUDIT) {
                p[0] = fpx[j] - dt * k4 * ffx[j];
                p[1] = fpy[j] - dt * k4 * ffy[j];
                p[2] = fpz[j] - dt * k4 * ffz[j];
            }
            float v[3], tb[3];
            int c[3];
            node_coords(P, (int)((size_t)(b - ip.boff) << 6) | l, c[0], c[1], c[2]);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                // trT holds a node's summed reactions; the signed impulse follows from
                // the face the node serves (node 2 the low face, inward +1; node
                // `cells` the high face, -1).  A grid two cells deep puts both faces
                // on node 2 and keeps the signed sum in trS.
                tb[a] = ftb[j][a];
                float D;
                if (P.cells[a] <= 2) {
                    const long long ds = wall_block ? B.trS[a * na + n] : 0;
                    D = (float)((double)ds * ut);
                    if (ds != 0) B.trS[a * na + n] = 0;
                } else {
                    D = c[a] == 2 ? tb[a] : -tb[a];
                }
                if (D != 0.0f) {
                    // Telescoped kinetic-energy jump of this node's deposits (grid.rs:265-295).
                    vals[8] += (double)D * ((double)p[a] / (double)m + (double)D / (2.0 * (double)m));
                    vals[3 + a] += D;
                }
                if (tb[a] != 0.0f) B.trT[a * na + n] = 0;
                v[a] = (p[a] + D) / m + P.gravity[a] * dt;
                if (AUDIT) {
                    const double md = (double)m;
                    const double pt = a == 0 ? (double)fpx[j] : (a == 1 ? (double)fpy[j] : (double)fpz[j]);
                    avals[0] += 0.5 * pt * pt / md;
                    avals[1] += 0.5 * (double)p[a] * (double)p[a] / md;
                    avals[2] += 0.5 * ((double)p[a] + (double)D) * ((double)p[a] + (double)D) / md;
                    avals[3] += 0.5 * md * (double)v[a] * (double)v[a];
                }
            }
            if (!(isfinite(v[0]) && isfinite(v[1]) && isfinite(v[2]))) {
                bad += 1;
                v[0] = v[1] = v[2] = 0.0f;
            } else {
                float before[3] = {v[0], v[1], v[2]};
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
                    vals[a] += ((double)v[a] - (double)before[a]) * (double)m;
                    if (AUDIT) avals[4] += 0.5 * (double)m * (double)v[a] * (double)v[a];
                }
            }
            B.gv[n] = v[0];
            B.gv[na + n] = v[1];
            B.gv[2 * na + n] = v[2];
        }
    }
    // The workgroup's partials.  A workgroup with no deposit in the trial's set
    // has nothing to sum (a uniform branch: every wavefront holds the same flag
    // mask): zeros, without the reduction and its barrier.  Otherwise each
    // wavefront sums by DPP into lane 63, and thread i combines value i over the
    // wavefronts in order.
    const unsigned long long amask = s.gcur ? fmask >> kGridWgBlocks : fmask;
    if ((amask & ((1ull << kGridWgBlocks) - 1ull)) == 0ull) {
        if (threadIdx.x < kGridVals) B.gpart[(size_t)threadIdx.x * B.ngwg + blockIdx.x] = 0.0;
        if (threadIdx.x == kGridVals) B.gbad[blockIdx.x] = 0;
        if (AUDIT && threadIdx.x < kAuditVals) B.apart[(size_t)threadIdx.x * B.ngwg + blockIdx.x] = 0.0;
        return;
    }
    const int nbad = wave_sum_i_last(bad);
#pragma unroll
    for (int i = 0; i < kGridVals; ++i) vals[i] = wave_sum_last(vals[i]);
    if (AUDIT)
#pragma unroll
        for (int i = 0; i < kAuditVals; ++i) avals[i] = wave_sum_last(avals[i]);
    if (l == 63) {
        sbad[wave] = nbad;
#pragma unroll
        for (int i = 0; i < kGridVals; ++i) scratch[wave * kGridVals + i] = vals[i];
        if (AUDIT)
#pragma unroll
            for (int i = 0; i < kAuditVals; ++i) ascratch[wave * kAuditVals + i] = avals[i];
    }
    __syncthreads();
    const int t = threadIdx.x;
    if (t < kGridVals)
        B.gpart[(size_t)t * B.ngwg + blockIdx.x] =
            (((0.0 + scratch[t]) + scratch[kGridVals + t]) + scratch[2 * kGridVals + t]) + scratch[3 * kGridVals + t];
    if (t == kGridVals) B.gbad[blockIdx.x] = sbad[0] + sbad[1] + sbad[2] + sbad[3];
    if (AUDIT && t >= 32 && t < 32 + kAuditVals) {
        const int i = t - 32;
        B.apart[(size_t)i * B.ngwg + blockIdx.x] =
            (((0.0 + ascratch[i]) + ascratch[kAuditVals + i]) + ascratch[2 * kAuditVals + i]) + ascratch[3 * kAuditVals + i];
    }
}

// ------------------------------------------------------------ pellet coupling

// couple_grid (rigid.rs:385-438) for the pellets of the node block's instance in
// pellet index order, so a node sees the pellets one after another as in the
// sequential loop over pellets.  A block outside the union of the pellets'
// boxes returns at once.  The wavefront tests 256 pellets at a time against the
// block's node box (the boxes' loads issued together), then walks the hits in
// index order.  f64 with the reference's operation order; node positions are
// (i - 1) dx exactly.  Each pellet's impulse, angular impulse and energy
// change over the block's constrained nodes are summed across the wavefront in
// lane order (a fixed order) and added to the pellet's exact fixed-point sums
// by one lane, so the totals do not depend on the order of the blocks.
__global__ __launch_bounds__(64) void couple_kernel(BatchDev B) {
    const int b = blockIdx.x, l = threadIdx.x;
    const int inst = B.blk_inst[b];
    const InstState& s = B.is[inst];
    const InstParams& ip = B.ip[inst];
    if (!s.active || ip.npel == 0) return;
    GridPtrs A = grid_set(B, s.gcur);
    if (!A.flag[b]) return;
    const double dx = ip.dx;
    // Node box of the block (node coordinates 4 bi .. 4 bi + 3 per axis).
    double blo[3], bhi[3];
    {
        int c0[3];
        node_coords(ip.P, (b - ip.boff) << 6, c0[0], c0[1], c0[2]);
        for (int a = 0; a < 3; ++a) {
            blo[a] = ((double)c0[a] - 1.0) * dx;
            bhi[a] = ((double)c0[a] + 2.0) * dx;
        }
    }
    if (bhi[0] < s.pel_box[0] || blo[0] > s.pel_box[3] || bhi[1] < s.pel_box[1] || blo[1] > s.pel_box[4] ||
        bhi[2] < s.pel_box[2] || blo[2] > s.pel_box[5])
        return;
    const size_t n = ((size_t)b << 6) | l;
    const size_t na = B.na;
    // The node mass as the grid update holds it (f32 of the exact sum).
    const double m = (double)(float)((double)A.v[n] * (1.0 / (double)s.sc_m));
    int ci, cj, ck;
    node_coords(ip.P, (int)((size_t)(b - ip.boff) << 6) | l, ci, cj, ck);
    const double node[3] = {((double)ci - 1.0) * dx, ((double)cj - 1.0) * dx, ((double)ck - 1.0) * dx};
    double v[3] = {B.gv[n], B.gv[na + n], B.gv[2 * na + n]};
    bool changed = false;
    constexpr int kGroups = 4;
    for (int c0 = 0; c0 < ip.npel; c0 += 64 * kGroups) {
        unsigned long long masks[kGroups];
        {
            double lo[kGroups][3], hi[kGroups][3];
            bool in[kGroups];
#pragma unroll
            for (int u = 0; u < kGroups; ++u) {
                const int k = c0 + 64 * u + l;
                in[u] = k < ip.npel;
                const PelletGeo& g = B.pgeo[ip.pel_off + (in[u] ? k : 0)];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    lo[u][a] = g.lo[a];
                    hi[u][a] = g.hi[a];
                }
            }
#pragma unroll
            for (int u = 0; u < kGroups; ++u) {
                const bool hit = in[u] && !(bhi[0] < lo[u][0] || blo[0] > hi[u][0] || bhi[1] < lo[u][1] ||
                                            blo[1] > hi[u][1] || bhi[2] < lo[u][2] || blo[2] > hi[u][2]);
                masks[u] = __ballot(hit);
            }
        }
        for (int u = 0; u < kGroups; ++u) {
            unsigned long long mask = masks[u];
            while (mask) {
                const int k = __ffsll((long long)mask) - 1;
                mask &= mask - 1;
                const int pi = ip.pel_off + c0 + 64 * u + k;
                const PelletGeo& g = B.pgeo[pi];
                bool act = !(m <= 0.0) && !(node[0] < g.lo[0] || node[0] > g.hi[0] || node[1] < g.lo[1] ||
                                        node[1] > g.hi[1] || node[2] < g.lo[2] || node[2] > g.hi[2]);
                double sums[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
                if (act) {
                    const double* offs = B.pel[pi].offsets;
                    double best = INFINITY, nrm[3] = {0.0, 0.0, 1.0};
                    for (int q = 0; q < g.spheres; ++q) {
                        double o = offs[q];
                        double dlt[3] = {node[0] - (g.pos[0] + g.axis[0] * o), node[1] - (g.pos[1] + g.axis[1] * o),
                                         node[2] - (g.pos[2] + g.axis[2] * o)};
                        double dist = sqrt(dlt[0] * dlt[0] + dlt[1] * dlt[1] + dlt[2] * dlt[2]);
                        double phi = dist - g.radius;
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
                    double r[3] = {node[0] - g.pos[0], node[1] - g.pos[1], node[2] - g.pos[2]};
                    double body[3] = {g.vel[0] + (g.omega[1] * r[2] - g.omega[2] * r[1]),
                                      g.vel[1] + (g.omega[2] * r[0] - g.omega[0] * r[2]),
                                      g.vel[2] + (g.omega[0] * r[1] - g.omega[1] * r[0])};
                    double rel[3] = {v[0] - body[0], v[1] - body[1], v[2] - body[2]};
                    double ns = rel[0] * nrm[0] + rel[1] * nrm[1] + rel[2] * nrm[2];
                    act = !(best > 0.0) && !(ns >= 0.0);  // rigid.rs:401-410: outside, or separating
                    if (act) {
                        double t[3] = {rel[0] - nrm[0] * ns, rel[1] - nrm[1] * ns, rel[2] - nrm[2] * ns};
                        double ts = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
                        double red = ip.wall_friction * (-ns);
                        double f = (ts <= red || ts < 1e-14) ? 0.0 : (ts - red) / ts;
                        double nv[3] = {body[0] + t[0] * f, body[1] + t[1] * f, body[2] + t[2] * f};
                        double dp[3] = {(nv[0] - v[0]) * m, (nv[1] - v[1]) * m, (nv[2] - v[2]) * m};
                        sums[0] = -dp[0];
                        sums[1] = -dp[1];
                        sums[2] = -dp[2];
                        sums[3] = -(r[1] * dp[2] - r[2] * dp[1]);
                        sums[4] = -(r[2] * dp[0] - r[0] * dp[2]);
                        sums[5] = -(r[0] * dp[1] - r[1] * dp[0]);
                        sums[6] = 0.5 * m *
                                  ((nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]) -
                                   (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
                        v[0] = nv[0];
                        v[1] = nv[1];
                        v[2] = nv[2];
                        changed = true;
                    }
                }
                const unsigned long long cm = __ballot(act);
                if (cm) {
#pragma unroll
                    for (int i = 0; i < 7; ++i) sums[i] = wave_sum_last(sums[i]);
                    if (l == 63) {
                        PelletAccum* acc = B.pacc + pi;
#pragma unroll
                        for (int i = 0; i < 7; ++i) {
                            const Fix128 add = fix_from(sums[i]);
                            atomicAdd(&acc->sum[i][0], add.lo & 0xffffffffull);
                            atomicAdd(&acc->sum[i][1], add.lo >> 32);
                            atomicAdd(&acc->sum[i][2], (unsigned long long)add.hi);
                        }
                        atomicAdd(&acc->constrained, (unsigned long long)__popcll(cm));
                    }
                }
            }
        }
    }
    if (changed) {
        B.gv[n] = (float)v[0];
        B.gv[na + n] = (float)v[1];
        B.gv[2 * na + n] = (float)v[2];
    }
}

// Pellet::apply_impulse and ::integrate (rigid.rs:251-334) on a trial copy,
// committed with the step.
__global__ void pellet_kernel(BatchDev B) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= B.npel) return;
    const PelletState& p = B.pel[k];
    if (p.inst < 0) return;
    const InstState& s = B.is[p.inst];
    if (!s.active) return;
    const InstParams& ip = B.ip[p.inst];
    const PelletAccum& acc = B.pacc[k];
    PelletTrial t;
    double vel[3], L[3], q[4], pos[3];
    for (int a = 0; a < 3; ++a) {
        vel[a] = p.vel[a];
        L[a] = p.L[a];
        pos[a] = p.pos[a];
    }
    for (int a = 0; a < 4; ++a) q[a] = p.q[a];
    double before = pellet_kinetic(p.mass, vel, q, L, p.inertia);
    for (int a = 0; a < 3; ++a) {
        vel[a] = vel[a] + pel_sum(acc, a) / p.mass;
        L[a] = L[a] + pel_sum(acc, 3 + a);
    }
    t.energy = pellet_kinetic(p.mass, vel, q, L, p.inertia) - before;
    // rigid.rs:279-334
    const double dt = s.dt;
    const double damping = ip.contact_damping * sqrt(ip.contact_k * p.mass);
    double force[3] = {0, 0, 0}, torque[3] = {0, 0, 0}, work = 0.0;
    double om[3], axis[3];
    pellet_omega(q, L, p.inertia, om);
    const double xh[3] = {1.0, 0.0, 0.0};
    qrot(q, xh, axis);
    const double ext[3] = {ip.P.cells[0] * ip.dx, ip.P.cells[1] * ip.dx, ip.P.cells[2] * ip.dx};
    for (int sph = 0; sph < p.spheres; ++sph) {
        double o = p.offsets[sph];
        double center[3] = {pos[0] + axis[0] * o, pos[1] + axis[1] * o, pos[2] + axis[2] * o};
        for (int a = 0; a < 3; ++a)
            for (int side = 0; side < 2; ++side) {
                double face = side == 0 ? 0.0 : ext[a];
                double sign = side == 0 ? 1.0 : -1.0;
                double pen = p.radius - (center[a] - face) * sign;
                if (pen <= 0.0) continue;
                double nrm[3] = {0, 0, 0};
                nrm[a] = sign;
                double lever[3] = {center[0] - nrm[0] * p.radius - pos[0], center[1] - nrm[1] * p.radius - pos[1],
                                   center[2] - nrm[2] * p.radius - pos[2]};
                double pv[3] = {vel[0] + (om[1] * lever[2] - om[2] * lever[1]),
                                vel[1] + (om[2] * lever[0] - om[0] * lever[2]),
                                vel[2] + (om[0] * lever[1] - om[1] * lever[0])};
                double vn = pv[0] * nrm[0] + pv[1] * nrm[1] + pv[2] * nrm[2];
                double fn = fmax(ip.contact_k * pen - damping * vn, 0.0);
                double tg[3] = {pv[0] - nrm[0] * vn, pv[1] - nrm[1] * vn, pv[2] - nrm[2] * vn};
                double ts = sqrt(tg[0] * tg[0] + tg[1] * tg[1] + tg[2] * tg[2]);
                double fr[3] = {0, 0, 0};
                if (ts > 1e-12) {
                    double lim = fmin(ip.contact_friction * fn, p.mass * ts / dt);
                    for (int c = 0; c < 3; ++c) fr[c] = -tg[c] / ts * lim;
                }
                double total[3] = {nrm[0] * fn + fr[0], nrm[1] * fn + fr[1], nrm[2] * fn + fr[2]};
                for (int c = 0; c < 3; ++c) force[c] += total[c];
                torque[0] += lever[1] * total[2] - lever[2] * total[1];
                torque[1] += lever[2] * total[0] - lever[0] * total[2];
                torque[2] += lever[0] * total[1] - lever[1] * total[0];
                work += (total[0] * pv[0] + total[1] * pv[1] + total[2] * pv[2]) * dt;
            }
    }
    for (int a = 0; a < 3; ++a) {
        double imp = force[a] * dt;
        t.wall_impulse[a] = imp;
        vel[a] = vel[a] + (ip.gravity[a] * dt + imp / p.mass);
        L[a] = L[a] + torque[a] * dt;
        pos[a] = pos[a] + vel[a] * dt;
    }
    t.wall_work = work;
    // orientation = UnitQuaternion::from_scaled_axis(omega dt) * orientation
    double omn[3];
    pellet_omega(q, L, p.inertia, omn);
    double half[3] = {omn[0] * dt * 0.5, omn[1] * dt * 0.5, omn[2] * dt * 0.5};
    double nn = half[0] * half[0] + half[1] * half[1] + half[2] * half[2];
    double inc[4] = {0.0, 0.0, 0.0, 1.0};
    const double eps = 2.220446049250313e-16;
    if (nn > eps * eps) {
        double nr = sqrt(nn);
        double f = sin(nr) / nr;
        inc[0] = half[0] * f;
        inc[1] = half[1] * f;
        inc[2] = half[2] * f;
        inc[3] = cos(nr);
    }
    double nq[4] = {inc[3] * q[0] + inc[0] * q[3] + inc[1] * q[2] - inc[2] * q[1],
                    inc[3] * q[1] - inc[0] * q[2] + inc[1] * q[3] + inc[2] * q[0],
                    inc[3] * q[2] + inc[0] * q[1] - inc[1] * q[0] + inc[2] * q[3],
                    inc[3] * q[3] - inc[0] * q[0] - inc[1] * q[1] - inc[2] * q[2]};
    bool finite = isfinite(t.energy);
    for (int a = 0; a < 3; ++a) {
        t.pos[a] = pos[a];
        t.vel[a] = vel[a];
        t.L[a] = L[a];
        finite = finite && isfinite(pos[a]) && isfinite(vel[a]) && isfinite(L[a]);
    }
    for (int a = 0; a < 4; ++a) {
        t.q[a] = nq[a];
        finite = finite && isfinite(nq[a]);
    }
    t.bad = finite ? 0 : 1;
    t.pad = 0;
    B.ptrial[k] = t;
}

// ------------------------------------------------------------ fused G2P2G

// Appends the near-wall particles of one work item to the instance's list:
// bit r of a thread's `bits` flags particle first + r * kBlock + threadIdx.x.
// One atomic on the list counter per workgroup (wavefront appends on a single
// counter serialise); the entries follow round, wavefront and lane order.
// Every thread of the workgroup must call it.
template <int R>
__device__ inline void block_append_rounds(unsigned bits, int first, int* list, int* count) {
    __shared__ int swc[kBlock / 64], sbase;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int wcount = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) wcount += __popcll(__ballot((bits >> r) & 1u));
    if (lane == 0) swc[wave] = wcount;
    __syncthreads();
    if (threadIdx.x == 0) {
        int total = 0;
        for (int w = 0; w < kBlock / 64; ++w) total += swc[w];
        sbase = total ? atomicAdd(count, total) : 0;
    }
    __syncthreads();
    int at = sbase;
    for (int w = 0; w < wave; ++w) at += swc[w];
    const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const bool on = (bits >> r) & 1u;
        const unsigned long long mask = __ballot(on);
        if (on) list[at + __popcll(mask & below)] = first + r * kBlock + (int)threadIdx.x;
        at += __popcll(mask);
    }
}

// The slab reads of one particle's G2P update: position, mass, V0, the plastic
// strain (paste) or J - 1 (liquid) and, for a paste, H = F - I.  Nothing is
// read for p >= end.
struct G2pIn {
    float x, y, z, pm, pv0, pe, H[9];
};
__device__ inline void g2p_load(const float* __restrict__ src, const float* __restrict__ shared, size_t cap, int kind,
                                int p, int end, G2pIn& o) {
    o.x = o.y = o.z = o.pm = o.pv0 = o.pe = 0.0f;
#pragma unroll
    for (int i = 0; i < 9; ++i) o.H[i] = 0.0f;
    if (p >= end) return;
    o.x = src[PX * cap + p];
    o.y = src[PY * cap + p];
    o.z = src[PZ * cap + p];
    o.pm = shared[SMASS * cap + p];
    o.pv0 = shared[SVOL0 * cap + p];
    o.pe = src[(kind == 0 ? EPS : JM1) * cap + p];
    if (kind == 0) {
#pragma unroll
        for (int i = 0; i < 9; ++i) o.H[i] = src[(H00 + i) * cap + p];
    }
}

// grid_to_particle with the constitutive update (solver.rs:895-977,
// constitutive.rs:456-589) into the trial slab.  One workgroup per work item:
// the node velocities of the item's stencils (the box of the committed slab's
// set-up record) are staged in LDS, and each particle gathers v and C,
// advects, and updates its deformation and stress.  The step's reductions go
// to per-item partials: the largest speed, displacement and deposit bounds,
// the plastic dissipation as an exact fixed-point sum, the first failure by
// particle id.  The set-up record and near-wall list of the trial slab are
// written for the next trial.  MLP: the return-map surrogate is on.  WGS:
// workgroups per CU the register budget allows (the launch shape picks it).
// LOG: the log of exact returns may be on (it keeps C live through the update).
template <bool MLP, int WGS, bool LOG>
__global__ __launch_bounds__(kBlock, WGS) void g2p_kernel(BatchDev B) {
    constexpr int R = kChunk / kBlock;
    __shared__ float4 sg[kTileCap];
    __shared__ int snbox[4][6];
    __shared__ float sred[4][4];
    __shared__ Fix128 sdiss[4];
    __shared__ unsigned long long sfail[4];
    __shared__ int sesc[4];
    if (B.ktime) kernel_start(B.ktime, 2);
    const int4 item = B.work[blockIdx.x];
    const int inst = item.x;
    InstState& is = B.is[inst];
    if (!is.active) return;
    const InstParams& ip = B.ip[inst];
    // A private copy: read through the reference, the compiler reloads every
    // field after each store to the slabs (it cannot rule out aliasing), one
    // memory round trip per use inside the particle loop.  The surrogate's
    // instance keeps the reference: its tiles need the registers the copy holds.
    const Params Pv = ip.P;
    const Params& P = MLP ? ip.P : Pv;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int chunk0 = item.y;
    const int chunk_end = item.y + item.z;
    const float dt = (float)is.dt;
    const size_t cap = B.cap, na = B.na;
    const float* __restrict__ src = B.slab[is.cur];
    float* __restrict__ dst = B.slab[is.cur ^ 1];
    const float* __restrict__ shared = B.shared[B.sh];
    const size_t noff = (size_t)ip.boff * 64;
    const float* __restrict__ gv = B.gv + noff;
    int* nw_list = B.nw_list[is.cur ^ 1] + ip.poff;
    int* nw_count = &is.nw_count[is.cur ^ 1];

    // The committed slab's set-up record: the box of the stencil bases of the
    // item's particles, whose nodes the gather tile holds.
    const ItemPre pre = B.wpre[is.cur][blockIdx.x];
    Tile t;
    {
        bool ok = true;
        int gsz = 1;
#pragma unroll
        for (int a = 0; a < 3; ++a) ok = ok && pre.lo[a] <= pre.hi[a];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            t.org[a] = ok ? pre.lo[a] : 0;
            t.gd[a] = ok ? pre.hi[a] - pre.lo[a] + 3 : 0;
            gsz *= t.gd[a];
        }
        t.tiled = ok && gsz <= kTileCap;
    }

    // The first round's particle reads go out before the tile is staged, so
    // their memory latency overlaps the staging; each later round's reads go
    // out a round ahead (below).  The surrogate's instance has no registers to
    // spare for the reads in flight: it reads each round's at its start.
    G2pIn cur;
    if (!MLP) g2p_load(src, shared, cap, P.kind, chunk0 + tid, chunk_end, cur);

    // Phase A: the gather tile, four nodes per thread at a time with all their
    // loads issued before the first LDS store (one memory round trip per four).
    if (t.tiled) {
        int gsz = t.gd[0] * t.gd[1] * t.gd[2];
        const float ig2 = 1.0f / (float)t.gd[2], ig1 = 1.0f / (float)t.gd[1];
        for (int q0 = tid; q0 < gsz; q0 += 4 * kBlock) {
            float gx[4], gy[4], gz[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int q = q0 + u * kBlock;
                gx[u] = gy[u] = gz[u] = 0.0f;
                if (q < gsz) {
                    const int rest = div_small(q, ig2), k = q - rest * t.gd[2];
                    const int i = div_small(rest, ig1), j = rest - i * t.gd[1];
                    const int ni = node_index(P, t.org[0] + i, t.org[1] + j, t.org[2] + k);
                    gx[u] = gv[ni];
                    gy[u] = gv[na + ni];
                    gz[u] = gv[2 * na + ni];
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int q = q0 + u * kBlock;
                if (q < gsz) sg[q] = make_float4(gx[u], gy[u], gz[u], 0.0f);
            }
        }
    }
    __syncthreads();

    // Phase B: per particle, G2P, advection, constitutive update and writes to
    // the trial slab.
    float speed = 0.0f, disp = 0.0f, bnd1 = 0.0f, bnd2 = 0.0f;
    Fix128 diss = {0ull, 0ll};
    unsigned long long fail = ~0ull;
    int escaped = 0;
    unsigned long long mlp_tried = 0, mlp_taken = 0;
    // The trial slab's set-up record, gathered round by round as uniform values.
    int nlo[3] = {INT_MAX, INT_MAX, INT_MAX}, nhi[3] = {INT_MIN, INT_MIN, INT_MIN};
    unsigned nwbits = 0u;  // near-wall flags of the thread's updated particles, bit r: round r
#pragma unroll 1
    for (int r = 0; r < R; ++r) {
        const int p = chunk0 + r * kBlock + tid;
        // The next round's reads go out now: their latency overlaps this round's
        // gather and update (the registers for them fit in the occupancy's budget).
        G2pIn nxt;
        if (MLP)
            g2p_load(src, shared, cap, P.kind, p, chunk_end, cur);
        else
            g2p_load(src, shared, cap, P.kind, p + kBlock, r + 1 < R ? chunk_end : 0, nxt);
        bool nw = false;
        // Trial: gather, advection, trial deformation (or the whole liquid update).
        bool live = false, yielded = false;
        int rc = F_OK;
        float x = 0.0f, y = 0.0f, z = 0.0f;
        float v[3] = {0.0f, 0.0f, 0.0f}, C[9], Ht[9], tau[6];
        bool dep = false;  // the updated particle lies inside the padded grid (stencil s2)
        Stencil s2;
        float jn = 0.0f;
        PasteTrial ts;
        // This round's reads, issued a round earlier (the barrier keeps the
        // scheduler from sinking the next round's reads to their first use).
        float Hp[9];
        x = cur.x;
        y = cur.y;
        z = cur.z;
        const float pm = cur.pm, pv0 = cur.pv0, pe = cur.pe;
#pragma unroll
        for (int i = 0; i < 9; ++i) Hp[i] = cur.H[i];
        if (!MLP) cur = nxt;
        __builtin_amdgcn_sched_barrier(0);
        if (p < chunk_end) {
            Stencil s;
            if (!make_stencil(P, x, y, z, s)) {
                // Outside the padded grid: frozen for this trial (solver.rs:905-917).
                for (int f = 0; f < NPF; ++f) dst[f * cap + p] = src[f * cap + p];
                escaped += 1;
            } else {
                live = true;
                float Bm[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                bool in_box = t.tiled;
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const int li = s.base[a] - t.org[a];
                    in_box = in_box && li >= 0 && li + 2 < t.gd[a];
                }
                if (in_box)
                    gather_lds(sg, t, s, v, Bm);
                else
                    gather_global(P, gv, na, s, v, Bm);
                const float k4 = 4.0f * P.inv_dx;
#pragma unroll
                for (int i = 0; i < 9; ++i) C[i] = Bm[i] * k4;
                float vn = particle_speed(v);
                speed = fmaxf(speed, vn);
                disp = fmaxf(disp, vn * dt * P.inv_dx);
                if (P.kind == 0) {
                    const float* H = Hp;
                    // (I + dt C)(I + H) - I = H + dt C + dt C H
#pragma unroll
                    for (int i = 0; i < 3; ++i)
#pragma unroll
                        for (int j = 0; j < 3; ++j)
                            Ht[3 * i + j] = H[3 * i + j] + dt * C[3 * i + j] +
                                            dt * (C[3 * i] * H[j] + C[3 * i + 1] * H[3 + j] + C[3 * i + 2] * H[6 + j]);
                }
                // Advection and every write and reduction that needs no
                // constitutive result go out before the update, so that x, v and
                // C are dead through it (fewer registers live at the peak).  A
                // failed update rejects the whole trial, so these writes, the
                // set-up box and the near-wall flag are then never used.
                const float nx = x + v[0] * dt, ny = y + v[1] * dt, nz = z + v[2] * dt;
                dst[PX * cap + p] = nx;
                dst[PY * cap + p] = ny;
                dst[PZ * cap + p] = nz;
                dst[VX * cap + p] = v[0];
                dst[VY * cap + p] = v[1];
                dst[VZ * cap + p] = v[2];
#pragma unroll
                for (int i = 0; i < 9; ++i) dst[(C00 + i) * cap + p] = C[i];
                bnd1 = fmaxf(bnd1, particle_bound_motion(P, pm, v, C));
                if (make_stencil(P, nx, ny, nz, s2)) {
                    dep = true;
                    nw = near_wall(P, nx, ny, nz);
                } else {
                    escaped += 1;
                }
                if (P.kind == 0) {
                    rc = paste_trial(P, Ht, ts);
                    yielded = rc == F_OK && paste_yields(P, ts);
                } else {
                    rc = liquid_update(P, dt, pe, C, jn, tau);
                }
            }
        }
        // Surrogate return for the wavefront's yielded particles.  Everything in
        // this block is uniform across the wavefront, so the MFMAs run with all
        // lanes; the samples reach the MFMA operand lanes by wave shuffles and the
        // weights come from global memory (21 KB, cache resident), so the surrogate
        // needs no LDS.
        float q = 0.0f;
        bool settled = false;
        if (MLP && P.kind == 0) {
            const bool want = yielded && P.K > 0.0f;
            const unsigned long long mask = __ballot(want);
            if (mask) {
                float f[kMlpIn];
                if (want) {
                    surrogate_features(P, dt, ts, C, f);
                } else {
#pragma unroll
                    for (int i = 0; i < kMlpIn; ++i) f[i] = 0.0f;
                }
                // Features as six words of two bf16 each.
                int w[6];
#pragma unroll
                for (int j = 0; j < 6; ++j) w[j] = (int)bf16_bits(f[2 * j]) | ((int)bf16_bits(f[2 * j + 1]) << 16);
                const int g = lane >> 4;  // operand lanes of group g take features 4g .. 4g + 3
                // The wanted samples are packed: sample r (the r-th wanted lane)
                // goes to column r % 16 of tile r / 16, so only ceil(count / 16)
                // tiles run.  Columns past the count read any lane (finite
                // features) and are ignored.
                const int count = __popcll(mask);
                const int rank = __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                                           __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
                float yhat = 0.0f;
#pragma unroll 1
                for (int mt = 0; mt * 16 < count; ++mt) {
                    const int r = mt * 16 + (lane & 15);
                    const int srcl = r < count ? nth_set_bit(mask, r) : lane;
                    int ww[6];
#pragma unroll
                    for (int j = 0; j < 6; ++j) ww[j] = __shfl(w[j], srcl, 64);
                    const int lo = g == 0 ? ww[0] : (g == 1 ? ww[2] : (g == 2 ? ww[4] : 0));
                    const int hi = g == 0 ? ww[1] : (g == 1 ? ww[3] : (g == 2 ? ww[5] : 0));
                    bf16x4 xb;
                    xb[0] = (short)(lo & 0xffff);
                    xb[1] = (short)((unsigned)lo >> 16);
                    xb[2] = (short)(hi & 0xffff);
                    xb[3] = (short)((unsigned)hi >> 16);
                    MlpTile tl;
                    mlp_tile_forward<false>(B.mlp_fwd, B.mlp_bias, xb, lane, tl);
                    // Output 6 (the retained overstress fraction) of column j
                    // sits in lane 16 + j, element 2.
                    const float y6 = __shfl(tl.y[2], 16 + (rank & 15), 64);
                    if (want && (rank >> 4) == mt) yhat = y6;
                }
                if (want) settled = surrogate_polish(P, dt, ts.qtr, yhat, B.mlp_polish, q) <= B.mlp_tol;
                mlp_tried += __popcll(mask);
                mlp_taken += __popcll(__ballot(settled));
            }
        }
        // Finish: exact return where the surrogate was not used or not accepted,
        // and the trial-slab writes.
        if (live) {
            if (P.kind == 0) {
                float inc = 0.0f, Hn[9];
                if (rc == F_OK) {
                    if (!yielded)
                        q = ts.qtr;
                    else if (settled)
                        inc = (ts.qtr - q) / (3.0f * P.G);
                    else
                        rc = hb_return(P, dt, ts.qtr, q, inc);
                }
                if (rc == F_OK) rc = paste_finish(P, ts, Ht, q, yielded, Hn, tau);
                if (rc == F_OK) {
#pragma unroll
                    for (int i = 0; i < 9; ++i) dst[(H00 + i) * cap + p] = Hn[i];
                    // Only a yielded particle dissipates; a wavefront with none skips
                    // the f64 fixed-point conversion (a zero adds nothing).
                    if (inc != 0.0f) fix_add(diss, fix_from((double)(q * inc) * (double)pv0));
                    dst[EPS * cap + p] = pe + inc;
                }
            } else if (rc == F_OK) {
                dst[JM1 * cap + p] = jn;
            }
            if (rc != F_OK) {
                unsigned long long f = ((unsigned long long)B.ids[p] << 8) | (unsigned long long)rc;
                fail = f < fail ? f : fail;
            } else {
#pragma unroll
                for (int i = 0; i < 6; ++i) dst[(TXX + i) * cap + p] = tau[i];
                bnd2 = fmaxf(bnd2, particle_bound_stress(P, pv0, tau));
            }
        }
        nwbits |= nw ? (1u << r) : 0u;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            nlo[a] = min(nlo[a], wave_min_uniform(dep ? s2.base[a] : INT_MAX));
            nhi[a] = max(nhi[a], wave_max_uniform(dep ? s2.base[a] : INT_MIN));
        }
        // Log of exact returns (training data of the surrogate).
        if (LOG && B.rlog) {
            const bool lg = live && P.kind == 0 && yielded && rc == F_OK;
            const unsigned long long mk = __ballot(lg);
            if (mk) {
                const int leader = __ffsll((long long)mk) - 1;
                int base = 0;
                if (lane == leader) base = atomicAdd(B.rlog_count, (int)__popcll(mk));
                base = __shfl(base, leader, 64);
                const int slot = base + (int)__popcll(mk & ((1ull << lane) - 1ull));
                if (lg && slot < B.rlog_cap) {
                    float f[kMlpIn], tg[kMlpOut];
                    surrogate_features(P, dt, ts, C, f);
                    surrogate_targets(P, ts, f, q, tg);
                    float* o = B.rlog + (size_t)slot * kLogStride;
#pragma unroll
                    for (int i = 0; i < kMlpIn; ++i) o[i] = f[i];
#pragma unroll
                    for (int i = 0; i < kMlpOut; ++i) o[kMlpIn + i] = tg[i];
                    o[19] = ts.qtr;
                    o[20] = P.ys;
                    o[21] = 3.0f * P.G * dt;
                    o[22] = P.K;
                    o[23] = 1.0f / P.inv_n;
                    o[24] = q;
                }
            }
        }
    }
    // Surrogate counters: summed over the workgroup first, so one pair of
    // atomics per workgroup reaches the two global counters.
    __shared__ unsigned long long smlp[4][2];
    if (MLP && lane == 0) {
        smlp[wave][0] = mlp_tried;
        smlp[wave][1] = mlp_taken;
    }
    __syncthreads();
    if (MLP && tid == 0) {
        const unsigned long long tried = smlp[0][0] + smlp[1][0] + smlp[2][0] + smlp[3][0];
        if (tried) {
            atomicAdd(B.mlp_stats, tried);
            atomicAdd(B.mlp_stats + 1, smlp[0][1] + smlp[1][1] + smlp[2][1] + smlp[3][1]);
        }
    }
    block_append_rounds<R>(nwbits, chunk0, nw_list, nw_count);
    // Per-work-item partials of the step reductions.
    speed = wave_fmax_uniform(speed);
    disp = wave_fmax_uniform(disp);
    bnd1 = wave_fmax_uniform(bnd1);
    bnd2 = wave_fmax_uniform(bnd2);
    diss = wave_fix_sum(diss);
    fail = wave_min_u64(fail);
    escaped = wave_sum_i(escaped);
    if (lane == 0) {
        sdiss[wave] = diss;
        sred[wave][0] = speed;
        sred[wave][1] = disp;
        sred[wave][2] = bnd1;
        sred[wave][3] = bnd2;
        sfail[wave] = fail;
        sesc[wave] = escaped;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            snbox[wave][a] = nlo[a];
            snbox[wave][3 + a] = nhi[a];
        }
    }
    __syncthreads();
    if (tid == 0) {
        // Set-up record of the trial slab for the next step (used if it commits).
        ItemPre rec;
        for (int a = 0; a < 3; ++a) {
            rec.lo[a] = min(min(snbox[0][a], snbox[1][a]), min(snbox[2][a], snbox[3][a]));
            rec.hi[a] = max(max(snbox[0][3 + a], snbox[1][3 + a]), max(snbox[2][3 + a], snbox[3][3 + a]));
        }
        B.wpre[is.cur ^ 1][blockIdx.x] = rec;
        float red[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        Fix128 d = {0ull, 0ll};
        unsigned long long f = ~0ull;
        int e = 0;
        for (int w = 0; w < 4; ++w) {
            for (int k = 0; k < 4; ++k) red[k] = fmaxf(red[k], sred[w][k]);
            fix_add(d, sdiss[w]);
            f = sfail[w] < f ? sfail[w] : f;
            e += sesc[w];
        }
        // Into one of the instance's step-reduction slots, by non-returning
        // atomics; identities are skipped (a maximum of +0 or -0 among them:
        // the maxima start at +0).
        ItemSlot* sl = B.islot + inst * kItemSlots + blockIdx.x % kItemSlots;
        for (int k = 0; k < 4; ++k)
            if (red[k] > 0.0f) atomicMax(&sl->max_bits[k], __float_as_uint(red[k]));
        if (f != ~0ull) atomicMin(&sl->fail, f);
        if ((d.lo | (unsigned long long)d.hi) != 0ull) {
            atomicAdd(&sl->diss[0], d.lo & 0xffffffffull);
            atomicAdd(&sl->diss[1], d.lo >> 32);
            atomicAdd(&sl->diss[2], (unsigned long long)d.hi);
        }
        if (e) atomicAdd(&sl->escaped, e);
    }
    // The instance's first kGridChunks work items also sum the grid update's
    // partials, chunk by chunk (the grid update ended before this launch), so
    // the finalize kernel reads kGridChunks records per value instead of one
    // per grid workgroup: wavefront w takes the values w, w + 4, ..., lane l
    // the chunk's workgroups l, l + 64, ... in order, then a fixed-order
    // wavefront sum.
    {
        const int j = (int)blockIdx.x - is.w_off;
        const int K = min(is.w_cnt, kGridChunks);
        if (j < K) {
            const int g0 = ip.boff / kGridWgBlocks, ng = ip.nblk / kGridWgBlocks;
            const int cs = (ng + K - 1) / K;
            const int c0 = g0 + j * cs, c1 = min(c0 + cs, g0 + ng);
            for (int v = wave; v <= kGridVals; v += 4) {
                if (v < kGridVals) {
                    double a = 0.0;
                    for (int q = c0 + lane; q < c1; q += 64) a += B.gpart[(size_t)v * B.ngwg + q];
                    a = wave_sum_last(a);
                    if (lane == 63) B.gchunk[((size_t)inst * kGridVals + v) * kGridChunks + j] = a;
                } else {
                    int nb = 0;
                    for (int q = c0 + lane; q < c1; q += 64) nb += B.gbad[q];
                    nb = wave_sum_i(nb);
                    if (lane == 0) B.gchunk_bad[inst * kGridChunks + j] = nb;
                }
            }
        }
    }
    if (B.ktime) kernel_end(B.ktime, 3);
}

// ------------------------------------------------------------ reductions and commit

// Accept or reject one instance's trial (solver.rs:444-467, 593-611, 980-998):
// failures in the order the reference checks them, then the commit and the
// ledger sums.  `pel` holds the instance's pellet sums: total mass, wall impulse
// (3), coupling impulse (3) and angular impulse (3), wall work, coupling grid
// energy and the pellets' kinetic-energy change from the coupling.
__device__ inline void commit_instance(InstState& s, const InstParams& ip, const double* pel, double pel_vmax) {
    const StepStatus& st = s.st;
    double* L = s.ledger;
    int code = 0;
    if (!isfinite(st.traction_energy) || !isfinite(L[23] + st.traction_energy))
        code = F_TRACTION;
    else if (st.grid_nonfinite)
        code = F_GRID;
    else if (st.pellet_bad)
        code = F_PELLET;
    else if (st.fail != ~0ull)
        code = (int)(st.fail & 0xff);
    else if ((double)st.max_disp > 1.0)
        code = F_DISPLACEMENT;
    s.last_failure = code;
    s.refsort = 0;
    // Every trial deposited into its own accumulator set; the next one takes
    // the other (cleared by this trial's grid update).
    s.gcur ^= 1;
    const bool trial_only = s.trial_dt > 0.0;
    if (code) {
        if (trial_only) {
            s.trial_dt = 0.0;
            s.mode = M_DONE;
            return;
        }
        s.rejected_steps += 1;
        s.dt_scale *= 0.5;
        if (s.dt_scale * s.limit < ip.min_dt) {
            s.mode = M_ERROR;
            s.err = E_MIN_AFTER_REJECT;
        }
        return;
    }
    const double dt = s.dt;
    s.cur ^= 1;
    s.committed = 1;
    s.commit_clock = (unsigned long long)__builtin_amdgcn_s_memrealtime();
    s.bound[0] = st.bound[0];
    s.bound[1] = st.bound[1];
    // commit, solver.rs:992-997
    L[27] += st.dissipation;
    for (int a = 0; a < 3; ++a) L[9 + a] += st.traction_impulse[a] + st.wall_impulse[a];
    L[21] += st.projection;
    L[22] += st.friction;
    L[23] += st.traction_energy;
    for (int a = 0; a < 3; ++a) L[6 + a] += ip.gravity[a] * (ip.particle_mass * dt);
    // pellets, solver.rs:600-609 (summed over the pellets first)
    if (ip.npel > 0) {
        for (int c = 0; c < 3; ++c) {
            L[6 + c] += ip.gravity[c] * (pel[0] * dt);
            L[12 + c] += pel[1 + c];
            L[15 + c] += pel[4 + c];
            L[18 + c] += pel[7 + c];
        }
        L[24] += pel[10];
        L[25] += pel[11];
        L[26] += pel[12];
        s.pel_vmax = pel_vmax;
    }
    s.max_speed = (double)st.max_speed;
    if (st.escaped) {
        s.mode = M_PAUSE;
        s.pause |= P_ESCAPED;
    }
    if (trial_only) {
        s.trial_dt = 0.0;
        if (s.mode == M_RUN) s.mode = M_DONE;
        return;
    }
    // advance_while bookkeeping, solver.rs:446-454
    s.time += dt;
    s.step_count += 1;
    s.limited_steps += s.limited ? 1 : 0;
    s.min_dt_used = fmin(s.min_dt_used, dt);
    s.max_dt_used = fmax(s.max_dt_used, dt);
    s.dt_scale = fmin(s.dt_scale * kDtRecovery, 1.0);
    s.budget -= 1;
    // The next trial's G2P2G (or the host, if no trial follows) writes the order keys.
    if (s.step_count % kSortInterval == 0) s.refsort = 1;
}

constexpr int kPelSums = 13;

// Block-wide fixed-order sum of one double (every thread gets the total).
__device__ inline double block_total(double v, double* sd) {
    const int tid = threadIdx.x;
    v = wave_sum(v);
    if ((tid & 63) == 0) sd[tid >> 6] = v;
    __syncthreads();
    double t = sd[0] + sd[1] + sd[2] + sd[3];
    __syncthreads();
    return t;
}

// The last finalize workgroup to finish (a counter of finished workgroups,
// with fences around it) writes every instance's range in the concatenated
// near-wall list, as the control kernel's scan does, and resets the counter.
// Every thread of every finalize workgroup calls it once.
__device__ inline void near_wall_prefix_last(const BatchDev& B) {
    __shared__ int slast;
    __shared__ int scarry;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        const int done = atomicAdd(&B.counters[0], 1);
        slast = done == B.ninst - 1;
        scarry = 0;
    }
    __syncthreads();
    if (!slast) return;
    __threadfence();
    __shared__ int sscan[256];
    const int tid = threadIdx.x;
    for (int t0 = 0; t0 < B.ninst; t0 += 256) {
        const int t = t0 + tid;
        int c = 0;
        if (t < B.ninst) {
            const InstState& s = B.is[t];
            c = s.active ? s.nw_count[s.cur] : 0;
        }
        sscan[tid] = c;
        __syncthreads();
        for (int o = 1; o < 256; o <<= 1) {
            const int v = tid >= o ? sscan[tid - o] : 0;
            __syncthreads();
            sscan[tid] += v;
            __syncthreads();
        }
        if (t < B.ninst) B.is[t].nw_base = scarry + sscan[tid];
        __syncthreads();
        if (tid == 255) scarry += sscan[255];
        __syncthreads();
    }
    if (tid == 0) {
        B.counters[1] = scarry;
        B.counters[0] = 0;
    }
}

// One workgroup per running instance: the step reductions of its trial, then
// the accept/reject.  Everything the kernel reads comes in one memory round
// trip: the instance's state and parameters, staged in LDS by all threads,
// and, in wavefront 0, the grid chunk sums (lane c: chunk c) and G2P's
// step-reduction slots (lane s < kItemSlots: slot s), both at addresses that
// depend on the instance index alone.  Wavefront 0 reduces them (sums by DPP
// into lane 63, in a fixed order; extrema and integer sums by DPP and
// shuffles) and clears them for the next trial; lane 63 commits the instance
// in LDS (commit_instance) and all threads write the state back.  The
// instance's pellets are then committed (rigid.rs:334, solver.rs:600-609) and
// their geometry for the next trial computed from the committed state, with
// cleared coupling sums.
__global__ __launch_bounds__(256) void finalize_kernel(BatchDev B) {
    __shared__ double sd[4];
    __shared__ InstState ls;
    __shared__ InstParams lp;
    static_assert(sizeof(InstState) % 4 == 0 && sizeof(InstParams) % 4 == 0, "staged as 32-bit words");
    static_assert(kGridChunks == 64, "one grid chunk per lane of wavefront 0");
    constexpr int kWords = sizeof(InstState) / 4, kPWords = sizeof(InstParams) / 4;
    const int inst = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    InstState& gs = B.is[inst];
    {
        const unsigned* src = reinterpret_cast<const unsigned*>(&gs);
        unsigned* dst = reinterpret_cast<unsigned*>(&ls);
        for (int w = tid; w < kWords; w += 256) dst[w] = src[w];
        const unsigned* psrc = reinterpret_cast<const unsigned*>(&B.ip[inst]);
        unsigned* pdst = reinterpret_cast<unsigned*>(&lp);
        for (int w = tid; w < kPWords; w += 256) pdst[w] = psrc[w];
    }
    double acc[kGridVals];
    int bad = 0;
    ItemSlot sl = {{0u, 0u, 0u, 0u}, ~0ull, {0ull, 0ull, 0ull}, 0, {0, 0, 0}};
    if (wave == 0) {
#pragma unroll
        for (int v = 0; v < kGridVals; ++v) acc[v] = B.gchunk[((size_t)inst * kGridVals + v) * kGridChunks + lane];
        bad = B.gchunk_bad[inst * kGridChunks + lane];
        if (lane < kItemSlots) sl = B.islot[inst * kItemSlots + lane];
    }
    __syncthreads();
    InstState& s = ls;
    const InstParams& ip = lp;
    if (!s.active) {
        if (B.fuse_control && B.ninst > 1) near_wall_prefix_last(B);
        return;
    }
    // Pellet sums (the instance's pellets, in parallel), reduced across the
    // workgroup into ssum; skipped without pellets (a uniform branch).
    __shared__ double ssum[4][kPelSums];
    __shared__ double svm[4];
    __shared__ int spbad[4];
    if (ip.npel > 0) {
        double pel[kPelSums];
        for (int i = 0; i < kPelSums; ++i) pel[i] = 0.0;
        double vmax = 0.0;
        int pbad = 0;
        for (int k = tid; k < ip.npel; k += 256) {
            const int pi = ip.pel_off + k;
            const PelletTrial& t = B.ptrial[pi];
            const PelletAccum& a = B.pacc[pi];
            pel[0] += B.pel[pi].mass;
            for (int c = 0; c < 3; ++c) {
                pel[1 + c] += t.wall_impulse[c];
                pel[4 + c] += pel_sum(a, c);
                pel[7 + c] += pel_sum(a, 3 + c);
            }
            pel[10] += t.wall_work;
            pel[11] += pel_sum(a, 6);
            pel[12] += t.energy;
            pbad |= t.bad;
            vmax = fmax(vmax, sqrt(t.vel[0] * t.vel[0] + t.vel[1] * t.vel[1] + t.vel[2] * t.vel[2]));
        }
#pragma unroll
        for (int i = 0; i < kPelSums; ++i) pel[i] = wave_sum_last(pel[i]);
        pbad = wave_sum_i(pbad != 0 ? 1 : 0);
        vmax = wave_max_d(vmax);
        if (lane == 63)
            for (int i = 0; i < kPelSums; ++i) ssum[wave][i] = pel[i];
        if (lane == 0) {
            svm[wave] = vmax;
            spbad[wave] = pbad;
        }
        __syncthreads();
    }
    const int g0 = ip.boff / kGridWgBlocks, g1 = (ip.boff + ip.nblk) / kGridWgBlocks;
    if (B.audit) {
        // Audit stage sums of this instance, in a fixed order.
        for (int i = 0; i < kAuditVals; ++i) {
            double a = 0.0;
            for (int q = g0 + tid; q < g1; q += 256) a += B.apart[(size_t)i * B.ngwg + q];
            a = block_total(a, sd);
            if (tid == 0) B.audit[(size_t)inst * kAuditOut + i] = a;
        }
    }
    if (wave == 0) {
#pragma unroll
        for (int v = 0; v < kGridVals; ++v) acc[v] = wave_sum_last(acc[v]);
        bad = wave_sum_i_last(bad);
        float mx[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) mx[k] = wave_fmax_uniform(__uint_as_float(sl.max_bits[k]));
        const unsigned long long fl = wave_min_u64_last(sl.fail);
        const unsigned long long l0 = wave_sum_u64_last(sl.diss[0]), l1 = wave_sum_u64_last(sl.diss[1]),
                                 lh = wave_sum_u64_last(sl.diss[2]);
        const int esc = wave_sum_i_last(sl.escaped);
        // Cleared for the next trial (the values are in registers).
#pragma unroll
        for (int v = 0; v < kGridVals; ++v) B.gchunk[((size_t)inst * kGridVals + v) * kGridChunks + lane] = 0.0;
        B.gchunk_bad[inst * kGridChunks + lane] = 0;
        if (lane < kItemSlots) B.islot[inst * kItemSlots + lane] = ItemSlot{{0u, 0u, 0u, 0u}, ~0ull, {0ull, 0ull, 0ull}, 0, {0, 0, 0}};
        if (lane == 63) {
            const double diss = fix_value(fix_from_limbs(l0, l1, lh));
            if (B.audit) B.audit[(size_t)inst * kAuditOut + kAuditVals] = diss;
            double pel[kPelSums];
            double pv = 0.0;
            int pb = 0;
            for (int i = 0; i < kPelSums; ++i) pel[i] = 0.0;
            if (ip.npel > 0) {
                for (int i = 0; i < kPelSums; ++i) pel[i] = ((ssum[0][i] + ssum[1][i]) + ssum[2][i]) + ssum[3][i];
                pv = fmax(fmax(svm[0], svm[1]), fmax(svm[2], svm[3]));
                pb = spbad[0] + spbad[1] + spbad[2] + spbad[3];
            }
            StepStatus& st = s.st;
            for (int a = 0; a < 3; ++a) {
                st.wall_impulse[a] = acc[a];
                st.traction_impulse[a] = acc[3 + a];
            }
            st.projection = acc[6];
            st.friction = acc[7];
            st.traction_energy = acc[8];
            st.dissipation = diss;
            st.max_speed = mx[0];
            st.max_disp = mx[1];
            st.bound[0] = mx[2];
            st.bound[1] = mx[3];
            st.fail = fl;
            st.grid_nonfinite = bad;
            st.escaped = esc;
            st.pellet_bad = pb != 0;
            commit_instance(s, ip, pel, pv);
            // Without pellets the next trial's control follows at once, in the
            // same thread (a batch of one is its own near-wall prefix).
            if (B.fuse_control && ip.npel == 0) {
                const int nwc = control_instance(s, ip);
                if (B.ninst == 1) {
                    s.nw_base = nwc;
                    B.counters[1] = nwc;
                }
            }
        }
    }
    __syncthreads();
    if (ip.npel > 0) {
        // The instance's pellets: commit an accepted trial, then the next
        // trial's geometry and cleared sums, and the union of the pellets' boxes.
        const bool accepted = s.last_failure == 0;
        double box[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
        for (int k = tid; k < ip.npel; k += 256) {
            const int pi = ip.pel_off + k;
            PelletState& p = B.pel[pi];
            double pos[3], vel[3], L[3], q[4];
            if (accepted) {
                const PelletTrial& t = B.ptrial[pi];
                for (int c = 0; c < 3; ++c) {
                    pos[c] = t.pos[c];
                    vel[c] = t.vel[c];
                    L[c] = t.L[c];
                }
                for (int c = 0; c < 4; ++c) q[c] = t.q[c];
                for (int c = 0; c < 3; ++c) {
                    p.pos[c] = pos[c];
                    p.vel[c] = vel[c];
                    p.L[c] = L[c];
                }
                for (int c = 0; c < 4; ++c) p.q[c] = q[c];
            } else {
                for (int c = 0; c < 3; ++c) {
                    pos[c] = p.pos[c];
                    vel[c] = p.vel[c];
                    L[c] = p.L[c];
                }
                for (int c = 0; c < 4; ++c) q[c] = p.q[c];
            }
            const PelletGeo g = pellet_geometry(B, pi, p, pos, vel, L, q);
            for (int a = 0; a < 3; ++a) {
                box[a] = fmin(box[a], g.lo[a]);
                box[3 + a] = fmax(box[3 + a], g.hi[a]);
            }
        }
        __shared__ double sbox[4][6];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            box[a] = -wave_max_d(-box[a]);
            box[3 + a] = wave_max_d(box[3 + a]);
        }
        if (lane == 0)
            for (int a = 0; a < 6; ++a) sbox[wave][a] = box[a];
        __syncthreads();
        if (tid < 3) {
            s.pel_box[tid] = fmin(fmin(sbox[0][tid], sbox[1][tid]), fmin(sbox[2][tid], sbox[3][tid]));
            s.pel_box[3 + tid] =
                fmax(fmax(sbox[0][3 + tid], sbox[1][3 + tid]), fmax(sbox[2][3 + tid], sbox[3][3 + tid]));
        }
        __syncthreads();
    }
    // The next trial of this instance (what the control kernel would choose),
    // here after the pellets' commit.
    if (ip.npel > 0) {
        if (B.fuse_control && tid == 0) {
            const int nwc = control_instance(s, ip);
            if (B.ninst == 1) {
                s.nw_base = nwc;
                B.counters[1] = nwc;
            }
        }
        __syncthreads();
    }
    {
        const unsigned* src = reinterpret_cast<const unsigned*>(&ls);
        unsigned* dst = reinterpret_cast<unsigned*>(&gs);
        for (int w = tid; w < kWords; w += 256) dst[w] = src[w];
    }
    if (B.fuse_control && B.ninst > 1) near_wall_prefix_last(B);
}

// Order keys at the 32-step cadence, for the instances whose accepted step hit
// it and that took no trial step since.
__global__ __launch_bounds__(kBlock) void refkey_kernel(BatchDev B) {
    const int4 item = B.work[blockIdx.x];
    const InstState& s = B.is[item.x];
    if (!s.refsort) return;
    const InstParams& ip = B.ip[item.x];
    const float* slab = B.slab[s.cur];
    const size_t cap = B.cap;
    for (int p = item.y + threadIdx.x; p < item.y + item.z; p += kBlock)
        B.refkey[p] = (ref_cell(ip, slab[PX * cap + p], slab[PY * cap + p], slab[PZ * cap + p]) << 32) |
                      (uint64_t)B.ids[p];
}

// ------------------------------------------------------------ set-up

// Set-up of every instance's committed slab after a state upload or a sort,
// one workgroup per work item: the near-wall list, the item's set-up record
// (ItemPre), and the largest deposit bounds and speed of the particles inside
// the padded grid, with the same arithmetic as the trial step that writes a
// state (particle_bounds, particle_speed), so an uploaded state is stepped
// exactly as the same state reached by stepping.  The maxima go to the
// per-item partials (speed, -, b1, b2), which setup_reduce_kernel reduces per
// instance; the caller clears the list counters first.
__global__ __launch_bounds__(kBlock) void setup_kernel(BatchDev B) {
    __shared__ int sbox[4][6];
    const int4 item = B.work[blockIdx.x];
    InstState& s = B.is[item.x];
    const InstParams& ip = B.ip[item.x];
    const Params P = ip.P;
    const float* slab = B.slab[s.cur];
    const float* shared = B.shared[B.sh];
    const size_t cap = B.cap;
    int* list = B.nw_list[s.cur] + ip.poff;
    int* count = &s.nw_count[s.cur];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
    float b1m = 0.0f, b2m = 0.0f, spd = 0.0f;
    // Near-wall flags of the thread's particles (bit r: round r), appended once
    // per workgroup below: one atomic on the instance's counter per work item.
    unsigned nwbits = 0u;
    constexpr int R = kChunk / kBlock;
    for (int base = item.y, r = 0; base < item.y + item.z; base += kBlock, ++r) {
        const int p = base + threadIdx.x;
        bool nw = false;
        if (p < item.y + item.z) {
            const float x = slab[PX * cap + p], y = slab[PY * cap + p], z = slab[PZ * cap + p];
            nw = near_wall(P, x, y, z);
            Stencil st;
            if (make_stencil(P, x, y, z, st)) {
                for (int a = 0; a < 3; ++a) {
                    lo[a] = min(lo[a], st.base[a]);
                    hi[a] = max(hi[a], st.base[a]);
                }
                float v[3], C[9], tau[6];
                for (int a = 0; a < 3; ++a) v[a] = slab[(VX + a) * cap + p];
                for (int i = 0; i < 9; ++i) C[i] = slab[(C00 + i) * cap + p];
                for (int i = 0; i < 6; ++i) tau[i] = slab[(TXX + i) * cap + p];
                float b1, b2;
                particle_bounds(P, shared[SMASS * cap + p], shared[SVOL0 * cap + p], v, C, tau, b1, b2);
                b1m = fmaxf(b1m, b1);
                b2m = fmaxf(b2m, b2);
                spd = fmaxf(spd, particle_speed(v));
            }
        }
        nwbits |= nw ? (1u << r) : 0u;
    }
    block_append_rounds<R>(nwbits, item.y, list, count);
    for (int a = 0; a < 3; ++a) {
        lo[a] = wave_min_uniform(lo[a]);
        hi[a] = wave_max_uniform(hi[a]);
    }
    b1m = wave_fmax_uniform(b1m);
    b2m = wave_fmax_uniform(b2m);
    spd = wave_fmax_uniform(spd);
    __shared__ float smx[4][3];
    if (lane == 0) {
        for (int a = 0; a < 3; ++a) {
            sbox[wave][a] = lo[a];
            sbox[wave][3 + a] = hi[a];
        }
        smx[wave][0] = spd;
        smx[wave][1] = b1m;
        smx[wave][2] = b2m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        ItemPre r;
        for (int a = 0; a < 3; ++a) {
            r.lo[a] = min(min(sbox[0][a], sbox[1][a]), min(sbox[2][a], sbox[3][a]));
            r.hi[a] = max(max(sbox[0][3 + a], sbox[1][3 + a]), max(sbox[2][3 + a], sbox[3][3 + a]));
        }
        B.wpre[s.cur][blockIdx.x] = r;
        float mx[3];
        for (int k = 0; k < 3; ++k) mx[k] = fmaxf(fmaxf(smx[0][k], smx[1][k]), fmaxf(smx[2][k], smx[3][k]));
        reinterpret_cast<float4*>(B.pmax)[blockIdx.x] = make_float4(mx[0], 0.0f, mx[1], mx[2]);
    }
}

// The set-up maxima of every instance from its items' partials: one
// workgroup per instance.
__global__ __launch_bounds__(256) void setup_reduce_kernel(BatchDev B) {
    InstState& s = B.is[blockIdx.x];
    float m[3] = {0.0f, 0.0f, 0.0f};
    for (int q = s.w_off + threadIdx.x; q < s.w_off + s.w_cnt; q += 256) {
        const float4 v = reinterpret_cast<const float4*>(B.pmax)[q];
        m[0] = fmaxf(m[0], v.x);
        m[1] = fmaxf(m[1], v.z);
        m[2] = fmaxf(m[2], v.w);
    }
    __shared__ float sm[4][3];
    for (int k = 0; k < 3; ++k) m[k] = wave_max(m[k]);
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 3; ++k) sm[threadIdx.x >> 6][k] = m[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < 3; ++k) m[k] = fmaxf(fmaxf(sm[0][k], sm[1][k]), fmaxf(sm[2][k], sm[3][k]));
        // The committed speed as the trial steps compute it (particle_speed in
        // f32), so that a state set from outside takes the same dt as the same
        // state reached by stepping.
        s.setup_speed = m[0];
        s.max_speed = (double)m[0];
        s.bound[0] = m[1];
        s.bound[1] = m[2];
    }
}

// Device order: 4x4x4-cell blocks (compact workgroup tiles), instances in
// range order; entries past an instance's particle count go to the end of its
// range.  Radix sort is stable, so equal keys keep their order.  K: 32-bit keys
// when (instance, block, cell) fits in 32 bits (half the sort's traffic), else 64.
template <typename K>
__global__ void sort_keys_kernel(BatchDev B, int cell_bits, K* keys, uint32_t* order) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (int)B.cap) return;
    const int inst = instance_of(B, p);
    const InstState& s = B.is[inst];
    const Params& P = B.ip[inst].P;
    uint64_t key = ((uint64_t)1 << cell_bits) - 1;
    if (p - B.poffs[inst] < s.n) {
        const float* slab = B.slab[s.cur];
        uint64_t c[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float sc = floorf(slab[(PX + a) * B.cap + p] / P.dx);
            sc = isnan(sc) ? 0.0f : fmaxf(sc, 0.0f);
            c[a] = sc >= (float)P.cells[a] ? (uint64_t)P.cells[a] : (uint64_t)sc;
        }
        uint64_t nb1 = (uint64_t)(P.cells[1] + 4) / 4, nb2 = (uint64_t)(P.cells[2] + 4) / 4;
        uint64_t blk = ((c[0] >> 2) * nb1 + (c[1] >> 2)) * nb2 + (c[2] >> 2);
        uint64_t within = ((c[0] & 3) << 4) | ((c[1] & 3) << 2) | (c[2] & 3);
        key = blk * 64 + within;
    }
    keys[p] = (K)(((uint64_t)inst << cell_bits) | key);
    order[p] = (uint32_t)p;
}

// (instance, cell block) of a sorted key: the run-length encoding's input.
struct BlockOfKey {
    template <typename K>
    __host__ __device__ uint64_t operator()(K k) const {
        return (uint64_t)(k >> 6);
    }
};

// Gather every slab entry into sorted order: from each instance's committed
// slab into its other slab (the host then flips the instance's selector), and
// the shared slab pair, ids and order keys likewise.
__global__ void permute_kernel(BatchDev B, const uint32_t* __restrict__ order, const float* __restrict__ ssrc,
                               float* __restrict__ sdst, const uint32_t* __restrict__ isrc,
                               uint32_t* __restrict__ idst, const uint64_t* __restrict__ rsrc,
                               uint64_t* __restrict__ rdst) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (int)B.cap) return;
    const size_t cap = B.cap;
    const int cur = B.is[instance_of(B, k)].cur;
    const float* src = B.slab[cur];
    float* dst = B.slab[cur ^ 1];
    const uint32_t o = order[k];
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

inline int blocks_for(size_t n) { return (int)((n + kBlock - 1) / kBlock); }

}  // namespace

int g2p_chunk() { return kChunk; }

void launch_step(hipStream_t s, const BatchDev& B, int traction_wgs, bool pellets, bool control) {
    if (control) control_kernel<<<1, 256, 0, s>>>(B);
    if (B.nwork > 0) {
        if (B.apart)
            p2g_kernel<kAuditFields, 256><<<B.nwork, 256, 0, s>>>(B);
        else if (B.p2g_block == 512)
            p2g_kernel<kDepositFields, 512><<<B.nwork, 512, 0, s>>>(B);
        else
            p2g_kernel<kDepositFields, 256><<<B.nwork, 256, 0, s>>>(B);
    }
    traction_kernel<<<traction_wgs, kBlock, 0, s>>>(B);
    if (B.apart)
        grid_update_kernel<true><<<B.ngwg, 256, 0, s>>>(B);
    else
        grid_update_kernel<false><<<B.ngwg, 256, 0, s>>>(B);
    if (pellets) couple_kernel<<<B.ngwg * kGridWgBlocks, 64, 0, s>>>(B);
    if (B.nwork > 0) {
        if (B.mlp_fwd && B.rlog)
            g2p_kernel<true, 2, true><<<B.nwork, kBlock, 0, s>>>(B);
        else if (B.mlp_fwd)
            g2p_kernel<true, 2, false><<<B.nwork, kBlock, 0, s>>>(B);
        else if (B.rlog)
            g2p_kernel<false, 2, true><<<B.nwork, kBlock, 0, s>>>(B);
        else if (B.g2p_occupancy >= 3)
            g2p_kernel<false, 3, false><<<B.nwork, kBlock, 0, s>>>(B);
        else
            g2p_kernel<false, 2, false><<<B.nwork, kBlock, 0, s>>>(B);
    }
    if (pellets) pellet_kernel<<<(B.npel + 63) / 64, 64, 0, s>>>(B);
    finalize_kernel<<<B.ninst, 256, 0, s>>>(B);
}

__global__ void clock_kernel(unsigned long long* out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *out = (unsigned long long)__builtin_amdgcn_s_memrealtime();
}

void launch_clock(hipStream_t s, unsigned long long* out) { clock_kernel<<<1, 64, 0, s>>>(out); }

void launch_refkeys(hipStream_t s, const BatchDev& B) {
    if (B.nwork > 0) refkey_kernel<<<B.nwork, kBlock, 0, s>>>(B);
}

void launch_setup(hipStream_t s, const BatchDev& B) {
    if (B.nwork > 0) setup_kernel<<<B.nwork, kBlock, 0, s>>>(B);
    setup_reduce_kernel<<<B.ninst, 256, 0, s>>>(B);
}

void launch_sort_keys(hipStream_t s, const BatchDev& B, int cell_bits, int end_bit, uint64_t* keys, uint32_t* order) {
    if (end_bit <= 32)
        sort_keys_kernel<uint32_t>
            <<<blocks_for(B.cap), kBlock, 0, s>>>(B, cell_bits, reinterpret_cast<uint32_t*>(keys), order);
    else
        sort_keys_kernel<uint64_t><<<blocks_for(B.cap), kBlock, 0, s>>>(B, cell_bits, keys, order);
}

void launch_permute(hipStream_t s, const BatchDev& B, const uint32_t* order, const float* shared_src,
                    float* shared_dst, const uint32_t* id_src, uint32_t* id_dst, const uint64_t* ref_src,
                    uint64_t* ref_dst) {
    permute_kernel<<<blocks_for(B.cap), kBlock, 0, s>>>(B, order, shared_src, shared_dst, id_src, id_dst, ref_src,
                                                        ref_dst);
}

// rocPRIM's radix sort picks a merge sort at up to 2^20 keys, whatever the key's
// width: one pass per doubling of the sorted runs, 20 launches at 1M particles.
// A merge-sort limit of 0 takes the onesweep radix sort instead, one pass per
// 8 bits of the sorted key range.  (Its gfx950 tiles of 6144 keys give a 1M-key
// pass 163 workgroups; tiles of 2048 keys and 7-bit digits measured 41 us per
// pass against 28 us.)
using OnesweepOnly =
    rocprim::radix_sort_config<rocprim::default_config, rocprim::default_config, rocprim::default_config, 0>;

template <typename K>
hipError_t radix_sort_as(hipStream_t s, void* temp, size_t& bytes, const uint64_t* keys_in, uint64_t* keys_out,
                         const uint32_t* vals_in, uint32_t* vals_out, int n, int begin_bit, int end_bit) {
    return rocprim::radix_sort_pairs<OnesweepOnly>(temp, bytes, reinterpret_cast<const K*>(keys_in),
                                                   reinterpret_cast<K*>(keys_out), vals_in, vals_out, n,
                                                   (unsigned)begin_bit, (unsigned)end_bit, s);
}

template <typename K>
hipError_t block_runs_as(hipStream_t s, void* temp, size_t& bytes, int n, const uint64_t* sorted_keys,
                         uint64_t* unique, uint32_t* counts, int* nruns) {
    auto blocks = rocprim::make_transform_iterator(reinterpret_cast<const K*>(sorted_keys), BlockOfKey());
    return rocprim::run_length_encode(temp, bytes, blocks, (unsigned)n, unique, counts, nruns, s);
}

hipError_t sort_temp_bytes(int n, int end_bit, size_t* bytes) {
    *bytes = 0;
    if (end_bit <= 32)
        return radix_sort_as<uint32_t>(nullptr, nullptr, *bytes, nullptr, nullptr, nullptr, nullptr, n, 0, end_bit);
    return radix_sort_as<uint64_t>(nullptr, nullptr, *bytes, nullptr, nullptr, nullptr, nullptr, n, 0, end_bit);
}

hipError_t rle_temp_bytes(int n, int end_bit, size_t* bytes) {
    *bytes = 0;
    if (end_bit <= 32) return block_runs_as<uint32_t>(nullptr, nullptr, *bytes, n, nullptr, nullptr, nullptr, nullptr);
    return block_runs_as<uint64_t>(nullptr, nullptr, *bytes, n, nullptr, nullptr, nullptr, nullptr);
}

hipError_t launch_radix_sort(hipStream_t s, void* temp, size_t temp_bytes, const uint64_t* keys_in,
                             uint64_t* keys_out, const uint32_t* vals_in, uint32_t* vals_out, int n, int begin_bit,
                             int end_bit) {
    if (end_bit <= 32)
        return radix_sort_as<uint32_t>(s, temp, temp_bytes, keys_in, keys_out, vals_in, vals_out, n, begin_bit,
                                       end_bit);
    return radix_sort_as<uint64_t>(s, temp, temp_bytes, keys_in, keys_out, vals_in, vals_out, n, begin_bit, end_bit);
}

hipError_t launch_block_runs(hipStream_t s, void* temp, size_t temp_bytes, int n, int end_bit,
                             const uint64_t* sorted_keys, uint64_t* unique, uint32_t* counts, int* nruns) {
    if (n <= 0) return hipSuccess;
    if (end_bit <= 32) return block_runs_as<uint32_t>(s, temp, temp_bytes, n, sorted_keys, unique, counts, nruns);
    return block_runs_as<uint64_t>(s, temp, temp_bytes, n, sorted_keys, unique, counts, nruns);
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
