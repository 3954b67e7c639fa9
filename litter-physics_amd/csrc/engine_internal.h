// Internal types shared by the MLS-MPM HIP kernels and the host engine.
//
// The engine steps a batch of independent instances (one instance is the
// reference's Simulation); every kernel covers all instances of the batch, and
// each instance carries its own adaptive dt, accept/reject state and ledgers on
// the device, so one launch sequence advances every instance by one trial step.
//
// Device layout (HBM):
//  * particles: two f32 slabs of NPF fields x total capacity (structure of
//    arrays, field f of particle p at slab[f * cap + p]); instance i owns the
//    range [poff_i, poff_i + pcap_i) of both.  Per instance, one slab holds the
//    committed state and the other receives the trial step; acceptance flips
//    the instance's selector.  The deformation gradient is held as H = F - I and
//    the liquid volume ratio as J - 1 so that small elastic strains keep f32
//    precision; Kirchhoff stress is symmetric (six components).  Mass, reference
//    volume, a liquid's plastic strain, id and the reference-order key never
//    change within a step and live in a separate slab pair swapped by the sort.
//  * grid: each instance's dense padded node array (grid.rs:12-21) re-laid in
//    4x4x4-node blocks of 64 consecutive nodes (one wavefront per block); the
//    instances' blocks are concatenated.  The particle-to-grid deposit of a
//    trial step (solver.rs:632-661) runs once the step's dt is known and sums
//    mass and momentum w (m v + (m C - dt (4/dx^2) V0 tau) d) as exact 64-bit
//    integers in units of a per-instance power of two (DepositScale), so the
//    sums do not depend on the order of the atomics or on how the particles
//    are grouped into workgroups.  Two accumulator sets alternate from trial
//    to trial: the grid update reads one and clears the other.  The energy
//    audit deposits seven fields instead: mass, the APIC momentum
//    w (m v + m C d) and the stress sum w V0 tau d.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

struct mpm_mlp;

namespace mpm {

enum PField {
    PX, PY, PZ,
    VX, VY, VZ,
    C00, C01, C02, C10, C11, C12, C20, C21, C22,
    H00, H01, H02, H10, H11, H12, H20, H21, H22,
    JM1,
    TXX, TYY, TZZ, TXY, TXZ, TYZ,
    EPS,
    NPF
};

enum SField { SMASS, SVOL0, SEPS, NSF };  // SEPS: plastic strain of a liquid (never changes)

enum Fail {
    F_OK = 0,
    F_INVERTED = 1,
    F_EIGEN = 2,
    F_STRETCH = 3,
    F_RETURN = 4,
    F_NONFINITE = 5,
    F_GRID = 6,
    F_DISPLACEMENT = 7,
    F_PELLET = 8,
    F_TRACTION = 9
};

// Instance modes.
enum Mode { M_RUN = 0, M_DONE = 1, M_PAUSE = 2, M_ERROR = 3 };
// Pause reasons (bit flags).
enum { P_ESCAPED = 1, P_BUDGET = 2 };
// Error kinds of an instance in M_ERROR.
enum { E_MIN_STEP = 1, E_MIN_AFTER_REJECT = 2 };

constexpr int kMaxSpheres = 64;
// Accumulator fields of the deposit: mass and momentum, or (audit) mass, APIC
// momentum and the stress sum.
constexpr int kDepositFields = 4, kAuditFields = 7;
constexpr int kLedger = 31;  // doubles of mpm_ledger
constexpr int kGridVals = 9;
// Node blocks per grid-update workgroup; instance block ranges are multiples of it.
constexpr int kGridWgBlocks = 4;
// Kernel timing stamps: partial minima and maxima per timed quantity.
constexpr int kStampSlots = 16;
// Energy audit (audit.rs): grid kinetic energies at five stages of one trial
// (transport, after stress, after traction, after gravity, after walls), and
// per instance those five plus the trial's plastic-work sum.
constexpr int kAuditVals = 5, kAuditOut = 6;

// Launch-constant parameters of one instance (f32 copies of the f64 host values).
struct Params {
    float dx, inv_dx;
    int cells[3];
    int nodes[3];
    int nb[3];  // node blocks per axis
    float extent[3];
    float gravity[3];
    int kind;  // 0 paste, 1 liquid
    float G, lam, ys, K, inv_n, bulk, gamma, visc;
    float face_mu[6];
};

// Constant per instance (uploaded when the instance is configured).
struct InstParams {
    Params P;
    double dx, cfl, max_dt, min_dt, wave, diff;
    double gravity[3];
    double wall_friction, contact_k, contact_damping, contact_friction;
    double particle_mass;   // sum of the stored particle masses (f64)
    float mass_max;         // largest stored particle mass (f32, as the kernels hold it)
    double pel_contact_dt;  // min over the instance's pellets of 0.2 sqrt(m / k)
    int poff, pcap;         // particle range of the slabs
    int boff, nblk;         // node-block range (nblk a multiple of kGridWgBlocks)
    int pel_off, npel;      // pellet range
};

// Accumulators of one trial step.
struct StepStatus {
    double wall_impulse[3];      // sum over nodes of (v' - v) m, walls and budget
    double traction_impulse[3];  // sum of signed traction deposits
    double projection;           // wall_normal_projection_energy increment
    double friction;             // wall_friction_dissipation increment
    double traction_energy;      // wall_normal_traction_energy increment
    double dissipation;          // plastic dissipation increment
    float max_speed;
    float max_disp;
    float bound[2];              // deposit bounds of the trial state (particle_bounds)
    unsigned long long fail;     // (particle << 8) | code of the first failure
    int grid_nonfinite;
    int escaped;
    int pellet_bad;
    int pad;
};

// Device-resident state of one instance (solver.rs:324-352 plus the driver's loop state).
struct InstState {
    double ledger[kLedger];
    double time, target, dt, limit, dt_scale, max_speed, min_dt_used, max_dt_used;
    double trial_dt;  // > 0: one bare trial step of this size (try_step), no time advance
    double pel_vmax;  // largest pellet speed of the committed state
    // Deposit scales of the running trial (powers of two; the integer sums hold
    // value x scale): mass, momentum with the stress impulse, and for the audit
    // the APIC momentum and the stress sum; wall-traction reactions.
    float sc_m, sc_p, sc_mom, sc_f, sc_t;
    // Largest particle_bounds of the committed state (set from the accepted
    // trial, or by the set-up kernel after a state upload or sort) and largest
    // particle speed found by the set-up kernel.
    float bound[2];
    float setup_speed;
    unsigned long long step_count, limited_steps, rejected_steps;
    StepStatus st;
    int n;            // particle count
    int cur, gcur;    // slab / accumulator-set selectors
    int nw_count[2];  // near-wall list length per slab
    int mode, pause, err, last_failure;
    int active, limited;
    int budget;       // accepted steps left before a pause (host-set)
    int w_off, w_cnt; // work items
    int nw_base;      // inclusive end of the instance's range in the near-wall prefix (traction)
    unsigned long long commit_clock;  // device constant clock at the last accepted step
    int refsort;      // the last accepted step hit the 32-step cadence; its order keys are pending
    int committed;    // a trial step was accepted since the state was set
    int geo_stale;    // the host changed the state: pellet geometry and coupling sums need recomputing
    double pel_box[6];  // union of the pellets' boxes at the start of the trial (lo, hi)
};

// Exact fixed-point sum of f64 values in units of 2^-kFixBits, as a two's
// complement 128-bit integer in two words: (hi, lo) with lo carrying into hi.
// Integer sums do not depend on the order of the additions, so atomics and
// partial sums give the same total whatever the scheduling.
constexpr int kFixBits = 100;
struct Fix128 {
    unsigned long long lo;
    long long hi;
};

// G2P's step reductions of one instance, reached by order-independent integer
// atomics: work item w adds into slot w % kItemSlots of its instance, so no
// address takes every workgroup's atomic.  The maxima are of non-negative
// floats, as their bit patterns; the dissipation is a Fix128 sum kept in three
// 64-bit limbs (bits 0-31 and 32-63 of the low word, the high word modulo
// 2^64), so no carry crosses a word below 2^32 additions.  The finalize kernel
// reads an instance's slots and leaves them at the identity (0, except fail:
// all ones) for the next trial.
constexpr int kItemSlots = 16;
struct ItemSlot {
    unsigned max_bits[4];     // speed, displacement, the two deposit bounds
    unsigned long long fail;  // (particle << 8) | code of the first failure
    unsigned long long diss[3];
    int escaped;
    int pad[3];
};
// Grid partials per instance after G2P's first-level reduction: the first
// kGridChunks work items of an instance each sum one chunk of its grid-update
// workgroups' partials (unused chunks stay zero).
constexpr int kGridChunks = 64;

// Per-pellet coupling accumulators (rigid.rs:350-363): impulse (3), angular
// impulse (3) and grid energy, exact fixed point; constrained node count.
struct PelletAccum {
    // Fix128 sums in three 64-bit limbs each (bits 0-31 and 32-63 of the low
    // word, the high word modulo 2^64, as ItemSlot::diss): added with
    // non-returning atomics, so a block need not wait for a carry's round trip
    // before its next pellet.
    unsigned long long sum[7][3];
    unsigned long long constrained;
};

// Pellet state (rigid.rs:30-49), f64; orientation quaternion (i, j, k, w).
struct PelletState {
    double radius, length, mass;
    double inertia[3];
    double pos[3], q[4], vel[3], L[3];
    int inst, spheres;
    double offsets[kMaxSpheres];
    double half;  // largest |offset|: the pellet's half-extent along its axis, less the radius
};

// Start-of-step geometry of a pellet for the coupling kernel.
struct PelletGeo {
    double pos[3], axis[3], vel[3], omega[3];
    double lo[3], hi[3];
    double radius;
    int spheres, inst;
};

// Trial pellet after coupling and integration (committed on acceptance).
struct PelletTrial {
    double pos[3], q[4], vel[3], L[3];
    double energy;  // kinetic-energy change from the coupling impulse
    double wall_impulse[3];
    double wall_work;
    int bad;
    int pad;
};

// One accumulator set: field f of node n at v[f * na + n] (0 mass, 1-3 momentum,
// 4-6 the audit's stress sum).
struct GridPtrs {
    long long* v;
    int* flag;
};

// Set-up record of one work item for one slab, written by the step that
// produced the slab (or by setup_kernel after a state upload or sort): the box
// of the stencil bases of the item's particles.  The deposit and the gather of
// the next trial lay out their LDS tiles from it without a pass over the
// particles first; a particle outside the box takes the untiled path.
struct ItemPre {
    int lo[3], hi[3];  // stencil-base box of the particles inside the padded grid (lo > hi: none)
};

// Device pointers of the whole batch.
struct BatchDev {
    int ninst;
    size_t cap;  // slab stride (total particle capacity)
    size_t na;   // total node count (all instances)
    InstParams* ip;
    InstState* is;
    const int* poffs;  // particle-range starts of the instances (ninst + 1 entries, the last = cap)
    float* slab[2];
    float* shared[2];
    int sh;       // shared-slab selector (host controlled)
    int g2p_occupancy;  // G2P workgroups resident per CU (2 or 3: its register budget)
    int p2g_block;      // threads of a deposit workgroup (256 or 512)
    int fuse_control;   // finalize also chooses the next trial (control_instance) for every instance
    // Kernel timing (null: off): the deposit's and the G2P's earliest workgroup
    // start and latest workgroup end on the device's constant clock, as
    // kStampSlots partial minima or maxima each (workgroup w stamps slot
    // w % kStampSlots, so no address takes every workgroup's atomic):
    // {p2g start, p2g end, g2p start, g2p end} x kStampSlots, set up as
    // {max, 0, max, 0}.
    unsigned long long* ktime;
    long long* gmem[2];  // kAuditFields * na integer sums each
    int* gflag[2];       // one per node block
    float* gv;           // 3 * na
    long long* trS;      // 3 * na, units of the traction scale
    long long* trT;      // 3 * na, units of the traction scale
    uint32_t* ids;     // current shared-slab ids and reference-order keys
    uint64_t* refkey;
    int* nw_list[2];  // per slab, instance ranges like the particles
    int* blk_inst;    // node block -> instance
    int4* work;       // (instance, start, count, 0)
    ItemPre* wpre[2];  // per slab, one per work item
    int nwork;        // number of work items
    int ngwg;         // grid-update workgroups (all instances)
    int* counters;    // [0] finalize workgroups done (the last resets it), [1] near-wall entries of the running instances
    double* gpart;    // grid partials, field-major: gpart[v * ngwg + workgroup]
    double* apart;    // audit partials, field-major like gpart (NULL: no audit)
    double* audit;    // audit sums per instance, kAuditOut each (NULL: no audit)
    int* gbad;
    // Chunk sums of gpart, field-major per instance:
    // gchunk[(inst * kGridVals + v) * kGridChunks + chunk], and of gbad.
    double* gchunk;
    int* gchunk_bad;  // [inst * kGridChunks + chunk]
    ItemSlot* islot;  // [inst * kItemSlots + slot]
    float* pmax;      // set-up partials per work item: speed, -, the two deposit bounds
    PelletState* pel;
    PelletGeo* pgeo;
    PelletAccum* pacc;
    PelletTrial* ptrial;
    int npel;
    // Return-map surrogate (null: exact return only): packed bf16 forward
    // fragments and biases, acceptance tolerance on the residual, Newton
    // corrections after the prediction, counters (considered, accepted).
    const uint16_t* mlp_fwd;
    const float* mlp_bias;
    float mlp_tol;
    int mlp_polish;
    unsigned long long* mlp_stats;
    // Log of exact returns (null: off): kLogStride floats per yielded particle.
    float* rlog;
    int* rlog_count;
    int rlog_cap;
};

// Logged exact return: 12 surrogate inputs, 7 targets, then q_tr, sigma_y,
// 3 G dt, K, n and the returned q.
constexpr int kLogStride = 25;

__host__ __device__ inline GridPtrs grid_set(const BatchDev& B, int set) {
    GridPtrs g;
    g.v = B.gmem[set];
    g.flag = B.gflag[set];
    return g;
}

__host__ __device__ inline int node_index(const Params& P, int i, int j, int k) {
    int b = ((i >> 2) * P.nb[1] + (j >> 2)) * P.nb[2] + (k >> 2);
    return (b << 6) | ((i & 3) << 4) | ((j & 3) << 2) | (k & 3);
}

__host__ __device__ inline void node_coords(const Params& P, int index, int& i, int& j, int& k) {
    int b = index >> 6, l = index & 63;
    int bk = b % P.nb[2];
    int bj = (b / P.nb[2]) % P.nb[1];
    int bi = b / (P.nb[2] * P.nb[1]);
    i = bi * 4 + (l >> 4);
    j = bj * 4 + ((l >> 2) & 3);
    k = bk * 4 + (l & 3);
}

// Device buffers of a network (mlp.hip).
const uint16_t* mlp_device_fragments(const mpm_mlp* mlp);
const float* mlp_device_biases(const mpm_mlp* mlp);

// Host launchers (kernels.hip).
int g2p_chunk();
// One trial step of every running instance: control, deposit, traction, grid
// update, pellet coupling, G2P, pellet integration, reductions and commit.
// One trial step.  control: launch the control kernel first (the first step
// after the host touched the instance states); otherwise the previous step's
// finalize chose this trial (launched with B.fuse_control set).
void launch_step(hipStream_t s, const BatchDev& B, int traction_wgs, bool pellets, bool control);
// The device's constant clock into *out (the start of a timed run).
void launch_clock(hipStream_t s, unsigned long long* out);
// Order keys of the instances whose refresh is pending (the flags are left set).
void launch_refkeys(hipStream_t s, const BatchDev& B);
// Near-wall lists, set-up records, deposit bounds and largest speed of every
// instance's committed slab (list counters cleared by the caller).
void launch_setup(hipStream_t s, const BatchDev& B);
// Device-order keys of every slab entry: (instance << cell_bits) | 4^3-cell block
// key; entries past an instance's particle count sort to the end of its range.
// Keys of at most 32 bits (end_bit <= 32) are stored as 32-bit words in `keys`.
void launch_sort_keys(hipStream_t s, const BatchDev& B, int cell_bits, int end_bit, uint64_t* keys, uint32_t* order);
// Gather every slab entry into the sorted order (each instance from its
// committed slab into the other one; shared slabs, ids and order keys likewise).
void launch_permute(hipStream_t s, const BatchDev& B, const uint32_t* order, const float* shared_src,
                    float* shared_dst, const uint32_t* id_src, uint32_t* id_dst, const uint64_t* ref_src,
                    uint64_t* ref_dst);
// Temporary storage of the radix sort and of the run-length encoding (the
// size is written to *bytes); the sort and encode calls return rocPRIM's status.
// Keys are 32-bit words when end_bit <= 32, 64-bit otherwise.  The sort orders
// by bits [begin_bit, end_bit) of the key, stably.
hipError_t sort_temp_bytes(int n, int end_bit, size_t* bytes);
hipError_t rle_temp_bytes(int n, int end_bit, size_t* bytes);
hipError_t launch_radix_sort(hipStream_t s, void* temp, size_t temp_bytes, const uint64_t* keys_in,
                             uint64_t* keys_out, const uint32_t* vals_in, uint32_t* vals_out, int n, int begin_bit,
                             int end_bit);
// Runs of equal (instance, cell block) in the sorted keys.
hipError_t launch_block_runs(hipStream_t s, void* temp, size_t temp_bytes, int n, int end_bit,
                             const uint64_t* sorted_keys, uint64_t* unique, uint32_t* counts, int* nruns);
void launch_paste_batch(hipStream_t s, const Params& P, float dt, int n, const float* h_in, float* h_out,
                        float* tau, float* inc, float* dd, int* codes);
void launch_liquid_batch(hipStream_t s, const Params& P, float dt, int n, const float* jm1, const float* c,
                         float* jm1_out, float* tau, int* codes);

}  // namespace mpm
