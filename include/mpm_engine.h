/*
 * mpm_engine.h -- C ABI of the MI355X (gfx950) MLS-MPM substep engine.
 *
 * The engine replaces the inner stepping of the reference crate's
 * `Simulation` (core/src/mpm/solver.rs:324-611) and keeps the same state,
 * ledgers, step semantics and errors.  Particle state lives on the GPU in
 * f32 structure-of-arrays form between calls; host arrays are borrowed only
 * for the duration of a call and are f64, laid out as the reference stores
 * them (one array per field; vectors as 3 doubles, matrices as 9 doubles in
 * row-major order).
 *
 * Status codes: 0 = ok; > 0 = a rejected trial step, numbered like the
 * reference's StepFailure / ConstitutiveError variants (solver.rs:249-259,
 * constitutive.rs:41-66); < 0 = SolverError or API misuse (solver.rs:128-142).
 * No C++ exception crosses this ABI.  One engine per host thread; an engine
 * is not re-entrant.
 */
#ifndef MPM_ENGINE_H
#define MPM_ENGINE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPM_ABI_VERSION 1

enum {
    MPM_OK = 0,
    MPM_FAIL_INVERTED = 1,        /* ConstitutiveError::InvertedDeformation */
    MPM_FAIL_EIGEN = 2,           /* ConstitutiveError::EigenFailed */
    MPM_FAIL_STRETCH = 3,         /* ConstitutiveError::DegenerateStretch */
    MPM_FAIL_RETURN = 4,          /* ConstitutiveError::ReturnNotConverged */
    MPM_FAIL_NONFINITE = 5,       /* ConstitutiveError::NonFinite */
    MPM_FAIL_NONFINITE_GRID = 6,  /* StepFailure::NonFiniteGrid */
    MPM_FAIL_DISPLACEMENT = 7,    /* StepFailure::Displacement */
    MPM_FAIL_PELLET = 8,          /* StepFailure::PelletNonFinite */
    MPM_FAIL_TRACTION = 9,        /* StepFailure::NonFiniteTractionEnergy */
    MPM_ERR_MIN_STEP = -1,        /* SolverError::MinimumStepReached */
    MPM_ERR_INVALID = -2,         /* SolverError::InvalidState / bad argument */
    MPM_ERR_DEVICE = -3,          /* HIP runtime failure */
    MPM_ERR_CAPACITY = -4         /* more particles or pellets than allocated */
};

/* Material, constitutive.rs:111-131.  kind: 0 paste, 1 liquid.  tait_gamma is
 * an extension for the liquid: p = (K/g)(J^-g - 1); 1.0 is the reference's
 * linear law p = K(1/J - 1) (constitutive.rs:574). */
typedef struct {
    int kind;
    double density;
    double shear_modulus;
    double lame_lambda;
    double bulk_modulus;
    double yield_stress;
    double consistency;
    double flow_index;
    double tait_gamma;
} mpm_material;

/* SolverConfig, solver.rs:49-62.  face_friction extends the single wall
 * coefficient to one per face (x-, x+, y-, y+, z-, z+); +inf is a sticky
 * face, 0 a free-slip face; six copies of wall_friction is the reference. */
typedef struct {
    double gravity[3];
    double wall_friction;
    double cfl;
    double max_dt;
    double min_dt;
    double contact_stiffness;
    double contact_restitution;
    double contact_friction;
    double face_friction[6];
} mpm_config;

/* GridLayout::new arguments, grid.rs:37-86. */
typedef struct {
    double domain[3];
    double spacing;
    size_t max_nodes;
} mpm_layout;

/* ParticleSet, particles.rs:11-32 (SoA, f64, borrowed). */
typedef struct {
    size_t n;
    uint32_t* id;
    double* position;       /* n x 3 */
    double* velocity;       /* n x 3 */
    double* affine;         /* n x 9 */
    double* deformation;    /* n x 9 */
    double* volume_ratio;   /* n */
    double* kirchhoff;      /* n x 9 */
    double* plastic_strain; /* n */
    double* mass;           /* n */
    double* volume0;        /* n */
} mpm_particles;

/* Ledger, solver.rs:67-124 (field order as in the reference). */
typedef struct {
    double initial_mass;
    double outflow_mass;
    double initial_momentum[3];
    double initial_mechanical_energy;
    double gravity_impulse[3];
    double wall_impulse[3];
    double pellet_wall_impulse[3];
    double coupling_impulse[3];
    double coupling_angular_impulse[3];
    double wall_normal_projection_energy;
    double wall_friction_dissipation;
    double wall_normal_traction_energy;
    double pellet_wall_work;
    double coupling_grid_energy;
    double coupling_pellet_energy;
    double plastic_dissipation;
    double outflow_momentum[3];
} mpm_ledger;

/* Integration history, solver.rs:338-351 (min_dt_used = +inf before the first step). */
typedef struct {
    double time;
    uint64_t step_count;
    uint64_t limited_steps;
    uint64_t rejected_steps;
    double dt_scale;
    double min_dt_used;
    double max_dt_used;
} mpm_history;

/* Pellet, rigid.rs:30-49; orientation as the quaternion (i, j, k, w). */
typedef struct {
    double radius, length, mass;
    double inertia_body[3];
    double position[3];
    double orientation[4];
    double velocity[3];
    double angular_momentum[3];
    int sphere_count;
    double sphere_offsets[64];
} mpm_pellet;

/* One trial step's outcome (solver.rs:593-611). */
typedef struct {
    int accepted;
    int failure;      /* MPM_FAIL_* when rejected */
    double max_cells; /* largest particle displacement in cells */
    double dt;
} mpm_step_report;

/* FixtureSpec, fixtures.rs:94-115 (seeded inputs; jitter is the placement
 * jitter as a fraction of the sub-cell spacing, 0.1 in the reference). */
typedef struct {
    int fixture; /* 0 slump, 1 hydrostatic, 2 dam_break, 3 coupled_patch */
    mpm_material material;
    double grid_spacing;
    double domain[3];
    double initial_size[3];
    double initial_velocity[3];
    uint64_t seed;
    double gravity_norm;
    double jitter;
} mpm_fixture;

typedef struct mpm_engine mpm_engine;

/* ABI version of the loaded library. */
int mpm_abi_version(void);

/* Material constructors: Material::paste, ::paste_from_shear_rheology and
 * ::liquid, constitutive.rs:146-223. */
void mpm_material_paste(double density, double young, double poisson, double yield_stress,
                        double consistency, double flow_index, mpm_material* out);
void mpm_material_paste_shear(double density, double young, double poisson, double shear_yield,
                              double shear_consistency, double flow_index, mpm_material* out);
void mpm_material_liquid(double density, double young, double poisson, double viscosity,
                         mpm_material* out);

/* FixtureSpec::particle_count and ::place_particles with SplitMix64,
 * fixtures.rs:120-133, 272-349 and rng.rs:35-47.  `out` holds caller-allocated
 * arrays of mpm_fixture_particle_count() entries. */
size_t mpm_fixture_particle_count(const mpm_fixture* spec);
int mpm_fixture_place(const mpm_fixture* spec, mpm_particles* out);

/* Simulation::new, solver.rs:357-390: allocate device storage for `capacity`
 * particles and `max_pellets` pellets on the current HIP device. */
int mpm_create(const mpm_material* material, const mpm_config* config, const mpm_layout* layout,
               size_t capacity, size_t max_pellets, mpm_engine** out);
void mpm_destroy(mpm_engine* engine);
const char* mpm_last_error(const mpm_engine* engine);

/* Replace the state (restore from a snapshot, research.rs:324-368).  A NULL
 * ledger seeds it as Simulation::new does; a NULL history starts at t = 0. */
int mpm_set_state(mpm_engine* engine, const mpm_particles* particles, const mpm_pellet* pellets,
                  size_t n_pellets, const mpm_ledger* ledger, const mpm_history* history);
/* Copy the state out (snapshot, research.rs:308-321), in the reference's
 * particle order.  Arrays are caller-allocated with mpm_particle_count() entries. */
int mpm_get_state(mpm_engine* engine, mpm_particles* particles, mpm_pellet* pellets,
                  size_t pellet_capacity, mpm_ledger* ledger, mpm_history* history);
size_t mpm_particle_count(const mpm_engine* engine);
size_t mpm_pellet_count(const mpm_engine* engine);
int mpm_set_config(mpm_engine* engine, const mpm_config* config);

/* Simulation::try_step, solver.rs:535-611: one trial step of size dt;
 * commits only if every check passes.  Returns 0 or the failure code. */
int mpm_try_step(mpm_engine* engine, double dt, mpm_step_report* report);
/* Simulation::advance_while / advance_to, solver.rs:402-472.  keep_going may
 * be NULL; *reached is 0 when it stopped the loop at a resumable boundary. */
int mpm_advance_while(mpm_engine* engine, double target, int (*keep_going)(void*), void* ctx,
                      int* reached);
int mpm_advance_to(mpm_engine* engine, double target);
/* Simulation::stability_limit, solver.rs:477-498. */
double mpm_stability_limit(mpm_engine* engine);

/* Residual and energy diagnostics, solver.rs:1002-1099. */
double mpm_mass_residual(mpm_engine* engine);
int mpm_momentum_residual(mpm_engine* engine, double* out3);
int mpm_energies(mpm_engine* engine, double* kinetic, double* potential, double* elastic);

/* Batched constitutive update on the device for parity tests: update_paste
 * (constitutive.rs:456-531) for n trial deformation gradients (n x 9, f64 in
 * and out), or update_liquid (557-589) for n (J, C) pairs.  codes[i] is 0 or
 * the failure code of particle i. */
int mpm_paste_update_batch(const mpm_material* material, double dt, size_t n, const double* trial,
                           double* deformation, double* kirchhoff, double* increment,
                           double* dissipation_density, int* codes);
int mpm_liquid_update_batch(const mpm_material* material, double dt, size_t n,
                            const double* volume_ratio, const double* gradient, double* new_ratio,
                            double* kirchhoff, int* codes);

/* Per-kernel timing since profiling was last enabled, for the benchmark: the
 * average device time in ms of the grid-to-particle kernel (G2P with the
 * constitutive update) and of the particle-to-grid deposit kernel, and the
 * number of steps timed.  mpm_enable_profiling with on = N > 0 times the
 * kernels of every N-th trial step; on = 0 stops.  The kernels stamp their
 * first workgroup start and last workgroup end on the device clock, which
 * costs two atomics per workgroup and no idle time between kernels.  Either
 * call resets the sums.  Timing turns the graph replay off. */
int mpm_profile_kernels(mpm_engine* engine, double* g2p_avg_ms, double* p2g_avg_ms, uint64_t* launches);
void mpm_enable_profiling(mpm_engine* engine, int on);

/* ---- Batched instances.  No counterpart in the reference, whose sweeps run
 * one Simulation per job in sequence (python/litter_physics/sweep.py:170-261).
 * One engine holds n independent instances, each with its own material,
 * config, grid, particles, pellets, adaptive dt, accept/reject state and
 * ledgers; one launch sequence steps all of them.  An engine from mpm_create
 * is a batch of one, and the single-instance calls above act on instance 0. */
int mpm_batch_create(size_t n_instances, const mpm_material* materials, const mpm_config* solver_configs,
                     const mpm_layout* layouts, const size_t* capacities, const size_t* max_pellets,
                     mpm_engine** out);
size_t mpm_instance_count(const mpm_engine* engine);
size_t mpm_instance_particle_count(const mpm_engine* engine, size_t instance);
/* mpm_set_state / mpm_get_state for one instance. */
int mpm_instance_set_state(mpm_engine* engine, size_t instance, const mpm_particles* particles,
                           const mpm_pellet* pellets, size_t n_pellets, const mpm_ledger* ledger,
                           const mpm_history* history);
int mpm_instance_get_state(mpm_engine* engine, size_t instance, mpm_particles* particles,
                           mpm_pellet* pellets, size_t pellet_capacity, mpm_ledger* ledger,
                           mpm_history* history);
/* Advance every instance towards targets[i] (advance_to, solver.rs:402-472,
 * per instance) taking at most max_steps accepted steps each (0: no limit).
 * status[i] (may be NULL): 0 target reached, 1 step limit reached,
 * MPM_ERR_MIN_STEP when the instance aborted (its state stays consistent, as in
 * the reference).  done_ms[i] (may be NULL): time in ms from the start of the
 * call to the end of instance i's last accepted step, on the device clock (an
 * instance that took no step: the host time at which it was found stopped).
 * Returns MPM_OK, or a negative code for misuse or a device failure. */
int mpm_batch_advance(mpm_engine* engine, const double* targets, uint64_t max_steps, int* status,
                      double* done_ms);

/* Launch shape of the stepping loop (the autotuner's search space). */
typedef struct {
    int steps_per_sync;  /* trial steps enqueued between host synchronisations (1..256) */
    int use_graph;       /* 1: replay the steps as one captured HIP graph */
    int merge_blocks;    /* cell blocks per work item along z (1 or 2) */
    int traction_wgs;    /* workgroups of the wall-traction kernel; 0: sized from the near-wall lists */
    int g2p_occupancy;   /* G2P workgroups per CU its register budget allows (2 or 3) */
    int p2g_block;       /* threads of a deposit workgroup (256 or 512) */
} mpm_launch;
int mpm_get_launch(const mpm_engine* engine, mpm_launch* out);
int mpm_set_launch(mpm_engine* engine, const mpm_launch* launch);
/* Pick the launch shape for the current state: time max(steps, 64) trial
 * steps (two of the longest synchronisation intervals) per candidate from the
 * same state, restoring it each time.  First the kernel shapes (cell blocks per
 * work item, G2P occupancy, deposit workgroup size) by the device time of the
 * deposit and G2P kernels, then the synchronisation interval and graph replay
 * by wall time.  Keeps the fastest and returns it in *chosen (may be NULL).
 * The state, ledgers, histories and the caller's kernel timing are unchanged
 * afterwards. */
int mpm_autotune(mpm_engine* engine, int steps, mpm_launch* chosen);

/* ---- Learned surrogate of the Herschel-Bulkley return map.  No counterpart
 * in the reference; the exact return (solve_herschel_bulkley_return,
 * constitutive.rs:280-362) stays the fallback and the acceptance test of every
 * prediction.  The network is 12 -> 64 -> 64 -> 64 -> 7 with SiLU, bf16
 * weights on MFMA with f32 accumulation; its f32 master parameters are, in
 * order, W1[12][64] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[64][7] b4[7]
 * (W[k][n] maps input k to output n).  Inputs and outputs are the encoding
 * documented with the log below. */
typedef struct mpm_mlp mpm_mlp;
size_t mpm_mlp_param_count(void);
/* params NULL: Xavier-uniform weights from SplitMix64(seed), zero biases. */
int mpm_mlp_create(const float* params, uint64_t seed, mpm_mlp** out);
void mpm_mlp_destroy(mpm_mlp* mlp);
const char* mpm_mlp_last_error(const mpm_mlp* mlp);
int mpm_mlp_get_params(mpm_mlp* mlp, float* out);
int mpm_mlp_set_params(mpm_mlp* mlp, const float* in);
/* y (n x 7) = network(x (n x 12)), on the device. */
int mpm_mlp_forward(mpm_mlp* mlp, size_t n, const float* x, float* y);
/* Loss sum_i sum_u w_u (y_iu - t_iu)^2 / n (w NULL: all ones) and its gradient
 * with respect to the master parameters, from the MFMA backward pass. */
int mpm_mlp_loss_grad(mpm_mlp* mlp, size_t n, const float* x, const float* t,
                      const float* loss_weights, double* loss, float* grad);
/* mpm_mlp_loss_grad for callers that cannot allocate device memory: code under
 * HIP graph capture, training loops that forbid allocation.  No counterpart in
 * the reference.  x (n x 12), t (n x 7), loss_weights (7; NULL: all ones),
 * loss (1 double; may be NULL) and grad (mpm_mlp_param_count() floats; may be
 * NULL) are device pointers on the network's device, which must be current;
 * the caller's writes to x, t and loss_weights must be complete.  The call
 * allocates nothing: the per-workgroup partial gradients and losses live in a
 * fixed array of the library on each GPU (64 partials, 2.46 MB, held until
 * the process ends).  Every call on one GPU shares that array, so calls that
 * overlap on one GPU race.  n is at most INT32_MAX / 16 (MPM_ERR_CAPACITY
 * beyond).  Returns when the results are written. */
int mpm_mlp_loss_grad_noalloc(mpm_mlp* mlp, size_t n, const float* x, const float* t, const float* loss_weights,
                              double* loss, float* grad);
/* Adam on the device (cosine decay from lr to lr/100), minibatches of `batch`
 * samples in an order shuffled per epoch from SplitMix64(seed); losses[e]
 * (may be NULL) receives the mean loss of epoch e. */
int mpm_mlp_train(mpm_mlp* mlp, size_t n, const float* x, const float* t, const float* loss_weights,
                  int epochs, int batch, double lr, uint64_t seed, double* losses);
/* Use the network in the paste return of every instance (NULL: exact return
 * only).  A prediction is refined by polish_steps Newton corrections and
 * accepted when the return equation's residual |f(q)| / max(q_tr, 1) is at
 * most `tolerance`; otherwise the exact return runs.  The network must stay
 * alive while it is in use.  Resets the counters of mpm_surrogate_stats. */
int mpm_set_surrogate(mpm_engine* engine, const mpm_mlp* mlp, double tolerance, int polish_steps);
/* Yielded particles offered to the surrogate, and those whose prediction was accepted. */
int mpm_surrogate_stats(mpm_engine* engine, uint64_t* considered, uint64_t* accepted);
/* Log exact returns of yielded paste particles during stepping (capacity
 * samples; 0 turns logging off).  A sample is 12 inputs, 7 targets and 6
 * auxiliary values:
 *   inputs  0-5 trial deviatoric Kirchhoff stress / q_tr (xx yy zz xy xz yz),
 *           6 log10((q_tr - sigma_y) / (3 G dt)) / 4,
 *           7 log1p(1e4 dt eps_dot) / 5 (eps_dot: equivalent deviatoric rate of sym C),
 *           8 asinh(p / sigma_y) / 4, 9 sigma_y / q_tr, 10 log10(K / (3 G dt)) / 4, 11 n;
 *   targets 0-5 returned deviatoric stress / q_tr, 6 (q - sigma_y) / (q_tr - sigma_y);
 *   aux     q_tr, sigma_y, 3 G dt, K, n, q. */
int mpm_log_returns(mpm_engine* engine, size_t capacity);
/* Copy out up to `capacity` logged samples (arrays n x 12, n x 7, n x 6; any
 * may be NULL) and clear the log. */
int mpm_take_returns(mpm_engine* engine, float* inputs, float* targets, float* aux, size_t capacity,
                     size_t* n);

/* ------------------------------------------------------------------ validation and energy audit */

/* Simulation::validate, solver.rs:1106-1140 (with ParticleSet::validate,
 * particles.rs:159-211, and Ledger::validate, solver.rs:182-243): 0, or
 * MPM_ERR_INVALID with the reason in mpm_last_error.  Single-instance engines. */
int mpm_validate(mpm_engine* engine);

/* EnergyAudit, audit.rs:17-250.
 *   before/after: translation, affine (h^2/8 sum m |C|^2), elastic, potential, pellet;
 *   stages: grid_transport, grid_after_stress, grid_after_traction, grid_after_gravity,
 *     grid_after_walls, grid_after_coupling, pellet_after_coupling, plastic_dissipation;
 *   increments: plastic_dissipation, wall_friction_dissipation, wall_normal_projection_energy,
 *     wall_normal_traction_energy, coupling_grid_energy, coupling_pellet_energy, pellet_wall_work;
 *   terms: p2g_transfer_loss, g2p_transfer_loss, stress_storage_plastic_mismatch, gravity
 *     (uncoupled_gravity_potential_mismatch when coupled == 0, else
 *     coupled_grid_gravity_total_potential_change), pellet_integrator_combined_change,
 *     traction_change, walls_change, coupling_change, augmented_energy_change,
 *     telescoping_sum, telescoping_error. */
typedef struct {
    double time, dt;
    double before[5], after[5];
    double stages[8];
    double increments[7];
    double terms[11];
    int coupled;
} mpm_energy_audit;
/* Simulation::audit_step, audit.rs:275-345: measure one trial step of size dt
 * without changing the simulation (the device state is snapshotted and
 * restored).  The grid stage energies are reduced on the device in f64 from
 * the f32 node values.  Errors (invalid state, dt outside
 * [min_dt, min(max_dt, dt_scale * stability_limit)], a rejected trial,
 * nonfinite energies) return MPM_ERR_INVALID with the reason in mpm_last_error.
 * Single-instance engines. */
int mpm_audit_step(mpm_engine* engine, double dt, mpm_energy_audit* out);

/* ------------------------------------------------------------------ research wire layer
 * Strings returned through a char** are allocated by the library (also on
 * failure, where they hold the error message) and released with
 * mpm_string_free. */

/* protocol_version, core/src/lib.rs:11-16. */
int mpm_protocol_version(void);
void mpm_string_free(char* s);
/* run_json, core/src/lib.rs:24-49, for mode "research" (research::run,
 * core/src/research.rs:72-132): validate the request, build the fixture on
 * the GPU, advance record interval by record interval under the request's
 * wall-clock budget, and return the output JSON (frames, metrics,
 * observables, diagnostics, checkpoint).  resume_json (may be NULL) is a
 * checkpoint of an earlier segment (restore, research.rs:324-368).  Returns
 * 0, or MPM_ERR_INVALID (invalid input; household mode), MPM_ERR_MIN_STEP
 * (numerical failure) or MPM_ERR_DEVICE, with the message in *output (the
 * reference's ValueError text). */
int mpm_run_json(const char* request_json, const char* resume_json, char** output);
/* run_json over many research requests at once.  The reference's sweep and
 * refinement-study runners call run_json once per job, in sequence
 * (python/litter_physics/sweep.py:206-261, verification.py:645-680); here all
 * valid jobs step together as the instances of one batch engine, each to its
 * own record boundaries under its own wall budget, and each ends with the
 * output run_json would give it.  requests_json: a JSON array of requests;
 * resumes_json: NULL, or an array of the same length of checkpoints or nulls.
 * *output: a JSON array with one element per request, {"ok": true, "output":
 * <run_json output>} or {"ok": false, "error": "<message>"}.  Returns 0 unless
 * the arrays are malformed or the device fails. */
int mpm_run_json_batch(const char* requests_json, const char* resumes_json, char** output);
/* specification and FixtureSpec::validate, research.rs:170-232 and
 * fixtures.rs:152-236, without building anything: 0 or MPM_ERR_INVALID with
 * the message in *error. */
int mpm_research_validate(const char* request_json, char** error);
/* FixtureSpec::observables, fixtures.rs:369-615, of the request's fixture for
 * a given state (f64, in the reference's particle order): a JSON object. */
int mpm_research_observables(const char* request_json, const mpm_particles* particles, const mpm_pellet* pellets,
                             size_t n_pellets, const mpm_ledger* ledger, const mpm_history* history,
                             char** output);
/* The checkpoint codec (Checkpoint and Snapshot, research.rs:42-65, 308-321):
 * encode a state as checkpoint JSON, and decode one back.  Every f64 survives
 * the round trip bit for bit.  decode fails with MPM_ERR_CAPACITY, setting
 * particles->n, when the arrays are too small; mpm_checkpoint_particle_count
 * gives the count first. */
int mpm_checkpoint_encode(const char* request_json, const mpm_particles* particles, const mpm_pellet* pellets,
                          size_t n_pellets, const mpm_ledger* ledger, const mpm_history* history,
                          uint32_t next_record_index, char** output);
int mpm_checkpoint_particle_count(const char* checkpoint_json, size_t* n);
int mpm_checkpoint_decode(const char* checkpoint_json, mpm_particles* particles, mpm_pellet* pellets,
                          size_t pellet_capacity, size_t* n_pellets, mpm_ledger* ledger, mpm_history* history,
                          uint32_t* next_record_index, char** error);

#ifdef __cplusplus
}
#endif
#endif
