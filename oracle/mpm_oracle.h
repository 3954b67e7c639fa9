/*
 * CPU f64 oracle for the MLS-MPM substep -- TEST INFRASTRUCTURE ONLY.
 *
 * This library is a plain C++ restatement of the reference crate's CPU path
 * (core/src/mpm/{grid,particles,constitutive,rigid,solver,fixtures,rng}.rs),
 * kept in the reference's operation order so that it reproduces the
 * reference's published numbers.  It is the parity checker for the HIP
 * engine and the timed CPU baseline.  Only tests, the smoke check and the
 * benchmark's CPU-baseline leg may load it; the engine never does.
 *
 * Matrices cross this ABI as 9 doubles, row-major (m[3*i+j] = M(i,j)).
 * Extensions beyond the reference (each defaults to reference behaviour):
 *   - tait_gamma   : Tait EOS p = (K/g)(J^-g - 1); g == 1 is the reference law.
 *   - face_friction: per-face Coulomb coefficient (x-,x+,y-,y+,z-,z+);
 *                    +inf is a sticky face.  All equal to wall_friction
 *                    reproduces the reference.
 *   - jitter       : placement jitter as a fraction of the sub-cell spacing
 *                    (reference: 0.1).
 *   - several pellets, coupled to the grid one after another in index order
 *     (one pellet is the reference).
 */
#ifndef MPM_ORACLE_H
#define MPM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Material (constitutive.rs:111-131) plus the Tait exponent extension. */
typedef struct {
    int kind; /* 0 paste, 1 liquid */
    double density;
    double shear_modulus;
    double lame_lambda;
    double bulk_modulus;
    double yield_stress;
    double consistency;
    double flow_index;
    double tait_gamma; /* liquid only; 1.0 = reference linear EOS */
} oracle_material;

/* SolverConfig (solver.rs:49-62) plus per-face friction. */
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
} oracle_config;

/* FixtureSpec (fixtures.rs:94-115). */
typedef struct {
    int fixture; /* 0 slump, 1 hydrostatic, 2 dam_break, 3 coupled_patch */
    oracle_material material;
    double grid_spacing;
    double domain[3];
    double initial_size[3];
    double initial_velocity[3];
    uint64_t seed;
    int has_pellet;
    double pellet_radius, pellet_length, pellet_density;
    oracle_config config;
    size_t max_particles;
    size_t max_nodes;
    double jitter; /* fraction of the sub-cell spacing; reference 0.1 */
} oracle_fixture;

/* Ledger (solver.rs:67-124), field order as in the reference. */
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
} oracle_ledger;

/* Integration history (solver.rs:338-351). */
typedef struct {
    double time;
    uint64_t step_count;
    uint64_t limited_steps;
    uint64_t rejected_steps;
    double dt_scale;
    double min_dt_used;
    double max_dt_used;
} oracle_history;

/* Pellet state (rigid.rs:30-49); quaternion as (i, j, k, w). */
typedef struct {
    double radius, length, mass;
    double inertia_body[3];
    double position[3];
    double orientation[4];
    double velocity[3];
    double angular_momentum[3];
    int sphere_count;
    double sphere_offsets[64];
} oracle_pellet;

typedef struct oracle_sim oracle_sim;

/* EnergyAudit, audit.rs:17-250, flattened.
 *   before/after: translation, affine (h^2/8 sum m |C|^2), elastic, potential, pellet;
 *   stages: grid_transport, grid_after_stress, grid_after_traction, grid_after_gravity,
 *           grid_after_walls, grid_after_coupling, pellet_after_coupling, plastic_dissipation;
 *   increments: plastic_dissipation, wall_friction_dissipation, wall_normal_projection_energy,
 *           wall_normal_traction_energy, coupling_grid_energy, coupling_pellet_energy, pellet_wall_work;
 *   terms: p2g_transfer_loss, g2p_transfer_loss, stress_storage_plastic_mismatch, gravity
 *           (uncoupled_gravity_potential_mismatch when coupled == 0, else
 *           coupled_grid_gravity_total_potential_change), pellet_integrator_combined_change,
 *           traction_change, walls_change, coupling_change, augmented_energy_change,
 *           telescoping_sum, telescoping_error. */
typedef struct {
    double time, dt;
    double before[5], after[5];
    double stages[8];
    double increments[7];
    double terms[11];
    int coupled;
} oracle_audit;

/* Step failure codes (solver.rs:249-259, constitutive.rs:41-66). */
enum {
    ORACLE_OK = 0,
    ORACLE_FAIL_INVERTED = 1,
    ORACLE_FAIL_EIGEN = 2,
    ORACLE_FAIL_STRETCH = 3,
    ORACLE_FAIL_RETURN = 4,
    ORACLE_FAIL_NONFINITE_CONST = 5,
    ORACLE_FAIL_NONFINITE_GRID = 6,
    ORACLE_FAIL_DISPLACEMENT = 7,
    ORACLE_FAIL_PELLET = 8,
    ORACLE_FAIL_TRACTION = 9,
    ORACLE_ERR_MIN_STEP = -1,
    ORACLE_ERR_INVALID = -2
};

/* Materials (constitutive.rs:146-223). */
void oracle_material_paste(double density, double young, double poisson, double yield_stress,
                           double consistency, double flow_index, oracle_material* out);
void oracle_material_paste_shear(double density, double young, double poisson,
                                 double shear_yield, double shear_consistency, double flow_index,
                                 oracle_material* out);
void oracle_material_liquid(double density, double young, double poisson, double viscosity,
                            oracle_material* out);
double oracle_wave_speed(const oracle_material* m);
void oracle_default_config(oracle_config* c);

/* Constitutive kernels (constitutive.rs:280-595). */
int oracle_hb_return(double trial, double yield_stress, double consistency, double flow_index,
                     double shear_modulus, double dt, double* stress, double* increment,
                     int* iterations);
int oracle_eigen3(const double* m, double* values, double* vectors);
int oracle_update_paste(const double* trial, const oracle_material* m, double dt,
                        double* deformation, double* kirchhoff, double* increment,
                        double* dissipation_density, double* equivalent_stress);
int oracle_update_liquid(double volume_ratio, const double* gradient, const oracle_material* m,
                         double dt, double* new_ratio, double* kirchhoff, double* pressure);
double oracle_liquid_pressure(const oracle_material* m, double volume_ratio);

/* SplitMix64 (rng.rs:35-47). */
uint64_t oracle_splitmix_next(uint64_t* state);
double oracle_splitmix_unit(uint64_t* state);

/* Simulations. */
int oracle_fixture_build(const oracle_fixture* spec, oracle_sim** out, char* err, size_t errlen);
size_t oracle_fixture_particle_count(const oracle_fixture* spec);
int oracle_sim_create(const oracle_material* m, const oracle_config* c, const double* domain,
                      double spacing, size_t max_nodes, size_t n, const double* pos,
                      const double* vel, const double* mass, const double* volume0,
                      oracle_sim** out);
void oracle_sim_destroy(oracle_sim* s);
oracle_sim* oracle_sim_clone(const oracle_sim* s);
size_t oracle_sim_len(const oracle_sim* s);
/* Bulk state access; any pointer may be NULL. */
void oracle_sim_get(const oracle_sim* s, uint32_t* id, double* pos, double* vel, double* affine,
                    double* deformation, double* volume_ratio, double* kirchhoff,
                    double* plastic, double* mass, double* volume0);
int oracle_sim_set(oracle_sim* s, size_t n, const uint32_t* id, const double* pos,
                   const double* vel, const double* affine, const double* deformation,
                   const double* volume_ratio, const double* kirchhoff, const double* plastic,
                   const double* mass, const double* volume0);
void oracle_sim_get_ledger(const oracle_sim* s, oracle_ledger* l);
void oracle_sim_set_ledger(oracle_sim* s, const oracle_ledger* l);
void oracle_sim_get_history(const oracle_sim* s, oracle_history* h);
void oracle_sim_set_history(oracle_sim* s, const oracle_history* h);
void oracle_sim_get_config(const oracle_sim* s, oracle_config* c);
void oracle_sim_set_config(oracle_sim* s, const oracle_config* c);
void oracle_sim_get_layout(const oracle_sim* s, double* spacing, int* cells);
/* Pellets. */
int oracle_pellet_new(double radius, double length, double density, const double* position,
                      const double* orientation, oracle_pellet* out);
int oracle_sim_add_pellet(oracle_sim* s, const oracle_pellet* p);
int oracle_sim_pellet_count(const oracle_sim* s);
void oracle_sim_get_pellet(const oracle_sim* s, int index, oracle_pellet* p);
void oracle_sim_set_pellet(oracle_sim* s, int index, const oracle_pellet* p);
double oracle_pellet_kinetic_energy(const oracle_pellet* p);
void oracle_pellet_signed_distance(const oracle_pellet* p, const double* point, double* phi,
                                   double* normal);
/* Stepping (solver.rs:402-611).  advance returns 0 or a negative error code;
 * try_step returns 0 or a failure code and leaves the state untouched on failure. */
int oracle_sim_advance_to(oracle_sim* s, double target, double* fail_dt);
int oracle_sim_try_step(oracle_sim* s, double dt);
double oracle_sim_stability_limit(const oracle_sim* s);
void oracle_sim_sort(oracle_sim* s);
/* Grid-level probes for tests: P2G (+ traction) on a scratch copy.  Returns the
 * number of active nodes; fills traction totals and the full-scan kinetic-energy
 * jump of the traction transfer. */
size_t oracle_sim_probe_transfer(const oracle_sim* s, double dt, double* traction_impulse,
                                 double* traction_energy, double* full_scan_jump);
/* Dense node arrays of the scratch grid after a P2G probe (sized nodes^3). */
size_t oracle_sim_probe_grid(const oracle_sim* s, double dt, int with_traction, double* mass,
                             double* momentum);
/* P2G, wall traction and grid_update (solver.rs:632-892) on a scratch copy.
 * Dense node arrays (sized nodes^3; vectors 3 per node): mass, the velocity
 * the walls act on (p/m + g dt, solver.rs:833), the velocity after the walls
 * and the traction budget.  *code is 0 or the grid failure code.  Returns the
 * number of active nodes (0 on failure). */
size_t oracle_sim_probe_update(const oracle_sim* s, double dt, double* mass, double* before, double* after,
                               double* budget, int* code);
/* couple_grid (rigid.rs:385-438) of n_pellets pellets, one after another in
 * index order (one pellet is the reference's step), on a grid given by its
 * active nodes in order: dense node index, mass and velocity (3 per node,
 * updated in place).  results: 8 per pellet -- impulse (3), angular impulse
 * about the pellet centre (3), grid energy, constrained node count. */
int oracle_couple_nodes(const double* domain, double spacing, size_t n_nodes, const size_t* index, const double* mass,
                        double* velocity, const oracle_pellet* pellets, int n_pellets, double friction,
                        double* results);
/* Diagnostics (solver.rs:1002-1099). */
double oracle_sim_mass_residual(const oracle_sim* s);
void oracle_sim_momentum_residual(const oracle_sim* s, double* out);
double oracle_sim_kinetic_energy(const oracle_sim* s);
double oracle_sim_potential_energy(const oracle_sim* s);
double oracle_sim_elastic_energy(const oracle_sim* s);
double oracle_sim_mechanical_energy(const oracle_sim* s);
/* Simulation::audit_step, audit.rs:275-345: 0, or ORACLE_ERR_INVALID without touching the simulation. */
int oracle_sim_audit_step(const oracle_sim* s, double dt, oracle_audit* out);
void oracle_sim_get_material(const oracle_sim* s, oracle_material* out);
double oracle_sim_energy_residual(const oracle_sim* s);
/* Thread count used by the node- and particle-parallel phases. */
void oracle_set_threads(int threads);

#ifdef __cplusplus
}
#endif
#endif
