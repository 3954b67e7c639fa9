"""MI355X (gfx950) MLS-MPM substep engine: Python host binding.

`Simulation` mirrors the reference crate's `Simulation` (core/src/mpm/solver.rs:324-1140):
`advance_to`, `advance_while`, `try_step`, `stability_limit`, the residual and energy
getters and the public state.  Every call goes through the C ABI of
include/mpm_engine.h into libmpm_amd.so; the particle and grid work runs only in the
HIP kernels.  There is no CPU path: without the built library or a GPU the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmpm_amd.so")

dp = C.POINTER(C.c_double)
fp = C.POINTER(C.c_float)
u32p = C.POINTER(C.c_uint32)


class MpmError(RuntimeError):
    """A SolverError (negative status) or API misuse."""


class StepRejected(RuntimeError):
    """A trial step failed (StepFailure, positive status)."""

    def __init__(self, code):
        super().__init__(FAILURES.get(code, str(code)))
        self.code = code


FAILURES = {1: "inverted deformation", 2: "eigen-decomposition failed", 3: "degenerate stretch",
            4: "return not converged", 5: "non-finite constitutive quantity", 6: "non-finite grid velocity",
            7: "displacement over one cell", 8: "pellet non-finite", 9: "non-finite traction energy"}


class Material(C.Structure):
    _fields_ = [("kind", C.c_int), ("density", C.c_double), ("shear_modulus", C.c_double),
                ("lame_lambda", C.c_double), ("bulk_modulus", C.c_double), ("yield_stress", C.c_double),
                ("consistency", C.c_double), ("flow_index", C.c_double), ("tait_gamma", C.c_double)]


class Config(C.Structure):
    _fields_ = [("gravity", C.c_double * 3), ("wall_friction", C.c_double), ("cfl", C.c_double),
                ("max_dt", C.c_double), ("min_dt", C.c_double), ("contact_stiffness", C.c_double),
                ("contact_restitution", C.c_double), ("contact_friction", C.c_double),
                ("face_friction", C.c_double * 6)]


class Layout(C.Structure):
    _fields_ = [("domain", C.c_double * 3), ("spacing", C.c_double), ("max_nodes", C.c_size_t)]


class Particles(C.Structure):
    _fields_ = [("n", C.c_size_t), ("id", u32p), ("position", dp), ("velocity", dp), ("affine", dp),
                ("deformation", dp), ("volume_ratio", dp), ("kirchhoff", dp), ("plastic_strain", dp),
                ("mass", dp), ("volume0", dp)]


class Ledger(C.Structure):
    _fields_ = [("initial_mass", C.c_double), ("outflow_mass", C.c_double),
                ("initial_momentum", C.c_double * 3), ("initial_mechanical_energy", C.c_double),
                ("gravity_impulse", C.c_double * 3), ("wall_impulse", C.c_double * 3),
                ("pellet_wall_impulse", C.c_double * 3), ("coupling_impulse", C.c_double * 3),
                ("coupling_angular_impulse", C.c_double * 3),
                ("wall_normal_projection_energy", C.c_double), ("wall_friction_dissipation", C.c_double),
                ("wall_normal_traction_energy", C.c_double), ("pellet_wall_work", C.c_double),
                ("coupling_grid_energy", C.c_double), ("coupling_pellet_energy", C.c_double),
                ("plastic_dissipation", C.c_double), ("outflow_momentum", C.c_double * 3)]

    def ledgered_energy(self):
        """Ledger::ledgered_energy, solver.rs:168-175."""
        return (self.wall_normal_projection_energy - self.wall_friction_dissipation
                + self.wall_normal_traction_energy + self.pellet_wall_work + self.coupling_grid_energy
                + self.coupling_pellet_energy - self.plastic_dissipation)


class History(C.Structure):
    _fields_ = [("time", C.c_double), ("step_count", C.c_uint64), ("limited_steps", C.c_uint64),
                ("rejected_steps", C.c_uint64), ("dt_scale", C.c_double), ("min_dt_used", C.c_double),
                ("max_dt_used", C.c_double)]


class Pellet(C.Structure):
    _fields_ = [("radius", C.c_double), ("length", C.c_double), ("mass", C.c_double),
                ("inertia_body", C.c_double * 3), ("position", C.c_double * 3),
                ("orientation", C.c_double * 4), ("velocity", C.c_double * 3),
                ("angular_momentum", C.c_double * 3), ("sphere_count", C.c_int),
                ("sphere_offsets", C.c_double * 64)]


class StepReport(C.Structure):
    _fields_ = [("accepted", C.c_int), ("failure", C.c_int), ("max_cells", C.c_double), ("dt", C.c_double)]


class Launch(C.Structure):
    """mpm_launch: launch shape of the stepping loop (the autotuner's search space)."""
    _fields_ = [("steps_per_sync", C.c_int), ("use_graph", C.c_int), ("merge_blocks", C.c_int),
                ("traction_wgs", C.c_int), ("g2p_occupancy", C.c_int), ("p2g_block", C.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class EnergyAudit(C.Structure):
    """EnergyAudit (audit.rs:17-250) as mpm_energy_audit in include/mpm_engine.h."""
    _fields_ = [("time", C.c_double), ("dt", C.c_double), ("before", C.c_double * 5), ("after", C.c_double * 5),
                ("stages", C.c_double * 8), ("increments", C.c_double * 7), ("terms", C.c_double * 11),
                ("coupled", C.c_int)]

    STATE = ("translation", "affine", "elastic", "potential", "pellet")
    STAGES = ("grid_transport", "grid_after_stress", "grid_after_traction", "grid_after_gravity", "grid_after_walls",
              "grid_after_coupling", "pellet_after_coupling", "plastic_dissipation")
    INCREMENTS = ("plastic_dissipation", "wall_friction_dissipation", "wall_normal_projection_energy",
                  "wall_normal_traction_energy", "coupling_grid_energy", "coupling_pellet_energy", "pellet_wall_work")
    TERMS = ("p2g_transfer_loss", "g2p_transfer_loss", "stress_storage_plastic_mismatch", "gravity",
             "pellet_integrator_combined_change", "traction_change", "walls_change", "coupling_change",
             "augmented_energy_change", "telescoping_sum", "telescoping_error")

    def as_dict(self):
        """The reference's serialised layout (the Option terms as None)."""
        terms = dict(zip(self.TERMS, self.terms))
        g = terms.pop("gravity")
        terms["uncoupled_gravity_potential_mismatch"] = None if self.coupled else g
        terms["coupled_grid_gravity_total_potential_change"] = g if self.coupled else None
        return {"time": self.time, "dt": self.dt, "before": dict(zip(self.STATE, self.before)),
                "after": dict(zip(self.STATE, self.after)), "stages": dict(zip(self.STAGES, self.stages)),
                "increments": dict(zip(self.INCREMENTS, self.increments)), "terms": terms}


class FixtureSpec(C.Structure):
    _fields_ = [("fixture", C.c_int), ("material", Material), ("grid_spacing", C.c_double),
                ("domain", C.c_double * 3), ("initial_size", C.c_double * 3),
                ("initial_velocity", C.c_double * 3), ("seed", C.c_uint64), ("gravity_norm", C.c_double),
                ("jitter", C.c_double)]


KEEP_GOING = C.CFUNCTYPE(C.c_int, C.c_void_p)

_lib = None


def build(verbose=False):
    from importlib import util
    spec = util.spec_from_file_location("_mpm_build", os.path.join(HERE, "build.py"))
    mod = util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.build(verbose=verbose)


def lib():
    """Load libmpm_amd.so (built by build.py); raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MpmError(f"{LIB_PATH} is not built; run litter-physics_amd/build.py")
    L = C.CDLL(LIB_PATH)
    sig = {
        "mpm_abi_version": (C.c_int, []),
        "mpm_material_paste": (None, [C.c_double] * 6 + [C.POINTER(Material)]),
        "mpm_material_paste_shear": (None, [C.c_double] * 6 + [C.POINTER(Material)]),
        "mpm_material_liquid": (None, [C.c_double] * 4 + [C.POINTER(Material)]),
        "mpm_fixture_particle_count": (C.c_size_t, [C.POINTER(FixtureSpec)]),
        "mpm_fixture_place": (C.c_int, [C.POINTER(FixtureSpec), C.POINTER(Particles)]),
        "mpm_create": (C.c_int, [C.POINTER(Material), C.POINTER(Config), C.POINTER(Layout), C.c_size_t,
                                 C.c_size_t, C.POINTER(C.c_void_p)]),
        "mpm_destroy": (None, [C.c_void_p]),
        "mpm_last_error": (C.c_char_p, [C.c_void_p]),
        "mpm_set_state": (C.c_int, [C.c_void_p, C.POINTER(Particles), C.POINTER(Pellet), C.c_size_t,
                                    C.POINTER(Ledger), C.POINTER(History)]),
        "mpm_get_state": (C.c_int, [C.c_void_p, C.POINTER(Particles), C.POINTER(Pellet), C.c_size_t,
                                    C.POINTER(Ledger), C.POINTER(History)]),
        "mpm_particle_count": (C.c_size_t, [C.c_void_p]),
        "mpm_pellet_count": (C.c_size_t, [C.c_void_p]),
        "mpm_set_config": (C.c_int, [C.c_void_p, C.POINTER(Config)]),
        "mpm_try_step": (C.c_int, [C.c_void_p, C.c_double, C.POINTER(StepReport)]),
        "mpm_advance_while": (C.c_int, [C.c_void_p, C.c_double, KEEP_GOING, C.c_void_p, C.POINTER(C.c_int)]),
        "mpm_advance_to": (C.c_int, [C.c_void_p, C.c_double]),
        "mpm_stability_limit": (C.c_double, [C.c_void_p]),
        "mpm_mass_residual": (C.c_double, [C.c_void_p]),
        "mpm_momentum_residual": (C.c_int, [C.c_void_p, dp]),
        "mpm_energies": (C.c_int, [C.c_void_p, dp, dp, dp]),
        "mpm_paste_update_batch": (C.c_int, [C.POINTER(Material), C.c_double, C.c_size_t, dp, dp, dp, dp, dp,
                                             C.POINTER(C.c_int)]),
        "mpm_liquid_update_batch": (C.c_int, [C.POINTER(Material), C.c_double, C.c_size_t, dp, dp, dp, dp,
                                              C.POINTER(C.c_int)]),
        "mpm_profile_kernels": (C.c_int, [C.c_void_p, dp, dp, C.POINTER(C.c_uint64)]),
        "mpm_enable_profiling": (None, [C.c_void_p, C.c_int]),
        "mpm_batch_create": (C.c_int, [C.c_size_t, C.POINTER(Material), C.POINTER(Config), C.POINTER(Layout),
                                       C.POINTER(C.c_size_t), C.POINTER(C.c_size_t), C.POINTER(C.c_void_p)]),
        "mpm_instance_count": (C.c_size_t, [C.c_void_p]),
        "mpm_instance_particle_count": (C.c_size_t, [C.c_void_p, C.c_size_t]),
        "mpm_instance_set_state": (C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(Particles), C.POINTER(Pellet),
                                             C.c_size_t, C.POINTER(Ledger), C.POINTER(History)]),
        "mpm_instance_get_state": (C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(Particles), C.POINTER(Pellet),
                                             C.c_size_t, C.POINTER(Ledger), C.POINTER(History)]),
        "mpm_batch_advance": (C.c_int, [C.c_void_p, dp, C.c_uint64, C.POINTER(C.c_int), dp]),
        "mpm_get_launch": (C.c_int, [C.c_void_p, C.POINTER(Launch)]),
        "mpm_set_launch": (C.c_int, [C.c_void_p, C.POINTER(Launch)]),
        "mpm_autotune": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(Launch)]),
        "mpm_mlp_param_count": (C.c_size_t, []),
        "mpm_mlp_create": (C.c_int, [fp, C.c_uint64, C.POINTER(C.c_void_p)]),
        "mpm_mlp_destroy": (None, [C.c_void_p]),
        "mpm_mlp_last_error": (C.c_char_p, [C.c_void_p]),
        "mpm_mlp_get_params": (C.c_int, [C.c_void_p, fp]),
        "mpm_mlp_set_params": (C.c_int, [C.c_void_p, fp]),
        "mpm_mlp_forward": (C.c_int, [C.c_void_p, C.c_size_t, fp, fp]),
        "mpm_mlp_loss_grad": (C.c_int, [C.c_void_p, C.c_size_t, fp, fp, fp, dp, fp]),
        "mpm_mlp_loss_grad_noalloc": (C.c_int, [C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p,
                                                C.c_void_p, C.c_void_p]),
        "mpm_mlp_train": (C.c_int, [C.c_void_p, C.c_size_t, fp, fp, fp, C.c_int, C.c_int, C.c_double, C.c_uint64,
                                    dp]),
        "mpm_set_surrogate": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double, C.c_int]),
        "mpm_surrogate_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
        "mpm_log_returns": (C.c_int, [C.c_void_p, C.c_size_t]),
        "mpm_take_returns": (C.c_int, [C.c_void_p, fp, fp, fp, C.c_size_t, C.POINTER(C.c_size_t)]),
        "mpm_validate": (C.c_int, [C.c_void_p]),
        "mpm_audit_step": (C.c_int, [C.c_void_p, C.c_double, C.POINTER(EnergyAudit)]),
        "mpm_protocol_version": (C.c_int, []),
        "mpm_string_free": (None, [C.c_void_p]),
        "mpm_run_json": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(C.c_void_p)]),
        "mpm_run_json_batch": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(C.c_void_p)]),
        "mpm_research_validate": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
        "mpm_research_observables": (C.c_int, [C.c_char_p, C.POINTER(Particles), C.POINTER(Pellet), C.c_size_t,
                                               C.POINTER(Ledger), C.POINTER(History), C.POINTER(C.c_void_p)]),
        "mpm_checkpoint_encode": (C.c_int, [C.c_char_p, C.POINTER(Particles), C.POINTER(Pellet), C.c_size_t,
                                            C.POINTER(Ledger), C.POINTER(History), C.c_uint32,
                                            C.POINTER(C.c_void_p)]),
        "mpm_checkpoint_particle_count": (C.c_int, [C.c_char_p, C.POINTER(C.c_size_t)]),
        "mpm_checkpoint_decode": (C.c_int, [C.c_char_p, C.POINTER(Particles), C.POINTER(Pellet), C.c_size_t,
                                            C.POINTER(C.c_size_t), C.POINTER(Ledger), C.POINTER(History),
                                            C.POINTER(C.c_uint32), C.POINTER(C.c_void_p)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


EXPORTED = [
    "mpm_abi_version", "mpm_material_paste", "mpm_material_paste_shear", "mpm_material_liquid",
    "mpm_fixture_particle_count", "mpm_fixture_place", "mpm_create", "mpm_destroy", "mpm_last_error",
    "mpm_set_state", "mpm_get_state", "mpm_particle_count", "mpm_pellet_count", "mpm_set_config",
    "mpm_try_step", "mpm_advance_while", "mpm_advance_to", "mpm_stability_limit", "mpm_mass_residual",
    "mpm_momentum_residual", "mpm_energies", "mpm_paste_update_batch", "mpm_liquid_update_batch",
    "mpm_profile_kernels", "mpm_enable_profiling", "mpm_batch_create", "mpm_instance_count",
    "mpm_instance_particle_count", "mpm_instance_set_state", "mpm_instance_get_state", "mpm_batch_advance",
    "mpm_get_launch", "mpm_set_launch", "mpm_autotune", "mpm_mlp_param_count", "mpm_mlp_create", "mpm_mlp_destroy",
    "mpm_mlp_last_error", "mpm_mlp_get_params", "mpm_mlp_set_params", "mpm_mlp_forward", "mpm_mlp_loss_grad",
    "mpm_mlp_loss_grad_noalloc",
    "mpm_mlp_train", "mpm_set_surrogate", "mpm_surrogate_stats", "mpm_log_returns", "mpm_take_returns",
    "mpm_validate", "mpm_audit_step", "mpm_protocol_version", "mpm_string_free", "mpm_run_json", "mpm_run_json_batch", "mpm_research_validate", "mpm_research_observables",
    "mpm_checkpoint_encode", "mpm_checkpoint_particle_count", "mpm_checkpoint_decode",
]


def _p(a):
    if a is None:
        return None
    if a.dtype == np.float64:
        return a.ctypes.data_as(dp)
    if a.dtype == np.float32:
        return a.ctypes.data_as(fp)
    if a.dtype == np.uint32:
        return a.ctypes.data_as(u32p)
    raise TypeError(a.dtype)


# ------------------------------------------------------------------ inputs

def paste(density, young, poisson, yield_stress, consistency, flow_index, shear_units=False):
    """Material::paste / ::paste_from_shear_rheology, constitutive.rs:146-197."""
    m = Material()
    fn = lib().mpm_material_paste_shear if shear_units else lib().mpm_material_paste
    fn(density, young, poisson, yield_stress, consistency, flow_index, C.byref(m))
    return m


def liquid(density, young, poisson, viscosity, tait_gamma=1.0):
    """Material::liquid, constitutive.rs:211-223, with the Tait exponent."""
    m = Material()
    lib().mpm_material_liquid(density, young, poisson, viscosity, C.byref(m))
    m.tait_gamma = tait_gamma
    return m


def config(gravity=(0.0, 0.0, -9.81), wall_friction=0.4, cfl=0.3, max_dt=1e-3, min_dt=1e-9, stiffness=1000.0,
           restitution=0.2, contact_friction=0.4, face_friction=None):
    """SolverConfig, solver.rs:49-62 (face_friction: x-, x+, y-, y+, z-, z+)."""
    c = Config()
    c.gravity[:] = list(gravity)
    c.wall_friction = wall_friction
    c.cfl = cfl
    c.max_dt = max_dt
    c.min_dt = min_dt
    c.contact_stiffness = stiffness
    c.contact_restitution = restitution
    c.contact_friction = contact_friction
    c.face_friction[:] = [wall_friction] * 6 if face_friction is None else list(face_friction)
    return c


@dataclass
class State:
    """ParticleSet (particles.rs:11-32) as f64 numpy arrays."""
    id: np.ndarray
    pos: np.ndarray
    vel: np.ndarray
    affine: np.ndarray
    deformation: np.ndarray
    volume_ratio: np.ndarray
    kirchhoff: np.ndarray
    plastic: np.ndarray
    mass: np.ndarray
    volume0: np.ndarray

    @classmethod
    def empty(cls, n):
        return cls(np.zeros(n, np.uint32), np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3, 3)),
                   np.zeros((n, 3, 3)), np.zeros(n), np.zeros((n, 3, 3)), np.zeros(n), np.zeros(n), np.zeros(n))

    def __len__(self):
        return self.pos.shape[0]

    def to_c(self):
        self._keep = [np.ascontiguousarray(a, dtype=np.float64) for a in
                      (self.pos, self.vel, self.affine, self.deformation, self.volume_ratio, self.kirchhoff,
                       self.plastic, self.mass, self.volume0)]
        self._ids = np.ascontiguousarray(self.id, dtype=np.uint32)
        k = self._keep
        return Particles(len(self), _p(self._ids), _p(k[0]), _p(k[1]), _p(k[2]), _p(k[3]), _p(k[4]), _p(k[5]),
                         _p(k[6]), _p(k[7]), _p(k[8]))

    def by_id(self):
        """Copy sorted by particle id (for comparisons across engines)."""
        o = np.argsort(self.id, kind="stable")
        return State(*(getattr(self, f)[o] for f in
                       ("id", "pos", "vel", "affine", "deformation", "volume_ratio", "kirchhoff", "plastic",
                        "mass", "volume0")))


FIXTURES = {"slump": 0, "hydrostatic": 1, "dam_break": 2, "coupled_patch": 3}


def place_particles(kind, material, spacing, domain, initial_size, initial_velocity=(0.0, 0.0, 0.0), seed=0,
                    jitter=0.1, gravity=9.81) -> State:
    """FixtureSpec::place_particles with SplitMix64 jitter (fixtures.rs:272-349)."""
    f = FixtureSpec()
    f.fixture = FIXTURES[kind]
    f.material = material
    f.grid_spacing = spacing
    f.domain[:] = list(domain)
    f.initial_size[:] = list(initial_size)
    f.initial_velocity[:] = list(initial_velocity)
    f.seed = seed
    f.gravity_norm = gravity
    f.jitter = jitter
    n = lib().mpm_fixture_particle_count(C.byref(f))
    st = State.empty(n)
    cp = st.to_c()
    rc = lib().mpm_fixture_place(C.byref(f), C.byref(cp))
    if rc != 0:
        raise MpmError(f"fixture placement failed ({rc})")
    for name, arr in zip(("pos", "vel", "affine", "deformation", "volume_ratio", "kirchhoff", "plastic", "mass",
                          "volume0"), st._keep):
        setattr(st, name, arr)
    st.id = st._ids
    return st


def new_pellet(radius, length, density, position, orientation=(0.0, 0.0, 0.0, 1.0)):
    """Pellet::new, rigid.rs:115-182 (multisphere proxy of the cylinder)."""
    if not (radius > 0 and length >= 2 * radius and density > 0):
        raise ValueError("invalid pellet geometry")
    p = Pellet()
    mass = density * np.pi * radius * radius * length
    p.radius, p.length, p.mass = radius, length, mass
    perp = mass * (3.0 * radius * radius + length * length) / 12.0
    p.inertia_body[:] = [0.5 * mass * radius * radius, perp, perp]
    p.position[:] = list(position)
    p.orientation[:] = list(orientation)
    half = 0.5 * length - radius
    count = 1 if half <= 0 else int(np.ceil(2.0 * half / radius)) + 1
    if count > 64:
        raise ValueError("pellet proxy exceeds the sphere cap")
    p.sphere_count = count
    for i in range(count):
        p.sphere_offsets[i] = 0.0 if count == 1 else -half + 2.0 * half * i / (count - 1)
    return p


# -------------------------------------------------------------- simulation

class Simulation:
    """Device-resident MLS-MPM state behind the reference's step API."""

    def __init__(self, material, cfg, domain, spacing, state: State, pellets=(), capacity=None, max_pellets=None,
                 max_nodes=1 << 40, ledger=None, history=None):
        L = lib()
        layout = Layout()
        layout.domain[:] = list(domain)
        layout.spacing = spacing
        layout.max_nodes = max_nodes
        self.material = material
        self.cfg = cfg
        self.domain = tuple(domain)
        self.spacing = spacing
        cap = max(len(state), 1) if capacity is None else capacity
        mp = len(pellets) if max_pellets is None else max_pellets
        h = C.c_void_p()
        rc = L.mpm_create(C.byref(material), C.byref(cfg), C.byref(layout), cap, mp, C.byref(h))
        if rc != 0:
            raise MpmError(f"mpm_create failed ({rc})")
        self.h = h
        self.set_state(state, pellets, ledger, history)

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value and _lib is not None:
            _lib.mpm_destroy(h)
            self.h = None

    def _check(self, rc):
        if rc < 0:
            raise MpmError(lib().mpm_last_error(self.h).decode())
        return rc

    def set_state(self, state: State, pellets=(), ledger=None, history=None):
        cp = state.to_c()
        arr = (Pellet * max(len(pellets), 1))(*pellets) if pellets else None
        self._check(lib().mpm_set_state(self.h, C.byref(cp), arr, len(pellets),
                                        C.byref(ledger) if ledger is not None else None,
                                        C.byref(history) if history is not None else None))

    def set_config(self, cfg):
        self.cfg = cfg
        self._check(lib().mpm_set_config(self.h, C.byref(cfg)))

    def __len__(self):
        return lib().mpm_particle_count(self.h)

    def state(self) -> State:
        st = State.empty(len(self))
        cp = st.to_c()
        self._check(lib().mpm_get_state(self.h, C.byref(cp), None, 0, None, None))
        for name, arr in zip(("pos", "vel", "affine", "deformation", "volume_ratio", "kirchhoff", "plastic",
                              "mass", "volume0"), st._keep):
            setattr(st, name, arr)
        st.id = st._ids
        return st

    def pellets(self):
        n = lib().mpm_pellet_count(self.h)
        arr = (Pellet * max(n, 1))()
        self._check(lib().mpm_get_state(self.h, None, arr, n, None, None))
        return [arr[i] for i in range(n)]

    def ledger(self) -> Ledger:
        led = Ledger()
        self._check(lib().mpm_get_state(self.h, None, None, 0, C.byref(led), None))
        return led

    def history(self) -> History:
        hist = History()
        self._check(lib().mpm_get_state(self.h, None, None, 0, None, C.byref(hist)))
        return hist

    @property
    def time(self):
        return self.history().time

    def try_step(self, dt) -> StepReport:
        """Simulation::try_step (solver.rs:535-611): commit or leave the state untouched."""
        rep = StepReport()
        self._check(lib().mpm_try_step(self.h, dt, C.byref(rep)))
        return rep

    def advance_to(self, target):
        """Simulation::advance_to, solver.rs:402-404."""
        self._check(lib().mpm_advance_to(self.h, target))

    def advance_while(self, target, should_continue) -> bool:
        """Simulation::advance_while, solver.rs:413-472; False at a resumable budget stop."""
        cb = KEEP_GOING(lambda _ctx: 1 if should_continue() else 0)
        reached = C.c_int()
        self._check(lib().mpm_advance_while(self.h, target, cb, None, C.byref(reached)))
        return bool(reached.value)

    def advance_steps(self, steps, target=1e9) -> int:
        """Advance towards `target` taking at most `steps` accepted steps, with no
        host callback between steps; returns the number of accepted steps."""
        h0 = self.history().step_count
        _advance(self, [target], steps)
        return self.history().step_count - h0

    def launch(self) -> Launch:
        return _get_launch(self)

    def set_launch(self, **kw):
        _set_launch(self, **kw)

    def autotune(self, steps=8) -> Launch:
        return _autotune(self, steps)

    def set_surrogate(self, net, tolerance=1e-4, polish=2):
        """Use `net` for the paste return (None: exact only); see mpm_set_surrogate."""
        _set_surrogate(self, net, tolerance, polish)

    def surrogate_stats(self):
        """(yielded particles offered to the surrogate, predictions accepted)."""
        return _surrogate_stats(self)

    def log_returns(self, capacity):
        _log_returns(self, capacity)

    def take_returns(self, capacity):
        """(inputs n x 12, targets n x 7, aux n x 6) of the logged exact returns."""
        return _take_returns(self, capacity)

    def stability_limit(self):
        return lib().mpm_stability_limit(self.h)

    def validate(self):
        """Simulation::validate (solver.rs:1106-1140); raises MpmError with the reason."""
        self._check(lib().mpm_validate(self.h))

    def audit_step(self, dt) -> dict:
        """Simulation::audit_step (audit.rs:275-345): one measured trial of size dt that
        leaves the simulation unchanged; raises MpmError where the reference errs."""
        a = EnergyAudit()
        self._check(lib().mpm_audit_step(self.h, dt, C.byref(a)))
        return a.as_dict()

    def mass_residual(self):
        return lib().mpm_mass_residual(self.h)

    def momentum_residual(self):
        out = np.zeros(3)
        self._check(lib().mpm_momentum_residual(self.h, _p(out)))
        return out

    def energies(self):
        k, p, e = C.c_double(), C.c_double(), C.c_double()
        self._check(lib().mpm_energies(self.h, C.byref(k), C.byref(p), C.byref(e)))
        return k.value, p.value, e.value

    def mechanical_energy(self):
        return sum(self.energies())

    def energy_residual(self):
        """Simulation::energy_residual, solver.rs:1095-1099."""
        led = self.ledger()
        return self.mechanical_energy() - led.initial_mechanical_energy - led.ledgered_energy()

    def enable_profiling(self, on=True, every=1):
        """Time the G2P and deposit kernels of every `every`-th trial step (on=False stops)."""
        lib().mpm_enable_profiling(self.h, max(int(every), 1) if on else 0)

    def kernel_profile(self):
        """(G2P average ms, deposit average ms, steps timed) since profiling was enabled."""
        g = C.c_double()
        d = C.c_double()
        n = C.c_uint64()
        self._check(lib().mpm_profile_kernels(self.h, C.byref(g), C.byref(d), C.byref(n)))
        return g.value, d.value, n.value


def _f32(a, cols=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a if cols is None else a.reshape(-1, cols)


class Network:
    """Return-map surrogate network on the device (mpm_mlp_*): 12 -> 64 -> 64 -> 64 -> 7, SiLU,
    bf16 MFMA with f32 accumulation; f32 master parameters in the layout of mpm_engine.h."""

    def __init__(self, params=None, seed=0):
        h = C.c_void_p()
        p = None if params is None else _f32(params)
        rc = lib().mpm_mlp_create(_p(p), seed, C.byref(h))
        if rc != 0:
            raise MpmError(f"mpm_mlp_create failed ({rc})")
        self.h = h

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value and _lib is not None:
            _lib.mpm_mlp_destroy(h)
            self.h = None

    def _check(self, rc):
        if rc < 0:
            raise MpmError(lib().mpm_mlp_last_error(self.h).decode())
        return rc

    @staticmethod
    def param_count():
        return lib().mpm_mlp_param_count()

    def params(self):
        out = np.zeros(self.param_count(), np.float32)
        self._check(lib().mpm_mlp_get_params(self.h, _p(out)))
        return out

    def set_params(self, p):
        p = _f32(p)
        assert p.size == self.param_count()
        self._check(lib().mpm_mlp_set_params(self.h, _p(p)))

    def forward(self, x):
        x = _f32(x, 12)
        y = np.zeros((x.shape[0], 7), np.float32)
        self._check(lib().mpm_mlp_forward(self.h, x.shape[0], _p(x), _p(y)))
        return y

    def loss_grad(self, x, t, weights=None):
        x, t = _f32(x, 12), _f32(t, 7)
        w = None if weights is None else _f32(weights)
        g = np.zeros(self.param_count(), np.float32)
        loss = C.c_double()
        self._check(lib().mpm_mlp_loss_grad(self.h, x.shape[0], _p(x), _p(t), _p(w), C.byref(loss), _p(g)))
        return loss.value, g

    def loss_grad_device(self, n, x_ptr, t_ptr, loss_ptr=None, grad_ptr=None, weights_ptr=None):
        """mpm_mlp_loss_grad_noalloc on device pointers (integers): inputs x (n x 12) and
        t (n x 7) f32, optional loss weights (7 f32), outputs loss (1 f64) and grad (f32)."""
        return self._check(lib().mpm_mlp_loss_grad_noalloc(self.h, n, x_ptr, t_ptr, weights_ptr, loss_ptr, grad_ptr))

    def train(self, x, t, weights=None, epochs=1, batch=4096, lr=1e-3, seed=0):
        x, t = _f32(x, 12), _f32(t, 7)
        w = None if weights is None else _f32(weights)
        losses = np.zeros(epochs)
        self._check(lib().mpm_mlp_train(self.h, x.shape[0], _p(x), _p(t), _p(w), epochs, batch, lr, seed,
                                        _p(losses)))
        return losses


def _set_surrogate(eng, net, tolerance=1e-4, polish=2):
    eng._surrogate = net  # the network must outlive its use by the engine
    eng._check(lib().mpm_set_surrogate(eng.h, net.h if net is not None else None, tolerance, polish))


def _surrogate_stats(eng):
    a, b = C.c_uint64(), C.c_uint64()
    eng._check(lib().mpm_surrogate_stats(eng.h, C.byref(a), C.byref(b)))
    return a.value, b.value


def _log_returns(eng, capacity):
    eng._check(lib().mpm_log_returns(eng.h, capacity))


def _take_returns(eng, capacity):
    x = np.zeros((capacity, 12), np.float32)
    t = np.zeros((capacity, 7), np.float32)
    aux = np.zeros((capacity, 6), np.float32)
    n = C.c_size_t()
    eng._check(lib().mpm_take_returns(eng.h, _p(x), _p(t), _p(aux), capacity, C.byref(n)))
    k = n.value
    return x[:k], t[:k], aux[:k]


def _get_launch(eng) -> Launch:
    out = Launch()
    eng._check(lib().mpm_get_launch(eng.h, C.byref(out)))
    return out


def _set_launch(eng, **kw):
    cur = _get_launch(eng)
    for k, v in kw.items():
        setattr(cur, k, int(v))
    eng._check(lib().mpm_set_launch(eng.h, C.byref(cur)))


def _autotune(eng, steps) -> Launch:
    out = Launch()
    eng._check(lib().mpm_autotune(eng.h, steps, C.byref(out)))
    return out


def _advance(eng, targets, max_steps=0):
    t = np.ascontiguousarray(targets, dtype=np.float64)
    status = np.zeros(len(t), np.int32)
    done = np.zeros(len(t))
    eng._check(lib().mpm_batch_advance(eng.h, _p(t), int(max_steps), status.ctypes.data_as(C.POINTER(C.c_int)),
                                       _p(done)))
    return status, done


class Batch:
    """Many independent instances stepped by one launch sequence (mpm_batch_*).

    Each instance has its own material, config, grid, particles, pellets, adaptive
    dt and ledgers; `advance` moves every instance towards its own target."""

    def __init__(self, specs):
        """specs: list of dicts with material, cfg, domain, spacing, state, and
        optionally pellets, capacity, max_pellets."""
        L = lib()
        n = len(specs)
        mats = (Material * n)(*[s["material"] for s in specs])
        cfgs = (Config * n)(*[s["cfg"] for s in specs])
        lays = (Layout * n)()
        caps = (C.c_size_t * n)()
        mps = (C.c_size_t * n)()
        for i, s in enumerate(specs):
            lays[i].domain[:] = list(s["domain"])
            lays[i].spacing = s["spacing"]
            lays[i].max_nodes = s.get("max_nodes", 1 << 40)
            caps[i] = s.get("capacity") or max(len(s["state"]), 1)
            mps[i] = s.get("max_pellets", len(s.get("pellets", ())))
        h = C.c_void_p()
        rc = L.mpm_batch_create(n, mats, cfgs, lays, caps, mps, C.byref(h))
        if rc != 0:
            raise MpmError(f"mpm_batch_create failed ({rc})")
        self.h = h
        self.n = n
        for i, s in enumerate(specs):
            self.set_state(i, s["state"], s.get("pellets", ()))

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value and _lib is not None:
            _lib.mpm_destroy(h)
            self.h = None

    def __len__(self):
        return self.n

    def _check(self, rc):
        if rc < 0:
            raise MpmError(lib().mpm_last_error(self.h).decode())
        return rc

    def set_state(self, i, state: State, pellets=(), ledger=None, history=None):
        cp = state.to_c()
        arr = (Pellet * max(len(pellets), 1))(*pellets) if pellets else None
        self._check(lib().mpm_instance_set_state(self.h, i, C.byref(cp), arr, len(pellets),
                                                 C.byref(ledger) if ledger is not None else None,
                                                 C.byref(history) if history is not None else None))

    def particle_count(self, i):
        return lib().mpm_instance_particle_count(self.h, i)

    def state(self, i) -> State:
        st = State.empty(self.particle_count(i))
        cp = st.to_c()
        self._check(lib().mpm_instance_get_state(self.h, i, C.byref(cp), None, 0, None, None))
        for name, arr in zip(("pos", "vel", "affine", "deformation", "volume_ratio", "kirchhoff", "plastic",
                              "mass", "volume0"), st._keep):
            setattr(st, name, arr)
        st.id = st._ids
        return st

    def ledger(self, i) -> Ledger:
        led = Ledger()
        self._check(lib().mpm_instance_get_state(self.h, i, None, None, 0, C.byref(led), None))
        return led

    def history(self, i) -> History:
        hist = History()
        self._check(lib().mpm_instance_get_state(self.h, i, None, None, 0, None, C.byref(hist)))
        return hist

    def advance(self, targets, max_steps=0):
        """Advance instance i towards targets[i]; returns (status, done_ms) arrays."""
        if np.isscalar(targets):
            targets = [targets] * self.n
        return _advance(self, targets, max_steps)

    def launch(self) -> Launch:
        return _get_launch(self)

    def set_launch(self, **kw):
        _set_launch(self, **kw)

    def autotune(self, steps=8) -> Launch:
        return _autotune(self, steps)

    def set_surrogate(self, net, tolerance=1e-4, polish=2):
        """Use `net` for the paste return (None: exact only); see mpm_set_surrogate."""
        _set_surrogate(self, net, tolerance, polish)

    def surrogate_stats(self):
        """(yielded particles offered to the surrogate, predictions accepted)."""
        return _surrogate_stats(self)

    def log_returns(self, capacity):
        _log_returns(self, capacity)

    def take_returns(self, capacity):
        """(inputs n x 12, targets n x 7, aux n x 6) of the logged exact returns."""
        return _take_returns(self, capacity)

    def enable_profiling(self, on=True, every=1):
        """Time the G2P and deposit kernels of every `every`-th trial step (on=False stops)."""
        lib().mpm_enable_profiling(self.h, max(int(every), 1) if on else 0)

    def kernel_profile(self):
        """(G2P average ms, deposit average ms, steps timed) since profiling was enabled."""
        g = C.c_double()
        d = C.c_double()
        n = C.c_uint64()
        self._check(lib().mpm_profile_kernels(self.h, C.byref(g), C.byref(d), C.byref(n)))
        return g.value, d.value, n.value


def paste_update_batch(material, dt, trial):
    """update_paste (constitutive.rs:456-531) on the device for a batch of trial F."""
    trial = np.ascontiguousarray(trial, dtype=np.float64).reshape(-1, 9)
    n = trial.shape[0]
    f = np.zeros((n, 9))
    tau = np.zeros((n, 9))
    inc = np.zeros(n)
    dd = np.zeros(n)
    codes = np.zeros(n, np.int32)
    rc = lib().mpm_paste_update_batch(C.byref(material), dt, n, _p(trial), _p(f), _p(tau), _p(inc), _p(dd),
                                      codes.ctypes.data_as(C.POINTER(C.c_int)))
    if rc != 0:
        raise MpmError(f"paste batch failed ({rc})")
    return f.reshape(n, 3, 3), tau.reshape(n, 3, 3), inc, dd, codes


def liquid_update_batch(material, dt, volume_ratio, gradient):
    """update_liquid (constitutive.rs:557-589) on the device for a batch of (J, C)."""
    j = np.ascontiguousarray(volume_ratio, dtype=np.float64).reshape(-1)
    g = np.ascontiguousarray(gradient, dtype=np.float64).reshape(-1, 9)
    n = j.shape[0]
    nj = np.zeros(n)
    tau = np.zeros((n, 9))
    codes = np.zeros(n, np.int32)
    rc = lib().mpm_liquid_update_batch(C.byref(material), dt, n, _p(j), _p(g), _p(nj), _p(tau),
                                       codes.ctypes.data_as(C.POINTER(C.c_int)))
    if rc != 0:
        raise MpmError(f"liquid batch failed ({rc})")
    return nj, tau.reshape(n, 3, 3), codes


# ------------------------------------------------------------------ research wire layer
# core/src/lib.rs:11-49 and core/src/research.rs: JSON in, JSON out.  Failures
# raise ValueError with the message, as the reference's extension does.

def _take_string(ptr):
    if not ptr.value:
        return ""
    try:
        return C.string_at(ptr.value).decode()
    finally:
        lib().mpm_string_free(ptr)


def protocol_version() -> int:
    """protocol_version (core/src/lib.rs:11-16)."""
    return lib().mpm_protocol_version()


def run_json(request_json: str, resume_json: str | None = None) -> str:
    """run_json (core/src/lib.rs:24-49) for research requests, on the GPU engine."""
    out = C.c_void_p()
    rc = lib().mpm_run_json(request_json.encode(), None if resume_json is None else resume_json.encode(),
                            C.byref(out))
    text = _take_string(out)
    if rc != 0:
        raise ValueError(text)
    return text


def run_json_batch(requests_json, resumes_json=None) -> list:
    """run_json for many research requests stepped together on one batch engine.
    Returns one dict per request: {"ok": True, "output": <output dict>} or
    {"ok": False, "error": <message>}."""
    import json
    reqs = "[" + ",".join(requests_json) + "]"
    res = None if resumes_json is None else "[" + ",".join("null" if r is None else r for r in resumes_json) + "]"
    out = C.c_void_p()
    rc = lib().mpm_run_json_batch(reqs.encode(), None if res is None else res.encode(), C.byref(out))
    text = _take_string(out)
    if rc != 0:
        raise ValueError(text)
    return json.loads(text)


def research_validate(request_json: str) -> None:
    """Request and fixture validation only (research.rs:135-270, fixtures.rs:152-236)."""
    err = C.c_void_p()
    rc = lib().mpm_research_validate(request_json.encode(), C.byref(err))
    text = _take_string(err)
    if rc != 0:
        raise ValueError(text)


def _pellet_array(pellets):
    pellets = list(pellets)
    arr = (Pellet * max(len(pellets), 1))(*pellets)
    return arr, len(pellets)


def research_observables(request_json: str, state: State, pellets, ledger: Ledger, history: History) -> dict:
    """FixtureSpec::observables (fixtures.rs:369-615) of a given state."""
    import json
    parts = state.to_c()
    arr, n = _pellet_array(pellets)
    out = C.c_void_p()
    rc = lib().mpm_research_observables(request_json.encode(), C.byref(parts), arr, n, C.byref(ledger),
                                        C.byref(history), C.byref(out))
    text = _take_string(out)
    if rc != 0:
        raise ValueError(text)
    return json.loads(text)


def checkpoint_encode(request_json: str, state: State, pellets, ledger: Ledger, history: History,
                      next_record_index: int) -> str:
    """Checkpoint JSON of a state (research.rs:42-65, 308-321)."""
    parts = state.to_c()
    arr, n = _pellet_array(pellets)
    out = C.c_void_p()
    rc = lib().mpm_checkpoint_encode(request_json.encode(), C.byref(parts), arr, n, C.byref(ledger),
                                     C.byref(history), next_record_index, C.byref(out))
    text = _take_string(out)
    if rc != 0:
        raise ValueError(text)
    return text


def checkpoint_decode(text: str):
    """(State, pellets, Ledger, History, next_record_index) of checkpoint JSON."""
    raw = text.encode()
    n = C.c_size_t(0)
    if lib().mpm_checkpoint_particle_count(raw, C.byref(n)) != 0:
        raise ValueError("checkpoint has no particle id array")
    st = State.empty(n.value)
    parts = st.to_c()
    pel = (Pellet * 1)()
    npel = C.c_size_t(0)
    led, hist, nri = Ledger(), History(), C.c_uint32(0)
    err = C.c_void_p()
    rc = lib().mpm_checkpoint_decode(raw, C.byref(parts), pel, 1, C.byref(npel), C.byref(led), C.byref(hist),
                                     C.byref(nri), C.byref(err))
    msg = _take_string(err)
    if rc != 0:
        raise ValueError(msg)
    k = st._keep
    out = State(st._ids, *k)
    return out, [pel[i] for i in range(npel.value)], led, hist, nri.value
