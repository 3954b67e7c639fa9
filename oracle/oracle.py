"""ctypes binding of the f64 oracle (test infrastructure only).

The oracle restates the reference crate's CPU path (core/src/mpm/*.rs) in C++;
this module exposes it to the parity tests, the smoke check and the
benchmark's CPU-baseline leg.  The engine never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

dp = C.POINTER(C.c_double)
u32p = C.POINTER(C.c_uint32)


class Material(C.Structure):
    _fields_ = [("kind", C.c_int), ("density", C.c_double), ("shear_modulus", C.c_double),
                ("lame_lambda", C.c_double), ("bulk_modulus", C.c_double), ("yield_stress", C.c_double),
                ("consistency", C.c_double), ("flow_index", C.c_double), ("tait_gamma", C.c_double)]


class Config(C.Structure):
    _fields_ = [("gravity", C.c_double * 3), ("wall_friction", C.c_double), ("cfl", C.c_double),
                ("max_dt", C.c_double), ("min_dt", C.c_double), ("contact_stiffness", C.c_double),
                ("contact_restitution", C.c_double), ("contact_friction", C.c_double),
                ("face_friction", C.c_double * 6)]


class Fixture(C.Structure):
    _fields_ = [("fixture", C.c_int), ("material", Material), ("grid_spacing", C.c_double),
                ("domain", C.c_double * 3), ("initial_size", C.c_double * 3),
                ("initial_velocity", C.c_double * 3), ("seed", C.c_uint64), ("has_pellet", C.c_int),
                ("pellet_radius", C.c_double), ("pellet_length", C.c_double),
                ("pellet_density", C.c_double), ("config", Config), ("max_particles", C.c_size_t),
                ("max_nodes", C.c_size_t), ("jitter", C.c_double)]


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


FIXTURES = {"slump": 0, "hydrostatic": 1, "dam_break": 2, "coupled_patch": 3}
FAILURES = {1: "inverted", 2: "eigen", 3: "stretch", 4: "return", 5: "nonfinite", 6: "nonfinite_grid",
            7: "displacement", 8: "pellet", 9: "traction", -1: "minimum_step", -2: "invalid"}


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


class Audit(C.Structure):
    """EnergyAudit (audit.rs:17-250), flattened as oracle_audit in mpm_oracle.h."""
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
        """The reference's serialised layout (Option terms as None)."""
        terms = dict(zip(self.TERMS, self.terms))
        g = terms.pop("gravity")
        terms["uncoupled_gravity_potential_mismatch"] = None if self.coupled else g
        terms["coupled_grid_gravity_total_potential_change"] = g if self.coupled else None
        return {"time": self.time, "dt": self.dt, "before": dict(zip(self.STATE, self.before)),
                "after": dict(zip(self.STATE, self.after)), "stages": dict(zip(self.STAGES, self.stages)),
                "increments": dict(zip(self.INCREMENTS, self.increments)), "terms": terms}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        sig = {
            "oracle_material_paste": (None, [C.c_double] * 6 + [C.POINTER(Material)]),
            "oracle_material_paste_shear": (None, [C.c_double] * 6 + [C.POINTER(Material)]),
            "oracle_material_liquid": (None, [C.c_double] * 4 + [C.POINTER(Material)]),
            "oracle_wave_speed": (C.c_double, [C.POINTER(Material)]),
            "oracle_default_config": (None, [C.POINTER(Config)]),
            "oracle_hb_return": (C.c_int, [C.c_double] * 6 + [dp, dp, C.POINTER(C.c_int)]),
            "oracle_eigen3": (C.c_int, [dp, dp, dp]),
            "oracle_update_paste": (C.c_int, [dp, C.POINTER(Material), C.c_double, dp, dp, dp, dp, dp]),
            "oracle_update_liquid": (C.c_int, [C.c_double, dp, C.POINTER(Material), C.c_double, dp, dp, dp]),
            "oracle_liquid_pressure": (C.c_double, [C.POINTER(Material), C.c_double]),
            "oracle_splitmix_next": (C.c_uint64, [C.POINTER(C.c_uint64)]),
            "oracle_splitmix_unit": (C.c_double, [C.POINTER(C.c_uint64)]),
            "oracle_fixture_build": (C.c_int, [C.POINTER(Fixture), C.POINTER(C.c_void_p), C.c_char_p,
                                               C.c_size_t]),
            "oracle_fixture_particle_count": (C.c_size_t, [C.POINTER(Fixture)]),
            "oracle_sim_create": (C.c_int, [C.POINTER(Material), C.POINTER(Config), dp, C.c_double,
                                            C.c_size_t, C.c_size_t, dp, dp, dp, dp, C.POINTER(C.c_void_p)]),
            "oracle_sim_destroy": (None, [C.c_void_p]),
            "oracle_sim_clone": (C.c_void_p, [C.c_void_p]),
            "oracle_sim_len": (C.c_size_t, [C.c_void_p]),
            "oracle_sim_get": (None, [C.c_void_p, u32p, dp, dp, dp, dp, dp, dp, dp, dp, dp]),
            "oracle_sim_set": (C.c_int, [C.c_void_p, C.c_size_t, u32p, dp, dp, dp, dp, dp, dp, dp, dp, dp]),
            "oracle_sim_get_ledger": (None, [C.c_void_p, C.POINTER(Ledger)]),
            "oracle_sim_set_ledger": (None, [C.c_void_p, C.POINTER(Ledger)]),
            "oracle_sim_get_history": (None, [C.c_void_p, C.POINTER(History)]),
            "oracle_sim_set_history": (None, [C.c_void_p, C.POINTER(History)]),
            "oracle_sim_get_config": (None, [C.c_void_p, C.POINTER(Config)]),
            "oracle_sim_set_config": (None, [C.c_void_p, C.POINTER(Config)]),
            "oracle_sim_get_layout": (None, [C.c_void_p, dp, C.POINTER(C.c_int)]),
            "oracle_pellet_new": (C.c_int, [C.c_double, C.c_double, C.c_double, dp, dp, C.POINTER(Pellet)]),
            "oracle_sim_add_pellet": (C.c_int, [C.c_void_p, C.POINTER(Pellet)]),
            "oracle_sim_pellet_count": (C.c_int, [C.c_void_p]),
            "oracle_sim_get_pellet": (None, [C.c_void_p, C.c_int, C.POINTER(Pellet)]),
            "oracle_sim_set_pellet": (None, [C.c_void_p, C.c_int, C.POINTER(Pellet)]),
            "oracle_pellet_kinetic_energy": (C.c_double, [C.POINTER(Pellet)]),
            "oracle_pellet_signed_distance": (None, [C.POINTER(Pellet), dp, dp, dp]),
            "oracle_sim_advance_to": (C.c_int, [C.c_void_p, C.c_double, dp]),
            "oracle_sim_try_step": (C.c_int, [C.c_void_p, C.c_double]),
            "oracle_sim_stability_limit": (C.c_double, [C.c_void_p]),
            "oracle_sim_sort": (None, [C.c_void_p]),
            "oracle_sim_probe_transfer": (C.c_size_t, [C.c_void_p, C.c_double, dp, dp, dp]),
            "oracle_sim_probe_grid": (C.c_size_t, [C.c_void_p, C.c_double, C.c_int, dp, dp]),
            "oracle_sim_probe_update": (C.c_size_t, [C.c_void_p, C.c_double, dp, dp, dp, dp, C.POINTER(C.c_int)]),
            "oracle_couple_nodes": (C.c_int, [dp, C.c_double, C.c_size_t, C.POINTER(C.c_size_t), dp, dp,
                                              C.POINTER(Pellet), C.c_int, C.c_double, dp]),
            "oracle_sim_mass_residual": (C.c_double, [C.c_void_p]),
            "oracle_sim_momentum_residual": (None, [C.c_void_p, dp]),
            "oracle_sim_kinetic_energy": (C.c_double, [C.c_void_p]),
            "oracle_sim_potential_energy": (C.c_double, [C.c_void_p]),
            "oracle_sim_elastic_energy": (C.c_double, [C.c_void_p]),
            "oracle_sim_mechanical_energy": (C.c_double, [C.c_void_p]),
            "oracle_sim_energy_residual": (C.c_double, [C.c_void_p]),
            "oracle_set_threads": (None, [C.c_int]),
            "oracle_sim_audit_step": (C.c_int, [C.c_void_p, C.c_double, C.POINTER(Audit)]),
            "oracle_sim_get_material": (None, [C.c_void_p, C.POINTER(Material)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def ptr(a):
    if a is None:
        return None
    if a.dtype == np.float64:
        return a.ctypes.data_as(dp)
    if a.dtype == np.uint32:
        return a.ctypes.data_as(u32p)
    raise TypeError(a.dtype)


def paste(density, young, poisson, yield_stress, consistency, flow_index, shear_units=False):
    m = Material()
    fn = lib().oracle_material_paste_shear if shear_units else lib().oracle_material_paste
    fn(density, young, poisson, yield_stress, consistency, flow_index, C.byref(m))
    return m


def liquid(density, young, poisson, viscosity, tait_gamma=1.0):
    m = Material()
    lib().oracle_material_liquid(density, young, poisson, viscosity, C.byref(m))
    m.tait_gamma = tait_gamma
    return m


def config(gravity=(0.0, 0.0, -9.81), wall_friction=0.4, cfl=0.3, max_dt=1e-3, min_dt=1e-9,
           stiffness=1000.0, restitution=0.2, contact_friction=0.4, face_friction=None):
    c = Config()
    lib().oracle_default_config(C.byref(c))
    c.gravity[:] = list(gravity)
    c.wall_friction = wall_friction
    c.cfl = cfl
    c.max_dt = max_dt
    c.min_dt = min_dt
    c.contact_stiffness = stiffness
    c.contact_restitution = restitution
    c.contact_friction = contact_friction
    ff = [wall_friction] * 6 if face_friction is None else list(face_friction)
    c.face_friction[:] = ff
    return c


def fixture(kind, material, spacing, domain, initial_size, cfg, initial_velocity=(0.0, 0.0, 0.0), seed=0,
            pellet=None, max_particles=100_000_000, max_nodes=100_000_000, jitter=0.1):
    f = Fixture()
    f.fixture = FIXTURES[kind]
    f.material = material
    f.grid_spacing = spacing
    f.domain[:] = list(domain)
    f.initial_size[:] = list(initial_size)
    f.initial_velocity[:] = list(initial_velocity)
    f.seed = seed
    if pellet is not None:
        f.has_pellet = 1
        f.pellet_radius, f.pellet_length, f.pellet_density = pellet
    f.config = cfg
    f.max_particles = max_particles
    f.max_nodes = max_nodes
    f.jitter = jitter
    return f


@dataclass
class State:
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


class Sim:
    """Owning handle of one oracle simulation."""

    def __init__(self, handle):
        self.h = C.c_void_p(handle)

    @classmethod
    def from_fixture(cls, f):
        h = C.c_void_p()
        err = C.create_string_buffer(512)
        rc = lib().oracle_fixture_build(C.byref(f), C.byref(h), err, 512)
        if rc != 0:
            raise ValueError(err.value.decode())
        return cls(h.value)

    @classmethod
    def create(cls, material, cfg, domain, spacing, pos, vel, mass, volume0, max_nodes=100_000_000):
        h = C.c_void_p()
        dom = np.asarray(domain, dtype=np.float64)
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        n = pos.shape[0]
        vel = np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 3)
        mass = np.ascontiguousarray(mass, dtype=np.float64).reshape(n)
        volume0 = np.ascontiguousarray(volume0, dtype=np.float64).reshape(n)
        rc = lib().oracle_sim_create(C.byref(material), C.byref(cfg), ptr(dom), spacing, max_nodes, n, ptr(pos),
                                     ptr(vel), ptr(mass), ptr(volume0), C.byref(h))
        if rc != 0:
            raise ValueError("invalid layout")
        return cls(h.value)

    def __del__(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().oracle_sim_destroy(self.h)
            self.h = None

    def clone(self):
        return Sim(lib().oracle_sim_clone(self.h))

    def __len__(self):
        return lib().oracle_sim_len(self.h)

    def state(self) -> State:
        n = len(self)
        s = State(np.zeros(n, np.uint32), np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3, 3)),
                  np.zeros((n, 3, 3)), np.zeros(n), np.zeros((n, 3, 3)), np.zeros(n), np.zeros(n), np.zeros(n))
        lib().oracle_sim_get(self.h, ptr(s.id), ptr(s.pos), ptr(s.vel), ptr(s.affine), ptr(s.deformation),
                             ptr(s.volume_ratio), ptr(s.kirchhoff), ptr(s.plastic), ptr(s.mass), ptr(s.volume0))
        return s

    def set_state(self, s: State):
        n = s.pos.shape[0]
        arrs = [np.ascontiguousarray(a) for a in
                (s.pos, s.vel, s.affine, s.deformation, s.volume_ratio, s.kirchhoff, s.plastic, s.mass,
                 s.volume0)]
        ids = np.ascontiguousarray(s.id, dtype=np.uint32)
        lib().oracle_sim_set(self.h, n, ptr(ids), *[ptr(a.astype(np.float64)) for a in arrs])

    def ledger(self) -> Ledger:
        l = Ledger()
        lib().oracle_sim_get_ledger(self.h, C.byref(l))
        return l

    def set_ledger(self, l):
        lib().oracle_sim_set_ledger(self.h, C.byref(l))

    def history(self) -> History:
        h = History()
        lib().oracle_sim_get_history(self.h, C.byref(h))
        return h

    def set_history(self, h):
        lib().oracle_sim_set_history(self.h, C.byref(h))

    def config(self) -> Config:
        c = Config()
        lib().oracle_sim_get_config(self.h, C.byref(c))
        return c

    def set_config(self, c):
        lib().oracle_sim_set_config(self.h, C.byref(c))

    def layout(self):
        sp = C.c_double()
        cells = (C.c_int * 3)()
        lib().oracle_sim_get_layout(self.h, C.byref(sp), cells)
        return sp.value, tuple(cells)

    def add_pellet(self, p: Pellet):
        return lib().oracle_sim_add_pellet(self.h, C.byref(p))

    def pellets(self):
        out = []
        for i in range(lib().oracle_sim_pellet_count(self.h)):
            p = Pellet()
            lib().oracle_sim_get_pellet(self.h, i, C.byref(p))
            out.append(p)
        return out

    def set_pellet(self, i, p):
        lib().oracle_sim_set_pellet(self.h, i, C.byref(p))

    def advance_to(self, target):
        fail = C.c_double()
        rc = lib().oracle_sim_advance_to(self.h, target, C.byref(fail))
        if rc != 0:
            raise RuntimeError(f"oracle advance failed: {FAILURES.get(rc, rc)} at dt={fail.value}")

    def try_step(self, dt):
        return lib().oracle_sim_try_step(self.h, dt)

    def stability_limit(self):
        return lib().oracle_sim_stability_limit(self.h)

    def sort(self):
        lib().oracle_sim_sort(self.h)

    def probe_transfer(self, dt):
        imp = np.zeros(3)
        e = C.c_double()
        j = C.c_double()
        lib().oracle_sim_probe_transfer(self.h, dt, ptr(imp), C.byref(e), C.byref(j))
        return imp, e.value, j.value

    def probe_grid(self, dt, with_traction=True):
        sp, cells = self.layout()
        n = (cells[0] + 3) * (cells[1] + 3) * (cells[2] + 3)
        m = np.zeros(n)
        p = np.zeros((n, 3))
        lib().oracle_sim_probe_grid(self.h, dt, int(with_traction), ptr(m), ptr(p))
        return m, p

    def probe_update(self, dt):
        """P2G, traction and grid_update on a scratch copy: dense node mass, velocity
        before and after the walls, traction budget (solver.rs:632-892)."""
        sp, cells = self.layout()
        n = (cells[0] + 3) * (cells[1] + 3) * (cells[2] + 3)
        m, vb, va, tb = np.zeros(n), np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        code = C.c_int()
        lib().oracle_sim_probe_update(self.h, dt, ptr(m), ptr(vb), ptr(va), ptr(tb), C.byref(code))
        if code.value != 0:
            raise RuntimeError(f"grid update failed: {FAILURES.get(code.value, code.value)}")
        return m, vb, va, tb

    def mass_residual(self):
        return lib().oracle_sim_mass_residual(self.h)

    def momentum_residual(self):
        out = np.zeros(3)
        lib().oracle_sim_momentum_residual(self.h, ptr(out))
        return out

    def kinetic_energy(self):
        return lib().oracle_sim_kinetic_energy(self.h)

    def potential_energy(self):
        return lib().oracle_sim_potential_energy(self.h)

    def elastic_energy(self):
        return lib().oracle_sim_elastic_energy(self.h)

    def mechanical_energy(self):
        return lib().oracle_sim_mechanical_energy(self.h)

    def energy_residual(self):
        return lib().oracle_sim_energy_residual(self.h)

    def audit_step(self, dt) -> dict:
        """Simulation::audit_step (audit.rs:275-345); raises ValueError when the reference would err."""
        a = Audit()
        if lib().oracle_sim_audit_step(self.h, dt, C.byref(a)) != 0:
            raise ValueError("audit rejected")
        return a.as_dict()


def hb_return(trial, ys, k, n, mu, dt):
    s = C.c_double()
    inc = C.c_double()
    it = C.c_int()
    rc = lib().oracle_hb_return(trial, ys, k, n, mu, dt, C.byref(s), C.byref(inc), C.byref(it))
    return rc, s.value, inc.value, it.value


def eigen3(m):
    m = np.ascontiguousarray(m, dtype=np.float64).reshape(9)
    vals = np.zeros(3)
    vecs = np.zeros(9)
    rc = lib().oracle_eigen3(ptr(m), ptr(vals), ptr(vecs))
    return rc, vals, vecs.reshape(3, 3)


def update_paste(trial, material, dt):
    trial = np.ascontiguousarray(trial, dtype=np.float64).reshape(9)
    f = np.zeros(9)
    tau = np.zeros(9)
    inc = C.c_double()
    dd = C.c_double()
    eq = C.c_double()
    rc = lib().oracle_update_paste(ptr(trial), C.byref(material), dt, ptr(f), ptr(tau), C.byref(inc),
                                   C.byref(dd), C.byref(eq))
    return rc, f.reshape(3, 3), tau.reshape(3, 3), inc.value, dd.value, eq.value


def update_liquid(j, gradient, material, dt):
    g = np.ascontiguousarray(gradient, dtype=np.float64).reshape(9)
    nj = C.c_double()
    tau = np.zeros(9)
    p = C.c_double()
    rc = lib().oracle_update_liquid(j, ptr(g), C.byref(material), dt, C.byref(nj), ptr(tau), C.byref(p))
    return rc, nj.value, tau.reshape(3, 3), p.value


def liquid_pressure(material, j):
    return lib().oracle_liquid_pressure(C.byref(material), j)


def splitmix(seed, count):
    st = C.c_uint64(seed)
    return [lib().oracle_splitmix_next(C.byref(st)) for _ in range(count)]


def new_pellet(radius, length, density, position, orientation=(0.0, 0.0, 0.0, 1.0)):
    p = Pellet()
    pos = np.asarray(position, dtype=np.float64)
    q = np.asarray(orientation, dtype=np.float64)
    rc = lib().oracle_pellet_new(radius, length, density, ptr(pos), ptr(q), C.byref(p))
    if rc != 0:
        raise ValueError("invalid pellet geometry")
    return p


def couple_nodes(domain, spacing, index, mass, velocity, pellets, friction):
    """couple_grid (rigid.rs:385-438) of `pellets` in order on the given active
    nodes (dense index, mass, velocity).  Returns the new velocities and, per
    pellet, impulse (3), angular impulse about its centre (3), grid energy and
    constrained node count."""
    dom = np.asarray(domain, dtype=np.float64)
    idx = np.ascontiguousarray(index, dtype=np.uintp)
    m = np.ascontiguousarray(mass, dtype=np.float64)
    v = np.ascontiguousarray(velocity, dtype=np.float64).reshape(-1, 3).copy()
    arr = (Pellet * max(len(pellets), 1))(*pellets)
    res = np.zeros((max(len(pellets), 1), 8))
    rc = lib().oracle_couple_nodes(ptr(dom), spacing, len(idx), idx.ctypes.data_as(C.POINTER(C.c_size_t)), ptr(m),
                                   ptr(v), arr, len(pellets), friction, ptr(res))
    if rc != 0:
        raise ValueError("invalid coupling layout")
    return v, res[:len(pellets)]


def set_threads(n):
    lib().oracle_set_threads(int(n))
