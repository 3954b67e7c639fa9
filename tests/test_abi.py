"""CPU-side checks of the engine's C ABI and host logic (no GPU calls)."""
import os
import re

import numpy as np

import cases
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "mpm_engine.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mpm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(engine):
    lib = engine.lib()
    names = declared_symbols()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(engine.EXPORTED)
    assert lib.mpm_abi_version() == 1


def test_fixture_placement_is_bit_identical_to_oracle(engine):
    # fixtures.rs:272-319 with SplitMix64: both builders must agree bit for bit.
    for spec, kind in ((cases.paste_column_32(), "slump"), (cases.water_column([0.02] * 3, 0.01, 0.002, 5e-5),
                                                       "hydrostatic"),
                       (cases.dam_break_block(cells=(16, 16, 16), block=(8, 16, 12), dx=0.001), "dam_break")):
        ref = O.Sim.from_fixture(spec).state()
        mat = engine.Material(*[getattr(spec.material, n) for n, _ in spec.material._fields_])
        got = engine.place_particles(kind, mat, spec.grid_spacing, list(spec.domain), list(spec.initial_size),
                                     seed=spec.seed, jitter=spec.jitter)
        assert len(got) == ref.pos.shape[0]
        for field in ("pos", "vel", "mass", "volume0", "volume_ratio", "deformation", "kirchhoff"):
            assert np.array_equal(getattr(got, field), getattr(ref, field)), field
        assert np.array_equal(got.id, ref.id)


def test_materials_match_reference_constructors(engine):
    # constitutive.rs:146-223
    a = engine.paste(1000.0, 1000.0, 0.2, 30.0, 10.0, 0.6, shear_units=True)
    b = O.paste(1000.0, 1000.0, 0.2, 30.0, 10.0, 0.6, shear_units=True)
    for name, _ in a._fields_:
        assert getattr(a, name) == getattr(b, name), name
    a = engine.liquid(1000.0, 1e5, 0.2, 1e-3)
    b = O.liquid(1000.0, 1e5, 0.2, 1e-3)
    for name, _ in a._fields_:
        assert getattr(a, name) == getattr(b, name), name


def test_pellet_constructor_matches_oracle(engine):
    a = engine.new_pellet(0.0015, 0.008, 650.0, [0.01, 0.02, 0.03])
    b = O.new_pellet(0.0015, 0.008, 650.0, [0.01, 0.02, 0.03])
    assert a.mass == b.mass and a.sphere_count == b.sphere_count == 5
    assert list(a.inertia_body) == list(b.inertia_body)
    assert list(a.sphere_offsets[:5]) == list(b.sphere_offsets[:5])


def test_struct_layouts_match_header(engine):
    import ctypes as C
    # Sizes follow from include/mpm_engine.h field lists (all doubles except noted ints).
    assert C.sizeof(engine.Ledger) == 8 * 31
    assert C.sizeof(engine.History) == 8 * 7
    assert C.sizeof(engine.Config) == 8 * 16
    assert C.sizeof(engine.Material) == 8 * 9
