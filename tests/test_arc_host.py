"""Field-emission cathode and the vacuum-arc PIC window, host side (no GPU):
the oracle's restatement of FieldEmissionBoundaryCondition::processParticle
against the reference's own test and closed forms, the window's probe
seeding and plate mesh, and the parameter defaults of the C ABI.

The reference's FieldEmissionBoundaryCondition.cpp needs the particle
classes of ParticleDefinitions.cpp, which do not build with this image's
standard library, so it is not compiled here; its unit test's numbers are
the golden case."""
import math

import numpy as np
import pytest

import oracle
import scdat_amd as S
from scdat_amd import arc

QE = S.ELEMENTARY_CHARGE


def _params(**kw):
    p = dict(emission_current_density_a_per_m2=0.0, emission_area_m2=1.0e-8, emitted_electron_speed_m_per_s=2.0e6,
             max_emitted_particles_per_event=64, enable_secondary_electron_emission=False,
             secondary_electron_yield_per_ion=0.0, secondary_electron_speed_m_per_s=8.0e5,
             secondary_electron_weight=1.0, max_secondary_particles_per_event=64)
    p.update(kw)
    return p


def test_oracle_sub_particle_budget_matches_reference_test():
    # VacuumArc_presets_smoke_test.cpp:197-244: 0.25 expected electrons per
    # impact; three impacts emit nothing, the fourth emits one field electron
    dt = 1.0e-9
    p = _params(emission_area_m2=1.0e-8, emitted_electron_speed_m_per_s=1.0e6, max_emitted_particles_per_event=16)
    p["emission_current_density_a_per_m2"] = 0.25 * QE / (p["emission_area_m2"] * dt)
    pending = (0.0, 0.0)
    total = 0
    for _ in range(3):
        out, pending = oracle.field_emission(p, [((0, 0, 0), (0, 0, 1), 1.0)], dt, pending)
        assert out == []
    out, pending = oracle.field_emission(p, [((0, 0, 0), (0, 0, 1), 1.0)], dt, pending)
    total += len(out)
    assert total == 1
    eid, x, v, kind = out[0]
    assert kind == "field"
    np.testing.assert_array_equal(v, [0, 0, 1.0e6])
    np.testing.assert_allclose(x, [0, 0, 1.0e-10], rtol=0, atol=1e-25)


def test_oracle_counts_conserve_the_budget():
    # every impact adds `expect`; what is not emitted stays pending, in [0, 1)
    # while the cap does not bind
    dt = 2.0e-10
    p = _params(emission_current_density_a_per_m2=0.474, emission_area_m2=4.0e-9, max_emitted_particles_per_event=1000)
    expect = p["emission_current_density_a_per_m2"] * p["emission_area_m2"] * dt / QE
    rng = np.random.default_rng(3)
    hits = [(rng.uniform(size=3), (0, 0, 1), 1.0) for _ in range(57)]
    out, (pf, _) = oracle.field_emission(p, hits, dt)
    assert 0.0 <= pf < 1.0
    assert abs(len(out) + pf - 57 * expect) < 1e-9 * 57 * expect
    assert len(out) == math.floor(57 * expect + 1e-9)
    assert 2.0 < expect < 3.0
    # a binding cap: every impact emits the cap (2) and the rest waits
    p["max_emitted_particles_per_event"] = 2
    out2, (pf2, _) = oracle.field_emission(p, hits, dt)
    assert len(out2) == 2 * 57
    assert abs(pf2 - (57 * expect - 2 * 57)) < 1e-9 * 57 * expect


def test_oracle_secondaries_and_offsets():
    # yield 0.75 per ion of weight 1: pending 0.75, 1.5 -> 1, 1.25 -> 1, 1.0 -> 1
    p = _params(enable_secondary_electron_emission=True, secondary_electron_yield_per_ion=0.75,
                secondary_electron_speed_m_per_s=5.0e5, max_secondary_particles_per_event=64)
    hits = [((1.0, 2.0, 0.0), (0, 0, 2.0), 1.0)] * 4
    out, (_, ps) = oracle.field_emission(p, hits, 1e-9)
    assert [k for _, _, _, k in out] == ["secondary"] * 3
    assert ps == 0.0
    for eid, x, v, _ in out:
        np.testing.assert_array_equal(v, [0, 0, 5.0e5])  # unit normal x speed
        np.testing.assert_allclose(x, [1.0, 2.0, 1.0e-10], atol=1e-25)
    assert [e[0] for e in out] == [0, 1, 2]
    # field then secondary within one impact share the offset sequence
    p.update(emission_current_density_a_per_m2=3.0 * QE / (1e-8 * 1e-9), secondary_electron_yield_per_ion=2.0)
    out, _ = oracle.field_emission(p, [((0, 0, 0), (3.0, 0, 0), 1.0)], 1e-9)
    kinds = [k for _, _, _, k in out]
    assert kinds.count("secondary") == 2 and kinds[-2:] == ["secondary"] * 2
    offs = [x[0] for _, x, _, _ in out]
    np.testing.assert_allclose(offs, 1.0e-10 * np.arange(1, len(out) + 1), rtol=1e-12)
    # a zero normal emits along +z
    out, _ = oracle.field_emission(p, [((0, 0, 0), (0, 0, 0), 1.0)], 1e-9)
    assert all(v[2] > 0 and v[0] == 0 and v[1] == 0 for _, _, v, _ in out)


def test_probe_count_closed_forms():
    # SurfaceDischargeArcAlgorithm.cpp:541-549
    area, dt = 1.6e-9, 1.0e-10
    per = area * dt / QE  # electrons per unit current density
    assert arc.probe_count(0.0, area, dt) == 1
    assert arc.probe_count(-5.0, area, dt) == 1
    assert arc.probe_count(400.0 / per, area, dt) == 1        # 0.8 -> 1
    assert arc.probe_count(1250.0 / per, area, dt) == 3       # 2.5 -> 3
    assert arc.probe_count(1.0e9 / per, area, dt) == 4        # capped
    assert arc.probe_count(1.0, area, 0.0) == 1               # dt floored at 1e-12


def test_plate_mesh_and_top_element():
    cfg = arc.WindowConfig()
    L, W, T = arc.plate_dimensions(cfg)
    assert (L, W, T) == (1.0e-4, 1.0e-4, 5.0e-4)  # max(2 r, 1e-4), max(gap, 1e-4)
    m = arc.plate_mesh(L, W, T)
    z = m["xyz"][:, 2]
    on = (np.abs(z) < 1e-12) | (np.abs(z - T) < 1e-12)
    np.testing.assert_array_equal(m["kind"] == S.NODE_DIRICHLET, on)
    assert len(m["tets"]) == 6 * 6
    t = arc.top_tet(m)
    cz = m["xyz"][m["tets"], 2].mean(axis=1)
    assert cz[t] == cz.max() and np.all(cz[:t] < cz[t])
    # the z faces carry groups 4 (cathode) and 5 (anode)
    for g, z0 in ((4, 0.0), (5, T)):
        ft = np.argwhere(m["group"] == g)
        assert len(ft) > 0
        for tt, f in ft:
            face = [m["tets"][tt][k] for k in range(4) if k != f]
            assert np.allclose(m["xyz"][face, 2], z0)


def test_reference_five_tet_plate_is_not_conforming():
    # MeshParsing.cpp:1248-1254 split of a 1 x 1 x 2 plate: the faces between
    # the two layers do not pair up, which is why the window uses a
    # conforming split of the same hexahedra
    def idx(i, j, k):
        return i * 2 * 3 + j * 3 + k

    tets = []
    for k in range(2):
        v = {(a, b, c): idx(a, b, k + c) for a in (0, 1) for b in (0, 1) for c in (0, 1)}
        tets += [(v[0, 0, 0], v[1, 0, 0], v[0, 1, 0], v[0, 0, 1]), (v[1, 0, 0], v[1, 1, 0], v[0, 1, 0], v[1, 1, 1]),
                 (v[1, 0, 0], v[0, 1, 0], v[0, 0, 1], v[1, 1, 1]), (v[1, 0, 0], v[0, 0, 1], v[1, 0, 1], v[1, 1, 1]),
                 (v[0, 1, 0], v[0, 0, 1], v[0, 1, 1], v[1, 1, 1])]
    faces = {}
    for t in tets:
        for f in range(4):
            key = tuple(sorted(t[:f] + t[f + 1:]))
            faces[key] = faces.get(key, 0) + 1
    mid = [k for k in faces if all(n % 3 == 1 for n in k)]  # faces on the plane z = dz
    assert len(mid) == 4 and all(faces[k] == 1 for k in mid)


def test_field_emission_defaults_match_reference():
    # FieldEmissionBoundaryCondition.h:17-31
    p = S.field_emission()
    assert p.emission_current_density_a_per_m2 == 0.0
    assert p.emission_area_m2 == 1.0e-8
    assert p.emitted_electron_speed_m_per_s == 2.0e6
    assert p.emitted_electron_weight == 1.0
    assert p.max_emitted_particles_per_event == 64
    assert p.enable_secondary_electron_emission == 0
    assert p.secondary_electron_yield_per_ion == 0.0
    assert p.secondary_electron_speed_m_per_s == 8.0e5
    assert p.secondary_electron_weight == 1.0
    assert p.max_secondary_particles_per_event == 64
