"""Lattice location in the FP32 push (SCDAT_SHAPE_LATTICE_PUSH) against the
tet-face walk, through the C ABI.  The lattice locates only end points at
least 1e-3 (barycentric) inside a tet, clear of boundary faces, where the
walk reaches the same tet; so from the same state a push with each gives
the same end tets, positions, velocities, ledger and fixed-point deposit
(which also equals the oracle's from the pushed particles), and a
full-cycle run gives the same fields.  The field is strong enough that
electrons take several adaptive sub-steps."""
import numpy as np
import pytest

import oracle
import scdat_amd as S
from helpers import make_ctx, plasma_setup, smooth_field

pytestmark = pytest.mark.gpu


def _ctx(st, lattice, precision=0, sort_interval=0):
    c = make_ctx(st, precision=precision, tol=1e-8, sort_interval=sort_interval)
    c.lattice_push = lattice
    c.initialize()
    return c


def test_lattice_push_is_used_on_lattice_meshes_only():
    st = plasma_setup(n=10, npart=2000, seed=3)
    c = _ctx(st, 1)
    assert c.lattice_push == 1
    c.lattice_push = 0
    assert c.lattice_push == 0
    with pytest.raises(S.ScdatError):
        c.lattice_push = 2
    c.close()
    c64 = _ctx(st, 1, precision=1)
    assert c64.lattice_push == 0  # FP64 walks
    c64.close()
    # interior nodes moved off the grid lines: not a lattice
    m = st["mesh"]
    xyz = m["xyz"].copy()
    rng = np.random.default_rng(1)
    inner = m["kind"] == 0
    xyz[inner] += rng.uniform(-1e-3, 1e-3, size=(int(inner.sum()), 3))
    c2 = S.Context(precision=0, dt=st["dt"], pcg_tol=1e-8)
    c2.set_mesh(xyz, m["tets"], m["kind"], group=m["group"])
    assert c2.lattice_push == 0
    c2.close()


def _by_id(p):
    x, v, cell, ids = p
    o = np.argsort(ids, kind="stable")
    return x[o], v[o], cell[o], ids[o]


def _one_push(st, lattice, E):
    c = _ctx(st, lattice)
    c.set_efield(E)
    c.push()
    # slots: the initial cell sort orders a cell's particles by atomic arrival
    out = dict(parts=[_by_id(c.get_particles(s)) for s in (0, 1)], ledger=c.ledger(), rho=c.rho_fixed()[0],
               maps=c.tet_maps(), cs=[c.species_charge_fixed(s) for s in (0, 1)], n_nodes=c.n_nodes)
    c.close()
    return out


@pytest.mark.parametrize("seed", [5, 6])
def test_lattice_push_equals_walk(seed):
    st = plasma_setup(n=16, npart=60000, seed=seed)
    E = smooth_field(st["mesh"]["xyz"])
    lat = _one_push(st, 1, E)
    wlk = _one_push(st, 0, E)
    tets = st["mesh"]["tets"]
    for s in (0, 1):
        for ql, qw in zip(lat["parts"][s], wlk["parts"][s]):
            np.testing.assert_array_equal(ql, qw)
    np.testing.assert_array_equal(lat["rho"], wlk["rho"])
    for la, lw in zip(lat["ledger"], wlk["ledger"]):
        # the incident energy is an FP64 sum of atomics: equal to rounding
        assert la["incident_energy_j"] == pytest.approx(lw["incident_energy_j"], rel=1e-12, abs=0.0)
        assert {k: v for k, v in la.items() if k != "incident_energy_j"} == \
            {k: v for k, v in lw.items() if k != "incident_energy_j"}
    # deposit of the lattice run equals the oracle's from its particles
    ref = np.zeros(lat["n_nodes"], np.int64)
    for s in (0, 1):
        x, _, cell, _ = lat["parts"][s]
        oracle.deposit_fixed_f32(lat["maps"], tets, x.astype(np.float32), cell, lat["cs"][s], lat["n_nodes"], ref)
    np.testing.assert_array_equal(lat["rho"], ref)


def test_lattice_cycle_equals_walk_cycle():
    st = plasma_setup(n=16, npart=40000, seed=9)
    res = []
    for lattice in (1, 0):
        c = _ctx(st, lattice, sort_interval=5)
        c.step(30)
        _, phi, E = c.fields()
        res.append((phi, E, c.stats(), c.ledger()))
        c.close()
    (pl, El, sl, ll), (pw, Ew, sw, lw) = res
    np.testing.assert_array_equal(pl, pw)
    np.testing.assert_array_equal(El, Ew)
    assert sl["floating_potential"] == sw["floating_potential"]
    for g in range(12):
        assert ll[g]["absorbed_electron_particles"] == lw[g]["absorbed_electron_particles"]
        assert ll[g]["absorbed_ion_particles"] == lw[g]["absorbed_ion_particles"]
