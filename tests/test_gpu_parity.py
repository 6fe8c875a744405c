"""Parity of the gfx950 path against the FP64 oracle, through the C ABI."""
import numpy as np
import pytest

from helpers import make_ctx, make_oracle, plasma_setup, rel_err, smooth_field

pytestmark = pytest.mark.gpu


def test_device_present(S):
    name, ncu = S.device_info()
    print("device:", name, ncu)
    assert "gfx950" in name


@pytest.mark.parametrize("precision", [0, 1])
def test_push_trajectories_100_steps(S, O, precision):
    """Collisionless trajectories in a prescribed field agree with the FP64
    oracle to 1e-5 relative over 100 steps."""
    st = plasma_setup(n=8, npart=5000, seed=3)
    m = st["mesh"]
    # weak field: the electron kick per step stays near 1% of v_th
    E = smooth_field(m["xyz"], amp=0.2)
    o = make_oracle(st)
    o.set_efield(E)
    c = make_ctx(st, precision=precision, sort_interval=0)
    c.initialize()
    c.set_efield(E)
    for _ in range(100):
        o.push()
        c.push()
    for s in range(2):
        xo, vo, co = o.get_particles(s)
        X, V, Cc = c.get_particles_by_id(s, len(xo))
        alive = (co >= 0) & (Cc >= 0)
        # A particle that ends a step within FP32 resolution of an absorbing
        # face can be absorbed one step apart on the two paths; allow at most
        # 0.1% such grazing cases and compare everything else.
        mismatch = int(np.sum((co >= 0) != (Cc >= 0)))
        print(f"species {s}: absorption mismatches {mismatch} of {len(co)}")
        assert mismatch <= max(2, len(co) // 1000)
        ex = np.max(np.linalg.norm(X[alive] - xo[alive], axis=1) / np.linalg.norm(xo[alive], axis=1))
        # velocity error relative to max(|v|, species RMS speed): a particle that
        # is momentarily near rest has no meaningful own-scale relative error
        vs = np.linalg.norm(vo[alive], axis=1)
        vscale = np.maximum(vs, np.sqrt(np.mean(vs ** 2)))
        ev = np.max(np.linalg.norm(V[alive] - vo[alive], axis=1) / vscale)
        print(f"species {s} precision {precision}: pos rel {ex:.2e} vel rel {ev:.2e} alive {alive.sum()}")
        assert alive.sum() > 50
        assert ex < 1e-5 and ev < 1e-5


def test_deposit_bit_exact(S, O):
    """Device fixed-point deposition equals the integer checker bit for bit."""
    st = plasma_setup(n=8, npart=5000, seed=5)
    c = make_ctx(st, precision=0, sort_interval=0)
    c.initialize()
    c.set_efield(smooth_field(st["mesh"]["xyz"]))
    c.push()
    rho_dev, unit = c.rho_fixed()
    maps = c.tet_maps()
    ref = np.zeros(c.n_nodes, np.int64)
    for s in range(2):
        x, v, cell, ids = c.get_particles(s)
        O.deposit_fixed_f32(maps, st["mesh"]["tets"], x.astype(np.float32), cell,
                            c.species_charge_fixed(s), c.n_nodes, ref)
    assert np.array_equal(rho_dev, ref)


@pytest.mark.parametrize("precision,tol", [(1, 1e-6), (0, 1e-4)])
def test_poisson_floating_conductor(S, O, precision, tol):
    """Potential with a floating conductor matches the FP64 oracle's node
    collapse formulation."""
    st = plasma_setup(n=10, npart=4000, seed=9)
    o = make_oracle(st, tol=1e-13)
    o.deposit()
    o.set_conductor_charge(3e-13)
    o.solve()
    rho_o, phi_o, _ = o.fields()
    c = make_ctx(st, precision=precision, tol=1e-10 if precision else 1e-8)
    c.set_conductor_charge(3e-13)
    c.set_rho(rho_o)
    c.solve()
    _, phi_c, _ = c.fields()
    e = rel_err(phi_c, phi_o)
    print("phi rel err", e, "vfloat", o.floating_potential(), c.stats()["floating_potential"])
    assert e < tol
