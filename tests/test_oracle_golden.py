"""Pin the FP64 oracle to golden vectors produced by the compiled reference
(tests/golden/gen_golden.py runs the reference's own routines).

Each test feeds the oracle the fixture's inputs and compares with the
reference's outputs; tolerances are FP64 round-off unless stated.
"""
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
EPS0 = 8.8541878128e-12


def load(name):
    return dict(np.load(os.path.join(GOLD, name)))


def node_share(o, tets, nn):
    """sum over elements of V_e/4 per node (the reference's lumped load)."""
    _, _, vol = o.geometry()
    share = np.zeros(nn)
    np.add.at(share, tets.ravel(), np.repeat(vol / 4, 4))
    return share


def test_barycentric_matches_reference():
    g = load("bary_kuhn4.npz")
    o = O.Oracle(g["xyz"], g["tets"], np.zeros(len(g["xyz"]), np.int32))
    lam = np.array([o.barycentric(int(c), p) for c, p in zip(g["cells"], g["pts"])])
    assert np.max(np.abs(lam - g["lam"])) < 1e-13
    assert g["inside"].all()
    assert np.allclose(lam.sum(1), 1.0, atol=1e-14)


def test_deposit_matches_reference():
    g = load("deposit_kuhn4.npz")
    o = O.Oracle(g["xyz"], g["tets"], np.zeros(len(g["xyz"]), np.int32))
    s = o.add_species(float(g["charge"]), 1.0, float(g["weight"]))
    o.set_particles(s, g["pts"], np.zeros_like(g["pts"]), g["cells"])
    o.deposit()
    rho = o.fields()[0]
    scale = np.max(np.abs(g["rho"]))
    assert np.max(np.abs(rho - g["rho"])) < 1e-12 * scale
    # charge conservation: total deposited = N q w
    assert abs(rho.sum() - len(g["pts"]) * g["charge"] * g["weight"]) < 1e-12 * abs(rho.sum())


@pytest.mark.parametrize("name,mode", [("cross_nearest.npz", 2), ("cross_segcloud.npz", 1),
                                       ("cross_single_substep.npz", 0)])
def test_trajectory_matches_reference(name, mode):
    """One step of CrossTetrahedron::crossTetrahedron: Boris, adaptive
    sub-steps, relocation, loss on leaving the mesh, deposition."""
    g = load(name)
    o = O.Oracle(g["xyz"], g["tets"], g["kind"], group=g["group"])
    o.set_poisson(0, EPS0, 1, 1e-10, 100)
    o.set_push(float(g["dt"]), 1, 1e-3, 100000, mode, 10)
    s = o.add_species(float(g["charge"]), float(g["mass"]), float(g["weight"]))
    st = g["state"]
    o.set_particles(s, st[:, :3], st[:, 3:], g["cells"])
    o.set_efield(g["E"])
    o.push()
    x, v, cell = o.get_particles(s)
    dead = cell < 0
    assert np.array_equal(dead, g["lost"] == 1), "loss events differ"
    alive = ~dead
    ref = g["out"]
    assert alive.sum() > 100
    assert np.max(np.abs(x[alive] - ref[alive, :3])) < 1e-12
    vs = np.max(np.abs(ref[alive, 3:]))
    assert np.max(np.abs(v[alive] - ref[alive, 3:])) < 1e-12 * vs
    rho = o.fields()[0]
    scale = np.max(np.abs(g["rho"]))
    assert np.max(np.abs(rho - g["rho"])) < 1e-9 * scale


def test_uniform_e_b_boris_rotation_matches_reference():
    g = load("cross_uniform_eb.npz")
    nn = len(g["xyz"])
    o = O.Oracle(g["xyz"], g["tets"], g["kind"], group=g["group"])
    o.set_poisson(0, EPS0, 1, 1e-10, 100)
    o.set_push(float(g["dt"]), 1, 1e-3, 100000, 2, 10)
    s = o.add_species(float(g["charge"]), float(g["mass"]), float(g["weight"]))
    st = g["state"]
    o.set_particles(s, st[:, :3], st[:, 3:], g["cells"])
    o.set_efield(np.tile(g["uE"], (nn, 1)))
    o.set_bfield(np.tile(g["uB"], (nn, 1)))
    o.push()
    x, v, cell = o.get_particles(s)
    dead = cell < 0
    assert np.array_equal(dead, g["lost"] == 1)
    alive = ~dead
    ref = g["out"]
    assert alive.sum() > 50
    assert np.max(np.abs(x[alive] - ref[alive, :3])) < 1e-12
    assert np.max(np.abs(v[alive] - ref[alive, 3:])) < 1e-12 * np.max(np.abs(ref[:, 3:]))


@pytest.mark.parametrize("n", [4, 6])
def test_fem_floating_conductor_matches_reference(n):
    """P1 FEM + Dirichlet + FLOATING node collapse vs
    MultiBoundaryElectricFieldSolver (direct).  The reference's load is
    -rho V_e/4 per element node; the oracle takes the lumped nodal charge
    with load_sign = -1."""
    g = load(f"mb_floating_n{n}.npz")
    nn = len(g["xyz"])
    o = O.Oracle(g["xyz"], g["tets"], g["kind"], dirichlet=g["dval"], group=g["group"])
    o.set_poisson(0, float(g["permittivity"]), -1, 1e-14, 100000)
    o.set_rho(g["density"] * node_share(o, g["tets"], nn))
    o.set_conductor_charge(float(g["q_float"]))
    it, rel, ok = o.solve()
    assert ok
    phi = o.fields()[1]
    assert np.max(np.abs(phi - g["phi"])) < 1e-12 * np.max(np.abs(g["phi"]))
    cond = g["kind"] == 2
    assert np.ptp(phi[cond]) == 0.0  # the conductor is one equipotential
    assert abs(o.floating_potential() - g["phi"][cond][0]) < 1e-12 * abs(g["phi"][cond][0])


def test_graph_laplacian_and_lsq_field_match_reference():
    g = load("poisson_graph_n6.npz")
    nn = len(g["xyz"])
    o = O.Oracle(g["xyz"], g["tets"], g["kind"], dirichlet=g["dval"], group=g["group"])
    o.set_poisson(1, EPS0, -1, 1e-14, 100000)
    o.set_rho(g["density"] * node_share(o, g["tets"], nn))
    it, rel, ok = o.solve()
    assert ok
    o.efield()
    _, phi, E = o.fields()
    assert np.max(np.abs(phi - g["phi"])) < 1e-12 * np.max(np.abs(g["phi"]))
    assert np.max(np.abs(E - g["E"])) < 1e-11 * np.max(np.abs(g["E"]))
