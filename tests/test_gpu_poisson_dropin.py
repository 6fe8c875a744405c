"""The drop-in PoissonSolver on the device.

scdat_amd/facade/PoissonSolver_device.cpp defines the members of the
reference's SCDAT::FieldSolver::PoissonSolver (its unchanged header) over the
C ABI.  oracle/ref/Makefile links it, in place of the reference's
PoissonSolver.cpp, with the reference's geometry and mesh sources and the
same harness (oracle/ref/ref_harness_poisson.cpp) that libscdat_ref.so links
with the reference's own solver.  Both libraries are called on the same
inputs: the graph-Laplacian fixture made by the reference (direct solve),
the reference's plate test (PoissonSolver_test.cpp:48-117) and a second
assembly with new charges, Dirichlet values and Neumann fluxes.  The device
solves to a relative residual of 1e-12 where the reference eliminates
directly, so potentials agree to 1e-10 of their largest value and fields to
1e-9 of theirs.  Both libraries are built in this repository from the
reference's sources (skipped where they were not built)."""
import ctypes as C
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libscdat_ref.so")
DROPIN_LIB = os.path.join(ROOT, "oracle", "_ref", "libscdat_poisson_dropin.so")
GOLDEN = os.path.join(ROOT, "tests", "golden", "poisson_graph_n6.npz")

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not (os.path.exists(REF_LIB) and os.path.exists(DROPIN_LIB)),
                                 reason="reference-side libraries not built (oracle/ref/Makefile)")]

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int32)


def _lib(path):
    L = C.CDLL(path)
    L.ref_poisson.argtypes = [C.c_int, _D, C.c_int, _I, _I, _D, _D, _D, _D]
    L.ref_poisson_twice.argtypes = [C.c_int, _D, C.c_int, _I, _I, _D, _D, _D, _D, _D, _D, _D]
    L.ref_poisson_plate.argtypes = [C.c_int, C.POINTER(C.c_int), _D, _D, _D, _D]
    return L


def _p(a, t=_D):
    return a.ctypes.data_as(t)


def _poisson(L, g, dens):
    nn = len(g["xyz"])
    xyz = np.ascontiguousarray(g["xyz"], np.float64)
    tets = np.ascontiguousarray(g["tets"], np.int32)
    kind = np.ascontiguousarray(g["kind"], np.int32)
    dval = np.ascontiguousarray(g["dval"], np.float64)
    phi = np.zeros(nn)
    E = np.zeros(3 * nn)
    rc = L.ref_poisson(nn, _p(xyz), len(tets), _p(tets, _I), _p(kind, _I), _p(dval), _p(dens), _p(phi), _p(E))
    return rc, phi, E.reshape(-1, 3)


def _close(a, b, rel):
    scale = max(np.max(np.abs(b)), 1e-300)
    return np.max(np.abs(a - b)) <= rel * scale


def test_graph_laplacian_fixture_matches_reference_direct_solve():
    g = np.load(GOLDEN)
    rc, phi, E = _poisson(_lib(DROPIN_LIB), g, np.ascontiguousarray(g["density"], np.float64))
    assert rc == 0
    assert _close(phi, g["phi"], 1e-10)
    assert _close(E, g["E"], 1e-9)
    cond = g["kind"] == 1
    assert _close(phi[cond], g["dval"][cond], 1e-13)  # Dirichlet nodes hold their values


def test_second_assembly_with_new_charge_dirichlet_and_neumann_matches_reference():
    g = np.load(GOLDEN)
    nn = len(g["xyz"])
    rng = np.random.default_rng(17)
    kind = np.ascontiguousarray(g["kind"], np.int32)
    dens2 = rng.normal(size=nn) * 2e-9
    dval2 = np.where(kind == 1, rng.normal(size=nn), 0.0)
    flux2 = np.where((kind == 0) & (rng.random(nn) < 0.1), rng.normal(size=nn) * 1e-10, 0.0)
    out = []
    for path in (REF_LIB, DROPIN_LIB):
        L = _lib(path)
        phi = np.zeros(nn)
        E = np.zeros(3 * nn)
        args = [np.ascontiguousarray(a, np.float64) for a in (g["xyz"], g["dval"], g["density"], dval2, dens2, flux2)]
        tets = np.ascontiguousarray(g["tets"], np.int32)
        rc = L.ref_poisson_twice(nn, _p(args[0]), len(tets), _p(tets, _I), _p(kind, _I), _p(args[1]), _p(args[2]),
                                 _p(args[3]), _p(args[4]), _p(args[5]), _p(phi), _p(E))
        assert rc == 0, path
        out.append((phi, E.reshape(-1, 3)))
    (phi_r, E_r), (phi_d, E_d) = out
    assert flux2.any()
    assert _close(phi_d, phi_r, 1e-10)
    assert _close(E_d, E_r, 1e-9)
    assert not _close(phi_r, g["phi"], 1e-3)  # the second solve saw the new inputs


def _plate(path):
    L = _lib(path)
    cap = 1000
    n = C.c_int(0)
    z, phi, E, tot = np.zeros(cap), np.zeros(cap), np.zeros(3 * cap), np.zeros(3)
    rc = L.ref_poisson_plate(cap, C.byref(n), _p(z), _p(phi), _p(E), _p(tot))
    k = n.value
    return rc, z[:k], phi[:k], E[:3 * k].reshape(-1, 3), tot


def test_reference_plate_test_on_the_device():
    rc, z, phi, E, tot = _plate(DROPIN_LIB)
    assert rc == 0 and len(z) > 0
    # the reference test's own expectations (PoissonSolver_test.cpp:84-116)
    lo, hi = z.min(), z.max()
    bottom, top = np.abs(z - lo) < 1e-12, np.abs(z - hi) < 1e-12
    assert np.all(np.abs(phi[bottom] - 10.0) < 1e-8) and np.all(np.abs(phi[top]) < 1e-8)
    inner = ~(bottom | top)
    assert inner.any() and np.all(phi[inner] >= -1e-8) and np.all(phi[inner] <= 10.0 + 1e-8)
    assert np.all(np.isfinite(E)) and np.max(np.linalg.norm(E, axis=1)) > 0.0
    assert tot[0] >= 0.0 and tot[1] >= tot[2]
    # and the reference's own numbers on the same mesh
    rc_r, z_r, phi_r, E_r, tot_r = _plate(REF_LIB)
    assert rc_r == 0
    assert np.array_equal(z, z_r)
    assert _close(phi, phi_r, 1e-10)
    assert _close(E, E_r, 1e-9)
    assert abs(tot[0] - tot_r[0]) <= 1e-8 * tot_r[0]
    assert abs(tot[1] - tot_r[1]) <= 1e-12 * tot_r[1] and abs(tot[2] - tot_r[2]) <= 1e-10 * tot_r[1]
