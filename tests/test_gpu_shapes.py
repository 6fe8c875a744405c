"""Launch shapes of the push kernel: every shape gives the same bits, and the
tuner picks one without touching the simulation state."""
import numpy as np
import pytest

import scdat_amd as S
from helpers import make_ctx, plasma_setup

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [0, 1])
def test_push_shapes_give_identical_results(precision):
    st = plasma_setup(n=10, npart=20000, seed=5)
    out = []
    for chunks in (1, 2, 4, 8):
        c = make_ctx(st, precision=precision, tol=1e-10, sort_interval=2)
        c.push_chunks = chunks
        assert c.push_chunks == chunks
        c.initialize()
        c.step(5)
        out.append((c.fields(), c.get_particles_by_id(0, 20000), c.stats()))
    for f, p, s in out[1:]:
        for a, b in zip(f, out[0][0]):
            np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(p[0], out[0][1][0])
        assert s["floating_potential"] == out[0][2]["floating_potential"]
        assert s["active_particles"] == out[0][2]["active_particles"]


@pytest.mark.parametrize("precision", [0, 1])
def test_deposit_buckets_give_identical_results(precision):
    """Push blocks handing their charge tables to per-range buckets give the
    charge of one global atomic per node, bit for bit."""
    st = plasma_setup(n=12, npart=30000, seed=7)
    out = []
    for mode in (1, 0):
        c = make_ctx(st, precision=precision, tol=1e-10, sort_interval=3)
        c.deposit_buckets = mode
        assert c.deposit_buckets == mode
        c.initialize()
        c.push()
        rho1 = c.rho_fixed()[0]
        c.step(6)
        out.append((rho1, c.rho_fixed()[0], c.fields(), c.get_particles_by_id(0, 30000), c.stats()))
    a, b = out
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])
    for x, y in zip(a[2], b[2]):
        np.testing.assert_array_equal(x, y)
    np.testing.assert_array_equal(a[3][0], b[3][0])
    assert a[4]["floating_potential"] == b[4]["floating_potential"]
    with pytest.raises(S.ScdatError):
        c.deposit_buckets = 2


def test_autotune_keeps_state():
    st = plasma_setup(n=10, npart=20000, seed=6)
    c = make_ctx(st, precision=0, tol=1e-10, sort_interval=0)
    c.initialize()
    c.step(2)
    before = (c.fields(), c.get_particles(0), c.stats(), c.ledger())
    k = c.autotune()
    assert k in (1, 2, 4, 8) and c.push_chunks == k
    after = (c.fields(), c.get_particles(0), c.stats(), c.ledger())
    for a, b in zip(before[0], after[0]):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(before[1], after[1]):
        np.testing.assert_array_equal(a, b)
    assert before[2] == after[2] and before[3] == after[3]
    with pytest.raises(S.ScdatError):
        c.push_chunks = 3


def test_row_patterns_give_identical_results():
    """The FP32 inner iteration reading row patterns (2- and 4-byte ids)
    gives the ELL form's bits."""
    st = plasma_setup(n=12, npart=20000, seed=8)
    out = []
    for mode in (1, 2, 0):
        c = make_ctx(st, precision=0, tol=1e-10, sort_interval=2)
        c.row_patterns = mode
        assert c.row_patterns == mode
        c.initialize()
        info = c.solver_info()
        if mode:
            assert 0 < info["row_patterns"] < info["rows"]
        else:
            assert info["row_patterns"] == 0
        c.step(6)
        out.append((c.fields(), c.get_particles_by_id(0, 20000), c.stats()))
        c.close()
    f0, p0, s0 = out[-1]
    for f1, p1, s1 in out[:-1]:
        for a, b in zip(f0, f1):
            np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(p0[0], p1[0])
        assert s0["floating_potential"] == s1["floating_potential"]
        assert s0["pcg_iterations"] == s1["pcg_iterations"]


def test_pcg_block_sizes_agree_and_overrides_are_reported():
    """256- and 512-row PCG blocks solve the same system to the tolerance
    (their reductions differ in order, so the bits may not); other sizes are
    refused; the two-level preconditioner reports the 512 it runs on."""
    st = plasma_setup(n=12, npart=20000, seed=8)
    res = {}
    for pb in (256, 512):
        c = make_ctx(st, precision=0, tol=1e-10, sort_interval=2)
        c.pcg_block = pb
        assert c.pcg_block == pb
        c.initialize()
        assert c.solver_info()["pcg_block"] == pb
        c.step(4)
        res[pb] = (c.fields(), c.stats())
        c.close()
    (_, phi_a, E_a), sa = res[256]
    (_, phi_b, E_b), sb = res[512]
    scale = np.abs(phi_a).max()
    assert np.abs(phi_a - phi_b).max() <= 1e-6 * scale
    assert sa["floating_potential"] == pytest.approx(sb["floating_potential"], rel=1e-6, abs=1e-9)
    assert sa["active_particles"] == sb["active_particles"]
    c = make_ctx(st, precision=0, tol=1e-10, sort_interval=2)
    with pytest.raises(S.ScdatError):
        c.pcg_block = 384
    c.close()
    st2 = plasma_setup(n=18, npart=8000, seed=9)
    c = make_ctx(st2, precision=0, tol=1e-10, sort_interval=2, order=S.ORDER_TILED)
    c.pcg_block = 256
    c.set_preconditioner(S.PRECOND_TWO_LEVEL)
    c.initialize()
    info = c.solver_info()
    assert info["coarse_unknowns"] > 0 and info["pcg_block"] == 512
    c.step(2)
    assert c.stats()["pcg_converged"] == 1
    c.close()
