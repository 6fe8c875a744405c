"""Two-level Poisson preconditioner (row-block aggregates of the Morton or
tiled node order; in mixed precision with the tiled order the coarse step
runs inside the iteration kernel) through the C ABI: the same fields as Jacobi-PCG and as the oracle's
FP64 cycle, fewer iterations, bitwise reproducible solves; node arrays come
back in the caller's order under the internal renumbering."""
import numpy as np
import pytest

import scdat_amd as S
from helpers import make_ctx, make_oracle, plasma_setup, rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("order", [S.ORDER_MORTON, S.ORDER_TILED])
@pytest.mark.parametrize("precision", [1, 0])
def test_two_level_matches_jacobi_and_needs_fewer_iterations(precision, order):
    st = plasma_setup(n=24, npart=40000, seed=8)
    res = {}
    for pc in (S.PRECOND_JACOBI, S.PRECOND_TWO_LEVEL):
        c = make_ctx(st, precision=precision, tol=1e-10, sort_interval=0, order=order)
        c.set_preconditioner(pc)
        c.initialize()
        info = c.solver_info()
        its = []
        for _ in range(3):
            c.step(1)
            its.append(c.stats()["pcg_iterations"])
        res[pc] = (c.fields(), c.stats(), info, its)
    jac, two = res[S.PRECOND_JACOBI], res[S.PRECOND_TWO_LEVEL]
    print("coarse", two[2], "iterations jacobi", jac[3], "two-level", two[3])
    assert jac[2]["preconditioner"] == S.PRECOND_JACOBI
    assert two[2]["preconditioner"] == S.PRECOND_TWO_LEVEL and two[2]["coarse_unknowns"] >= 8
    # mixed precision in the tiled order: the coarse step inside the iteration kernel
    assert two[2]["coarse_in_kernel"] == int(precision == 0 and order == S.ORDER_TILED)
    assert rel_err(two[0][1], jac[0][1]) < 1e-7
    assert two[1]["floating_potential"] == pytest.approx(jac[1]["floating_potential"], rel=1e-7)
    assert all(c == 1 for c in (two[1]["pcg_converged"], jac[1]["pcg_converged"]))
    assert sum(two[3]) < 0.85 * sum(jac[3])  # 27 aggregates at this size


def test_two_level_full_cycle_fp64_matches_oracle_and_is_reproducible():
    st = plasma_setup(n=20, npart=20000, seed=23)
    o = make_oracle(st, tol=1e-12)
    o.deposit()
    o.solve()
    o.efield()
    runs = []
    for _ in range(2):
        c = make_ctx(st, precision=1, tol=1e-12, sort_interval=3, order=S.ORDER_MORTON)
        c.set_preconditioner(S.PRECOND_TWO_LEVEL)
        c.initialize()
        assert c.solver_info()["preconditioner"] == S.PRECOND_TWO_LEVEL
        c.step(6)
        runs.append(c.fields())
    for k in range(6):
        o.step(1)
    _, phi_o, E_o = o.fields()
    _, phi_c, E_c = runs[0]
    print("phi rel", rel_err(phi_c, phi_o), "E rel", rel_err(E_c, E_o))
    assert rel_err(phi_c, phi_o) < 1e-6 and rel_err(E_c, E_o) < 1e-6
    for a, b in zip(runs[0], runs[1]):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("precond", [S.PRECOND_JACOBI, S.PRECOND_TWO_LEVEL])
def test_tile_staging_gives_the_row_pattern_bits(precond):
    """In the tiled order the iteration that stages each tile's neighbour
    values in LDS reads the same operator entries in the same slot order as
    the row-pattern iteration: same bits, Jacobi and two-level."""
    st = plasma_setup(n=24, npart=40000, seed=8)
    res = {}
    for staged in (1, 0):
        c = make_ctx(st, precision=0, tol=1e-10, sort_interval=0, order=S.ORDER_TILED)
        c.set_preconditioner(precond)
        c.tile_staging = staged
        c.initialize()
        info = c.solver_info()
        c.step(3)
        res[staged] = (c.fields(), c.stats(), info)
    on, off = res[1], res[0]
    print("staged", on[2], "row patterns", off[2])
    assert on[2]["tile_staged"] == 1 and on[2]["box_patterns"] > 0
    assert off[2]["tile_staged"] == 0
    assert on[2]["coarse_in_kernel"] == int(precond == S.PRECOND_TWO_LEVEL)
    for a, b in zip(on[0], off[0]):
        np.testing.assert_array_equal(a, b)
    assert on[1]["floating_potential"] == off[1]["floating_potential"]
    assert on[1]["pcg_iterations_total"] == off[1]["pcg_iterations_total"]
