"""Poisson-failure policy of the device cycle.

PICCycle::solvePoissonEquation keeps the previous potential when every
solver fails (Tools/PICcore/src/PICCycle.cpp:731-767) and
updateElectricField rebuilds the same E from it (:769-800), while the step
itself (push, deposit, boundaries) still runs.  The failure is forced from
the host with an unreachable tolerance and a two-iteration cap.
"""
import numpy as np
import pytest

import scdat_amd as S
from helpers import make_ctx, plasma_setup

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [0, 1])
def test_failed_solves_keep_previous_field_and_are_latched(precision):
    st = plasma_setup(n=8, npart=4000, seed=31)
    c = make_ctx(st, precision=precision, tol=1e-8 if precision == 0 else 1e-11, sort_interval=0)
    c.initialize()
    c.step(2)
    rho0, phi0, E0 = c.fields()
    s0 = c.stats()
    x0, _, _ = c.get_particles_by_id(0, len(st["e"][0]))
    assert s0["pcg_failed_solves"] == 0 and s0["pcg_converged"] == 1
    c.set_pcg_controls(1e-15, 2)
    with pytest.raises(S.SolverFailure) as e:
        c.step(3)
    assert "3 of 3" in str(e.value) and "first in step 3" in str(e.value)
    rho1, phi1, E1 = c.fields()
    s1 = c.stats()
    # the field is the previous one, bit for bit; the floating potential too
    assert np.array_equal(phi1, phi0)
    assert np.array_equal(E1, E0)
    assert s1["floating_potential"] == s0["floating_potential"]
    # the steps themselves ran: particles moved, new charge was deposited
    assert s1["completed_steps"] == s0["completed_steps"] + 3
    x1, _, _ = c.get_particles_by_id(0, len(st["e"][0]))
    ok = np.isfinite(x0[:, 0]) & np.isfinite(x1[:, 0])
    assert np.max(np.abs(x1[ok] - x0[ok])) > 0
    assert not np.array_equal(rho1, rho0)
    assert s1["pcg_failed_solves"] == 3 and s1["pcg_converged"] == 0
    # back to a reachable tolerance: the next solve converges and moves phi
    c.set_pcg_controls(1e-8 if precision == 0 else 1e-11, 50000)
    c.step(1)
    s2 = c.stats()
    assert s2["pcg_converged"] == 1 and s2["pcg_failed_solves"] == 3
    assert not np.array_equal(c.fields()[1], phi0)
    c.close()


def test_piccycle_mirror_reports_failed_step():
    st = plasma_setup(n=8, npart=2000, seed=32)
    m = st["mesh"]
    cyc = S.PICCycle(S.piccycle.Parameters(time_step=st["dt"], precision=0, sort_interval=0))
    cyc.setMesh(m["xyz"], m["tets"], m["kind"], face_group=m["group"])
    for g in range(6, 12):
        cyc.setBoundaryCondition(g, S.BC_CONDUCTOR)
    cyc.addElectrons(st["e"][0], st["e"][1], st["w"], cell=st["e"][2])
    cyc.addIons(st["i"][0], st["i"][1], st["w"], cell=st["i"][2])
    assert cyc.initialize()
    assert cyc.executeTimeStep()
    phi = cyc.getPotential()
    cyc.context.set_pcg_controls(1e-15, 2)
    assert cyc.executeTimeStep() is False
    assert np.array_equal(cyc.getPotential(), phi)
    assert cyc.getStatistics()["pcg_failed_solves"] == 1
    cyc.context.close()
