"""Captured steps (SCDAT_SHAPE_STEP_GRAPH) against eager launches through the
C ABI: the same kernels with the same arguments in the same order, the slot
counts and the step number read from the device, so fields, floating
potentials, ledgers, event counts and particles are equal bit for bit, for
a single case with collisions, secondary emission and photoemission and for
a batch of cases; with the sorts that change the particle buffers falling
inside the captured calls."""
import numpy as np
import pytest

import scdat_amd as S
import scdat_amd.batch as B
import scdat_amd.geo as G

pytestmark = pytest.mark.gpu


def _state(c, n_species):
    rho, phi, E = c.fields()
    parts = []
    for s in range(n_species):
        x, v, cell, ids = c.get_particles(s)
        live = cell >= 0
        o = np.lexsort((x[live][:, 0], ids[live]))
        parts.append((ids[live][o], x[live][o], v[live][o], cell[live][o]))
    return dict(rho=rho, phi=phi, E=E, parts=parts, ledger=c.ledger(), events=c.event_counts(), stats=c.stats())


def _assert_same(a, b):
    np.testing.assert_array_equal(a["rho"], b["rho"])
    np.testing.assert_array_equal(a["phi"], b["phi"])
    np.testing.assert_array_equal(a["E"], b["E"])
    # the kinetic energy sums per-block partials, and the particles' order
    # inside a cell after a sort follows atomic arrival: equal to rounding
    assert a["stats"]["total_energy"] == pytest.approx(b["stats"]["total_energy"], rel=1e-12)
    assert {k: v for k, v in a["stats"].items() if k != "total_energy"} == \
        {k: v for k, v in b["stats"].items() if k != "total_energy"}
    # the incident energy is an FP64 sum of atomics (its rounding follows
    # the arrival order); every other ledger field is exact
    for la, lb in zip(a["ledger"], b["ledger"]):
        assert la["incident_energy_j"] == pytest.approx(lb["incident_energy_j"], rel=1e-12, abs=0.0)
        assert {k: v for k, v in la.items() if k != "incident_energy_j"} == \
            {k: v for k, v in lb.items() if k != "incident_energy_j"}
    assert a["events"] == b["events"]
    for pa, pb in zip(a["parts"], b["parts"]):
        for qa, qb in zip(pa, pb):
            np.testing.assert_array_equal(qa, qb)


@pytest.mark.parametrize("precision", [1, 0])
def test_captured_steps_equal_eager_steps_geo_case(precision):
    # GEO-like case at a small size, with a dense background so that
    # collisions and ionisation happen every step
    st = G.inputs(edge=8, n_particles=40000, neutral_density=1e19, seed=7)
    runs = []
    for graph in (0, 1):
        c = G.context(st, precision=precision, tol=1e-10, sort_interval=4, steps=40)
        c.step_graph = graph
        assert c.step_graph == graph
        print("case graph", graph, flush=True)
        c.step(9)    # three calls, so recordings restart per call
        c.step(1)
        c.step(10)
        runs.append(_state(c, 3))
        c.close()
    e = runs[1]["events"]
    print("events", e)
    assert sum(e["collisions"][:3]) > 0 and e["photo_emitted"] > 0 and e["see_emitted"] >= 0
    _assert_same(runs[0], runs[1])


@pytest.mark.parametrize("precision", [1, 0])
def test_captured_steps_equal_eager_steps_batched(precision):
    cases = [dict(edge=6, n_particles=3000, te_ev=2.0, density=1e10, delta_max=1.5, seed=11,
                  neutral_density=1e19, photo_j=1e-3),
             dict(edge=5, n_particles=1500, te_ev=400.0, density=1e7, delta_max=1.0, seed=13, photo_j=2e-5),
             dict(edge=7, n_particles=2500, te_ev=8.0, density=2e9, delta_max=3.0, seed=14)]
    inputs = [B.case_inputs(sc) for sc in cases]
    runs = []
    for graph in (0, 1):
        bc, u = B.batch_context(cases, inputs, precision=precision, tol=1e-10, sort_interval=3)
        bc.step_graph = graph
        print("batch graph", graph, flush=True)
        bc.step(8)
        st = _state(bc, 3)
        st["scen"] = [bc.scenario_stats(k) for k in range(len(cases))]
        runs.append(st)
        bc.close()
    _assert_same(runs[0], runs[1])
    for a, b in zip(runs[0]["scen"], runs[1]["scen"]):
        assert a == b


def test_step_graph_shape_is_checked():
    c = S.Context(precision=0)
    with pytest.raises(S.ScdatError):
        c.step_graph = 2
    c.close()
