"""Host side of the batched charging scenarios (no GPU): the sweep's
parameter draws and the concatenation of the cases into one mesh."""
import math

import numpy as np

import scdat_amd as S
import scdat_amd.batch as B


def test_sweep_is_deterministic_and_within_ranges():
    a = B.sweep(64, 20261206)
    b = B.sweep(64, 20261206)
    assert a == b
    assert B.sweep(64, 7) != a
    for sc in a:
        assert 8 <= sc["edge"] <= 32
        assert 2e4 <= sc["n_particles"] <= 2e6
        assert 1.0 <= sc["te_ev"] <= 2e4
        assert 1e6 <= sc["density"] <= 1e11
        assert 1.0 <= sc["delta_max"] <= 3.0
    # log-uniform: the mean of log(n) sits near the middle of the range
    ln = np.log([sc["density"] for sc in B.sweep(400, 3)])
    assert abs(ln.mean() - 0.5 * (math.log(1e6) + math.log(1e11))) < 0.6


def test_concat_aligns_cases_and_offsets_tets_and_cells():
    cases = [dict(edge=e, n_particles=500, te_ev=1.0, density=1e10, delta_max=2.0, seed=e) for e in (5, 6, 4)]
    inputs = [B.case_inputs(sc) for sc in cases]
    u = B.concat(inputs)
    ns = u["node_start"]
    assert ns[0] == 0 and ns[-1] == len(u["xyz"])
    assert np.all(ns % S.SEGMENT_ALIGN == 0)
    for k, st in enumerate(inputs):
        n = len(st["mesh"]["xyz"])
        t0, t1 = u["tet_offset"][k], u["tet_offset"][k + 1]
        np.testing.assert_array_equal(u["tets"][t0:t1] - ns[k], st["mesh"]["tets"])
        np.testing.assert_array_equal(u["xyz"][ns[k]:ns[k] + n], st["mesh"]["xyz"])
        # padding nodes: used by no tet, Dirichlet
        pad = np.arange(ns[k] + n, ns[k + 1])
        assert not np.isin(pad, u["tets"]).any()
        assert np.all(u["kind"][pad] == S.NODE_DIRICHLET)
    # cells of each case shifted by its tet offset, in case order
    off = 0
    for k, st in enumerate(inputs):
        c = st["e"][2]
        np.testing.assert_array_equal(u["e"][2][off:off + len(c)], c + u["tet_offset"][k])
        off += len(c)


def test_case_time_step_rule():
    sc = dict(edge=10, n_particles=1000, te_ev=1.0, density=1e10, delta_max=2.0, seed=1)
    st = B.case_inputs(sc)
    wpe = math.sqrt(1e10 * S.ELEMENTARY_CHARGE ** 2 / (S.VACUUM_PERMITTIVITY * S.ELECTRON_MASS))
    assert st["dt"] == min(0.2 / wpe, 0.2 * 0.1 / math.sqrt(S.ELEMENTARY_CHARGE / S.ELECTRON_MASS))
    assert st["w"] == 1e10 * (1 - 0.2 ** 3) / 1000
