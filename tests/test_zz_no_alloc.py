"""scdat_get_stats_noalloc, the allocation-free form of scdat_get_stats,
against scdat_get_stats through the C ABI: equal bit for bit on two calls in
a row, the second with a smaller input (a batched context of four scenarios
with secondary emission, then one small FP64 context), so that nothing of
the first call's scratch contents reaches the second's results."""
import pytest

import scdat_amd.batch as B
from helpers import make_ctx, plasma_setup

pytestmark = pytest.mark.gpu

CASES = [
    dict(edge=6, n_particles=3000, te_ev=2.0, density=1e10, delta_max=1.5, seed=11),
    dict(edge=8, n_particles=5000, te_ev=60.0, density=3e8, delta_max=2.5, seed=12),
    dict(edge=5, n_particles=1500, te_ev=400.0, density=1e7, delta_max=1.0, seed=13),
    dict(edge=7, n_particles=2500, te_ev=8.0, density=2e9, delta_max=3.0, seed=14),
]


def _same(a, b):
    assert a.keys() == b.keys()
    for k in a:
        assert a[k] == b[k], (k, a[k], b[k])


def test_stats_noalloc_equals_stats_on_two_calls_in_a_row():
    bc, _ = B.batch_context(CASES, precision=0, tol=1e-8, sort_interval=3)
    bc.step(4)
    first = bc.stats_noalloc()
    _same(bc.stats(), first)
    assert first["active_particles"] > 0 and first["surface_deposited_charge"] != 0.0
    bc.close()
    st = plasma_setup(n=4, npart=500, seed=5)
    c = make_ctx(st, precision=1, tol=1e-10, sort_interval=0)
    c.initialize()
    c.step(2)
    second = c.stats_noalloc()
    _same(c.stats(), second)
    assert second["active_particles"] < first["active_particles"]
    # and again on the same context: the scratch keeps nothing between calls
    c.step(1)
    _same(c.stats(), c.stats_noalloc())
    c.close()
