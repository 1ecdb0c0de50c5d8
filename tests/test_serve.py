"""Served stream (abeille_amd.serve): continuous batching under an HBM
budget, in one lane or in several (fixed-source lanes and a chain lane,
each stepped by a thread of its own).  Each request's result in the mix equals its result
alone (same precision, abx_run on a fresh problem) up to tally summation
order, and the budget is honoured (requests wait and are admitted as
memory frees up)."""
import numpy as np
import pytest

from abeille_amd import Problem, abi, problems
from abeille_amd.serve import Request, Server, serial_latencies


@pytest.mark.gpu
@pytest.mark.parametrize("lanes", [3, 1])
def test_served_results_equal_alone(lanes):
    rng = np.random.default_rng(99)
    reqs = []
    t = 0.0
    while len(reqs) < 14:
        m, nh, kind = problems.served_request(rng)
        if nh > 400_000:
            continue                      # keep the test short
        if len(reqs) % 2:
            m.set(precision=abi.SPLIT)
        t += rng.exponential(0.005)
        reqs.append(Request(m, 2, t, name=f"{kind}{len(reqs)}"))
    assert any(r.name.startswith("carter") for r in reqs)
    foot = max(Problem(r.model).footprint() for r in reqs)
    srv = Server(budget=3 * foot, lanes=lanes)
    srv.run(reqs)
    waited = sum(r.admitted > r.arrival + 1e-3 for r in reqs)
    assert waited > 0, "the budget never made a request wait"
    for r in reqs:
        print(f"[serve] {r.name}", flush=True)
        assert r.finished >= r.admitted >= r.arrival - 1e-9
        p = Problem(r.model)
        st, _ = p.run(r.nbatches, 0)
        assert r.stats.diag.collisions == st.diag.collisions, r.name
        assert r.stats.histories == st.histories, r.name
        assert r.stats.leak_avg == pytest.approx(st.leak_avg, rel=1e-9, abs=1e-14), r.name
        for t_idx, (a, e) in enumerate(r.tallies):
            ao, eo = p.tally_stats(t_idx)
            np.testing.assert_allclose(a, ao, rtol=1e-8, atol=1e-14, err_msg=r.name)
            np.testing.assert_allclose(e, eo, rtol=1e-6, atol=1e-12, err_msg=r.name)


def test_serial_latencies_first_come_first_served():
    arrivals = np.array([0.0, 0.1, 0.15, 2.0])
    service = np.array([0.5, 0.2, 0.1, 0.3])
    lat = serial_latencies(arrivals, service)
    # finishes at 0.5, 0.7, 0.8, 2.3
    np.testing.assert_allclose(lat, [0.5, 0.6, 0.65, 0.3])
