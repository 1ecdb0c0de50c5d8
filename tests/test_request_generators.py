"""The batched-form and served-stream problem generators (host logic): the
problems they emit are valid descriptions, and every carter-tracking request
(0.8 x majorant, fissile, cancellation) has a fission chain that ends, which
the restatement checks on a bounded sample."""
import numpy as np

from abeille_amd import abi, problems
from oracle.pyoracle import Oracle


def test_served_requests_mix_and_sizes():
    rng = np.random.default_rng(3)
    kinds, hist = [], []
    for _ in range(200):
        m, nh, kind = problems.served_request(rng)
        kinds.append(kind)
        hist.append(nh)
        assert m.settings["nparticles"] * 2 == nh
        if kind == "carter":
            assert m.settings["tracking"] == abi.CARTER_TRACKING
            assert m.settings["mode"] == abi.MODIFIED_FIXED_SOURCE
            assert m.cancel is not None
    frac = kinds.count("carter") / len(kinds)
    assert 0.12 < frac < 0.28
    assert 1e4 <= min(hist) and max(hist) <= 5e6


def test_carter_request_chains_end():
    rng = np.random.default_rng(17)
    seen = 0
    while seen < 8:
        m, _, kind = problems.served_request(rng)
        if kind != "carter":
            continue
        m.set(nparticles=400)
        st, k = Oracle(m).run(1, 0, nthreads=2)
        assert np.isfinite(k).all()
        assert st.diag.collisions > 0
        seen += 1


def test_batched_form_problem_set():
    rng = np.random.default_rng(512)
    G = set()
    for _ in range(40):
        m, nh = problems.slab_problem(rng)
        G.add(m.ngroups)
        assert 1e3 <= nh <= 1e6
        assert len(m.tallies) == 1 and m.tallies[0]["shape"] == [16, 16, 16]
        m.desc()
    assert G <= {2, 4, 7, 16, 33} and len(G) >= 3
