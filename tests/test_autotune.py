"""The launch-shape autotuner: it picks one of its candidates, caches the
choice per problem class, and the tuning changes speed only - a tuned run
gives the untuned run's results up to tally summation order."""
import json

import numpy as np
import pytest

from abeille_amd import Problem, problems
from abeille_amd.autotune import CANDIDATES, autotune, problem_class


def test_problem_class_separates_configurations():
    a = problem_class(problems.pin_lattice_2d(100))
    b = problem_class(problems.pin_lattice_2d(100, tracking="surface"))
    c = problem_class(problems.pin_lattice_2d(5000))   # history count is not part of the class
    assert a != b and a == c


@pytest.mark.gpu
def test_autotune_caches_and_keeps_results(tmp_path):
    cache = tmp_path / "tuning.json"
    m = problems.pin_lattice_2d(20000)
    p = Problem(m)
    cands = CANDIDATES[:4]
    best, times = autotune(p, m, sample=20000, repeats=1, cache=str(cache), candidates=cands)
    assert best in cands and len(times) == len(cands)
    assert any(k.startswith(problem_class(m)) for k in json.loads(cache.read_text()))
    q = Problem(m)
    best2, times2 = autotune(q, m, cache=str(cache), candidates=cands)
    assert best2 == best and times2 == {}
    assert q.tuning() == {k: best[k] for k in q.tuning()}
    _, k_tuned = q.run(4, 1)
    r = Problem(m)
    _, k_plain = r.run(4, 1)
    np.testing.assert_allclose(k_tuned, k_plain, rtol=1e-9)


@pytest.mark.gpu
def test_autotune_batch_picks_and_keeps_results():
    """The batch-engine parameters change speed only: a batch run with the
    chosen chunk length and side-by-side cap gives each problem its results
    alone."""
    from abeille_amd import Batch
    from abeille_amd.autotune import BATCH_CANDIDATES, autotune_batch
    rng = np.random.default_rng(5)
    models = [problems.slab_problem(rng, nhist=int(rng.integers(500, 3000)))[0] for _ in range(6)]
    sample = [Problem(m) for m in models[:3]]
    cands = [BATCH_CANDIDATES[0], dict(chunk=300, multi_hist=1), dict(chunk=64, multi_hist=1 << 23),
             dict(chunk=-32, multi_hist=1 << 20)]
    b = Batch()
    params, times = autotune_batch(b, sample, nbatches=1, candidates=cands)
    assert params in cands and len(times) == len(cands)
    for c in cands:                      # every candidate keeps the results
        b = Batch()
        b.set_params(**c)
        probs = [Problem(m) for m in models]
        for p in probs:
            assert b.add(p, 2)
        while len(b):
            b.step()
        for m, p in zip(models, probs):
            q = Problem(m)
            st, _ = q.run(2, 0)
            assert p.stats().diag.collisions == st.diag.collisions
            for t in range(len(m.tallies)):
                np.testing.assert_allclose(p.tally_stats(t)[0], q.tally_stats(t)[0], rtol=1e-8,
                                           atol=1e-14)
