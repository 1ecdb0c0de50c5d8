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
    assert problem_class(m) in json.loads(cache.read_text())
    q = Problem(m)
    best2, times2 = autotune(q, m, cache=str(cache), candidates=cands)
    assert best2 == best and times2 == {}
    assert q.tuning() == {k: best[k] for k in q.tuning()}
    _, k_tuned = q.run(4, 1)
    r = Problem(m)
    _, k_plain = r.run(4, 1)
    np.testing.assert_allclose(k_tuned, k_plain, rtol=1e-9)
