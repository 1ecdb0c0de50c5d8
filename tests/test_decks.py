"""The reference's own multigroup decks (input_files/*.yaml), as flattened
fixtures written by tests/golden/make_deck_fixtures.py, on the CPU
restatement and through the HIP path.

Known answers: the LA-13511 analytic k of the infinite media (PUa-1-0-IN
2.612903, PUb-1-0-IN 2.290323, Ua-1-1-IN 2.25) and k = 1 for the critical
slabs, cylinder and two-group slab.  The GPU runs every deck - C5G7's
lattice included - bit for bit against the CPU restatement (per-history
traces, ordered fission banks) and as whole device-resident runs.
"""
import json
import os

import numpy as np
import pytest

from abeille_amd.model import Model
from helpers import assert_traces_equal, bins_close
from oracle.pyoracle import Oracle

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "decks")
REF = "/root/reference/input_files"
KNOWN = {"PUa-1-0-IN": 2.612903, "PUb-1-0-IN": 2.290323, "Ua-1-1-IN": 2.25,
         "PUa-1-0-SL": 1.0, "PUb-1-0-SL": 1.0, "PUa-1-1-SL": 1.0, "PUa-1-2-SL": 1.0,
         "Ua-1-1-CY": 1.0, "UD2O-2-1-SL": 1.0}
DECKS = sorted(KNOWN) + ["c5g7"]


def load(name, nparticles=None):
    with open(os.path.join(FIX, name + ".json")) as f:
        m = Model.from_dict(json.load(f))
    if nparticles:
        m.set(nparticles=nparticles)
    return m


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference decks not present")
@pytest.mark.parametrize("name", DECKS)
def test_fixture_is_the_deck(name):
    """The committed fixture is what the deck reader builds from the deck."""
    from abeille_amd.deck import model_from_deck
    d = model_from_deck(os.path.join(REF, name + ".yaml")).to_dict()
    d["deck"] = name + ".yaml"
    with open(os.path.join(FIX, name + ".json")) as f:
        assert json.loads(json.dumps(d)) == json.load(f)


@pytest.mark.parametrize("name", sorted(KNOWN))
def test_oracle_deck_known_k(name):
    m = load(name, 5000)
    st, _ = Oracle(m).run(40, 10)
    assert abs(st.kcol_avg - KNOWN[name]) < 4 * st.kcol_err, (name, st.kcol_avg, st.kcol_err)


@pytest.mark.gpu
@pytest.mark.parametrize("name", DECKS)
def test_gpu_deck_traces_and_banks(name):
    from abeille_amd import Problem
    print(f"[deck] traces {name}", flush=True)
    m = load(name, 2000)
    n, cap = 2000, 512
    o, g = Oracle(m), Problem(m)
    so = o.sample_sources(0, n)
    assert so.tobytes() == g.sample_sources(0, n).tobytes(), "source particles differ"
    ro = o.transport(so, kcol=0.95, nthreads=1, trace_cap=cap)
    rg = g.transport(so, kcol=0.95, trace_cap=cap)
    assert_traces_equal(ro, rg, cap)
    assert len(ro["fission"]) > 0
    assert ro["fission"].tobytes() == rg["fission"].tobytes(), "fission banks differ"


@pytest.mark.gpu
@pytest.mark.parametrize("name", DECKS)
def test_gpu_deck_run_matches_oracle(name):
    from abeille_amd import Problem
    print(f"[deck] run {name}", flush=True)
    m = load(name, 5000)
    o, g = Oracle(m), Problem(m)
    so, ko = o.run(4, 1, nthreads=0)
    sg, kg = g.run(4, 1)
    assert np.allclose(ko, kg, rtol=1e-9, atol=0), (ko, kg)
    assert so.diag.collisions == sg.diag.collisions
    for t in range(len(m.tallies)):
        ao, _ = o.tally_stats(t)
        ag, _ = g.tally_stats(t)
        assert bins_close(ao, ag, 1e-8)


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(KNOWN))
def test_gpu_deck_known_k(name):
    from abeille_amd import Problem
    m = load(name, 20000)
    st, _ = Problem(m).run(50, 10)
    assert abs(st.kcol_avg - KNOWN[name]) < 4 * st.kcol_err, (name, st.kcol_avg, st.kcol_err)
