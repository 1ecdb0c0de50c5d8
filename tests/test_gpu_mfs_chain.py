"""Modified-fixed-source runs whose fission chains end on the smaller bank.

abx_run swaps two device banks at every re-banked generation of a chain.
The second bank is sized for the chain's re-banked generations, which are
much smaller than a batch (a served carter request: 383,812 sources, then
8,181, 109 and 1 re-banked histories).  After an odd number of swaps the
next batch's sources must not be sampled into the smaller bank: the run
grows it back to a whole batch first, and refuses to launch over a bank that
is too small.  These runs once wrote a whole batch into the chain's small
bank; each now matches the CPU restatement's run.
"""
import numpy as np
import pytest

from abeille_amd import Problem, abi, problems
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu


def _carter_request(index):
    """The index-th request of the seeded request generator (the served
    stream's request mix, drawn without arrival times)."""
    rng = np.random.default_rng(2025)
    for i in range(index + 1):
        m, _, kind = problems.served_request(rng)
    assert kind == "carter"
    return m


# (request index, histories per batch): both chains have three re-banked
# generations in their first batch (odd), so the second batch used to land
# in the small bank.
CASES = {"r16": (16, 11240), "r3": (3, 383812)}


@pytest.mark.parametrize("name", sorted(CASES))
def test_mfs_carter_second_batch_after_odd_chain(name):
    idx, n = CASES[name]
    m = _carter_request(idx)
    assert m.settings["nparticles"] == n
    print(f"[mfs-chain] case={name} request={idx} histories/batch={n} engine=history",
          flush=True)
    o = Oracle(m)
    g = Problem(m)
    so, ko = o.run(2, 0, nthreads=0)
    sg, kg = g.run(2, 0)
    assert sg.diag.histories == so.diag.histories
    np.testing.assert_allclose(kg, ko, rtol=1e-9)
    assert sg.diag.collisions == so.diag.collisions
    assert sg.diag.secondaries == so.diag.secondaries
    for t in range(len(m.tallies)):
        ao, eo = o.tally_stats(t)
        ag, eg = g.tally_stats(t)
        np.testing.assert_allclose(ag, ao, rtol=1e-8, atol=1e-14)
        np.testing.assert_allclose(eg, eo, rtol=1e-6, atol=1e-12)
