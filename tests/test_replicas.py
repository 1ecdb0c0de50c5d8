"""Batched k-eigenvalue chains (abx_run_replicas): R chains of one problem
with their own seeds, run together in one bank and one launch per
generation, against each chain run alone through abx_run (itself checked
against the CPU restatement in the other tests) and against the CPU
restatement directly for one chain.

Per chain: k_col of every generation to 1e-9 (weight sums add in another
order).  Together: event counts exactly the alone runs' sums, and the mesh
tally mean exactly the mean of the alone runs' tally means (the generation
records the chains' mean, and means commute)."""
import numpy as np
import pytest

from abeille_amd import AbxError, Problem, abi, problems
from helpers import bins_close
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

SEEDS = [abi.DEFAULT_SEED, 987654321, 4242424242, 77]


def _alone(make, seed, ngen, nign):
    m = make()
    m.set(seed=seed)
    p = Problem(m)
    st, kh = p.run(ngen, nign)
    return st, kh, [p.tally_stats(t) for t in range(len(m.tallies))]


CASES = {
    "lattice_dt": lambda: problems.pin_lattice_2d(3000, tracking="delta"),
    "lattice_st": lambda: problems.pin_lattice_2d(3000, tracking="surface"),
    "lattice_dt_split": lambda: _split(problems.pin_lattice_2d(3000, tracking="delta")),
    "core_st": lambda: problems.core_3x3(3000, tracking="surface"),
}


def _split(m):
    m.set(precision=abi.SPLIT)
    return m


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("R", [1, 3])
def test_replicas_equal_chains_alone(name, R):
    print(f"[replicas] {name} R={R}", flush=True)
    make = CASES[name]
    ngen, nign = 6, 2
    seeds = SEEDS[:R]
    g = Problem(make())
    st, kh = g.run_replicas(R, ngen, nign, seeds=seeds)
    alone = [_alone(make, sd, ngen, nign) for sd in seeds]
    for r in range(R):
        assert np.allclose(kh[r], alone[r][1], rtol=1e-9, atol=0), (r, kh[r], alone[r][1])
    assert st.diag.collisions == sum(a[0].diag.collisions for a in alone)
    assert st.diag.histories == sum(a[0].diag.histories for a in alone)
    assert st.active_generations == ngen - nign
    # the recorded k is the chains' mean per generation
    kmean = np.mean([a[1] for a in alone], axis=0)
    assert st.kcol_avg == pytest.approx(kmean[nign:].mean(), rel=1e-9)
    np.testing.assert_allclose(g.gen_values(abi.GEN_KCOL), kmean, rtol=1e-9)
    for t in range(len(g.model.tallies)):
        avg, _ = g.tally_stats(t)
        mean_alone = np.mean([a[2][t][0] for a in alone], axis=0)
        assert bins_close(mean_alone, avg, 1e-8), t


def test_replicas_continue_equals_one_call():
    """Two calls (inactive generations, then active ones) give the chains and
    statistics of one call over all generations."""
    m = problems.pin_lattice_2d(3000, tracking="delta")
    a, b = Problem(m), Problem(m)
    sa, ka = a.run_replicas(3, 7, 2, seeds=SEEDS[:3])
    b.run_replicas(3, 2, 2, seeds=SEEDS[:3])
    sb, kb = b.run_replicas_continue(5, 0)
    assert np.allclose(ka[:, 2:], kb, rtol=1e-9, atol=0)
    assert sb.active_generations == sa.active_generations == 5
    assert sb.kcol_avg == pytest.approx(sa.kcol_avg, rel=1e-9)
    assert sb.kcol_err == pytest.approx(sa.kcol_err, rel=1e-6)
    for t in range(len(m.tallies)):
        assert bins_close(a.tally_stats(t)[0], b.tally_stats(t)[0], 1e-8)
    with pytest.raises(AbxError, match="no replica run"):
        Problem(m).run_replicas_continue(2, 0)


def test_replicas_from_bank():
    """Chains started from the bank of a converged run: deterministic, every
    chain's first generation transports that bank (on its own stream), and a
    problem without a bank is refused."""
    m = problems.pin_lattice_2d(3000, tracking="delta")
    runs = []
    for _ in range(2):
        p = Problem(m)
        p.run(4, 4)
        n0 = len(p.bank())
        st, kh = p.run_replicas(3, 4, 0, seeds=SEEDS[:3], from_bank=True)
        runs.append((st, kh))
    (s1, k1), (s2, k2) = runs
    assert np.allclose(k1, k2, rtol=1e-12, atol=0)   # weight sums add in any order
    assert len(set(k1[:, 0].tolist())) == 3      # same particles, different streams
    assert s1.diag.histories > 3 * n0 * 3        # four generations of three chains
    assert abs(s1.kcol_avg - 0.295) < 0.02
    with pytest.raises(AbxError, match="no k-eigenvalue bank"):
        Problem(m).run_replicas(2, 2, 0, from_bank=True)


def test_replica_chain_matches_oracle():
    """Chain 1 of a 2-chain run against the CPU restatement with its seed."""
    m = problems.pin_lattice_2d(2000, tracking="delta")
    g = Problem(m)
    _, kh = g.run_replicas(2, 4, 1, seeds=SEEDS[:2])
    mo = problems.pin_lattice_2d(2000, tracking="delta")
    mo.set(seed=SEEDS[1])
    _, ko = Oracle(mo).run(4, 1, nthreads=0)
    assert np.allclose(kh[1], ko, rtol=1e-9, atol=0), (kh[1], ko)


def test_replicas_refuse_what_they_do_not_run():
    g = Problem(problems.pin_lattice_2d(1000))
    with pytest.raises(AbxError, match="replicas"):
        g.run_replicas(17, 2, 0)
    with pytest.raises(AbxError, match="k-eigenvalue"):
        Problem(problems.fixed_source_cube(1000)).run_replicas(2, 2, 0)
    with pytest.raises(AbxError, match="carter"):
        m, _ = problems.la13511("PUa-1-0-SL", 1000, tracking="carter")
        m.set(sample_xs_ratio=[0.8])
        Problem(m).run_replicas(2, 2, 0)
