"""Batched problems (abx_batch_*): a problem advanced together with others
gives the results it gives alone through abx_run - same histories, same RNG
streams, same ordered fission banks - up to the summation order of tallies
and scores."""
import numpy as np
import pytest

from abeille_amd import Batch, Problem, abi, problems

pytestmark = pytest.mark.gpu


def _mix():
    rng = np.random.default_rng(4242)
    out = [("cube_dt", problems.fixed_source_cube(3000, "delta"), 3),
           ("cube_st", problems.fixed_source_cube(2000, "surface"), 2),
           ("mfs_carter_cancel", problems.subcritical_slab(2000, tracking="carter"), 3),
           ("mfs_delta", problems.subcritical_slab(1500, tracking="delta", cancel=False), 2),
           ("hex_fs", problems.hex_assembly(1500, tracking="surface", top="flat",
                                            mode=abi.FIXED_SOURCE)[0], 2)]
    for k in range(6):
        m, _ = problems.slab_problem(rng, nhist=int(rng.integers(500, 4000)))
        if k % 2:
            m.set(precision=abi.SPLIT)
        out.append((f"slab{k}", m, 2 + k % 2))
    return out


def _alone(m, nb):
    p = Problem(m)
    st, _ = p.run(nb, 0)
    return st, [p.tally_stats(t) for t in range(len(m.tallies))]


def test_batch_equals_alone():
    cases = _mix()
    b = Batch()
    probs = {}
    for name, m, nb in cases:
        probs[name] = Problem(m)
        assert b.add(probs[name], nb)
    finished = []
    steps = 0
    while len(b):
        finished += b.step()
        steps += 1
        assert steps < 1000
    assert len(finished) == len(cases)
    for name, m, nb in cases:
        print(f"[batch] {name}", flush=True)
        sa, ta = _alone(m, nb)
        sb = probs[name].stats()
        assert sb.active_generations == sa.active_generations == nb
        assert sb.diag.collisions == sa.diag.collisions, name
        assert sb.diag.flights == sa.diag.flights, name
        assert sb.histories == sa.histories, name
        for f in ("kcol_avg", "leak_avg", "mig_avg", "ktrk_avg"):
            assert getattr(sb, f) == pytest.approx(getattr(sa, f), rel=1e-9, abs=1e-14), (name, f)
        for t, (aa, ea) in enumerate(ta):
            ab, eb = probs[name].tally_stats(t)
            np.testing.assert_allclose(ab, aa, rtol=1e-8, atol=1e-14, err_msg=name)
            np.testing.assert_allclose(eb, ea, rtol=1e-6, atol=1e-12, err_msg=name)


def test_batch_budget_admission():
    m = problems.fixed_source_cube(1000)
    p = Problem(m)
    need = p.footprint()
    b = Batch(budget=need + need // 2)
    assert b.add(p, 1)
    q = Problem(problems.fixed_source_cube(1000))
    assert not b.add(q, 1)       # does not fit next to p
    while len(b):
        b.step()
    assert b.add(q, 1)           # p's bytes were released when it finished
    while len(b):
        b.step()


def test_batch_rejects_k_eigenvalue():
    b = Batch()
    with pytest.raises(Exception):
        b.add(Problem(problems.la13511("PUa-1-0-SL", 1000)[0]), 1)


def test_sliced_batches_equal_alone():
    """With a step cap, fixed-source batches larger than their share run over
    several steps; results are those of the whole batch run alone.  The
    modified-fixed-source problems in the mix run whole generations."""
    rng = np.random.default_rng(77)
    cases = [("cube_dt", problems.fixed_source_cube(30000, "delta"), 2),
             ("cube_st", problems.fixed_source_cube(20000, "surface"), 2),
             ("small", problems.fixed_source_cube(3000, "delta"), 3),
             ("mfs_carter_cancel", problems.subcritical_slab(2000, tracking="carter"), 2)]
    m, _ = problems.slab_problem(rng, nhist=25000)
    m.set(precision=abi.SPLIT)
    cases.append(("slab_split", m, 2))
    b = Batch(step_histories=16384)
    probs = {}
    for name, m, nb in cases:
        probs[name] = Problem(m)
        assert b.add(probs[name], nb)
    steps = 0
    while len(b):
        b.step()
        steps += 1
        assert steps < 1000
    # the 30000-history batches need at least 4 slices of at most 8192
    assert steps >= 8
    for name, m, nb in cases:
        print(f"[sliced] {name}", flush=True)
        sa, ta = _alone(m, nb)
        sb = probs[name].stats()
        assert sb.active_generations == sa.active_generations == nb
        assert sb.diag.collisions == sa.diag.collisions, name
        assert sb.histories == sa.histories, name
        for f in ("kcol_avg", "leak_avg", "mig_avg"):
            assert getattr(sb, f) == pytest.approx(getattr(sa, f), rel=1e-9, abs=1e-14), (name, f)
        for t, (aa, ea) in enumerate(ta):
            ab, eb = probs[name].tally_stats(t)
            np.testing.assert_allclose(ab, aa, rtol=1e-8, atol=1e-14, err_msg=name)
            np.testing.assert_allclose(eb, ea, rtol=1e-6, atol=1e-12, err_msg=name)


def test_batch_parameters_only_before_attaching():
    b = Batch()
    b.set_params(chunk=2048, multi_hist=1 << 20)
    b.set_step_histories(0)
    p = Problem(problems.fixed_source_cube(1000))
    assert b.add(p, 1)
    with pytest.raises(Exception):
        b.set_params(chunk=1024)
    with pytest.raises(Exception):
        b.set_step_histories(4096)
    while len(b):
        b.step()


@pytest.mark.parametrize("block", [64, 128])
@pytest.mark.parametrize("graphs", [True, False], ids=["graphs", "eager"])
def test_batch_block_size_and_graph_replay(block, graphs):
    """Steps that repeat the previous step's launches replay its captured
    transport graph; the results are those of the problems alone, with any
    block size of the batch kernel."""
    print(f"[batch] block={block} graphs={graphs}", flush=True)
    cases = [("cube_dt", problems.fixed_source_cube(2500, "delta"), 6),
             ("cube_st", problems.fixed_source_cube(1500, "surface"), 5)]
    b = Batch()
    b.set_params(chunk=512, multi_hist=1, block=block)  # one batch per problem and step
    b.set_graphs(graphs)
    probs = {}
    for name, m, nb in cases:
        probs[name] = Problem(m)
        assert b.add(probs[name], nb)
    steps = 0
    while len(b):
        b.step()
        steps += 1
        assert steps < 100
    captures, replays = b.graph_stats()
    if graphs:
        assert captures >= 1 and replays >= 3, (captures, replays)
    else:
        assert captures == replays == 0
    for name, m, nb in cases:
        sa, ta = _alone(m, nb)
        sb = probs[name].stats()
        assert sb.active_generations == nb
        assert sb.diag.collisions == sa.diag.collisions, name
        assert sb.histories == sa.histories, name
        assert sb.leak_avg == pytest.approx(sa.leak_avg, rel=1e-9, abs=1e-14), name
        for t, (aa, ea) in enumerate(ta):
            ab, eb = probs[name].tally_stats(t)
            np.testing.assert_allclose(ab, aa, rtol=1e-8, atol=1e-14, err_msg=name)
            np.testing.assert_allclose(eb, ea, rtol=1e-6, atol=1e-12, err_msg=name)
