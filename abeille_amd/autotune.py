"""Project autotuner.

Single problems: the transport launch shape (abx_tuning) - block size,
persistent-grid size and LDS staging of the tables - per problem class. A
candidate is timed with abx_time_transport on a sample of fresh source
histories (the same kernels and tables as the run, none of its state); the
fastest is applied and remembered in a JSON cache keyed by the problem's
class (tracking, mode, precision, groups, geometry and tally sizes), so a
class is searched once.

Batches (autotune_batch): the batch engine's chunk length (a block's unit
of work) and the cap on one problem's side-by-side batches, timed by
running a sample of the batch's problems through it.
"""
import hashlib
import json
import os
import time

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "abeille_amd", "tuning.json")

CANDIDATES = [dict(block_size=bs, blocks_per_cu=bpc, stage_tables=st, tail_threshold=0,
                   sort_every=0)
              for bs in (256, 128, 64) for bpc in (0, 4, 2) for st in (1, 0)]
# the event engine: sort cadence of the collision queue x tail threshold
EVENT_CANDIDATES = [dict(block_size=256, blocks_per_cu=0, stage_tables=1, tail_threshold=tt,
                         sort_every=se)
                    for se in (0, 1, 2, 4) for tt in (0, 65536)]


def problem_class(model):
    s = model.settings
    sig = dict(tracking=int(s["tracking"]), mode=int(s["mode"]), precision=int(s["precision"]),
               G=model.ngroups, surfaces=len(model.surfaces), cells=len(model.cells),
               universes=len(model.universes), materials=len(model.materials),
               tallies=[(t["estimator"], list(t["shape"])) for t in model.tallies])
    return hashlib.sha1(json.dumps(sig, sort_keys=True).encode()).hexdigest()[:16]


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def autotune(problem, model, sample=200_000, repeats=2, cache=DEFAULT_CACHE,
             candidates=CANDIDATES, force=False):
    """Pick and apply the fastest tuning for `problem`; returns (tuning,
    {candidate index: seconds})."""
    key = problem_class(model)
    if candidates is not CANDIDATES:   # another search space, another entry
        key += "-" + hashlib.sha1(json.dumps(candidates, sort_keys=True).encode()).hexdigest()[:8]
    db = _load(cache) if cache else {}
    if key in db and not force:
        problem.set_tuning(**db[key]["tuning"])
        return db[key]["tuning"], {}
    n = max(1, min(int(sample), int(model.settings["nparticles"])))
    times = {}
    for i, c in enumerate(candidates):
        problem.set_tuning(**c)
        problem.time_transport(min(n, 4096))          # warm the variant
        times[i] = min(problem.time_transport(n) for _ in range(repeats))
    best = min(times, key=times.get)
    problem.set_tuning(**candidates[best])
    if cache:
        db[key] = {"tuning": candidates[best], "seconds": times[best], "sample": n}
        os.makedirs(os.path.dirname(cache), exist_ok=True)
        tmp = cache + ".tmp"
        with open(tmp, "w") as f:
            json.dump(db, f, indent=1, sort_keys=True)
        os.replace(tmp, cache)
    return candidates[best], times


# chunk lengths as chunks per CU (negative), so that a choice made on a
# sample carries over to the full set of problems
BATCH_CANDIDATES = [dict(chunk=c, multi_hist=m, block=bs) for c in (0, -4, -16)
                    for m in (1 << 20, 1 << 23) for bs in (128, 256)]


def autotune_batch(batch, sample, nbatches=1, candidates=BATCH_CANDIDATES, repeats=1):
    """Pick and apply (batch.set_params) the fastest batch-engine parameters
    for problems like `sample` (a list of Problems, run `nbatches` batches
    each through a scratch Batch per candidate; their run statistics are
    reset when they are attached again).  Returns (params, {index: s})."""
    from . import Batch, synchronize
    times = {}
    for i, c in enumerate(candidates):
        best = None
        for _ in range(repeats):
            b = Batch()
            b.set_params(**c)
            synchronize()
            t0 = time.perf_counter()
            for p in sample:
                b.add(p, nbatches)
            while len(b):
                b.step()
            synchronize()
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        times[i] = best
    k = min(times, key=times.get)
    batch.set_params(**candidates[k])
    return candidates[k], times
