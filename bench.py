"""Benchmark of the transport hot path on one MI355X.

Default workload (the headline): the 17x17 pin-lattice k-eigenvalue problem
- pitch 1.26 cm, pin radius 0.54 cm, reflective, 7 groups, delta tracking,
1e6 particles per generation, 51x51x7 track-length flux tally - on two
paths that run the same number of active histories (so their tallies reach
the same per-bin relative error):
  * plain: FP64, one chain, the device-resident generation loop;
  * batched: `--chains` chains of the problem with their own seeds, split
    precision, one transport launch per generation for all chains.
`--warmup` inactive generations run first (untimed; they converge the
source), then exactly `--steps` batched generations are timed (the plain
leg runs chains x steps generations).  A short served stream (Poisson
arrivals, continuous batching) gives p50/p99 latency against serving the
same requests one at a time on the plain path.

Prints one JSON line: collisions/s of the batched path (value), the plain
path's rate beside it, each path's per-bin relative errors and the
histories and seconds each needs to bring 95 % of the bins to 1 %, the
latencies, the roofline of the dominant transport kernel (algorithmic bytes
per launch / HIP-event launch time, HBM-bound; issue rate from the committed
counter profile) and a CPU baseline (the CPU restatement under OpenMP on this
host's cores, on a bounded sample of the same workload).

Other workloads: --workload lattice|core|cube|hex (the plain path alone),
batch (512 slab problems), serve (the full served stream), exact (exact
regional cancellation of one generation's fission sites, GPU against the
CPU restatement, and the device loop with it),
--compare-precision (FP64 against split precision, per-bin deviations).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "collisions/s & histories/s vs plain GPU path at 1% tally err; p50/p99 latency"
HBM_PEAK_GBPS = 8000.0
# algorithmic bytes (per event = read + write of the 80 B FP64 particle
# record, 68 B in split precision with the direction in FP32; per tally
# update = one FP64 read-modify-write; per history = the 80 B source record)
BYTES_PER_EVENT = 160
BYTES_PER_EVENT_SPLIT = 136
BYTES_PER_TALLY_UPDATE = 16
BYTES_PER_HISTORY = 80
# the 17x17 lattice's inactive generations (they converge the fission source)
INACTIVE = 50


def workload(name, nparticles, tracking, precision):
    from abeille_amd import abi, problems
    if name == "lattice":
        m = problems.pin_lattice_2d(nparticles, tracking=tracking)
        desc = (f"17x17 pin lattice k-eigenvalue, 7 groups, {tracking} tracking, "
                f"{nparticles:.0e} particles/generation, TLE 51x51x7")
    elif name == "core":
        m = problems.core_3x3(nparticles, tracking=tracking)
        desc = (f"3x3-assembly core k-eigenvalue, 7 groups, {tracking} tracking, "
                f"{nparticles:.0e} particles/generation, TLE 153x153x30x7")
    elif name == "cube":
        m = problems.fixed_source_cube(nparticles, tracking=tracking)
        desc = (f"20 cm vacuum cube fixed source, 2 groups, {tracking} tracking, "
                f"{nparticles:.0e} histories/batch, collision tally 20^3x2")
    elif name == "hex":
        m, _ = problems.hex_assembly(nparticles, tracking=tracking, top="pointy", rings=9, nz=2)
        desc = (f"hexagonal assembly k-eigenvalue (9 rings, 217 pins, 2 layers), 7 groups, "
                f"{tracking} tracking, {nparticles:.0e} particles/generation, TLE 24x24x2x7")
    else:
        raise SystemExit(f"unknown workload {name}")
    if precision == "split":
        m.set(precision=abi.SPLIT)
    return m, desc


def measured_traffic(kernel, workload_name, precision):
    """HBM bytes per launch of `kernel` on this workload from the committed
    rocprofv3 counter summary (FETCH_SIZE and WRITE_SIZE per dispatch, each
    scaled by the copy-kernel calibration written in the same file), or
    None when no summary exists for it."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload_name}_{precision}_pmc.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1].split()[0])
    except OSError:
        pass
    return None


def algorithmic_bytes(diag, split=False):
    events = diag.flights + diag.collisions
    per_event = BYTES_PER_EVENT_SPLIT if split else BYTES_PER_EVENT
    return (events * per_event + diag.tally_updates * BYTES_PER_TALLY_UPDATE +
            diag.histories * BYTES_PER_HISTORY)


def cpu_baseline(name, tracking, seconds, nthreads):
    """Time the CPU restatement (oracle/) on a bounded sample: generations of
    a scaled-down bank of the same problem until about `seconds` of work.
    Threads are pinned close to each other (OMP_PROC_BIND=close, read when
    the OpenMP runtime loads with the oracle)."""
    os.environ.setdefault("OMP_PROC_BIND", "close")
    from oracle.pyoracle import Oracle
    n = 100000
    m, _ = workload(name, n, tracking, "fp64")
    o = Oracle(m, portable_math=False)
    o.run(2, 2, nthreads)  # warm up (threads, source)
    gens, coll, hist, t_used = 0, 0, 0, 0.0
    while t_used < seconds and gens < 400:
        st, _ = o.run(2, 1, nthreads)
        t_used += st.transport_seconds
        coll += st.diag.collisions
        hist += st.diag.histories
        gens += 2
    # one thread, on a shorter sample
    g1, c1, t1 = 0, 0, 0.0
    while t1 < max(1.0, seconds / 5) and g1 < 40:
        st, _ = o.run(1, 0, 1)
        t1 += st.transport_seconds
        c1 += st.diag.collisions
        g1 += 1
    return {"value": coll / t_used, "unit": "collisions/s", "cores": nthreads, "kind": "port",
            "sample": f"{gens} generations x {n} particles of the same problem "
                      f"({t_used:.1f} s of transport, glibc libm)",
            "histories_per_s": hist / t_used,
            "one_thread_collisions_per_s": c1 / t1 if t1 > 0 else None,
            "cpu_model": _cpu_model(), "omp_proc_bind": os.environ.get("OMP_PROC_BIND")}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def _common(args, metric_value, unit, ms_per_step, dtype, config, extra):
    out = {"metric": METRIC, "value": metric_value, "unit": unit, "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
           "data": "synthetic (seeded cross sections and geometry, no external data)",
           "config": config}
    out.update(extra)
    return out


def _c4_problems(count, precision, seed=512):
    """The batched-form problem set: `count` seeded slab problems, histories
    log-uniform in [1e3, 1e6] run as 2 batches each."""
    from abeille_amd import abi, problems
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        nhist = int(np.exp(rng.uniform(np.log(1e3), np.log(1e6))))
        m, _ = problems.slab_problem(rng, nhist=nhist, nbatch_hist=max(1, nhist // 2))
        m.set(precision=precision)
        out.append(m)
    return out


def rel_err_summary(avg, err, hist_active, hist_per_s, target=0.01):
    """Per-bin relative error of a tally (bins with a nonzero mean) and the
    active histories (and seconds at `hist_per_s`) that would bring 95 % of
    those bins to `target`: the error falls as 1/sqrt(histories)."""
    nz = avg > 0
    rel = err[nz] / avg[nz]
    q = {f"p{k}": float(np.percentile(rel, k)) for k in (50, 90, 95, 99)}
    need = hist_active * (q["p95"] / target) ** 2
    return {"bins": int(nz.sum()), "rel_err": q,
            "histories_to_1pct_p95": need, "seconds_to_1pct_p95": need / hist_per_s}


def bench_headline(args):
    """The headline: the 17x17 lattice k-eigenvalue problem (1e6 particles
    per generation, 7 groups, delta tracking, 51x51x7 track-length tally) on
    the plain FP64 path (one chain, the device-resident generation loop)
    and on the batched mixed-precision path (`--chains` chains of the same
    problem with their own seeds, split precision, one transport launch per
    generation for all chains).  Each path first converges its source with
    INACTIVE inactive generations of one chain (untimed; the batched chains
    all start from that bank), then `--warmup` untimed generations, then the
    timed ones: `--steps` generations of the batched path, chains x steps of
    the plain path, the same number of active histories, so that both
    tallies reach the same per-bin relative error (both reported).  A
    served-stream leg reports p50/p99 latency beside throughput."""
    from abeille_amd import Problem, abi, synchronize
    from abeille_amd.autotune import autotune
    N = int(args.particles)
    R = args.chains
    K = args.steps
    W = max(0, args.warmup)
    # ---- plain FP64 GPU path: one chain
    m, desc = workload("lattice", N, "delta", "fp64")
    plain = Problem(m)
    tuning = plain.tuning()
    if not args.no_autotune:
        tuning, _ = autotune(plain, m, sample=N, cache=None)
    plain.run(INACTIVE, nignored=INACTIVE)
    if W:
        plain.run_continue(W, nignored=W)
    synchronize()
    t0 = time.perf_counter()
    sp, _ = plain.run_continue(R * K, nignored=0)
    synchronize()
    wall_p = time.perf_counter() - t0
    hist_p = sp.histories
    avg_p, err_p = plain.tally_stats(0)
    plain_out = {"collisions_per_s": sp.diag.collisions / wall_p,
                 "histories_per_s": hist_p / wall_p,
                 "ms_per_generation": 1e3 * wall_p / (R * K), "generations": R * K,
                 "active_histories": hist_p, "kcol": [sp.kcol_avg, sp.kcol_err],
                 "roofline": _roofline(sp, False, "k_history")}
    plain_out.update(rel_err_summary(avg_p, err_p, hist_p, hist_p / wall_p))
    del plain
    # ---- batched split path: R chains started from one converged bank
    mb, _ = workload("lattice", N, "delta", "split")
    bat = Problem(mb)
    bat.set_tuning(**{k: tuning[k] for k in ("block_size", "blocks_per_cu", "stage_tables")})
    bat.run(INACTIVE, nignored=INACTIVE)
    bat.run_replicas(R, max(1, W), max(1, W), from_bank=True)
    synchronize()
    t0 = time.perf_counter()
    sb, _ = bat.run_replicas_continue(K, 0)
    synchronize()
    wall_b = time.perf_counter() - t0
    hist_b = sb.histories
    avg_b, err_b = bat.tally_stats(0)
    coll_b = sb.diag.collisions / wall_b
    bat_out = {"collisions_per_s": coll_b, "histories_per_s": hist_b / wall_b,
               "ms_per_generation_of_all_chains": 1e3 * wall_b / K, "generations": K,
               "chains": R, "active_histories": hist_b, "kcol": [sb.kcol_avg, sb.kcol_err]}
    bat_out.update(rel_err_summary(avg_b, err_b, hist_b, hist_b / wall_b))
    roof = _roofline(sb, True, "k_history (chains, split)")
    del bat
    # ---- served stream: p50 / p99 latency
    served = None if args.no_serve else serve_legs(args, args.requests)
    out = _common(args, coll_b, "collisions/s", 1e3 * wall_b / K, "f64/f32 split",
                  {"workload": f"{desc}; batched form: {R} chains (own seeds) in one launch "
                               "per generation, split precision; plain form: one chain, FP64",
                   "nparticles_per_chain": N, "chains": R, "inactive_generations": INACTIVE,
                   "tuning": tuning},
                  {"histories_per_s": hist_b / wall_b,
                   "speedup_vs_plain": coll_b / plain_out["collisions_per_s"],
                   "time_to_1pct_ratio": plain_out["seconds_to_1pct_p95"]
                   / bat_out["seconds_to_1pct_p95"],
                   "plain_fp64": plain_out, "batched_split": bat_out,
                   "latency_s": served,
                   "roofline": roof,
                   "issue": measured_issue("k_history", "lattice", "split_chains")})
    if not args.no_cpu:
        nthreads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        out["cpu_baseline"] = cpu_baseline("lattice", "delta", args.cpu_seconds, nthreads)
    print(json.dumps(out), flush=True)


def _roofline(st, split, kname):
    """HBM roofline of the transport kernel of a run: algorithmic bytes per
    launch (one launch per generation) over the HIP-event launch time."""
    launches = max(1, st.kernel_launches)
    kernel_avg = st.kernel_seconds / launches
    per_launch = algorithmic_bytes(st.diag, split) / launches
    achieved = per_launch / kernel_avg / 1e9 if kernel_avg > 0 else 0.0
    wl, prec = ("lattice", "split_chains") if split else ("lattice", "fp64")
    return {"bound": "hbm", "kernel": kname, "achieved": achieved, "peak": HBM_PEAK_GBPS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBPS,
            "traffic": measured_traffic("k_history", wl, prec),
            "avg_launch_ms": kernel_avg * 1e3, "launches": st.kernel_launches,
            "algorithmic_bytes_per_launch": per_launch}


def measured_issue(kernel, workload_name, precision):
    """Instruction-issue roofline of `kernel` from the committed rocprofv3
    counter summary: VALU instructions per event and the share of the
    chip's VALU issue slots the kernel uses (or None without a summary)."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload_name}_{precision}_pmc.md")
    out = {}
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("issue_") and ":" in line:
                    k, v = line.split(":", 1)
                    out[k[len("issue_"):].strip()] = float(v.split()[0])
    except OSError:
        return None
    return out or None


def bench_precision(args):
    """The workload once in FP64 and once in split precision: collisions/s of
    both and, per tally, the distribution of the per-bin differences in
    units of their combined standard error (z), and the mean relative
    deviation over the bins FP64 resolves to 1 %."""
    from abeille_amd import Problem, abi, synchronize
    nparticles = int(args.particles)
    res = {}
    # "fp64_seed2" is FP64 with another RNG seed: its z against FP64 is the
    # spread the statistics alone give (generation-to-generation correlation
    # of k-eigenvalue tallies makes it wider than 1), the yardstick for the
    # split run's z
    for prec in ("fp64", "split", "fp64_seed2"):
        m, desc = workload(args.workload, nparticles, args.tracking, prec.split("_")[0])
        if prec == "fp64_seed2":
            m.set(seed=m.settings["seed"] + 7919)
        p = Problem(m)
        p.run(max(1, args.warmup), nignored=max(1, args.warmup))
        synchronize()
        t0 = time.perf_counter()
        st, _ = p.run_continue(args.steps, nignored=0)
        synchronize()
        wall = time.perf_counter() - t0
        res[prec] = dict(rate=st.diag.collisions / wall, ms=1e3 * wall / args.steps,
                         k=(st.kcol_avg, st.kcol_err),
                         tallies=[p.tally_stats(t) for t in range(len(m.tallies))])
        del p
    def deviation(other):
        dev = []
        for (a64, e64), (a32, e32) in zip(res["fp64"]["tallies"], res[other]["tallies"]):
            sig = np.hypot(e64, e32)
            ok = sig > 0
            z = (a32[ok] - a64[ok]) / sig[ok]
            good = (a64 > 0) & (e64 < 0.01 * a64)
            rel = np.abs(a32[good] - a64[good]) / a64[good] if good.any() else None
            dev.append({"bins": int(ok.sum()), "z_mean": float(z.mean()),
                        "z_std": float(z.std()),
                        "frac_within_3sigma": float((np.abs(z) <= 3).mean()),
                        "bins_below_1pct_err": int(good.sum()),
                        "mean_rel_dev_those_bins": None if rel is None else float(rel.mean())})
        return dev
    dev = deviation("split")
    out = _common(args, res["split"]["rate"], "collisions/s", res["split"]["ms"], "f64/f32 split",
                  {"workload": desc + ", FP64 vs split precision", "nparticles": nparticles},
                  {"fp64_collisions_per_s": res["fp64"]["rate"],
                   "split_vs_fp64": res["split"]["rate"] / res["fp64"]["rate"],
                   "k_fp64": res["fp64"]["k"], "k_split": res["split"]["k"],
                   "k_fp64_seed2": res["fp64_seed2"]["k"],
                   "tally_deviation": dev,
                   "fp64_other_seed_deviation": deviation("fp64_seed2")})
    print(json.dumps(out), flush=True)


def bench_batch(args):
    """512 independent fixed-source slab problems: the batch engine (split
    precision, shared persistent launches) against the same problems run one
    at a time through the FP64 device loop."""
    from abeille_amd import Batch, Problem, abi, synchronize
    from abeille_amd.autotune import autotune_batch
    nb = 2
    split = [Problem(m) for m in _c4_problems(args.problems, abi.SPLIT)]
    # warm-up: a few problems through the batch (module load, grid sizing)
    w = Batch()
    for p in split[:8]:
        w.add(p, 1)
    while len(w):
        w.step()
    b = Batch()
    # chunk length and side-by-side cap from the autotuner, on a sample of
    # problems drawn from the same generator (not the timed set)
    params = None
    if not args.no_autotune:
        sample = [Problem(m) for m in _c4_problems(64, abi.SPLIT, seed=4096)]
        params, _ = autotune_batch(b, sample, nbatches=nb)
        del sample
    synchronize()
    t0 = time.perf_counter()
    for p in split:
        b.add(p, nb)
    steps = 0
    while len(b):
        b.step()
        steps += 1
    synchronize()
    wall = time.perf_counter() - t0
    coll = sum(p.stats().diag.collisions for p in split)
    hist = sum(p.stats().histories for p in split)
    kernel_s = b.kernel_seconds()
    del split
    # plain path: FP64, one problem at a time
    plain = [Problem(m) for m in _c4_problems(args.problems, abi.FP64)]
    plain[0].run(1, 0)
    t0 = time.perf_counter()
    pc = 0
    for p in plain:
        st, _ = p.run(nb, 0)
        pc += st.diag.collisions
    pwall = time.perf_counter() - t0
    out = _common(args, coll / wall, "collisions/s", 1e3 * wall / max(1, steps), "f64/f32 split",
                  {"workload": f"{args.problems} slab fixed-source problems (1-8 regions, G in "
                               "{2,4,7,16,33}, 1e3-1e6 histories, DT/ST, 16^3 collision "
                               "tally), 2 batches each, batched split precision",
                   "problems": args.problems, "batch_params": params},
                  {"histories_per_s": hist / wall, "wall_s": wall, "batch_steps": steps,
                   "transport_kernel_s": kernel_s,
                   "plain_fp64_one_at_a_time": {"collisions_per_s": pc / pwall, "wall_s": pwall},
                   "speedup_vs_plain": (coll / wall) / (pc / pwall)})
    print(json.dumps(out), flush=True)


def bench_exact(args):
    """Exact regional cancellation (ExactMGCancelator) of one k-eigenvalue
    generation: the three-group fuel | moderator | fuel slab under carter
    tracking (0.85 x the majorant, signed weights), --particles histories,
    8 x 2 x 2 regions x 2 energy regions, 10 Sobol points per region.  The
    same sites and parent records go through the GPU pass (abx_exact_cancel_bank:
    host bank in and out, so PCIe included) and the CPU restatement (one
    core); then the device-resident loop runs `--steps` generations with the
    cancellation in it."""
    from abeille_amd import Problem, problems, synchronize
    from oracle.pyoracle import Oracle
    n = int(args.particles)
    m = problems.exact_slab(n)
    m.exact_cancellation([-4., -5., -5.], [4., 5., 5.], [8, 2, 2], kind="exact",
                         group_bins=[[0], [1, 2]])
    g = Problem(m)
    src = g.sample_sources(0, n)
    r = g.transport(src, kcol=1.0, cap=4 * n + 4096)
    fis, par = r["fission"], g.parents()
    g.exact_cancel(fis, par)  # warm-up
    reps = max(1, args.steps // 3)
    synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        g.exact_cancel(fis, par)
    synchronize()
    gpu_s = (time.perf_counter() - t0) / reps
    g.rng_reset()  # the compared pass starts from settings::rng's seed, as the oracle's
    out = g.exact_cancel(fis, par)
    cpu = None
    if not args.no_cpu:
        o = Oracle(m)
        t0 = time.perf_counter()
        ref = o.exact_cancel(fis, par)
        cpu_s = time.perf_counter() - t0
        cpu = {"seconds": cpu_s, "cores": 1, "kind": "port",
               "sites_per_s": len(fis) / cpu_s, "same_bytes_as_gpu": ref.tobytes() == out.tobytes()}
    # the device loop with the cancellation in every generation
    p = Problem(m)
    t0 = time.perf_counter()
    st, kh = p.run(args.steps, min(2, args.steps - 1))
    synchronize()
    loop_s = time.perf_counter() - t0
    res = _common(args, len(fis) / gpu_s, "sites/s", 1e3 * gpu_s, "f64",
                  {"workload": f"exact regional cancellation of one generation of the "
                               f"3-group slab (carter 0.85), {n:.0e} histories, 8x2x2x2 "
                               "regions, 10 Sobol points"},
                  {"sites": int(len(fis)), "negative_sites": int((fis["wgt"] < 0).sum()),
                   "uniform_particles": int(len(out) - len(fis)),
                   "gpu_pass_s_pcie_inclusive": gpu_s, "cpu_baseline": cpu,
                   "device_loop": {"generations": args.steps, "wall_s": loop_s,
                                   "transport_s": st.transport_seconds,
                                   "histories": int(st.histories),
                                   "kcol_avg": st.kcol_avg}})
    print(json.dumps(res), flush=True)


def bench_cancel(args):
    """Approximate mesh cancellation (ApproximateMeshCancelator) of a
    synthetic bank of 2e6 sites (positions uniform in [-12, 12]^3, part of
    them outside the [-10, 10]^3 mesh; signed weights; wgt2 nonzero for 30%)
    on three meshes: one bin, 4 x 4 x 4 x 2 energy bins, 20 x 20 x 20.  Host
    bank in and out (abx_cancel_bank), so the pass time includes PCIe; the
    kernels' own durations come from a kernel trace of this command.  The
    CPU restatement (one core) cancels the same bank on every mesh."""
    import numpy as np
    from abeille_amd import Problem, abi, problems, synchronize
    n = 2_000_000
    rng = np.random.default_rng(5)
    b = np.zeros(n, abi.BANKED_DTYPE)
    b["r"] = rng.uniform(-12, 12, (n, 3))
    b["E"] = rng.choice([0.5, 1.5], n)
    b["wgt"] = rng.choice([-1.0, 1.0], n) * rng.uniform(0.5, 1.5, n)
    b["wgt2"] = np.where(rng.random(n) < 0.3, rng.normal(size=n), 0.0)
    b["parent_history_id"] = np.arange(n)
    meshes = {"1x1x1": ([1, 1, 1], None), "4x4x4x2": ([4, 4, 4], [0.0, 1.0, 2.0]),
              "20x20x20": ([20, 20, 20], None)}
    legs = {}
    for name, (shape, eb) in meshes.items():
        m = problems.fixed_source_cube(100)
        m.cancellation_mesh([-10, -10, -10], [10, 10, 10], shape, energy_bounds=eb)
        g = Problem(m)
        out = g.cancel_bank(b)  # warm-up
        reps = max(1, args.steps)
        synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            out = g.cancel_bank(b)
        synchronize()
        leg = {"gpu_pass_s_pcie_inclusive": (time.perf_counter() - t0) / reps}
        if not args.no_cpu:
            from oracle.pyoracle import Oracle
            t0 = time.perf_counter()
            ref = Oracle(m).cancel(b)
            leg["cpu_s"] = time.perf_counter() - t0
            leg["same_bytes_as_gpu"] = ref.tobytes() == out.tobytes()
        legs[name] = leg
    gpu_s = legs["4x4x4x2"]["gpu_pass_s_pcie_inclusive"]
    res = _common(args, n / gpu_s, "sites/s", 1e3 * gpu_s, "f64",
                  {"workload": "approximate mesh cancellation of 2e6 banked sites, "
                               "4x4x4 mesh x 2 energy bins (value); 1x1x1 and 20x20x20 "
                               "in meshes"},
                  {"meshes": legs, "cpu_baseline": None})
    print(json.dumps(res), flush=True)


def serve_legs(args, n_requests):
    """The served stream in short (the headline's latency legs): Poisson
    arrivals of `n_requests` requests of the served-stream generator, served
    with continuous batching in split precision, and the same requests
    served one at a time on the plain FP64 path, first come first served.
    Returns p50 / p99 latency and throughput of both."""
    from abeille_amd import Problem, abi, problems
    from abeille_amd.serve import Request, Server, latency_percentiles, serial_latencies
    rng = np.random.default_rng(2025)
    arrivals = np.cumsum(rng.exponential(1.0 / args.rate, n_requests))
    reqs, plain_models = [], []
    for i in range(n_requests):
        state = rng.bit_generator.state
        m, _, kind = problems.served_request(rng)
        rng2 = np.random.default_rng()
        rng2.bit_generator.state = state
        mp, _, _ = problems.served_request(rng2)
        m.set(precision=abi.SPLIT)
        reqs.append(Request(m, 2, float(arrivals[i]), name=f"{kind}{i}"))
        plain_models.append(mp)
    Server(keep_tallies=False).run([Request(reqs[0].model, 1, 0.0)])  # warm-up
    srv = Server(budget=64 << 30, keep_tallies=False)
    t0 = time.perf_counter()
    srv.run(reqs)
    wall = time.perf_counter() - t0
    coll = sum(r.stats.diag.collisions for r in reqs)
    service = np.zeros(n_requests)
    pc = 0
    for i, m in enumerate(plain_models):
        t = time.perf_counter()
        p = Problem(m)
        st, _ = p.run(2, 0)
        service[i] = time.perf_counter() - t
        pc += st.diag.collisions
        del p
    lat = serial_latencies(arrivals, service)
    makespan = float(np.max(arrivals + lat))
    return {"requests": n_requests, "rate_per_s": args.rate,
            "served_split": dict(latency_percentiles(reqs), collisions_per_s=coll / wall),
            "one_at_a_time_fp64": {"p50": float(np.percentile(lat, 50)),
                                   "p99": float(np.percentile(lat, 99)),
                                   "collisions_per_s": pc / makespan}}


def bench_serve(args):
    """Served stream: Poisson arrivals of slab requests (20 % carter tracking
    with a 0.8 x majorant sampling cross section and approximate mesh
    cancellation), continuous batching under an HBM budget, split precision;
    against the same requests served one at a time, FP64, first come first
    served."""
    from abeille_amd import Problem, abi, problems
    from abeille_amd.serve import Request, Server, latency_percentiles, serial_latencies
    rng = np.random.default_rng(2025)
    arrivals = np.cumsum(rng.exponential(1.0 / args.rate, args.requests))
    reqs, plain_models, kinds = [], [], []
    for i in range(args.requests):
        state = rng.bit_generator.state
        m, nh, kind = problems.served_request(rng)
        rng2 = np.random.default_rng()
        rng2.bit_generator.state = state
        mp, _, _ = problems.served_request(rng2)
        m.set(precision=abi.SPLIT)
        reqs.append(Request(m, 2, float(arrivals[i]), name=f"r{i}"))
        plain_models.append(mp)
        kinds.append(kind)
    # warm-up
    Server(keep_tallies=False).run([Request(reqs[0].model, 1, 0.0)])
    srv = Server(budget=64 << 30, keep_tallies=False, step_histories=args.step_histories)
    t0 = time.perf_counter()
    srv.run(reqs)
    wall = time.perf_counter() - t0
    coll = sum(r.stats.diag.collisions for r in reqs)
    hist = sum(r.stats.histories for r in reqs)
    pct = latency_percentiles(reqs)
    # One at a time on the plain FP64 path: the fixed-source requests here;
    # every request, carter chains included, in the headline's serve_legs
    plain_idx = [i for i in range(len(reqs)) if kinds[i] != "carter"]
    service = np.zeros(len(plain_idx))
    pc = 0
    for k, i in enumerate(plain_idx):
        m = plain_models[i]
        print(f"[serve] plain request {i} ({kinds[i]}, {m.settings['nparticles']} histories/batch)",
              file=sys.stderr, flush=True)
        t = time.perf_counter()
        p = Problem(m)          # staging the tables is part of serving a request
        st, _ = p.run(2, 0)
        service[k] = time.perf_counter() - t
        pc += st.diag.collisions
        del p
    arr_p = arrivals[plain_idx]
    lat = serial_latencies(arr_p, service)
    span = float(arrivals[-1])
    t = float(np.max(arr_p + lat))  # makespan of the one-at-a-time service
    served_fixed = [reqs[i] for i in plain_idx]
    pct_fixed = latency_percentiles(served_fixed)
    coll_fixed = sum(r.stats.diag.collisions for r in served_fixed)
    # the fixed-source requests served on their own (the stream the plain
    # leg serves): same arrivals, no carter chains sharing the steps
    alone = [Request(reqs[i].model, 2, float(arrivals[i]), name=f"r{i}") for i in plain_idx]
    srv2 = Server(budget=64 << 30, keep_tallies=False, step_histories=args.step_histories)
    t0 = time.perf_counter()
    srv2.run(alone)
    wall2 = time.perf_counter() - t0
    coll2 = sum(r.stats.diag.collisions for r in alone)
    out = _common(args, coll / wall, "collisions/s", 1e3 * wall / max(1, srv.steps),
                  "f64/f32 split",
                  {"workload": f"{args.requests} slab requests, Poisson {args.rate}/s, 1e4-5e6 "
                               "histories, 20% carter (0.8 x majorant) + 10^3 approximate "
                               "cancellation, continuous batching, 64 GiB budget, split precision",
                   "requests": args.requests, "step_histories": args.step_histories},
                  {"histories_per_s": hist / wall, "wall_s": wall, "arrival_span_s": span,
                   "latency_s": pct, "steps": srv.steps,
                   "step_wall_s": srv.step_seconds,
                   "free_s": {"total": srv.free_seconds, "max": srv.free_max},
                   "transport_kernel_s": srv.kernel_seconds(),
                   "carter_requests": int(sum(k == "carter" for k in kinds)),
                   "fixed_source_requests": {
                       "served_latency_s": pct_fixed, "served_collisions": coll_fixed,
                       "served_makespan_s": float(max(r.finished for r in served_fixed)),
                       "served_without_carter": {
                           "collisions_per_s": coll2 / wall2,
                           "makespan_s": float(max(r.finished for r in alone)),
                           "latency_s": latency_percentiles(alone)},
                       "plain_fp64_one_at_a_time": {
                           "collisions_per_s": pc / t,
                           "saturated_collisions_per_s": pc / float(service.sum()),
                           "makespan_s": t,
                           "latency_s": {"p50": float(np.percentile(lat, 50)),
                                         "p99": float(np.percentile(lat, 99))}}}})
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="headline",
                    choices=["headline", "lattice", "core", "cube", "hex", "batch", "serve",
                             "exact", "cancel"])
    ap.add_argument("--chains", type=int, default=12,
                    help="headline: chains of the batched k-eigenvalue form (1..16)")
    ap.add_argument("--compare-precision", action="store_true",
                    help="run the workload in FP64 and in split precision and report the "
                         "rates and every tally bin's deviation")
    ap.add_argument("--problems", type=int, default=512, help="batch: number of problems")
    ap.add_argument("--requests", type=int, default=None,
                    help="serve: number of requests (default 100; 40 in the headline)")
    ap.add_argument("--rate", type=float, default=50.0, help="serve: arrivals per second")
    ap.add_argument("--step-histories", type=int, default=0,
                    help="serve: histories per step over all requests (0 = whole batches)")
    ap.add_argument("--particles", type=float, default=1e6)
    ap.add_argument("--tracking", default="delta", choices=["delta", "surface"])
    ap.add_argument("--engine", default="history", choices=["event", "history"])
    ap.add_argument("--precision", default="fp64", choices=["fp64", "split"])
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shape instead of the autotuner's choice")
    ap.add_argument("--tuning", default=None,
                    help="fixed launch shape block_size,blocks_per_cu,stage_tables (no autotuning)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true", help="headline: skip the served-stream leg")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.copy_probe_mib:
        import torch
        n = args.copy_probe_mib * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n * 8}), flush=True)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is a "
                                   "design note only (histories split by id range, tally "
                                   "reduction), --gpus must be 1"}))
        sys.exit(2)

    if args.requests is None:
        args.requests = 40 if args.workload == "headline" else 100
    if args.compare_precision:
        return bench_precision(args)
    if args.workload == "headline":
        return bench_headline(args)
    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "serve":
        return bench_serve(args)
    if args.workload == "exact":
        return bench_exact(args)
    if args.workload == "cancel":
        return bench_cancel(args)
    from abeille_amd import Problem, abi, synchronize
    nparticles = int(args.particles)
    m, desc = workload(args.workload, nparticles, args.tracking, args.precision)
    engine = abi.ENGINE_EVENT if args.engine == "event" else abi.ENGINE_HISTORY
    prob = Problem(m, engine=engine)
    kmode = m.settings["mode"] == abi.K_EIGENVALUE
    tuning = prob.tuning()
    if args.tuning:
        bs, bpc, st = (int(x) for x in args.tuning.split(","))
        prob.set_tuning(block_size=bs, blocks_per_cu=bpc, stage_tables=st)
        tuning = prob.tuning()
    elif not args.no_autotune:
        from abeille_amd.autotune import EVENT_CANDIDATES, autotune
        if engine == abi.ENGINE_HISTORY:
            tuning, _ = autotune(prob, m, sample=nparticles, cache=None)
        else:   # the event engine: collision-queue sort cadence and tail threshold
            tuning, _ = autotune(prob, m, sample=nparticles, cache=None,
                                 candidates=EVENT_CANDIDATES)
    # warmup: inactive generations (also what converges the fission source)
    prob.run(max(1, args.warmup), nignored=max(1, args.warmup) if kmode else 0)
    synchronize()
    t0 = time.perf_counter()
    st, kh = prob.run_continue(args.steps, nignored=0)
    synchronize()
    wall = time.perf_counter() - t0
    d = st.diag
    coll_s = d.collisions / wall
    hist_s = st.histories / wall
    kernel_avg = st.kernel_seconds / max(1, st.kernel_launches)
    split = args.precision == "split"
    # algorithmic bytes of the dominant kernel per launch: the events it
    # processed (all flights) over its launches
    if args.engine == "history":
        # one launch transports one generation: all its events, tally updates
        # and histories
        bytes_per_launch = algorithmic_bytes(d, split) / max(1, st.kernel_launches)
    else:
        # the flight kernel: the events it processed
        bytes_per_launch = (st.kernel_items * (BYTES_PER_EVENT_SPLIT if split else BYTES_PER_EVENT)
                            / max(1, st.kernel_launches))
    achieved = bytes_per_launch / kernel_avg / 1e9 if kernel_avg > 0 else 0.0
    kname = "k_ev_flight" if args.engine == "event" else "k_history"
    traffic = measured_traffic(kname, args.workload, args.precision)
    out = {
        "metric": METRIC,
        "value": coll_s,
        "unit": "collisions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f64/f32 split",
        "data": "synthetic (seeded cross sections and geometry, no external data)",
        "config": {"workload": desc, "engine": args.engine, "nparticles": nparticles,
                   "generations_timed": args.steps, "tuning": tuning},
        "histories_per_s": hist_s,
        "events_per_s": (d.flights + d.collisions) / wall,
        "virtual_collisions_per_s": d.virtual_collisions / wall,
        "kcol_avg": st.kcol_avg,
        "kcol_err": st.kcol_err,
        "roofline": {
            "bound": "hbm",
            "kernel": kname,
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": traffic,
            "avg_launch_ms": kernel_avg * 1e3,
            "launches": st.kernel_launches,
            "algorithmic_bytes_per_launch": bytes_per_launch,
        },
        "whole_job_algorithmic_GBps": algorithmic_bytes(d, split) / wall / 1e9,
    }
    if not args.no_cpu:
        nthreads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        out["cpu_baseline"] = cpu_baseline(args.workload, args.tracking, args.cpu_seconds,
                                           nthreads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
