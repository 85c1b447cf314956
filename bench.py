#!/usr/bin/env python3
"""Throughput of the lattice-MC sampler: accepted MC moves/s, all replicas on one GPU.

Workload (the PTMC raster system): a 64-domain HalfTurn scaffold on an 8x8
boustrophedon raster in z=0, 32 two-domain staple types from a seeded domino
tiling, random 16-nt ACGT domain sequences, nearest-neighbour hybridization,
1 uM staples, 0.5 M cation, 32 replicas at 300-360 K (linear), temperature
exchange every 1000 steps.  One bench step is one exchange interval: every
replica runs `--interval` MC steps, then one even/odd exchange round.

Prints ONE JSON line.  `value` is accepted moves/s summed over all replicas
with the state resident in device memory; `roofline` prices the sweep
kernel's algorithmic bytes (counted events x fixed bytes per event) against
HBM peak, timed with HIP events on the sampler's stream; `cpu_baseline` is the
CPU oracle (a restatement of the reference's algorithm) timed on this host's
cores over a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "accepted MC moves/sec (all replicas, 1 GPU)"
HBM_PEAK_GBS = 8000.0
# algorithmic bytes per counted event: trial placement onto an empty site,
# trial placement that evaluates binding/stacking, domain commit, move attempt
BYTES_EMPTY, BYTES_BIND, BYTES_COMMIT, BYTES_MOVE = 24, 216, 32, 16


# LDS bandwidth: 128 B/clk/CU at 2.4 GHz; the chip's 256 CUs, or only the CUs
# a launch occupies (one workgroup per replica on a CU of its own)
LDS_PEAK_GBS_PER_CU = 128 * 2.4
N_CU = 256
# the rocprofv3 FETCH_SIZE / WRITE_SIZE passes of the raster sweep, scaled by
# a copy of known size; bench.py reads the measured HBM bytes per 1000-step
# launch from the summary's "traffic_bytes_per_launch" line
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "ldo_sweep_lds_kernel_raster64_ptmc32_pmc.md")
BATCH_TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "ldo_batch_sweep_kernel_batch4096.md")
MWUS3D_TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "ldo_sweep_lds_kernel_mwus3d.md")


def profiled_traffic(path=TRAFFIC_PROFILE):
    """HBM bytes per launch from a committed counter summary, or None."""
    try:
        with open(path) as f:
            for line in f:
                if line.strip().startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1].split()[0])
    except OSError:
        return None
    return None


def host_cores():
    """CPUs this process may run on (its affinity set), capped by the job
    limit the host sets for build and worker pools where it sets one."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    for var in ("MAX_JOBS", "OMP_NUM_THREADS"):
        v = os.environ.get(var, "")
        if v.isdigit() and int(v) > 0:
            n = min(n, int(v))
    return max(1, n)


def algorithmic_bytes(c0, c1):
    n_empty = c1.site_empty - c0.site_empty
    n_bind = c1.site_bind - c0.site_bind
    n_commit = c1.commits - c0.commits
    n_move = sum(c1.attempts) - sum(c0.attempts)
    return (BYTES_EMPTY * n_empty + BYTES_BIND * n_bind + BYTES_COMMIT * n_commit
            + BYTES_MOVE * n_move)


def _cpu_worker(args):
    """One process of the CPU baseline: the oracle on its share of the
    (replica, seed) items of the raster workload, each replica at its ladder
    temperature (as one MPI rank of the reference runs one replica)."""
    items, steps, untimed = args
    sys.path.insert(0, ROOT)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import dataclasses
    import pyoracle
    from latticednaorigami_amd import systems
    base = systems.raster_system()
    refs = []
    for r, seed in items:
        s = dataclasses.replace(base, seed=base.seed + 1000 * seed)
        refs.append(pyoracle.Oracle(s, temps=[base.temps[r]]))
    for ref in refs:
        ref.sweep(untimed)
    c0 = [ref.counters() for ref in refs]
    t0 = time.perf_counter()
    for ref in refs:
        ref.sweep(steps)
    dt = time.perf_counter() - t0
    acc = sum(sum(ref.counters().accepts) - sum(c.accepts) for ref, c in zip(refs, c0))
    att = sum(sum(ref.counters().attempts) - sum(c.attempts) for ref, c in zip(refs, c0))
    return acc, att, dt


def cpu_baseline(workers, steps, untimed, repeats, n_rep=32):
    """The oracle on `workers` concurrent processes over the same window of
    each replica's chain as the GPU line (`untimed` steps, then `steps`
    timed), for the 32 replicas x `repeats` seeds."""
    from concurrent.futures import ProcessPoolExecutor
    import multiprocessing as mp
    ctx = mp.get_context("spawn")
    items = [(r, k) for k in range(repeats) for r in range(n_rep)]
    shares = [items[w::workers] for w in range(workers)]
    with ProcessPoolExecutor(max_workers=workers, mp_context=ctx) as ex:
        # every worker builds its systems outside its own timed region; the
        # aggregate is the sum of the concurrent per-process rates
        res = list(ex.map(_cpu_worker, [(s, steps, untimed) for s in shares]))
    accepted = sum(r[0] for r in res)
    rate = sum(r[0] / r[2] for r in res)
    longest = max(r[2] for r in res)
    return {
        "value": rate,
        "unit": "moves/s",
        "cores": workers,
        "host_cpus": os.cpu_count(),
        "kind": "port",
        "sample": (f"{workers} concurrent processes over the 32 replicas of the raster workload "
                   f"x {repeats} seeds (each replica at its ladder temperature, as one MPI rank "
                   f"of the reference runs one replica; exchanges, one per 1000 steps, left "
                   f"out): steps {untimed}..{untimed + steps} of every chain, the GPU line's "
                   f"timed window ({accepted} accepted moves, longest timed region "
                   f"{longest:.1f} s)"),
    }


def _cpu_batch_worker(args):
    """One process of the batching CPU baseline: the oracle on every
    `stride`-th chain of the batch workload, `steps` MC steps each."""
    first, stride, n_chains, steps = args
    sys.path.insert(0, ROOT)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from latticednaorigami_amd import systems
    ss, ts = systems.batch_systems()
    refs = [pyoracle.Oracle(ss[k], temps=[ts[k]]) for k in range(first, len(ss), stride)][:n_chains]
    accepted = 0
    t0 = time.perf_counter()
    for ref in refs:
        ref.sweep(steps)
    dt = time.perf_counter() - t0
    for ref in refs:
        accepted += sum(ref.counters().accepts)
    return accepted, dt


def cpu_baseline_batch(workers, chains_per_worker, steps):
    from concurrent.futures import ProcessPoolExecutor
    import multiprocessing as mp
    ctx = mp.get_context("spawn")
    stride = 4096 // (workers * chains_per_worker)
    with ProcessPoolExecutor(max_workers=workers, mp_context=ctx) as ex:
        res = list(ex.map(_cpu_batch_worker,
                          [(w * stride, workers * stride, chains_per_worker, steps)
                           for w in range(workers)]))
    accepted = sum(r[0] for r in res)
    rate = sum(r[0] / r[1] for r in res)
    longest = max(r[1] for r in res)
    return {
        "value": rate,
        "unit": "moves/s",
        "cores": workers,
        "host_cpus": os.cpu_count(),
        "kind": "port",
        "sample": (f"{workers} concurrent processes, each {chains_per_worker} chains of the "
                   f"4096-chain batch (every {stride}-th chain, all lengths represented) for "
                   f"{steps} MC steps from the initial configuration ({accepted} accepted "
                   f"moves, longest timed region {longest:.1f} s)"),
    }


def gpu_leg_batch(args):
    """The batching workload: 4096 independent constant-T chains."""
    from latticednaorigami_amd import abi, systems
    from latticednaorigami_amd.sampler import BatchSampler
    ss, ts = systems.batch_systems()
    b = BatchSampler(ss, ts)
    chosen, tune_ms = b.tune(args.tune_steps)
    if args.batch_kernel is not None:
        b.set_kernel(args.batch_kernel)
        chosen = args.batch_kernel
    for _ in range(args.warmup):
        b.sweep(args.interval)
    c0 = b.counters()
    kernel_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        kernel_ms += b.sweep(args.interval)
    wall = time.perf_counter() - t0
    c1 = b.counters()
    accepted = sum(c1.accepts) - sum(c0.accepts)
    attempted = sum(c1.attempts) - sum(c0.attempts)
    nbytes = algorithmic_bytes(c0, c1)
    achieved = nbytes / (kernel_ms / 1e3) / 1e9 if kernel_ms > 0 else None
    traffic = (profiled_traffic(BATCH_TRAFFIC_PROFILE)
               if chosen == abi.BATCH_WAVE_SOLO and args.interval == 1000 else None)
    roofline = {
        "bound": "hbm",
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS if achieved is not None else None,
        # the counter profile's figure holds for the kernel it measured
        "traffic": traffic,
        "traffic_source": os.path.relpath(BATCH_TRAFFIC_PROFILE, ROOT) if traffic else None,
        "kernel": abi.BATCH_KERNEL_NAMES[chosen],
        "launches": args.steps,
        "avg_launch_ms": kernel_ms / max(1, args.steps),
        "algorithmic_bytes_per_launch": nbytes / max(1, args.steps),
    }
    extra = {
        "attempted_per_s": attempted / wall,
        "accepts_by_move": list(c1.accepts)[:4],
        "attempts_by_move": list(c1.attempts)[:4],
        "tune_ms": {abi.BATCH_KERNEL_NAMES[k]: v for k, v in tune_ms.items()},
    }
    b.close()
    return accepted / wall, wall, roofline, extra


def _cpu_mwus3d_worker(args):
    """One process of the 3D umbrella CPU baseline: the oracle on a slice of
    the 256 windows x temperatures, windows exchanged within each ladder."""
    first, n_rep, swaps, interval = args
    sys.path.insert(0, ROOT)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from latticednaorigami_amd import abi, systems
    s, wins, grid, glen = systems.mwus3d_system()
    temps = s.temps[first:first + n_rep]
    wins = wins[first:first + n_rep]
    s.temps = temps
    ref = pyoracle.Oracle(s, temps=temps, exchange_mode=abi.EXCHANGE_WINDOWS)
    ref.set_windows(wins, grid, glen)
    ref.run(1, interval)  # untimed: first-visit table builds
    a0 = sum(ref.counters().accepts)
    t0 = time.perf_counter()
    ref.run(swaps, interval)
    dt = time.perf_counter() - t0
    return sum(ref.counters().accepts) - a0, dt


def cpu_baseline_mwus3d(workers, reps_per_worker, swaps, interval):
    from concurrent.futures import ProcessPoolExecutor
    import multiprocessing as mp
    ctx = mp.get_context("spawn")
    stride = 256 // workers
    with ProcessPoolExecutor(max_workers=workers, mp_context=ctx) as ex:
        res = list(ex.map(_cpu_mwus3d_worker,
                          [(w * stride, reps_per_worker, swaps, interval)
                           for w in range(workers)]))
    accepted = sum(r[0] for r in res)
    rate = sum(r[0] / r[1] for r in res)
    longest = max(r[1] for r in res)
    return {
        "value": rate,
        "unit": "moves/s",
        "cores": workers,
        "host_cpus": os.cpu_count(),
        "kind": "port",
        "sample": (f"{workers} concurrent processes, each {reps_per_worker} consecutive windows "
                   f"of the 256 (64 windows x 4 T) for {swaps} exchange interval(s) of "
                   f"{interval} steps from the bare scaffold ({accepted} accepted moves, "
                   f"longest timed region {longest:.1f} s)"),
    }


def gpu_leg_mwus3d(args):
    """The 3D umbrella workload: 256 replicas (64 windows x 4 T), window
    exchange within each temperature and swaps of each window between
    neighbouring temperatures every `interval` steps."""
    from latticednaorigami_amd import abi, systems
    from latticednaorigami_amd.sampler import Sampler
    s, wins, grid, glen = systems.mwus3d_system()
    s.flags |= abi.FLAG_WINDOW_TEMPERATURE_SWAPS
    n_learned = args.mwus3d_learned
    train_info = None
    if n_learned > 0:
        # the learned proposal: samples from a plain run of every eighth
        # replica's temperature, a network over the 200 scaffold identities
        from latticednaorigami_amd import learned
        t0 = time.perf_counter()
        data, n_cfg = learned.collect_samples(s, s.temps[::8], 8, 500, seed_steps=5000)
        params, hist = learned.train(data, 200, args.learned_train_steps, batch=256, lr=1e-3)
        train_info = {"samples": int(len(data[1])), "train_steps": args.learned_train_steps,
                      "final_loss": float(hist[-1]) if len(hist) else None,
                      "collect_and_train_s": time.perf_counter() - t0}
        print(f"mwus3d: network trained on {len(data[1])} samples", file=sys.stderr, flush=True)
    smp = Sampler(s, temps=s.temps, exchange_mode=abi.EXCHANGE_WINDOWS)
    smp.set_windows(wins, grid, glen)
    if n_learned > 0:
        smp.learned_set_network(200, params, 0.1)

    def interval_round(n):
        for _ in range(n):
            if n_learned > 0:
                smp.learned_sweep(n_learned)
            smp.run(1, args.interval)
    if args.warmup > 0:
        interval_round(args.warmup)
    c0 = smp.counters()
    xa0, xc0 = smp.exchange_counts()
    smp.set_profiling(True)
    t0 = time.perf_counter()
    interval_round(args.steps)
    wall = time.perf_counter() - t0
    prof = smp.profile()
    smp.set_profiling(False)
    c1 = smp.counters()
    xa1, xc1 = smp.exchange_counts()
    accepted = sum(c1.accepts) - sum(c0.accepts)
    attempted = sum(c1.attempts) - sum(c0.attempts)
    nbytes = algorithmic_bytes(c0, c1)
    kernel_s = prof["sweep_ms"] / 1e3
    achieved = nbytes / kernel_s / 1e9 if kernel_s > 0 else None
    roofline = {
        "bound": "hbm",
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS if achieved is not None else None,
        # HBM bytes per launch from the committed counter summary of this
        # workload (500-step intervals)
        "traffic": (profiled_traffic(MWUS3D_TRAFFIC_PROFILE)
                    if args.interval == 500 and n_learned == 0 else None),
        "kernel": smp.lib.ldo_kernel_name(smp.h).decode(),
        "launches": prof["sweep_launches"],
        "avg_launch_ms": prof["sweep_ms"] / max(1, prof["sweep_launches"]),
        "algorithmic_bytes_per_launch": nbytes / max(1, prof["sweep_launches"]),
    }
    extra = {
        "attempted_per_s": attempted / wall,
        "accepts_by_move": list(c1.accepts)[:4],
        "attempts_by_move": list(c1.attempts)[:4],
        "exchange_attempts": c1.exchange_attempts - c0.exchange_attempts,
        "exchange_accepts": c1.exchange_accepts - c0.exchange_accepts,
        "exchange_kernel_ms": prof["exchange_ms"],
        "temperature_swap_attempts": int((xa1[1] - xa0[1]).sum()),
        "temperature_swap_accepts": int((xc1[1] - xc0[1]).sum()),
    }
    if n_learned > 0:
        # counter slots 5, 6, 7: learned staple exchange, CB fallback, learned regrowth
        extra["learned_attempts_by_slot"] = [c1.attempts[k] - c0.attempts[k] for k in (5, 6, 7)]
        extra["learned_accepts_by_slot"] = [c1.accepts[k] - c0.accepts[k] for k in (5, 6, 7)]
        extra["learned_training"] = train_info
    smp.close()
    return accepted / wall, wall, roofline, extra


def gpu_leg_learned(args):
    """The learned proposal on the PTMC raster system: collect samples from
    the plain sampler, train the network on the device, then compare
    effective samples per second of the plain and the learned form."""
    import dataclasses
    import numpy as np
    from latticednaorigami_amd import learned, systems
    from latticednaorigami_amd.sampler import Sampler
    if args.learned_system == "four_domain":
        # 345 K: the enumerable system's melting region, where the staples
        # bind and unbind (P(numfulldomains) = 0.29 / 0.53 / 0.17 at 0 / 2 / 4)
        s = systems.four_domain_system(max_total_staples=2)
        temps = [args.learned_eval_temp or 345.0] * 32
        n_id = 4
    else:
        s = systems.raster_system()
        temps = s.temps
        n_id = 64
    # collection may run more replicas than the evaluation (each takes the
    # temperatures of the evaluation ladder in turn)
    n_col = max(len(temps), args.learned_collect_replicas)
    col_temps = [temps[k % len(temps)] for k in range(n_col)]
    t0 = time.perf_counter()
    data, n_cfg = learned.collect_samples(s, col_temps, args.learned_rounds, 500,
                                          seed_steps=args.learned_seed_steps)
    t_collect = time.perf_counter() - t0
    n = len(data[1])
    print(f"learned: {n} samples from {n_cfg} configurations in {t_collect:.1f} s",
          file=sys.stderr, flush=True)
    rng = np.random.default_rng(0)
    perm = rng.permutation(n)
    hold = perm[: max(1, n // 10)]
    fit = perm[max(1, n // 10):]
    split = lambda idx: tuple(a[idx] for a in data)  # noqa: E731
    t0 = time.perf_counter()
    params, hist = learned.train(split(fit), n_id, args.learned_train_steps, batch=256, lr=1e-3,
                                 log_every=500)
    t_train = time.perf_counter() - t0
    held = learned.mean_loss(params, n_id, split(hold))
    uniform = float(np.mean(np.log(split(hold)[2].sum(axis=1))))
    keys = ["num_fully_bound_domain_pairs", "num_staples"]
    res = {}
    eval_temp = args.learned_eval_temp
    if eval_temp is None:
        eval_temp = 345.0 if args.learned_system == "four_domain" else 350.0
    eval_temps = [eval_temp] * len(temps)
    # the reference distribution of the accuracy protocol: the exact
    # enumeration for the four-domain system (tests/golden), else a long
    # plain run with its own seed
    exact = None
    if args.learned_system == "four_domain":
        with open(os.path.join(ROOT, "tests", "golden", "four_domain_enumeration.json")) as f:
            exact = json.load(f)["p_by_temperature"].get(f"{eval_temp:g}")
    t_ref = 0.0
    if exact is not None:
        ref_series = None
    else:
        ref = Sampler(dataclasses.replace(s, seed=s.seed + 7919), temps=eval_temps)
        ref.sweep(args.learned_seed_steps)
        t0 = time.perf_counter()
        ref_series, _ = learned.run_series(ref, args.learned_ref_rounds, args.interval, 0,
                                           keys[:1])
        t_ref = time.perf_counter() - t0
        ref.close()

    def f_error(vals):
        if exact is None:
            return learned.free_energy_gap(vals, ref_series[keys[0]], floor=0.01)
        v = np.concatenate([np.asarray(x) for x in vals]).astype(int)
        p = np.bincount(v, minlength=len(exact))[:len(exact)] / len(v)
        pe = np.asarray(exact)
        ok = pe >= 0.01
        if (p[ok] <= 0).any():
            return None
        fa, fb = -np.log(p[ok]), -np.log(pe[ok])
        return float(np.abs((fa - fa.min()) - (fb - fb.min())).max())
    for form in ("plain", "learned"):
        smp = Sampler(s, temps=eval_temps)
        smp.sweep(args.learned_seed_steps)
        if form == "learned":
            smp.learned_set_network(n_id, params, 0.1)
            la = args.learned_attempts
        else:
            la = 0
        # each form relaxes under its own moves before the measured window
        learned.run_series(smp, args.learned_eval_rounds // 4, args.interval, la, keys,
                           progress=f"learned: {form} form relaxing")
        series, wall = learned.run_series(smp, args.learned_eval_rounds, args.interval, la, keys,
                                          progress=f"learned: {form} form")
        c = smp.counters()
        print(f"learned: {form} form evaluated in {wall:.1f} s", file=sys.stderr, flush=True)
        # free-energy error of numfulldomains against the reference after
        # each twentieth of the run; the time to 0.1 kT is the first
        # checkpoint after which the error stays at or below it
        rounds = args.learned_eval_rounds
        marks = sorted({max(1, rounds * q // 20) for q in range(1, 21)})
        errs = [f_error([x[:k] for x in series[keys[0]]]) for k in marks]
        t_tol = None
        for q, k in enumerate(marks):
            if all(e is not None and e <= 0.1 for e in errs[q:]):
                t_tol = wall * k / rounds
                break
        res[form] = {
            "seconds": wall,
            "free_energy_error_kT": errs[-1],
            "error_curve_kT": errs,
            "seconds_to_0.1_kT": t_tol,
            "ess_per_s": {k: learned.ess(series[k]) / wall for k in keys},
            "series": series,
            "learned_regrowth": [c.accepts[7], c.attempts[7]],
            "learned_exchange": [c.accepts[5], c.attempts[5]],
            "cb_fallback": [c.accepts[6], c.attempts[6]],
            "met_exchange": [c.accepts[1], c.attempts[1]],
            "accepted_moves_per_s": sum(c.accepts) / wall,
        }
        smp.close()
    gap = learned.free_energy_gap(res["plain"]["series"][keys[0]],
                                  res["learned"]["series"][keys[0]])
    ratio = {k: res["learned"]["ess_per_s"][k] / res["plain"]["ess_per_s"][k]
             if res["plain"]["ess_per_s"][k] > 0 else None for k in keys}
    for f in res.values():
        f.pop("series")
    return res, ratio, gap, {"samples": n, "configurations": n_cfg,
                             "reference": ("exact enumeration" if exact is not None
                                           else f"plain run of {args.learned_ref_rounds} rounds"),
                             "reference_s": t_ref,
                             "collect_s": t_collect, "train_s": t_train,
                             "train_steps": args.learned_train_steps,
                             "first_loss": float(np.mean(hist[:20])),
                             "last_loss": float(np.mean(hist[-20:])),
                             "heldout_loss": held, "uniform_loss": uniform}


def gpu_leg(args, system):
    from latticednaorigami_amd import abi
    from latticednaorigami_amd.sampler import Sampler
    smp = Sampler(system, temps=system.temps, exchange_mode=abi.EXCHANGE_TEMPERATURE)
    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except ImportError:
        sync = lambda: None  # noqa: E731
    if args.warmup > 0:
        smp.run(args.warmup, args.interval)
    c0 = smp.counters()
    smp.set_profiling(True)
    sync()
    t0 = time.perf_counter()
    smp.run(args.steps, args.interval)
    sync()
    wall = time.perf_counter() - t0
    prof = smp.profile()
    smp.set_profiling(False)
    c1 = smp.counters()
    accepted = sum(c1.accepts) - sum(c0.accepts)
    attempted = sum(c1.attempts) - sum(c0.attempts)
    nbytes = algorithmic_bytes(c0, c1)
    kernel_s = prof["sweep_ms"] / 1e3
    achieved = nbytes / kernel_s / 1e9 if kernel_s > 0 else None
    traffic = profiled_traffic() if args.interval == 1000 else None
    n_cu = min(N_CU, len(system.temps))
    roofline = {
        "bound": "hbm",
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS if achieved is not None else None,
        "traffic": traffic,
        "traffic_source": os.path.relpath(TRAFFIC_PROFILE, ROOT) if traffic else None,
        "kernel": smp.lib.ldo_kernel_name(smp.h).decode(),
        "launches": prof["sweep_launches"],
        "avg_launch_ms": prof["sweep_ms"] / max(1, prof["sweep_launches"]),
        "algorithmic_bytes_per_launch": nbytes / max(1, prof["sweep_launches"]),
        # the replica state is LDS-resident: the same bytes priced against LDS
        # bandwidth, of the chip and of the CUs the launch occupies
        "lds": {
            "achieved": achieved,
            "peak_chip": LDS_PEAK_GBS_PER_CU * N_CU,
            "frac_chip": achieved / (LDS_PEAK_GBS_PER_CU * N_CU) if achieved else None,
            "peak_occupied_cus": LDS_PEAK_GBS_PER_CU * n_cu,
            "frac_occupied_cus": achieved / (LDS_PEAK_GBS_PER_CU * n_cu) if achieved else None,
            "occupied_cus": n_cu,
        },
    }
    extra = {
        "attempted_per_s": attempted / wall,
        "accepts_by_move": list(c1.accepts)[:4],
        "attempts_by_move": list(c1.attempts)[:4],
        "exchange_kernel_ms": prof["exchange_ms"],
    }
    smp.close()
    return accepted / wall, wall, roofline, extra


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed exchange intervals")
    ap.add_argument("--warmup", type=int, default=2, help="untimed exchange intervals")
    ap.add_argument("--interval", type=int, default=None,
                    help="MC steps per replica per step (default 1000; mwus3d: 500, its "
                         "exchange interval)")
    ap.add_argument("--cpu-repeats", type=int, default=16,
                    help="raster: seeds per replica in the CPU baseline (each runs the GPU "
                         "line's warm-up and timed steps)")
    ap.add_argument("--cpu-workers", type=int, default=0,
                    help="CPU-baseline processes (0: every available host core, at most one "
                         "per replica)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", default="raster64_ptmc32",
                    choices=["raster64_ptmc32", "batch4096", "mwus3d", "learned", "hbm_copy"],
                    help="raster64_ptmc32: the 32-replica PTMC raster run; batch4096: 4096 "
                         "independent constant-T chains of 8-128 domains; mwus3d: the "
                         "200-domain 3D scaffold, 64 umbrella windows x 4 T")
    ap.add_argument("--tune-steps", type=int, default=100,
                    help="batch4096: MC steps per chain in each autotuner trial")
    ap.add_argument("--batch-kernel", type=int, default=None, choices=[0, 1, 2],
                    help="batch4096: force a sweep kernel instead of the autotuner's choice")
    ap.add_argument("--cpu-chains", type=int, default=64,
                    help="batch4096: chains per CPU-baseline process")
    ap.add_argument("--cpu-batch-steps", type=int, default=60000,
                    help="batch4096: MC steps per chain in the CPU baseline")
    ap.add_argument("--cpu-swaps", type=int, default=1000,
                    help="mwus3d: exchange intervals in each CPU-baseline process")
    ap.add_argument("--learned-rounds", type=int, default=200,
                    help="learned: sample-collection rounds (500 steps each)")
    ap.add_argument("--learned-collect-replicas", type=int, default=32,
                    help="learned: replicas of the sample collection (configurations = "
                         "replicas x rounds)")
    ap.add_argument("--learned-seed-steps", type=int, default=100000,
                    help="learned: plain steps before collection and before each evaluation")
    ap.add_argument("--learned-eval-temp", type=float, default=None,
                    help="learned: temperature of every replica in the evaluation (default: "
                         "350 K on the raster, 345 K on the four-domain system)")
    ap.add_argument("--learned-system", default="raster", choices=["raster", "four_domain"],
                    help="learned: the 64-domain raster or the enumerable four-domain system")
    ap.add_argument("--learned-ref-rounds", type=int, default=800,
                    help="learned: rounds of the plain reference run of the accuracy protocol")
    ap.add_argument("--learned-train-steps", type=int, default=3000)
    ap.add_argument("--mwus3d-learned", type=int, default=0,
                    help="mwus3d: learned attempts per replica before each exchange interval "
                         "(0: the plain move set)")
    ap.add_argument("--learned-eval-rounds", type=int, default=200,
                    help="learned: evaluation rounds (--interval plain steps each)")
    ap.add_argument("--learned-attempts", type=int, default=100,
                    help="learned: learned attempts per replica per evaluation round")
    args = ap.parse_args()
    if args.interval is None:
        args.interval = 500 if args.workload == "mwus3d" else 1000

    if args.gpus != 1:
        print(f"bench.py measures one GPU: every replica of the run is resident on a single "
              f"device, so --gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    if args.workload == "batch4096":
        return main_batch(args)
    if args.workload == "mwus3d":
        return main_mwus3d(args)
    if args.workload == "learned":
        return main_learned(args)
    if args.workload == "hbm_copy":
        return main_hbm_copy(args)
    from latticednaorigami_amd import systems
    system = systems.raster_system()
    line = {
        "metric": METRIC,
        "value": None,
        "unit": "moves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "raster64_ptmc32",
            "scaffold_domains": 64,
            "staple_types": 32,
            "replicas": 32,
            "temps_K": [300.0, 360.0],
            "exchange_interval": args.interval,
            "hybridization": "NearestNeighbour",
            "step": "one exchange interval on every replica",
        },
        "roofline": None,
        "cpu_baseline": None,
    }
    status = 0
    try:
        value, wall, roofline, extra = gpu_leg(args, system)
        line["value"] = value
        line["ms_per_step"] = wall / args.steps * 1e3
        line["roofline"] = roofline
        line["config"].update(extra)
    except Exception as e:  # the engine refused or failed: report, do not fake a number
        line["error"] = f"{type(e).__name__}: {e}"
        status = 1
    if not args.no_cpu_baseline:
        workers = args.cpu_workers or min(32, host_cores())
        line["cpu_baseline"] = cpu_baseline(workers, args.steps * args.interval,
                                            args.warmup * args.interval, args.cpu_repeats)
    print(json.dumps(line), flush=True)
    sys.exit(status)


def main_batch(args):
    line = {
        "metric": "accepted MC moves/sec (all chains, 1 GPU)",
        "value": None,
        "unit": "moves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "batch4096",
            "chains": 4096,
            "scaffold_domains": [8, 128],
            "temps_K": [310.0, 350.0],
            "hybridization": "Uniform",
            "step": f"{args.interval} MC steps on every chain (one sweep launch)",
        },
        "roofline": None,
        "cpu_baseline": None,
    }
    status = 0
    try:
        value, wall, roofline, extra = gpu_leg_batch(args)
        line["value"] = value
        line["ms_per_step"] = wall / args.steps * 1e3
        line["roofline"] = roofline
        line["config"].update(extra)
    except Exception as e:
        line["error"] = f"{type(e).__name__}: {e}"
        status = 1
    if not args.no_cpu_baseline:
        workers = args.cpu_workers or min(32, host_cores())
        line["cpu_baseline"] = cpu_baseline_batch(workers, args.cpu_chains, args.cpu_batch_steps)
    print(json.dumps(line), flush=True)
    sys.exit(status)


def main_mwus3d(args):
    line = {
        "metric": "accepted MC moves/sec (all replicas, 1 GPU)",
        "value": None,
        "unit": "moves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "mwus3d",
            "scaffold_domains": 200,
            "domain_type": "ThreeQuarterTurn",
            "staple_types": 67,
            "replicas": 256,
            "windows": 64,
            "temps_K": [310.0, 315.0, 320.0, 325.0],
            "exchange_interval": args.interval,
            "exchange": "windows within each temperature, then each window between "
                        "neighbouring temperatures",
            "hybridization": "NearestNeighbour",
            "learned_proposal": args.mwus3d_learned > 0,
            "learned_attempts_per_interval": args.mwus3d_learned,
            "step": "one exchange interval on every replica (after its learned attempts)",
        },
        "roofline": None,
        "cpu_baseline": None,
    }
    status = 0
    try:
        value, wall, roofline, extra = gpu_leg_mwus3d(args)
        line["value"] = value
        line["ms_per_step"] = wall / args.steps * 1e3
        line["roofline"] = roofline
        line["config"].update(extra)
    except Exception as e:
        line["error"] = f"{type(e).__name__}: {e}"
        status = 1
    if not args.no_cpu_baseline:
        workers = args.cpu_workers or min(32, host_cores())
        line["cpu_baseline"] = cpu_baseline_mwus3d(workers, 8, args.cpu_swaps, args.interval)
    print(json.dumps(line), flush=True)
    sys.exit(status)


def main_learned(args):
    line = {
        "metric": "ESS/sec of numfulldomains (all replicas, 1 GPU), learned vs plain form",
        "value": None,
        "unit": "effective samples/s",
        "n_gpus": 1,
        "steps": args.learned_eval_rounds,
        "warmup": args.learned_seed_steps,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 chain, bf16/fp32 network",
        "data": "synthetic",
        "config": {
            "workload": ("learned_four_domain" if args.learned_system == "four_domain"
                         else "learned_raster64"),
            "system": ("four-domain system (Uniform), samples from and evaluation with 32 "
                       f"replicas at {args.learned_eval_temp or 345.0} K, no exchange"
                       if args.learned_system == "four_domain" else
                       "64-domain raster, NN; samples from 32 replicas at 300-360 K, "
                       f"evaluation with 32 replicas at {args.learned_eval_temp or 350.0} K, "
                       "no exchange"),
            "protocol": ("accuracy: free-energy error of numfulldomains (max |dF| over values "
                         "with P >= 0.01) against the exact enumeration (four-domain system) "
                         f"or a {args.learned_ref_rounds}-round plain run with its own seed "
                         "(raster); time until the error stays within 0.1 kT; ESS/s over the "
                         "whole evaluation"),
            "network": "7x7x7x13 one-hot + 16-dim identity embedding, 3 x 256 ReLU, 36 outputs",
            "fallback_cb": 0.1,
            "round": f"{args.learned_attempts} learned attempts + {args.interval} plain steps "
                     f"per replica (plain form: the plain steps only)",
        },
        "roofline": None,
        "cpu_baseline": None,
    }
    status = 0
    try:
        res, ratio, gap, train = gpu_leg_learned(args)
        line["value"] = res["learned"]["ess_per_s"]["num_fully_bound_domain_pairs"]
        line["ms_per_step"] = res["learned"]["seconds"] / args.learned_eval_rounds * 1e3
        line["config"].update({"forms": res, "ess_ratio_learned_over_plain": ratio,
                               "free_energy_gap_kT": gap, "training": train})
    except Exception as e:
        line["error"] = f"{type(e).__name__}: {e}"
        status = 1
    print(json.dumps(line), flush=True)
    sys.exit(status)


def main_hbm_copy(args):
    """Calibration for the HBM traffic counters: `--steps` device copies of a
    256 MiB buffer (256 MiB read and 256 MiB written per copy), so that
    rocprofv3's FETCH_SIZE / WRITE_SIZE per dispatch of this copy give the
    factor from counter units to bytes."""
    import torch
    n = 64 << 20
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dst.copy_(src)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    print(json.dumps({"workload": "hbm_copy", "bytes_read_per_copy": 4 * n,
                      "bytes_written_per_copy": 4 * n, "copies": args.steps,
                      "GB_per_s": 8 * n * args.steps / dt / 1e9}), flush=True)


if __name__ == "__main__":
    main()
