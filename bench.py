"""Benchmark: pair.slice action evaluations per second of the GPU-resident bisection chain.

Workload (the configuration the headline metric is quoted on): Lennard-Jones
particles of mass 4 amu, N=256, M=64 time slices, tau=0.0125 K^-1, level-4
bisection moves, run as R independent replicas (PIMC++ clones) on one GPU.
One step = `moves_per_step` bisection moves on every replica, in one kernel
launch (one block per replica).  A pair.slice evaluation is one U(q,z,s^2) for
one (moved particle, partner, link, OLD/NEW) -- an innermost iteration of
ShortRangeClass::SingleAction.  `value` is the mixed-precision throughput;
the FP64 path is measured the same way and reported beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import subprocess
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from pimcplusplus_amd import abi, native, workloads  # noqa: E402

METRIC = "pair·slice action evals/s, mixed vs FP64 on GPU at ΔS err ≤1e-6"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
LDS_PEAK_GBS = 256 * 128 * 2.4      # 256 CUs x 128 B/clk x 2.4 GHz
CALIB_LIB = os.path.join(ROOT, "pimcplusplus_amd", "_lib", "libpimc_calib.so")
FLOP_PER_EVAL = 180.0      # FP64 flops of one pair·slice evaluation (DistDisp, q/z/s², splines, sum)
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_bisection_lj_n256_m64_l4_traffic.md")
# Replicas of the headline workload: 4 blocks per CU x 256 CUs are resident at once, so 16384
# replicas run as 16 rounds of blocks (4096 replicas, 4 rounds, spend ~10% of a launch in its
# tail, where the last blocks finish one by one)
HEADLINE_REPLICAS = 16384


def traffic_profile(workload: str) -> str:
    """Committed FETCH_SIZE/WRITE_SIZE summary of the generic kernel on `workload`."""
    return os.path.join(ROOT, "profiles", f"k_bisection_generic_{workload}_traffic.md")


def measured_traffic(want: str, path: str = TRAFFIC_PROFILE):
    """HBM bytes per launch from a committed counter summary (FETCH_SIZE and WRITE_SIZE of
    the same workload, scaled by a copy kernel of known size), when its workload string
    holds `want` (replica count, moves per launch and, for a fixed plan, the plan)."""
    try:
        text = open(path).read()
    except OSError:
        return None
    block = text.split("```json", 1)[1].split("```", 1)[0] if "```json" in text else ""
    try:
        d = json.loads(block)
    except ValueError:
        return None
    return d.get("traffic_bytes_per_launch") if want in d.get("workload", "") else None


def bytes_per_launch(st, n: int) -> float:
    """Algorithmic HBM bytes of the launches summarised by `st`: every executed
    stage reads the K+1 slices of N+1 FP64 beads (partners, plus the moved
    particle's OLD and NEW copies counted once more) and returns two FP64
    actions: 24 (K+1)(N+1) + 16 bytes.  sum_stages K = pair_evals / (2 (N-1))."""
    sum_k = st.pair_evals / (2.0 * (n - 1))
    return 24.0 * (n + 1) * (sum_k + st.stages) + 16.0 * st.stages


def measured_peaks():
    """HBM copy and LDS read bandwidth measured on this GPU by the microbenchmarks of
    libpimc_calib.so (GB/s), or None where the library is missing or a call fails."""
    if not os.path.exists(CALIB_LIB):
        return None
    lib = ctypes.CDLL(CALIB_LIB)
    lib.pimc_calib_hbm.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    lib.pimc_calib_lds.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)]
    hbm, lds, cus = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
    out = {}
    if lib.pimc_calib_hbm(2 << 30, 10, ctypes.byref(hbm)) == 0:
        out["hbm_copy_GBs"] = hbm.value / 1e9
    if lib.pimc_calib_lds(8192, ctypes.byref(lds), ctypes.byref(cus)) == 0:
        out["lds_read_GBs"] = lds.value / 1e9
        out["cus"] = cus.value
    out["note"] = ("HBM: grid-stride copy of 2 GiB (read + write bytes); LDS: 16-byte lane-consecutive "
                   "reads, eight 256-thread blocks per CU")
    return out


def offdiag_fraction(paths, box, end: float) -> float:
    """Fraction of pair.slice distances below the table end, over sampled paths: the share of
    evaluations that look up the off-diagonal columns (q < End)."""
    L = np.asarray(box, dtype=float)
    inside = total = 0
    for p in paths:
        for s in range(0, p.shape[0], 4):
            d = p[s][:, None, :] - p[s][None, :, :]
            d -= L * np.rint(d / L)
            r = np.sqrt(np.sum(d * d, axis=-1))
            iu = np.triu_indices(p.shape[1], 1)
            inside += int(np.count_nonzero(r[iu] < end))
            total += iu[0].size
    return inside / max(1, total)


def lds_bytes_per_launch(st, steps: int, mixed: bool, f_off: float, nint: int, ncol: int, noff: int) -> float:
    """LDS bytes of the launches summarised by `st` (per launch): per evaluation the moved
    bead (24 B, one broadcast row), the u00 row of the new slice and, with probability
    f_off, the off-diagonal rows (16 B per row in the mixed path, 32 B in FP64); per
    executed stage the staged level table written into LDS."""
    row = 16.0 if mixed else 32.0
    per_eval = 24.0 + row * (1.0 + noff * f_off)
    return (st.pair_evals * per_eval + st.stages * nint * ncol * row) / steps


def load_system(lib, system):
    ids = workloads.load(lib, system)
    return ids


PROGRESS = False   # --progress: a line on stderr at the end of each phase


def progress(msg: str):
    """A line on stderr when a phase of the run ends (--progress), so that long set-ups of many
    replicas show that they advance; stdout keeps the one JSON line."""
    if PROGRESS:
        print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_gpu(system, precision, steps, warmup, moves_per_step, tune=False, graph=False, paths_out=None, plan=None,
            setup=None):
    """W untimed steps, then K steps queued back to back on the context's stream;
    the timed region is bracketed by stream synchronisations (the stats take).
    tune: the autotuner picks the launch plan first (on moves the chain then
    repeats); graph: each step is a replay of one captured sweep.  paths_out: a list
    that receives the paths of the first eight replicas after the run."""
    lib = native.open_context(0)
    ids = load_system(lib, system)
    progress(f"{system.name}: {len(ids)} replicas loaded")
    if setup:
        setup(lib, ids)
    L, seed = system.num_levels, system.seed
    if plan is not None:
        lib.set_plan(*plan)
    plan = lib.bisection_tune(ids, L, seed, 0, moves_per_step, precision) if tune else lib.get_plan()
    if tune:
        progress(f"{system.name}: tuned plan {plan.block_threads}, {plan.link_split}, {plan.persistent}, {plan.cluster}")
    move = 0
    if graph:
        g = lib.sweep_graph_create(ids, L, seed, 0, moves_per_step, precision)
        step = lambda: lib.sweep_graph_launch(g, 1)  # noqa: E731
    else:
        def step():
            nonlocal move
            lib.bisection_run(ids, L, seed, move, moves_per_step, precision, records=False, wait=False)
            move += moves_per_step
    for _ in range(warmup):
        step()
    lib.bisection_stats_take(ids)          # waits for the warmup; clears its counters
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    tot = lib.bisection_stats_take(ids)    # waits for the timed steps
    wall = time.perf_counter() - t0
    progress(f"{system.name}: precision {precision}, {steps} timed steps in {wall:.2f} s")
    if graph:
        lib.sweep_graph_destroy(g)
    if paths_out is not None:
        paths_out.extend(lib.path_get(i) for i in ids[:8])
    lib.close()
    return tot, wall, tot.device_ms, {"block_threads": plan.block_threads, "link_split": plan.link_split,
                                      "persistent": bool(plan.persistent), "cluster": plan.cluster}


def other_workload(name: str, replicas: int):
    """The configurations besides the headline one, for reports."""
    if name == "uneven":
        return workloads.config3(num_replicas=128)
    if name in ("hydrogen", "hydrogen15k"):
        from pimcplusplus_amd.tables import load_npz
        return workloads.config2(load_npz(os.path.join(ROOT, "tests", "golden", "testcoul_fit2.npz")),
                                 num_replicas=replicas, m=168 if name == "hydrogen15k" else 128)
    if name == "ocp":
        return workloads.config4(num_replicas=16)
    if name == "lj64":
        return workloads.config0(num_replicas=replicas)
    if name == "bose":
        return workloads.dilute_boson_system(num_replicas=replicas)
    raise ValueError(name)


# PermuteType TABLE settings of the boson workload
BOSE_GAMMA, BOSE_EPSILON = (1.0, 1.0, 1.0, 1.0), 1e-4


def permute_setup(workload: str):
    """The replicas' set-up beyond loading: TABLE permutations for the boson workload."""
    if workload != "bose":
        return None
    def setup(lib, ids):
        for r in ids:
            lib.set_permute(r, 0, BOSE_GAMMA, BOSE_EPSILON)
    return setup


def report_other(args):
    """One JSON line for a non-headline workload: mixed and FP64 throughput, tuned plan,
    one captured sweep per step."""
    system = other_workload(args.workload, args.replicas)
    out = {"metric": METRIC, "unit": "pair·slice evals/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "higher_is_better": True, "data": "synthetic",
           "config": {"workload": system.name, "replicas": len(system.replicas),
                      "moves_per_step": args.moves_per_step, "num_levels": system.num_levels,
                      "tau": system.replicas[0].tau, "temperature_k": workloads.temperature_k(system),
                      "driver": "tuned plan, captured sweep graph"}}
    nmean = float(np.mean([r.num_particles for r in system.replicas]))
    fixed = None
    setup = permute_setup(args.workload)
    if setup:   # TABLE moves: the tuner picks the link split; no sweep graphs
        out["config"]["driver"] = "PermuteType TABLE, tuned link split, eager launches"
    if args.plan:
        b, ls, pers, cl = (int(x) for x in args.plan.split(","))
        fixed = (b, ls, bool(pers), cl)
        out["config"]["driver"] = f"fixed plan {args.plan}, eager launches"
    runs = (("mixed", abi.MIXED), ("fp64", abi.FP64))
    if args.only:
        runs = tuple(r for r in runs if r[0] == args.only)
    for name, prec in runs:
        tot, wall, dev_ms, plan = run_gpu(system, prec, args.steps, args.warmup, args.moves_per_step,
                                          tune=fixed is None, graph=fixed is None and setup is None, plan=fixed,
                                          setup=setup)
        sr_bytes, lr_bytes = workload_bytes(tot, nmean)
        kms = dev_ms / args.steps
        out[name] = {"evals_per_s": tot.pair_evals / wall, "ms_per_step": 1e3 * wall / args.steps,
                     "kernel_ms_avg": kms, "k_evals_per_s": tot.k_evals / wall,
                     "acceptance": tot.accepted / max(1, tot.moves), "plan": plan,
                     "bytes_per_launch": (sr_bytes + lr_bytes) / args.steps,
                     "long_range_bytes_per_launch": lr_bytes / args.steps}
    if args.only:
        print(json.dumps({"only": args.only, **out[args.only]}))
        return
    out["value"] = out["mixed"]["evals_per_s"]
    out["mixed_over_fp64"] = out["mixed"]["evals_per_s"] / out["fp64"]["evals_per_s"]
    achieved = {k: out[k]["bytes_per_launch"] / (out[k]["kernel_ms_avg"] * 1e-3) / 1e9 for k in ("mixed", "fp64")}
    out["roofline"] = {"bound": "hbm", "achieved": achieved["mixed"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": achieved["mixed"] / HBM_PEAK_GBS,
                       "traffic": measured_traffic(f"{len(system.replicas)} replicas, {args.moves_per_step} "
                                                   f"moves/launch, plan {args.plan}", traffic_profile(args.workload))
                       if args.plan else None,
                       "fp64": {"achieved": achieved["fp64"], "frac": achieved["fp64"] / HBM_PEAK_GBS},
                       "kernel": "k_bisection (the workload's variant and tuned plan)",
                       "note": "algorithmic bytes: 24 (K+1)(N+1) + 16 per executed stage (short range; N the mean "
                               "particle count) plus 16 B per counted k-term (32 B per k.slice.species update of "
                               "the lowest stage, counted once per OLD/NEW pair), over the kernel's HIP-event time"}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds / 2, system=system, moves_per_call=4, setup=setup)
    if args.workload == "ocp" and not args.no_dserr:
        out["estimators_mixed_vs_fp64"] = estimators_check(system)
    print(json.dumps(out))


def estimators_check(system, blocks: int = 10, moves: int = 1024, bins: int = 40, drop: int = 2):
    """The OCP's acceptance check: the thermodynamic energy and the pair correlation g(r) of
    mixed chains against FP64 chains that share no random numbers (the mixed chains use
    another seed).  Each replica contributes its mean over the blocks after `drop` blocks of
    equilibration.  Replica r starts from the same configuration in both precisions, so the
    comparison is paired: the mean over replicas of (mixed - FP64) against the standard error
    of those differences, for the energy and for every g(r) bin (largest |z|, mean z^2)."""
    M = system.replicas[0].num_time_slices
    L = system.replicas[0].box[0]
    res = {}
    for prec, seed in ((abi.FP64, system.seed), (abi.MIXED, system.seed + 7919)):
        lib = native.open_context(0)
        ids = load_system(lib, system)
        lib.bisection_tune(ids, system.num_levels, seed, 0, 8, prec)
        e, g = [], []
        for b in range(blocks):
            lib.bisection_run(ids, system.num_levels, seed, moves * b, moves, prec, records=False, wait=False)
            e.append([x.total / M for x in lib.energy(ids)])
            g.append([lib.pair_correlation(r, 0, 0, 0.0, L / 2, bins + 1) for r in ids])
        lib.close()
        res[prec] = (np.array(e[drop:]).mean(axis=0), np.array(g[drop:], dtype=float).mean(axis=0))
    (ea, ga), (eb, gb) = res[abi.FP64], res[abi.MIXED]
    n = len(ea)
    de, dg = eb - ea, gb - ga
    e_err = float(de.std(ddof=1) / np.sqrt(n))
    g_err = dg.std(axis=0, ddof=1) / np.sqrt(n)
    ok = g_err > 0
    z = dg.mean(axis=0)[ok] / g_err[ok]
    return {"energy": {"fp64_per_slice": float(ea.mean()), "mixed_per_slice": float(eb.mean()),
                       "difference": float(de.mean()), "std_error": e_err,
                       "within_4_sigma": abs(float(de.mean())) < 4 * e_err},
            "g_of_r": {"bins": bins, "range": [0.0, L / 2], "max_abs_z": float(np.max(np.abs(z))),
                       "mean_z2": float(np.mean(z * z)), "bins_compared": int(ok.sum()),
                       "within_error_bars": bool(np.max(np.abs(z)) < 4.5)},
            "comparison": "paired per replica (same start, independent random numbers)",
            "sample": f"{n} replicas per precision, {blocks} blocks of {moves} moves, the first {drop} dropped; "
                      "mixed chains on another seed"}


def workload_bytes(tot, nmean: float):
    """Algorithmic HBM bytes of the summarised launches: the short-range stage bytes of
    bytes_per_launch with the mean N, and 32 B per k.slice.species term of the long-range
    stage (rho_k read and update), which the counters count once for OLD and once for NEW."""
    sr = bytes_per_launch(tot, max(2.0, nmean))
    lr = 16.0 * tot.k_evals
    return sr, lr


def report_dropin(args):
    """The drop-in mode: PIMC++ keeps its move loop and calls the library once per
    SingleAction, as StageClass::StageAction does (StageClass.cc:124-136).  Times the calls a
    level-4 bisection stage issues on one replica of the benchmark system (Kinetic and
    ShortRange, OLD then NEW, each with one moved particle) and, for the hydrogen system, the
    lowest stage's LongRange calls; reports calls per second and the short-range pair·slice
    evaluations per second of this mode.  Inputs are resident; each call returns its result
    to the host (one synchronisation per call)."""
    from pimcplusplus_amd.tables import load_npz
    out = {"metric": METRIC, "unit": "pair·slice evals/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "higher_is_better": True, "data": "synthetic",
           "config": {"workload": "drop-in SingleAction calls, one request per call", "driver": "host loop"}}
    lib = native.open_context(0)
    for name, system, lr in (("lj256", workloads.config1(num_replicas=1), False),
                             ("hydrogen", workloads.config2(load_npz(os.path.join(ROOT, "tests", "golden",
                                                                                  "testcoul_fit2.npz")),
                                                            num_replicas=1), True)):
        ids = workloads.load(lib, system)
        n = system.replicas[0].num_particles
        W = 1 << system.num_levels
        rng = np.random.default_rng(5)
        calls, evals, t_sr, t_kin, t_lr = 0, 0, 0.0, 0.0, 0.0
        for it in range(args.warmup + args.steps):
            s1 = int(rng.integers(0, system.replicas[0].num_time_slices - W))
            p = int(rng.integers(0, system.replicas[0].species_count[0]))
            for level in range(system.num_levels - 1, -1, -1):
                for mode in (abi.OLDMODE, abi.NEWMODE):
                    req = [(ids[0], mode, s1, s1 + W, level, [p])]
                    t0 = time.perf_counter()
                    lib.action_kinetic(req)
                    t1 = time.perf_counter()
                    lib.action_short_range(req, abi.FP64)
                    t2 = time.perf_counter()
                    if lr and level == 0:
                        lib.action_long_range(req, abi.FP64)
                    t3 = time.perf_counter()
                    if it >= args.warmup:
                        t_kin += t1 - t0
                        t_sr += t2 - t1
                        t_lr += t3 - t2
                        calls += 2 + (1 if lr and level == 0 else 0)
                        evals += (n - 1) * (W >> level)
        nsr = 2 * system.num_levels * args.steps
        out[name] = {"short_range_calls_per_s": nsr / t_sr,
                     "short_range_evals_per_s": evals / t_sr,
                     "kinetic_us_per_call": 1e6 * t_kin / nsr,
                     "short_range_us_per_call": 1e6 * t_sr / nsr,
                     "long_range_us_per_call": (1e6 * t_lr / (2 * args.steps)) if lr else None,
                     "calls": calls}
    lib.close()
    out["value"] = out["lj256"]["short_range_evals_per_s"]
    print(json.dumps(out))


def longrange_requests(system, ids, lib, seed: int = 7):
    """One lowest-level bisection window per replica (slices s1 .. s1 + 2^L, level 0) with one
    moved particle of species 0, whose NEW beads inside the window are displaced (set once on
    the NEW copy); returns the NEWMODE and OLDMODE request lists and the full-path OLDMODE list
    (every slice, level 0)."""
    W = 1 << system.num_levels
    rng = np.random.default_rng(seed)
    new, old, full = [], [], []
    for i, r in zip(ids, system.replicas):
        M = r.num_time_slices
        s1 = int(rng.integers(0, M - W))
        p = int(rng.integers(0, r.species_count[0]))
        x = np.asarray(r.beads, dtype=float)[s1 + 1:s1 + W, p] + rng.normal(scale=0.05, size=(W - 1, 3))
        lib.path_set_beads(i, abi.NEWMODE, p, s1 + 1, x)
        new.append((i, abi.NEWMODE, s1, s1 + W, 0, [p]))
        old.append((i, abi.OLDMODE, s1, s1 + W, 0, [p]))
        full.append((i, abi.OLDMODE, 0, M, 0, [p]))
    return new, old, full


def report_longrange(args):
    """The long-range entry point pimc_action_long_range (LongRangeClass::SingleAction with
    UpdateRho_ks) at batch size: one call holds a request for every replica of the hydrogen
    workload (32 e + 32 p, 242 k), as a clone-parallel driver would issue them.  Per step: a
    NEWMODE call over each replica's lowest-level window (rho_k NEW written from OLD plus the
    moved bead's phases), an OLDMODE call over the same windows, and an OLDMODE call over every
    slice.  Algorithmic HBM bytes: 16 B read per (k, slice, species), plus 16 B written in
    NEWMODE.  Device time from the library's HIP events around each call."""
    from pimcplusplus_amd.tables import load_npz
    system = workloads.config2(load_npz(os.path.join(ROOT, "tests", "golden", "testcoul_fit2.npz")),
                               num_replicas=args.replicas)
    r0 = system.replicas[0]
    out = {"metric": "k.slice.species long-range terms/s (drop-in LongRange SingleAction, batched)",
           "unit": "k.slice.species terms/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "higher_is_better": True, "data": "synthetic",
           "config": {"workload": f"{system.name}, pimc_action_long_range, one request per replica per call",
                      "replicas": len(system.replicas)}}
    lib = native.open_context(0)
    ids = load_system(lib, system)
    new, old, full = longrange_requests(system, ids, lib)
    nk, ns = lib.replicas[ids[0]].nk, len(r0.species_count)
    out["config"]["nk"], out["config"]["species"] = nk, ns
    W = (1 << system.num_levels) + 1
    calls = {"new_window": (new, 32.0 * W), "old_window": (old, 16.0 * W),
             "old_full_path": (full, 16.0 * (r0.num_time_slices + 1))}
    if args.lr_call:
        calls = {args.lr_call: calls[args.lr_call]}
    precs = (("fp64", abi.FP64), ("mixed", abi.MIXED))
    if args.only:
        precs = tuple(p for p in precs if p[0] == args.only)
    for prec_name, prec in precs:
        res = {}
        for name, (reqs, bytes_per_k) in calls.items():
            for _ in range(args.warmup):
                lib.action_long_range(reqs, prec)
            lib.bisection_stats_take(ids)
            t0 = time.perf_counter()
            for _ in range(args.steps):
                lib.action_long_range(reqs, prec)
            wall = time.perf_counter() - t0
            dev_ms = lib.bisection_stats_take(ids).device_ms / args.steps
            terms = len(reqs) * (bytes_per_k / (32.0 if name == "new_window" else 16.0)) * nk * ns
            nbytes = len(reqs) * bytes_per_k * nk * ns
            res[name] = {"kernel_ms_avg": dev_ms, "wall_ms_per_call": 1e3 * wall / args.steps,
                         "terms_per_s": terms / (dev_ms * 1e-3), "bytes_per_call": nbytes,
                         "achieved_GBs": nbytes / (dev_ms * 1e-3) / 1e9,
                         "frac": nbytes / (dev_ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
        out[prec_name] = res
    lib.close()
    if args.only or args.lr_call:
        print(json.dumps(out))
        return
    main_call = out["fp64"]["new_window"]
    out["value"] = main_call["terms_per_s"]
    out["dtype"] = "f64"
    out["roofline"] = {"bound": "hbm", "achieved": main_call["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": main_call["frac"],
                       "traffic": measured_traffic(f"{len(system.replicas)} NEWMODE window requests",
                                                   os.path.join(ROOT, "profiles", "k_lr_req_hydrogen_traffic.md")),
                       "kernel": "k_lr_req<FP64> (NEWMODE window requests)",
                       "note": "algorithmic bytes: 32 B per (k, slice, species) of a NEWMODE request (rho_k OLD "
                               "read, NEW written), 16 B of an OLDMODE one, over the call's HIP-event time"}
    if not args.no_cpu:
        out["cpu_baseline"] = longrange_cpu_baseline(system, new, old, min(args.cpu_seconds, 10.0))
    print(json.dumps(out))


def longrange_cpu_baseline(system, new, old, seconds: float):
    """The oracle's LongRange SingleAction on one host core: NEWMODE then OLDMODE requests of
    the same windows, replica by replica, for about `seconds`."""
    o = abi.PimcLib(oracle_lib())
    nrep = 64
    sub = workloads.System(system.name, system.tables, system.desired_tau, system.replicas[:nrep],
                           system.num_levels, system.seed)
    ids = workloads.load(o, sub)
    nq, oq, _ = longrange_requests(sub, ids, o)
    nk, ns = o.replicas[ids[0]].nk, len(sub.replicas[0].species_count)
    W = (1 << system.num_levels) + 1
    terms, t0, i = 0, time.perf_counter(), 0
    while time.perf_counter() - t0 < seconds:
        o.action_long_range([nq[i % nrep]])
        o.action_long_range([oq[i % nrep]])
        terms += 2 * W * nk * ns
        i += 1
    wall = time.perf_counter() - t0
    o.close()
    return {"value": terms / wall, "unit": "k.slice.species terms/s", "cores": 1, "kind": "port",
            "sample": f"{i} NEWMODE + OLDMODE window request pairs over {nrep} hydrogen replicas, one core",
            "cpu_model": cpu_model(), "build": "g++ -O3 -march=x86-64-v3 -ffp-contract=off (oracle/Makefile)"}


def delta_s_error(nrep: int = 4, nmoves: int = 2500, seed_offset: int = 0, replica_offset: int = 0):
    """Max |dS_mixed - dS_fp64| over stages attempted on identical paths (chains run
    until their first differing decision).  seed_offset / replica_offset select a sample
    other than the one the FP32 limit was chosen on."""
    system = workloads.config1(num_replicas=nrep + replica_offset)
    system = workloads.System(system.name, system.tables, system.desired_tau, system.replicas[replica_offset:],
                              system.num_levels, system.seed + seed_offset)
    out = {}
    for prec in (abi.FP64, abi.MIXED):
        lib = native.open_context(0)
        ids = load_system(lib, system)
        rec, _ = lib.bisection_run(ids, system.num_levels, system.seed, 0, nmoves, prec)
        out[prec] = rec
        lib.close()
    a, b = out[abi.FP64], out[abi.MIXED]
    err, n = 0.0, 0
    for r in range(nrep):
        for m in range(nmoves):
            if a["num_stages"][r, m] != b["num_stages"][r, m] or a["accepted"][r, m] != b["accepted"][r, m]:
                break
            k = a["num_stages"][r, m]
            da = a["new_action"][r, m, :k] - a["old_action"][r, m, :k]
            db = b["new_action"][r, m, :k] - b["old_action"][r, m, :k]
            err = max(err, float(np.max(np.abs(da - db))))
            n += int(k)
    return err, n


def energy_check(nrep: int = 256, blocks: int = 8, moves: int = 400):
    """Thermodynamic energy per time slice of the benchmark system from mixed and FP64 chains
    that share no random numbers (the mixed chains use another seed): each replica's mean over
    the blocks after two blocks of equilibration, the two precisions compared through the
    spread over their independent replicas.  Returns the two means, the standard error of
    their difference and whether they agree within four of it."""
    system = workloads.config1(num_replicas=nrep)
    m = system.replicas[0].num_time_slices
    res = {}
    for prec, seed in ((abi.FP64, system.seed), (abi.MIXED, system.seed + 7919)):
        lib = native.open_context(0)
        ids = load_system(lib, system)
        vals = []
        for b in range(blocks):
            lib.bisection_run(ids, system.num_levels, seed, moves * b, moves, prec, records=False, wait=False)
            vals.append([e.total / m for e in lib.energy(ids)])
        lib.close()
        res[prec] = np.array(vals[2:]).mean(axis=0)
    a, b = res[abi.FP64], res[abi.MIXED]
    err = float(np.sqrt(a.var(ddof=1) / nrep + b.var(ddof=1) / nrep))
    d = float(b.mean() - a.mean())
    return {"fp64_K_per_slice": float(a.mean()), "mixed_K_per_slice": float(b.mean()), "difference": d,
            "std_error": err, "within_4_sigma": abs(d) < 4.0 * err,
            "comparison": "independent chains: mixed with another seed, error from the spread over replicas",
            "sample": f"{nrep} replicas per precision, {blocks} blocks of {moves} moves, the first two blocks dropped"}


def oracle_lib():
    path = os.path.join(ROOT, "oracle", "_build", "liboracle_pimc.so")
    if not os.path.exists(path):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    return ctypes.CDLL(path)


def host_cores():
    """Cores this process may use: its CPU affinity, capped by the cgroup CPU quota."""
    n = len(os.sched_getaffinity(0))
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()
        if quota != "max":
            n = min(n, max(1, int(float(quota) / float(period))))
    except (OSError, ValueError):
        pass
    return n


def cpu_model():
    """The host CPU: lscpu's model name (the /proc/cpuinfo one where lscpu is missing), with
    the x86 family/model numbers, which name the core generation when the name is generic."""
    name, fam = None, {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                name = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        for line in open("/proc/cpuinfo"):
            key = line.split(":", 1)[0].strip()
            if key in ("model name", "cpu family", "model") and key not in fam:
                fam[key] = line.split(":", 1)[1].strip()
    except OSError:
        pass
    name = name or fam.get("model name", "unknown")
    if "cpu family" in fam and "model" in fam:
        name += f" (family {fam['cpu family']}, model {fam['model']})"
    return name


def cpu_baseline(seconds: float = 8.0, threads: int | None = None, system=None, moves_per_call: int = 20,
                 setup=None):
    """The oracle (the C++ restatement of the reference's CPU path) on host cores:
    one replica per thread, the reference's clone-parallel model; every core the process
    may use (affinity and cgroup quota).  system: the workload whose replicas the threads
    run (default: the benchmark's N=256, M=64 system)."""
    lib = oracle_lib()
    threads = threads or host_cores()
    if system is None:
        system = workloads.config1(num_replicas=threads)
    reps = [system.replicas[r % len(system.replicas)] for r in range(threads)]
    ctxs = []
    for r in range(threads):
        o = abi.PimcLib(lib)
        sub = workloads.System(system.name, system.tables, system.desired_tau, [reps[r]],
                               system.num_levels, system.seed)
        ids = workloads.load(o, sub)
        if setup:
            setup(o, ids)
        ctxs.append((o, ids))

    def chunk(o, ids, res, k):
        evals, kev, t0, move = 0, 0, time.perf_counter(), 0
        while time.perf_counter() - t0 < seconds:
            _, st = o.bisection_run(ids, system.num_levels, system.seed, move, moves_per_call, records=False)
            evals += st.pair_evals
            kev += st.k_evals
            move += moves_per_call
        res[k] = (evals, time.perf_counter() - t0, kev)

    # one thread alone first (the reference's per-rank model), then all threads together
    single = [None]
    chunk(ctxs[0][0], ctxs[0][1], single, 0)
    res = [None] * threads
    ths = [threading.Thread(target=chunk, args=(o, ids, res, k)) for k, (o, ids) in enumerate(ctxs)]
    t0 = time.perf_counter()
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    wall = time.perf_counter() - t0
    total = sum(r[0] for r in res)
    ktotal = sum(r[2] for r in res)
    for o, _ in ctxs:
        o.close()
    out = {"value": total / wall, "unit": "pair·slice evals/s", "cores": threads, "kind": "port",
           "sample": f"{system.name}, one replica per thread, {seconds:.0f} s per thread "
                     f"({total} evaluations); single thread {single[0][0] / single[0][1]:.4g} evals/s",
           "single_core_value": single[0][0] / single[0][1],
           "cpu_model": cpu_model(), "host_cores": host_cores(),
           "build": "g++ -O3 -march=x86-64-v3 -ffp-contract=off (oracle/Makefile)"}
    if ktotal:
        out["k_evals_per_s"] = ktotal / wall
    return out


def calibrate_copy(mib: int):
    """One elementwise kernel that reads and writes `mib` MiB each (out = in + 1, FP32): the
    FETCH_SIZE / WRITE_SIZE counters of this dispatch scale the kernels' counter readings."""
    import torch
    x = torch.ones(mib * (1 << 20) // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(3):
        torch.add(x, 1.0, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_mib": mib, "dispatches": 3, "check": float(y[-1].item())}))


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--replicas", type=int, default=None,
                    help="independent replicas (default: 16384 for the LJ workloads lj256 and lj64: 16 "
                         "rounds of the blocks the GPU holds at once, so the end of a launch, where the last "
                         "blocks finish one by one, is a small part of it; 4096 for the others, whose "
                         "replicas each upload their own long-range tables; uneven and ocp have their own "
                         "counts)")
    ap.add_argument("--moves-per-step", type=int, default=8)
    ap.add_argument("--cpu-seconds", type=float, default=8.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=("mixed", "fp64"), default=None,
                    help="time one precision only (profiling runs); the JSON line then lacks the other")
    ap.add_argument("--no-dserr", action="store_true", help="skip the mixed-vs-FP64 dS and energy comparisons")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run an elementwise copy of MIB MiB (read MIB, write MIB) for counter scaling")
    ap.add_argument("--progress", action="store_true", help="a line on stderr at the end of each phase")
    ap.add_argument("--plan", default=None, metavar="BLOCK,SPLIT,PERSISTENT,CLUSTER",
                    help="non-headline workloads: this launch plan with eager launches instead of the tuner "
                         "and a captured sweep graph (profiling runs)")
    ap.add_argument("--lr-call", default=None, choices=("new_window", "old_window", "old_full_path"),
                    help="--workload longrange: time this call only (profiling runs)")
    ap.add_argument("--workload", default="lj256",
                    choices=("lj256", "lj64", "uneven", "hydrogen", "hydrogen15k", "ocp", "dropin", "longrange",
                             "bose"),
                    help="lj256 (default) is the headline configuration; the others print a report line "
                         "(hydrogen15k: the hydrogen system at M=168, T=15,037 K; dropin: per-SingleAction "
                         "calls of the drop-in mode; longrange: batched LongRange SingleAction requests; bose: "
                         "dilute helium bosons with PermuteType TABLE moves)")
    args = ap.parse_args()
    global PROGRESS
    PROGRESS = args.progress
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.replicas is None:
        args.replicas = HEADLINE_REPLICAS if args.workload in ("lj256", "lj64") else 4096
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(replicas would shard across GPUs with no communication; see the design notes)",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "dropin":
        report_dropin(args)
        return
    if args.workload == "longrange":
        report_longrange(args)
        return
    if args.workload != "lj256":
        report_other(args)
        return
    system = workloads.config1(num_replicas=args.replicas)
    n = system.replicas[0].num_particles
    res = {}
    runs = (("mixed", abi.MIXED), ("fp64", abi.FP64))
    if args.only:
        runs = tuple(r for r in runs if r[0] == args.only)
    table = system.tables["lj"]
    end, nint, noff = float(table.grid_end), table.y.shape[1] - 1, table.num_terms * (table.num_terms + 3) // 2
    paths = []
    for name, prec in runs:
        tot, wall, dev_ms, _ = run_gpu(system, prec, args.steps, args.warmup, args.moves_per_step,
                                       paths_out=paths if not paths else None)
        f_off = offdiag_fraction(paths, system.replicas[0].box, end)
        res[name] = {
            "evals_per_s": tot.pair_evals / wall,
            "ms_per_step": 1e3 * wall / args.steps,
            "kernel_ms_avg": dev_ms / args.steps,
            "device_evals_per_s": tot.pair_evals / (dev_ms * 1e-3),
            "bytes_per_launch": bytes_per_launch(tot, n) / args.steps,
            "acceptance": tot.accepted / max(1, tot.moves),
            "stages_per_move": tot.stages / max(1, tot.moves),
            "pair_evals": tot.pair_evals,
            "offdiag_fraction": f_off,
            "lds_bytes_per_launch": lds_bytes_per_launch(tot, args.steps, prec == abi.MIXED, f_off, nint,
                                                         1 + noff, noff),
        }
    if args.only:
        print(json.dumps({"only": args.only, **res[args.only]}))
        return
    err, nstages = (None, 0) if args.no_dserr else delta_s_error()
    # the same check on a sample the FP32 limit was not chosen on (another seed and replicas)
    err2, nstages2 = (None, 0) if args.no_dserr else delta_s_error(seed_offset=104729, replica_offset=16)
    energies = None if args.no_dserr else energy_check()
    mixed = res["mixed"]
    achieved = mixed["bytes_per_launch"] / (mixed["kernel_ms_avg"] * 1e-3) / 1e9
    peaks = measured_peaks()

    lds_peak = (peaks or {}).get("lds_read_GBs") or LDS_PEAK_GBS

    def lds_roof(r):
        a = r["lds_bytes_per_launch"] / (r["kernel_ms_avg"] * 1e-3) / 1e9
        return {"bound": "lds", "achieved": a, "peak": lds_peak, "unit": "GB/s", "frac": a / lds_peak,
                "frac_of_128B_per_clk": a / LDS_PEAK_GBS,
                "bytes_per_launch": r["lds_bytes_per_launch"], "offdiag_fraction": r["offdiag_fraction"]}
    flops = mixed["device_evals_per_s"] * FLOP_PER_EVAL / 1e12
    line = {
        "metric": METRIC,
        "value": mixed["evals_per_s"],
        "unit": "pair·slice evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["ms_per_step"],
        "kernel_ms_avg": mixed["kernel_ms_avg"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32/f64 mixed",
        "data": "synthetic",
        "config": {"workload": "LJ N=256 M=64 tau=0.0125 level-4 bisection, independent replicas",
                   "replicas": args.replicas, "moves_per_step": args.moves_per_step,
                   "parallelism": "replicas (one block per replica)"},
        "fp64": {"value": res["fp64"]["evals_per_s"], "ms_per_step": res["fp64"]["ms_per_step"],
                 "kernel_ms_avg": res["fp64"]["kernel_ms_avg"]},
        "mixed_over_fp64": mixed["evals_per_s"] / res["fp64"]["evals_per_s"],
        "delta_s_max_err": err,
        "delta_s_stages_compared": nstages,
        "delta_s_fresh_sample": {"max_err": err2, "stages": nstages2,
                                 "sample": "4 replicas x 2500 moves, seed + 104729, replicas 16-19"},
        "energy_mixed_vs_fp64": energies,
        "acceptance": mixed["acceptance"],
        "stages_per_move": mixed["stages_per_move"],
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(f"{args.replicas} replicas, {args.moves_per_step} moves/launch"),
                     "kernel": "k_bisection<NT=2,MIXED>",
                     "note": "algorithmic bytes = 24 (K+1)(N+1) + 16 per executed stage, over the "
                             "kernel's HIP-event time; traffic from the committed FETCH_SIZE/WRITE_SIZE "
                             "summary. The binding roof is the vector ALU (see roofline_valu)"},
        "roofline_valu": {"bound": "valu", "achieved": flops, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                          "frac": flops / FP64_PEAK_TFLOPS,
                          "note": f"{FLOP_PER_EVAL:.0f} algorithmic FP64 flop per pair·slice evaluation "
                                  "over the kernel time; peak = MI355X FP64 vector"},
        "roofline_lds": {**lds_roof(mixed),
                         "fp64": {k: v for k, v in lds_roof(res["fp64"]).items() if k in ("achieved", "frac")},
                         "note": "LDS bytes read by the evaluations (moved bead, u00 row, off-diagonal rows "
                                 "where q < End) plus the staged tables, over the kernel time; peak = the LDS "
                                 "read bandwidth measured on this GPU (measured_peaks), which exceeds "
                                 "256 CUs x 128 B/clk x 2.4 GHz"},
        "measured_peaks": peaks,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
