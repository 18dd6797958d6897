#!/usr/bin/env python3
"""Benchmark of the per-step force-and-integrate loop on one MI355X.

Workload (the 10k-bead configuration the headline metric is quoted on):
40 chains x 250 beads in a periodic cube of 464.2 A (1e-4 residues/A^3),
full term set (Go LJ, Pauli, Debye-Hueckel, QA, PID, tether, chirality,
native angles and dihedrals), Langevin Gear predictor-corrector, dt 0.005.
`value` is the bead-steps/s of the batched mixed-precision form: --replicas
independent boxes (seeds base+k, temperatures 0.30..0.50) advanced by one
graph-captured step.  Beside it: the plain FP64 form on one box, the force
error of the mixed form against FP64 on identical coordinates, a roofline
entry for the dominant kernel (HIP-event durations on the engine stream)
and the CPU oracle (the reference algorithm, OpenMP) on a bounded sample.

  python bench.py [--steps K] [--warmup W] [--replicas B] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from xmd_amd import _abi, synth  # noqa: E402
from xmd_amd.engine import Engine, profile  # noqa: E402
from xmd_amd._host import XmdError  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


NONFINITE = 4  # XMD_ERR_NONFINITE: a replica's positions (or its noise) turned non-finite


def step_tolerant(eng, steps):
    """eng.step(steps), letting replicas halt: a replica whose positions turn
    non-finite (the model's own dynamics, or a zero noise draw) stops where
    it failed while the others take every step (the engine's per-replica
    errors); any other error is raised."""
    try:
        eng.step(steps)  # xmd_step synchronizes before returning
    except XmdError as err:
        if err.status != NONFINITE:
            raise
        log(f"halted: {err}")


def running_beads(eng, system):
    """Beads of the replicas that are not halted."""
    return sum(system.replica_offset[k + 1] - system.replica_offset[k] for k in range(system.n_replicas)
               if eng.replica_status(k)[0] == 0)


def timed_steps(eng, warmup, steps, system=None):
    """Seconds for `steps` steps after `warmup`; with `system`, also the
    beads that count for the window: those of the replicas still running at
    its end (a replica that halts inside the window counts for none of it)."""
    step_tolerant(eng, warmup)
    t0 = time.perf_counter()
    step_tolerant(eng, steps)
    dt = time.perf_counter() - t0
    return dt if system is None else (dt, running_beads(eng, system))


def tune_groups(params, system, candidates=(1, 2, 4), warm=150, steps=150):
    """Stream-group count of the batched engine (fixed at creation, so
    chosen by creating one engine per candidate): each warms up, autotunes
    its force tile and times `steps` steps from the same state; returns the
    count to use -- the engine's default (2) unless another is more than 3 %
    faster -- and the ms per step of each.  Per-replica results do not
    depend on the count."""
    ms = []
    for g in candidates:
        params.stream_groups = g
        e = Engine(params, system)
        step_tolerant(e, warm)
        e.autotune(reps=5)
        ms.append(1e3 * timed_steps(e, 0, steps) / steps)
        e.close()
    params.stream_groups = 0
    best, default = int(np.argmin(ms)), candidates.index(2)
    pick = candidates[best] if ms[best] < 0.97 * ms[default] else 2
    return pick, [round(x, 4) for x in ms]


def row_entries(eng):
    return eng.counters()["nl_entries"]


def force_kernel_bytes(eng, system):
    """Algorithmic bytes of one launch of the per-bead force kernel.

    Per neighbour-row entry: one word, partner and flags (4 B).  Per native
    row entry: int32 partner + fp64 distance + uint8 active (13 B).  Per
    bead: its position (32 B, AoS x/y/z/pad; partner gathers re-read
    positions that are already counted), bonded natives (5 fp64 = 40 B),
    row length + slice offset (8 B), flags + aa type (2 B), chain id + QA
    slot count (8 B), energy write (8 B), force write (24 B) and the PID
    spill to the chain neighbours (48 B)."""
    n = system.n_beads
    e = row_entries(eng)
    nnat = 2 * len(system.nat_i1)
    per_bead = 32 + 40 + 8 + 2 + 8 + 8 + 24 + 48
    return 4 * e + 13 * nnat + per_bead * n


def measured_traffic(kernel, replicas, groups):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (rocprofv3 FETCH_SIZE and WRITE_SIZE, each scaled by what it reads for a
    copy of known size), when it was taken on this workload and split."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    k = t.get("kernels", {}).get(kernel)
    if not k or t.get("replicas") != replicas or t.get("groups", 1) != groups:
        return None
    return int(k["hbm_bytes_per_launch"])


def valu_issue(kernel, replicas, groups, avg_ms):
    """VALU issue time of one launch (committed SQ_INSTS_VALU per launch x 4
    cycles per wave64 instruction over 1024 SIMDs at 2.4 GHz) and its
    fraction of the live launch time (a launch shares the chip with the other
    stream group, so this is a lower bound on the issue rate)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    k = t.get("kernels", {}).get(kernel)
    if not k or "valu_insts_per_launch" not in k or t.get("replicas") != replicas or t.get("groups", 1) != groups:
        return None
    us = k["valu_insts_per_launch"] * 4 / 1024 / 2.4e3
    return {"issue_us": round(us, 1), "frac_of_launch": round(us / (avg_ms * 1e3), 3)}


def host_cpu():
    """CPU model, logical CPUs this process may use, physical cores of the
    host (distinct (package, core) pairs)."""
    model, cores = None, set()
    phys = core = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model is None:
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not k and phys is not None:
                    cores.add((phys, core))
                    phys = core = None
    except OSError:
        pass
    if phys is not None:
        cores.add((phys, core))
    return model, len(os.sched_getaffinity(0)), len(cores) or None


def cpu_baseline(params, system, seconds):
    """The CPU baseline: the oracle in the reference's OpenMP structure
    (thread-private forces, atomic reduce, serial list rebuild and
    integrator) built with the reference's Release flags (-O2
    -march=native, xmd/CMakeLists.txt:32) on this host; 100 warm-up steps,
    then the median of 5 timed runs sized to ~`seconds` in total, with every
    thread this process may use (the OMP_NUM_THREADS share when set) and on
    one thread (3 runs)."""
    import subprocess
    import tempfile
    from oracle.oracle import Oracle
    tmp = tempfile.mkdtemp(prefix="xmd_cpu_baseline_")
    lib = os.path.join(tmp, "liboracle_native.so")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "native", f"OUT={lib}"], check=True)
    model, avail, phys = host_cpu()
    threads = avail
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        threads = min(threads, int(os.environ["OMP_NUM_THREADS"]))

    def median_rate(nthreads, warm, runs, budget):
        ora = Oracle(params, system, lib_path=lib)
        ora.step(1, nthreads=nthreads)  # first list build (O(N^2), serial as in the reference)
        t1 = time.perf_counter()
        ora.step(warm, nthreads=nthreads)
        per = (time.perf_counter() - t1) / warm
        k = int(max(10, budget / runs / max(per, 1e-9)))
        rates = []
        for _ in range(runs):
            t1 = time.perf_counter()
            ora.step(k, nthreads=nthreads)
            rates.append(system.n_beads * k / (time.perf_counter() - t1))
        ora.close()
        return float(np.median(rates)), k, [round(r, 1) for r in rates]

    import shutil
    try:
        value, k, rates = median_rate(threads, 100, 5, seconds)
        v1, k1, rates1 = median_rate(1, 20, 3, 0.5 * seconds)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return {"value": round(value, 1), "unit": "bead-steps/s", "cores": threads, "kind": "port",
            "sample": f"one 10k-bead box (40x250, PBC), FP64 forces, reference algorithm (O(N^2) serial list "
                      f"rebuild, per-term omp-for with thread-private forces and atomic reduce, serial Gear "
                      f"integrator), -O2 -march=native; median of 5 runs of {k} steps after 100 warm-up steps, "
                      f"{threads} OpenMP threads",
            "cpu_model": model, "logical_cpus_available": avail, "host_physical_cores": phys,
            "runs": rates,
            "one_thread": {"value": round(v1, 1), "steps_per_run": k1, "runs": rates1}}


def copy_calibration(mib):
    """Known-traffic kernels: a D2D copy and out = 2*x, each reading and
    writing `mib` MiB once."""
    import torch
    n = mib * (1 << 20) // 4
    x = torch.rand(n, device="cuda:0")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
        torch.mul(x, 2.0, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_mib": mib, "read_bytes": n * 4, "write_bytes": n * 4}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--replicas", type=int, default=16)
    ap.add_argument("--profile-steps", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-extra", action="store_true", help="skip the 512-chain and 100k-bead workloads")
    ap.add_argument("--no-step-roofline", action="store_true")
    ap.add_argument("--extra-steps", type=int, default=300)
    ap.add_argument("--long-steps", type=int, default=2000,
                    help="steps of the second timed window (0: none); the rate depends on how many list "
                         "rebuilds fall in a window, so it is reported beside the --steps one")
    ap.add_argument("--stream-groups", type=int, default=0,
                    help="stream groups of the batched engine instead of the sweep over 1, 2, 4 (0: sweep)")
    ap.add_argument("--force-lanes", type=int, default=0, choices=(0, 4, 8, 16),
                    help="lanes per bead of the force kernel instead of the autotuner's choice (0: autotune); "
                         "for counter runs, whose serialized kernels change the autotuner's timings")
    ap.add_argument("--seed-horizon", type=int, default=0, metavar="STEPS",
                    help="pick the batch's seeds free of zero log() draws over STEPS steps (0: the steps of this "
                         "run); the profile commands pass the default run's 4464 so that they step its replicas")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy and a scale kernel of MIB MiB each "
                         "(known HBM bytes, to calibrate FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU sharding is designed "
                                   "(independent replicas per GPU) but not built in this version"}))
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return

    t0 = time.time()
    # seeds whose noise would hit a zero log() draw inside the run are
    # skipped: the reference's noise is infinite there (rand_gen.h:45-46)
    # every engine step of the default run (warm-up, profiled, both windows);
    # a shorter run given the default run's horizon steps the same seeds
    horizon = args.seed_horizon or args.warmup + args.profile_steps + args.steps + args.long_steps + 64
    batch = synth.config_system(2, n_batch=args.replicas, horizon_steps=horizon)
    single = synth.config_system(2, n_batch=1, horizon_steps=horizon + max(200, args.steps // 4))
    log(f"systems generated in {time.time() - t0:.1f}s: batch {batch.n_beads} beads, single {single.n_beads}")
    p_mixed = synth.config_params(2, precision=_abi.PREC_MIXED)
    p_fp64 = synth.config_params(2, precision=_abi.PREC_FP64)

    # ---- batched mixed precision (the headline form)
    if args.force_lanes:
        p_mixed.force_lanes = args.force_lanes
    groups, groups_ms = None, None
    if args.stream_groups:
        p_mixed.stream_groups = args.stream_groups
        groups = args.stream_groups
    elif args.replicas > 1:
        groups, groups_ms = tune_groups(p_mixed, batch)
        p_mixed.stream_groups = groups
        log(f"stream groups: {groups} (ms per step for 1/2/4: {groups_ms})")
    eng = Engine(p_mixed, batch)
    step_tolerant(eng, args.warmup)
    if args.force_lanes:
        lanes, tune_ms = args.force_lanes, None
    else:
        lanes, tune_ms = eng.autotune(reps=10)
        log(f"autotune: lanes per bead {lanes} (ms per step for 4/8/16: {[round(x, 4) for x in tune_ms]})")
    # ---- roofline of the dominant kernel, HIP events on the engine stream,
    # over the steps right after the warm-up and the autotuner: the run is
    # deterministic, so the kernel trace and counter commands of profiles/
    # (same seeds, warm-up and profiled steps) see the same state
    try:
        prof = profile(eng, args.profile_steps)
    except XmdError as err:
        if err.status != NONFINITE:
            raise
        log(f"halted while profiling: {err}; profiling again")
        prof = profile(eng, args.profile_steps)
    dom = max(prof.items(), key=lambda kv: kv[1][1])
    name, (launches, total_ms) = dom
    avg_ms = total_ms / launches
    # the replicas are stepped as stream groups, one launch per group per
    # step: bytes of one launch = the batch's bytes / launches per step
    per_step = launches / args.profile_steps
    nbytes = force_kernel_bytes(eng, batch) / per_step if name.startswith("k_force") else None
    roofline = None
    if nbytes is not None:
        achieved = nbytes / (avg_ms * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": measured_traffic(name, args.replicas, round(per_step)),
                    "kernel": name, "avg_ms": round(avg_ms, 5), "bytes_per_launch": int(nbytes),
                    "valu_issue": valu_issue(name, args.replicas, round(per_step), avg_ms)}
    kernels = {k: {"launches": n, "avg_us": round(1e3 * ms / n, 2), "share": round(ms / sum(v[1] for v in prof.values()), 4)}
               for k, (n, ms) in sorted(prof.items(), key=lambda kv: -kv[1][1])}
    log("kernels:", json.dumps(kernels))

    rb0 = eng.counters()["rebuilds"]
    dt_mixed, beads_mixed = timed_steps(eng, 0, args.steps, batch)
    value = beads_mixed * args.steps / dt_mixed
    rb_window = eng.counters()["rebuilds"] - rb0
    log(f"mixed batched: {args.steps} steps in {dt_mixed:.3f}s -> {value:.3e} bead-steps/s "
        f"({rb_window} list rebuilds over the replicas)")
    # the rate depends on how many list rebuilds fall in the window (one per
    # replica every ~70 steps here): beside the K-step figure, a window of at
    # least 2000 steps
    long_window = None
    if args.long_steps > 0:
        long_steps = max(args.long_steps, args.steps)
        rb1 = eng.counters()["rebuilds"]
        dt_long, beads_long = timed_steps(eng, 0, long_steps, batch)
        long_window = {"steps": long_steps, "value": round(beads_long * long_steps / dt_long, 1),
                       "ms_per_step": round(1e3 * dt_long / long_steps, 4),
                       "rebuilds": eng.counters()["rebuilds"] - rb1}
        log(f"long window: {long_window}")
    counters = eng.counters()

    # ---- whole-step algorithmic bytes per bead-step (integrator 348 B in the
    # mixed form; per pair term P_t (8 B index pair + 4 B parameter) + 48 B
    # per bead; bonded 88 B per bead; rebuild (36 N + 8 P_nl) over the
    # measured rebuild interval) and the fraction of HBM peak it implies
    step_roof = None
    if not args.no_step_roofline:
        P = {name: 0 for name in ("go", "pauli", "es", "pid", "nl_all")}
        for k in range(batch.n_replicas):
            for name in P:
                P[name] += len(eng.get_list(k, name)[0])
        c = eng.counters()
        n = batch.n_beads
        qa_pairs = c["qa_contacts"]
        interval = c["steps"] * batch.n_replicas / max(c["rebuilds"], 1)
        pair = sum(P[t] * 12 + n * 48 for t in ("go", "pauli", "es", "pid")) + qa_pairs * 12 + n * 48
        per_bead = 348 + pair / n + 88 + (36 * n + 8 * P["nl_all"]) / interval / n
        step_roof = {"bytes_per_bead_step": round(per_bead, 1), "achieved": round(per_bead * value / 1e9, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(per_bead * value / 1e9 / HBM_PEAK_GBS, 4),
                     "pairs": {**P, "qa_contacts": qa_pairs}, "rebuild_interval_steps": round(interval, 1)}
        log(f"step roofline: {step_roof}")

    # ---- plain FP64 on one box, and the mixed force error against it
    fp64_value = None
    fp64_batched = None
    rel_rms = None
    rel_terms = {}
    if not args.no_fp64:
        eb = Engine(p_fp64, batch)
        kb = max(100, args.steps // 10)
        dtb, beads_b = timed_steps(eb, min(args.warmup, 50), kb, batch)
        fp64_batched = beads_b * kb / dtb
        log(f"fp64 batched: {kb} steps in {dtb:.3f}s -> {fp64_batched:.3e} bead-steps/s")
        eb.close()
        e64 = Engine(p_fp64, single)
        k64 = max(200, args.steps // 4)
        dt64 = timed_steps(e64, min(args.warmup, 100), k64)
        fp64_value = single.n_beads * k64 / dt64
        log(f"fp64 single box: {k64} steps in {dt64:.3f}s -> {fp64_value:.3e} bead-steps/s")
        # same coordinates into a mixed single-box engine
        em = Engine(p_mixed, single)
        y, _, rng, t = e64.get_state(0)
        cont, sync = e64.get_qa_state(0)
        em.set_state(0, y, rng, t)
        em.set_qa_state(0, cont, sync)
        e64.set_state(0, y, rng, t)
        e64.set_qa_state(0, cont, sync)
        F64, _ = e64.eval_forces()
        Fm, _ = em.eval_forces()
        rel_rms = float(np.sqrt(((Fm - F64) ** 2).sum() / (F64 ** 2).sum()))
        log(f"mixed vs fp64 force rel-RMS: {rel_rms:.3e}")
        for name in _abi.TERMS:
            a, _ = e64.eval_forces(1 << _abi.TERM[name])
            if np.abs(a).max() == 0:
                continue
            b, _ = em.eval_forces(1 << _abi.TERM[name])
            rel_terms[name] = float(np.sqrt(((b - a) ** 2).sum() / (a ** 2).sum()))
        log(f"per term: {rel_terms}")

    # ---- secondary workloads (mixed precision, autotuned): 512 single chains
    # of 30..400 beads in open boxes, and one 100k-bead periodic box
    extra = {}
    if not args.no_extra:
        c4 = None
        for key, cid in (("c3_512_single_chains", 3), ("c4_box_100k", 4), ("c4_batch_8x100k", 4)):
            t1 = time.time()
            if key == "c4_batch_8x100k":
                # 8 copies of the generated box at T = 0.30 + 0.2 k/7 with their own seeds
                seeds = synth._pick_seeds(synth.SEED_BASE + 4 + 1, 8, lambda _: c4.n_beads, args.extra_steps + 150)
                sysx = synth.replicate(c4, [0.30 + 0.2 * k / 7 for k in range(8)], seeds)
            else:
                sysx = synth.config_system(cid, horizon_steps=args.extra_steps + 150)
                if cid == 4:
                    c4 = sysx
            ex = Engine(synth.config_params(cid, precision=_abi.PREC_MIXED), sysx)
            step_tolerant(ex, 100)
            lx, _ = ex.autotune(reps=5)
            dtx, beads_x = timed_steps(ex, 0, args.extra_steps, sysx)
            extra[key] = {"value": round(beads_x * args.extra_steps / dtx, 1), "unit": "bead-steps/s",
                          "beads": int(sysx.n_beads), "replicas": int(sysx.n_replicas), "steps": args.extra_steps,
                          "ms_per_step": round(1e3 * dtx / args.extra_steps, 4), "lanes_per_bead": lx}
            log(f"{key}: {extra[key]} (setup {time.time() - t1:.1f}s)")
            ex.close()

    # ---- CPU baseline: the oracle in the reference's OpenMP structure
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(p_fp64, single, args.cpu_seconds)
        log(f"cpu oracle: {cpu}")

    out = {
        "metric": "bead-steps/s: batched mixed-prec vs plain FP64 on GPU, force rel-RMS <=1e-5",
        "value": round(value, 1),
        "unit": "bead-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * dt_mixed / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 forces / fp64 state",
        "data": "synthetic (seeded self-avoiding chains)",
        "config": {"workload": f"C2 box 40x250 beads (10k, PBC L=464.2 A, full term set) x {args.replicas} "
                               f"replicas, mixed precision, graph-captured step, replicas in "
                               f"{round(per_step)} concurrent stream groups",
                   "replicas": args.replicas, "beads_total": batch.n_beads, "parallelism": "replicas in one launch",
                   "seeds": [int(x) for x in batch.rng_state]},
        "fp64_single_box": None if fp64_value is None else round(fp64_value, 1),
        "mixed_over_fp64": None if fp64_value is None else round(value / fp64_value, 3),
        "fp64_batched": None if fp64_batched is None else round(fp64_batched, 1),
        "mixed_over_fp64_batched": None if fp64_batched is None else round(value / fp64_batched, 3),
        "force_rel_rms_mixed_vs_fp64": rel_rms,
        "force_rel_rms_mixed_vs_fp64_per_term": rel_terms,
        "long_window": long_window,
        "window_rebuilds": rb_window,
        # a window shorter than a few rebuild intervals per replica times a
        # rate that depends on whether rebuilds fall inside it: flagged, with
        # the long window beside it
        "window_note": (None if rb_window >= 3 * args.replicas else
                        f"timed window of {args.steps} steps holds {rb_window} list rebuilds over "
                        f"{args.replicas} replicas (fewer than 3 per replica): its rate is not representative; "
                        f"see long_window"),
        "halted_replicas": [k for k in range(batch.n_replicas) if eng.replica_status(k)[0] != 0],
        "rebuilds": counters["rebuilds"],
        "nl_entries": counters["nl_entries"],
        "qa_contacts": counters["qa_contacts"],
        "stream_groups": {"chosen": groups, "ms_per_step_1_2_4": groups_ms if groups_ms else "fixed by --stream-groups"},
        "autotune": {"lanes_per_bead": lanes,
                     "ms_per_step_4_8_16": [round(x, 4) for x in tune_ms] if tune_ms else "fixed by --force-lanes"},
        "secondary": extra,
        "roofline": roofline,
        "roofline_step": step_roof,
        "cpu_baseline": cpu,
        "kernels": kernels,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
