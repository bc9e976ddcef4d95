#!/usr/bin/env python3
"""Barnes-Hut throughput on one MI355X: timesteps/s at N = 1,048,576 Plummer bodies.

Workload (one "step" = one full pipeline pass): Plummer sphere (M = 1, a = 1,
equal masses, virial velocities, seed 1), theta = 0.5, eps = 0.01, G = 1,
dt = 1e-3, Euler; bounds -> reset -> build -> summarize -> sort -> force ->
update, all on the GPU with the state resident in HBM.

Prints one JSON line.  Besides the throughput it reports:
  * roofline of the dominant kernel (the force traversal): algorithmic FP32
    flops = 20 x I_alg per step, where I_alg is the number of interactions a
    per-body traversal under the same rules needs (estimated by the oracle on
    a sample of bodies); the kernel time comes from HIP events recorded around
    every force launch of the timed region on the simulation's stream;
  * cpu_baseline: the oracle (C, OpenMP) timed on this host on a bounded
    sample of the same step (full tree build, per-body traversal of a sample
    of bodies scaled to N);
  * the force error against an FP64 direct sum on 1024 seeded targets.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "BH timesteps/s @1M bodies θ=0.5; GInteractions/s; ensemble p50/p99 latency"
PEAK_FP32_TFLOPS = 157.3
FLOPS_PER_INTERACTION = 20


def host_cpu():
    """The host CPU the CPU baseline ran on: model name (from /proc/cpuinfo) and logical CPUs."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count()}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def calibrate_copy(mib):
    """A plain device-to-device copy of known size (HIP's blit kernel), so that the
    HBM byte counters of a profiling run can be scaled: read mib MiB, write mib MiB."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    nbytes = mib << 20
    src, dst = C.c_void_p(), C.c_void_p()
    for ptr in (src, dst):
        if hip.hipMalloc(C.byref(ptr), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, C.c_size_t(nbytes))
    for _ in range(3):
        if hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    log(f"bench: copy calibration {mib} MiB x3 (each copy reads and writes {mib} MiB)")


def parse_args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["bh1m", "ensemble", "precision", "serving", "bf4096"], default="bh1m",
                    help="bh1m: Barnes-Hut 1M Plummer (headline); ensemble: 512 systems batched vs one per launch; "
                         "precision: Kahan vs plain brute force against FP64; serving: continuous batching; "
                         "bf4096: plain brute force N=4096")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: bh1m 20, ensemble 20, bf4096 100, precision 200)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--rules", choices=["standard", "reference"], default="standard")
    ap.add_argument("--theta", type=float, default=0.5)
    ap.add_argument("--cpu-sample", type=int, default=32768,
                    help="bh1m: bodies whose traversal the one-thread CPU figure times")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-error", action="store_true")
    ap.add_argument("--no-shares", action="store_true", help="bh1m: skip the per-GPU target-share timing")
    ap.add_argument("--no-companion", action="store_true", help="bh1m: skip the timed run under the other rule set")
    ap.add_argument("--no-plain", action="store_true",
                    help="ensemble: skip the one-system-per-launch comparison legs (time the batched step only)")
    ap.add_argument("--requests", type=int, default=2000, help="serving: number of requests")
    ap.add_argument("--rate", type=float, default=0.0,
                    help="serving: Poisson arrival rate, requests/s (0: --load of the measured capacity)")
    ap.add_argument("--load", type=float, default=0.8, help="serving: offered fraction of the measured capacity")
    ap.add_argument("--capacity-requests", type=int, default=300,
                    help="serving: requests of the closed-loop saturation run")
    ap.add_argument("--budget-gib", type=float, default=96.0, help="serving: HBM budget of the server arena")
    ap.add_argument("--max-systems", type=int, default=512, help="serving: system slots")
    ap.add_argument("--calibrate-copy-mib", type=lambda t: [int(x) for x in t.split(",") if x], default=[],
                    help="first run plain device copies of these many MiB, comma-separated (scale for HBM "
                         "byte counters)")
    return ap.parse_args()


DEFAULT_STEPS = {"bh1m": 20, "ensemble": 20, "bf4096": 100, "precision": 200, "serving": 0}


def main():
    args = parse_args()
    if args.steps is None:
        args.steps = DEFAULT_STEPS[args.workload]
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each box has one MI355X); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    for mib in args.calibrate_copy_mib:
        calibrate_copy(mib)
    {"bh1m": bench_bh, "ensemble": bench_ensemble, "precision": bench_precision, "serving": bench_serving,
     "bf4096": bench_bf4096}[args.workload](args)


REFERENCE_WGS = 256  # the reference's work-group size on an AMD GPU (CL_DEVICE_MAX_WORK_GROUP_SIZE)


def run_bh(n, rules, theta, eps, G, dt, steps, warmup, state):
    """The timed Barnes-Hut workload under one rule set: replayed step graphs for the rate (the
    production path), the same steps launched stage by stage with HIP events on the simulation's
    stream for the per-kernel times.  Returns the rate, the stage times and the end state with
    one more force pass on it."""
    from nbodysimulation_amd import _lib as L
    from nbodysimulation_amd.engine import NBodySim, make_config

    pos, vel, mass = state
    with NBodySim(make_config(n, L.BARNES_HUT, rules=rules, theta=theta, eps=eps, G=G, dt=dt,
                              time_stages=True)) as timed:
        timed.upload(pos, vel, mass)
        if warmup:
            timed.step(warmup)
        timed.step(steps)
        stages = timed.stage_times()  # ms summed over the timed steps
    sim = NBodySim(make_config(n, L.BARNES_HUT, rules=rules, theta=theta, eps=eps, G=G, dt=dt))
    sim.upload(pos, vel, mass)
    if warmup:
        sim.step(warmup)
    sim.sync()
    t0 = time.perf_counter()
    sim.step(steps)  # returns after a stream synchronise
    elapsed = time.perf_counter() - t0
    out = {"steps_per_s": steps / elapsed, "ms_per_step": 1e3 * elapsed / steps, "elapsed": elapsed,
           "stages": stages, "exec_inter": sim.interactions(), "tinfo": sim.tree_info()}
    sim.force_pipeline()
    t_dl = time.perf_counter()
    out["end"] = sim.download()
    out["download_s"] = time.perf_counter() - t_dl
    out["sim"] = sim
    return out


def engine_force(n, rules, theta, eps, G, pos, vel, mass):
    """The engine's accelerations on a given state (one force pass) under a rule set."""
    from nbodysimulation_amd import _lib as L
    from nbodysimulation_amd.engine import NBodySim, make_config
    with NBodySim(make_config(n, L.BARNES_HUT, rules=rules, theta=theta, eps=eps, G=G)) as s:
        s.upload(pos, vel, mass)
        s.force_pipeline()
        return s.download()[2]


def reference_vote_error(pos, mass, theta, eps, G, engine_acc, groups=8, seed=99):
    """How far the engine's accelerations are from the reference kernels' own opening rule.
    calculate_force_local (kernels/calculate_forces.cl:84-145) walks the tree once per
    work-group and accepts a cell only if work_group_all says it is far enough for every body
    of the work-group: WGS consecutive bodies in index order (body_idx = get_global_id(0),
    WGS = CL_DEVICE_MAX_WORK_GROUP_SIZE, src/MyApp.cpp:176-181), under the reference's tree
    and thresholds (REFERENCE rules).  The oracle emulates exactly that on `groups` sampled
    work-groups; each engine result (rules -> accelerations) is compared with it and with an
    FP64 direct sum on the same bodies."""
    from oracle import oracle as orc

    n = mass.size
    tree = orc.Tree(pos, mass, 0)  # REFERENCE rules: the reference's tree and thresholds
    g = np.sort(np.random.default_rng(seed).choice(n // REFERENCE_WGS, groups, replace=False)).astype(np.int32)
    a_k, _, ni = tree.force(theta, eps, G, vote=REFERENCE_WGS, order=np.arange(n, dtype=np.int32), groups=g,
                            want_absacc=False)
    members = (g[:, None] * REFERENCE_WGS + np.arange(REFERENCE_WGS)[None, :]).reshape(-1).astype(np.int32)
    a64 = orc.direct_f64(pos, mass, orc.LAW_PLUMMER, eps, G, targets=members)
    n64 = np.linalg.norm(a64, axis=1)
    nk = np.linalg.norm(a_k[members], axis=1)

    def stats(rel):
        return {"median": float(np.median(rel)), "p99": float(np.quantile(rel, 0.99))}

    out = {"reference_kernels": {"vote": f"work_group_all over {REFERENCE_WGS} consecutive index-order bodies",
                                 "rules": "reference", "bodies": int(members.size),
                                 "interactions_per_body": float(ni[members].mean()),
                                 "error_vs_fp64": stats(np.linalg.norm(a_k[members] - a64, axis=1) / n64)}}
    for name, a in engine_acc.items():
        out[f"engine_{name}"] = {
            "rel_diff_vs_reference_kernels": stats(np.linalg.norm(a[members] - a_k[members], axis=1) / nk),
            "error_vs_fp64": stats(np.linalg.norm(a[members] - a64, axis=1) / n64)}
    return out


def cpu_step(pos, vel, mass, rules, theta, eps, G, dt, cap, sample1, one_thread=True):
    """The oracle's Barnes-Hut step timed on this host: bounds, the sequential tree build with
    summarize, sort and force order (serial code), the per-body traversal of every body (OpenMP
    over all threads) and the Euler update, each part measured whole.  Also the per-body
    interaction count I_alg (exact) and a one-thread figure: the serial parts measured again on one
    thread and the traversal of `sample1` bodies on one thread, scaled to N."""
    from oracle import oracle as orc

    n = mass.size
    threads = orc.threads()
    t = [time.perf_counter()]
    box, _ = orc.bbox(pos)
    t.append(time.perf_counter())
    tree = orc.Tree(pos, mass, rules, cap=cap, box=box)
    t.append(time.perf_counter())
    acc, _, ni = tree.force(theta, eps, G, vote=1, want_absacc=False)
    t.append(time.perf_counter())
    orc.euler(acc, vel, pos, dt)
    t.append(time.perf_counter())
    parts = {"bounds": t[1] - t[0], "build_summarize_sort_serial": t[2] - t[1], "traversal": t[3] - t[2],
             "update": t[4] - t[3]}
    full = t[4] - t[0]
    if not one_thread:
        return {"full_s": full, "parts_s": parts, "threads": threads, "i_alg": float(ni.sum())}
    orc.set_threads(1)
    try:
        u = [time.perf_counter()]
        box1, _ = orc.bbox(pos)
        tree1 = orc.Tree(pos, mass, rules, cap=cap, box=box1)
        u.append(time.perf_counter())
        g = np.sort(np.random.default_rng(7).choice(n, sample1, replace=False)).astype(np.int32)
        tree1.force(theta, eps, G, vote=1, groups=g, want_absacc=False)
        u.append(time.perf_counter())
        orc.euler(acc, vel, pos, dt)
        u.append(time.perf_counter())
    finally:
        orc.set_threads(threads)
    one = (u[1] - u[0]) + (u[2] - u[1]) * n / sample1 + (u[3] - u[2])
    return {"full_s": full, "parts_s": parts, "threads": threads, "i_alg": float(ni.sum()),
            "one_thread_s": one, "one_thread_sample": sample1}


def bench_bh(args):
    from nbodysimulation_amd import _lib as L
    from nbodysimulation_amd import device_info, ic

    info = device_info()
    if info is None:
        raise SystemExit("no HIP device")
    rules = L.RULES_STANDARD if args.rules == "standard" else L.RULES_REFERENCE
    n, theta, eps, G, dt = args.n, args.theta, 0.01, 1.0, 1e-3
    log(f"bench: device {info}, N={n}, rules={args.rules}, theta={theta}")
    pos, vel, mass = ic.plummer(n, 1)

    main = run_bh(n, rules, theta, eps, G, dt, args.steps, args.warmup, (pos, vel, mass))
    steps_per_s, ms_per_step, stages, tinfo = main["steps_per_s"], main["ms_per_step"], main["stages"], main["tinfo"]
    exec_inter = main["exec_inter"]
    log(f"bench: {steps_per_s:.2f} steps/s, {ms_per_step:.3f} ms/step, stages(ms/step) "
        + ", ".join(f"{k}={v / args.steps:.3f}" for k, v in stages.items() if v))
    p_end, v_end, a_end = main["end"]
    sim = main["sim"]
    # host round trip through the C ABI (pageable buffers): what a caller that hands over
    # host arrays every step would pay on top of the resident-state rate
    t_ul = time.perf_counter()
    sim.upload(p_end, v_end, mass)
    t_ul = time.perf_counter() - t_ul
    sim.close()
    pcie = {"upload_ms": 1e3 * t_ul, "download_ms": 1e3 * main["download_s"],
            "steps_per_s_with_host_round_trip": 1.0 / (main["elapsed"] / args.steps + t_ul + main["download_s"])}

    shares = None if args.no_shares else target_shares(p_end, v_end, mass, rules, theta, eps, G, dt)

    # the other rule set on the same workload: its rate, and its accelerations on the same end state
    other_rules = L.RULES_REFERENCE if rules == L.RULES_STANDARD else L.RULES_STANDARD
    other_name = "reference" if other_rules == L.RULES_REFERENCE else "standard"
    companion = None
    if not args.no_companion:
        oth = run_bh(n, other_rules, theta, eps, G, dt, args.steps, args.warmup, (pos, vel, mass))
        oth["sim"].close()
        companion = {"rules": other_name, "value": oth["steps_per_s"], "unit": "timesteps/s",
                     "ms_per_step": oth["ms_per_step"],
                     "stage_ms_per_step": {k: v / args.steps for k, v in oth["stages"].items()},
                     "executed_interactions_per_step": oth["exec_inter"]}
        log(f"bench: {other_name} rules {oth['steps_per_s']:.2f} steps/s")

    from oracle import oracle as orc

    out_err = ref_vote = None
    if not args.no_error:
        tg = np.random.default_rng(12345).choice(n, 1024, replace=False).astype(np.int32)
        a64 = orc.direct_f64(p_end, mass, orc.LAW_PLUMMER, eps, G, targets=tg)
        n64 = np.linalg.norm(a64, axis=1)
        rel = np.linalg.norm(a_end[tg] - a64, axis=1) / n64
        out_err = {"median": float(np.median(rel)), "p99": float(np.quantile(rel, 0.99)), "targets": 1024,
                   "against": "FP64 direct sum", "state": "end of the timed steps"}
        log(f"bench: force rel err vs FP64 median {out_err['median']:.3e} p99 {out_err['p99']:.3e}")
        a_other = engine_force(n, other_rules, theta, eps, G, p_end, v_end, mass)
        rel_o = np.linalg.norm(a_other[tg] - a64, axis=1) / n64
        if companion is not None:
            companion["force_error"] = {"median": float(np.median(rel_o)), "p99": float(np.quantile(rel_o, 0.99)),
                                        "targets": 1024, "against": "FP64 direct sum",
                                        "state": "the same end state as force_error"}
        ref_vote = reference_vote_error(p_end, mass, theta, eps, G, {args.rules: a_end, other_name: a_other})
        log("bench: vs the reference kernels' work-group vote: "
            + ", ".join(f"{k} median {v['rel_diff_vs_reference_kernels']['median']:.3e}"
                        for k, v in ref_vote.items() if k.startswith("engine_")))

    # the oracle's step on the same state: the CPU baseline, and the per-body interaction count I_alg
    c = cpu_step(p_end, v_end, mass, rules, theta, eps, G, dt, tinfo["capacity"], min(args.cpu_sample, n),
                 one_thread=not args.no_cpu_baseline)
    i_alg = c["i_alg"]
    cpu = None
    if not args.no_cpu_baseline:
        cpu = {"value": 1.0 / c["full_s"], "unit": "timesteps/s", "cores": c["threads"], "host": host_cpu(),
               "kind": "port",
               "sample": (f"one whole step of the same workload on the oracle, measured: bounds, the sequential "
                          f"(one-thread) tree build, summarize, sort and force order, the per-body traversal of "
                          f"all {n} bodies on {c['threads']} threads, Euler update; {c['full_s']:.2f} s"),
               "parts_s": c["parts_s"],
               "one_core": {"value": 1.0 / c["one_thread_s"], "unit": "timesteps/s", "cores": 1,
                            "sample": (f"serial parts measured on one thread; traversal of {c['one_thread_sample']} "
                                       f"sampled bodies on one thread scaled to {n}; {c['one_thread_s']:.1f} s per step")}}
        log(f"bench: cpu baseline {cpu['value']:.4f} steps/s ({c['full_s']:.2f} s/step, {c['threads']} threads); "
            f"one thread {1.0 / c['one_thread_s']:.4f} steps/s")

    force_s = stages["force"] * 1e-3 / args.steps
    achieved = FLOPS_PER_INTERACTION * i_alg / force_s / 1e12
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic_force_bh_plummer_1M.json")
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get(args.rules, {}).get("bytes_per_launch")
    line = {
        "metric": METRIC,
        "value": steps_per_s,
        "unit": "timesteps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded Plummer sphere (AHW 1974), generated by the library",
        "config": {"workload": f"barnes-hut Plummer N={n} theta={theta} eps={eps} G={G} dt={dt} Euler seed 1",
                   "rules": args.rules, "vote_group": L.vote_group(), "cell_capacity": tinfo["capacity"],
                   "cells_used": tinfo["num_cells"], "max_depth": tinfo["max_depth"],
                   "device": info},
        "ginteractions_per_s": i_alg * steps_per_s / 1e9,
        "ginteractions_executed_per_s": exec_inter * steps_per_s / 1e9,
        "interactions_per_step": {"algorithmic_per_body_traversal": i_alg, "executed_wave_vote": exec_inter},
        "stage_ms_per_step": {k: v / args.steps for k, v in stages.items()},
        "stage_bandwidth": stage_bandwidth(n, tinfo["num_cells"], stages, args.steps),
        "stage_timing": "HIP events around every stage of the same steps launched stage by stage; value and "
                        "ms_per_step come from replayed graphs of the whole step",
        "host_round_trip": pcie,
        "target_shares": shares,
        "force_error": out_err,
        "other_rules": companion,
        "reference_vote_error": ref_vote,
        "roofline": {"bound": "mfma", "note": "FP32 vector ALU (no MFMA); peak = MI355X FP32 157.3 TFLOP/s",
                     "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                     "kernel": "k_force_bh", "kernel_ms": force_s * 1e3,
                     "flops_per_launch": FLOPS_PER_INTERACTION * i_alg},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


# Algorithmic HBM bytes per step of the tree stages and the update (N bodies, C cells used):
# bounds 12N, reset 32(C+1), build 16N+20C, summarize 16N+72C, sort 4N+40C (one permutation out;
# per cell its child slots and a count), update 60N.  This design's sort reads each body's position
# and writes two permutations (tree order and force order) and reads 8 per-slot counts per cell:
# 24N+64C, reported beside the credited figure.
STAGE_BYTES = {"bounds": lambda n, c: 12 * n, "reset": lambda n, c: 32 * (c + 1),
               "build": lambda n, c: 16 * n + 20 * c, "summarize": lambda n, c: 16 * n + 72 * c,
               "sort": lambda n, c: 4 * n + 40 * c, "update": lambda n, c: 60 * n}
SORT_BYTES_THIS_DESIGN = lambda n, c: 24 * n + 64 * c  # noqa: E731
PEAK_HBM_GBS = 8000.0


def stage_bandwidth(n, cells, stages, steps):
    """Algorithmic bytes over each stage's event-timed duration, against MI355X's 8 TB/s."""
    out = {}
    for k, f in STAGE_BYTES.items():
        ms = stages.get(k, 0.0) / steps
        if ms > 0:
            gbs = f(n, cells) / (ms * 1e-3) / 1e9
            out[k] = {"algorithmic_bytes": f(n, cells), "ms": ms, "GB_s": gbs, "frac": gbs / PEAK_HBM_GBS}
    if "sort" in out:
        b = SORT_BYTES_THIS_DESIGN(n, cells)
        out["sort"]["bytes_this_design"] = b
        out["sort"]["GB_s_this_design"] = b / (out["sort"]["ms"] * 1e-3) / 1e9
    return out


def target_shares(pos, vel, mass, rules, theta, eps, G, dt, gpus=(2, 4, 8), reps=5):
    """One GPU's share of the force pass when the targets are split over G GPUs that each hold
    the whole tree: the force stage over [r*N/G, (r+1)*N/G) of the force order, timed with HIP
    events on the simulation's stream (mean of `reps` launches after one warm-up), for every rank
    r.  The all-gather of positions that such a split needs is not part of it.  The full pass is timed
    the same way; repeated without a new sort it has no dispatch-order keys, so like the shares it
    runs its workgroups in table order."""
    import ctypes as C

    from nbodysimulation_amd import _lib as L
    from nbodysimulation_amd.engine import NBodySim, make_config

    n = mass.size
    stream, hip = hip_stream()
    ev0, ev1 = C.c_void_p(), C.c_void_p()
    hip.hipEventCreate(C.byref(ev0))
    hip.hipEventCreate(C.byref(ev1))

    def timed(launch):
        launch()
        hip.hipEventRecord(ev0, C.c_void_p(stream))
        for _ in range(reps):
            launch()
        hip.hipEventRecord(ev1, C.c_void_p(stream))
        hip.hipEventSynchronize(ev1)
        ms = C.c_float()
        hip.hipEventElapsedTime(C.byref(ms), ev0, ev1)
        return ms.value / reps

    out = {"reps": reps, "order": "force order, rank r takes [r*N/G, (r+1)*N/G)"}
    with NBodySim(make_config(n, L.BARNES_HUT, rules=rules, theta=theta, eps=eps, G=G, dt=dt), stream=stream) as sim:
        sim.upload(pos, vel, mass)
        sim.force_pipeline()
        out["full_force_ms"] = timed(sim.force)
        for g in gpus:
            ms = [timed(lambda a=n * r // g, b=n * (r + 1) // g: sim.force_range(a, b - a)) for r in range(g)]
            out[str(g)] = {"share_force_ms": ms, "max_share_force_ms": max(ms), "sum_ms": sum(ms)}
            log(f"bench: force share at G={g}: max {max(ms):.3f} ms, sum {sum(ms):.3f} ms "
                f"(full pass {out['full_force_ms']:.3f} ms)")
        assert sim.status() == L.NB_OK
    hip.hipEventDestroy(ev0)
    hip.hipEventDestroy(ev1)
    hip.hipStreamDestroy(C.c_void_p(stream))
    return out


def hip_stream():
    """A HIP stream for the one-system-per-launch baseline (plumbing via ctypes)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    st = C.c_void_p()
    if hip.hipStreamCreate(C.byref(st)) != 0:
        raise RuntimeError("hipStreamCreate failed")
    return st.value, hip


def ensemble_force_error(specs, ics, systems=32, targets=128, seed=5):
    """FP64 force error of the ensemble workload: one force pass on the initial state of `systems`
    sampled systems (each system's arithmetic is the same alone and batched), against an FP64
    direct sum on `targets` sampled bodies of each."""
    from nbodysimulation_amd import _lib as L
    from nbodysimulation_amd.engine import NBodySim, make_config
    from oracle import oracle as orc

    r = np.random.default_rng(seed)
    rel = {L.BARNES_HUT: [], L.BRUTE_FORCE_GLOBAL: []}
    for i in np.sort(r.choice(len(specs), systems, replace=False)):
        sp, (p, v, m) = specs[i], ics[i]
        n = sp.num_bodies
        with NBodySim(make_config(n, sp.algorithm, rules=sp.rules, law=L.LAW_PLUMMER, theta=sp.theta,
                                  eps=sp.epsilon, G=sp.gravitational_constant,
                                  cell_capacity=2 * n + 1024)) as s:
            s.upload(p, v, m)
            if sp.algorithm == L.BARNES_HUT:
                s.force_pipeline()
            else:
                s.brute_force()
            a = s.download()[2]
        tg = np.sort(r.choice(n, min(targets, n), replace=False)).astype(np.int32)
        a64 = orc.direct_f64(p, m, orc.LAW_PLUMMER, sp.epsilon, sp.gravitational_constant, targets=tg)
        rel[sp.algorithm].append(np.linalg.norm(a[tg] - a64, axis=1) / np.linalg.norm(a64, axis=1))
    allr = np.concatenate(rel[L.BARNES_HUT] + rel[L.BRUTE_FORCE_GLOBAL])
    out = {"median": float(np.median(allr)), "p99": float(np.quantile(allr, 0.99)), "systems": systems,
           "targets_per_system": targets, "against": "FP64 direct sum"}
    for name, k in (("barnes_hut", L.BARNES_HUT), ("brute_force", L.BRUTE_FORCE_GLOBAL)):
        if rel[k]:
            x = np.concatenate(rel[k])
            out[name] = {"median": float(np.median(x)), "p99": float(np.quantile(x, 0.99)), "systems": len(rel[k])}
    return out


def bench_ensemble(args):
    """512 independent systems, N log-uniform in [1024, 65536], each a Plummer, Normal(0,1) or
    uniform-ball cloud (equal odds), m = 1/N, eps = 0.01, dt = 1e-3, seed 2; brute force below the
    autotuned crossover N, Barnes-Hut (STANDARD, theta = 0.5) from it on.  One batched, graph-captured
    step for all systems against the same systems stepped one system per launch."""
    from nbodysimulation_amd import _lib as L
    from nbodysimulation_amd import autotune, device_info
    from nbodysimulation_amd import workloads as W
    from nbodysimulation_amd.engine import NBodyEnsemble, NBodySim, make_config

    info = device_info()
    if info is None:
        raise SystemExit("no HIP device")
    tune = autotune.load_or_tune(0.5, 0.01, log=log)
    cross = int(tune["crossover_n"])
    count = W.ENSEMBLE_SYSTEMS
    sizes, kinds = W.ensemble_sizes_kinds()
    ics = [W.ensemble_ic(i, n, k) for i, (n, k) in enumerate(zip(sizes, kinds))]
    specs = W.ensemble_specs(sizes, cross)
    algos = [sp.algorithm for sp in specs]
    steps, warm = args.steps, args.warmup
    log(f"bench: ensemble of {count} systems, {int(sizes.sum())} bodies, crossover N={cross}, "
        f"{sum(a == L.BARNES_HUT for a in algos)} Barnes-Hut / {sum(a != L.BARNES_HUT for a in algos)} brute force")

    with NBodyEnsemble(specs, law=L.LAW_PLUMMER) as ens:
        for i, (p, v, m) in enumerate(ics):
            ens.upload(i, p, v, m)
        ens.step(max(warm, 1))
        t0 = time.perf_counter()
        ens.step(steps)
        t_batched = time.perf_counter() - t0
        hbm = ens.device_bytes()

    def plain(shared_stream):
        stream = hip = None
        if shared_stream:
            stream, hip = hip_stream()
        sims = []
        for sp, (p, v, m) in zip(specs, ics):
            s = NBodySim(make_config(sp.num_bodies, sp.algorithm, rules=sp.rules, law=L.LAW_PLUMMER, theta=sp.theta,
                                     eps=sp.epsilon, dt=sp.dt, G=1.0, cell_capacity=2 * sp.num_bodies + 1024),
                         stream=stream)
            s.upload(p, v, m)
            sims.append(s)
        for s in sims:
            s.step_async(max(warm, 1))
        for s in sims:
            s.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            for s in sims:
                s.step_async(1)
        for s in sims:
            s.sync()
        t = time.perf_counter() - t0
        for s in sims:
            s.close()
        if hip is not None:
            hip.hipStreamDestroy(__import__("ctypes").c_void_p(stream))
        return t

    # the comparison legs (one system per launch); --no-plain skips them (profiling the batched step)
    t_plain_shared = None if args.no_plain else plain(True)
    t_plain_own = None if args.no_plain else plain(False)
    err = ensemble_force_error(specs, ics)
    log(f"bench: ensemble force error vs FP64 median {err['median']:.3e} p99 {err['p99']:.3e}")
    bodies = int(sizes.sum())
    line = {
        "metric": "ensemble steps/s (512 systems, one batched graph-captured step)",
        "value": steps / t_batched, "unit": "ensemble steps/s", "n_gpus": 1, "steps": steps, "warmup": warm,
        "ms_per_step": 1e3 * t_batched / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic: seeded Plummer / Normal(0,1) / unit-ball clouds",
        "config": {"workload": "512 systems, N log-uniform [1024, 65536], seed 2", "bodies": bodies,
                   "crossover_n": cross, "barnes_hut_systems": int(sum(a == L.BARNES_HUT for a in algos)),
                   "hbm_bytes": hbm, "device": info},
        "bodies_steps_per_s": bodies * steps / t_batched,
        "plain_one_system_per_launch": None if args.no_plain else {
            "shared_stream_steps_per_s": steps / t_plain_shared,
            "per_system_streams_steps_per_s": steps / t_plain_own,
            "speedup_vs_shared_stream": t_plain_shared / t_batched,
            "speedup_vs_per_system_streams": t_plain_own / t_batched},
        "autotune": tune,
        "force_error": err,
    }
    if args.no_plain:
        log(f"bench: batched {steps / t_batched:.2f} ensemble steps/s")
    else:
        log(f"bench: batched {steps / t_batched:.2f} ensemble steps/s; plain shared stream "
            f"{steps / t_plain_shared:.2f}, own streams {steps / t_plain_own:.2f}")
    print(json.dumps(line), flush=True)


def bench_precision(args):
    """Brute force at N = 131,072: two Plummer spheres of 65,536 equal-mass bodies (a = 0.5,
    M = 1 each, centres +-(2,0,0), bulk velocities -+(0.3,0,0)), eps = 0.005, G = 1, dt = 1e-3,
    Euler, seed 3.  Plain and Kahan-compensated FP32 accumulation: force error against an FP64
    direct sum on 1024 seeded targets, then 200 steps each with FP64 energy / momentum /
    angular momentum computed on the GPU before and after."""
    from nbodysimulation_amd import _lib as L
    from nbodysimulation_amd import autotune, device_info, ic
    from nbodysimulation_amd.engine import NBodySim, make_config
    from oracle import oracle as orc

    info = device_info()
    if info is None:
        raise SystemExit("no HIP device")
    shape = autotune.load_or_tune(0.5, 0.01, log=log)["brute_shape"]
    n, eps, G, dt, steps = 131072, 0.005, 1.0, 1e-3, args.steps
    pos, vel, mass = ic.collision(n, 3)
    tg = np.random.default_rng(12345).choice(n, 1024, replace=False).astype(np.int32)
    a64 = orc.direct_f64(pos, mass, L.LAW_PLUMMER, eps, G, targets=tg)
    out = {}
    for name, accum in (("plain", L.ACCUM_PLAIN), ("kahan", L.ACCUM_KAHAN)):
        cfg = dict(law=L.LAW_PLUMMER, accumulation=accum, eps=eps, G=G, dt=dt)
        with NBodySim(make_config(n, L.BRUTE_FORCE_GLOBAL, **cfg)) as s:
            s.upload(pos, vel, mass)
            s.brute_force()
            _, _, a = s.download()
            rel = np.linalg.norm(a[tg] - a64, axis=1) / np.linalg.norm(a64, axis=1)
            e0 = s.energy()
            s.step(1)  # graph capture outside the timed run
            s.upload(pos, vel, mass)
            t0 = time.perf_counter()
            s.step(steps)
            wall = time.perf_counter() - t0
            e1 = s.energy()
        # kernel time from HIP events on the simulation's stream (a separate short run)
        with NBodySim(make_config(n, L.BRUTE_FORCE_GLOBAL, time_stages=True, **cfg)) as s:
            s.upload(pos, vel, mass)
            s.step(args.warmup or 1)
            s.step(10)
            brute_ms = s.stage_times()["brute"] / 10
        pscale = float(np.abs(mass[:, None] * vel).sum())
        lscale = float((mass * np.linalg.norm(pos, axis=1) * np.linalg.norm(vel, axis=1)).sum())
        out[name] = {
            "force_rel_err_vs_fp64": {"median": float(np.median(rel)), "p99": float(np.quantile(rel, 0.99)),
                                      "max": float(rel.max()), "targets": 1024},
            "steps_per_s": steps / wall, "brute_kernel_ms": brute_ms,
            "energy_start": e0["total"], "energy_end": e1["total"],
            "energy_rel_drift": abs(e1["total"] - e0["total"]) / abs(e0["total"]),
            "momentum_drift_rel": float(np.linalg.norm(e1["momentum"] - e0["momentum"]) / pscale),
            "angular_momentum_drift_rel": float(np.linalg.norm(e1["angular"] - e0["angular"]) / lscale),
        }
        log(f"bench: {name}: median err {out[name]['force_rel_err_vs_fp64']['median']:.3e}, "
            f"{steps / wall:.1f} steps/s, brute {brute_ms:.3f} ms, dE/E {out[name]['energy_rel_drift']:.3e}")

    # CPU baseline: the oracle's brute force on 2048 sampled targets, scaled to N
    cpu = None
    if not args.no_cpu_baseline:
        ts = np.random.default_rng(5).choice(n, 2048, replace=False).astype(np.int32)
        t0 = time.perf_counter()
        orc.force_bf(pos, mass, L.LAW_PLUMMER, eps, G, targets=ts)
        t_cpu = (time.perf_counter() - t0) * n / ts.size
        cpu = {"value": 1.0 / t_cpu, "unit": "timesteps/s", "cores": orc.threads(), "host": host_cpu(), "kind": "port",
               "sample": f"oracle brute-force force pass (plain FP32) on 2048 of {n} targets, scaled to {n}; "
                         f"{t_cpu:.2f} s per step"}
    flops = FLOPS_PER_INTERACTION * float(n) * n
    achieved = flops / (out["plain"]["brute_kernel_ms"] * 1e-3) / 1e12
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic_brute_bf_precision_131072.json")
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("plain", {}).get("bytes_per_launch")
    line = {
        "metric": "brute-force timesteps/s at N=131072 (plain FP32; Kahan and FP64 diagnostics beside)",
        "value": out["plain"]["steps_per_s"], "unit": "timesteps/s", "n_gpus": 1, "steps": steps,
        "warmup": 1, "ms_per_step": 1e3 / out["plain"]["steps_per_s"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: two seeded Plummer spheres (library generator)",
        "config": {"workload": f"brute force, two colliding Plummer spheres N={n} eps={eps} G={G} dt={dt} Euler "
                               f"seed 3", "device": info,
                   "brute_shape": {k: shape[k] for k in ("bf_targets_per_lane", "bf_min_waves")}},
        "ginteractions_per_s": float(n) * n * out["plain"]["steps_per_s"] / 1e9,
        "plain": out["plain"], "kahan": out["kahan"],
        "roofline": {"bound": "mfma", "note": "FP32 vector ALU (no MFMA); peak = MI355X FP32 157.3 TFLOP/s",
                     "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic, "kernel": "k_brute",
                     "kernel_ms": out["plain"]["brute_kernel_ms"], "flops_per_launch": flops},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def serving_run(sv, rq, idx, arrivals, log_every=10.0):
    """Drive the server over the requests `idx` of the stream `rq`: each is submitted at its
    arrival time (arrivals None: a backlog, every request submitted at time 0 with its initial
    conditions generated before the clock starts), the server ticks while it has work.  Open-loop
    initial conditions are generated by the client in a thread pool, ahead of use.  Returns
    per-request submit and finish times, per-tick step times and the span."""
    from concurrent.futures import ThreadPoolExecutor

    from nbodysimulation_amd import _lib as L
    from nbodysimulation_amd import workloads as W

    R = len(idx)
    pool = ThreadPoolExecutor(max_workers=8)  # the generator releases the GIL
    futs = {}

    def next_ic(k):  # the k-th request's state, generated up to 32 requests ahead
        for j in range(k, min(k + 32, R)):
            if j not in futs:
                futs[j] = pool.submit(W.serving_ic, int(idx[j]), int(rq.sizes[idx[j]]))
        return futs.pop(k).result()

    pre = [next_ic(k) for k in range(R)] if arrivals is None else None
    ids = {}
    submit = np.zeros(R)
    finish = np.zeros(R)
    tick_s = []
    failed = nxt = done = 0
    t0 = time.perf_counter()  # after the backlog's initial conditions exist
    last_log = t0
    while done < R:
        now = time.perf_counter() - t0
        while nxt < R and (arrivals is None or arrivals[nxt] <= now):
            p, v, m = next_ic(nxt) if pre is None else pre[nxt]
            if pre is not None:
                pre[nxt] = None
            i = idx[nxt]
            ids[sv.submit(rq.specs[i], int(rq.steps[i]), p, v, m)] = nxt
            submit[nxt] = time.perf_counter() - t0 if arrivals is None else arrivals[nxt]
            nxt += 1
            now = time.perf_counter() - t0
        st = sv.stats()
        if st["active"] == 0 and st["queued"] == 0 and st["finished"] == 0:
            if nxt < R and arrivals is not None:
                time.sleep(max(0.0, arrivals[nxt] - (time.perf_counter() - t0)))
            continue
        ticks_before = st["ticks"]
        tk = time.perf_counter()
        sv.tick()
        dt_tick = time.perf_counter() - tk
        if sv.stats()["ticks"] > ticks_before:  # a step ran in this tick
            tick_s.append(dt_tick)
        while (r := sv.poll()) is not None:
            k = ids.pop(r[0])
            finish[k] = time.perf_counter() - t0
            failed += r[1] != L.NB_OK
            sv.fetch(r[0])
            done += 1
        if time.perf_counter() - last_log > log_every:
            last_log = time.perf_counter()
            st = sv.stats()
            log(f"bench: t={last_log - t0:7.1f}s submitted {nxt} done {done} active {st['active']} "
                f"queued {st['queued']} ticks {st['ticks']}")
    pool.shutdown()
    return submit, finish, np.array(tick_s), float(finish.max()), failed


def serving_force_error(rq, cross, per_theta=2, targets=256, seed=11):
    """FP64 force error of served requests: one force pass on the initial state of `per_theta`
    requests per theta (each request's arithmetic is the same alone and served), against an FP64
    direct sum on `targets` sampled bodies of each."""
    from nbodysimulation_amd import _lib as L
    from nbodysimulation_amd import workloads as W
    from nbodysimulation_amd.engine import NBodySim, make_config
    from oracle import oracle as orc

    r = np.random.default_rng(seed)
    out = {}
    for th in W.SERVING_THETAS:
        cand = np.flatnonzero(np.isclose(rq.thetas, th))
        rel = []
        for i in r.choice(cand, per_theta, replace=False):
            sp = rq.specs[i]
            n = sp.num_bodies
            p, v, m = W.serving_ic(int(i), n)
            with NBodySim(make_config(n, sp.algorithm, rules=sp.rules, law=L.LAW_PLUMMER, theta=sp.theta,
                                      eps=sp.epsilon, G=sp.gravitational_constant, cell_capacity=2 * n + 1024)) as s:
                s.upload(p, v, m)
                if sp.algorithm == L.BARNES_HUT:
                    s.force_pipeline()
                else:
                    s.brute_force()
                a = s.download()[2]
            tg = np.sort(r.choice(n, min(targets, n), replace=False)).astype(np.int32)
            a64 = orc.direct_f64(p, m, orc.LAW_PLUMMER, sp.epsilon, sp.gravitational_constant, targets=tg)
            rel.append(np.linalg.norm(a[tg] - a64, axis=1) / np.linalg.norm(a64, axis=1))
        x = np.concatenate(rel)
        out[f"theta={th}"] = {"median": float(np.median(x)), "p99": float(np.quantile(x, 0.99)),
                              "requests": per_theta, "targets_per_request": targets}
    out["against"] = "FP64 direct sum"
    return out


def bench_serving(args):
    """Continuous-batching server on the request stream: each request a new Plummer system
    (M = 1, a = 1, seed 3-derived), N log-uniform in [4096, 2,097,152], steps uniform in
    [10, 200], theta from {0.3, 0.5, 0.7}, eps = 0.01, dt = 1e-3, Euler; brute force below the
    autotuned crossover N, Barnes-Hut (STANDARD) from it on.  Two runs on one server:
    (1) saturation: the first --capacity-requests requests as a backlog (all submitted at
    once), whose bodies x steps per second is the server's capacity, `value`; (2) the open-loop
    Poisson stream of --requests requests at --rate requests/s (default: the rate that offers
    --load of the measured capacity), with latency from the scheduled arrival to the tick that
    retires the request.  Both report per-tick step time percentiles (host wall time of the
    ticks that ran a step, admission uploads and retirement downloads included)."""
    from nbodysimulation_amd import autotune, device_info
    from nbodysimulation_amd import workloads as W
    from nbodysimulation_amd.engine import NBodyServer

    info = device_info()
    if info is None:
        raise SystemExit("no HIP device")
    tune = autotune.load_or_tune(0.5, 0.01, log=log)
    cross = int(tune["crossover_n"])
    R = args.requests
    rq0 = W.serving_requests(R, 1.0, cross)  # arrival times are rescaled below
    mean_work = float((rq0.sizes * rq0.steps).mean())
    budget = int(args.budget_gib * (1 << 30))

    def pct(x):
        return {"p50": float(np.quantile(x, 0.5)), "p99": float(np.quantile(x, 0.99)), "max": float(x.max()),
                "mean": float(x.mean()), "count": int(x.size)} if len(x) else None

    with NBodyServer(budget, args.max_systems) as sv:
        # (1) saturation throughput, closed loop
        K = min(args.capacity_requests, R)
        log(f"bench: serving capacity run, a backlog of {K} requests, budget {args.budget_gib} GiB, "
            f"{args.max_systems} slots, crossover N={cross}")
        s0 = sv.stats()
        _, _, ticks_c, span_c, failed_c = serving_run(sv, rq0, list(range(K)), None)
        s1 = sv.stats()
        capacity = (s1["body_steps"] - s0["body_steps"]) / span_c
        log(f"bench: capacity {capacity:.3e} bodies*steps/s over {span_c:.1f}s, tick p50 "
            f"{1e3 * np.quantile(ticks_c, 0.5):.1f} ms p99 {1e3 * np.quantile(ticks_c, 0.99):.1f} ms")
        # (2) open-loop Poisson arrivals at a stated fraction of the capacity
        rate = args.rate if args.rate else args.load * capacity / mean_work
        rq = W.serving_requests(R, rate, cross)
        log(f"bench: serving {R} requests open-loop at {rate:.2f} req/s "
            f"(offered {rate * mean_work:.3e} bodies*steps/s = {rate * mean_work / capacity:.2f} of capacity)")
        s2 = sv.stats()
        sub, fin, ticks_o, span_o, failed_o = serving_run(sv, rq, list(range(R)), rq.arrivals)
        s3 = sv.stats()
        arena = s3["arena_bytes"]
    lat = fin - rq.arrivals
    served = (s3["body_steps"] - s2["body_steps"]) / span_o
    err = serving_force_error(rq, cross)
    line = {
        "metric": "serving capacity, bodies*steps/s (continuous batching), with p50/p99 request and step latency",
        "value": capacity, "unit": "bodies*steps/s", "n_gpus": 1, "steps": int(rq.steps.sum()),
        "warmup": 0, "ms_per_step": 1e3 * float(np.quantile(ticks_o, 0.5)), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: seeded Plummer spheres generated by the client while the server runs",
        "config": {"workload": f"{R} requests, Poisson arrivals, N log-uniform [4096, 2097152], steps U[10,200], "
                               f"theta in {{0.3,0.5,0.7}}, seed 3", "hbm_budget_bytes": budget, "arena_bytes": arena,
                   "max_systems": args.max_systems, "crossover_n": cross, "device": info},
        "capacity": {"body_steps_per_s": capacity, "requests": K, "span_s": span_c, "failed": int(failed_c),
                     "tick_step_s": pct(ticks_c),
                     "how": "backlog: the first requests of the stream all submitted at time 0 (initial conditions "
                            "generated before the clock); bodies x steps over the time to finish them all"},
        "open_loop": {"rate_req_per_s": rate, "offered_body_steps_per_s": rate * mean_work,
                      "offered_fraction_of_capacity": rate * mean_work / capacity,
                      "served_body_steps_per_s": served, "span_s": span_o, "failed": int(failed_o),
                      "latency_s": pct(lat), "tick_step_s": pct(ticks_o)},
        "force_error": err,
    }
    log(f"bench: open loop served {served:.3e} bodies*steps/s, latency p50 {line['open_loop']['latency_s']['p50']:.2f}s "
        f"p99 {line['open_loop']['latency_s']['p99']:.2f}s; tick p50 {1e3 * np.quantile(ticks_o, 0.5):.1f} ms "
        f"p99 {1e3 * np.quantile(ticks_o, 0.99):.1f} ms")
    print(json.dumps(line), flush=True)


REPEATS_BF4096 = 7


def bench_bf4096(args):
    """Plain brute force, N = 4096: x ~ Normal(0,1) per axis, m ~ U[0.5,1.5]/N, v = 0, G = 1,
    dt = 1e-3, Euler, seed 0, the reference update kernel's law (unit vectors, e = 0.01); the
    Plummer-softened law (eps = 0.01) is timed beside it.  Steps replay captured HIP graphs with
    the integrator fused into the brute-force kernel; `value` is the median rate of 7 runs of the
    whole workload from the same initial state.  The per-kernel times come from a separate run
    with stage timing (every stage launched and timed on its own)."""
    from nbodysimulation_amd import _lib as L
    from nbodysimulation_amd import autotune, device_info, ic
    from nbodysimulation_amd.engine import NBodySim, make_config
    from oracle import oracle as orc

    info = device_info()
    if info is None:
        raise SystemExit("no HIP device")
    shape = autotune.load_or_tune(0.5, 0.01, log=log)["brute_shape"]
    n, dt, steps = 4096, 1e-3, args.steps
    pos, vel, mass = ic.normal_cloud(n, 0)
    res = {}
    for name, law in (("reference_unit", L.LAW_REFERENCE_UNIT), ("plummer", L.LAW_PLUMMER)):
        cfg = dict(law=law, eps=0.01, G=1.0, dt=dt)
        with NBodySim(make_config(n, L.BRUTE_FORCE_GLOBAL, **cfg)) as s:
            s.upload(pos, vel, mass)
            s.step(max(args.warmup, 1))
            walls = []
            for _ in range(REPEATS_BF4096):  # the whole run from the same initial state, median of the repeats
                s.upload(pos, vel, mass)
                t0 = time.perf_counter()
                s.step(steps)
                walls.append(time.perf_counter() - t0)
            wall = float(np.median(walls))
        with NBodySim(make_config(n, L.BRUTE_FORCE_GLOBAL, time_stages=True, **cfg)) as s:
            s.upload(pos, vel, mass)
            s.step(max(args.warmup, 1))
            s.step(steps)
            st = s.stage_times()
        # force error of one pass on the initial state against an FP64 direct sum of the same law
        with NBodySim(make_config(n, L.BRUTE_FORCE_GLOBAL, **cfg)) as s:
            s.upload(pos, vel, mass)
            s.brute_force()
            a32 = s.download()[2]
        a64 = orc.direct_f64(pos, mass, law, 0.01, 1.0)
        rel = np.linalg.norm(a32 - a64, axis=1) / np.linalg.norm(a64, axis=1)
        res[name] = {"steps_per_s": steps / wall, "brute_kernel_ms": st["brute"] / steps,
                     "update_kernel_ms": st["update"] / steps,
                     "steps_per_s_repeats": [steps / w for w in walls],
                     "force_error": {"median": float(np.median(rel)), "p99": float(np.quantile(rel, 0.99)),
                                     "targets": n, "against": "FP64 direct sum of the same law"}}
        log(f"bench: bf4096 {name}: {steps / wall:.1f} steps/s, brute {st['brute'] / steps * 1e3:.2f} us")
    cpu = None
    if not args.no_cpu_baseline:
        p, v = pos.copy(), vel.copy()
        k = 20
        t0 = time.perf_counter()
        for _ in range(k):
            a, _ = orc.force_bf(p, mass, orc.LAW_REFERENCE_UNIT, 0.01, 1.0)
            v, p = orc.euler(a, v, p, dt)
        t_cpu = (time.perf_counter() - t0) / k
        cpu = {"value": 1.0 / t_cpu, "unit": "timesteps/s", "cores": orc.threads(), "host": host_cpu(), "kind": "port",
               "sample": f"{k} full steps of the same workload on the oracle (all-pairs force + Euler)"}
    r = res["reference_unit"]
    flops = FLOPS_PER_INTERACTION * float(n) * n
    achieved = flops / (r["brute_kernel_ms"] * 1e-3) / 1e12
    line = {
        "metric": "brute-force timesteps/s at N=4096 (reference update-kernel law)",
        "value": r["steps_per_s"], "unit": "timesteps/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": 1e3 / r["steps_per_s"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic: seeded Normal(0,1) cloud, m ~ U[0.5,1.5]/N (library generator)",
        "config": {"workload": "plain brute force N=4096, Normal(0,1), G=1, dt=1e-3, 100 steps, seed 0",
                   "device": info, "brute_shape": {k: shape[k] for k in ("bf_targets_per_lane", "bf_min_waves")}},
        "laws": res,
        "roofline": {"bound": "mfma", "note": "FP32 vector ALU (no MFMA); peak = MI355X FP32 157.3 TFLOP/s; "
                                              "at N=4096 one launch is a few microseconds",
                     "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None, "kernel": "k_brute",
                     "kernel_ms": r["brute_kernel_ms"], "flops_per_launch": flops},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
