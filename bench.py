#!/usr/bin/env python3
"""Full P3T step throughput on one MI355X (particle-steps/s at N = 1,000,000).

Workload: planetesimal disk, N = 1,000,000, m = 1e-12, a in [0.5, 2] AU with surface
density ~ a^-1.5, Rayleigh e/i (rms e = 2 h, rms i = h), theta = 0.5 with quadrupole
super-particles, n_group_limit 256, n_leaf_limit 8, r_out = 3 mutual Hill radii
(R_cut0 = 3*2^(1/3)), r_in/r_out = 0.1, dt_tree = 2^-4 (the power-of-two soft step
nearest below 2pi/64), eta = 0.01, perfect-merging collisions on, seed 1.
A step is one full soft step (main_p3t.cpp:423-699): kick, clusters, batched Hermite +
Kepler drift, tree, FP32 EP-EP/EP-SP, FP64 neighbour finalize, kick, energy, merging and
boundary removal, and the Initial recompute after a collision.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--no-cpu]
    python bench.py --workload hard [--clusters 50000]     (batched hard part alone)
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s at N=1M (full P3T step); soft-force interactions/s vs FP64"
PEAK_FP32_TFLOPS = 157.3          # MI355X published dense FP32 vector peak
FLOP_EPEP = 28                    # per EP-EP pair (reference op list)
FLOP_EPSP_QUAD = 60               # per EP-SP quadrupole pair
FLOP_SPJ_QUAD = 12                # per staged quadrupole super-particle (hoisted terms)
FLOP_EPSP_MONO = 25


def params_dict():
    return dict(R_cut0=3.0 * 2 ** (1 / 3), R_cut1=8.0, R_search0=1.1, R_search1=6.0, R_search2=1.1,
                R_search3=2.0, gamma=0.1, eta=0.01, eta_sun=0.01, eta_0=0.002, eta_sun0=0.002,
                dt_tree=2.0 ** -4, dt_min=2.0 ** -30, n_group_limit=256, n_leaf_limit=8, theta=0.5,
                quadrupole=1, collision=1)


def make_disk(n, seed=1):
    from gplum_amd import ic
    return ic.disk(n, 1e-12, 0.5, 2.0, seed=seed)


def measured_traffic(name, key=None):
    """HBM bytes per launch of a kernel from its committed counter summary under
    profiles/ (FETCH_SIZE and WRITE_SIZE passes scaled by a plain copy), for the launch
    shape `key` (e.g. "n_group_limit=128") when the summary lists several; None if absent."""
    import re
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    tag = r"\[" + re.escape(key) + r"\]" if key else ""
    m = re.search(r"^traffic_bytes_per_launch" + tag + r":\s*(\d+)", open(path).read(), re.M)
    return int(m.group(1)) if m else None


def cpu_setup(threads):
    """The CPU legs run the oracle built on this host with -O3 -march=native (so that it
    uses the host's own vector ISA), OpenMP threads bound close to the cores.  Returns the
    host description recorded with the figures."""
    os.environ["OMP_NUM_THREADS"] = str(threads)
    os.environ["OMP_PROC_BIND"] = "close"
    os.environ["OMP_PLACES"] = "cores"
    from oracle import oracle as O
    native = O.use_native_build()
    model, isa = "unknown", "unknown"
    try:
        import subprocess
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in txt.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
            if line.startswith("Flags:"):
                fl = set(line.split(":", 1)[1].split())
                isa = "avx512f" if "avx512f" in fl else ("avx2" if "avx2" in fl else "sse")
    except (OSError, ValueError):
        pass
    return {"cpu_model": model, "isa": isa, "build": "gcc -O3 -march=native" if native else
            "gcc -O2 -march=x86-64-v3 (native build failed)", "omp": "OMP_PROC_BIND=close, OMP_PLACES=cores"}


def cpu_threads_default():
    """Host threads for the CPU legs: the cores this process may run on, at most 16 (the
    host share of one GPU on the measurement machines)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return max(1, min(16, avail))


def cpu_soft_force(n, seed, threads, pp=None):
    """Oracle FP32 soft force (PIKG-order kernels, OpenMP over groups) on the same disk:
    interactions per second on `threads` host threads."""
    from oracle import oracle as O
    p = O.params(**params_dict())
    pp = make_disk(n, seed) if pp is None else pp.copy()
    O.set_cutoff_radii(p, pp)
    tree = O.tree_build(p, pp)
    t0 = time.perf_counter()
    nep, nsp = O.soft_force(p, pp, tree)
    dt = time.perf_counter() - t0
    tree.free()
    return {"interactions_per_s": round((nep + nsp) / dt, 1), "epep_pairs": int(nep), "epsp_pairs": int(nsp),
            "seconds": round(dt, 3), "cores": threads}


def cpu_baseline(n, seed, threads, host):
    """Oracle full P3T step (CPU restatement, OpenMP) on the same workload: five steps
    after initialisation, timed on `threads` host threads."""
    from oracle import oracle as O
    p = O.params(**params_dict())
    pp = make_disk(n, seed)
    run = O.P3T(p, pp)
    steps = 5
    t0 = time.perf_counter()
    for _ in range(steps):
        run.step()
    dt = time.perf_counter() - t0
    run.close()
    out = {"value": n * steps / dt, "unit": "particle-steps/s", "cores": threads, "kind": "port",
           "sample": "%d full soft steps of the same N=%d disk after initialisation, "
                     "oracle C restatement (PIKG-order FP32 soft kernels, FP64 correction/Hermite/"
                     "Kepler, OpenMP over groups and clusters), %.2f s" % (steps, n, dt)}
    out.update(host)
    return out


PEAK_FP64_TFLOPS = 78.6           # MI355X published dense FP64 vector peak
FLOP_HARD_PAIR = 75               # FP64 flops per hard_pair evaluation (op list in the kernel docs)


def hard_params():
    # perfect merging on: without it, bound pairs of 1e-8 bodies in these dense groups
    # shrink their block step towards dt_min = 2^-30 (up to 2^26 block steps per cluster)
    return dict(eta=0.01, eta_sun=0.01, eta_0=0.002, eta_sun0=0.002, dt_tree=2.0 ** -4, dt_min=2.0 ** -30,
                gamma=0.1, collision=1)


def bench_hard(args):
    """Batched hard part: 50,000 clusters (sizes log-uniform in [2, 256], Gaussian members
    with sigma = 2 Hill radii of m = 1e-10 around circular 1 AU orbits, masses log-uniform
    in [1e-12, 1e-8]) over one soft step 2^-4 with perfect merging; persistent launch,
    compared with one launch per cluster on a sample of the clusters; the oracle on a
    bounded sample of the same clusters as the CPU baseline."""
    from gplum_amd import gpu, ic
    n_cl = args.clusters
    p = gpu.params(**hard_params())
    pp, off = ic.clusters(n_cl, seed=2)
    ic.cluster_acc0(pp, off)
    n = len(pp)
    print("# %d clusters, %d particles" % (n_cl, n), file=sys.stderr, flush=True)
    warm = pp.copy()
    gpu.hard_integrate(p, warm, off, persistent=True, prepare=True)
    a = pp.copy()
    ed, nb, ms_e, npair = gpu.hard_integrate(p, a, off, persistent=True, prepare=True, counts=True)
    print("# persistent (eager launch): %.2f ms" % ms_e, file=sys.stderr, flush=True)
    # the same launch captured into a hipGraph and replayed (each replay restores the
    # uploaded state inside the graph, so every replay integrates the same step)
    g = pp.copy()
    ed_g, nb_g, ms_p, mism = gpu.hard_integrate_graph(p, g, off, replays=args.replays, prepare=True)
    graph_same = bool(mism == 0 and nb_g == nb and np.array_equal(g["pos"], a["pos"]) and
                      np.array_equal(g["vel"], a["vel"]))
    print("# persistent (graph replay, mean of %d): %.2f ms, identical %s" % (args.replays, ms_p, graph_same),
          file=sys.stderr, flush=True)
    # per-cluster launches run the clusters one after another: compare on a sample
    k = min(args.compare_clusters, n_cl)
    per_cluster = None
    if k > 0:
        s_off = off[: k + 1]
        sub = pp[: s_off[-1]]
        sa, sb = sub.copy(), sub.copy()
        _, nb_a, ms_sp = gpu.hard_integrate(p, sa, s_off, persistent=True, prepare=True)
        _, nb_b, ms_c = gpu.hard_integrate(p, sb, s_off, persistent=False, prepare=True)
        print("# sample of %d clusters: persistent %.2f ms, per-cluster launches %.2f ms" % (k, ms_sp, ms_c),
              file=sys.stderr, flush=True)
        same = bool(np.array_equal(sa["pos"], sb["pos"]) and np.array_equal(sa["vel"], sb["vel"]) and nb_a == nb_b)
        per_cluster = {"sample_clusters": k, "ms": round(ms_c, 3), "persistent_ms": round(ms_sp, 3),
                       "speedup_persistent": round(ms_c / max(ms_sp, 1e-9), 3), "identical_results": same}
    value = n / (ms_p * 1e-3)
    achieved = FLOP_HARD_PAIR * npair / (ms_p * 1e-3) / 1e12
    out = {
        "metric": "batched hard part: particle-steps/s (one soft step, block-timestep Hermite)",
        "value": round(value, 1), "unit": "particle-steps/s", "n_gpus": 1, "steps": args.replays, "warmup": 1,
        "ms_per_step": round(ms_p, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "batched hard: %d clusters, sizes log-uniform 2-256, sigma 2 r_H(1e-10) at 1 AU, "
                               "m log-uniform 1e-12..1e-8, dt_tree 2^-4, eta 0.01, perfect merging" % n_cl,
                   "clusters": n_cl, "particles": n, "seed": 2},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP64_TFLOPS, 4), "traffic": None,
                     "compute": "fp64 VALU, no MFMA", "kernel": "k_hermite size classes (persistent, graph replay)",
                     "flops_per_launch": FLOP_HARD_PAIR * npair, "avg_launch_ms": round(ms_p, 3)},
        "per_cluster_launches": per_cluster,
        "graph": {"replays": args.replays, "ms_per_replay": round(ms_p, 3), "eager_launch_ms": round(ms_e, 3),
                  "identical_to_eager_and_across_replays": graph_same},
        "block_steps": nb, "pair_evaluations": npair,
    }
    if not args.no_cpu:
        from oracle import oracle as O
        O.build()
        po = O.params(**hard_params())
        k = args.cpu_clusters
        sub_off = off[: k + 1]
        sub = pp[: sub_off[-1]].copy()
        O.cluster_prepare(po, sub, sub_off)
        t0 = time.perf_counter()
        O.hermite_clusters(po, sub, sub_off)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(int(sub_off[-1]) / dt, 1), "unit": "particle-steps/s", "cores": 1,
                               "kind": "port",
                               "sample": "first %d of the same clusters (%d particles), oracle C restatement "
                                         "of timeIntegrate_multi, one thread, %.2f s" % (k, int(sub_off[-1]), dt)}
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


def bench_copy(args):
    """A plain device-to-device copy of known size (torch's copy kernel): the reference
    point for the FETCH_SIZE / WRITE_SIZE counters (a copy of B bytes reads B and writes
    B bytes of HBM)."""
    import torch
    nbytes = args.copy_mib << 20
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(max(1, args.warmup)):
        b.copy_(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    print(json.dumps({"metric": "device copy bandwidth (read + write bytes / s)", "value": round(2 * nbytes / dt / 1e9, 1),
                      "unit": "GB/s", "bytes_copied": nbytes, "steps": args.steps, "ms_per_copy": round(dt * 1e3, 4)}))


def neighbour_pairs(s):
    """Sorted (id_i << 32 | id_j) of a system's neighbour graph (order-free comparison)."""
    pp = s.particles()
    off, idx = s.neighbors()
    ids = pp["id"].astype(np.int64)
    rows = np.repeat(np.arange(len(pp)), np.diff(off))
    return np.sort((ids[rows] << 32) | ids[idx])


def bench_large(args):
    """Configs 4 and 5: the full step on a larger disk.  config4: N = 4M over 0.5-3 AU,
    forecaster on; the neighbour graph after the timed steps is recomputed with the
    forecaster off and must be identical.  config5: N = 16M over 0.3-5 AU plus 10
    embryos (m = 1e-6, circular, evenly spaced in a); energy drift of the FP32 path and
    of the FP64-soft path over the same steps from the same state."""
    from gplum_amd import gpu, ic
    big = args.workload == "config5"
    n = args.n if args.n != 1_000_000 else (16_000_000 if big else 4_000_000)
    a_in, a_out = (0.3, 5.0) if big else (0.5, 3.0)
    p = gpu.params(**params_dict())
    t0 = time.perf_counter()
    pp = ic.disk(n, 1e-12, a_in, a_out, seed=args.seed)
    if big:
        pp = ic.add_embryos(pp, 10, 1e-6, a_in, a_out)
    n_tot = len(pp)
    print("# %s: initial conditions for N=%d in %.1f s" % (args.workload, n_tot, time.perf_counter() - t0),
          file=sys.stderr, flush=True)
    s = gpu.System(p, pp)
    del pp
    s.set_forecast("network")
    if args.giant_min:
        s.set_tuning(hard_giant_min=args.giant_min)
    t0 = time.perf_counter()
    s.initialize()
    print("# initialised in %.1f s" % (time.perf_counter() - t0), file=sys.stderr, flush=True)
    tuning = None if args.no_tune else s.autotune(trials=1)
    print("# %s: N=%d, init + tune %.1f s, tuning %s" % (args.workload, n_tot, time.perf_counter() - t0, tuning),
          file=sys.stderr, flush=True)
    for k in range(args.warmup):
        s.step(1)
    start_state = s.particles()
    start_time = s.stats()["time_sys"]
    s.reset_stats(timing=True)
    e0 = s.energy()
    t0 = time.perf_counter()
    for k in range(args.steps):
        s.step(1)
        print("# step %d done" % k, file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t0
    e1 = s.energy()
    st = s.stats()
    drift32 = (e1[0] - e0[0] - (e1[5] - e0[5])) / e0[0]
    calls = max(1, st["fc_calls"])
    hit = 1.0 - st["fc_ovf_groups"] / float(max(1, st["n_groups"]) * calls)
    out = {
        "metric": "particle-steps/s (full P3T step, %s)" % args.workload, "value": round(n_tot * args.steps / elapsed, 1),
        "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 soft force, f64 hard/correction", "data": "synthetic",
        "config": {"workload": "%s: disk N=%d, m=1e-12, a=%g-%g AU%s, theta=0.5 quadrupole, dt_tree=2^-4, "
                               "collisions on" % (args.workload, n, a_in, a_out,
                                                   " + 10 embryos m=1e-6" if big else ""),
                   "n": n_tot, "seed": args.seed},
        "autotune": tuning,
        "breakdown_ms_per_step": {k: round(st[k] / args.steps, 3) for k in ("force_ms", "tree_ms", "finalize_ms",
                                                                             "hard_ms")},
        "hard": {"clusters": st["n_clusters"], "max_cluster": st["max_cluster"], "block_steps": st["block_steps"],
                 "max_block_steps_one_cluster": st["max_block_steps"], "members_of_that_cluster": st["max_block_cluster"], "collisions": st["n_collisions"]},
        "forecast": {"calls": st["fc_calls"], "groups_per_call": st["n_groups"],
                     "overflowed_groups": st["fc_ovf_groups"], "hit_rate": round(hit, 5),
                     "ms_per_step": round(st["fc_ms"] / args.steps, 3)},
        "energy_error_fp32": drift32,
    }
    if not big:
        # the exact search stays the judge: the same state with the forecaster off
        s.soft_force(initial=False)
        a = neighbour_pairs(s)
        s_off = gpu.System(p, s.particles(), time_sys=st["time_sys"])
        s_off.set_tuning(**{k: tuning[k] for k in ("n_group_limit", "soft_tile", "hard_wg_per_cu",
                                                    "hard_lds_max")} if tuning else {})
        s_off.soft_force(initial=False)
        b = neighbour_pairs(s_off)
        out["neighbours_identical_forecast_on_off"] = bool(np.array_equal(a, b))
        out["neighbour_pairs"] = int(len(a))
        s_off.close()
    else:
        # the same steps with the FP64 soft force (persistent hard part: identical
        # results to per-cluster launches) from the same state
        s64 = gpu.System(p, start_state, time_sys=start_time)
        s64.set_variant(fp64_soft=True, per_cluster_hard=False)
        s64.initialize()
        f0 = s64.energy()
        for k in range(args.steps):
            s64.step(1)
            print("# FP64-soft step %d done" % k, file=sys.stderr, flush=True)
        f1 = s64.energy()
        out["energy_error_fp64_soft"] = (f1[0] - f0[0] - (f1[5] - f0[5])) / f0[0]
        s64.close()
    s.close()
    out["cpu_baseline"] = None
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--mode-steps", type=int, default=6, help="steps per forecaster mode compared (0: skip)")
    ap.add_argument("--workload", choices=["step", "hard", "copy", "config4", "config5"], default="step",
                    help="step: full P3T step at N (default); hard: batched hard part alone; "
                         "copy: plain device copy of --copy-mib MiB (calibrates the HBM byte counters); "
                         "config4: N=4M disk over 0.5-3 AU; config5: N=16M over 0.3-5 AU + 10 embryos")
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--clusters", type=int, default=5000)
    ap.add_argument("--cpu-clusters", type=int, default=20)
    ap.add_argument("--compare-clusters", type=int, default=100)
    ap.add_argument("--replays", type=int, default=3, help="graph replays (hard workload)")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (default launch shapes)")
    ap.add_argument("--giant-min", type=int, default=0,
                    help="clusters above this size on the whole-GPU path (config4/5; 0: library default)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "hard":
        bench_hard(args)
        return
    if args.workload == "copy":
        bench_copy(args)
        return
    if args.workload in ("config4", "config5"):
        bench_large(args)
        return

    from gplum_amd import gpu
    name, n_cu, clk = gpu.device_info()
    p = gpu.params(**params_dict())
    pp = make_disk(args.n, args.seed)
    s = gpu.System(p, pp)
    s.set_forecast("network")          # learned candidate sizing, trained online
    t0 = time.perf_counter()
    s.initialize()
    t_init = time.perf_counter() - t0
    print("# device %s, %d CUs; init %.2f s" % (name, n_cu, t_init), file=sys.stderr, flush=True)
    tuning = None
    if not args.no_tune:
        # the project's autotuner picks i-group size, soft j-tile, Hermite grid and LDS
        # binning threshold on this state (untimed, like initialisation)
        tuning = s.autotune(trials=2)
        print("# autotune: %s" % tuning, file=sys.stderr, flush=True)
    for k in range(args.warmup):
        s.step(1)
        print("# warmup step %d done" % k, file=sys.stderr, flush=True)
    warm_state = s.particles()
    warm_time = s.stats()["time_sys"]
    warm_fc = s.get_forecast()      # the network as trained over the warm-up steps
    s.reset_stats(timing=True)
    e0 = s.energy()
    t0 = time.perf_counter()
    for k in range(args.steps):
        s.step(1)
    t1 = time.perf_counter()
    st = s.stats()
    e1 = s.energy()
    elapsed = t1 - t0
    ms_per_step = 1e3 * elapsed / args.steps
    value = args.n * args.steps / elapsed

    # FP32 soft force (EP-EP + EP-SP), HIP-event timed on the system stream
    ncalls = max(1, st["n_force_calls"])
    force_ms = st["force_ms"] / ncalls
    pairs_ep = st["n_epep_near"] + st["n_epep_far"]
    pairs_sp = st["n_epsp"]
    flops = FLOP_EPEP * pairs_ep + FLOP_EPSP_QUAD * pairs_sp + FLOP_SPJ_QUAD * st["n_sp_entries"]
    achieved = flops / (force_ms * 1e-3) / 1e12
    # compute-bound kernel; "mfma" names the schema's compute roof: the arithmetic is FP32
    # VALU (pairwise gravity is not a contraction), priced against the FP32 vector peak
    # counters of the launch shape in effect (the autotuner's i-group size, else the default)
    ng_lim = tuning["n_group_limit"] if tuning else params_dict()["n_group_limit"]
    traffic = (measured_traffic("k_soft_f32_n1m_traffic.md", "n_group_limit=%d" % ng_lim)
               if args.n == 1_000_000 else None)
    roofline_soft = {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": traffic,
                     "compute": "fp32 VALU, no MFMA",
                     "kernel": "k_soft_f32 (EP-EP near/far + EP-SP quadrupole, one launch per force)",
                     "flops_per_launch": flops, "avg_launch_ms": round(force_ms, 4)}
    # the kernel with the most time in the step: the batched FP64 Hermite (one launch set
    # per hard part, HIP events on its stream); algorithmic flops = 75 per pair-force
    # evaluation; it is bound by the block-step chain of the longest cluster (one wave),
    # not by the chip's FP64 rate
    n_hl = max(1, st["n_hard_launches"])
    hk_ms = st["hard_kernel_ms"] / n_hl
    hk_flops = FLOP_HARD_PAIR * st["hard_pairs"] / n_hl
    hk_ach = hk_flops / (hk_ms * 1e-3) / 1e12
    roofline = {"bound": "mfma", "achieved": round(hk_ach, 4), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": round(hk_ach / PEAK_FP64_TFLOPS, 6),
                "traffic": measured_traffic("k_hermite_n1m_traffic.md") if args.n == 1_000_000 else None,
                "compute": "fp64 VALU, no MFMA; latency-bound on the longest cluster's block-step chain",
                "kernel": "k_hermite size classes (batched FP64 Hermite, one launch set per step)",
                "flops_per_launch": round(hk_flops), "avg_launch_ms": round(hk_ms, 4),
                "us_per_block_step_longest_cluster": round(1e3 * hk_ms / max(1, st["max_block_steps"]), 3)}

    # FP64 twin on the same state: interactions/s and the FP32 error
    q32 = s.particles()
    s64 = gpu.System(p, q32, time_sys=st["time_sys"])
    s64.reset_stats(timing=True)
    s64.soft_force(initial=False, use_f64=True)
    st64 = s64.stats()
    q64 = s64.particles()
    s32 = gpu.System(p, q32, time_sys=st["time_sys"])
    s32.reset_stats(timing=True)
    s32.soft_force(initial=False, use_f64=False)
    st32 = s32.stats()
    q32b = s32.particles()
    a = q32b[np.argsort(q32b["id"])]["acc"]
    b = q64[np.argsort(q64["id"])]["acc"]
    rel = np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)
    acc_rms = float(np.sqrt(np.mean(rel ** 2)))
    pa = q32b[np.argsort(q32b["id"])]["phi"]
    pb = q64[np.argsort(q64["id"])]["phi"]
    phi_max = float(np.max(np.abs(pa - pb) / np.abs(pb)))
    ips32 = (pairs_ep + pairs_sp) / (force_ms * 1e-3)
    ips64 = (st64["n_epep_near"] + st64["n_epep_far"] + st64["n_epsp"]) / (st64["force_ms"] * 1e-3)
    s64.close()
    s32.close()

    # plain comparator: FP64 soft force, one launch per cluster, no forecaster, from the
    # same warmed-up state
    plain = None
    if args.plain_steps > 0:
        sp_ = gpu.System(p, warm_state, time_sys=warm_time)
        sp_.set_variant(fp64_soft=True, per_cluster_hard=True)
        sp_.initialize()
        sp_.step(1)
        sp_.reset_stats(timing=True)
        tp0 = time.perf_counter()
        sp_.step(args.plain_steps)
        tp1 = time.perf_counter()
        vp = args.n * args.plain_steps / (tp1 - tp0)
        plain = {"value": round(vp, 1), "unit": "particle-steps/s", "steps": args.plain_steps,
                 "ms_per_step": round(1e3 * (tp1 - tp0) / args.plain_steps, 3),
                 "speedup": round(value / vp, 3),
                 "variant": "FP64 soft force, per-cluster Hermite launches, global candidate buffer"}
        sp_.close()
        print("# plain variant done", file=sys.stderr, flush=True)

    # the same steps with the candidate buffers sized by the network (the timed run), by
    # a uniform per-particle rate, and by the global buffer without any forecast, each
    # from the warmed-up state (one untimed step, then the timed ones)
    # (results are bit-identical across the modes, so each runs exactly the same work)
    fc_modes = None
    if args.mode_steps > 0:
        fc_modes = {}
        for mode in ("network", "uniform", "off"):
            sm = gpu.System(p, warm_state, time_sys=warm_time)
            if tuning:
                sm.set_tuning(**tuning)
            sm.set_forecast(mode)
            if mode == "network" and warm_fc[3] > 0:
                sm.put_forecast(*warm_fc[:3])
            sm.initialize()
            sm.step(1)
            sm.reset_stats(timing=True)
            tm0 = time.perf_counter()
            sm.step(args.mode_steps)
            fc_modes[mode] = round(1e3 * (time.perf_counter() - tm0) / args.mode_steps, 3)
            stm = sm.stats()
            fc_modes[mode + "_breakdown"] = {k: round(stm[k] / args.mode_steps, 3) for k in (
                "force_ms", "tree_ms", "finalize_ms", "hard_ms", "fc_ms")}
            sm.close()
        fc_modes["unit"] = "ms per step"
        fc_modes["steps"] = args.mode_steps
        print("# forecaster modes: %s" % fc_modes, file=sys.stderr, flush=True)

    de = (e1[0] - e0[0] - (e1[5] - e0[5])) / e0[0]
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32 soft force, f64 hard/correction",
        "data": "synthetic",
        "config": {"workload": "full P3T step: planetesimal disk N=%d, m=1e-12, a=0.5-2 AU, "
                               "theta=0.5 quadrupole, n_group<=256, dt_tree=2^-4, collisions on" % args.n,
                   "n": args.n, "seed": args.seed, "device": name},
        "roofline": roofline,
        "roofline_soft_f32": roofline_soft,
        "autotune": tuning,
        "soft_force": {"interactions_per_s_fp32": round(ips32, 1), "interactions_per_s_fp64": round(ips64, 1),
                       "fp32_over_fp64": round(ips32 / ips64, 3), "acc_rms_rel_err_fp32_vs_fp64": acc_rms,
                       "phi_max_rel_err_fp32_vs_fp64": phi_max,
                       "epep_pairs": pairs_ep, "epsp_pairs": pairs_sp, "groups": st["n_groups"]},
        "breakdown_ms_per_step": {k: round(st[k] / args.steps, 3) for k in ("force_ms", "tree_ms",
                                                                             "finalize_ms", "hard_ms")},
        "hard": {"clusters": st["n_clusters"], "cluster_particles": st["n_cluster_particles"],
                 "max_cluster": st["max_cluster"], "block_steps": st["block_steps"],
                 "max_block_steps_one_cluster": st["max_block_steps"], "members_of_that_cluster": st["max_block_cluster"], "collisions": st["n_collisions"]},
        "energy_error_over_timed_steps": de,
        "plain_fp64_per_cluster": plain,
        "forecast_modes": fc_modes,
        "forecast": {"calls": st["fc_calls"], "slots": st["fc_slots"], "overflowed_groups": st["fc_ovf_groups"],
                     "redone_candidates": st["fc_ovf_candidates"], "loss": st["fc_loss"],
                     "ms_per_step": round(st["fc_ms"] / args.steps, 3)},
    }
    # list building (neighbour finalize): HBM-bound; bytes from the k_finalize op list,
    # 116 B per particle + 105 B per FP32 candidate, over the finalize stage's event time
    lb_bytes = 116 * args.n + 105 * st["n_candidates"]
    fin_ms = st["finalize_ms"] / max(1, st["n_force_calls"])
    out["list_build"] = {"bytes_per_call": lb_bytes, "finalize_ms_per_call": round(fin_ms, 4),
                         "achieved_GBps": round(lb_bytes / (fin_ms * 1e-3) / 1e9, 1), "peak_GBps": 8000.0,
                         "note": "finalize stage (candidate sort, k_finalize, CSR build) timed as a whole"}
    s.close()
    if not args.no_cpu:
        threads = args.cpu_threads or cpu_threads_default()
        host = cpu_setup(threads)
        out["soft_force"]["interactions_per_s_cpu_fp32"] = cpu_soft_force(args.n, args.seed, threads, warm_state)
        out["cpu_baseline"] = cpu_baseline(args.n, args.seed, threads, host)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
