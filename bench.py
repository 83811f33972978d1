#!/usr/bin/env python3
"""Throughput benchmark of the simulation core on one MI355X.

Workload (the reference's headline, README.rst:40-48): rWW, N=100 nodes,
32,768 simulations on a 32x32x32 grid of G in [0.001, 10], w_p in [0, 2],
J_N in [0.001, 0.5] (itertools.product order, as GridOptimizer builds it,
optimize.py:1220-1223), 60 s at dt 0.1 ms / BW dt 1 ms, TR 1 s, FCD window 10
TRs step 2, 30 s burn-in, analytical FIC off.  SC: seeded lognormal(0,1),
symmetrised, zero diagonal, max 1.  Empirical target: seeded AR(1) BOLD
100x60 (phi 0.7, Cholesky-correlated innovations from a Wishart covariance).

The grid is processed in batches of --batch simulations (default 8,192, a
quarter of the grid); one step = one batch: integrate its simulations (BOLD,
states), FC + FCD, and the FC-Pearson and FCD-KS scores against the empirical
target, with inputs and outputs resident in HBM (cubnm_run_device).  Step i
runs batch i mod (grid / batch), so every run of a multiple of 4 steps covers
the whole grid an equal number of times; every batch is a full 60 s
integration of its simulations.  Prints ONE JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "sim-node-seconds/wall-s (rWW N=100, 60 s); grad-fit vs CMA-ES time-to-loss"
A100_PUBLISHED = 585000.0        # sim-node-s/s: the reference's published A100 run of this grid, 32,768 x 100 x 60 s in 5.6 min
FP64_PEAK_TFLOPS = 78.6          # MI355X FP64 (vector and matrix)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def grid_params(n_sims, nodes):
    g = np.linspace(0.001, 10.0, 32)
    wp = np.linspace(0.0, 2.0, 32)
    jn = np.linspace(0.001, 0.5, 32)
    GG, WW, JJ = np.meshgrid(g, wp, jn, indexing="ij")
    G, W, J = GG.ravel(), WW.ravel(), JJ.ravel()
    idx = np.arange(n_sims) % G.size
    return G[idx], W[idx], J[idx]


def empirical_bold(nodes, vols, seed=0, phi=0.7):
    rng = np.random.default_rng(seed)
    df = nodes + 10
    A = rng.standard_normal((nodes, df))
    L = np.linalg.cholesky(A @ A.T / df)
    x = np.empty((nodes, vols))
    x[:, 0] = L @ rng.standard_normal(nodes)
    for t in range(1, vols):
        x[:, t] = phi * x[:, t - 1] + np.sqrt(1 - phi ** 2) * (L @ rng.standard_normal(nodes))
    return x


def ops_per_sim_node_second(nodes, dt_ms, bw_dt_ms, TR_s):
    """Algorithmic FP64 ops (exp, pow, div count as one): 2N+42 per node per
    model step (N coupling MACs + rWW step), 24 per BW step, 10 per BOLD."""
    return (1000.0 / dt_ms) * (2 * nodes + 42) + (1000.0 / bw_dt_ms) * 24 + 10.0 / TR_s


def host_cpu():
    """Threads the host lets this process run, and the CPU model.  The thread
    count is the affinity mask, capped by the cgroup CPU quota when one is set
    (a quota below the mask means more threads would only be throttled)."""
    n = len(os.sched_getaffinity(0))
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            n = max(1, min(n, int(-(-int(q) // int(period)))))
    except (OSError, ValueError):
        pass
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return n, model


def cpu_baseline(n_cores, nodes, duration_s, G, W, J, sc, cfg, per_core=2, repeats=3):
    """The oracle (FP64 C++ restatement of the reference CPU engine, same
    flags, OpenMP over simulations) on a bounded sample of the grid:
    per_core full-length simulations per thread, run phase only (noise
    generation excluded), median of ``repeats`` runs."""
    import cases
    from oracle import oracle as O
    O.lib()
    n = n_cores * per_core
    pick = np.linspace(0, G.size - 1, n).round().astype(int)
    gp, rp = cases.params("rWW", n, nodes, G=G[pick], w_p=W[pick][:, None], J_N=J[pick][:, None],
                          do_fic=False, wIE=1.0)
    args = cases.sim_args("rWW", sc, gp, rp, cfg, n, duration_s=duration_s)
    rts = [cases.run_oracle(args, n_threads=n_cores)["run_time"] for _ in range(repeats)]
    rt = float(np.median(rts))
    _, model = host_cpu()
    return {
        "value": round(n * nodes * duration_s / rt, 1),
        "unit": "sim-node-s/s",
        "cores": n_cores,
        "kind": "port",
        "cpu_model": model,
        "sample": f"{n} of the workload's simulations (evenly spaced), full {duration_s:g} s each, "
                  f"{per_core} per OpenMP thread on {n_cores} threads; run phase only, noise generation "
                  f"excluded; median of {repeats} runs ({', '.join(f'{r:.2f}' for r in rts)} s)",
    }


def read_traffic(path):
    """Bytes per integrator launch measured with rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, calibrated against a copy kernel), if committed."""
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("bytes_per_launch="):
            return float(line.split("=", 1)[1])
    return None


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--sims", type=int, default=32768)
    ap.add_argument("--batch", type=int, default=8192,
                    help="simulations per step (a divisor of --sims; the grid is run batch by batch)")
    ap.add_argument("--nodes", type=int, default=100)
    ap.add_argument("--duration", type=float, default=60.0)
    ap.add_argument("--cpu-cores", type=int, default=0,
                    help="threads for the CPU baseline (0: every core the process may use)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tpw", type=int, default=0, help="batched integrator tiles per wave (0: auto)")
    ap.add_argument("--autotune", type=int, default=0, choices=[-1, 0, 1],
                    help="launch-shape autotuner: -1 never, 0 large launches (default), 1 always")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="run one device-to-device copy of this many MiB first (PMC byte-counter calibration)")
    ap.add_argument("--workload", choices=["c2", "c3", "c4", "c5"], default="c2",
                    help="c2: the grid throughput (default); c3: heterogeneous batch; "
                         "c4: gradient fit vs CMA-ES time-to-loss; c5: N=400 FC/FCD with split-BF16 Gram products")
    ap.add_argument("--c3-sims", type=int, default=4096, help="simulations of the heterogeneous batch (c3)")
    ap.add_argument("--c4-gens", type=int, default=60, help="CMA-ES generations (c4)")
    ap.add_argument("--c4-steps", type=int, default=40, help="maximum Adam steps (c4)")
    ap.add_argument("--c4-duration", type=float, default=120.0, help="simulated seconds (c4)")
    ap.add_argument("--c4-lr", type=float, default=0.12,
                    help="Adam learning rate in the unit cube (c4; 0.12 was the best of 0.02, 0.05, 0.1, 0.12, "
                         "0.15, 0.2, 0.3)")
    ap.add_argument("--c5-sims", type=int, default=8192, help="simulations (c5)")
    ap.add_argument("--c5-batch", type=int, default=4096, help="simulations per run_simulations call (c5)")
    ap.add_argument("--c5-check-sims", type=int, default=4,
                    help="simulations whose FC/FCD are recomputed in FP64 on the host (c5)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box in this setup has one MI355X, and the simulations shard "
              "across GPUs with no exchange between them; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    import cases
    from cubnm_amd import capi, utils

    if args.workload == "c4":
        return run_c4(args)
    if args.workload == "c3":
        return run_c3(args)
    if args.workload == "c5":
        return run_c5(args)

    dev = torch.device("cuda:0")
    S, N, dur, B = args.sims, args.nodes, args.duration, args.batch
    if B <= 0 or S % B != 0:
        raise SystemExit(f"--batch {B} must divide --sims {S}")
    n_batches = S // B
    TR_s, dt, bw_dt = 1.0, 0.1, 1.0
    sc = cases.synthetic_sc(N, seed=0)
    G, W, J = grid_params(S, N)
    cfg = cases.model_config("rWW", TR=TR_s, window_size=10, window_step=2, bold_remove_s=30, do_fic=False)
    emp = empirical_bold(N, int(dur / TR_s), seed=0)
    emp_fc = utils.calculate_fc(emp)
    emp_fcd = np.sort(utils.calculate_fcd(emp, 10, 2))

    capi.set_tuning("batched_tpw", args.tpw)
    capi.set_tuning("autotune", args.autotune)
    capi.force_integrator(-1)
    d = capi.prepare("rWW", cfg, B, N, int(dur * 1000), int(TR_s * 1000), int(TR_s * 1000), 0, dt, bw_dt)
    t = lambda a, dtype=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=dev)
    # every batch's parameters resident in HBM ([param][sim] and
    # [param][sim][node] per batch), outputs for the whole grid
    d_gp, d_rp = [], []
    for b in range(n_batches):
        sl = slice(b * B, (b + 1) * B)
        gp, rp = cases.params("rWW", B, N, G=G[sl], w_p=W[sl][:, None], J_N=J[sl][:, None], do_fic=False, wIE=1.0)
        d_gp.append(t(gp))
        d_rp.append(t(rp))
    d_sc, d_sci = t(sc.reshape(1, -1)), t(np.zeros(B), torch.int32)
    d_dist, d_v = t(np.zeros(N * N)), t(np.zeros(B))
    d_bold = torch.empty((S, d.bold_len * N), dtype=torch.float64, device=dev)
    d_fc = torch.empty((S, d.n_pairs), dtype=torch.float64, device=dev)
    d_fcd = torch.empty((S, d.n_window_pairs), dtype=torch.float64, device=dev)
    d_st = torch.empty((n_batches, d.n_state_vars, B, N), dtype=torch.float64, device=dev)
    d_gb = torch.zeros((n_batches, d.n_global_out_bool, B), dtype=torch.uint8, device=dev)
    d_gi = torch.zeros((n_batches, d.n_global_out_int, B), dtype=torch.int32, device=dev)
    d_emp_fc, d_emp_fcd = t(emp_fc), t(emp_fcd)
    d_corr = torch.full((S,), float("nan"), dtype=torch.float64, device=dev)
    d_ks = torch.full((S,), float("nan"), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    counter = [0]

    def step():
        b = counter[0] % n_batches
        counter[0] += 1
        o = b * B
        capi.run_device(d_sc.data_ptr(), d_sci.data_ptr(), d_dist.data_ptr(), d_gp[b].data_ptr(),
                        d_rp[b].data_ptr(), d_v.data_ptr(), d_bold[o].data_ptr(), d_fc[o].data_ptr(),
                        d_fcd[o].data_ptr(), d_st[b].data_ptr(), d_gb[b].data_ptr(), d_gi[b].data_ptr())
        capi.gof_fc_corr_device(d_fc[o].data_ptr(), B, d.n_pairs, d_emp_fc.data_ptr(), d_corr[o].data_ptr())
        capi.gof_fcd_ks_device(d_fcd[o].data_ptr(), B, d.n_window_pairs, d_emp_fcd.data_ptr(), emp_fcd.size,
                               d_ks[o].data_ptr())
        return capi.last_timings()

    if args.copy_calibration_mib:
        n = args.copy_calibration_mib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        torch.mul(src, 1.0, out=dst)  # one elementwise kernel: reads and writes n*8 bytes
        torch.cuda.synchronize()
        log(f"calibration copy: {args.copy_calibration_mib} MiB read + {args.copy_calibration_mib} MiB written")
        del src, dst
    for i in range(args.warmup):
        tm = step()
        log(f"warmup {i}: {tm}")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    integ_ms, fc_ms = [], []
    for i in range(args.steps):
        tm = step()
        integ_ms.append(tm["integrate_ms"])
        fc_ms.append(tm["fc_fcd_ms"])
        log(f"step {i}: {tm}")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1000.0 / args.steps
    value = B * N * dur / (ms_per_step / 1000.0)

    shape = capi.last_launch_shape()
    log(f"launch shape: {shape}")
    corr, ks = d_corr.cpu().numpy(), d_ks.cpu().numpy()
    scored = min(S, counter[0] * B)
    log(f"scores of the first {scored} simulations: fc_corr mean {np.nanmean(corr[:scored]):.4f}, "
        f"fcd_ks mean {np.nanmean(ks[:scored]):.4f}, "
        f"nan {int(np.isnan(corr[:scored]).sum())}/{int(np.isnan(ks[:scored]).sum())}")
    ops = ops_per_sim_node_second(N, dt, bw_dt, TR_s) * B * N * dur
    k_ms = float(np.mean(integ_ms))
    achieved = ops / (k_ms / 1000.0) / 1e12
    traffic = read_traffic(os.path.join(ROOT, "profiles", "integrate_batched_rww_n100_c2_traffic.txt"))
    roofline = {"bound": "mfma", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "traffic": traffic,
                "kernel": ("integrate_batched_kernel" if shape["integrator"] == 1 else "integrate_general_kernel"),
                "kernel_ms": round(k_ms, 2),
                "fc_fcd_ms": round(float(np.mean(fc_ms)), 2)}

    cpu = None
    if not args.no_cpu_baseline:
        n_cores = args.cpu_cores or host_cpu()[0]
        cpu = cpu_baseline(n_cores, N, dur, G, W, J, sc, cfg)
        log(f"cpu baseline: {cpu}")

    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "sim-node-s/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True, "scaling": "weak",
        "vs_baseline": round(value / A100_PUBLISHED, 3),
        "dtype": "f64", "data": "synthetic (seeded lognormal SC, seeded AR(1) empirical BOLD, host-generated noise)",
        "config": {"workload": f"rWW grid {S} sims x {N} nodes x {dur:g} s (dt 0.1 ms, BW dt 1 ms, TR 1 s, "
                               f"FCD window 10 step 2, FIC off), FC + FCD + FC-Pearson + FCD-KS on GPU, "
                               f"{B} simulations per step ({n_batches} steps per pass over the grid)",
                   "sims": S, "sims_per_step": B, "nodes": N, "duration_s": dur,
                   "integrator": ["per-simulation", "batched MFMA FP64", "many-node"][shape["integrator"]],
                   "launch_shape": {"tiles_per_wave": shape["tiles_per_wave"], "sc_in_lds": shape["sc_in_lds"],
                                    "ksplit": shape["ksplit"],
                                    "chosen_by": ["rule", "autotuner trials", "autotuner cache"][shape["chosen_by"]]}},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def run_c3(args):
    """Heterogeneous batch through one call of the grouped entry point:
    --c3-sims simulations (4096 in the full configuration), node count
    uniform in {68, 100, 200, 400}, duration uniform in {60, 120, 300, 600} s,
    models 70% rWW / 15% JR / 15% Kuramoto, each simulation with its own
    seeded lognormal SC and uniform G in [0.001, 5] (other parameters at their
    defaults, rWW FIC off), TR 0.72 s, FCD windows 30 TRs step 5, 30 s burn-in.
    Simulations are bucketed by (model, nodes, duration), one request per
    bucket; value = total sim-node-seconds / wall seconds of the call."""
    import torch  # noqa: F401
    import cases
    from cubnm_amd import _core
    rng = np.random.default_rng(3)
    S = args.c3_sims
    Ns = rng.choice([68, 100, 200, 400], S)
    durs = rng.choice([60, 120, 300, 600], S)
    models = rng.choice(["rWW", "JR", "Kuramoto"], S, p=[0.7, 0.15, 0.15])
    reqs, work, sc_bytes, coupling_flops = [], 0.0, 0.0, 0.0
    for m in ["rWW", "JR", "Kuramoto"]:
        for n in [68, 100, 200, 400]:
            for d in [60, 120, 300, 600]:
                idx = np.where((models == m) & (Ns == n) & (durs == d))[0]
                if idx.size == 0:
                    continue
                k = idx.size
                scs = np.stack([cases.synthetic_sc(n, seed=int(1000 + i)) for i in idx])
                gp, rp = cases.params(m, k, n, G=rng.uniform(0.001, 5.0, k), do_fic=False)
                cfg = cases.model_config(m, TR=0.72, bold_remove_s=30, do_fic=False)
                cfg.update({"window_size": "30", "window_step": "5", "verbose": "1"})
                reqs.append(cases.sim_args(m, scs, gp, rp, cfg, k, duration_s=d, TR_s=0.72,
                                           sc_indices=np.arange(k, dtype=np.intc)))
                work += k * n * d
                # coupling FP64 ops: N multiply-adds per node and model step
                coupling_flops += 2.0 * n * n * (d * 10000) * k
                # SC bytes that must stream every model step: an SC of N=200
                # (320 KB) or N=400 (1.28 MB) exceeds a CU's 160 KB of LDS, so
                # no design keeps it on chip between steps; N=68 and N=100
                # (37 and 80 KB) could stay in LDS and are not counted
                if n >= 200:
                    sc_bytes += 8.0 * n * n * (d * 10000) * k
    log(f"c3: {S} simulations in {len(reqs)} requests, {work:.3e} sim-node-s")
    _core.core.run_groups(reqs[:1])  # session and noise set-up for the first bucket (untimed)
    t0 = time.time()
    outs = _core.core.run_groups(reqs)
    wall = time.time() - t0
    assert len(outs) == len(reqs)
    setup = sum(o["init_time"] for o in outs)  # host noise generation and session set-up
    int_ms = max(o["integrate_ms"] for o in outs)  # the one per-simulation launch
    achieved = sc_bytes / (int_ms / 1000) / 1e9
    info = _core.core.run_groups_info()
    log(f"c3: launch shape {info}")
    line = {
        "metric": METRIC, "value": round(work / wall, 1), "unit": "sim-node-s/s", "n_gpus": 1, "steps": 1,
        "warmup": 0, "ms_per_step": round(1000 * wall, 1), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded lognormal SC per simulation, host-generated noise)",
        "config": {"workload": f"heterogeneous batch: {S} sims, N in {{68,100,200,400}}, 60-600 s, rWW/JR/Kuramoto "
                               f"70/15/15, own SC each, TR 0.72 s, FCD 30/5 TRs, {len(reqs)} requests in one call",
                   "sim_node_seconds": work, "host_setup_s": round(setup, 2),
                   "value_excluding_host_setup": round(work / (wall - setup), 1),
                   "launch_shape": {"ksplit": int(info["ksplit"]), "pair": int(info["pair"]),
                                    "chosen_by": ["rule", "timed trials", "on-disk cache"][int(info["tuned"])],
                                    "trial_ms": round(info["trial_ms"], 1), "shapes_timed": int(info["shapes_timed"]),
                                    "graph": ["none", "captured", "replayed"][int(info["graph"])]},
                   "coupling_fp64_tflops": round(coupling_flops / (int_ms / 1000) / 1e12, 2)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": 8000.0, "unit": "GB/s",
                     "frac": round(achieved / 8000.0, 4), "traffic": None,
                     "kernel": "integrate_general_multi_kernel", "kernel_ms": round(int_ms, 1),
                     "algorithmic_bytes": sc_bytes,
                     "bytes_counted": "SC of the N=200 and N=400 simulations, read once per model step: each "
                                      "simulation has its own SC of 320 KB or 1.28 MB, larger than a CU's 160 KB "
                                      "of LDS, and all of them (1.6 GB) exceed the 256 MB last-level cache, so "
                                      "every step streams it from HBM; N=68 and N=100 SCs stay in LDS and count "
                                      "zero"},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def run_c5(args):
    """N=400 rWW (one shared seeded lognormal SC, G uniform in [0.5, 3],
    other parameters at their defaults, FIC off), 300 s at dt 0.1 ms, TR 0.72 s
    (416 volumes), FC and FCD over windows of 30 TRs with step 1 and no burn-in
    (the reference's window placement, utils.cpp:21-72: windows centred on a
    volume span window_size + 1 = 31 volumes, 386 of them; 30-volume windows
    are not expressible in that rule), FC/FCD as split-BF16 Gram products
    (fc_mode 2), integration FP64 on the batched MFMA integrator for
    257..512 nodes (SC fragments streamed from L2).  --c5-sims simulations
    (8192 in the full configuration) in run_simulations calls of --c5-batch
    (4096: 256 workgroups of 16 simulations, one per CU).  value =
    sim-node-seconds per wall second of the calls.  For the first
    --c5-check-sims simulations the FC and FCD are recomputed on the host in
    FP64 (NumPy Pearson correlations of the returned BOLD, cubnm_amd.utils)
    and the largest differences are reported."""
    import torch  # noqa: F401
    import cases
    from cubnm_amd import _core, utils
    N, dur, TR = 400, 300.0, 0.72
    sc = cases.synthetic_sc(N, seed=0)
    S, B = args.c5_sims, min(args.c5_batch, args.c5_sims)
    G = np.random.default_rng(5).uniform(0.5, 3.0, S)

    def args_for(g, d):
        n = g.size
        gp, rp = cases.params("rWW", n, N, G=g, do_fic=False, wIE=1.0)
        cfg = cases.model_config("rWW", TR=TR, bold_remove_s=0, do_fic=False)
        cfg.update({"window_size": "30", "window_step": "1", "verbose": "0"})
        return cases.sim_args("rWW", sc, gp, rp, cfg, n, duration_s=d, TR_s=TR)

    _core.core.set_tuning("fc_mode", 2)
    wall = integ_ms = fc_ms = 0.0
    C = min(args.c5_check_sims, S)
    fc_err = fcd_err = 0.0
    try:
        log(f"c5: warm-up (noise segment, workspaces): 16 simulations of 30 s")
        _core.run_simulations(*args_for(G[:16], 30.0))
        for b0 in range(0, S, B):
            g = G[b0:b0 + B]
            t0 = time.time()
            out = _core.run_simulations(*args_for(g, dur))
            dt_call = time.time() - t0
            wall += dt_call
            t = _core.last_timings()
            assert t["fc_path"] == 2
            integ_ms += t["integrate_ms"]
            fc_ms += t["fc_fcd_ms"]
            integrator = ["per-simulation", "batched MFMA (wide)", "many-node"][int(t["integrator"])]
            log(f"c5: simulations {b0}..{b0 + g.size - 1} in {dt_call:.1f} s (integration {t['integrate_ms'] / 1000:.1f} s, "
                f"FC/FCD {t['fc_fcd_ms'] / 1000:.2f} s, {integrator})")
            vols = out["sim_bold"].shape[1] // N
            n_fcd = out["sim_fcd_trils"].shape[1]
            for i in range(min(C - b0, g.size)):
                bb = out["sim_bold"][i].reshape(vols, N).T
                fc_err = max(fc_err, float(np.abs(utils.calculate_fc(bb) - out["sim_fc_trils"][i]).max()))
                fcd_err = max(fcd_err, float(np.abs(utils.calculate_fcd(bb, 30, 1) - out["sim_fcd_trils"][i]).max()))
                log(f"c5: check {b0 + i + 1}/{C}: FC err {fc_err:.2e}, FCD err {fcd_err:.2e}")
            del out
    finally:
        _core.core.set_tuning("fc_mode", 0)
    n_w = int(round((1 + np.sqrt(1 + 8 * n_fcd)) / 2))
    pairs = N * (N - 1) // 2
    # FCD Gram products alone, lower triangle (the FCD tril), 3 BF16 terms: a
    # lower bound on the FC/FCD stage's rate
    fcd_flops = 2.0 * (n_w * (n_w - 1) // 2) * pairs * 3 * S
    work = S * N * dur
    int_flops = work * ops_per_sim_node_second(N, 0.1, 1.0, TR)
    line = {
        "metric": METRIC, "value": round(work / wall, 1), "unit": "sim-node-s/s", "n_gpus": 1,
        "steps": (S + B - 1) // B, "warmup": 1, "ms_per_step": round(1000 * wall / ((S + B - 1) // B), 1),
        "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64 integration, split-bf16 FC/FCD products",
        "data": "synthetic (seeded lognormal SC, host-generated noise)",
        "config": {"workload": f"rWW N={N}, {S} sims in calls of {B}, {dur:g} s, TR {TR} s ({vols} volumes), "
                               f"FC + FCD over {n_w} windows of 31 volumes (window_size 30) step 1, split-BF16 "
                               f"Gram products",
                   "integrate_ms": round(integ_ms, 1), "fc_fcd_ms": round(fc_ms, 1),
                   "fc_fcd_sims_per_s": round(S / (fc_ms / 1000), 1),
                   "fcd_bf16x3_tflops_lower_bound": round(fcd_flops / (fc_ms / 1000) / 1e12, 1),
                   "check_sims": C, "fc_max_abs_err_vs_fp64_host": fc_err,
                   "fcd_max_abs_err_vs_fp64_host": fcd_err},
        # the dominant kernel is the integrator: algorithmic FP64 ops (rWW,
        # N=400: coupling on the matrix cores, model and BW steps on the
        # vector units, one FP64 datapath and peak for both) over its launch
        # time
        "roofline": {"bound": "mfma" if integrator.startswith("batched") else "fp64-vector",
                     "fp64_datapath": "FP64 MFMA and FP64 VALU share one datapath: one 78.6 TFLOP/s peak for both",
                     "achieved": round(int_flops / (integ_ms / 1000) / 1e12, 3),
                     "peak": 78.6, "unit": "TFLOP/s",
                     "frac": round(int_flops / (integ_ms / 1000) / 1e12 / 78.6, 4), "traffic": None,
                     "kernel": ("integrate_batched_wide_kernel" if integrator.startswith("batched")
                                else "integrate_general_kernel"),
                     "integrator": integrator, "kernel_ms": round(integ_ms, 1)},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def run_c4(args):
    """Regional fit of rWW (G + w_p[100] + J_N[100], unit-cube bounds G in
    [0.001, 10], w_p in [0, 2], J_N in [0.001, 0.5]; FIC off, wIE = 1) to a
    seeded AR(1) empirical BOLD of 100 x 166 volumes at TR 0.72 s, 120 s
    simulations at dt 0.1 ms, loss (1 - FC corr) + mean |FCD quantile
    differences| (64 quantiles; FCD windows 42 volumes, step 7, 41 volumes of
    burn-in).  CMA-ES (popsize 64, sigma 0.5) runs for --c4-gens generations;
    Adam on 256 starts (one adjoint pass per step, checkpoint every 1,000
    integration steps = 100 BW steps; loss and its gradient by the loss
    kernel) runs until it reaches CMA-ES's best loss.  value = CMA-ES wall time
    to its best loss / gradient wall time to the same loss (higher is better)."""
    import torch
    import cases
    from cubnm_amd import fit, grad
    N, TR_s, dur = 100, 0.72, args.c4_duration
    sc = cases.synthetic_sc(N, seed=0)
    emp = empirical_bold(N, int(dur * 1000) // int(TR_s * 1000), seed=4)
    bounds = {"G": (0.001, 10.0), "w_p": (0.0, 2.0), "J_N": (0.001, 0.5)}
    nvr = int(30 * 1000 / (TR_s * 1000))
    mk = lambda n: grad.RwwProblem(sc, np.ones((n, N)), 0.01, dur, TR_s, seed=0, checkpoint_bw_steps=100)
    fit_cma = fit.RegionalFit(mk(64), emp, bounds, 42, 7, nvr)
    log("c4: CMA-ES")
    h_cma = fit.cmaes_fit(fit_cma, args.c4_gens, seed=1, log=log)
    best = h_cma[-1][1]
    t_cma = next(t for t, b in h_cma if b <= best)
    del fit_cma
    torch.cuda.empty_cache()
    fit_gd = fit.RegionalFit(mk(256), emp, bounds, 42, 7, nvr)
    log("c4: Adam")
    h_gd = fit.adam_fit(fit_gd, args.c4_steps, lr=args.c4_lr, seed=2, target=best, log=log)
    reached = [t for t, b in h_gd if b <= best]
    t_gd = reached[0] if reached else None
    # one forward and one backward pass of the 256-start batch, timed alone:
    # the adjoint kernels' share of a gradient step and their FP64 rate
    # (algorithmic operations: the forward's 2.444e6-style count per
    # sim-node-second, twice for the backward = recomputation + adjoint sweep
    # of about the forward's size)
    prob = fit_gd.p
    rng = np.random.default_rng(7)
    Gt = torch.tensor(rng.uniform(0.5, 3.0, 256), dtype=torch.float64, device="cuda")
    wpt = torch.tensor(rng.uniform(0.5, 1.5, (256, N)), dtype=torch.float64, device="cuda")
    jnt = torch.tensor(rng.uniform(0.1, 0.3, (256, N)), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    t0 = time.time()
    bold = prob.forward(Gt, wpt, jnt)
    torch.cuda.synchronize()
    t1 = time.time()
    prob.backward(Gt, wpt, jnt, torch.ones_like(bold))
    torch.cuda.synchronize()
    t2 = time.time()
    fwd_ops = 256 * N * dur * ops_per_sim_node_second(N, 0.1, 1.0, TR_s)
    bwd_tflops = 2 * fwd_ops / (t2 - t1) / 1e12
    log(f"c4: forward {t1 - t0:.2f} s, backward {t2 - t1:.2f} s for 256 starts")
    cpu = None
    if not args.no_cpu_baseline:
        n_cores = args.cpu_cores or host_cpu()[0]
        g = np.random.default_rng(8)
        cfg = cases.model_config("rWW", TR=TR_s, window_size=42, window_step=7, bold_remove_s=30, do_fic=False)
        cpu = cpu_baseline(n_cores, N, dur, g.uniform(0.5, 3.0, 64), g.uniform(0.5, 1.5, 64), g.uniform(0.1, 0.3, 64),
                           sc, cfg, per_core=1, repeats=1)
        cpu["cmaes_generation_s"] = round(64 * N * dur / cpu["value"], 1)
        cpu["note"] = ("the CMA-ES generation's forward work (64 simulations of 120 s) on the host engine; "
                       "the GPU's generation takes cmaes_seconds_to_best / generations")
        log(f"cpu baseline: {cpu}")
    line = {
        "metric": METRIC, "value": round(t_cma / t_gd, 3) if t_gd else 0.0, "unit": "x (CMA-ES time / gradient time to CMA-ES best loss)",
        "n_gpus": 1, "steps": len(h_gd), "warmup": 0, "ms_per_step": round(1000 * h_gd[-1][0] / len(h_gd), 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded lognormal SC, seeded AR(1) empirical BOLD 100x166 at TR 0.72 s, host-generated noise)",
        "config": {"workload": f"rWW regional fit, 201 parameters, N={N}, {dur:g} s, TR 0.72 s; CMA-ES popsize 64 "
                               f"x {args.c4_gens} generations vs Adam (lr {args.c4_lr:g}) on 256 starts with adjoint "
                               f"gradients",
                   "cmaes_best_loss": round(best, 6), "cmaes_seconds_to_best": round(t_cma, 2),
                   "grad_seconds_to_cmaes_best": round(t_gd, 2) if t_gd else None,
                   "grad_best_loss": round(h_gd[-1][1], 6), "grad_steps": len(h_gd),
                   "forward_s_256": round(t1 - t0, 3), "backward_s_256": round(t2 - t1, 3)},
        # the adjoint kernels issue no MFMA: FP64 vector peak
        "roofline": {"bound": "fp64", "achieved": round(bwd_tflops, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(bwd_tflops / FP64_PEAK_TFLOPS, 4), "traffic": None,
                     "kernel": "rww_adj_kernel (backward pass)", "kernel_ms": round(1000 * (t2 - t1), 1),
                     "note": "one simulation per workgroup, 256 simulations on 256 CUs: bound by the latency "
                             "of each step's dependent chain, not by issue"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
