#!/usr/bin/env python3
"""Benchmark on one MI355X.

Headline (`value`): the plain GPU path — Monte Carlo model build (move + LDS radix
sort/count-by-key + CSR) followed by fused SpMV value iteration to residual 1e-6 — on
the 100x100 x nt=120, 16-heading, 5-mode, 2500-realization workload. Each sweep
recomputes the slices whose values are not yet final (slice t is final from sweep nt - t
on): the same sweeps, residuals and bits as sweeps over every state, which "vi.full_sweeps"
times on a model built apart. One step = one
end-to-end build + solve with the fields already resident in HBM; `value` = Monte Carlo
transitions per second of the whole step.

Also reported on the same line:
  * "batched": 64 goal columns on the 25x25 x nt=20, 8-heading model, solved as one
    SpMM-VI (autotuned, hipGraph sweeps) vs 64 sequential plain SpMV-VI solves of the
    per-goal models, with the largest column difference (bar: 1e-5 relative), and the
    same columns by one exact backward pass (goal variant -3); and 16 goal columns on the
    100x100 x nt=120 model: the batched sweep's cost in plain sweeps and its roofline;
  * "serving": a Poisson(20/s) stream of planning requests over environments with
    grids {50^2, 100^2, 150^2}, nt in U[60,150], 8 or 16 headings, 1000-5000
    realizations: p50/p99 latency and solves/s of the continuous-batching server vs a
    one-request-at-a-time queue, and of the server in exact-pass mode (one backward pass
    per scheduler step);
  * "paths": build cost of C2 with obstacle islands and with reward method 2 over a
    scalar field;
  * "alpha_sweep": the 11-alpha sweep (src/my_script.sh:10-19) from one build and one
    batched solve vs 11 builds and 11 solves;
  * "rollout": the C2 policy's trajectories (100 start cells x 2500 realizations,
    interpolated actions), with the oracle's CPU rate beside it;
  * "vi_backward": the same model solved by the exact backward pass (one launch per
    slice), bit-identical to the sweeps once they have run nt times;
  * "envgen": input generation at the C2 size (realizations, 5 leading modes per slice
    from the Gram matrix, clouds, obstacles), with one numpy/LAPACK slice beside it.
Per-kernel times come from HIP events recorded inside the library on the stream the
kernels run on.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_HBM_GBPS = 8000.0
PEAK_FP32_TFLOPS = 157.3


def _pmc_traffic(workload: str, kernel: str):
    """HBM bytes per launch of `kernel` as measured with rocprofv3 PMC counters on this
    workload (profiles/hbm_traffic.json, commands in profiles/*_hbm_traffic.md); None
    when no measurement for this workload and kernel is on file."""
    f = ROOT / "profiles" / "hbm_traffic.json"
    try:
        d = json.loads(f.read_text())
    except (OSError, ValueError):
        return None
    k = d.get(workload, {}).get(kernel)
    return k["traffic_bytes"] if k else None


def _build_kernel(info: dict) -> str:
    """The kernel that carries a build: the counting kernel when it took the cells, else
    the per-row wavefront kernel, else the block-level kernel."""
    if info["counted_cells"] > 0:
        return "mc_count"
    return "mc_wave" if info["wave_blocks"] > 0 else "mc_move_count"


def _build_roofline(kernel: str, tflops: float, flops_per_lane: int, traffic):
    """The build's rate in the reference's per-lane flops (the velocity sample's 2 nm
    mul+add, 2 mean adds and the move's flops, for every (state, action, realization)
    lane), over the move/count kernels' time. The counting kernel reaches the same counts
    without executing those flops lane by lane (one sample per realization and cell,
    placed among all actions' landing boundaries), so for it this is a lane-equivalent
    rate; its executed VALU issue is in profiles/c2_count_kernel_counters.md."""
    r = {"kernel": kernel, "bound": "valu", "achieved": tflops, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
         "frac": tflops / PEAK_FP32_TFLOPS, "traffic": traffic, "flops_per_lane": flops_per_lane}
    if kernel == "mc_count":
        r["flops"] = "lane-equivalent: the per-lane flops of the moves the counting replaces"
    return r


def _threads() -> int:
    """Host threads of every CPU leg: OMP_NUM_THREADS when set (the CPU share a GPU box
    gives one GPU), else every core this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def _host_info() -> dict:
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "threads_used": _threads()}


def _cpu_vi(np, p, coo, thresh, budget_s: float, gpu_bytes_per_sweep: float):
    """The oracle's OpenMP float32 sweeps (the reference's solve loop, src/spvi_2.cu:650-726,
    on the host cores) over the same model: the whole solve to `thresh` when it fits the
    budget, else a sample of sweeps extrapolated to the GPU's sweep count (labelled)."""
    import oracle
    th = _threads()
    S1, S2, val, R = coo
    ns, na = p.ns, p.na
    t0 = time.perf_counter()
    oracle.vi_sweeps_f32(ns, na, S1, S2, val, R, thresh, max_sweeps=1, nthreads=th)
    t1 = time.perf_counter() - t0
    t0 = time.perf_counter()
    oracle.vi_sweeps_f32(ns, na, S1, S2, val, R, thresh, max_sweeps=2, nthreads=th)
    per = max(time.perf_counter() - t0 - t1, 1e-6)  # the row-pointer set-up cancels
    full = per * (p.nt + 2) <= budget_s
    if full:
        t0 = time.perf_counter()
        _, _, it, _ = oracle.vi_sweeps_f32(ns, na, S1, S2, val, R, thresh, max_sweeps=p.nt + 2, nthreads=th)
        solve_s = time.perf_counter() - t0
        sample = f"whole solve to {thresh:g}: {it} sweeps in {solve_s:.2f} s"
    else:
        it = p.nt
        solve_s = t1 - per + per * it
        sample = (f"runs of 1 and 2 sweeps timed ({per * 1e3:.0f} ms per sweep), extrapolated to {it} sweeps: "
                  f"{solve_s:.1f} s (not run in full)")
    return {"ms_per_sweep": per * 1e3, "sweeps": it, "solve_s": solve_s, "extrapolated": not full,
            "gbps_equiv": gpu_bytes_per_sweep / per / 1e9, "cores": th, "kind": "port",
            "sample": f"oracle float32 sweeps (OpenMP, {th} threads): {sample}"}


def _cpu_baseline(w, fields, budget_s: float):
    """The oracle's OpenMP build on the host cores over the first slices of the same
    workload (bounded sample)."""
    import oracle
    threads = _threads()
    p = w.params
    lanes_per_slice = p.nc * p.na * p.nrzns
    t0 = time.perf_counter()
    oracle.build(p, fields, 0, 1, threads)
    t1 = time.perf_counter() - t0
    n = max(1, min(p.nt - 2, int(budget_s / max(t1, 1e-3))))
    t0 = time.perf_counter()
    oracle.build(p, fields, 1, 1 + n, threads)
    dt = time.perf_counter() - t0
    return {
        "value": lanes_per_slice * n / dt, "unit": "transitions/s", "cores": threads,
        "kind": "port",
        "sample": f"oracle build (OpenMP, {threads} threads) of time slices 1..{n} of the same "
                  f"workload: {lanes_per_slice * n:.3e} lanes in {dt:.1f} s",
    }


def _batched(e2e, np, steps: int, cpu: bool = True):
    """C1 model x 64 goals: batched SpMM-VI vs 64 sequential plain SpMV-VI solves."""
    from e2e_gpu_dp_amd.configs import goal_cells, workload
    w = workload("C1")
    f = w.fields()
    p = w.params
    goals = goal_cells(p.gsize, 64, 3456)
    shared = e2e.build_model(p.replace(term_subgrid_size=0), f, keep_rsum=True)
    shared.solve_goals(goals, thresh=1e-6, use_graph=True)  # autotune + graph capture
    t0 = time.perf_counter()
    for _ in range(steps):
        pol_b, val_b, sw_b, st_b = shared.solve_goals(goals, thresh=1e-6, use_graph=True)
    t_batch = (time.perf_counter() - t0) / steps
    # the same 64 columns by one exact backward pass (each column at its fixed point)
    shared.solve_goals(goals, variant=-3)
    t0 = time.perf_counter()
    for _ in range(steps):
        pol_x, val_x, _, st_x = shared.solve_goals(goals, variant=-3)
    t_back = (time.perf_counter() - t0) / steps
    plains = [e2e.build_model(p.replace(i_term=int(g) // p.gsize, j_term=int(g) % p.gsize,
                                        term_subgrid_size=1), f) for g in goals]
    for m in plains[:2]:
        m.solve(thresh=1e-6, use_graph=True)
    t0 = time.perf_counter()
    outs = [m.solve(thresh=1e-6, use_graph=True) for m in plains]
    t_plain = time.perf_counter() - t0
    err = err_x = 0.0
    for j, (pp, vp, sp) in enumerate(outs):
        v64 = vp.astype(np.float64)
        err = max(err, float((np.abs(val_b[j] - v64) / np.maximum(np.abs(v64), 1.0)).max()))
        err_x = max(err_x, float((np.abs(val_x[j] - v64) / np.maximum(np.abs(v64), 1.0)).max()))
    coos = [m.export_coo() for m in plains] if cpu else []
    for m in plains:
        m.close()
    shared.close()
    cpu_leg = None
    if cpu:  # the same 64 per-goal models, solved one after another on the host cores
        import oracle
        th = _threads()
        t0 = time.perf_counter()
        for (S1, S2, v, R) in coos:
            oracle.vi_sweeps_f32(p.ns, p.na, S1, S2, v, R, 1e-6, max_sweeps=p.nt + 2, nthreads=th)
        dt = time.perf_counter() - t0
        cpu_leg = {"value": 64 / dt, "unit": "solves/s", "cores": th, "kind": "port",
                   "sample": f"oracle float32 sweeps (OpenMP, {th} threads) of the 64 per-goal models one "
                             f"after another, to 1e-6: {dt * 1e3:.1f} ms"}
    return {
        "workload": "25x25 x nt=20, 8 headings, 4 modes, 1000 realizations; 64 goal cells (seed 3456)",
        "solves_per_s": 64 / t_batch, "plain_solves_per_s": 64 / t_plain,
        "speedup_vs_plain": t_plain / t_batch, "batched_ms": t_batch * 1e3, "plain_64_ms": t_plain * 1e3,
        "max_rel_diff_vs_plain": err, "variant": st_b["variant"], "cols_per_lane": st_b["cols_per_lane"],
        "states_per_wave": st_b["states_per_wave"], "sweeps": st_b["sweeps"],
        "sweep_kernel_ms": st_b["sweep_kernel_ms"] / max(st_b["launches"], 1),
        "bytes_per_sweep": st_b["bytes_per_sweep"],
        "backward_ms": t_back * 1e3, "backward_solves_per_s": 64 / t_back,
        "backward_speedup_vs_plain": t_plain / t_back, "backward_max_rel_diff_vs_plain": err_x,
        "backward_launches": st_x["launches"], "cpu_baseline": cpu_leg,
    }


def _batched_c2(e2e, np, torch, k: int = 16):
    """k goal columns on the shared 100x100 x nt=120 model (no terminal cell, reward sums
    kept): the batched sweep over every state against the plain sweep over every state,
    and its algorithmic bytes (entries and row data once, k value columns read and
    written, k policies) over HBM peak."""
    from e2e_gpu_dp_amd.configs import goal_cells, workload
    w = workload("C2")
    p = w.params.replace(term_subgrid_size=0)
    f = w.fields()
    dev = {kk: torch.from_numpy(np.ascontiguousarray(v)).cuda() for kk, v in f.items()}
    goals = goal_cells(p.gsize, k, 3456)
    with e2e.build_model(p, dev, keep_rsum=True) as m:
        _, _, sp = m.solve(thresh=1e-6, full_sweeps=True)
        m.solve_goals(goals, thresh=1e-6, full_sweeps=True)  # warm (kernel choice, buffers)
        _, _, sw, st = m.solve_goals(goals, thresh=1e-6, full_sweeps=True)
        nnz, rows, ns = m.info["nnz"], p.ns * p.na, p.ns
    del dev
    per = st["sweep_kernel_ms"] / max(st["launches"], 1)
    plain = sp["sweep_kernel_ms_per_sweep"]
    bytes_ = 8.0 * nnz + 8.0 * rows + 4.0 + 20.0 * k * ns
    gbps = bytes_ / (per * 1e-3) / 1e9
    return {"workload": f"C2 shared model (100x100 x nt=120, 16 headings, 2500 realizations), {k} goal columns, "
                        f"sweeps over every state", "columns": k, "variant": st["variant"],
            "sweep_kernel_ms": per, "plain_sweep_kernel_ms": plain, "plain_sweeps_per_batched_sweep": per / plain,
            "per_column_vs_plain": per / plain / k,
            "roofline": {"kernel": "spmm_vi (batched sweep)", "bound": "hbm", "achieved": gbps,
                         "peak": PEAK_HBM_GBPS, "unit": "GB/s", "frac": gbps / PEAK_HBM_GBPS,
                         "bytes_per_sweep": bytes_}}


def _serving(e2e, np, n_requests: int, seed: int = 5678, rate: float = 20.0, only=None):
    """Poisson(rate) planning requests over mixed environments; server vs serial queue.
    only: run just the server with these options and return its figures."""
    from e2e_gpu_dp_amd.configs import Params
    from e2e_gpu_dp_amd.fields import make_fields
    rng = np.random.default_rng(seed)
    envs = {}
    for eid in range(6):
        g = int(rng.choice([50, 100, 150]))
        nt = int(rng.integers(60, 151))
        na = int(rng.choice([8, 16]))
        nr = int(rng.integers(1000, 5001))
        nm = 5
        p = Params(gsize=g, nt=nt, nrzns=nr, nmodes=nm, num_ac_angles=na, term_subgrid_size=0)
        envs[eid] = (p, make_fields(g, nt, nm, nr, seed + 1 + eid))
    arrivals = np.cumsum(rng.exponential(1 / rate, n_requests))
    reqs = []
    for t in arrivals:
        eid = int(rng.integers(0, len(envs)))
        nc = envs[eid][0].nc
        reqs.append((float(t), eid, int(rng.integers(0, nc)), int(rng.integers(0, nc))))

    def run(**opts):
        with e2e.Server(thresh=1e-6, **opts) as srv:
            for eid, (p, f) in envs.items():
                srv.add_env(eid, p, f)
            t0 = time.perf_counter()
            ids = []
            for (t, eid, g, s0) in reqs:
                dt = t - (time.perf_counter() - t0)
                if dt > 0:
                    time.sleep(dt)
                ids.append(srv.submit(eid, g, s0))
            srv.drain(timeout_ms=600000)
            st = srv.stats()
            res = [srv.poll(i) for i in ids]
        return st, res

    if only is not None:
        st, res = run(**only)
        slow = sorted(range(len(res)), key=lambda i: -res[i]["latency_ms"])[:6]
        return {"p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "solves_per_s": st["solves_per_s"],
                "completed": st["completed"], "wall_ms": st["wall_ms"],
                "slowest": [(i, reqs[i][1], round(reqs[i][0], 3), round(res[i]["latency_ms"], 1),
                             round(res[i]["queue_ms"], 1), res[i]["cache_hit"]) for i in slow]}
    st, res = run(max_cols=16)
    st_q, res_q = run(max_cols=1, serial=True)
    st_x, res_x = run(max_cols=16, exact_pass=True)
    st_a, _ = run(max_cols=16, max_models=-1)  # every model with running requests steps at once
    st_w, _ = run(max_cols=64)  # batches of up to 64 columns
    same = sum(1 for a, b in zip(res, res_q) if a["value_at_start"] == b["value_at_start"])
    heavy = None
    if rate == 20.0:  # a stream the one-at-a-time queue cannot keep up with
        h = {name: _serving(e2e, np, 2 * n_requests, seed, rate=60.0, only=opts)
             for name, opts in (("server", dict(max_cols=16)), ("queue", dict(max_cols=1, serial=True)),
                                ("exact_pass", dict(max_cols=16, exact_pass=True)))}
        heavy = {"workload": f"{2 * n_requests} requests, Poisson 60/s, the same 6 environments",
                 **{k: {kk: v[kk] for kk in ("p50_ms", "p99_ms", "solves_per_s", "completed")} for k, v in h.items()}}
    return {
        "workload": f"{n_requests} requests, Poisson 20/s, 6 environments "
                    f"(grids {{50,100,150}}^2, nt U[60,150], 8/16 headings, 1000-5000 realizations)",
        "envs": [{"g": p.gsize, "nt": p.nt, "na": p.na, "nr": p.nrzns} for p, _ in envs.values()],
        "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "solves_per_s": st["solves_per_s"],
        "completed": st["completed"], "models_built": st["models_built"],
        "queue_p50_ms": st_q["p50_ms"], "queue_p99_ms": st_q["p99_ms"],
        "queue_solves_per_s": st_q["solves_per_s"],
        "same_value_at_start_as_queue": f"{same}/{len(res)}",
        "exact_pass": {"p50_ms": st_x["p50_ms"], "p99_ms": st_x["p99_ms"], "solves_per_s": st_x["solves_per_s"],
                       "completed": st_x["completed"]},
        "all_models_at_once": {"p50_ms": st_a["p50_ms"], "p99_ms": st_a["p99_ms"],
                               "solves_per_s": st_a["solves_per_s"], "completed": st_a["completed"]},
        "batches_of_64": {"p50_ms": st_w["p50_ms"], "p99_ms": st_w["p99_ms"],
                          "solves_per_s": st_w["solves_per_s"], "completed": st_w["completed"]},
        "heavy": heavy,
    }


def _paths(e2e, np, torch):
    """Build cost of the paths C2 leaves unexercised, on the same grid, horizon and
    sample: obstacle islands (the per-lane obstacle test) and reward method 2 over a
    scalar field (non-integer rewards: the canonical float64 chunk sums)."""
    from e2e_gpu_dp_amd.configs import workload
    res = {}
    for name, key in (("C2-obs", "obstacles"), ("C2-rad", "scalar_reward")):
        w = workload(name)
        f = w.fields()
        dev = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in f.items()}
        torch.cuda.synchronize()
        infos = []
        for _ in range(3):  # the first is a warmup
            m = e2e.build_model(w.params, dev)
            infos.append(m.info)
            m.close()
        i = infos[-1]
        mv = float(np.mean([x["move_kernel_ms"] for x in infos[1:]]))
        res[key] = {"workload": name, "build_ms": float(np.mean([x["build_ms"] for x in infos[1:]])),
                    "move_kernel_ms": mv, "transitions_per_s": w.lanes / (mv * 1e-3),
                    "nnz": i["nnz"], "deferred_rows": i["n_deferred_rows"],
                    "obstacle_cells": int(f["mask"].sum()) if name == "C2-obs" else 0}
        del dev
    return res


ROLLOUT_STARTS = 100  # a 10 x 10 block of start cells next to the reference's startpos


def _rollout_starts(p):
    """10 x 10 start cells ending at the reference's startpos (int(0.5 g), int(0.95 g)),
    Post_processing/plot_e2eGPU_policy.py:875."""
    si, sj = int(0.5 * p.gsize), int(0.95 * p.gsize)
    return [(i, j) for i in range(si - 5, si + 5) for j in range(sj - 10, sj)]


def _rollouts(e2e, np, p, dev, pol):
    """Policy rollouts (plot_exact_trajectory_set_DP with the interpolated policy, the
    run the reference logs, :949-952) of the C2 policy: 100 start cells x all
    realizations, fields resident in HBM."""
    starts = _rollout_starts(p)
    runs = []
    for _ in range(3):  # the first is a warmup
        runs.append(e2e.rollout(p, dev, pol, starts, interp=True))
    st = [r["stats"] for r in runs[1:]]
    kms = float(np.mean([x["kernel_ms"] for x in st]))
    met = e2e.rollout_metrics(runs[-1])
    return {"workload": f"C2 policy, {len(starts)} start cells x {p.nrzns} realizations, interpolated actions",
            "trajectories": st[-1]["trajectories"], "moves": st[-1]["moves"], "kernel_ms": kms,
            "trajectories_per_s": st[-1]["trajectories"] / (kms * 1e-3),
            "moves_per_s": st[-1]["moves"] / (kms * 1e-3),
            "success_rate": met["success_rate"], "time_mean": met["time_mean"],
            "reqd_energy_mean": met["reqd_energy_mean"]}


def _rollout_cpu_baseline(np, p, fields, pol, budget_s):
    """The oracle's OpenMP rollout of the same trajectories on the host cores, over as
    many start cells as fit the budget."""
    import oracle
    starts = _rollout_starts(p)
    nth = _threads()
    t0 = time.perf_counter()
    oracle.rollout(p, fields, pol, starts[:1], interp=True, nthreads=nth)
    one = time.perf_counter() - t0
    k = int(max(1, min(len(starts), budget_s / max(one, 1e-6))))
    t0 = time.perf_counter()
    r = oracle.rollout(p, fields, pol, starts[:k], interp=True, nthreads=nth)
    dt = time.perf_counter() - t0
    n = len(r["success"])
    return {"value": n / dt, "unit": "trajectories/s", "cores": nth, "kind": "port",
            "sample": f"oracle rollout (OpenMP, {nth} threads) of {k} start cells x {p.nrzns} "
                      f"realizations: {n} trajectories in {dt:.2f} s"}


ENVGEN = dict(gsize=100, nt=120, n_wsamples=2500, n_modes=5)  # the C2 grid, horizon, sample, modes


def _envgen_cfg(np):
    return dict(**ENVGEN, dt=10 / ENVGEN["nt"], dxy=2 / ENVGEN["gsize"], A=0.5, eps=0.1, wx=np.pi,
                wy=0.5 * np.pi, w_range=(np.pi / 10, 8 * np.pi / 10), A_sc=3.0)


def _envgen(e2e, np):
    """Input generation (src/data_input/build_environment_DG_with_westward_clouds.py) at the
    C2 size: realizations, means, the 5 leading modes per slice, clouds, obstacles."""
    cfg = _envgen_cfg(np)
    runs = [e2e.generate_environment(**cfg, obstacles="static") for _ in range(2)]  # the first warms up
    st = runs[-1]["stats"]
    m, nw, nt = 2 * cfg["gsize"] ** 2, cfg["n_wsamples"], cfg["nt"]
    gram_flops = 2.0 * min(nw, m) ** 2 * max(nw, m) * nt  # one Gram GEMM per slice
    return {"workload": f"{cfg['gsize']}x{cfg['gsize']} grid, nt={nt}, {nw} forcing frequencies, "
                        f"{cfg['n_modes']} modes, static obstacle",
            "total_ms": st["total_ms"], "modes_ms": st["modes_ms"], "slices_per_s": nt / (st["total_ms"] * 1e-3),
            "gram_tflops": gram_flops / (st["modes_ms"] * 1e-3) / 1e12, "subspace_iterations": st["max_iterations"],
            "leading_singular_value_t60": float(runs[-1]["singular_values"][60, 0])}


def _envgen_cpu_baseline(np):
    """The oracle's numpy/LAPACK path (np.linalg.svd of the centred realizations, as the
    reference) for one slice of the same workload."""
    from oracle import envgen
    cfg = _envgen_cfg(np)
    t0 = time.perf_counter()
    vx, vy, _ = envgen.realizations(cfg["gsize"], cfg["nt"], cfg["n_wsamples"], cfg["dt"], cfg["dxy"], cfg["A"],
                                    cfg["eps"], cfg["wx"], cfg["wy"], *cfg["w_range"], t_slices=[60])
    R = np.concatenate((vx[0], vy[0]), axis=1)
    envgen.modes(R - R.mean(axis=0), cfg["n_modes"])
    dt = time.perf_counter() - t0
    return {"value": 1.0 / dt, "unit": "slices/s", "cores": _threads(), "kind": "port",
            "sample": f"one slice (t = 60) with numpy realizations and LAPACK SVD of the "
                      f"{cfg['n_wsamples']} x {2 * cfg['gsize'] ** 2} matrix: {dt:.1f} s"}


def _backward(e2e, np, p, dev, thresh):
    """The exact backward pass (kernel_variant 7: one launch per slice, last to first) on
    the C2 model, against the sweeps to the bench's residual; reports whether the two
    return the same bits (they do once the sweeps have run nt times)."""
    m = e2e.build_model(p, dev)
    pj, vj, sj = m.solve(thresh=thresh, max_sweeps=p.nt + 2)
    runs = [m.solve(kernel_variant=7) for _ in range(3)]  # the first warms up
    pb, vb, sb = runs[-1]
    m.close()
    kms = float(np.mean([r[2]["sweep_kernel_ms"] for r in runs[1:]]))
    sms = float(np.mean([r[2]["solve_ms"] for r in runs[1:]]))
    return {"launches": sb["launches"], "kernel_ms": kms, "solve_ms": sms,
            "sweeps_solve_ms": sj["solve_ms"], "sweeps": sj["sweeps"],
            "speedup_vs_sweeps": sj["solve_ms"] / sms,
            "bit_identical_to_sweeps": bool(np.array_equal(vj.view(np.uint32), vb.view(np.uint32))
                                            and np.array_equal(pj, pb))}


def _alpha_sweep(e2e, np, torch):
    """The reference's alpha sweep (src/my_script.sh:10-19: build + solve for alpha =
    0, 0.1, .., 1) on the C2 grid with reward method 4 and 2 speeds x 8 headings: one build with 11 alpha
    columns + one batched solve, against 11 builds + 11 plain solves, every column
    compared bit for bit."""
    from e2e_gpu_dp_amd.configs import workload
    w = workload("C2-alpha")
    p = w.params
    alphas = (np.arange(11, dtype=np.float32) * np.float32(10)) / np.float32(100)
    f = w.fields()
    dev = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in f.items()}
    torch.cuda.synchronize()

    def batched():
        t0 = time.perf_counter()
        m = e2e.build_model(p, dev, alphas=alphas)
        b_ms = m.info["build_ms"]
        pol, val, sw, st = m.solve_alphas(thresh=1e-6)
        m.close()
        return (time.perf_counter() - t0) * 1e3, b_ms, st, pol, val

    batched()  # warmup
    total_ms, build_ms, st, pol, val = batched()
    t0 = time.perf_counter()
    same = 0
    for j, a in enumerate(alphas):
        m = e2e.build_model(p.replace(alpha=float(a)), dev)
        pj, vj, _ = m.solve(thresh=1e-6)
        m.close()
        same += int(np.array_equal(vj.view(np.uint32), val[j].view(np.uint32)) and np.array_equal(pj, pol[j]))
    plain_ms = (time.perf_counter() - t0) * 1e3
    # the same sweep with every column by the exact backward pass (variant -3)
    t0 = time.perf_counter()
    m = e2e.build_model(p, dev, alphas=alphas)
    pol_b, val_b, _, st_b = m.solve_alphas(variant=-3)
    m.close()
    bw_ms = (time.perf_counter() - t0) * 1e3
    same_b = sum(int(np.array_equal(val_b[j].view(np.uint32), val[j].view(np.uint32))
                     and np.array_equal(pol_b[j], pol[j])) for j in range(len(alphas)))
    return {"workload": "C2-alpha: 100x100 x nt=120, 2 speeds x 8 headings, reward method 4, 11 alphas",
            "backward_sweep_ms": bw_ms, "backward_solve_ms": st_b["solve_ms"],
            "backward_speedup": plain_ms / bw_ms,
            "backward_columns_equal_sweeps": f"{same_b}/{len(alphas)}",
            "sweep_ms": total_ms, "build_ms": build_ms, "solve_ms": st["solve_ms"],
            "per_alpha_builds_and_solves_ms": plain_ms, "speedup": plain_ms / total_ms,
            "columns_bit_identical": f"{same}/{len(alphas)}", "sweeps": st["sweeps"],
            "variant": st["variant"], "cols_per_lane": st["cols_per_lane"]}


def _c1(e2e, np, torch, thresh, cpu: bool):
    """The small workload (25x25 x nt=20, 8 headings, 4 modes, 1000 realizations) end to
    end: build + solve to `thresh` on the GPU (fields resident in HBM), and the same on
    the host cores (the oracle's OpenMP build, then its float32 sweeps)."""
    from e2e_gpu_dp_amd.configs import workload
    w = workload("C1")
    p = w.params
    f = w.fields()
    dev = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in f.items()}
    torch.cuda.synchronize()

    def step():
        with e2e.build_model(p, dev) as m:
            _, _, st = m.solve(thresh=thresh, max_sweeps=p.nt + 2)
        return st

    step()
    t0 = time.perf_counter()
    for _ in range(5):
        st = step()
    gpu_ms = (time.perf_counter() - t0) * 1e3 / 5
    out = {"workload": "C1: 25x25 x nt=20, 8 headings, 4 modes, 1000 realizations; build + solve to "
                       f"{thresh:g}", "gpu_ms": gpu_ms, "sweeps": st["sweeps"],
           "transitions_per_s": w.lanes / (gpu_ms * 1e-3)}
    if cpu:
        import oracle
        th = _threads()
        t0 = time.perf_counter()
        r = oracle.build(p, f, nthreads=th)
        tb = time.perf_counter() - t0
        _, _, it, _ = oracle.vi_sweeps_f32(p.ns, p.na, r["S1"], r["S2"], r["val"], r["R"], thresh,
                                           max_sweeps=p.nt + 2, nthreads=th)
        tt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": w.lanes / tt, "unit": "transitions/s", "cores": th, "kind": "port",
                               "build_s": tb, "solve_s": tt - tb, "sweeps": it,
                               "sample": f"the whole workload: oracle build (OpenMP, {th} threads) {tb:.2f} s + "
                                         f"float32 sweeps {tt - tb:.3f} s"}
        out["gpu_vs_cpu"] = tt * 1e3 / gpu_ms
    return out


def _c4(e2e, np, torch, thresh, cpu: bool, cpu_budget: float):
    """The large single MDP (200x200 x nt=200, 16 headings, 6 modes, 5000 realizations):
    build + solve on the GPU (fields resident in HBM) with the build and sweep rooflines,
    and the oracle on the host cores over bounded samples, extrapolated and labelled."""
    from e2e_gpu_dp_amd.configs import workload
    w = workload("C4")
    p = w.params
    print("[bench] C4 fields", file=sys.stderr, flush=True)
    f = w.fields()
    dev = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in f.items()}
    torch.cuda.synchronize()

    def step():
        with e2e.build_model(p, dev) as m:
            _, _, st = m.solve(thresh=thresh, max_sweeps=p.nt + 2)
            return m.info, st

    step()
    runs = []
    t0 = time.perf_counter()
    for _ in range(2):
        runs.append(step())
    ms = (time.perf_counter() - t0) * 1e3 / 2
    info, st = runs[-1]
    move_ms = float(np.mean([i["move_kernel_ms"] for i, _ in runs]))
    fl = 4 * p.nmodes + 18
    tflops = w.lanes * fl / (move_ms * 1e-3) / 1e12
    with e2e.build_model(p, dev) as m:
        _, vf, sf = m.solve(thresh=thresh, max_sweeps=p.nt + 2, full_sweeps=True)
        _, va, sa = m.solve(thresh=thresh, max_sweeps=p.nt + 2)
        coo = m.export_coo() if cpu else None
    del dev
    pf = sf["sweep_kernel_ms_per_sweep"]
    gbps = sf["bytes_per_sweep"] / (pf * 1e-3) / 1e9
    out = {
        "workload": f"C4: {p.gsize}x{p.gsize} grid x nt={p.nt}, {p.na} headings, {p.nmodes} DO modes, "
                    f"{p.nrzns} realizations",
        "states": p.ns, "nnz": info["nnz"], "lanes": w.lanes, "ms_per_step": ms,
        "transitions_per_s": w.lanes / (ms * 1e-3), "build_ms": float(np.mean([i["build_ms"] for i, _ in runs])),
        "move_kernel_ms": move_ms, "move_launches": info["move_launches"], "deferred_rows": info["n_deferred_rows"],
        "solve_ms": float(np.mean([s["solve_ms"] for _, s in runs])), "sweeps": st["sweeps"],
        "step_sweep_kernel_ms": st["sweep_kernel_ms_per_sweep"],
        "step_sweeps_gbps": st["bytes_per_sweep"] / (st["sweep_kernel_ms_per_sweep"] * 1e-3) / 1e9,
        "full_sweeps": {"sweep_kernel_ms": pf, "bytes_per_sweep": sf["bytes_per_sweep"], "gbps": gbps,
                        "same_bits": bool(np.array_equal(vf.view(np.uint32), va.view(np.uint32))
                                          and sf["sweeps"] == sa["sweeps"])},
        "counted_cells": info["counted_cells"],
        "roofline": _build_roofline(_build_kernel(info), tflops, fl, _pmc_traffic("C4", _build_kernel(info))),
        "roofline_vi": {"kernel": "spmv_vi_lean", "bound": "hbm", "achieved": gbps, "peak": PEAK_HBM_GBPS,
                        "unit": "GB/s", "frac": gbps / PEAK_HBM_GBPS,
                        "traffic": _pmc_traffic("C4", "spmv_vi_lean")},
    }
    if cpu:
        import oracle
        th = _threads()
        lanes_slice = p.nc * p.na * p.nrzns
        t0 = time.perf_counter()
        oracle.build(p, f, 1, 2, th)
        one = time.perf_counter() - t0
        n = int(max(1, min(5, cpu_budget // max(one, 1e-3) + 1)))
        if n > 1:
            t0 = time.perf_counter()
            oracle.build(p, f, 2, 2 + n - 1, th)
            one = (one + time.perf_counter() - t0) / n
        out["cpu_baseline"] = {
            "value": lanes_slice / one, "unit": "transitions/s", "cores": th, "kind": "port",
            "extrapolated_build_s": one * (p.nt - 1),
            "sample": f"oracle build (OpenMP, {th} threads) of {n} time slice(s) from t = 1, "
                      f"{lanes_slice:.3e} lanes each, {one:.1f} s per slice; the {p.nt - 1}-slice build "
                      f"extrapolated from them: {one * (p.nt - 1):.0f} s (not run in full)"}
        S1, S2, val, R = coo
        out["cpu_baseline"]["vi"] = _cpu_vi(np, p, (S1, S2, val, R), thresh, 0.0, sf["bytes_per_sweep"])
        del coo, S1, S2, val, R
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="C2")
    ap.add_argument("--thresh", type=float, default=1e-6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--skip-vi", action="store_true", help="profiling aid: build only")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serving-requests", type=int, default=120)
    ap.add_argument("--no-paths", action="store_true",
                    help="skip the obstacle and scalar-reward build variants")
    ap.add_argument("--no-c4", action="store_true", help="skip the large-MDP (200^2 x 200) section")
    ap.add_argument("--no-c1", action="store_true", help="skip the small-workload end-to-end section")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run 3 device-to-device copies of BYTES (HBM counter calibration)")
    args = ap.parse_args(argv)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU; multi-GPU runs are not part of it "
              "(run with --gpus 1)", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    if args.copy_calibration:
        n = args.copy_calibration // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n * 4, "copies": 3}))
        return 0

    import e2e_gpu_dp_amd as e2e
    from e2e_gpu_dp_amd.configs import workload

    if e2e.device_count() < 1:
        print("bench.py: no HIP device", file=sys.stderr)
        return 1
    w = workload(args.workload)
    p = w.params
    print(f"[bench] generating fields for {w.name}", file=sys.stderr, flush=True)
    fields = w.fields()
    dev = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in fields.items()}
    torch.cuda.synchronize()

    policy = {}

    def step():
        m = e2e.build_model(p, dev)
        st = None
        if not args.skip_vi:
            pol, _, st = m.solve(thresh=args.thresh, max_sweeps=p.nt + 2)
            policy["last"] = pol
        info = m.info
        m.close()
        return info, st

    for i in range(args.warmup):
        step()
        print(f"[bench] warmup {i} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    infos = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        infos.append(step())
        print(f"[bench] step {i} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1e3 / args.steps
    lanes = infos[0][0]["lanes"]

    # summed HIP-event time of every move/count launch of a build (the counting kernel,
    # then the per-row kernel over the cells it left): all lanes over all of that time
    move_ms = float(np.mean([i["move_kernel_ms"] for i, _ in infos]))
    move_launches = infos[0][0]["move_launches"]
    build_kernel = _build_kernel(infos[0][0])
    flops_per_lane = 4 * p.nmodes + 18
    build_tflops = lanes * flops_per_lane / (move_ms * 1e-3) / 1e12
    build_ms = float(np.mean([i["build_ms"] for i, _ in infos]))
    out = {
        "metric": "MC transitions/s, VI GB/s, batched solves/s + p99 latency vs plain at 1e-5",
        "value": lanes * args.steps / elapsed,
        "unit": "transitions/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 inputs and model (probabilities, rewards), f64 values",
        "data": "synthetic",
        "config": {"workload": f"{w.name}: plain build+SpVI, {p.gsize}x{p.gsize} grid x nt={p.nt}, "
                               f"{p.na} headings, {p.nmodes} DO modes, {p.nrzns} realizations",
                   "gsize": p.gsize, "nt": p.nt, "na": p.na, "nrzns": p.nrzns, "nmodes": p.nmodes,
                   "states": p.ns, "nnz": infos[0][0]["nnz"], "lanes": lanes},
        "build": {"transitions_per_s": lanes / (build_ms * 1e-3), "build_ms": build_ms,
                  "move_kernel_ms": move_ms, "move_launches": move_launches, "scan_ms": infos[0][0]["scan_ms"],
                  "csr_ms": infos[0][0]["csr_ms"], "counted_cells": infos[0][0]["counted_cells"],
                  "cells": (p.nt - 1) * p.nc},
        "roofline": _build_roofline(build_kernel, build_tflops, flops_per_lane,
                                    _pmc_traffic(args.workload, build_kernel)),
    }
    if not args.skip_vi:
        st = infos[-1][1]
        per = float(np.mean([s["sweep_kernel_ms_per_sweep"] for _, s in infos]))
        gbps = st["bytes_per_sweep"] / (per * 1e-3) / 1e9
        out["vi"] = {"sweeps": st["sweeps"], "solve_ms": float(np.mean([s["solve_ms"] for _, s in infos])),
                     "sweep_kernel_ms": per, "bytes_per_sweep": st["bytes_per_sweep"],
                     "gbps": gbps, "loads_in_flight": st["loads_in_flight"]}
        out["vi"]["kernel_variant"] = st["kernel_variant"]
        out["vi"]["blocks"] = st["blocks"]
        # the bytes the variant's layout moves (compressed entries, float64 values), and
        # the rate that makes; bytes_per_sweep / gbps are in the reference's CSR terms
        out["vi"]["moved_bytes_per_sweep"] = st["moved_bytes_per_sweep"]
        out["vi"]["moved_gbps"] = st["moved_bytes_per_sweep"] / (per * 1e-3) / 1e9
        out["vi"]["compressed_windows"] = st["compressed_windows"]
        kname = {1: "spmv_vi_sweep", 2: "spmv_vi_sweep", 3: "spmv_vi_lean", 4: "spmv_vi_tile",
                 5: "spmv_vi_jagged", 6: "spmv_vi_lean"}.get(st["kernel_variant"], "?")
        # the same solve with every slice recomputed on every sweep (what the step's
        # solve skips: the slices already at their final values), on a model built apart
        with e2e.build_model(p, dev) as mf:
            _, vfull, sf = mf.solve(thresh=args.thresh, max_sweeps=p.nt + 2, full_sweeps=True)
            _, vact, sa = mf.solve(thresh=args.thresh, max_sweeps=p.nt + 2)
        pf = sf["sweep_kernel_ms_per_sweep"]
        gbps_full = sf["bytes_per_sweep"] / (pf * 1e-3) / 1e9
        out["vi"]["full_sweeps"] = {
            "sweeps": sf["sweeps"], "sweep_kernel_ms": pf, "solve_kernel_ms": sf["sweep_kernel_ms"],
            "bytes_per_sweep": sf["bytes_per_sweep"], "gbps": gbps_full,
            "active_solve_kernel_ms": sa["sweep_kernel_ms"],
            "same_bits": bool(np.array_equal(vfull.view(np.uint32), vact.view(np.uint32)) and
                              sf["sweeps"] == sa["sweeps"])}
        # the sweep kernel at its full shape (every state, one launch = one sweep of
        # bytes_per_sweep), whose HBM traffic per launch the PMC profile measured; the
        # step's sweeps over the active slices run at vi.gbps
        out["roofline_vi"] = {"kernel": kname, "bound": "hbm", "achieved": gbps_full,
                              "peak": PEAK_HBM_GBPS, "unit": "GB/s", "frac": gbps_full / PEAK_HBM_GBPS,
                              "traffic": _pmc_traffic(args.workload, kname),
                              "step_sweeps_gbps": gbps, "step_sweeps_frac": gbps / PEAK_HBM_GBPS}
    if not args.no_paths and args.workload == "C2":
        print("[bench] obstacle and scalar-reward builds", file=sys.stderr, flush=True)
        out["paths"] = _paths(e2e, np, torch)
        print("[bench] alpha sweep", file=sys.stderr, flush=True)
        out["alpha_sweep"] = _alpha_sweep(e2e, np, torch)
        print("[bench] input generation", file=sys.stderr, flush=True)
        out["envgen"] = _envgen(e2e, np)
    if not args.skip_vi:
        print("[bench] backward pass", file=sys.stderr, flush=True)
        out["vi_backward"] = _backward(e2e, np, p, dev, args.thresh)
    pol32 = None
    if not args.skip_vi and args.workload == "C2":
        print("[bench] rollouts", file=sys.stderr, flush=True)
        pol32 = policy["last"].astype(np.int64).astype(np.int32)
        out["rollout"] = _rollouts(e2e, np, p, dev, pol32)
    if not args.no_batched:
        print("[bench] batched goals", file=sys.stderr, flush=True)
        out["batched"] = _batched(e2e, np, max(args.steps, 3), cpu=not args.no_cpu)
        if args.workload == "C2":
            out["batched"]["c2_16_goals"] = _batched_c2(e2e, np, torch)
    if not args.no_serving:
        print("[bench] serving", file=sys.stderr, flush=True)
        out["serving"] = _serving(e2e, np, args.serving_requests)
    if not args.no_c1 and args.workload == "C2":
        print("[bench] C1 end to end", file=sys.stderr, flush=True)
        out["c1"] = _c1(e2e, np, torch, args.thresh, not args.no_cpu)
    if not args.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        out["host"] = _host_info()
        out["cpu_baseline"] = _cpu_baseline(w, fields, args.cpu_budget)
        if "vi" in out:
            with e2e.build_model(p, dev) as mc:
                coo = mc.export_coo()
            out["vi"]["cpu_baseline"] = _cpu_vi(np, p, coo, args.thresh, 2 * args.cpu_budget,
                                                out["vi"]["full_sweeps"]["bytes_per_sweep"])
            del coo
        if pol32 is not None:
            out["rollout"]["cpu_baseline"] = _rollout_cpu_baseline(np, p, fields, pol32, 3.0)
        if "envgen" in out:
            out["envgen"]["cpu_baseline"] = _envgen_cpu_baseline(np)
    if not args.no_c4 and args.workload == "C2":
        del dev
        torch.cuda.empty_cache()
        print("[bench] C4", file=sys.stderr, flush=True)
        out["c4"] = _c4(e2e, np, torch, args.thresh, not args.no_cpu, args.cpu_budget)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
