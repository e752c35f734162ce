"""Continuous-batching server: every request's result is the same alone and in the mix;
HBM-budget admission evicts idle models or rejects what cannot fit."""
import time

import numpy as np
import pytest

import e2e_gpu_dp_amd as e2e
from e2e_gpu_dp_amd.configs import Params, small
from e2e_gpu_dp_amd.fields import make_fields
from helpers import assert_close_to_f64, assert_slice_matches, goal_params, oracle_model

pytestmark = pytest.mark.gpu


def _envs():
    a = small(g=16, nt=12, nr=300, nm=2, seed=81, na=8)
    b = small(g=20, nt=15, nr=500, nm=3, seed=82, na=16)
    return {1: (a.params, a.fields()), 2: (b.params, b.fields())}


def _solo(params, fields, goal):
    with e2e.build_model(params.replace(term_subgrid_size=0), fields, keep_rsum=True) as m:
        pol, val, sw, _ = m.solve_goals([goal], thresh=1e-6)
    return pol[0], val[0], sw[0]


@pytest.mark.parametrize("max_models", [0, -1])  # the oldest requests' model first; every model at once
def test_server_mix_equals_solo(gpu, max_models):
    envs = _envs()
    rng = np.random.default_rng(7)
    with e2e.Server(keep_full_results=True, thresh=1e-6, sweeps_per_step=3, max_models=max_models) as srv:
        for eid, (p, f) in envs.items():
            srv.add_env(eid, p, f)
        reqs = []
        for i in range(40):
            eid = 1 + (i % 2)
            nc = envs[eid][0].nc
            g, s0 = (int(x) for x in rng.integers(0, nc, 2))
            reqs.append((srv.submit(eid, g, s0), eid, g, s0))
        srv.drain(timeout_ms=120000)
        st = srv.stats()
        assert st["completed"] == 40 and st["models_built"] == 2
        assert st["p99_ms"] >= st["p50_ms"] > 0
        for rid, eid, g, s0 in reqs[:6] + reqs[-4:]:
            r = srv.wait(rid)
            assert r["status_name"] == "done"
            pol, val = srv.columns(rid, eid)
            spol, sval, ssw = _solo(*envs[eid], g)
            assert r["sweeps"] == ssw
            assert np.array_equal(val.view(np.uint32), sval.view(np.uint32))
            assert np.array_equal(pol, spol)
            assert r["value_at_start"] == val[s0] and r["action_at_start"] == int(np.int32(pol[s0]))
            # the FP64 backward induction of the oracle's build of the goal's model
            p, f = envs[eid]
            assert_close_to_f64(oracle_model(goal_params(p, g), f), pol, val, 1e-5, f"env {eid} goal {g}")


def test_server_long_steps_with_converged_column_blocks(gpu):
    """Many sweeps per scheduler step: column blocks whose requests have all converged
    keep being launched until the step's harvest, beside blocks still running; every
    result still equals its solo solve."""
    envs = _envs()
    p, f = envs[2]
    rng = np.random.default_rng(11)
    goals = [int(x) for x in rng.integers(0, p.nc, 12)]
    with e2e.Server(keep_full_results=True, thresh=1e-6, sweeps_per_step=64) as srv:
        srv.add_env(2, p, f)
        ids = [srv.submit(2, g, 0) for g in goals]
        srv.drain(timeout_ms=120000)
        assert srv.stats()["completed"] == len(goals)
        for rid, g in zip(ids, goals):
            r = srv.wait(rid)
            assert r["status_name"] == "done"
            pol, val = srv.columns(rid, 2)
            spol, sval, ssw = _solo(p, f, g)
            assert r["sweeps"] == ssw
            assert np.array_equal(val.view(np.uint32), sval.view(np.uint32))
            assert np.array_equal(pol, spol)


def test_server_budget_evicts_and_rejects(gpu):
    envs = _envs()
    with e2e.Server(hbm_budget_bytes=120 << 20, keep_full_results=False) as srv:
        for eid, (p, f) in envs.items():
            srv.add_env(eid, p, f)
        ids = [srv.submit(1 + (i % 2), 5 + i, 3) for i in range(8)]
        srv.drain(timeout_ms=120000)
        st = srv.stats()
        assert st["completed"] == 8
        assert st["peak_reserved_bytes"] <= 120 << 20
        assert all(srv.poll(i)["status_name"] == "done" for i in ids)
    with e2e.Server(hbm_budget_bytes=1 << 20) as srv:
        p, f = envs[2]
        srv.add_env(2, p, f)
        rid = srv.submit(2, 10, 3)
        r = srv.wait(rid, timeout_ms=30000)
        assert r["status_name"] == "rejected"


def test_server_serial_queue_same_answers(gpu):
    envs = _envs()
    p, f = envs[1]
    with e2e.Server(serial=True, max_cols=1, keep_full_results=True) as srv:
        srv.add_env(1, p, f)
        ids = [srv.submit(1, g, 0) for g in (7, 100, 200)]
        srv.drain(timeout_ms=60000)
        for rid, g in zip(ids, (7, 100, 200)):
            pol, val = srv.columns(rid, 1)
            spol, sval, _ = _solo(p, f, g)
            assert np.array_equal(val.view(np.uint32), sval.view(np.uint32))


def _fixed_point(params, fields, goal):
    """The goal column's fixed point: batched sweeps run to a zero residual."""
    with e2e.build_model(params.replace(term_subgrid_size=0), fields, keep_rsum=True) as m:
        pol, val, sw, _ = m.solve_goals([goal], thresh=1e-30, max_sweeps=params.nt + 4)
    return pol[0], val[0]


def test_server_exact_pass_mix_equals_fixed_point(gpu):
    """exact_pass: each scheduler step is one backward pass (one launch per slice, last to
    first) over a model's running columns; every result, in the mix or alone, is its goal
    column's fixed point bit for bit, after one pass."""
    envs = _envs()
    rng = np.random.default_rng(11)
    with e2e.Server(keep_full_results=True, exact_pass=True) as srv:
        for eid, (p, f) in envs.items():
            srv.add_env(eid, p, f)
        reqs = []
        for i in range(40):
            eid = 1 + (i % 2)
            nc = envs[eid][0].nc
            g, s0 = (int(x) for x in rng.integers(0, nc, 2))
            reqs.append((srv.submit(eid, g, s0), eid, g, s0))
        srv.drain(timeout_ms=120000)
        st = srv.stats()
        assert st["completed"] == 40 and st["models_built"] == 2
        for rid, eid, g, s0 in reqs[:6] + reqs[-4:]:
            r = srv.wait(rid)
            assert r["status_name"] == "done" and r["sweeps"] == 1
            pol, val = srv.columns(rid, eid)
            fpol, fval = _fixed_point(*envs[eid], g)
            assert np.array_equal(val.view(np.uint32), fval.view(np.uint32))
            assert np.array_equal(pol, fpol)
            assert r["value_at_start"] == val[s0] and r["action_at_start"] == int(np.int32(pol[s0]))
        for rid, eid, g, s0 in reqs[:2]:
            pol, val = srv.columns(rid, eid)
            p, f = envs[eid]
            assert_close_to_f64(oracle_model(goal_params(p, g), f), pol, val, 1e-5, f"env {eid} goal {g}")


# The serving workload's shapes: grids 50^2, 100^2, 150^2, horizons in [60, 150], 8 or 16
# headings, 1000-5000 realizations; (gsize, nt, headings, realizations, field seed)
C5_ENVS = {1: (50, 64, 8, 1200, 5701), 2: (100, 76, 16, 3024, 5702), 3: (150, 90, 8, 1000, 5703)}


def _gpu_goal_model(p, f, goal, slices):
    """The goal's model built on the GPU, its sampled slices checked bit for bit against
    the oracle's build of those slices, as a COO dict for the FP64 backward induction."""
    q = goal_params(p, goal)
    with e2e.build_model(q, f) as m:
        S1, S2, val, R = m.export_coo(np.int32)
    for t in slices:
        assert_slice_matches(q, f, t, S1, S2, val, R)
    return {"S1": S1, "S2": S2, "val": val, "R": R, "ns": q.ns, "na": q.na, "nc": q.nc}


def test_server_c5_shapes(gpu):
    """Requests at the serving workload's shapes arrive about every 50 ms (20 per second)
    while others are on the GPU. Every one completes; each checked result equals its solo
    solve bit for bit and lies within 1e-5 of the FP64 backward induction of its goal's
    model (the oracle's build for the 50^2 environment; for the larger ones the GPU build,
    whose first and a late slice equal the oracle's bit for bit)."""
    envs = {}
    for eid, (g, nt, na, nr, seed) in C5_ENVS.items():
        p = Params(gsize=g, nt=nt, nrzns=nr, nmodes=5, num_ac_angles=na, term_subgrid_size=0)
        envs[eid] = (p, make_fields(g, nt, 5, nr, seed))
    rng = np.random.default_rng(5678)
    gaps = rng.exponential(0.05, 20)
    reqs = []
    with e2e.Server(keep_full_results=True, thresh=1e-6) as srv:
        for eid, (p, f) in envs.items():
            srv.add_env(eid, p, f)
        for i, gap in enumerate(gaps):
            time.sleep(float(gap))
            eid = 1 + i % 3
            nc = envs[eid][0].nc
            g, s0 = (int(x) for x in rng.integers(0, nc, 2))
            reqs.append((srv.submit(eid, g, s0), eid, g, s0))
        srv.drain(timeout_ms=240000)
        st = srv.stats()
        print(f"[c5] completed {st['completed']} models {st['models_built']} p50 {st['p50_ms']:.1f} ms "
              f"p99 {st['p99_ms']:.1f} ms", flush=True)
        assert st["completed"] == len(reqs) and st["models_built"] == 3
        checked = {}
        for rid, eid, g, s0 in reqs:
            if checked.get(eid, 0) >= (2 if eid == 1 else 1):
                continue
            checked[eid] = checked.get(eid, 0) + 1
            r = srv.wait(rid)
            assert r["status_name"] == "done"
            pol, val = srv.columns(rid, eid)
            p, f = envs[eid]
            spol, sval, ssw = _solo(p, f, g)
            assert r["sweeps"] == ssw
            assert np.array_equal(val.view(np.uint32), sval.view(np.uint32))
            assert np.array_equal(pol, spol)
            assert r["value_at_start"] == val[s0]
            ref = oracle_model(goal_params(p, g), f) if eid == 1 else _gpu_goal_model(p, f, g, (0, p.nt - 5))
            e64 = assert_close_to_f64(ref, pol, val, 1e-5, f"env {eid} goal {g}")
            print(f"[c5] env {eid} (g={p.gsize}, nt={p.nt}) goal {g}: sweeps {r['sweeps']}, "
                  f"max rel err vs FP64 {e64:.2e}", flush=True)
        assert len(checked) == 3
