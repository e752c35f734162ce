"""Full-size parity on the large single MDP: 200^2 grid x nt=200, 16 headings, 6 modes,
5000 realizations (8e6 states, about 1.3e9 transitions).

The CPU oracle needs about 20 s per time slice here, so it cannot rebuild all 199
slices within a test. What is checked:
* two whole slices, t = 0 and the late-penalty slice t = 185 (T > nt - 20), bit-exact
  against the oracle's build of exactly those slices (src/build_model.cu:1142-1165
  builds slice by slice, so a slice's rows do not depend on the others);
* over every row: probabilities are exactly count/nrzns with counts summing to nrzns,
  next states lie in slice t+1, strictly ascending per row, rows of the last slice
  empty (:1142-1176);
* the solve to 1e-6: values within 1e-5 relative of float64 backward induction over
  the exported model, policies equal except at ties, sweep count nt.
"""
import numpy as np
import pytest

import e2e_gpu_dp_amd as e2e
import oracle
from e2e_gpu_dp_amd.configs import workload
from helpers import policy_ties_ok, q_values_f64, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-5
SLICES = (0, 185)


def test_large_model_slices_invariants_and_solve(gpu):
    import torch
    w = workload("C4")
    p = w.params
    ns, na, nc, nr = p.ns, p.na, p.nc, p.nrzns
    f = w.fields()
    dev = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in f.items()}
    print(f"[large] fields ready: {ns} states x {na} actions", flush=True)
    with e2e.build_model(p, dev) as m:
        info = m.info
        print(f"[large] built: nnz {info['nnz']} in {info['build_ms']:.0f} ms, "
              f"deferred rows {info['n_deferred_rows']}", flush=True)
        pol, vgpu, st = m.solve(thresh=1e-6, sweeps_per_check=16)
        # the float32-value sweeps (1: compensated row sums, 2: plain fma), measured
        # against float64 below
        f32_runs = {kv: m.solve(thresh=1e-6, kernel_variant=kv, raise_on_nonconv=False) for kv in (1, 2)}
        S1, S2, val, R = m.export_coo(np.int32)  # the reference's action-major order
    del dev
    print(f"[large] solved in {st['sweeps']} sweeps, {st['sweep_kernel_ms_per_sweep']:.3f} ms/sweep", flush=True)
    assert st["converged"] and st["sweeps"] == p.nt

    # two slices against the oracle, bit for bit
    for t in SLICES:
        ref = oracle.build(p, f, t, t + 1)
        sel = (S1 % ns) // nc == t
        assert sel.sum() == len(ref["S1"]), t
        assert np.array_equal(S1[sel], ref["S1"].astype(np.int32)), t
        assert np.array_equal(S2[sel], ref["S2"].astype(np.int32)), t
        assert np.array_equal(val[sel].view(np.uint32), ref["val"].view(np.uint32)), t
        rows = (np.arange(na)[:, None] * ns + t * nc + np.arange(nc)[None, :]).ravel()
        assert np.array_equal(R[rows].view(np.uint32), ref["R"][rows].view(np.uint32)), t
        print(f"[large] slice {t}: {sel.sum()} entries bit-exact", flush=True)
        del sel, ref

    # invariants over every row, one action (one contiguous block of rows) at a time
    starts = np.searchsorted(S1, np.arange(na + 1, dtype=np.int64) * ns)
    for a in range(na):
        lo, hi = starts[a], starts[a + 1]
        s1 = S1[lo:hi].astype(np.int64) - a * ns
        s2 = S2[lo:hi].astype(np.int64)
        v = val[lo:hi]
        cnt = np.rint(v.astype(np.float64) * nr).astype(np.int64)
        assert np.array_equal((cnt.astype(np.float32) / np.float32(nr)).view(np.uint32), v.view(np.uint32))
        first = np.r_[True, s1[1:] != s1[:-1]]
        heads = np.flatnonzero(first)
        assert (np.add.reduceat(cnt, heads) == nr).all(), a
        assert len(heads) == ns - nc  # every state but the last slice has a row
        assert ((s2 // nc) == (s1 // nc) + 1).all(), a
        assert (first[1:] | (s2[1:] > s2[:-1])).all(), a
        del s1, s2, v, cnt, first, heads
    assert (R.reshape(na, ns)[:, (p.nt - 1) * nc:] == 0).all()
    print("[large] invariants hold", flush=True)

    V64, pi64 = oracle.vi_backward_f64(ns, na, nc, S1, S2, val, R)
    err = rel_err(vgpu, V64)
    print(f"[large] max rel err vs float64 {err.max():.3e}", flush=True)
    assert err.max() <= TOL
    for kv, (_, v32, s32) in f32_runs.items():
        e32 = rel_err(v32, V64)
        print(f"[large] float32 values, kernel_variant {kv}: max rel err {e32.max():.3e} "
              f"(mean {e32.mean():.2e}), {s32['sweeps']} sweeps, {s32['sweep_kernel_ms_per_sweep']:.3f} ms/sweep",
              flush=True)
        assert np.isfinite(v32).all() and e32.max() < 1e-3
    Q64 = q_values_f64(ns, na, S1.astype(np.int64), S2.astype(np.int64), val, R, V64)
    policy_ties_ok(pol, pi64, Q64, 4 * TOL)
