"""Shared helpers for the parity tests (oracle side = checker)."""
from __future__ import annotations

import numpy as np

import oracle


def oracle_model(params, fields, nthreads=0):
    return oracle.build(params, fields, nthreads=nthreads)


def assert_same_model(gpu, ref, what=""):
    """Bit-exact comparison of GPU export (S1, S2, val, R) with the oracle's."""
    S1, S2, val, R = gpu
    assert len(S1) == len(ref["S1"]), f"{what}: nnz {len(S1)} != {len(ref['S1'])}"
    bad = np.flatnonzero((S1 != ref["S1"]) | (S2 != ref["S2"]))
    assert bad.size == 0, (f"{what}: {bad.size} index mismatches, first at {bad[0]}: "
                           f"gpu ({S1[bad[0]]},{S2[bad[0]]}) ref ({ref['S1'][bad[0]]},{ref['S2'][bad[0]]})")
    vb = np.flatnonzero(val.view(np.uint32) != ref["val"].view(np.uint32))
    assert vb.size == 0, f"{what}: {vb.size} prob mismatches, first at {vb[0]}: {val[vb[0]]} vs {ref['val'][vb[0]]}"
    rb = np.flatnonzero(R.view(np.uint32) != ref["R"].view(np.uint32))
    assert rb.size == 0, f"{what}: {rb.size} R mismatches, first at {rb[0]}: {R[rb[0]]} vs {ref['R'][rb[0]]}"


def rel_err(v, v64):
    """|v - v64| / max(|v64|, 1): relative error with a floor of one step cost."""
    return np.abs(v.astype(np.float64) - v64) / np.maximum(np.abs(v64), 1.0)


def q_values_f64(ns, na, S1, S2, val, R, V, discount=1.0):
    """Q64[a, s] = R + discount * P V (V float64), rows a*ns + s."""
    acc = np.bincount(S1, weights=val.astype(np.float64) * V[S2], minlength=na * ns)
    return (R.astype(np.float64) + discount * acc).reshape(na, ns)


def policy_ties_ok(pol_gpu, pi64, Q64, tol):
    """Policies must agree except where the FP64 Q of both choices are within tol."""
    pol = pol_gpu.astype(np.int64)
    pol[pol == 0xFFFFFFFF] = -1
    diff = np.flatnonzero(pol != pi64)
    if diff.size == 0:
        return 0
    s = diff
    a_g = pol[s]
    a_r = pi64[s]
    assert (a_g >= 0).all() and (a_r >= 0).all(), "policy -1 mismatch"
    qg = Q64[a_g, s]
    qr = Q64[a_r, s]
    gap = np.abs(qg - qr) / np.maximum(np.abs(qr), 1.0)
    assert (gap <= tol).all(), f"{(gap > tol).sum()} policy differences beyond ties (max gap {gap.max()})"
    return diff.size


def assert_slice_matches(params, fields, t, S1, S2, val, R):
    """Rows of time slice t in an exported reference-order model (S1, S2, val, R as
    int32/float32 arrays) equal, bit for bit, the oracle's build of that slice alone
    (src/build_model.cu:1142-1165 builds slice by slice, so a slice's rows do not depend
    on the others). Returns the number of entries compared."""
    import oracle
    p = params
    ns, na, nc = p.ns, p.na, p.nc
    ref = oracle.build(p, fields, t, t + 1)
    sel = (S1 % ns) // nc == t
    assert sel.sum() == len(ref["S1"]), t
    assert np.array_equal(S1[sel], ref["S1"].astype(np.int32)), t
    assert np.array_equal(S2[sel], ref["S2"].astype(np.int32)), t
    assert np.array_equal(val[sel].view(np.uint32), ref["val"].view(np.uint32)), t
    rows = (np.arange(na)[:, None] * ns + t * nc + np.arange(nc)[None, :]).ravel()
    assert np.array_equal(R[rows].view(np.uint32), ref["R"][rows].view(np.uint32)), t
    return int(sel.sum())


def goal_params(params, goal):
    """The reference's per-goal model: terminal subgrid of side 1 at cell `goal`
    (src/input_to_build_model.h:33-40)."""
    g = params.gsize
    return params.replace(i_term=int(goal) // g, j_term=int(goal) % g, term_subgrid_size=1)


def assert_close_to_f64(ref, pol, val, tol=1e-5, what=""):
    """Values within tol * max(|v64|, 1) of the FP64 backward induction of the model
    `ref` (an oracle build or a COO dict with S1, S2, val, R, ns, na, nc); policies equal
    except where the FP64 Q values of both choices tie within 4 tol. Returns the max
    relative error."""
    ns, na, nc = ref["ns"], ref["na"], ref["nc"]
    V64, pi64 = oracle.vi_backward_f64(ns, na, nc, ref["S1"], ref["S2"], ref["val"], ref["R"])
    err = rel_err(val, V64)
    assert err.max() <= tol, f"{what}: max rel err {err.max()} at state {err.argmax()}"
    Q64 = q_values_f64(ns, na, ref["S1"], ref["S2"], ref["val"], ref["R"], V64)
    policy_ties_ok(pol, pi64, Q64, 4 * tol)
    return float(err.max())
