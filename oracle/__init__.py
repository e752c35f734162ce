"""Python handle on the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py. The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
SO = HERE / "liboracle.so"

_lib = None


def compile_lib(force: bool = False):
    """Compile liboracle.so from oracle.c (make in this folder; force: rebuild even when
    up to date)."""
    r = subprocess.run(["make", "-C", str(HERE), "-s"] + (["-B"] if force else []), capture_output=True,
                       text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")


def lib():
    global _lib
    if _lib is None:
        if not SO.exists():
            compile_lib()
        L = C.CDLL(str(SO))
        VP = C.c_void_p
        i64 = C.c_int64
        L.orc_actions.argtypes = [VP, VP, VP, VP, VP]
        L.orc_grid.argtypes = [VP, VP, VP]
        L.orc_lane.argtypes = [VP, VP, C.c_int, C.c_int, C.c_int, C.c_int, VP, VP]
        L.orc_lanes.argtypes = [VP, VP, C.c_int, C.c_int, VP, VP]
        L.orc_build.argtypes = [VP, VP, C.c_int, C.c_int, C.c_int]
        L.orc_build.restype = VP
        for n in ("orc_model_nnz", "orc_model_out_of_window", "orc_model_max_row_nnz"):
            getattr(L, n).argtypes = [VP]
            getattr(L, n).restype = i64
        L.orc_model_copy.argtypes = [VP, VP, VP, VP, VP, VP]
        L.orc_model_free.argtypes = [VP]
        L.orc_vi_backward_f64.argtypes = [i64, i64, i64, i64, VP, VP, VP, VP, C.c_double, VP, VP]
        L.orc_vi_sweeps_f64.argtypes = [i64, i64, i64, VP, VP, VP, VP, C.c_double, C.c_double,
                                        C.c_int, VP, VP, VP]
        L.orc_vi_sweeps_f32.argtypes = [i64, i64, i64, VP, VP, VP, VP, C.c_float, C.c_float,
                                        C.c_int, C.c_int, VP, VP, VP]
        L.orc_bellman_residual_f64.argtypes = [i64, i64, i64, VP, VP, VP, VP, C.c_double, VP]
        L.orc_bellman_residual_f64.restype = C.c_double
        L.orc_rollout.argtypes = [VP, VP, VP, C.c_int, VP, C.c_int, C.c_double, C.c_double, C.c_int,
                                  VP, VP, VP, VP]
        L.orc_rollout.restype = C.c_int
        L.orc_rollout_actions.argtypes = [VP, VP, VP, VP, VP]
        L.orc_set_mutation.argtypes = [C.c_int]
        L.orc_max_threads.restype = C.c_int
        _lib = L
    return _lib


def _cp(params):
    from e2e_gpu_dp_amd import c_params
    return c_params(params)


class _FieldHold:
    def __init__(self, fields):
        from e2e_gpu_dp_amd._abi import Fields
        self.keep = []
        d = {}
        for k, dt in (("u", np.float32), ("v", np.float32), ("ui", np.float32), ("vi", np.float32),
                      ("yi", np.float32), ("s", np.float32), ("mask", np.int32)):
            x = fields.get(k)
            if x is None:
                d[k] = None
                continue
            a = np.ascontiguousarray(x, dtype=dt)
            self.keep.append(a)
            d[k] = a.ctypes.data
        self.f = Fields(**d)


def set_mutation(code: int):
    lib().orc_set_mutation(code)


def max_threads() -> int:
    return lib().orc_max_threads()


def actions(params):
    p = _cp(params)
    na = p.num_ac_speeds * p.num_ac_angles
    out = [np.zeros(na, np.float32) for _ in range(4)]
    lib().orc_actions(C.byref(p), *[o.ctypes.data for o in out])
    return out  # speed, angle, F*cos, F*sin


def grid(params):
    p = _cp(params)
    xs = np.zeros(p.gsize, np.float32)
    ys = np.zeros(p.gsize, np.float32)
    lib().orc_grid(C.byref(p), xs.ctypes.data, ys.ctypes.data)
    return xs, ys


def lane(params, fields, t, a, sp, rzn):
    p = _cp(params)
    h = _FieldHold(fields)
    sp2 = C.c_int32()
    r = C.c_float()
    rc = lib().orc_lane(C.byref(p), C.byref(h.f), t, a, sp, rzn, C.byref(sp2), C.byref(r))
    if rc:
        raise ValueError("non-finite velocity sample")
    return sp2.value, r.value


def lanes(params, fields, t, a):
    """All lanes of (t, a): sp2[sp, rzn], r[sp, rzn]."""
    p = _cp(params)
    h = _FieldHold(fields)
    nc = p.gsize * p.gsize
    sp2 = np.zeros((nc, p.nrzns), np.int32)
    r = np.zeros((nc, p.nrzns), np.float32)
    rc = lib().orc_lanes(C.byref(p), C.byref(h.f), t, a, sp2.ctypes.data, r.ctypes.data)
    if rc:
        raise ValueError("non-finite velocity sample")
    return sp2, r


def build(params, fields, t_begin=0, t_end=None, nthreads=0):
    """Reference-order model: dict with S1, S2 (int64), val, R, Rsum (float32), oow,
    max_row_nnz."""
    p = _cp(params)
    if t_end is None:
        t_end = p.nt - 1
    h = _FieldHold(fields)
    m = lib().orc_build(C.byref(p), C.byref(h.f), t_begin, t_end, nthreads)
    if not m:
        raise ValueError("oracle build failed (bad params or non-finite field)")
    try:
        nnz = lib().orc_model_nnz(m)
        na = p.num_ac_speeds * p.num_ac_angles
        ns = p.gsize * p.gsize * p.nt
        out = {
            "S1": np.empty(nnz, np.int64), "S2": np.empty(nnz, np.int64),
            "val": np.empty(nnz, np.float32), "R": np.empty(na * ns, np.float32),
            "Rsum": np.empty(na * ns, np.float32),
        }
        lib().orc_model_copy(m, out["S1"].ctypes.data, out["S2"].ctypes.data, out["val"].ctypes.data,
                             out["R"].ctypes.data, out["Rsum"].ctypes.data)
        out["oow"] = lib().orc_model_out_of_window(m)
        out["max_row_nnz"] = lib().orc_model_max_row_nnz(m)
        out["ns"], out["na"], out["nc"] = ns, na, p.gsize * p.gsize
        return out
    finally:
        lib().orc_model_free(m)


def rollout_actions(params):
    """speed, angle, speed*cos(angle), speed*sin(angle) of the post-processing's action
    table (float64)."""
    p = _cp(params)
    na = p.num_ac_speeds * p.num_ac_angles
    out = [np.zeros(na, np.float64) for _ in range(4)]
    lib().orc_rollout_actions(C.byref(p), *[o.ctypes.data for o in out])
    return out


def rollout(params, fields, policy, starts, interp=True, Cf=1.0, Cr=1.0, nthreads=0, want_xy=False):
    """Trajectories n = k*nrzns + r from start cells starts[k] = (i, j): dict with success
    (int8: 1, 0, -1 rejected interpolation), travel (int32), energy (float64 [n, 3]:
    energy_cons, energy_col, net_energy_cons), xy (float64 [n, nt+1, 2] or None), bad."""
    p = _cp(params)
    h = _FieldHold(fields)
    pol = np.ascontiguousarray(np.asarray(policy).ravel().astype(np.int64).astype(np.int32))
    st = np.ascontiguousarray(np.asarray(starts, np.int32).reshape(-1, 2))
    n = len(st) * p.nrzns
    out = {"success": np.empty(n, np.int8), "travel": np.empty(n, np.int32),
           "energy": np.empty((n, 3), np.float64),
           "xy": np.empty((n, p.nt + 1, 2), np.float64) if want_xy else None}
    bad = lib().orc_rollout(C.byref(p), C.byref(h.f), pol.ctypes.data, len(st), st.ctypes.data, int(interp),
                            float(Cf), float(Cr), nthreads, out["success"].ctypes.data,
                            out["travel"].ctypes.data, out["energy"].ctypes.data,
                            out["xy"].ctypes.data if want_xy else None)
    if bad < 0:
        raise ValueError("oracle rollout: bad parameters")
    out["bad"] = bad
    return out


def _coo(S1, S2, val, R):
    return (np.ascontiguousarray(S1, np.int64), np.ascontiguousarray(S2, np.int64),
            np.ascontiguousarray(val, np.float32), np.ascontiguousarray(R, np.float32))


def vi_backward_f64(ns, na, nc, S1, S2, val, R, discount=1.0):
    S1, S2, val, R = _coo(S1, S2, val, R)
    V = np.zeros(ns, np.float64)
    pi = np.zeros(ns, np.int32)
    rc = lib().orc_vi_backward_f64(ns, na, nc, len(S1), S1.ctypes.data, S2.ctypes.data, val.ctypes.data,
                                   R.ctypes.data, discount, V.ctypes.data, pi.ctypes.data)
    if rc:
        raise ValueError("model is not a t -> t+1 DAG")
    return V, pi


def vi_sweeps_f64(ns, na, S1, S2, val, R, thresh, max_sweeps=100000, discount=1.0):
    S1, S2, val, R = _coo(S1, S2, val, R)
    V = np.zeros(ns, np.float64)
    pi = np.zeros(ns, np.int32)
    res = C.c_double()
    it = lib().orc_vi_sweeps_f64(ns, na, len(S1), S1.ctypes.data, S2.ctypes.data, val.ctypes.data,
                                 R.ctypes.data, discount, thresh, max_sweeps, V.ctypes.data,
                                 pi.ctypes.data, C.byref(res))
    return V, pi, it, res.value


def vi_sweeps_f32(ns, na, S1, S2, val, R, thresh, max_sweeps=100000, discount=1.0, nthreads=0):
    S1, S2, val, R = _coo(S1, S2, val, R)
    V = np.zeros(ns, np.float32)
    pi = np.zeros(ns, np.int32)
    res = C.c_float()
    it = lib().orc_vi_sweeps_f32(ns, na, len(S1), S1.ctypes.data, S2.ctypes.data, val.ctypes.data,
                                 R.ctypes.data, discount, thresh, max_sweeps, nthreads, V.ctypes.data,
                                 pi.ctypes.data, C.byref(res))
    return V, pi, it, res.value


def bellman_residual_f64(ns, na, S1, S2, val, R, V, discount=1.0):
    S1, S2, val, R = _coo(S1, S2, val, R)
    V = np.ascontiguousarray(V, np.float32)
    return lib().orc_bellman_residual_f64(ns, na, len(S1), S1.ctypes.data, S2.ctypes.data,
                                          val.ctypes.data, R.ctypes.data, discount, V.ctypes.data)
