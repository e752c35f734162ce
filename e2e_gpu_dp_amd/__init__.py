"""MI355X-native end-to-end GPU dynamic-programming planner (Python mirror of the C ABI).

The reference pipeline is two executables (src/build_model.cu, src/spvi_2.cu) joined by
.npy files. Here the same two stages are library calls on one GPU:

    m = build_model(params, fields)          # build_sparse_transition_model_at_T_at_a,
                                             # all t and a (src/build_model.cu:696-892)
    S1, S2, val, R = m.export_coo()          # save_master_Coos_to_file (:1246-1314)
    policy, value, stats = m.solve(thresh)   # solver_spvi_solve (src/spvi_2.cu:650-726)

Fields are numpy arrays (host) or torch tensors already on the GPU (device pointers are
passed through; nothing is copied back). The HIP library must be built; there is no CPU
fallback.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from ._abi import E2EError, check, lib
from .configs import Params as PyParams

__all__ = ["Model", "Server", "build_model", "import_coo", "c_params", "device_count", "E2EError",
           "pack_params", "actions", "rollout", "rollout_metrics", "generate_environment"]

_PFIELDS = [f[0] for f in _abi.Params._fields_]


def c_params(p) -> _abi.Params:
    """Accepts a configs.Params, a dict, or an _abi.Params."""
    if isinstance(p, _abi.Params):
        return p
    d = p if isinstance(p, dict) else p.__dict__
    cp = _abi.Params()
    for k in _PFIELDS:
        if k in d:
            setattr(cp, k, d[k])
    return cp


def device_count() -> int:
    return int(lib().e2e_device_count())


def cached_device_bytes() -> int:
    """Device memory kept mapped for reuse after models closed."""
    return int(lib().e2e_cached_device_bytes())


def release_cached_memory() -> None:
    """Returns the device memory kept for reuse to the HIP runtime."""
    lib().e2e_release_cached_memory()


def pack_params(p) -> np.ndarray:
    """The reference's 32-float prob_params vector (src/build_model.cu:1032-1055)."""
    out = np.zeros(32, np.float32)
    check(lib().e2e_params_pack(C.byref(c_params(p)), out.ctypes.data))
    return out


def actions(p):
    cp = c_params(p)
    na = cp.num_ac_speeds * cp.num_ac_angles
    sp = np.zeros(na, np.float32)
    an = np.zeros(na, np.float32)
    check(lib().e2e_actions(C.byref(cp), sp.ctypes.data, an.ctypes.data))
    return sp, an


def _ptr(x, keep, dtype):
    if x is None:
        return None, False
    if hasattr(x, "data_ptr") and getattr(x, "is_cuda", False):
        assert x.is_contiguous(), "device field tensors must be contiguous"
        keep.append(x)
        return x.data_ptr(), True
    a = np.ascontiguousarray(x, dtype=dtype)
    keep.append(a)
    return a.ctypes.data, False


class Model:
    """A device-resident transition model (state-major CSR + R) with its solve buffers."""

    def __init__(self, handle: int, params=None):
        self._h = C.c_void_p(handle)
        self.params = params
        self.n_alpha = 0  # alpha columns built with the model

    # -- lifetime
    def close(self):
        if self._h:
            lib().e2e_model_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self) -> int:
        return self._h.value

    # -- info
    @property
    def info(self) -> dict:
        i = _abi.ModelInfo()
        check(lib().e2e_model_info_get(self._h, C.byref(i)))
        return {k: getattr(i, k) for k, _ in i._fields_}

    # -- export
    def export_coo(self, index_dtype=np.int64):
        """Reference on-disk order (rows a*Ns+s sorted, S2 ascending), R of Na*Ns."""
        inf = self.info
        nnz, rows = inf["nnz"], inf["ns"] * inf["na"]
        it = np.dtype(index_dtype)
        S1 = np.empty(nnz, it)
        S2 = np.empty(nnz, it)
        val = np.empty(nnz, np.float32)
        R = np.empty(rows, np.float32)
        fn = lib().e2e_model_export_coo if it == np.int64 else lib().e2e_model_export_coo_i32
        check(fn(self._h, S1.ctypes.data, S2.ctypes.data, val.ctypes.data, R.ctypes.data))
        return S1, S2, val, R

    def export_csr(self):
        inf = self.info
        nnz, rows = inf["nnz"], inf["ns"] * inf["na"]
        rp = np.empty(rows + 1, np.uint32)
        col = np.empty(nnz, np.int32)
        val = np.empty(nnz, np.float32)
        R = np.empty(rows, np.float32)
        check(lib().e2e_model_export_csr(self._h, rp.ctypes.data, col.ctypes.data, val.ctypes.data,
                                         R.ctypes.data))
        return rp, col, val, R

    # -- solve
    def solve(self, thresh: float = 0.1, discount: float = 1.0, max_sweeps: int = 0,
              sweeps_per_check: int = 0, use_graph: bool = False, loads_in_flight: int = 0,
              timeout_ms: float = 0.0, raise_on_nonconv: bool = True, kernel_variant: int = 0,
              blocks_per_cu: int = 0, comp_span: int = 0, full_sweeps: bool = False,
              noalloc: bool = False):
        """Sparse VI. Returns (policy uint32[Ns], value float32[Ns], stats dict).
        full_sweeps: recompute every slice on every sweep (by default a built model's
        sweeps skip the slices that already hold their final values; same results).
        noalloc: run e2e_vi_solve_noalloc, which allocates no device memory (the same
        results; at most 2,097,152 states, no graph)."""
        ns = self.info["ns"]
        o = _abi.SolveOpts(thresh, discount, max_sweeps, sweeps_per_check, int(use_graph),
                           loads_in_flight, timeout_ms, kernel_variant, blocks_per_cu, comp_span,
                           int(full_sweeps))
        st = _abi.SolveStats()
        pol = np.empty(ns, np.uint32)
        val = np.empty(ns, np.float32)
        fn = lib().e2e_vi_solve_noalloc if noalloc else lib().e2e_vi_solve
        code = fn(self._h, C.byref(o), pol.ctypes.data, val.ctypes.data, C.byref(st))
        stats = {k: getattr(st, k) for k, _ in st._fields_}
        stats["status"] = code
        if code != 0 and (raise_on_nonconv or code not in (4, 5)):
            check(code)
        return pol, val, stats


    def solve_goals(self, goal_cells, thresh: float = 0.1, discount: float = 1.0,
                    max_sweeps: int = 0, sweeps_per_check: int = 0, variant: int = -1,
                    use_graph: bool = False, timeout_ms: float = 0.0, raise_on_nonconv: bool = True,
                    full_sweeps: bool = False):
        """Batched multi-goal VI (CSR x dense-block SpMM). Returns (policy uint32[k, Ns],
        value float32[k, Ns], per-column sweeps int32[k], stats dict). full_sweeps: every
        sweep recomputes every slice (same results)."""
        goals = np.ascontiguousarray(goal_cells, np.int32).ravel()
        k = len(goals)
        ns = self.info["ns"]
        o = _abi.SolveOpts(thresh, discount, max_sweeps, sweeps_per_check, 0, 0, timeout_ms,
                           full_sweeps=int(full_sweeps))
        bo = _abi.BatchOpts(variant, int(use_graph))
        st = _abi.BatchStats()
        pol = np.empty((k, ns), np.uint32)
        val = np.empty((k, ns), np.float32)
        sw = np.empty(k, np.int32)
        code = lib().e2e_vi_solve_goals(self._h, goals.ctypes.data, k, C.byref(o), C.byref(bo),
                                        pol.ctypes.data, val.ctypes.data, sw.ctypes.data, C.byref(st))
        stats = {kk: getattr(st, kk) for kk, _ in st._fields_}
        stats["status"] = code
        if code != 0 and (raise_on_nonconv or code != 4):
            check(code)
        return pol, val, sw, stats

    def export_R_alpha(self, j: int) -> np.ndarray:
        """R of alpha column j in the reference order (float32[na*ns])."""
        R = np.empty(self.info["na"] * self.info["ns"], np.float32)
        check(lib().e2e_model_export_R_alpha(self._h, j, R.ctypes.data))
        return R

    def solve_alphas(self, thresh: float = 0.1, discount: float = 1.0, max_sweeps: int = 0,
                     sweeps_per_check: int = 0, variant: int = -1, use_graph: bool = False,
                     timeout_ms: float = 0.0, raise_on_nonconv: bool = True, full_sweeps: bool = False):
        """Alpha-sweep VI over the model's alpha columns (one SpMM per sweep). Returns
        (policy uint32[n_alpha, Ns], value float32[n_alpha, Ns], per-column sweeps, stats).
        full_sweeps: every sweep recomputes every slice (same results)."""
        k = self.n_alpha
        if k <= 0:
            raise ValueError("the model was built without alpha columns")
        ns = self.info["ns"]
        o = _abi.SolveOpts(thresh, discount, max_sweeps, sweeps_per_check, 0, 0, timeout_ms,
                           full_sweeps=int(full_sweeps))
        bo = _abi.BatchOpts(variant, int(use_graph))
        st = _abi.BatchStats()
        pol = np.empty((k, ns), np.uint32)
        val = np.empty((k, ns), np.float32)
        sw = np.empty(k, np.int32)
        code = lib().e2e_vi_solve_alphas(self._h, C.byref(o), C.byref(bo), pol.ctypes.data, val.ctypes.data,
                                         sw.ctypes.data, C.byref(st))
        stats = {kk: getattr(st, kk) for kk, _ in st._fields_}
        stats["status"] = code
        if code != 0 and (raise_on_nonconv or code != 4):
            check(code)
        return pol, val, sw, stats


def build_model(params, fields: dict, keep_rsum: bool = False, slot_k: int = 0,
                radix_bits: int = 0, slices_per_launch: int = 0, block_threads: int = 0,
                stream: int = 0, count_kernel: int = 0, paired_rounds: int = 0, alphas=None,
                lane_moves: int = 0, comp_span: int = 0) -> Model:
    """Monte Carlo build on the GPU. `fields` holds u, v, ui, vi, yi and optionally s,
    mask (numpy arrays, or contiguous torch CUDA tensors). `alphas` (reward method 4):
    also build the reward vector of each of these alpha values (float32, at most 64),
    for export_R_alpha and solve_alphas. `lane_moves` 1: move every lane even where a
    cell can be counted at once (the same model either way). `comp_span`: column span
    limit of the compressed entry stream the build writes (0: 65536)."""
    keep = []
    ptrs = {}
    on_dev = None
    for k, dt in (("u", np.float32), ("v", np.float32), ("ui", np.float32), ("vi", np.float32),
                  ("yi", np.float32), ("s", np.float32), ("mask", np.int32)):
        p, dev = _ptr(fields.get(k), keep, dt)
        ptrs[k] = p
        if p is not None:
            if on_dev is None:
                on_dev = dev
            elif on_dev != dev:
                raise ValueError("fields must be all on the host or all on the device")
    f = _abi.Fields(**ptrs)
    al = None if alphas is None else np.ascontiguousarray(alphas, np.float32).ravel()
    o = _abi.BuildOpts(int(bool(on_dev)), int(keep_rsum), slot_k, radix_bits, slices_per_launch,
                       block_threads, stream or None, count_kernel, paired_rounds,
                       0 if al is None else len(al), lane_moves, None if al is None else al.ctypes.data,
                       comp_span)
    h = C.c_void_p()
    check(lib().e2e_model_build(C.byref(c_params(params)), C.byref(f), C.byref(o), C.byref(h)))
    m = Model(h.value, params)
    m.n_alpha = 0 if al is None else len(al)
    return m


def import_coo(ns: int, na: int, S1, S2, val, R, stream: int = 0) -> Model:
    """Model from the reference COO layout (what src/spvi_2.cu:393-504 loads)."""
    S1 = np.ascontiguousarray(S1, np.int32).ravel()
    S2 = np.ascontiguousarray(S2, np.int32).ravel()
    val = np.ascontiguousarray(val, np.float32).ravel()
    R = np.ascontiguousarray(R, np.float32).ravel()
    h = C.c_void_p()
    check(lib().e2e_model_import_coo(ns, na, len(S1), S1.ctypes.data, S2.ctypes.data, val.ctypes.data,
                                     R.ctypes.data, stream or None, C.byref(h)))
    return Model(h.value)


def rollout(params, fields: dict, policy, starts, interp: bool = True, cf: float = 1.0, cr: float = 1.0,
            want_xy: bool = False, stream: int = 0):
    """Policy rollouts on the GPU: plot_exact_trajectory_set_DP's realization loop
    (Post_processing/plot_e2eGPU_policy.py:436-516) for every start cell in `starts`
    ((i, j) pairs) and every realization. `policy` is the solver's policy (nt*g*g, e.g.
    policy.npy). Returns a dict: success (int8 [n]: 1 reached, 0 failed), travel (int32),
    energy (float64 [n, 3]: energy_cons, energy_col, net_energy_cons), xy (float64
    [n, nt+1, 2] or None), stats; trajectory n = k*nrzns + r."""
    keep = []
    ptrs = {}
    on_dev = None
    for k, dt in (("u", np.float32), ("v", np.float32), ("ui", np.float32), ("vi", np.float32),
                  ("yi", np.float32), ("s", np.float32), ("mask", np.int32)):
        ptr, dev = _ptr(fields.get(k), keep, dt)
        ptrs[k] = ptr
        if ptr is not None:
            if on_dev is None:
                on_dev = dev
            elif on_dev != dev:
                raise ValueError("fields must be all on the host or all on the device")
    f = _abi.Fields(**ptrs)
    cp = c_params(params)
    pol = np.ascontiguousarray(np.asarray(policy).ravel().astype(np.int64).astype(np.int32))
    if len(pol) != cp.nt * cp.gsize * cp.gsize:
        raise ValueError(f"policy has {len(pol)} entries, expected nt*gsize^2 = {cp.nt * cp.gsize ** 2}")
    st = np.ascontiguousarray(np.asarray(starts, np.int32).reshape(-1, 2))
    n = len(st) * cp.nrzns
    succ = np.empty(n, np.int8)
    trav = np.empty(n, np.int32)
    en = np.empty((n, 3), np.float64)
    xy = np.empty((n, cp.nt + 1, 2), np.float64) if want_xy else None
    o = _abi.RolloutOpts(int(interp), int(bool(on_dev)), float(cf), float(cr), stream or None)
    stats = _abi.RolloutStats()
    check(lib().e2e_rollout(C.byref(cp), C.byref(f), pol.ctypes.data, len(st), st.ctypes.data, C.byref(o),
                            succ.ctypes.data, trav.ctypes.data, en.ctypes.data,
                            xy.ctypes.data if want_xy else None, C.byref(stats)))
    return {"success": succ, "travel": trav, "energy": en, "xy": xy,
            "stats": {k: getattr(stats, k) for k, _ in stats._fields_}}


def rollout_metrics(r: dict) -> dict:
    """get_metrics (Post_processing/plot_e2eGPU_policy.py:776-797) and the statistics
    write_log puts in log.csv (:799-824): success rate, then min / max / mean of travel
    time, required energy, collected energy and net energy over the successful
    trajectories."""
    ok = r["success"] == 1
    out = {"success_rate": float(np.count_nonzero(ok)) / len(ok)}
    for name, col in (("time", r["travel"].astype(np.float64)), ("reqd_energy", r["energy"][:, 0]),
                      ("coll_energy", r["energy"][:, 1]), ("net_energy", r["energy"][:, 2])):
        v = col[ok]
        out[f"{name}_min"] = float(v.min()) if len(v) else float("nan")
        out[f"{name}_max"] = float(v.max()) if len(v) else float("nan")
        out[f"{name}_mean"] = float(v.mean()) if len(v) else float("nan")
    return out


def generate_environment(gsize: int = 25, nt: int = 10, n_wsamples: int = 1000, n_modes: int = 3,
                         dt: float = None, dxy: float = None, A: float = 0.5, eps: float = 0.1,
                         wx: float = np.pi, wy: float = 0.5 * np.pi, w_range=(np.pi / 10, 8 * np.pi / 10),
                         A_sc: float = 3.0, obstacles=None, phi: bool = False):
    """The double-gyre environment of
    src/data_input/build_environment_DG_with_westward_clouds.py on the GPU, at refinement
    degree 1 (defaults: the commented parameter block :424-437, dt = 10/nt, dxy =
    2/gsize). obstacles: None, "static", or ("dynamic", upper_row, left_col, width,
    (speed_x, speed_y)). Returns a dict of float32 / int32 arrays under the builder's
    field names (u, v, ui, vi, yi, s, mask; with phi: phi, phi_i, phi_yi), the singular
    values, and stats."""
    dt = 10.0 / nt if dt is None else dt
    dxy = 2.0 / gsize if dxy is None else dxy
    mode, ru, cl, wd, sp = 0, 0, 0, 0, (0.0, 0.0)
    if obstacles == "static":
        mode = 1
    elif obstacles is not None:
        _, ru, cl, wd, sp = obstacles
        mode = 2
    ep = _abi.EnvParams(gsize, nt, n_wsamples, n_modes, dt, dxy, A, eps, wx, wy, w_range[0], w_range[1], A_sc,
                        mode, ru, cl, wd, sp[0], sp[1])
    g, k, nw = gsize, n_modes, n_wsamples
    out = {"u": np.empty((nt, g, g), np.float32), "v": np.empty((nt, g, g), np.float32),
           "ui": np.empty((nt, k, g, g), np.float32), "vi": np.empty((nt, k, g, g), np.float32),
           "yi": np.empty((nt, nw, k), np.float32), "s": np.empty((nt, g, g), np.float32),
           "mask": np.empty((nt, g, g), np.int32), "singular_values": np.empty((nt, k), np.float64)}
    if phi:
        out.update(phi=np.empty((nt, g, g), np.float32), phi_i=np.empty((nt, k, g, g), np.float32),
                   phi_yi=np.empty((nt, nw, k), np.float32))
    eo = _abi.EnvOut(**{name: (arr.ctypes.data if name in out else None)
                        for name, arr in [(n, out.get(n)) for n, _ in _abi.EnvOut._fields_]})
    st = _abi.EnvStats()
    check(lib().e2e_env_generate(C.byref(ep), C.byref(eo), C.byref(st)))
    out["stats"] = {kk: getattr(st, kk) for kk, _ in st._fields_}
    return out


REQUEST_STATUS = {0: "queued", 1: "running", 2: "done", 3: "rejected", 4: "failed"}


class Server:
    """Continuous-batching planning server (C++ scheduler + builder threads in libe2edp)."""

    def __init__(self, hbm_budget_bytes: int = 0, max_cols: int = 16, sweeps_per_step: int = 4,
                 keep_full_results: bool = False, serial: bool = False, thresh: float = 1e-6,
                 exact_pass: bool = False, max_models: int = 0):
        o = _abi.ServerOpts(hbm_budget_bytes, max_cols, sweeps_per_step, int(keep_full_results),
                            int(serial), thresh, int(exact_pass), max_models)
        h = C.c_void_p()
        check(lib().e2e_server_create(C.byref(o), C.byref(h)))
        self._h = h
        self._ns = {}

    def add_env(self, env_id: int, params, fields: dict):
        keep = []
        ptrs = {k: _ptr(fields.get(k), keep, np.int32 if k == "mask" else np.float32)[0]
                for k in ("u", "v", "ui", "vi", "yi", "s", "mask")}
        check(lib().e2e_server_add_env(self._h, env_id, C.byref(c_params(params)),
                                       C.byref(_abi.Fields(**ptrs))))
        cp = c_params(params)
        self._ns[env_id] = cp.gsize * cp.gsize * cp.nt

    def submit(self, env_id: int, goal_cell: int, start_cell: int) -> int:
        rid = C.c_uint64()
        check(lib().e2e_server_submit(self._h, env_id, int(goal_cell), int(start_cell), C.byref(rid)))
        return rid.value

    @staticmethod
    def _res(r):
        d = {k: getattr(r, k) for k, _ in r._fields_}
        d["status_name"] = REQUEST_STATUS.get(d["status"], "?")
        return d

    def poll(self, req_id: int) -> dict:
        r = _abi.RequestResult()
        check(lib().e2e_server_poll(self._h, req_id, C.byref(r)))
        return self._res(r)

    def wait(self, req_id: int, timeout_ms: float = 0.0) -> dict:
        r = _abi.RequestResult()
        check(lib().e2e_server_wait(self._h, req_id, timeout_ms, C.byref(r)))
        return self._res(r)

    def columns(self, req_id: int, env_id: int):
        ns = self._ns[env_id]
        pol = np.empty(ns, np.uint32)
        val = np.empty(ns, np.float32)
        check(lib().e2e_server_result_columns(self._h, req_id, pol.ctypes.data, val.ctypes.data))
        return pol, val

    def drain(self, timeout_ms: float = 0.0):
        check(lib().e2e_server_drain(self._h, timeout_ms))

    def stats(self) -> dict:
        st = _abi.ServerStats()
        check(lib().e2e_server_stats_get(self._h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def close(self):
        if self._h:
            lib().e2e_server_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
