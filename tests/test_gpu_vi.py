"""GPU parity of the fused SpMV value iteration (d) against the FP64 oracle."""
import numpy as np
import pytest

import e2e_gpu_dp_amd as e2e
import oracle
from e2e_gpu_dp_amd.configs import small, workload
from helpers import assert_slice_matches, oracle_model, policy_ties_ok, q_values_f64, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-5  # |v - v64| <= 1e-5 * max(|v64|, 1)


def _check_solve(ref, pol, val, stats, tol=TOL):
    ns, na, nc = ref["ns"], ref["na"], ref["nc"]
    V64, pi64 = oracle.vi_backward_f64(ns, na, nc, ref["S1"], ref["S2"], ref["val"], ref["R"])
    err = rel_err(val, V64)
    assert err.max() <= tol, f"max rel err {err.max()} at {err.argmax()}"
    Q64 = q_values_f64(ns, na, ref["S1"], ref["S2"], ref["val"], ref["R"], V64)
    policy_ties_ok(pol, pi64, Q64, 4 * tol)
    # final residual of the GPU values under an FP64 backup
    r64 = oracle.bellman_residual_f64(ns, na, ref["S1"], ref["S2"], ref["val"], ref["R"], val)
    assert r64 <= tol * max(1.0, np.abs(V64).max())
    return err.max()


@pytest.mark.parametrize("lif", [0, 2, 4, 8])
def test_vi_small(gpu, lif):
    w = small(g=9, nt=12, nr=300, nm=2, seed=41)
    f = w.fields()
    ref = oracle_model(w.params, f)
    with e2e.build_model(w.params, f) as m:
        pol, val, st = m.solve(thresh=1e-6, loads_in_flight=lif)
        assert st["converged"] and st["sweeps"] <= w.params.nt + 1
        _check_solve(ref, pol, val, st)


def test_vi_c1_and_sweep_count(gpu):
    w = workload("C1")
    f = w.fields()
    ref = oracle_model(w.params, f)
    _, _, it64, _ = oracle.vi_sweeps_f64(ref["ns"], ref["na"], ref["S1"], ref["S2"], ref["val"],
                                         ref["R"], 1e-6)
    with e2e.build_model(w.params, f) as m:
        pol, val, st = m.solve(thresh=1e-6)
        assert st["converged"]
        assert st["sweeps"] == it64
        assert st["final_residual"] == 0.0
        _check_solve(ref, pol, val, st)
        # eager and graph-replayed sweeps give identical bits
        pol2, val2, st2 = m.solve(thresh=1e-6, use_graph=True, sweeps_per_check=4)
        assert st2["sweeps"] == st["sweeps"]
        assert np.array_equal(val.view(np.uint32), val2.view(np.uint32))
        assert np.array_equal(pol, pol2)


def test_vi_reference_threshold_stops_early(gpu):
    """The drop-in default thresh 0.1 (spvi_2.cu:757) stops at the first sweep with
    residual < 0.1, as the reference loop does (spvi_2.cu:669-698)."""
    w = small(g=9, nt=15, nr=300, nm=2, seed=43)
    f = w.fields()
    ref = oracle_model(w.params, f)
    V64, pi64, it64, res64 = oracle.vi_sweeps_f64(ref["ns"], ref["na"], ref["S1"], ref["S2"],
                                                  ref["val"], ref["R"], 0.1)
    with e2e.build_model(w.params, f) as m:
        pol, val, st = m.solve(thresh=0.1, sweeps_per_check=3)
        assert st["sweeps"] == it64
        assert rel_err(val, V64).max() <= TOL


def test_vi_imported_coo_matches_built(gpu):
    w = small(g=10, nt=8, nr=250, nm=2, seed=47)
    f = w.fields()
    with e2e.build_model(w.params, f) as m:
        S1, S2, val, R = m.export_coo(np.int32)
        pol, v, st = m.solve(thresh=1e-6)
    with e2e.import_coo(w.params.ns, w.params.na, S1, S2, val, R) as mi:
        pol2, v2, st2 = mi.solve(thresh=1e-6)
    assert st2["sweeps"] == st["sweeps"]
    assert np.array_equal(v.view(np.uint32), v2.view(np.uint32))
    assert np.array_equal(pol, pol2)


def test_vi_not_converged_status(gpu):
    w = small(g=8, nt=10, nr=100, nm=1, seed=49)
    with e2e.build_model(w.params, w.fields()) as m:
        pol, val, st = m.solve(thresh=1e-6, max_sweeps=3, sweeps_per_check=1, raise_on_nonconv=False)
        assert st["status"] == 4 and st["sweeps"] == 3 and not st["converged"]


# every sweep kernel with the loads-in-flight settings it supports: 1 chunked +
# compensated float, 2 chunked + fma float, 3 lean streaming, 4 tiled with an LDS value
# band, 5 jagged-diagonal (3-5 in float64). Variant 2 rounds each product, so it gets a
# wider bound on these short horizons.
VARIANTS = [(1, 2), (1, 8), (2, 4), (3, 2), (3, 4), (3, 8), (4, 2), (4, 4), (4, 8), (5, 2), (5, 4), (5, 8),
            (6, 4)]


@pytest.mark.parametrize("variant,lif", VARIANTS)
def test_vi_kernel_variants(gpu, variant, lif):
    w = small(g=11, nt=14, nr=400, nm=2, na=12, seed=53)  # 12 actions: 4 idle lanes per state
    f = w.fields()
    ref = oracle_model(w.params, f)
    with e2e.build_model(w.params, f) as m:
        pol, val, st = m.solve(thresh=1e-6, kernel_variant=variant, loads_in_flight=lif)
        assert st["kernel_variant"] == variant and st["converged"]
        _check_solve(ref, pol, val, st, tol=TOL if variant != 2 else 4 * TOL)


def _random_coo(ns, na, seed, empty_frac=0.2):
    """Random row-sorted COO with empty rows, and dead states whose rows are all empty
    and whose rewards are all -2e6."""
    rng = np.random.default_rng(seed)
    S1, S2, val = [], [], []
    dead = set(rng.choice(ns, size=max(1, ns // 10), replace=False).tolist())
    for a in range(na):
        for s in range(ns):
            if s in dead or rng.random() < empty_frac:
                continue
            k = int(rng.integers(1, 40))
            cols = np.unique(rng.integers(0, ns, size=k))
            p = rng.random(len(cols)).astype(np.float32)
            S1 += [a * ns + s] * len(cols)
            S2 += cols.tolist()
            val += (p / p.sum()).astype(np.float32).tolist()
    R = (-rng.random(na * ns) * 3).astype(np.float32)
    for s in dead:  # every Q of a dead state is below -1e6: value -1e6, policy -1
        R[np.arange(na) * ns + s] = -2e6
    return (np.array(S1, np.int32), np.array(S2, np.int32), np.array(val, np.float32), R)


@pytest.mark.parametrize("ns,na", [(3001, 1), (3001, 3), (3001, 16), (3001, 33), (20011, 3)])
def test_vi_float64_kernels_agree_bitwise(gpu, ns, na):
    """The lean, tiled and jagged sweeps form the same exact float64 products and add
    each row's in its entry order, so their values and policies are bit-identical;
    empty rows and dead states (every Q below -1e6: value -1e6, policy -1, as in
    spvi_2.cu:93-122) included. Columns are uniform over all states, so with 20011
    states the tiles' column bands exceed the LDS window and the tiled kernel gathers
    from global memory."""
    S1, S2, val, R = _random_coo(ns, na, seed=60 + na)
    with e2e.import_coo(ns, na, S1, S2, val, R) as m:
        pa, va, sa = m.solve(thresh=1e-3, max_sweeps=6, kernel_variant=3, raise_on_nonconv=False)
        for b_ in (4, 5):
            pb, vb, sb = m.solve(thresh=1e-3, max_sweeps=6, kernel_variant=b_, raise_on_nonconv=False)
            assert sa["sweeps"] == sb["sweeps"]
            assert np.array_equal(va.view(np.uint32), vb.view(np.uint32)), b_
            assert np.array_equal(pa, pb), b_
        pc, vc, sc = m.solve(thresh=1e-3, max_sweeps=6, kernel_variant=1, raise_on_nonconv=False)
        assert rel_err(vc, va.astype(np.float64)).max() <= TOL
        V64, pi64, it64, _ = oracle.vi_sweeps_f64(ns, na, S1, S2, val, R, 1e-3, 6)
        assert rel_err(va, V64).max() <= TOL
        assert (pi64 == -1).sum() >= ns // 10
        assert (pa.astype(np.int64)[pi64 == -1] == 0xFFFFFFFF).all()
        assert (va[pi64 == -1] == -1e6).all()


def test_vi_c2_accuracy(gpu):
    """Full-size workload (100x100 grid, nt=120, 16 headings, 2500 realizations,
    1.34e8 nnz) with the default kernel: values within 1e-5 relative of the float64
    backward induction, policies equal except at ties, sweep count as in float64."""
    w = workload("C2")
    f = w.fields()
    with e2e.build_model(w.params, f) as m:
        pol, val, st = m.solve(thresh=1e-6, sweeps_per_check=8)
        f32_runs = {kv: m.solve(thresh=1e-6, kernel_variant=kv, raise_on_nonconv=False) for kv in (1, 2)}
        S1, S2, pv, R = m.export_coo(np.int32)
    # the plain build itself, bit for bit against the oracle on three slices (the first,
    # a middle one and one with the late penalty, T > nt - 20)
    for t in (0, 60, 110):
        n = assert_slice_matches(w.params, f, t, S1, S2, pv, R)
        print(f"[c2] slice {t}: {n} entries bit-exact", flush=True)
    ns, na = w.params.ns, w.params.na
    V64, pi64 = oracle.vi_backward_f64(ns, na, w.params.nc, S1, S2, pv, R)
    err = rel_err(val, V64)
    print(f"[c2] sweeps={st['sweeps']} max rel err {err.max():.3e} ms/sweep "
          f"{st['sweep_kernel_ms_per_sweep']:.4f}", flush=True)
    assert st["converged"] and st["sweeps"] == w.params.nt
    assert err.max() <= TOL
    Q64 = q_values_f64(ns, na, S1, S2, pv, R, V64)
    policy_ties_ok(pol, pi64, Q64, 4 * TOL)
    # the float32-value sweeps (1: compensated row sums, 2: plain fma), measured
    for kv, (_, v32, s32) in f32_runs.items():
        e32 = rel_err(v32, V64)
        print(f"[c2] float32 values, kernel_variant {kv}: max rel err {e32.max():.3e} (mean {e32.mean():.2e}), "
              f"{s32['sweeps']} sweeps, {s32['sweep_kernel_ms_per_sweep']:.4f} ms/sweep", flush=True)
        assert np.isfinite(v32).all() and e32.max() < 1e-3


def test_vi_unknown_variant_is_an_error(gpu):
    w = small(g=6, nt=4, nr=50, nm=1, seed=71)
    with e2e.build_model(w.params, w.fields()) as m:
        for bad in ({"kernel_variant": 8}, {"comp_span": 65537}):
            with pytest.raises(e2e.E2EError) as ei:
                m.solve(**bad)
            assert ei.value.code == 1


def test_vi_compressed_needs_a_built_model(gpu):
    S1, S2, val, R = _random_coo(301, 2, seed=77)
    with e2e.import_coo(301, 2, S1, S2, val, R) as m:
        _, _, st = m.solve(thresh=1e-3, max_sweeps=4, raise_on_nonconv=False)
        assert st["kernel_variant"] == 3  # the default for an imported model
        with pytest.raises(e2e.E2EError) as ei:
            m.solve(kernel_variant=6)
        assert ei.value.code == 1


@pytest.mark.parametrize("case,span", [("small", 0), ("small", 24), ("c1", 0), ("c1", 200),
                                       ("wide_flow", 0), ("wide_flow", 64)])
def test_vi_compressed_matches_lean_bitwise(gpu, case, span):
    """The compressed sweep (4-byte entries: count and column offset) decodes every
    probability to the stored float and adds the same float64 products in the same
    order as the lean sweep, so values and policies are bit-identical. A small column
    span limit keeps part of the windows at 8 bytes, so both window kinds mix."""
    if case == "small":
        w = small(g=11, nt=14, nr=400, nm=2, na=12, seed=53)
    elif case == "c1":
        w = workload("C1")
    else:
        w = small(g=30, nt=4, nr=400, nm=2, seed=13)
    f = w.fields()
    if case == "wide_flow":
        from e2e_gpu_dp_amd.fields import make_fields
        p = w.params
        f = make_fields(p.gsize, p.nt, p.nmodes, p.nrzns, 13, F=p.F, mean_speed_frac=9.0, mode_rms_frac=6.0)
    with e2e.build_model(w.params, f) as m:
        pa, va, sa = m.solve(thresh=1e-6, kernel_variant=3)
        pb, vb, sb = m.solve(thresh=1e-6, kernel_variant=6, comp_span=span)
        print(f"[{case} span {span}] windows compressed {sb['compressed_windows']}", flush=True)
        assert sb["kernel_variant"] == 6 and sa["sweeps"] == sb["sweeps"]
        assert np.array_equal(va.view(np.uint32), vb.view(np.uint32))
        assert np.array_equal(pa, pb)
        nwin = -(-m.info["nnz"] // 256)
        if span == 0:
            assert sb["compressed_windows"] == nwin  # every window of a built model fits
        else:
            assert 0 <= sb["compressed_windows"] < nwin
        assert sb["moved_bytes_per_sweep"] < sa["moved_bytes_per_sweep"] or span != 0
        # the default for a built model is the compressed sweep
        pc, vc, sc = m.solve(thresh=1e-6)
        assert sc["kernel_variant"] == 6 and np.array_equal(vc.view(np.uint32), va.view(np.uint32))


@pytest.mark.parametrize("case,span", [("small", 0), ("small", 60), ("c1", 0), ("c1", 150)])
def test_vi_compressed_stream_from_the_build(gpu, case, span):
    """The build writes the compressed entries in the pass that writes the CSR entries.
    They give the same solve, bit for bit, as the 8-byte entries, and keep the same
    windows as the solve's own conversion from the CSR entries (forced by a column-span
    change and back). A small span limit leaves part of the windows at 8 bytes (on
    these models the windows' column spans have medians of 58 and 146)."""
    w = small(g=13, nt=10, nr=300, nm=2, na=12, seed=71) if case == "small" else workload("C1")
    with e2e.build_model(w.params, w.fields(), comp_span=span) as m:
        nwin = -(-m.info["nnz"] // 256)
        pa, va, sa = m.solve(thresh=1e-6, kernel_variant=3)
        pb, vb, sb = m.solve(thresh=1e-6, kernel_variant=6, comp_span=span)   # the build's stream
        print(f"[{case} span {span}] windows {nwin}, kept by the build {sb['compressed_windows']}", flush=True)
        assert sb["kernel_variant"] == 6 and sa["sweeps"] == sb["sweeps"]
        assert np.array_equal(va.view(np.uint32), vb.view(np.uint32)) and np.array_equal(pa, pb)
        if span == 0:
            assert sb["compressed_windows"] == nwin
        else:
            assert 0 < sb["compressed_windows"] < nwin
        m.solve(thresh=1e-6, kernel_variant=6, comp_span=(span or 65536) - 1)
        pc, vc, sc = m.solve(thresh=1e-6, kernel_variant=6, comp_span=span)  # converted again
        assert sc["compressed_windows"] == sb["compressed_windows"]
        assert np.array_equal(vc.view(np.uint32), va.view(np.uint32)) and np.array_equal(pc, pa)
    with pytest.raises(e2e.E2EError) as ei:
        e2e.build_model(w.params, w.fields(), comp_span=65537)
    assert ei.value.code == 1


@pytest.mark.parametrize("case", ["small", "terminal", "obstacles", "c1", "wide_flow"])
def test_vi_backward_pass_equals_converged_sweeps(gpu, case):
    """Transitions of a built model go from each slice to the next, so the sweeps reach
    their fixed point after nt of them; the backward pass (kernel_variant 7: one launch
    per slice, last to first) computes each slice once from its final successor, with the
    same per-row float64 arithmetic, so it returns the same bits as sweeps run to a zero
    residual."""
    from e2e_gpu_dp_amd.configs import Workload
    if case == "small":
        w = small(g=11, nt=14, nr=400, nm=2, na=12, seed=53)
    elif case == "terminal":
        w = small(g=14, nt=20, nr=300, nm=3, seed=57, term_subgrid_size=3)
    elif case == "obstacles":
        b = small(g=16, nt=18, nr=300, nm=2, seed=59, term_subgrid_size=2)
        w = Workload("obs", b.params, b.seed, b.sigmas, obstacles=[(0, 18, 5, 8, 5, 8)])
    elif case == "c1":
        w = workload("C1")
    else:
        w = small(g=30, nt=4, nr=400, nm=2, seed=13)
    f = w.fields()
    if case == "wide_flow":
        from e2e_gpu_dp_amd.fields import make_fields
        p = w.params
        f = make_fields(p.gsize, p.nt, p.nmodes, p.nrzns, 13, F=p.F, mean_speed_frac=9.0, mode_rms_frac=6.0)
    with e2e.build_model(w.params, f) as m:
        for variant in (6, 3):  # compressed and 8-byte entries
            pj, vj, sj = m.solve(thresh=1e-30, kernel_variant=variant, max_sweeps=w.params.nt + 4)
            pb, vb, sb = m.solve(kernel_variant=7)
            print(f"[{case}] sweeps {sj['sweeps']} ({sj['sweep_kernel_ms']:.3f} ms) vs backward "
                  f"{sb['launches']} launches ({sb['sweep_kernel_ms']:.3f} ms)", flush=True)
            assert sj["converged"] and sj["sweeps"] <= w.params.nt + 1
            assert sb["kernel_variant"] == 7 and sb["sweeps"] == 1 and sb["launches"] == w.params.nt
            assert np.array_equal(vj.view(np.uint32), vb.view(np.uint32))
            assert np.array_equal(pj, pb)


@pytest.mark.parametrize("case", ["small", "terminal", "obstacles", "c1"])
def test_vi_active_slice_sweeps_equal_full_sweeps(gpu, case):
    """A built model's lean sweeps skip the slices already at their final values (slice t
    is final after sweep nt - t); values, policies, the sweep count and the residual of
    the stopping sweep equal those of sweeps over every slice, bit for bit, for early
    stops (thresh 0.1, 1e-3), the bench's 1e-6 and a zero residual."""
    from e2e_gpu_dp_amd.configs import Workload
    if case == "small":
        w = small(g=11, nt=14, nr=400, nm=2, na=12, seed=53)
    elif case == "terminal":
        w = small(g=14, nt=20, nr=300, nm=3, seed=57, term_subgrid_size=3)
    elif case == "obstacles":
        b = small(g=16, nt=18, nr=300, nm=2, seed=59, term_subgrid_size=2)
        w = Workload("obs", b.params, b.seed, b.sigmas, obstacles=[(0, 18, 5, 8, 5, 8)])
    else:
        w = workload("C1")
    f = w.fields()
    with e2e.build_model(w.params, f) as m:
        rp = m.export_csr()[0].astype(np.float64)
        nt, nc, na = w.params.nt, w.params.gsize ** 2, m.info["na"]

        def sweep_bytes(T):  # entries, row pointers, R, V read, V write, policy of slices [0, T)
            rows = T * nc * na
            return 8 * rp[rows] + 4 * (rows + 1) + 4 * rows + 12 * T * nc

        for variant in (6, 3):
            for thresh, graph in ((0.1, False), (1e-3, True), (1e-6, False), (1e-30, True)):
                kw = dict(thresh=thresh, kernel_variant=variant, max_sweeps=w.params.nt + 4, use_graph=graph,
                          sweeps_per_check=3, raise_on_nonconv=False)
                pf, vf, sf = m.solve(full_sweeps=True, **kw)
                pa, va, sa = m.solve(**kw)
                print(f"[{case} v{variant} {thresh:g}] sweeps {sa['sweeps']}: {sa['sweep_kernel_ms']:.3f} ms "
                      f"vs full {sf['sweep_kernel_ms']:.3f} ms", flush=True)
                assert sa["status"] == sf["status"] == 0
                assert sa["sweeps"] == sf["sweeps"] and sa["final_residual"] == sf["final_residual"]
                assert np.array_equal(va.view(np.uint32), vf.view(np.uint32))
                assert np.array_equal(pa, pf)
                # the bytes reported: every slice per full sweep, the recomputed ones per active sweep
                n = sa["sweeps"]
                assert sf["bytes_per_sweep"] == pytest.approx(sweep_bytes(nt), rel=1e-12)
                acts = [nt if k <= 2 else max(nt - k + 2, 1) for k in range(1, n + 1)]
                assert sa["bytes_per_sweep"] == pytest.approx(np.mean([sweep_bytes(T) for T in acts]), rel=1e-12)


def test_vi_backward_pass_needs_a_built_model(gpu):
    w = small(g=8, nt=6, nr=100, nm=1, seed=61)
    ref = oracle_model(w.params, w.fields())
    with e2e.import_coo(ref["ns"], ref["na"], ref["S1"], ref["S2"], ref["val"], ref["R"]) as m:
        with pytest.raises(e2e.E2EError):
            m.solve(kernel_variant=7)


def test_vi_work_pool_layouts_agree_bitwise(gpu):
    """The lean sweep's work pool: at 1 and 2 blocks per CU the 60x60 x 40 model's waves
    hold 140 and 70 states, so every long sweep hands the tails of their ranges to the
    per-XCD pools (pieces of different sizes taken by different waves); at 7 blocks per CU
    (20 states per wave) no sweep uses the pool. Each state is still swept once with its
    entries in the same order: values, policies, sweep counts and the stopping residual
    agree bit for bit, over every slice and over the active slices, and stay within 1e-5
    of FP64."""
    w = small(g=60, nt=40, nr=300, nm=3, seed=61)
    f = w.fields()
    with e2e.build_model(w.params, f) as m:
        for full in (True, False):
            runs = {}
            for bpc in (1, 2, 7):
                pol, val, st = m.solve(thresh=1e-6, max_sweeps=w.params.nt + 4, blocks_per_cu=bpc, full_sweeps=full)
                print(f"full={full} bpc={bpc}: blocks {st['blocks']}, sweeps {st['sweeps']}", flush=True)
                assert st["blocks"] == bpc * 256 or st["blocks"] < bpc * 256
                runs[bpc] = (pol, val, st)
            p0, v0, s0 = runs[1]
            for bpc in (2, 7):
                p1, v1, s1 = runs[bpc]
                assert s1["sweeps"] == s0["sweeps"] and s1["final_residual"] == s0["final_residual"], bpc
                assert np.array_equal(v1.view(np.uint32), v0.view(np.uint32)), bpc
                assert np.array_equal(p1, p0), bpc
        ref = oracle_model(w.params, f)
        _check_solve(ref, p0, v0, s0)
