"""GPU checks of e2e_vi_solve_noalloc, the solve that allocates no device memory: the
same values, policies, sweep counts and residuals as e2e_vi_solve, bit for bit, on
consecutive calls over models of decreasing size (its fixed device array keeps the
previous call's contents), plus its limits and refusals."""
import numpy as np
import pytest

import e2e_gpu_dp_amd as e2e
from e2e_gpu_dp_amd.configs import small
from helpers import oracle_model, rel_err

import oracle

pytestmark = pytest.mark.gpu

E2E_ERR_INVALID, E2E_ERR_CAPACITY = 1, 7
MAX_STATES = 1 << 21


def _same(a, b):
    pa, va, sa = a
    pb, vb, sb = b
    assert np.array_equal(va.view(np.uint32), vb.view(np.uint32))
    assert np.array_equal(pa, pb)
    assert sa["sweeps"] == sb["sweeps"] and sa["converged"] == sb["converged"]
    assert np.float32(sa["final_residual"]).view(np.uint32) == np.float32(sb["final_residual"]).view(np.uint32)


@pytest.mark.parametrize("thresh", [1e-6, 0.1])
def test_noalloc_equals_solve_on_consecutive_models(gpu, thresh):
    # the first model is larger than the second: the second call runs over what the
    # first left in the array
    cases = [small(g=30, nt=40, nr=400, nm=3, seed=61), small(g=9, nt=12, nr=300, nm=2, seed=62)]
    for w in cases:
        p = w.params
        print(f"noalloc case g={p.gsize} nt={p.nt} thresh={thresh}", flush=True)
        f = w.fields()
        with e2e.build_model(p, f) as m:
            bytes0 = m.info["device_bytes"]
            cached0 = e2e.cached_device_bytes()
            na = m.solve(thresh=thresh, noalloc=True)
            # no buffer of the model's own solve was made, and the pool was not touched
            assert m.info["device_bytes"] == bytes0
            assert e2e.cached_device_bytes() == cached0
            assert na[2]["kernel_variant"] == 6   # the build made the compressed entries
            ref = m.solve(thresh=thresh)           # variant 6, allocates the model's buffers
            assert ref[2]["kernel_variant"] == 6
            _same(na, ref)
            # over the 8-byte entries: same bits
            na3 = m.solve(thresh=thresh, noalloc=True, kernel_variant=3)
            assert na3[2]["kernel_variant"] == 3
            _same(na3, ref)
            _same(m.solve(thresh=thresh, noalloc=True), ref)
            # every state and slice on every sweep: same bits again
            _same(m.solve(thresh=thresh, noalloc=True, full_sweeps=True), ref)
        if thresh == 1e-6:
            r = oracle_model(p, f)
            V64, _ = oracle.vi_backward_f64(r["ns"], r["na"], r["nc"], r["S1"], r["S2"], r["val"], r["R"])
            assert rel_err(na[1], V64).max() <= 1e-5


def test_noalloc_imported_model(gpu):
    """An imported model (no slice structure): every state on every sweep, the 8-byte
    entries; the same bits as e2e_vi_solve's default for imported models."""
    w = small(g=11, nt=9, nr=200, nm=2, seed=63)
    r = oracle_model(w.params, w.fields())
    with e2e.import_coo(r["ns"], r["na"], r["S1"], r["S2"], r["val"], r["R"]) as m:
        ref = m.solve(thresh=1e-6)
        na = m.solve(thresh=1e-6, noalloc=True)
        assert ref[2]["kernel_variant"] == 3 and na[2]["kernel_variant"] == 3
        _same(na, ref)
        # no compressed entries: an imported model has no slice structure
        assert _code(lambda: m.solve(thresh=1e-6, noalloc=True, kernel_variant=6)) == E2E_ERR_INVALID


def _code(fn, msg=None):
    with pytest.raises(e2e.E2EError) as ei:
        fn()
    if msg is not None:
        assert msg in str(ei.value), str(ei.value)
    return ei.value.code


def test_noalloc_refusals(gpu):
    w = small(g=9, nt=8, nr=100, nm=1, seed=64)
    with e2e.build_model(w.params, w.fields()) as m:
        assert _code(lambda: m.solve(thresh=1e-6, noalloc=True, use_graph=True)) == E2E_ERR_INVALID
        assert _code(lambda: m.solve(thresh=1e-6, noalloc=True, kernel_variant=1)) == E2E_ERR_INVALID
        ok = m.solve(thresh=1e-6, noalloc=True)
        _same(ok, m.solve(thresh=1e-6))
        _same(m.solve(thresh=1e-6, noalloc=True, kernel_variant=6), ok)
    # one state more than the array holds
    ns = MAX_STATES + 1
    R = np.zeros(ns, np.float32)
    with e2e.import_coo(ns, 1, np.array([0]), np.array([1]), np.array([1.0], np.float32), R) as big:
        assert _code(lambda: big.solve(thresh=1e-6, noalloc=True), "2097152") == E2E_ERR_CAPACITY
