"""GPU parity of the Monte Carlo build (a)(b)(c): bit-exact against the oracle.

Every case exports the device CSR in the reference COO order through the C ABI and
compares S1, S2, prob and R bit for bit with the oracle's restatement of
src/build_model.cu.
"""
import numpy as np
import pytest

import e2e_gpu_dp_amd as e2e
from e2e_gpu_dp_amd.configs import small, workload
from e2e_gpu_dp_amd.fields import make_fields
from helpers import assert_same_model, oracle_model

pytestmark = pytest.mark.gpu


def _case(name):
    if name == "tiny":
        w = small(g=5, nt=3, nr=5, nm=1, na=8, seed=3)
        return w.params, w.fields(), {}
    if name == "small":
        w = small()
        return w.params, w.fields(), {}
    if name == "obstacles_tsg2":
        w = small(g=12, nt=6, nr=300, nm=3, seed=11, term_subgrid_size=2)
        p = w.params
        f = make_fields(p.gsize, p.nt, p.nmodes, p.nrzns, 11, F=p.F, mean_speed_frac=1.5,
                        obstacles=[(0, 3, 4, 7, 5, 6), (2, 6, 8, 9, 2, 8)])
        return p, f, {}
    if name.startswith("reward"):
        rt = int(name[-1])
        w = small(g=9, nt=25, nr=200, nm=2, seed=20 + rt, reward_type=rt, alpha=0.3,
                  num_ac_speeds=2, F=2.0)
        p = w.params
        f = make_fields(p.gsize, p.nt, p.nmodes, p.nrzns, 20 + rt, F=p.F, with_scalar=True)
        return p, f, {}
    if name == "multipass":
        w = small(g=10, nt=4, nr=700, nm=3, seed=5)
        return w.params, w.fields(), {"radix_bits": 1}
    if name == "radix2_overflow":
        w = small(g=10, nt=4, nr=700, nm=3, seed=6)
        return w.params, w.fields(), {"radix_bits": 2, "slot_k": 2}
    if name == "overflow_k1":
        w = small(g=8, nt=3, nr=257, nm=2, seed=9)
        return w.params, w.fields(), {"slot_k": 1}
    if name == "fast_flow_oow":
        w = small(g=30, nt=4, nr=400, nm=2, seed=13)
        p = w.params
        f = make_fields(p.gsize, p.nt, p.nmodes, p.nrzns, 13, F=p.F, mean_speed_frac=9.0,
                        mode_rms_frac=6.0)
        return p, f, {}
    if name == "cells_0p7":  # general remainderf path (cells of width 0.7)
        w = small(g=11, nt=5, nr=300, nm=2, seed=19, dx=0.7, dy=0.7, x0=0.35, y0=0.35)
        return w.params, w.fields(), {}
    if name == "cells_2_wg512":  # width 2 cells, forced 512-thread blocks
        w = small(g=9, nt=4, nr=600, nm=2, seed=23, dx=2.0, dy=2.0, x0=1.0, y0=1.0, F=2.0)
        return w.params, w.fields(), {"block_threads": 512}
    if name == "wg1024":
        w = small(g=8, nt=4, nr=1500, nm=3, seed=29)
        return w.params, w.fields(), {"block_threads": 1024}
    if name == "many_rounds":
        w = small(g=6, nt=3, nr=5000, nm=6, seed=17)
        return w.params, w.fields(), {}
    raise KeyError(name)


CASES = ["tiny", "small", "obstacles_tsg2", "reward0", "reward1", "reward2", "reward3",
         "reward4", "multipass", "radix2_overflow", "overflow_k1", "fast_flow_oow", "many_rounds",
         "cells_0p7", "cells_2_wg512", "wg1024"]


# count_kernel 0: counted cells (every cell the fast unit-cell pass can take, all its
# rows from one count over its samples), the rest one wavefront per row, rows with wide
# S2 spreads handed to the block-level kernel; lane_moves 1: no counted cells, every
# lane moved, the per-row kernel's rounds one at a time (paired_rounds 1) or in pairs
# (2); count_kernel 1: the block-level kernel (radix sort, forced block sizes) for all
KERNELS = {"count": (0, 1, 0), "row": (0, 1, 1), "row_paired": (0, 2, 1), "block": (1, 0, 0)}
# cases whose cells the counting kernel takes (unit cells, integer rewards)
COUNTED = {"tiny", "small", "obstacles_tsg2", "reward0", "multipass", "radix2_overflow", "overflow_k1",
           "fast_flow_oow", "many_rounds", "wg1024"}


@pytest.mark.parametrize("kname", list(KERNELS))
@pytest.mark.parametrize("name", CASES)
def test_build_bit_exact(gpu, name, kname):
    kernel, paired, lanes = KERNELS[kname]
    p, f, opts = _case(name)
    print(f"[case {name} kernel {kname}] g={p.gsize} nt={p.nt} nr={p.nrzns} opts={opts}", flush=True)
    ref = oracle_model(p, f)
    with e2e.build_model(p, f, count_kernel=kernel, paired_rounds=paired, lane_moves=lanes, **opts) as m:
        assert_same_model(m.export_coo(), ref, name)
        info = m.info
        assert info["n_out_of_window"] == ref["oow"]
        assert info["max_row_nnz"] == ref["max_row_nnz"]
        assert (info["wave_blocks"] > 0) == (kernel == 0)
        assert info["paired_rounds"] == (1 if paired == 2 else 0)
        if kernel == 1 and name == "multipass":
            assert info["n_multipass_rows"] > 0
        if name in ("radix2_overflow", "overflow_k1"):
            assert info["n_overflow_rows"] > 0
        if name == "fast_flow_oow":
            assert info["n_out_of_window"] > 0
            if kernel == 0 and lanes:
                assert info["n_deferred_rows"] > 0  # spreads beyond the per-row kernel
        assert (info["counted_cells"] > 0) == (kname == "count" and name in COUNTED)
        assert info["unit_cells"] == (p.dx == 1.0 and p.dy == 1.0)
        if "block_threads" in opts:
            if kernel == 1:
                assert info["block_threads"] == opts["block_threads"]
            else:
                assert info["wave_blocks"] == opts["block_threads"] // 64
        # the int32 export (src/convert_model_datatype.py) carries the same values
        S1, S2, val, R = m.export_coo(np.int32)
        assert (S1 == ref["S1"]).all() and (S2 == ref["S2"]).all()


@pytest.mark.parametrize("lanes", [0, 1])
def test_build_wide_grid_8byte_staging(gpu, lanes):
    """A 260 x 260 grid has more cells than 16 bits hold, so the staging slots and the
    overflow arena keep 8-byte (cell, count) entries: the same model as the oracle's, bit
    for bit, through both the counting and the per-row kernels (2 slots a row, so most
    rows overflow), and the build's compressed entries give the 8-byte entries' solve."""
    w = small(g=260, nt=3, nr=40, nm=1, na=4, seed=37)
    p, f = w.params, w.fields()
    print(f"[wide grid lane_moves {lanes}] g={p.gsize} nt={p.nt} nr={p.nrzns}", flush=True)
    ref = oracle_model(p, f)
    with e2e.build_model(p, f, lane_moves=lanes, slot_k=2) as m:
        assert_same_model(m.export_coo(), ref, "wide grid")
        assert m.info["n_overflow_rows"] > 0
        assert (m.info["counted_cells"] > 0) == (lanes == 0)
        pa, va, sa = m.solve(thresh=1e-6, kernel_variant=3)
        pb, vb, sb = m.solve(thresh=1e-6, kernel_variant=6)
        assert sb["kernel_variant"] == 6 and sb["compressed_windows"] == -(-m.info["nnz"] // 256)
        assert np.array_equal(va.view(np.uint32), vb.view(np.uint32)) and np.array_equal(pa, pb)


@pytest.mark.parametrize("kernel,lanes", [(0, 0), (0, 1), (1, 0)])
def test_build_c1_bit_exact(gpu, kernel, lanes):
    w = workload("C1")
    f = w.fields()
    ref = oracle_model(w.params, f)
    with e2e.build_model(w.params, f, count_kernel=kernel, lane_moves=lanes) as m:
        assert_same_model(m.export_coo(), ref, "C1")
        assert m.info["lanes"] == w.lanes
        if kernel == 0:
            assert m.info["n_deferred_rows"] == 0
        # every C1 cell but the terminal one is counted
        assert m.info["counted_cells"] == (0 if kernel or lanes else (w.params.nt - 1) * (w.params.nc - 1))


def test_build_deterministic_and_device_fields(gpu):
    import torch
    w = small(g=16, nt=6, nr=900, nm=3, seed=31)
    f = w.fields()
    with e2e.build_model(w.params, f) as m1:
        a = m1.export_coo()
    fd = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in f.items()}
    with e2e.build_model(w.params, fd) as m2:
        b = m2.export_coo()
    for x, y in zip(a, b):
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8))


def test_nonfinite_field_is_an_error(gpu):
    w = small(g=6, nt=3, nr=64, nm=1, seed=2)
    f = w.fields()
    f["u"][1, 2, 3] = np.nan
    with pytest.raises(e2e.E2EError) as ei:
        e2e.build_model(w.params, f)
    assert ei.value.code == 8
