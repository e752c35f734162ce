"""Batched multi-goal SpMM-VI: each column equals the plain solve of the model built
with its terminal cell at that goal (the reference's one-terminal-per-build workflow,
src/input_to_build_model.h:33-40, then src/spvi_2.cu:650-726)."""
import numpy as np
import pytest

import e2e_gpu_dp_amd as e2e
from e2e_gpu_dp_amd.configs import goal_cells, small, workload
from e2e_gpu_dp_amd.fields import make_fields
from helpers import assert_close_to_f64, goal_params, oracle_model, rel_err

import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _shared(w, f):
    return e2e.build_model(w.params.replace(term_subgrid_size=0), f, keep_rsum=True)


def _plain(w, f, goal, thresh=1e-6):
    g = w.params.gsize
    p = w.params.replace(i_term=int(goal) // g, j_term=int(goal) % g, term_subgrid_size=1)
    with e2e.build_model(p, f) as m:
        return m.solve(thresh=thresh)


def test_batched_c1_64_goals_match_plain_solves(gpu):
    w = workload("C1")
    f = w.fields()
    goals = goal_cells(w.params.gsize, 64, 3456)
    with _shared(w, f) as m:
        pol, val, sw, st = m.solve_goals(goals, thresh=1e-6)
    assert st["converged"] == 64
    print(f"variant {st['variant']} cpl={st['cols_per_lane']} spw={st['states_per_wave']} "
          f"sweeps={st['sweeps']} solve_ms={st['solve_ms']:.2f}")
    for j in range(0, 64, 9):
        pp, vp, sp = _plain(w, f, goals[j])
        err = rel_err(val[j], vp.astype(np.float64))
        assert err.max() <= TOL, (j, err.max())
        assert sw[j] == sp["sweeps"], (j, sw[j], sp["sweeps"])
        same = np.mean(val[j].view(np.uint32) == vp.view(np.uint32))
        assert np.mean(pol[j] == pp) > 0.99
        # and the FP64 backward induction of the oracle's build of that goal's model
        e64 = assert_close_to_f64(oracle_model(goal_params(w.params, goals[j]), f), pol[j], val[j],
                                  TOL, f"goal {goals[j]}")
        print(f"goal {goals[j]}: max rel err vs plain {err.max():.2e} (bit-identical {same:.4f}), "
              f"vs FP64 oracle {e64:.2e}")


def test_batched_reference_threshold(gpu):
    """thresh 0.1 (spvi_2.cu:757): each column stops at its own first sweep < 0.1."""
    w = small(g=12, nt=25, nr=300, nm=2, seed=71)
    f = w.fields()
    goals = goal_cells(12, 20, 5)
    with _shared(w, f) as m:
        pol, val, sw, st = m.solve_goals(goals, thresh=0.1)
    for j in (0, 7, 19):
        pp, vp, sp = _plain(w, f, goals[j], thresh=0.1)
        assert sw[j] == sp["sweeps"]
        assert rel_err(val[j], vp.astype(np.float64)).max() <= TOL
        # the oracle's FP64 sweeps of that goal's model stop at the same sweep with the
        # same values (the early stop of src/spvi_2.cu:669-698, in FP64)
        r = oracle_model(goal_params(w.params, goals[j]), f)
        V64, _, it64, _ = oracle.vi_sweeps_f64(r["ns"], r["na"], r["S1"], r["S2"], r["val"], r["R"], 0.1)
        assert sw[j] == it64, (j, sw[j], it64)
        assert rel_err(val[j], V64).max() <= TOL


def test_batched_columns_independent_of_batch_and_variant(gpu):
    w = small(g=14, nt=10, nr=400, nm=3, seed=73)
    f = w.fields()
    goals = goal_cells(14, 40, 9)
    with _shared(w, f) as m:
        ref_pol, ref_val, ref_sw, st = m.solve_goals(goals, thresh=1e-6, variant=0)
        for j in (0, 17, 39):
            assert_close_to_f64(oracle_model(goal_params(w.params, goals[j]), f), ref_pol[j], ref_val[j], TOL,
                                f"goal {goals[j]}")
        for variant in range(1, N_VARIANTS):
            pol, val, sw, st = m.solve_goals(goals, thresh=1e-6, variant=variant)
            assert np.array_equal(val.view(np.uint32), ref_val.view(np.uint32)), variant
            assert np.array_equal(pol, ref_pol) and np.array_equal(sw, ref_sw)
        # one goal alone gives the same bits as inside the batch
        for j in (3, 31):
            pol1, val1, sw1, _ = m.solve_goals(goals[j:j + 1], thresh=1e-6)
            assert np.array_equal(val1[0].view(np.uint32), ref_val[j].view(np.uint32))
            assert sw1[0] == ref_sw[j]
        # graph-replayed sweeps
        polg, valg, swg, _ = m.solve_goals(goals, thresh=1e-6, use_graph=True, sweeps_per_check=4)
        assert np.array_equal(valg.view(np.uint32), ref_val.view(np.uint32))


# SpMM variants 0-8: one lane group per state walks every row; 9-20: one lane group per
# row (spmm_vi_rows), as (column lanes, columns per lane); 27-32: the same over a tile's
# LDS halo of the next slice's values (spmm_vi_halo); 33-37: one lane per row over all
# the block's columns, with the halo and a row group's entries in LDS (spmm_vi_tile), by
# column block
N_VARIANTS = 38
HALO_VARIANTS = {27: (4, 4), 28: (2, 4), 29: (4, 2), 30: (1, 4), 31: (2, 8), 32: (2, 2)}
TILE_VARIANTS_KB = {33: 8, 34: 16, 35: 8, 36: 16, 37: 16}
ROWS_VARIANTS = {9: (4, 1), 10: (8, 1), 11: (4, 2), 12: (8, 2), 13: (4, 4), 14: (16, 1), 15: (8, 4),
                 16: (16, 2), 17: (1, 4), 18: (1, 8), 19: (2, 4), 20: (2, 8)}


# column block of variants 0-8 (the state-parallel kernel fits any action count) and of
# 21-26 (spmm_vi_stream: one lane per row, transposed gathers; any action count)
STATE_VARIANTS_KB = {0: 64, 1: 64, 2: 64, 3: 32, 4: 32, 5: 32, 6: 16, 7: 16, 8: 16}
STREAM_VARIANTS_KB = {21: 4, 22: 4, 23: 8, 24: 8, 25: 1, 26: 2}


@pytest.mark.parametrize("na,k,g,fast", [(8, 5, 12, 0), (8, 16, 12, 0), (16, 40, 12, 0), (12, 32, 12, 0),
                                         (4, 64, 12, 0), (16, 3, 12, 0), (8, 16, 37, 0), (16, 16, 23, 1)])
def test_batched_row_kernels_bit_identical(gpu, na, k, g, fast):
    """Every variant that fits the model (row and halo variants: column lanes x actions
    rounded up to a power of two <= 64; all: column block dividing the padded width, 4 or
    8 columns for k <= 4 or 8, else a multiple of 16) runs as asked and gives the same
    bits; na = 12 leaves padding action lanes. g = 37 and 23 leave partial halo tiles at
    the grid's edges; the fast flow moves lanes up to 5+ cells, outside the halo."""
    w = small(g=g, nt=9, nr=300, nm=3, na=na, seed=91 + na)
    p = w.params
    f = make_fields(p.gsize, p.nt, p.nmodes, p.nrzns, 91 + na, mean_speed_frac=3.5) if fast else w.fields()
    goals = goal_cells(g, k, 17 + k)
    kp = 4 if k <= 4 else 8 if k <= 8 else -(-k // 16) * 16
    nap = 1 << (na - 1).bit_length()
    with _shared(w, f) as m:
        if fast:  # some entries land more than 4 cells away: outside the halo
            S1, S2, _, _ = m.export_coo()
            sp1, sp2 = (S1 % p.ns) % p.nc, S2 % p.nc
            assert (np.maximum(np.abs(sp2 // g - sp1 // g), np.abs(sp2 % g - sp1 % g)) > 4).any()
        ref_pol, ref_val, ref_sw, ref_st = m.solve_goals(goals, thresh=1e-6, variant=9)
        assert ref_st["variant"] == 9
        ran = set()
        for variant in range(N_VARIANTS):
            if variant in ROWS_VARIANTS or variant in HALO_VARIANTS:
                lc, cpl = ROWS_VARIANTS.get(variant) or HALO_VARIANTS[variant]
                fits = lc * nap <= 64 and kp % (lc * cpl) == 0
            elif variant in STREAM_VARIANTS_KB:
                fits = kp % STREAM_VARIANTS_KB[variant] == 0
            elif variant in TILE_VARIANTS_KB:
                fits = nap <= 64 and kp % TILE_VARIANTS_KB[variant] == 0
            else:
                fits = kp % STATE_VARIANTS_KB[variant] == 0
            pol, val, sw, st = m.solve_goals(goals, thresh=1e-6, variant=variant)
            print(f"na={na} k={k} variant {variant} fits={fits} ran={st['variant']} "
                  f"spw={st['states_per_wave']}", flush=True)
            if fits:
                assert st["variant"] == variant
                ran.add(variant)
            assert np.array_equal(val.view(np.uint32), ref_val.view(np.uint32)), variant
            assert np.array_equal(pol, ref_pol) and np.array_equal(sw, ref_sw), variant
        assert len(ran & set(ROWS_VARIANTS)) >= 2
        assert len(ran & set(STREAM_VARIANTS_KB)) >= 2
        assert len(ran & set(HALO_VARIANTS)) >= 2
        if kp >= 8:
            assert len(ran & set(TILE_VARIANTS_KB)) >= (2 if kp >= 16 else 1)
        if kp >= 16:
            assert ran & set(STATE_VARIANTS_KB)


def test_batched_tiles_from_counters_and_strides(gpu):
    """64 goal columns on a model with more tiles than resident blocks: the tile and halo
    sweeps' blocks of the first 8 column blocks take their tiles from per-XCD counters,
    those of later column blocks (the 4-column halo variants have 16) stride their XCD's
    band of tiles; every variant gives the bits, sweep counts and policies of the row
    kernel, over every state and over the active slices."""
    g, nt = 64, 12
    w = small(g=g, nt=nt, nr=300, nm=3, na=8, seed=101)
    f = w.fields()
    goals = goal_cells(g, 64, 41)
    with _shared(w, f) as m:
        ref_pol, ref_val, ref_sw, ref_st = m.solve_goals(goals, thresh=1e-6, variant=9)
        assert ref_st["variant"] == 9
        for variant in sorted(set(HALO_VARIANTS) | set(TILE_VARIANTS_KB)):
            for full in (False, True):
                pol, val, sw, st = m.solve_goals(goals, thresh=1e-6, variant=variant, full_sweeps=full)
                print(f"variant {variant} full={full}", flush=True)
                assert st["variant"] == variant
                assert np.array_equal(val.view(np.uint32), ref_val.view(np.uint32)), (variant, full)
                assert np.array_equal(pol, ref_pol) and np.array_equal(sw, ref_sw), (variant, full)


@pytest.mark.parametrize("k", [3, 40])
def test_batched_backward_pass_equals_converged_columns(gpu, k):
    """Variant -3: every goal column by one exact backward pass, bit for bit the columns of
    the batched sweeps run to a zero residual."""
    w = small(g=14, nt=12, nr=400, nm=3, seed=93)
    f = w.fields()
    goals = goal_cells(14, k, 29)
    with _shared(w, f) as m:
        pj, vj, sj, stj = m.solve_goals(goals, thresh=1e-30, max_sweeps=w.params.nt + 4)
        pb, vb, sb, stb = m.solve_goals(goals, variant=-3)
        assert stj["converged"] == k and stb["variant"] == -3 and (sb == 1).all()
        assert np.array_equal(vb.view(np.uint32), vj.view(np.uint32))
        assert np.array_equal(pb, pj)
    for j in sorted({0, k - 1}):  # the fixed point is the FP64 backward induction's, within 1e-5
        assert_close_to_f64(oracle_model(goal_params(w.params, goals[j]), f), pb[j], vb[j], TOL, f"goal {goals[j]}")


@pytest.mark.parametrize("na,k", [(8, 16), (16, 40), (8, 3)])
def test_batched_active_slices_equal_full_sweeps(gpu, na, k):
    """Every sweep of a batched solve recomputes only the slices not yet at their final
    values; every variant gives the bits, per-column sweep counts and residuals of sweeps
    over every state, for columns stopping at different sweeps (thresh 1e-3) and at the
    fixed point (1e-30)."""
    g, nt = 12, 9
    w = small(g=g, nt=nt, nr=300, nm=3, na=na, seed=97 + na)
    f = w.fields()
    goals = goal_cells(g, k, 5 + k)
    with _shared(w, f) as m:
        for variant in range(N_VARIANTS):
            for thresh in (1e-3, 1e-30):
                kw = dict(thresh=thresh, variant=variant, max_sweeps=nt + 4, raise_on_nonconv=False)
                pf, vf, swf, sf = m.solve_goals(goals, full_sweeps=True, **kw)
                pa, va, swa, sa = m.solve_goals(goals, **kw)
                assert sa["variant"] == sf["variant"] and sa["status"] == sf["status"] == 0, variant
                assert np.array_equal(swa, swf), (variant, thresh)
                assert np.array_equal(va.view(np.uint32), vf.view(np.uint32)), (variant, thresh)
                assert np.array_equal(pa, pf), (variant, thresh)
                if sa["sweeps"] > 2:
                    assert sa["bytes_per_sweep"] < sf["bytes_per_sweep"]
        # graph-replayed sweeps over the active slices
        pg, vg, swg, _ = m.solve_goals(goals, thresh=1e-3, use_graph=True, sweeps_per_check=4)
        pf, vf, swf, _ = m.solve_goals(goals, thresh=1e-3, full_sweeps=True)
        assert np.array_equal(vg.view(np.uint32), vf.view(np.uint32)) and np.array_equal(swg, swf)


def test_batched_rejects_models_it_cannot_serve(gpu):
    w = small(g=8, nt=5, nr=64, nm=1, seed=75)
    f = w.fields()
    with e2e.build_model(w.params, f, keep_rsum=True) as m:  # has a terminal cell
        with pytest.raises(e2e.E2EError):
            m.solve_goals([3, 4])
    with e2e.build_model(w.params.replace(term_subgrid_size=0), f) as m:  # no Rsum
        with pytest.raises(e2e.E2EError):
            m.solve_goals([3, 4])
