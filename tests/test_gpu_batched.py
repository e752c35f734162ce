"""Batched multi-goal SpMM-VI: each column equals the plain solve of the model built
with its terminal cell at that goal (the reference's one-terminal-per-build workflow,
src/input_to_build_model.h:33-40, then src/spvi_2.cu:650-726)."""
import numpy as np
import pytest

import e2e_gpu_dp_amd as e2e
from e2e_gpu_dp_amd.configs import goal_cells, small, workload
from helpers import rel_err

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
        print(f"goal {goals[j]}: max rel err {err.max():.2e}, bit-identical {same:.4f}")
        assert np.mean(pol[j] == pp) > 0.99


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


def test_batched_columns_independent_of_batch_and_variant(gpu):
    w = small(g=14, nt=10, nr=400, nm=3, seed=73)
    f = w.fields()
    goals = goal_cells(14, 40, 9)
    with _shared(w, f) as m:
        ref_pol, ref_val, ref_sw, st = m.solve_goals(goals, thresh=1e-6, variant=0)
        for variant in range(1, 9):
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


def test_batched_rejects_models_it_cannot_serve(gpu):
    w = small(g=8, nt=5, nr=64, nm=1, seed=75)
    f = w.fields()
    with e2e.build_model(w.params, f, keep_rsum=True) as m:  # has a terminal cell
        with pytest.raises(e2e.E2EError):
            m.solve_goals([3, 4])
    with e2e.build_model(w.params.replace(term_subgrid_size=0), f) as m:  # no Rsum
        with pytest.raises(e2e.E2EError):
            m.solve_goals([3, 4])
