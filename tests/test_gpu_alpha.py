"""GPU parity of the alpha sweep (src/my_script.sh:10-19): one build with alpha columns
against one oracle build per alpha, and one batched solve against one plain solve per
alpha.

The reference runs build_model and spvi once per alpha = k*10 / 100 (float, as
src/build_model.cu:937-939 forms it); only reward method 4 depends on alpha
(calculate_one_step_reward :256-260). The cases have two speeds and F = 2, so the
rewards are not integers; obstacles, so that lanes cross them (:444-477); and horizons
long enough to hold slices with and without the late penalty (T > nt - 20).
"""
import numpy as np
import pytest

import e2e_gpu_dp_amd as e2e
from e2e_gpu_dp_amd.configs import small
from e2e_gpu_dp_amd.fields import make_fields
from helpers import assert_close_to_f64, assert_same_model, oracle_model

pytestmark = pytest.mark.gpu

ALPHAS = (np.arange(11, dtype=np.float32) * np.float32(10)) / np.float32(100)


def _case(name):
    if name == "two_speeds":
        w = small(g=10, nt=24, nr=300, nm=2, na=8, seed=101, reward_type=4, F=2.0, num_ac_speeds=2)
        p = w.params
        f = make_fields(p.gsize, p.nt, p.nmodes, p.nrzns, 101, F=p.F, mean_speed_frac=0.8)
    else:  # obstacles, one speed
        w = small(g=12, nt=22, nr=260, nm=2, na=8, seed=103, reward_type=4, F=1.5)
        p = w.params
        boxes = [(0, 22, 3, 5, 3, 6), (10, 22, 7, 9, 2, 4)]
        boxes = [b for b in boxes if not (b[2] <= p.i_term < b[3] and b[4] <= p.j_term < b[5])]
        f = make_fields(p.gsize, p.nt, p.nmodes, p.nrzns, 103, F=p.F, mean_speed_frac=0.8, obstacles=boxes)
    return p, f


@pytest.mark.parametrize("name", ["two_speeds", "obstacles"])
def test_alpha_rewards_match_per_alpha_builds(gpu, name):
    p, f = _case(name)
    with e2e.build_model(p, f, alphas=ALPHAS) as m:
        coo = m.export_coo()
        assert m.n_alpha == len(ALPHAS)
        for j, a in enumerate(ALPHAS):
            ref = oracle_model(p.replace(alpha=float(a)), f)
            if j == 0:
                assert (ref["R"] != np.round(ref["R"])).any()  # non-integer rewards
                # the transitions are the same for every alpha
                assert np.array_equal(coo[1], ref["S2"]) and np.array_equal(coo[2].view(np.uint32),
                                                                             ref["val"].view(np.uint32))
            Ra = m.export_R_alpha(j)
            assert np.array_equal(Ra.view(np.uint32), ref["R"].view(np.uint32)), (name, j, float(a))
        # the model's own reward is that of its alpha (params.alpha = 0 here)
        assert np.array_equal(coo[3].view(np.uint32), m.export_R_alpha(0).view(np.uint32))


# the batched SpMM (autotuned, and the tile kernel over 16 columns); one plain sweep chain
# per column
@pytest.mark.parametrize("variant", [-1, 33, -2])
def test_alpha_solve_matches_plain_solves(gpu, variant):
    p, f = _case("two_speeds")
    with e2e.build_model(p, f, alphas=ALPHAS) as m:
        pol, val, sw, st = m.solve_alphas(thresh=1e-6, variant=variant)
        assert st["converged"] == len(ALPHAS)
        assert (st["variant"] == -2) == (variant == -2)
        if variant >= 0:
            assert st["variant"] == variant
    for j, a in enumerate(ALPHAS):
        with e2e.build_model(p.replace(alpha=float(a)), f) as mj:
            pj, vj, sj = mj.solve(thresh=1e-6)
        assert sw[j] == sj["sweeps"], j
        assert np.array_equal(val[j].view(np.uint32), vj.view(np.uint32)), j
        assert np.array_equal(pol[j], pj), j
        if variant == -1 and j in (0, 3, 10):  # the FP64 backward induction of the oracle's build
            assert_close_to_f64(oracle_model(p.replace(alpha=float(a)), f), pol[j], val[j], 1e-5, f"alpha {a}")


def test_alpha_backward_columns_match_converged_plain_solves(gpu):
    """Variant -3: each alpha column by the exact backward pass, the fixed point of the
    plain sweeps of a separate build with that alpha, bit for bit."""
    p, f = _case("two_speeds")
    with e2e.build_model(p, f, alphas=ALPHAS) as m:
        pol, val, sw, st = m.solve_alphas(variant=-3)
        assert st["variant"] == -3 and st["converged"] == len(ALPHAS) and (sw == 1).all()
    for j, a in enumerate(ALPHAS):
        with e2e.build_model(p.replace(alpha=float(a)), f) as mj:
            pj, vj, sj = mj.solve(thresh=1e-30, max_sweeps=p.nt + 4)
        assert sj["converged"]
        assert np.array_equal(val[j].view(np.uint32), vj.view(np.uint32)), j
        assert np.array_equal(pol[j], pj), j
        if j in (1, 7):
            assert_close_to_f64(oracle_model(p.replace(alpha=float(a)), f), pol[j], val[j], 1e-5, f"alpha {a}")


def test_alpha_columns_errors(gpu):
    p, f = _case("two_speeds")
    with pytest.raises(e2e.E2EError) as ei:  # only reward method 4 depends on alpha
        e2e.build_model(p.replace(reward_type=0), f, alphas=ALPHAS)
    assert ei.value.code == 1
    with pytest.raises(e2e.E2EError) as ei:  # the per-row kernel is required
        e2e.build_model(p, f, alphas=ALPHAS, count_kernel=1)
    assert ei.value.code == 7
    # rewards spanning too many binades for an exact sum over classes: refused, not rounded
    q = p.replace(r_outbound=-3.0e6, F=0.0173)
    with pytest.raises(e2e.E2EError) as ei:
        e2e.build_model(q, f, alphas=ALPHAS)
    assert ei.value.code == 7
    with e2e.build_model(p, f) as m:
        with pytest.raises(ValueError):
            m.solve_alphas()
