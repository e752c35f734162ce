"""Policy rollouts on the GPU (e2e_rollout) against the oracle's restatement of
Post_processing/plot_e2eGPU_policy.py:376-520 and Post_processing/grid_world.py:160-230,
on policies solved by this library. Outcomes (success, travel time) must agree exactly;
positions and energies to 1e-12 (the interpolated action's atan2 / cos / sin run on the
device, the oracle's in glibc)."""
import numpy as np
import pytest

import e2e_gpu_dp_amd as e2e
import oracle
from e2e_gpu_dp_amd.configs import Workload, small

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def _solved(w):
    f = w.fields()
    with e2e.build_model(w.params, f) as m:
        pol, _, _ = m.solve(thresh=1e-6)
    return f, pol.astype(np.int64).astype(np.int32)


def _compare(w, f, pol, starts, interp, want_xy=True):
    g = e2e.rollout(w.params, f, pol, starts, interp=interp, want_xy=want_xy)
    o = oracle.rollout(w.params, f, pol, starts, interp=interp, want_xy=want_xy)
    print(f"{w.name} interp={interp}: {len(g['success'])} trajectories, reached {g['stats']['reached']}, "
          f"moves {g['stats']['moves']}, kernel {g['stats']['kernel_ms']:.3f} ms", flush=True)
    np.testing.assert_array_equal(g["success"], o["success"])
    np.testing.assert_array_equal(g["travel"], o["travel"])
    np.testing.assert_allclose(g["energy"], o["energy"], rtol=RTOL, atol=1e-12)
    if want_xy:
        np.testing.assert_allclose(g["xy"], o["xy"], rtol=0, atol=1e-12)
    assert g["stats"]["moves"] == int(o["travel"].sum())
    assert g["stats"]["reached"] == int((o["success"] == 1).sum())
    return g


@pytest.mark.parametrize("interp", [True, False])
def test_rollout_matches_oracle(gpu, interp):
    w = small(g=16, nt=30, nr=300, nm=3, seed=41, term_subgrid_size=2)
    f, pol = _solved(w)
    g = _compare(w, f, pol, [(8, 3), (12, 12), (3, 9)], interp)
    assert g["stats"]["reached"] > 0


def test_rollout_obstacles_scalar_field(gpu):
    """Obstacle islands (in_obstacle ends a run, :473-474) and a scalar field (the
    energy_col terms, :460-467)."""
    base = small(g=20, nt=40, nr=200, nm=2, seed=43, term_subgrid_size=2, reward_type=2)
    boxes = [(0, 40, 6, 9, 6, 9), (20, 40, 12, 15, 3, 6)]
    w = Workload("small-obs-rad", base.params, base.seed, base.sigmas, obstacles=boxes, with_scalar=True)
    f, pol = _solved(w)
    for interp in (True, False):
        g = _compare(w, f, pol, [(10, 2), (15, 15), (2, 10), (7, 12)], interp)
        assert np.any(g["energy"][:, 1] != 0)


def test_rollout_bench_size_subset(gpu):
    """A 50x50 x 60 model, 2000 realizations, 6 start cells, no positions kept."""
    w = small(g=50, nt=60, nr=2000, nm=4, seed=47, term_subgrid_size=2)
    f, pol = _solved(w)
    _compare(w, f, pol, [(25, 47), (10, 10), (40, 20), (30, 30), (5, 44), (44, 5)], True, want_xy=False)


def test_rollout_metrics_and_errors(gpu):
    w = small(g=12, nt=20, nr=100, nm=2, seed=45, term_subgrid_size=2)
    f, pol = _solved(w)
    r = e2e.rollout(w.params, f, pol, [(6, 2)])
    met = e2e.rollout_metrics(r)
    ok = r["success"] == 1
    assert met["success_rate"] == ok.mean()
    if ok.any():
        assert met["time_mean"] == r["travel"][ok].mean()
        assert met["net_energy_max"] == r["energy"][ok, 2].max()
    # a start outside the grid and a short policy are argument errors
    with pytest.raises(e2e.E2EError):
        e2e.rollout(w.params, f, pol, [(12, 0)])
    with pytest.raises(ValueError):
        e2e.rollout(w.params, f, pol[:-1], [(6, 2)])
    # no terminal subgrid: a vehicle still in the endpos cell when its next action is
    # interpolated meets a support pattern the reference rejects (:158-161)
    p0 = w.params.replace(term_subgrid_size=0)
    st = [(p0.i_term, p0.j_term)]
    o = oracle.rollout(p0, f, pol, st, interp=True)
    assert 0 < o["bad"] == int((o["success"] == -1).sum())
    with pytest.raises(e2e.E2EError):
        e2e.rollout(p0, f, pol, st, interp=True)
