"""Input generation on the GPU (e2e_env_generate) against the oracle's restatement of
src/data_input/build_environment_DG_with_westward_clouds.py (numpy, with LAPACK's SVD).

Bars: means within float32 rounding; singular values within 1e-9 relative (2e-8 of the
largest in absolute terms, the Gram method's floor); modes and
coefficients (same sign convention) within 2e-6 of the largest entry, which float32
outputs and a Gram-matrix eigensolve allow; the cloud field and obstacle mask exact. The
generated fields feed the model build."""
import math

import numpy as np
import pytest

import e2e_gpu_dp_amd as e2e
from e2e_gpu_dp_amd.configs import Params
from oracle import envgen

pytestmark = pytest.mark.gpu


def _cfg(gsize, nt, nw, k):
    return dict(gsize=gsize, nt=nt, n_wsamples=nw, n_modes=k, dt=10 / nt, dxy=2 / gsize, A=0.5, eps=0.1,
                wx=math.pi, wy=0.5 * math.pi, w_range=(math.pi / 10, 8 * math.pi / 10), A_sc=3.0)


def _compare(cfg, obstacle=None, obs_mode=0, obs=(0, 0, 0, (0.0, 0.0)), phi=True):
    g = e2e.generate_environment(**cfg, obstacles=obstacle, phi=phi)
    o = envgen.environment(**cfg, obstacle_mode=obs_mode, obs=obs, phi=phi)
    print(f"g={cfg['gsize']} nt={cfg['nt']} nw={cfg['n_wsamples']} k={cfg['n_modes']}: {g['stats']}, "
          f"sv[0]={o['singular_values'][0]}", flush=True)
    for name in ("u", "v"):
        np.testing.assert_allclose(g[name], o[name], rtol=1e-6, atol=1e-6 * np.abs(o[name]).max())
    # singular values from the Gram matrix: relative 1e-9 for the leading ones, and an
    # absolute floor of about sqrt(eps) s_max (2e-8 s_max) below which s = sqrt(lambda)
    # carries no digits (such modes hold no energy a float32 field can show)
    np.testing.assert_allclose(g["singular_values"], o["singular_values"], rtol=1e-9,
                               atol=2e-8 * o["singular_values"].max())
    # modes are defined where the spectrum is: at t = 0 every forcing frequency gives the
    # same field (sin(w t) = 0), the centred matrix is zero and any basis is a valid SVD,
    # in the reference too; a mode is compared when its singular value stands apart from
    # its neighbours by 1e-4 of the slice's largest
    k = cfg["n_modes"]
    smax = o["s_full"].max()
    compared = 0
    for t in range(cfg["nt"]):
        sf = o["s_full"][t]
        for c in range(k):
            gap = min(sf[c - 1] - sf[c] if c > 0 else np.inf, sf[c] - sf[c + 1] if c + 1 < len(sf) else sf[c])
            if sf[c] < 1e-6 * smax or gap < 1e-4 * sf[0]:
                continue
            compared += 1
            # the sign rule (first largest-magnitude entry positive) can tie on symmetric
            # fields, where last-bit differences pick another entry: align by the dot product
            mg = np.concatenate([g["ui"][t, c].ravel(), g["vi"][t, c].ravel()]).astype(np.float64)
            mo = np.concatenate([o["ui"][t, c].ravel(), o["vi"][t, c].ravel()]).astype(np.float64)
            sg = 1.0 if mg @ mo >= 0 else -1.0
            np.testing.assert_allclose(sg * mg, mo, rtol=0, atol=2e-6 * np.abs(mo).max(), err_msg=f"mode t={t} c={c}")
            np.testing.assert_allclose(sg * g["yi"][t, :, c], o["yi"][t, :, c], rtol=0,
                                       atol=2e-6 * np.abs(o["yi"][t, :, c]).max(), err_msg=f"yi t={t} c={c}")
    assert compared >= (cfg["nt"] - 1) * min(k, 2)
    if phi:
        np.testing.assert_allclose(g["phi"], o["phi"], rtol=1e-6, atol=1e-6 * np.abs(o["phi"]).max())
        rp = [(g["phi_yi"][t].astype(np.float64) @ g["phi_i"][t].reshape(k, -1),
               o["phi_yi"][t].astype(np.float64) @ o["phi_i"][t].reshape(k, -1)) for t in range(cfg["nt"])]
        scale = max(np.abs(b).max() for _, b in rp)  # the field's scale over all slices
        for rp_g, rp_o in rp:
            np.testing.assert_allclose(rp_g, rp_o, rtol=0, atol=1e-5 * scale)
    np.testing.assert_array_equal(g["s"], o["s"])
    np.testing.assert_array_equal(g["mask"], o["mask"])
    # the velocity each realization sees, mean + sum_k Yi_k mode_k, agrees as a whole
    n = cfg["gsize"] ** 2
    for t in range(cfg["nt"]):
        rec_g = g["u"][t].ravel() + g["yi"][t].astype(np.float64) @ g["ui"][t].reshape(cfg["n_modes"], n)
        rec_o = o["u"][t].ravel() + o["yi"][t].astype(np.float64) @ o["ui"][t].reshape(cfg["n_modes"], n)
        np.testing.assert_allclose(rec_g, rec_o, rtol=0, atol=1e-5 * np.abs(rec_o).max())
    return g


def test_envgen_matches_oracle_left_gram():
    """n_wsamples < 2 g^2: the Gram matrix of the realizations (rows)."""
    _compare(_cfg(12, 6, 60, 3), obstacle="static", obs_mode=1)


def test_envgen_matches_oracle_right_gram():
    """n_wsamples > 2 g^2: the Gram matrix of the grid columns."""
    _compare(_cfg(4, 5, 200, 4), phi=False)


def test_envgen_moving_obstacle_and_more_modes():
    cfg = _cfg(16, 8, 120, 6)
    _compare(cfg, obstacle=("dynamic", 5, 3, 4, (1.0, 0.5)), obs_mode=2, obs=(5, 3, 4, (1.0, 0.5)))


def test_envgen_feeds_the_builder():
    """The generated fields build and solve as any other input."""
    cfg = _cfg(20, 12, 200, 3)
    env = e2e.generate_environment(**cfg)
    p = Params(gsize=20, nt=12, nrzns=200, nmodes=3, num_ac_angles=8, F=0.2, dt=10 / 12, dx=2 / 20, dy=2 / 20,
               x0=0.05, y0=0.05, i_term=6, j_term=14, term_subgrid_size=2)
    fields = {k: env[k] for k in ("u", "v", "ui", "vi", "yi", "s", "mask")}
    with e2e.build_model(p, fields) as m:
        assert m.info["nnz"] > 0
        pol, val, st = m.solve(thresh=1e-6)
        assert st["converged"]


def test_envgen_rejects_bad_arguments():
    with pytest.raises(e2e.E2EError):
        e2e.generate_environment(**_cfg(6, 3, 1, 1))  # n_wsamples - 1 = 0 divides sample_w
    with pytest.raises(e2e.E2EError):
        e2e.generate_environment(**_cfg(6, 3, 20, 40))  # more modes than realizations
