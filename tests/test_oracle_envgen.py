"""Pins the input-generation oracle (oracle/envgen.py, the restatement of
src/data_input/build_environment_DG_with_westward_clouds.py) without running the
reference: the vectorized realizations against the reference's scalar formula, closed
forms (eps = 0 leaves one realization: no modes, the mean is the steady double gyre), the
truncation reproducing the realizations at full rank, and the cloud and obstacle fields.
No GPU."""
import math

import numpy as np

from oracle import envgen

P = dict(gsize=6, nt=4, n_wsamples=9, dt=10 / 4, dxy=2 / 6, A=0.5, eps=0.1, wx=math.pi, wy=0.5 * math.pi,
         w_range=(math.pi / 10, 8 * math.pi / 10), A_sc=3.0)


def test_vectorized_realizations_equal_the_scalar_formula():
    g, nt, nw = P["gsize"], P["nt"], P["n_wsamples"]
    vx, vy, ph = envgen.realizations(g, nt, nw, P["dt"], P["dxy"], P["A"], P["eps"], P["wx"], P["wy"],
                                     *P["w_range"])
    for (tid, k, i, j) in [(0, 0, 0, 0), (3, 8, 5, 5), (2, 4, 1, 3), (1, 7, 4, 0)]:
        sx, sy, sp = envgen.realization_scalar(g, P["dt"], P["dxy"], P["A"], P["eps"], P["wx"], P["wy"],
                                               *P["w_range"], nw, tid, k, i, j)
        col = g * i + j
        np.testing.assert_allclose([vx[tid, k, col], vy[tid, k, col], ph[tid, k, col]], [sx, sy, sp],
                                   rtol=1e-14, atol=1e-15)


def test_unperturbed_gyre_has_no_modes():
    """eps = 0: f = x, fx = 1 for every forcing frequency, so all realizations equal the
    steady gyre vx = -(A pi / 2) sin(wx x) cos(wy y), vy = A pi sin(wy y) cos(wx x); the
    centred matrix is zero."""
    e = envgen.environment(**{**P, "eps": 0.0}, n_modes=2)
    g, dxy, A = P["gsize"], P["dxy"], P["A"]
    xs = (0.5 + np.arange(g)) * dxy
    X, Y = xs[None, :], xs[::-1][:, None]
    vx = -(0.5 * A * math.pi * np.sin(math.pi * X) * np.cos(0.5 * math.pi * Y))
    vy = A * math.pi * np.sin(0.5 * math.pi * Y) * np.cos(math.pi * X)
    for t in range(P["nt"]):
        np.testing.assert_allclose(e["u"][t], vx.astype(np.float32), rtol=1e-6, atol=1e-7)
        np.testing.assert_allclose(e["v"][t], vy.astype(np.float32), rtol=1e-6, atol=1e-7)
    assert np.abs(e["singular_values"]).max() < 1e-12


def test_full_rank_modes_reproduce_the_realizations():
    """With n_modes = rank, mean + C M is every realization (the truncation keeps it all),
    which pins the mode/coefficient assembly and the [vx | vy] layout."""
    nw = P["n_wsamples"]
    e = envgen.environment(**P, n_modes=nw)
    vx, vy, _ = envgen.realizations(P["gsize"], P["nt"], nw, P["dt"], P["dxy"], P["A"], P["eps"], P["wx"],
                                    P["wy"], *P["w_range"])
    n = P["gsize"] ** 2
    for t in range(P["nt"]):
        rec = e["mean"][t] + e["C"][t] @ e["M"][t]
        np.testing.assert_allclose(rec[:, :n], vx[t], rtol=0, atol=1e-12)
        np.testing.assert_allclose(rec[:, n:], vy[t], rtol=0, atol=1e-12)
        # float32 fields carry the same numbers
        np.testing.assert_array_equal(e["ui"][t].reshape(nw, n), e["M"][t][:, :n].astype(np.float32))
        np.testing.assert_array_equal(e["yi"][t], e["C"][t].astype(np.float32))
        # singular values descending, modes orthonormal, signs as documented
        sv = e["singular_values"][t]
        assert np.all(np.diff(sv) <= 1e-12)
        np.testing.assert_allclose(e["M"][t] @ e["M"][t].T, np.eye(nw), atol=1e-10)
        for c in range(nw):
            assert e["M"][t][c, np.argmax(np.abs(e["M"][t][c]))] > 0


def test_clouds_closed_form():
    """generate_phi: A_sc left of the front x + T dxy <= l, 0 beyond the ramp of width
    0.2 dom_size, linear in between; T = t/2 moves the front half a cell per slice."""
    g, dxy = 20, 0.1
    s = envgen.clouds(g, 12, dxy, 3.0)
    xs = (0.5 + np.arange(g)) * dxy
    l, dom = xs[-1] / 2, xs[-1] + dxy
    for t in range(12):
        xt = xs + (t / 2) * dxy
        want = np.where(xt <= l, 3.0, np.where(xt < l + 0.2 * dom, 3.0 * (1 - (xt - l) / (0.2 * dom)), 0.0))
        np.testing.assert_allclose(s[t, 7], want, rtol=1e-6, atol=1e-6)
        assert (s[t] == s[t, 0][None, :]).all()  # no y dependence


def test_obstacle_boxes():
    m = envgen.obstacles(20, 5, 1)
    assert m.sum() == 5 * 2 * 2 and m[0, 8:10, 10:12].all()
    m = envgen.obstacles(20, 8, 2, 5, 3, 4, (2.0, 0.5))
    # t = 3: cl = floor(3 + 6) = 9, ru = floor(5 + 1.5) = 6 -> rows 6..9, cols 9..12
    assert m[3, 6:10, 9:13].all() and m[3].sum() == 16
    # t = 6: cl = 15, cr = 19 = g - 1 -> outside the interior, no box
    assert m[6].sum() == 0
