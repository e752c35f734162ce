"""Seeded synthetic DO ("dynamically orthogonal") velocity fields.

The reference reads its fields from ``data_input/<prob_name>/*.npy`` written by
``src/data_input/build_environment_DG_with_westward_clouds.py`` (double-gyre
realizations reduced by SVD). That data is not shipped, so inputs here come from a
deterministic recipe with the same array layouts (src/build_model.cu:978-985,
index maps :380-398):

    all_u_mat, all_v_mat, all_s_mat : float32 [nt, g, g]
    all_ui_mat, all_vi_mat          : float32 [nt, nmodes, g, g]
    all_Yi                          : float32 [nt, nrzns, nmodes]
    obstacle_mask                   : int32   [nt, g, g]

Recipe (all arithmetic in float64, cast to float32 at the end):
  * streamfunction psi(x, y, t) = sum_{k<6} A_k exp(-|p - c_k(t)|^2 / (2 L_k^2)) with
    A_k ~ U[-1, 1], L_k ~ U[g/12, g/5], centres c_k uniform in the domain drifting
    linearly with t by at most g/4 over the horizon;
  * velocity u = d psi / d y, v = -d psi / d x (analytic derivatives at cell centres,
    y measured upward, row 0 = top as in get_xypos_from_ij, src/build_model.cu:199-205);
  * the mean field is scaled so max |(u, v)| over space-time is ``mean_speed`` (0.5 F);
  * mode k uses the same construction with seed + 1 + k and is scaled so that its
    root-mean-square speed over each time slice is ``mode_rms`` (0.5 F);
  * Yi[t, w, k] ~ N(0, sigma_k^2).
"""
from __future__ import annotations

import numpy as np

__all__ = ["vortex_field", "make_fields", "zero_fields"]


def _cell_centres(g: int, dx: float = 1.0, dy: float = 1.0, x0: float = 0.5, y0: float = 0.5):
    xs = x0 + np.arange(g) * dx          # column j -> x
    ys_top = y0 + (g - 1 - np.arange(g)) * dy  # row i -> y (row 0 at the top)
    return xs, ys_top


def vortex_field(g: int, nt: int, seed: int, n_vortex: int = 6):
    """Returns (u, v) float64 arrays [nt, g, g] of a drifting Gaussian-vortex flow."""
    rng = np.random.default_rng(seed)
    A = rng.uniform(-1.0, 1.0, n_vortex)
    L = rng.uniform(g / 12.0, g / 5.0, n_vortex)
    cx = rng.uniform(0.0, g, n_vortex)
    cy = rng.uniform(0.0, g, n_vortex)
    drift = rng.uniform(-1.0, 1.0, (n_vortex, 2)) * (g / 4.0) / max(nt - 1, 1)
    xs, ys = _cell_centres(g)
    u = np.zeros((nt, g, g))
    v = np.zeros((nt, g, g))
    t = np.arange(nt, dtype=np.float64)
    for k in range(n_vortex):
        cxt = cx[k] + drift[k, 0] * t            # [nt]
        cyt = cy[k] + drift[k, 1] * t
        dxk = xs[None, :] - cxt[:, None]          # [nt, g] over columns
        dyk = ys[None, :] - cyt[:, None]          # [nt, g] over rows
        gx = np.exp(-dxk ** 2 / (2 * L[k] ** 2))  # separable Gaussian
        gy = np.exp(-dyk ** 2 / (2 * L[k] ** 2))
        G = gy[:, :, None] * gx[:, None, :]       # [nt, row, col]
        # psi = A G ; u = dpsi/dy, v = -dpsi/dx
        u += A[k] * G * (-dyk[:, :, None] / L[k] ** 2)
        v -= A[k] * G * (-dxk[:, None, :] / L[k] ** 2)
    return u, v


def make_fields(g: int, nt: int, nmodes: int, nrzns: int, seed: int, sigmas=None,
                F: float = 1.0, mean_speed_frac: float = 0.5, mode_rms_frac: float = 0.5,
                with_scalar: bool = False, obstacles=None):
    """Builds the seven input arrays of the reference builder.

    ``sigmas`` defaults to 1, 1/2, 1/4, ... (sigma halving from 1.0).
    ``obstacles`` is an optional list of (t0, t1, i0, i1, j0, j1) boxes set to 1 in
    obstacle_mask for t0 <= t < t1.
    """
    if sigmas is None:
        sigmas = [2.0 ** -k for k in range(nmodes)]
    sigmas = np.asarray(sigmas, dtype=np.float64)
    assert len(sigmas) == nmodes
    u, v = vortex_field(g, nt, seed)
    vmax = np.sqrt(u ** 2 + v ** 2).max()
    scale = (mean_speed_frac * F) / vmax if vmax > 0 else 0.0
    u *= scale
    v *= scale
    ui = np.zeros((nt, nmodes, g, g))
    vi = np.zeros((nt, nmodes, g, g))
    for k in range(nmodes):
        mu, mv = vortex_field(g, nt, seed + 1 + k)
        rms = np.sqrt((mu ** 2 + mv ** 2).mean(axis=(1, 2)))  # per slice
        rms[rms == 0] = 1.0
        s = (mode_rms_frac * F) / rms
        ui[:, k] = mu * s[:, None, None]
        vi[:, k] = mv * s[:, None, None]
    rng = np.random.default_rng(seed + 1000)
    yi = rng.standard_normal((nt, nrzns, nmodes)) * sigmas[None, None, :]
    if with_scalar:
        su, sv = vortex_field(g, nt, seed + 500)
        sfield = np.sqrt(su ** 2 + sv ** 2)
        sfield /= max(sfield.max(), 1e-30)
    else:
        sfield = np.zeros((nt, g, g))
    mask = np.zeros((nt, g, g), dtype=np.int32)
    for (t0, t1, i0, i1, j0, j1) in (obstacles or []):
        mask[t0:t1, i0:i1, j0:j1] = 1
    return {
        "u": u.astype(np.float32), "v": v.astype(np.float32),
        "ui": ui.astype(np.float32), "vi": vi.astype(np.float32),
        "yi": yi.astype(np.float32), "s": sfield.astype(np.float32),
        "mask": mask,
    }


def zero_fields(g: int, nt: int, nmodes: int, nrzns: int, yi=None, u=0.0, v=0.0):
    """Uniform mean field (u, v) with zero modes (closed-form test fields)."""
    f = {
        "u": np.full((nt, g, g), u, np.float32), "v": np.full((nt, g, g), v, np.float32),
        "ui": np.zeros((nt, nmodes, g, g), np.float32),
        "vi": np.zeros((nt, nmodes, g, g), np.float32),
        "yi": np.zeros((nt, nrzns, nmodes), np.float32) if yi is None else np.asarray(yi, np.float32),
        "s": np.zeros((nt, g, g), np.float32), "mask": np.zeros((nt, g, g), np.int32),
    }
    return f
