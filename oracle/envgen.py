"""CPU restatement of the reference's input generation
(src/data_input/build_environment_DG_with_westward_clouds.py) at refinement degree 1.

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's cpu_baseline leg. The
product never imports it.

Arithmetic is numpy float64 as in the reference; the truncated SVD is numpy's LAPACK
(np.linalg.svd, the reference's own call :227-236); modes are signed as the product signs
them (first largest-magnitude entry positive), since LAPACK's sign is arbitrary.
"""
from __future__ import annotations

import math

import numpy as np


def realizations(gsize, nt, n_wsamples, dt, dxy, A, eps, wx, wy, w_lo, w_hi, t_slices=None):
    """generate_R (:156-175) vectorized: R_vx, R_vy, phi of shape (len(t_slices), nw, n)."""
    ts = np.arange(nt) if t_slices is None else np.asarray(t_slices)
    g = gsize
    xs = (0.5 + np.arange(g)) * dxy                      # get_xs_ys :131-132
    col = np.arange(g * g)
    i, j = col // g, col % g
    x = xs[j]                                            # get_ij_from_xy :152-155
    y = xs[g - 1 - i]
    k = np.arange(n_wsamples)
    w = w_lo + (k / (n_wsamples - 1)) * (w_hi - w_lo)    # sample_w :144-150
    t = (ts * dt)[:, None, None]
    swt = np.sin(w[None, :, None] * t)
    f = eps * swt * (x ** 2) + (1 - 2 * eps * swt) * x   # f :123-126
    fx = (2 * eps * swt * (x - 1)) + 1                   # fx :128-130
    vx = -(0.5 * A * np.pi * np.sin(wx * f) * np.cos(wy * y))     # vx_vy :137-142
    vy = A * np.pi * np.sin(wy * y) * fx * np.cos(wx * f)
    phi = A * np.sin(np.pi * f) * np.sin(wy * y)         # phi :134-135
    return vx, vy, phi


def realization_scalar(gsize, dt, dxy, A, eps, wx, wy, w_lo, w_hi, n_wsamples, tid, k, i, j):
    """One element of generate_R with the reference's scalar math (math.sin/cos), for
    checking the vectorized form."""
    t = tid * dt
    r = k / (n_wsamples - 1)
    w = w_lo + r * (w_hi - w_lo)
    x = (0.5 + j) * dxy
    y = (0.5 + (gsize - 1 - i)) * dxy
    f = eps * math.sin(w * t) * (x ** 2) + (1 - 2 * eps * math.sin(w * t)) * x
    fx = (2 * eps * math.sin(w * t) * (x - 1)) + 1
    vx = -(0.5 * A * math.pi * math.sin(wx * f) * math.cos(wy * y))
    vy = A * math.pi * math.sin(wy * y) * fx * math.cos(wx * f)
    phi = A * math.sin(math.pi * f) * math.sin(wy * y)
    return vx, vy, phi


def _mean_rows(R):
    """np.mean over realizations, added in realization order (:217)."""
    s = np.zeros(R.shape[1:], np.float64)
    for k in range(R.shape[0]):
        s += R[k]
    return s / R.shape[0]


def sign_modes(C, M):
    """Each mode's first largest-magnitude entry made positive, its coefficients with it."""
    C, M = C.copy(), M.copy()
    for c in range(M.shape[0]):
        if M[c, np.argmax(np.abs(M[c]))] < 0:
            M[c] = -M[c]
            C[:, c] = -C[:, c]
    return C, M


def modes(X, n_modes):
    """compute_modes_and_coeffs per slice (:209-236): np.linalg.svd, then rank_reduction
    (:199-207): C = U[:, :k] diag(s[:k]), M = Vh[:k]. Returns C, M, s (all float64)."""
    u, s, vh = np.linalg.svd(X, full_matrices=False)
    C = u[:, :n_modes] * s[:n_modes]
    M = vh[:n_modes]
    C, M = sign_modes(C, M)
    return C, M, s


def clouds(gsize, nt, dxy, A_sc):
    """DG_scalar_field.generate_phi (:59-99) at refinement degree 1, float32."""
    xs = [(0.5 + i) * dxy for i in range(gsize)]
    out = np.empty((nt, gsize, gsize), np.float32)
    for tid in range(nt):
        T = tid * 1 / 2
        l = xs[-1] / 2
        dom = xs[-1] + dxy
        row = []
        for x in xs:
            if (x + (T * dxy)) <= l:
                f = 1
            elif l < (x + (T * dxy)) < l + (0.2 * dom):
                slope = (1 - 0) / (-0.2 * dom)
                f = 1 + (slope * (x - (l - T * dxy)))
            else:
                f = 0
            row.append(A_sc * f)
        out[tid] = np.asarray(row, np.float32)[None, :]
    return out


def obstacles(gsize, nt, mode, ru_i=0, cl_i=0, width=0, speed=(0.0, 0.0)):
    """fill_obstacles (:470-505)."""
    m = np.zeros((nt, gsize, gsize), np.int32)
    if mode == 1:
        m[:, int(0.4 * gsize):int(0.5 * gsize), int(0.5 * gsize):int(0.6 * gsize)] = 1
    elif mode == 2:
        for t in range(nt):
            cl = math.floor(cl_i + (t * speed[0]))
            cr = int(cl + width)
            ru = math.floor(ru_i + (t * speed[1]))
            rd = int(ru + width)
            if (0 < cl < gsize - 1) and (0 < cr < gsize - 1) and (0 < ru < gsize - 1) and (0 < rd < gsize - 1):
                m[t, ru:rd, cl:cr] = 1
    return m


def environment(gsize, nt, n_wsamples, n_modes, dt, dxy, A, eps, wx, wy, w_range, A_sc, obstacle_mode=0,
                obs=(0, 0, 0, (0.0, 0.0)), phi=False):
    """The whole generator: dict of float32 fields in the builder's layouts plus the
    float64 singular values, C and M per slice (for checks)."""
    g, n, k = gsize, gsize * gsize, n_modes
    vx, vy, ph = realizations(gsize, nt, n_wsamples, dt, dxy, A, eps, wx, wy, w_range[0], w_range[1])
    out = {"u": np.empty((nt, g, g), np.float32), "v": np.empty((nt, g, g), np.float32),
           "ui": np.empty((nt, k, g, g), np.float32), "vi": np.empty((nt, k, g, g), np.float32),
           "yi": np.empty((nt, n_wsamples, k), np.float32), "singular_values": np.empty((nt, k)),
           "C": np.empty((nt, n_wsamples, k)), "M": np.empty((nt, k, 2 * n)), "mean": np.empty((nt, 2 * n)),
           "s_full": np.zeros((nt, min(n_wsamples, 2 * n)))}
    if phi:
        out.update(phi=np.empty((nt, g, g), np.float32), phi_i=np.empty((nt, k, g, g), np.float32),
                   phi_yi=np.empty((nt, n_wsamples, k), np.float32))
    for t in range(nt):
        R = np.concatenate((vx[t], vy[t]), axis=1)        # :211-212
        mu = _mean_rows(R)
        C, M, s = modes(R - mu, k)
        out["mean"][t], out["C"][t], out["M"][t] = mu, C, M
        out["singular_values"][t] = s[:k]
        out["s_full"][t] = s
        out["u"][t] = mu[:n].reshape(g, g)                # reshape_matrices_for_square_grid :262-281
        out["v"][t] = mu[n:].reshape(g, g)
        out["ui"][t] = M[:, :n].reshape(k, g, g)
        out["vi"][t] = M[:, n:].reshape(k, g, g)
        out["yi"][t] = C
        if phi:
            pm = _mean_rows(ph[t])
            Cp, Mp, _ = modes(ph[t] - pm, k)
            out["phi"][t] = pm.reshape(g, g)
            out["phi_i"][t] = Mp.reshape(k, g, g)
            out["phi_yi"][t] = Cp
    out["s"] = clouds(gsize, nt, dxy, A_sc)
    ru, cl, wd, sp = obs
    out["mask"] = obstacles(gsize, nt, obstacle_mode, ru, cl, wd, sp)
    return out
