"""Named workloads: problem parameters plus the seeded field recipe.

Parameter names follow src/input_to_build_model.h:3-48. Every workload uses dx = dy =
dt = 1, x0 = y0 = 0.5, one speed F = 1, reward method 0 ("time", -dt),
r_outbound = -100, r_terminal = +100, window m = 21, and one terminal cell
(term_subgrid_size 1) at a seeded interior location.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from .fields import make_fields


@dataclasses.dataclass
class Params:
    gsize: int
    nt: int
    nrzns: int
    nmodes: int
    num_ac_speeds: int = 1
    num_ac_angles: int = 8
    i_term: int = 0
    j_term: int = 0
    term_subgrid_size: int = 1
    reward_type: int = 0
    neighb_gsize: int = 21
    is_stationary: int = 0
    F: float = 1.0
    dt: float = 1.0
    dx: float = 1.0
    dy: float = 1.0
    x0: float = 0.5
    y0: float = 0.5
    r_outbound: float = -100.0
    r_terminal: float = 100.0
    alpha: float = 0.0
    z: float = -9999.0

    @property
    def na(self) -> int:
        return self.num_ac_speeds * self.num_ac_angles

    @property
    def nc(self) -> int:
        return self.gsize * self.gsize

    @property
    def ns(self) -> int:
        return self.nc * self.nt

    def replace(self, **kw) -> "Params":
        return dataclasses.replace(self, **kw)


def _goal(g: int, seed: int):
    rng = np.random.default_rng(seed + 77)
    return int(rng.integers(g // 4, 3 * g // 4)), int(rng.integers(g // 4, 3 * g // 4))


@dataclasses.dataclass
class Workload:
    name: str
    params: Params
    seed: int
    sigmas: list
    obstacles: list = None   # (t0, t1, i0, i1, j0, j1) boxes of obstacle_mask
    with_scalar: bool = False  # a scalar field (all_s_mat) for reward methods 2 and 3

    def fields(self):
        p = self.params
        return make_fields(p.gsize, p.nt, p.nmodes, p.nrzns, self.seed, self.sigmas, F=p.F,
                           with_scalar=self.with_scalar, obstacles=self.obstacles)

    @property
    def lanes(self) -> int:
        p = self.params
        return (p.nt - 1) * p.nc * p.na * p.nrzns


def _islands(g: int, nt: int, seed: int):
    """Obstacle boxes for the obstacle variants: three fixed islands of about g/10 cells
    a side, and one that is present only in the second half of the horizon, away from
    the goal."""
    rng = np.random.default_rng(seed + 91)
    side = max(2, g // 10)
    boxes = []
    for k in range(4):
        i0 = int(rng.integers(1, g - side - 1))
        j0 = int(rng.integers(1, g - side - 1))
        t0, t1 = (0, nt) if k < 3 else (nt // 2, nt)
        boxes.append((t0, t1, i0, i0 + side, j0, j0 + side))
    return boxes


def workload(name: str) -> Workload:
    """C1 (small, CPU-runnable), C2 (100^2 x 120, the single-GPU bench case),
    C4 (200^2 x 200, large); C3 and C5 reuse these builders. "C2-obs" is C2 with obstacle
    islands (the goes_through_obstacle path, src/build_model.cu:444-477) and "C2-rad" is
    C2 with reward method 2 over a scalar field (:244-250): the two build paths the
    named workloads leave unexercised. "C2-alpha" is C2 with reward method 4 and 2 speeds
    x 8 headings (16 actions: F/2 and F), for the alpha sweep of src/my_script.sh:10-19."""
    if name == "C2-alpha":  # the alpha sweep: reward method 4, 2 speeds x 8 headings
        w = workload("C2")
        return Workload(name, w.params.replace(reward_type=4, num_ac_speeds=2, num_ac_angles=8), w.seed,
                        w.sigmas)
    if name in ("C2-obs", "C2-rad"):
        w = workload("C2")
        p = w.params
        if name == "C2-obs":
            boxes = _islands(p.gsize, p.nt, w.seed)
            gi, gj = p.i_term, p.j_term
            boxes = [b for b in boxes if not (b[2] <= gi < b[3] and b[4] <= gj < b[5])]
            return Workload(name, p, w.seed, w.sigmas, obstacles=boxes)
        return Workload(name, p.replace(reward_type=2), w.seed, w.sigmas, with_scalar=True)
    if name == "C1":
        g, nt, na, nr, nm, seed = 25, 20, 8, 1000, 4, 1234
        sig = [1.0, 0.5, 0.25, 0.125]
    elif name == "C2":
        g, nt, na, nr, nm, seed = 100, 120, 16, 2500, 5, 2345
        sig = [2.0 ** -k for k in range(nm)]
    elif name == "C4":
        g, nt, na, nr, nm, seed = 200, 200, 16, 5000, 6, 4567
        sig = [2.0 ** -k for k in range(nm)]
    else:
        raise KeyError(name)
    it, jt = _goal(g, seed)
    p = Params(gsize=g, nt=nt, nrzns=nr, nmodes=nm, num_ac_angles=na, i_term=it, j_term=jt)
    return Workload(name, p, seed, sig)


def small(g=7, nt=4, nr=40, nm=2, na=8, seed=7, **kw) -> Workload:
    """Tiny workloads for brute-force parity (the reference's own debug size is
    5x5, nt=3, nr=5, 1 mode: prel_tests/build_model.cu:944-974)."""
    it, jt = _goal(g, seed)
    base = dict(gsize=g, nt=nt, nrzns=nr, nmodes=nm, num_ac_angles=na, i_term=it, j_term=jt)
    base.update(kw)
    return Workload(f"small{g}x{nt}", Params(**base), seed, [2.0 ** -k for k in range(nm)])


def goal_cells(g: int, k: int = 64, seed: int = 3456):
    """k distinct goal cells drawn uniformly from the grid (cell index i*g + j)."""
    rng = np.random.default_rng(seed)
    return rng.choice(g * g, size=k, replace=False).astype(np.int32)
