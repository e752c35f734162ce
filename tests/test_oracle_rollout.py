"""Pins the rollout oracle (oracle/oracle.c orc_rollout, the restatement of
Post_processing/plot_e2eGPU_policy.py:376-520 and Post_processing/grid_world.py:160-230)
without the reference's Python: closed forms on dyadic grids where every operation is
exact, failure modes at known steps, and the interpolation checked against independent
linear solves. No GPU."""
import math

import numpy as np
import pytest

import oracle
from e2e_gpu_dp_amd.configs import Params

G, NT, NR = 12, 20, 3
EAST, NORTH, WEST, SOUTH = 0, 1, 2, 3  # num_ac_angles = 4: angles 0, pi/2, pi, 3pi/2


def _params(F=0.25, nt=NT, term=(5, 7), tsg=1, na=4):
    # dx = dy = 0.25, dt = 1: F = 0.25 moves one cell per step, F = 0.125 half a cell
    return Params(gsize=G, nt=nt, nrzns=NR, nmodes=1, num_ac_speeds=1, num_ac_angles=na, F=F, dt=1.0,
                  dx=0.25, dy=0.25, x0=0.125, y0=0.125, i_term=term[0], j_term=term[1],
                  term_subgrid_size=tsg)


def _fields(nt=NT, s=0.5, mask=None):
    z = np.zeros((nt, G, G), np.float32)
    return {"u": z, "v": z, "ui": np.zeros((nt, 1, G, G), np.float32), "vi": np.zeros((nt, 1, G, G), np.float32),
            "yi": np.zeros((nt, NR, 1), np.float32), "s": np.full((nt, G, G), s, np.float32),
            "mask": np.zeros((nt, G, G), np.int32) if mask is None else mask}


def _policy(a, nt=NT):
    return np.full(nt * G * G, a, np.int32)


def _cell(x, y):
    """(i, j) of a position on the 0.25 grid (cell centres 0.125 + 0.25 k), ties to the
    upper index as the snap of move_exact does (grid_world.py:207-214)."""
    return G - 1 - int(math.floor((y - 0.125) / 0.25 + 0.5)), int(math.floor((x - 0.125) / 0.25 + 0.5))


@pytest.mark.parametrize("interp", [False, True])
def test_straight_run_closed_form(interp):
    """Zero field, every cell heading east, one cell per step: the terminal 5 cells
    east is reached after exactly 5 moves; energy_cons = 5 F^2, energy_col = 5 c (s = c
    everywhere), net = 5 (F^2 - c); the positions are x0 + k dx exactly."""
    p, f = _params(), _fields()
    r = oracle.rollout(p, f, _policy(EAST), [(5, 2)], interp=interp, want_xy=True)
    assert (r["success"] == 1).all() and (r["travel"] == 5).all() and r["bad"] == 0
    np.testing.assert_array_equal(r["energy"], np.tile([5 * 0.0625, 5 * 0.5, 5 * (0.0625 - 0.5)], (NR, 1)))
    xs = 0.125 + 0.25 * 2 + 0.25 * np.minimum(np.arange(NT + 1), 5)
    np.testing.assert_array_equal(r["xy"][:, :, 0], np.tile(xs, (NR, 1)))
    np.testing.assert_array_equal(r["xy"][:, :, 1], 0.125 + 0.25 * (G - 1 - 5))


def test_half_cell_steps_round_half_up():
    """F dt = dx/2: a position exactly half-way between two centres snaps to the upper
    cell (remx >= 0.5 dj, grid_world.py:207-214), so the cells run j0, j0+1, j0+1,
    j0+2: a terminal 2 cells east is reached after 3 moves."""
    p = _params(F=0.125, term=(5, 4))
    r = oracle.rollout(p, _fields(), _policy(EAST), [(5, 2)], interp=False, want_xy=True)
    assert (r["success"] == 1).all() and (r["travel"] == 3).all()
    cells = [_cell(x, y)[1] for x, y in r["xy"][0, :4]]
    assert cells == [2, 3, 3, 4]


def test_edge_obstacle_and_time_failures():
    p, f = _params(term=(5, 10)), _fields()
    # heading north from row 3: rows 2, 1, 0 -> row 0 is an edge state after 3 moves
    r = oracle.rollout(p, f, _policy(NORTH), [(3, 4)], interp=False)
    assert (r["success"] == 0).all() and (r["travel"] == 3).all()
    # an obstacle cell 2 east of the start at t = 2 ends the run there
    mask = np.zeros((NT, G, G), np.int32)
    mask[2, 5, 4] = 1
    r = oracle.rollout(p, _fields(mask=mask), _policy(EAST), [(5, 2)], interp=False)
    assert (r["success"] == 0).all() and (r["travel"] == 2).all()
    # the same obstacle one step later is never met: the run reaches the terminal
    mask = np.zeros((NT, G, G), np.int32)
    mask[3, 5, 4] = 1
    r = oracle.rollout(p, _fields(mask=mask), _policy(EAST), [(5, 2)], interp=False)
    assert (r["success"] == 1).all() and (r["travel"] == 8).all()
    # out of time: nt = 6 allows 5 moves, the terminal is 8 cells away
    p6 = _params(nt=6, term=(5, 10))
    r = oracle.rollout(p6, _fields(nt=6), _policy(EAST, nt=6), [(5, 2)], interp=False)
    assert (r["success"] == 0).all() and (r["travel"] == 5).all()
    # a start inside the terminal subgrid succeeds without moving (one counted step)
    r = oracle.rollout(p, f, _policy(EAST), [(5, 10)], interp=False)
    assert (r["success"] == 1).all() and (r["travel"] == 1).all()


def test_uniform_policy_interpolation_is_the_cell_action():
    """Equal action vectors at the four supports interpolate to that vector exactly
    here (dyadic positions), so interpolated and grid-policy runs coincide."""
    p = _params(F=0.125, term=(2, 9), tsg=2)
    f = _fields()
    for a in (EAST, NORTH):
        r0 = oracle.rollout(p, f, _policy(a), [(6, 3), (8, 8)], interp=False, want_xy=True)
        r1 = oracle.rollout(p, f, _policy(a), [(6, 3), (8, 8)], interp=True, want_xy=True)
        np.testing.assert_array_equal(r0["travel"], r1["travel"])
        np.testing.assert_allclose(r0["xy"], r1["xy"], rtol=0, atol=1e-15)


def _bilinear_by_solve(corners, vals, x, y):
    """The unique a + b x + c y + d x y through four corner values, evaluated at (x, y)."""
    A = np.array([[1, cx, cy, cx * cy] for cx, cy in corners], np.float64)
    coef = np.linalg.solve(A, np.asarray(vals, np.float64))
    return coef @ np.array([1, x, y, x * y])


def _plane_by_solve(corners, vals, x, y):
    A = np.array([[1, cx, cy] for cx, cy in corners], np.float64)
    coef = np.linalg.solve(A, np.asarray(vals, np.float64))
    return coef @ np.array([1, x, y])


def _second_step_velocity(p, pol, start, flow_v=0.0, interp=True):
    """(x, y) after the first (grid-policy) move and the vehicle's own velocity on the
    second move, which uses the interpolated action; the flow is a uniform (0, flow_v)."""
    f = _fields()
    f["v"] = np.full((NT, G, G), flow_v, np.float32)
    r = oracle.rollout(p, f, pol, [start], interp=interp, want_xy=True)
    xy = r["xy"][0]
    return xy[1], (xy[2] - xy[1]) / p.dt - np.array([0.0, flow_v])


@pytest.mark.parametrize("flow_v", [0.046875, -0.046875])
def test_bilinear_interpolation_matches_linear_solve(flow_v):
    """After a 3/8-cell move east from (6, 3) with a vertical flow of 3/16 cell per step,
    the vehicle is off-centre in x (supports in column 4) and in y (supports in row 5 for
    an upward offset, row 7 for a downward one, f() :31-35). The four supports hold four
    different actions; the next move's velocity is their bilinear interpolant, found
    here by solving for a + bx + cy + dxy."""
    p = _params(F=0.09375, term=(0, 0), tsg=1)
    pol = _policy(EAST).reshape(NT, G, G)
    pol[:, :, 4:] = NORTH
    pol[:, 5, 3], pol[:, 5, 4] = SOUTH, WEST
    pol[:, 7, 3], pol[:, 7, 4] = NORTH, SOUTH
    (x, y), (vx, vy) = _second_step_velocity(p, pol.ravel(), (6, 3), flow_v)
    assert _cell(x, y) == (6, 3)  # both offsets below half a cell
    xs = lambda jj: 0.125 + 0.25 * jj
    ys = lambda ii: 0.125 + 0.25 * (G - 1 - ii)
    row = 5 if flow_v > 0 else 7
    F = 0.09375
    vec = {EAST: (F, 0.0), NORTH: (0.0, F), WEST: (-F, 0.0), SOUTH: (0.0, -F)}
    sup = [(6, 3), (row, 3), (6, 4), (row, 4)]
    corners = [(xs(j), ys(i)) for i, j in sup]
    vals = [vec[int(pol[1, i, j])] for i, j in sup]
    ex = _bilinear_by_solve(corners, [v[0] for v in vals], x, y)
    ey = _bilinear_by_solve(corners, [v[1] for v in vals], x, y)
    # the move applies v cos(theta), v sin(theta) of the interpolated (v, theta); the
    # table's cos(pi/2) etc. are 6e-17-sized, hence the tolerance
    np.testing.assert_allclose([vx, vy], [ex, ey], rtol=0, atol=1e-15)
    assert abs(ex) > 1e-3 and abs(ey) > 1e-3  # both components exercised


def test_triangle_rule_next_to_a_diagonal_terminal():
    """When only the diagonal support is the terminal cell (term_status [0,0,0,1]) the
    reference interpolates on the triangle of the other three supports (:152-154): the
    plane through them, found by a 3x3 solve."""
    # start (6, 3) moving east by 3/8 cell and up by 0: diagonal support (5, 4) terminal
    p = _params(F=0.09375, term=(5, 4), tsg=1)
    pol = _policy(EAST).reshape(NT, G, G)
    pol[:, 5, :] = NORTH  # the row above heads north
    (x, y), (vx, vy) = _second_step_velocity(p, pol.ravel(), (6, 3))
    xs = lambda jj: 0.125 + 0.25 * jj
    ys = lambda ii: 0.125 + 0.25 * (G - 1 - ii)
    tri = [(xs(3), ys(6)), (xs(3), ys(5)), (xs(4), ys(6))]  # (i,j), (i-1,j), (i,j+1)
    ax = [0.09375, 0.0, 0.09375]
    ay = [0.0, 0.09375, 0.0]
    ex = _plane_by_solve(tri, ax, x, y)
    ey = _plane_by_solve(tri, ay, x, y)
    np.testing.assert_allclose([vx, vy], [ex, ey], rtol=0, atol=1e-15)


def test_rejected_support_pattern_is_reported():
    """With no terminal subgrid (tsg 0) a vehicle whose own cell is endpos meets the
    pattern [1,0,0,0] the reference rejects (:158-161)."""
    p = _params(F=0.09375, term=(6, 3), tsg=0)
    r = oracle.rollout(p, _fields(), _policy(EAST), [(6, 3)], interp=True)
    assert r["bad"] == NR and (r["success"] == -1).all()


def test_action_table_angles_are_python_rounded():
    """angle_list: round(i*2*pi/na, 14) (grid_world.py:351-352), products with glibc."""
    for na in (4, 8, 16):
        p = _params(na=na)
        speed, angle, ax, ay = oracle.rollout_actions(p)
        want = [round(i * 2 * math.pi / na, 14) for i in range(na)]
        assert list(angle) == want
        np.testing.assert_array_equal(ax, speed * np.cos(angle))
        np.testing.assert_array_equal(ay, speed * np.sin(angle))
    # two speeds: (i+1) * F / 2 in float32, speed-major (itertools.product)
    p = Params(gsize=G, nt=NT, nrzns=NR, nmodes=1, num_ac_speeds=2, num_ac_angles=4, F=0.3, dt=1.0,
               dx=0.25, dy=0.25)
    speed = oracle.rollout_actions(p)[0]
    half = np.float32(0.3) / np.float32(2)
    np.testing.assert_array_equal(speed, [half] * 4 + [np.float32(2) * half] * 4)
