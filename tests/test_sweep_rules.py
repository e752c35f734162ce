"""Pins, on the CPU, two integer rules the kernels rely on (restated here, checked
exhaustively or against a brute-force recursion; no GPU):

- the key split of the per-row build kernel's emit: for a key kk < 256 and a key width
  1 <= Wj <= 256, q = (kk * ceil(2^16 / Wj)) >> 16 equals kk // Wj, with both products
  inside the 24-bit multiplier's range;
- the active slices of a sweep (active_slices): at sweep k a built model's sweep may
  skip slices t >= active_slices(k, nt), because those hold their final values in both
  value buffers. Checked against the Jacobi recursion of a chain of slices, where slice
  t's value at sweep k depends on slice t+1's at sweep k-1."""
import numpy as np


def active_slices(k, nt):
    if k <= 2:
        return nt
    if k >= nt + 1:
        return 1
    return nt - k + 2


def test_key_split_is_exact_for_every_key_and_width():
    for wj in range(1, 257):
        rec = (65536 + wj - 1) // wj
        assert rec < 1 << 24
        kk = np.arange(256)
        q = (kk * rec) >> 16
        assert (kk * rec < 1 << 24).all()
        assert np.array_equal(q, kk // wj) and np.array_equal(kk - q * wj, kk % wj)


def test_active_slices_stay_in_range():
    for nt in (1, 2, 3, 7, 20, 120):
        for k in list(range(0, 3 * nt + 5)) + [2 ** 31, 2 ** 32 - 1]:
            assert 1 <= active_slices(k, nt) <= nt


def test_skipped_slices_are_final_in_both_buffers():
    """Jacobi sweeps over a chain: V_k(t) = f(t, V_{k-1}(t+1)) for t < nt-1, V_k(nt-1) =
    f(nt-1) (no entries). Sweep k writes buffer k % 2. A slice skipped at sweep k must hold
    V_k(t) in buffer k % 2 already: it was last written at sweep k-2 or before with the
    same value, which is what the rule guarantees."""
    rng = np.random.default_rng(3)
    for nt in (1, 2, 3, 5, 9, 16):
        a = rng.normal(size=nt)
        f = lambda t, nxt: a[t] + 0.5 * nxt  # any function of the successor
        full = [np.zeros(nt)]  # V_0
        for k in range(1, 3 * nt + 4):
            prev = full[-1]
            v = np.array([f(t, prev[t + 1]) if t < nt - 1 else a[t] for t in range(nt)])
            full.append(v)
        buf = [np.zeros(nt), np.full(nt, np.nan)]  # buffer 1 is unwritten before sweep 1
        for k in range(1, 3 * nt + 4):
            act = active_slices(k, nt)
            rd, wr = buf[(k - 1) % 2], buf[k % 2]
            for t in range(act):
                wr[t] = f(t, rd[t + 1]) if t < nt - 1 else a[t]
            assert np.array_equal(wr, full[k]), (nt, k)
