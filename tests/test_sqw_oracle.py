"""Pins of the S(q, omega) oracle (oracle/sqw_oracle.py) on the CPU.

The reference's n_of_q_w / s_of_q_w_avg_mat (s_of_q_w.py:3-20,
s_of_q_w_avg.py:5-20) are restated with numpy.fft.fft.  Checked here against
what the mathematics fixes, chosen so that a dropped dt or N T factor, a sign
of the exponent, a wrong q column or a wrong axis mean fails:
* the O(n^2) DFT definition on small random series;
* a pure tone exp(2 pi i f t / n): one line of height |A n dt|^2 / (N T) at
  frequency index f, zero elsewhere (exponent sign and normalisation);
* Parseval: sum_w S = dt^2 n sum_t |x_t|^2 / (N T);
* the column 3 i and axis mean of s_of_q_w_avg_mat on tones placed only in
  those columns.
"""
import numpy as np
import pytest

from oracle import sqw_oracle as O


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 12, 17, 30])
def test_fft_matches_dft_definition(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
    N, dt, T = 100.0, 0.05, n * 0.05
    S = O.n_of_q_w(x, N, dt, T)
    for a in range(3):
        ref = np.abs(O.dft(x[:, a]) * dt) ** 2 / (N * T)
        assert np.allclose(S[:, a], ref, rtol=1e-12, atol=1e-14 * ref.max())


def test_pure_tone_closed_form():
    n, N, dt = 64, 1000.0, 0.1
    T = n * dt
    t = np.arange(n)
    f = [3, 17, 60]
    A = [2.0, 0.5 - 1.0j, 1.5j]
    x = np.stack([A[a] * np.exp(2j * np.pi * f[a] * t / n) for a in range(3)], axis=1)
    S = O.n_of_q_w(x, N, dt, T)
    for a in range(3):
        line = abs(A[a] * n * dt) ** 2 / (N * T)
        assert S[f[a], a] == pytest.approx(line, rel=1e-12)
        rest = np.delete(S[:, a], f[a])
        assert np.all(rest < 1e-20 * line)


def test_parseval():
    rng = np.random.default_rng(7)
    n, N, dt = 1000, 500.0, 0.05
    T = n * dt
    x = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
    S = O.n_of_q_w(x, N, dt, T)
    for a in range(3):
        assert S[:, a].sum() == pytest.approx(dt * dt * n * np.sum(np.abs(x[:, a]) ** 2) / (N * T), rel=1e-12)


def test_avg_takes_column_3i_and_axis_mean():
    n_avg, Nq, n, N, dt = 2, 3, 32, 50.0, 0.1
    T = n * dt
    t = np.arange(n)
    rng = np.random.default_rng(3)
    nqt = np.zeros((n_avg, n, 3 * Nq, 3), dtype=np.complex128)
    nqt[:, :, :, :] = 1e3 * (rng.standard_normal(nqt.shape) + 1j * rng.standard_normal(nqt.shape))
    expect = np.zeros((n_avg, Nq, n))
    for j in range(n_avg):
        for i in range(Nq):
            for a in range(3):
                f = (j * 7 + i * 3 + a) % n
                amp = 1.0 + i + a
                nqt[j, :, 3 * i, a] = amp * np.exp(2j * np.pi * f * t / n)
                expect[j, i, f] += abs(amp * n * dt) ** 2 / (N * T) / 3.0
    out = O.s_of_q_w_avg_mat(nqt, Nq, np.zeros((n_avg, Nq, n)), n_avg, N, dt, T)
    assert np.allclose(out, expect, rtol=1e-12, atol=1e-12 * expect.max())
