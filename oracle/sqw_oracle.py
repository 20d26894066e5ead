"""CPU oracle of S(q, omega) (test infrastructure; never imported by the product).

NumPy restatement of the reference's post-processing:
* ``n_of_q_w``: s_of_q_w.py:3-20 -- per axis column, n(w) = FFT_t(n) * dt
  (:5-7), S = |n(w)|^2 / (N T) (:9-11), stacked as (n_w, 3) (:13-18);
* ``s_of_q_w_avg_mat``: s_of_q_w_avg.py:5-20 -- for run j and i < Nq the
  series n_q_t[:, 3 i, :] (:13), its S averaged over the axes (:15-16), into
  sqw_mat[j, i, :] (:18);
* ``dft``: the O(n^2) definition of the transform numpy.fft.fft computes,
  for pinning the FFT path on small inputs.
Pinned in tests/test_sqw_oracle.py against the O(n^2) DFT, a pure tone's
closed form and Parseval's identity.
"""
from __future__ import annotations

import numpy as np


def n_of_q_w(n_q_t, N, dt, T):
    n_q_t = np.asarray(n_q_t)
    S = np.zeros((n_q_t.shape[0], 3))
    for a in range(3):
        nw = np.fft.fft(n_q_t[:, a]) * dt
        S[:, a] = np.abs(nw) ** 2 / (N * T)
    return S


def s_of_q_w_avg_mat(n_q_t_array, Nq, sqw_mat, N_avg, N, dt, T):
    for j in range(N_avg):
        n_q_t = n_q_t_array[j]
        for i in range(Nq):
            s_q_w = n_of_q_w(n_q_t[:, i * 3, :], N, dt, T)
            sqw_mat[j, i, :] = np.mean(s_q_w, axis=1)
    return sqw_mat


def dft(x):
    x = np.asarray(x, dtype=np.complex128)
    n = len(x)
    t = np.arange(n)
    return np.array([np.sum(x * np.exp(-2j * np.pi * ((k * t) % n) / n)) for k in range(n)])
