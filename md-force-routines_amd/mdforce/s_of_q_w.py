"""Drop-in for the reference's s_of_q_w (s_of_q_w.py:3-20).

``n_of_q_w(n_q_t, N, dt, T)``: n_q_t is one q's (Npd, 3) complex series (one
column per axis); returns S(q, omega) per axis, (Npd, 3) float64, with
S = |dt FFT_t n|^2 / (N T).  Computed by the device FFT of ``mdforce.sqw``;
NumPy in, NumPy out.
"""
from __future__ import annotations

import numpy as np
import torch

from .sqw import sqw


def n_of_q_w(n_q_t, N, dt, T):
    x = n_q_t if isinstance(n_q_t, torch.Tensor) else torch.as_tensor(np.asarray(n_q_t, dtype=np.complex128))
    npd = x.shape[0]
    s = sqw(x.reshape(1, npd, 1, 3), N, dt, T, q_stride=1, nq=1, mean_axes=False)[0, 0]
    return s if isinstance(n_q_t, torch.Tensor) else s.cpu().numpy()
