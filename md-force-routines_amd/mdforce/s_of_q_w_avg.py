"""Drop-in for the reference's s_of_q_w_avg (s_of_q_w_avg.py:5-20).

``s_of_q_w_avg_mat(n_q_t_array, Nq, sqw_mat, N_avg, N, dt, T)`` fills
sqw_mat[j, i, :] with S(q_i, omega) of run j averaged over the three axes,
q_i being column 3 i of run j's n_q_t (Npd, 3 Nq, 3), and returns it.  All
N_avg * Nq * 3 series are transformed in one device call (``mdforce.sqw``).
"""
from __future__ import annotations

import numpy as np
import torch

from .sqw import sqw


def s_of_q_w_avg_mat(n_q_t_array, Nq, sqw_mat, N_avg, N, dt, T):
    x = n_q_t_array if isinstance(n_q_t_array, torch.Tensor) else torch.as_tensor(
        np.asarray(n_q_t_array[:N_avg], dtype=np.complex128))
    s = sqw(x[:N_avg], N, dt, T, q_stride=3, nq=int(Nq), mean_axes=True)
    if isinstance(sqw_mat, torch.Tensor):
        sqw_mat[:N_avg, :Nq].copy_(s)
    else:
        sqw_mat[:N_avg, :Nq, :] = s.cpu().numpy()
    return sqw_mat
