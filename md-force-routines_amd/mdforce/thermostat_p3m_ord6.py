"""Drop-in for the reference's thermostat_p3m_ord6.vscale
(thermostat_p3m_ord6.py:5-25).

Same 42 positional arguments and return (pos, vel, acc, U): one Verlet step,
then K = kf sum v^2, T = K/(kf N) and the velocity rescale (hard rescale for
it <= 1999, sqrt((20 T_d/T - 1)/20) after), all on the GPU.  The
reference's per-step print of (dt*it, T) is available as ``last_T()``.
"""
from __future__ import annotations

import torch

from . import _dropin as D
from .velocity_verlet_p3m_ord6 import _update_dev

_last_T = None


def vscale(pos, vel, acc, T_desired, PBC, it, dt, N, d, Lv, Lmax_v, Lmin_v, kappa, G, rc, af, uf, kf, Z, G_k,
           kx_v, ky_v, kz_v, Mx, My, Mz, hx, hy, hz, Lx, Ly, Lz, p, mx_max, my_max, mz_max, acc_s_r, acc_fft,
           rho_r, E_x_p, E_y_p, E_z_p):
    global _last_T
    pos_d, vel_d, acc_d, U, e = _update_dev(pos, vel, acc, dt, PBC, kappa, N, Lv, Lmax_v, Lmin_v, G, rc, af, uf, Z,
                                           G_k, Mx, My, Mz, Lx, Ly, Lz,
                                           (acc_s_r, acc_fft, rho_r, E_x_p, E_y_p, E_z_p))
    ksum = torch.sum(vel_d * vel_d).reshape(1)
    t_out = torch.zeros(1, dtype=torch.float64, device=e.device)
    e.thermostat(vel_d, ksum, T_desired, int(it), t_out)
    _last_T = t_out
    return D.like_input(pos_d, pos), D.like_input(vel_d, vel), D.like_input(acc_d, pos), U


def last_T():
    """Temperature T = K/(kf N) before the rescale of the last vscale call."""
    return None if _last_T is None else float(_last_T.item())
