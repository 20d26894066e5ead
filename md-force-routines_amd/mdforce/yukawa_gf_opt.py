"""Drop-in for the reference's yukawa_gf_opt.gf_opt (yukawa_gf_opt.py:5-98).

Returns (G_k, kx_v, ky_v, kz_v, A_pm) as NumPy arrays in the reference's
shapes and fftshift order; G_k and A_pm are computed on the GPU.
"""
from __future__ import annotations

import numpy as np

from . import _dropin as D


def _mid(M):
    return M // 2 if M % 2 == 0 else (M - 1) // 2


def gf_opt(kappa, Gew, p, mx_max, my_max, mz_max, Mx, My, Mz, hx, hy, hz, Lx, Ly, Lz):
    Mx, My, Mz = int(Mx), int(My), int(Mz)
    rc = min(float(Lx), float(Ly), float(Lz)) / 3.0
    e = D.engine(1, Mx, My, Mz, Lx, Ly, Lz, kappa, Gew, rc)
    G, apm = e.build_green(p=int(p), m_max=(int(mx_max), int(my_max), int(mz_max)), want_full=True)
    e._green_fp = None  # this engine's Green function came from here
    G_k = G.cpu().numpy()
    A_pm = float(apm.item())
    kx_v = 2 * np.pi * (np.arange(Mx).reshape((1, Mx)) - _mid(Mx)) / Lx
    ky_v = 2 * np.pi * (np.arange(My).reshape((My, 1)) - _mid(My)) / Ly
    kz_v = 2 * np.pi * (np.arange(Mz).reshape((Mz, 1, 1)) - _mid(Mz)) / Lz
    return G_k, kx_v, ky_v, kz_v, A_pm
