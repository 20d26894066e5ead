"""Drop-in for the reference's yukawa_pp (yukawa_pp.py:6-139).

``particle_particle(kappa, G, rc, Lv, pos, acc_s_r)`` returns
(U_s_r, acc_s_r) with acc_s_r filled in place (unit charges, as the
reference).  Runs the HIP sub-cell PP kernel.
"""
from __future__ import annotations

from . import _dropin as D


def particle_particle(kappa, G, rc, Lv, pos, acc_s_r):
    N = len(pos)
    Lx, Ly, Lz = (float(v) for v in Lv)
    # PP-only engines carry a minimal mesh (the mesh is not used here)
    e = D.engine(N, 8, 8, 8, Lx, Ly, Lz, kappa, G, rc)
    pos_d = D.to_dev(pos).reshape(N, 3)
    acc, u = e.pp(pos_d)
    U = float(u.item())
    D.copy_into(acc_s_r, acc)
    return U, acc_s_r
