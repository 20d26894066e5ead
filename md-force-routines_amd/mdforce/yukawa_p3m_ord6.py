"""Drop-in for the reference's yukawa_p3m_ord6 (yukawa_p3m_ord6.py:6-29).

``force_pot`` keeps the reference's 33 positional arguments and returns
(U, acc) with a new acc array; the scratch arguments acc_s_r, acc_fft,
rho_r and E_x_p / E_y_p / E_z_p are left holding the PP acceleration, the
PM acceleration, the charge mesh and the interpolated field, as in the
reference (when ``_dropin.MIRROR_SCRATCH`` is on).  The work runs on the
GPU through libmdforce; the reference's per-call print (:27) is omitted.
"""
from __future__ import annotations

from . import _dropin as D


def force_pot(kappa, G, N, rc, Lv, af, uf, pos, acc, Z, G_k, kx_v, ky_v, kz_v, Mx, My, Mz, hx, hy, hz,
              Lx, Ly, Lz, p, mx_max, my_max, mz_max, acc_s_r, acc_fft, rho_r, E_x_p, E_y_p, E_z_p):
    e = D.engine(N, Mx, My, Mz, Lx, Ly, Lz, kappa, G, rc)
    D.ensure_green(e, G_k)
    pos_d = D.to_dev(pos).reshape(int(N), 3)
    Z_d = D.charges(e, Z, int(N))
    mirror = D.MIRROR_SCRATCH
    import torch
    dev = e.device
    n = int(N)
    acc_pp = torch.empty((n, 3), dtype=torch.float64, device=dev) if mirror and acc_s_r is not None else None
    acc_pm = torch.empty((n, 3), dtype=torch.float64, device=dev) if mirror and acc_fft is not None else None
    ex = torch.empty(n, dtype=torch.float64, device=dev) if mirror and E_x_p is not None else None
    ey = torch.empty(n, dtype=torch.float64, device=dev) if mirror and E_y_p is not None else None
    ez = torch.empty(n, dtype=torch.float64, device=dev) if mirror and E_z_p is not None else None
    rho = torch.empty((int(Mz), int(My), int(Mx)), dtype=torch.float64, device=dev) \
        if mirror and rho_r is not None else None
    acc_new, en = e.force(pos_d, Z_d, af=af, uf=uf, acc_pp=acc_pp, acc_pm=acc_pm, e_x=ex, e_y=ey, e_z=ez, rho=rho)
    U = float(en[0].item())
    if mirror:
        D.copy_into(acc_s_r, acc_pp)
        D.copy_into(acc_fft, acc_pm)
        D.copy_into(E_x_p, ex)
        D.copy_into(E_y_p, ey)
        D.copy_into(E_z_p, ez)
        D.copy_into(rho_r, rho)
    return U, D.like_input(acc_new, pos)


def energy_parts():
    """[U, U_short, U_fft, U_self] of the last force_pot call of any engine
    (the values the reference prints at yukawa_p3m_ord6.py:27)."""
    if not D._engines:
        return None
    e = list(D._engines.values())[-1]
    return [float(v) for v in e.energies.cpu()]
