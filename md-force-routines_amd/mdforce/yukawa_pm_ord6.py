"""Drop-in for the reference's yukawa_pm_ord6.particle_mesh_fft_r
(yukawa_pm_ord6.py:439-484).

Same 27 positional arguments; returns (U_f, acc_f) with acc_f filled in
place and rho_r, E_x_p, E_y_p, E_z_p left holding the charge mesh and the
interpolated field (when ``_dropin.MIRROR_SCRATCH`` is on).  p, kappa's role
in the mesh part and the m_max arguments are unused here as in the
reference; the Green function G_k carries them.
"""
from __future__ import annotations

from . import _dropin as D


def particle_mesh_fft_r(kappa, G_ew, pos, Z, N, G_k, kx_v, ky_v, kz_v, Mx, My, Mz, hx, hy, hz, Lx, Ly, Lz,
                        p, mx_max, my_max, mz_max, rho_r, acc_f, E_x_p, E_y_p, E_z_p):
    import torch
    n = int(N)
    # the particle grouping of the mesh kernels uses the PP cell list; any
    # rc <= L/3 gives the same mesh result
    rc = min(float(Lx), float(Ly), float(Lz)) / 3.0
    e = D.engine(n, Mx, My, Mz, Lx, Ly, Lz, kappa, G_ew, rc)
    D.ensure_green(e, G_k)
    pos_d = D.to_dev(pos).reshape(n, 3)
    Z_d = D.charges(e, Z, n)
    dev = e.device
    mirror = D.MIRROR_SCRATCH
    ex = torch.empty(n, dtype=torch.float64, device=dev) if mirror and E_x_p is not None else None
    ey = torch.empty(n, dtype=torch.float64, device=dev) if mirror and E_y_p is not None else None
    ez = torch.empty(n, dtype=torch.float64, device=dev) if mirror and E_z_p is not None else None
    rho = torch.empty((int(Mz), int(My), int(Mx)), dtype=torch.float64, device=dev) \
        if mirror and rho_r is not None else None
    acc, u = e.pm(pos_d, Z_d, e_x=ex, e_y=ey, e_z=ez, rho=rho)
    U_f = float(u.item())
    D.copy_into(acc_f, acc)
    if mirror:
        D.copy_into(E_x_p, ex)
        D.copy_into(E_y_p, ey)
        D.copy_into(E_z_p, ez)
        D.copy_into(rho_r, rho)
    return U_f, acc_f
