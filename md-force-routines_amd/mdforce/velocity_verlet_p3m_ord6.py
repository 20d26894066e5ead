"""Drop-in for the reference's velocity_verlet_p3m_ord6.update
(velocity_verlet_p3m_ord6.py:5-26).

Same 39 positional arguments and return (pos, vel, acc, U) with new arrays;
kick, drift, single wrap, force and kick all run on the GPU.
"""
from __future__ import annotations

from . import _dropin as D
from . import yukawa_p3m_ord6


def _update_dev(pos, vel, acc, dt, PBC, kappa, N, Lv, Lmax_v, Lmin_v, G, rc, acc_fact, U_fact, Z, G_k, Mx, My,
                Mz, Lx, Ly, Lz, scratch):
    n = int(N)
    e = D.engine(n, Mx, My, Mz, Lx, Ly, Lz, kappa, G, rc)
    D.ensure_green(e, G_k)
    pos_d = D.to_dev(pos).reshape(n, 3).clone()
    vel_d = D.to_dev(vel).reshape(n, 3).clone()
    acc_d = D.to_dev(acc).reshape(n, 3)
    lmin = [float(v) for v in Lmin_v]
    lmax = [float(v) for v in Lmax_v]
    e.kick_drift(pos_d, vel_d, acc_d, dt, pbc=1 if int(PBC) == 1 else 0, lmin=lmin, lmax=lmax)
    acc_s_r, acc_fft, rho_r, E_x_p, E_y_p, E_z_p = scratch
    # force through the force_pot drop-in so the scratch mirrors behave the same
    U, acc_new = yukawa_p3m_ord6.force_pot(kappa, G, N, rc, Lv, acc_fact, U_fact, pos_d, None, Z, G_k, None, None,
                                          None, Mx, My, Mz, None, None, None, Lx, Ly, Lz, None, None, None, None,
                                          acc_s_r, acc_fft, rho_r, E_x_p, E_y_p, E_z_p)
    e.kick(vel_d, acc_new, dt)
    return pos_d, vel_d, acc_new, U, e


def update(pos, vel, acc, dt, PBC, kappa, N, d, Lv, Lmax_v, Lmin_v, G, rc, acc_fact, U_fact, Z, G_k, kx_v, ky_v,
           kz_v, Mx, My, Mz, hx, hy, hz, Lx, Ly, Lz, p, mx_max, my_max, mz_max, acc_s_r, acc_fft, rho_r, E_x_p,
           E_y_p, E_z_p):
    pos_d, vel_d, acc_d, U, _ = _update_dev(pos, vel, acc, dt, PBC, kappa, N, Lv, Lmax_v, Lmin_v, G, rc, acc_fact,
                                           U_fact, Z, G_k, Mx, My, Mz, Lx, Ly, Lz,
                                           (acc_s_r, acc_fft, rho_r, E_x_p, E_y_p, E_z_p))
    return D.like_input(pos_d, pos), D.like_input(vel_d, vel), D.like_input(acc_d, pos), U
