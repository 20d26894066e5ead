"""NumPy restatement of the reference P3M Yukawa force path (FP64).

Test infrastructure only (see ``oracle/__init__.py``): it is the checker the
HIP path is compared with, never a fallback for it.

Each function follows the reference file and lines named in its docstring.
Vectorisation changes the summation order only; every formula, index map and
convention (shifted k-grid, +2h stencil offset, Coulomb self-energy constant,
single-period wrap) is the reference's.

Pinning (the checks that make this oracle trustworthy, all in ``tests/``):

* ``particle_particle``: golden accelerations of the compiled reference C
  ``particle_particle`` (``yukawa_p3m.c:182-416``, built under ``oracle/_ref``),
  an O(N^2) minimum-image brute force, Newton's third law, the kappa -> 0
  Ewald limit ``erfc(alpha r)/r`` and a two-particle closed form.
* ``bspline6_weights``: the cardinal B-spline M6 from the Cox-de Boor
  recursion, partition of unity.
* ``charge_assign``: sum(rho) == sum(Z) (the reference's own check,
  ``yukawa_p3m.c:163-175``) and single-particle tensor structure.
* ``pm_solve``: the full-spectrum FFT path equals an explicit DFT-free plane
  wave known answer, Parseval for the energy.
* ``field_gather``: exact adjointness with ``charge_assign`` and partition of
  unity on a constant field.
* ``gf_opt``: m_max = 0 reduces to R/U^2, G(k) = G(-k), and the h -> 0 limit.
* ``initial``: NumPy's legacy ``RandomState`` stream is frozen (NEP 19).
* ``density_modes``: direct sum on lattice positions with a known answer.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erfc

# --------------------------------------------------------------------------
# Seeded initial state
# --------------------------------------------------------------------------


def maxwell_velocities(rng, N, T_desired):
    """Box-Muller Maxwellian velocities with the mean removed.

    initialize_pos_vel.py:8-28 (sigma = sqrt(T/3), u1..u4, mean removal).
    """
    sig = np.sqrt(T_desired / 3)
    u1 = rng.random_sample(N)
    u2 = rng.random_sample(N)
    u3 = rng.random_sample(N)
    u4 = rng.random_sample(N)
    vel = np.empty((N, 3))
    vel[:, 0] = sig * np.sqrt(-2 * np.log(u1)) * np.cos(2 * np.pi * u2)
    vel[:, 1] = sig * np.sqrt(-2 * np.log(u1)) * np.sin(2 * np.pi * u2)
    vel[:, 2] = sig * np.sqrt(-2 * np.log(u3)) * np.cos(2 * np.pi * u4)
    # the mean of each strided column, as :21-28 take it (np.mean of vel[:, a]:
    # pairwise summation; vel.mean(axis=0) accumulates row by row and rounds
    # differently)
    for a in range(3):
        vel[:, a] = vel[:, a] - np.mean(vel[:, a])
    return vel


def uniform_positions(rng, N, Lv):
    """Uniform positions, one column at a time: initialize_pos_vel.py:31-33."""
    pos = np.empty((N, 3))
    pos[:, 0] = Lv[0] * rng.random_sample(N)
    pos[:, 1] = Lv[1] * rng.random_sample(N)
    pos[:, 2] = Lv[2] * rng.random_sample(N)
    return pos


def initial(seed_int, N, Lv, T_desired):
    """initialize_pos_vel.py:4-35: one RandomState stream, velocities first."""
    rng = np.random.RandomState(seed_int)
    vel = maxwell_velocities(rng, N, T_desired)
    pos = uniform_positions(rng, N, Lv)
    return pos, vel


def initial_split(pos_seed, vel_seed, N, Lv, T_desired):
    """Positions from RandomState(pos_seed), velocities from RandomState(vel_seed).

    Same draws as initialize_pos_vel.py:4-35, with the single stream split
    in two so positions and velocities can be seeded independently.
    """
    pos = uniform_positions(np.random.RandomState(pos_seed), N, Lv)
    vel = maxwell_velocities(np.random.RandomState(vel_seed), N, T_desired)
    return pos, vel


# --------------------------------------------------------------------------
# Optimal Green function
# --------------------------------------------------------------------------


def _mid(M):
    """yukawa_gf_opt.py:12-25: n_mid = M/2 (even) or (M-1)/2 (odd)."""
    return M // 2 if M % 2 == 0 else (M - 1) // 2


def gf_opt(kappa, Gew, p, mx_max, my_max, mz_max, Mx, My, Mz, hx, hy, hz, Lx, Ly, Lz):
    """Optimal influence function for ik-differentiation.

    yukawa_gf_opt.py:5-98.  Returns (G_k, kx_v, ky_v, kz_v, A_pm) with G_k of
    shape (Mz, My, Mx) in fftshift order (index n <-> k = 2 pi (n - n_mid)/L)
    and G(k=0) = 0.  The alias sum runs mz, my, mx from -m_max to m_max in
    the reference's order, so every mode accumulates its 343 terms in the same
    sequence as the reference loop.
    """
    kappa_sq = kappa ** 2
    Gew_sq = Gew ** 2
    nx_mid, ny_mid, nz_mid = _mid(Mx), _mid(My), _mid(Mz)
    nx_sh = (np.arange(Mx) - nx_mid).reshape(1, 1, Mx)
    ny_sh = (np.arange(My) - ny_mid).reshape(1, My, 1)
    nz_sh = (np.arange(Mz) - nz_mid).reshape(Mz, 1, 1)
    kx_v = 2 * np.pi * (np.arange(Mx).reshape((1, Mx)) - nx_mid) / Lx
    ky_v = 2 * np.pi * (np.arange(My).reshape((My, 1)) - ny_mid) / Ly
    kz_v = 2 * np.pi * (np.arange(Mz).reshape((Mz, 1, 1)) - nz_mid) / Lz
    kx = 2 * np.pi * nx_sh / Lx
    ky = 2 * np.pi * ny_sh / Ly
    kz = 2 * np.pi * nz_sh / Lz
    k_sq = kx ** 2 + ky ** 2 + kz ** 2

    def sinc1(kM, h):
        arg = 0.5 * kM * h
        safe = np.where(kM != 0.0, arg, 1.0)
        return np.where(kM != 0.0, np.sin(safe) / safe, 1.0)

    U_G_k = np.zeros((Mz, My, Mx))
    U_k_sq = np.zeros((Mz, My, Mx))
    for mz in range(-mz_max, mz_max + 1):
        kz_M = 2 * np.pi * (nz_sh + mz * Mz) / Lz
        Uz = sinc1(kz_M, hz)
        for my in range(-my_max, my_max + 1):
            ky_M = 2 * np.pi * (ny_sh + my * My) / Ly
            Uy = sinc1(ky_M, hy)
            for mx in range(-mx_max, mx_max + 1):
                kx_M = 2 * np.pi * (nx_sh + mx * Mx) / Lx
                Ux = sinc1(kx_M, hx)
                k_M_sq = kx_M ** 2 + ky_M ** 2 + kz_M ** 2
                U_k_M = (Ux * Uy * Uz) ** p
                U_k_M_sq = U_k_M ** 2
                G_k_M = 4 * np.pi * np.exp(-0.25 * (kappa_sq + k_M_sq) / Gew_sq) / (kappa_sq + k_M_sq)
                k_dot_k_M = kx * kx_M + ky * ky_M + kz * kz_M
                U_G_k = U_G_k + U_k_M_sq * G_k_M * k_dot_k_M
                U_k_sq = U_k_sq + U_k_M_sq
    nz = k_sq != 0.0
    G_k = np.zeros((Mz, My, Mx))
    G_k[nz] = U_G_k[nz] / ((U_k_sq[nz] ** 2) * k_sq[nz])
    A_pm = float(np.sum(U_G_k[nz] ** 2 / ((U_k_sq[nz] ** 2) * k_sq[nz])))
    return G_k, kx_v, ky_v, kz_v, A_pm


# --------------------------------------------------------------------------
# Order-6 charge assignment and field gather
# --------------------------------------------------------------------------


def bspline6_weights(x):
    """Order-6 assignment weights, yukawa_pm_ord6.py:29-48.

    ``x`` is the offset from the cell centre in units of h, in [-1/2, 1/2).
    Returns an array of shape x.shape + (6,).
    """
    x = np.asarray(x, dtype=np.float64)
    w = np.empty(x.shape + (6,))
    w[..., 0] = (1 - 10 * x + 40 * x ** 2 - 80 * x ** 3 + 80 * x ** 4 - 32 * x ** 5) / 3840
    w[..., 1] = (237 - 750 * x + 840 * x ** 2 - 240 * x ** 3 - 240 * x ** 4 + 160 * x ** 5) / 3840
    w[..., 2] = (841 - 770 * x - 440 * x ** 2 + 560 * x ** 3 + 80 * x ** 4 - 160 * x ** 5) / 1920
    w[..., 3] = (841 + 770 * x - 440 * x ** 2 - 560 * x ** 3 + 80 * x ** 4 + 160 * x ** 5) / 1920
    w[..., 4] = (237 + 750 * x + 840 * x ** 2 + 240 * x ** 3 - 240 * x ** 4 - 160 * x ** 5) / 3840
    w[..., 5] = (1 + 10 * x + 40 * x ** 2 + 80 * x ** 3 + 80 * x ** 4 + 32 * x ** 5) / 3840
    return w


def _wrap_index(idx, M):
    """Single-period index wrap, yukawa_pm_ord6.py:56-61 (and :67-72, :78-83)."""
    idx = np.where(idx < 0, idx + M, idx)
    return np.where(idx > M - 1, idx - M, idx)


def _stencil(pos, Mx, My, Mz, hx, hy, hz):
    """Base cell, weights and wrapped node indices: yukawa_pm_ord6.py:17-48,52-93.

    The six nodes run from the base index ix to ix+5 (the stencil is offset
    by +2h; the gather uses the same offset, so it cancels).
    """
    out = []
    for d, (M, h) in enumerate(((Mx, hx), (My, hy), (Mz, hz))):
        xd = pos[:, d]
        i0 = np.floor(xd / h).astype(np.int64)
        frac = (xd - (i0 + 0.5) * h) / h
        w = bspline6_weights(frac)
        nodes = _wrap_index(i0[:, None] + np.arange(6)[None, :], M)
        out.append((nodes, w))
    return out


def charge_assign(pos, Z, Mx, My, Mz, hx, hy, hz):
    """charge_assgn_6_numba, yukawa_pm_ord6.py:5-95: rho of shape (Mz, My, Mx)."""
    (ixn, wx), (iyn, wy), (izn, wz) = _stencil(pos, Mx, My, Mz, hx, hy, hz)
    # value order Z*wz*wy*wx as in :87
    val = ((Z[:, None, None, None] * wz[:, :, None, None]) * wy[:, None, :, None]) * wx[:, None, None, :]
    flat = (izn[:, :, None, None] * My + iyn[:, None, :, None]) * Mx + ixn[:, None, None, :]
    rho = np.bincount(flat.ravel(), weights=val.ravel(), minlength=Mx * My * Mz)
    return rho.reshape(Mz, My, Mx)


def field_gather(Ex, Ey, Ez, pos, Z, Mx, My, Mz, hx, hy, hz):
    """field_assgn_6_numba, yukawa_pm_ord6.py:212-312.

    Returns (acc, E_p) with E_p of shape (N, 3) and acc = Z * E_p (:308-310).
    """
    (ixn, wx), (iyn, wy), (izn, wz) = _stencil(pos, Mx, My, Mz, hx, hy, hz)
    w = (wz[:, :, None, None] * wy[:, None, :, None]) * wx[:, None, None, :]
    flat = (izn[:, :, None, None] * My + iyn[:, None, :, None]) * Mx + ixn[:, None, None, :]
    E_p = np.empty((pos.shape[0], 3))
    for a, Em in enumerate((Ex, Ey, Ez)):
        E_p[:, a] = np.sum(Em.ravel()[flat] * w, axis=(1, 2, 3))
    acc = Z[:, None] * E_p
    return acc, E_p


def pm_solve(rho, G_k, kx_v, ky_v, kz_v, Lx, Ly, Lz):
    """Mesh solve of particle_mesh_fft_r, yukawa_pm_ord6.py:442-480.

    numpy.fft stands in for pyFFTW, as in the reference's own earlier version
    (yukawa_pm_ord6.py:400-436).  Returns (U_f, E_x, E_y, E_z) with the real
    meshes of :478-480.
    """
    Mz, My, Mx = rho.shape
    V = Lx * Ly * Lz
    M_V = Mx * My * Mz
    rho_k = np.fft.fftshift(np.fft.fftn(rho))
    phi_k = G_k * rho_k
    rho_k_sq = rho_k.real * rho_k.real + rho_k.imag * rho_k.imag
    U_f = 0.5 * np.sum(rho_k_sq * G_k) / V
    fields = []
    for kv in (kx_v, ky_v, kz_v):
        E_k = -1j * kv * phi_k
        E = M_V * np.fft.ifftn(np.fft.ifftshift(E_k)) / V
        fields.append(np.real(E))
    return float(U_f), fields[0], fields[1], fields[2]


def particle_mesh_fft_r(kappa, G_ew, pos, Z, N, G_k, kx_v, ky_v, kz_v, Mx, My, Mz,
                        hx, hy, hz, Lx, Ly, Lz):
    """yukawa_pm_ord6.py:439-484.  Returns (U_f, acc_f, rho, E_p)."""
    rho = charge_assign(pos, Z, Mx, My, Mz, hx, hy, hz)
    U_f, Ex, Ey, Ez = pm_solve(rho, G_k, kx_v, ky_v, kz_v, Lx, Ly, Lz)
    acc_f, E_p = field_gather(Ex, Ey, Ez, pos, Z, Mx, My, Mz, hx, hy, hz)
    return U_f, acc_f, rho, E_p


# --------------------------------------------------------------------------
# Short-range particle-particle part
# --------------------------------------------------------------------------


def pp_pair(r, kappa, G):
    """Pair energy and radial force of yukawa_pp.py:115-124, as written."""
    a = 0.5 * kappa / G
    phi = (0.5 / r) * (np.exp(kappa * r) * erfc(G * r + a) + np.exp(-kappa * r) * erfc(G * r - a))
    f1 = (0.5 / r ** 2) * np.exp(kappa * r) * erfc(G * r + a) * (1 - kappa * r)
    f2 = (0.5 / r ** 2) * np.exp(-kappa * r) * erfc(G * r - a) * (1 + kappa * r)
    f3 = (G / np.sqrt(np.pi) / r) * (np.exp(-(G * r + a) ** 2) * np.exp(kappa * r)
                                     + np.exp(-(G * r - a) ** 2) * np.exp(-kappa * r))
    return phi, f1 + f2 + f3


def particle_particle(kappa, G, rc, Lv, pos, chunk_pairs=4_000_000, Z=None):
    """Linked-cell PP sum with i < j (Newton's third law), yukawa_pp.py:6-139.

    Cells: Lxd = floor(L/rc), width L/Lxd (:17-23); c = cx + cy*Lxd +
    cz*Lxd*Lyd (:46-49); the 27 neighbour cells with the periodic image shift
    of :60-94; pair test i < j and r < rc (:104-112).  Pairs are gathered
    per neighbour offset over all cells at once instead of walking the
    linked list; the pair set is identical.  A cell index of Lxd (a position
    exactly at L) is clamped into range, which the reference leaves
    unchecked.  Returns (U_s_r, acc_s_r).

    With ``Z`` given, each pair is weighted by Z_i Z_j as in the reference's
    C twin (yukawa_p3m.c:376-383); the numba path (the default) uses unit
    charges.
    """
    pos = np.asarray(pos, dtype=np.float64)
    N = pos.shape[0]
    Lx, Ly, Lz = (float(v) for v in Lv)
    nd = np.array([int(np.floor(Lx / rc)), int(np.floor(Ly / rc)), int(np.floor(Lz / rc))])
    width = np.array([Lx / nd[0], Ly / nd[1], Lz / nd[2]])
    cidx = np.floor(pos / width).astype(np.int64)
    cidx = np.clip(cidx, 0, nd - 1)
    c = cidx[:, 0] + cidx[:, 1] * nd[0] + cidx[:, 2] * nd[0] * nd[1]
    ncell = int(nd.prod())
    counts = np.bincount(c, minlength=ncell)
    maxc = int(counts.max()) if N else 0
    order = np.argsort(c, kind="stable")
    starts = np.concatenate(([0], np.cumsum(counts)[:-1]))
    rank = np.arange(N) - starts[c[order]]
    members = np.full((ncell, max(maxc, 1)), -1, dtype=np.int64)
    members[c[order], rank] = order

    cells = np.arange(ncell)
    ccx = cells % nd[0]
    ccy = (cells // nd[0]) % nd[1]
    ccz = cells // (nd[0] * nd[1])
    L3 = (Lx, Ly, Lz)

    U = 0.0
    acc = np.zeros((N, 3))
    per_cell = max(maxc, 1) ** 2
    cstep = max(1, chunk_pairs // per_cell)
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                nb = []
                shift = []
                for cc, dd, n, Ld in ((ccx, dx, nd[0], L3[0]), (ccy, dy, nd[1], L3[1]), (ccz, dz, nd[2], L3[2])):
                    cn = cc + dd
                    sh = np.where(cn < 0, -Ld, np.where(cn >= n, Ld, 0.0))
                    cn = np.where(cn < 0, cn + n, np.where(cn >= n, cn - n, cn))
                    nb.append(cn)
                    shift.append(sh)
                cN = nb[0] + nb[1] * nd[0] + nb[2] * nd[0] * nd[1]
                for c0 in range(0, ncell, cstep):
                    sl = slice(c0, min(ncell, c0 + cstep))
                    I = members[sl][:, :, None]
                    J = members[cN[sl]][:, None, :]
                    valid = (I >= 0) & (J >= 0) & (I < J)
                    if not valid.any():
                        continue
                    ci, ai, bi = np.nonzero(valid)
                    ii = members[sl][ci, ai]
                    jj = members[cN[sl]][ci, bi]
                    rx = pos[ii, 0] - (pos[jj, 0] + shift[0][sl][ci])
                    ry = pos[ii, 1] - (pos[jj, 1] + shift[1][sl][ci])
                    rz = pos[ii, 2] - (pos[jj, 2] + shift[2][sl][ci])
                    r = np.sqrt(rx ** 2 + ry ** 2 + rz ** 2)
                    m = r < rc
                    if not m.any():
                        continue
                    ii, jj, rx, ry, rz, r = ii[m], jj[m], rx[m], ry[m], rz[m], r[m]
                    phi, fr = pp_pair(r, kappa, G)
                    if Z is not None:
                        zz = Z[ii] * Z[jj]
                        phi = phi * zz
                        fr = fr * zz
                    U += float(np.sum(phi))
                    for a, comp in enumerate((rx, ry, rz)):
                        fa = fr * comp / r
                        acc[:, a] += np.bincount(ii, weights=fa, minlength=N)
                        acc[:, a] -= np.bincount(jj, weights=fa, minlength=N)
    return U, acc


def particle_particle_brute(kappa, G, rc, Lv, pos):
    """O(N^2) minimum-image sum of the same pair terms (independent check).

    Equal to ``particle_particle`` whenever rc <= L/3 per axis (no pair has
    two images within rc).
    """
    pos = np.asarray(pos, dtype=np.float64)
    L = np.asarray(Lv, dtype=np.float64)
    N = pos.shape[0]
    U = 0.0
    acc = np.zeros((N, 3))
    for i in range(N - 1):
        d = pos[i] - pos[i + 1:]
        d -= L * np.round(d / L)
        r = np.sqrt(np.sum(d * d, axis=1))
        m = r < rc
        if not m.any():
            continue
        phi, fr = pp_pair(r[m], kappa, G)
        U += float(np.sum(phi))
        f = (fr / r[m])[:, None] * d[m]
        acc[i] += f.sum(axis=0)
        acc[i + 1:][m] -= f
    return U, acc


# --------------------------------------------------------------------------
# Force API, integrator, thermostat, density modes
# --------------------------------------------------------------------------


def force_pot(kappa, G, N, rc, Lv, af, uf, pos, Z, G_k, kx_v, ky_v, kz_v, Mx, My, Mz,
              hx, hy, hz, Lx, Ly, Lz):
    """yukawa_p3m_ord6.py:6-29.

    Returns (U, acc, parts) where parts holds the intermediate buffers the
    reference leaves in its scratch arguments: acc_s_r, acc_fft, rho_r, E_p
    and the energy split (U_short, U_fft, U_self).
    """
    U_short, acc_s_r = particle_particle(kappa, G, rc, Lv, pos)
    U_fft, acc_fft, rho, E_p = particle_mesh_fft_r(kappa, G, pos, Z, N, G_k, kx_v, ky_v, kz_v,
                                                   Mx, My, Mz, hx, hy, hz, Lx, Ly, Lz)
    acc = af * (acc_s_r + acc_fft)
    U_self = -N * G / np.sqrt(np.pi)
    U = uf * (U_short + U_fft + U_self)
    parts = dict(acc_s_r=acc_s_r, acc_fft=acc_fft, rho_r=rho, E_p=E_p,
                 U_short=U_short, U_fft=U_fft, U_self=U_self)
    return U, acc, parts


class P3MParams:
    """The scalar and mesh arguments the reference threads through every call
    (yukawa_MD_p3m_ord6.py:41-79), bundled for the oracle's own callers."""

    def __init__(self, N, L, kappa, alpha, rc, M, p=6, m_max=3, af=1.0 / 3.0, uf=1.0, kf=1.5,
                 with_green=True):
        self.N = int(N)
        self.L = float(L)
        self.Lv = np.array([L, L, L], dtype=np.float64)
        self.kappa = float(kappa)
        self.G = float(alpha)
        self.rc = float(rc)
        self.M = int(M)
        self.h = self.L / self.M
        self.p = p
        self.m_max = m_max
        self.af, self.uf, self.kf = af, uf, kf
        if with_green:
            self.G_k, self.kx_v, self.ky_v, self.kz_v, self.A_pm = gf_opt(
                kappa, alpha, p, m_max, m_max, m_max, M, M, M, self.h, self.h, self.h, L, L, L)

    def force(self, pos, Z=None):
        if Z is None:
            Z = np.ones(pos.shape[0])
        M, h, L = self.M, self.h, self.L
        return force_pot(self.kappa, self.G, pos.shape[0], self.rc, self.Lv, self.af, self.uf, pos, Z,
                         self.G_k, self.kx_v, self.ky_v, self.kz_v, M, M, M, h, h, h, L, L, L)


def wrap_positions(pos, Lv, Lmax_v=None, Lmin_v=None):
    """Single wrap per component, velocity_verlet_p3m_ord6.py:12-19."""
    Lv = np.asarray(Lv)
    Lmax_v = Lv if Lmax_v is None else np.asarray(Lmax_v)
    Lmin_v = np.zeros(3) if Lmin_v is None else np.asarray(Lmin_v)
    pos = pos.copy()
    pos = np.where(pos > Lmax_v, pos - Lv, pos)
    pos = np.where(pos < Lmin_v, pos + Lv, pos)
    return pos


def update(pos, vel, acc, dt, params: P3MParams, Z=None):
    """Kick-drift-kick step, velocity_verlet_p3m_ord6.py:5-26."""
    vel = vel + 0.5 * acc * dt
    pos = pos + vel * dt
    pos = wrap_positions(pos, params.Lv)
    U, acc, parts = params.force(pos, Z)
    vel = vel + 0.5 * acc * dt
    return pos, vel, acc, U


def vscale(pos, vel, acc, T_desired, it, dt, params: P3MParams, Z=None):
    """Velocity-rescale thermostat around one step, thermostat_p3m_ord6.py:5-25."""
    pos, vel, acc, U = update(pos, vel, acc, dt, params, Z)
    K = params.kf * np.sum(vel ** 2)
    T = K / params.kf / float(pos.shape[0])
    if it <= 1999:
        fact = np.sqrt(T_desired / T)
    else:
        fact = np.sqrt((20.0 * T_desired / T - 1.0) / 20.0)
    vel = vel * fact
    return pos, vel, acc, U


def density_modes(pos, q):
    """n(q) = sum_j exp(-i q x_j) per axis, yukawa_MD_p3m_ord6.py:214-219.

    Returns a complex array of shape (len(q), 3).
    """
    q = np.asarray(q, dtype=np.float64)
    out = np.empty((q.size, 3), dtype=np.complex128)
    for iq, q_p in enumerate(q):
        for a in range(3):
            out[iq, a] = np.sum(np.exp(-1j * q_p * pos[:, a]))
    return out


def reference_q_list(L, q_max=2.5):
    """q list of yukawa_MD_p3m_ord6.py:121-137: (i+1)dq, sqrt2 (i+1)dq, sqrt3 (i+1)dq."""
    dq = 2.0 * np.pi / L
    Nq = 3 * int(q_max / dq)
    qv = np.zeros(Nq)
    for iqv in range(0, Nq, 3):
        iq = iqv / 3.0
        qv[iqv] = (iq + 1.0) * dq
        qv[iqv + 1] = (iq + 1.0) * np.sqrt(2.0) * dq
        qv[iqv + 2] = (iq + 1.0) * np.sqrt(3.0) * dq
    return qv


def rms_force_error(F, F_ref):
    """sqrt(mean |F - F_ref|^2) / sqrt(mean |F_ref|^2)."""
    num = np.sqrt(np.mean(np.sum((F - F_ref) ** 2, axis=1)))
    den = np.sqrt(np.mean(np.sum(F_ref ** 2, axis=1)))
    return float(num / den)


def box_length(N):
    """L = (4 pi N / 3)^(1/3), yukawa_MD_p3m_ord6.py:42."""
    return (4.0 * np.pi * N / 3.0) ** (1.0 / 3.0)
