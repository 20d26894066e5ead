"""Pins of the CPU oracle (no GPU).

Every oracle function is checked against something other than a retyping of
its own formula: the reference's compiled C particle_particle (golden
fixtures), brute force, closed forms, limits and invariants.
"""
import glob
import math
import os

import numpy as np
import pytest
from scipy.special import erfc

from oracle import c_oracle
from oracle import p3m_oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = sorted(glob.glob(os.path.join(HERE, "golden", "pp_ref_*.npz")))


def rel_rms(a, b):
    return float(np.sqrt(np.mean(np.sum((a - b) ** 2, axis=-1)) / np.mean(np.sum(b ** 2, axis=-1))))


# ---------------------------------------------------------------- PP ------

@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_pp_matches_reference_c_golden(path):
    """Numpy oracle and C restatement vs the reference's own compiled C PP."""
    g = np.load(path)
    pos, Z, L = g["pos"], g["Z"], float(g["L"])
    kappa, alpha, rc = float(g["kappa"]), float(g["alpha"]), float(g["rc"])
    unit = bool(np.all(Z == 1.0))
    U, acc = O.particle_particle(kappa, alpha, rc, [L, L, L], pos, Z=None if unit else Z)
    assert rel_rms(acc, g["acc"]) < 1e-13
    assert np.max(np.abs(acc - g["acc"])) < 1e-11 * np.max(np.abs(g["acc"]))
    if unit:
        _, acc_c = c_oracle.pp(pos, [L, L, L], kappa, alpha, rc)
        assert rel_rms(acc_c, g["acc"]) < 1e-13


def test_golden_fixtures_present():
    assert len(GOLDEN) >= 4


def test_pp_cell_list_equals_brute_force():
    N = 600
    L = O.box_length(N)
    pos, _ = O.initial_split(5, 6, N, [L] * 3, 0.01)
    for kappa, alpha, rc in ((2.0, 0.85, L / 3.1), (0.0, 1.1, L / 3.5), (3.0, 0.5, L / 3.0)):
        U1, a1 = O.particle_particle(kappa, alpha, rc, [L] * 3, pos)
        U2, a2 = O.particle_particle_brute(kappa, alpha, rc, [L] * 3, pos)
        assert abs(U1 - U2) <= 1e-13 * abs(U2)
        assert rel_rms(a1, a2) < 1e-14


def test_pp_newton_third_law():
    N = 1500
    L = O.box_length(N)
    pos, _ = O.initial_split(7, 8, N, [L] * 3, 0.01)
    _, acc = O.particle_particle(2.0, 0.8, L / 3.2, [L] * 3, pos)
    assert np.max(np.abs(acc.sum(axis=0))) < 1e-10 * np.abs(acc).max()


def test_pp_kappa_zero_is_ewald_real_space():
    r = np.linspace(0.05, 5.0, 400)
    alpha = 0.9
    phi, f = O.pp_pair(r, 0.0, alpha)
    np.testing.assert_allclose(phi, erfc(alpha * r) / r, rtol=1e-14)
    fr = erfc(alpha * r) / r ** 2 + 2 * alpha / math.sqrt(math.pi) * np.exp(-(alpha * r) ** 2) / r
    np.testing.assert_allclose(f, fr, rtol=1e-13)


def test_pp_pair_force_is_minus_derivative():
    # f(r) = -dphi/dr, checked by a 4th-order central difference
    r = np.linspace(0.3, 6.0, 300)
    for kappa, alpha in ((2.0, 0.85), (0.5, 0.4), (3.0, 1.2)):
        h = 1e-4
        p = lambda x: O.pp_pair(x, kappa, alpha)[0]  # noqa: E731
        d = (-p(r + 2 * h) + 8 * p(r + h) - 8 * p(r - h) + p(r - 2 * h)) / (12 * h)
        _, f = O.pp_pair(r, kappa, alpha)
        np.testing.assert_allclose(f, -d, rtol=1e-8, atol=1e-12)


def test_pp_pair_large_alpha_limit_is_bare_yukawa():
    # alpha -> 0 sends the erfc factors to 2 and erfc-weighted terms to the
    # bare Yukawa pair e^{-kappa r}/r (the short-range part becomes the whole)
    r = np.linspace(0.5, 4.0, 50)
    kappa = 1.3
    phi, f = O.pp_pair(r, kappa, 1e-6)
    np.testing.assert_allclose(phi, np.exp(-kappa * r) / r, rtol=1e-5)
    np.testing.assert_allclose(f, np.exp(-kappa * r) * (1 + kappa * r) / r ** 2, rtol=1e-5)


def test_pp_two_particles_closed_form():
    L = 20.0
    pos = np.array([[1.0, 2.0, 3.0], [3.5, 2.0, 3.0]])
    kappa, alpha, rc = 1.0, 0.7, 6.0
    U, acc = O.particle_particle(kappa, alpha, rc, [L] * 3, pos)
    phi, f = O.pp_pair(np.array([2.5]), kappa, alpha)
    assert U == pytest.approx(phi[0], rel=1e-15)
    np.testing.assert_allclose(acc[0], [-f[0], 0, 0], rtol=1e-15)
    np.testing.assert_allclose(acc[1], [f[0], 0, 0], rtol=1e-15)
    # across the periodic boundary
    pos = np.array([[0.5, 2.0, 3.0], [19.0, 2.0, 3.0]])
    U, acc = O.particle_particle(kappa, alpha, rc, [L] * 3, pos)
    phi, f = O.pp_pair(np.array([1.5]), kappa, alpha)
    assert U == pytest.approx(phi[0], rel=1e-14)
    np.testing.assert_allclose(acc[0], [f[0], 0, 0], rtol=1e-14)


# --------------------------------------------------------- assignment ----

def _cox_de_boor(order, t):
    """Cardinal B-spline M_order(t) on [0, order] by the Cox-de Boor recursion."""
    t = np.asarray(t, dtype=np.float64)
    b = ((t >= np.arange(order)[:, None]) & (t < np.arange(1, order + 1)[:, None])).astype(float)
    for k in range(2, order + 1):
        nb = np.zeros((order - k + 1, t.size))
        for i in range(order - k + 1):
            nb[i] = ((t - i) * b[i] + (i + k - t) * b[i + 1]) / (k - 1)
        b = nb
    return b[0]


def test_weights_are_cardinal_bspline_order6():
    # weight g at offset x equals the cardinal B-spline M6 at (5 - g) + 1/2 + x
    x = np.linspace(-0.5, 0.5, 101, endpoint=False)
    w = O.bspline6_weights(x)
    for g in range(6):
        np.testing.assert_allclose(w[:, g], _cox_de_boor(6, (5 - g) + 0.5 + x), atol=3e-16)
    np.testing.assert_allclose(w.sum(axis=1), 1.0, atol=1e-15)


def test_charge_assignment_conserves_charge():
    N = 3000
    L = O.box_length(N)
    pos, _ = O.initial_split(3, 4, N, [L] * 3, 0.01)
    Z = np.random.RandomState(0).choice([1.0, 2.0], N)
    for M in (16, 32):
        rho = O.charge_assign(pos, Z, M, M, M, L / M, L / M, L / M)
        assert rho.sum() == pytest.approx(Z.sum(), rel=1e-14)
        rho_c = c_oracle.charge_assign(pos, Z, M, L / M)
        np.testing.assert_allclose(rho, rho_c, atol=1e-14)


def test_single_particle_assignment_is_tensor_product():
    M, L = 16, 10.0
    h = L / M
    pos = np.array([[3.3, 7.1, 0.2]])
    rho = O.charge_assign(pos, np.array([2.0]), M, M, M, h, h, h)
    ix = np.floor(pos[0] / h).astype(int)
    fr = (pos[0] - (ix + 0.5) * h) / h
    w = [O.bspline6_weights(f) for f in fr]
    exp = np.zeros((M, M, M))
    for a in range(6):
        for b in range(6):
            for c in range(6):
                exp[(ix[2] + a) % M, (ix[1] + b) % M, (ix[0] + c) % M] += 2.0 * w[2][a] * w[1][b] * w[0][c]
    np.testing.assert_allclose(rho, exp, atol=1e-16)


def test_gather_is_adjoint_of_assignment():
    N, M = 400, 16
    L = O.box_length(N)
    h = L / M
    rng = np.random.RandomState(9)
    pos = L * rng.random_sample((N, 3))
    Z = rng.random_sample(N) + 0.5
    E = [rng.standard_normal((M, M, M)) for _ in range(3)]
    acc, Ep = O.field_gather(*E, pos, Z, M, M, M, h, h, h)
    rho = O.charge_assign(pos, Z, M, M, M, h, h, h)
    for a in range(3):
        assert np.sum(Z * Ep[:, a]) == pytest.approx(np.sum(rho * E[a]), rel=1e-12)
    # constant field: partition of unity
    acc, Ep = O.field_gather(np.full((M, M, M), 2.5), np.zeros((M, M, M)), np.full((M, M, M), -1.0), pos, Z,
                             M, M, M, h, h, h)
    np.testing.assert_allclose(Ep[:, 0], 2.5, rtol=1e-14)
    np.testing.assert_allclose(Ep[:, 2], -1.0, rtol=1e-14)
    np.testing.assert_allclose(acc[:, 0], 2.5 * Z, rtol=1e-14)


# ----------------------------------------------------------- Green fn ----

def test_green_m0_reduces_to_R_over_U2():
    M, L, kappa, alpha, p = 8, 9.0, 1.5, 0.8, 6
    h = L / M
    G, kx, ky, kz, _ = O.gf_opt(kappa, alpha, p, 0, 0, 0, M, M, M, h, h, h, L, L, L)
    K2 = kx ** 2 + ky ** 2 + kz ** 2
    R = 4 * np.pi * np.exp(-(kappa ** 2 + K2) / (4 * alpha ** 2)) / (kappa ** 2 + K2)
    U = (np.sinc(kx * h / 2 / np.pi) * np.sinc(ky * h / 2 / np.pi) * np.sinc(kz * h / 2 / np.pi)) ** p
    exp = R / U ** 2
    exp[K2 == 0] = 0.0
    np.testing.assert_allclose(G, exp, rtol=1e-13, atol=0)


def test_green_symmetric_and_matches_c():
    M, L = 16, 12.0
    h = L / M
    G, *_ , A = O.gf_opt(2.0, 0.85, 6, 3, 3, 3, M, M, M, h, h, h, L, L, L)
    # k -> -k: shifted index n -> 2*mid - n (for n > 0)
    Gi = G[1:, 1:, 1:]
    np.testing.assert_allclose(Gi, Gi[::-1, ::-1, ::-1], rtol=1e-13)
    Gc, Ac = c_oracle.gf_opt(2.0, 0.85, 6, 3, M, L)
    np.testing.assert_allclose(G, Gc, rtol=1e-13, atol=1e-300)
    assert A == pytest.approx(Ac, rel=1e-12)


def test_green_fine_mesh_limit():
    # aliases vanish as h -> 0, leaving G = R / U(k)^2, and U -> 1 as kh -> 0
    L, kappa, alpha = 10.0, 1.0, 0.6
    k = 2 * np.pi / L
    R = 4 * np.pi * np.exp(-(kappa ** 2 + k ** 2) / (4 * alpha ** 2)) / (kappa ** 2 + k ** 2)
    dev = []
    for M in (32, 64):
        h = L / M
        G, *_ = O.gf_opt(kappa, alpha, 6, 1, 1, 1, M, M, M, h, h, h, L, L, L)
        mid = M // 2
        U = np.sinc(k * h / 2 / np.pi) ** 6
        assert G[mid, mid, mid + 1] == pytest.approx(R / U ** 2, rel=1e-9)
        dev.append(abs(G[mid, mid, mid + 1] / R - 1))
    assert dev[1] < dev[0] / 3.9  # O(h^2) approach to R


# ----------------------------------------------------------------- FFT / PM --

def test_c_fft_matches_numpy():
    rng = np.random.RandomState(1)
    for shape in ((8, 16, 32), (12, 8, 10)):
        x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        np.testing.assert_allclose(c_oracle.fft3d(x, -1), np.fft.fftn(x), rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(c_oracle.fft3d(x, +1), np.fft.ifftn(x) * x.size, rtol=1e-12, atol=1e-12)


def test_pm_plane_wave_known_answer():
    # rho = cos(2 pi x / L): E_x = (1/V) sum_k -i k G rho_k e^{ikx} = G(k1) (N/V)... analytic
    M, L = 16, 8.0
    h = L / M
    x = (np.arange(M) * h)
    rho = np.broadcast_to(np.cos(2 * np.pi * x / L)[None, None, :], (M, M, M)).copy()
    G = np.zeros((M, M, M))
    mid = M // 2
    G[mid, mid, mid + 1] = 3.0
    G[mid, mid, mid - 1] = 3.0
    kx = 2 * np.pi * (np.arange(M).reshape(1, M) - mid) / L
    ky = 2 * np.pi * (np.arange(M).reshape(M, 1) - mid) / L
    kz = 2 * np.pi * (np.arange(M).reshape(M, 1, 1) - mid) / L
    U, Ex, Ey, Ez = O.pm_solve(rho, G, kx, ky, kz, L, L, L)
    k1 = 2 * np.pi / L
    V = L ** 3
    # rho_k at +-k1 is M^3/2; E_x(x) = (1/V) * 2 * Re(-i k1 * 3 * M^3/2 * e^{i k1 x}) = 3 k1 M^3/V sin(k1 x)
    exp = 3.0 * k1 * M ** 3 / V * np.sin(k1 * x)
    np.testing.assert_allclose(Ex[0, 0, :], exp, atol=1e-12)
    np.testing.assert_allclose(Ey, 0.0, atol=1e-12)
    # Parseval: U = 1/2 sum |rho_k|^2 G / V = 1/2 * 2 * (M^3/2)^2 * 3 / V
    assert U == pytest.approx(0.5 * 2 * (M ** 3 / 2) ** 2 * 3.0 / V, rel=1e-12)


def test_pm_oracles_agree():
    N, M = 1500, 32
    L = O.box_length(N)
    h = L / M
    pos, _ = O.initial_split(1, 2, N, [L] * 3, 0.01)
    Z = np.ones(N)
    G, kx, ky, kz, _ = O.gf_opt(2.0, 0.85, 6, 3, 3, 3, M, M, M, h, h, h, L, L, L)
    U1, a1, rho1, E1 = O.particle_mesh_fft_r(2.0, 0.85, pos, Z, N, G, kx, ky, kz, M, M, M, h, h, h, L, L, L)
    U2, a2, rho2, E2 = c_oracle.pm(pos, Z, G, M, L)
    assert U1 == pytest.approx(U2, rel=1e-13)
    assert rel_rms(a1, a2) < 1e-13
    np.testing.assert_allclose(rho1, rho2, atol=1e-14)


def test_pm_converges_to_ewald_reciprocal_sum():
    # method check: P3M long-range force -> explicit Yukawa-Ewald k-sum
    N, L, kappa, alpha = 20, 8.0, 1.0, 0.9
    rng = np.random.RandomState(2)
    pos = L * rng.random_sample((N, 3))
    Z = np.ones(N)
    kmax = 9
    m = np.arange(-kmax, kmax + 1)
    KX, KY, KZ = np.meshgrid(m, m, m, indexing="ij")
    K = 2 * np.pi / L * np.stack([KX.ravel(), KY.ravel(), KZ.ravel()], 1)
    K = K[np.any(K != 0, axis=1)]
    K2 = np.sum(K * K, axis=1)
    R = 4 * np.pi * np.exp(-(kappa ** 2 + K2) / (4 * alpha ** 2)) / (kappa ** 2 + K2)
    S = np.exp(-1j * pos @ K.T).sum(axis=0)  # sum_j e^{-ik x_j}
    # E(x_i) = (1/V) sum_k -i k R S(k) e^{ik x_i}
    ph = np.exp(1j * pos @ K.T)
    E_ref = np.real((ph * (R * S)[None, :]) @ (-1j * K)) / L ** 3
    M = 32
    h = L / M
    G, kx, ky, kz, _ = O.gf_opt(kappa, alpha, 6, 3, 3, 3, M, M, M, h, h, h, L, L, L)
    _, acc, _, _ = O.particle_mesh_fft_r(kappa, alpha, pos, Z, N, G, kx, ky, kz, M, M, M, h, h, h, L, L, L)
    assert rel_rms(acc, E_ref) < 1e-5


# ------------------------------------------------------- init / modes / MD --

def test_initial_matches_reference_stream():
    N, L, T = 50, 5.0, 0.01
    pos, vel = O.initial(103, N, [L] * 3, T)
    rng = np.random.RandomState(103)
    u = [rng.random_sample(N) for _ in range(4)]
    s = np.sqrt(T / 3)
    vx = s * np.sqrt(-2 * np.log(u[0])) * np.cos(2 * np.pi * u[1])
    np.testing.assert_allclose(vel[:, 0], vx - vx.mean(), rtol=0, atol=1e-17)
    np.testing.assert_allclose(pos[:, 0], L * rng.random_sample(N), rtol=0, atol=0)
    assert np.abs(vel.mean(axis=0)).max() < 1e-17
    # NumPy's legacy stream is frozen (NEP 19): first draw of seed 1
    assert np.random.RandomState(1).random_sample() == 0.417022004702574


def test_density_modes_lattice_known_answer():
    L = 10.0
    n = 10
    x = np.arange(n) * L / n
    pos = np.stack([x, np.zeros(n), np.zeros(n)], 1)
    dq = 2 * np.pi / L
    out = O.density_modes(pos, np.array([dq, 10 * dq]))
    assert abs(out[0, 0]) < 1e-12          # sum of n-th roots of unity
    assert out[1, 0] == pytest.approx(n)   # q on the lattice's reciprocal vector
    assert out[0, 1] == pytest.approx(n)   # all y = 0


def test_reference_q_list():
    L = O.box_length(10000)
    q = O.reference_q_list(L)
    dq = 2 * np.pi / L
    assert q.size == 3 * int(2.5 / dq)
    assert q[0] == dq and q[1] == pytest.approx(np.sqrt(2) * dq) and q[3] == pytest.approx(2 * dq)


def test_c_md_steps_match_numpy_oracle():
    N = 400
    L = O.box_length(N)
    P = O.P3MParams(N, L, 2.0, 0.9, L / 3.1, 16)
    pos, vel = O.initial_split(1, 2, N, P.Lv, 0.01)
    Z = np.ones(N)
    _, acc, _ = P.force(pos, Z)
    p1, v1, a1 = pos.copy(), vel.copy(), acc.copy()
    for _ in range(3):
        p1, v1, a1, U1 = O.update(p1, v1, a1, 0.1, P, Z)
    p2, v2, a2 = pos.copy(), vel.copy(), acc.copy()
    en, _ = c_oracle.md_steps(p2, v2, a2, Z, 3, 0.1, L, 2.0, 0.9, L / 3.1, P.G_k, 16)
    np.testing.assert_allclose(p2, p1, rtol=0, atol=1e-12)
    assert rel_rms(a2, a1) < 1e-12
    assert en[-1, 2] == pytest.approx(U1, rel=1e-12)


def test_vscale_branches():
    N = 200
    L = O.box_length(N)
    P = O.P3MParams(N, L, 2.0, 0.9, L / 3.1, 16)
    pos, vel = O.initial_split(1, 2, N, P.Lv, 0.01)
    _, acc, _ = P.force(pos)
    _, v, _, _ = O.vscale(pos, vel, acc, 0.02, 5, 0.1, P)
    assert np.sum(v ** 2) / N == pytest.approx(0.02, rel=1e-12)
    _, v2, _, _ = O.vscale(pos, vel, acc, 0.02, 2500, 0.1, P)
    _, vv, _, _ = O.update(pos, vel, acc, 0.1, P)
    T = np.sum(vv ** 2) / N
    np.testing.assert_allclose(v2, vv * np.sqrt((20 * 0.02 / T - 1) / 20), rtol=1e-14)


class _StubForce:
    """A stand-in for P3MParams whose force is a fixed field (no P3M)."""

    def __init__(self, L, a):
        self.Lv = np.array([L, L, L])
        self.a = a

    def force(self, pos, Z=None):
        return 0.0, np.broadcast_to(self.a, pos.shape).copy(), None


def test_update_free_flight_zero_force():
    """Pins update (velocity_verlet_p3m_ord6.py:7-24) on a zero force field:
    the step is then free flight, x1 = wrap(x0 + v0 dt) and v1 = v0 exactly,
    with the single strict wrap of :12-19 (x > L -> x - L, x < 0 -> x + L)."""
    rng = np.random.RandomState(7)
    N, L, dt = 300, 10.0, 0.1
    pos = rng.random_sample((N, 3)) * L
    vel = (rng.random_sample((N, 3)) - 0.5) * 30.0  # up to 1.5 L per step: some leave the box
    P = _StubForce(L, np.zeros(3))
    p1, v1, a1, _ = O.update(pos, vel, np.zeros((N, 3)), dt, P)
    assert np.array_equal(v1, vel)
    free = pos + vel * dt
    want = free.copy()
    want[free > L] -= L
    want[want < 0] += L
    assert np.array_equal(p1, want)
    assert ((free > L) | (free < 0)).any()  # the wrap branch was exercised
    assert not np.any(a1)


def test_update_constant_force_closed_form():
    """Velocity Verlet integrates a constant acceleration a exactly:
    x(t) = x0 + v0 t + a t^2 / 2, v(t) = v0 + a t, to rounding.  A drift before
    the first half kick, or a full kick in place of a half kick, gives
    x0 + v0 t + a t^2 or x0 + v0 t, off by a t^2 / 2 (here ~1e-2 of L)."""
    rng = np.random.RandomState(11)
    N, L, dt, nstep = 100, 1.0e6, 0.1, 20  # box large enough that nothing wraps
    pos = rng.random_sample((N, 3)) * 10.0 + L / 2
    vel = rng.standard_normal((N, 3))
    a = np.array([0.3, -0.7, 1.1])
    P = _StubForce(L, a)
    p, v, acc = pos, vel, np.broadcast_to(a, pos.shape).copy()
    for _ in range(nstep):
        p, v, acc, _ = O.update(p, v, acc, dt, P)
    t = nstep * dt
    np.testing.assert_allclose(v, vel + a * t, rtol=0, atol=1e-13)
    np.testing.assert_allclose(p - L / 2, pos - L / 2 + vel * t + 0.5 * a * t * t, rtol=0, atol=1e-9)


def test_update_time_reversal_with_p3m_forces():
    """Velocity Verlet is time reversible: 20 steps forward, negate the
    velocities, 20 steps back return every ion to its start (N = 400 with the
    oracle's own P3M forces).  A mis-ordered kick and drift breaks this."""
    N = 400
    L = O.box_length(N)
    P = O.P3MParams(N, L, 2.0, 0.9, L / 3.1, 16)
    # a warm start (T = 1) so that ions move a good fraction of a spacing
    pos, vel = O.initial_split(5, 6, N, P.Lv, 1.0)
    _, acc, _ = P.force(pos)
    p, v, a = pos.copy(), vel.copy(), acc.copy()
    for _ in range(20):
        p, v, a, _ = O.update(p, v, a, 0.05, P)
    moved = np.abs(p - pos)
    moved = np.minimum(moved, L - moved)
    assert moved.max() > 0.3  # the trajectory went somewhere
    v = -v
    for _ in range(20):
        p, v, a, _ = O.update(p, v, a, 0.05, P)
    d = np.abs(p - pos)
    d = np.minimum(d, L - d)
    assert d.max() < 1e-10 * L
    np.testing.assert_allclose(-v, vel, rtol=0, atol=1e-10)


def test_update_energy_error_is_second_order():
    """The energy error of velocity Verlet over a fixed time falls as dt^2:
    halving dt quarters it.  An update with the force sign flipped, or a
    first-order splitting, fails this (no conservation, or a factor 2)."""
    N = 200
    L = O.box_length(N)
    P = O.P3MParams(N, L, 2.0, 0.9, L / 3.1, 16)
    pos0, vel0 = O.initial_split(8, 9, N, P.Lv, 0.5)
    U0, acc0, _ = P.force(pos0)
    E0 = U0 + P.kf * np.sum(vel0 ** 2)
    errs = []
    for dt, nstep in ((0.04, 10), (0.02, 20), (0.01, 40)):
        p, v, a = pos0.copy(), vel0.copy(), acc0.copy()
        for _ in range(nstep):
            p, v, a, U = O.update(p, v, a, dt, P)
        errs.append(abs(U + P.kf * np.sum(v ** 2) - E0))
    assert errs[0] > 1e-9 * abs(E0)  # well above rounding
    for e1, e2 in zip(errs, errs[1:]):
        assert 3.0 < e1 / e2 < 5.5


def test_dropin_initial_matches_oracle_and_global_stream():
    """The drop-in initial (initialize_pos_vel.py:4-35 signature: fills pos
    and vel in place) draws the same numbers as the oracle, and both equal
    the reference's own sequence of global-state calls: np.random.seed, four
    np.random.random(N) for Box-Muller, the per-column mean removal of
    :21-28, then three position columns (:31-33)."""
    from mdforce.initialize_pos_vel import initial as dropin_initial

    N, T = 1000, 0.01
    Lv = np.array([30.0, 31.0, 32.5])
    pos = np.empty((N, 3))
    vel = np.empty((N, 3))
    p_ret, v_ret = dropin_initial(17, N, Lv, pos, vel, T)
    assert p_ret is pos and v_ret is vel
    po, vo = O.initial(17, N, Lv, T)
    assert np.array_equal(pos, po) and np.array_equal(vel, vo)
    state = np.random.get_state()
    try:
        np.random.seed(seed=17)
        sig = np.sqrt(T / 3)
        u1, u2, u3, u4 = (np.random.random(N) for _ in range(4))
        want_v = np.empty((N, 3))
        want_v[:, 0] = sig * np.sqrt(-2 * np.log(u1)) * np.cos(2 * np.pi * u2)
        want_v[:, 1] = sig * np.sqrt(-2 * np.log(u1)) * np.sin(2 * np.pi * u2)
        want_v[:, 2] = sig * np.sqrt(-2 * np.log(u3)) * np.cos(2 * np.pi * u4)
        means = [np.mean(want_v[:, a]) for a in range(3)]
        for a in range(3):
            want_v[:, a] = want_v[:, a] - means[a]
        want_p = np.stack([Lv[a] * np.random.random(N) for a in range(3)], 1)
    finally:
        np.random.set_state(state)
    assert np.array_equal(vel, want_v)
    assert np.array_equal(pos, want_p)


def test_openmp_restatement_matches_serial():
    """The all-core CPU baseline (the C restatement built with OpenMP) computes
    what the serial checker does: 2 NVE steps of a cfg1-sized system agree to
    1e-12 in positions, forces and energies."""
    N = 3000
    L = O.box_length(N)
    P = O.P3MParams(N, L, 2.0, 0.85, L / 3.1, 32)
    pos, vel = O.initial_split(3, 4, N, P.Lv, 0.01)
    _, acc, _ = P.force(pos)
    runs = []
    for threads in (None, 4):
        p, v, a = pos.copy(), vel.copy(), acc.copy()
        en, nqt = c_oracle.md_steps(p, v, a, np.ones(N), 2, 0.1, L, 2.0, 0.85, L / 3.1, P.G_k, 32,
                                    q=np.array([0.3, 0.6]), threads=threads)
        runs.append((p, v, a, en, nqt))
    (p1, v1, a1, e1, q1), (p2, v2, a2, e2, q2) = runs
    assert c_oracle.lib_omp(0)[1] >= 1
    np.testing.assert_allclose(p2, p1, rtol=0, atol=1e-12)
    assert rel_rms(a2, a1) < 1e-12
    np.testing.assert_allclose(e2, e1, rtol=1e-12)
    np.testing.assert_allclose(q2, q1, rtol=1e-12, atol=1e-9)
