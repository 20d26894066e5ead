"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Bar (plain FP64): relative RMS force error <= 1e-12 and |dU|/|U| <= 1e-12
against the oracle on identical seeded positions.
"""
import glob
import math
import os

import numpy as np
import pytest

from oracle import c_oracle
from oracle import p3m_oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TOL = 1e-12
HERE = os.path.dirname(os.path.abspath(__file__))


def rel_rms(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.sqrt(np.mean(np.sum((a - b) ** 2, axis=-1)) / np.mean(np.sum(b ** 2, axis=-1))))


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


@pytest.fixture(scope="module")
def P3M():
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    from mdforce.engine import P3M as cls
    return cls


@pytest.fixture(scope="module")
def cfg1():
    """cfg1: N = 1e4, kappa = 2, alpha = 0.85, rc = 6.2, 64^3, seeds 1/2."""
    N = 10000
    L = O.box_length(N)
    P = O.P3MParams(N, L, 2.0, 0.85, 6.2, 64)
    pos, vel = O.initial_split(1, 2, N, P.Lv, 1.0 / 100)
    Z = np.ones(N)
    U, acc, parts = P.force(pos, Z)
    return dict(P=P, pos=pos, vel=vel, Z=Z, U=U, acc=acc, parts=parts)


def test_force_cfg1_parity(P3M, cfg1):
    P = cfg1["P"]
    N = P.N
    e = P3M(N, 64, P.L, P.kappa, P.G, P.rc)
    e.set_green(dev(P.G_k))
    Z = dev(cfg1["Z"])
    acc_pp = torch.empty((N, 3), dtype=torch.float64, device="cuda")
    acc_pm = torch.empty_like(acc_pp)
    ex, ey, ez = (torch.empty(N, dtype=torch.float64, device="cuda") for _ in range(3))
    rho = torch.empty((64, 64, 64), dtype=torch.float64, device="cuda")
    acc, en = e.force(dev(cfg1["pos"]), Z, af=P.af, uf=P.uf, acc_pp=acc_pp, acc_pm=acc_pm, e_x=ex, e_y=ey, e_z=ez,
                      rho=rho)
    torch.cuda.synchronize()
    parts = cfg1["parts"]
    en = en.cpu().numpy()
    print("rel rms acc", rel_rms(acc.cpu().numpy(), cfg1["acc"]), "U", en[0], cfg1["U"])
    assert rel_rms(acc.cpu().numpy(), cfg1["acc"]) <= TOL
    assert abs(en[0] - cfg1["U"]) <= TOL * abs(cfg1["U"])
    assert abs(en[1] - parts["U_short"]) <= TOL * abs(parts["U_short"])
    assert abs(en[2] - parts["U_fft"]) <= TOL * abs(parts["U_fft"])
    assert en[3] == pytest.approx(parts["U_self"], rel=1e-15)
    assert rel_rms(acc_pp.cpu().numpy(), parts["acc_s_r"]) <= TOL
    assert rel_rms(acc_pm.cpu().numpy(), parts["acc_fft"]) <= TOL
    Ep = np.stack([ex.cpu().numpy(), ey.cpu().numpy(), ez.cpu().numpy()], 1)
    assert rel_rms(Ep, parts["E_p"]) <= TOL
    np.testing.assert_allclose(rho.cpu().numpy(), parts["rho_r"], atol=1e-13)
    assert float(rho.sum()) == pytest.approx(N, rel=1e-13)
    assert e.stats()["pp_multipass_tiles"] == 0
    e.close()


def test_force_is_repeatable_and_reuses_context(P3M, cfg1):
    P = cfg1["P"]
    e = P3M(P.N, 64, P.L, P.kappa, P.G, P.rc)
    e.set_green(dev(P.G_k))
    pos = dev(cfg1["pos"])
    a1, en1 = e.force(pos)
    a1 = a1.clone()
    u1 = float(en1[0])
    a2, en2 = e.force(pos)
    torch.cuda.synchronize()
    assert rel_rms(a2.cpu().numpy(), a1.cpu().numpy()) < 1e-14
    assert float(en2[0]) == pytest.approx(u1, rel=1e-14)
    e.close()


GOLDEN = [p for p in sorted(glob.glob(os.path.join(HERE, "golden", "pp_ref_*.npz"))) if "mixedZ" not in p]


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_pp_matches_reference_golden(P3M, path):
    g = np.load(path)
    pos, L = g["pos"], float(g["L"])
    N = pos.shape[0]
    e = P3M(N, 8, L, float(g["kappa"]), float(g["alpha"]), float(g["rc"]))
    acc, u = e.pp(dev(pos))
    torch.cuda.synchronize()
    assert rel_rms(acc.cpu().numpy(), g["acc"]) <= TOL
    U_ref, _ = O.particle_particle(float(g["kappa"]), float(g["alpha"]), float(g["rc"]), [L] * 3, pos)
    assert float(u) == pytest.approx(U_ref, rel=TOL)
    e.close()


@pytest.mark.parametrize("M", [16, 32, 64])
def test_green_build_matches_oracle(P3M, M):
    N = 10000
    L = O.box_length(N)
    h = L / M
    G, *_, A = O.gf_opt(2.0, 0.85, 6, 3, 3, 3, M, M, M, h, h, h, L, L, L)
    e = P3M(N, M, L, 2.0, 0.85, 6.2)
    full, apm = e.build_green(6, 3, want_full=True)
    torch.cuda.synchronize()
    Gd = full.cpu().numpy()
    assert np.max(np.abs(Gd - G)) <= 1e-13 * np.max(np.abs(G))
    assert float(apm) == pytest.approx(A, rel=1e-12)
    e.close()


def test_built_green_gives_same_force_as_uploaded(P3M, cfg1):
    P = cfg1["P"]
    e = P3M(P.N, 64, P.L, P.kappa, P.G, P.rc)
    e.build_green(6, 3)
    acc, en = e.force(dev(cfg1["pos"]))
    torch.cuda.synchronize()
    assert rel_rms(acc.cpu().numpy(), cfg1["acc"]) <= TOL
    assert float(en[0]) == pytest.approx(cfg1["U"], rel=TOL)
    e.close()


def test_pm_and_pp_parts_alone(P3M, cfg1):
    P = cfg1["P"]
    e = P3M(P.N, 64, P.L, P.kappa, P.G, P.rc)
    e.set_green(dev(P.G_k))
    pos = dev(cfg1["pos"])
    app, upp = e.pp(pos)
    app = app.cpu().numpy()
    upp = float(upp)
    apm, upm = e.pm(pos, dev(cfg1["Z"]))
    torch.cuda.synchronize()
    parts = cfg1["parts"]
    assert rel_rms(app, parts["acc_s_r"]) <= TOL
    assert upp == pytest.approx(parts["U_short"], rel=TOL)
    assert rel_rms(apm.cpu().numpy(), parts["acc_fft"]) <= TOL
    assert float(upm) == pytest.approx(parts["U_fft"], rel=TOL)
    e.close()


def test_dropin_force_pot_numpy_semantics(cfg1):
    """force_pot with the reference's 33 positional NumPy arguments."""
    from mdforce import yukawa_p3m_ord6 as yp
    P = cfg1["P"]
    N, M, L, h = P.N, P.M, P.L, P.h
    pos = cfg1["pos"].copy()
    acc_in = np.zeros((N, 3))
    acc_s_r = np.zeros((N, 3))
    acc_fft = np.zeros((N, 3))
    rho_r = np.zeros((M, M, M))
    E_x_p, E_y_p, E_z_p = np.zeros(N), np.zeros(N), np.zeros(N)
    U, acc = yp.force_pot(P.kappa, P.G, N, P.rc, P.Lv, P.af, P.uf, pos, acc_in, cfg1["Z"], P.G_k, P.kx_v, P.ky_v,
                          P.kz_v, M, M, M, h, h, h, L, L, L, 6, 3, 3, 3, acc_s_r, acc_fft, rho_r, E_x_p, E_y_p, E_z_p)
    assert isinstance(acc, np.ndarray) and acc.shape == (N, 3)
    assert not np.shares_memory(acc, acc_in) and np.all(acc_in == 0)
    assert rel_rms(acc, cfg1["acc"]) <= TOL
    assert U == pytest.approx(cfg1["U"], rel=TOL)
    parts = cfg1["parts"]
    assert rel_rms(acc_s_r, parts["acc_s_r"]) <= TOL
    assert rel_rms(acc_fft, parts["acc_fft"]) <= TOL
    np.testing.assert_allclose(rho_r, parts["rho_r"], atol=1e-13)
    assert rel_rms(np.stack([E_x_p, E_y_p, E_z_p], 1), parts["E_p"]) <= TOL


def test_dropin_pp_pm_gf(cfg1):
    from mdforce import yukawa_gf_opt, yukawa_pm_ord6, yukawa_pp
    P = cfg1["P"]
    N, M, L, h = P.N, P.M, P.L, P.h
    acc_s_r = np.zeros((N, 3))
    U, out = yukawa_pp.particle_particle(P.kappa, P.G, P.rc, P.Lv, cfg1["pos"], acc_s_r)
    assert out is acc_s_r
    assert rel_rms(acc_s_r, cfg1["parts"]["acc_s_r"]) <= TOL
    assert U == pytest.approx(cfg1["parts"]["U_short"], rel=TOL)
    acc_f = np.zeros((N, 3))
    rho = np.zeros((M, M, M))
    U_f, out = yukawa_pm_ord6.particle_mesh_fft_r(P.kappa, P.G, cfg1["pos"], cfg1["Z"], N, P.G_k, P.kx_v, P.ky_v,
                                                  P.kz_v, M, M, M, h, h, h, L, L, L, 6, 3, 3, 3, rho, acc_f,
                                                  np.zeros(N), np.zeros(N), np.zeros(N))
    assert out is acc_f
    assert rel_rms(acc_f, cfg1["parts"]["acc_fft"]) <= TOL
    assert U_f == pytest.approx(cfg1["parts"]["U_fft"], rel=TOL)
    G_k, kx_v, ky_v, kz_v, A = yukawa_gf_opt.gf_opt(P.kappa, P.G, 6, 3, 3, 3, M, M, M, h, h, h, L, L, L)
    assert np.max(np.abs(G_k - P.G_k)) <= 1e-13 * np.max(np.abs(P.G_k))
    np.testing.assert_array_equal(kx_v, P.kx_v)
    np.testing.assert_array_equal(kz_v, P.kz_v)
    assert A == pytest.approx(P.A_pm, rel=1e-12)


def test_dropin_update_and_vscale(cfg1):
    from mdforce import thermostat_p3m_ord6, velocity_verlet_p3m_ord6
    P = cfg1["P"]
    N, M, L, h = P.N, P.M, P.L, P.h
    pos, vel, acc = cfg1["pos"].copy(), cfg1["vel"].copy(), cfg1["acc"].copy()
    scratch = [np.zeros((N, 3)), np.zeros((N, 3)), np.zeros((M, M, M)), np.zeros(N), np.zeros(N), np.zeros(N)]
    dt = 0.1
    p1, v1, a1, U1 = velocity_verlet_p3m_ord6.update(pos, vel, acc, dt, 1, P.kappa, N, 3, P.Lv, P.Lv, np.zeros(3),
                                                     P.G, P.rc, P.af, P.uf, cfg1["Z"], P.G_k, P.kx_v, P.ky_v, P.kz_v,
                                                     M, M, M, h, h, h, L, L, L, 6, 3, 3, 3, *scratch)
    p0, v0, a0, U0 = O.update(pos, vel, acc, dt, P, cfg1["Z"])
    np.testing.assert_array_equal(pos, cfg1["pos"])  # inputs untouched
    np.testing.assert_allclose(p1, p0, rtol=0, atol=0)  # same accelerations in -> same bits out
    assert rel_rms(a1, a0) <= TOL
    assert rel_rms(v1, v0) <= TOL
    assert U1 == pytest.approx(U0, rel=TOL)
    T_d = 0.01
    p2, v2, a2, U2 = thermostat_p3m_ord6.vscale(pos, vel, acc, T_d, 1, 3, dt, N, 3, P.Lv, P.Lv, np.zeros(3),
                                                P.kappa, P.G, P.rc, P.af, P.uf, 1.5, cfg1["Z"], P.G_k, P.kx_v, P.ky_v,
                                                P.kz_v, M, M, M, h, h, h, L, L, L, 6, 3, 3, 3, *scratch)
    _, v3, _, _ = O.vscale(pos, vel, acc, T_d, 3, dt, P, cfg1["Z"])
    assert rel_rms(v2, v3) <= TOL
    assert np.sum(v2 ** 2) / N == pytest.approx(T_d, rel=1e-12)


def test_md_run_short_horizon_vs_c_oracle(P3M):
    N, M = 3000, 32
    L = O.box_length(N)
    kappa, alpha, rc = 2.0, 0.85, L / 3.05
    P = O.P3MParams(N, L, kappa, alpha, rc, M)
    pos, vel = O.initial_split(1, 2, N, P.Lv, 0.01)
    Z = np.ones(N)
    U0, acc, _ = P.force(pos, Z)
    q = 2 * np.pi / L * np.arange(1, 21)
    steps = 5
    pc, vc, ac = pos.copy(), vel.copy(), acc.copy()
    en_c, nqt_c = c_oracle.md_steps(pc, vc, ac, Z, steps, 0.1, L, kappa, alpha, rc, P.G_k, M, q=q)
    e = P3M(N, M, L, kappa, alpha, rc)
    e.set_green(dev(P.G_k))
    pd, vd, ad = dev(pos), dev(vel), dev(acc)
    en, nqt = e.md_run(pd, vd, ad, steps, 0.1, Z=dev(Z), q=dev(q))
    torch.cuda.synchronize()
    # the same run with n(q, t) on the q grid by recurrence
    pg, vg, ag = dev(pos), dev(vel), dev(acc)
    en_g, nqt_g = e.md_run(pg, vg, ag, steps, 0.1, Z=dev(Z), q_step=2 * np.pi / L, nq=20)
    torch.cuda.synchronize()
    np.testing.assert_allclose(nqt_g.cpu().numpy(), nqt_c, rtol=1e-9, atol=1e-9)
    en = en.cpu().numpy()
    assert np.max(np.abs(pd.cpu().numpy() - pc)) < 1e-11
    assert rel_rms(ad.cpu().numpy(), ac) < 1e-11
    np.testing.assert_allclose(en[:, 2], en_c[:, 2], rtol=1e-11)  # U
    np.testing.assert_allclose(en[:, 1], en_c[:, 1], rtol=1e-11)  # K
    np.testing.assert_allclose(nqt.cpu().numpy(), nqt_c, rtol=1e-9, atol=1e-9)


def test_density_modes_vs_oracle(P3M, cfg1):
    P = cfg1["P"]
    e = P3M(P.N, 64, P.L, P.kappa, P.G, P.rc)
    q = O.reference_q_list(P.L)
    out = e.density_modes(dev(cfg1["pos"]), torch.from_numpy(q))
    torch.cuda.synchronize()
    ref = O.density_modes(cfg1["pos"], q)
    np.testing.assert_allclose(out.cpu().numpy(), ref, rtol=1e-11, atol=1e-9)
    e.close()


def test_edge_positions_wrap_like_periodic_images(P3M):
    """x = L, x = 0 and slightly negative coordinates: same force as the
    wrapped positions (the reference leaves x = L out of range)."""
    N = 2000
    L = O.box_length(N)
    P = O.P3MParams(N, L, 1.0, 0.8, L / 3.3, 32)
    pos, _ = O.initial_split(21, 22, N, P.Lv, 0.01)
    pos[0] = [L, 0.0, L]
    pos[1] = [0.0, L, 0.5 * L]
    pos[2] = [-1e-13, 0.3 * L, L]
    wrapped = np.where(pos >= L, pos - L, np.where(pos < 0, pos + L, pos))
    U_ref, acc_ref, _ = P.force(wrapped)
    e = P3M(N, 32, L, 1.0, 0.8, L / 3.3)
    e.set_green(dev(P.G_k))
    acc, en = e.force(dev(pos))
    torch.cuda.synchronize()
    assert rel_rms(acc.cpu().numpy(), acc_ref) <= 1e-11
    assert float(en[0]) == pytest.approx(U_ref, rel=1e-11)
    e.close()


@pytest.fixture(scope="module")
def faces_case():
    """Coordinates exactly on reference-cell faces (which are mesh nodes here:
    4 cells of 16 nodes per axis), one ulp either side of them, and on other
    mesh nodes, with the oracle's result."""
    N, M = 10000, 64
    L = O.box_length(N)
    rc = L / 4.2  # floor(L / rc) = 4 reference cells of L/4 = 16 h per axis
    P = O.P3MParams(N, L, 1.5, 0.7, rc, M)
    rng = np.random.default_rng(31)
    pos = rng.uniform(0.0, L, (N, 3))
    face = (L / 4) * rng.integers(0, 4, N)
    node = (L / M) * rng.integers(0, M, N)
    ax = rng.integers(0, 3, N)
    kind = rng.integers(0, 5, N)
    rows = np.arange(N)
    pos[rows[kind == 0], ax[kind == 0]] = face[kind == 0]
    pos[rows[kind == 1], ax[kind == 1]] = np.nextafter(face[kind == 1], -np.inf)
    pos[rows[kind == 2], ax[kind == 2]] = np.nextafter(face[kind == 2], np.inf)
    pos[rows[kind == 3], ax[kind == 3]] = node[kind == 3]
    pos = np.where(pos < 0, pos + L, pos)
    pos = np.where(pos >= L, pos - L, pos)
    U_ref, acc_ref, parts = P.force(pos)
    return dict(P=P, pos=pos, U=U_ref, acc=acc_ref, rho=parts["rho_r"])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_positions_on_cell_faces_and_mesh_nodes(P3M, faces_case, precision):
    """x/h and the cell index round at their edges on these coordinates, and
    a charge stencil can start one node outside its block's brick (it then
    takes the global-memory path).  Charge mesh, forces and energy match the
    oracle to the plain or mixed bar."""
    P, pos = faces_case["P"], faces_case["pos"]
    e = P3M(P.N, P.M, P.L, P.kappa, P.G, P.rc, precision=precision)
    e.set_green(dev(P.G_k))
    rho = torch.empty((P.M, P.M, P.M), dtype=torch.float64, device="cuda")
    acc, en = e.force(dev(pos), None, rho=rho)
    torch.cuda.synchronize()
    err = rel_rms(acc.cpu().numpy(), faces_case["acc"])
    print(f"[faces] {precision}: rel rms {err:.3g}, U {float(en[0])} vs {faces_case['U']}, {e.stats()}")
    tol = TOL if precision == "fp64" else 2e-6
    assert err <= tol
    assert float(en[0]) == pytest.approx(faces_case["U"], rel=tol)
    # mixed: FP32 stencil weights (~1e-7 relative each), FP64 accumulation
    rho_ref = faces_case["rho"]
    atol = 1e-13 if precision == "fp64" else 1e-6 * float(np.abs(rho_ref).max())
    np.testing.assert_allclose(rho.cpu().numpy(), rho_ref, atol=atol)
    e.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_anisotropic_box_and_mesh(P3M, precision):
    """Box 30 x 34 x 40 with a 32 x 32 x 64 mesh: the reference takes Lx, Ly, Lz
    and Mx, My, Mz separately (yukawa_p3m_ord6.py:6).  3 x 4 x 5 reference
    cells at rc 8.  Forces, U and the charge mesh against the oracle, and the
    device Green function against gf_opt."""
    N, M = 6000, (32, 32, 64)
    Lv = np.array([30.0, 34.0, 40.0])
    kappa, alpha, rc = 1.0, 0.6, 8.0
    h = Lv / np.array(M)
    G_k, kx_v, ky_v, kz_v, _ = O.gf_opt(kappa, alpha, 6, 3, 3, 3, *M, *h, *Lv)
    pos = np.random.default_rng(41).uniform(0.0, 1.0, (N, 3)) * Lv
    U_ref, acc_ref, parts = O.force_pot(kappa, alpha, N, rc, Lv, 1.0 / 3.0, 1.0, pos, np.ones(N), G_k, kx_v, ky_v,
                                        kz_v, *M, *h, *Lv)
    e = P3M(N, M, tuple(Lv), kappa, alpha, rc, precision=precision)
    full, _ = e.build_green(6, 3, want_full=True)
    torch.cuda.synchronize()
    assert np.max(np.abs(full.cpu().numpy() - G_k)) <= 1e-13 * np.max(np.abs(G_k))
    rho = torch.empty((M[2], M[1], M[0]), dtype=torch.float64, device="cuda")
    acc, en = e.force(dev(pos), None, af=1.0 / 3.0, uf=1.0, rho=rho)
    torch.cuda.synchronize()
    err = rel_rms(acc.cpu().numpy(), acc_ref)
    print(f"[aniso] {precision}: rel rms {err:.3g}, U {float(en[0])} vs {U_ref}")
    tol = TOL if precision == "fp64" else 2e-6
    assert err <= tol
    assert float(en[0]) == pytest.approx(U_ref, rel=tol)
    atol = 1e-13 if precision == "fp64" else 1e-6 * float(np.abs(parts["rho_r"]).max())
    np.testing.assert_allclose(rho.cpu().numpy(), parts["rho_r"], atol=atol)
    e.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_nonunit_charges(P3M, precision):
    """Charges Z uniform in [0.5, 1.5): the reference's PM weights the deposit
    and the field by Z (yukawa_pm_ord6.py:52-93, :308-310) while its PP pair
    terms are for unit charges (yukawa_pp.py:115-124).  Forces, U and the
    charge mesh against the oracle."""
    N, M = 4000, 32
    L = O.box_length(N)
    P = O.P3MParams(N, L, 2.0, 0.7, L / 3.5, M)
    rng = np.random.default_rng(43)
    pos = rng.uniform(0.0, L, (N, 3))
    Z = rng.uniform(0.5, 1.5, N)
    U_ref, acc_ref, parts = P.force(pos, Z)
    e = P3M(N, M, L, P.kappa, P.G, P.rc, precision=precision)
    e.set_green(dev(P.G_k))
    rho = torch.empty((M, M, M), dtype=torch.float64, device="cuda")
    acc, en = e.force(dev(pos), dev(Z), rho=rho)
    torch.cuda.synchronize()
    err = rel_rms(acc.cpu().numpy(), acc_ref)
    print(f"[charges] {precision}: rel rms {err:.3g}, U {float(en[0])} vs {U_ref}")
    tol = TOL if precision == "fp64" else 2e-6
    assert err <= tol
    assert float(en[0]) == pytest.approx(U_ref, rel=tol)
    atol = 1e-13 if precision == "fp64" else 1e-6 * float(np.abs(parts["rho_r"]).max())
    np.testing.assert_allclose(rho.cpu().numpy(), parts["rho_r"], atol=atol)
    assert float(rho.sum()) == pytest.approx(float(Z.sum()), rel=1e-12 if precision == "fp64" else 1e-6)
    e.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_coulomb_limit_kappa_zero(P3M, precision):
    """kappa = 0 (the Coulomb limit, b = kappa / 2 alpha = 0 in the pair terms
    and R(k) = 4 pi exp(-k^2 / 4 alpha^2) / k^2): full force, U and the device
    Green function against the oracle."""
    N, M = 3000, 32
    L = O.box_length(N)
    with np.errstate(divide="ignore", invalid="ignore"):  # gf_opt's k = 0 alias term, zeroed after (as the reference)
        P = O.P3MParams(N, L, 0.0, 0.8, L / 3.2, M)
    pos = np.random.default_rng(47).uniform(0.0, L, (N, 3))
    U_ref, acc_ref, _ = P.force(pos)
    e = P3M(N, M, L, 0.0, P.G, P.rc, precision=precision)
    full, _ = e.build_green(6, 3, want_full=True)
    torch.cuda.synchronize()
    assert np.max(np.abs(full.cpu().numpy() - P.G_k)) <= 1e-13 * np.max(np.abs(P.G_k))
    acc, en = e.force(dev(pos))
    torch.cuda.synchronize()
    err = rel_rms(acc.cpu().numpy(), acc_ref)
    print(f"[coulomb] {precision}: rel rms {err:.3g}, U {float(en[0])} vs {U_ref}")
    tol = TOL if precision == "fp64" else 2e-6
    assert err <= tol
    assert float(en[0]) == pytest.approx(U_ref, rel=tol)
    e.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_positions_several_periods_out(P3M, precision):
    """Ions displaced by whole periods (up to +-5 L, one by 1000 L), as a
    drift faster than L/dt leaves them: the force is that of the in-box
    positions (the sorted copy reduces them with floor(x/L))."""
    N = 2000
    L = O.box_length(N)
    P = O.P3MParams(N, L, 1.0, 0.8, L / 3.3, 32)
    pos, _ = O.initial_split(23, 24, N, P.Lv, 0.01)
    U_ref, acc_ref, _ = P.force(pos)
    k = np.random.RandomState(5).randint(-5, 6, size=pos.shape)
    k[0] = [1000, -1000, 7]
    far = pos + k * L
    print(f"[wrap] {precision}: {int((k != 0).any(1).sum())} of {N} ions outside the box, max |k| {np.abs(k).max()}")
    e = P3M(N, 32, L, 1.0, 0.8, L / 3.3, precision=precision)
    e.set_green(dev(P.G_k))
    acc, en = e.force(dev(far))
    torch.cuda.synchronize()
    tol = 1e-9 if precision == "fp64" else 2e-6
    assert np.isfinite(acc.cpu().numpy()).all()
    assert rel_rms(acc.cpu().numpy(), acc_ref) <= tol
    assert float(en[0]) == pytest.approx(U_ref, rel=tol)
    e.close()


def test_ragged_and_tiny_inputs(P3M):
    L = 30.0
    e = P3M(1000, 16, L, 1.0, 0.7, 9.0)
    G, *_ = O.gf_opt(1.0, 0.7, 6, 3, 3, 3, 16, 16, 16, L / 16, L / 16, L / 16, L, L, L)
    e.set_green(dev(G))
    for N in (1, 2, 7, 999):
        pos = np.random.RandomState(N).random_sample((N, 3)) * L
        P = O.P3MParams(N, L, 1.0, 0.7, 9.0, 16, with_green=False)
        P.G_k = G
        P.kx_v = 2 * np.pi * (np.arange(16).reshape(1, 16) - 8) / L
        P.ky_v = P.kx_v.reshape(16, 1)
        P.kz_v = P.kx_v.reshape(16, 1, 1)
        U_ref, acc_ref, _ = P.force(pos)
        acc, en = e.force(dev(pos))
        torch.cuda.synchronize()
        a = acc.cpu().numpy()
        assert np.max(np.abs(a - acc_ref)) <= 1e-12 * max(1.0, np.max(np.abs(acc_ref)))
        assert float(en[0]) == pytest.approx(U_ref, rel=1e-12, abs=1e-12)
    with pytest.raises(Exception):
        e.force(dev(np.zeros((1001, 3))))
    e.close()


def test_n1e6_parity_vs_c_oracle(P3M):
    """N = 1e6 on a 128^3 mesh (the large single-GPU workload): full force
    against the C restatement."""
    N, M = 1_000_000, 128
    L = O.box_length(N)
    kappa, alpha, rc = 2.0, 0.5, 6.5
    pos, _ = O.initial_split(1, 2, N, [L] * 3, 0.01)
    Z = np.ones(N)
    e = P3M(N, M, L, kappa, alpha, rc)
    G_full, _ = e.build_green(6, 3, want_full=True)
    acc, en = e.force(dev(pos), dev(Z))
    torch.cuda.synchronize()
    G = G_full.cpu().numpy()
    U, acc_ref, app, apm, (upp, upm, uself) = c_oracle.force_pot(pos, Z, L, kappa, alpha, rc, G, M)
    en = en.cpu().numpy()
    print("N=1e6 rel rms", rel_rms(acc.cpu().numpy(), acc_ref), "U", en[0], U)
    assert rel_rms(acc.cpu().numpy(), acc_ref) <= TOL
    assert abs(en[0] - U) <= TOL * abs(U)
    assert e.stats()["pp_multipass_tiles"] == 0
    e.close()


def test_large_rc_parity_vs_c_oracle(P3M):
    """rc = 10 with ~330 particles per cell: the PP tile exceeds 64 KB (16-wave
    workgroups) and the LDS capacity (several passes over the tile rows);
    plain FP64 against the C restatement, and the same force in the mixed mode
    to its FP32 bar."""
    N, M = 20000, 32
    L = O.box_length(N)
    kappa, alpha, rc = 0.5, 0.45, 10.0
    pos, _ = O.initial_split(21, 22, N, [L] * 3, 0.02)
    e = P3M(N, M, L, kappa, alpha, rc)
    G_full, _ = e.build_green(6, 3, want_full=True)
    acc, en = e.force(dev(pos), None)
    torch.cuda.synchronize()
    U, acc_ref, app, apm, (upp, upm, uself) = c_oracle.force_pot(pos, np.ones(N), L, kappa, alpha, rc,
                                                                  G_full.cpu().numpy(), M)
    en = en.cpu().numpy()
    st = e.stats()
    print("large rc rel rms", rel_rms(acc.cpu().numpy(), acc_ref), "U", en[0], U, st)
    assert rel_rms(acc.cpu().numpy(), acc_ref) <= TOL
    assert abs(en[0] - U) <= TOL * abs(U)
    assert st["pp_multipass_tiles"] > 0
    e.close()
    em = P3M(N, M, L, kappa, alpha, rc, precision="mixed")
    em.build_green(6, 3)
    accm, enm = em.force(dev(pos), None)
    torch.cuda.synchronize()
    assert rel_rms(accm.cpu().numpy(), acc_ref) <= 2e-6
    em.close()
