"""The benchmark configurations at their full sizes, and cross-checks of the
FFT passes and the Green function at the N = 1e6 mesh.

* cfg3: the 64 replicas of the replica ensemble (RandomState(3) draws of N in
  2,000..50,000, Gamma and kappa), stepped together by the segmented graph,
  each against itself stepped alone; three replicas' FP64 forces against the
  C restatement at their own N.
* cfg5: N = 8e6, Gamma 50, kappa 0.5 on its stated 256^3 mesh and on 512^3,
  both forms, on a liquid snapshot: sum(rho) = N, zero net PP force, the
  FP32-mixed form within 2e-6 of FP64, FP64 within 1e-5 of the mean force of
  a tight FP64 reference (rc 15, alpha 0.32, 512^3).
* 128^3: the device G(k) against the C restatement of gf_opt, and the device
  FFT / Green / ik / inverse passes against rocFFT (torch.fft) on the same
  charge mesh.
"""
import math

import numpy as np
import pytest

from oracle import c_oracle
from oracle import p3m_oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def rel_rms(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.sqrt(np.mean(np.sum((a - b) ** 2, -1)) / np.mean(np.sum(b ** 2, -1))))


@pytest.fixture(scope="module")
def mods():
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    from mdforce import autotune, batch, engine, mlp_train
    return batch, engine, autotune, mlp_train


def cfg3_systems(precision):
    from mdforce.initialize_pos_vel import initial_split
    rs = np.random.RandomState(3)
    R = 64
    Ns = rs.randint(2000, 50001, R)
    gammas = rs.choice([10.0, 50.0, 100.0, 175.0], R)
    kappas = rs.choice([0.5, 1.0, 2.0, 3.0], R)
    systems, states = [], []
    for r in range(R):
        N = int(Ns[r])
        L = (4.0 * math.pi * N / 3.0) ** (1.0 / 3.0)
        rc = min(5.0, L / 3.0 * (1.0 - 1e-9))
        systems.append(dict(n=N, L=L, kappa=float(kappas[r]), mesh=32 if N <= 10000 else 64, alpha=3.2 / rc,
                            rc=rc, precision=precision))
        states.append(initial_split(1000 + r, 2000 + r, N, [L, L, L], 1.0 / gammas[r]))
    return systems, states


def test_cfg3_64_replicas_graph_matches_each_alone(mods):
    batch, engine, _, _ = mods
    systems, states = cfg3_systems("fp64")
    assert len(systems) == 64 and sum(s["n"] for s in systems) > 1e6
    nsteps, K, nq, dt = 6, 3, 20, 0.1
    singles = []
    for s, (pos, vel) in zip(systems, states):
        e = engine.P3M(s["n"], s["mesh"], s["L"], s["kappa"], s["alpha"], s["rc"])
        e.build_green()
        p = torch.from_numpy(pos).cuda()
        v = torch.from_numpy(vel).cuda()
        a, _ = e.force(p)
        a = a.clone()
        en, nqt = e.md_run(p, v, a, nsteps, dt, q_step=2 * math.pi / s["L"], nq=nq)
        torch.cuda.synchronize()
        singles.append((p, a, en, nqt))
        e.close()
    torch.cuda.synchronize()
    b = batch.ReplicaBatch(systems, n_streams=8)
    b.set_mode(0)
    P, V, A = [], [], []
    for rep, (pos, vel) in zip(b.replicas, states):
        p = torch.from_numpy(pos).cuda()
        a, _ = rep.force(p)
        P.append(p)
        V.append(torch.from_numpy(vel).cuda())
        A.append(a.clone())
    ens, nqts = b.run(P, V, A, nsteps, dt, nq=nq, graph_steps=K)
    torch.cuda.synchronize()
    assert b.graph_mode == 0
    worst = 0.0
    for r, s in enumerate(systems):
        p, a, en, nqt = (t.cpu().numpy() for t in singles[r])
        d = np.abs(P[r].cpu().numpy() - p)
        dx = float(np.minimum(d, s["L"] - d).max() / s["L"])
        da = float(np.abs(A[r].cpu().numpy() - a).max() / np.abs(a).max())
        worst = max(worst, dx, da)
        # the charge mesh sums its FP64 brick flushes with atomics, in an order
        # that differs between runs; six steps of the random start's release
        # (close pairs thrown apart) carry that rounding to ~2e-11 of max|a|
        assert dx <= 1e-11 and da <= 1e-10, (r, s["n"], dx, da)
        np.testing.assert_allclose(ens[r].cpu().numpy(), en, rtol=1e-11, atol=1e-11 * np.abs(en).max())
        np.testing.assert_allclose(nqts[r].cpu().numpy(), nqt, rtol=0, atol=1e-11 * s["n"])
    print(f"[cfg3] 64 replicas, sum N = {sum(s['n'] for s in systems)}: worst relative difference {worst:.2e}")
    b.close()


def test_cfg3_replica_forces_against_c_restatement(mods):
    _, engine, _, _ = mods
    systems, states = cfg3_systems("fp64")
    order = np.argsort([s["n"] for s in systems])
    for r in (order[0], order[len(order) // 2], order[-1]):
        s, (pos, _) = systems[r], states[r]
        e = engine.P3M(s["n"], s["mesh"], s["L"], s["kappa"], s["alpha"], s["rc"])
        G_full, _ = e.build_green(want_full=True)
        a, en = e.force(torch.from_numpy(pos).cuda())
        torch.cuda.synchronize()
        U_c, a_c, _, _, _ = c_oracle.force_pot(pos, np.ones(s["n"]), s["L"], s["kappa"], s["alpha"], s["rc"],
                                               G_full.cpu().numpy(), s["mesh"])
        err = rel_rms(a.cpu().numpy(), a_c)
        dU = abs(float(en[0]) - U_c) / abs(U_c)
        print(f"[cfg3] replica {r}: N={s['n']} err {err:.2e} dU {dU:.2e}")
        assert err <= 1e-12 and dU <= 1e-12
        e.close()


@pytest.fixture(scope="module")
def cfg5_liquid(mods):
    """N = 8e6, Gamma 50, kappa 0.5: the seeded start equilibrated (100 steps
    at dt/10, 300 rescaled steps at dt 0.05 with a speed clip), and the tight
    FP64 reference force of that state."""
    _, engine, autotune, mlp_train = mods
    from mdforce.initialize_pos_vel import initial_split
    N, kappa, gamma, dt = 8_000_000, 0.5, 50.0, 0.05
    L = (4.0 * math.pi * N / 3.0) ** (1.0 / 3.0)
    pos, vel = initial_split(1, 2, N, [L, L, L], 1.0 / gamma)
    p = torch.from_numpy(pos).cuda()
    v = torch.from_numpy(vel).cuda()
    del pos, vel
    eq = engine.P3M(N, 256, L, kappa, 0.5, 6.0, precision="mixed")
    eq.build_green()
    mlp_train.equilibrate(eq, p, v, 1.0 / gamma, 300, dt, vcap=4.0, soft_steps=100)
    eq.close()
    del v
    ref = autotune.reference_force(engine.P3M, p, L, kappa)
    torch.cuda.synchronize()
    return dict(N=N, L=L, kappa=kappa, pos=p, ref=ref)


@pytest.mark.parametrize("mesh,alpha,rc", [(256, 0.25, 14.0), (512, 0.42, 8.5)])
def test_cfg5_n8e6_properties_and_error(mods, cfg5_liquid, mesh, alpha, rc):
    _, engine, autotune, _ = mods
    S = cfg5_liquid
    N, L, kappa, p, ref = S["N"], S["L"], S["kappa"], S["pos"], S["ref"]
    acc = {}
    for prec in ("fp64", "mixed"):
        e = engine.P3M(N, mesh, L, kappa, alpha, rc, precision=prec)
        e.build_green()
        rho = torch.empty((mesh, mesh, mesh), dtype=torch.float64, device="cuda")
        a_pm, _ = e.pm(p, rho=rho)
        torch.cuda.synchronize()
        dq = float(rho.sum()) - N
        a_pp, _ = e.pp(p)
        torch.cuda.synchronize()
        net = a_pp.sum(0).abs().max().item()
        scale = a_pp.abs().sum().item() / N
        print(f"[cfg5] {mesh}^3 {prec}: sum(rho) - N = {dq:.3e}; net PP force {net:.3e} "
              f"(mean |component| {scale:.3e})", flush=True)
        assert abs(dq) <= 1e-9 * N  # sum(rho) = sum(Z)
        # every pair enters from both ends with opposite sign: FP64 leaves only
        # rounding; FP32 sums leave a random walk of the per-ion errors (2e-6 bar)
        bar = 1e-12 * scale * N if prec == "fp64" else 10 * 2e-6 * scale * math.sqrt(N)
        assert net <= bar, (prec, net, bar)
        a, _ = e.force(p, None, af=1.0)
        acc[prec] = a.clone()
        e.close()
        del rho, a_pm, a_pp
    err64 = autotune.force_error(acc["fp64"], ref)
    d = acc["mixed"] - acc["fp64"]
    mix = float(torch.sqrt(torch.mean(torch.sum(d * d, 1)) / torch.mean(torch.sum(acc["fp64"] ** 2, 1))))
    print(f"[cfg5] {mesh}^3 alpha {alpha} rc {rc}: FP64 vs tight {err64:.2e} of the mean force; "
          f"mixed vs FP64 {mix:.2e}")
    assert err64 <= 1e-5
    assert mix <= 2e-6


def test_green_128_against_c_restatement(mods):
    _, engine, _, _ = mods
    N, M, kappa, alpha = 1_000_000, 128, 2.0, 0.42
    L = (4.0 * math.pi * N / 3.0) ** (1.0 / 3.0)
    e = engine.P3M(1000, M, L, kappa, alpha, 7.0)
    G_dev, A_dev = e.build_green(6, 3, want_full=True)
    torch.cuda.synchronize()
    G_c, A_c = c_oracle.gf_opt(kappa, alpha, 6, 3, M, L, threads=0)
    g = G_dev.cpu().numpy()
    assert np.max(np.abs(g - G_c)) <= 1e-13 * np.max(np.abs(G_c))
    assert abs(float(A_dev[0]) - A_c) <= 1e-11 * abs(A_c)
    e.close()


def test_fft_passes_against_rocfft_128(mods):
    """E_p of the device PM solve against a PM solve whose transforms are
    rocFFT's (torch.fft on the GPU): the same charge mesh (the device's own
    assignment), G(k) shifted back to natural order, E_k = -i k G rho_k, the
    reference's real(ifftn) and M^3 / V scaling (yukawa_pm_ord6.py:447-480), and
    the oracle's gather on the host."""
    _, engine, _, _ = mods
    N, M, kappa, alpha, rc = 200_000, 128, 2.0, 0.42, 7.0
    L = (4.0 * math.pi * N / 3.0) ** (1.0 / 3.0)
    pos, _ = O.initial_split(11, 12, N, [L, L, L], 0.01)
    e = engine.P3M(N, M, L, kappa, alpha, rc)
    G_full, _ = e.build_green(6, 3, want_full=True)
    p = torch.from_numpy(pos).cuda()
    rho = torch.empty((M, M, M), dtype=torch.float64, device="cuda")
    ex = torch.empty(N, dtype=torch.float64, device="cuda")
    ey = torch.empty_like(ex)
    ez = torch.empty_like(ex)
    e.pm(p, rho=rho, e_x=ex, e_y=ey, e_z=ez)
    torch.cuda.synchronize()
    # rocFFT route
    rk = torch.fft.fftn(rho)
    G_nat = torch.fft.ifftshift(G_full)
    n = torch.fft.fftfreq(M, d=1.0 / M, device="cuda", dtype=torch.float64)
    k = 2.0 * math.pi * n / L
    V = L ** 3
    fields = []
    for ax in range(3):  # mesh layout (z, y, x)
        shape = [1, 1, 1]
        shape[2 - ax] = M
        kk = k.reshape(shape)
        Ek = -1j * kk * G_nat * rk
        fields.append((torch.fft.ifftn(Ek).real * (M ** 3 / V)).cpu().numpy())
    h = L / M
    Ep = O.field_gather(fields[0], fields[1], fields[2], pos, np.ones(N), M, M, M, h, h, h)[1]
    got = np.stack([ex.cpu().numpy(), ey.cpu().numpy(), ez.cpu().numpy()], 1)
    err = rel_rms(got, Ep)
    print(f"[rocfft] 128^3 E_p device vs rocFFT route: {err:.2e}")
    assert err <= 1e-12
    e.close()
