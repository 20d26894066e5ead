"""FP32-mixed P3M (MDF_PREC_MIXED) against the FP64 oracle.

The mixed mode runs the PP pair math, the charge-assignment weights and the
FFT/Green/ik passes in FP32, and keeps the charge-mesh atomics, the energy
reductions and the positions in FP64.  It is not bit-comparable with the
oracle; the bar is an error far below the 1e-5 force-error budget of the
method: relative RMS force error <= 2e-6 and |dU|/|U| <= 1e-6 against the
oracle on identical inputs.
"""
import numpy as np
import pytest

from oracle import p3m_oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

FORCE_TOL = 2e-6
ENERGY_TOL = 1e-6


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
    N = 10000
    L = O.box_length(N)
    P = O.P3MParams(N, L, 2.0, 0.85, 6.2, 64)
    pos, vel = O.initial_split(1, 2, N, P.Lv, 1.0 / 100)
    Z = np.ones(N)
    U, acc, parts = P.force(pos, Z)
    return dict(P=P, pos=pos, vel=vel, Z=Z, U=U, acc=acc, parts=parts)


def test_mixed_force_close_to_oracle(P3M, cfg1):
    P = cfg1["P"]
    N = P.N
    e = P3M(N, 64, P.L, P.kappa, P.G, P.rc, precision="mixed")
    e.set_green(dev(P.G_k))
    acc_pp = torch.empty((N, 3), dtype=torch.float64, device="cuda")
    acc_pm = torch.empty_like(acc_pp)
    rho = torch.empty((64, 64, 64), dtype=torch.float64, device="cuda")
    acc, en = e.force(dev(cfg1["pos"]), dev(cfg1["Z"]), af=P.af, uf=P.uf, acc_pp=acc_pp, acc_pm=acc_pm, rho=rho)
    torch.cuda.synchronize()
    parts = cfg1["parts"]
    en = en.cpu().numpy()
    e_acc = rel_rms(acc.cpu().numpy(), cfg1["acc"])
    e_pp = rel_rms(acc_pp.cpu().numpy(), parts["acc_s_r"])
    e_pm = rel_rms(acc_pm.cpu().numpy(), parts["acc_fft"])
    dU = abs(en[0] - cfg1["U"]) / abs(cfg1["U"])
    dUpp = abs(en[1] - parts["U_short"]) / abs(parts["U_short"])
    dUpm = abs(en[2] - parts["U_fft"]) / abs(parts["U_fft"])
    print(f"mixed cfg1: rel rms acc {e_acc:.3e} pp {e_pp:.3e} pm {e_pm:.3e}; dU {dU:.3e} pp {dUpp:.3e} pm {dUpm:.3e}")
    assert e_acc <= FORCE_TOL
    assert e_pp <= FORCE_TOL
    assert e_pm <= FORCE_TOL
    assert dU <= ENERGY_TOL
    assert dUpp <= ENERGY_TOL
    assert dUpm <= ENERGY_TOL
    assert en[3] == pytest.approx(parts["U_self"], rel=1e-15)
    # FP64 mesh accumulation of FP32 weights: total charge to FP32 weight precision
    assert float(rho.sum()) == pytest.approx(N, rel=1e-6)
    np.testing.assert_allclose(rho.cpu().numpy(), parts["rho_r"], atol=1e-6 * float(np.abs(parts["rho_r"]).max()))


def test_mixed_momentum_and_repeatability(P3M, cfg1):
    """Net PP force vanishes to FP32 rounding; repeated calls agree."""
    P = cfg1["P"]
    N = P.N
    e = P3M(N, 64, P.L, P.kappa, P.G, P.rc, precision="mixed")
    e.build_green()
    pos = dev(cfg1["pos"])
    acc_pp = torch.empty((N, 3), dtype=torch.float64, device="cuda")
    a1, u1 = e.force(pos, None, acc_pp=acc_pp)
    a1 = a1.clone()
    u1 = u1.clone()
    a2, u2 = e.force(pos, None)
    torch.cuda.synchronize()
    s = acc_pp.sum(0).abs().max().item()
    scale = acc_pp.abs().mean().item()
    print(f"net PP force {s:.3e} (mean |a_pp| {scale:.3e})")
    assert s <= 1e-4 * scale * np.sqrt(N)
    # LDS float atomics make the assignment order-dependent at FP32 rounding
    assert rel_rms(a2.cpu().numpy(), a1.cpu().numpy()) <= 1e-6
    assert abs(u2[0].item() - u1[0].item()) <= 1e-7 * abs(u1[0].item())


def test_mixed_n1e5_error_vs_fp64(P3M):
    """N = 1e5 at cfg2's density, 64^3 mesh: mixed vs the FP64 GPU path."""
    N = 100000
    L = O.box_length(N)
    kappa, alpha, rc, M = 2.0, 0.48, 5.5, 64
    rng = np.random.default_rng(7)
    pos = dev(rng.random((N, 3)) * L)
    outs = {}
    for prec in ("fp64", "mixed"):
        e = P3M(N, M, L, kappa, alpha, rc, precision=prec)
        e.build_green()
        acc, en = e.force(pos, None)
        torch.cuda.synchronize()
        outs[prec] = (acc.cpu().numpy(), en.cpu().numpy())
    a64, u64 = outs["fp64"]
    a32, u32 = outs["mixed"]
    err = rel_rms(a32, a64)
    mean_f = float(np.mean(np.linalg.norm(a64, axis=1)))
    err_mean = float(np.sqrt(np.mean(np.sum((a32 - a64) ** 2, axis=1)))) / mean_f
    print(f"N=1e5 mixed vs fp64: rel rms {err:.3e}, rms / mean|F| {err_mean:.3e}, dU {abs(u32[0]-u64[0])/abs(u64[0]):.3e}")
    assert err <= FORCE_TOL
    assert err_mean <= FORCE_TOL
    assert abs(u32[0] - u64[0]) <= ENERGY_TOL * abs(u64[0])


@pytest.mark.parametrize("rc, waves", [(5.5, 4), (5.5, 8), (7.0, 12), (8.5, 16)])
def test_mixed_pp_wave_counts_vs_fp64(P3M, rc, waves):
    """Each FP32 PP kernel variant against the FP64 GPU path at unit density:
    4 and 8 waves, 12 waves with the packed 24-byte tile, and 16 waves with
    tiles over the LDS capacity (several passes), the wave count set with
    mdf_p3m_set_launch."""
    kappa = 2.0
    alpha = 2.8 / rc
    N = 60000
    L = O.box_length(N)
    e_mixed = P3M(N, 64, L, kappa, alpha, rc, precision="mixed")
    e_mixed.set_launch(waves)
    assert e_mixed.config()["pp_waves"] == waves
    rng = np.random.default_rng(11)
    pos = dev(rng.random((N, 3)) * L)
    outs = {}
    for prec in ("fp64", "mixed"):
        e = e_mixed if prec == "mixed" else P3M(N, 64, L, kappa, alpha, rc, precision=prec)
        e.build_green()
        acc, en = e.force(pos, None)
        torch.cuda.synchronize()
        outs[prec] = (acc.cpu().numpy(), en.cpu().numpy(), e.stats())
        e.close()
    a64, u64, _ = outs["fp64"]
    a32, u32, st = outs["mixed"]
    err = rel_rms(a32, a64)
    print(f"rc {rc}, N {N}: {waves} waves, multipass tiles {st['pp_multipass_tiles']}: rel rms {err:.3e}, "
          f"dU {abs(u32[0] - u64[0]) / abs(u64[0]):.3e}")
    if waves == 16:
        assert st["pp_multipass_tiles"] > 0
    assert err <= FORCE_TOL
    assert abs(u32[0] - u64[0]) <= ENERGY_TOL * abs(u64[0])


def test_mixed_md_run_tracks_fp64(P3M, cfg1):
    """20 NVE steps from the same state: mixed trajectory stays close to FP64.

    The uniform random start is far from equilibrium (close pairs, large
    forces), so the 1e-7 force differences grow chaotically; after 20 steps
    positions differ by ~1e-5.  The bound 1e-4 is a tenth of a percent of the
    interparticle spacing.
    """
    P = cfg1["P"]
    N = P.N
    res = {}
    for prec in ("fp64", "mixed"):
        e = P3M(N, 64, P.L, P.kappa, P.G, P.rc, precision=prec)
        e.set_green(dev(P.G_k))
        pos = dev(cfg1["pos"])
        vel = dev(cfg1["vel"])
        acc, _ = e.force(pos, None, af=P.af)
        acc = acc.clone()
        en = torch.zeros((20, 4), dtype=torch.float64, device="cuda")
        e.md_run(pos, vel, acc, 20, 0.1, energies=en)
        torch.cuda.synchronize()
        res[prec] = (pos.cpu().numpy(), en.cpu().numpy())
    p64, e64 = res["fp64"]
    p32, e32 = res["mixed"]
    dpos = np.abs(p32 - p64)
    dpos = np.minimum(dpos, P.L - dpos).max()
    dE = np.abs(e32[:, 0] - e64[:, 0]).max() / (N / 100.0)  # per N T
    print(f"20 steps: max |dx| {dpos:.3e}, max |dU|/(N T) {dE:.3e}")
    assert dpos <= 1e-4
    assert dE <= 1e-4
