"""BF16-MFMA pair network on the GPU against the NumPy oracle (oracle/mlp_oracle.py).

On the GPU's own pair list (same r, u, density features), the oracle
restates forward, correction force, loss and gradient with the same BF16
rounding points; the GPU accumulates in FP32 where the oracle sums in FP64,
and an FP32 sum can move a pre-activation across a BF16 rounding boundary,
so the bars are relative-RMS: outputs and forces <= 2e-3, loss <= 1e-4 and
gradient (norm-wise) <= 1e-2.
"""
import math

import numpy as np
import pytest

from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def setup():
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    from mdforce import mlp
    from mdforce.engine import P3M
    N = 3000
    L = (4 * math.pi * N / 3) ** (1 / 3)
    kappa, alpha, rc = 2.0, 0.86, 3.5
    e = P3M(N, 32, L, kappa, alpha, rc)
    e.build_green()
    meta = mlp.default_meta(kappa, alpha, rc, out_scale=0.5)
    net = mlp.PairMLP(N, 400000, meta)
    p = mlp.init_params(1, out_std=1.0)
    net.set_params(p)
    rs = np.random.RandomState(11)
    pos = torch.from_numpy(rs.random_sample((N, 3)) * L).cuda()
    return dict(mlp=mlp, e=e, net=net, p=p.astype(np.float64), pos=pos, N=N, meta=meta, rs=rs)


def _oracle_inputs(d, meta):
    pi = d["i"].cpu().numpy()
    pj = d["j"].cpu().numpy()
    r = d["r"].double().cpu().numpy()
    u = d["u"].double().cpu().numpy()
    rho = d["rho"].double().cpu().numpy()
    x = M.features(r, M.pair_density(rho, pi, pj), meta)
    return pi, pj, r, u, x


def test_forward_and_correction_match_oracle(setup):
    S = setup
    d = S["net"].pairs(S["e"], S["pos"])
    P = d["npairs"]
    print(f"[mlp] N={S['N']}: {P} ordered pairs")
    assert 40 * S["N"] < P < 50 * S["N"]  # about rc^3 = 42.9 per particle at unit Wigner-Seitz density
    pi, pj, r, u, x = _oracle_inputs(d, S["meta"])
    assert (r < 3.5).all() and (r > 0).all()
    np.testing.assert_allclose(np.linalg.norm(u, axis=1), 1.0, atol=1e-6)
    # every pair appears in both orders
    key = set(zip(pi.tolist(), pj.tolist()))
    assert all((b, a) in key for a, b in list(key)[:2000])
    g_o, _ = M.forward(S["p"], x, S["meta"], bf16_mode=True)
    g_gpu = d["g"].double().cpu().numpy()
    print(f"[mlp] forward rel rms {rel(g_gpu, g_o):.2e}")
    assert rel(g_gpu, g_o) <= 2e-3
    F = S["net"].correction(S["e"], S["pos"]).cpu().numpy()
    perm = d["perm"].cpu().numpy()
    Fs = M.correction(g_o, pi, u, S["N"])
    Fo = np.empty_like(Fs)
    Fo[perm] = Fs
    print(f"[mlp] correction rel rms {rel(F, Fo):.2e}")
    assert rel(F, Fo) <= 2e-3


def test_gradient_matches_oracle(setup):
    S = setup
    N = S["N"]
    target = S["rs"].randn(N, 3) * 0.3
    t_dev = torch.from_numpy(target).cuda()
    d = S["net"].pairs(S["e"], S["pos"])
    g, loss = S["net"].grad(S["e"], S["pos"], t_dev)
    torch.cuda.synchronize()
    g = g.double().cpu().numpy()
    pi, pj, r, u, x = _oracle_inputs(d, S["meta"])
    perm = d["perm"].cpu().numpy()
    L_o, G_o, _ = M.loss_and_grad(S["p"], x, pi, u, target[perm], S["meta"], bf16_mode=True)
    print(f"[mlp] loss gpu {float(loss[0]):.6e} oracle {L_o:.6e}; grad rel {rel(g, G_o):.2e}")
    assert abs(float(loss[0]) - L_o) <= 1e-4 * L_o
    assert rel(g, G_o) <= 1e-2
    for name, (o, s) in M.OFF.items():
        assert rel(g[o:o + s], G_o[o:o + s]) <= 3e-2, name


def test_training_fits_a_teacher(setup):
    """Adam on GPU gradients recovers the correction of a teacher network."""
    S = setup
    mlp = S["mlp"]
    N = S["N"]
    teacher = mlp.PairMLP(N, 400000, S["meta"])
    teacher.set_params(mlp.init_params(7, out_std=1.0))
    target = teacher.correction(S["e"], S["pos"]).clone()
    teacher.close()
    S["net"].set_params(mlp.init_params(3, out_std=0.1))
    losses = []
    for it in range(150):
        _, loss = S["net"].grad(S["e"], S["pos"], target, build=(it == 0))
        S["net"].adam(3e-3)
        losses.append(float(loss[0]))
    torch.cuda.synchronize()
    base = float((target * target).sum(1).mean())
    print(f"[mlp] teacher fit: loss {losses[0]:.3e} -> {losses[-1]:.3e} (|target|^2 {base:.3e})")
    assert losses[-1] < 0.1 * losses[0]
    S["net"].set_params(S["p"].astype(np.float32))


def test_trained_network_reduces_the_residual():
    """Train on equilibrated snapshots (Gamma 100, kappa 2) of the reduced
    setting rc 3.5 / 32^3 mesh against tight FP64 P3M; on a held-out snapshot
    the correction removes part of the residual.  (At this setting the
    residual is ~1e-2 of the mean force and mostly configuration noise; the
    network removes its systematic pair-distance part, 10-20%.)"""
    from mdforce import mlp, mlp_train as T
    N = 27000
    red = lambda k: (32, 0.6, 3.5)
    tight = lambda k: (128, 0.55, 8.0)
    prop = lambda k: (32, 0.7, 4.6)
    train = T.make_snapshots(N, [(100.0, 2.0)], 6, 20, 200, red=red, tight=tight, prop=prop, seed0=300)
    test = T.make_snapshots(N, [(100.0, 2.0)], 1, 20, 200, red=red, tight=tight, prop=prop, seed0=700)
    rms = T.residual_rms(train)
    eng = T.Setting(N, 32, 0.6, 3.5, 2.0).eng
    net = mlp.PairMLP(N, 60 * N, mlp.default_meta(2.0, 0.6, 3.5, out_scale=rms / 5.0))
    hist = T.train(net, {2.0: eng}, train, 15, 2e-3, lr_end=2e-4)
    b, a = T.evaluate(net, eng, test[0]["pos"], test[0]["target"], 2.0, 0.6)
    print(f"[mlp] held-out residual rms {b:.3e} -> {a:.3e} (x{a / b:.3f}); loss {hist[0]:.3e} -> {hist[-1]:.3e}")
    assert hist[-1] < hist[0]
    assert a < 0.95 * b
    # fallback: at a 1e-5 budget the corrected reduced setting is judged against plain P3M and refused
    plain = T.Setting(N, 64, 0.7, 6.0, 2.0).eng
    form, est = T.choose_form(net, eng, plain, test[0]["pos"], 2.0, 0.6, 1e-5)
    print(f"[mlp] judged error {est:.3e} -> {form}")
    assert form == "plain" and est > 1e-5
    form2, _ = T.choose_form(net, eng, plain, test[0]["pos"], 2.0, 0.6, 1.0)
    assert form2 == "learned"


def test_fallback_switch_routes_to_the_plain_engine(setup):
    """LearnedForce on the device engines: a network far outside the budget is
    refused, the network is detached and the driver's forces are the plain
    engine's.  Two plain contexts are not bitwise alike (the binning slots and
    the FP64 charge deposits are atomics taken in arrival order), so the bar
    is the plain path's own run-to-run spread: 1e-13 relative."""
    from mdforce import mlp, mlp_train as T
    from mdforce.engine import P3M
    S = setup
    N, pos = S["N"], S["pos"]
    L = (4 * math.pi * N / 3) ** (1 / 3)
    red = S["e"]
    net = mlp.PairMLP(N, 400000, S["meta"])
    net.set_params(mlp.init_params(5, out_std=1.0))
    plain = P3M(N, 32, L, 2.0, 0.7, 6.0)
    plain.build_green()
    plain2 = P3M(N, 32, L, 2.0, 0.7, 6.0)
    plain2.build_green()
    drv = T.LearnedForce(net, red, plain, 2.0, 0.86, budget=1e-5)
    form, est = drv.decide(pos)
    print(f"[fallback] judged {est:.3e} -> {form}")
    assert form == "plain" and red._mlp is None
    a_drv = drv.force(pos, None, af=1.0)[0].clone()
    a_ref = plain2.force(pos, None, af=1.0)[0].clone()
    torch.cuda.synchronize()
    d = float(torch.sqrt(((a_drv - a_ref) ** 2).sum(1).mean() / (a_ref ** 2).sum(1).mean()))
    assert d <= 1e-13
    # three NVE steps through the driver follow the plain context's steps
    p1, p2 = pos.clone(), pos.clone()
    v1 = torch.zeros_like(p1)
    v2 = torch.zeros_like(p2)
    e1, _ = drv.md_run(p1, v1, a_drv, 3, 0.05)
    e2, _ = plain2.md_run(p2, v2, a_ref, 3, 0.05)
    torch.cuda.synchronize()
    dp = (p1 - p2).abs()
    assert float(torch.minimum(dp, L - dp).max()) <= 1e-12 * L
    np.testing.assert_allclose(e1.cpu().numpy(), e2.cpu().numpy(), rtol=1e-12, atol=1e-12)
    # within a loose budget the network is attached and its correction enters
    drv.budget = 1e9
    assert drv.decide(pos)[0] == "learned" and red._mlp is net
    a_l = drv.force(pos, None, af=1.0)[0].clone()
    mlp.PairMLP.detach(red)
    a_r = red.force(pos, None, af=1.0)[0].clone()
    corr = net.correction(red, pos)
    torch.cuda.synchronize()
    assert float(((a_l - a_r - corr) ** 2).sum().sqrt() / (corr ** 2).sum().sqrt()) <= 1e-5
    net.close()
    plain.close()
    plain2.close()
