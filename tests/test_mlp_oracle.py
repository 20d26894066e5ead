"""CPU checks of the MLP oracle: finite-difference gradients and BF16 rounding."""
import numpy as np

from oracle import mlp_oracle as M

META = dict(mu0=0.25, dmu=0.3, inv_sigma=1.0 / 0.3, f_kappa=2.0 / 3.0, f_alpha=0.86, f_rc=3.5 / 8.0, out_scale=0.5)


def _problem(seed=0, P=300, n=40):
    rs = np.random.RandomState(seed)
    r = rs.uniform(0.4, 3.5, P)
    rho = rs.uniform(-0.3, 0.3, n)
    pi = rs.randint(0, n, P)
    pj = rs.randint(0, n, P)
    u = rs.randn(P, 3)
    u /= np.linalg.norm(u, axis=1)[:, None]
    x = M.features(r, M.pair_density(rho, pi, pj), META)
    t = rs.randn(n, 3) * 0.3
    return x, pi, u, t


def test_bf16_known_answers():
    v = np.array([1.0, -2.5, 3.0e38, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, 0.0, -0.0], dtype=np.float32)
    b = M.bf16(v)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: ties go to the even mantissa (1.0);
    # 1 + 3*2^-8 ties between 1 + 2^-7 and 1 + 2^-6: even is 1 + 2^-6
    assert b[0] == 1.0 and b[1] == -2.5
    assert b[3] == 1.0
    assert b[4] == np.float32(1.0 + 2 ** -6)
    assert b[5] == 1.0
    assert b[6] == 0.0 and np.signbit(b[7])
    assert (M.bf16(v).view(np.uint32) & 0xFFFF == 0).all()
    # 3e38 = 1.7632 * 2^127: the 7-bit mantissa 97.69 rounds to 98
    assert M.bf16(np.float32(3.0e38)) == np.float32((1 + 98 / 128) * 2.0 ** 127)


def test_fp64_gradient_matches_finite_differences():
    x, pi, u, t = _problem()
    p = M.init_params(3, scale_out=1.0)
    L, g, _ = M.loss_and_grad(p, x, pi, u, t, META, bf16_mode=False)
    rs = np.random.RandomState(5)
    idx = np.concatenate([rs.choice(M.N_PARAM - 1, 40, replace=False), [M.N_PARAM - 1]])
    h = 1e-6
    for k in idx:
        pp = p.copy()
        pm = p.copy()
        pp[k] += h
        pm[k] -= h
        Lp = M.loss_and_grad(pp, x, pi, u, t, META, bf16_mode=False)[0]
        Lm = M.loss_and_grad(pm, x, pi, u, t, META, bf16_mode=False)[0]
        fd = (Lp - Lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_correction_is_a_pair_sum():
    x, pi, u, t = _problem(1)
    p = M.init_params(2)
    g, _ = M.forward(p, x, META, bf16_mode=False)
    F = M.correction(g, pi, u, t.shape[0])
    k = 7
    sel = pi == k
    np.testing.assert_allclose(F[k], (g[sel, None] * u[sel]).sum(0), rtol=1e-14, atol=1e-14)
