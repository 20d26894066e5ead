"""S(q, omega) on the device against the oracle (oracle/sqw_oracle.py).

The device path (mdf_sqw: FP64 mixed-radix Stockham FFT along t, then
|dt X|^2 / (N T)) is compared with numpy.fft on the same series.  Both are
FP64 transforms of the same data, so S agrees to rounding: the bar is
|S_gpu - S_oracle| <= 1e-11 of the largest S of that series (a transform
error of ~1e-15 relative, squared into S).  Lengths cover one factor of each
radix (4, 2, 3, 5), a prime above 5 (997, 13), mixed products, length 1 and
cfg1's 16000 = 4^3 2 5^3 at the full ensemble shape (20 runs, 13 q).
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import sqw_oracle as O  # noqa: E402


def _dev():
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    return torch.device("cuda", 0)


def _series(nrun, npd, ncol, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((nrun, npd, ncol, 3)) + 1j * rng.standard_normal((nrun, npd, ncol, 3))


def _close(got, ref, axis):
    scale = np.max(ref, axis=axis, keepdims=True)
    assert np.all(np.abs(got - ref) <= 1e-11 * scale + 1e-300), np.max(np.abs(got - ref) / scale)


@pytest.mark.parametrize("npd", [1, 2, 3, 4, 5, 13, 64, 100, 210, 997, 1000, 4096, 2 * 997])
def test_avg_matches_oracle(npd):
    from mdforce.sqw import sqw
    dev = _dev()
    nrun, Nq, N, dt = 2, 3, 2000.0, 0.05
    T = npd * dt
    x = _series(nrun, npd, 3 * Nq, npd)
    got = sqw(torch.from_numpy(x).to(dev), N, dt, T, q_stride=3, nq=Nq).cpu().numpy()
    ref = O.s_of_q_w_avg_mat(x, Nq, np.zeros((nrun, Nq, npd)), nrun, N, dt, T)
    _close(got, ref, axis=2)


@pytest.mark.parametrize("npd", [7, 48, 250])
def test_per_axis_matches_oracle(npd):
    from mdforce.sqw import sqw
    dev = _dev()
    N, dt = 300.0, 0.1
    T = npd * dt
    x = _series(1, npd, 5, 11)
    got = sqw(torch.from_numpy(x).to(dev), N, dt, T, q_stride=1, mean_axes=False).cpu().numpy()
    for i in range(5):
        _close(got[0, i], O.n_of_q_w(x[0, :, i, :], N, dt, T), axis=0)


def test_dropins_numpy_in_numpy_out():
    _dev()
    from mdforce.s_of_q_w import n_of_q_w
    from mdforce.s_of_q_w_avg import s_of_q_w_avg_mat
    npd, Nq, n_avg, N, dt = 160, 4, 3, 1000.0, 0.05
    T = npd * dt
    x = _series(n_avg, npd, 3 * Nq, 5)
    s1 = n_of_q_w(x[0, :, 0, :], N, dt, T)
    assert isinstance(s1, np.ndarray) and s1.shape == (npd, 3)
    _close(s1, O.n_of_q_w(x[0, :, 0, :], N, dt, T), axis=0)
    mat = np.zeros((n_avg, Nq, npd))
    ret = s_of_q_w_avg_mat(x, Nq, mat, n_avg, N, dt, T)
    assert ret is mat
    _close(mat, O.s_of_q_w_avg_mat(x, Nq, np.zeros((n_avg, Nq, npd)), n_avg, N, dt, T), axis=2)


def test_full_ensemble_shape_and_parseval():
    """cfg1's post-processing shape: 20 runs x 13 q (39 columns) x 16000 steps."""
    from mdforce.sqw import sqw
    dev = _dev()
    n_avg, Nq, npd, N, dt = 20, 13, 16000, 10000.0, 0.05
    T = npd * dt
    g = torch.Generator(device=dev).manual_seed(0)
    x = torch.randn((n_avg, npd, 3 * Nq, 3), dtype=torch.complex128, device=dev, generator=g)
    per_axis = sqw(x, N, dt, T, q_stride=3, nq=Nq, mean_axes=False)
    avg = sqw(x, N, dt, T, q_stride=3, nq=Nq)
    # Parseval on every series, at full size
    e_t = (x[:, :, 0::3, :].abs() ** 2).sum(dim=1)                    # (run, q, axis)
    e_w = per_axis.sum(dim=2) * (N * T) / (dt * dt * npd)              # (run, q, axis)
    assert torch.allclose(e_w, e_t, rtol=1e-12, atol=0)
    assert torch.allclose(avg, per_axis.mean(dim=3), rtol=1e-15, atol=0)
    # the oracle on two runs of the same data
    xs = x[:2].cpu().numpy()
    ref = O.s_of_q_w_avg_mat(xs, Nq, np.zeros((2, Nq, npd)), 2, N, dt, T)
    _close(avg[:2].cpu().numpy(), ref, axis=2)


def test_md_run_modes_to_sqw():
    """n(q, t) recorded by mdf_md_run on the device, transformed in place."""
    from mdforce.engine import P3M
    from mdforce.initialize_pos_vel import initial_split
    from mdforce.sqw import sqw
    dev = _dev()
    n, gamma, kappa, steps, dt = 2000, 50.0, 1.0, 256, 0.1
    L = (4.0 * math.pi * n / 3.0) ** (1.0 / 3.0)
    pos, vel = initial_split(1, 2, n, [L, L, L], 1.0 / gamma)
    e = P3M(n, 32, L, kappa, 0.7, L / 3.0 * 0.999)
    e.build_green()
    p = torch.from_numpy(pos).to(dev)
    v = torch.from_numpy(vel).to(dev)
    a, _ = e.force(p)
    a = a.clone()
    Nq = 4
    _, nqt = e.md_run(p, v, a, steps, dt, q_step=2 * math.pi / L, nq=3 * Nq)
    T = steps * dt
    got = sqw(nqt, n, dt, T, q_stride=3, nq=Nq).cpu().numpy()
    x = nqt.cpu().numpy()[None]
    ref = O.s_of_q_w_avg_mat(x, Nq, np.zeros((1, Nq, steps)), 1, n, dt, T)
    _close(got, ref, axis=2)
    e.close()


def test_bad_arguments_raise():
    from mdforce._lib import MdforceError
    from mdforce.sqw import sqw
    dev = _dev()
    x = torch.zeros((4, 2, 3), dtype=torch.complex128, device=dev)
    with pytest.raises(MdforceError):
        sqw(x, 100.0, 0.1, 0.0)          # N T = 0
    with pytest.raises(MdforceError):
        sqw(x, 100.0, 0.1, 0.4, q_stride=3, nq=2)   # column 3 beyond nq_total 2
