"""Error paths of the HIP engine that the host turns into errors: a PP tile
row denser than the context's LDS tile, and caller buffers of the wrong shape
or type.  Every kernel here exits normally; the error is raised on the host."""
import numpy as np
import pytest

from oracle import p3m_oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


@pytest.fixture(scope="module")
def P3M():
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    from mdforce.engine import P3M as cls
    return cls


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_tile_row_over_capacity_raises(P3M, precision):
    """600 ions packed into one sub-cell of a 2,000-ion context (tile capacity
    512): the PP kernel skips that row and flags it; force, md_run and the
    PP-only call then raise instead of returning incomplete forces, and the
    context computes normally again on a regular configuration."""
    from mdforce._lib import MdforceError

    N, L, rc, M = 2000, 30.0, 3.0, 32
    e = P3M(N, M, L, 1.0, 1.2, rc, precision=precision)
    assert e.config()["tile_max"] < 600
    e.build_green()
    rng = np.random.RandomState(3)
    pos = rng.random_sample((N, 3)) * L
    # one sub-cell of width L / 20 = 1.5: 600 ions inside [15.1, 15.9)^3
    pos[:600] = 15.1 + 0.8 * rng.random_sample((600, 3))
    p = dev(pos)
    with pytest.raises(MdforceError, match="tile capacity"):
        e.force(p)
    with pytest.raises(MdforceError, match="tile capacity"):
        e.pp(p)
    vel = torch.zeros_like(p)
    acc = torch.zeros_like(p)
    with pytest.raises(MdforceError, match="tile capacity"):
        e.md_run(p.clone(), vel, acc, 2, 0.01)
    # a regular configuration on the same context: no error, forces match the oracle
    pos2 = rng.random_sample((N, 3)) * L
    P = O.P3MParams(N, L, 1.0, 1.2, rc, M)
    U_ref, acc_ref, _ = P.force(pos2)
    e2 = P3M(N, M, L, 1.0, 1.2, rc, precision=precision)
    e2.set_green(dev(P.G_k))
    a, _ = e2.force(dev(pos2))
    e.set_green(dev(P.G_k))
    a1, en = e.force(dev(pos2))
    torch.cuda.synchronize()
    # the mesh sums use atomics: equal to rounding (FP64) or to the FP32
    # rounding of the mesh copy (mixed), not bit for bit
    tol = 1e-12 if precision == "fp64" else 1e-6
    assert torch.allclose(a1, a, rtol=0, atol=tol * float(a.abs().max()))
    err = np.sqrt(np.mean(np.sum((a1.cpu().numpy() - acc_ref) ** 2, 1)) / np.mean(np.sum(acc_ref ** 2, 1)))
    assert err < (1e-12 if precision == "fp64" else 2e-6)
    e.close()
    e2.close()


def test_engine_rejects_bad_buffers(P3M):
    N, L, M = 500, 20.0, 16
    e = P3M(N, M, L, 1.0, 1.0, 5.0)
    e.build_green()
    rng = np.random.RandomState(4)
    pos = dev(rng.random_sample((N, 3)) * L)
    vel = torch.zeros_like(pos)
    acc = torch.zeros_like(pos)
    with pytest.raises(ValueError, match="energies"):
        e.md_run(pos, vel, acc, 3, 0.01, energies=torch.empty((2, 4), dtype=torch.float64, device="cuda"))
    with pytest.raises(TypeError, match="nqt"):
        e.md_run(pos, vel, acc, 3, 0.01, q_step=0.3, nq=4,
                 nqt=torch.empty((3, 4, 3), dtype=torch.float64, device="cuda"))
    with pytest.raises(ValueError, match="nqt"):
        e.md_run(pos, vel, acc, 3, 0.01, q_step=0.3, nq=4,
                 nqt=torch.empty((3, 2, 3), dtype=torch.complex128, device="cuda"))
    with pytest.raises(ValueError, match="Z"):
        e.md_run(pos, vel, acc, 3, 0.01, Z=torch.ones(N - 1, dtype=torch.float64, device="cuda"))
    with pytest.raises(ValueError, match="out"):
        e.density_modes(pos, torch.tensor([0.1, 0.2]), out=torch.empty((1, 3), dtype=torch.complex128, device="cuda"))
    with pytest.raises(ValueError, match="rho"):
        e.pm(pos, rho=torch.empty((M, M), dtype=torch.float64, device="cuda"))
    with pytest.raises(ValueError, match="e_x"):
        e.pm(pos, e_x=torch.empty(N - 3, dtype=torch.float64, device="cuda"))
    with pytest.raises(ValueError, match="acc"):
        e.kick_drift(pos, vel, acc[:10], 0.01)
    e.close()
