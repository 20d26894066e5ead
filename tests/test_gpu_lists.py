"""Verlet pair lists (mdf_p3m_set_skin) against the list-free engine and the
oracle.  With a skin the PP stage tests candidates only when some ion has
moved more than skin / 2 since the last build; every other step evaluates the
listed pairs.  The forces must be those of the list-free path up to summation
order: relative RMS <= 1e-12 in FP64 (the oracle's bar) and <= 2e-6 in the
FP32-mixed form (two FP32 summation orders)."""
import numpy as np
import pytest

from oracle import p3m_oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def rel_rms(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.sqrt(np.mean(np.sum((a - b) ** 2, -1)) / np.mean(np.sum(b ** 2, -1))))


@pytest.fixture(scope="module")
def P3M():
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    from mdforce.engine import P3M as cls
    return cls


@pytest.fixture(scope="module")
def system():
    # a jittered simple-cubic lattice (no close pairs: the lists live for many
    # steps, as in the liquid)
    n1 = 27
    N = n1 ** 3
    L = O.box_length(N)
    kappa, alpha, rc, M = 1.0, 0.7, 4.5, 32
    P = O.P3MParams(N, L, kappa, alpha, rc, M)
    g = (np.arange(n1) + 0.5) * (L / n1)
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    pos = pos + np.random.RandomState(5).uniform(-0.25, 0.25, pos.shape) * (L / n1)
    _, vel = O.initial_split(5, 6, N, P.Lv, 0.01)
    # a common drift carries every ion across the periodic boundary within a
    # list's lifetime, which the frozen order has to follow
    vel = vel + np.array([0.6, -0.45, 0.3])
    return dict(P=P, pos=pos, vel=vel, N=N, L=L, M=M, kappa=kappa, alpha=alpha, rc=rc)


@pytest.mark.parametrize("precision,skin", [("fp64", 0.5), ("mixed", 0.5), ("fp64", 0.25)])
def test_md_run_with_lists_matches_list_free_forces(P3M, system, precision, skin):
    S = system
    P, N = S["P"], S["N"]
    e = P3M(N, S["M"], S["L"], S["kappa"], S["alpha"], S["rc"], precision=precision)
    e.set_green(dev(P.G_k))
    e.set_skin(skin)
    ref = P3M(N, S["M"], S["L"], S["kappa"], S["alpha"], S["rc"], precision=precision)
    ref.set_green(dev(P.G_k))
    pos, vel = dev(S["pos"]), dev(S["vel"])
    acc, _ = e.force(pos)
    acc = acc.clone()
    # FP32: each path is ~1.1e-6 (RMS, relative) from the FP64 force on this
    # system, and the list path rounds in the frozen order (its pair sums and
    # its charge deposits' dithered rounding differ from the sorted order's),
    # so the two differ by 1.85e-6 - 2.0e-6 from run to run
    tol = 1e-12 if precision == "fp64" else 3e-6
    total_list_steps = 0
    for chunk in range(3):
        en, _ = e.md_run(pos, vel, acc, 7, 0.02)
        torch.cuda.synchronize()
        st = e.list_stats()
        # acc holds the force of the final positions, from the lists unless the
        # last step built; either way it must equal the list-free force
        a_ref, en_ref = ref.force(pos)
        torch.cuda.synchronize()
        err = rel_rms(acc.cpu().numpy(), a_ref.cpu().numpy())
        assert err <= tol, (chunk, err, st)
        U = float(en[-1, 2])
        assert abs(U - float(en_ref[0])) <= tol * abs(float(en_ref[0])), (chunk, U, float(en_ref[0]))
        total_list_steps = st["list_steps"]
    st = e.list_stats()
    assert st["builds"] >= 3 and total_list_steps >= 6, st
    # positions left [0, L) and came back by the drift's wrap; the oracle agrees on the end state
    if precision == "fp64":
        U_o, a_o, _ = P.force(pos.cpu().numpy())
        assert rel_rms(acc.cpu().numpy(), a_o) <= 1e-12
    e.close()
    ref.close()


def test_list_build_on_large_displacement_and_skin_off(P3M, system):
    """A caller that moves ions between calls gets a build at the first step;
    skin 0 turns the lists off and keeps the context usable."""
    S = system
    P, N = S["P"], S["N"]
    e = P3M(N, S["M"], S["L"], S["kappa"], S["alpha"], S["rc"], precision="fp64")
    e.set_green(dev(P.G_k))
    e.set_skin(0.4)
    pos, vel = dev(S["pos"]), dev(S["vel"]) * 0.0
    acc, _ = e.force(pos)
    acc = acc.clone()
    e.md_run(pos, vel, acc, 3, 0.01)
    rng = np.random.RandomState(9)
    moved = np.mod(S["pos"] + rng.uniform(-2.0, 2.0, S["pos"].shape), S["L"])  # in [0, L), as the oracle needs
    pos.copy_(dev(moved))
    U_o, a_o, _ = P.force(pos.cpu().numpy())
    a1, _ = e.force(pos)  # force() always builds
    torch.cuda.synchronize()
    assert rel_rms(a1.cpu().numpy(), a_o) <= 1e-12
    acc = a1.clone()
    e.md_run(pos, vel, acc, 2, 1e-6)  # first step builds; the second uses the lists
    torch.cuda.synchronize()
    U_o, a_o, _ = P.force(pos.cpu().numpy())
    assert rel_rms(acc.cpu().numpy(), a_o) <= 1e-12
    e.set_skin(0.0)
    assert e.config()["pp_waves"] > 0
    a2, _ = e.force(pos)
    torch.cuda.synchronize()
    assert rel_rms(a2.cpu().numpy(), a_o) <= 1e-12
    e.close()


def test_skin_too_large_for_the_box_is_refused(P3M, system):
    from mdforce._lib import MdforceError
    S = system
    e = P3M(S["N"], S["M"], S["L"], S["kappa"], S["alpha"], S["rc"], precision="mixed")
    with pytest.raises(MdforceError, match="L/\\(rc \\+ skin\\) >= 3"):
        e.set_skin(S["L"] / 3.0 - S["rc"] + 0.5)
    e.build_green()  # the context is unchanged and still works
    a, _ = e.force(dev(S["pos"]))
    torch.cuda.synchronize()
    assert torch.isfinite(a).all()
    e.close()


def _dense_blob(S, factor, radius, seed):
    """The lattice system with the ions inside a sphere pulled towards its
    centre by ``factor`` (local density x factor^-3)."""
    pos = S["pos"].copy()
    c = np.full(3, S["L"] / 2)
    d = pos - c
    r = np.linalg.norm(d, axis=1)
    inside = r < radius
    pos[inside] = c + d[inside] * factor
    rng = np.random.RandomState(seed)
    pos[inside] += rng.uniform(-0.05, 0.05, (inside.sum(), 3))
    return np.mod(pos, S["L"]), int(inside.sum())


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_list_overflow_builds_every_step_with_correct_forces(P3M, system, precision):
    """Ions whose neighbour count exceeds the list capacity (1.4x the mean +
    64, rounded up to 64): the build marks itself incomplete, every step builds again and the
    forces stay those of the list-free path."""
    S = system
    P, N = S["P"], S["N"]
    pos0, n_in = _dense_blob(S, 0.6, 9.0, 21)  # density x 4.6 in the blob (radius 5.4)
    e = P3M(N, S["M"], S["L"], S["kappa"], S["alpha"], S["rc"], precision=precision)
    e.set_green(dev(P.G_k))
    e.set_skin(0.5)
    ref = P3M(N, S["M"], S["L"], S["kappa"], S["alpha"], S["rc"], precision=precision)
    ref.set_green(dev(P.G_k))
    pos, vel = dev(pos0), dev(S["vel"]) * 0.1
    acc, _ = e.force(pos)
    acc = acc.clone()
    e.md_run(pos, vel, acc, 4, 0.002)
    torch.cuda.synchronize()
    st = e.list_stats()
    a_ref, _ = ref.force(pos)
    torch.cuda.synchronize()
    tol = 1e-12 if precision == "fp64" else 2e-6
    err = rel_rms(acc.cpu().numpy(), a_ref.cpu().numpy())
    print(f"[lists] {precision}: {n_in} ions in the blob, cap {st['list_cap']}: {st}, err {err:.2e}")
    assert st["list_steps"] == 0 and st["builds"] >= 5, st
    assert err <= tol


def test_list_steps_after_pm_only_and_charge_calls_rebuild(P3M, system):
    """A PM-only call between two md_run calls bins the positions outside a
    step and so replaces the lists' frozen order; md_run builds on entry and
    then steps on its lists, with the list-free forces, under non-unit
    charges (the frozen order carries them from the build)."""
    S = system
    P, N = S["P"], S["N"]
    e = P3M(N, S["M"], S["L"], S["kappa"], S["alpha"], S["rc"], precision="fp64")
    e.set_green(dev(P.G_k))
    e.set_skin(0.4)
    ref = P3M(N, S["M"], S["L"], S["kappa"], S["alpha"], S["rc"], precision="fp64")
    ref.set_green(dev(P.G_k))
    pos, vel = dev(S["pos"]), dev(S["vel"])
    Z = dev(np.random.RandomState(4).uniform(0.5, 1.5, N))
    acc, _ = e.force(pos, Z)
    acc = acc.clone()
    e.md_run(pos, vel, acc, 3, 0.02, Z=Z)
    b0 = e.list_stats()["builds"]
    e.pm(pos, Z)  # bins outside a step
    e.md_run(pos, vel, acc, 3, 0.02, Z=Z)
    torch.cuda.synchronize()
    st = e.list_stats()
    a_ref, _ = ref.force(pos, Z)
    torch.cuda.synchronize()
    assert st["builds"] > b0 and st["list_steps"] >= 2, st
    assert rel_rms(acc.cpu().numpy(), a_ref.cpu().numpy()) <= 1e-12


def test_list_skin_cap(P3M, system):
    """mdf_p3m_set_list_skin caps the lists' skin below the largest the cells
    allow: fewer entries a list, more builds, the same forces."""
    S = system
    P, N = S["P"], S["N"]
    e = P3M(N, S["M"], S["L"], S["kappa"], S["alpha"], S["rc"], precision="fp64")
    e.set_green(dev(P.G_k))
    e.set_skin(0.5)
    s_max = e.set_list_skin(0.0)
    cap_max = e.list_stats()["list_cap"]
    cells = e.list_stats()["cells_x"]
    assert s_max >= 0.5 and abs(s_max - (S["L"] / cells - S["rc"])) < 1e-12
    s = e.set_list_skin(0.2)
    st0 = e.list_stats()
    assert s == 0.2 and st0["list_cap"] < cap_max and st0["cells_x"] == cells
    ref = P3M(N, S["M"], S["L"], S["kappa"], S["alpha"], S["rc"], precision="fp64")
    ref.set_green(dev(P.G_k))
    pos, vel = dev(S["pos"]), dev(S["vel"])
    acc, _ = e.force(pos)
    acc = acc.clone()
    e.md_run(pos, vel, acc, 12, 0.02)
    torch.cuda.synchronize()
    st = e.list_stats()
    a_ref, _ = ref.force(pos)
    torch.cuda.synchronize()
    print(f"[lists] cap 0.2 (largest {s_max:.3f}): {st}")
    assert st["list_steps"] >= 1 and st["builds"] >= 2
    assert rel_rms(acc.cpu().numpy(), a_ref.cpu().numpy()) <= 1e-12
    assert e.set_list_skin(10.0) == s_max  # a cap above the largest skin: the largest
    e.set_skin(0.0)
    assert e.set_list_skin(0.2) == 0.0  # no lists
    with pytest.raises(Exception):
        e.set_list_skin(-1.0)
    e.close()
    ref.close()
