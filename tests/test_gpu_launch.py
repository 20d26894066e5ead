"""Launch choices (mdf_p3m_set_launch) and the launch tuner.

The PP waves per workgroup and the PM block shapes change only how the work
is split over workgroups, so every accepted choice gives the forces of the
context's own choice up to summation order: relative RMS <= 1e-12 in FP64,
<= 2e-6 in the FP32-mixed form.  Choices whose LDS does not fit are refused
and leave the context as it was."""
import math

import numpy as np
import pytest

from oracle import p3m_oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def rel_rms(a, b):
    d = (a - b).double()
    return float(torch.sqrt((d * d).sum(1).mean() / (b.double() ** 2).sum(1).mean()))


@pytest.fixture(scope="module")
def system():
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    N = 40 ** 3
    L = O.box_length(N)
    pos, vel = O.initial_split(21, 22, N, [L, L, L], 0.01)
    return dict(N=N, L=L, pos=torch.from_numpy(pos).cuda(), vel=torch.from_numpy(vel).cuda())


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_every_launch_choice_gives_the_same_forces(system, precision):
    from mdforce._lib import MDF_EINVAL, MdforceError
    from mdforce.engine import P3M
    S = system
    e = P3M(S["N"], 32, S["L"], 2.0, 0.7, 5.5, precision=precision)
    e.build_green()
    ref, _ = e.force(S["pos"])
    ref = ref.clone()
    auto = e.launch_config()
    print(f"[launch] {precision} automatic: {auto}")
    tol = 1e-12 if precision == "fp64" else 2e-6
    tried = []
    for waves in (4, 8, 12, 16):
        for ga in ((0, 0), (4, 3), (3, 5), (2, 7), (1, 1), (3, 8), (4, 9), (3, 10)):
            try:
                e.set_launch(waves, *ga)
            except MdforceError as err:
                assert err.code == MDF_EINVAL
                tried.append((waves, ga, None))
                continue
            cfg = e.launch_config()
            assert cfg["pp_waves"] == waves
            if ga[0]:
                assert (cfg["assign_shape"], cfg["gather_shape"]) == ga
            a, _ = e.force(S["pos"])
            torch.cuda.synchronize()
            err = rel_rms(a, ref)
            tried.append((waves, ga, err))
            assert err <= tol, (waves, ga, err)
    print(f"[launch] {precision}: {tried}")
    assert sum(1 for t in tried if t[2] is not None) >= 12
    # refused choices leave the context unchanged
    e.set_launch(0, 0, 0)
    for bad in ((6, 0, 0), (0, 5, 0), (0, 0, 11), (-4, 0, 0)):
        with pytest.raises(MdforceError):
            e.set_launch(*bad)
    assert e.launch_config() == auto
    a, _ = e.force(S["pos"])
    torch.cuda.synchronize()
    assert rel_rms(a, ref) <= tol
    e.close()


def test_launch_choices_survive_a_skin_change_and_lists_agree(system):
    from mdforce.engine import P3M
    S = system
    e = P3M(S["N"], 32, S["L"], 2.0, 0.7, 5.5, precision="fp64")
    e.build_green()
    plain = P3M(S["N"], 32, S["L"], 2.0, 0.7, 5.5, precision="fp64")
    plain.build_green()
    e.set_launch(16, 4, 5)
    e.set_skin(0.4)
    cfg = e.launch_config()
    assert (cfg["pp_waves"], cfg["assign_shape"], cfg["gather_shape"]) == (16, 4, 5)
    # a jittered lattice (no close pairs, so the lists live for several steps)
    n1 = round(S["N"] ** (1 / 3))
    assert n1 ** 3 == S["N"]
    g = (np.arange(n1) + 0.5) * (S["L"] / n1)
    lat = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    lat = lat + np.random.RandomState(8).uniform(-0.2, 0.2, lat.shape) * (S["L"] / n1)
    p = torch.from_numpy(np.mod(lat, S["L"])).cuda()
    v = S["vel"].clone()
    acc, _ = e.force(p)
    acc = acc.clone()
    e.md_run(p, v, acc, 6, 0.05)
    st = e.list_stats()
    a_ref, _ = plain.force(p)
    torch.cuda.synchronize()
    assert st["list_steps"] >= 1, st
    assert rel_rms(acc, a_ref) <= 1e-12
    e.close()
    plain.close()


def test_tune_launch_picks_a_valid_setting(system):
    from mdforce import autotune
    from mdforce.engine import P3M
    S = system
    e = P3M(S["N"], 32, S["L"], 2.0, 0.7, 5.5, precision="mixed")
    e.build_green()
    skins = autotune.skin_candidates(S["L"], 5.5)
    assert skins[0] == 0.0 and all(math.floor(S["L"] / (5.5 + s)) == math.floor(S["L"] / 5.5) - k
                                   for k, s in enumerate(skins[1:]))
    res = autotune.tune_launch(e, S["pos"], S["vel"], 0.05, steps=6)
    print(f"[tune_launch] chosen {res['requested']} in use {res['in_use']}: {res['ms_per_step']:.4f} ms/step; "
          f"{len(res['timings'])} timings")
    measured = [ms for _, ms in res["timings"] if ms is not None]
    assert res["ms_per_step"] == min(measured) or res["ms_per_step"] <= min(measured) / 0.99
    ref = P3M(S["N"], 32, S["L"], 2.0, 0.7, 5.5, precision="mixed")
    ref.build_green()
    a, _ = e.force(S["pos"])
    a = a.clone()
    b, _ = ref.force(S["pos"])
    torch.cuda.synchronize()
    assert rel_rms(a, b) <= 2e-6
    e.close()
    ref.close()


def test_tune_measured_meets_the_budget_and_times_its_candidates():
    """alpha refinement at the tuned rc and its neighbours, each candidate
    timed with its launch choices; the choice meets the budget against an
    independent tight force."""
    from mdforce import autotune
    from mdforce.engine import P3M
    N = 8000
    L = O.box_length(N)
    pos, vel = O.initial_split(31, 32, N, [L, L, L], 0.01)
    p, v = torch.from_numpy(pos).cuda(), torch.from_numpy(vel).cuda()
    ref = autotune.reference_force(P3M, p, L, 2.0)
    ch = autotune.tune_measured(p, v, L, 2.0, 0.05, budget=1e-5, ref=ref, steps=5)
    print(f"[tune_measured] mesh {ch.mesh} alpha {ch.alpha:.4f} rc {ch.rc} error {ch.error:.2e} "
          f"{ch.cost:.4f} ms/step; timed {ch.tried}")
    assert ch.error <= 0.8e-5
    assert all(t["error"] <= 0.8e-5 and t["ms_per_step"] > 0 for t in ch.tried)
    assert ch.cost == min(t["ms_per_step"] for t in ch.tried)
    e = P3M(N, ch.mesh, L, 2.0, ch.alpha, ch.rc, precision="mixed")
    e.build_green()
    acc, _ = e.force(p, None, af=1.0)
    assert autotune.force_error(acc, ref) <= 1e-5
    e.close()
