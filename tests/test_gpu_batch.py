"""Replica ensembles (graph-captured batch) and the autotuner.

The batch runs the same kernels per replica as a single context's
mdf_md_run, in the same order, so in FP64 a replica stepped inside the
graph must match the same replica stepped alone up to the run-to-run
rounding of the charge assignment (FP64 atomics add in arrival order):
positions, velocities, accelerations, per-step energies and n(q, t) agree to
1e-11 relative after 10 steps.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _system(n, seed, gamma):
    from mdforce.initialize_pos_vel import initial_split
    L = (4.0 * math.pi * n / 3.0) ** (1.0 / 3.0)
    pos, vel = initial_split(1000 + seed, 2000 + seed, n, [L, L, L], 1.0 / gamma)
    return L, pos, vel


@pytest.fixture(scope="module")
def mods():
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    from mdforce import batch, engine
    return batch, engine


@pytest.mark.parametrize("mode,skins", [(0, None), (1, None), (0, (0.3, 0.0, 0.5)), (1, (0.3, 0.0, 0.5))],
                         ids=["segmented", "streams", "segmented-lists", "streams-lists"])
def test_batch_graph_matches_single_replicas(mods, mode, skins):
    """With skins, replicas 0 and 2 step on pair lists (in the segmented
    graph one list-aware PP launch, replica 1 in the other), as do their
    single contexts."""
    batch, engine = mods
    specs = [(2100, 32, 2.0, 0.8, 4.0, 100.0), (5000, 32, 0.5, 0.6, 5.0, 10.0), (3300, 64, 3.0, 0.9, 3.9, 175.0)]
    systems, states = [], []
    for r, (n, M, kappa, alpha, rc, gamma) in enumerate(specs):
        L, pos, vel = _system(n, r, gamma)
        systems.append(dict(n=n, L=L, kappa=kappa, mesh=M, alpha=alpha, rc=rc, precision="fp64"))
        states.append((pos, vel))
    nsteps, K, nq, dt = 10, 4, 20, 0.1
    # singles
    singles = []
    for s, (pos, vel) in zip(systems, states):
        e = engine.P3M(s["n"], s["mesh"], s["L"], s["kappa"], s["alpha"], s["rc"])
        e.build_green()
        if skins:
            e.set_skin(skins[len(singles)])
        p = torch.from_numpy(pos).cuda()
        v = torch.from_numpy(vel).cuda()
        a, _ = e.force(p)
        a = a.clone()
        en, nqt = e.md_run(p, v, a, nsteps, dt, q_step=2 * math.pi / s["L"], nq=nq)
        torch.cuda.synchronize()
        singles.append((p.cpu().numpy(), v.cpu().numpy(), a.cpu().numpy(), en.cpu().numpy(), nqt.cpu().numpy(),
                        e.list_stats()))
        e.close()
    # batch (2 graph launches of 4 steps + 2 eager steps)
    b = batch.ReplicaBatch(systems, n_streams=2)
    b.set_mode(mode)
    if skins:
        b.set_skins(skins)
    P, V, A = [], [], []
    for rep, (pos, vel) in zip(b.replicas, states):
        p = torch.from_numpy(pos).cuda()
        v = torch.from_numpy(vel).cuda()
        a, _ = rep.force(p)
        P.append(p)
        V.append(v)
        A.append(a.clone())
    ens, nqts = b.run(P, V, A, nsteps, dt, nq=nq, graph_steps=K)
    torch.cuda.synchronize()
    assert b.graph_mode == mode
    for r in range(len(systems)):
        p, v, a, en, nqt, st1 = singles[r]
        st = b.replicas[r].list_stats()
        print(f"[batch] mode {mode} replica {r}: max|dx| {np.abs(P[r].cpu().numpy() - p).max():.3e} "
              f"lists {st} alone {st1}")
        if skins and skins[r] > 0:
            # the same builds and list steps as alone (the force call builds too)
            assert st["list_steps"] >= 1 and (st["builds"], st["list_steps"]) == (st1["builds"], st1["list_steps"])
        else:
            assert st["list_steps"] == 0 and st["list_cap"] == 0
        L = systems[r]["L"]
        dpos = np.abs(P[r].cpu().numpy() - p)
        assert np.minimum(dpos, L - dpos).max() <= 1e-11 * L
        np.testing.assert_allclose(V[r].cpu().numpy(), v, rtol=0, atol=1e-11 * np.abs(v).max())
        np.testing.assert_allclose(A[r].cpu().numpy(), a, rtol=0, atol=1e-11 * np.abs(a).max())
        np.testing.assert_allclose(ens[r].cpu().numpy(), en, rtol=1e-11, atol=1e-11 * np.abs(en).max())
        np.testing.assert_allclose(nqts[r].cpu().numpy(), nqt, rtol=0, atol=1e-11 * systems[r]["n"])
    # a second run replays the captured graph (same arguments) and continues
    ens2, _ = b.run(P, V, A, 8, dt, nq=nq, graph_steps=K)
    torch.cuda.synchronize()
    assert np.isfinite(ens2[0].cpu().numpy()).all()
    # a replica rebuilt in place (new skin: new buffers) makes the next run
    # capture again; its energies continue those of the run before
    b.replicas[1].set_skin(0.4)
    b.replicas[1].set_list_skin(0.4)
    ens3, _ = b.run(P, V, A, 8, dt, nq=nq, graph_steps=K)
    torch.cuda.synchronize()
    e2, e3 = ens2[1].cpu().numpy(), ens3[1].cpu().numpy()
    st = b.replicas[1].list_stats()  # (a hot replica: it may build every step)
    assert np.isfinite(e3).all() and st["list_cap"] > 0 and st["builds"] + st["list_steps"] >= 8
    # the total energy carries on: its first change no larger than the run before's largest step-to-step change
    assert abs(e3[0, 0] - e2[-1, 0]) <= 3.0 * np.abs(np.diff(e2[:, 0])).max() + 1e-12 * abs(e2[-1, 0])
    b.close()


def test_autotuner_meets_budget(mods):
    from mdforce import autotune
    batch, engine = mods
    n, kappa = 4000, 2.0
    L, pos, _ = _system(n, 7, 100.0)
    p = torch.from_numpy(pos).cuda()
    ref = autotune.reference_force(engine.P3M, p, L, kappa)
    ch = autotune.tune(p, L, kappa, budget=1e-5, ref=ref)
    print(f"[autotune] N={n} L={L:.2f}: mesh {ch.mesh} alpha {ch.alpha} rc {ch.rc} err {ch.error:.2e} "
          f"({len(ch.tried)} candidates)")
    assert ch.error <= 1e-5
    # independent check of the chosen setting
    e = engine.P3M(n, ch.mesh, L, kappa, ch.alpha, ch.rc, precision="mixed")
    e.build_green()
    acc, _ = e.force(p, None, af=1.0)
    assert autotune.force_error(acc, ref) <= 1e-5
    # the tight reference itself: a still tighter setting agrees with it
    e2 = engine.P3M(n, 128, L, kappa, 0.7, min(9.0, L / 3 * 0.999))
    e2.build_green()
    acc2, _ = e2.force(p, None, af=1.0)
    assert autotune.force_error(acc2, ref) <= 1e-7


def test_batch_without_modes_and_mixed_precisions_fall_back(mods):
    """nq = 0 drops the n(q) stages of the segmented graph; a batch whose
    replicas differ in precision cannot be segmented and runs as per-replica
    streams.  Both still match the replicas stepped alone."""
    batch, engine = mods
    specs = [(2500, 32, 2.0, 0.8, 4.0, 100.0, "fp64"), (3100, 32, 1.0, 0.7, 4.5, 50.0, "mixed")]
    for uniform in (True, False):
        systems, states = [], []
        for r, (n, M, kappa, alpha, rc, gamma, prec) in enumerate(specs):
            L, pos, vel = _system(n, 10 + r, gamma)
            systems.append(dict(n=n, L=L, kappa=kappa, mesh=M, alpha=alpha, rc=rc,
                                precision="fp64" if uniform else prec))
            states.append((pos, vel))
        singles = []
        for s, (pos, vel) in zip(systems, states):
            e = engine.P3M(s["n"], s["mesh"], s["L"], s["kappa"], s["alpha"], s["rc"], precision=s["precision"])
            e.build_green()
            p = torch.from_numpy(pos).cuda()
            v = torch.from_numpy(vel).cuda()
            a, _ = e.force(p)
            a = a.clone()
            en, _ = e.md_run(p, v, a, 6, 0.1)
            torch.cuda.synchronize()
            singles.append((p.cpu().numpy(), en.cpu().numpy()))
            e.close()
        b = batch.ReplicaBatch(systems, n_streams=2)
        P, V, A = [], [], []
        for rep, (pos, vel) in zip(b.replicas, states):
            p = torch.from_numpy(pos).cuda()
            v = torch.from_numpy(vel).cuda()
            a, _ = rep.force(p)
            P.append(p)
            V.append(v)
            A.append(a.clone())
        ens, nqts = b.run(P, V, A, 6, 0.1, nq=0, graph_steps=3)
        torch.cuda.synchronize()
        print(f"[batch] uniform={uniform}: graph mode {b.graph_mode}")
        assert nqts is None
        assert b.graph_mode == (0 if uniform else 1)
        for r, s in enumerate(systems):
            p, en = singles[r]
            tol = 1e-11 if s["precision"] == "fp64" else 1e-6
            d = np.abs(P[r].cpu().numpy() - p)
            assert np.minimum(d, s["L"] - d).max() <= tol * s["L"]
            np.testing.assert_allclose(ens[r].cpu().numpy(), en, rtol=tol, atol=tol * np.abs(en).max())
        b.close()
