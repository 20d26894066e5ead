"""Autotuner: P3M splitting parameters for a force-error budget.

For one system (N ions in a cubic box of side L at screening kappa, with its
own configuration on the GPU) the autotuner chooses the mesh M, the Ewald
parameter alpha and the PP cutoff rc with the lowest modelled step cost whose
measured RMS force error stays within the budget.  The error of a candidate
is measured, not estimated: the candidate's P3M force on the system's own
positions is compared with a tight FP64 P3M force (reference_settings:
alpha rc = 4.8, alpha h <= 0.2) computed once per system.  The error measure is the one the force
budget is stated in: sqrt(mean_i |F_i - F_i^ref|^2) / mean_i |F_i^ref|.

Cost model (seconds per step), fitted to the N = 1e6, 128^3, rc 7 stage
times of the FP32-mixed engine on MI355X (PP 1.60 ms, assignment + gather
0.37 ms, FFT solve 0.106 ms):
    t = a_pp * N * rc^3 + a_pm * N + a_fft * M^3 * log2(M^3) / 21 + a_mesh * M^3 + t_0
with a_pp = 4.7e-12 (PP at unit density: rc^3 in-cutoff pairs per
particle), a_pm = 3.7e-10 (assignment + gather), a_fft = 5.0e-11 (FFT
solve per mesh node at M = 128, scaled by log2(M^3)/21), a_mesh = 3e-11
(the assignment's and gather's mesh bricks: clearing, staging and adding
per node, which dominate for few ions on a fine mesh; fitted to the 64-replica
ensemble, where moving ten replicas of N = 25,000-50,000 from 64^3 to 128^3
added 0.6 ms of assignment and gather for 1.8e7 nodes), t_0 = 2e-5 (launch
overheads per step of a replica).

The PP cell list needs L/rc >= 3, so rc <= L/3.

tune_launch then picks, by timing NVE steps of the chosen setting on the
system's own state, what the error does not depend on: the Verlet skin (and
with it the PP cell size: the cells of rc + skin), the PP waves per
workgroup and the PM block shapes of the assignment and the gather.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import torch

from ._lib import MDF_EINVAL, MdforceError


class NoCandidateError(RuntimeError):
    """No (mesh, alpha, rc) of the search meets the force budget."""


A_PP = 4.7e-12
A_PM = 3.7e-10
A_FFT = 5.0e-11
A_MESH = 3.0e-11
T_0 = 2e-5


def model_cost(N, M, rc):
    """Modelled seconds per NVE step of one system."""
    M3 = M ** 3
    return A_PP * N * rc ** 3 + A_PM * N + A_FFT * M3 * math.log2(M3) / 21.0 + A_MESH * M3 + T_0


def force_error(acc, ref):
    """RMS force error relative to the mean reference force magnitude."""
    d = acc - ref
    rms = torch.sqrt(torch.mean(torch.sum(d * d, 1)))
    mean_f = torch.mean(torch.sqrt(torch.sum(ref * ref, 1)))
    return float(rms / mean_f)


@dataclass
class Choice:
    mesh: int
    alpha: float
    rc: float
    error: float
    cost: float
    tried: list = field(default_factory=list)


def reference_settings(L):
    """(mesh, alpha, rc) of the tight reference for a cubic box of side L:
    rc = min(15, L/3), alpha = 4.8/rc (PP truncation ~exp(-23)), and the
    smallest power-of-two mesh (32 .. 512) with alpha h <= 0.2.  Against
    rc 20 / alpha 0.24 / 512^3 on equilibrated liquids this differs by 2e-7
    of the mean force at N = 8e6 (kappa 0.5, h = 0.63) and 2e-13 at N = 1e6
    (kappa 2); a spacing-only rule (alpha h = 0.32 at h = 0.63) was off by
    1.2e-5 in the large box."""
    rc = min(15.0, L / 3.0 * (1.0 - 1e-9))
    alpha = 4.8 / rc
    mesh = 32
    while mesh < 512 and alpha * L / mesh > 0.2:
        mesh *= 2
    return mesh, alpha, rc


def reference_force(P3M, pos, L, kappa, device=None):
    """Tight FP64 P3M force of a configuration (unit charges, af = 1)."""
    mesh, alpha, rc = reference_settings(L)
    e = P3M(pos.shape[0], mesh, L, kappa, alpha, rc, device=device)
    e.build_green()
    acc, _ = e.force(pos, None, af=1.0)
    acc = acc.clone()
    e.close()
    return acc


def tune(pos, L, kappa, budget=1e-5, meshes=(32, 64), rcs=None, alphas=None, precision="mixed", margin=0.8,
         ref=None, device=None, fallback_meshes=(128,), prescreen=4.0):
    """Cheapest (mesh, alpha, rc) with measured error <= margin * budget.

    prescreen: a candidate whose a-priori error estimate (mdforce.error_estimate,
    random positions) exceeds prescreen * budget of the reference's mean force
    is recorded with error None and not built or measured; None measures all.

    pos: (N, 3) float64 GPU tensor of the system's positions.  The meshes in
    ``fallback_meshes`` are tried only when none of ``meshes`` meets the
    budget (weak screening in a large box).  Returns a Choice; raises
    NoCandidateError if no candidate meets the budget.  Any other failure
    (a device allocation, a failed force call) propagates as raised.
    """
    try:
        return _tune(pos, L, kappa, budget, meshes, rcs, alphas, precision, margin, ref, device, prescreen)
    except NoCandidateError:
        if not fallback_meshes:
            raise
        return _tune(pos, L, kappa, budget, fallback_meshes, rcs, alphas, precision, margin, ref, device, prescreen)


def _tune(pos, L, kappa, budget, meshes, rcs, alphas, precision, margin, ref, device, prescreen=None):
    from . import error_estimate
    from .engine import P3M

    N = pos.shape[0]
    if ref is None:
        ref = reference_force(P3M, pos, L, kappa, device=device)
    rc_max = L / 3.0 * (1.0 - 1e-9)
    if rcs is None:
        rcs = [r for r in (3.0 + 0.5 * k for k in range(19)) if r <= rc_max]
        if not rcs:
            rcs = [rc_max]
    if alphas is None:
        alphas = (0.25, 0.3, 0.35, 0.42, 0.5, 0.6, 0.72, 0.86, 1.03)
    cands = sorted(((model_cost(N, M, rc), M, rc) for M in meshes for rc in rcs))
    tried = []
    mean_f = None
    for cost, M, rc in cands:
        best = None
        for a in alphas:
            # PP truncation alone is about exp(-(alpha rc)^2): skip hopeless pairs
            if a * rc < 2.6 or a * rc > 5.0:
                continue
            try:
                e = P3M(N, M, L, kappa, a, rc, precision=precision, device=device)
            except MdforceError as err:
                if err.code != MDF_EINVAL:  # only a refused geometry (too many cells for a small rc) is skipped
                    raise
                break
            if prescreen is not None:
                if mean_f is None:
                    mean_f = float(torch.mean(torch.sqrt(torch.sum(ref * ref, 1))))
                if error_estimate.estimate(e, N)["total"] > prescreen * budget * mean_f:
                    e.close()
                    tried.append((M, a, rc, None))
                    continue
            e.build_green()
            acc, _ = e.force(pos, None, af=1.0)
            err = force_error(acc, ref)
            e.close()
            tried.append((M, a, rc, err))
            if best is None or err < best[1]:
                best = (a, err)
        if best is not None and best[1] <= margin * budget:
            return Choice(M, best[0], rc, best[1], cost, tried)
    raise NoCandidateError(f"no (mesh, alpha, rc) meets the force budget {budget} (N={N}, L={L:.3f}, kappa={kappa})")


def skin_candidates(L, rc, cell_counts=3):
    """0 (no pair lists) and, for the cell counts k = floor(L/rc), k - 1, ...,
    the largest skin of k cells per axis: L/k - rc (cells of rc + skin)."""
    Lm = min(L) if isinstance(L, (tuple, list)) else L
    k0 = int(math.floor(Lm / rc))
    out = [0.0]
    for k in range(k0, max(k0 - cell_counts, 2), -1):
        s = Lm / k - rc
        if s > 1e-6:
            out.append(s * (1.0 - 1e-9))
    return out


def time_steps(eng, pos, vel, dt, steps=20, warmup=2):
    """ms per NVE step (md_run) of eng from a copy of (pos, vel)."""
    p, v = pos.clone(), vel.clone()
    acc, _ = eng.force(p)
    acc = acc.clone()
    if warmup:
        eng.md_run(p, v, acc, warmup, dt)
    s0 = torch.cuda.Event(enable_timing=True)
    s1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    s0.record()
    eng.md_run(p, v, acc, steps, dt)
    s1.record()
    torch.cuda.synchronize()
    return s0.elapsed_time(s1) / steps


def tune_launch(eng, pos, vel, dt, steps=20, skins=None, min_gain=0.01, log=None):
    """Launch choices of eng by measurement: the skin (skin_candidates; the
    cell size follows) and the PP waves per workgroup together (every pair:
    the best wave count differs between skins, and one knob at a time could
    settle on a skin whose best waves it never timed), then the assignment
    block shape, then the gather block shape.  A choice replaces the current
    one only when it is faster by more than min_gain.  Choices the context
    refuses (LDS) are skipped.  eng is left with the chosen settings; returns
    a dict of them and every timing."""
    if skins is None:
        skins = skin_candidates(eng.box, eng.rc)
    timings = []

    def measure(what):
        ms = time_steps(eng, pos, vel, dt, steps)
        timings.append((what, ms))
        if log:
            log(f"[tune_launch] {what}: {ms:.4f} ms/step")
        return ms

    cur = {"skin": getattr(eng, "skin", 0.0), "pp_waves": 0, "assign_shape": 0, "gather_shape": 0}
    eng.set_skin(cur["skin"])
    eng.set_launch(0, 0, 0)
    best = measure(dict(cur))

    def try_knob(name, values, apply):
        nonlocal best
        in_use = eng.launch_config().get(name)
        for val in values:
            if val == cur[name] or (cur[name] == 0 and val == in_use):
                continue
            trial = dict(cur, **{name: val})
            try:
                apply(trial)
            except MdforceError as err:
                if err.code != MDF_EINVAL:
                    raise
                timings.append((trial, None))
                continue
            ms = measure(trial)
            if ms < best * (1.0 - min_gain):
                best = ms
                cur.update(trial)
        apply(cur)

    launch = lambda c: eng.set_launch(c["pp_waves"], c["assign_shape"], c["gather_shape"])
    start = dict(cur)
    for sk in [start["skin"]] + [x for x in skins if x != start["skin"]]:
        for w in (4, 8, 12, 16):
            trial = dict(start, skin=sk, pp_waves=w)
            try:
                eng.set_skin(sk)
                launch(trial)
            except MdforceError as err:
                if err.code != MDF_EINVAL:
                    raise
                timings.append((trial, None))
                continue
            ms = measure(trial)
            if ms < best * (1.0 - min_gain):
                best = ms
                cur.update(trial)
    eng.set_skin(cur["skin"])
    launch(cur)
    try_knob("assign_shape", (1, 2, 3, 4), launch)
    try_knob("gather_shape", (1, 2, 8, 9, 3, 4, 5, 6, 7, 10), launch)
    cfg = eng.launch_config()
    return {"skin": cur["skin"], "requested": dict(cur), "in_use": cfg, "ms_per_step": best,
            "timings": [(w, ms) for w, ms in timings]}


def refine_alpha(pos, L, kappa, mesh, rc, ref, precision="mixed", device=None, lo=None, hi=None, iters=8):
    """alpha of least measured error at (mesh, rc): golden-section search on
    [lo, hi] (default alpha rc in [2.4, 4.0]; the error is the sum of a PP
    truncation part falling and a mesh part rising with alpha).  Returns
    (alpha, error, [(alpha, error), ...])."""
    from .engine import P3M
    N = pos.shape[0]
    lo = 2.4 / rc if lo is None else lo
    hi = 4.0 / rc if hi is None else hi
    seen = []

    def err_at(a):
        e = P3M(N, mesh, L, kappa, a, rc, precision=precision, device=device)
        try:
            e.build_green()
            acc, _ = e.force(pos, None, af=1.0)
            er = force_error(acc, ref)
        finally:
            e.close()
        seen.append((a, er))
        return er

    g = (math.sqrt(5.0) - 1.0) / 2.0
    x1, x2 = hi - g * (hi - lo), lo + g * (hi - lo)
    f1, f2 = err_at(x1), err_at(x2)
    for _ in range(iters - 2):
        if f1 <= f2:
            hi, x2, f2 = x2, x1, f1
            x1 = hi - g * (hi - lo)
            f1 = err_at(x1)
        else:
            lo, x1, f1 = x1, x2, f2
            x2 = lo + g * (hi - lo)
            f2 = err_at(x2)
    a, er = min(seen, key=lambda t: t[1])
    return a, er, seen


def tune_measured(pos, vel, L, kappa, dt, budget=1e-5, meshes=(32, 64), margin=0.8, ref=None, device=None,
                  fallback_meshes=(128,), steps=30, log=None):
    """The setting of least measured step time among those whose measured
    error meets margin * budget: tune() gives the cheapest (mesh, rc) of the
    0.5-wide rc grid by the cost model; rc - 0.25, rc and rc + 0.25 on that
    mesh then get alpha by refine_alpha (and, while the smallest of them meets
    the budget, rc - 0.5, rc - 0.75, ... until one does not), and each that
    meets the budget is timed with its launch choices tuned (tune_launch: the
    skin and cell size, PP waves, PM blocks), NVE steps from (pos, vel).
    Returns a Choice whose cost is the measured ms per step, with the
    candidates' timings in tried."""
    from .engine import P3M
    N = pos.shape[0]
    if ref is None:
        ref = reference_force(P3M, pos, L, kappa, device=device)
    ch = tune(pos, L, kappa, budget=budget, meshes=meshes, margin=margin, ref=ref, device=device,
              fallback_meshes=fallback_meshes)
    rc_max = L / 3.0 * (1.0 - 1e-9)
    timed = []
    rcs = [ch.rc + 0.25, ch.rc, ch.rc - 0.25]
    while rcs:
        rc = rcs.pop()
        if rc < 2.0 or rc > rc_max:
            continue
        a, er, _ = refine_alpha(pos, L, kappa, ch.mesh, rc, ref, device=device)
        if er > margin * budget:
            if log:
                log(f"[tune_measured] mesh {ch.mesh} rc {rc}: best alpha {a:.4f} error {er:.3e} over the budget")
            continue
        if rc <= ch.rc - 0.25 and rc - 0.25 >= 2.0:
            rcs.append(rc - 0.25)  # the smallest so far meets the budget: try below it
        e = P3M(N, ch.mesh, L, kappa, a, rc, precision="mixed", device=device)
        try:
            e.build_green()
            tl = tune_launch(e, pos, vel, dt, steps=steps)
        finally:
            e.close()
        timed.append((tl["ms_per_step"], rc, a, er, tl["requested"]))
        if log:
            log(f"[tune_measured] mesh {ch.mesh} rc {rc} alpha {a:.4f}: error {er:.3e}, "
                f"{tl['ms_per_step']:.4f} ms/step with {tl['requested']}")
    if not timed:
        return ch
    ms, rc, a, er, launch = min(timed)
    return Choice(ch.mesh, a, rc, er, ms, tried=[dict(rc=t[1], alpha=t[2], error=t[3], ms_per_step=t[0],
                                                      launch=t[4]) for t in timed])
