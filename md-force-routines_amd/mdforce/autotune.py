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
    t = a_pp * N * rc^3 + a_pm * N + a_fft * M^3 * log2(M^3) / 21 + t_0
with a_pp = 4.7e-12 (PP at unit density: rc^3 in-cutoff pairs per
particle), a_pm = 3.7e-10 (assignment + gather), a_fft = 5.0e-11 (FFT
solve per mesh node at M = 128, scaled by log2(M^3)/21), t_0 = 2e-5 (launch
overheads per step of a replica).

The PP cell list needs L/rc >= 3, so rc <= L/3.
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
T_0 = 2e-5


def model_cost(N, M, rc):
    """Modelled seconds per NVE step of one system."""
    M3 = M ** 3
    return A_PP * N * rc ** 3 + A_PM * N + A_FFT * M3 * math.log2(M3) / 21.0 + T_0


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
