"""A-priori RMS force error of a P3M setting for N random unit charges.

The two parts of the error add in quadrature, each as
<|dF|^2> = (N / V) * (integral of the squared pair-force error over space):
* PP, the pairs dropped beyond rc: dF_pp^2 = (N / V) 4 pi int_rc^inf f(r)^2 r^2 dr
  with f = g2(r) / r^2 the short-range force magnitude (yukawa_pp.py:121-124;
  g2 as the PP kernel tabulates it), integrated numerically;
* PM, the mesh force: the Hockney-Eastwood error functional of the optimal
  influence function, dF_pm^2 = N Q_sum / V^2, with
  Q_sum = sum_{k != 0} [sum_m |k_m|^2 G(k_m)^2 - A_k] summed on the device
  over the same alias ranges as gf_opt (yukawa_gf_opt.py:51-96; A_k is the k
  term of its A_pm, which the reference computes and leaves unused).
Units are those of the engine's accelerations at af = 1 (acc_pp + acc_pm).
Random positions are the estimate's assumption; liquids with short-range
order have a smaller mean force, so a budget relative to their mean force
is tighter than this absolute estimate suggests.
"""
from __future__ import annotations

import ctypes as ct
import math

import torch

from ._lib import check


def _g2(r, kappa, alpha):
    from scipy.special import erfcx
    b = kappa / (2.0 * alpha)
    # e^{kr} erfc(ar + b) = erfcx(ar + b) e^{kr - (ar + b)^2}; the same with -k
    x1, x2 = alpha * r + b, alpha * r - b
    t1 = erfcx(x1) * math.exp(kappa * r - x1 * x1) * (1.0 - kappa * r)
    t2 = erfcx(x2) * math.exp(-kappa * r - x2 * x2) * (1.0 + kappa * r) if x2 > -5.0 else \
        math.erfc(x2) * math.exp(-kappa * r) * (1.0 + kappa * r)
    return 0.5 * (t1 + t2) + 2.0 * alpha / math.sqrt(math.pi) * r * math.exp(-(alpha * r) ** 2 - b * b)


def pp_error(N, V, kappa, alpha, rc):
    """RMS force error of the PP cutoff (absolute, af = 1)."""
    from scipy.integrate import quad
    upper = rc + 40.0 / alpha
    val, _ = quad(lambda r: (_g2(r, kappa, alpha) / r) ** 2, rc, upper, limit=400, epsabs=0.0, epsrel=1e-10)
    return math.sqrt(N / V * 4.0 * math.pi * val)


def pm_error(engine, N, p=6, m_max=3):
    """RMS mesh force error of the engine's mesh, order p and alias range (absolute, af = 1)."""
    q = torch.zeros(1, dtype=torch.float64, device=engine.device)
    check(engine._lib.mdf_p3m_pm_error_sum(engine._ctx, int(p), int(m_max), int(m_max), int(m_max),
                                           ct.c_void_p(q.data_ptr()), engine._stream()), "mdf_p3m_pm_error_sum")
    V = engine.box[0] * engine.box[1] * engine.box[2]
    return math.sqrt(max(float(q.item()), 0.0) * N) / V


def estimate(engine, N, p=6, m_max=3):
    """{'pp', 'pm', 'total'}: the engine's RMS force error estimate for N random unit charges."""
    V = engine.box[0] * engine.box[1] * engine.box[2]
    pp = pp_error(N, V, engine.kappa, engine.alpha, engine.rc)
    pm = pm_error(engine, N, p, m_max)
    return {"pp": pp, "pm": pm, "total": math.hypot(pp, pm)}
