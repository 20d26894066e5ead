"""CPU oracle of the P3M error estimate (test infrastructure only).

* ``pm_error_sum``: Q_sum = sum_{k != 0} [sum_m |k_m|^2 G(k_m)^2 - A_k] over
  the shifted k grid and alias range m of gf_opt (yukawa_gf_opt.py:40-96); the
  A_pm part is the pinned oracle gf_opt's (p3m_oracle.gf_opt), the first sum a
  direct loop over the same k_m.
Pinned in tests/test_error_oracle.py: each k's term is >= 0 (the mesh
functional is a minimum of a squared error), Q_sum falls as the mesh refines
at fixed alpha, and m_max = 0 gives exactly zero (one alias: the optimal G
reproduces R(k) and the functional vanishes).
"""
from __future__ import annotations

import numpy as np

from .p3m_oracle import gf_opt


def pm_error_terms(kappa, alpha, p, m_max, M, L):
    h = L / M
    mid = M // 2
    nsh = np.arange(M) - mid
    k1 = 2 * np.pi * nsh / L
    kx, ky, kz = np.meshgrid(k1, k1, k1, indexing="ij")
    ms = np.arange(-m_max, m_max + 1)
    S1 = np.zeros((M, M, M))
    num = np.zeros((M, M, M))
    den = np.zeros((M, M, M))
    k_sq = kx * kx + ky * ky + kz * kz
    for mx in ms:
        for my in ms:
            for mz in ms:
                ax, ay, az = kx + 2 * np.pi * mx * M / L, ky + 2 * np.pi * my * M / L, kz + 2 * np.pi * mz * M / L
                km2 = ax * ax + ay * ay + az * az
                U = (np.sinc(ax * h / (2 * np.pi)) * np.sinc(ay * h / (2 * np.pi)) * np.sinc(az * h / (2 * np.pi))) ** p
                G = 4 * np.pi * np.exp(-0.25 * (kappa ** 2 + km2) / alpha ** 2) / (kappa ** 2 + km2)
                S1 += km2 * G * G
                num += U * U * G * (kx * ax + ky * ay + kz * az)
                den += U * U
    with np.errstate(divide="ignore", invalid="ignore"):
        A = np.where(k_sq > 0, num * num / (den * den * k_sq), 0.0)
    return np.where(k_sq > 0, S1 - A, 0.0), A


def pm_error_sum(kappa, alpha, p, m_max, M, L):
    terms, A = pm_error_terms(kappa, alpha, p, m_max, M, L)
    # the subtracted part is the pinned gf_opt's A_pm
    h = L / M
    A_pm = gf_opt(kappa, alpha, p, m_max, m_max, m_max, M, M, M, h, h, h, L, L, L)[-1]
    assert abs(A.sum() - A_pm) <= 1e-10 * abs(A_pm)
    return float(terms.sum())
