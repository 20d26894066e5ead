"""The a-priori error estimate (mdforce.error_estimate) on the device.

* The device mesh functional (mdf_p3m_pm_error_sum) against the oracle's
  numpy sum at 8^3 and 16^3, to 1e-9 relative (the per-k difference of two
  close sums loses ~6 digits of FP64 at small k).
* The estimate against the measured RMS force error of FP64 P3M on N = 4000
  uniform random unit charges (the estimate's assumption), against the tight
  reference force of the autotuner, in three regimes: mesh-dominated,
  cutoff-dominated and mixed.  The estimate must land within a factor 1.25 of
  the measurement (measured: 0.99, 1.01, 1.00).
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import error_oracle as EO  # noqa: E402


def _dev():
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    return torch.device("cuda", 0)


@pytest.mark.parametrize("M", [8, 16])
def test_pm_functional_matches_oracle(M):
    from mdforce import error_estimate as EE
    from mdforce.engine import P3M
    _dev()
    L, kappa, alpha, N = 12.0, 1.0, 0.8, 100
    e = P3M(N, M, L, kappa, alpha, 3.9)
    got = EE.pm_error(e, N) ** 2 * L ** 6 / N
    ref = EO.pm_error_sum(kappa, alpha, 6, 3, M, L)
    assert got == pytest.approx(ref, rel=1e-9)
    e.close()


@pytest.mark.parametrize("M,alpha,rc,regime", [(16, 0.9, 5.3, "mesh"), (64, 0.8, 3.0, "cutoff"),
                                                 (32, 0.75, 4.0, "mixed")])
def test_estimate_tracks_measured_error(M, alpha, rc, regime):
    from mdforce import autotune, error_estimate as EE
    from mdforce.engine import P3M
    dev = _dev()
    N, kappa = 4000, 1.0
    L = (4.0 * math.pi * N / 3.0) ** (1.0 / 3.0)
    rng = np.random.RandomState(21)
    pos = torch.from_numpy(L * rng.random_sample((N, 3))).to(dev)
    ref = autotune.reference_force(P3M, pos, L, kappa)
    e = P3M(N, M, L, kappa, alpha, rc)
    e.build_green()
    acc, _ = e.force(pos, None, af=1.0)
    d = acc - ref
    measured = float(torch.sqrt(torch.mean(torch.sum(d * d, 1))))
    est = EE.estimate(e, N)
    print(f"{regime}: measured {measured:.3e}, estimate {est['total']:.3e} (pp {est['pp']:.2e}, pm {est['pm']:.2e}), "
          f"ratio {est['total'] / measured:.2f}")
    assert 1 / 1.25 <= est["total"] / measured <= 1.25
    e.close()
