"""Writes golden PP accelerations from the reference's own C
particle_particle (yukawa_p3m.c:182-416), compiled from the reference tree
into oracle/_ref by ``make -C oracle ref``.

    python tests/golden/make_pp_golden.py

Each fixture holds the inputs (positions, charges, L, kappa, alpha, rc) and
the reference's output accelerations (the C routine returns no energy).
Positions are seeded uniform draws in [0, L) plus a few edge placements
(x = 0, x = L (1 - 1e-12), a close pair).  A coordinate within an ulp of L
is left out: the reference's cell index floor(x / (L/Lxd)) can round to Lxd
there and index past its head array (undefined behaviour, yukawa_p3m.c:267-284).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import c_oracle  # noqa: E402
from oracle.p3m_oracle import box_length  # noqa: E402

CASES = [
    # name, N, kappa, alpha, rc, seed, charges
    ("pp_ref_n1000_k2", 1000, 2.0, 0.85, 5.0, 11, "ones"),
    ("pp_ref_n2000_k05", 2000, 0.5, 0.6, 6.0, 12, "ones"),
    ("pp_ref_n500_k0", 500, 0.0, 1.0, 4.0, 13, "ones"),
    ("pp_ref_n800_k1_mixedZ", 800, 1.0, 0.7, 4.5, 14, "mixed"),
]


def make():
    c_oracle.build(ref=True)
    for name, N, kappa, alpha, rc, seed, zkind in CASES:
        L = box_length(N)
        rng = np.random.RandomState(seed)
        pos = L * rng.random_sample((N, 3))
        pos[0] = [0.0, 0.0, 0.0]
        pos[1] = [L * (1 - 1e-12), 0.5 * L, L * (1 - 1e-12)]
        pos[2] = pos[3] + np.array([0.05, -0.02, 0.01])
        pos[2] = np.mod(pos[2], L)
        Z = np.ones(N) if zkind == "ones" else rng.choice([1.0, 2.0, 3.0], N)
        acc = c_oracle.ref_particle_particle(pos, Z, L, kappa, alpha, rc)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), pos=pos, Z=Z, L=L, kappa=kappa, alpha=alpha, rc=rc,
                            acc=acc)
        print(name, N, float(np.abs(acc).max()))


if __name__ == "__main__":
    make()
