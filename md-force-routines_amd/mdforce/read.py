"""Drop-in for the reference's read (read.py:3-74): resume from a text file.

``initL(pos, vel, f_input)`` reads columns x y z vx vy vz (ax ay az) of a
file in the ``pos_vel_acc_final.out`` / ``init.out`` layout (read.py:44),
keeps the first N = len(pos) rows and returns (pos, vel) as new (N, 3)
arrays; ``init0`` reads the older layout with the velocities first
(vx vy vz x y z ..., read.py:7).  The values round-trip the writer's %.18e
text exactly.  Host-side file parsing (numpy.loadtxt / genfromtxt, as the
reference); the arrays then go to the device with the engine's calls.
"""
from __future__ import annotations

import numpy as np


def _cols(data, N, idx):
    return np.stack([data[:N, i] for i in idx], axis=1)


def initL(pos, vel, f_input):
    N = len(pos[:, 0])
    data = np.atleast_2d(np.loadtxt(f_input, usecols=(0, 1, 2, 3, 4, 5, 6, 7, 8)))
    return _cols(data, N, (0, 1, 2)), _cols(data, N, (3, 4, 5))


def init0(pos, vel, f_input):
    N = len(pos[:, 0])
    data = np.atleast_2d(np.genfromtxt(f_input, usecols=(0, 1, 2, 3, 4, 5, 6, 7, 8)))
    return _cols(data, N, (3, 4, 5)), _cols(data, N, (0, 1, 2))
