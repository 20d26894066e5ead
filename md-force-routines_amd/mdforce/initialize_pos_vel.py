"""Seeded initial state, same draws as the reference's initialize_pos_vel
(initialize_pos_vel.py:4-35): NumPy's legacy RandomState stream, Box-Muller
Maxwellian velocities (sigma = sqrt(T/3)) with the mean removed, then uniform
positions.  Host-side input generation, as in the reference."""
from __future__ import annotations

import numpy as np


def _velocities(rng, N, T_desired):
    sig = np.sqrt(T_desired / 3)
    u1 = rng.random_sample(N)
    u2 = rng.random_sample(N)
    u3 = rng.random_sample(N)
    u4 = rng.random_sample(N)
    vel = np.empty((N, 3))
    vel[:, 0] = sig * np.sqrt(-2 * np.log(u1)) * np.cos(2 * np.pi * u2)
    vel[:, 1] = sig * np.sqrt(-2 * np.log(u1)) * np.sin(2 * np.pi * u2)
    vel[:, 2] = sig * np.sqrt(-2 * np.log(u3)) * np.cos(2 * np.pi * u4)
    # per-column means of the strided columns (:21-28), not vel.mean(axis=0),
    # whose row-wise accumulation rounds differently from NumPy's pairwise sum
    for a in range(3):
        vel[:, a] = vel[:, a] - np.mean(vel[:, a])
    return vel


def _positions(rng, N, Lv):
    pos = np.empty((N, 3))
    for a in range(3):
        pos[:, a] = Lv[a] * rng.random_sample(N)
    return pos


def initial(seed_int, N, Lv, pos, vel, T_desired):
    """initialize_pos_vel.py:4-35 (one stream: velocities, then positions)."""
    rng = np.random.RandomState(seed_int)
    vel[...] = _velocities(rng, N, T_desired)
    pos[...] = _positions(rng, N, Lv)
    return pos, vel


def initial_split(pos_seed, vel_seed, N, Lv, T_desired):
    """Positions from RandomState(pos_seed), velocities from RandomState(vel_seed)."""
    return _positions(np.random.RandomState(pos_seed), N, Lv), _velocities(np.random.RandomState(vel_seed), N,
                                                                             T_desired)
