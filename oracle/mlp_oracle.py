"""CPU oracle of the learned pair correction (test infrastructure only).

Only tests/, the smoke check and bench.py's CPU leg import this module; the
product path (md-force-routines_amd/csrc/mlp.hip) never does.  It restates
the network of the HIP kernels in NumPy so the GPU forward pass, correction
force and BF16-MFMA backward pass can be checked on the same pair lists:

  features  f_k = exp(-((r - mu0 - k dmu) inv_sigma)^2), k = 0..11;
            f_12 = (rho_i + rho_j) / 2; f_13..f_15 = f_kappa, f_alpha, f_rc
  forward   z1 = x W1 + b1, h1 = silu(z1), z2 = h1 W2 + b2, h2 = silu(z2),
            z3 = h2 W3 + b3, h3 = silu(z3), g = out_scale (h3 . w4 + b4)
  force     F_i = sum over pairs p of i of g_p u_p
  loss      L = (1/n) sum_i |F_i - t_i|^2

With ``bf16=True`` the inputs, weights, activations and backward deltas are
rounded to BF16 (round to nearest even) at the points where the kernels round
them; the sums are then exact (FP64) while the kernels accumulate in FP32,
which bounds the expected difference.  With ``bf16=False`` everything is
FP64: that form is pinned by finite differences (tests/test_mlp_oracle.py).
The network has no counterpart in the reference, so parity here is with the
restatement of the build's own definition, pinned by a finite-difference
gradient check and BF16 known answers.
"""
from __future__ import annotations

import numpy as np

N_IN, N_H = 16, 64
OFF = {}
_o = 0
for _name, _size in (("W1", N_IN * N_H), ("b1", N_H), ("W2", N_H * N_H), ("b2", N_H), ("W3", N_H * N_H),
                     ("b3", N_H), ("w4", N_H), ("b4", 1)):
    OFF[_name] = (_o, _size)
    _o += _size
N_PARAM = _o  # 9473


def bf16(x):
    """Round float32 values to the nearest BF16 (ties to even), returned as float32."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    return np.where(np.isfinite(a), out, a).astype(np.float32)


def unpack(p):
    p = np.asarray(p)
    d = {}
    for k, (o, s) in OFF.items():
        v = p[o:o + s]
        if k == "W1":
            v = v.reshape(N_IN, N_H)
        elif k in ("W2", "W3"):
            v = v.reshape(N_H, N_H)
        d[k] = v
    return d


def pack(d):
    out = np.zeros(N_PARAM, dtype=np.float64)
    for k, (o, s) in OFF.items():
        out[o:o + s] = np.asarray(d[k]).reshape(-1)
    return out


def init_params(seed=0, scale_out=0.1):
    rs = np.random.RandomState(seed)
    d = {
        "W1": rs.randn(N_IN, N_H) * np.sqrt(2.0 / N_IN),
        "b1": np.zeros(N_H),
        "W2": rs.randn(N_H, N_H) * np.sqrt(2.0 / N_H),
        "b2": np.zeros(N_H),
        "W3": rs.randn(N_H, N_H) * np.sqrt(2.0 / N_H),
        "b3": np.zeros(N_H),
        "w4": rs.randn(N_H) * scale_out / np.sqrt(N_H),
        "b4": np.zeros(1),
    }
    return pack(d)


def features(r, rho_pair, meta):
    """(P, 16) FP64 features of pairs with distances r and pair densities."""
    r = np.asarray(r, dtype=np.float64)
    k = np.arange(12)
    mu = meta["mu0"] + k * meta["dmu"]
    f = np.empty((r.shape[0], N_IN))
    f[:, :12] = np.exp(-((r[:, None] - mu[None, :]) * meta["inv_sigma"]) ** 2)
    f[:, 12] = rho_pair
    f[:, 13] = meta["f_kappa"]
    f[:, 14] = meta["f_alpha"]
    f[:, 15] = meta["f_rc"]
    return f


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, x, meta, bf16_mode=True):
    """Pair outputs g (P,) and the cache of the backward pass."""
    d = unpack(np.asarray(p, dtype=np.float64))
    R = (lambda a: bf16(a).astype(np.float64)) if bf16_mode else (lambda a: np.asarray(a, dtype=np.float64))
    f32 = (lambda a: np.asarray(a, dtype=np.float32).astype(np.float64)) if bf16_mode else (lambda a: a)
    W1, W2, W3, w4 = R(d["W1"]), R(d["W2"]), R(d["W3"]), R(d["w4"])
    b1, b2, b3, b4 = f32(d["b1"]), f32(d["b2"]), f32(d["b3"]), f32(d["b4"])
    xb = R(x)
    z1 = xb @ W1 + b1
    h1 = R(_silu(z1))
    z2 = h1 @ W2 + b2
    h2 = R(_silu(z2))
    z3 = h2 @ W3 + b3
    h3 = R(_silu(z3))
    out = h3 @ w4 + b4[0]
    g = meta["out_scale"] * out
    cache = dict(xb=xb, z1=z1, h1=h1, z2=z2, h2=h2, z3=z3, h3=h3, W1=W1, W2=W2, W3=W3, w4=w4, R=R)
    return g, cache


def correction(g, pi, u, n):
    """F (n,3): per particle sum of g_p u_p over its pairs (pi: particle of each pair)."""
    F = np.zeros((n, 3))
    np.add.at(F, pi, g[:, None] * u)
    return F


def loss_and_grad(p, x, pi, u, target, meta, bf16_mode=True):
    """Loss (1/n) sum_i |F_i - t_i|^2 and its gradient w.r.t. the packed parameters."""
    n = target.shape[0]
    g, c = forward(p, x, meta, bf16_mode)
    F = correction(g, pi, u, n)
    err = F - target
    if bf16_mode:
        err = err.astype(np.float32).astype(np.float64)
    L = float(np.sum(err * err) / n)
    R = c["R"]
    dg = (2.0 / n) * np.sum(err[pi] * u, axis=1)        # dL/dg per pair
    dout = dg * meta["out_scale"]
    grads = {"b4": np.array([dout.sum()]), "w4": c["h3"].T @ dout}
    dz3 = R(dout[:, None] * c["w4"][None, :] * _dsilu(c["z3"]))
    grads["W3"] = c["h2"].T @ dz3
    grads["b3"] = dz3.sum(0)
    dz2 = R((dz3 @ c["W3"].T) * _dsilu(c["z2"]))
    grads["W2"] = c["h1"].T @ dz2
    grads["b2"] = dz2.sum(0)
    dz1 = R((dz2 @ c["W2"].T) * _dsilu(c["z1"]))
    grads["W1"] = c["xb"].T @ dz1
    grads["b1"] = dz1.sum(0)
    return L, pack(grads), F


def pair_density(rho, pi, pj):
    return 0.5 * (np.asarray(rho)[pi] + np.asarray(rho)[pj])
