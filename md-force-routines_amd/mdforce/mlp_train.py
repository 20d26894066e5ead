"""Training data and training loop of the learned pair correction.

Data: for each state (Gamma, kappa) a system of N ions is equilibrated from
the seeded uniform start by NVE steps with a velocity rescale to T = 1/Gamma
every ``rescale_every`` steps (the hard-rescale branch of
thermostat_p3m_ord6.py:15-18), then snapshots are taken every ``spacing``
steps.  The target of a snapshot is the force residual
    F_tight - F_reduced
between a tight FP64 P3M (rc = 8, fine mesh) and the reduced setting
(rc_red, coarse mesh, alpha_red), both FP64 on the GPU, with unit charges
and af = 1.  The network, applied with the reduced setting's cutoff, learns
to predict that residual from the pairs inside rc_red.

Training: Adam on the full per-snapshot gradients (one snapshot per step,
snapshots shuffled each epoch), all passes on the GPU.
"""
from __future__ import annotations

import math
import time

import numpy as np
import torch

from .engine import P3M
from .initialize_pos_vel import initial_split
from .mlp import PairMLP, default_meta


def box_length(n):
    return (4.0 * math.pi * n / 3.0) ** (1.0 / 3.0)


def equilibrate(eng, pos, vel, T, steps, dt=0.1, rescale_every=10, vcap=None, soft_steps=0):
    """NVE steps with a velocity rescale to T every rescale_every steps (in place).

    A uniform random start holds a few pairs at distances ~N^(-1/3); at large
    N their release at dt = 0.1 throws single ions out at tens of thermal
    speeds, and a global rescale does not slow them down.  vcap (in units of
    sqrt(3 T)) clips each ion's speed at every rescale, and the first
    soft_steps steps run at dt / 10 with a rescale and clip after every step,
    so that the release itself is integrated accurately and its energy
    removed.  Both default off (the plain rescale protocol)."""
    def clip():
        if vcap is not None:
            s = vel.norm(dim=1, keepdim=True)
            vel.mul_(torch.clamp(vcap * math.sqrt(3.0 * T) / s.clamp_min(1e-300), max=1.0))

    acc, _ = eng.force(pos)
    acc = acc.clone()
    for _ in range(soft_steps):
        en, _ = eng.md_run(pos, vel, acc, 1, dt / 10.0)
        clip()
        ksum = (vel * vel).sum().reshape(1)
        eng.thermostat(vel, ksum, T, 0)
    done = 0
    while done < steps:
        k = min(rescale_every, steps - done)
        en, _ = eng.md_run(pos, vel, acc, k, dt)
        clip()
        ksum = (vel * vel).sum().reshape(1) if vcap is not None else (en[-1, 1:2] / 1.5).contiguous()
        eng.thermostat(vel, ksum, T, 0)
        done += k
    return acc


def forces(eng, pos):
    acc, _ = eng.force(pos, None, af=1.0)
    return acc.clone()


class Setting:
    """A P3M setting (mesh, alpha, rc) for systems of N ions."""

    def __init__(self, N, mesh, alpha, rc, kappa, precision="fp64", device=None):
        self.N, self.mesh, self.alpha, self.rc, self.kappa = N, mesh, alpha, rc, kappa
        self.L = box_length(N)
        self.eng = P3M(N, mesh, self.L, kappa, alpha, rc, precision=precision, device=device)
        self.eng.build_green()

    def close(self):
        self.eng.close()


def make_snapshots(N, states, n_snap, spacing, eq_steps, red, tight, prop, dt=0.1, seed0=100, device=None,
                   log=None):
    """List of dicts (pos, target, kappa, gamma) on the device.

    red, tight, prop: dicts {kappa: (mesh, alpha, rc)} or callables kappa ->
    (mesh, alpha, rc) giving the reduced, tight and propagation settings.
    """
    get = (lambda s, k: s(k) if callable(s) else s[k])
    L = box_length(N)
    out = []
    for si, (gamma, kappa) in enumerate(states):
        T = 1.0 / gamma
        pos0, vel0 = initial_split(seed0 + 2 * si, seed0 + 2 * si + 1, N, [L, L, L], T)
        pos = torch.from_numpy(pos0).to(device or "cuda")
        vel = torch.from_numpy(vel0).to(device or "cuda")
        Sp = Setting(N, *get(prop, kappa), kappa, precision="mixed", device=device)
        St = Setting(N, *get(tight, kappa), kappa, device=device)
        Sr = Setting(N, *get(red, kappa), kappa, device=device)
        t0 = time.perf_counter()
        acc = equilibrate(Sp.eng, pos, vel, T, eq_steps, dt)
        for k in range(n_snap):
            if k:
                acc = equilibrate(Sp.eng, pos, vel, T, spacing, dt)
            target = forces(St.eng, pos) - forces(Sr.eng, pos)
            out.append(dict(pos=pos.clone(), target=target, kappa=kappa, gamma=gamma,
                            alpha_red=Sr.alpha, rc_red=Sr.rc))
        torch.cuda.synchronize()
        if log:
            log(f"state Gamma={gamma:g} kappa={kappa:g}: {n_snap} snapshots in {time.perf_counter() - t0:.2f} s")
        for S in (Sp, St, Sr):
            S.close()
    return out


def residual_rms(snaps):
    return float(torch.sqrt(torch.stack([(s["target"] ** 2).sum(1).mean() for s in snaps]).mean()))


def train(net, pair_engines, snaps, epochs, lr, lr_end=None, seed=0, log=None):
    """Adam over snapshots.  pair_engines: {kappa: P3M} with the reduced rc (pair lists).
    Returns the per-epoch mean loss."""
    rs = np.random.RandomState(seed)
    hist = []
    nstep = epochs * len(snaps)
    step = 0
    for ep in range(epochs):
        order = rs.permutation(len(snaps))
        tot = torch.zeros(1, dtype=torch.float64, device=net.device)
        for k in order:
            s = snaps[k]
            net.set_meta(f_kappa=s["kappa"] / 3.0, f_alpha=s["alpha_red"])
            _, loss = net.grad(pair_engines[s["kappa"]], s["pos"], s["target"], build=True)
            tot += loss
            frac = step / max(1, nstep - 1)
            cur = lr if lr_end is None else lr * (lr_end / lr) ** frac
            net.adam(cur)
            step += 1
        hist.append(float(tot) / len(snaps))
        if log:
            log(f"epoch {ep}: mean loss {hist[-1]:.4e}")
    return hist


def evaluate(net, eng_red, pos, target, kappa, alpha_red):
    """RMS of the residual before and after the correction (force units)."""
    net.set_meta(f_kappa=kappa / 3.0, f_alpha=alpha_red)
    corr = net.correction(eng_red, pos)
    before = float(torch.sqrt((target ** 2).sum(1).mean()))
    after = float(torch.sqrt(((target - corr) ** 2).sum(1).mean()))
    return before, after


def judge(net, eng_red, eng_plain, pos, kappa, alpha_red):
    """Error estimate of the corrected reduced force: RMS distance to the plain
    P3M force over the mean plain force magnitude (plain P3M is the judge)."""
    net.set_meta(f_kappa=kappa / 3.0, f_alpha=alpha_red)
    fr = forces(eng_red, pos) + net.correction(eng_red, pos)
    fp = forces(eng_plain, pos)
    d = fr - fp
    return float(torch.sqrt((d * d).sum(1).mean()) / torch.sqrt((fp * fp).sum(1)).mean())


def choose_form(net, eng_red, eng_plain, pos, kappa, alpha_red, budget):
    """'learned' if the judged error of the corrected reduced setting is within
    the budget, else 'plain' (the fallback); returns (form, estimate)."""
    est = judge(net, eng_red, eng_plain, pos, kappa, alpha_red)
    return ("learned" if est <= budget else "plain"), est


class LearnedForce:
    """The host driver's switch between the learned form (the reduced setting
    with the network's correction attached) and plain P3M.

    ``decide(pos)`` judges the corrected reduced force against the plain one
    (choose_form).  Within the budget the network is attached to the reduced
    engine and force() / md_run() run there; otherwise (the fallback) the
    network is detached and every call is the plain engine's own call, with
    the plain engine's arguments and buffers, so the caller gets the plain
    path's forces.  Until decide() has run, calls are refused."""

    def __init__(self, net, eng_red, eng_plain, kappa, alpha_red, budget=1e-5):
        self.net, self.eng_red, self.eng_plain = net, eng_red, eng_plain
        self.kappa, self.alpha_red, self.budget = float(kappa), float(alpha_red), float(budget)
        self.form = None
        self.estimate = None

    def decide(self, pos):
        self.net.detach(self.eng_red)  # the judge adds the correction itself
        form, est = choose_form(self.net, self.eng_red, self.eng_plain, pos, self.kappa, self.alpha_red,
                                self.budget)
        if form == "learned":
            self.net.attach(self.eng_red)
        self.form, self.estimate = form, est
        return form, est

    @property
    def engine(self):
        if self.form is None:
            raise RuntimeError("LearnedForce: call decide(pos) before stepping")
        return self.eng_red if self.form == "learned" else self.eng_plain

    def force(self, pos, *args, **kw):
        return self.engine.force(pos, *args, **kw)

    def md_run(self, pos, vel, acc, nsteps, dt, *args, **kw):
        return self.engine.md_run(pos, vel, acc, nsteps, dt, *args, **kw)


def radial_design(net, eng, pos, centers, width, chunk=4_000_000):
    """Design matrix of a radial pair correction f(r) = sum_k c_k b_k(r) with
    Gaussian b_k (centres ``centers``, width ``width``): G[i, k] = sum_j
    b_k(r_ij) u_ij, the force on ion i (caller order) of basis k, over the pairs
    r_ij < eng.rc from the network's pair list (u_ij the unit vector from j to
    i).  Returns (G of shape (N, K, 3) float64, pair count)."""
    d = net.pairs(eng, pos)
    P, perm = d["npairs"], d["perm"]
    K = centers.numel()
    G = torch.zeros((pos.shape[0], K, 3), dtype=torch.float64, device=pos.device)
    for s in range(0, P, chunk):
        i = perm[d["i"][s:s + chunk]]
        r = d["r"][s:s + chunk].double()
        u = d["u"][s:s + chunk].double()
        b = torch.exp(-0.5 * ((r[:, None] - centers[None, :]) / width) ** 2)
        G.index_add_(0, i, b[:, :, None] * u[:, None, :])
    return G, P


def fit_radial(G, residual, ridge=1e-10):
    """Least-squares coefficients c of sum_k c_k G[:, k] ~ residual (N, 3), with
    a ridge of ``ridge`` times the mean diagonal of the normal matrix."""
    A = torch.einsum("nka,nla->kl", G, G)
    b = torch.einsum("nka,na->k", G, residual)
    K = A.shape[0]
    lam = ridge * float(torch.trace(A)) / K
    return torch.linalg.solve(A + lam * torch.eye(K, dtype=A.dtype, device=A.device), b)
