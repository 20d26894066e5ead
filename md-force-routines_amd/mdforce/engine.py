"""Device-side P3M engine: one libmdforce context plus torch-ROCm buffers.

Everything here runs on the GPU through the C ABI; torch provides device
memory and the current HIP stream only.  All tensors passed in must be
contiguous float64 CUDA (ROCm) tensors on the engine's device.
"""
from __future__ import annotations

import ctypes as ct
import math

import torch

from ._lib import PREC_FP64, PREC_MIXED, P3MParams, check, lib

_PREC = {"fp64": PREC_FP64, "plain": PREC_FP64, "mixed": PREC_MIXED}


def _ptr(t):
    """Device pointer of a tensor, or None."""
    if t is None:
        return None
    return ct.c_void_p(t.data_ptr())


def _need(t, name, shape=None, device=None, dtype=torch.float64):
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch tensor")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_cuda:
        raise ValueError(f"{name} must be a GPU tensor")
    if device is not None and t.device != device:
        raise ValueError(f"{name} is on {t.device}, engine on {device}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {tuple(shape)}")
    return t


class P3M:
    """A P3M force context for a fixed box, mesh, kappa, alpha and rc.

    Mirrors the parameters the reference threads through force_pot
    (yukawa_p3m_ord6.py:6): N (as n_max), Lv, (Mx, My, Mz), kappa, G (alpha),
    rc.  The Green function is set from the reference's gf_opt array
    (``set_green``) or computed on the device (``build_green``).

    precision: "fp64" (also "plain"; the reference's arithmetic) or "mixed"
    (FP32 pair math, assignment weights and FFT passes; FP64 charge-mesh
    accumulation, energies, reductions and positions).
    """

    def __init__(self, n_max, mesh, box, kappa, alpha, rc, precision="fp64", device=None):
        if not torch.cuda.is_available():
            raise RuntimeError("mdforce needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        if isinstance(mesh, int):
            mesh = (mesh, mesh, mesh)
        if isinstance(box, (int, float)):
            box = (box, box, box)
        self.n_max = int(n_max)
        self.mesh = tuple(int(m) for m in mesh)
        self.box = tuple(float(b) for b in box)
        self.kappa, self.alpha, self.rc = float(kappa), float(alpha), float(rc)
        self.precision = precision
        p = P3MParams()
        p.n_max = self.n_max
        for d in range(3):
            p.mesh[d] = self.mesh[d]
            p.box[d] = self.box[d]
        p.kappa, p.alpha, p.rc = self.kappa, self.alpha, self.rc
        p.order = 6
        p.precision = _PREC[precision]
        self._lib = lib()
        self._ctx = ct.c_void_p()
        with torch.cuda.device(self.device):
            check(self._lib.mdf_p3m_create(ct.byref(p), ct.byref(self._ctx)), "mdf_p3m_create")
        self.energies = torch.zeros(4, dtype=torch.float64, device=self.device)
        self._scratch = torch.zeros(4, dtype=torch.float64, device=self.device)
        self.has_green = False

    # ------------------------------------------------------------ plumbing --
    def _stream(self):
        return ct.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def close(self):
        if getattr(self, "_ctx", None) and self._ctx.value:
            self._lib.mdf_p3m_destroy(self._ctx)
            self._ctx = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def volume(self):
        return self.box[0] * self.box[1] * self.box[2]

    def stats(self):
        out = (ct.c_int * 2)()
        check(self._lib.mdf_p3m_stats(self._ctx, out), "mdf_p3m_stats")
        return {"pp_multipass_tiles": out[0], "mesh_brick_outside": out[1]}

    def config(self):
        """PP launch configuration chosen at create: waves per workgroup, LDS
        tile capacity (particles), pair-table pieces, LDS bytes."""
        out = (ct.c_int * 4)()
        check(self._lib.mdf_p3m_config(self._ctx, out), "mdf_p3m_config")
        return {"pp_waves": out[0], "tile_max": out[1], "npiece": out[2], "pp_lds_bytes": out[3]}

    def set_skin(self, skin):
        """Verlet pair lists of skin ``skin`` for the PP stage (0 = off): the
        cells become L / floor(L / (rc + skin)) wide and candidate tests run
        only when some particle has moved more than skin / 2 since the last
        build.  Forces are unchanged up to summation order."""
        check(self._lib.mdf_p3m_set_skin(self._ctx, float(skin)), "mdf_p3m_set_skin")
        self.skin = float(skin)

    def set_list_skin(self, list_skin):
        """Caps the lists' skin at ``list_skin`` (0: the largest the cells
        allow, the default); the cells stay.  Returns the list skin in use
        (0 without lists)."""
        eff = ct.c_double(0.0)
        check(self._lib.mdf_p3m_set_list_skin(self._ctx, float(list_skin), ct.byref(eff)), "mdf_p3m_set_list_skin")
        return eff.value

    def set_launch(self, pp_waves=0, assign_shape=0, gather_shape=0):
        """Launch choices (0 = the context's own): PP waves per workgroup (4, 8,
        12, 16) and the PM block shapes (1 = 8^3 .. 4 = 2^3 sub-cells, gather
        also 5 = 1x2x2, 6 = 1x1x2, 7 = 1x1x1).  Forces are unchanged up to
        summation order."""
        check(self._lib.mdf_p3m_set_launch(self._ctx, int(pp_waves), int(assign_shape), int(gather_shape)),
              "mdf_p3m_set_launch")

    def launch_config(self):
        out = (ct.c_int * 6)()
        check(self._lib.mdf_p3m_launch_config(self._ctx, out), "mdf_p3m_launch_config")
        return {"pp_waves": out[0], "assign_shape": out[1], "gather_shape": out[2], "assign_blocks": out[3],
                "gather_blocks": out[4], "tile_max": out[5]}

    def list_stats(self):
        out = (ct.c_int * 4)()
        check(self._lib.mdf_p3m_list_stats(self._ctx, out), "mdf_p3m_list_stats")
        return {"builds": out[0], "list_steps": out[1], "list_cap": out[2], "cells_x": out[3]}

    STAGES = ("bin", "pp", "assign", "fft", "gather", "drift", "modes", "finalize")

    def profile_start(self, events=65536, pp_only=False):
        """Stage timing by HIP events; pp_only: two events a step around the PP
        stage (only the pp figure is then a stage time)."""
        check(self._lib.mdf_p3m_profile(self._ctx, 3 if pp_only else 1, None, int(events)), "mdf_p3m_profile")

    def profile_read(self):
        """Per-stage (total ms, intervals) since the last read; synchronises."""
        out = (ct.c_double * 16)()
        check(self._lib.mdf_p3m_profile(self._ctx, 2, ct.cast(out, ct.c_void_p), 16), "mdf_p3m_profile")
        return {name: (out[i], int(out[8 + i])) for i, name in enumerate(self.STAGES)}

    def profile_stop(self):
        check(self._lib.mdf_p3m_profile(self._ctx, 0, None, 0), "mdf_p3m_profile")

    # --------------------------------------------------------------- Green --
    def set_green(self, G_k):
        """Upload G_k of shape (Mz, My, Mx) in fftshift order (gf_opt output)."""
        Mx, My, Mz = self.mesh
        if not isinstance(G_k, torch.Tensor):
            G_k = torch.as_tensor(G_k, dtype=torch.float64)
        G_k = G_k.to(self.device, torch.float64).contiguous()
        _need(G_k, "G_k", (Mz, My, Mx))
        check(self._lib.mdf_p3m_set_green(self._ctx, _ptr(G_k), self._stream()), "mdf_p3m_set_green")
        self._green_ref = G_k  # keep alive until the async copy ran
        self.has_green = True

    def build_green(self, p=6, m_max=3, want_full=False):
        """Optimal G(k) on the device (yukawa_gf_opt.py:5-98).

        Returns (G_k or None, A_pm tensor); G_k is (Mz, My, Mx) in fftshift order.
        """
        Mx, My, Mz = self.mesh
        full = torch.empty((Mz, My, Mx), dtype=torch.float64, device=self.device) if want_full else None
        apm = torch.zeros(1, dtype=torch.float64, device=self.device)
        mm = (m_max, m_max, m_max) if isinstance(m_max, int) else tuple(m_max)
        check(self._lib.mdf_p3m_build_green(self._ctx, int(p), mm[0], mm[1], mm[2], _ptr(full), _ptr(apm),
                                            self._stream()), "mdf_p3m_build_green")
        self.has_green = True
        return full, apm

    # --------------------------------------------------------------- force --
    def force(self, pos, Z=None, af=1.0 / 3.0, uf=1.0, acc=None, acc_pp=None, acc_pm=None,
              e_x=None, e_y=None, e_z=None, rho=None):
        """acc = af (acc_pp + acc_pm); returns (acc, energies) with energies a
        device tensor [U, U_pp, U_pm, U_self] (yukawa_p3m_ord6.py:6-29)."""
        n = pos.shape[0]
        _need(pos, "pos", (n, 3), self.device)
        if Z is not None:
            _need(Z, "Z", (n,), self.device)
        if acc is None:
            acc = torch.empty((n, 3), dtype=torch.float64, device=self.device)
        _need(acc, "acc", (n, 3), self.device)
        for t, name, shp in ((acc_pp, "acc_pp", (n, 3)), (acc_pm, "acc_pm", (n, 3)), (e_x, "e_x", (n,)),
                             (e_y, "e_y", (n,)), (e_z, "e_z", (n,)),
                             (rho, "rho", (self.mesh[2], self.mesh[1], self.mesh[0]))):
            if t is not None:
                _need(t, name, shp, self.device)
        check(self._lib.mdf_p3m_force(self._ctx, _ptr(pos), _ptr(Z), n, float(af), float(uf), _ptr(acc),
                                      _ptr(self.energies), _ptr(acc_pp), _ptr(acc_pm), _ptr(e_x), _ptr(e_y),
                                      _ptr(e_z), _ptr(rho), self._stream()), "mdf_p3m_force")
        return acc, self.energies

    def pp(self, pos, acc_pp=None):
        """Short-range part alone (yukawa_pp.py:6): returns (acc_pp, U_pp tensor)."""
        n = pos.shape[0]
        _need(pos, "pos", (n, 3), self.device)
        if acc_pp is None:
            acc_pp = torch.empty((n, 3), dtype=torch.float64, device=self.device)
        _need(acc_pp, "acc_pp", (n, 3), self.device)
        u = self._scratch[0:1]
        check(self._lib.mdf_pp_force(self._ctx, _ptr(pos), n, _ptr(acc_pp), _ptr(u), self._stream()),
              "mdf_pp_force")
        return acc_pp, u

    def pm(self, pos, Z=None, acc_pm=None, e_x=None, e_y=None, e_z=None, rho=None):
        """Long-range part alone (yukawa_pm_ord6.py:440): returns (acc_pm, U_pm tensor)."""
        n = pos.shape[0]
        _need(pos, "pos", (n, 3), self.device)
        if Z is not None:
            _need(Z, "Z", (n,), self.device)
        if acc_pm is None:
            acc_pm = torch.empty((n, 3), dtype=torch.float64, device=self.device)
        for t, name, shp in ((acc_pm, "acc_pm", (n, 3)), (e_x, "e_x", (n,)), (e_y, "e_y", (n,)), (e_z, "e_z", (n,)),
                             (rho, "rho", (self.mesh[2], self.mesh[1], self.mesh[0]))):
            if t is not None:
                _need(t, name, shp, self.device)
        u = self._scratch[1:2]
        check(self._lib.mdf_pm_force(self._ctx, _ptr(pos), _ptr(Z), n, _ptr(acc_pm), _ptr(u), _ptr(e_x),
                                     _ptr(e_y), _ptr(e_z), _ptr(rho), self._stream()), "mdf_pm_force")
        return acc_pm, u

    # ---------------------------------------------------------- integrator --
    def kick_drift(self, pos, vel, acc, dt, pbc=1, lmin=None, lmax=None):
        n = pos.shape[0]
        for t, name in ((pos, "pos"), (vel, "vel"), (acc, "acc")):
            _need(t, name, (n, 3), self.device)
        lo = (ct.c_double * 3)(*(lmin if lmin is not None else (0.0, 0.0, 0.0)))
        hi = (ct.c_double * 3)(*(lmax if lmax is not None else self.box))
        check(self._lib.mdf_verlet_kick_drift(self._ctx, _ptr(pos), _ptr(vel), _ptr(acc), n, float(dt), int(pbc),
                                              ct.cast(lo, ct.c_void_p), ct.cast(hi, ct.c_void_p), self._stream()),
              "mdf_verlet_kick_drift")

    def kick(self, vel, acc, dt, k_sum=None):
        n = vel.shape[0]
        _need(vel, "vel", (n, 3), self.device)
        _need(acc, "acc", (n, 3), self.device)
        if k_sum is not None:
            _need(k_sum, "k_sum", None, self.device)
            if k_sum.numel() < 1:
                raise ValueError("k_sum must hold one double")
        check(self._lib.mdf_verlet_kick(self._ctx, _ptr(vel), _ptr(acc), n, float(dt), _ptr(k_sum),
                                        self._stream()), "mdf_verlet_kick")

    def thermostat(self, vel, k_sum, T_desired, it, t_out=None):
        n = vel.shape[0]
        _need(vel, "vel", (n, 3), self.device)
        for t, name in ((k_sum, "k_sum"), (t_out, "t_out")):
            if t is not None:
                _need(t, name, None, self.device)
                if t.numel() < 1:
                    raise ValueError(f"{name} must hold one double")
        check(self._lib.mdf_thermostat_scale(self._ctx, _ptr(vel), n, _ptr(k_sum), float(T_desired), int(it),
                                             _ptr(t_out), self._stream()), "mdf_thermostat_scale")

    def density_modes(self, pos, q, out=None):
        """n(q) = sum_j exp(-i q x_j) per axis: complex128 tensor (nq, 3)."""
        n = pos.shape[0]
        _need(pos, "pos", (n, 3), self.device)
        q = q.to(self.device, torch.float64).contiguous()
        nq = q.shape[0]
        if out is None:
            out = torch.empty((nq, 3), dtype=torch.complex128, device=self.device)
        _need(out, "out", (nq, 3), self.device, torch.complex128)
        check(self._lib.mdf_density_modes(self._ctx, _ptr(pos), n, _ptr(q), nq, _ptr(out), self._stream()),
              "mdf_density_modes")
        return out

    def md_run(self, pos, vel, acc, nsteps, dt, Z=None, af=1.0 / 3.0, uf=1.0, kf=1.5, q=None, q_step=None,
               nq=None, energies=None, nqt=None):
        """nsteps fused NVE steps in place; returns (energies (nsteps,4): E, K, U, T; nqt or None).

        n(q, t) is recorded for the wave numbers ``q`` (tensor), or for the
        grid q_m = m * q_step, m = 1..nq, when ``q_step`` is given.
        """
        n = pos.shape[0]
        nsteps = int(nsteps)
        for t, name in ((pos, "pos"), (vel, "vel"), (acc, "acc")):
            _need(t, name, (n, 3), self.device)
        if Z is not None:
            _need(Z, "Z", (n,), self.device)
        if energies is None:
            energies = torch.empty((nsteps, 4), dtype=torch.float64, device=self.device)
        _need(energies, "energies", (nsteps, 4), self.device)
        nqv = 0
        step = 0.0
        if q_step is not None:
            step = float(q_step)
            nqv = int(nq)
            q = None
        elif q is not None:
            q = q.to(self.device, torch.float64).contiguous()
            nqv = q.shape[0]
        if nqv and nqt is None:
            nqt = torch.empty((nsteps, nqv, 3), dtype=torch.complex128, device=self.device)
        if nqv:
            _need(nqt, "nqt", (nsteps, nqv, 3), self.device, torch.complex128)
        check(self._lib.mdf_md_run(self._ctx, _ptr(pos), _ptr(vel), _ptr(acc), _ptr(Z), n, int(nsteps), float(dt),
                                   float(af), float(uf), float(kf), _ptr(q), nqv, step, _ptr(energies),
                                   _ptr(nqt) if nqv else None, self._stream()), "mdf_md_run")
        return energies, (nqt if nqv else None)


def self_energy(n, alpha):
    """Coulomb self term of yukawa_p3m_ord6.py:23: -N alpha / sqrt(pi)."""
    return -n * alpha / math.sqrt(math.pi)
