"""Learned short-range correction: the pair network of csrc/mlp.hip.

A PairMLP holds the network's parameters (FP32 master copy, Adam state) and
its pair-list scratch on the GPU.  Forward, correction force, loss and
gradient all run in the HIP kernels (BF16 MFMA with FP32 accumulation);
this module only moves pointers.  Attached to a P3M engine (``attach``) the
correction is added to the PP force of every evaluation of that engine,
including the fused NVE steps of ``md_run``.
"""
from __future__ import annotations

import ctypes as ct
import math

import numpy as np
import torch

from ._lib import MLPMeta, MdforceError, check, last_error, lib

N_IN, N_H = 16, 64


def n_param():
    return lib().mdf_mlp_nparam()


def default_meta(kappa, alpha, rc, out_scale=1e-3):
    """Radial basis of 12 Gaussians spread over [0.25, rc]; constant features scaled to O(1)."""
    mu0 = 0.25
    dmu = (rc - mu0) / 11.0
    return dict(mu0=mu0, dmu=dmu, inv_sigma=1.0 / dmu, f_kappa=kappa / 3.0, f_alpha=alpha, f_rc=rc / 8.0,
                out_scale=out_scale)


def init_params(seed=0, out_std=0.1):
    """He-normal hidden weights, small output weights, zero biases (packed FP32 vector)."""
    rs = np.random.RandomState(seed)
    parts = [rs.randn(N_IN, N_H) * math.sqrt(2.0 / N_IN), np.zeros(N_H),
             rs.randn(N_H, N_H) * math.sqrt(2.0 / N_H), np.zeros(N_H),
             rs.randn(N_H, N_H) * math.sqrt(2.0 / N_H), np.zeros(N_H),
             rs.randn(N_H) * out_std / math.sqrt(N_H), np.zeros(1)]
    return np.concatenate([p.reshape(-1) for p in parts]).astype(np.float32)


def _meta_struct(meta):
    m = MLPMeta()
    for k in ("mu0", "dmu", "inv_sigma", "f_kappa", "f_alpha", "f_rc", "out_scale"):
        setattr(m, k, float(meta[k]))
    return m


def _ptr(t):
    return None if t is None else ct.c_void_p(t.data_ptr())


class PairMLP:
    def __init__(self, n_max, max_pairs, meta, device=None, seed=0):
        if not torch.cuda.is_available():
            raise RuntimeError("mdforce needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        self.n_max, self.max_pairs = int(n_max), int(max_pairs)
        self.meta = dict(meta)
        self._lib = lib()
        self._m = ct.c_void_p()
        ms = _meta_struct(self.meta)
        with torch.cuda.device(self.device):
            check(self._lib.mdf_mlp_create(self.n_max, self.max_pairs, ct.byref(ms), ct.byref(self._m)),
                  "mdf_mlp_create")
        self.nparam = n_param()
        self.grad_buf = torch.zeros(self.nparam, dtype=torch.float32, device=self.device)
        self.loss_buf = torch.zeros(1, dtype=torch.float64, device=self.device)
        self.set_params(init_params(seed))

    def _stream(self):
        return ct.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def close(self):
        if getattr(self, "_m", None) and self._m.value:
            self._lib.mdf_mlp_destroy(self._m)
            self._m = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_meta(self, **kw):
        self.meta.update(kw)
        check(self._lib.mdf_mlp_set_meta(self._m, ct.byref(_meta_struct(self.meta))), "mdf_mlp_set_meta")

    def set_params(self, p, reset_adam=True):
        t = torch.as_tensor(np.asarray(p, dtype=np.float32) if not isinstance(p, torch.Tensor) else p)
        t = t.to(self.device, torch.float32).contiguous()
        if t.numel() != self.nparam:
            raise ValueError(f"expected {self.nparam} parameters, got {t.numel()}")
        check(self._lib.mdf_mlp_params(self._m, _ptr(t), None, int(bool(reset_adam)), self._stream()),
              "mdf_mlp_params")
        torch.cuda.current_stream(self.device).synchronize()

    def params(self):
        out = torch.empty(self.nparam, dtype=torch.float32, device=self.device)
        check(self._lib.mdf_mlp_params(self._m, None, _ptr(out), 0, self._stream()), "mdf_mlp_params")
        return out

    # ------------------------------------------------------------ passes --
    def pairs(self, engine, pos, want=True):
        """Bin pos with the engine's context, build the pair list (r < engine.rc), run the
        network.  Returns dict(npairs, i, j, r, u, g, rho, perm) (sorted indices)."""
        n = pos.shape[0]
        words = None
        g = rho = perm = None
        if want:
            words = torch.empty(1 + 6 * self.max_pairs, dtype=torch.int32, device=self.device)
            g = torch.empty(self.max_pairs, dtype=torch.float32, device=self.device)
            rho = torch.empty(n, dtype=torch.float32, device=self.device)
            perm = torch.empty(n, dtype=torch.int32, device=self.device)
        rc = self._lib.mdf_mlp_pairs(self._m, engine._ctx, _ptr(pos), n, _ptr(words), _ptr(g), _ptr(rho),
                                     _ptr(perm), self._stream())
        if rc < 0:
            raise MdforceError(f"mdf_mlp_pairs failed (code {rc}): {last_error()}")
        P = rc
        if not want:
            return {"npairs": P}
        rec = words[1:1 + 6 * P].view(P, 6)
        f = rec[:, 2:].contiguous().view(torch.float32)
        return {"npairs": P, "i": rec[:, 0].long(), "j": rec[:, 1].long(), "r": f[:, 0], "u": f[:, 1:4],
                "g": g[:P], "rho": rho, "perm": perm.long()}

    def correction(self, engine, pos, out=None):
        n = pos.shape[0]
        if out is None:
            out = torch.empty((n, 3), dtype=torch.float64, device=self.device)
        check(self._lib.mdf_mlp_correction(self._m, engine._ctx, _ptr(pos), n, _ptr(out), self._stream()),
              "mdf_mlp_correction")
        return out

    def grad(self, engine, pos, target, build=True):
        """(gradient (nparam,) FP32, loss (1,) FP64) of (1/n) sum |F_i - target_i|^2."""
        n = pos.shape[0]
        check(self._lib.mdf_mlp_grad(self._m, engine._ctx, _ptr(pos), _ptr(target), n, _ptr(self.grad_buf),
                                     _ptr(self.loss_buf), int(bool(build)), self._stream()), "mdf_mlp_grad")
        return self.grad_buf, self.loss_buf

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        check(self._lib.mdf_mlp_adam(self._m, float(lr), float(beta1), float(beta2), float(eps), self._stream()),
              "mdf_mlp_adam")

    def attach(self, engine):
        check(self._lib.mdf_p3m_attach_mlp(engine._ctx, self._m), "mdf_p3m_attach_mlp")
        engine._mlp = self

    @staticmethod
    def detach(engine):
        check(lib().mdf_p3m_attach_mlp(engine._ctx, None), "mdf_p3m_attach_mlp")
        engine._mlp = None

    def stats(self, n):
        out = (ct.c_int * 2)()
        check(self._lib.mdf_mlp_stats(self._m, int(n), out), "mdf_mlp_stats")
        return {"npairs": out[0], "overflow": out[1]}
