"""Shared plumbing of the drop-in modules: engine cache and array transfer.

The reference passes NumPy arrays and a long positional argument list on
every call.  The drop-ins accept the same arguments (NumPy arrays or float64
GPU tensors) and keep one P3M engine per (N, mesh, box, kappa, alpha, rc),
re-uploading G_k and Z only when a different array is passed.
"""
from __future__ import annotations

import numpy as np
import torch

from .engine import P3M

# Mirror the reference's scratch side effects (acc_s_r, acc_fft, rho_r, E_*_p
# are left holding intermediates, yukawa_p3m_ord6.py:8-19) when they are NumPy
# arrays.  Costs one device-to-host copy each per call; switch off for speed.
MIRROR_SCRATCH = True

_engines: dict = {}


def device():
    if not torch.cuda.is_available():
        raise RuntimeError("mdforce drop-ins need a ROCm GPU")
    return torch.device("cuda", torch.cuda.current_device())


def to_dev(a):
    """float64 contiguous GPU tensor view/copy of a NumPy array or tensor."""
    if isinstance(a, torch.Tensor):
        return a.to(device(), torch.float64).contiguous()
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(device())


def like_input(t, ref):
    """Return t as the same kind of array as ref (NumPy in, NumPy out)."""
    if isinstance(ref, torch.Tensor):
        return t
    return t.cpu().numpy()


def copy_into(dst, src):
    """In-place update of a caller scratch buffer (NumPy or tensor)."""
    if dst is None:
        return
    if isinstance(dst, torch.Tensor):
        dst.copy_(src.reshape(dst.shape))
    else:
        dst[...] = src.cpu().numpy().reshape(dst.shape)


def _fingerprint(a):
    if isinstance(a, torch.Tensor):
        return ("t", a.data_ptr(), tuple(a.shape), a._version)
    a = np.asarray(a)
    flat = a.reshape(-1)
    probe = tuple(float(flat[i]) for i in (0, flat.size // 3, flat.size // 2, flat.size - 1)) if flat.size else ()
    return ("n", id(a), a.shape, probe)


def engine(N, Mx, My, Mz, Lx, Ly, Lz, kappa, alpha, rc):
    key = (int(N), int(Mx), int(My), int(Mz), float(Lx), float(Ly), float(Lz), float(kappa), float(alpha),
           float(rc))
    e = _engines.get(key)
    if e is None:
        e = P3M(int(N), (int(Mx), int(My), int(Mz)), (float(Lx), float(Ly), float(Lz)), kappa, alpha, rc)
        e._green_fp = None
        e._z_fp = None
        e._z_dev = None
        _engines[key] = e
    return e


def ensure_green(e, G_k):
    fp = _fingerprint(G_k)
    if e._green_fp != fp:
        e.set_green(to_dev(G_k))
        e._green_fp = fp


def charges(e, Z, N):
    if Z is None:
        return None
    fp = _fingerprint(Z)
    if e._z_fp != fp:
        e._z_dev = to_dev(Z).reshape(N)
        e._z_fp = fp
    return e._z_dev


def clear_cache():
    for e in _engines.values():
        e.close()
    _engines.clear()
