"""S(q, omega) of density-mode series on the GPU (the ``mdf_sqw`` C ABI).

``sqw(nqt, N, dt, T, q_stride=3, nq=None, mean_axes=True)`` takes the
reference's n_q_t arrays -- complex128 (npd, nq_total, 3), or a stack
(nrun, npd, nq_total, 3) of runs, as ``P3M.md_run`` and ``Batch.run`` return
them -- and returns S = |dt FFT_t n|^2 / (N T) for the columns
``i * q_stride``, i < nq: (nrun, nq, npd) means over the axes, or
(nrun, nq, npd, 3) per axis (s_of_q_w.py:3-20, s_of_q_w_avg.py:5-20).
Everything runs on the device; inputs on the host are copied over.
"""
from __future__ import annotations

import ctypes as ct

import torch

from ._lib import check, lib


def sqw(nqt, N, dt, T, q_stride=3, nq=None, mean_axes=True, out=None, stream=None):
    if not torch.cuda.is_available():
        raise RuntimeError("mdforce.sqw needs a ROCm GPU")
    dev = torch.device("cuda", torch.cuda.current_device())
    x = nqt if isinstance(nqt, torch.Tensor) else torch.as_tensor(nqt)
    x = x.to(dev, torch.complex128).contiguous()
    if x.dim() == 3:
        x = x.unsqueeze(0)
    if x.dim() != 4 or x.shape[3] != 3:
        raise ValueError(f"n_q_t must be (npd, nq_total, 3) or (nrun, npd, nq_total, 3), got {tuple(x.shape)}")
    nrun, npd, nq_total = (int(v) for v in x.shape[:3])
    if nq is None:
        nq = (nq_total + q_stride - 1) // q_stride
    shape = (nrun, nq, npd) if mean_axes else (nrun, nq, npd, 3)
    if out is None:
        out = torch.empty(shape, dtype=torch.float64, device=dev)
    elif tuple(out.shape) != shape or out.dtype != torch.float64 or not out.is_contiguous():
        raise ValueError(f"out must be a contiguous float64 tensor of shape {shape}")
    L = lib()
    need = ct.c_size_t(0)
    check(L.mdf_sqw_workspace(nrun, npd, int(nq), ct.byref(need)), "mdf_sqw_workspace")
    work = torch.empty(max(int(need.value), 16), dtype=torch.uint8, device=dev)
    s = stream if stream is not None else torch.cuda.current_stream(dev)
    check(L.mdf_sqw(ct.c_void_p(x.data_ptr()), nrun, npd, nq_total, int(q_stride), int(nq), float(N), float(dt),
                    float(T), 1 if mean_axes else 0, ct.c_void_p(out.data_ptr()), ct.c_void_p(work.data_ptr()),
                    int(need.value), ct.c_void_p(s.cuda_stream)), "mdf_sqw")
    return out
