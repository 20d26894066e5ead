"""The allocation-free S(q, omega) entry point (mdf_sqw_noalloc) against
mdf_sqw with a caller workspace, through the C ABI: two calls in a row, the
second on a smaller input, must give the same S bit for bit (the same stage
kernels run, on a library-owned static scratch instead of the workspace).
Inputs over the scratch's capacity are refused with MDF_EINVAL."""
import ctypes as ct

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _run_pair(L, x, N, dt, q_stride, nq, mean_axes):
    dev = x.device
    nrun, npd, nq_total = (int(v) for v in x.shape[:3])
    shape = (nrun, nq, npd) if mean_axes else (nrun, nq, npd, 3)
    out_ws = torch.full(shape, np.nan, dtype=torch.float64, device=dev)
    out_na = torch.full(shape, np.nan, dtype=torch.float64, device=dev)
    need = ct.c_size_t(0)
    assert L.mdf_sqw_workspace(nrun, npd, nq, ct.byref(need)) == 0
    work = torch.empty(int(need.value), dtype=torch.uint8, device=dev)
    s = ct.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    T = npd * dt
    rc = L.mdf_sqw(ct.c_void_p(x.data_ptr()), nrun, npd, nq_total, q_stride, nq, float(N), float(dt), float(T),
                   int(mean_axes), ct.c_void_p(out_ws.data_ptr()), ct.c_void_p(work.data_ptr()), int(need.value), s)
    assert rc == 0, L.mdf_last_error()
    rc = L.mdf_sqw_noalloc(ct.c_void_p(x.data_ptr()), nrun, npd, nq_total, q_stride, nq, float(N), float(dt),
                           float(T), int(mean_axes), ct.c_void_p(out_na.data_ptr()), s)
    assert rc == 0, L.mdf_last_error()
    torch.cuda.synchronize()
    return out_ws.cpu().numpy(), out_na.cpu().numpy()


def test_noalloc_matches_workspace_entry_twice():
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    from mdforce._lib import lib
    L = lib()
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(21)
    # first call: 4 runs x 13 q (39 columns, every third taken) x 2000 steps
    x1 = rng.standard_normal((4, 2000, 39, 3)) + 1j * rng.standard_normal((4, 2000, 39, 3))
    a, b = _run_pair(L, torch.from_numpy(x1).to(dev), 1000.0, 0.05, 3, 13, True)
    assert np.isfinite(a).all() and np.array_equal(a, b)
    # second call, smaller: 1 run x 5 columns x 210 steps, per axis
    x2 = rng.standard_normal((1, 210, 5, 3)) + 1j * rng.standard_normal((1, 210, 5, 3))
    a2, b2 = _run_pair(L, torch.from_numpy(x2).to(dev), 300.0, 0.1, 1, 5, False)
    assert np.isfinite(a2).all() and np.array_equal(a2, b2)
    # and its S is the oracle's (the same check as the workspace entry's own tests)
    from oracle import sqw_oracle as O
    ref = O.n_of_q_w(x2[0, :, 2, :], 300.0, 0.1, 210 * 0.1)
    assert np.max(np.abs(b2[0, 2] - ref)) <= 1e-11 * np.max(ref)


def test_noalloc_refuses_more_than_its_scratch():
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    from mdforce._lib import SQW_STATIC_BYTES, lib
    L = lib()
    dev = torch.device("cuda", 0)
    # 32 B of scratch per series sample: one run of 3 * nq series just over the capacity
    npd = 4096
    nq = SQW_STATIC_BYTES // (32 * 3 * npd) + 1
    x = torch.zeros((1, npd, nq, 3), dtype=torch.complex128, device=dev)
    out = torch.zeros((1, nq, npd), dtype=torch.float64, device=dev)
    s = ct.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    rc = L.mdf_sqw_noalloc(ct.c_void_p(x.data_ptr()), 1, npd, nq, 1, nq, 1.0, 0.1, 409.6, 1,
                           ct.c_void_p(out.data_ptr()), s)
    assert rc == -1 and b"static scratch" in L.mdf_last_error()
    torch.cuda.synchronize()
