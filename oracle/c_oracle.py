"""ctypes bindings for the C restatement (``c/liboracle_p3m.so``) and for the
compiled reference particle_particle (``_ref/libyukawa_p3m_ref.so``).

Test infrastructure only; see ``oracle/__init__.py``.
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "c", "liboracle_p3m.so")
REF_SO = os.path.join(HERE, "_ref", "libyukawa_p3m_ref.so")
REF_DIR = "/root/reference"

_dp = ct.POINTER(ct.c_double)


def _ptr(a):
    return a.ctypes.data_as(_dp)


def build(ref: bool | None = None):
    """Compile the C restatement, and the reference checker when the
    reference tree is present (``ref=None``) or when asked (``ref=True``)."""
    targets = ["oracle"]
    if ref or (ref is None and os.path.isfile(os.path.join(REF_DIR, "yukawa_p3m.c"))):
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", HERE] + targets, check=True)


_lib = None
_ref = None
_omp = None
ORACLE_OMP_SO = os.path.join(HERE, "c", "liboracle_p3m_omp.so")


def lib_omp(threads: int = 0):
    """The same restatement built with OpenMP (all-core CPU baseline);
    threads > 0 sets the thread count.  Returns (lib, threads in use)."""
    global _omp
    if _omp is None:
        if not os.path.isfile(ORACLE_OMP_SO):
            build(ref=False)
        L = ct.CDLL(ORACLE_OMP_SO)
        base = lib()
        for name in ("orc_pp", "orc_charge_assign", "orc_field_gather", "orc_fft3d", "orc_gf_opt",
                     "orc_pm", "orc_force_pot", "orc_md_steps"):
            fn = getattr(L, name)
            fn.argtypes = getattr(base, name).argtypes
            fn.restype = ct.c_int
        L.orc_threads.argtypes = [ct.c_int]
        L.orc_threads.restype = ct.c_int
        _omp = L
    return _omp, _omp.orc_threads(int(threads))


def lib():
    global _lib
    if _lib is None:
        if not os.path.isfile(ORACLE_SO):
            build(ref=False)
        L = ct.CDLL(ORACLE_SO)
        i, d = ct.c_int, ct.c_double
        L.orc_pp.argtypes = [_dp, i, _dp, d, d, d, _dp, _dp]
        L.orc_charge_assign.argtypes = [_dp, _dp, i, i, i, i, d, d, d, _dp]
        L.orc_field_gather.argtypes = [_dp, _dp, _dp, _dp, _dp, i, i, i, i, d, d, d, _dp, _dp]
        L.orc_fft3d.argtypes = [_dp, i, i, i, i]
        L.orc_gf_opt.argtypes = [d, d, i, i, i, i, i, i, i, d, d, d, d, d, d, _dp, _dp]
        L.orc_pm.argtypes = [_dp, _dp, i, _dp, i, i, i, d, d, d, d, d, d, _dp, _dp, _dp, _dp]
        L.orc_force_pot.argtypes = [_dp, _dp, i, _dp, d, d, d, d, d, _dp, i, _dp, _dp, _dp, _dp]
        L.orc_md_steps.argtypes = [_dp, _dp, _dp, _dp, i, i, d, _dp, d, d, d, d, d, d, _dp, i,
                                   _dp, i, _dp, _dp]
        for name in ("orc_pp", "orc_charge_assign", "orc_field_gather", "orc_fft3d", "orc_gf_opt",
                     "orc_pm", "orc_force_pot", "orc_md_steps"):
            getattr(L, name).restype = ct.c_int
        _lib = L
    return _lib


def ref_available():
    return os.path.isfile(REF_SO)


def ref_lib():
    global _ref
    if _ref is None:
        L = ct.CDLL(REF_SO)
        L.ref_particle_particle.argtypes = [_dp, _dp, ct.c_int, ct.c_double, ct.c_double,
                                            ct.c_double, ct.c_double, _dp]
        L.ref_particle_particle.restype = ct.c_int
        _ref = L
    return _ref


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with code {rc}")


def ref_particle_particle(pos, Z, L, kappa, G, rc):
    """Accelerations from the reference's own C particle_particle (Z_i Z_j
    weighted, yukawa_p3m.c:374-391).  The reference returns no energy."""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    Z = np.ascontiguousarray(Z, dtype=np.float64)
    acc = np.zeros_like(pos)
    _check(ref_lib().ref_particle_particle(_ptr(pos), _ptr(Z), pos.shape[0], L, kappa, G, rc, _ptr(acc)),
           "ref_particle_particle")
    return acc


def pp(pos, Lv, kappa, G, rc):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    Lv = np.ascontiguousarray(Lv, dtype=np.float64)
    acc = np.zeros_like(pos)
    U = np.zeros(1)
    _check(lib().orc_pp(_ptr(pos), pos.shape[0], _ptr(Lv), kappa, G, rc, _ptr(acc), _ptr(U)), "orc_pp")
    return float(U[0]), acc


def charge_assign(pos, Z, M, h):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    Z = np.ascontiguousarray(Z, dtype=np.float64)
    rho = np.zeros((M, M, M))
    _check(lib().orc_charge_assign(_ptr(pos), _ptr(Z), pos.shape[0], M, M, M, h, h, h, _ptr(rho)),
           "orc_charge_assign")
    return rho


def fft3d(c, sign):
    c = np.ascontiguousarray(c, dtype=np.complex128).copy()
    Mz, My, Mx = c.shape
    _check(lib().orc_fft3d(c.ctypes.data_as(_dp), Mx, My, Mz, sign), "orc_fft3d")
    return c


def gf_opt(kappa, Gew, p, m_max, M, L, threads=None):
    """threads: None runs the serial checker; an int the OpenMP build (same
    G_k; A_pm summed in another order)."""
    h = L / M
    G_k = np.zeros((M, M, M))
    A = np.zeros(1)
    L_ = lib() if threads is None else lib_omp(threads)[0]
    _check(L_.orc_gf_opt(kappa, Gew, p, m_max, m_max, m_max, M, M, M, h, h, h, L, L, L,
                            _ptr(G_k), _ptr(A)), "orc_gf_opt")
    return G_k, float(A[0])


def pm(pos, Z, G_k, M, L):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    Z = np.ascontiguousarray(Z, dtype=np.float64)
    G_k = np.ascontiguousarray(G_k, dtype=np.float64)
    h = L / M
    acc = np.zeros_like(pos)
    E_p = np.zeros_like(pos)
    rho = np.zeros((M, M, M))
    U = np.zeros(1)
    _check(lib().orc_pm(_ptr(pos), _ptr(Z), pos.shape[0], _ptr(G_k), M, M, M, h, h, h, L, L, L,
                        _ptr(acc), _ptr(E_p), _ptr(rho), _ptr(U)), "orc_pm")
    return float(U[0]), acc, rho, E_p


def force_pot(pos, Z, L, kappa, G, rc, G_k, M, af=1.0 / 3.0, uf=1.0):
    """Returns (U, acc, acc_pp, acc_pm, (U_pp, U_pm, U_self))."""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    Z = np.ascontiguousarray(Z, dtype=np.float64)
    G_k = np.ascontiguousarray(G_k, dtype=np.float64)
    Lv = np.array([L, L, L])
    acc = np.zeros_like(pos)
    app = np.zeros_like(pos)
    apm = np.zeros_like(pos)
    Up = np.zeros(4)
    _check(lib().orc_force_pot(_ptr(pos), _ptr(Z), pos.shape[0], _ptr(Lv), kappa, G, rc, af, uf,
                               _ptr(G_k), M, _ptr(acc), _ptr(app), _ptr(apm), _ptr(Up)), "orc_force_pot")
    return float(Up[0]), acc, app, apm, (float(Up[1]), float(Up[2]), float(Up[3]))


def md_steps(pos, vel, acc, Z, nsteps, dt, L, kappa, G, rc, G_k, M, q=None,
             af=1.0 / 3.0, uf=1.0, kf=1.5, threads=None):
    """Runs nsteps NVE steps in place; returns (energies (nsteps,4), nqt).
    threads: None runs the serial checker; an int runs the OpenMP build on
    that many threads (0: all)."""
    L_ = lib() if threads is None else lib_omp(threads)[0]
    Lv = np.array([L, L, L])
    q = np.zeros(0) if q is None else np.ascontiguousarray(q, dtype=np.float64)
    en = np.zeros((nsteps, 4))
    nqt = np.zeros((nsteps, q.size, 3), dtype=np.complex128)
    G_k = np.ascontiguousarray(G_k, dtype=np.float64)
    _check(L_.orc_md_steps(_ptr(pos), _ptr(vel), _ptr(acc), _ptr(np.ascontiguousarray(Z, dtype=np.float64)),
                              pos.shape[0], nsteps, dt, _ptr(Lv), kappa, G, rc, af, uf, kf, _ptr(G_k), M,
                              _ptr(q), q.size, _ptr(en), nqt.ctypes.data_as(_dp)), "orc_md_steps")
    return en, nqt
