"""Replica ensembles on one GPU (mdf_batch_* of the C ABI).

Many independent seeded systems - the runs whose n(q, t) the reference
averages into S(q, w) (s_of_q_w_avg.py) - advance together: each replica has
its own context (box, mesh, kappa, alpha, rc), and the NVE steps of all
replicas are captured into one hipGraph of ``graph_steps`` steps that is
replayed.  By default each stage of the step is one segmented launch for all
replicas (``set_mode(0)``); ``set_mode(1)`` captures each replica's own
launches on worker streams instead.  n(q, t) on the grid q_m = m * 2 pi / L_r
(m = 1..nq, along x, y and z) and the per-step (E, K, U, T) are written to
device arrays for every step.
"""
from __future__ import annotations

import ctypes as ct

import torch

from ._lib import P3MParams, check, lib
from .engine import _PREC, P3M, _need


class _Replica(P3M):
    """A P3M view of a context owned by a batch (not destroyed with the view)."""

    def __init__(self, ctx, n_max, mesh, box, kappa, alpha, rc, precision, device):
        self.device = device
        self.n_max = n_max
        self.mesh = (mesh, mesh, mesh)
        self.box = (box, box, box)
        self.kappa, self.alpha, self.rc = kappa, alpha, rc
        self.precision = precision
        self._lib = lib()
        self._ctx = ctx
        self.energies = torch.zeros(4, dtype=torch.float64, device=device)
        self._scratch = torch.zeros(4, dtype=torch.float64, device=device)
        self.has_green = False

    def close(self):
        self._ctx = ct.c_void_p()


class ReplicaBatch:
    """systems: list of dicts with keys n, L, kappa, mesh, alpha, rc (and
    optionally precision, default "mixed")."""

    def __init__(self, systems, n_streams=8, device=None):
        if not torch.cuda.is_available():
            raise RuntimeError("mdforce needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        self.systems = [dict(s) for s in systems]
        R = len(self.systems)
        arr = (P3MParams * R)()
        for r, s in enumerate(self.systems):
            p = arr[r]
            p.n_max = int(s["n"])
            for d in range(3):
                p.mesh[d] = int(s["mesh"])
                p.box[d] = float(s["L"])
            p.kappa, p.alpha, p.rc = float(s["kappa"]), float(s["alpha"]), float(s["rc"])
            p.order = 6
            p.precision = _PREC[s.get("precision", "mixed")]
        self._lib = lib()
        self._b = ct.c_void_p()
        with torch.cuda.device(self.device):
            check(self._lib.mdf_batch_create(R, arr, int(n_streams), ct.byref(self._b)), "mdf_batch_create")
        self.replicas = []
        for r, s in enumerate(self.systems):
            c = ct.c_void_p()
            check(self._lib.mdf_batch_context(self._b, r, ct.byref(c)), "mdf_batch_context")
            rep = _Replica(c, int(s["n"]), int(s["mesh"]), float(s["L"]), float(s["kappa"]), float(s["alpha"]),
                           float(s["rc"]), s.get("precision", "mixed"), self.device)
            rep.build_green()
            self.replicas.append(rep)
        self._keep = None

    def set_skins(self, skins):
        """Pair-list skin per replica (0: none; see P3M.set_skin).  In the
        segmented graph the replicas with lists share one list-aware PP
        launch and those without another."""
        if len(skins) != len(self.replicas):
            raise ValueError("one skin per replica")
        for rep, s in zip(self.replicas, skins):
            rep.set_skin(float(s))

    def apply_skin(self, r, skin):
        """Pair lists of skin ``skin`` for replica r (0: none): cells of
        floor(L / (rc + skin)) per axis and the list skin capped at ``skin``.
        Returns the list skin in use (0 when the box is too small for
        L / (rc + skin) >= 3)."""
        from ._lib import MdforceError
        rep = self.replicas[r]
        if not skin > 0.0:
            rep.set_skin(0.0)
            return 0.0
        try:
            rep.set_skin(float(skin))
        except MdforceError:
            rep.set_skin(0.0)
            return 0.0
        return rep.set_list_skin(float(skin))

    def tune_skins(self, pos, vel, acc, dt, targets=(0.25, 0.5, 1.0), steps=20, nq=0, log=None):
        """Pair-list skins by measurement, in two passes:
        1. each replica stepped alone (md_run from copies of its state) with
           no lists and with each skin of ``targets`` (apply_skin); the
           fastest is its choice;
        2. the whole ensemble (graph of ``steps`` steps, run twice from copies
           of (pos, vel, acc), the second run timed) with those choices and
           with no lists at all; the faster is kept.
        Returns the list skin of each replica."""
        import time
        from .autotune import time_steps
        R = len(self.replicas)
        chosen = []
        for r in range(R):
            best = None
            for s in (0.0,) + tuple(targets):
                eff = self.apply_skin(r, s)
                if s > 0.0 and eff == 0.0:
                    continue
                t = time_steps(self.replicas[r], pos[r], vel[r], dt, steps=steps)
                if log:
                    log(f"replica {r}: skin {eff:.3f} alone {t:.4f} ms/step")
                if best is None or t < best[1]:
                    best = (s, t)
            chosen.append(best[0])
        P = [t.clone() for t in pos]
        V = [t.clone() for t in vel]
        A = [t.clone() for t in acc]

        def trial():
            for _ in range(2):
                for r in range(R):
                    P[r].copy_(pos[r])
                    V[r].copy_(vel[r])
                    A[r].copy_(acc[r])
                torch.cuda.synchronize(self.device)
                t0 = time.perf_counter()
                self.run(P, V, A, steps, dt, nq=nq, graph_steps=steps)
                torch.cuda.synchronize(self.device)
                t = time.perf_counter() - t0
            return t * 1e3 / steps

        for r in range(R):
            self.apply_skin(r, 0.0)
        t_none = trial()
        skins = [self.apply_skin(r, s) for r, s in enumerate(chosen)]
        t_lists = trial()
        if log:
            log(f"ensemble: no lists {t_none:.4f} ms/step, chosen skins {t_lists:.4f} ms/step "
                f"({sum(1 for x in skins if x > 0)} replicas with lists)")
        if t_none < t_lists:
            skins = [self.apply_skin(r, 0.0) for r in range(R)]
        return skins

    def set_mode(self, mode):
        """0: segmented launches (one per stage for all replicas); 1: per-replica streams."""
        check(self._lib.mdf_batch_set_mode(self._b, int(mode)), "mdf_batch_set_mode")

    @property
    def graph_mode(self):
        return self._lib.mdf_batch_mode(self._b)

    def __len__(self):
        return len(self.systems)

    def close(self):
        if getattr(self, "_b", None) and self._b.value:
            for rep in self.replicas:
                rep.close()
            self._lib.mdf_batch_destroy(self._b)
            self._b = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, pos, vel, acc, nsteps, dt, nq=20, energies=None, nqt=None, graph_steps=50, af=1.0 / 3.0,
            uf=1.0, kf=1.5):
        """nsteps NVE steps of every replica in place.

        pos/vel/acc: lists of (n_r, 3) float64 GPU tensors (acc holds the
        force of pos on entry).  Returns (energies, nqt): lists of (nsteps, 4)
        and (nsteps, nq, 3) complex tensors (allocated if not given).
        """
        R = len(self.systems)
        if not (len(pos) == len(vel) == len(acc) == R):
            raise ValueError("pos, vel and acc need one tensor per replica")
        ns = []
        for r in range(R):
            n = pos[r].shape[0]
            for t, name in ((pos[r], "pos"), (vel[r], "vel"), (acc[r], "acc")):
                _need(t, f"{name}[{r}]", (n, 3), self.device)
            ns.append(n)
        if energies is None:
            energies = [torch.empty((nsteps, 4), dtype=torch.float64, device=self.device) for _ in range(R)]
        if nq and nqt is None:
            nqt = [torch.empty((nsteps, nq, 3), dtype=torch.complex128, device=self.device) for _ in range(R)]
        P = ct.c_void_p * R
        pp = P(*[t.data_ptr() for t in pos])
        vp = P(*[t.data_ptr() for t in vel])
        ap = P(*[t.data_ptr() for t in acc])
        ep = P(*[t.data_ptr() for t in energies])
        qp = P(*[t.data_ptr() for t in nqt]) if nq else None
        narr = (ct.c_int * R)(*ns)
        stream = ct.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
        check(self._lib.mdf_batch_run(self._b, pp, vp, ap, narr, int(nsteps), float(dt), float(af), float(uf),
                                      float(kf), int(nq), ep, qp, int(graph_steps), stream), "mdf_batch_run")
        self._keep = (pos, vel, acc, energies, nqt)
        return energies, (nqt if nq else None)
