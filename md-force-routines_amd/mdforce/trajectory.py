"""Trajectory, energy and n(q, t) files in the reference's formats, written
from the device loop without stalling it.

Formats (yukawa_MD_p3m_ord6.py:189-242):
* ``pos_vel_acc.out``: every snapshot, the (N, 9) rows x y z vx vy vz ax ay
  az written by ``numpy.savetxt`` (``%.18e``, space separated) (:224-225);
* ``time_temp_totalEnergy_kinEnergy_potEnergy.out``: one line per snapshot,
  t Tp E K U with t = dt * it, Tp = K / (kf N) (:209-212, :226);
* ``pos_vel_acc.xyz``: per snapshot the line N, the line
  ``x y z vx vy vz ax ay az`` and the same rows (:228-231);
* ``n_q_t_1.npy``: the (Nt, Nq, 3) complex128 modes, ``numpy.save`` (:233);
* ``pos_vel_acc_final.out``: the last (N, 9) rows (:241-242), which the
  reference reads back as ``init.out`` through ``read.initL`` to resume
  (:147-148; ``mdforce.read``).
A snapshot is taken after step it when (it + 1) is a multiple of snap_int.

A snapshot costs the device loop one gather of pos, vel and acc into a
staging slot on the loop's stream.  The device-to-host copy runs on a
second stream into pinned host memory, and a writer thread formats and
writes the text, so the loop continues while a snapshot is written.  Two
slots: a third snapshot waits for the oldest one's copy to finish, and the
writer thread for its file writes (text formatting of N = 1e6 rows takes
seconds, as in the reference).
"""
from __future__ import annotations

import os
import queue
import threading

import numpy as np
import torch

XYZ_HEADER = "x y z vx vy vz ax ay az\n"


def write_snapshot_text(f_output, f_output_E, f_xyz, irp, row5, write_xyz=True):
    """The text of one snapshot, in the reference's order (:224-231)."""
    np.savetxt(f_output, irp)
    np.savetxt(f_output_E, np.asarray(row5, dtype=np.float64).reshape(1, 5))
    if write_xyz and f_xyz is not None:
        f_xyz.writelines("{0:d}\n".format(irp.shape[0]))
        f_xyz.writelines(XYZ_HEADER)
        np.savetxt(f_xyz, irp)


class TrajectoryWriter:
    """Asynchronous writer of the reference's production-run files.

    ``snapshot(pos, vel, acc, it, dt, energies_row, kf=1.5)`` takes device
    (N, 3) float64 tensors and the step's (E, K, U, T) row of ``md_run``'s
    energies; ``close()`` drains the writer and closes the files.
    """

    def __init__(self, N, directory=".", write_xyz=True, device=None, slots=2):
        if not torch.cuda.is_available():
            raise RuntimeError("TrajectoryWriter needs a ROCm GPU")
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.N = int(N)
        self.write_xyz = bool(write_xyz)
        os.makedirs(directory, exist_ok=True)
        self.f_output = open(os.path.join(directory, "pos_vel_acc.out"), "w")
        self.f_output_E = open(os.path.join(directory, "time_temp_totalEnergy_kinEnergy_potEnergy.out"), "w")
        self.f_xyz = open(os.path.join(directory, "pos_vel_acc.xyz"), "w") if self.write_xyz else None
        self.stage = torch.empty((slots, self.N, 9), dtype=torch.float64, device=self.device)
        self.stage_e = torch.empty((slots, 3), dtype=torch.float64, device=self.device)
        self.host = torch.empty((slots, self.N, 9), dtype=torch.float64, pin_memory=True)
        self.host_e = torch.empty((slots, 3), dtype=torch.float64, pin_memory=True)
        self.copy_stream = torch.cuda.Stream(self.device)  # default priority
        self.copied = [None] * slots                        # D2H done events
        self.free = [threading.Event() for _ in range(slots)]
        for ev in self.free:
            ev.set()
        self.slot = 0
        self.q = queue.Queue()
        self.err = None
        self.worker = threading.Thread(target=self._run, daemon=True)
        self.worker.start()

    def _run(self):
        while True:
            item = self.q.get()
            if item is None:
                return
            slot, done, t, kf = item
            try:
                done.synchronize()
                E, K, U = (float(x) for x in self.host_e[slot, :3].numpy())
                # t and Tp on the host, with the reference's float64 arithmetic (:202, :209)
                write_snapshot_text(self.f_output, self.f_output_E, self.f_xyz, self.host[slot].numpy(),
                                    [t, K / kf / float(self.N), E, K, U], self.write_xyz)
            except Exception as e:  # reported by the next call on the loop's thread
                self.err = e
            finally:
                self.free[slot].set()

    def snapshot(self, pos, vel, acc, it, dt, energies_row, kf=1.5):
        if self.err is not None:
            raise RuntimeError(f"trajectory writer failed: {self.err}")
        s = self.slot
        self.slot = (s + 1) % len(self.free)
        self.free[s].wait()             # the slot's previous text is written
        self.free[s].clear()
        main = torch.cuda.current_stream(self.device)
        if self.copied[s] is not None:  # and its staging is no longer being copied
            main.wait_event(self.copied[s])
        st = self.stage[s]
        st[:, 0:3].copy_(pos)
        st[:, 3:6].copy_(vel)
        st[:, 6:9].copy_(acc)
        self.stage_e[s, :3].copy_(energies_row.reshape(4)[:3])  # E, K, U
        ready = torch.cuda.Event()
        ready.record(main)
        with torch.cuda.stream(self.copy_stream):
            self.copy_stream.wait_event(ready)
            self.host[s].copy_(st, non_blocking=True)
            self.host_e[s].copy_(self.stage_e[s], non_blocking=True)
            done = torch.cuda.Event()
            done.record(self.copy_stream)
        self.copied[s] = done
        self.q.put((s, done, dt * it, kf))

    def close(self):
        self.q.put(None)
        self.worker.join()
        for f in (self.f_output, self.f_output_E, self.f_xyz):
            if f is not None:
                f.close()
        if self.err is not None:
            raise RuntimeError(f"trajectory writer failed: {self.err}")


def save_n_q_t(nqt, path="n_q_t_1"):
    """numpy.save of the (Nt, Nq, 3) complex128 modes (:233)."""
    np.save(path, nqt.cpu().numpy() if isinstance(nqt, torch.Tensor) else np.asarray(nqt))


def write_final(pos, vel, acc, path="pos_vel_acc_final.out"):
    """The last positions, velocities and accelerations (:241-242)."""
    def h(a):
        return a.cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    np.savetxt(path, np.hstack((h(pos), h(vel), h(acc))))
