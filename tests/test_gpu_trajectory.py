"""The asynchronous trajectory writer against synchronous NumPy copies.

A device MD run (mdf_md_run in chunks of snap_int steps) writes snapshots
through mdforce.trajectory.TrajectoryWriter while the next chunk runs.  The
files must be byte-identical to what numpy.savetxt writes from host copies
of the same tensors taken synchronously at each snapshot (the reference's
own writes, yukawa_MD_p3m_ord6.py:222-242), the final file must resume
bit-exactly through read.initL, and n_q_t_1.npy must hold the device modes.
"""
import io
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def test_writer_matches_synchronous_numpy(tmp_path):
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    from mdforce import read
    from mdforce.engine import P3M
    from mdforce.initialize_pos_vel import initial_split
    from mdforce.trajectory import TrajectoryWriter, save_n_q_t, write_final, write_snapshot_text
    dev = torch.device("cuda", 0)
    n, gamma, kappa, dt, snap_int, nsnap = 1000, 50.0, 1.0, 0.1, 5, 4
    L = (4.0 * math.pi * n / 3.0) ** (1.0 / 3.0)
    pos, vel = initial_split(3, 4, n, [L, L, L], 1.0 / gamma)
    e = P3M(n, 16, L, kappa, 0.9, L / 3.0 * 0.999)
    e.build_green()
    p = torch.from_numpy(pos).to(dev)
    v = torch.from_numpy(vel).to(dev)
    a, _ = e.force(p)
    a = a.clone()
    w = TrajectoryWriter(n, str(tmp_path / "async"))
    fo, fe, fx = io.StringIO(), io.StringIO(), io.StringIO()
    nqts = []
    for k in range(nsnap):
        en, nqt = e.md_run(p, v, a, snap_int, dt, q_step=2 * math.pi / L, nq=6)
        nqts.append(nqt)
        it = (k + 1) * snap_int - 1
        w.snapshot(p, v, a, it, dt, en[-1])
        torch.cuda.synchronize()
        irp = np.hstack((p.cpu().numpy(), v.cpu().numpy(), a.cpu().numpy()))
        E, K, U, _ = en[-1].cpu().numpy()
        write_snapshot_text(fo, fe, fx, irp, [dt * it, K / 1.5 / n, E, K, U])
    w.close()
    d = tmp_path / "async"
    assert (d / "pos_vel_acc.out").read_text() == fo.getvalue()
    assert (d / "time_temp_totalEnergy_kinEnergy_potEnergy.out").read_text() == fe.getvalue()
    assert (d / "pos_vel_acc.xyz").read_text() == fx.getvalue()
    assert fo.getvalue().count("\n") == nsnap * n
    final = str(tmp_path / "pos_vel_acc_final.out")
    write_final(p, v, a, final)
    p2, v2 = read.initL(np.zeros((n, 3)), np.zeros((n, 3)), final)
    assert np.array_equal(p2, p.cpu().numpy()) and np.array_equal(v2, v.cpu().numpy())
    allq = torch.cat(nqts)
    save_n_q_t(allq, str(tmp_path / "n_q_t_1"))
    back = np.load(str(tmp_path / "n_q_t_1.npy"))
    assert back.shape == (nsnap * snap_int, 6, 3) and back.dtype == np.complex128
    assert np.array_equal(back, allq.cpu().numpy())
    e.close()
    assert os.path.getsize(d / "pos_vel_acc.xyz") > 0
