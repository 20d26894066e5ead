"""Text formats of the trajectory writer and the resume reader (CPU).

The reference writes with numpy.savetxt's defaults (yukawa_MD_p3m_ord6.py:
224-231, 241-242: '%.18e', one space, '\\n') and reads back with
read.initL / read.init0 (read.py:3-74).  Pinned here by the literal text of
known rows (not by calling savetxt again), the .xyz frame header, the
column order of both readers, the first-N-rows rule, and an exact round
trip of random doubles through the text.
"""
import io

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from mdforce import read  # noqa: E402
from mdforce.trajectory import write_snapshot_text  # noqa: E402


def test_snapshot_text_literal():
    irp = np.array([[1.0, -2.5, 0.1, 0.0, 3.0, -0.0, 1e-300, 6.02e23, -7.0],
                    [0.5, 0.25, 0.125, 1.0, 2.0, 3.0, 4.0, 5.0, 6.0]])
    row5 = [0.1 * 9, 0.01, -380.5, 12.25, -392.75]
    fo, fe, fx = io.StringIO(), io.StringIO(), io.StringIO()
    write_snapshot_text(fo, fe, fx, irp, row5)
    first = fo.getvalue().split("\n")[0]
    assert first == ("1.000000000000000000e+00 -2.500000000000000000e+00 1.000000000000000056e-01 "
                     "0.000000000000000000e+00 3.000000000000000000e+00 -0.000000000000000000e+00 "
                     "1.000000000000000025e-300 6.019999999999999958e+23 -7.000000000000000000e+00")
    assert fo.getvalue().count("\n") == 2
    assert fe.getvalue() == ("9.000000000000000222e-01 1.000000000000000021e-02 -3.805000000000000000e+02 "
                             "1.225000000000000000e+01 -3.927500000000000000e+02\n")
    lines = fx.getvalue().split("\n")
    assert lines[0] == "2" and lines[1] == "x y z vx vy vz ax ay az"
    assert lines[2] == first and len(lines) == 5 and lines[4] == ""


def test_round_trip_and_column_order(tmp_path):
    rng = np.random.default_rng(0)
    N = 50
    pos, vel, acc = (rng.standard_normal((N + 3, 3)) * s for s in (30.0, 0.1, 1e-3))
    f = tmp_path / "pos_vel_acc_final.out"
    np.savetxt(f, np.hstack((pos, vel, acc)))
    p, v = read.initL(np.zeros((N, 3)), np.zeros((N, 3)), str(f))
    assert p.shape == (N, 3) and np.array_equal(p, pos[:N]) and np.array_equal(v, vel[:N])
    # init0: velocities in the first three columns
    g = tmp_path / "old_layout.out"
    np.savetxt(g, np.hstack((vel, pos, acc)))
    p0, v0 = read.init0(np.zeros((N, 3)), np.zeros((N, 3)), str(g))
    assert np.array_equal(p0, pos[:N]) and np.array_equal(v0, vel[:N])
