"""The command-line driver (csrc/cli/mdf_bench.cpp), the stand-in for the
reference C driver's main (yukawa_p3m.c:418-560): run as a child process on
the kappa = 0 golden fixture, its printed PP accelerations (particles 0..9,
467 and the last ten, %.8lf as the reference prints them, yukawa_p3m.c:530-556)
must equal the reference C's own accelerations stored in the fixture to the
print rounding (5e-9; bar 1e-8)."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "md-force-routines_amd", "mdf_bench")


def test_cli_prints_reference_accelerations(tmp_path):
    if not os.path.isfile(EXE):
        pytest.fail("mdf_bench is not built (run __graft_entry__.build())")
    g = np.load(os.path.join(ROOT, "tests", "golden", "pp_ref_n500_k0.npz"))
    pos, acc = g["pos"], g["acc"]
    N = len(pos)
    f = tmp_path / "pos.txt"
    np.savetxt(f, pos)
    print(f"[cli] mdf_bench N={N} alpha={float(g['alpha'])} rc={float(g['rc'])}", flush=True)
    r = subprocess.run([EXE, str(N), str(f), "%.17g" % float(g["alpha"]), "%.17g" % float(g["rc"])],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().split("\n")
    vals = {}
    for line in lines:
        if line.count(",") == 3:
            i, a, b, c = line.split(",")
            vals[int(i)] = [float(a), float(b), float(c)]
    for k, line in enumerate(lines[-10:]):
        vals[N - 10 + k] = [float(x) for x in line.split()]
    assert sorted(vals) == list(range(10)) + [467] + list(range(N - 10, N))
    err = max(abs(vals[i][j] - acc[i][j]) for i in vals for j in range(3))
    print(f"[cli] 21 particles, max |printed - reference C| = {err:.2e}")
    assert err <= 1e-8
