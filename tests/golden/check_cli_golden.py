"""Runs mdf_bench (csrc/cli/mdf_bench.cpp) on the kappa = 0 golden fixture
and compares the PP accelerations it prints (particles 0..9, 467 and the last
ten, %.8lf as the reference C driver prints them, yukawa_p3m.c:530-556) with
the reference C's own accelerations stored in the fixture.  Needs a GPU;
run from the repository root as its own process (not under pytest, whose
process may already hold the GPU):

    python tests/golden/check_cli_golden.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def main():
    g = np.load(os.path.join(ROOT, "tests", "golden", "pp_ref_n500_k0.npz"))
    pos, acc = g["pos"], g["acc"]
    N = len(pos)
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, "pos.txt")
        np.savetxt(f, pos)
        r = subprocess.run([os.path.join(ROOT, "md-force-routines_amd", "mdf_bench"), str(N), f,
                            "%.17g" % float(g["alpha"]), "%.17g" % float(g["rc"])],
                           capture_output=True, text=True, timeout=120)
    print(r.stdout, end="")
    if r.returncode != 0:
        print(r.stderr, file=sys.stderr)
        return 1
    lines = r.stdout.strip().split("\n")
    vals = {}
    for line in lines:
        if line.count(",") == 3:
            i, a, b, c = line.split(",")
            vals[int(i)] = [float(a), float(b), float(c)]
    tail = [line.split() for line in lines[-10:]]
    for k, t in enumerate(tail):
        vals[N - 10 + k] = [float(x) for x in t]
    err = max(abs(vals[i][j] - acc[i][j]) for i in vals for j in range(3))
    ok = len(vals) == 21 and err <= 1e-8
    print(f"{len(vals)} particles, max |printed - reference C| = {err:.2e} (print rounding 5e-9): "
          f"{'ok' if ok else 'MISMATCH'}")
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
