"""The standalone timing program mdf_bench (csrc/cli/mdf_bench.cpp), CPU side.

Its argument handling fails before any HIP call, so these run without a
GPU: the usage message and status 2 for missing or bad arguments, and a
clear error for a positions file with too few lines.  Its output on a GPU
is checked against the reference C's PP accelerations (the kappa = 0
golden fixture) by running it on the GPU host, outside pytest.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "md-force-routines_amd", "mdf_bench")


@pytest.fixture(scope="module")
def cli():
    if not os.path.isfile(CLI):
        pytest.skip("mdf_bench is not built (python -c 'import __graft_entry__ as g; g.build()')")
    return CLI


def run(args):
    return subprocess.run(args, capture_output=True, text=True, timeout=60)


def test_usage(cli):
    r = run([cli])
    assert r.returncode == 2 and "usage: mdf_bench N positions_file Gew rc" in r.stderr
    r = run([cli, "10", "x", "0.5", "3", "0", "64", "fp16"])
    assert r.returncode == 2


def test_short_positions_file(cli, tmp_path):
    f = tmp_path / "pos.txt"
    f.write_text("0 0 0\n1 1 1\n")
    r = run([cli, "3", str(f), "0.5", "3.0"])
    assert r.returncode == 1 and "fewer than 3 x y z lines" in r.stderr
    r = run([cli, "3", str(tmp_path / "missing.txt"), "0.5", "3.0"])
    assert r.returncode == 1 and "cannot open" in r.stderr
