"""The drop-in headers used the way the reference's harness uses tcfft_half.h:
host gen_rev by plan.radices, half staging, tcfftExec in place, sync
(tests/cpp/cli_1d.cpp, cli_2d.cpp).  Inputs are the reference's own
generator (glibc rand, U[0, 1e-4], doit.h:7-19); the FP64 standard is the
checker's transform of the same double data, as FFTW double is in
accuracy.cpp.  Both the reference's statistic (accuracy.cpp:25-41) and
rel-L2 are recorded; the bound is on rel-L2 (the reference asserts nothing).
A large share of these inputs are fp16 subnormals, so this also exercises
the input quantisation the reference's flow has."""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_build")
# rel-L2 vs FP64 of the unquantised inputs: the fp16 input rounding of values
# in [0, 1e-4] (many subnormal) dominates; bound stated here.
REL_L2_BOUND = 1e-3


def run(args):
    print("[gpu] " + " ".join(args), flush=True)
    res = subprocess.run(args, capture_output=True, text=True, timeout=240)
    print(res.stdout, res.stderr, flush=True)
    assert res.returncode == 0, res.stderr
    return res.stdout


def parse(out):
    rows = []
    lines = out.strip().splitlines()
    for i, l in enumerate(lines):
        if l.startswith("Accuracy test"):
            err = float(lines[i + 1])
            rel = float(lines[i + 2].split()[1])
            rows.append((l, err, rel))
    return rows


def test_cli_1d_sweep():
    rows = parse(run([os.path.join(BIN, "cli_1d"), "sweep"]))
    assert len(rows) == 6
    for label, err, rel in rows:
        assert rel <= REL_L2_BOUND, (label, rel)
        assert err < 0.02, (label, err)


def test_cli_2d_sweep():
    rows = parse(run([os.path.join(BIN, "cli_2d"), "sweep"]))
    assert len(rows) == 7
    for label, err, rel in rows:
        assert rel <= REL_L2_BOUND, (label, rel)
        assert err < 0.02, (label, err)


def test_cli_speed_lines():
    out = run([os.path.join(BIN, "cli_1d"), "speed", "-n", "65536", "-b", "256", "-t", "0.2", "-m", "4096"])
    assert re.match(r"n: 65536, n_batch: 256, iter: \d+, time per iter: [0-9.]+", out.strip())
    out = run([os.path.join(BIN, "cli_2d"), "speed", "-x", "1024", "-y", "1024", "-b", "8", "-t", "0.2", "-m", "4096"])
    assert re.match(r"nx: 1024, ny: 1024, n_batch: 8, iter: \d+, time per iter: [0-9.]+", out.strip())
