"""The reference's test harness rebuilt over this library and over hipFFT
(tests/cpp/: tcfft_half_*, hipfft_half_*, one main per mode over the
setup/doit/finalize seam of doit.h:2-4).  These tests run the accuracy half
of the reference's test.sh (test.sh:25-66) unchanged apart from the binary
prefix (cufft_ -> hipfft_): the same sizes, batches and flags, inputs from
the reference's glibc generator (U[0, 1e-4], doit.h:7-19), the FP64
checker's transform standing in for FFTW double (accuracy.cpp:11-23).

Each binary prints the reference's statistic (accuracy.cpp:25-41) on stdout
and rel-L2 vs FP64 on stderr.  The reference asserts nothing; the bounds
here are stated per backend:
  tcfft : rel-L2 <= 1e-3 for n <= 2^16 (the stated FP16 bound), <= 1.5e-3
          above (one extra FP16 rounding per pass); statistic < 0.02
  hipfft: rel-L2 <= 4e-3, statistic < 0.05 (the vendor's FP16 path, checked
          so that the comparator line in the speed tables is a correct FFT)
Many of these inputs are fp16 subnormals, so the input quantisation of the
reference's flow is part of what is measured."""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_build")

# test.sh:25-46 (1D accuracy) and test.sh:48-66 (2D accuracy)
ACC_1D = [(256, 32), (512, 32), (1024, 32), (131072, 1), (262144, 1), (524288, 1),
          (16777216, 1), (33554432, 1), (67108864, 1), (134217728, 1)]
ACC_2D = [(x, y) for x in (256, 512) for y in (256, 512, 1024)]
BOUND = {"tcfft": (1e-3, 1.5e-3, 0.02), "hipfft": (4e-3, 4e-3, 0.05)}


def run(args, timeout=600):
    print("[gpu] " + " ".join(args), flush=True)
    res = subprocess.run(args, capture_output=True, text=True, timeout=timeout)
    print(res.stdout.strip(), "|", res.stderr.strip(), flush=True)
    assert res.returncode == 0, res.stderr
    return res.stdout, res.stderr


def accuracy(cmd):
    out, err = run(cmd)
    stat = float(out.strip().splitlines()[-1])
    rel = float(re.search(r"rel_l2 ([0-9.e+-]+)", err).group(1))
    return stat, rel


@pytest.mark.parametrize("backend", ["tcfft", "hipfft"])
@pytest.mark.parametrize("n,batch", ACC_1D)
def test_accuracy_1d(backend, n, batch):
    cmd = [os.path.join(BIN, f"{backend}_half_accuracy")]
    if batch != 1:
        cmd += ["-b", str(batch)]
    stat, rel = accuracy(cmd + ["-n", str(n)])
    small, large, stat_bound = BOUND[backend]
    assert rel <= (small if n <= 65536 else large), rel
    assert stat < stat_bound, stat


@pytest.mark.parametrize("backend", ["tcfft", "hipfft"])
@pytest.mark.parametrize("nx,ny", ACC_2D)
def test_accuracy_2d(backend, nx, ny):
    stat, rel = accuracy([os.path.join(BIN, f"{backend}_half_2d_accuracy"), "-b", "1", "-x", str(nx), "-y", str(ny)])
    small, _, stat_bound = BOUND[backend]
    assert rel <= small, rel
    assert stat < stat_bound, stat


def test_accuracy_seed_and_batch_flags():
    """-s changes the inputs (a different statistic), -b covers several signals."""
    a = accuracy([os.path.join(BIN, "tcfft_half_accuracy"), "-n", "4096", "-b", "3", "-s", "7"])
    b = accuracy([os.path.join(BIN, "tcfft_half_accuracy"), "-n", "4096", "-b", "3", "-s", "8"])
    assert a != b
    assert max(a[1], b[1]) <= 1e-3


@pytest.mark.parametrize("backend", ["tcfft", "hipfft"])
def test_speed_lines(backend):
    out, _ = run([os.path.join(BIN, f"{backend}_half_speed"), "-n", "65536", "-b", "256", "-t", "0.2", "-m", "4096"])
    assert re.fullmatch(r"n: 65536, n_batch: 256, iter: \d+, time per iter: [0-9.]+", out.strip())
    out, _ = run([os.path.join(BIN, f"{backend}_half_2d_speed"), "-x", "1024", "-y", "1024", "-b", "8", "-t", "0.2",
                  "-m", "4096"])
    assert re.fullmatch(r"nx: 1024, ny: 1024, n_batch: 8, iter: \d+, time per iter: [0-9.]+", out.strip())


def test_unsupported_size_fails_loudly():
    """A size the library cannot plan is an error exit, not a silently wrong
    result (the reference runs a wrong plan for unlisted sizes)."""
    res = subprocess.run([os.path.join(BIN, "tcfft_half_accuracy"), "-n", "100"], capture_output=True, text=True,
                         timeout=60)
    assert res.returncode != 0
    assert "tcfftCreate" in res.stderr
