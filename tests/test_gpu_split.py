"""Split-precision path (tcfft_exec_split): FP32 complex in/out, FP16 hi+lo on
MFMA.  Inputs follow the configured distribution, N(0,1) * 10^U(-3,3) per
signal.  Bound: rel-L2 <= 1e-6 against FP64 for every signal (the stated
split-precision target), checked per signal so large signals cannot hide
small ones."""
import sys

import numpy as np
import pytest

import oracle
import tcfft_amd

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

N = 65536
TOL = 1e-6


def note(m):
    print(m, flush=True)
    sys.stdout.flush()


def inputs(batch, seed):
    rng = np.random.default_rng(seed)
    scale = 10.0 ** rng.uniform(-3, 3, (batch, 1))
    x = (rng.standard_normal((batch, N)) + 1j * rng.standard_normal((batch, N))) * scale
    return x.astype(np.complex64)


def run_split(x, direction, chunk=None):
    batch = x.shape[0]
    plan = tcfft_amd.Plan1D(N, batch)
    if chunk is not None:
        plan.set_option(tcfft_amd.OPT_SPLIT_CHUNK, chunk)
    src = torch.from_numpy(np.ascontiguousarray(x).view(np.float32)).cuda()
    dst = torch.empty_like(src)
    plan.exec_split(src.data_ptr(), dst.data_ptr(), direction)
    torch.cuda.synchronize()
    out = dst.cpu().numpy().view(np.complex64).astype(np.complex128)
    plan.destroy()
    return out


def per_signal_err(got, ref):
    return [oracle.rel_l2(got[i], ref[i]) for i in range(got.shape[0])]


@pytest.mark.parametrize("direction,sign", [(tcfft_amd.FORWARD, -1), (tcfft_amd.INVERSE, +1)])
def test_split_vs_fp64(direction, sign):
    note(f"[gpu] split n=65536 batch=6 dir={direction}")
    x = inputs(6, 1 if sign < 0 else 2)
    got = run_split(x, direction)
    ref = oracle.fft(x.astype(np.complex128), sign)
    errs = per_signal_err(got, ref)
    note(f"  per-signal rel_l2 max={max(errs):.3e} mean={np.mean(errs):.3e}")
    assert max(errs) <= TOL


def test_split_matches_numpy_and_chunking_is_exact():
    note("[gpu] split chunking")
    x = inputs(20, 3)
    a = run_split(x, tcfft_amd.FORWARD, chunk=0)
    b = run_split(x, tcfft_amd.FORWARD, chunk=7)
    np.testing.assert_array_equal(a, b)
    ref = np.fft.fft(x.astype(np.complex128), axis=-1)
    assert max(per_signal_err(a, ref)) <= TOL


def test_split_rejects_bad_arguments():
    plan = tcfft_amd.Plan1D(N, 1)
    t = torch.zeros(N * 2, dtype=torch.float32, device="cuda")
    with pytest.raises(tcfft_amd.TcfftError):
        plan.exec_split(t.data_ptr(), t.data_ptr())  # in == out
    plan.destroy()
    p2 = tcfft_amd.Plan1D(4096, 1)
    u = torch.zeros(4096 * 2, dtype=torch.float32, device="cuda")
    with pytest.raises(tcfft_amd.TcfftError):
        p2.exec_split(t.data_ptr(), u.data_ptr())  # size not supported
    p2.destroy()


def test_split_start_delay_bit_identical():
    """The split passes' start schedule (TCFFT_OPT_SPLIT_DELAY) never changes
    results; out-of-range values are rejected."""
    rng = np.random.default_rng(41)
    n, batch = 1 << 16, 24
    x = torch.from_numpy(rng.standard_normal((batch, n, 2)).astype(np.float32)).cuda()
    plan = tcfft_amd.Plan1D(n, batch)
    plan.set_option(tcfft_amd.OPT_SPLIT_DELAY, 0)
    ref = torch.empty_like(x)
    plan.exec_split(x.data_ptr(), ref.data_ptr())
    for d in (1, 2, 5, 64):
        plan.set_option(tcfft_amd.OPT_SPLIT_DELAY, d)
        out = torch.empty_like(x)
        plan.exec_split(x.data_ptr(), out.data_ptr())
        torch.cuda.synchronize()
        assert torch.equal(out, ref), d
    for bad in (-1, 65):
        with pytest.raises(tcfft_amd.TcfftError):
            plan.set_option(tcfft_amd.OPT_SPLIT_DELAY, bad)
    plan.destroy()


def test_split_tiny_inputs():
    """Signals whose largest component is below 2^-114 take the clamped wave
    scale (2^126); scale and inverse must still multiply to exactly 1, so
    the result keeps the same per-signal bound as ordinary inputs."""
    note("[gpu] split tiny inputs n=65536 batch=4")
    rng = np.random.default_rng(43)
    scale = np.array([1e-36, 3e-37, 5e-36, 1.0])[:, None]
    x = ((rng.standard_normal((4, N)) + 1j * rng.standard_normal((4, N))) * scale).astype(np.complex64)
    assert np.abs(x[:3]).max() < 2.0 ** -114
    for direction, sign in ((tcfft_amd.FORWARD, -1), (tcfft_amd.INVERSE, +1)):
        got = run_split(x, direction)
        errs = per_signal_err(got, oracle.fft(x.astype(np.complex128), sign))
        note(f"  dir={direction} per-signal rel_l2 {[f'{e:.2e}' for e in errs]}")
        assert max(errs) <= TOL

