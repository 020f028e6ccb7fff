"""Split precision at every size (tcfft_plan_prepare_split + tcfft_exec_split):
FP32 complex in/out, natural order, 1D n = 2^8..2^27 and 2D shapes, both
directions, out of place and in place.  Inputs follow the split-precision
configuration, N(0,1) * 10^U(-3,3) per signal, so that signals a million
times apart in scale share waves and passes.  Bound: rel-L2 <= 1e-6 against
FP64 (numpy's pocketfft) for every signal (2D: every image); above 2^24 the
whole-signal FP64 transform is replaced by size-independent checks: the
forward -> inverse round trip and exact FP64 sums for sampled bins."""
import sys

import numpy as np
import pytest

import tcfft_amd

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-6


def note(m):
    print(m, flush=True)
    sys.stdout.flush()


def inputs(shape, seed):
    rng = np.random.default_rng(seed)
    scale = 10.0 ** rng.uniform(-3, 3, (shape[0],) + (1,) * (len(shape) - 1))
    x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * scale
    return x.astype(np.complex64)


def to_dev(x):
    return torch.from_numpy(np.ascontiguousarray(x).view(np.float32)).cuda()


def from_dev(t, shape):
    return t.cpu().numpy().view(np.complex64).astype(np.complex128).reshape(shape)


def per_signal_err(got, ref):
    b = got.shape[0]
    g, r = got.reshape(b, -1), ref.reshape(b, -1)
    return np.linalg.norm(g - r, axis=1) / np.linalg.norm(r, axis=1)


def run(plan, x, direction, inplace):
    src = to_dev(x)
    if inplace:
        plan.exec_split(src.data_ptr(), src.data_ptr(), direction)
        torch.cuda.synchronize()
        return from_dev(src, x.shape)
    dst = torch.empty_like(src)
    plan.exec_split(src.data_ptr(), dst.data_ptr(), direction)
    torch.cuda.synchronize()
    return from_dev(dst, x.shape)


SIZES_1D = [(1 << 8, 33), (1 << 9, 7), (1 << 10, 5), (1 << 11, 3), (1 << 12, 9), (1 << 13, 2), (1 << 14, 3),
            (1 << 15, 2), (1 << 16, 3), (1 << 17, 2), (1 << 18, 3), (1 << 19, 1), (1 << 20, 2), (1 << 21, 1),
            (1 << 22, 1), (1 << 24, 1)]


@pytest.mark.parametrize("n,batch", SIZES_1D, ids=[f"n2^{n.bit_length() - 1}-b{b}" for n, b in SIZES_1D])
def test_split_1d_vs_fp64(n, batch):
    note(f"[gpu] split 1d n={n} batch={batch}")
    plan = tcfft_amd.Plan1D(n, batch)
    plan.prepare_split()
    x = inputs((batch, n), n + batch)
    x64 = x.astype(np.complex128)
    for direction, inplace in ((tcfft_amd.FORWARD, False), (tcfft_amd.INVERSE, True)):
        got = run(plan, x, direction, inplace)
        ref = np.fft.fft(x64, axis=-1) if direction < 0 else np.fft.ifft(x64, axis=-1) * n
        errs = per_signal_err(got, ref)
        note(f"  dir={direction} inplace={inplace} per-signal rel_l2 max={errs.max():.3e}")
        assert errs.max() <= TOL
    plan.destroy()


SHAPES_2D = [(2, 256, 5), (4, 512, 3), (8, 256, 2), (16, 256, 2), (256, 256, 3), (512, 1024, 1),
             (1024, 1024, 2), (4096, 256, 1), (32, 8192, 1)]


@pytest.mark.parametrize("nx,ny,batch", SHAPES_2D, ids=[f"{a}x{b}-b{c}" for a, b, c in SHAPES_2D])
def test_split_2d_vs_fp64(nx, ny, batch):
    note(f"[gpu] split 2d {nx}x{ny} batch={batch}")
    plan = tcfft_amd.Plan2D(nx, ny, batch)
    plan.prepare_split()
    x = inputs((batch, nx, ny), nx * 7 + ny)
    x64 = x.astype(np.complex128)
    for direction, inplace in ((tcfft_amd.FORWARD, True), (tcfft_amd.INVERSE, False)):
        got = run(plan, x, direction, inplace)
        ref = np.fft.fft2(x64) if direction < 0 else np.fft.ifft2(x64) * (nx * ny)
        errs = per_signal_err(got, ref)
        note(f"  dir={direction} inplace={inplace} per-image rel_l2 max={errs.max():.3e}")
        assert errs.max() <= TOL
    plan.destroy()


@pytest.mark.parametrize("k", [25, 27])
def test_split_largest_round_trip_and_bins(k):
    """n = 2^25 and 2^27, one signal: inverse(forward(x)) = n x within 2e-6,
    and 4 sampled bins of the forward transform against exact FP64 sums."""
    n = 1 << k
    note(f"[gpu] split 1d n=2^{k} round trip + sampled bins")
    plan = tcfft_amd.Plan1D(n, 1)
    plan.prepare_split()
    rng = np.random.default_rng(k)
    x = (rng.standard_normal(n, dtype=np.float32) + 1j * rng.standard_normal(n, dtype=np.float32)).astype(np.complex64)
    src = to_dev(x)
    spec = torch.empty_like(src)
    plan.exec_split(src.data_ptr(), spec.data_ptr(), tcfft_amd.FORWARD)
    back = torch.empty_like(src)
    plan.exec_split(spec.data_ptr(), back.data_ptr(), tcfft_amd.INVERSE)
    torch.cuda.synchronize()
    X = spec.cpu().numpy().view(np.complex64).astype(np.complex128)
    y = back.cpu().numpy().view(np.complex64).astype(np.complex128) / n
    x64 = x.astype(np.complex128)
    rt = np.linalg.norm(y - x64) / np.linalg.norm(x64)
    note(f"  round trip rel_l2={rt:.3e}")
    assert rt <= 2e-6
    j = np.arange(n, dtype=np.int64)
    bins = [0, 1, n // 2 + 3, int(rng.integers(0, n))]
    worst = 0.0
    scale = np.sqrt(n) * np.sqrt(np.mean(np.abs(x64) ** 2))  # rms bin magnitude
    for b in bins:
        ph = np.exp(-2j * np.pi * ((j * int(b)) % n) / n)
        ref = np.sum(x64 * ph)
        worst = max(worst, abs(X[b] - ref) / scale)
    note(f"  sampled bins: worst |err| / rms bin = {worst:.3e}")
    assert worst <= 1e-5
    plan.destroy()


@pytest.mark.parametrize("n", [256, 4096, 1 << 17])
def test_split_tiny_and_mixed_scales(n):
    """Signals below 2^-114 (the clamped unit scale, whose inverse must still
    undo it exactly) beside signals of magnitude 1 and 1e30: every unit of 16
    columns takes its own scale, so each signal keeps the same bound; at
    n = 256 a wave's 16 units are 16 different signals."""
    note(f"[gpu] split tiny/mixed scales n={n}")
    rng = np.random.default_rng(n + 7)
    scales = np.array([1e-36, 3e-37, 1.0, 1e30, 5e-36, 1e-3] * 3)[:, None]
    b = scales.shape[0]
    x = ((rng.standard_normal((b, n)) + 1j * rng.standard_normal((b, n))) * scales).astype(np.complex64)
    plan = tcfft_amd.Plan1D(n, b)
    plan.prepare_split()
    for direction in (tcfft_amd.FORWARD, tcfft_amd.INVERSE):
        got = run(plan, x, direction, False)
        x64 = x.astype(np.complex128)
        ref = np.fft.fft(x64, axis=-1) if direction < 0 else np.fft.ifft(x64, axis=-1) * n
        errs = per_signal_err(got, ref)
        note(f"  dir={direction} per-signal rel_l2 max={errs.max():.3e}")
        assert errs.max() <= TOL
    plan.destroy()


def test_split_prepared_65536_in_place_and_identical():
    """The 65536-point kernel pair on a prepared plan: in place through the
    workspace, bit-identical to the out-of-place call."""
    n, batch = 65536, 5
    plan = tcfft_amd.Plan1D(n, batch)
    plan.prepare_split()
    x = inputs((batch, n), 99)
    a = run(plan, x, tcfft_amd.FORWARD, False)
    b = run(plan, x, tcfft_amd.FORWARD, True)
    np.testing.assert_array_equal(a, b)
    errs = per_signal_err(a, np.fft.fft(x.astype(np.complex128), axis=-1))
    assert errs.max() <= TOL
    plan.destroy()


def test_split_in_place_matches_out_of_place_bits():
    for plan, shape in ((tcfft_amd.Plan1D(1 << 20, 2), (2, 1 << 20)), (tcfft_amd.Plan2D(512, 1024, 1), (1, 512, 1024)),
                        (tcfft_amd.Plan1D(4096, 7), (7, 4096))):
        plan.prepare_split()
        x = inputs(shape, 5)
        for d in (tcfft_amd.FORWARD, tcfft_amd.INVERSE):
            np.testing.assert_array_equal(run(plan, x, d, False), run(plan, x, d, True))
        plan.destroy()


def test_split_unprepared_and_bad_arguments():
    p = tcfft_amd.Plan1D(4096, 1)
    t = torch.zeros(4096 * 2, dtype=torch.float32, device="cuda")
    u = torch.zeros(4096 * 2, dtype=torch.float32, device="cuda")
    with pytest.raises(tcfft_amd.TcfftError):
        p.exec_split(t.data_ptr(), u.data_ptr())  # not prepared
    p.prepare_split()
    p.prepare_split()  # idempotent
    p.exec_split(t.data_ptr(), u.data_ptr())
    big = torch.zeros(4096 * 2 + 64, dtype=torch.float32, device="cuda")
    with pytest.raises(tcfft_amd.TcfftError):
        p.exec_split(big.data_ptr(), big.data_ptr() + 64)  # partial overlap
    with pytest.raises(tcfft_amd.TcfftError):
        p.exec_split(t.data_ptr() + 4, u.data_ptr())  # misaligned
    p.destroy()
    r = tcfft_amd.PlanR2C(8192, 1)
    with pytest.raises(tcfft_amd.TcfftError):
        r.prepare_split()
    r.destroy()
