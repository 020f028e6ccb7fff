"""GPU parity of the HIP path against the FP64 oracle (through the C ABI).

Tolerance (plain FP16 path, fp16-exact uniform[-1,1) inputs):
  rel-L2 <= 1e-3 for n <= 2^16 (the stated FP16 bound), <= 1.5e-3 above.
"""
import os
import sys

import numpy as np
import pytest

import oracle
import tcfft_amd

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-3
TOL_BIG = 1.5e-3


def tol_for(n):
    return TOL if n <= (1 << 16) else TOL_BIG


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def fp16_input(shape, seed):
    rng = np.random.default_rng(seed)
    re = rng.uniform(-1, 1, shape).astype(np.float16)
    im = rng.uniform(-1, 1, shape).astype(np.float16)
    return re.astype(np.float64) + 1j * im.astype(np.float64)


def to_device(x):
    """complex128 numpy -> interleaved fp16 cuda tensor (exact for fp16 values)."""
    h = np.ascontiguousarray(np.stack([x.real, x.imag], axis=-1).astype(np.float16))
    t = torch.from_numpy(h).cuda()
    assert t.is_contiguous()  # raw pointers are handed to the C ABI
    return t


def from_device(t):
    h = t.cpu().numpy().astype(np.float64)
    return h[..., 0] + 1j * h[..., 1]


def run_1d(x, direction=tcfft_amd.FORWARD, mode="dropin", inplace=True):
    b, n = x.shape
    note(f"[gpu] 1d n={n} batch={b} dir={direction} mode={mode} inplace={inplace}")
    plan = tcfft_amd.Plan1D(n, b)
    if mode == "dropin":
        rev = tcfft_amd.gen_rev(n, plan.radices)
        d = to_device(x[:, rev])
        plan.exec(d.data_ptr(), direction)
        torch.cuda.synchronize()
        out = from_device(d)
    else:
        src = to_device(x)
        dst = src if inplace else torch.empty_like(src)
        plan.exec_natural(src.data_ptr(), dst.data_ptr(), direction)
        torch.cuda.synchronize()
        out = from_device(dst)
    plan.destroy()
    return out


def run_2d(x, direction=tcfft_amd.FORWARD, mode="dropin", inplace=True):
    b, nx, ny = x.shape
    note(f"[gpu] 2d nx={nx} ny={ny} batch={b} dir={direction} mode={mode} inplace={inplace}")
    plan = tcfft_amd.Plan2D(nx, ny, b)
    if mode == "dropin":
        rx = tcfft_amd.gen_rev(nx, plan.radices_x)
        ry = tcfft_amd.gen_rev(ny, plan.radices_y)
        d = to_device(x[:, rx][:, :, ry])
        plan.exec(d.data_ptr(), direction)
    else:
        src = to_device(x)
        d = src if inplace else torch.empty_like(src)
        plan.exec_natural(src.data_ptr(), d.data_ptr(), direction)
    torch.cuda.synchronize()
    out = from_device(d)
    plan.destroy()
    return out


@pytest.mark.parametrize("n,batch", [(256, 1), (256, 37), (512, 3), (1024, 32), (2048, 5), (4096, 2), (4096, 17)])
def test_small_forward_vs_oracle(n, batch):
    x = fp16_input((batch, n), seed=n + batch)
    got = run_1d(x)
    err = oracle.rel_l2(got, oracle.fft(x))
    note(f"  rel_l2={err:.3e}")
    assert err <= TOL


@pytest.mark.parametrize("n,batch", [(8192, 2), (16384, 1), (32768, 2), (65536, 3), (1 << 17, 1), (1 << 18, 1)])
def test_multipass_forward_vs_oracle(n, batch):
    x = fp16_input((batch, n), seed=n)
    got = run_1d(x)
    err = oracle.rel_l2(got, oracle.fft(x))
    note(f"  rel_l2={err:.3e}")
    assert err <= tol_for(n)


@pytest.mark.parametrize("n,batch,direction", [(1 << 19, 1, tcfft_amd.FORWARD), (1 << 20, 2, tcfft_amd.FORWARD),
                                               (1 << 20, 1, tcfft_amd.INVERSE), (1 << 21, 1, tcfft_amd.FORWARD),
                                               (1 << 22, 1, tcfft_amd.FORWARD), (1 << 23, 1, tcfft_amd.INVERSE),
                                               (1 << 24, 1, tcfft_amd.FORWARD), (1 << 25, 1, tcfft_amd.INVERSE),
                                               (1 << 26, 1, tcfft_amd.FORWARD), (1 << 27, 1, tcfft_amd.INVERSE)])
def test_large_vs_numpy(n, batch, direction):
    """Sizes above 2^18 (every radix of the pass after the 65536-point
    first pass, including the single-pass radices 512..2048 whose twiddle the
    first pass folds over three digits) against numpy's FP64 FFT; the oracle
    is pinned to numpy at smaller sizes by the CPU tests."""
    x = fp16_input((batch, n), seed=n + 11)
    got = run_1d(x, direction)
    ref = np.fft.fft(x, axis=-1) if direction == tcfft_amd.FORWARD else np.fft.ifft(x, axis=-1) * n
    err = oracle.rel_l2(got, ref)
    note(f"  rel_l2={err:.3e}")
    assert err <= tol_for(n)


@pytest.mark.parametrize("n,batch", [(256, 4), (1024, 2), (4096, 3), (65536, 1), (1 << 17, 1)])
def test_inverse_vs_oracle(n, batch):
    x = fp16_input((batch, n), seed=3 * n)
    got = run_1d(x, tcfft_amd.INVERSE)
    err = oracle.rel_l2(got, oracle.fft(x, +1))
    note(f"  rel_l2={err:.3e}")
    assert err <= tol_for(n)


@pytest.mark.parametrize("n,batch,inplace", [(256, 3, False), (4096, 2, True), (65536, 1, False), (1 << 17, 1, True)])
def test_natural_order_vs_oracle(n, batch, inplace):
    x = fp16_input((batch, n), seed=5 * n)
    got = run_1d(x, mode="natural", inplace=inplace)
    err = oracle.rel_l2(got, oracle.fft(x))
    note(f"  rel_l2={err:.3e}")
    assert err <= tol_for(n)


# natural-order input gives the drop-in path's result bit for bit: lines of
# up to 4096 points do the digit reversal in the first pass's load, longer
# ones through the LDS-tiled reversal kernel before the passes
@pytest.mark.parametrize("n,batch,inplace,direction", [
    (256, 3, False, tcfft_amd.FORWARD), (256, 37, True, tcfft_amd.INVERSE), (256, 65, False, tcfft_amd.FORWARD),
    (512, 5, True, tcfft_amd.FORWARD), (1024, 7, False, tcfft_amd.INVERSE), (2048, 9, True, tcfft_amd.FORWARD),
    (4096, 3, False, tcfft_amd.FORWARD), (4096, 17, True, tcfft_amd.INVERSE),
    (8192, 3, False, tcfft_amd.FORWARD), (16384, 2, True, tcfft_amd.INVERSE), (32768, 3, False, tcfft_amd.FORWARD),
    (65536, 2, True, tcfft_amd.FORWARD), (1 << 18, 1, False, tcfft_amd.INVERSE), (1 << 20, 2, False, tcfft_amd.FORWARD)])
def test_natural_order_equals_dropin(n, batch, inplace, direction):
    x = fp16_input((batch, n), seed=9 * n + batch)
    got = run_1d(x, direction, mode="natural", inplace=inplace)
    ref = run_1d(x, direction)
    assert np.array_equal(got, ref)
    sign = -1 if direction == tcfft_amd.FORWARD else +1
    want = (np.fft.fft(x, axis=-1) if sign < 0 else np.fft.ifft(x, axis=-1) * n) if n > 65536 else oracle.fft(x, sign)
    err = oracle.rel_l2(got, want)
    note(f"  rel_l2={err:.3e}")
    assert err <= tol_for(n)


@pytest.mark.parametrize("shape,batch,direction", [((2, 256), 3, tcfft_amd.FORWARD), ((16, 512), 2, tcfft_amd.INVERSE),
                                                   ((512, 1024), 1, tcfft_amd.FORWARD),
                                                   ((1024, 1024), 2, tcfft_amd.INVERSE),
                                                   ((256, 4096), 1, tcfft_amd.FORWARD),
                                                   ((4096, 256), 1, tcfft_amd.FORWARD),
                                                   ((4, 8192), 2, tcfft_amd.FORWARD),
                                                   ((8, 32768), 1, tcfft_amd.INVERSE)])
def test_2d_natural_out_of_place(shape, batch, direction):
    """2D natural order out of place: the y reversal in the rows pass's load,
    the x reversal as the line order of its stores; equal to the drop-in."""
    nx, ny = shape
    x = fp16_input((batch, nx, ny), seed=nx + 3 * ny)
    got = run_2d(x, direction, mode="natural", inplace=False)
    ref = run_2d(x, direction)
    assert np.array_equal(got, ref)
    sign = -1 if direction == tcfft_amd.FORWARD else +1
    err = oracle.rel_l2(got, oracle.fft_2d(x, sign))
    note(f"  rel_l2={err:.3e}")
    assert err <= TOL


@pytest.mark.parametrize("n", [256, 512, 1024, 4096])
def test_golden_vectors(golden_dir, n):
    d = np.load(os.path.join(golden_dir, f"fft1d_n{n}.npz"))
    x = d["x_re"].astype(np.float64) + 1j * d["x_im"].astype(np.float64)
    got = run_1d(x)
    err = oracle.rel_l2(got, d["forward"])
    note(f"  rel_l2={err:.3e}")
    assert err <= TOL
    if "adjoint" in d:
        assert oracle.rel_l2(run_1d(x, tcfft_amd.INVERSE), d["adjoint"]) <= TOL


# nx = 512, 1024, 2048: the column axis in one pass of radix 256*W; nx = 128,
# 256: the 64-column LDS-tile pass of radix 128 / 256
@pytest.mark.parametrize("shape,batch", [((256, 256), 2), ((512, 1024), 1), ((1024, 1024), 1), ((16, 256), 3),
                                         ((2, 512), 2), ((256, 8192), 1), ((2048, 256), 2), ((512, 4096), 1),
                                         ((4096, 256), 1), ((128, 256), 3), ((128, 4096), 1)])
def test_2d_vs_oracle(shape, batch):
    nx, ny = shape
    x = fp16_input((batch, nx, ny), seed=nx * 7 + ny)
    got = run_2d(x)
    err = oracle.rel_l2(got, oracle.fft_2d(x))
    note(f"  rel_l2={err:.3e}")
    assert err <= TOL
    got_inv = run_2d(x, tcfft_amd.INVERSE, mode="natural")
    err_inv = oracle.rel_l2(got_inv, oracle.fft_2d(x, +1))
    note(f"  inverse natural rel_l2={err_inv:.3e}")
    assert err_inv <= TOL


def test_reference_harness_flow():
    """accuracy.cpp's flow: glibc data in [0, 1e-4] (doit.h:7-19), gen_rev on
    the host, exec, compare with the FP64 transform of the same fp16 inputs."""
    n, b = 1024, 32
    data = oracle.generate_data(n, b, 42)
    x = data.real.astype(np.float16).astype(np.float64) + 1j * data.imag.astype(np.float16).astype(np.float64)
    got = run_1d(x)
    ref = oracle.fft(x)
    err = oracle.rel_l2(got, ref)
    note(f"  rel_l2={err:.3e} reference-metric={oracle.get_error(got, ref):.3e}")
    assert err <= TOL


def test_full_size_configs_against_numpy():
    """The benchmark shapes: N=2^20 x 16 and 2D 1024x1024 x 8 (plus a round trip)."""
    x = fp16_input((16, 1 << 20), seed=11)
    got = run_1d(x)
    err = oracle.rel_l2(got, np.fft.fft(x, axis=-1))
    note(f"  2^20x16 rel_l2={err:.3e}")
    assert err <= TOL_BIG
    y = fp16_input((8, 1024, 1024), seed=12)
    got2 = run_2d(y)
    err2 = oracle.rel_l2(got2, np.fft.fft2(y, axes=(-2, -1)))
    note(f"  2D 1024^2x8 rel_l2={err2:.3e}")
    assert err2 <= TOL


def test_linearity_and_round_trip_2_16():
    n, b = 1 << 16, 4
    x = fp16_input((b, n), seed=21)
    X = run_1d(x, mode="natural", inplace=False)
    # forward then adjoint, divided by n, returns the input (to fp16 accuracy)
    Xq = X.real.astype(np.float16).astype(np.float64) + 1j * X.imag.astype(np.float16).astype(np.float64)
    scale = 1.0 / 256.0  # keep the adjoint's output inside fp16 range
    back = run_1d(Xq * scale, tcfft_amd.INVERSE, mode="natural", inplace=False) / (n * scale)
    assert oracle.rel_l2(back, x) <= 2 * TOL
    # a delta at index 0 transforms to all ones
    d = np.zeros((1, n), complex)
    d[0, 0] = 1.0
    np.testing.assert_allclose(run_1d(d, mode="natural", inplace=False), np.ones((1, n)), atol=2e-3)
