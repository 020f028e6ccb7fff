"""Natural-order DIF transform (tcfft_exec_dif) and the fused FFT convolution
(tcfft_conv_exec, tcfft_conv_filter_grad) against numpy's FP64 FFT.

Spectra come back in the drop-in's digit-reversed order: position p holds
X[rev[p]], rev = gen_rev(n, plan.radices).  Tolerance: rel-L2 <= 1e-3 (fp16
storage) for transforms, 2e-3 for the convolution (two transforms)."""
import sys

import numpy as np
import pytest

import tcfft_amd

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def note(m):
    print(m, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def to_dev(z):
    a = np.ascontiguousarray(np.stack([z.real, z.imag], axis=-1).astype(np.float16))
    t = torch.from_numpy(a).cuda()
    assert t.is_contiguous()
    return t


def from_dev(t):
    v = t.float().cpu().numpy().astype(np.float64)
    return v[..., 0] + 1j * v[..., 1]


def fp16c(rng, shape, scale=1.0):
    re = (rng.uniform(-1, 1, shape) * scale).astype(np.float16).astype(np.float64)
    im = (rng.uniform(-1, 1, shape) * scale).astype(np.float16).astype(np.float64)
    return re + 1j * im


@pytest.mark.parametrize("n,batch,direction", [(4096, 3, -1), (8192, 2, -1), (8192, 2, 1), (16384, 2, -1),
                                               (4096, 1, 1)])
def test_dif_natural_to_digit_reversed(n, batch, direction):
    note(f"[gpu] dif n={n} batch={batch} dir={direction}")
    rng = np.random.default_rng(n + batch)
    x = fp16c(rng, (batch, n))
    plan = tcfft_amd.Plan1D(n, batch)
    rev = tcfft_amd.gen_rev(n, plan.radices)
    src = to_dev(x)
    dst = torch.empty_like(src)
    plan.exec_dif(src.data_ptr(), dst.data_ptr(), direction)
    torch.cuda.synchronize()
    got = from_dev(dst)
    ref = np.fft.fft(x, axis=-1) if direction < 0 else np.fft.ifft(x, axis=-1) * n
    err = rel(got, ref[:, rev])
    note(f"  rel_l2={err:.3e}")
    assert err <= 1e-3
    # in place gives the same bits
    plan.exec_dif(src.data_ptr(), src.data_ptr(), direction)
    torch.cuda.synchronize()
    assert torch.equal(src, dst)
    # and the drop-in exec inverts it back to natural order (round trip)
    plan.exec(dst.data_ptr(), -direction)
    torch.cuda.synchronize()
    assert rel(from_dev(dst) / n, x) <= 2e-3
    plan.destroy()


def conv_ref(x, h):
    return np.fft.ifft(np.fft.fft(x, axis=-1) * h, axis=-1)


@pytest.mark.parametrize("n,B,C", [(4096, 2, 3), (8192, 3, 4), (16384, 2, 2)])
def test_conv_forward_and_conj(n, B, C):
    note(f"[gpu] conv n={n} B={B} C={C}")
    rng = np.random.default_rng(7 * n + C)
    x = fp16c(rng, (B, C, n))
    h = (rng.standard_normal((C, n)) + 1j * rng.standard_normal((C, n))) * 0.02
    plan = tcfft_amd.Plan1D(n, B * C)
    rev = tcfft_amd.gen_rev(n, plan.radices)
    h16 = h.astype(np.complex64).astype(np.complex128)
    hrev = to_dev(h16[:, rev])
    hq = from_dev(hrev)[:, np.argsort(rev)]  # the fp16 filter the kernel sees, natural order
    xd = to_dev(x)
    y = torch.empty_like(xd)
    spec = torch.empty_like(xd)
    plan.conv_exec(C, xd.data_ptr(), hrev.data_ptr(), y.data_ptr(), False, spec.data_ptr())
    torch.cuda.synchronize()
    err = rel(from_dev(y), conv_ref(x, hq[None]))
    note(f"  forward rel_l2={err:.3e}")
    assert err <= 2e-3
    err_s = rel(from_dev(spec), np.fft.fft(x, axis=-1)[..., rev])
    note(f"  spectrum rel_l2={err_s:.3e}")
    assert err_s <= 1e-3
    plan.conv_exec(C, xd.data_ptr(), hrev.data_ptr(), y.data_ptr(), True)
    torch.cuda.synchronize()
    err_c = rel(from_dev(y), conv_ref(x, np.conj(hq)[None]))
    note(f"  conj rel_l2={err_c:.3e}")
    assert err_c <= 2e-3
    # in place
    plan.conv_exec(C, xd.data_ptr(), hrev.data_ptr(), xd.data_ptr(), True)
    torch.cuda.synchronize()
    assert torch.equal(xd, y)
    plan.destroy()


def test_conv_filter_grad():
    n, B, C = 8192, 5, 3
    note(f"[gpu] conv filter grad n={n} B={B} C={C}")
    rng = np.random.default_rng(11)
    x = fp16c(rng, (B, C, n))
    g = fp16c(rng, (B, C, n))
    plan = tcfft_amd.Plan1D(n, B * C)
    rev = tcfft_amd.gen_rev(n, plan.radices)
    xs = torch.empty(B, C, n, 2, dtype=torch.float16, device="cuda")
    gs = torch.empty_like(xs)
    plan.exec_dif(to_dev(x).data_ptr(), xs.data_ptr())
    plan.exec_dif(to_dev(g).data_ptr(), gs.data_ptr())
    gh = torch.empty(C, n, 2, dtype=torch.float32, device="cuda")
    plan.conv_filter_grad(C, xs.data_ptr(), gs.data_ptr(), gh.data_ptr())
    torch.cuda.synchronize()
    got = gh.double().cpu().numpy()
    got = (got[..., 0] + 1j * got[..., 1])[:, np.argsort(rev)]
    ref = (np.conj(np.fft.fft(x, axis=-1)) * np.fft.fft(g, axis=-1)).sum(0) / n
    err = rel(got, ref)
    note(f"  rel_l2={err:.3e}")
    assert err <= 2e-3
    plan.destroy()


def test_conv_rejects_bad_arguments():
    plan = tcfft_amd.Plan1D(65536, 2)
    t = torch.zeros(65536 * 2, 2, dtype=torch.float16, device="cuda")
    with pytest.raises(tcfft_amd.TcfftError):
        plan.exec_dif(t.data_ptr(), t.data_ptr())
    plan.destroy()
    p2 = tcfft_amd.Plan1D(8192, 6)
    u = torch.zeros(8192 * 6, 2, dtype=torch.float16, device="cuda")
    with pytest.raises(tcfft_amd.TcfftError):
        p2.conv_exec(4, u.data_ptr(), u.data_ptr(), u.data_ptr())  # batch 6 not a multiple of 4 channels
    with pytest.raises(tcfft_amd.TcfftError):
        p2.conv_exec(3, u.data_ptr(), u.data_ptr(), u.data_ptr(), False, u.data_ptr())  # spec aliases x
    p2.destroy()
