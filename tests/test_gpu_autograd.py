"""Differentiable FFT (tcfft_amd.torch_fft) against FP64 references.

Gradients of the fp16 HIP path are compared with (a) torch.fft's autograd in
complex128 and (b) central finite differences of the FP64 reference function.
Tolerances (relative, fp16 storage): 2e-3 on values and gradients.
"""
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from tcfft_amd import torch_fft  # noqa: E402

TOL = 2e-3


def note(m):
    print(m, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def to_c(t):
    t = t.detach().float().cpu().double()
    return (t[..., 0] + 1j * t[..., 1]).numpy()


def fp16_pair(shape, seed, scale=1.0):
    g = torch.Generator().manual_seed(seed)
    return ((torch.rand(*shape, 2, generator=g) * 2 - 1) * scale).half()


@pytest.mark.parametrize("shape", [(3, 256), (2, 2, 4096), (1, 65536), (2, 8192)])
def test_fft_and_adjoint_values(shape):
    note(f"[gpu] fft values {shape}")
    x = fp16_pair(shape, 1).cuda()
    xc = to_c(x)
    X = torch_fft.fft(x)
    assert rel(to_c(X), np.fft.fft(xc, axis=-1)) <= 1e-3
    Y = torch_fft.fft_adjoint(x)
    n = shape[-1]
    assert rel(to_c(Y), np.fft.ifft(xc, axis=-1) * n) <= 1e-3


@pytest.mark.parametrize("n", [512, 8192])
def test_rfft_c2r_values_and_gradients_vs_torch_fft(n):
    """rfft / c2r / irfft and their backward passes against torch.fft's
    autograd in float64 on the CPU, for the loss sum(out * w)."""
    note(f"[gpu] rfft/c2r autograd n={n}")
    g = torch.Generator().manual_seed(n)
    x = ((torch.rand(3, n, generator=g) * 2 - 1).half()).cuda().requires_grad_(True)
    w = torch.randn(3, n // 2 + 1, 2, generator=g)
    X = torch_fft.rfft(x)
    (X.float() * w.cuda()).sum().backward()
    xd = x.detach().double().cpu().requires_grad_(True)
    Xd = torch.view_as_real(torch.fft.rfft(xd, dim=-1))
    (Xd * w.double()).sum().backward()
    assert rel(to_c(X), torch.view_as_complex(Xd.detach()).numpy()) <= 1e-3
    assert rel(x.grad.double().cpu().numpy(), xd.grad.numpy()) <= TOL
    # c2r = n * irfft, and its gradient
    Xh = ((torch.rand(3, n // 2 + 1, 2, generator=g) * 2 - 1).half()).cuda().requires_grad_(True)
    v = torch.randn(3, n, generator=g)
    y = torch_fft.c2r(Xh, n)
    (y.float() * v.cuda()).sum().backward()
    Xd = Xh.detach().double().cpu().requires_grad_(True)
    yd = torch.fft.irfft(torch.view_as_complex(Xd), n=n, dim=-1) * n
    (yd * v.double()).sum().backward()
    assert rel(y.detach().double().cpu().numpy(), yd.detach().numpy()) <= 1e-3
    # torch's irfft ignores the imaginary parts of bins 0 and n/2: so does c2r
    gd = Xd.grad.clone()
    gd[:, 0, 1] = 0
    gd[:, -1, 1] = 0
    gg = Xh.grad.double().cpu()
    assert float(gg[:, 0, 1].abs().max()) == 0.0 and float(gg[:, -1, 1].abs().max()) == 0.0
    assert rel(to_c(gg), to_c(gd)) <= TOL
    # the normalised inverse round-trips
    back = torch_fft.irfft(torch_fft.rfft(x.detach()), n)
    assert rel(back.double().cpu().numpy(), x.detach().double().cpu().numpy()) <= 1.5e-3


@pytest.mark.parametrize("n", [512, 8192])
def test_fft_backward_matches_torch_fft(n):
    note(f"[gpu] fft backward n={n}")
    x = fp16_pair((4, n), 2).cuda().requires_grad_(True)
    w = torch.randn(4, n, 2, device="cuda", dtype=torch.float32)
    X = torch_fft.fft(x)
    loss = (X.float() * w).sum()
    loss.backward()
    # reference: complex128 torch.fft autograd of the same real functional
    xt = torch.tensor(to_c(x), dtype=torch.complex128, requires_grad=True)
    Xt = torch.fft.fft(xt, dim=-1)
    wc = w.double().cpu()
    lt = (Xt.real * wc[..., 0] + Xt.imag * wc[..., 1]).sum()
    lt.backward()
    got = to_c(x.grad)
    assert rel(got, xt.grad.numpy()) <= TOL
    # the adjoint's backward is the forward kernel
    x2 = fp16_pair((4, n), 3).cuda().requires_grad_(True)
    Y = torch_fft.fft_adjoint(x2)
    (Y.float() * w).sum().backward()
    xt2 = torch.tensor(to_c(x2), dtype=torch.complex128, requires_grad=True)
    Yt = torch.fft.ifft(xt2, dim=-1) * n
    (Yt.real * wc[..., 0] + Yt.imag * wc[..., 1]).sum().backward()
    assert rel(to_c(x2.grad), xt2.grad.numpy()) <= TOL


@pytest.mark.parametrize("n", [256, 4096, 1 << 17])
def test_fft32_values_and_gradients_vs_torch_fft(n):
    """Split-precision fft32 / fft32_adjoint on complex64 and float32 [..., n, 2]
    tensors: values and gradients against complex128 torch.fft autograd
    within 1e-6 (near-FP32: the bound is three orders below the fp16 one)."""
    note(f"[gpu] fft32 n={n}")
    g = torch.Generator().manual_seed(n)
    xc = torch.complex(torch.randn(3, n, generator=g), torch.randn(3, n, generator=g)).cuda().requires_grad_(True)
    w = torch.complex(torch.randn(3, n, generator=g), torch.randn(3, n, generator=g)).cuda()
    X = torch_fft.fft32(xc)
    assert X.dtype == torch.complex64 and X.shape == xc.shape
    (X * w.conj()).real.sum().backward()
    xt = xc.detach().cpu().to(torch.complex128).requires_grad_(True)
    Xt = torch.fft.fft(xt, dim=-1)
    (Xt * w.cpu().to(torch.complex128).conj()).real.sum().backward()
    assert rel(X.detach().cpu().numpy(), Xt.detach().numpy()) <= 1e-6
    assert rel(xc.grad.cpu().numpy(), xt.grad.numpy()) <= 1e-6
    # float32 [..., n, 2] layout and the adjoint, whose backward is fft32
    xr = torch.view_as_real(xc.detach().clone()).requires_grad_(True)
    Y = torch_fft.fft32_adjoint(xr)
    assert Y.dtype == torch.float32 and Y.shape == xr.shape
    wr = torch.view_as_real(w)
    (Y * wr).sum().backward()
    xt2 = xc.detach().cpu().to(torch.complex128).requires_grad_(True)
    Yt = torch.fft.ifft(xt2, dim=-1) * n
    wc = w.cpu().to(torch.complex128)
    (Yt.real * wc.real + Yt.imag * wc.imag).sum().backward()
    assert rel(to_c(Y.detach()), Yt.detach().numpy()) <= 1e-6
    assert rel(to_c(xr.grad), xt2.grad.numpy()) <= 1e-6


def test_fft32_rejects_fp16():
    with pytest.raises(RuntimeError):
        torch_fft.fft32(torch.zeros(4, 256, 2, dtype=torch.float16, device="cuda"))


def reference_conv(x_c, h_c, n):
    """FP64 y = F^H (H * F x) / n."""
    return np.fft.ifft(np.fft.fft(x_c, axis=-1) * h_c, axis=-1)


@pytest.mark.parametrize("n", [256, 4096])
def test_fftconv_gradients_vs_fp64_finite_differences(n):
    """n = 256 composes the transforms; n = 4096 runs the fused kernel."""
    note(f"[gpu] fftconv finite differences n={n}")
    torch.manual_seed(0)
    B, C = 2, 3
    layer = torch_fft.FFTConv1d(C, n, std=0.5, device="cuda")
    x = fp16_pair((B, C, n), 4).cuda().requires_grad_(True)
    w = torch.randn(B, C, n, 2, device="cuda")
    y = layer(x)
    (y.float() * w).sum().backward()
    gx = to_c(x.grad)
    gh = to_c(layer.weight.grad)
    xc = to_c(x)
    hc = to_c(layer.weight)
    wc = to_c(w)

    def loss64(xv, hv):
        yv = reference_conv(xv, hv, n)
        return float(np.sum(yv.real * wc.real + yv.imag * wc.imag))

    rng = np.random.default_rng(5)
    eps = 1e-6
    for _ in range(6):
        b, c, k, part = rng.integers(B), rng.integers(C), rng.integers(n), rng.integers(2)
        d = np.zeros_like(xc)
        d[b, c, k] = eps if part == 0 else 1j * eps
        fd = (loss64(xc + d, hc) - loss64(xc - d, hc)) / (2 * eps)
        an = gx[b, c, k].real if part == 0 else gx[b, c, k].imag
        scale = np.abs(gx).max()
        assert abs(fd - an) <= TOL * scale, (fd, an, scale)
    for _ in range(6):
        c, k, part = rng.integers(C), rng.integers(n), rng.integers(2)
        d = np.zeros_like(hc)
        d[c, k] = eps if part == 0 else 1j * eps
        fd = (loss64(xc, hc + d) - loss64(xc, hc - d)) / (2 * eps)
        an = gh[c, k].real if part == 0 else gh[c, k].imag
        scale = np.abs(gh).max()
        assert abs(fd - an) <= TOL * scale, (fd, an, scale)


def test_fftconv_matches_torch_fft_autograd():
    note("[gpu] fftconv vs torch.fft autograd")
    torch.manual_seed(1)
    B, C, n = 4, 8, 8192
    layer = torch_fft.FFTConv1d(C, n, std=0.02, device="cuda")
    x = fp16_pair((B, C, n), 6).cuda().requires_grad_(True)
    w = torch.randn(B, C, n, 2, device="cuda")
    y = layer(x)
    (y.float() * w).sum().backward()
    xt = torch.tensor(to_c(x), dtype=torch.complex128, requires_grad=True)
    ht = torch.tensor(to_c(layer.weight), dtype=torch.complex128, requires_grad=True)
    yt = torch.fft.ifft(torch.fft.fft(xt, dim=-1) * ht, dim=-1)
    wt = torch.tensor(to_c(w))
    (yt.real * wt.real + yt.imag * wt.imag).sum().backward()
    assert rel(to_c(y), yt.detach().numpy()) <= TOL
    assert rel(to_c(x.grad), xt.grad.numpy()) <= TOL
    assert rel(to_c(layer.weight.grad), ht.grad.numpy()) <= TOL


def test_fftconv_full_shape_subset():
    """The configured layer: x [64, 256, 8192] real-as-complex fp16, filter
    [256, 8192] ~ N(0, 0.02); forward + backward run, checked on 3 signals."""
    note("[gpu] fftconv [64,256,8192]")
    torch.manual_seed(2)
    B, C, n = 64, 256, 8192
    layer = torch_fft.FFTConv1d(C, n, std=0.02, device="cuda")
    xr = (torch.rand(B, C, n, device="cuda") * 2 - 1).half()
    x = torch.stack([xr, torch.zeros_like(xr)], dim=-1).requires_grad_(True)
    y = layer(x)
    g = torch.randn_like(y)
    y.backward(g)
    torch.cuda.synchronize()
    hc = to_c(layer.weight)
    for (b, c) in ((0, 0), (17, 130), (63, 255)):
        xs = to_c(x[b, c])
        ref = reference_conv(xs, hc[c], n)
        assert rel(to_c(y[b, c]), ref) <= TOL
        gs = to_c(g[b, c])
        # gx = F^H (conj(H) * F g) / n
        ref_gx = np.fft.ifft(np.fft.fft(gs) * np.conj(hc[c]))
        assert rel(to_c(x.grad[b, c]), ref_gx) <= TOL
    # filter gradient at the full shape: gH[c] = (1/n) sum_b F(g_b,c) conj(F(x_b,c)),
    # summed over all 64 batch entries, against FP64 for sampled channels
    gw = layer.weight.grad.detach().cpu().double()
    for c in (0, 130, 255):
        X = np.fft.fft(to_c(x[:, c]), axis=-1)
        G = np.fft.fft(to_c(g[:, c]), axis=-1)
        ref_gh = (G * np.conj(X)).sum(axis=0) / n
        got = gw[c, :, 0].numpy() + 1j * gw[c, :, 1].numpy()
        err = rel(got, ref_gh)
        note(f"  filter grad channel {c}: rel_l2 {err:.2e}")
        assert err <= TOL


def test_fftconv_rejects_mismatched_layout():
    """The fused layer pairs signal [..., c, :] with filter c, so an input whose
    channel axis does not match the filter is an error, not a silent pairing."""
    layer = torch_fft.FFTConv1d(4, 4096, device="cuda")
    for shape in ((2, 8, 4096), (8, 4096), (2, 4, 8192)):
        x = torch.zeros(*shape, 2, dtype=torch.float16, device="cuda")
        with pytest.raises(RuntimeError):
            layer(x)
