"""Optional PyTorch-ROCm binding over the C ABI: differentiable FP16 FFTs.

Complex fp16 tensors are carried as real fp16 tensors of shape [..., n, 2]
(interleaved re/im, the kernels' layout).  Transforms run along the n axis,
in natural order, on the current HIP stream, through libtcfft.so.

    fft(x)           X = F x          (forward, exp(-2*pi*i*nk/n), unnormalised)
    fft_adjoint(x)   Y = F^H x        (conjugate-twiddle kernel, unnormalised)

Autograd: both are complex-linear maps, and for a complex-linear A viewed as a
real-linear map on [..., 2] the transpose is the complex adjoint A^H.  So the
backward of fft is the hand-written adjoint kernel (F^H) and the backward of
fft_adjoint is the forward kernel (F).  No CPU fallback: these raise if the
library or a GPU is missing.

Real input (n = 2^9..2^27):

    rfft(x)          bins 0..n/2 of F x for real fp16 x [..., n] -> [..., n/2+1, 2]
    c2r(X, n)        y = sum over the Hermitian extension of X, n * irfft
    irfft(X, n)      c2r(X, n) / n (the scale is applied in fp32)

Near-FP32 accuracy (split precision, complex64 [..., n] or float32 [..., n, 2],
n = 2^8..2^27):

    fft32(x)         X = F x
    fft32_adjoint(x) Y = F^H x

with the same adjoint pairing as fft / fft_adjoint in their backward passes.

The real-input transforms' backward passes are each other's kernels: viewed as real-linear maps,
the transpose of rfft is c2r applied to the gradient with bins 1..n/2-1
halved (c2r counts those bins twice), and the transpose of c2r is rfft of the
gradient with bins 1..n/2-1 doubled.

FFTConv1d implements y = F^H (H * F x) / n with a learnable filter spectrum H
broadcast over leading dimensions (circular convolution by the impulse
response F^H H / n); for n = 4096, 8192 and 16384 forward and backward run
the fused single-pass convolution kernel (tcfft_conv_exec), which keeps the
spectrum in registers and in the drop-in's digit-reversed order.
"""
from __future__ import annotations

from collections import OrderedDict

import torch

from . import FORWARD, INVERSE, Plan1D, PlanR2C, gen_rev

# Plans are cached per (n, batch, device), least recently used first out:
# a plan holds device tables (and, for some layouts, a workspace as large as
# its data), so a caller cycling through batch sizes must not grow device
# memory without bound.
PLAN_CACHE_SIZE = 8
_PLANS: OrderedDict = OrderedDict()
_RPLANS: OrderedDict = OrderedDict()


def _cached(cache: OrderedDict, key, make):
    p = cache.get(key)
    if p is not None:
        cache.move_to_end(key)
        return p
    with torch.cuda.device(key[2]):
        p = make()
    cache[key] = p
    while len(cache) > PLAN_CACHE_SIZE:
        _, old = cache.popitem(last=False)
        # work queued with the evicted plan may still be running on any stream
        torch.cuda.synchronize(old.device_index)
        old.destroy()
    return p


def _plan(n: int, batch: int, device: torch.device) -> Plan1D:
    def make():
        p = Plan1D(n, batch)
        p.device_index = device.index
        return p
    return _cached(_PLANS, (n, batch, device.index), make)


def _check(x: torch.Tensor):
    if x.device.type != "cuda":
        raise RuntimeError("tcfft_amd.torch_fft needs a CUDA (HIP) tensor")
    if x.dtype != torch.float16 or x.dim() < 2 or x.shape[-1] != 2:
        raise RuntimeError("expected an fp16 tensor of shape [..., n, 2] (interleaved complex)")


def _exec(x: torch.Tensor, direction: int) -> torch.Tensor:
    _check(x)
    x = x.contiguous()
    n = x.shape[-2]
    batch = x.numel() // (2 * n)
    out = torch.empty_like(x)
    plan = _plan(n, batch, x.device)
    stream = torch.cuda.current_stream(x.device).cuda_stream
    plan.exec_natural(x.data_ptr(), out.data_ptr(), direction, stream)
    return out


class _FFT(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        return _exec(x, FORWARD)

    @staticmethod
    def backward(ctx, g):
        return _exec(g, INVERSE)


class _FFTAdjoint(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        return _exec(x, INVERSE)

    @staticmethod
    def backward(ctx, g):
        return _exec(g, FORWARD)


def fft(x: torch.Tensor) -> torch.Tensor:
    """Differentiable forward FFT of an fp16 [..., n, 2] tensor."""
    return _FFT.apply(x)


def fft_adjoint(x: torch.Tensor) -> torch.Tensor:
    """Differentiable unnormalised inverse (adjoint) FFT, F^H x."""
    return _FFTAdjoint.apply(x)


def ifft(x: torch.Tensor) -> torch.Tensor:
    """Normalised inverse, F^H x / n (the scale is applied in fp32)."""
    n = x.shape[-2]
    return (fft_adjoint(x).float() / n).half()


_SPLANS: OrderedDict = OrderedDict()


def _splan(n: int, batch: int, device: torch.device) -> Plan1D:
    def make():
        p = Plan1D(n, batch)
        p.device_index = device.index
        p.prepare_split()  # tables and workspace here, not in the exec
        return p
    return _cached(_SPLANS, (n, batch, device.index), make)


def _exec_split(x: torch.Tensor, direction: int) -> torch.Tensor:
    """Split-precision transform along the n axis of an FP32 [..., n, 2] or
    complex64 [..., n] CUDA tensor; returns the same dtype and shape."""
    if x.device.type != "cuda":
        raise RuntimeError("tcfft_amd.torch_fft needs a CUDA (HIP) tensor")
    cplx = x.dtype == torch.complex64
    v = torch.view_as_real(x) if cplx else x
    if v.dtype != torch.float32 or v.dim() < 2 or v.shape[-1] != 2:
        raise RuntimeError("expected a complex64 [..., n] or float32 [..., n, 2] tensor")
    v = v.contiguous()
    n = v.shape[-2]
    batch = v.numel() // (2 * n)
    out = torch.empty_like(v)
    _splan(n, batch, v.device).exec_split(v.data_ptr(), out.data_ptr(), direction,
                                          torch.cuda.current_stream(v.device).cuda_stream)
    return torch.view_as_complex(out) if cplx else out


class _FFT32(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        return _exec_split(x, FORWARD)

    @staticmethod
    def backward(ctx, g):
        return _exec_split(g.contiguous(), INVERSE)


class _FFT32Adjoint(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        return _exec_split(x, INVERSE)

    @staticmethod
    def backward(ctx, g):
        return _exec_split(g.contiguous(), FORWARD)


def fft32(x: torch.Tensor) -> torch.Tensor:
    """Differentiable forward FFT at near-FP32 accuracy (split precision, FP16
    hi+lo on MFMA) along the last complex axis: complex64 [..., n] or float32
    [..., n, 2], n = 2^8..2^27."""
    return _FFT32.apply(x)


def fft32_adjoint(x: torch.Tensor) -> torch.Tensor:
    """Differentiable unnormalised inverse (adjoint) of fft32, F^H x."""
    return _FFT32Adjoint.apply(x)


def _rplan(n: int, batch: int, device: torch.device) -> PlanR2C:
    def make():
        p = PlanR2C(n, batch)
        p.device_index = device.index
        return p
    return _cached(_RPLANS, (n, batch, device.index), make)


def _r2c(x: torch.Tensor) -> torch.Tensor:
    if x.device.type != "cuda" or x.dtype != torch.float16 or x.dim() < 1:
        raise RuntimeError("expected a real fp16 CUDA (HIP) tensor of shape [..., n]")
    x = x.contiguous()
    n = x.shape[-1]
    batch = x.numel() // n
    out = torch.empty(x.shape[:-1] + (n // 2 + 1, 2), dtype=torch.float16, device=x.device)
    _rplan(n, batch, x.device).exec_r2c(x.data_ptr(), out.data_ptr(), torch.cuda.current_stream(x.device).cuda_stream)
    return out


def _c2r(X: torch.Tensor, n: int) -> torch.Tensor:
    _check(X)
    if X.shape[-2] != n // 2 + 1:
        raise RuntimeError(f"expected {n // 2 + 1} bins for n = {n}, got {X.shape[-2]}")
    X = X.contiguous()
    batch = X.numel() // (2 * (n // 2 + 1))
    out = torch.empty(X.shape[:-2] + (n,), dtype=torch.float16, device=X.device)
    _rplan(n, batch, X.device).exec_c2r(X.data_ptr(), out.data_ptr(), torch.cuda.current_stream(X.device).cuda_stream)
    return out


def _scale_interior(t: torch.Tensor, f: float) -> torch.Tensor:
    v = t.float()
    v[..., 1:-1, :] *= f
    return v.half()


class _RFFT(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        ctx.n = x.shape[-1]
        return _r2c(x)

    @staticmethod
    def backward(ctx, g):
        return _c2r(_scale_interior(g, 0.5), ctx.n)


class _C2R(torch.autograd.Function):
    @staticmethod
    def forward(ctx, X, n):
        return _c2r(X, n)

    @staticmethod
    def backward(ctx, g):
        return _scale_interior(_r2c(g), 2.0), None


def rfft(x: torch.Tensor) -> torch.Tensor:
    """Differentiable real-input FFT: fp16 [..., n] -> bins 0..n/2 as [..., n/2+1, 2]."""
    return _RFFT.apply(x)


def c2r(X: torch.Tensor, n: int) -> torch.Tensor:
    """Differentiable unnormalised Hermitian inverse (n times irfft); the
    imaginary parts of bins 0 and n/2 are ignored."""
    return _C2R.apply(X, n)


def irfft(X: torch.Tensor, n: int) -> torch.Tensor:
    """Normalised inverse of rfft, c2r(X, n) / n (the scale is applied in fp32)."""
    return (c2r(X, n).float() / n).half()


def complex_mul(a: torch.Tensor, h: torch.Tensor) -> torch.Tensor:
    """Pointwise complex product of [..., n, 2] tensors, computed in fp32,
    returned in a's dtype; h broadcasts over a's leading dimensions."""
    ar, ai = a[..., 0].float(), a[..., 1].float()
    hr, hi = h[..., 0].float(), h[..., 1].float()
    return torch.stack([ar * hr - ai * hi, ar * hi + ai * hr], dim=-1).to(a.dtype)


FUSED_N = (4096, 8192, 16384)
_REV: dict = {}


def _rev(n: int, device: torch.device):
    """(rev, inverse rev) index tensors of the digit-reversed spectrum order."""
    key = (n, device.index)
    r = _REV.get(key)
    if r is None:
        plan = _plan(n, 1, device)
        rev = torch.as_tensor(gen_rev(n, plan.radices).astype("int64"), device=device)
        r = (rev, torch.argsort(rev))
        _REV[key] = r
    return r


class _FusedConv(torch.autograd.Function):
    """y = F^H (H * F x) / n with the single-pass HIP kernel; backward:
    gx = F^H (conj(H) * F gy) / n (same kernel, conjugated filter) and
    gH = (1/n) sum_b conj(F x_b) * F gy_b from the spectra both passes store."""

    @staticmethod
    def forward(ctx, x, weight, want_filter_grad: bool):
        _check(x)
        if weight.dim() != 3 or weight.shape[-1] != 2:
            raise RuntimeError("expected a filter spectrum of shape [C, n, 2]")
        if x.dim() < 3 or x.shape[-3] != weight.shape[0] or x.shape[-2] != weight.shape[1]:
            raise RuntimeError(f"expected x of shape [..., C={weight.shape[0]}, n={weight.shape[1]}, 2], "
                               f"got {tuple(x.shape)}")
        x = x.contiguous()
        n, C = x.shape[-2], weight.shape[0]
        signals = x.numel() // (2 * n)
        plan = _plan(n, signals, x.device)
        stream = torch.cuda.current_stream(x.device).cuda_stream
        # the filter into the digit-reversed order on the device (LDS-tiled,
        # copy speed; a torch gather reads 16x the bytes for this permutation)
        h16 = weight.detach().half().contiguous()
        h_rev = torch.empty_like(h16)
        plan.permute_to_dropin(h16.data_ptr(), h_rev.data_ptr(), C, stream)
        y = torch.empty_like(x)
        # the spectrum of x is kept only when the filter gradient will be needed
        # (grad mode is off inside forward, so the caller passes it in)
        spec = torch.empty_like(x) if want_filter_grad else None
        plan.conv_exec(C, x.data_ptr(), h_rev.data_ptr(), y.data_ptr(), False,
                       spec.data_ptr() if spec is not None else 0, stream)
        ctx.save_for_backward(h_rev, spec if spec is not None else torch.empty(0, device=x.device))
        ctx.shape = (n, C, signals, spec is not None)
        return y

    @staticmethod
    def backward(ctx, gy):
        h_rev, spec = ctx.saved_tensors
        n, C, signals, want_h = ctx.shape
        gy = gy.to(torch.float16).contiguous()
        plan = _plan(n, signals, gy.device)
        stream = torch.cuda.current_stream(gy.device).cuda_stream
        gx = torch.empty_like(gy)
        gspec = torch.empty_like(gy) if want_h else None
        plan.conv_exec(C, gy.data_ptr(), h_rev.data_ptr(), gx.data_ptr(), True,
                       gspec.data_ptr() if gspec is not None else 0, stream)
        gw = None
        if want_h:
            gh_rev = torch.empty(C, n, 2, dtype=torch.float32, device=gy.device)
            plan.conv_filter_grad(C, spec.data_ptr(), gspec.data_ptr(), gh_rev.data_ptr(), stream)
            gw = gh_rev.index_select(1, _rev(n, gy.device)[1])
        return gx, gw, None


class FFTConv1d(torch.nn.Module):
    """y = F^H (H * F x) / n along the n axis of an fp16 [..., C, n, 2] input.

    H has shape [C, n, 2] (a complex filter spectrum per channel) and is kept
    in fp32 as a parameter.  For n = 4096, 8192, 16384 the layer runs the
    fused single-pass convolution kernel (forward and both gradients);
    other sizes compose the differentiable transforms with an fp32 product.
    """

    def __init__(self, channels: int, n: int, std: float = 0.02, device=None):
        super().__init__()
        self.n = n
        self.weight = torch.nn.Parameter(torch.randn(channels, n, 2, device=device) * std)

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        if self.n in FUSED_N:
            return _FusedConv.apply(x, self.weight, torch.is_grad_enabled() and self.weight.requires_grad)
        X = fft(x)
        Y = complex_mul(X, self.weight)
        y = fft_adjoint(Y)
        return (y.float() / self.n).half()
