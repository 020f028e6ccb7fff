"""The configured workloads at their full sizes, checked through properties
that do not depend on the size: Parseval per signal (sum |X|^2 = n sum |x|^2),
the forward -> adjoint round trip (F^H F = n I), linearity, and numpy FP64
on a few sampled signals.  Tolerances: FP16 path rel-L2 <= 1e-3 (per-signal
energy within 2e-3), split precision <= 1e-6."""
import sys

import numpy as np
import pytest

import tcfft_amd

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def note(m):
    print(m, flush=True)
    sys.stdout.flush()


def to_c(t):
    h = t.detach().cpu().numpy().astype(np.float64)
    return h[..., 0] + 1j * h[..., 1]


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def dropin(plan, x):
    """x: complex128 [batch, n] -> fp16 device buffer in the plan's digit-reversed order."""
    n = x.shape[-1]
    xr = x[:, tcfft_amd.gen_rev(n, plan.radices)]
    return torch.from_numpy(np.ascontiguousarray(np.stack([xr.real, xr.imag], -1).astype(np.float16))).cuda()


@pytest.mark.parametrize("n,batch", [(1 << 16, 256), (4096, 1024)])
def test_plain_full_size_properties(n, batch):
    note(f"[gpu] full size n={n} batch={batch}")
    rng = np.random.default_rng(n + batch)
    x = (rng.uniform(-1, 1, (batch, n)).astype(np.float16).astype(np.float64)
         + 1j * rng.uniform(-1, 1, (batch, n)).astype(np.float16).astype(np.float64))
    plan = tcfft_amd.Plan1D(n, batch)
    d = dropin(plan, x)
    plan.exec(d.data_ptr())
    torch.cuda.synchronize()
    X = to_c(d)
    # Parseval on every signal
    ex = np.sum(np.abs(x) ** 2, axis=1) * n
    eX = np.sum(np.abs(X) ** 2, axis=1)
    worst = float(np.max(np.abs(eX / ex - 1)))
    note(f"  Parseval worst per-signal relative deviation {worst:.2e}")
    assert worst <= 2e-3
    # sampled signals against numpy
    idx = [0, batch // 3, batch - 1]
    err = rel(X[idx], np.fft.fft(x[idx], axis=1))
    note(f"  sampled rel_l2 {err:.2e}")
    assert err <= 1e-3
    # round trip: the adjoint of the fp16 spectrum scaled by 1/sqrt(n) (a power
    # of two, exact) returns sqrt(n) x; unscaled, n x would leave fp16's range
    r = float(np.sqrt(n))
    Xs = X / r
    Xq = Xs.real.astype(np.float16).astype(np.float64) + 1j * Xs.imag.astype(np.float16).astype(np.float64)
    d2 = dropin(plan, Xq)
    plan.exec(d2.data_ptr(), tcfft_amd.INVERSE)
    torch.cuda.synchronize()
    back = to_c(d2) / r
    err = rel(back, x)
    note(f"  round trip rel_l2 {err:.2e}")
    assert err <= 1.5e-3
    plan.destroy()


def test_plain_full_size_linearity():
    """F(a x + b y) = a F x + b F y over the whole 2^16 x 256 batch, with
    a = b = 1/2 (exact in fp16), against the two transforms run separately."""
    n, batch = 1 << 16, 256
    rng = np.random.default_rng(3)
    x = rng.uniform(-1, 1, (batch, n, 2)).astype(np.float16)
    y = rng.uniform(-1, 1, (batch, n, 2)).astype(np.float16)
    s = ((x.astype(np.float32) + y.astype(np.float32)) * 0.5).astype(np.float16)
    plan = tcfft_amd.Plan1D(n, batch)
    outs = []
    for a in (x, y, s):
        t = torch.from_numpy(a).cuda()
        plan.exec(t.data_ptr())  # linearity holds in any input order
        outs.append(to_c(t))
    torch.cuda.synchronize()
    err = rel(outs[2], 0.5 * (outs[0] + outs[1]))
    note(f"[gpu] linearity 2^16 x 256 rel_l2 {err:.2e}")
    assert err <= 1e-3
    plan.destroy()


def test_split_full_size_round_trip():
    n, batch = 1 << 16, 256
    note(f"[gpu] split precision full size n={n} batch={batch}")
    rng = np.random.default_rng(17)
    scale = 10.0 ** rng.uniform(-3, 3, (batch, 1))
    x = ((rng.standard_normal((batch, n)) + 1j * rng.standard_normal((batch, n))) * scale).astype(np.complex64)
    xin = torch.from_numpy(x.view(np.float32).reshape(batch, n, 2)).cuda()
    X = torch.empty_like(xin)
    back = torch.empty_like(xin)
    plan = tcfft_amd.Plan1D(n, batch)
    plan.exec_split(xin.data_ptr(), X.data_ptr())
    plan.exec_split(X.data_ptr(), back.data_ptr(), tcfft_amd.INVERSE)
    torch.cuda.synchronize()
    Xc = X.cpu().numpy().view(np.complex64).reshape(batch, n).astype(np.complex128)
    bc = back.cpu().numpy().view(np.complex64).reshape(batch, n).astype(np.complex128) / n
    x64 = x.astype(np.complex128)
    ex = np.sum(np.abs(x64) ** 2, axis=1) * n
    worst_p = float(np.max(np.abs(np.sum(np.abs(Xc) ** 2, axis=1) / ex - 1)))
    worst_rt = max(rel(bc[i], x64[i]) for i in range(batch))
    note(f"  Parseval worst {worst_p:.2e}, round trip worst per-signal rel_l2 {worst_rt:.2e}")
    assert worst_p <= 2e-6 and worst_rt <= 2e-6
    plan.destroy()


def test_uneven_full_size_round_trip():
    """The configured uneven batch (4096 signals of 2^U{8..16}): forward then
    adjoint through the persistent launch returns every signal."""
    rng = np.random.default_rng(0)
    sizes = (2 ** rng.integers(8, 17, 4096)).astype(np.int32)
    note(f"[gpu] uneven full size: {len(sizes)} signals, {int(sizes.sum())} elements")
    mp = tcfft_amd.MultiPlan(sizes)
    total = int(sizes.sum())
    x = rng.uniform(-1, 1, (total, 2)).astype(np.float16)
    d = torch.from_numpy(x).cuda()
    mp.exec(d.data_ptr())
    torch.cuda.synchronize()
    X = d.cpu().numpy().astype(np.float64)
    # the spectrum is in natural order; the adjoint takes the digit-reversed order
    parts, o = [], 0
    for n in sizes:
        n = int(n)
        seg = X[o:o + n, 0] + 1j * X[o:o + n, 1]
        rad, _ = tcfft_amd.query_1d(n, 1)
        seg = seg / np.sqrt(n)  # keeps the adjoint's result, sqrt(n) x, inside fp16's range
        segq = seg.real.astype(np.float16).astype(np.float64) + 1j * seg.imag.astype(np.float16).astype(np.float64)
        parts.append(segq[tcfft_amd.gen_rev(n, rad)])
        o += n
    flat = np.concatenate(parts)
    d2 = torch.from_numpy(np.ascontiguousarray(np.stack([flat.real, flat.imag], -1).astype(np.float16))).cuda()
    mp.exec(d2.data_ptr(), tcfft_amd.INVERSE)
    torch.cuda.synchronize()
    back = d2.cpu().numpy().astype(np.float64)
    worst, o = 0.0, 0
    for n in sizes:
        n = int(n)
        # input of the forward exec was in the drop-in order: compare in that order
        rad, _ = tcfft_amd.query_1d(n, 1)
        xin = x[o:o + n, 0].astype(np.float64) + 1j * x[o:o + n, 1].astype(np.float64)
        nat = np.empty_like(xin)
        nat[tcfft_amd.gen_rev(n, rad)] = xin  # the natural-order signal the forward transformed
        b = (back[o:o + n, 0] + 1j * back[o:o + n, 1]) / np.sqrt(n)
        worst = max(worst, rel(b, nat))
        o += n
    note(f"  round trip worst per-signal rel_l2 {worst:.2e}")
    assert worst <= 1.5e-3
    mp.destroy()
