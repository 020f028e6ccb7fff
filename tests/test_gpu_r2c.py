"""Real-input transforms (tcfft_exec_r2c / tcfft_exec_c2r) against the FP64
oracle: bins 0..n/2 of the DFT of real fp16 signals and the unnormalised
Hermitian inverse, n = 512..65536 (row kernels up to 8192, one workgroup per
signal above), ragged batches (a row-kernel wave holds 4096/(n/2) signals),
the reference's own glibc inputs, a round trip, agreement with the
complex transform of the same data fed as complex, and the error paths.
Tolerance: rel-L2 <= 1e-3 (round trip 1.5e-3)."""
import sys

import numpy as np
import pytest

import oracle
import tcfft_amd

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

SIZES = [512, 1024, 2048, 4096, 8192, 16384, 32768, 65536]


def note(m):
    print(m, flush=True)
    sys.stdout.flush()


def real_dev(x):
    return torch.from_numpy(np.ascontiguousarray(x.astype(np.float16))).cuda()


def cplx_dev(X):
    return torch.from_numpy(np.ascontiguousarray(np.stack([X.real, X.imag], -1).astype(np.float16))).cuda()


def to_c(t):
    h = t.cpu().numpy().astype(np.float64)
    return h[..., 0] + 1j * h[..., 1]


def r2c(plan, x):
    """x: float64 [batch, n] (fp16-exact) -> complex128 [batch, n/2+1]."""
    b, n = x.shape
    xin = real_dev(x)
    out = torch.empty(b, n // 2 + 1, 2, dtype=torch.float16, device="cuda")
    plan.exec_r2c(xin.data_ptr(), out.data_ptr())
    torch.cuda.synchronize()
    return to_c(out)


def c2r(plan, X, n):
    b = X.shape[0]
    xin = cplx_dev(X)
    out = torch.empty(b, n, dtype=torch.float16, device="cuda")
    plan.exec_c2r(xin.data_ptr(), out.data_ptr())
    torch.cuda.synchronize()
    return out.cpu().numpy().astype(np.float64)


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("batch", [1, 3, 17, 64])
def test_r2c_matches_oracle(n, batch):
    note(f"[gpu] r2c n={n} batch={batch}")
    rng = np.random.default_rng(n + batch)
    x = rng.uniform(-1, 1, (batch, n)).astype(np.float16).astype(np.float64)
    plan = tcfft_amd.PlanR2C(n, batch)
    got = r2c(plan, x)
    err = oracle.rel_l2(got, oracle.rfft(x))
    note(f"  rel_l2={err:.3e}")
    assert err <= 1e-3
    # bins 0 and n/2 of a real signal are real
    assert np.all(got[:, 0].imag == 0) and np.all(got[:, -1].imag == 0)
    plan.destroy()


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("batch", [1, 5, 33])
def test_c2r_matches_oracle(n, batch):
    note(f"[gpu] c2r n={n} batch={batch}")
    rng = np.random.default_rng(3 * n + batch)
    X = (rng.uniform(-1, 1, (batch, n // 2 + 1)) + 1j * rng.uniform(-1, 1, (batch, n // 2 + 1)))
    X = X.real.astype(np.float16).astype(np.float64) + 1j * X.imag.astype(np.float16).astype(np.float64)
    plan = tcfft_amd.PlanR2C(n, batch)
    got = c2r(plan, X, n)
    err = float(np.linalg.norm(got - oracle.irfft(X, n)) / np.linalg.norm(oracle.irfft(X, n)))
    note(f"  rel_l2={err:.3e}")
    assert err <= 1e-3
    plan.destroy()


@pytest.mark.parametrize("n", [1024, 8192, 16384])
def test_r2c_reference_inputs_and_round_trip(n):
    """The reference's glibc generator (doit.h:7-19), real parts as the real
    signal; then c2r(r2c(x)) = n x."""
    batch = 16
    x = oracle.generate_data(n, batch, seed=42).real
    x = (x * 1e4).astype(np.float16).astype(np.float64)  # generator range 1e-4 scaled into fp16's normal range
    plan = tcfft_amd.PlanR2C(n, batch)
    X = r2c(plan, x)
    assert oracle.rel_l2(X, oracle.rfft(x)) <= 1e-3
    back = c2r(plan, X, n) / n
    err = float(np.linalg.norm(back - x) / np.linalg.norm(x))
    note(f"[gpu] r2c round trip n={n}: rel_l2={err:.3e}")
    assert err <= 1.5e-3
    plan.destroy()


@pytest.mark.parametrize("n", [2048, 8192, 32768])
def test_r2c_agrees_with_complex_transform_of_real_data(n):
    batch = 8
    rng = np.random.default_rng(n)
    x = rng.uniform(-1, 1, (batch, n)).astype(np.float16).astype(np.float64)
    rp = tcfft_amd.PlanR2C(n, batch)
    X = r2c(rp, x)
    cp = tcfft_amd.Plan1D(n, batch)
    d = cplx_dev(x + 0j)
    o = torch.empty_like(d)
    cp.exec_natural(d.data_ptr(), o.data_ptr())
    torch.cuda.synchronize()
    full = to_c(o)
    assert oracle.rel_l2(X, full[:, : n // 2 + 1]) <= 1.5e-3
    rp.destroy()
    cp.destroy()


def test_r2c_error_paths():
    n, batch = 1024, 4
    plan = tcfft_amd.PlanR2C(n, batch)
    assert plan.bins == 513 and plan.radices == tcfft_amd.query_1d(512, batch)[0]
    buf = torch.zeros(4 * batch * n, dtype=torch.float16, device="cuda")
    p = buf.data_ptr()
    with pytest.raises(tcfft_amd.TcfftError):   # overlapping in / out
        plan.exec_r2c(p, p + 1024)
    with pytest.raises(tcfft_amd.TcfftError):   # real side not 16-byte aligned
        plan.exec_r2c(p + 8, p + 2 * batch * n * 2 + 64)
    with pytest.raises(tcfft_amd.TcfftError):   # complex side not 4-byte aligned
        plan.exec_c2r(p + 2, p + 2 * batch * n * 2 + 64)
    cp = tcfft_amd.Plan1D(512, batch)           # a complex plan is not a real-input plan
    with pytest.raises(tcfft_amd.TcfftError):
        tcfft_amd._check(tcfft_amd.lib().tcfft_exec_r2c(cp._h, tcfft_amd._vp(p), tcfft_amd._vp(p + 65536), None),
                         "tcfft_exec_r2c")
    cp.destroy()
    plan.destroy()


@pytest.mark.parametrize("n", [16384, 65536])
def test_r2c_large_round_trip(n):
    """c2r(r2c(x)) = n x for the workgroup kernels (x scaled so that n x
    stays inside fp16's range)."""
    batch = 6
    rng = np.random.default_rng(n + 5)
    x = (rng.uniform(-1, 1, (batch, n)) * (0.25 if n > 32768 else 1.0)).astype(np.float16).astype(np.float64)
    plan = tcfft_amd.PlanR2C(n, batch)
    back = c2r(plan, r2c(plan, x), n) / n
    err = float(np.linalg.norm(back - x) / np.linalg.norm(x))
    note(f"[gpu] r2c round trip n={n}: rel_l2={err:.3e}")
    assert err <= 1.5e-3
    plan.destroy()


@pytest.mark.parametrize("n,batch", [(1 << 17, 1), (1 << 17, 3), (1 << 18, 2), (1 << 20, 1)])
def test_real_above_65536_matches_oracle(n, batch):
    """Lengths above 65536: the n/2-point transform as the plan's passes plus
    the separate even/odd pass, both directions and a round trip."""
    note(f"[gpu] r2c/c2r two-step n={n} batch={batch}")
    rng = np.random.default_rng(n + 7 * batch)
    plan = tcfft_amd.PlanR2C(n, batch)
    x = rng.uniform(-1, 1, (batch, n)).astype(np.float16).astype(np.float64)
    X = r2c(plan, x)
    err = oracle.rel_l2(X, oracle.rfft(x))
    note(f"  r2c rel_l2={err:.3e}")
    assert err <= 1.5e-3
    Y = (rng.uniform(-1, 1, (batch, n // 2 + 1)) + 1j * rng.uniform(-1, 1, (batch, n // 2 + 1)))
    Y = Y.real.astype(np.float16).astype(np.float64) + 1j * Y.imag.astype(np.float16).astype(np.float64)
    y = c2r(plan, Y, n)
    want = oracle.irfft(Y, n)
    err = float(np.linalg.norm(y - want) / np.linalg.norm(want))
    note(f"  c2r rel_l2={err:.3e}")
    assert err <= 1.5e-3
    # round trip on a scaled signal (n x must stay inside fp16's range)
    xs = (x / 64).astype(np.float16).astype(np.float64) if n > (1 << 16) else x
    back = c2r(plan, r2c(plan, xs), n) / n
    err = float(np.linalg.norm(back - xs) / np.linalg.norm(xs))
    note(f"  round trip rel_l2={err:.3e}")
    assert err <= 2e-3
    plan.destroy()
