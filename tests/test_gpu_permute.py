"""tcfft_permute_to_dropin: the caller-side gen_rev staging on the device.
Bit-exact against the host permutation (tcfft_doit_half.cpp:6-29,40-46 via
gen_rev) for several sizes and line counts; the drop-in exec of the permuted
data equals the natural-order exec bit for bit; bad arguments are rejected."""
import sys

import numpy as np
import pytest

import tcfft_amd

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def note(m):
    print(m, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("n,lines", [(256, 3), (4096, 7), (8192, 256), (1 << 16, 2), (1 << 20, 1)])
def test_permute_matches_host_gen_rev(n, lines):
    note(f"[gpu] permute to drop-in order n={n} lines={lines}")
    rng = np.random.default_rng(n + lines)
    x = rng.integers(0, 2 ** 32, (lines, n), dtype=np.uint64).astype(np.uint32)
    plan = tcfft_amd.Plan1D(n, 1)
    d = torch.from_numpy(x.view(np.int32)).cuda()
    o = torch.empty_like(d)
    plan.permute_to_dropin(d.data_ptr(), o.data_ptr(), lines)
    torch.cuda.synchronize()
    want = x[:, tcfft_amd.gen_rev(n, plan.radices)]
    np.testing.assert_array_equal(o.cpu().numpy().view(np.uint32), want)
    plan.destroy()


def test_permute_then_exec_equals_natural():
    n, batch = 8192, 16
    rng = np.random.default_rng(5)
    x = torch.from_numpy(rng.uniform(-1, 1, (batch, n, 2)).astype(np.float16)).cuda()
    plan = tcfft_amd.Plan1D(n, batch)
    a = torch.empty_like(x)
    plan.permute_to_dropin(x.data_ptr(), a.data_ptr(), batch)
    plan.exec(a.data_ptr())
    b = torch.empty_like(x)
    plan.exec_natural(x.data_ptr(), b.data_ptr())
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    plan.destroy()


def test_permute_rejects_bad_arguments():
    plan = tcfft_amd.Plan1D(4096, 4)
    buf = torch.zeros(3 * 4096 * 4, dtype=torch.int32, device="cuda")
    p = buf.data_ptr()
    for args in ((p, p + 4096 * 4, 4),       # overlap
                 (p + 4, p + 8 * 4096 * 4, 1),  # misaligned
                 (p, p + 8 * 4096 * 4, 0)):     # no lines
        with pytest.raises(tcfft_amd.TcfftError):
            plan.permute_to_dropin(*args)
    p2 = tcfft_amd.Plan2D(256, 256, 1)
    with pytest.raises(tcfft_amd.TcfftError):
        p2.permute_to_dropin(p, p + 8 * 4096 * 4, 1)
    p2.destroy()
    plan.destroy()
