"""tcfft_exec_natural_noalloc (the natural-order entry point that allocates
nothing) against tcfft_exec_natural, through the C ABI: bit-identical
results on two calls in a row, the second with a smaller input, for the
in-place cases that stage through the library's fixed per-GPU scratch
(2D plans; 1D lines above 65536 points) and for an out-of-place call (no
staging); a call whose staging exceeds the scratch is rejected."""
import sys

import numpy as np
import pytest

import tcfft_amd

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def note(m):
    print(m, flush=True)
    sys.stdout.flush()


def make_plan(shape):
    if len(shape) == 3:
        return tcfft_amd.Plan2D(shape[1], shape[2], shape[0])
    return tcfft_amd.Plan1D(shape[1], shape[0])


def check_in_place(shape, seed):
    note(f"[gpu] noalloc in place {shape}")
    rng = np.random.default_rng(seed)
    x = torch.from_numpy(rng.uniform(-1, 1, shape + (2,)).astype(np.float16)).cuda()
    plan = make_plan(shape)
    want = x.clone()
    plan.exec_natural(want.data_ptr(), want.data_ptr())
    got = x.clone()
    plan.exec_natural_noalloc(got.data_ptr(), got.data_ptr())
    torch.cuda.synchronize()
    assert torch.equal(got, want), shape
    plan.destroy()


@pytest.mark.parametrize("shapes", [((8, 1024, 1024), (2, 256, 512)), ((4, 1 << 20), (2, 1 << 17))])
def test_noalloc_matches_exec_natural_twice(shapes):
    """Two calls in a row share the scratch; the second (smaller) one must
    not see anything of the first."""
    big, small = shapes
    check_in_place(big, 1)
    check_in_place(small, 2)


def test_noalloc_out_of_place_and_adjoint():
    note("[gpu] noalloc out of place, adjoint, 2D 512 x 1024 x 2")
    rng = np.random.default_rng(3)
    x = torch.from_numpy(rng.uniform(-1, 1, (2, 512, 1024, 2)).astype(np.float16)).cuda()
    plan = tcfft_amd.Plan2D(512, 1024, 2)
    for d in (tcfft_amd.FORWARD, tcfft_amd.INVERSE):
        want, got = torch.empty_like(x), torch.empty_like(x)
        plan.exec_natural(x.data_ptr(), want.data_ptr(), d)
        plan.exec_natural_noalloc(x.data_ptr(), got.data_ptr(), d)
        torch.cuda.synchronize()
        assert torch.equal(got, want), d
    plan.destroy()


def test_noalloc_rejects_staging_beyond_scratch():
    """2D 1024 x 1024 x 32 in place stages 128 MiB > 64 MiB of scratch:
    TCFFT_INVALID_SIZE, nothing launched; out of place it stages nothing."""
    note("[gpu] noalloc size limit")
    plan = tcfft_amd.Plan2D(1024, 1024, 32)
    x = torch.zeros(32, 1024, 1024, 2, dtype=torch.float16, device="cuda")
    assert x.numel() * 2 > tcfft_amd.NOALLOC_SCRATCH_BYTES
    with pytest.raises(tcfft_amd.TcfftError) as e:
        plan.exec_natural_noalloc(x.data_ptr(), x.data_ptr())
    assert e.value.status == 1
    y = torch.empty_like(x)
    plan.exec_natural_noalloc(x.data_ptr(), y.data_ptr())
    torch.cuda.synchronize()
    assert torch.count_nonzero(y).item() == 0
    plan.destroy()
