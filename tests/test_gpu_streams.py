"""Reentrancy and caller-chosen streams: one plan used from two HIP streams
at once (different buffers), and two plans on two streams, must give results
bit-identical to the same calls run one after another on the null stream.
Covers the drop-in exec, out-of-place natural order, the real-input
transforms and the uneven batch.  Streams keep the default priority."""
import sys

import numpy as np
import pytest

import tcfft_amd

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def note(m):
    print(m, flush=True)
    sys.stdout.flush()


def fp16(rng, shape):
    return torch.from_numpy(rng.uniform(-1, 1, shape).astype(np.float16)).cuda()


def test_one_plan_two_streams():
    note("[gpu] one plan, two streams: 2^16 x 8 exec and natural")
    rng = np.random.default_rng(21)
    plan = tcfft_amd.Plan1D(1 << 16, 8)
    xs = [fp16(rng, (8, 1 << 16, 2)) for _ in range(4)]
    want_exec, want_nat = [], []
    for x in xs:
        a = x.clone()
        plan.exec(a.data_ptr())
        b = torch.empty_like(x)
        plan.exec_natural(x.data_ptr(), b.data_ptr())
        want_exec.append(a)
        want_nat.append(b)
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    bufs = [x.clone() for x in xs]
    nats = [torch.empty_like(x) for x in xs]
    torch.cuda.synchronize()
    for rep in range(3):
        for i, x in enumerate(xs):
            st = streams[i % 2]
            plan.exec(bufs[i].data_ptr(), tcfft_amd.FORWARD, st.cuda_stream)
            plan.exec_natural(x.data_ptr(), nats[i].data_ptr(), tcfft_amd.FORWARD, st.cuda_stream)
        torch.cuda.synchronize()
        for i in range(len(xs)):
            assert torch.equal(bufs[i], want_exec[i]), (rep, i)
            assert torch.equal(nats[i], want_nat[i]), (rep, i)
            bufs[i].copy_(xs[i])
        torch.cuda.synchronize()
    plan.destroy()


def test_two_plans_two_streams():
    note("[gpu] two plans, two streams: 2D 1024^2 x 2 and R2C 8192 x 64, uneven batch")
    rng = np.random.default_rng(22)
    p2 = tcfft_amd.Plan2D(1024, 1024, 2)
    pr = tcfft_amd.PlanR2C(8192, 64)
    sizes = (2 ** rng.integers(8, 17, 40)).astype(np.int32)
    mp = tcfft_amd.MultiPlan(sizes)
    x2 = fp16(rng, (2, 1024, 1024, 2))
    xr = fp16(rng, (64, 8192))
    xm = fp16(rng, (int(sizes.sum()), 2))
    w2, wr, wm = x2.clone(), torch.empty(64, 4097, 2, dtype=torch.float16, device="cuda"), xm.clone()
    p2.exec(w2.data_ptr())
    pr.exec_r2c(xr.data_ptr(), wr.data_ptr())
    mp.exec(wm.data_ptr())
    torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    for rep in range(3):
        a, r, m = x2.clone(), torch.empty_like(wr), xm.clone()
        torch.cuda.synchronize()
        p2.exec(a.data_ptr(), tcfft_amd.FORWARD, s1.cuda_stream)
        pr.exec_r2c(xr.data_ptr(), r.data_ptr(), s2.cuda_stream)
        mp.exec(m.data_ptr(), tcfft_amd.FORWARD, s2.cuda_stream)
        torch.cuda.synchronize()
        assert torch.equal(a, w2) and torch.equal(r, wr) and torch.equal(m, wm), rep
    p2.destroy()
    pr.destroy()
    mp.destroy()
