"""hipGraph capture of the entry points that allocate nothing: the drop-in
exec (1D and 2D), natural order (out of place, and in place through the
plan's workspace), transforms run in batch chunks, the uneven batch, split
precision, the real-input transforms and the fused convolution.  Each graph is replayed three times with
eager work (input restore, an eager transform of another buffer) between the
replays, and every replay must equal the eager result bit for bit."""
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


def replay_check(launch, reset, out, want, other):
    """Capture launch(stream) once; then three times: reset inputs eagerly,
    run an eager transform of `other`, replay, compare `out` with `want`."""
    g = torch.cuda.CUDAGraph()
    cs = torch.cuda.Stream()
    reset()
    cs.wait_stream(torch.cuda.current_stream())
    with torch.cuda.graph(g, stream=cs):
        launch(cs.cuda_stream)
    torch.cuda.synchronize()
    for r in range(3):
        reset()
        other()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, want), f"replay {r}"


def test_graph_dropin_1d_and_2d():
    rng = np.random.default_rng(1)
    note("[gpu] graph: drop-in 1D 2^16 x 8 and 2D 512 x 1024 x 2")
    p1 = tcfft_amd.Plan1D(1 << 16, 8)
    p2 = tcfft_amd.Plan2D(512, 1024, 2)
    x1, x2 = fp16(rng, (8, 1 << 16, 2)), fp16(rng, (2, 512, 1024, 2))
    w1, w2 = x1.clone(), x2.clone()
    p1.exec(w1.data_ptr())
    p2.exec(w2.data_ptr())
    torch.cuda.synchronize()
    b1, b2 = x1.clone(), x2.clone()
    scratch = x1.clone()

    def reset():
        b1.copy_(x1)
        b2.copy_(x2)

    def launch(st):
        p1.exec(b1.data_ptr(), tcfft_amd.FORWARD, st)
        p2.exec(b2.data_ptr(), tcfft_amd.FORWARD, st)

    def other():
        scratch.copy_(x1)
        p1.exec(scratch.data_ptr())

    replay_check(launch, reset, b1, w1, other)
    assert torch.equal(b2, w2)
    p1.destroy()
    p2.destroy()


@pytest.mark.parametrize("n,batch", [(4096, 16), (1 << 16, 4)])
def test_graph_natural_out_of_place(n, batch):
    note(f"[gpu] graph: natural order n={n} batch={batch}")
    rng = np.random.default_rng(n)
    plan = tcfft_amd.Plan1D(n, batch)
    x = fp16(rng, (batch, n, 2))
    want = torch.empty_like(x)
    plan.exec_natural(x.data_ptr(), want.data_ptr())
    torch.cuda.synchronize()
    src, dst = x.clone(), torch.empty_like(x)

    def reset():
        src.copy_(x)
        dst.zero_()

    replay_check(lambda st: plan.exec_natural(src.data_ptr(), dst.data_ptr(), tcfft_amd.FORWARD, st), reset, dst,
                 want, lambda: plan.exec(torch.empty_like(x).copy_(x).data_ptr()))
    plan.destroy()


@pytest.mark.parametrize("shape", [(2, 512, 1024), (2, 1 << 20)])
def test_graph_natural_in_place(shape):
    """In-place natural order stages the input through a workspace that the
    plan allocated at creation (2D, and 1D lines above 65536 points), so the
    call allocates nothing and captures into a graph; every replay equals
    the out-of-place result."""
    note(f"[gpu] graph: natural order in place {shape}")
    rng = np.random.default_rng(sum(shape))
    plan = tcfft_amd.Plan2D(*shape[1:], shape[0]) if len(shape) == 3 else tcfft_amd.Plan1D(shape[1], shape[0])
    x = fp16(rng, shape + (2,))
    want = torch.empty_like(x)
    plan.exec_natural(x.data_ptr(), want.data_ptr())
    torch.cuda.synchronize()
    b = x.clone()
    scratch = torch.empty_like(x)
    replay_check(lambda st: plan.exec_natural(b.data_ptr(), b.data_ptr(), tcfft_amd.FORWARD, st),
                 lambda: b.copy_(x), b, want,
                 lambda: plan.exec_natural(scratch.copy_(x).data_ptr(), scratch.data_ptr()))
    plan.destroy()


@pytest.mark.parametrize("shape", [(5, 1 << 20), (2, 65536, 256)], ids=["1d-2^20x5", "2d-65536x256x2"])
def test_graph_batch_chunks(shape):
    """A transform run in batch chunks (TCFFT_OPT_MALL_CHUNK at 16 MiB: 3 + 2
    signals of 2^20, one 2D image of 64 MiB per chunk at 64 MiB) captures as
    one graph of all chunks' launches, drop-in and natural order; every
    replay equals the eager whole-batch result."""
    note(f"[gpu] graph: batch chunks {shape}")
    rng = np.random.default_rng(sum(shape))
    if len(shape) == 2:
        plan, mib = tcfft_amd.Plan1D(shape[1], shape[0]), 16
    else:
        plan, mib = tcfft_amd.Plan2D(shape[1], shape[2], shape[0]), 64
    x = fp16(rng, shape + (2,))
    plan.set_option(tcfft_amd.OPT_MALL_CHUNK, 0)
    want = x.clone()
    plan.exec(want.data_ptr())
    want_nat = torch.empty_like(x)
    plan.exec_natural(x.data_ptr(), want_nat.data_ptr())
    torch.cuda.synchronize()
    plan.set_option(tcfft_amd.OPT_MALL_CHUNK, mib)
    b, nat = x.clone(), torch.empty_like(x)
    scratch = torch.empty_like(x)

    def reset():
        b.copy_(x)
        nat.zero_()

    def launch(st):
        plan.exec(b.data_ptr(), tcfft_amd.FORWARD, st)
        plan.exec_natural(x.data_ptr(), nat.data_ptr(), tcfft_amd.FORWARD, st)

    replay_check(launch, reset, b, want, lambda: plan.exec(scratch.copy_(x).data_ptr()))
    assert torch.equal(nat, want_nat)
    plan.destroy()


def test_graph_uneven_batch():
    note("[gpu] graph: uneven batch")
    rng = np.random.default_rng(5)
    sizes = (2 ** rng.integers(8, 17, 120)).astype(np.int32)
    mp = tcfft_amd.MultiPlan(sizes)
    total = int(sizes.sum())
    x = fp16(rng, (total, 2))
    want = x.clone()
    mp.exec(want.data_ptr())
    torch.cuda.synchronize()
    b = x.clone()
    scratch = x.clone()
    replay_check(lambda st: mp.exec(b.data_ptr(), tcfft_amd.FORWARD, st), lambda: b.copy_(x), b, want,
                 lambda: mp.exec(scratch.copy_(x).data_ptr()))
    mp.destroy()


def test_graph_uneven_batch_natural():
    note("[gpu] graph: uneven batch, natural order")
    rng = np.random.default_rng(6)
    sizes = (2 ** rng.integers(8, 17, 90)).astype(np.int32)
    mp = tcfft_amd.MultiPlan(sizes)
    total = int(sizes.sum())
    x = fp16(rng, (total, 2))
    want = torch.empty_like(x)
    mp.exec_natural(x.data_ptr(), want.data_ptr())
    torch.cuda.synchronize()
    out = torch.empty_like(x)
    scratch = torch.empty_like(x)
    replay_check(lambda st: mp.exec_natural(x.data_ptr(), out.data_ptr(), tcfft_amd.FORWARD, st),
                 lambda: out.zero_(), out, want, lambda: mp.exec_natural(x.data_ptr(), scratch.data_ptr()))
    mp.destroy()


@pytest.mark.parametrize("chunk", [0, 3])
def test_graph_split_precision(chunk):
    """chunk 3 of 8 signals: three pass pairs (the last one ragged);
    chunk 0: the whole batch in one pass pair."""
    note(f"[gpu] graph: split precision 2^16 x 8, chunk {chunk}")
    rng = np.random.default_rng(7)
    plan = tcfft_amd.Plan1D(1 << 16, 8)
    plan.set_option(tcfft_amd.OPT_SPLIT_CHUNK, chunk)
    x = torch.from_numpy(rng.standard_normal((8, 1 << 16, 2)).astype(np.float32)).cuda()
    want = torch.empty_like(x)
    plan.exec_split(x.data_ptr(), want.data_ptr())
    torch.cuda.synchronize()
    out = torch.empty_like(x)
    replay_check(lambda st: plan.exec_split(x.data_ptr(), out.data_ptr(), tcfft_amd.FORWARD, st),
                 lambda: out.zero_(), out, want, lambda: plan.exec_split(x.data_ptr(), torch.empty_like(x).data_ptr()))
    plan.destroy()


@pytest.mark.parametrize("shape", [(1, 1 << 20, 0, 3), (2, 512, 1024, 2)], ids=["1d-2^20x3", "2d-512x1024x2"])
def test_graph_split_precision_prepared(shape):
    """Prepared split plans (three or four column passes through the plan's
    workspace), captured in place: every replay restores the input first."""
    dims, a, b, batch = shape
    note(f"[gpu] graph: prepared split {shape}")
    rng = np.random.default_rng(8)
    plan = tcfft_amd.Plan1D(a, batch) if dims == 1 else tcfft_amd.Plan2D(a, b, batch)
    plan.prepare_split()
    total = a * (b or 1) * batch
    x = torch.from_numpy(rng.standard_normal((total, 2)).astype(np.float32)).cuda()
    want = x.clone()
    plan.exec_split(want.data_ptr(), want.data_ptr())
    torch.cuda.synchronize()
    buf = torch.empty_like(x)
    replay_check(lambda st: plan.exec_split(buf.data_ptr(), buf.data_ptr(), tcfft_amd.FORWARD, st),
                 lambda: buf.copy_(x), buf, want, lambda: plan.exec_split(x.data_ptr(), torch.empty_like(x).data_ptr()))
    plan.destroy()


def test_graph_fused_convolution():
    note("[gpu] graph: fused convolution n=8192, 4 x 8 channels")
    rng = np.random.default_rng(9)
    n, ch, b = 8192, 8, 4
    plan = tcfft_amd.Plan1D(n, b * ch)
    x = fp16(rng, (b, ch, n, 2))
    h = (fp16(rng, (ch, n, 2)) * 0.02).contiguous()
    want, spec_w = torch.empty_like(x), torch.empty_like(x)
    plan.conv_exec(ch, x.data_ptr(), h.data_ptr(), want.data_ptr(), False, spec_w.data_ptr())
    torch.cuda.synchronize()
    y, spec = torch.empty_like(x), torch.empty_like(x)

    def reset():
        y.zero_()
        spec.zero_()

    replay_check(lambda st: plan.conv_exec(ch, x.data_ptr(), h.data_ptr(), y.data_ptr(), False, spec.data_ptr(), st),
                 reset, y, want, lambda: plan.exec_dif(x.data_ptr(), torch.empty_like(x).data_ptr()))
    assert torch.equal(spec, spec_w)
    plan.destroy()


def test_graph_real_input():
    note("[gpu] graph: r2c then c2r, n=8192 x 12")
    rng = np.random.default_rng(11)
    n, batch = 8192, 12
    plan = tcfft_amd.PlanR2C(n, batch)
    x = fp16(rng, (batch, n))
    spec_w = torch.empty(batch, n // 2 + 1, 2, dtype=torch.float16, device="cuda")
    back_w = torch.empty_like(x)
    plan.exec_r2c(x.data_ptr(), spec_w.data_ptr())
    plan.exec_c2r(spec_w.data_ptr(), back_w.data_ptr())
    torch.cuda.synchronize()
    spec, back = torch.empty_like(spec_w), torch.empty_like(x)

    def reset():
        spec.zero_()
        back.zero_()

    def launch(st):
        plan.exec_r2c(x.data_ptr(), spec.data_ptr(), st)
        plan.exec_c2r(spec.data_ptr(), back.data_ptr(), st)

    replay_check(launch, reset, back, back_w,
                 lambda: plan.exec_r2c(x.data_ptr(), torch.empty_like(spec_w).data_ptr()))
    assert torch.equal(spec, spec_w)
    plan.destroy()


def test_graph_real_input_two_step():
    """Real input above 65536 points (passes + separate split through the
    plan's workspace): captured R2C then C2R replays bit-identically."""
    note("[gpu] graph: r2c then c2r, n=2^17 x 2")
    rng = np.random.default_rng(12)
    n, batch = 1 << 17, 2
    plan = tcfft_amd.PlanR2C(n, batch)
    x = fp16(rng, (batch, n)) * 0.125
    spec_w = torch.empty(batch, n // 2 + 1, 2, dtype=torch.float16, device="cuda")
    back_w = torch.empty_like(x)
    plan.exec_r2c(x.data_ptr(), spec_w.data_ptr())
    plan.exec_c2r(spec_w.data_ptr(), back_w.data_ptr())
    torch.cuda.synchronize()
    spec, back = torch.empty_like(spec_w), torch.empty_like(x)

    def reset():
        spec.zero_()
        back.zero_()

    def launch(st):
        plan.exec_r2c(x.data_ptr(), spec.data_ptr(), st)
        plan.exec_c2r(spec.data_ptr(), back.data_ptr(), st)

    replay_check(launch, reset, back, back_w, lambda: None)
    assert torch.equal(spec, spec_w)
    plan.destroy()
