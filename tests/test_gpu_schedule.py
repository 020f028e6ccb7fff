"""Start schedule of the first-pass kernels (plan options TCFFT_OPT_BIG_DELAY /
TCFFT_OPT_BIG_GROUPS and tcfft_plan_autotune): every schedule gives output
bit-identical to the unstaggered launch, the default and tuned schedules match
FP64 numpy (rel-L2 <= 1e-3 up to 2^16, 1.5e-3 above), and bad option values
are rejected."""
import sys

import numpy as np
import pytest

import tcfft_amd

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def note(m):
    print(m, flush=True)
    sys.stdout.flush()


def dropin_input(n, batch, seed):
    rng = np.random.default_rng(seed)
    x = (rng.uniform(-1, 1, (batch, n)).astype(np.float16).astype(np.float64)
         + 1j * rng.uniform(-1, 1, (batch, n)).astype(np.float16).astype(np.float64))
    rad, _ = tcfft_amd.query_1d(n, batch)
    xr = x[:, tcfft_amd.gen_rev(n, rad)]
    h = np.ascontiguousarray(np.stack([xr.real, xr.imag], -1).astype(np.float16))
    return x, torch.from_numpy(h).cuda()


def rel_l2(t, want):
    h = t.cpu().numpy().astype(np.float64)
    got = h[..., 0] + 1j * h[..., 1]
    return float(np.linalg.norm(got - want) / np.linalg.norm(want))


# sizes whose first pass is the single-pass kernel of each width, a longer
# transform whose first pass folds the next pass's twiddle, and contiguous-row
# first passes (one workgroup per CU and more)
@pytest.mark.parametrize("n,batch", [(8192, 300), (16384, 260), (32768, 270), (65536, 259), (1 << 20, 3),
                                     (4096, 1024), (1024, 300), (256, 4099)])
def test_schedules_bit_identical(n, batch):
    note(f"[gpu] schedule n={n} batch={batch}")
    x, d = dropin_input(n, batch, n + batch)
    plan = tcfft_amd.Plan1D(n, batch)
    plan.set_option(tcfft_amd.OPT_BIG_DELAY, 0)
    ref = d.clone()
    plan.exec(ref.data_ptr())
    torch.cuda.synchronize()
    tol = 1e-3 if n <= 65536 else 1.5e-3
    assert rel_l2(ref, np.fft.fft(x, axis=1)) <= tol
    for groups, delay in [(3, -1), (3, 1), (3, 3), (5, 2), (7, 1), (2, 4), (1, 5), (8, 64)]:
        plan.set_option(tcfft_amd.OPT_BIG_GROUPS, groups)
        plan.set_option(tcfft_amd.OPT_BIG_DELAY, delay)
        for direction in (tcfft_amd.FORWARD, tcfft_amd.INVERSE):
            t = d.clone()
            plan.exec(t.data_ptr(), direction)
            torch.cuda.synchronize()
            if direction == tcfft_amd.FORWARD:
                assert torch.equal(t, ref), (groups, delay)
            else:
                assert rel_l2(t, np.fft.ifft(x, axis=1) * n) <= tol
    plan.destroy()


@pytest.mark.parametrize("n,batch", [(65536, 256), (4096, 64)])
def test_autotune_then_parity(n, batch):
    note(f"[gpu] plan autotune n={n} batch={batch}")
    x, d = dropin_input(n, batch, 11)
    plan = tcfft_amd.Plan1D(n, batch)
    scratch = d.clone()
    groups, delay, ms = plan.autotune(scratch.data_ptr())
    torch.cuda.synchronize()
    assert ms > 0
    assert (groups, delay) in {(3, 0)} | {(g, dl) for g in (3, 5, 7) for dl in (1, 2, 3, 4)}
    t = d.clone()
    plan.exec(t.data_ptr())
    torch.cuda.synchronize()
    assert rel_l2(t, np.fft.fft(x, axis=1)) <= 1e-3
    plan.destroy()


@pytest.mark.parametrize("shape,batch", [((1024, 1024), 2), ((2048, 256), 3), ((512, 512), 4)])
def test_column_pass_delay_bit_identical(shape, batch):
    """Start delay of the radix-256*W column pass (2D nx = 512..2048): every
    value gives the same result bit for bit; the 2D autotuner keeps parity."""
    nx, ny = shape
    note(f"[gpu] column pass delay {nx}x{ny}x{batch}")
    rng = np.random.default_rng(nx + ny)
    x = torch.from_numpy(rng.uniform(-1, 1, (batch, nx, ny, 2)).astype(np.float16)).cuda()
    plan = tcfft_amd.Plan2D(nx, ny, batch)
    ref = x.clone()
    plan.set_option(tcfft_amd.OPT_COLS_DELAY, 0)
    plan.exec(ref.data_ptr())
    for d in (1, 2, 5, 64):
        plan.set_option(tcfft_amd.OPT_COLS_DELAY, d)
        t = x.clone()
        plan.exec(t.data_ptr())
        torch.cuda.synchronize()
        assert torch.equal(t, ref), d
    scratch = x.clone()
    plan.autotune(scratch.data_ptr())
    t = x.clone()
    plan.exec(t.data_ptr())
    torch.cuda.synchronize()
    assert torch.equal(t, ref)
    for bad in (-1, 65):
        with pytest.raises(tcfft_amd.TcfftError):
            plan.set_option(tcfft_amd.OPT_COLS_DELAY, bad)
    plan.destroy()


@pytest.mark.parametrize("n,batch", [(4096, 1024), (4096, 37), (4096, 8192), (1024, 300), (256, 4099), (1 << 20, 3)])
def test_rows_workgroup_size_bit_identical(n, batch):
    """Waves per workgroup of a contiguous-row first pass (TCFFT_OPT_ROWS_WAVES
    0 = by the pass, 1, 2, 4; 4096 x 8192 is a batch whose default is 4; 2^20
    with a 4096-point first pass) never change results; other values are
    rejected."""
    note(f"[gpu] rows workgroup size n={n} batch={batch}")
    x, d = dropin_input(n, batch, n + 3 * batch)
    plan = tcfft_amd.Plan1D(n, batch, first_log2=12) if n > 65536 else tcfft_amd.Plan1D(n, batch)
    if n > 65536:
        rad = plan.radices
        xr = x[:, tcfft_amd.gen_rev(n, rad)]
        d = torch.from_numpy(np.ascontiguousarray(np.stack([xr.real, xr.imag], -1).astype(np.float16))).cuda()
    ref = d.clone()
    plan.exec(ref.data_ptr())
    torch.cuda.synchronize()
    assert rel_l2(ref, np.fft.fft(x, axis=1)) <= (1e-3 if n <= 65536 else 1.5e-3)
    for wg in (1, 2, 4, 0):
        plan.set_option(tcfft_amd.OPT_ROWS_WAVES, wg)
        for delay in (0, 1):
            plan.set_option(tcfft_amd.OPT_BIG_DELAY, delay)
            t = d.clone()
            plan.exec(t.data_ptr())
            torch.cuda.synchronize()
            assert torch.equal(t, ref), (wg, delay)
    for bad in (-1, 3, 8):
        with pytest.raises(tcfft_amd.TcfftError):
            plan.set_option(tcfft_amd.OPT_ROWS_WAVES, bad)
    plan.destroy()


def test_bad_schedule_options():
    plan = tcfft_amd.Plan1D(65536, 1)
    for opt, val in [(tcfft_amd.OPT_BIG_DELAY, -2), (tcfft_amd.OPT_BIG_DELAY, 65), (tcfft_amd.OPT_BIG_GROUPS, 0),
                     (tcfft_amd.OPT_BIG_GROUPS, 9), (99, 1)]:
        with pytest.raises(tcfft_amd.TcfftError):
            plan.set_option(opt, val)
    d = torch.zeros(65536, 2, dtype=torch.float16, device="cuda")
    with pytest.raises(tcfft_amd.TcfftError):
        plan.autotune(d.data_ptr() + 4)
    plan.destroy()


# radix splits: a 4096-point first pass (row kernel with the folded twiddle of
# the strided passes that follow) against the default, and the tuned plan
@pytest.mark.parametrize("n,batch", [(8192, 17), (1 << 16, 5), (1 << 17, 3), (1 << 20, 2), (1 << 22, 1)])
def test_radix_split_parity(n, batch):
    note(f"[gpu] radix split n={n} batch={batch}")
    x, d = dropin_input(n, batch, 31 + n)
    tol = 1e-3 if n <= 65536 else 1.5e-3
    want = np.fft.fft(x, axis=1)
    for first_log2 in (12, 10):
        plan = tcfft_amd.Plan1D(n, batch, first_log2=first_log2)
        assert plan.radices == tcfft_amd.query_1d_split(n, batch, first_log2)[0]
        # the caller's digit reversal follows the plan's own radix list
        xr = x[:, tcfft_amd.gen_rev(n, plan.radices)]
        t = torch.from_numpy(np.ascontiguousarray(np.stack([xr.real, xr.imag], -1).astype(np.float16))).cuda()
        plan.exec(t.data_ptr())
        torch.cuda.synchronize()
        err = rel_l2(t, want)
        note(f"  first pass 2^{first_log2}: {plan.num_passes} passes, rel_l2={err:.3e}")
        assert err <= tol
        plan.destroy()
    plan = tcfft_amd.Plan1D(n, batch, tuned=True)
    note(f"  tuned: first pass 2^{plan.first_log2}, {plan.tune_ms:.4f} ms")
    assert plan.first_log2 in (12, 16) and plan.tune_ms > 0
    assert plan.radices == tcfft_amd.query_1d(n, batch)[0]  # both candidates share the default digit list
    t = d.clone()
    plan.exec(t.data_ptr())
    torch.cuda.synchronize()
    assert rel_l2(t, want) <= tol
    plan.destroy()



@pytest.mark.parametrize("n,batch", [(4096, 1024), (1024, 300)])
def test_natural_rows_schedule_bit_identical(n, batch):
    """The start schedule of the natural-order rows pass (TCFFT_OPT_BIG_DELAY)
    never changes results."""
    rng = np.random.default_rng(n)
    x = torch.from_numpy(rng.uniform(-1, 1, (batch, n, 2)).astype(np.float16)).cuda()
    plan = tcfft_amd.Plan1D(n, batch)
    plan.set_option(tcfft_amd.OPT_BIG_DELAY, 0)
    ref = torch.empty_like(x)
    plan.exec_natural(x.data_ptr(), ref.data_ptr())
    for d in (-1, 1, 3):
        plan.set_option(tcfft_amd.OPT_BIG_DELAY, d)
        out = torch.empty_like(x)
        plan.exec_natural(x.data_ptr(), out.data_ptr())
        torch.cuda.synchronize()
        assert torch.equal(out, ref), d
    plan.destroy()


@pytest.mark.parametrize("shape", [("1d", 1 << 20, 4, 16), ("1d", 1 << 18, 8, 12), ("2d", 1024, 512, 2),
                                   ("2d", 4096, 256, 1), ("1d", 65536, 4, 16)])
def test_first_pass_cached_bit_identical(shape):
    """Cached or non-temporal first-pass stores (TCFFT_OPT_FIRST_CACHED -1 =
    by buffer size, 0, 1) give bit-identical output, on digit-reversed and on
    natural-order input; 1D natural output matches FP64 numpy; values outside
    -1..1 are rejected."""
    kind, a, b, c = shape
    note(f"[gpu] first pass cached {shape}")
    rng = np.random.default_rng(a + b)
    if kind == "1d":
        plan = tcfft_amd.Plan1D(a, b, first_log2=c) if a > 65536 else tcfft_amd.Plan1D(a, b)
        total = a * b
    else:
        plan = tcfft_amd.Plan2D(a, b, c)
        total = a * b * c
    x = torch.from_numpy(rng.uniform(-1, 1, (total, 2)).astype(np.float16)).cuda()
    outs = {}
    for v in (-1, 0, 1):
        plan.set_option(tcfft_amd.OPT_FIRST_CACHED, v)
        t = x.clone()
        plan.exec(t.data_ptr())
        o = torch.empty_like(x)
        plan.exec_natural(x.data_ptr(), o.data_ptr())
        torch.cuda.synchronize()
        outs[v] = (t, o)
    for v in (0, 1):
        assert torch.equal(outs[v][0], outs[-1][0]), v
        assert torch.equal(outs[v][1], outs[-1][1]), v
    if kind == "1d":
        h = x.cpu().numpy().astype(np.float64).reshape(b, a, 2)
        want = np.fft.fft(h[..., 0] + 1j * h[..., 1], axis=1)
        assert rel_l2(outs[-1][1].view(b, a, 2), want) <= (1e-3 if a <= 65536 else 1.5e-3)
    for bad in (-2, 2):
        with pytest.raises(tcfft_amd.TcfftError):
            plan.set_option(tcfft_amd.OPT_FIRST_CACHED, bad)
    plan.destroy()


# (kind, n or nx, ny, batch, chunk MiB[, first pass log2]): ragged last chunks
# (2^20 x 5 in 3 + 2 signals, 2D 256^2 x 300 in five chunks of 60 images), one
# signal per chunk (2^22 x 3), a 2D shape with a k_cols_big pass (never
# chunked), one whose rows are the 65536-point kernel, and three-pass plans
# whose middle column pass folds the next pass's twiddle (2D 65536 x 256: two
# radix-256 column passes; 1D 2^24 split at 4096 points), one signal per chunk
CHUNK_CASES = [("1d", 1 << 20, 0, 5, 16), ("1d", 1 << 22, 0, 3, 16), ("2d", 256, 256, 300, 16),
               ("2d", 1024, 1024, 20, 16), ("2d", 64, 65536, 9, 16), ("1d", 1 << 17, 0, 40, 16),
               ("2d", 65536, 256, 2, 64), ("1d", 1 << 24, 0, 3, 64, 12)]


@pytest.mark.parametrize("case", CHUNK_CASES)
def test_mall_chunks_bit_identical(case):
    """Batch chunks (TCFFT_OPT_MALL_CHUNK) give output bit-identical to the
    whole batch at once, for exec (both directions) and exec_natural (in and
    out of place); 0 and 16..65536 are accepted, other values rejected."""
    kind, a, b, batch, mib = case[:5]
    note(f"[gpu] mall chunks {case}")
    if kind == "1d":
        plan = tcfft_amd.Plan1D(a, batch, first_log2=case[5]) if len(case) > 5 else tcfft_amd.Plan1D(a, batch)
    else:
        plan = tcfft_amd.Plan2D(a, b, batch)
    if len(case) > 5 or a == 65536:
        assert plan.num_passes == 3, plan.num_passes
    total = a * (b or 1) * batch
    rng = np.random.default_rng(total % 1000)
    x = torch.from_numpy(rng.uniform(-1, 1, (total, 2)).astype(np.float16)).cuda()
    outs = []
    for chunk in (0, mib):
        plan.set_option(tcfft_amd.OPT_MALL_CHUNK, chunk)
        res = []
        for direction in (tcfft_amd.FORWARD, tcfft_amd.INVERSE):
            t = x.clone()
            plan.exec(t.data_ptr(), direction)
            res.append(t)
        o = torch.empty_like(x)
        plan.exec_natural(x.data_ptr(), o.data_ptr())
        res.append(o)
        t = x.clone()
        plan.exec_natural(t.data_ptr(), t.data_ptr(), tcfft_amd.INVERSE)
        res.append(t)
        torch.cuda.synchronize()
        outs.append(res)
    for i, (u, v) in enumerate(zip(*outs)):
        assert torch.equal(u, v), (case, i)
    # tcfft_exec_passes runs each pass over the whole batch: pass by pass it
    # equals the chunked exec bit for bit
    t = x.clone()
    for p in range(plan.num_passes):
        plan.exec_passes(t.data_ptr(), p, 1)
    torch.cuda.synchronize()
    assert torch.equal(t, outs[1][0]), case
    if kind == "1d" and a <= (1 << 20):
        h = x.cpu().numpy().astype(np.float64).reshape(batch, a, 2)
        want = np.fft.fft(h[..., 0] + 1j * h[..., 1], axis=1)
        assert rel_l2(outs[1][2].view(batch, a, 2), want) <= 1.5e-3
    for bad in (-1, 1, 15, 65537):
        with pytest.raises(tcfft_amd.TcfftError):
            plan.set_option(tcfft_amd.OPT_MALL_CHUNK, bad)
    plan.destroy()


@pytest.mark.parametrize("n,batch", [(1024, 1024), (256, 300), (2048, 64), (512, 4096), (1 << 18, 3),
                                     (4096, 1024), (4096, 3), (4096, 2048)])
def test_rows_tiles_per_wave_bit_identical(n, batch):
    """Tiles per wave of a contiguous-row first pass (TCFFT_OPT_ROWS_TILES 0 =
    by the grid, 4, 8, 16; 2048-point lines take at least 8; 4096-point lines
    16, or 8 / 4 = two / four waves per line, k_rows_half, with one or two
    lines per workgroup and an odd line count; a first pass with a folded twiddle, 2^18
    split at 1024 points, always 16) never change results, with any
    workgroup size, and after tcfft_plan_autotune (which puts the caller's
    data back); other values are rejected."""
    note(f"[gpu] rows tiles per wave n={n} batch={batch}")
    x, d = dropin_input(n, batch, n + 7 * batch)
    plan = tcfft_amd.Plan1D(n, batch, first_log2=10) if n > 65536 else tcfft_amd.Plan1D(n, batch)
    if n > 65536:
        xr = x[:, tcfft_amd.gen_rev(n, plan.radices)]
        d = torch.from_numpy(np.ascontiguousarray(np.stack([xr.real, xr.imag], -1).astype(np.float16))).cuda()
    ref = d.clone()
    plan.exec(ref.data_ptr())
    torch.cuda.synchronize()
    assert rel_l2(ref, np.fft.fft(x, axis=1)) <= (1e-3 if n <= 65536 else 1.5e-3)
    for tiles in (16, 8, 4, 0):
        plan.set_option(tcfft_amd.OPT_ROWS_TILES, tiles)
        for wg in (1, 4):
            plan.set_option(tcfft_amd.OPT_ROWS_WAVES, wg)
            for direction in (tcfft_amd.FORWARD, tcfft_amd.INVERSE):
                t = d.clone()
                plan.exec(t.data_ptr(), direction)
                torch.cuda.synchronize()
                if direction == tcfft_amd.FORWARD:
                    assert torch.equal(t, ref), (tiles, wg)
                else:
                    if tiles == 16 and wg == 1:
                        inv_ref = t
                    assert torch.equal(t, inv_ref), (tiles, wg)
    plan.set_option(tcfft_amd.OPT_ROWS_WAVES, 0)
    plan.set_option(tcfft_amd.OPT_ROWS_TILES, 0)
    scratch = d.clone()
    plan.autotune(scratch.data_ptr())
    assert torch.equal(scratch, d)  # the tuner puts the caller's data back
    t = d.clone()
    plan.exec(t.data_ptr())
    torch.cuda.synchronize()
    assert torch.equal(t, ref)
    for bad in (-1, 2, 12, 32):
        with pytest.raises(tcfft_amd.TcfftError):
            plan.set_option(tcfft_amd.OPT_ROWS_TILES, bad)
    plan.destroy()


def test_rows_half_cached_first_pass_bit_identical():
    """Two or four waves per 4096-point line as the first pass of a 2D plan
    (256 x 4096 x 3: its output stores go through the caches for the column
    pass that follows), one or two lines per workgroup, equal one wave per
    line, forward and inverse."""
    note("[gpu] rows half, 2D 256 x 4096 x 3")
    rng = np.random.default_rng(4096)
    x = torch.from_numpy(rng.uniform(-1, 1, (3, 256, 4096, 2)).astype(np.float16)).cuda()
    plan = tcfft_amd.Plan2D(256, 4096, 3)
    for direction in (tcfft_amd.FORWARD, tcfft_amd.INVERSE):
        outs = []
        for tiles in (16, 8, 4):
            plan.set_option(tcfft_amd.OPT_ROWS_TILES, tiles)
            for wg in (1, 4):
                plan.set_option(tcfft_amd.OPT_ROWS_WAVES, wg)
                t = x.clone()
                plan.exec(t.data_ptr(), direction)
                torch.cuda.synchronize()
                outs.append(t)
        for t in outs[1:]:
            assert torch.equal(t, outs[0]), direction
    plan.destroy()


@pytest.mark.parametrize("shape", [(4096, 0, 5), (4096, 0, 1024), (8, 4096, 3)], ids=["1d-4096x5", "1d-4096x1024",
                                                                                   "2d-8x4096x3"])
def test_rows_half_natural_bit_identical(shape):
    """Natural-order input on 4096-point lines with two or four waves per
    line (k_rows_half_nat: TCFFT_OPT_ROWS_TILES 8 or 4; the default picks
    one of them) equals one wave per line (16) and the drop-in path on host-permuted
    input, bit for bit: 1D out of place and in place, 2D out of place and in
    place (through the plan's staging copy), forward and inverse."""
    a, b, batch = shape
    note(f"[gpu] rows half natural {shape}")
    rng = np.random.default_rng(a + b + batch)
    if b:
        plan = tcfft_amd.Plan2D(a, b, batch)
        x = rng.uniform(-1, 1, (batch, a, b, 2)).astype(np.float16)
        xr = x[:, tcfft_amd.gen_rev(a, plan.radices_x)][:, :, tcfft_amd.gen_rev(b, plan.radices_y)]
    else:
        plan = tcfft_amd.Plan1D(a, batch)
        x = rng.uniform(-1, 1, (batch, a, 2)).astype(np.float16)
        xr = x[:, tcfft_amd.gen_rev(a, plan.radices)]
    src = torch.from_numpy(x).cuda()
    for direction in (tcfft_amd.FORWARD, tcfft_amd.INVERSE):
        want = torch.from_numpy(np.ascontiguousarray(xr)).cuda()
        plan.exec(want.data_ptr(), direction)
        for tiles in (16, 8, 4, 0):
            plan.set_option(tcfft_amd.OPT_ROWS_TILES, tiles)
            out = torch.empty_like(src)
            plan.exec_natural(src.data_ptr(), out.data_ptr(), direction)
            inp = src.clone()
            plan.exec_natural(inp.data_ptr(), inp.data_ptr(), direction)
            torch.cuda.synchronize()
            assert torch.equal(out, want), (tiles, direction)
            assert torch.equal(inp, want), (tiles, direction)
    plan.set_option(tcfft_amd.OPT_ROWS_TILES, 0)
    plan.destroy()
