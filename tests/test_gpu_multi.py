"""Uneven batch in one persistent launch (tcfft_multi_*): signals of sizes
2^U{8..16} packed back to back, each digit-reversed by its own radix list,
checked per signal against FP64 (rel-L2 <= 1e-3), every schedule giving
bit-identical output."""
import sys

import numpy as np
import pytest

import oracle
import tcfft_amd

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def note(m):
    print(m, flush=True)
    sys.stdout.flush()


def make(n_signals, seed):
    rng = np.random.default_rng(seed)
    sizes = (2 ** rng.integers(8, 17, n_signals)).astype(np.int32)
    sigs = []
    for n in sizes:
        re = rng.uniform(-1, 1, n).astype(np.float16)
        im = rng.uniform(-1, 1, n).astype(np.float16)
        sigs.append(re.astype(np.float64) + 1j * im.astype(np.float64))
    return sizes, sigs


def pack_dropin(sizes, sigs):
    parts = []
    for n, x in zip(sizes, sigs):
        rad, _ = tcfft_amd.query_1d(int(n), 1)
        parts.append(x[tcfft_amd.gen_rev(int(n), rad)])
    flat = np.concatenate(parts)
    h = np.ascontiguousarray(np.stack([flat.real, flat.imag], -1).astype(np.float16))
    return torch.from_numpy(h).cuda()


def unpack(t, sizes):
    h = t.cpu().numpy().astype(np.float64)
    c = h[:, 0] + 1j * h[:, 1]
    out, o = [], 0
    for n in sizes:
        out.append(c[o:o + n])
        o += n
    return out


@pytest.mark.parametrize("direction,sign", [(tcfft_amd.FORWARD, -1), (tcfft_amd.INVERSE, +1)])
def test_multi_per_signal_parity(direction, sign):
    sizes, sigs = make(300, 7 if sign < 0 else 8)
    note(f"[gpu] multi {len(sizes)} signals, {int(sizes.sum())} elements, dir={direction}")
    plan = tcfft_amd.MultiPlan(sizes)
    d = pack_dropin(sizes, sigs)
    plan.exec(d.data_ptr(), direction)
    torch.cuda.synchronize()
    got = unpack(d, sizes)
    worst = 0.0
    for n, x, y in zip(sizes, sigs, got):
        ref = np.fft.fft(x) if sign < 0 else np.fft.ifft(x) * n
        worst = max(worst, oracle.rel_l2(y, ref))
    note(f"  worst per-signal rel_l2={worst:.3e} tasks={plan.n_tasks}")
    assert worst <= 1e-3
    plan.destroy()


def test_multi_schedules_identical_and_autotune():
    sizes, sigs = make(200, 9)
    plan = tcfft_amd.MultiPlan(sizes)
    ref = None
    for grid, order, delay in ((0, 0, 0), (64, 1, 0), (256, 2, 0), (7, 0, 0), (256, 0, 2), (512, 1, 5)):
        note(f"[gpu] multi schedule grid={grid} order={order} delay={delay}")
        plan.set_schedule(grid, order)
        plan.set_start_delay(delay)
        d = pack_dropin(sizes, sigs)
        plan.exec(d.data_ptr())
        torch.cuda.synchronize()
        out = d.cpu().numpy()
        if ref is None:
            ref = out
        else:
            np.testing.assert_array_equal(out, ref)
    scratch = pack_dropin(sizes, sigs)
    g, o, ms = plan.autotune(scratch.data_ptr())
    note(f"  autotune -> grid={g} order={o} delay={plan.schedule()[2]} {ms:.4f} ms")
    assert plan.schedule()[:2] == (g, o) and 0 <= plan.schedule()[2] <= 2
    d = pack_dropin(sizes, sigs)
    plan.exec(d.data_ptr())
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d.cpu().numpy(), ref)
    plan.destroy()


def test_multi_rejects_bad_start_delay():
    plan = tcfft_amd.MultiPlan([256, 4096])
    for bad in (-1, 65):
        with pytest.raises(tcfft_amd.TcfftError):
            plan.set_start_delay(bad)
    plan.destroy()


def test_multi_rejects_bad_sizes():
    with pytest.raises(tcfft_amd.TcfftError):
        tcfft_amd.MultiPlan([256, 100])
    with pytest.raises(tcfft_amd.TcfftError):
        tcfft_amd.MultiPlan([1 << 17])


def pack_natural(sizes, sigs):
    flat = np.concatenate(sigs)
    h = np.ascontiguousarray(np.stack([flat.real, flat.imag], -1).astype(np.float16))
    return torch.from_numpy(h).cuda()


@pytest.mark.parametrize("direction,sign", [(tcfft_amd.FORWARD, -1), (tcfft_amd.INVERSE, +1)])
def test_multi_natural_matches_dropin(direction, sign):
    """tcfft_multi_exec_natural on natural-order signals gives the same bits
    as tcfft_multi_exec on each signal permuted by its own gen_rev, out of
    place and in place, with every size class present (including a ragged
    last task per class); a sample of signals is also checked against FP64."""
    rng = np.random.default_rng(21 if sign < 0 else 22)
    sizes = np.concatenate([2 ** np.arange(8, 17), 2 ** rng.integers(8, 17, 150)]).astype(np.int32)
    rng.shuffle(sizes)
    sigs = []
    for n in sizes:
        sigs.append(rng.uniform(-1, 1, n).astype(np.float16).astype(np.float64)
                    + 1j * rng.uniform(-1, 1, n).astype(np.float16).astype(np.float64))
    note(f"[gpu] multi natural {len(sizes)} signals, {int(sizes.sum())} elements, dir={direction}")
    plan = tcfft_amd.MultiPlan(sizes)
    want = pack_dropin(sizes, sigs)
    plan.exec(want.data_ptr(), direction)
    nat = pack_natural(sizes, sigs)
    out = torch.empty_like(nat)
    plan.exec_natural(nat.data_ptr(), out.data_ptr(), direction)
    inplace = nat.clone()
    plan.exec_natural(inplace.data_ptr(), inplace.data_ptr(), direction)
    torch.cuda.synchronize()
    assert torch.equal(out, want)
    assert torch.equal(inplace, want)
    got = unpack(out, sizes)
    for i in range(0, len(sizes), 17):
        ref = oracle.fft(sigs[i][None, :], sign)[0]
        assert oracle.rel_l2(got[i], ref) <= 1e-3, (i, int(sizes[i]))
    for grid, order in ((64, 0), (0, 2)):
        plan.set_schedule(grid, order)
        o2 = torch.empty_like(nat)
        plan.exec_natural(nat.data_ptr(), o2.data_ptr(), direction)
        torch.cuda.synchronize()
        assert torch.equal(o2, want), (grid, order)
    plan.destroy()


def test_multi_natural_rejects_bad_pointers():
    plan = tcfft_amd.MultiPlan(np.array([256, 512], dtype=np.int32))
    t = torch.zeros(2 * 768 + 8, dtype=torch.float16, device="cuda")
    with pytest.raises(tcfft_amd.TcfftError):
        plan.exec_natural(t.data_ptr(), t.data_ptr() + 16)  # partial overlap
    with pytest.raises(tcfft_amd.TcfftError):
        plan.exec_natural(t.data_ptr() + 4, t.data_ptr())  # misaligned
    plan.destroy()
