"""CPU-only checks of the C ABI: the library loads, exports every symbol that
include/*.h declares, and its host-side logic (validation, plan radix lists,
gen_rev) is right.  No kernel runs here."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle
import tcfft_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    syms = set()
    for name in os.listdir(os.path.join(ROOT, "include")):
        if not name.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", name)).read()
        if 'extern "C"' not in text:
            continue
        body = text.split('extern "C"', 1)[1]
        syms |= set(re.findall(r"\b(tcfft_[a-z0-9_]+)\s*\(", body))
    return syms


def test_library_loads_and_exports_declared_symbols():
    L = tcfft_amd.lib()
    decl = declared_symbols()
    assert decl, "no extern \"C\" declarations found"
    assert decl == set(tcfft_amd.EXPORTS)
    for s in decl:
        assert hasattr(L, s), s
    assert tcfft_amd.version() == 100
    assert L.tcfft_status_string(1) == b"TCFFT_INVALID_SIZE"


@pytest.mark.parametrize("n", [0, 100, 128, 255, 3 << 10, 1 << 28])
def test_create_rejects_bad_sizes_without_device(n):
    h = ctypes.c_void_p()
    assert tcfft_amd.lib().tcfft_plan_create_1d(ctypes.byref(h), n, 1) == 1
    assert h.value is None


def test_create_rejects_bad_batch_and_pointers():
    L = tcfft_amd.lib()
    h = ctypes.c_void_p()
    assert L.tcfft_plan_create_1d(ctypes.byref(h), 256, 0) == 2
    assert L.tcfft_plan_create_1d(ctypes.byref(h), 1 << 20, 1 << 11) == 2  # 2^31 elements
    assert L.tcfft_plan_create_1d(None, 256, 1) == 3
    assert L.tcfft_plan_create_2d(ctypes.byref(h), 256, 128, 1) == 1   # ny < 256
    assert L.tcfft_plan_create_2d(ctypes.byref(h), 3, 256, 1) == 1
    assert L.tcfft_exec(None, ctypes.c_void_p(0x1000), -1, None) == 4
    assert L.tcfft_exec_natural(None, None, None, -1, None) == 4
    assert L.tcfft_plan_destroy(None) == 4
    # a non-plan pointer is rejected by the magic check
    junk = (ctypes.c_uint32 * 64)()
    assert L.tcfft_exec(ctypes.cast(junk, ctypes.c_void_p), ctypes.c_void_p(0x1000), -1, None) == 4


def test_newer_entry_points_reject_bad_arguments_without_device():
    """Radix-split creation, tuning, per-pass execution, schedules and the
    uneven-batch setters validate their arguments before any device call."""
    L = tcfft_amd.lib()
    h = ctypes.c_void_p()
    assert L.tcfft_plan_create_1d_split(ctypes.byref(h), 1 << 20, 1, 14) == 5   # no such first pass
    assert L.tcfft_plan_create_1d_split(ctypes.byref(h), 100, 1, 16) == 1
    assert L.tcfft_plan_create_1d_split(None, 1 << 20, 1, 16) == 3
    assert L.tcfft_plan_create_1d_tuned(ctypes.byref(h), 3 << 12, 1, None, None) == 1
    assert L.tcfft_plan_create_1d_tuned(ctypes.byref(h), 1 << 20, 0, None, None) == 2
    assert h.value is None
    assert L.tcfft_exec_passes(None, ctypes.c_void_p(0x1000), -1, 0, 1, None) == 4
    assert L.tcfft_plan_autotune(None, ctypes.c_void_p(0x1000), None, None, None, None) == 4
    assert L.tcfft_plan_set_option(None, 2, 1) == 4
    assert L.tcfft_multi_set_start_delay(None, 1) == 4
    assert L.tcfft_multi_get_schedule(None, None, None, None) == 4
    assert L.tcfft_multi_set_schedule(None, 256, 0) == 4
    assert L.tcfft_multi_exec(None, ctypes.c_void_p(0x1000), -1, None) == 4
    assert L.tcfft_multi_autotune(None, ctypes.c_void_p(0x1000), None, None, None, None) == 4
    rad = (ctypes.c_int * 32)()
    nr, npass = ctypes.c_int(), ctypes.c_int()
    assert L.tcfft_query_1d_split(1 << 20, 1, 13, rad, 32, ctypes.byref(nr), ctypes.byref(npass)) == 5
    assert L.tcfft_query_1d_split(1 << 20, 1, 12, rad, 32, ctypes.byref(nr), ctypes.byref(npass)) == 0
    assert npass.value == 2 and list(rad[: nr.value]) == [16] * 5


@pytest.mark.parametrize("n", [256, 512, 1024, 2048, 4096, 8192, 1 << 16, 1 << 17, 1 << 20, 1 << 24, 1 << 27])
def test_query_radix_lists_reproduce_the_dft(n):
    rad, npass = tcfft_amd.query_1d(n, 1)
    assert int(np.prod(rad)) == n and len(rad) <= 9   # fits tcfftHandle::radices[9]
    assert all(r in (2, 4, 8, 16) for r in rad)
    if n <= (1 << 17):
        # gen_rev(plan.radices) + DIT over that list = natural DFT (oracle, FP64)
        x = np.random.default_rng(n).standard_normal((1, n)) + 0j
        rev = oracle.gen_rev(n, rad)
        got = oracle.dit(x[:, rev], rad)
        assert oracle.rel_l2(got, np.fft.fft(x, axis=-1)) < 1e-12


@pytest.mark.parametrize("n", [8192, 1 << 16, 1 << 17, 1 << 20, 1 << 24, 1 << 27])
@pytest.mark.parametrize("first_log2", [8, 10, 12, 16])
def test_radix_splits(n, first_log2):
    """Every radix split of a 1D plan: its list reproduces the DFT through
    gen_rev (oracle, FP64) and at most 9 radices; a first pass of 2^first_log2
    points plus strided passes."""
    rad, npass = tcfft_amd.query_1d_split(n, 1, first_log2)
    assert int(np.prod(rad)) == n and len(rad) <= 9 and all(r in (2, 4, 8, 16) for r in rad)
    L = min(n, 1 << first_log2)
    rest = n.bit_length() - L.bit_length()
    assert npass == 1 + (0 if rest == 0 else 1 if rest <= 11 else -(-rest // 8))
    if n <= (1 << 17):
        x = np.random.default_rng(first_log2).standard_normal((1, n)) + 0j
        got = oracle.dit(x[:, oracle.gen_rev(n, rad)], rad)
        assert oracle.rel_l2(got, np.fft.fft(x, axis=-1)) < 1e-12


def test_radix_split_rejects_unsupported_first_pass():
    for bad in (7, 13, 14, 15, 17):
        with pytest.raises(tcfft_amd.TcfftError):
            tcfft_amd.query_1d_split(1 << 20, 1, bad)


@pytest.mark.parametrize("shape", [(256, 256), (512, 1024), (1024, 1024), (16, 256), (2, 512), (4096, 4096)])
def test_query_2d_radix_lists(shape):
    nx, ny = shape
    rx, ry, npass = tcfft_amd.query_2d(nx, ny, 1)
    assert int(np.prod(rx)) == nx and int(np.prod(ry)) == ny
    if nx * ny <= 512 * 1024:
        x = np.random.default_rng(1).standard_normal((1, nx, ny)) + 0j
        revx, revy = oracle.gen_rev(nx, rx), oracle.gen_rev(ny, ry)
        got = oracle.dit_2d(x[:, revx][:, :, revy], rx, ry)
        assert oracle.rel_l2(got, np.fft.fft2(x, axes=(-2, -1))) < 1e-12


def test_reference_lists_where_they_exist():
    """For the sizes the reference lists with 256/512/1024 first passes the
    2D per-axis lists coincide with tcfft_half_2d.cu's."""
    for nx in (256, 512):
        rx, _, _ = tcfft_amd.query_2d(nx, 256, 1)
        ok, ref_rx, _, _ = oracle.ref_plan_2d(nx, 256)
        assert ok and rx == ref_rx


@pytest.mark.parametrize("radices", [[16, 16], [16, 16, 2], [16, 16, 16, 16, 8], [2, 4, 8], [16, 16, 4, 16, 16, 2]])
def test_gen_rev_matches_oracle(radices):
    n = int(np.prod(radices))
    np.testing.assert_array_equal(tcfft_amd.gen_rev(n, radices), oracle.gen_rev(n, radices))


def test_gen_rev_rejects_bad_lists():
    L = tcfft_amd.lib()
    r = (ctypes.c_int * 2)(16, 16)
    out = (ctypes.c_int * 256)()
    assert L.tcfft_gen_rev(512, r, 2, out) == 1
    assert L.tcfft_gen_rev(256, None, 2, out) == 3


def test_real_input_entry_points_validate_without_device():
    L = tcfft_amd.lib()
    h = ctypes.c_void_p()
    for bad in (0, 256, 1000, 3 << 12, 1 << 28):
        assert L.tcfft_plan_create_r2c_1d(ctypes.byref(h), bad, 1) == 1
    assert L.tcfft_plan_create_r2c_1d(ctypes.byref(h), 1024, 0) == 2
    assert L.tcfft_plan_create_r2c_1d(None, 1024, 1) == 3
    assert h.value is None
    assert L.tcfft_exec_r2c(None, ctypes.c_void_p(0x1000), ctypes.c_void_p(0x2000), None) == 4
    assert L.tcfft_exec_c2r(None, ctypes.c_void_p(0x1000), ctypes.c_void_p(0x2000), None) == 4


def test_permute_to_dropin_validates_without_device():
    L = tcfft_amd.lib()
    assert L.tcfft_permute_to_dropin(None, ctypes.c_void_p(0x1000), ctypes.c_void_p(0x2000), 1, None) == 4
