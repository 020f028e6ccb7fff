"""tcfft_amd -- MI355X (gfx950) half-precision FFT, Python view of the C ABI.

The product is libtcfft.so (HIP kernels + C++ plan builder behind the C ABI in
include/tcfft.h).  This module binds that ABI with ctypes and offers:

  * Plan1D / Plan2D: the tcfftCreate / tcfftExec pair of tcfft_half.h and
    tcfft_half_2d.h (same arguments: sizes, batch, an in-place device buffer
    whose input is digit-reversed by plan.radices), plus natural-order and
    adjoint execution;
  * gen_rev: the caller-side digit reversal of tcfft_doit_half.cpp:6-29;
  * torch helpers (tcfft_amd.torch_fft) that pass tensor pointers and the
    current HIP stream.

There is no CPU fallback: if libtcfft.so is missing or fails to load, every
entry point raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# TCFFT_LIBRARY selects another build of the same library (the host-sanitizer
# build of tests/run_sanitized.py); by default the in-tree libtcfft.so.
LIB_PATH = os.environ.get("TCFFT_LIBRARY") or os.path.join(_HERE, "libtcfft.so")

FORWARD = -1
INVERSE = +1

STATUS = {
    0: "TCFFT_SUCCESS", 1: "TCFFT_INVALID_SIZE", 2: "TCFFT_INVALID_BATCH", 3: "TCFFT_INVALID_POINTER",
    4: "TCFFT_INVALID_PLAN", 5: "TCFFT_INVALID_VALUE", 6: "TCFFT_HIP_ERROR", 7: "TCFFT_ALLOC_FAILED",
}

# Every extern "C" symbol declared in include/tcfft.h.
EXPORTS = (
    "tcfft_version", "tcfft_status_string", "tcfft_last_hip_error", "tcfft_plan_create_1d",
    "tcfft_plan_create_2d", "tcfft_plan_destroy", "tcfft_plan_radices", "tcfft_query_1d", "tcfft_query_2d",
    "tcfft_exec", "tcfft_exec_natural", "tcfft_gen_rev", "tcfft_plan_num_passes", "tcfft_exec_passes",
    "tcfft_exec_split", "tcfft_plan_set_option",
    "tcfft_multi_create", "tcfft_multi_destroy", "tcfft_multi_info", "tcfft_multi_exec",
    "tcfft_multi_set_schedule", "tcfft_multi_autotune", "tcfft_multi_set_start_delay", "tcfft_multi_get_schedule",
    "tcfft_exec_dif", "tcfft_conv_exec", "tcfft_conv_filter_grad", "tcfft_plan_autotune",
    "tcfft_plan_create_1d_split", "tcfft_plan_create_1d_tuned", "tcfft_query_1d_split",
    "tcfft_plan_create_r2c_1d", "tcfft_exec_r2c", "tcfft_exec_c2r", "tcfft_permute_to_dropin",
    "tcfft_exec_natural_noalloc", "tcfft_multi_exec_natural", "tcfft_plan_prepare_split", "tcfft_query_split",
)
NOALLOC_SCRATCH_BYTES = 64 << 20  # TCFFT_NOALLOC_SCRATCH_BYTES
OPT_SPLIT_CHUNK = 1
OPT_BIG_DELAY = 2
OPT_BIG_GROUPS = 3
OPT_COLS_DELAY = 4
OPT_SPLIT_DELAY = 5
OPT_ROWS_WAVES = 6
OPT_FIRST_CACHED = 7
OPT_MALL_CHUNK = 8
OPT_ROWS_TILES = 9


class TcfftError(RuntimeError):
    def __init__(self, status: int, what: str):
        self.status = status
        name = STATUS.get(status, f"status {status}")
        extra = ""
        if status == 6 and _lib is not None:
            extra = f" (hip error {_lib.tcfft_last_hip_error()})"
        super().__init__(f"{what}: {name}{extra}")


_lib = None
_i32p = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p


def lib():
    """Load libtcfft.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m tcfft_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        L.tcfft_version.restype = ctypes.c_int
        L.tcfft_status_string.argtypes = [ctypes.c_int]
        L.tcfft_status_string.restype = ctypes.c_char_p
        L.tcfft_plan_create_1d.argtypes = [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int]
        L.tcfft_plan_create_2d.argtypes = [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.tcfft_plan_destroy.argtypes = [_vp]
        L.tcfft_plan_radices.argtypes = [_vp, ctypes.c_int, _i32p, ctypes.c_int, _i32p]
        L.tcfft_query_1d.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, ctypes.c_int, _i32p, _i32p]
        L.tcfft_query_2d.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, _i32p, _i32p, _i32p,
                                     _i32p, ctypes.c_int]
        L.tcfft_exec.argtypes = [_vp, _vp, ctypes.c_int, _vp]
        L.tcfft_exec_natural.argtypes = [_vp, _vp, _vp, ctypes.c_int, _vp]
        L.tcfft_exec_natural_noalloc.argtypes = [_vp, _vp, _vp, ctypes.c_int, _vp]
        L.tcfft_gen_rev.argtypes = [ctypes.c_int, _i32p, ctypes.c_int, _i32p]
        L.tcfft_plan_num_passes.argtypes = [_vp, _i32p]
        L.tcfft_exec_passes.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp]
        L.tcfft_exec_split.argtypes = [_vp, _vp, _vp, ctypes.c_int, _vp]
        L.tcfft_plan_prepare_split.argtypes = [_vp]
        L.tcfft_query_split.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, ctypes.c_int,
                                        _i32p]
        L.tcfft_plan_set_option.argtypes = [_vp, ctypes.c_int, ctypes.c_int]
        L.tcfft_plan_autotune.argtypes = [_vp, _vp, _vp, _i32p, _i32p, ctypes.POINTER(ctypes.c_float)]
        L.tcfft_plan_create_1d_split.argtypes = [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.tcfft_plan_create_1d_tuned.argtypes = [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, _i32p,
                                                 ctypes.POINTER(ctypes.c_float)]
        L.tcfft_query_1d_split.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, ctypes.c_int, _i32p, _i32p]
        L.tcfft_plan_create_r2c_1d.argtypes = [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int]
        L.tcfft_exec_r2c.argtypes = [_vp, _vp, _vp, _vp]
        L.tcfft_permute_to_dropin.argtypes = [_vp, _vp, _vp, ctypes.c_int, _vp]
        L.tcfft_exec_c2r.argtypes = [_vp, _vp, _vp, _vp]
        L.tcfft_exec_dif.argtypes = [_vp, _vp, _vp, ctypes.c_int, _vp]
        L.tcfft_conv_exec.argtypes = [_vp, ctypes.c_int, _vp, _vp, _vp, ctypes.c_int, _vp, _vp]
        L.tcfft_conv_filter_grad.argtypes = [_vp, ctypes.c_int, _vp, _vp, _vp, _vp]
        L.tcfft_multi_create.argtypes = [ctypes.POINTER(_vp), ctypes.c_int, _i32p]
        L.tcfft_multi_destroy.argtypes = [_vp]
        L.tcfft_multi_info.argtypes = [_vp, _i32p, ctypes.POINTER(ctypes.c_longlong)]
        L.tcfft_multi_exec.argtypes = [_vp, _vp, ctypes.c_int, _vp]
        L.tcfft_multi_exec_natural.argtypes = [_vp, _vp, _vp, ctypes.c_int, _vp]
        L.tcfft_multi_set_schedule.argtypes = [_vp, ctypes.c_int, ctypes.c_int]
        L.tcfft_multi_autotune.argtypes = [_vp, _vp, _vp, _i32p, _i32p, ctypes.POINTER(ctypes.c_float)]
        L.tcfft_multi_set_start_delay.argtypes = [_vp, ctypes.c_int]
        L.tcfft_multi_get_schedule.argtypes = [_vp, _i32p, _i32p, _i32p]
        _lib = L
    return _lib


def _check(rc: int, what: str):
    if rc != 0:
        raise TcfftError(rc, what)


def version() -> int:
    return lib().tcfft_version()


def gen_rev(n: int, radices) -> np.ndarray:
    """Source index of every position (tcfft_doit_half.cpp:6-29 semantics)."""
    r = np.ascontiguousarray(radices, dtype=np.int32)
    out = np.empty(n, dtype=np.int32)
    _check(lib().tcfft_gen_rev(n, r.ctypes.data_as(_i32p), len(r), out.ctypes.data_as(_i32p)), "tcfft_gen_rev")
    return out


def query_1d(n: int, batch: int = 1):
    """Host-only: (radices, n_passes) the 1D plan would use."""
    rad = np.zeros(32, dtype=np.int32)
    nr, npass = ctypes.c_int(), ctypes.c_int()
    _check(lib().tcfft_query_1d(n, batch, rad.ctypes.data_as(_i32p), 32, ctypes.byref(nr), ctypes.byref(npass)),
           f"tcfft_query_1d(n={n}, batch={batch})")
    return [int(v) for v in rad[: nr.value]], npass.value


def query_1d_split(n: int, batch: int, first_log2: int):
    """Host-only: (radices, n_passes) of the 1D plan whose first pass covers
    2^first_log2 points (tcfft_plan_create_1d_split)."""
    rad = np.zeros(32, dtype=np.int32)
    nr, npass = ctypes.c_int(), ctypes.c_int()
    _check(lib().tcfft_query_1d_split(n, batch, first_log2, rad.ctypes.data_as(_i32p), 32, ctypes.byref(nr),
                                      ctypes.byref(npass)), f"tcfft_query_1d_split(n={n}, first_log2={first_log2})")
    return [int(v) for v in rad[: nr.value]], npass.value


def query_2d(nx: int, ny: int, batch: int = 1):
    rx = np.zeros(32, dtype=np.int32)
    ry = np.zeros(32, dtype=np.int32)
    nrx, nry, npass = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _check(lib().tcfft_query_2d(nx, ny, batch, rx.ctypes.data_as(_i32p), ctypes.byref(nrx),
                                ry.ctypes.data_as(_i32p), ctypes.byref(nry), ctypes.byref(npass), 32),
           f"tcfft_query_2d(nx={nx}, ny={ny}, batch={batch})")
    return [int(v) for v in rx[: nrx.value]], [int(v) for v in ry[: nry.value]], npass.value


def query_split(dims: int, n_or_nx: int, ny: int = 0, batch: int = 1):
    """Split-precision pass list of a shape (tcfft_query_split): one dict per
    pass with R, q_log2, qlo_log2, m_log2 (0 = no twiddle) and units."""
    desc = np.zeros(5 * 16, dtype=np.int32)
    npass = ctypes.c_int()
    _check(lib().tcfft_query_split(dims, n_or_nx, ny, batch, desc.ctypes.data_as(_i32p), 16, ctypes.byref(npass)),
           f"tcfft_query_split(dims={dims}, n_or_nx={n_or_nx}, ny={ny}, batch={batch})")
    keys = ("R", "q_log2", "qlo_log2", "m_log2", "units")
    return [dict(zip(keys, (int(v) for v in desc[5 * i: 5 * i + 5]))) for i in range(npass.value)]


class _Plan:
    dims = 0

    def __init__(self):
        self._h = _vp()

    @property
    def handle(self):
        return self._h.value

    def _radices(self, axis: int):
        rad = np.zeros(32, dtype=np.int32)
        nr = ctypes.c_int()
        _check(lib().tcfft_plan_radices(self._h, axis, rad.ctypes.data_as(_i32p), 32, ctypes.byref(nr)),
               "tcfft_plan_radices")
        return [int(v) for v in rad[: nr.value]]

    def exec(self, data_ptr: int, direction: int = FORWARD, stream: int = 0):
        """In place; input digit-reversed by the plan's radices (tcfftExec)."""
        _check(lib().tcfft_exec(self._h, _vp(data_ptr), direction, _vp(stream)), "tcfft_exec")

    @property
    def num_passes(self) -> int:
        n = ctypes.c_int()
        _check(lib().tcfft_plan_num_passes(self._h, ctypes.byref(n)), "tcfft_plan_num_passes")
        return n.value

    def exec_passes(self, data_ptr: int, first: int, count: int, direction: int = FORWARD, stream: int = 0):
        """Passes [first, first+count) only (profiling of individual pass kernels)."""
        _check(lib().tcfft_exec_passes(self._h, _vp(data_ptr), direction, first, count, _vp(stream)),
               "tcfft_exec_passes")

    def exec_natural(self, in_ptr: int, out_ptr: int, direction: int = FORWARD, stream: int = 0):
        _check(lib().tcfft_exec_natural(self._h, _vp(in_ptr), _vp(out_ptr), direction, _vp(stream)),
               "tcfft_exec_natural")

    def exec_natural_noalloc(self, in_ptr: int, out_ptr: int, direction: int = FORWARD, stream: int = 0):
        """exec_natural that allocates nothing: in-place staging goes to the
        library's fixed per-GPU scratch (NOALLOC_SCRATCH_BYTES), shared by
        all calls on that GPU."""
        _check(lib().tcfft_exec_natural_noalloc(self._h, _vp(in_ptr), _vp(out_ptr), direction, _vp(stream)),
               "tcfft_exec_natural_noalloc")

    def permute_to_dropin(self, in_ptr: int, out_ptr: int, n_lines: int, stream: int = 0):
        """out[b, p] = in[b, rev[p]] for n_lines lines of the plan's n (1D):
        natural-order data into the order exec consumes, on the device."""
        _check(lib().tcfft_permute_to_dropin(self._h, _vp(in_ptr), _vp(out_ptr), n_lines, _vp(stream)),
               "tcfft_permute_to_dropin")

    def exec_split(self, in_ptr: int, out_ptr: int, direction: int = FORWARD, stream: int = 0):
        """FP32 complex in/out (natural order), FP16 hi/lo MFMA split precision:
        every size and in place after prepare_split(), else 1D n = 65536 out
        of place only."""
        _check(lib().tcfft_exec_split(self._h, _vp(in_ptr), _vp(out_ptr), direction, _vp(stream)), "tcfft_exec_split")

    def prepare_split(self):
        """Pass list, twiddle tables and the 8-byte-per-element workspace of
        exec_split at this plan's size (allocated here, not in exec)."""
        _check(lib().tcfft_plan_prepare_split(self._h), "tcfft_plan_prepare_split")

    def exec_dif(self, in_ptr: int, out_ptr: int, direction: int = FORWARD, stream: int = 0):
        """Natural-order input -> spectrum in the digit-reversed order of
        self.radices (the order exec consumes); n = 4096, 8192 or 16384."""
        _check(lib().tcfft_exec_dif(self._h, _vp(in_ptr), _vp(out_ptr), direction, _vp(stream)), "tcfft_exec_dif")

    def conv_exec(self, channels: int, x_ptr: int, h_rev_ptr: int, y_ptr: int, conj_filter: bool = False,
                  spec_ptr: int = 0, stream: int = 0):
        """y = F^H (H * F x) / n per [.., channels, n] signal, H given as a
        digit-reversed fp16 spectrum per channel; spec_ptr (optional)
        receives F x, digit-reversed."""
        _check(lib().tcfft_conv_exec(self._h, channels, _vp(x_ptr), _vp(h_rev_ptr), _vp(y_ptr), int(conj_filter),
                                     _vp(spec_ptr) if spec_ptr else None, _vp(stream)), "tcfft_conv_exec")

    def conv_filter_grad(self, channels: int, x_spec_ptr: int, g_spec_ptr: int, gh_rev_ptr: int, stream: int = 0):
        """gh_rev[c] = (1/n) sum_b conj(X[b, c]) * G[b, c] (FP32, digit-reversed)."""
        _check(lib().tcfft_conv_filter_grad(self._h, channels, _vp(x_spec_ptr), _vp(g_spec_ptr), _vp(gh_rev_ptr),
                                            _vp(stream)), "tcfft_conv_filter_grad")

    def set_option(self, option: int, value: int):
        _check(lib().tcfft_plan_set_option(self._h, option, value), "tcfft_plan_set_option")

    def autotune(self, data_ptr: int, stream: int = 0):
        """Picks the fastest start schedule of the first pass (then its row
        workgroup size, its tiles per wave for lines up to 2048 points, and
        the column-pass delay) by
        timing forward transforms of the buffer at data_ptr, each from the
        buffer's original contents, which it holds again on return.  Returns
        (groups, delay, ms)."""
        g, d, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_float()
        _check(lib().tcfft_plan_autotune(self._h, _vp(data_ptr), _vp(stream), ctypes.byref(g), ctypes.byref(d),
                                         ctypes.byref(ms)), "tcfft_plan_autotune")
        return g.value, d.value, ms.value

    def destroy(self):
        if self._h.value:
            _check(lib().tcfft_plan_destroy(self._h), "tcfft_plan_destroy")
            self._h = _vp()

    def __del__(self):
        try:
            if self._h.value and _lib is not None:
                _lib.tcfft_plan_destroy(self._h)
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()


class Plan1D(_Plan):
    """tcfftCreate(&plan, n, n_batch) of tcfft_half.h."""

    dims = 1

    def __init__(self, n: int, batch: int = 1, first_log2: int | None = None, tuned: bool = False):
        """first_log2: radix split (length of the first pass, 8..12 or 16);
        tuned: time the candidate splits and keep the fastest (the plan's
        radices follow the choice, reported in self.first_log2)."""
        super().__init__()
        self.n, self.batch = n, batch
        self.first_log2, self.tune_ms = first_log2, None
        if tuned:
            fl, ms = ctypes.c_int(), ctypes.c_float()
            _check(lib().tcfft_plan_create_1d_tuned(ctypes.byref(self._h), n, batch, ctypes.byref(fl),
                                                    ctypes.byref(ms)), f"tcfft_plan_create_1d_tuned(n={n})")
            self.first_log2, self.tune_ms = fl.value, ms.value
        elif first_log2 is not None:
            _check(lib().tcfft_plan_create_1d_split(ctypes.byref(self._h), n, batch, first_log2),
                   f"tcfft_plan_create_1d_split(n={n}, first_log2={first_log2})")
        else:
            _check(lib().tcfft_plan_create_1d(ctypes.byref(self._h), n, batch),
                   f"tcfft_plan_create_1d(n={n}, batch={batch})")
        self.radices = self._radices(0)
        self.n_radices = len(self.radices)

    @property
    def numel(self) -> int:
        return self.n * self.batch


class PlanR2C(_Plan):
    """Real-input transforms of batch signals of n = 2^9..2^27 real points
    (no reference counterpart).  exec_r2c: n fp16 reals -> bins 0..n/2
    (n/2 + 1 fp16 complex per signal); exec_c2r: the unnormalised inverse
    (n times numpy's irfft).  Out of place.  As a complex plan (exec,
    exec_natural, ...) it is the plan of n/2 points."""

    dims = 1

    def __init__(self, n: int, batch: int = 1):
        super().__init__()
        self.n, self.batch = n, batch
        _check(lib().tcfft_plan_create_r2c_1d(ctypes.byref(self._h), n, batch),
               f"tcfft_plan_create_r2c_1d(n={n}, batch={batch})")
        self.radices = self._radices(0)
        self.n_radices = len(self.radices)

    @property
    def bins(self) -> int:
        return self.n // 2 + 1

    def exec_r2c(self, in_ptr: int, out_ptr: int, stream: int = 0):
        _check(lib().tcfft_exec_r2c(self._h, _vp(in_ptr), _vp(out_ptr), _vp(stream)), "tcfft_exec_r2c")

    def exec_c2r(self, in_ptr: int, out_ptr: int, stream: int = 0):
        _check(lib().tcfft_exec_c2r(self._h, _vp(in_ptr), _vp(out_ptr), _vp(stream)), "tcfft_exec_c2r")


class Plan2D(_Plan):
    """tcfftCreate(&plan, nx, ny, n_batch) of tcfft_half_2d.h."""

    dims = 2

    def __init__(self, nx: int, ny: int, batch: int = 1):
        super().__init__()
        self.nx, self.ny, self.batch = nx, ny, batch
        _check(lib().tcfft_plan_create_2d(ctypes.byref(self._h), nx, ny, batch),
               f"tcfft_plan_create_2d(nx={nx}, ny={ny}, batch={batch})")
        self.radices_x = self._radices(0)
        self.radices_y = self._radices(1)
        self.n_radices_x = len(self.radices_x)
        self.n_radices_y = len(self.radices_y)

    @property
    def numel(self) -> int:
        return self.nx * self.ny * self.batch


class MultiPlan:
    """Uneven batch: signals of sizes[i] (2^8..2^16) packed back to back,
    each digit-reversed by its own 1D radix list, one persistent launch."""

    def __init__(self, sizes):
        self.sizes = np.ascontiguousarray(sizes, dtype=np.int32)
        self._h = _vp()
        _check(lib().tcfft_multi_create(ctypes.byref(self._h), len(self.sizes), self.sizes.ctypes.data_as(_i32p)),
               "tcfft_multi_create")
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes, dtype=np.int64)[:-1]])
        self.total = int(self.sizes.sum())

    @property
    def n_tasks(self) -> int:
        n = ctypes.c_int()
        _check(lib().tcfft_multi_info(self._h, ctypes.byref(n), None), "tcfft_multi_info")
        return n.value

    def exec(self, data_ptr: int, direction: int = FORWARD, stream: int = 0):
        _check(lib().tcfft_multi_exec(self._h, _vp(data_ptr), direction, _vp(stream)), "tcfft_multi_exec")

    def exec_natural(self, in_ptr: int, out_ptr: int, direction: int = FORWARD, stream: int = 0):
        """Natural-order signals in and out (in == out allowed); the same
        bits as exec on each signal permuted by its own gen_rev."""
        _check(lib().tcfft_multi_exec_natural(self._h, _vp(in_ptr), _vp(out_ptr), direction, _vp(stream)),
               "tcfft_multi_exec_natural")

    def set_schedule(self, grid: int, order: int):
        _check(lib().tcfft_multi_set_schedule(self._h, grid, order), "tcfft_multi_set_schedule")

    def autotune(self, data_ptr: int, stream: int = 0):
        g, o, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_float()
        _check(lib().tcfft_multi_autotune(self._h, _vp(data_ptr), _vp(stream), ctypes.byref(g), ctypes.byref(o),
                                          ctypes.byref(ms)), "tcfft_multi_autotune")
        return g.value, o.value, ms.value

    def set_start_delay(self, delay: int):
        _check(lib().tcfft_multi_set_start_delay(self._h, delay), "tcfft_multi_set_start_delay")

    def schedule(self):
        """(grid, order, start delay) currently used."""
        g, o, d = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib().tcfft_multi_get_schedule(self._h, ctypes.byref(g), ctypes.byref(o), ctypes.byref(d)),
               "tcfft_multi_get_schedule")
        return g.value, o.value, d.value

    def destroy(self):
        if self._h.value:
            _check(lib().tcfft_multi_destroy(self._h), "tcfft_multi_destroy")
            self._h = _vp()

    def __del__(self):
        try:
            if self._h.value and _lib is not None:
                _lib.tcfft_multi_destroy(self._h)
        except Exception:
            pass
