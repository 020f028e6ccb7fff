"""Run the CPU test suite (pytest -m "not gpu") against host-sanitized builds.

    python tests/run_sanitized.py [extra pytest args]

Builds libtcfft.so with AddressSanitizer + UndefinedBehaviorSanitizer on its
host code (tcfft_amd/_asan/libtcfft.so; device code unchanged, every
-fsanitize behind -Xarch_host) and the FP64 checker with the same clang
runtime (oracle/_build/asan/), then runs the CPU tests in a child Python
with that runtime preloaded and TCFFT_LIBRARY / ORACLE_LIBRARY pointing at
the sanitized builds.  The CPU tests drive the host code this covers: plan
validation and size checks (n * batch < 2^31), radix queries and splits,
gen_rev, the uneven-batch task builder (tcfft_multi_create builds its task
lists before it needs a device), argument checks of every entry point, and
the checker itself.  Any sanitizer report aborts the run (halt_on_error,
-fno-sanitize-recover); leak checking is off because the Python interpreter
itself holds allocations until exit.

Test infrastructure: needs no GPU and is never run on one.
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def asan_runtime() -> str:
    hits = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    if not hits:
        raise RuntimeError("clang's AddressSanitizer runtime was not found under /opt/rocm/llvm")
    return hits[-1]


def main() -> int:
    from tcfft_amd import build as b
    lib = b.build(sanitize=True)
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True)
    env = dict(os.environ)
    env.update({
        "LD_PRELOAD": asan_runtime(),
        "ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1:abort_on_error=1",
        "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1",
        "TCFFT_LIBRARY": lib,
        "ORACLE_LIBRARY": os.path.join(ROOT, "oracle", "_build", "asan", "libtcfft_oracle.so"),
    })
    cmd = [sys.executable, "-m", "pytest", os.path.join(ROOT, "tests"), "-q", "-m", "not gpu", "-p", "no:cacheprovider",
           *sys.argv[1:]]
    print("running:", " ".join(cmd), flush=True)
    return subprocess.run(cmd, env=env, cwd=ROOT).returncode


if __name__ == "__main__":
    sys.exit(main())
