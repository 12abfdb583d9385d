"""The host code of the library (TIFF decoder, argument checks, the C ABI's
error paths) under AddressSanitizer and UndefinedBehaviorSanitizer, on the
CPU: the CPU tests of those parts run in a child process against the
host-sanitized build (bone_morphometry_amd/build.py build_sanitized(), made by
__graft_entry__.build()), with the sanitizer runtime preloaded.  Any report
aborts the child, so the test fails on the first finding."""
import os
import subprocess
import sys

import pytest

from bone_morphometry_amd import build

HERE = os.path.dirname(os.path.abspath(__file__))


def test_host_code_is_clean_under_asan_and_ubsan():
    if not os.path.exists(build.SANITIZED_LIB):
        pytest.skip("the host-sanitized library is not built (__graft_entry__.build())")
    rt = build.sanitizer_runtime()
    if not rt:
        pytest.skip("no clang ASan runtime found")
    env = dict(os.environ, BM_LIBRARY=build.SANITIZED_LIB, LD_PRELOAD=rt,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider",
                        os.path.join(HERE, "test_tiff.py"), os.path.join(HERE, "test_abi.py")],
                       env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stdout[-4000:]
    assert "runtime error" not in r.stdout and "AddressSanitizer" not in r.stdout
