"""The host Matrix Market reader (host.cpp) built with AddressSanitizer and
UndefinedBehaviorSanitizer and fed random mutations of valid files
(tests/cpp/mtx_fuzz.cpp): no memory error, no undefined behaviour, no leak,
and every accepted file gives a consistent CSR.  Such a fuzz found that
non-square symmetric files and overflowing size tokens were accepted.  No
GPU; skipped when the host compiler has no sanitizer runtime."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent


@pytest.fixture(scope="module")
def fuzz_exe(tmp_path_factory):
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("no g++")
    exe = tmp_path_factory.mktemp("fuzz") / "mtx_fuzz"
    cmd = [gxx, "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer",
           "-I", str(REPO / "include"), "-I", str(REPO / "accspmm_amd" / "csrc"),
           str(REPO / "tests" / "cpp" / "mtx_fuzz.cpp"),
           str(REPO / "accspmm_amd" / "csrc" / "host.cpp"), "-lpthread", "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        if "asan" in r.stderr.lower() or "ubsan" in r.stderr.lower():
            pytest.skip("no sanitizer runtime: " + r.stderr[-300:])
        raise AssertionError(r.stderr[-2000:])
    return exe


@pytest.mark.parametrize("threads", ["1", "4"])
def test_mtx_reader_fuzz_under_sanitizers(fuzz_exe, tmp_path, threads):
    env = dict(os.environ, ACCSPMM_HOST_THREADS=threads,
               ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(fuzz_exe), "3000", str(tmp_path / "f.mtx")], capture_output=True,
                       text=True, env=env, timeout=600)
    assert r.returncode == 0, (r.stdout[-1000:], r.stderr[-3000:])
    assert "fuzz done" in r.stdout
    assert "runtime error" not in r.stderr
