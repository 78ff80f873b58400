"""Host sanitizer run (CPU): the C host layer built with AddressSanitizer and
UndefinedBehaviorSanitizer (`make -C gnn-coloring_amd sanitize`; device code
is not sanitized) runs the host-facing CPU tests -- file parsers and writers,
ID / SD / SDA walks, generators, model loader and tuner records, the
exported-symbol checks -- in a child process with libasan preloaded. Any
sanitizer report ends the child with a non-zero status."""
import os
import subprocess
import sys

import pytest

from conftest import ROOT

PKG = os.path.join(ROOT, "gnn-coloring_amd")
ASAN_LIB = os.path.join(PKG, "_build", "libgcx_asan.so")


def _libasan():
    try:
        path = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True,
                              text=True, timeout=30).stdout.strip()
    except (OSError, subprocess.SubprocessError):
        return None
    return path if os.path.isabs(path) and os.path.exists(path) else None


@pytest.mark.skipif(_libasan() is None, reason="gcc's libasan is not installed")
def test_host_code_under_asan_and_ubsan():
    subprocess.run(["make", "-C", PKG, "sanitize", "-j8"], check=True, capture_output=True,
                   timeout=900)
    env = dict(os.environ, LD_PRELOAD=_libasan(), GCX_LIB_PATH=ASAN_LIB,
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    probe = ("import sys; sys.path[:0] = [%r, %r]; import gcx; "
             "maps = open('/proc/self/maps').read(); "
             "print('libgcx_asan.so' in maps, 'libasan' in maps)" % (ROOT, PKG))
    r = subprocess.run([sys.executable, "-c", probe], env=env, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0 and r.stdout.split() == ["True", "True"], (r.stdout, r.stderr)
    tests = [os.path.join(ROOT, "tests", f) for f in
             ("test_formats.py", "test_seqorder.py", "test_abi.py")]
    r = subprocess.run([sys.executable, "-m", "pytest", *tests, "-q", "-x", "-m", "not gpu",
                        "-p", "no:cacheprovider"], env=env, capture_output=True, text=True,
                       cwd=ROOT, timeout=1200)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])
