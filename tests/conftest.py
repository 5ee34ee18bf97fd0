import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "meshfreetrixi.jl_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def gpu_available():
    try:
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        return hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


@pytest.fixture(scope="session")
def O():
    import oracle

    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def mf():
    import mfrbf

    mfrbf._lib.lib()
    return mfrbf


def rel_inf(a, b):
    """Norm-wise relative difference max|a-b| / max|b|."""
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def build_c_client(out_dir):
    """Compiles tests/c_client/mfrbf_client.c as plain C99 (warnings are
    errors) against include/mfrbf.h and the built libmfrbf.so; returns the
    executable's path."""
    import shutil
    import subprocess

    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    lib_dir = os.path.join(ROOT, "meshfreetrixi.jl_amd", "mfrbf")
    if not os.path.exists(os.path.join(lib_dir, "libmfrbf.so")):
        pytest.fail("libmfrbf.so is not built")
    exe = os.path.join(str(out_dir), "mfrbf_client")
    cmd = [cc, "-std=c99", "-O2", "-Wall", "-Wextra", "-Werror", "-I" + os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "c_client", "mfrbf_client.c"), "-L" + lib_dir, "-lmfrbf",
           "-Wl,-rpath," + lib_dir, "-Wl,-rpath,/opt/rocm/lib", "-lm", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe
