"""ctypes wrapper of the CPU port (oracle/cpu_port.cpp) that bench.py times
as its CPU baseline.

TEST AND BASELINE INFRASTRUCTURE ONLY: loaded by the cpu_baseline leg of
``bench.py`` and by tests, never by the product package.
"""
import ctypes
import os
import platform
import subprocess
import tempfile

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PREBUILT = os.path.join(_HERE, "_build", "libcpu_port.so")

_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C")


def _bind(path):
    L = ctypes.CDLL(path)
    i64, i = ctypes.c_int64, ctypes.c_int
    L.cpu_query_vv_vf.restype = ctypes.c_double
    L.cpu_query_vv_vf.argtypes = [i64, i64, i64, _u32p, _u32p, i, i, _i32p, _i32p, _i32p, _i32p]
    L.cpu_smooth.restype = ctypes.c_double
    L.cpu_smooth.argtypes = [i64, _i32p, _i32p, _f32p, ctypes.c_float, i, i, i, _f32p, _f32p]
    L.cpu_max_threads.restype = i
    return L


def load(native=True):
    """Returns (library, build flags). With native=True the port is first
    compiled for this host (-O3 -march=native -fopenmp) into a temporary
    directory; when that fails the prebuilt x86-64-v3 library is used."""
    if native:
        out = tempfile.mkdtemp(prefix="rxm_cpu_port_")
        try:
            subprocess.run(["make", "-s", "-C", _HERE, "OUT=" + out, "PORT_ARCH=-march=native",
                            out + "/libcpu_port.so"], check=True, capture_output=True, timeout=120)
            return _bind(os.path.join(out, "libcpu_port.so")), "g++ -O3 -march=native -fopenmp"
        except (OSError, subprocess.SubprocessError):
            pass
    if not os.path.exists(_PREBUILT):
        subprocess.check_call(["make", "-s", "-C", _HERE])
    return _bind(_PREBUILT), "g++ -O3 -march=x86-64-v3 -fopenmp"


def cpu_model():
    """The host CPU's model name (lscpu), or the platform's processor string."""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return platform.processor() or "unknown"


def query_vv_vf(L, nV, ev, fv, threads, reps):
    """VV + VF CSRs over global ids; returns (median seconds, vv, vf)."""
    ev = np.ascontiguousarray(ev, np.uint32).reshape(-1)
    fv = np.ascontiguousarray(fv, np.uint32).reshape(-1)
    nE, nF = len(ev) // 2, len(fv) // 3
    vv_rp, vv_col = np.zeros(nV + 1, np.int32), np.zeros(2 * nE, np.int32)
    vf_rp, vf_col = np.zeros(nV + 1, np.int32), np.zeros(3 * nF, np.int32)
    t = L.cpu_query_vv_vf(nV, nE, nF, ev, fv, threads, reps, vv_rp, vv_col, vf_rp, vf_col)
    return t, (vv_rp, vv_col), (vf_rp, vf_col)


def smooth(L, vv, x, lam, iters, threads, reps):
    """`iters` Jacobi smoothing steps of fp32 SoA positions (3, V); returns
    (median seconds, result)."""
    x = np.ascontiguousarray(x, np.float32)
    out, tmp = np.empty_like(x), np.empty_like(x)
    t = L.cpu_smooth(x.shape[1], vv[0], vv[1], x, lam, iters, threads, reps, out, tmp)
    return t, out
