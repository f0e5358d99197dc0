"""Builds the C++ API test library (tests/cpp/_build/libapi_test.so) from
api_test.hip against the in-tree librxmesh_amd.so. Test infrastructure."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "_build", "libapi_test.so")


def build(verbose=False):
    sys.path.insert(0, ROOT)
    from rxmesh_amd.build import ARCH, HIPCC, LIBDIR, NO_PK_FP32
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "-O2", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}", *NO_PK_FP32,
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "rxmesh_amd", "include"),
           "-I", os.path.join(ROOT, "rxmesh_amd", "csrc"), os.path.join(HERE, "api_test.hip"),
           "-L", LIBDIR, "-lrxmesh_amd", "-Wl,-rpath," + LIBDIR, "-Wl,-rpath,$ORIGIN/../../../rxmesh_amd/_lib",
           "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
