"""Build recipe of the engine's shared libraries (in-tree):

  libfemh.so       HIP kernels + solver + C ABI (hipcc, gfx950; packed FP32
                   instructions are disabled for all device code)
  libfemh_host.so  C++ host facade with the reference's class names, over the
                   C ABI only (g++)
  _build/test_host C++ test program of the facade (run by a GPU test)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfemh.so")
HOST_LIB = os.path.join(HERE, "libfemh_host.so")
CXX = os.environ.get("CXX", "g++")

SOURCES = ["kernels.hip", "ops.hip", "tables.cpp", "solver.cpp", "capi.cpp", "noalloc.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FEMH_ARCH", "gfx950")
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-command-line-argument", "-Wno-unknown-cuda-version"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = (r.stdout + r.stderr)
    # the host pass does not know the device feature flag: drop that warning
    out = "\n".join(l for l in out.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + out)
    return out


def build(verbose=False):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src.rsplit(".", 1)[0] + ".o")
        cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd = [HIPCC] + FLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
        out = _run(cmd)
        if verbose and out:
            print(out)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    build_host(objdir)
    return LIB


def build_host(objdir):
    """The facade links only against libfemh.so's C ABI."""
    host = os.path.join(HERE, "host")
    tmp = HOST_LIB + ".tmp"
    _run([CXX, "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-o", tmp,
          os.path.join(host, "femh_host.cpp"), "-L" + HERE, "-lfemh", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, HOST_LIB)
    test_src = os.path.join(ROOT, "tests", "cpp", "test_host.cpp")
    if os.path.exists(test_src):
        _run([CXX, "-O2", "-std=c++17", "-Wall", "-o", os.path.join(objdir, "test_host"), test_src,
              "-L" + HERE, "-lfemh_host", "-lfemh", "-Wl,-rpath," + HERE, "-Wl,-rpath,$ORIGIN/.."])


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
