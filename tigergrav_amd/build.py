"""Build the in-tree shared library tigergrav_amd/lib/libtigergrav_amd.so for gfx950.

    python -m tigergrav_amd.build        (or tigergrav_amd.build.build())

hipcc cross-compiles without a GPU.  All device code is built with packed-FP32 instruction
selection disabled (-target-feature -packed-fp32-ops).
"""
import os
import shlex
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
BUILDDIR = os.path.join(PKG, "_build")
LIBNAME = "libtigergrav_amd.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TG_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-fPIC", "-std=c++17", f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}", "-Wall",
          "-Wno-unused-function", "-Wno-unused-variable"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-fgpu-flush-denormals-to-zero", "-munsafe-fp-atomics"]


def sources():
    out = []
    for name in sorted(os.listdir(CSRC)):
        if name.endswith(".hip") or name.endswith(".cpp"):
            out.append(os.path.join(CSRC, name))
    return out


# host-only C++ (the tree restatement): g++, contraction off so every product and sum rounds
# where the reference's does; HIP host API users get the ROCm include path
HOSTCXX = os.environ.get("CXX", "g++")
HOST = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-D__HIP_PLATFORM_AMD__",
        f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}", "-I/opt/rocm/include", "-Wall"]


def _compile(src):
    obj = os.path.join(BUILDDIR, os.path.basename(src) + ".o")
    if src.endswith(".cpp"):
        cmd = [HOSTCXX, "-c"] + HOST + [src, "-o", obj]
    else:
        cmd = [HIPCC, "-c", "-x", "hip"] + COMMON + DEVICE + [src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {shlex.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=False):
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(BUILDDIR, exist_ok=True)
    srcs = sources()
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    out = os.path.join(LIBDIR, LIBNAME)
    tmp = out + ".tmp"
    cmd = [HOSTCXX, "-shared", "-fopenmp", "-o", tmp] + objs + ["-L/opt/rocm/lib", "-lamdhip64",
                                                                 "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {shlex.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, out)
    if verbose:
        print(out)
    return out


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
