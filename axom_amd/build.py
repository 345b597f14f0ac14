"""Builds libaxsq.so (gfx950) in-tree: axom_amd/lib/libaxsq.so.

Every HIP translation unit is compiled with -ffp-contract=off (no fused
multiply-add, so FP64 results round like the reference's CPU code) and with
packed FP32 instructions disabled.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "lib", "libaxsq.so")
SOURCES = ["axsq_prims.hip", "axsq_build.hip", "axsq_query.hip", "axsq_sdf.hip", "axsq_mesh.hip",
           "axsq_api.hip", "axsq_host.hip"]
ARCH = os.environ.get("AXSQ_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas", "-Wno-unused-value",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.exists(c) or c == "hipcc"):
            return c
    raise RuntimeError("hipcc not found")


def _compile(src, defines=(), objdir=OBJ):
    out = os.path.join(objdir, src.replace(".hip", ".o"))
    cmd = [_hipcc()] + FLAGS + [f"-D{d}" for d in defines] + ["-c", os.path.join(CSRC, src),
                                                              "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return out, r.stderr


def build(verbose=False, defines=(), lib=LIB, objdir=OBJ):
    """Compiles every source (with extra -D `defines`) and links `lib`."""
    os.makedirs(objdir, exist_ok=True)
    os.makedirs(os.path.dirname(lib), exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        results = list(ex.map(lambda src: _compile(src, defines, objdir), SOURCES))
    for _, err in results:
        if verbose and err.strip():
            sys.stderr.write(err)
    tmp = lib + ".tmp"
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-o", tmp] + [o for o, _ in results]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, lib)
    return lib


CPP_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "test_facade.cpp")
CPP_TEST_BIN = os.path.join(ROOT, "tests", "cpp", "test_facade")


def build_cpp_tests():
    """C++ test of include/axom_mi355x.hpp (with kernels of its own that use
    the device traverser), linked to libaxsq.so."""
    cmd = [_hipcc(), "-O2", "-std=c++17", "--offload-arch=gfx950", "-Xclang", "-target-feature",
           "-Xclang", "-packed-fp32-ops", "-I", os.path.join(ROOT, "include"), CPP_TEST_SRC,
           "-o", CPP_TEST_BIN + ".tmp", "-L", os.path.dirname(LIB), "-laxsq",
           "-Wl,-rpath,$ORIGIN/../../axom_amd/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"C++ facade test build failed:\n{r.stderr}")
    os.replace(CPP_TEST_BIN + ".tmp", CPP_TEST_BIN)
    return CPP_TEST_BIN


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_cpp_tests())
