"""Build recipe for the native parts of gpu_mm_amd.

Products (all in-tree, next to this file):
  libgpu_mm_amd.so            HIP kernels for gfx950 + the C ABI (include/gpu_mm_amd.h)
  _gpu_mm_amd_py<ext>.so      pybind11 module over the C ABI (arrays in via DLPack)

Run as ``python gpu_mm_amd/_build.py`` or through ``__graft_entry__.build()`` (the package
itself cannot be imported before the extension exists).
"""

import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
CSRC = os.path.join(PKG_DIR, "csrc")
BUILD_DIR = os.path.join(REPO_DIR, "_build", "gpu_mm_amd")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
OFFLOAD_ARCH = "gfx950"

# -packed-fp32-ops: keep the compiler from emitting packed FP32 instructions.
# -munsafe-fp-atomics: native float atomics on ordinary (coarse-grained) device memory.
HIP_FLAGS = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={OFFLOAD_ARCH}",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]

LIB_SOURCES = [
    ("capi.cpp", ["-x", "hip"]),
    ("capi_batch.cpp", ["-x", "hip"]),
    ("host_ref.cpp", ["-x", "hip", "-ffp-contract=off"]),
    ("k_planner.hip", []),
    ("k_planned.hip", []),
    ("k_unplanned.hip", []),
    ("k_tod16.hip", []),
    ("k_dynmap.hip", []),
    ("k_cells.hip", []),
    ("k_debug.hip", []),
    ("k_probe.hip", []),
]

LIB_NAME = "libgpu_mm_amd.so"


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n  " + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def lib_path():
    return os.path.join(PKG_DIR, LIB_NAME)


def ext_path():
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    return os.path.join(PKG_DIR, "_gpu_mm_amd_py" + suffix)


def build_lib(verbose=False):
    os.makedirs(BUILD_DIR, exist_ok=True)

    def compile_one(item):
        src, extra = item
        obj = os.path.join(BUILD_DIR, src + ".o")
        cmd = [HIPCC] + HIP_FLAGS + extra + [f"-I{os.path.join(REPO_DIR, 'include')}", "-c", os.path.join(CSRC, src), "-o", obj]
        out = _run(cmd)
        if verbose and out.strip():
            print(out)
        return obj

    with ThreadPoolExecutor(max_workers=min(8, len(LIB_SOURCES))) as ex:
        objs = list(ex.map(compile_one, LIB_SOURCES))

    tmp = lib_path() + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={OFFLOAD_ARCH}", "-o", tmp] + objs)
    os.replace(tmp, lib_path())
    return lib_path()


def build_ext(verbose=False):
    import pybind11

    src = os.path.join(CSRC, "pybind_module.cpp")
    inc = [sysconfig.get_paths()["include"], pybind11.get_include(), os.path.join(REPO_DIR, "include")]
    out = ext_path()
    tmp = out + ".tmp"
    cmd = ["g++", "-O2", "-shared", "-fPIC", "-std=c++17", "-fvisibility=hidden", "-Wall"]
    cmd += [f"-I{d}" for d in inc]
    cmd += [src, "-o", tmp, f"-L{PKG_DIR}", "-l:" + LIB_NAME, "-Wl,-rpath,$ORIGIN"]
    o = _run(cmd)
    if verbose and o.strip():
        print(o)
    os.replace(tmp, out)
    return out


def build(verbose=False):
    build_lib(verbose)
    build_ext(verbose)


if __name__ == "__main__":
    build(verbose=("-v" in sys.argv))
    print("built", lib_path(), ext_path())
