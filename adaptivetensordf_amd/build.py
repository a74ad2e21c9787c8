"""Build the native parts in-tree.

* ``lib/libatdf.so``  -- hand-written gfx950 kernels behind include/atdf.h
  (hipcc, recipe in ``csrc/Makefile``).
* ``atdf_torch.so``   -- libtorch host layer + Python bindings, linked
  against ``lib/libatdf.so`` (g++ with torch's ABI flags).

Run ``python -m adaptivetensordf_amd.build``.  Every call compiles.
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
LIB_DIR = os.path.join(HERE, "lib")
EXT_PATH = os.path.join(HERE, "atdf_torch.so")
OBJ_DIR = os.path.join(REPO, "_build", "atdf_torch")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


def _run(cmd, cwd=None):
    proc = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}")
    return proc.stdout


def build_hip(jobs: int = 3) -> str:
    """Compile libatdf.so for gfx950 (always recompiles)."""
    _run(["make", "-C", os.path.join(HERE, "csrc"), "clean"])
    _run(["make", "-C", os.path.join(HERE, "csrc"), f"-j{jobs}"])
    return os.path.join(LIB_DIR, "libatdf.so")


def _torch_flags():
    import torch
    from torch.utils import cpp_extension as ce

    torch_lib = os.path.join(os.path.dirname(torch.__file__), "lib")
    inc = ce.include_paths() + [sysconfig.get_paths()["include"], os.path.join(ROCM, "include"),
                                os.path.join(REPO, "include")]
    abi = [f"-D_GLIBCXX_USE_CXX11_ABI={int(torch._C._GLIBCXX_USE_CXX11_ABI)}"]
    try:
        abi += ce._get_pybind11_abi_build_flags()
    except AttributeError:  # older torch
        pass
    cflags = ["-O2", "-std=c++17", "-fPIC", "-DTORCH_API_INCLUDE_EXTENSION_H",
              "-DTORCH_EXTENSION_NAME=atdf_torch", "-D__HIP_PLATFORM_AMD__=1", "-DUSE_ROCM=1",
              "-Wno-deprecated-declarations"] + abi + [f"-I{p}" for p in inc]
    ldflags = [f"-L{LIB_DIR}", "-latdf", "-Wl,-rpath,$ORIGIN/lib", f"-L{torch_lib}", "-lc10", "-ltorch",
               "-ltorch_cpu", "-ltorch_python", "-lc10_hip", "-ltorch_hip", f"-Wl,-rpath,{torch_lib}",
               f"-L{ROCM}/lib", "-lamdhip64", "-lz"]
    return cflags, ldflags


def build_torch_layer() -> str:
    """Compile and link atdf_torch.so (always recompiles)."""
    cflags, ldflags = _torch_flags()
    os.makedirs(OBJ_DIR, exist_ok=True)
    srcs = [os.path.join(HERE, "csrc", "torch", f) for f in ("atdf_torch.cpp", "atdf_tree_opt.cpp", "atdf_volume.cpp", "atdf_scene.cpp", "atdf_regularize.cpp",
                                                                    "atdf_vm_fused.cpp", "atdf_bindings.cpp")]
    objs = [os.path.join(OBJ_DIR, os.path.basename(s) + ".o") for s in srcs]
    cxx = os.environ.get("CXX", "g++")
    with ThreadPoolExecutor(max_workers=2) as ex:
        list(ex.map(lambda so: _run([cxx, *cflags, "-c", so[0], "-o", so[1]]), zip(srcs, objs)))
    tmp = EXT_PATH + ".tmp"
    _run([cxx, "-shared", "-o", tmp, *objs, *ldflags])
    os.replace(tmp, EXT_PATH)
    return EXT_PATH


def build_all() -> None:
    build_hip()
    build_torch_layer()


if __name__ == "__main__":
    build_all()
    print("built", os.path.join(LIB_DIR, "libatdf.so"), "and", EXT_PATH)
