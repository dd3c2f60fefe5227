"""Build recipe for the in-tree native pieces of psacking_amd.

* libspk.so   -- HIP kernels + C-ABI (include/spk.h), hipcc for gfx950
* _core*.so   -- pybind11 module over the C-ABI, host C++ (g++), rpath $ORIGIN

Run ``python psacking_amd/build.py`` (or ``__graft_entry__.build()``); it imports
nothing from the package, so it works on a tree that has never been built.
Always recompiles.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
ARCH = os.environ.get("SPK_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-value", "-Wno-unused-result", "-Wno-return-type-c-linkage",
    # a non-void function that can end without a return is undefined
    # behaviour the optimiser exploits: refused at build time
    "-Werror=return-type", "-Wno-error=return-type-c-linkage",
]


def _run(cmd):
    print("[build]", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        print(out)
        raise RuntimeError(f"build step failed: {cmd[0]} (exit {r.returncode})")
    return out


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def lib_path() -> str:
    return os.path.join(HERE, "libspk.so")


def core_path() -> str:
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    return os.path.join(HERE, "_core" + suffix)


def build_lib() -> str:
    out = lib_path()
    tmp = out + ".tmp"
    _run([hipcc(), *HIPCC_FLAGS, "-shared", "-o", tmp, os.path.join(CSRC, "spk_api.hip")])
    os.replace(tmp, out)
    return out


def build_core() -> str:
    import pybind11
    out = core_path()
    tmp = out + ".tmp"
    inc = sysconfig.get_paths()["include"]
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden", "-Werror=return-type",
          f"-I{inc}", f"-I{pybind11.get_include()}", f"-I{os.path.join(ROOT, 'include')}",
          os.path.join(CSRC, "core_bindings.cpp"), "-o", tmp,
          f"-L{HERE}", "-lspk", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, out)
    return out


def build_all() -> None:
    build_lib()
    build_core()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
