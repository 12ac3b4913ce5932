"""In-tree build of the native parts (no JIT, nothing installed).

* ``libtmhip.so`` - the gfx950 HIP engine and its C ABI (``include/tmhip.h``);
* ``_core`` - the pybind11 host module with the reference's object model
  (Mesh, Material, CyclicSymmetrySolver, ModalResult, ...), linked against
  ``libtmhip.so`` through the C ABI.

Run ``python -m turbomodal_amd.build`` or ``__graft_entry__.build()``.
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

ENGINE_SOURCES = ["capi.hip", "driver.hip", "factor.hip", "fem.hip", "vec.hip", "plan.cpp",
                  "symbolic.cpp"]
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             # packed FP32 instructions are kept out of every kernel
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + os.path.join(ROOT, "include")]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_engine(jobs: int = 8):
    out = os.path.join(HERE, "libtmhip.so")
    objs = []
    procs = []
    for src in ENGINE_SOURCES:
        obj = os.path.join(HERE, "_build", src + ".o")
        os.makedirs(os.path.dirname(obj), exist_ok=True)
        cmd = [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd = [HIPCC, "-x", "hip", *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        print("+", " ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        objs.append(obj)
        if len(procs) >= jobs:
            for p in procs:
                if p.wait() != 0:
                    raise RuntimeError("engine compile failed")
            procs = []
    for p in procs:
        if p.wait() != 0:
            raise RuntimeError("engine compile failed")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_core():
    import pybind11
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    out = os.path.join(HERE, "_core" + suffix)
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared",
           "-I" + pybind11.get_include(), "-I" + sysconfig.get_paths()["include"],
           "-I" + os.path.join(ROOT, "include"),
           os.path.join(CSRC, "core_bindings.cpp"), "-o", out + ".tmp",
           "-L" + HERE, "-ltmhip", "-Wl,-rpath,$ORIGIN"]
    _run(cmd)
    os.replace(out + ".tmp", out)
    return out


def build_all():
    build_engine()
    if os.path.exists(os.path.join(CSRC, "core_bindings.cpp")):
        build_core()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
