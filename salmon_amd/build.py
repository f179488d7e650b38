"""In-tree build of the HIP library (gfx950) and the Fortran binding module."""
from __future__ import annotations

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
ARCH = "gfx950"

HIP_SOURCES = ["salmon_gpu.hip"]
HIP_DEPS = ["common.hpp", "streamed.hpp", "fused.hpp", "cg.hpp", "gs.hpp", "scf.hpp", "pc.hpp", "pcnet.hpp",
            "batch.hpp", "batch_host.hpp"]
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
             # keep hipcc from forming packed-FP32 instructions
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function"]
# roctx ranges of the region timers (rocprofv3 --marker-trace)
HIP_LIBS = ["-L/opt/rocm/lib", "-lrocprofiler-sdk-roctx", "-Wl,-rpath,/opt/rocm/lib"]


def lib_path():
    return os.path.join(HERE, "lib", "libsalmon_gpu.so")


def build_hip(verbose=False):
    os.makedirs(os.path.join(HERE, "lib"), exist_ok=True)
    out = lib_path()
    tmp = out + ".tmp"
    cmd = [HIPCC] + HIP_FLAGS + ["-I", os.path.join(ROOT, "include"), "-o", tmp,
                                 os.path.join(HERE, "csrc", HIP_SOURCES[0])] + HIP_LIBS
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stdout + res.stderr)
    if verbose:
        print(res.stderr)
    os.replace(tmp, out)
    return out


def build_fortran():
    """Compiles the ISO_C_BINDING module and the reference-named shims."""
    if not os.path.exists(FLANG):
        return None
    src = os.path.join(HERE, "fortran")
    outdir = os.path.join(HERE, "lib", "fortran")
    os.makedirs(outdir, exist_ok=True)
    files = ["salmon_gpu_mod.f90"]
    objs = []
    for f in files:
        p = os.path.join(src, f)
        if not os.path.exists(p):
            continue
        o = os.path.join(outdir, f.replace(".f90", ".o"))
        subprocess.check_call([FLANG, "-c", "-O2", "-fPIC", "-module-dir", outdir, p, "-o", o])
        objs.append(o)
    return objs


def build_all():
    build_hip()
    build_fortran()
