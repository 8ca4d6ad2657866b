"""Build the native parts of fastfermion_amd in place.

    python -m fastfermion_amd.build

Products (next to this file):
  libffamd.so                       HIP kernels + C ABI (hipcc, gfx950)
  _ffcore.cpython-*.so              pybind11 drop-in module linked against libffamd.so
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
ARCH = os.environ.get("FFX_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", f"-I{INCLUDE}"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return r


def _newer(target, sources):
    if not os.path.exists(target):
        return False
    t = os.path.getmtime(target)
    return all(os.path.getmtime(s) <= t for s in sources)


def hip_sources():
    return [os.path.join(CSRC, f) for f in sorted(os.listdir(CSRC)) if f.endswith((".hip", ".h"))] + \
        [os.path.join(INCLUDE, "fastfermion_amd.h")]


def build_lib(force=True, jobs=None):
    """Compile every csrc/*.hip translation unit in parallel, then link libffamd.so."""
    out = os.path.join(HERE, "libffamd.so")
    if not force and _newer(out, hip_sources()):
        return out
    units = [os.path.join(CSRC, f) for f in sorted(os.listdir(CSRC)) if f.endswith(".hip")]
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = [os.path.join(objdir, os.path.basename(u)[:-4] + ".o") for u in units]
    jobs = jobs or min(len(units), max(1, min(16, os.cpu_count() or 1)))
    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(lambda uo: _run([HIPCC, *HIP_FLAGS, "-c", "-o", uo[1], uo[0]]), zip(units, objs)))
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", tmp, *objs])
    os.replace(tmp, out)
    return out


def ext_suffix():
    return sysconfig.get_config_var("EXT_SUFFIX") or ".so"


def build_ffcore(force=True):
    src = os.path.join(CSRC, "ffcore.cpp")
    if not os.path.exists(src):
        return None
    out = os.path.join(HERE, "_ffcore" + ext_suffix())
    if not force and _newer(out, [src, os.path.join(CSRC, "ffcore_types.h"), os.path.join(INCLUDE, "fastfermion_amd.h")]):
        return out
    import pybind11
    tmp = out + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden",
           f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}", f"-I{INCLUDE}",
           "-o", tmp, src, f"-L{HERE}", "-lffamd", "-Wl,-rpath,$ORIGIN"]
    _run(cmd)
    os.replace(tmp, out)
    return out


def build_all(force=True):
    # the pybind module links against libffamd.so, so the library goes first
    lib = build_lib(force)
    core = build_ffcore(force)
    return lib, core


if __name__ == "__main__":
    print(build_all(force="--if-stale" not in sys.argv))
