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


def compile_units():
    """(source, object name, extra flags) of every translation unit of libffamd.so:
    each plain csrc/*.hip file once, and ffx_inst.hip once per key width (1..4
    words), coefficient arithmetic (FP64, FP32, FP32 compensated) and algebra
    (Pauli, Majorana); the FP64 units, which also hold the logs, go first."""
    units = []
    for f in sorted(os.listdir(CSRC)):
        if f.endswith(".hip") and f != "ffx_inst.hip":
            units.append((os.path.join(CSRC, f), f[:-4] + ".o", []))
    inst = os.path.join(CSRC, "ffx_inst.hip")
    combos = [(w, p, a) for p in range(3) for w in (4, 3, 2, 1) for a in (0, 1)]
    return [(inst, f"ffx_inst_w{w}_p{p}_a{a}.o", [f"-DFFX_INST_W={w}", f"-DFFX_INST_P={p}",
                                                   f"-DFFX_INST_ALG={a}"]) for w, p, a in combos] + units


def build_lib(force=True, jobs=None):
    """Compile every translation unit in parallel, then link libffamd.so."""
    out = os.path.join(HERE, "libffamd.so")
    if not force and _newer(out, hip_sources()):
        return out
    units = compile_units()
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = [os.path.join(objdir, o) for _, o, _ in units]
    jobs = jobs or max(1, min(16, os.cpu_count() or 1))
    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(lambda u: _run([HIPCC, *HIP_FLAGS, *u[2], "-c", "-o", os.path.join(objdir, u[1]), u[0]]),
                    units))
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
