"""Build recipe for the native libraries (run by __graft_entry__.build()).

libpmx.so        HIP kernels + C ABI, gfx950 only (hipcc).
libpmx_scene.so  host-only synthetic scene renderer (g++ + OpenMP).
libpmx_drop_in.so / pmx_cli
                 the reference-compatible C++ classes on top of the C ABI.
All outputs are written next to their sources' package directory.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
CPP = os.path.join(HERE, "cpp")
ROOT = os.path.dirname(HERE)
INCLUDE = os.path.join(ROOT, "include")

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed FP32 (v_pk_*_f32) is kept off for every HIP object.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
CXX = os.environ.get("CXX", "g++")
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", "-Wall", "-Wextra"]

LIB_PMX = os.path.join(HERE, "libpmx.so")
LIB_SCENE = os.path.join(HERE, "libpmx_scene.so")
LIB_DROPIN = os.path.join(HERE, "libpmx_drop_in.so")
CLI = os.path.join(HERE, "pmx_cli")


def _run(cmd, quiet=False):
    if not quiet:
        print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd[:3])} ...")
    # hipcc's host pass does not know the packed-fp32 target feature; that
    # warning is expected, anything else is shown.
    noise = [ln for ln in (r.stdout + r.stderr).splitlines()
             if ln.strip() and "packed-fp32-ops" not in ln and "not a recognized feature" not in ln]
    if noise and not quiet:
        print("\n".join(noise[:40]))


def _atomic(out, cmd_builder):
    tmp = out + ".tmp"
    _run(cmd_builder(tmp))
    os.replace(tmp, out)


def build_pmx(extra=()):
    """Compiles the HIP sources to objects in parallel, then links libpmx.so."""
    names = ("pmx_kernels.hip", "pmx_capi.hip", "pmx_cnn.hip", "pmx_net.hip")
    objdir = os.path.join(ROOT, "_build")
    os.makedirs(objdir, exist_ok=True)
    jobs = []
    for n in names:
        obj = os.path.join(objdir, n.replace(".hip", ".o"))
        cmd = [HIPCC, *HIP_FLAGS, *extra, "-c", "-o", obj, os.path.join(CSRC, n)]
        print("+", " ".join(cmd), flush=True)
        jobs.append((cmd, obj, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                                stderr=subprocess.STDOUT, text=True)))
    objs = []
    for cmd, obj, pr in jobs:
        out, _ = pr.communicate()
        if pr.returncode != 0:
            sys.stderr.write(out)
            raise RuntimeError(f"build step failed ({pr.returncode}): {' '.join(cmd[:3])} ...")
        objs.append(obj)
    _atomic(LIB_PMX, lambda o: [HIPCC, *HIP_FLAGS, "-shared", "-o", o, *objs])


def build_scene():
    _atomic(LIB_SCENE, lambda o: [CXX, *CXX_FLAGS, "-shared", "-o", o,
                                  os.path.join(CSRC, "pmx_scene.cpp")])


def build_drop_in():
    """Reference-compatible C++ host classes, linked against libpmx.so."""
    if not os.path.isdir(CPP):
        return
    srcs = sorted(os.path.join(CPP, f) for f in os.listdir(CPP) if f.endswith(".cpp") and f != "pmx_cli.cpp")
    if not srcs:
        return
    inc = ["-I", INCLUDE, "-I", CPP, "-I", "/opt/rocm/include", "-D__HIP_PLATFORM_AMD__"]
    link = [f"-L{HERE}", "-lpmx", "-lz", f"-Wl,-rpath,{HERE}", "-Wl,-rpath,$ORIGIN"]
    _atomic(LIB_DROPIN, lambda o: [CXX, "-O2", "-std=c++17", "-fPIC", "-Wall", "-shared", *inc,
                                   "-o", o, *srcs, *link])
    cli = os.path.join(CPP, "pmx_cli.cpp")
    if os.path.exists(cli):
        _atomic(CLI, lambda o: [CXX, "-O2", "-std=c++17", "-Wall", *inc, "-o", o, cli,
                                f"-L{HERE}", "-lpmx_drop_in", "-lpmx", f"-Wl,-rpath,{HERE}",
                                "-Wl,-rpath,$ORIGIN"])
    # the C++ test program of the helper classes (run by tests/test_dropin_helpers_gpu.py)
    prog = os.path.join(ROOT, "tests", "dropin_helpers.cpp")
    if os.path.exists(prog):
        _atomic(prog[:-4], lambda o: [CXX, "-O2", "-std=c++17", "-Wall", *inc, "-o", o, prog,
                                      f"-L{HERE}", "-lpmx_drop_in", "-lpmx",
                                      f"-Wl,-rpath,{HERE}", "-Wl,-rpath,$ORIGIN/../patchmatch_cuda_amd"])


def build_oracle():
    oracle_dir = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-B", "-C", oracle_dir])


def build_all():
    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        raise RuntimeError("hipcc not found: the HIP engine cannot be built")
    build_pmx()
    build_scene()
    build_drop_in()
    build_oracle()


if __name__ == "__main__":
    build_all()
