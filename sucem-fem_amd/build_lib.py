"""Build libsucem_hip.so (HIP for gfx950) in-tree.

Usage: ``python sucem-fem_amd/build_lib.py [--jobs N]``.  Output:
``sucem-fem_amd/sucemfem/_native/libsucem_hip.so``.
"""
import argparse
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
OUT_DIR = os.path.join(HERE, "sucemfem", "_native")
OBJ_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libsucem_hip.so")

SOURCES = ["refdata.cpp", "capi.cpp", "mesh_space.hip", "assemble.hip", "ops.hip", "krylov.hip", "eigen.hip",
           "hiptmair.hip", "batch.hip", "postproc.hip", "matfree.hip", "patch.hip", "sweep.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SUCEM_OFFLOAD_ARCH", "gfx950")
COMMON = ["-O3", "-std=c++17", "-fPIC", "-I", INCLUDE, "-I", CSRC, "-I", OBJ_DIR, "-Wall",
          "-Wno-unused-function", "-Wno-unused-variable"]
# No packed-FP32 instructions in device code (v_pk_*_f32).
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ_DIR, src + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC, "-x", "hip", *DEVICE, *COMMON, "-c", path, "-o", obj]
    else:
        cmd = [HIPCC, *COMMON, "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, r.returncode, r.stdout + r.stderr, cmd


def _gen_tables():
    """Reference-tensor tables for the matrix-free kernel, generated from refdata.cpp."""
    exe = os.path.join(OBJ_DIR, "gen_mf_tables")
    cmd = ["g++", "-O2", "-std=c++17", "-I", CSRC, os.path.join(CSRC, "gen_mf_tables.cpp"),
           os.path.join(CSRC, "refdata.cpp"), "-o", exe, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("table generator build failed")
    out = os.path.join(OBJ_DIR, "mf_tables.inc")
    r = subprocess.run([exe, out], capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("table generation failed")


def build(jobs=None, verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    os.makedirs(OBJ_DIR, exist_ok=True)
    _gen_tables()
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 4)
    objs = []
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        for src, obj, rc, log, cmd in ex.map(_compile, SOURCES):
            if rc != 0:
                sys.stderr.write(log)
                raise RuntimeError("compilation of %s failed: %s" % (src, " ".join(cmd)))
            if verbose and log.strip():
                sys.stderr.write(log)
            objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.v))
