"""Build libpastel.so in-tree: host C++ with g++, HIP sources with hipcc for gfx950.

Usage: python pastel-palettes_amd/build.py [--jobs N]
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpastel.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HOST_SOURCES = [
    "rts/ops.cpp",
    "rts/dependency.cpp",
    "rts/h_ops_dag.cpp",
    "rts/scheduler.cpp",
    "matrix/Dense.cpp",
    "matrix/Hierarchical.cpp",
    "h_ops/h_ops.cpp",
    "instructions/instructions_manager.cpp",
    "capi_host.cpp",
    "runtime/runtime.cpp",
    "runtime/solver.cpp",
    "io/wire.cpp",
    "runtime/build.cpp",
    "runtime/pivot.cpp",
]
HIP_SOURCES = ["kernels/vm.hip", "kernels/build.hip", "kernels/pivot.hip"]

CXXFLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", f"-I{ROCM}/include",
            "-D__HIP_PLATFORM_AMD__"]
HIPFLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _obj(src):
    return os.path.join(OBJ, src.replace("/", "__") + ".o")


def _newest_header():
    t = 0.0
    for root in (CSRC, os.path.join(os.path.dirname(HERE), "include")):
        for d, _, files in os.walk(root):
            for f in files:
                if f.endswith(".h"):
                    t = max(t, os.path.getmtime(os.path.join(d, f)))
    return t


def _compile(src, hip, incremental=False, headers=0.0):
    out = _obj(src)
    path = os.path.join(CSRC, src)
    if incremental and os.path.exists(out) and os.path.getmtime(out) >= max(os.path.getmtime(path), headers):
        return out
    if hip:
        cmd = [os.path.join(ROCM, "bin", "hipcc")] + HIPFLAGS + ["-c", path, "-o", out]
    else:
        cmd = ["g++"] + CXXFLAGS + ["-c", path, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{err}")
    return out


def build(jobs=8, verbose=False, incremental=False):
    """Compiles every source (incremental: only sources newer than their
    object or older than the newest header) and links libpastel.so."""
    os.makedirs(OBJ, exist_ok=True)
    headers = _newest_header() if incremental else 0.0
    work = [(s, True) for s in HIP_SOURCES] + [(s, False) for s in HOST_SOURCES]
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(lambda a: _compile(a[0], a[1], incremental, headers), work))
    objs = objs[len(HIP_SOURCES):] + objs[:len(HIP_SOURCES)]
    tmp = LIB + ".tmp"
    cmd = [os.path.join(ROCM, "bin", "hipcc"), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + [
        f"-L{ROCM}/lib", "-lamdhip64", "-Wl,-rpath," + f"{ROCM}/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("--incremental", action="store_true", help="recompile only what changed")
    a = ap.parse_args()
    try:
        build(a.jobs, verbose=True, incremental=a.incremental)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
