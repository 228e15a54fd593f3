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
HIP_SOURCES = ["kernels/vm_fp64.hip", "kernels/vm_mixed.hip", "kernels/vm_fp32.hip", "kernels/vm.hip", "kernels/build.hip",
               "kernels/pivot.hip", "kernels/noalloc.hip"]

CXXFLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", f"-I{ROCM}/include",
            "-D__HIP_PLATFORM_AMD__"]
HIPFLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _obj(src):
    return os.path.join(OBJ, src.replace("/", "__") + ".o")


def _deps(out):
    """Files the object was built from (its -MMD dependency list), or None."""
    try:
        with open(out + ".d") as f:
            text = f.read().replace("\\\n", " ")
    except OSError:
        return None
    return text.split(":", 1)[1].split() if ":" in text else None


def _up_to_date(out):
    deps = _deps(out)
    if not deps or not os.path.exists(out):
        return False
    t = os.path.getmtime(out)
    return all(os.path.exists(d) and os.path.getmtime(d) <= t for d in deps)


def _compile(src, hip, incremental=False):
    out = _obj(src)
    path = os.path.join(CSRC, src)
    if incremental and _up_to_date(out):
        return out
    dep = ["-MMD", "-MF", out + ".d"]
    if hip:
        cmd = [os.path.join(ROCM, "bin", "hipcc")] + HIPFLAGS + dep + ["-c", path, "-o", out]
    else:
        cmd = ["g++"] + CXXFLAGS + dep + ["-c", path, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{err}")
    return out


def build(jobs=8, verbose=False, incremental=False):
    """Compiles every source (incremental: only those whose object is older
    than a file it was built from) and links libpastel.so."""
    os.makedirs(OBJ, exist_ok=True)
    work = [(s, True) for s in HIP_SOURCES] + [(s, False) for s in HOST_SOURCES]
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(lambda a: _compile(a[0], a[1], incremental), work))
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
