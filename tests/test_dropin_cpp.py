"""Source-level drop-in on the CPU (build container only; nothing here runs
on the GPU):

  * the reference's own property test (main.cpp, checkOverlap against a
    brute-force scan, rts/ops.cpp:7-40) compiled against this library's
    rts/ops.h and ops.cpp instead of the reference's: every case passes;
  * a C++ program written like the reference launcher's host stages
    (src/launcher.cu:46-90,117-145: getRootIndex -> generateOps_GETRF ->
    h_ops_dag(tree, start, length) -> instructions_scheduler -> getLengths
    -> instructions_manager), compiled against the library's headers and
    host sources, produces the same instructions, schedule entries, words,
    FLOPs and DAG density as the C ABI's programs, whole and windowed.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "pastel-palettes_amd", "csrc")
HOST = ["rts/ops.cpp", "rts/dependency.cpp", "rts/h_ops_dag.cpp", "rts/scheduler.cpp", "matrix/Dense.cpp",
        "matrix/Hierarchical.cpp", "h_ops/h_ops.cpp", "instructions/instructions_manager.cpp"]


@pytest.mark.skipif(not os.path.exists("/root/reference/main.cpp"), reason="reference sources not present")
def test_reference_main_against_dropin_ops():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "dropin"], check=True)
    exe = os.path.join(ROOT, "oracle", "_ref", "dropin_main")
    for _ in range(3):  # main.cpp seeds from the clock: three draws of 500 cases
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
        assert "FAIL" not in out, out[:2000]
        assert out.count("P") == 500


@pytest.fixture(scope="module")
def flow_exe(tmp_path_factory):
    d = tmp_path_factory.mktemp("flow")
    exe = str(d / "launcher_flow")
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    # scheduler.h declares analyzeClocks with hipError_t (the reference's
    # cudaError_t): HIP's runtime header and library, no device code
    cmd = ["g++", "-O1", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I" + CSRC, f"-I{rocm}/include", "-o", exe,
           os.path.join(ROOT, "tests", "cpp", "launcher_flow.cpp")]
    cmd += [os.path.join(CSRC, s) for s in HOST] + [f"-L{rocm}/lib", "-lamdhip64", f"-Wl,-rpath,{rocm}/lib"]
    subprocess.run(cmd, check=True)
    return exe


def _run(exe, *args):
    out = subprocess.run([exe] + [str(a) for a in args], check=True, capture_output=True, text=True).stdout
    wins = []
    for line in out.splitlines():
        if line.startswith("window"):
            f = line.split()
            wins.append({"insts": int(f[3]), "entries": int(f[5]), "words": int(f[7]), "flops": int(f[9]),
                         "density": float(f[11])})
    return out, wins


def _entries(prog):
    info = prog.info()
    return sum(len(prog.schedule(w)) for w in range(info["workers"]))


@pytest.mark.parametrize("n,levels,rank,workers", [(2048, 3, 16, 8), (1000, 2, 20, 5)])
def test_launcher_flow_matches_the_abi(pp, flow_exe, n, levels, rank, workers):
    out, wins = _run(flow_exe, n, levels, rank, workers, 0)
    assert f"matrix {n} x {n}" in out
    hm = pp.HMatrix("hodlr", n, levels=levels, rank=rank)
    prog = pp.Program.factor(hm, workers)
    info = prog.info()
    assert len(wins) == 1
    w = wins[0]
    assert (w["insts"], w["words"], w["flops"]) == (info["n_insts"], info["total_words"], info["flops"])
    assert w["entries"] == _entries(prog)
    assert abs(w["density"] - info["dag_density"]) < 1e-6
    # windowed: the kernel_size loop of dev_hierarchical_GETRF
    out, wins = _run(flow_exe, n, levels, rank, workers, 9)
    progs = pp.Program.factor_windows(hm, workers, 9)
    assert len(wins) == len(progs) > 1
    for w, p in zip(wins, progs):
        i = p.info()
        assert (w["insts"], w["words"], w["flops"], w["entries"]) == (i["n_insts"], i["total_words"], i["flops"],
                                                                        _entries(p))
