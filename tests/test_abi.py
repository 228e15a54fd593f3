"""The C-ABI boundary: libpastel.so loads without a GPU and exports every
function include/*.h declares (no compute call is made here)."""
import ctypes
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for path in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(path).read()
        text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s+\**\s*(pp_[a-z0-9_]+)\s*\(", text,
                             re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_the_boundary():
    names = declared_functions()
    assert len(names) > 80
    for must in ("pp_stream_upload", "pp_launch", "pp_wait", "pp_read_trace", "pp_solver_factor", "pp_solver_solve",
                 "pp_program_factor", "pp_schedule_create", "pp_rts_check_overlap", "pp_dep_table_update"):
        assert must in names


def test_library_exports_every_declared_function(pp):
    lib = ctypes.CDLL(pp.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_version_and_error_without_gpu(pp):
    L = pp.lib()
    assert L.pp_version().decode()
    # a call that fails on its arguments leaves a message and touches no device
    assert L.pp_hmatrix_create_hodlr(0, 1, 1, None) < 0
    assert L.pp_last_error().decode()
