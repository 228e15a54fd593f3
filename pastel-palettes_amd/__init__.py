"""Python host layer of the MI355X H-LU factor + solve library.

Thin ctypes binding over libpastel.so (include/pastel.h) plus the H-matrix
input generators.  The C++ host layer (rts scheduler, Hierarchical, h-op
generation, instruction encoding) and the HIP kernels do all the work; this
module only moves pointers and sizes.  There is no CPU fallback: every entry
point raises if the shared library is missing.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpastel.so")

FP64, FP32, MIXED = 0, 1, 2
SOLVE_HINV, SOLVE_LEAFINV, SOLVE_SUBST = 0, 1, 2  # solver modes (pp_solver_create_batch_mode)
KERNEL_EXP1D, KERNEL_INV2D = 0, 1  # kernels of the device builder (pp_hmatrix_build_kernel)
PRECISIONS = {"fp64": FP64, "fp32": FP32, "mixed": MIXED}

_lib = None


class PastelError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"pastel error {code}: {msg}")
        self.code = code


def lib():
    """Load libpastel.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PastelError(-5, f"{LIB_PATH} is missing; run pastel-palettes_amd/build.py")
    L = C.CDLL(LIB_PATH)
    i32, i64, u64, vp, dp = C.c_int32, C.c_int64, C.c_uint64, C.c_void_p, C.c_double
    P = C.POINTER
    sig = {
        "pp_last_error": (C.c_char_p, []),
        "pp_version": (C.c_char_p, []),
        "pp_min_inst_flops": (i64, []),
        "pp_rts_check_overlap": (C.c_int, [u64] * 8),
        "pp_rts_check_dependency": (C.c_int, [C.c_int, P(u64), P(i32), C.c_int, P(u64), P(i32)]),
        "pp_dep_table_create": (C.c_int, [P(vp)]),
        "pp_dep_table_destroy": (None, [vp]),
        "pp_dep_table_update": (C.c_int, [vp, i64, i64, C.c_int]),
        "pp_dep_table_lookup": (C.c_int, [vp, i64, i64]),
        "pp_dep_table_list": (C.c_int, [vp, i64, C.c_int, P(i64), C.c_int]),
        "pp_dep_table_count": (i64, [vp, i64, C.c_int]),
        "pp_dep_table_clear": (C.c_int, [vp]),
        "pp_schedule_create": (C.c_int, [C.c_int, P(i64), C.c_int, P(i32), P(i32), C.c_int, P(vp)]),
        "pp_schedule_worker_length": (C.c_int, [vp, C.c_int]),
        "pp_schedule_worker_entries": (C.c_int, [vp, C.c_int, P(i32), P(i32), P(i64)]),
        "pp_schedule_comm_waits": (C.c_int, [vp]),
        "pp_schedule_destroy": (None, [vp]),
        "pp_hmatrix_create_hodlr": (C.c_int, [C.c_int, C.c_int, C.c_int, P(vp)]),
        "pp_hmatrix_create_dense": (C.c_int, [C.c_int, C.c_int, P(vp)]),
        "pp_hmatrix_destroy": (None, [vp]),
        "pp_hmatrix_size": (C.c_int, [vp]),
        "pp_hmatrix_storage": (i64, [vp]),
        "pp_hmatrix_inverse_base": (i64, [vp]),
        "pp_hmatrix_inverse_elems": (i64, [vp]),
        "pp_hmatrix_inverse_identity": (C.c_int, [vp, P(dp)]),
        "pp_hmatrix_inverse_count": (C.c_int, [vp, P(C.c_int), P(C.c_int)]),
        "pp_hmatrix_inverse_leaf": (C.c_int, [vp, C.c_int, P(i64)]),
        "pp_hmatrix_inverse_node": (C.c_int, [vp, C.c_int, P(i64)]),
        "pp_hmatrix_num_leaves": (C.c_int, [vp]),
        "pp_hmatrix_leaf": (C.c_int, [vp, C.c_int, P(i64)]),
        "pp_program_factor": (C.c_int, [vp, C.c_int, P(vp)]),
        "pp_program_solve": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, P(vp)]),
        "pp_program_factor_batch": (C.c_int, [P(vp), C.c_int, C.c_int, P(vp)]),
        "pp_program_solve_batch": (C.c_int, [P(vp), C.c_int, P(i32), P(i32), C.c_int, C.c_int, P(vp)]),
        "pp_program_factor_inv_batch": (C.c_int, [P(vp), C.c_int, C.c_int, P(vp)]),
        "pp_program_factor_num_windows": (C.c_int, [vp, C.c_int]),
        "pp_program_factor_window": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, P(vp)]),
        "pp_program_solve_inv_batch": (C.c_int, [P(vp), C.c_int, P(i32), P(i32), C.c_int, C.c_int, P(vp)]),
        "pp_program_destroy": (None, [vp]),
        "pp_program_get_info": (C.c_int, [vp, vp]),
        "pp_program_temp_elems": (i64, [vp, C.c_int]),
        "pp_program_worker_length": (C.c_int, [vp, C.c_int]),
        "pp_program_worker_words": (C.c_int, [vp, C.c_int, P(i32), C.c_int]),
        "pp_program_inst_words": (C.c_int, [vp, C.c_int, P(i32), C.c_int]),
        "pp_program_inst_flops": (i64, [vp, C.c_int]),
        "pp_program_inst_bytes": (i64, [vp, C.c_int]),
        "pp_program_inst_preds": (C.c_int, [vp, C.c_int, P(i32), P(i32), C.c_int]),
        "pp_program_schedule_length": (C.c_int, [vp, C.c_int]),
        "pp_program_schedule_entries": (C.c_int, [vp, C.c_int, P(i32), P(i32), P(i64)]),
        "pp_ctx_create": (C.c_int, [C.c_int, u64, P(vp)]),
        "pp_ctx_destroy": (C.c_int, [vp]),
        "pp_ctx_cu_count": (C.c_int, [vp]),
        "pp_ctx_bytes_in_use": (u64, [vp]),
        "pp_alloc": (C.c_int, [vp, u64, P(vp)]),
        "pp_free": (C.c_int, [vp, vp]),
        "pp_copy_h2d": (C.c_int, [vp, vp, vp, u64]),
        "pp_copy_d2h": (C.c_int, [vp, vp, vp, u64]),
        "pp_stream_upload": (C.c_int, [vp, P(P(i32)), P(i32), C.c_int, P(vp), C.c_int, u64, C.c_int, C.c_int, P(vp)]),
        "pp_stream_upload_program": (C.c_int, [vp, vp, P(vp), C.c_int, C.c_int, P(vp)]),
        "pp_stream_set_slot": (C.c_int, [vp, vp, C.c_int, vp]),
        "pp_launch": (C.c_int, [vp, vp, C.c_int, vp]),
        "pp_wait": (C.c_int, [vp, vp, C.c_int, P(i32)]),
        "pp_stream_enable_trace": (C.c_int, [vp, vp, C.c_int]),
        "pp_read_trace": (i64, [vp, vp, P(u64), i64]),
        "pp_stream_destroy": (C.c_int, [vp, vp]),
        "pp_stream_set_watchdog": (C.c_int, [vp, vp, C.c_int]),
        "pp_stream_last_ms": (dp, [vp]),
        "pp_stream_qr_fallbacks": (C.c_int, [vp]),
        "pp_max_workers": (C.c_int, [vp, C.c_int, C.c_int]),
        "pp_exec_single": (C.c_int, [vp, C.c_int, P(i32), C.c_int, P(vp), C.c_int, u64, u64]),
        "pp_exec_single_timed": (C.c_int, [vp, C.c_int, P(i32), C.c_int, P(vp), C.c_int, u64, u64, C.c_int, C.c_int,
                                           P(dp)]),
        "pp_fill_sketch": (C.c_int, [vp, vp, i64, C.c_int, u64]),
        "pp_exec_single_noalloc": (C.c_int, [vp, C.c_int, P(i32), C.c_int, P(vp), C.c_int, u64, u64, vp]),
        "pp_exec_single_noalloc_max_bytes": (u64, []),
        "pp_mfma_probe": (C.c_int, [vp, C.c_int, P(dp), C.c_int, P(dp), C.c_int, P(dp)]),
        "pp_solver_create": (C.c_int, [vp, vp, P(dp), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, P(vp)]),
        "pp_solver_create_device": (C.c_int, [vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, P(vp)]),
        "pp_solver_factor": (C.c_int, [vp, vp]),
        "pp_solver_solve": (C.c_int, [vp, vp, vp, C.c_int, C.c_int, dp, C.c_int, vp, vp]),
        "pp_solver_residual": (C.c_int, [vp, vp, vp, vp, C.c_int, C.c_int, vp]),
        "pp_solver_factor_arena": (C.c_int, [vp, P(vp), P(C.c_int)]),
        "pp_solver_destroy": (C.c_int, [vp]),
        "pp_solver_factor_stream": (vp, [vp]),
        "pp_solver_factor_flops": (dp, [vp]),
        "pp_solver_solve_flops": (dp, [vp, C.c_int]),
        "pp_solver_create_batch": (C.c_int, [vp, P(vp), C.c_int, P(vp), C.c_int, C.c_int, C.c_int, C.c_int,
                                             C.c_int, C.c_int, P(vp)]),
        "pp_solver_create_batch_mode": (C.c_int, [vp, P(vp), C.c_int, P(vp), C.c_int, C.c_int, C.c_int, C.c_int,
                                                  C.c_int, C.c_int, C.c_int, P(vp)]),
        "pp_solver_num_members": (C.c_int, [vp]),
        "pp_solver_rows": (i64, [vp]),
        "pp_solver_member": (C.c_int, [vp, C.c_int, P(i64), P(C.c_int), P(dp)]),
        "pp_solver_member_arena": (C.c_int, [vp, C.c_int, P(i64)]),
        "pp_solver_storage_elems": (i64, [vp]),
        "pp_solver_use_graph": (C.c_int, [vp, C.c_int]),
        "pp_solver_last_factor_ms": (dp, [vp]),
        "pp_hmatrix_build_kernel": (C.c_int, [vp, vp, C.c_int, vp, dp, dp, vp, vp, vp]),
        "pp_hmatrix_build_pivots": (i64, [vp]),
        "pp_dense_getrf_pivoted": (C.c_int, [vp, C.c_int, vp, C.c_int, C.c_int, C.c_int, C.c_int, i64, vp, vp, vp]),
        "pp_dense_apply_pivot": (C.c_int, [vp, C.c_int, vp, C.c_int, C.c_int, C.c_int, vp, C.c_int, vp]),
        "pp_hmatrix_encode": (i64, [vp, P(dp), vp, u64]),
        "pp_hmatrix_decode": (C.c_int, [vp, u64, P(dp), u64, P(C.c_int), P(vp)]),
        "pp_program_encode": (i64, [vp, vp, u64]),
        "pp_program_decode": (C.c_int, [vp, u64, P(vp)]),
        "pp_solver_encode_factors": (i64, [vp, vp, u64]),
        "pp_solver_decode_factors": (C.c_int, [vp, vp, u64]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc is not None and rc < 0:
        raise PastelError(rc, lib().pp_last_error().decode())
    return rc


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _encode(fn, *args):
    """Runs a size-then-fill encoder of the byte-string format -> bytes."""
    n = check(fn(*args, None, 0))
    buf = C.create_string_buffer(n)
    check(fn(*args, buf, n))
    return buf.raw


class ProgramInfo(C.Structure):
    _fields_ = [("n_insts", C.c_int32), ("workers", C.c_int32), ("n_slots", C.c_int32), ("pad", C.c_int32),
                ("n_edges", C.c_int64), ("flops", C.c_int64), ("temp_elems", C.c_int64),
                ("scratch_elems", C.c_int64), ("total_words", C.c_int64), ("comm_waits", C.c_int64),
                ("dag_density", C.c_double), ("schedule_ms", C.c_double), ("dag_ms", C.c_double),
                ("encode_ms", C.c_double)]


class SolveReport(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32), ("pad0", C.c_int32), ("pad1", C.c_int32),
                ("residual", C.c_double), ("factor_ms", C.c_double), ("solve_ms", C.c_double),
                ("total_ms", C.c_double)]


# ---------------------------------------------------------------- rts

def check_overlap(a, b):
    """DataMap2D(ptr, pitch, width, height).checkOverlap (bytes)."""
    return bool(check(lib().pp_rts_check_overlap(*[int(x) for x in a], *[int(x) for x in b])))


def check_dependency(ops_from, ops_to):
    """Mask of op_to depending on op_from; each op is [(read_only, (ptr, pitch, width, height)), ...]."""
    def pack(ops):
        m = np.array([list(map(int, d)) for _, d in ops] or [[0, 0, 0, 0]], dtype=np.uint64)
        r = np.array([1 if ro else 0 for ro, _ in ops] or [0], dtype=np.int32)
        return m, r
    mf, rf = pack(ops_from)
    mt, rt = pack(ops_to)
    return check(lib().pp_rts_check_dependency(len(ops_from), _p(mf, C.c_uint64), _p(rf, C.c_int32), len(ops_to),
                                               _p(mt, C.c_uint64), _p(rt, C.c_int32)))


def min_inst_flops():
    return int(lib().pp_min_inst_flops())


class DepTable:
    """dependency_table: edges (from, to) with a 3-bit mask each."""

    def __init__(self):
        h = C.c_void_p()
        check(lib().pp_dep_table_create(C.byref(h)))
        self.h = h

    def update(self, f, t, dep):
        check(lib().pp_dep_table_update(self.h, f, t, dep))

    def lookup(self, f, t):
        return check(lib().pp_dep_table_lookup(self.h, f, t))

    def _list(self, inst, to):
        n = check(lib().pp_dep_table_list(self.h, inst, to, None, 0))
        a = np.zeros(max(n, 1), np.int64)
        check(lib().pp_dep_table_list(self.h, inst, to, _p(a, C.c_int64), n))
        return a[:n].tolist()

    def dep_from(self, f):
        return self._list(f, 0)

    def dep_to(self, t):
        return self._list(t, 1)

    def count_from(self, f):
        return int(check(lib().pp_dep_table_count(self.h, f, 0)))

    def count_to(self, t):
        return int(check(lib().pp_dep_table_count(self.h, t, 1)))

    def clear(self):
        check(lib().pp_dep_table_clear(self.h))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.pp_dep_table_destroy(self.h)
            self.h = None


class Schedule:
    """instructions_scheduler over an explicit DAG (edges (from, to), masks)."""

    def __init__(self, flops, edges, masks, workers):
        L = lib()
        f = np.ascontiguousarray(flops, dtype=np.int64)
        e = np.ascontiguousarray(np.asarray(edges, dtype=np.int32).reshape(-1, 2))
        m = np.ascontiguousarray(masks, dtype=np.int32)
        h = C.c_void_p()
        check(L.pp_schedule_create(len(f), _p(f, C.c_int64), len(m), _p(e if e.size else np.zeros(2, np.int32),
                                                                          C.c_int32),
                                   _p(m if m.size else np.zeros(1, np.int32), C.c_int32), workers, C.byref(h)))
        self.h, self.workers = h, workers

    def worker(self, w):
        L = lib()
        n = check(L.pp_schedule_worker_length(self.h, w))
        ins = np.zeros(max(n, 1), np.int32)
        ex = np.zeros(max(n, 1), np.int32)
        fl = np.zeros(max(n, 1), np.int64)
        check(L.pp_schedule_worker_entries(self.h, w, _p(ins, C.c_int32), _p(ex, C.c_int32), _p(fl, C.c_int64)))
        return [(int(i), bool(x), int(f)) for i, x, f in zip(ins[:n], ex[:n], fl[:n])]

    def text(self):
        return [" ".join(("i" if x else "w") + str(i) for i, x, _ in self.worker(w)) for w in range(self.workers)]

    def comm_waits(self):
        return check(lib().pp_schedule_comm_waits(self.h))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.pp_schedule_destroy(self.h)
            self.h = None


# ---------------------------------------------------------------- H-matrix structure

class HMatrix:
    """Hierarchical structure + storage plan (no data)."""

    def __init__(self, kind, n, **kw):
        L = lib()
        h = C.c_void_p()
        if kind == "hodlr":
            check(L.pp_hmatrix_create_hodlr(n, kw["levels"], kw["rank"], C.byref(h)))
        elif kind == "dense":
            check(L.pp_hmatrix_create_dense(n, kw["leaf"], C.byref(h)))
        else:
            raise ValueError(kind)
        self.h, self.kind, self.n, self.params = h, kind, n, kw

    def build_on_device(self, ctx, kind, points_ptr, param, shift, arena_ptr, pivots_ptr=None, hip_stream=None):
        """Fills the device arena at arena_ptr (storage FP64 elements) from the
        kernel `kind` (KERNEL_EXP1D / KERNEL_INV2D) of the device points at
        points_ptr: dense leaves evaluated, low-rank leaves by fixed-rank
        cross approximation.  pivots_ptr: optional device int32 array of
        build_pivots entries."""
        check(lib().pp_hmatrix_build_kernel(ctx.h, self.h, kind, C.c_void_p(int(points_ptr)), param, shift,
                                            C.c_void_p(int(arena_ptr)),
                                            C.c_void_p(int(pivots_ptr)) if pivots_ptr else None, hip_stream))

    @property
    def build_pivots(self):
        return int(lib().pp_hmatrix_build_pivots(self.h))

    def to_bytes(self, arena=None):
        """Byte string of the structure and, when given, the FP64 arena."""
        if arena is not None:
            arena = np.ascontiguousarray(arena, dtype=np.float64)
            if arena.size != self.storage:
                raise ValueError(f"arena has {arena.size} elements, the structure {self.storage}")
            return _encode(lib().pp_hmatrix_encode, self.h, _p(arena, C.c_double))
        return _encode(lib().pp_hmatrix_encode, self.h, None)

    @classmethod
    def from_bytes(cls, data):
        """(HMatrix, arena or None) from a string made by to_bytes."""
        L = lib()
        data = bytes(data)
        h, has = C.c_void_p(), C.c_int()
        check(L.pp_hmatrix_decode(data, len(data), None, 0, C.byref(has), C.byref(h)))
        self = cls.__new__(cls)
        self.h, self.kind, self.n, self.params = h, "decoded", int(L.pp_hmatrix_size(h)), {}
        arena = None
        if has.value:
            arena = np.empty(self.storage)
            h2 = C.c_void_p()
            check(L.pp_hmatrix_decode(data, len(data), _p(arena, C.c_double), arena.size, None, C.byref(h2)))
            L.pp_hmatrix_destroy(h2)
        return self, arena

    @property
    def storage(self):
        return int(lib().pp_hmatrix_storage(self.h))

    @property
    def inverse_base(self):
        """Start (elements) of the leaf-inverse area in the matrix slot."""
        return int(lib().pp_hmatrix_inverse_base(self.h))

    @property
    def solver_elems(self):
        """Arena elements a Solver keeps per matrix (arena + inverse area)."""
        return self.inverse_base + int(lib().pp_hmatrix_inverse_elems(self.h))

    def inverse_layout(self):
        """(leaves, nodes) of the inverse area: leaves as dicts row0, nb, inv_off,
        leaf_off; nodes as dicts row0, na, nb, rl, ru, l_u, l_v, u_u, u_v."""
        L = lib()
        nl, nn = C.c_int(), C.c_int()
        check(L.pp_hmatrix_inverse_count(self.h, C.byref(nl), C.byref(nn)))
        a4, a9 = (C.c_int64 * 4)(), (C.c_int64 * 9)()
        leaves, nodes = [], []
        for i in range(nl.value):
            check(L.pp_hmatrix_inverse_leaf(self.h, i, a4))
            leaves.append(dict(zip(("row0", "nb", "inv_off", "leaf_off"), list(a4))))
        for i in range(nn.value):
            check(L.pp_hmatrix_inverse_node(self.h, i, a9))
            nodes.append(dict(zip(("row0", "na", "nb", "rl", "ru", "l_u", "l_v", "u_u", "u_v"), list(a9))))
        return leaves, nodes

    def inverse_identity(self):
        """Initial contents of the leaf-inverse area (identity blocks)."""
        a = np.empty(int(lib().pp_hmatrix_inverse_elems(self.h)))
        check(lib().pp_hmatrix_inverse_identity(self.h, _p(a, C.c_double)))
        return a

    def with_inverse_area(self, arena):
        """The arena extended by the leaf-inverse area, ready for
        Program.factor_inv_batch."""
        inv = self.inverse_identity()
        out = np.zeros(self.inverse_base + inv.size)
        out[:arena.size] = arena
        out[self.inverse_base:] = inv
        return out

    def leaves(self):
        L = lib()
        out = []
        info = (C.c_int64 * 8)()
        for i in range(check(L.pp_hmatrix_num_leaves(self.h))):
            check(L.pp_hmatrix_leaf(self.h, i, info))
            t, r0, c0, ny, nx, rank, off, offv = list(info)
            out.append(dict(type="dense" if t == 1 else "lowrank", row0=r0, col0=c0, ny=ny, nx=nx, rank=rank,
                            off=off, off_v=offv))
        return out

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.pp_hmatrix_destroy(self.h)
            self.h = None


class Program:
    """Instruction streams of one launch (factor or solve, single or batched)."""

    def __init__(self, handle, members):
        self.h, self.members = handle, members

    @classmethod
    def factor(cls, hm, workers):
        h = C.c_void_p()
        check(lib().pp_program_factor(hm.h, workers, C.byref(h)))
        return cls(h, [hm])

    @classmethod
    def solve(cls, hm, nrhs, workers, ld=None, rhs_split=1):
        h = C.c_void_p()
        check(lib().pp_program_solve(hm.h, nrhs, ld or nrhs, rhs_split, workers, C.byref(h)))
        return cls(h, [hm])

    @classmethod
    def factor_batch(cls, hms, workers):
        arr = (C.c_void_p * len(hms))(*[m.h for m in hms])
        h = C.c_void_p()
        check(lib().pp_program_factor_batch(arr, len(hms), workers, C.byref(h)))
        return cls(h, list(hms))

    @classmethod
    def solve_batch(cls, hms, nrhs, workers, rhs_split=1):
        arr = (C.c_void_p * len(hms))(*[m.h for m in hms])
        nr = np.ascontiguousarray(nrhs, dtype=np.int32)
        h = C.c_void_p()
        check(lib().pp_program_solve_batch(arr, len(hms), _p(nr, C.c_int32), _p(nr, C.c_int32), rhs_split, workers,
                                           C.byref(h)))
        return cls(h, list(hms))

    @classmethod
    def factor_inv_batch(cls, hms, workers):
        """Factorization that also forms the explicit leaf inverses."""
        arr = (C.c_void_p * len(hms))(*[m.h for m in hms])
        h = C.c_void_p()
        check(lib().pp_program_factor_inv_batch(arr, len(hms), workers, C.byref(h)))
        return cls(h, list(hms))

    @classmethod
    def solve_inv_batch(cls, hms, nrhs, workers, rhs_split=1):
        """Solve applying the leaf inverses formed by factor_inv_batch."""
        arr = (C.c_void_p * len(hms))(*[m.h for m in hms])
        nr = np.ascontiguousarray(nrhs, dtype=np.int32)
        h = C.c_void_p()
        check(lib().pp_program_solve_inv_batch(arr, len(hms), _p(nr, C.c_int32), _p(nr, C.c_int32), rhs_split,
                                               workers, C.byref(h)))
        return cls(h, list(hms))

    def info(self):
        inf = ProgramInfo()
        check(lib().pp_program_get_info(self.h, C.byref(inf)))
        return {k: getattr(inf, k) for k, _ in ProgramInfo._fields_ if k != "pad"}

    def temp_elems(self, member=0):
        return int(check(lib().pp_program_temp_elems(self.h, member)))

    def streams(self):
        L = lib()
        out = []
        for w in range(self.info()["workers"]):
            n = check(L.pp_program_worker_length(self.h, w))
            a = np.zeros(n, np.int32)
            check(L.pp_program_worker_words(self.h, w, _p(a, C.c_int32), n))
            out.append(a)
        return out

    def inst_words(self, i):
        a = np.zeros(32, np.int32)
        n = check(lib().pp_program_inst_words(self.h, i, _p(a, C.c_int32), 32))
        return a[:n].copy()

    def inst_flops(self, i):
        return int(check(lib().pp_program_inst_flops(self.h, i)))

    def inst_bytes(self, i):
        """Algorithmic bytes at 8 per element (written operands twice)."""
        return int(check(lib().pp_program_inst_bytes(self.h, i)))

    def preds(self, i):
        L = lib()
        n = check(L.pp_program_inst_preds(self.h, i, None, None, 0))
        p = np.zeros(max(n, 1), np.int32)
        m = np.zeros(max(n, 1), np.int32)
        check(L.pp_program_inst_preds(self.h, i, _p(p, C.c_int32), _p(m, C.c_int32), n))
        return list(zip(p[:n].tolist(), m[:n].tolist()))

    def schedule(self, w):
        L = lib()
        n = check(L.pp_program_schedule_length(self.h, w))
        ins = np.zeros(max(n, 1), np.int32)
        ex = np.zeros(max(n, 1), np.int32)
        fl = np.zeros(max(n, 1), np.int64)
        check(L.pp_program_schedule_entries(self.h, w, _p(ins, C.c_int32), _p(ex, C.c_int32), _p(fl, C.c_int64)))
        return [(int(i), bool(x), int(f)) for i, x, f in zip(ins[:n], ex[:n], fl[:n])]

    @classmethod
    def factor_windows(cls, hm, workers, window_ops):
        """The factorization cut into windows of window_ops h-ops, each its own
        program; launched in order they factor the matrix.  Every window's
        stream must get the same temp buffer (slot 1)."""
        L = lib()
        out = []
        for i in range(check(L.pp_program_factor_num_windows(hm.h, window_ops))):
            h = C.c_void_p()
            check(L.pp_program_factor_window(hm.h, workers, window_ops, i, C.byref(h)))
            out.append(cls(h, [hm]))
        return out

    def to_bytes(self):
        """Byte string of the instructions, dependencies, schedule and streams."""
        return _encode(lib().pp_program_encode, self.h)

    @classmethod
    def from_bytes(cls, data):
        data = bytes(data)
        h = C.c_void_p()
        check(lib().pp_program_decode(data, len(data), C.byref(h)))
        return cls(h, None)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.pp_program_destroy(self.h)
            self.h = None


# ---------------------------------------------------------------- device

class Context:
    def __init__(self, device=0, budget_bytes=0):
        h = C.c_void_p()
        check(lib().pp_ctx_create(device, budget_bytes, C.byref(h)))
        self.h = h

    @property
    def cu_count(self):
        return check(lib().pp_ctx_cu_count(self.h))

    def bytes_in_use(self):
        return int(lib().pp_ctx_bytes_in_use(self.h))

    def max_workers(self, precision=FP64, wavefronts=4):
        return check(lib().pp_max_workers(self.h, precision, wavefronts))

    def exec_single(self, precision, words, ptrs, scratch_bytes=1 << 20, seed=0x5EED, wavefronts=4):
        w = np.ascontiguousarray(words, dtype=np.int32)
        arr = (C.c_void_p * len(ptrs))(*[int(p) for p in ptrs])
        check(lib().pp_exec_single_timed(self.h, precision, _p(w, C.c_int32), len(w), arr, len(ptrs), scratch_bytes,
                                         seed, wavefronts, 1, None))

    def time_single(self, precision, words, ptrs, scratch_bytes=1 << 22, wavefronts=4, reps=20, seed=0x5EED):
        """Mean kernel time (ms) of one instruction executed `reps` times in place."""
        w = np.ascontiguousarray(words, dtype=np.int32)
        arr = (C.c_void_p * len(ptrs))(*[int(p) for p in ptrs])
        ms = C.c_double()
        check(lib().pp_exec_single_timed(self.h, precision, _p(w, C.c_int32), len(w), arr, len(ptrs), scratch_bytes,
                                         seed, wavefronts, reps, C.byref(ms)))
        return ms.value

    def mfma_probe(self, mode, a, b):
        a = np.ascontiguousarray(a, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        out = np.zeros(256, np.float64)
        check(lib().pp_mfma_probe(self.h, mode, _p(a, C.c_double), a.size, _p(b, C.c_double), b.size,
                                  _p(out, C.c_double)))
        return out

    def getrf_pivoted(self, precision, a_ptr, ld, ny, nx, pivots_ptr, info_ptr, count=1, stride=0, stream=None):
        """Partial-pivoted LU of `count` row-major ny x nx blocks in device
        memory (pp_dense_getrf_pivoted); asynchronous on `stream`."""
        check(lib().pp_dense_getrf_pivoted(self.h, precision, C.c_void_p(int(a_ptr)), ld, ny, nx, count, stride,
                                           C.c_void_p(int(pivots_ptr)), C.c_void_p(int(info_ptr)), stream))

    def apply_pivot(self, precision, m_ptr, ld, nx, ny, pivots_ptr, recover=False, stream=None):
        """Rows of a device block permuted by a pivot vector (pp_dense_apply_pivot)."""
        check(lib().pp_dense_apply_pivot(self.h, precision, C.c_void_p(int(m_ptr)), ld, nx, ny,
                                         C.c_void_p(int(pivots_ptr)), int(bool(recover)), stream))

    def close(self):
        if getattr(self, "h", None):
            lib().pp_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Stream:
    """An uploaded program bound to device buffers."""

    def __init__(self, ctx, program, ptrs, precision=FP64):
        arr = (C.c_void_p * len(ptrs))(*[int(p) if p else None for p in ptrs])
        h = C.c_void_p()
        check(lib().pp_stream_upload_program(ctx.h, program.h, arr, len(ptrs), precision, C.byref(h)))
        self.ctx, self.h, self.program = ctx, h, program

    def launch(self, wavefronts=4, hip_stream=None):
        check(lib().pp_launch(self.ctx.h, self.h, wavefronts, hip_stream))

    def wait(self, timeout_ms=0):
        first = C.c_int32(-1)
        rc = lib().pp_wait(self.ctx.h, self.h, timeout_ms, C.byref(first))
        if rc < 0:
            raise PastelError(rc, lib().pp_last_error().decode() + f" (first unsignalled {first.value})")

    def set_watchdog(self, ms):
        check(lib().pp_stream_set_watchdog(self.ctx.h, self.h, ms))

    def enable_trace(self, on=True):
        check(lib().pp_stream_enable_trace(self.ctx.h, self.h, 1 if on else 0))

    def trace(self):
        n = check(lib().pp_read_trace(self.ctx.h, self.h, None, 0))
        a = np.zeros(n, np.uint64)
        lib().pp_read_trace(self.ctx.h, self.h, _p(a, C.c_uint64), n)
        return a

    def qr_fallbacks(self):
        """accum operations of the last waited launch that fell back to
        Householder QR."""
        return int(check(lib().pp_stream_qr_fallbacks(self.h)))

    def last_ms(self):
        return float(lib().pp_stream_last_ms(self.h))

    def close(self):
        if getattr(self, "h", None):
            lib().pp_stream_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Solver:
    """Factor + solve + FP64 iterative refinement on one H-matrix."""

    def __init__(self, ctx, hm, storage, precision=FP64, workers=None, wavefronts=4, nrhs_max=16, rhs_split=1,
                 solve_mode=None):
        if workers is None:
            workers = min(ctx.cu_count, 256)
        h = C.c_void_p()
        if solve_mode not in (None, SOLVE_HINV):
            b = Solver.batch(ctx, [hm], [storage], precision=precision, workers=workers, wavefronts=wavefronts,
                             nrhs_max=nrhs_max, rhs_split=rhs_split, solve_mode=solve_mode)
            self.ctx, self.hm, self.h, self.precision, self.workers = ctx, hm, b.h, precision, workers
            b.h = None
            return
        if isinstance(storage, np.ndarray):
            st = np.ascontiguousarray(storage, np.float64)
            assert st.size == hm.storage
            check(lib().pp_solver_create(ctx.h, hm.h, _p(st, C.c_double), precision, workers, wavefronts, nrhs_max,
                                         rhs_split, C.byref(h)))
        else:  # device pointer (int)
            check(lib().pp_solver_create_device(ctx.h, hm.h, C.c_void_p(int(storage)), precision, workers,
                                                wavefronts, nrhs_max, rhs_split, C.byref(h)))
        self.ctx, self.hm, self.h, self.precision, self.workers = ctx, hm, h, precision, workers

    @classmethod
    def batch(cls, ctx, hms, storages, precision=FP64, workers=None, wavefronts=4, nrhs_max=4, rhs_split=1,
              solve_mode=None):
        """One solver over many H-matrices (block-diagonal batch): one launch
        factors all, one launch solves all.  storages: host arrays or device
        pointers (ints)."""
        self = cls.__new__(cls)
        if workers is None:
            workers = min(ctx.cu_count, 256)
        on_dev = not isinstance(storages[0], np.ndarray)
        if on_dev:
            keep = None
            arr = (C.c_void_p * len(storages))(*[int(p) for p in storages])
        else:
            keep = [np.ascontiguousarray(a, np.float64) for a in storages]
            arr = (C.c_void_p * len(keep))(*[a.ctypes.data for a in keep])
        hs = (C.c_void_p * len(hms))(*[m.h for m in hms])
        h = C.c_void_p()
        if solve_mode is None:
            check(lib().pp_solver_create_batch(ctx.h, hs, len(hms), arr, 1 if on_dev else 0, precision, workers,
                                               wavefronts, nrhs_max, rhs_split, C.byref(h)))
        else:
            check(lib().pp_solver_create_batch_mode(ctx.h, hs, len(hms), arr, 1 if on_dev else 0, precision, workers,
                                                    wavefronts, nrhs_max, rhs_split, solve_mode, C.byref(h)))
        del keep
        self.ctx, self.hm, self.h, self.precision, self.workers = ctx, list(hms), h, precision, workers
        return self

    @property
    def rows(self):
        return int(lib().pp_solver_rows(self.h))

    def members(self):
        out = []
        ro, n, r, ao = C.c_int64(), C.c_int(), C.c_double(), C.c_int64()
        for i in range(check(lib().pp_solver_num_members(self.h))):
            check(lib().pp_solver_member(self.h, i, C.byref(ro), C.byref(n), C.byref(r)))
            check(lib().pp_solver_member_arena(self.h, i, C.byref(ao)))
            out.append(dict(row_off=ro.value, n=n.value, residual=r.value, arena_off=ao.value))
        return out

    def use_graph(self, on=True):
        check(lib().pp_solver_use_graph(self.h, 1 if on else 0))

    def last_factor_ms(self):
        return float(lib().pp_solver_last_factor_ms(self.h))

    def factor(self, hip_stream=None):
        check(lib().pp_solver_factor(self.h, hip_stream))

    def solve(self, b_ptr, x_ptr, nrhs, ld=None, tol=1e-10, max_iter=30, hip_stream=None):
        rep = SolveReport()
        check(lib().pp_solver_solve(self.h, C.c_void_p(int(b_ptr)), C.c_void_p(int(x_ptr)), nrhs, ld or nrhs, tol,
                                    max_iter, hip_stream, C.byref(rep)))
        return {k: getattr(rep, k) for k, _ in SolveReport._fields_ if not k.startswith("pad")}

    def residual(self, b_ptr, x_ptr, r_ptr, nrhs, ld=None):
        check(lib().pp_solver_residual(self.h, C.c_void_p(int(b_ptr)), C.c_void_p(int(x_ptr)),
                                       C.c_void_p(int(r_ptr)), nrhs, ld or nrhs, None))

    def factor_arena(self):
        p = C.c_void_p()
        eb = C.c_int()
        check(lib().pp_solver_factor_arena(self.h, C.byref(p), C.byref(eb)))
        return p.value, eb.value

    def factor_arena_host(self):
        """The whole factor arena (all members, inverse areas included) as
        float64 on the host."""
        import ctypes as _C
        ptr, eb = self.factor_arena()
        total = int(lib().pp_solver_storage_elems(self.h))
        raw = np.empty(total, np.float64 if eb == 8 else np.float32)
        check(lib().pp_copy_d2h(self.ctx.h, raw.ctypes.data_as(_C.c_void_p), _C.c_void_p(ptr), eb * total))
        return raw.astype(np.float64)

    def factor_stream_ms(self):
        return float(lib().pp_stream_last_ms(lib().pp_solver_factor_stream(self.h)))

    def factors_to_bytes(self):
        """Byte string of the factor arenas (after factor()), loadable with
        factors_from_bytes into a solver over the same matrices, precision and
        solve mode."""
        return _encode(lib().pp_solver_encode_factors, self.h)

    def factors_from_bytes(self, data):
        """Loads factors made by factors_to_bytes; solve() then runs without
        factor()."""
        data = bytes(data)
        check(lib().pp_solver_decode_factors(self.h, data, len(data)))

    @property
    def factor_flops(self):
        return float(lib().pp_solver_factor_flops(self.h))

    def solve_flops(self, nrhs):
        return float(lib().pp_solver_solve_flops(self.h, nrhs))

    def close(self):
        if getattr(self, "h", None):
            lib().pp_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


from .hmatrix import *  # noqa: E402,F401,F403
