"""CPU oracle of the H-LU factor + solve path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this package, and only as the checker (or the timed CPU baseline); the
product never imports it.

  liboracle.so      plain-C FP64 restatement of the reference routines
                    (oracle.c cites src/kernel.cu line by line), the
                    instruction-stream executor and a dense no-pivot LU;
  _ref/libref_rts.so the reference's own rts/ops.cpp + rts/dependency.cpp,
                    compiled from /root/reference (built only where that
                    tree exists; tests skip it elsewhere);
  scheduler.py      pure-Python restatement of rts/scheduler.cpp.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref_rts.so")

_lib = None
_ref = None


def build(ref=True):
    """Compile liboracle.so (and the reference rts library when the reference
    tree is present)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if ref and os.path.isdir("/root/reference/rts"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build(ref=False)
        L = C.CDLL(LIB)
        dp, ip, i32 = C.POINTER(C.c_double), C.c_int, C.c_int32
        L.or_getrf.argtypes = [dp, ip, ip, ip]
        L.or_trsml.argtypes = [dp, dp, ip, ip, ip, ip, ip]
        L.or_trsmr.argtypes = [dp, dp, ip, ip, ip, ip, ip]
        L.or_trsmu.argtypes = [dp, dp, ip, ip, ip, ip]
        L.or_trmm.argtypes = [dp, dp, ip, ip, ip, ip]
        L.or_gemm.argtypes = [C.c_double, C.c_double, dp, dp, dp, ip, ip, ip, ip, ip, ip, ip, ip]
        L.or_givens_qr.argtypes = [dp, ip, ip, ip, dp, ip]
        L.or_accum.argtypes = [dp, dp, dp, dp, ip, ip, ip, ip, ip, ip, ip, ip, dp]
        L.or_exec.argtypes = [C.POINTER(i32), C.POINTER(dp), dp, dp]
        L.or_exec.restype = ip
        L.or_run_streams.argtypes = [C.POINTER(C.POINTER(i32)), ip, C.POINTER(dp), dp, C.c_int64, dp, ip, ip]
        L.or_run_streams.restype = ip
        L.or_getrf_pivoted.argtypes = [dp, ip, ip, ip, C.POINTER(i32)]
        L.or_getrf_pivoted.restype = ip
        L.or_getrf_pivoted_f32.argtypes = [C.POINTER(C.c_float), ip, ip, ip, C.POINTER(i32)]
        L.or_getrf_pivoted_f32.restype = ip
        L.or_dense_lu.argtypes = [dp, ip, ip]
        L.or_dense_lu_solve.argtypes = [dp, ip, dp, ip, ip]
        L.or_dense_residual.argtypes = [dp, ip, dp, dp, ip, ip]
        L.or_dense_residual.restype = C.c_double
        L.or_sketch.argtypes = [dp, C.c_int64, C.c_uint64]
        L.or_kernel_entry.argtypes = [ip, dp, ip, ip, C.c_double, C.c_double]
        L.or_kernel_entry.restype = C.c_double
        L.or_kernel_block.argtypes = [ip, dp, ip, ip, ip, ip, C.c_double, C.c_double, dp]
        L.or_aca.argtypes = [ip, dp, ip, ip, ip, ip, C.c_double, C.c_double, ip, dp, dp, C.POINTER(C.c_int)]
        L.or_aca_dense.argtypes = [dp, ip, ip, ip, dp, dp, C.POINTER(C.c_int)]
        L.or_truncate.argtypes = [dp, dp, ip, ip, ip, ip, dp, dp]
        L.or_truncate.restype = ip
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(REF_LIB)


def ref():
    global _ref
    if _ref is None:
        R = C.CDLL(REF_LIB)
        u = C.c_uint64
        R.ref_check_overlap.argtypes = [u] * 8
        R.ref_check_overlap.restype = C.c_int
        R.ref_table_new.restype = C.c_void_p
        R.ref_table_free.argtypes = [C.c_void_p]
        R.ref_table_update.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_int]
        R.ref_table_lookup.argtypes = [C.c_void_p, C.c_int64, C.c_int64]
        R.ref_table_lookup.restype = C.c_int
        for f in (R.ref_table_from, R.ref_table_to):
            f.argtypes = [C.c_void_p, C.c_int64, C.POINTER(C.c_int64), C.c_int]
            f.restype = C.c_int
        for f in (R.ref_table_count_from, R.ref_table_count_to):
            f.argtypes = [C.c_void_p, C.c_int64]
            f.restype = C.c_int64
        _ref = R
    return _ref


def _d(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def getrf(m):
    m = np.ascontiguousarray(m, np.float64).copy()
    lib().or_getrf(_d(m), m.shape[1], m.shape[0], m.shape[1])
    return m


def trsml(b, l, b_t=False):
    b = np.ascontiguousarray(b, np.float64).copy()
    l = np.ascontiguousarray(l, np.float64)
    ny, nx = (b.shape[1], b.shape[0]) if b_t else b.shape
    lib().or_trsml(_d(b), _d(l), nx, ny, b.shape[1], l.shape[1], int(b_t))
    return b


def trsmr(b, u, b_t=False):
    b = np.ascontiguousarray(b, np.float64).copy()
    u = np.ascontiguousarray(u, np.float64)
    ny, nx = (b.shape[1], b.shape[0]) if b_t else b.shape
    lib().or_trsmr(_d(b), _d(u), nx, ny, b.shape[1], u.shape[1], int(b_t))
    return b


def trsmu(b, u):
    b = np.ascontiguousarray(b, np.float64).copy()
    u = np.ascontiguousarray(u, np.float64)
    lib().or_trsmu(_d(b), _d(u), b.shape[1], b.shape[0], b.shape[1], u.shape[1])
    return b


def trmm(b, t):
    b = np.ascontiguousarray(b, np.float64).copy()
    t = np.ascontiguousarray(t, np.float64)
    lib().or_trmm(_d(b), _d(t), b.shape[1], b.shape[0], b.shape[1], t.shape[1])
    return b


def gemm(alpha, beta, m, a, b, a_t=False, b_t=False):
    m = np.ascontiguousarray(m, np.float64).copy()
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    k = a.shape[0] if a_t else a.shape[1]
    lib().or_gemm(alpha, beta, _d(m), _d(a), _d(b), m.shape[0], m.shape[1], k, m.shape[1], a.shape[1], b.shape[1],
                  int(a_t), int(b_t))
    return m


def givens_qr(y):
    y = np.ascontiguousarray(y, np.float64).copy()
    ny, k = y.shape
    q = np.zeros((ny, k))
    lib().or_givens_qr(_d(y), k, ny, k, _d(q), k)
    return q, np.triu(y[:k])


def sketch(count, seed=0x5EED):
    out = np.zeros(max(count, 1))
    lib().or_sketch(_d(out), count, seed)
    return out[:count]


def accum(u1, v1, u2, v2, rnd=None):
    u1 = np.ascontiguousarray(u1, np.float64).copy()
    v1 = np.ascontiguousarray(v1, np.float64).copy()
    u2 = np.ascontiguousarray(u2, np.float64)
    v2 = np.ascontiguousarray(v2, np.float64)
    ny, k1 = u1.shape
    nx, k2 = v1.shape[0], u2.shape[1]
    if rnd is None:
        rnd = sketch(nx * k1)
    rnd = np.ascontiguousarray(rnd, np.float64)
    lib().or_accum(_d(u1), _d(v1), _d(u2), _d(v2), nx, ny, k1, k2, k1, k1, k2, k2, _d(rnd))
    return u1, v1


_ENTRY_WORDS = {0: 3, 1: 8, 2: 13, 3: 13, 4: 17, 5: 17, 6: 23, 7: 29, 8: 21, 9: 13, 10: 11}


def sketch_need(streams):
    """Sketch values the streams' accum entries read: max of nx * k1."""
    need = 1
    for st in streams:
        i, n = 0, len(st)
        while i < n:
            op = int(st[i])
            if op == 0:
                break
            if op == 2:
                i += 3
                continue
            o = int(st[i + 2])
            if o == 8:
                need = max(need, int(st[i + 11]) * int(st[i + 13]))
            i += _ENTRY_WORDS[o]
    return need


def run_streams(streams, slots, n_insts, scratch_elems, rnd=None, threads=1):
    """Execute per-worker instruction streams (lists of int32 arrays) on host
    arrays `slots` (float64, modified in place).  rnd: the sketch values
    (default: the library's sketch, as many values as the accum entries
    read); a shorter array is refused rather than read past its end."""
    W = len(streams)
    sarr = [np.ascontiguousarray(s, np.int32) for s in streams]
    need = sketch_need(sarr)
    if rnd is None:
        rnd = sketch(need)
    rnd = np.ascontiguousarray(rnd, np.float64)
    if rnd.size < need:
        raise ValueError(f"sketch of {rnd.size} values; the accum entries read {need}")
    sp = (C.POINTER(C.c_int32) * W)(*[s.ctypes.data_as(C.POINTER(C.c_int32)) for s in sarr])
    pp = (C.POINTER(C.c_double) * len(slots))(*[_d(s) if s is not None else None for s in slots])
    stride = max(int(scratch_elems), 64)
    scratch = np.zeros(stride * W)
    rc = lib().or_run_streams(sp, W, pp, _d(scratch), stride, _d(rnd), n_insts, threads)
    if rc != 0:
        raise RuntimeError("oracle stream run failed (deadlock or bad entry)")


def kernel_block(kind, pts, r0, ny, c0, nx, param, shift=0.0):
    """ny x nx kernel entries at global (r0, c0) (kind 0: exp1d, 1: inv2d)."""
    pts = np.ascontiguousarray(pts, np.float64)
    out = np.zeros((ny, nx))
    lib().or_kernel_block(kind, _d(pts), r0, ny, c0, nx, param, shift, _d(out))
    return out


def aca(kind, pts, r0, ny, c0, nx, param, shift, rank):
    """(U, V, pivots) of the fixed-rank cross approximation of a kernel block."""
    pts = np.ascontiguousarray(pts, np.float64)
    u, v = np.zeros((ny, rank)), np.zeros((nx, rank))
    piv = np.zeros(2 * rank, np.int32)
    lib().or_aca(kind, _d(pts), r0, ny, c0, nx, param, shift, rank, _d(u), _d(v),
                 piv.ctypes.data_as(C.POINTER(C.c_int)))
    return u, v, piv.reshape(rank, 2)


def aca_dense(a, rank):
    """(U, V, pivots) of the fixed-rank cross approximation of a matrix."""
    a = np.ascontiguousarray(a, np.float64)
    ny, nx = a.shape
    u, v = np.zeros((ny, rank)), np.zeros((nx, rank))
    piv = np.zeros(2 * rank, np.int32)
    lib().or_aca_dense(_d(a), ny, nx, rank, _d(u), _d(v), piv.ctypes.data_as(C.POINTER(C.c_int)))
    return u, v, piv.reshape(rank, 2)


def truncate(u, v, r):
    """(Uo, Vo): rank-r truncation of U V^T (U ny x k, V nx x k) by the
    device builder's algorithm (pitch padded to a multiple of 32)."""
    ny, k = u.shape
    nx = v.shape[0]
    kp = (k + 31) // 32 * 32
    up, vp = np.zeros((ny, kp)), np.zeros((nx, kp))
    up[:, :k], vp[:, :k] = u, v
    uo, vo = np.zeros((ny, r)), np.zeros((nx, r))
    if lib().or_truncate(_d(up), _d(vp), ny, nx, kp, r, _d(uo), _d(vo)) != 0:
        raise RuntimeError("truncation failed (Cholesky pivot)")
    return uo, vo


def aca_rank(ny, nx, r):
    """Rank of the device builder's cross approximation of a low-rank leaf."""
    k = min(max(4 * r, 64, min(ny, nx) // 16), 256, ny, nx)
    return k if k > r else r


def build_arena(leaves, storage, kind, pts, param, shift):
    """(arena, pivots) an H-matrix structure gets from the device builder:
    leaves as HMatrix.leaves() dicts; pivots ((row, column) per step of each
    low-rank leaf's cross approximation) in leaf order."""
    arena = np.zeros(storage)
    pivs = []
    for lf in leaves:
        ny, nx = lf["ny"], lf["nx"]
        if lf["type"] == "dense":
            arena[lf["off"]:lf["off"] + ny * nx] = kernel_block(kind, pts, lf["row0"], ny, lf["col0"], nx, param,
                                                                shift).ravel()
        else:
            r = lf["rank"]
            k = aca_rank(ny, nx, r)
            u, v, piv = aca(kind, pts, lf["row0"], ny, lf["col0"], nx, param, shift, k)
            if k > r:
                u, v = truncate(u, v, r)
            arena[lf["off"]:lf["off"] + ny * r] = u.ravel()
            arena[lf["off_v"]:lf["off_v"] + nx * r] = v.ravel()
            pivs.append(piv)
    return arena, (np.concatenate(pivs) if pivs else np.zeros((0, 2), np.int32))


def getrf_pivoted(a):
    """(lu, p, info): partial-pivoted LU of the row-major block a (float64 or
    float32, computed in that precision); row i of P a is row p[i] of a."""
    a = np.array(a, copy=True, order="C")
    ny, nx = a.shape
    p = np.zeros(max(ny, 1), np.int32)
    pp = p.ctypes.data_as(C.POINTER(C.c_int32))
    if a.dtype == np.float32:
        info = lib().or_getrf_pivoted_f32(a.ctypes.data_as(C.POINTER(C.c_float)), ny, nx, nx, pp)
    else:
        a = a.astype(np.float64)
        info = lib().or_getrf_pivoted(_d(a), ny, nx, nx, pp)
    return a, p[:ny], info


def dense_lu(a, threads=8):
    a = np.ascontiguousarray(a, np.float64).copy()
    lib().or_dense_lu(_d(a), a.shape[0], threads)
    return a


def dense_lu_solve(lu, b):
    x = np.ascontiguousarray(b, np.float64).copy()
    if x.ndim == 1:
        x = x[:, None].copy()
    lib().or_dense_lu_solve(_d(np.ascontiguousarray(lu)), lu.shape[0], _d(x), x.shape[1], x.shape[1])
    return x


def dense_residual(a, x, b):
    a = np.ascontiguousarray(a, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return lib().or_dense_residual(_d(a), a.shape[0], _d(x), _d(b), x.shape[1], x.shape[1])


def hmatvec(leaves, arena, x):
    """y = A_H x from the leaf list of an H-matrix and its FP64 arena (dense
    leaves D, low-rank leaves U V^T), in numpy."""
    y = np.zeros_like(x)
    for lf in leaves:
        r0, c0, ny, nx = lf["row0"], lf["col0"], lf["ny"], lf["nx"]
        if lf["type"] == "dense":
            y[r0:r0 + ny] += arena[lf["off"]:lf["off"] + ny * nx].reshape(ny, nx) @ x[c0:c0 + nx]
        else:
            r = lf["rank"]
            u = arena[lf["off"]:lf["off"] + ny * r].reshape(ny, r)
            v = arena[lf["off_v"]:lf["off_v"] + nx * r].reshape(nx, r)
            y[r0:r0 + ny] += u @ (v.T @ x[c0:c0 + nx])
    return y
