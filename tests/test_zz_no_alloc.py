"""The allocation-free single-op entry point (pp_exec_single_noalloc) against
the allocating one (pp_exec_single) through the C ABI: same kernel, same
inputs, so the results must be bit-identical -- on two calls in a row, the
second smaller than the first (the static scratch keeps the first call's
contents), for an accum (worker scratch and sketch) and a getrf (scratch for
the L11^{-1} blocks), in FP64 and mixed precision.  Also: a call whose
scratch exceeds the maximum is refused, and no device memory is taken."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _call(pp, ctx, fn, prec, words, bufs, scratch_bytes):
    L = pp.lib()
    w = np.ascontiguousarray(words, np.int32)
    ptrs = (C.c_void_p * len(bufs))(*[b.data_ptr() for b in bufs])
    wp = w.ctypes.data_as(C.POINTER(C.c_int32))
    if fn == "noalloc":
        rc = L.pp_exec_single_noalloc(ctx.h, prec, wp, len(w), ptrs, len(bufs), scratch_bytes, 0x5EED, None)
    else:
        rc = L.pp_exec_single(ctx.h, prec, wp, len(w), ptrs, len(bufs), scratch_bytes, 0x5EED)
    pp.check(rc)
    torch.cuda.synchronize()


def _accum_case(rng, nx, ny, k1, k2, prec):
    dt = torch.float64 if prec == 0 else torch.float32
    u1 = rng.standard_normal((ny, k1))
    v1 = rng.standard_normal((nx, k1)) * (0.5 ** np.arange(k1))
    u2 = rng.standard_normal((ny, k2)) * 1e-3
    v2 = rng.standard_normal((nx, k2))
    o1 = (ny * k1 + 31) // 32 * 32
    o2 = (o1 + nx * k1 + 31) // 32 * 32
    words = [1, 0, 8, 0, 1, 2, 3, 0, 0, 0, 0, nx, ny, k1, k2, k1, k1, k2, k2, o1, o2]
    host = [u1, v1, u2, v2]
    scratch = (o2 + ny * k1 + 4 * k1 * max(k1, k2) + 4096) * (8 if prec == 0 else 4)
    return words, host, dt, scratch


def _getrf_case(rng, n, prec):
    dt = torch.float64 if prec == 0 else torch.float32
    a = rng.standard_normal((n, n)) + 2.0 * n * np.eye(n)
    return [1, 0, 1, 0, 0, n, n, n], [a], dt, 1024 * 8


@pytest.mark.parametrize("prec", [0, 2])
def test_noalloc_matches_exec_single(pp, ctx, prec):
    rng = np.random.default_rng(17 + prec)
    cases = [_accum_case(rng, 1024, 2048, 32, 32, prec), _accum_case(rng, 300, 500, 16, 24, prec),
             _getrf_case(rng, 256, prec), _getrf_case(rng, 100, prec)]
    for k, (words, host, dt, scratch) in enumerate(cases):
        print(f"[case] noalloc prec={prec} case={k} op={words[2]}", flush=True)
        ref = [torch.from_numpy(h).to(dt).cuda() for h in host]
        got = [torch.from_numpy(h).to(dt).cuda() for h in host]
        _call(pp, ctx, "alloc", prec, words, ref, scratch)
        used = ctx.bytes_in_use()
        _call(pp, ctx, "noalloc", prec, words, got, scratch)
        assert ctx.bytes_in_use() == used
        for r, g in zip(ref, got):
            assert torch.equal(r, g), f"case {k} differs"


def test_noalloc_refuses_oversized_scratch(pp, ctx):
    L = pp.lib()
    cap = int(L.pp_exec_single_noalloc_max_bytes())
    assert cap == 4 << 20
    buf = torch.zeros(256 * 256, dtype=torch.float64, device="cuda")
    w = np.array([1, 0, 1, 0, 0, 256, 256, 256], np.int32)
    ptrs = (C.c_void_p * 1)(buf.data_ptr())
    rc = L.pp_exec_single_noalloc(ctx.h, 0, w.ctypes.data_as(C.POINTER(C.c_int32)), len(w), ptrs, 1, cap, 1, None)
    assert rc == -1 and "maximum" in L.pp_last_error().decode()
