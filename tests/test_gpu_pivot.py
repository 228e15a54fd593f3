"""Partial-pivoted LU and row permutation on the device against the CPU
restatement (oracle): bit-identical factors, permutations and info values
(both keep one correctly rounded division per multiplier and one fused
multiply-add per trailing element and step, and the same pivot rule), on
blocks that are not diagonally dominant, square and ragged, padded pitch,
batched, FP64 and FP32, with a zero pivot column.  The permutation kernel is
checked against numpy's row gather / scatter, and a bad permutation is
rejected before any row moves."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DT = {0: (np.float64, torch.float64), 1: (np.float32, torch.float32)}


def run_getrf(pp, ctx, prec, blocks, ld):
    npt, tt = DT[prec]
    count = len(blocks)
    ny, nx = blocks[0].shape
    stride = ny * ld
    host = np.zeros((count, ny, ld), npt)
    for b, a in enumerate(blocks):
        host[b, :, :nx] = a
    d = torch.from_numpy(host).cuda()
    piv = torch.zeros(count * ny, dtype=torch.int32, device="cuda")
    info = torch.full((count,), -7, dtype=torch.int32, device="cuda")
    # the library's stream is not torch's: the fills above finish first
    torch.cuda.synchronize()
    ctx.getrf_pivoted(prec, d.data_ptr(), ld, ny, nx, piv.data_ptr(), info.data_ptr(), count=count, stride=stride)
    torch.cuda.synchronize()
    return d.cpu().numpy()[:, :, :nx], piv.cpu().numpy().reshape(count, ny), info.cpu().numpy()


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("ny,nx,ld", [(1, 1, 1), (7, 7, 9), (64, 64, 64), (256, 256, 256), (100, 60, 61),
                                      (60, 100, 100), (300, 200, 203)])
def test_getrf_pivoted_bit_exact(pp, orc, ctx, prec, ny, nx, ld):
    print(f"[case] getrf_pivoted prec={prec} ny={ny} nx={nx} ld={ld}", flush=True)
    npt, _ = DT[prec]
    a = np.random.default_rng(ny * 31 + nx).standard_normal((ny, nx)).astype(npt)
    g, gp, gi = run_getrf(pp, ctx, prec, [a], ld)
    o, op, oi = orc.getrf_pivoted(a)
    assert gi[0] == oi == 0
    assert np.array_equal(gp[0], op)
    assert np.array_equal(g[0], o)


@pytest.mark.parametrize("prec", [0, 1])
def test_getrf_pivoted_batch_and_zero_column(pp, orc, ctx, prec):
    print(f"[case] getrf_pivoted batch prec={prec}", flush=True)
    npt, _ = DT[prec]
    rng = np.random.default_rng(77)
    blocks = [rng.standard_normal((96, 96)).astype(npt) for _ in range(3)]
    blocks[1][:, 5] = 0  # zero at and below the diagonal after any interchange
    g, gp, gi = run_getrf(pp, ctx, prec, blocks, 101)
    for b, a in enumerate(blocks):
        o, op, oi = orc.getrf_pivoted(a)
        assert gi[b] == oi
        assert np.array_equal(gp[b], op)
        assert np.array_equal(g[b], o)
    assert gi.tolist() == [0, 6, 0]


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("ny,nx,ld", [(256, 32, 32), (1000, 70, 75), (5, 300, 300)])
def test_apply_pivot(pp, ctx, prec, ny, nx, ld):
    print(f"[case] apply_pivot prec={prec} ny={ny} nx={nx}", flush=True)
    npt, _ = DT[prec]
    rng = np.random.default_rng(ny + nx)
    m = np.zeros((ny, ld), npt)
    m[:, :nx] = rng.standard_normal((ny, nx))
    p = rng.permutation(ny).astype(np.int32)
    d = torch.from_numpy(m.copy()).cuda()
    dp = torch.from_numpy(p).cuda()
    torch.cuda.synchronize()
    ctx.apply_pivot(prec, d.data_ptr(), ld, nx, ny, dp.data_ptr())
    torch.cuda.synchronize()
    fwd = d.cpu().numpy()
    assert np.array_equal(fwd[:, :nx], m[p, :nx])
    assert np.array_equal(fwd[:, nx:], m[:, nx:])  # the pitch padding is not touched
    ctx.apply_pivot(prec, d.data_ptr(), ld, nx, ny, dp.data_ptr(), recover=True)
    torch.cuda.synchronize()
    assert np.array_equal(d.cpu().numpy(), m)


def test_apply_pivot_rejects_non_permutation(pp, ctx):
    print("[case] apply_pivot bad permutation", flush=True)
    d = torch.zeros((8, 8), dtype=torch.float64, device="cuda")
    bad = torch.tensor([0, 1, 2, 3, 4, 5, 6, 6], dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    with pytest.raises(pp.PastelError):
        ctx.apply_pivot(0, d.data_ptr(), 8, 8, 8, bad.data_ptr())
    out = torch.tensor([0, 1, 2, 3, 4, 5, 6, 8], dtype=torch.int32, device="cuda")
    with pytest.raises(pp.PastelError):
        ctx.apply_pivot(0, d.data_ptr(), 8, 8, 8, out.data_ptr())


def test_pivoted_factor_solves(pp, orc, ctx):
    """The device factors of a block with small diagonal solve A x = b."""
    print("[case] getrf_pivoted solve", flush=True)
    rng = np.random.default_rng(3)
    n = 200
    a = rng.standard_normal((n, n))
    a[np.arange(n), np.arange(n)] *= 1e-8  # no-pivot LU would lose all accuracy
    g, gp, gi = run_getrf(pp, ctx, 0, [a], n)
    lu, p = g[0], gp[0]
    b = rng.standard_normal(n)
    y = b[p].copy()
    for i in range(n):
        y[i] -= lu[i, :i] @ y[:i]
    x = y.copy()
    for i in range(n - 1, -1, -1):
        x[i] = (y[i] - lu[i, i + 1:] @ x[i + 1:]) / lu[i, i]
    assert np.linalg.norm(a @ x - b) / np.linalg.norm(b) < 1e-10


@pytest.mark.parametrize("prec", [0, 1])
def test_getrf_pivoted_nan_columns(pp, orc, ctx, prec):
    """NaN entries never win the pivot search; a column that is NaN at and
    below the diagonal keeps the diagonal row (the pivot stays a row of the
    block), as in the sequential scan of the oracle."""
    print(f"[case] getrf_pivoted NaN columns prec={prec}", flush=True)
    npt, _ = DT[prec]
    rng = np.random.default_rng(21)
    a = rng.standard_normal((48, 48)).astype(npt)
    a[30:, 30:] = np.nan  # NaN from column 30 on, through the updates
    a[5, 2] = np.nan      # a lone NaN below a finite diagonal
    g, gp, gi = run_getrf(pp, ctx, prec, [a], 48)
    o, op, oi = orc.getrf_pivoted(a)
    assert sorted(gp[0].tolist()) == list(range(48))
    assert np.array_equal(gp[0], op)
    assert gi[0] == oi
    assert np.array_equal(g[0], o, equal_nan=True)
