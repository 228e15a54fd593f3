"""Pins of the oracle's H-matrix construction (CPU).

The device builder has no reference counterpart; its CPU restatement
(oracle.c: or_kernel_block, or_aca, or_truncate) is pinned here by facts
that do not depend on how it is written:
  * kernel entries equal the seeded input generator's (hmatrix.Kernel) bit
    for bit for the 2-D kernel;
  * cross approximation reproduces an exactly low-rank matrix, and its
    residual vanishes on every pivot row and column (the interpolation
    property of cross approximation);
  * pivots follow the stated rule: the first row pivot is row 0, each column
    pivot is the largest residual entry of its row;
  * truncation equals the best rank-r approximation (Eckart-Young, via
    numpy's SVD) of the cross approximation;
  * a zero block gives zero factors and empty pivots.
"""
import numpy as np


def test_kernel_entries_match_input_generator(pp, orc):
    k = pp.config_kernel(1, 2048)
    assert np.array_equal(orc.kernel_block(1, k.pts, 100, 64, 1500, 80, 1e-3, k.shift), k.block(100, 64, 1500, 80))
    assert np.array_equal(orc.kernel_block(1, k.pts, 0, 64, 0, 64, 1e-3, k.shift), k.block(0, 64, 0, 64))
    k0 = pp.config_kernel(0, 512)
    a, b = orc.kernel_block(0, k0.pts, 0, 128, 128, 128, 0.1, k0.shift), k0.block(0, 128, 128, 128)
    assert np.max(np.abs(a - b) / np.abs(b)) < 1e-15


def test_aca_exact_on_low_rank(orc):
    rng = np.random.default_rng(1)
    a = rng.standard_normal((60, 5)) @ rng.standard_normal((5, 70))
    u, v, piv = orc.aca_dense(a, 8)
    assert np.linalg.norm(a - u @ v.T) / np.linalg.norm(a) < 1e-14
    assert piv[0, 0] == 0 and piv[0, 1] == int(np.argmax(np.abs(a[0])))


def test_aca_interpolates_pivot_crosses(pp, orc):
    k = pp.config_kernel(1, 2048)
    a = k.block(0, 512, 1024, 384)
    u, v, piv = orc.aca(1, k.pts, 0, 512, 1024, 384, 1e-3, k.shift, 40)
    r = a - u @ v.T
    scale = np.abs(a).max()
    assert np.abs(r[piv[:, 0]]).max() < 1e-14 * scale
    assert np.abs(r[:, piv[:, 1]]).max() < 1e-14 * scale
    assert len(set(piv[:, 0].tolist())) == 40 and len(set(piv[:, 1].tolist())) == 40
    # the same block through the explicit-matrix entry point
    u2, v2, piv2 = orc.aca_dense(a, 40)
    assert np.array_equal(piv, piv2) and np.array_equal(u, u2) and np.array_equal(v, v2)


def test_truncation_is_eckart_young(pp, orc):
    k = pp.config_kernel(1, 2048)
    u, v, _ = orc.aca(1, k.pts, 0, 512, 512, 512, 1e-3, k.shift, 96)
    uo, vo = orc.truncate(u, v, 24)
    w, s, zt = np.linalg.svd(u @ v.T, full_matrices=False)
    best = (w[:, :24] * s[:24]) @ zt[:24]
    assert np.linalg.norm(uo @ vo.T - best) / np.linalg.norm(best) < 1e-12
    # U carries the singular values, V has orthonormal columns
    assert np.allclose(vo.T @ vo, np.eye(24), atol=1e-12)
    assert np.allclose(np.linalg.norm(uo, axis=0), s[:24], rtol=1e-12)


def test_zero_block(orc):
    u, v, piv = orc.aca_dense(np.zeros((40, 30)), 6)
    assert not u.any() and not v.any() and (piv == -1).all()


def test_build_quality_matches_randomised_svd(pp, orc):
    n = 2048
    hm = pp.HMatrix("hodlr", n, levels=3, rank=16)
    k = pp.config_kernel(1, n)
    arena, piv = orc.build_arena(hm.leaves(), hm.storage, 1, k.pts, 1e-3, k.shift)
    d = k.dense()
    e = np.linalg.norm(d - pp.densify(hm, arena)) / np.linalg.norm(d)
    e_svd = np.linalg.norm(d - pp.densify(hm, pp.fill_arena(hm, k))) / np.linalg.norm(d)
    assert e < 1.1 * e_svd, (e, e_svd)
    assert piv.shape == (hm.build_pivots // 2, 2)
