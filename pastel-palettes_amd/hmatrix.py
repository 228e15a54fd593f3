"""Seeded inputs and H-matrix construction for the factor + solve path.

Everything here is input preparation (not timed, not on the device path):
  * a portable splitmix64 generator (53-bit uniforms, Box-Muller normals);
  * the point sets and kernel matrices of the benchmark configurations;
  * filling an H-matrix arena laid out by the library (dense leaves copied,
    low-rank leaves compressed to U V^T by a randomised SVD).

Seeds: points use 1000 + cfg, right-hand side j uses 2000 + 64 cfg + j.
"""
import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, n):
    """First n outputs of splitmix64 started at `seed`."""
    with np.errstate(over="ignore"):
        s = np.uint64(seed) + GOLDEN * np.arange(1, n + 1, dtype=np.uint64)
        z = (s ^ (s >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
        return z ^ (z >> np.uint64(31))


def uniform(seed, n):
    """n uniforms in [0, 1) with 53 random bits."""
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def normal(seed, n):
    """n standard normals by Box-Muller on consecutive uniform pairs."""
    m = (n + 1) // 2
    u = uniform(seed, 2 * m)
    u1, u2 = 1.0 - u[0::2], u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.empty(2 * m)
    z[0::2] = r * np.cos(2 * np.pi * u2)
    z[1::2] = r * np.sin(2 * np.pi * u2)
    return z[:n]


def sketch(seed, n):
    """The accum sketch matrix entries (uniform in [-1, 1)), same stream as libpastel."""
    return 2.0 * uniform(seed, n) - 1.0


SKETCH_SEED = 0x5EED


def rhs(cfg, n, nrhs):
    """Right-hand sides: column j is N(0,1) from seed 2000 + 64 cfg + j."""
    return np.stack([normal(2000 + 64 * cfg + j, n) for j in range(nrhs)], axis=1)


# ---------------------------------------------------------------- point sets / kernels

def points_1d(cfg, n):
    return np.sort(uniform(1000 + cfg, n))


def _spread16(v):
    v = v.astype(np.uint64) & np.uint64(0xFFFF)
    v = (v | (v << np.uint64(8))) & np.uint64(0x00FF00FF)
    v = (v | (v << np.uint64(4))) & np.uint64(0x0F0F0F0F)
    v = (v | (v << np.uint64(2))) & np.uint64(0x33333333)
    v = (v | (v << np.uint64(1))) & np.uint64(0x55555555)
    return v


def points_2d_morton(cfg, n):
    """n points uniform in the unit square, ordered along a Morton curve
    (16-bit quantisation per axis, x bits in even positions, stable sort)."""
    p = uniform(1000 + cfg, 2 * n).reshape(n, 2)
    q = np.minimum((p * 65536.0).astype(np.int64), 65535)
    code = _spread16(q[:, 0]) | (_spread16(q[:, 1]) << np.uint64(1))
    order = np.argsort(code, kind="stable")
    return p[order]


class Kernel:
    """A_ij = f(points i, j) + shift * delta_ij, evaluated block by block."""

    def __init__(self, pts, kind, shift):
        self.pts, self.kind, self.shift = np.asarray(pts), kind, float(shift)
        self.n = len(pts)

    def block(self, r0, ny, c0, nx):
        a = self.pts[r0:r0 + ny]
        b = self.pts[c0:c0 + nx]
        if self.kind == "exp1d":
            d = np.abs(a[:, None] - b[None, :])
            m = np.exp(-d / 0.1)
        elif self.kind == "inv2d":
            d = np.sqrt((a[:, None, 0] - b[None, :, 0]) ** 2 + (a[:, None, 1] - b[None, :, 1]) ** 2)
            m = 1.0 / (d + 1e-3)
        else:
            raise ValueError(self.kind)
        if self.shift != 0.0:
            lo, hi = max(r0, c0), min(r0 + ny, c0 + nx)
            if lo < hi:
                idx = np.arange(lo, hi)
                m[idx - r0, idx - c0] += self.shift
        return m

    def dense(self):
        return self.block(0, self.n, 0, self.n)


def config_kernel(cfg, n):
    """Kernel of a benchmark configuration: 0 = 1-D exp (shift n... 1024 at
    n = 1024, i.e. n), 1 = 2-D 1/(r + 1e-3) Morton (shift n / 4)."""
    if cfg == 0:
        return Kernel(points_1d(0, n), "exp1d", float(n))
    return Kernel(points_2d_morton(cfg, n), "inv2d", n / 4.0)


# ---------------------------------------------------------------- compression

def compress_block(m, rank, seed, oversample=10, power=1):
    """Rank-`rank` U V^T of a dense block by a randomised SVD (U carries the
    singular values)."""
    ny, nx = m.shape
    k = min(rank + oversample, ny, nx)
    rng = np.random.default_rng(seed)
    om = rng.standard_normal((nx, k))
    y = m @ om
    for _ in range(power):
        y, _ = np.linalg.qr(y)
        y = m @ (m.T @ y)
    q, _ = np.linalg.qr(y)
    bsmall = q.T @ m
    ub, s, vt = np.linalg.svd(bsmall, full_matrices=False)
    r = min(rank, len(s))
    u = (q @ ub[:, :r]) * s[:r]
    v = vt[:r].T
    if r < rank:  # pad with zero columns to the fixed rank
        u = np.hstack([u, np.zeros((ny, rank - r))])
        v = np.hstack([v, np.zeros((nx, rank - r))])
    return u, v


def fill_arena(hm, kernel, seed=7):
    """FP64 arena of H-matrix structure `hm` for `kernel`."""
    arena = np.zeros(hm.storage, np.float64)
    for i, lf in enumerate(hm.leaves()):
        blk = kernel.block(lf["row0"], lf["ny"], lf["col0"], lf["nx"])
        if lf["type"] == "dense":
            arena[lf["off"]:lf["off"] + lf["ny"] * lf["nx"]] = blk.ravel()
        else:
            u, v = compress_block(blk, lf["rank"], seed + i)
            r = lf["rank"]
            arena[lf["off"]:lf["off"] + lf["ny"] * r] = u.ravel()
            arena[lf["off_v"]:lf["off_v"] + lf["nx"] * r] = v.ravel()
    return arena


def densify(hm, arena, n=None):
    """Dense matrix represented by an arena (A_H, or the packed factors)."""
    n = n or hm.n
    a = np.zeros((n, n))
    for lf in hm.leaves():
        r0, c0, ny, nx = lf["row0"], lf["col0"], lf["ny"], lf["nx"]
        if lf["type"] == "dense":
            a[r0:r0 + ny, c0:c0 + nx] = arena[lf["off"]:lf["off"] + ny * nx].reshape(ny, nx)
        else:
            r = lf["rank"]
            u = arena[lf["off"]:lf["off"] + ny * r].reshape(ny, r)
            v = arena[lf["off_v"]:lf["off_v"] + nx * r].reshape(nx, r)
            a[r0:r0 + ny, c0:c0 + nx] = u @ v.T
    return a


def lu_factors(dense_lu):
    """Split a packed no-pivot LU into unit-lower L and upper U."""
    l = np.tril(dense_lu, -1) + np.eye(dense_lu.shape[0])
    u = np.triu(dense_lu)
    return l, u


def build_config(cfg, n=None, **over):
    """(structure, kernel, arena) for a benchmark configuration.

    cfg 0: n = 1024, 1-D exp kernel, all-dense binary partition, leaf 128.
    cfg 1: n = 16384, 2-D Morton 1/(r + 1e-3), HODLR 6 levels, rank 32.
    Keyword overrides: leaf, levels, rank.
    """
    from . import HMatrix
    if cfg == 0:
        n = n or 1024
        hm = HMatrix("dense", n, leaf=over.get("leaf", 128))
    else:
        n = n or 16384
        hm = HMatrix("hodlr", n, levels=over.get("levels", 6), rank=over.get("rank", 32))
    k = config_kernel(cfg, n)
    return hm, k, fill_arena(hm, k)


def dense_lu_flops(n):
    """n(n-1)/2 + n(n-1)(2n-1)/3 (reference closed form)."""
    return n * (n - 1) // 2 + n * (n - 1) * (2 * n - 1) // 3
