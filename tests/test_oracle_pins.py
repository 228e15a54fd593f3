"""Pins of the CPU oracle's own checkers against independent mathematics.

The GPU tests decide "residual <= 1e-10" and "error against the FP64 oracle"
with or_dense_residual / or_dense_lu / or_dense_lu_solve, and compare accum
results with or_accum (the reference's Givens-QR projection,
src/kernel.cu:1160-1280).  Each is checked here against numpy / scipy or a
closed form, with cases chosen so that a plausible mistake (a dropped term,
a transposed operand, a wrong norm, a missing column) fails:

  * or_dense_residual: max over columns of ||b - A x|| / ||b|| against numpy,
    on a planted residual whose per-column norms are known exactly;
  * or_dense_lu / or_dense_lu_solve: factors against scipy's LU (which keeps
    the diagonal pivots of a column diagonally dominant matrix), solutions
    against numpy.linalg.solve, several right-hand sides with a pitch;
  * or_givens_qr: Q^T Q = I, Q R = Y, R upper triangular;
  * or_accum: exact when rank(U1 V1^T + U2 V2^T) <= k1 (U2 = U1 T), and in
    general equal to Q Q^T (U1 V1^T + U2 V2^T) with Q from numpy's QR of the
    sketch (A1 + A2) Omega, Omega the reference's sketch (pitch k1).
"""
import numpy as np
import pytest
import scipy.linalg


def test_dense_residual_matches_numpy(orc):
    rng = np.random.default_rng(1)
    n, k = 97, 5
    a = rng.standard_normal((n, n))
    x = rng.standard_normal((n, k))
    b = rng.standard_normal((n, k))
    want = np.max(np.linalg.norm(b - a @ x, axis=0) / np.linalg.norm(b, axis=0))
    assert abs(orc.dense_residual(a, x, b) - want) <= 1e-13 * want


def test_dense_residual_planted(orc):
    """b = A x + e with column c of e of norm 10^-c ||b_c|| exactly: the
    result is the largest ratio, 1 (column 0), not an average or a sum."""
    rng = np.random.default_rng(2)
    n, k = 64, 4
    a = rng.standard_normal((n, n)) + n * np.eye(n)
    x = rng.standard_normal((n, k))
    b0 = a @ x
    e = rng.standard_normal((n, k))
    b = b0.copy()
    for c in range(k):
        # scale e_c so that ||e_c|| = 10^-c ||b0_c + e_c||: solve for s in
        # ||s e|| = t ||b0 + s e|| by bisection on s
        t = 10.0 ** -(c + 1)
        lo, hi = 0.0, 1e6
        for _ in range(200):
            s = 0.5 * (lo + hi)
            if np.linalg.norm(s * e[:, c]) < t * np.linalg.norm(b0[:, c] + s * e[:, c]):
                lo = s
            else:
                hi = s
        b[:, c] = b0[:, c] + lo * e[:, c]
    got = orc.dense_residual(a, x, b)
    assert abs(got - 0.1) < 1e-9
    # exact solution: residual at roundoff level
    assert orc.dense_residual(a, x, b0) < 1e-14


def test_dense_lu_matches_scipy(orc):
    rng = np.random.default_rng(3)
    n = 150
    a = rng.standard_normal((n, n))
    a += np.diag(np.abs(a).sum(axis=0) + 1.0)  # column diagonally dominant: partial pivoting keeps the diagonal
    p, l, u = scipy.linalg.lu(a)
    assert np.array_equal(p, np.eye(n))
    lu = orc.dense_lu(a)
    assert np.max(np.abs(np.tril(lu, -1) - np.tril(l, -1))) < 1e-13
    assert np.max(np.abs(np.triu(lu) - u)) < 1e-12 * np.max(np.abs(u))


@pytest.mark.parametrize("n,k", [(1, 1), (33, 1), (200, 7), (513, 16)])
def test_dense_lu_solve_matches_numpy(orc, n, k):
    rng = np.random.default_rng(n + k)
    a = rng.standard_normal((n, n)) + 2.0 * n * np.eye(n)
    b = rng.standard_normal((n, k))
    x = orc.dense_lu_solve(orc.dense_lu(a), b)
    want = np.linalg.solve(a, b)
    assert x.shape == want.shape
    assert np.max(np.abs(x - want)) <= 1e-13 * np.max(np.abs(want))


def test_dense_lu_threads_agree(orc):
    rng = np.random.default_rng(4)
    a = rng.standard_normal((300, 300)) + 600 * np.eye(300)
    assert np.array_equal(orc.dense_lu(a, threads=1), orc.dense_lu(a, threads=8))


@pytest.mark.parametrize("ny,k", [(40, 8), (256, 32), (1000, 48), (64, 64)])
def test_givens_qr(orc, ny, k):
    rng = np.random.default_rng(ny)
    y = rng.standard_normal((ny, k))
    q, r = orc.givens_qr(y)
    assert np.max(np.abs(q.T @ q - np.eye(k))) < 1e-13
    assert np.max(np.abs(q @ r - y)) < 1e-13 * np.max(np.abs(y))
    assert np.array_equal(r, np.triu(r))


def _sketch_omega(orc, nx, k1):
    return orc.sketch(nx * k1).reshape(nx, k1)


@pytest.mark.parametrize("ny,nx,k1,k2", [(256, 256, 16, 16), (512, 256, 32, 32), (300, 700, 24, 8)])
def test_accum_exact_within_rank(orc, ny, nx, k1, k2):
    """U2 = U1 T keeps rank(U1 V1^T + U2 V2^T) <= k1: the recompression
    reproduces the sum to roundoff."""
    rng = np.random.default_rng(ny + nx)
    u1 = rng.standard_normal((ny, k1))
    v1 = rng.standard_normal((nx, k1))
    t = rng.standard_normal((k1, k2))
    u2 = u1 @ t
    v2 = rng.standard_normal((nx, k2))
    want = u1 @ v1.T + u2 @ v2.T
    uo, vo = orc.accum(u1, v1, u2, v2)
    assert np.max(np.abs(uo @ vo.T - want)) < 1e-10 * np.max(np.abs(want))
    # the new U is orthonormal (the reference stores Q in U1)
    assert np.max(np.abs(uo.T @ uo - np.eye(k1))) < 1e-12


@pytest.mark.parametrize("ny,nx,k1,k2", [(256, 256, 16, 16), (400, 300, 32, 32), (128, 512, 8, 24)])
def test_accum_is_the_sketched_projection(orc, ny, nx, k1, k2):
    """General ranks: U1 V1^T <- Q Q^T (A1 + A2) with Q an orthonormal basis
    of (A1 + A2) Omega (numpy QR), Omega the nx x k1 sketch."""
    rng = np.random.default_rng(7 * ny + nx)
    u1, v1 = rng.standard_normal((ny, k1)), rng.standard_normal((nx, k1))
    u2, v2 = rng.standard_normal((ny, k2)), rng.standard_normal((nx, k2))
    a = u1 @ v1.T + u2 @ v2.T
    om = _sketch_omega(orc, nx, k1)
    q, _ = np.linalg.qr(a @ om)
    want = q @ (q.T @ a)
    uo, vo = orc.accum(u1, v1, u2, v2)
    assert np.max(np.abs(uo @ vo.T - want)) < 1e-9 * np.max(np.abs(want))
    # a different sketch gives a different projection (the sketch matters)
    om2 = rng.uniform(-1, 1, (nx, k1))
    q2, _ = np.linalg.qr(a @ om2)
    assert np.max(np.abs(q2 @ (q2.T @ a) - want)) > 1e-6 * np.max(np.abs(want))


def test_sketch_values(orc):
    """The sketch is splitmix64 (seed 0x5EED) mapped to [-1, 1) with 53-bit
    uniforms: the first values follow from the published splitmix64 steps."""
    def splitmix(state):
        state = (state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return state, z ^ (z >> 31)
    st, want = 0x5EED, []
    for _ in range(8):
        st, z = splitmix(st)
        want.append(2.0 * ((z >> 11) * 2.0 ** -53) - 1.0)
    assert np.array_equal(orc.sketch(8), np.array(want))
