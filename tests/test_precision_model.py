"""How much refinement a low-precision factor of the local systems needs.

CPU emulation on the interior matrices K_II of cells of the headline field
(3-D, 8^3 fine cells, a = exp(g), g lognormal sigma 1.5, correlation length
1/64): a right-looking Cholesky of the Jacobi-equilibrated K_II with every
stored value rounded to the precision named, then iterative refinement with
FP64 residuals and the same factor.  It backs the precision choice of the
mixed solver: an FP32 factor needs one refinement step to reach 1e-10 on the
8 lifting solutions; BF16 products of the panel over an FP32 trailing matrix
need 2-3, BF16 storage 6-8 (the scaled condition numbers are 5-16, so the
refinement converges for all of them).  No GPU involved.
"""
import numpy as np
import pytest
import scipy.linalg as sl

from msfec_amd import fields as F


def to_bf16(x):
    """Round to bfloat16 (nearest, ties to even) and back to float64."""
    b = np.asarray(x, dtype=np.float32).view(np.uint32)
    r = ((b >> 16) & 1) + np.uint32(0x7FFF)
    return ((b + r) & np.uint32(0xFFFF0000)).view(np.float32).astype(np.float64)


def to_f32(x):
    return np.asarray(x, dtype=np.float32).astype(np.float64)


def cell_system(a, h, N=8):
    """K_II and the 8 lifting right-hand sides -K_IB g_B of one cube of side
    N h with a per fine cell (closed-form Q1 stiffness (a h / 12) S)."""
    S = np.array([[{0: 4.0, 1: 0.0, 2: -1.0, 3: -1.0}[bin(i ^ j).count("1")] for j in range(8)]
                  for i in range(8)])
    nn = (N + 1) ** 3
    K = np.zeros((nn, nn))
    for z in range(N):
        for y in range(N):
            for x in range(N):
                nodes = [(x + (v & 1)) + (N + 1) * ((y + ((v >> 1) & 1)) + (N + 1) * (z + (v >> 2)))
                         for v in range(8)]
                K[np.ix_(nodes, nodes)] += a[x + N * (y + N * z)] * h / 12.0 * S
    inner = [(i + 1) + (N + 1) * ((j + 1) + (N + 1) * (k + 1))
             for k in range(N - 1) for j in range(N - 1) for i in range(N - 1)]
    bnd = sorted(set(range(nn)) - set(inner))
    c = np.array([[n % (N + 1), (n // (N + 1)) % (N + 1), n // (N + 1) ** 2] for n in range(nn)]) / N
    G = np.stack([np.prod([c[:, d] if (k >> d) & 1 else 1 - c[:, d] for d in range(3)], axis=0)
                  for k in range(8)], 1)
    return K[np.ix_(inner, inner)], -K[np.ix_(inner, bnd)] @ G[bnd]


def chol_rounded(A, store, product=None):
    """Right-looking Cholesky; the trailing matrix and the panel are rounded
    by `store` after every step, the panel entering the update by `product`."""
    A = store(A.copy())
    n = A.shape[0]
    L = np.zeros_like(A)
    for k in range(n):
        d = np.sqrt(A[k, k])
        L[k, k] = d
        col = store(A[k + 1:, k] / d)
        L[k + 1:, k] = col
        v = product(col) if product else col
        A[k + 1:, k + 1:] = store(A[k + 1:, k + 1:] - np.outer(v, v))
    return L


def refinement_steps(K, R, L, dsc, target=1e-10, max_steps=20):
    xs = np.linalg.solve(K, R)

    def solve(b):
        return sl.solve_triangular(L.T, sl.solve_triangular(L, b / dsc[:, None], lower=True),
                                   lower=False) / dsc[:, None]

    x = solve(R)
    first = np.linalg.norm(x - xs) / np.linalg.norm(xs)
    for step in range(max_steps + 1):
        if np.linalg.norm(x - xs) / np.linalg.norm(xs) < target:
            return first, step
        x = x + solve(R - K @ x)
    return first, None


@pytest.fixture(scope="module")
def cells():
    coef = F.lognormal_coefficient(3, 16, 3, 1.5, 1.0 / 64, 20261202)
    pick = np.random.default_rng(11).choice(16 ** 3, 3, replace=False)
    return [cell_system(coef[c], 1.0 / 128) for c in pick]


def test_bf16_rounding_helper():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -3.140625, 1e-3])
    y = to_bf16(x)
    assert y[0] == 1.0
    assert y[1] == 1.0                # halfway between 1 and 1 + 2^-7: to the even one
    assert y[2] == 1.0 + 2 ** -7      # nearest
    assert y[3] == -3.140625          # exactly representable (8 significant bits)
    assert abs(y[4] - 1e-3) <= 2 ** -8 * 1e-3


def test_refinement_steps_by_factor_precision(cells):
    for K, R in cells:
        dsc = np.sqrt(np.diag(K))
        Ks = K / dsc[:, None] / dsc[None, :]
        assert np.linalg.cond(Ks) < 20.0  # the scaled systems are well conditioned
        first32, s32 = refinement_steps(K, R, chol_rounded(Ks, to_f32), dsc)
        firstb, sb = refinement_steps(K, R, chol_rounded(Ks, to_f32, to_bf16), dsc)
        firstbb, sbb = refinement_steps(K, R, chol_rounded(Ks, to_bf16), dsc)
        assert first32 < 1e-6 and s32 == 1
        assert 1e-5 < firstb < 1e-3 and sb in (2, 3)
        assert firstbb > 1e-3 and sbb is not None and 5 <= sbb <= 9


def jacobi_pcg_iterations(K, R, x0, target, maxit=200):
    """Jacobi-preconditioned CG on K x = R (8 columns at once) from x0: the
    iterations until every column is within `target` of the exact solution."""
    xs = np.linalg.solve(K, R)
    dinv = (1.0 / np.diag(K))[:, None]
    x = x0.copy()
    r = R - K @ x
    z = dinv * r
    p = z.copy()
    rz = np.sum(r * z, 0)
    for it in range(maxit + 1):
        if (np.linalg.norm(x - xs, axis=0) / np.linalg.norm(xs, axis=0)).max() < target:
            return it
        q = K @ p
        al = rz / np.sum(p * q, 0)
        x += al * p
        r -= al * q
        z = dinv * r
        rz1 = np.sum(r * z, 0)
        p = z + (rz1 / rz) * p
        rz = rz1
    return None


def test_cg_correction_iterations(cells):
    """The alternative to the two refinement sweeps: FP64 Jacobi-PCG on the
    cell, started from the FP32 solution -- still 10+ iterations of a
    27-point product with 8 right-hand sides, against 20+ from zero."""
    for K, R in cells:
        dsc = np.sqrt(np.diag(K))
        L = chol_rounded(K / dsc[:, None] / dsc[None, :], to_f32)
        x32 = sl.solve_triangular(L.T, sl.solve_triangular(L, R / dsc[:, None], lower=True),
                                  lower=False) / dsc[:, None]
        from32 = jacobi_pcg_iterations(K, R, x32, 1e-11)
        from0 = jacobi_pcg_iterations(K, R, np.zeros_like(R), 1e-11)
        assert 8 <= from32 <= 16 and 20 <= from0 <= 30, (from32, from0)
