"""Pins the 2-D Raviart-Thomas x DQ0 oracle (oracle/rt_oracle.c) without the GPU.

The oracle restates RTDQBasis (source/RT_DQ/rt_dq_basis.cc) in two
dimensions: a saddle-point solve per coarse face basis.  What fixes its
results independently of its own code:
  - constant coefficient: the coarse Raviart-Thomas function is the fine
    basis (closed form), so the flux block of M is its mass matrix over a;
  - coefficient a(x): the x bases stay the coarse functions (their flux is a
    gradient field for any a(x));
  - any coefficient: the flux block equals that of a stream-function solve
    written here (a different algorithm: a Q1 problem for the divergence-free
    correction), every fine cell balances its fluxes against -+|cell|/|K|,
    boundary fluxes are those of the coarse function, the saddle-point
    equations hold for the returned (sigma, u), and the complementary energy
    does not exceed that of the coarse function.
"""
import numpy as np
import pytest

from oracle import oracle as O

VERTS = np.array([[0.1, 0.2], [0.5, 0.2], [0.1, 0.45], [0.5, 0.45]])
HX, HY = 0.4, 0.25


def coarse_fluxes(l):
    """Fluxes of the 4 coarse Raviart-Thomas functions through the fine edges."""
    m = 1 << l
    nv, ne, _ = O.rt_sizes(l)
    F = np.zeros((4, ne))
    for q in range(m):
        for p in range(m + 1):
            F[0, q * (m + 1) + p] = (1 - p / m) / m
            F[1, q * (m + 1) + p] = (p / m) / m
    for q in range(m + 1):
        for p in range(m):
            F[2, nv + q * m + p] = (1 - q / m) / m
            F[3, nv + q * m + p] = (q / m) / m
    return F


GAUSS = np.array([0.5 - 0.5 / np.sqrt(3), 0.5 + 0.5 / np.sqrt(3)])


def ainv_at(a, c, qx, qy):
    """A^-1 at Gauss point (qx, qy) of fine cell c: a scalar per cell, or the
    inverse of the tensor (xx, xy, yy) given per point (a of shape (nf, 4, 3))."""
    a = np.asarray(a)
    if a.ndim == 1:
        return np.eye(2) / a[c]
    xx, xy, yy = a[c, qx + 2 * qy]
    return np.linalg.inv(np.array([[xx, xy], [xy, yy]]))


def rt_mass(l, a, S):
    """sum over fine cells of int sigma_i . A^-1 sigma_j, by 2-point Gauss."""
    m = 1 << l
    nv = m * (m + 1)
    hx, hy = HX / m, HY / m
    M = np.zeros((S.shape[0], S.shape[0]))
    for q in range(m):
        for p in range(m):
            fl, fr = S[:, q * (m + 1) + p], S[:, q * (m + 1) + p + 1]
            fb, ft = S[:, nv + q * m + p], S[:, nv + (q + 1) * m + p]
            for qy, t in enumerate(GAUSS):
                for qx, s in enumerate(GAUSS):
                    sig = np.stack([(fl * (1 - s) + fr * s) / hy, (fb * (1 - t) + ft * t) / hx], 1)
                    M += 0.25 * hx * hy * sig @ ainv_at(a, q * m + p, qx, qy) @ sig.T
    return M


def stream_function_fluxes(l, a):
    """sigma_n = coarse_n + curl psi, psi in Q1 with zero boundary values minimising
    int (coarse_n + curl psi) . A^-1 (coarse_n + curl psi) (the divergence-free
    correction)."""
    m = 1 << l
    hx, hy = HX / m, HY / m
    nn = (m + 1) ** 2
    K = np.zeros((nn, nn))
    R = np.zeros((nn, 4))
    for q in range(m):
        for p in range(m):
            nodes = [q * (m + 1) + p, q * (m + 1) + p + 1, (q + 1) * (m + 1) + p,
                     (q + 1) * (m + 1) + p + 1]
            for qy, t in enumerate(GAUSS):
                for qx, s in enumerate(GAUSS):
                    w = 0.25 * hx * hy
                    B = ainv_at(a, q * m + p, qx, qy)
                    phx = np.array([-(1 - t), 1 - t, -t, t]) / hx
                    phy = np.array([-(1 - s), -s, 1 - s, s]) / hy
                    curl = np.stack([phy, -phx], axis=1)  # (d_y phi, -d_x phi)
                    K[np.ix_(nodes, nodes)] += w * curl @ B @ curl.T
                    X = (p + s) / m
                    Y = (q + t) / m
                    sc = np.array([[(1 - X) / HY, 0], [X / HY, 0], [0, (1 - Y) / HX],
                                   [0, Y / HX]])
                    R[nodes, :] -= w * curl @ B @ sc.T
    inner = [j * (m + 1) + i for j in range(1, m) for i in range(1, m)]
    psi = np.zeros((nn, 4))
    if inner:
        psi[inner] = np.linalg.solve(K[np.ix_(inner, inner)], R[inner])
    S = coarse_fluxes(l).copy()
    nv = m * (m + 1)
    for q in range(m):
        for p in range(m + 1):
            S[:, q * (m + 1) + p] += psi[(q + 1) * (m + 1) + p] - psi[q * (m + 1) + p]
    for q in range(m + 1):
        for p in range(m):
            S[:, nv + q * m + p] -= psi[q * (m + 1) + p + 1] - psi[q * (m + 1) + p]
    return S


def lognormal(l, seed, sigma=1.0):
    return np.exp(sigma * np.random.default_rng(seed).normal(size=(1 << l) ** 2))


@pytest.mark.parametrize("l", [1, 2, 3])
def test_constant_coefficient_closed_form(l):
    a = 2.5
    M, b, S, U, st = O.rt_basis_cell(l, VERTS, a, 3.0)
    assert st == 0
    np.testing.assert_allclose(S, coarse_fluxes(l), rtol=0, atol=1e-13)
    ref = np.array([[HX / (3 * HY), HX / (6 * HY), 0, 0], [HX / (6 * HY), HX / (3 * HY), 0, 0],
                    [0, 0, HY / (3 * HX), HY / (6 * HX)], [0, 0, HY / (6 * HX), HY / (3 * HX)]]) / a
    np.testing.assert_allclose(M[:4, :4], ref, rtol=0, atol=1e-14)
    np.testing.assert_allclose(M[:4, 4], [1, -1, 1, -1], atol=1e-13)
    np.testing.assert_allclose(M[4, :4], [-1, 1, -1, 1], atol=1e-13)
    assert abs(M[4, 4]) < 1e-14
    np.testing.assert_allclose(b, [0, 0, 0, 0, 3.0 * HX * HY], rtol=1e-14, atol=0)


def test_layered_x_coefficient_keeps_x_bases():
    l = 3
    m = 1 << l
    ax = np.exp(np.random.default_rng(4).normal(size=m))
    a = np.tile(ax, m)  # depends on the column only
    M, b, S, U, st = O.rt_basis_cell(l, VERTS, a)
    assert st == 0
    F = coarse_fluxes(l)
    np.testing.assert_allclose(S[:2], F[:2], rtol=0, atol=1e-13)
    assert np.abs(S[2:] - F[2:]).max() > 1e-3  # the y bases do change
    # M_00 = int (1 - X)^2 / (Hy^2 a(x)) over the box, X = (x - x0)/Hx
    g = np.array([0.5 - 0.5 / np.sqrt(3), 0.5 + 0.5 / np.sqrt(3)])
    m00 = sum(0.5 * (1 - (p + s) / m) ** 2 / ax[p] for p in range(m) for s in g) * HX / m * HY / HY ** 2
    assert abs(M[0, 0] - m00) < 1e-14 * m00


@pytest.mark.parametrize("l,seed", [(1, 0), (2, 1), (3, 2), (4, 3)])
def test_matches_stream_function_solve(l, seed):
    a = lognormal(l, seed, 1.5)
    M, b, S, U, st = O.rt_basis_cell(l, VERTS, a)
    assert st == 0
    Sref = stream_function_fluxes(l, a)
    np.testing.assert_allclose(S, Sref, rtol=0, atol=1e-11 * np.abs(Sref).max())
    np.testing.assert_allclose(M[:4, :4], rt_mass(l, a, Sref), rtol=1e-11, atol=1e-14)


@pytest.mark.parametrize("l,seed", [(2, 5), (3, 6), (4, 7)])
def test_balance_boundary_and_equations(l, seed):
    m = 1 << l
    a = lognormal(l, seed, 1.5)
    f = np.random.default_rng(seed + 100).uniform(0.5, 2.0, size=m * m)
    M, b, S, U, st = O.rt_basis_cell(l, VERTS, a, f)
    assert st == 0
    nv, ne, nc = O.rt_sizes(l)
    F = coarse_fluxes(l)
    # boundary fluxes
    bnd = [q * (m + 1) + p for q in range(m) for p in (0, m)] + \
          [nv + q * m + p for q in (0, m) for p in range(m)]
    np.testing.assert_allclose(S[:, bnd], F[:, bnd], atol=1e-15)
    # per fine cell: out-flux = -+ |cell| / |K| = -+ 1/m^2
    sgn = np.array([-1.0, 1.0, -1.0, 1.0])
    for q in range(m):
        for p in range(m):
            out = (S[:, q * (m + 1) + p + 1] - S[:, q * (m + 1) + p]
                   + S[:, nv + (q + 1) * m + p] - S[:, nv + q * m + p])
            np.testing.assert_allclose(out, sgn / m ** 2, atol=1e-13)
    # the saddle-point equations with the recomputed assembled matrix
    A, Fl = O.rt_assemble(l, VERTS, a, f)
    inner = [e for e in range(ne) if e not in set(bnd)]
    for n in range(4):
        x = np.concatenate([S[n], U[n]])
        r = A @ x
        assert np.abs(r[inner]).max() < 1e-12 * np.abs(A).max()
        np.testing.assert_allclose(r[ne:], sgn[n] / m ** 2, atol=1e-12)
        assert abs(U[n].mean()) < 1e-13 * max(1.0, np.abs(U[n]).max())
    # element matrix from the assembled blocks, u := 1 (set_u_to_std)
    one = np.ones(nc)
    for i in range(4):
        assert abs(M[i, 4] - S[i] @ A[:ne, ne:] @ one) < 1e-13
        assert abs(M[4, i] - one @ A[ne:, :ne] @ S[i]) < 1e-13
    assert abs(b[4] - Fl[ne:].sum()) < 1e-14
    np.testing.assert_allclose(b[4], (f * HX * HY / m ** 2).sum(), rtol=1e-13)
    np.testing.assert_allclose(b[:4], 0.0)


@pytest.mark.parametrize("l,seed", [(2, 8), (3, 9)])
def test_energy_minimisation_and_symmetry(l, seed):
    a = lognormal(l, seed, 2.0)
    M, b, S, U, st = O.rt_basis_cell(l, VERTS, a)
    assert st == 0
    Mc = rt_mass(l, a, coarse_fluxes(l))
    Ms = M[:4, :4]
    np.testing.assert_allclose(Ms, Ms.T, rtol=1e-12, atol=1e-15)
    assert np.linalg.eigvalsh(Ms).min() > 0
    assert np.all(np.diag(Ms) <= np.diag(Mc) * (1 + 1e-12))
    np.testing.assert_allclose(M[:4, 4], -M[4, :4], atol=1e-13)


def test_rejects_non_box():
    v = VERTS.copy()
    v[3, 0] += 0.01
    assert O.rt_basis_cell(2, v, 1.0)[4] == 1


def random_tensors(l, seed, sigma=1.0):
    """Symmetric positive definite (xx, xy, yy) at every Gauss point: lognormal
    eigenvalues, random rotation per point."""
    rng = np.random.default_rng(seed)
    nf = (1 << l) ** 2
    lam = np.exp(sigma * rng.normal(size=(nf, 4, 2)))
    th = rng.uniform(0, np.pi, size=(nf, 4))
    c, s = np.cos(th), np.sin(th)
    xx = lam[..., 0] * c * c + lam[..., 1] * s * s
    yy = lam[..., 0] * s * s + lam[..., 1] * c * c
    xy = (lam[..., 0] - lam[..., 1]) * c * s
    return np.stack([xx, xy, yy], axis=-1)


@pytest.mark.parametrize("l", [1, 3])
def test_tensor_mode_with_isotropic_tensor_equals_scalar(l):
    a = lognormal(l, 20 + l, 1.5)
    t = np.zeros(((1 << l) ** 2, 4, 3))
    t[:, :, 0] = a[:, None]
    t[:, :, 2] = a[:, None]
    Ms, bs, Ss, Us, st0 = O.rt_basis_cell(l, VERTS, a)
    Mt, bt, St, Ut, st1 = O.rt_basis_cell(l, VERTS, t, tensor=True)
    assert st0 == 0 and st1 == 0
    np.testing.assert_allclose(St, Ss, rtol=0, atol=1e-13 * np.abs(Ss).max())
    np.testing.assert_allclose(Mt, Ms, rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("l,seed", [(2, 30), (3, 31), (4, 32)])
def test_tensor_mode_matches_stream_function_solve(l, seed):
    t = random_tensors(l, seed)
    M, b, S, U, st = O.rt_basis_cell(l, VERTS, t, tensor=True)
    assert st == 0
    Sref = stream_function_fluxes(l, t)
    np.testing.assert_allclose(S, Sref, rtol=0, atol=1e-11 * np.abs(Sref).max())
    np.testing.assert_allclose(M[:4, :4], rt_mass(l, t, Sref), rtol=1e-11, atol=1e-14)
    np.testing.assert_allclose(M[:4, :4], M[:4, :4].T, rtol=1e-12, atol=1e-15)
    Mc = rt_mass(l, t, coarse_fluxes(l))
    assert np.all(np.diag(M[:4, :4]) <= np.diag(Mc) * (1 + 1e-12))
    # the saddle-point equations with the recomputed assembled matrix
    m = 1 << l
    nv, ne, nc = O.rt_sizes(l)
    A, Fl = O.rt_assemble(l, VERTS, t, tensor=True)
    bnd = set([q * (m + 1) + p for q in range(m) for p in (0, m)] +
              [nv + q * m + p for q in (0, m) for p in range(m)])
    inner = [e for e in range(ne) if e not in bnd]
    sgn = np.array([-1.0, 1.0, -1.0, 1.0])
    for n in range(4):
        r = A @ np.concatenate([S[n], U[n]])
        assert np.abs(r[inner]).max() < 1e-12 * np.abs(A).max()
        np.testing.assert_allclose(r[ne:], sgn[n] / m ** 2, atol=1e-12)
