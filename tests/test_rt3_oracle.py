"""The 3-D RT_DQ oracle (oq_rt3_basis_cell in oracle/nrt_oracle.c): the
reference's saddle-point system (source/RT_DQ/rt_dq_basis.cc:488-538) on a
box, pinned by closed forms for constant coefficients, the balance each fine
cell must satisfy, the prescribed boundary fluxes, and the complementary
energy principle."""
import numpy as np
import pytest

from oracle import oracle as O
from test_nrt_oracle import Dofs, H, VERTS, coarse_face_values, rand_coefs


def flux_block_closed_form(a):
    """(A^-1 psi_i, psi_j) of the coarse Raviart-Thomas functions for A^-1 = a I:
    a H_d / (H_o1 H_o2) times 1/3 (same side) or 1/6 (opposite sides) for two
    faces of one axis d, zero across axes."""
    Mf = np.zeros((6, 6))
    for d in range(3):
        o = [k for k in range(3) if k != d]
        base = a * H[d] / (H[o[0]] * H[o[1]])
        for s in range(2):
            for t in range(2):
                Mf[2 * d + s, 2 * d + t] = base * (1 / 3 if s == t else 1 / 6)
    return Mf


@pytest.mark.parametrize("l", [1, 2])
def test_constant_coefficients_give_the_coarse_functions(l):
    a, f = 1.7, 0.8
    M, b, S, st = O.rt3_basis_cell(l, VERTS, [a, 0, 0, a, 0, a], f)
    assert st == 0
    np.testing.assert_allclose(M[:6, :6], flux_block_closed_form(a), rtol=0, atol=1e-12)
    sgn = np.array([-1.0, 1, -1, 1, -1, 1])        # net outflow of the coarse function of face i
    np.testing.assert_allclose(M[:6, 6], -sgn, rtol=0, atol=1e-12)
    np.testing.assert_allclose(M[6, :6], sgn, rtol=0, atol=1e-12)
    assert abs(M[6, 6]) < 1e-14
    np.testing.assert_allclose(b, [0, 0, 0, 0, 0, 0, f * np.prod(H)], rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(S, coarse_face_values(Dofs(l)), rtol=0, atol=1e-12)


def cell_net_outflow(S, l):
    """Net outflow of every fine cell (x fastest) for face fluxes S (+axis)."""
    m = 1 << l
    nf1 = m * m * (m + 1)
    fx = S[..., :nf1].reshape(S.shape[:-1] + (m, m, m + 1))
    fy = S[..., nf1:2 * nf1].reshape(S.shape[:-1] + (m, m + 1, m))
    fz = S[..., 2 * nf1:].reshape(S.shape[:-1] + (m + 1, m, m))
    out = (fx[..., 1:] - fx[..., :-1]) + (fy[..., 1:, :] - fy[..., :-1, :]) + \
          (fz[..., 1:, :, :] - fz[..., :-1, :, :])
    return out.reshape(S.shape[:-1] + (m ** 3,))


@pytest.mark.parametrize("l", [1, 2, 3])
def test_random_coefficients_balance_traces_and_energy(l):
    ainv, _, _, _ = rand_coefs(l, 80 + l, reaction=False)
    f = np.random.default_rng(5).normal(size=((1 << l) ** 3, 8))
    M, b, S, st = O.rt3_basis_cell(l, VERTS, ainv, f)
    assert st == 0
    m = 1 << l
    # every fine cell: net outflow -+ |c| / |K| (the u equation, lower faces minus)
    sgn = np.array([-1.0, 1, -1, 1, -1, 1])
    np.testing.assert_allclose(cell_net_outflow(S, l), np.outer(sgn, np.full(m ** 3, 1 / m ** 3)),
                               rtol=0, atol=1e-11)
    # boundary fluxes: the coarse functions' traces
    D = Dofs(l)
    C = coarse_face_values(D)
    for f_, t, p in D.faces():
        if D.face_on_boundary(t, p):
            np.testing.assert_allclose(S[:, f_ - D.ne], C[:, f_ - D.ne], rtol=0, atol=1e-15)
    # the flux block is symmetric positive definite, the divergence entries +-1
    Mf = M[:6, :6]
    np.testing.assert_allclose(Mf, Mf.T, rtol=0, atol=1e-12 * np.abs(Mf).max())
    assert np.linalg.eigvalsh(Mf).min() > 0
    np.testing.assert_allclose(M[6, :6], sgn, rtol=0, atol=1e-11)
    np.testing.assert_allclose(M[:6, 6], -sgn, rtol=0, atol=1e-11)
    # complementary energy: sigma minimises (A^-1 tau, tau) over the fields
    # with its divergence and boundary flux, the coarse function among them
    A, _ = O.rt3_assemble(l, VERTS, ainv, f)
    nf = O.rt3_sizes(l)[1]
    Af = A[:nf, :nf]
    for i in range(6):
        assert M[i, i] <= C[i] @ Af @ C[i] * (1 + 1e-12)
        np.testing.assert_allclose(M[i, i], S[i] @ Af @ S[i], rtol=1e-12)
    # the load pairs the DQ0 basis with f
    w = np.prod(H) / m ** 3 / 8
    np.testing.assert_allclose(b[6], f.sum() * w, rtol=1e-12)
    assert not b[:6].any()


def test_rejects_bad_cells_and_levels():
    v = VERTS.copy()
    v[5, 1] += 0.02
    assert O.rt3_basis_cell(1, v, [1, 0, 0, 1, 0, 1])[3] == 1
    assert O.rt3_basis_cell(4, VERTS, [1, 0, 0, 1, 0, 1])[3] == 1


@pytest.mark.parametrize("l", [1, 2, 3])
def test_sparse_solve_matches_dense(l):
    """rt3_basis_cell_sparse (SuperLU on the triplet assembly) against the
    dense LU of oq_rt3_basis_cell."""
    ainv, _, _, _ = rand_coefs(l, 120 + l, reaction=False)
    f = np.random.default_rng(6).normal(size=((1 << l) ** 3, 8))
    M, b, S, st = O.rt3_basis_cell(l, VERTS, ainv, f)
    Ms, bs, Ss = O.rt3_basis_cell_sparse(l, VERTS, ainv, f)
    assert st == 0
    np.testing.assert_allclose(Ms, M, rtol=0, atol=1e-13 * np.abs(M).max())
    np.testing.assert_allclose(Ss, S, rtol=0, atol=1e-13 * np.abs(S).max())
    np.testing.assert_allclose(bs, b, rtol=1e-13, atol=1e-15)
    A, F = O.rt3_assemble(l, VERTS, ainv, f)
    As, Fs = O.rt3_assemble_sparse(l, VERTS, ainv, f)
    np.testing.assert_allclose(As.toarray(), A, rtol=0, atol=1e-14 * np.abs(A).max())
    np.testing.assert_allclose(Fs, F, rtol=0, atol=1e-15)


def test_sparse_solve_l4_constant_coefficients():
    """At l = 4 (beyond the dense LU) the sparse solve reproduces the closed
    forms for A^-1 = a I."""
    a = 0.9
    M, b, S = O.rt3_basis_cell_sparse(4, VERTS, [a, 0, 0, a, 0, a], 1.0)
    np.testing.assert_allclose(M[:6, :6], flux_block_closed_form(a), rtol=0, atol=1e-12)
    np.testing.assert_allclose(S, coarse_face_values(Dofs(4)), rtol=0, atol=1e-12)
    np.testing.assert_allclose(b[6], np.prod(H), rtol=1e-12)
