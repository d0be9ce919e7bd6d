"""Parity of the HIP 3-D RT_DQ basis (msfec_rt3_* entry points) with the FP64
oracle (oq_rt3_basis_cell in oracle/nrt_oracle.c).

The oracle solves the reference's saddle-point system (fluxes and cell
values, u pinned in one cell, LU with partial pivoting and refinement); the
GPU writes the flux as the coarse Raviart-Thomas function plus the curl of
an interior Nedelec field and solves the curl-curl system by Jacobi-PCG to
|r| <= 1e-13 |b|.  Tolerance TOL = 1e-9 relative to the largest entry of
each basis's fluxes and of M.  At l = 4 (the reference examples' local
refinement), beyond the oracle's dense LU, the fluxes are checked through
the balance of every fine cell, the boundary traces, the symmetric positive
flux block and the complementary energy bound, against the oracle's sparse
direct solve of the saddle-point system, and against the closed form for
constant coefficients.
"""
import numpy as np
import pytest

from test_gpu_nrt import boxes, rand_coefs

pytestmark = pytest.mark.gpu

TOL = 1e-9


def run_gpu(verts, l, ainv, f=None, **kw):
    from msfec_amd import RTDQ3BasisBatch
    with RTDQ3BasisBatch(verts, n_refine_local=l, **kw) as rb:
        rb.set_coefficients(ainv, f)
        rb.run()
        return (rb.get_global_element_matrix(), rb.get_global_element_rhs(), rb.get_basis(),
                rb.get_cell_status(), rb.get_iterations())


def coarse_energy(v, l, ainv):
    """(A^-1 c_n, c_n) of the 6 coarse Raviart-Thomas functions by the
    2-point Gauss rule in every fine cell (ainv: fine cells x 8 x 6)."""
    m = 1 << l
    Hc = np.array([v[1, 0] - v[0, 0], v[2, 1] - v[0, 1], v[4, 2] - v[0, 2]])
    g = np.array([0.5 - 0.5 / np.sqrt(3), 0.5 + 0.5 / np.sqrt(3)])
    w = np.prod(Hc) / m ** 3 / 8
    comp = [0, 3, 5]
    E = np.zeros(6)
    idx = np.arange(m ** 3)
    cell = np.stack([idx % m, (idx // m) % m, idx // (m * m)], -1)
    for q in range(8):
        for d in range(3):
            X = (cell[:, d] + g[(q >> d) & 1]) / m
            o = [k for k in range(3) if k != d]
            for s in range(2):
                val = (X if s else 1 - X) / (Hc[o[0]] * Hc[o[1]])
                E[2 * d + s] += w * np.sum(ainv[:, q, comp[d]] * val ** 2)
    return E


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("l,n", [(1, 4), (2, 4), (3, 2)])
def test_rt3_matches_oracle(oracle_mod, l, n):
    print(f"case rt3 random l={l} n={n}", flush=True)
    rng = np.random.default_rng(500 + l)
    verts = boxes(rng, n)
    ainv, _, _, _ = rand_coefs(rng, n, l, reaction=False, load=False)
    f = rng.normal(size=(n, (1 << l) ** 3, 8))
    M, b, S, st, its = run_gpu(verts, l, ainv, f)
    assert (st == 0).all(), st
    print("iterations", its.min(), its.max(), flush=True)
    errs = []
    for c in range(n):
        Mo, bo, So, sto = oracle_mod.rt3_basis_cell(l, verts[c], ainv[c], f[c])
        assert sto == 0
        es = max(rel(S[c, i], So[i]) for i in range(6))
        em, eb = rel(M[c], Mo), rel(b[c], bo)
        errs.append((es, em, eb))
        assert es <= TOL and em <= TOL and eb <= 1e-12, (c, es, em, eb)
    print("max rel errors sigma / M / b:", np.max(errs, axis=0), flush=True)


def test_rt3_constant_coefficients_l4():
    """A^-1 = a I at l = 4: the fluxes are the coarse Raviart-Thomas functions
    and M the closed-form coarse element matrix."""
    print("case rt3 constant l=4", flush=True)
    import test_nrt_oracle as T
    from test_rt3_oracle import flux_block_closed_form
    a, fv = 1.7, 0.8
    M, b, S, st, its = run_gpu(T.VERTS[None], 4, [a, 0, 0, a, 0, a], fv)
    assert st[0] == 0
    print("iterations", its.min(), its.max(), flush=True)
    assert rel(M[0, :6, :6], flux_block_closed_form(a)) <= TOL
    sgn = np.array([-1.0, 1, -1, 1, -1, 1])
    np.testing.assert_allclose(M[0, :6, 6], -sgn, rtol=0, atol=1e-12)
    np.testing.assert_allclose(M[0, 6, :6], sgn, rtol=0, atol=1e-12)
    np.testing.assert_allclose(b[0, 6], fv * np.prod(T.H), rtol=1e-12)
    C = T.coarse_face_values(T.Dofs(4))
    assert np.abs(S[0] - C).max() <= TOL * np.abs(C).max()


def test_rt3_l4_reference_coefficients_invariants(oracle_mod):
    """Two cells of the reference example's 4^3 mesh at l = 4 with its
    rotated DiffusionInverse_A: every fine cell's net outflow is -+|c|/|K|,
    the boundary fluxes are the coarse traces, the flux block is symmetric
    positive definite, each basis's energy lies below that of the coarse
    function (the complementary energy principle), and the fluxes and M equal
    the oracle's sparse direct solve."""
    print("case rt3 reference coefficients l=4", flush=True)
    from msfec_amd import fields as F
    import test_nrt_oracle as T
    from test_rt3_oracle import cell_net_outflow
    l = 4
    verts = F.box_grid_3d(4)[[0, 21]]
    pts = F.gauss_points_3d(verts, l)
    ainv = F.diffusion_inverse_A_tensor(pts)
    M, b, S, st, its = run_gpu(verts, l, ainv, None)
    assert (st == 0).all(), st
    print("iterations", its.min(), its.max(), flush=True)
    m = 1 << l
    sgn = np.array([-1.0, 1, -1, 1, -1, 1])
    D = T.Dofs(l)
    C = T.coarse_face_values(D)
    bnd = np.array([D.face_on_boundary(t, p) for _, t, p in D.faces()])
    for c in range(2):
        out = cell_net_outflow(S[c], l)
        np.testing.assert_allclose(out, np.outer(sgn, np.full(m ** 3, 1 / m ** 3)),
                                   rtol=0, atol=1e-12)
        np.testing.assert_allclose(S[c][:, bnd], C[:, bnd], rtol=0, atol=1e-15)
        Mf = M[c, :6, :6]
        np.testing.assert_allclose(Mf, Mf.T, rtol=0, atol=1e-12 * np.abs(Mf).max())
        assert np.linalg.eigvalsh(Mf).min() > 0
        np.testing.assert_allclose(M[c, 6, :6], sgn, rtol=0, atol=1e-11)
        # complementary energy: M_ii <= the coarse function's (A^-1 c_i, c_i)
        E = coarse_energy(verts[c], l, ainv[c])
        assert (np.diag(Mf) <= E * (1 + 1e-12)).all(), (np.diag(Mf), E)
        assert (np.diag(Mf) < E).any()
        # and the oracle's sparse direct solve of the saddle-point system
        Ms, bs, Ss = oracle_mod.rt3_basis_cell_sparse(l, verts[c], ainv[c])
        es = max(rel(S[c, i], Ss[i]) for i in range(6))
        print(f"cell {c} vs sparse direct: sigma {es:.2e} M {rel(M[c], Ms):.2e}", flush=True)
        assert es <= TOL and rel(M[c], Ms) <= TOL


def test_rt3_iteration_limit_flags_the_cell():
    print("case rt3 iteration limit", flush=True)
    from msfec_amd import _lib as L
    rng = np.random.default_rng(551)
    verts = boxes(rng, 2)
    ainv, _, _, _ = rand_coefs(rng, 2, 3, reaction=False, load=False)
    M, b, S, st, its = run_gpu(verts, 3, ainv, None, max_iter=3)
    assert (st == L.CELL_NOT_CONVERGED).all()
    assert (its == 3).all()


def test_rt3_nonfinite_cell_is_flagged_alone(oracle_mod):
    """A NaN in one cell's A^-1 marks that cell MSFEC_CELL_NONFINITE; the
    other cells of the batch still match the oracle."""
    print("case rt3 non-finite cell l=2 n=4", flush=True)
    from msfec_amd import _lib as L
    rng = np.random.default_rng(561)
    n, l = 4, 2
    verts = boxes(rng, n)
    ainv, _, _, _ = rand_coefs(rng, n, l, reaction=False, load=False)
    ainv[2, 5, 3, 0] = np.nan
    M, b, S, st, its = run_gpu(verts, l, ainv, None)
    assert st[2] == L.CELL_NONFINITE
    for c in (0, 1, 3):
        assert st[c] == 0
        Mo, bo, So, sto = oracle_mod.rt3_basis_cell(l, verts[c], ainv[c])
        assert rel(M[c], Mo) <= TOL
        assert max(rel(S[c, i], So[i]) for i in range(6)) <= TOL
