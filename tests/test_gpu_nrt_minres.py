"""The matrix-free MINRES path of the HIP Ned_RT basis (n_refine_local 3 and
4, or any l with MSFEC_SOLVER_CG) against the FP64 oracle (oracle/nrt_oracle.c).

MINRES stops on the preconditioned residual (relative tolerance 1e-14 by
default), so its bases carry an iteration error the direct solves do not;
the tolerance TOL_IT = 1e-8 (relative to the largest entry of each basis,
and of M) holds it.  At l = 2 and 3 the oracle's direct solve is the
reference; at l = 4 (the reference's example refinement, 22 320 interior
unknowns, beyond the oracle's dense LU) against the oracle's sparse direct
solve, against its assembled system (interior residual, boundary traces,
the element matrix by the reference's block rule) and, for constant
coefficients, against the closed forms (the bases are the coarse
functions).
"""
import numpy as np
import pytest

from test_gpu_nrt import boxes, rand_coefs

pytestmark = pytest.mark.gpu

TOL_IT = 1e-8


def run_minres(verts, l, ainv, bq, rq=None, fq=None, **kw):
    from msfec_amd import NedRTBasisBatch, _lib as L
    kw.setdefault("solver", L.SOLVER_CG)
    with NedRTBasisBatch(verts, n_refine_local=l, **kw) as nb:
        nb.set_coefficients(ainv, bq, rq, fq)
        nb.run()
        return (nb.get_global_element_matrix(), nb.get_global_element_rhs(), nb.get_basis(),
                nb.get_cell_status(), nb.get_iterations())


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("l,n", [(1, 4), (2, 4), (3, 2)])
def test_minres_matches_oracle(oracle_mod, l, n):
    print(f"case nrt minres random l={l} n={n}", flush=True)
    rng = np.random.default_rng(900 + l)
    verts = boxes(rng, n)
    ainv, bq, rq, fq = rand_coefs(rng, n, l)
    M, b, X, st, its = run_minres(verts, l, ainv, bq, rq, fq)
    assert (st == 0).all(), st
    assert (its > 0).all()
    print("iterations", its.min(), its.max(), flush=True)
    o = oracle_mod.nrt_basis_batch(l, verts, ainv, bq, rq, fq, n_threads=n)
    assert (o["status"] == 0).all()
    ex = max(rel(X[c, i], o["X"][c, i]) for c in range(n) for i in range(18))
    em = max(rel(M[c], o["M"][c]) for c in range(n))
    eb = max(rel(b[c], o["b"][c]) for c in range(n))
    print(f"max rel errors basis {ex:.2e} M {em:.2e} b {eb:.2e}", flush=True)
    assert ex <= TOL_IT and em <= TOL_IT and eb <= TOL_IT


def test_minres_matches_direct_l2(oracle_mod):
    """The same cells through the dense direct path and through MINRES."""
    print("case nrt minres vs dense l=2", flush=True)
    from msfec_amd import _lib as L
    rng = np.random.default_rng(931)
    n = 16
    verts = boxes(rng, n)
    ainv, bq, rq, fq = rand_coefs(rng, n, 2)
    Md, bd, Xd, std, _ = run_minres(verts, 2, ainv, bq, rq, fq, solver=L.SOLVER_CG)
    from test_gpu_nrt import run_gpu
    M, b, X, st = run_gpu(verts, 2, ainv, bq, rq, fq)
    assert (st == 0).all() and (std == 0).all()
    for c in range(n):
        assert rel(Md[c], M[c]) <= TOL_IT
        assert max(rel(Xd[c, i], X[c, i]) for i in range(18)) <= TOL_IT


def test_minres_l4_constant_coefficients():
    """A^-1 = a I, constant B, no reaction at l = 4: the sigma block of the
    edge bases and the u block of the face bases are the coarse functions'
    fine values, the sigma block of the face bases vanishes, and M and b are
    the closed-form coarse element matrix and load."""
    print("case nrt minres constant l=4", flush=True)
    import test_nrt_oracle as T
    l, a, beta, fvec = 4, 1.7, 0.6, np.array([0.3, -1.1, 0.8])
    D = T.Dofs(l)
    M, b, X, st, its = run_minres(T.VERTS[None], l, [a, 0, 0, a, 0, a], beta, None, fvec)
    assert st[0] == 0
    print("iterations", its.min(), its.max(), flush=True)
    Mw, bw = T._coarse_element_matrix(a, beta, fvec)
    print(f"M {rel(M[0], Mw):.2e} b {rel(b[0], bw):.2e}", flush=True)
    assert rel(M[0], Mw) <= TOL_IT and rel(b[0], bw) <= TOL_IT
    E, Fv = T.coarse_edge_values(D), T.coarse_face_values(D)
    assert np.abs(X[0, :12, :D.ne] - E).max() <= TOL_IT * np.abs(E).max()
    assert np.abs(X[0, 12:, D.ne:] - Fv).max() <= TOL_IT * np.abs(Fv).max()
    assert np.abs(X[0, 12:, :D.ne]).max() <= TOL_IT * np.abs(E).max()


def test_minres_l4_reference_example_against_assembled_system(oracle_mod):
    """One cell of the 4 x 4 x 4 coarse grid of prm_ned_rt_test-01 at its
    local refinement 4: the bases satisfy the oracle's assembled interior
    equations (loads (v, curl N_n) for the edge bases), carry the coarse
    traces on the boundary, and M and b follow from the assembled blocks by
    the reference's block rule (ned_rt_basis.cc:850-948)."""
    print("case nrt minres reference example l=4", flush=True)
    from msfec_amd import fields as F
    import test_nrt_oracle as T
    l = 4
    verts = F.box_grid_3d(4)[[21]]
    pts = F.gauss_points_3d(verts, l)
    ainv, bq, fq = F.diffusion_inverse_A_tensor(pts), F.diffusion_B(pts), F.ned_rt_rhs(pts)
    M, b, X, st, its = run_minres(verts, l, ainv, bq, None, fq)
    assert st[0] == 0
    print("iterations", its.min(), its.max(), flush=True)
    A, Fl = oracle_mod.nrt_assemble(l, verts[0], ainv[0], bq[0], None, fq[0])
    Fc = oracle_mod.nrt_curl_loads(l, verts[0])
    D = T.Dofs(l)
    ne, N = D.ne, D.N
    bnd = np.zeros(N, dtype=bool)
    for e, t, p in D.edges():
        bnd[e] = D.edge_on_boundary(t, p)
    for f, t, p in D.faces():
        bnd[f] = D.face_on_boundary(t, p)
    I = np.flatnonzero(~bnd)
    Xc = X[0]
    # the assembled matrix times the edge block and the face block of every basis
    Ye = A[:, :ne] @ Xc[:, :ne].T
    Yf = A[:, ne:] @ Xc[:, ne:].T
    # the block rule
    Mw = np.zeros((18, 18))
    for i in range(18):
        ri = slice(0, ne) if i < 12 else slice(ne, N)
        for j in range(18):
            Mw[i, j] = Xc[i, ri] @ (Ye if j < 12 else Yf)[ri, j]
    bw = np.array([0.0] * 12 + [Xc[i, ne:] @ Fl[ne:] for i in range(12, 18)])
    print(f"M {rel(M[0], Mw):.2e} b {rel(b[0], bw):.2e}", flush=True)
    assert rel(M[0], Mw) <= 1e-12 and rel(b[0], bw) <= 1e-12
    # interior equations, each row against the size of its terms
    R = (Ye + Yf)[I]
    R[:, :12] -= Fc[:, I].T
    np.abs(A, out=A)
    scale = (A @ np.abs(Xc.T))[I]
    scale[:, :12] += np.abs(Fc[:, I].T)
    res = np.abs(R).max(axis=0) / scale.max(axis=0)
    print("relative interior residual per basis", res.max(), flush=True)
    assert res.max() <= 1e-8
    # boundary traces: the coarse functions' values on the boundary dofs
    E, Fv = T.coarse_edge_values(D), T.coarse_face_values(D)
    Bv = np.zeros((18, N))
    Bv[:12, :ne] = E
    Bv[12:, ne:] = Fv
    np.testing.assert_allclose(Xc[:, bnd], Bv[:, bnd], rtol=0, atol=1e-15)
    # and the sparse direct solve of the same system
    Ms, bs, Xs = oracle_mod.nrt_basis_cell_sparse(l, verts[0], ainv[0], bq[0], None, fq[0])
    ex = max(rel(Xc[i], Xs[i]) for i in range(18))
    print(f"vs sparse direct: basis {ex:.2e} M {rel(M[0], Ms):.2e} b {rel(b[0], bs):.2e}", flush=True)
    assert ex <= TOL_IT and rel(M[0], Ms) <= TOL_IT and rel(b[0], bs) <= TOL_IT


def test_minres_l4_random_coefficients_against_sparse_direct(oracle_mod):
    """l = 4 with random boxes, rotated SPD A^-1 tensors, B, a reaction rate
    and loads at the Gauss points (the element additive Schwarz and coarse
    correction on coefficients unlike the reference example's) against the
    oracle's sparse direct solve."""
    print("case nrt minres random l=4 n=2", flush=True)
    rng = np.random.default_rng(961)
    n, l = 2, 4
    verts = boxes(rng, n)
    ainv, bq, rq, fq = rand_coefs(rng, n, l)
    M, b, X, st, its = run_minres(verts, l, ainv, bq, rq, fq)
    assert (st == 0).all(), st
    print("iterations", its.min(), its.max(), flush=True)
    for c in range(n):
        Ms, bs, Xs = oracle_mod.nrt_basis_cell_sparse(l, verts[c], ainv[c], bq[c], rq[c], fq[c])
        ex = max(rel(X[c, i], Xs[i]) for i in range(18))
        print(f"cell {c}: basis {ex:.2e} M {rel(M[c], Ms):.2e} b {rel(b[c], bs):.2e}", flush=True)
        assert ex <= TOL_IT and rel(M[c], Ms) <= TOL_IT and rel(b[c], bs) <= TOL_IT


def test_minres_iteration_limit_flags_the_cell():
    """A limit below what the solve needs leaves MSFEC_CELL_NOT_CONVERGED."""
    print("case nrt minres iteration limit", flush=True)
    from msfec_amd import _lib as L
    rng = np.random.default_rng(941)
    verts = boxes(rng, 2)
    ainv, bq, rq, fq = rand_coefs(rng, 2, 3)
    M, b, X, st, its = run_minres(verts, 3, ainv, bq, rq, fq, max_iter=5)
    assert (st == L.CELL_NOT_CONVERGED).all()
    assert (its == 5).all()
