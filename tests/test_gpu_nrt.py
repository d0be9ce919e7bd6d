"""Parity of the HIP Ned_RT basis (libmsfec_hip.so, msfec_nrt_* entry points)
with the FP64 oracle (oracle/nrt_oracle.c) on seeded inputs.

The oracle solves each basis's mixed system over all interior edges and faces
by LU with partial pivoting; the GPU factors the same system in its symmetric
indefinite form (edge dofs first) by L D L^T without pivoting on the matrix
cores, and forms the element matrix from per-fine-cell local matrices.  Both
are FP64 direct solves; the tolerances are TOL_X = 1e-11 for the basis vectors
(relative to the largest entry of the basis) and TOL_M = 1e-11 for the element
matrix and load (relative to their largest entry).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL_X = 1e-11
TOL_M = 1e-11


def boxes(rng, n):
    v = np.empty((n, 8, 3))
    for c in range(n):
        x0 = rng.uniform(-1, 1, 3)
        H = rng.uniform(0.05, 0.6, 3)
        v[c] = [[x0[d] + ((k >> d) & 1) * H[d] for d in range(3)] for k in range(8)]
    return v


def rand_coefs(rng, n, l, reaction=True, load=True):
    nc = (1 << l) ** 3
    Lm = rng.normal(size=(n, nc, 8, 3, 3)) * 0.3 + np.eye(3)
    G = Lm @ np.swapaxes(Lm, -1, -2) + 0.1 * np.eye(3)
    ainv = np.stack([G[..., 0, 0], G[..., 0, 1], G[..., 0, 2], G[..., 1, 1], G[..., 1, 2],
                     G[..., 2, 2]], -1) * np.exp(rng.normal(size=(n, nc, 1, 1)))
    bq = np.exp(rng.normal(size=(n, nc, 8)))
    rq = np.exp(rng.normal(size=(n, nc, 8))) * 0.3 if reaction else None
    fq = rng.normal(size=(n, nc, 8, 3)) if load else None
    return ainv, bq, rq, fq


def run_gpu(verts, l, ainv, bq, rq=None, fq=None):
    from msfec_amd import NedRTBasisBatch
    with NedRTBasisBatch(verts, n_refine_local=l) as nb:
        nb.set_coefficients(ainv, bq, rq, fq)
        nb.run()
        return (nb.get_global_element_matrix(), nb.get_global_element_rhs(), nb.get_basis(),
                nb.get_cell_status())


def compare(M, b, X, o, cells):
    errs = []
    for k, c in enumerate(cells):
        Mo, bo, Xo = o["M"][k], o["b"][k], o["X"][k]
        ex = max(np.abs(X[c, i] - Xo[i]).max() / np.abs(Xo[i]).max() for i in range(18))
        em = np.abs(M[c] - Mo).max() / np.abs(Mo).max()
        eb = np.abs(b[c] - bo).max() / max(np.abs(bo).max(), 1e-300)
        errs.append((ex, em, eb))
        assert ex <= TOL_X, f"cell {c} basis {ex:.2e}"
        assert em <= TOL_M, f"cell {c} M {em:.2e}"
        assert eb <= TOL_M or np.abs(bo).max() == 0, f"cell {c} b {eb:.2e}"
    print("max rel errors basis / M / b:", np.max(errs, axis=0), flush=True)


@pytest.mark.parametrize("l", [1, 2])
def test_nrt_matches_oracle_random(oracle_mod, l):
    print(f"case nrt random l={l}", flush=True)
    rng = np.random.default_rng(700 + l)
    n = 6
    verts = boxes(rng, n)
    ainv, bq, rq, fq = rand_coefs(rng, n, l)
    M, b, X, st = run_gpu(verts, l, ainv, bq, rq, fq)
    assert (st == 0).all(), st
    o = oracle_mod.nrt_basis_batch(l, verts, ainv, bq, rq, fq, n_threads=6)
    assert (o["status"] == 0).all()
    compare(M, b, X, o, range(n))


def test_nrt_reference_example_coefficients(oracle_mod):
    """The coefficients of example_parameters/prm_ned_rt_test-01.prm (rotated
    DiffusionInverse_A with frequencies 17/16/15 and contrast 0.9, B with
    frequency 14, the gradient right-hand side, no reaction) on the 4^3
    coarse cells of the unit cube at l = 2; 6 cells against the oracle."""
    print("case nrt reference example 4^3 x l=2", flush=True)
    from msfec_amd import fields as F
    l = 2
    verts = F.box_grid_3d(4)
    pts = F.gauss_points_3d(verts, l)
    ainv = F.diffusion_inverse_A_tensor(pts)
    bq = F.diffusion_B(pts)
    fq = F.ned_rt_rhs(pts)
    M, b, X, st = run_gpu(verts, l, ainv, bq, None, fq)
    assert (st == 0).all()
    cells = [0, 5, 21, 42, 58, 63]
    o = oracle_mod.nrt_basis_batch(l, verts[cells], ainv[cells], bq[cells], None, fq[cells],
                                   n_threads=6)
    compare(M, b, X, o, cells)
    # the mixed blocks pair through the curl: M[face, edge] = -M[edge, face]
    np.testing.assert_allclose(M[:, 12:, :12], -np.swapaxes(M[:, :12, 12:], 1, 2), rtol=0,
                               atol=1e-12 * np.abs(M).max())


def test_nrt_constant_coefficients_reproduce_the_coarse_element(oracle_mod):
    """A^-1 = a I, constant B, no reaction: the bases are the coarse functions
    and M the standard element matrix, equal on every cell of a uniform grid."""
    print("case nrt constant coefficients", flush=True)
    from msfec_amd import fields as F
    verts = F.box_grid_3d(3)
    n = verts.shape[0]
    M, b, X, st = run_gpu(verts, 2, [2.0, 0, 0, 2.0, 0, 2.0], 0.5, None, [1.0, -2.0, 0.5])
    assert (st == 0).all()
    Mo, bo, Xo, sto = oracle_mod.nrt_basis_cell(2, verts[0], [2.0, 0, 0, 2.0, 0, 2.0], 0.5, None,
                                                [1.0, -2.0, 0.5])
    assert sto == 0
    for c in range(n):
        np.testing.assert_allclose(M[c], Mo, rtol=0, atol=TOL_M * np.abs(Mo).max())
        np.testing.assert_allclose(b[c], bo, rtol=0, atol=TOL_M * np.abs(bo).max())


def test_nrt_set_global_weights(oracle_mod):
    print("case nrt reconstruct", flush=True)
    from msfec_amd import NedRTBasisBatch
    rng = np.random.default_rng(77)
    n, l = 5, 2
    verts = boxes(rng, n)
    ainv, bq, rq, fq = rand_coefs(rng, n, l)
    with NedRTBasisBatch(verts, n_refine_local=l) as nb:
        nb.set_coefficients(ainv, bq, rq, fq)
        nb.run()
        X = nb.get_basis()
        w = rng.normal(size=(n, 18))
        sig, u = nb.set_global_weights(w)
        ne = nb.ne
    np.testing.assert_allclose(sig, np.einsum("ci,cij->cj", w[:, :12], X[:, :12, :ne]),
                               rtol=1e-13, atol=1e-13 * np.abs(sig).max())
    np.testing.assert_allclose(u, np.einsum("ci,cij->cj", w[:, 12:], X[:, 12:, ne:]),
                               rtol=1e-13, atol=1e-13 * np.abs(u).max())


def test_nrt_device_inputs_and_outputs(oracle_mod):
    """Vertices and coefficients in device memory (DeviceBuffer, the HIP
    runtime of the library), element matrix and bases written to device
    memory: bit for bit the host-pointer results."""
    print("case nrt device pointers", flush=True)
    from msfec_amd import DeviceBuffer, NedRTBasisBatch, _lib as L
    rng = np.random.default_rng(91)
    n, l = 4, 2
    verts = boxes(rng, n)
    ainv, bq, rq, fq = rand_coefs(rng, n, l)
    M_h, b_h, X_h, _ = run_gpu(verts, l, ainv, bq, rq, fq)
    bufs = [DeviceBuffer(array=a) for a in (verts, ainv, bq, rq, fq)]
    try:
        with NedRTBasisBatch(bufs[0], n_refine_local=l) as nb:
            nb.set_coefficients(*bufs[1:])
            nb.run()
            Md, bd, Xd = (DeviceBuffer((n, 18, 18)), DeviceBuffer((n, 18)),
                          DeviceBuffer((n, 18, nb.N)))
            bufs += [Md, bd, Xd]
            L.check(nb._lib.msfec_nrt_get_element(nb._h, Md.ptr, bd.ptr, L.DEVICE))
            L.check(nb._lib.msfec_nrt_get_basis(nb._h, Xd.ptr, L.DEVICE))
            np.testing.assert_array_equal(Md.numpy(), M_h)
            np.testing.assert_array_equal(bd.numpy(), b_h)
            np.testing.assert_array_equal(Xd.numpy(), X_h)
    finally:
        for bf in bufs:
            bf.free()


def test_nrt_flags_bad_cells():
    """A negative definite A^-1 makes the first edge pivot negative; a NaN
    coefficient gives a non-finite matrix.  Good cells are unaffected."""
    print("case nrt status flags", flush=True)
    rng = np.random.default_rng(5)
    n, l = 3, 1
    verts = boxes(rng, n)
    ainv, bq, rq, fq = rand_coefs(rng, n, l)
    ainv[1] = -ainv[1]
    bq[2, 0, 0] = np.nan
    M, b, X, st = run_gpu(verts, l, ainv, bq, rq, fq)
    assert st[0] == 0
    assert st[1] == 2
    assert st[2] in (2, 3)
