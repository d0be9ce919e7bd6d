"""Parity of the HIP Q_Ned basis (libmsfec_hip.so, msfec_ned_* entry points)
with the FP64 oracle (oracle/ned_oracle.c) on seeded inputs.

Both sides solve in FP64 directly (LDL^T of the equilibrated band matrix on
the GPU, equilibrated LU with partial pivoting plus refinement in the oracle).
The sigma components of the vertex bases sit on a Q1 mass block of size
~h^2 beside curl-curl entries of size ~a / h^2, so their FP64 floor is about
1e-11 relative at l = 3 (the oracle against an 80-bit elimination) and
grows with the 4x wider scale gap at l = 4: on the 1024-cell l = 4 batch
the GPU's sigma is 1.35e-10 from the oracle's on the worst of the 8 cells
test_ned_headline_batch checks (7.5e-11 over the bench's 256), while u and M
agree to 1e-14.  The tolerances are TOL_SIGMA = 3e-10 for sigma, TOL_U = 1e-11
for u (relative to the largest entry of the basis part) and TOL_M = 1e-10 per
block of the element matrix (relative to the block's largest entry).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL_SIGMA = 3e-10
TOL_U = 1e-11
TOL_M = 1e-10


def boxes(rng, n):
    v = np.empty((n, 4, 2))
    for c in range(n):
        x0, y0 = rng.uniform(-1, 1, 2)
        hx, hy = rng.uniform(0.02, 0.5, 2)
        v[c] = [[x0, y0], [x0 + hx, y0], [x0, y0 + hy], [x0 + hx, y0 + hy]]
    return v


def run_gpu(verts, l, a, binv, r, f, full_rhs=True):
    from msfec_amd import NedBasisBatch
    with NedBasisBatch(verts, n_refine_local=l, full_rhs=full_rhs) as nb:
        nb.set_coefficients(a, binv, r, f)
        nb.run()
        return (nb.get_global_element_matrix(), nb.get_global_element_rhs(), nb.get_basis(),
                nb.get_cell_status())


def compare(l, M, b, P, o):
    ns = (2 ** l + 1) ** 2
    for c in range(M.shape[0]):
        for i in (0, 4):
            for j in (0, 4):
                ref = o["M"][c, i:i + 4, j:j + 4]
                err = np.abs(M[c, i:i + 4, j:j + 4] - ref).max() / np.abs(ref).max()
                assert err < TOL_M, (c, i, j, err)
        assert np.abs(b[c] - o["b"][c]).max() <= TOL_M * np.abs(o["b"][c]).max() + 1e-300
        es = np.abs(P[c, :, :ns] - o["Psi"][c, :, :ns]).max() / np.abs(o["Psi"][c, :, :ns]).max()
        eu = np.abs(P[c, :, ns:] - o["Psi"][c, :, ns:]).max() / np.abs(o["Psi"][c, :, ns:]).max()
        assert es < TOL_SIGMA, (c, es)
        assert eu < TOL_U, (c, eu)


@pytest.mark.parametrize("l", [1, 2, 3, 4])
def test_ned_matches_oracle_random(oracle_mod, l):
    print(f"case ned_random l={l}", flush=True)
    rng = np.random.default_rng(50 + l)
    n = 6
    nf = 4 ** l
    verts = boxes(rng, n)
    a = np.exp(rng.standard_normal((n, nf)))
    binv = np.exp(0.5 * rng.standard_normal((n, nf)))
    r = np.exp(rng.standard_normal((n, nf)))
    f = rng.standard_normal((n, nf, 2))
    M, b, P, st = run_gpu(verts, l, a, binv, r, f)
    assert (st == 0).all(), st
    o = oracle_mod.ned_basis_batch(l, verts, a, binv, r, f)
    assert (o["status"] == 0).all()
    compare(l, M, b, P, o)


@pytest.mark.parametrize("full_rhs", [True, False])
def test_ned_constant_coefficients_closed_form(oracle_mod, full_rhs):
    """Constant a, no reaction: the edge bases are the coarse Nedelec
    functions (sigma = 0, u = N_E) and the edge block is a s s^T / |K|."""
    print(f"case ned_const full_rhs={full_rhs}", flush=True)
    l = 3
    verts = np.array([[[0.1, 0.2], [0.7, 0.2], [0.1, 0.55], [0.7, 0.55]]])
    nf = 4 ** l
    M, b, P, st = run_gpu(verts, l, np.full((1, nf), 2.5), np.full((1, nf), 0.7),
                          np.zeros((1, nf)), np.ones((1, nf, 2)), full_rhs)
    assert st[0] == 0
    o = oracle_mod.ned_basis_batch(l, verts, 2.5, 0.7, 0.0, (1.0, 1.0), full_rhs)
    compare(l, M, b, P, o)
    ns = (2 ** l + 1) ** 2
    # sigma of the edge bases is 0 in exact arithmetic: what is left is the
    # rounding of the FP64 band solve, a few kappa eps with kappa ~ 600 for
    # the equilibrated system, on the scale 1 of sigma (the vertex bases'
    # sigma = phi_n); 1.3e-13 measured
    assert np.abs(P[0, 4:, :ns]).max() < 1e-12
    s = np.array([-1.0, 1.0, 1.0, -1.0])
    assert np.abs(M[0, 4:, 4:] - 2.5 / (0.6 * 0.35) * np.outer(s, s)).max() < 1e-10


def test_ned_headline_batch(oracle_mod):
    """32 x 32 coarse cells of the unit square, 16 x 16 fine cells each,
    a = exp(g), g iid N(0, 1) per fine cell, B = 1, R = 0, f = (1, 1):
    every cell against the structural invariants, 8 seeded cells against the
    oracle."""
    print("case ned_headline 1024 cells l=4", flush=True)
    from msfec_amd import fields as F
    l, nc = 4, 32
    verts = F.coarse_vertices(2, nc)
    a = F.lognormal_coefficient(2, nc, l, 1.0, 0.0, seed=20261202)
    n, nf = verts.shape[0], 4 ** l
    M, b, P, st = run_gpu(verts, l, a, np.ones((n, nf)), np.zeros((n, nf)), np.ones((n, nf, 2)))
    assert (st == 0).all()
    scale = np.abs(M).max(axis=(1, 2))[:, None, None]
    assert np.abs(M[:, :4, 4:] + M[:, 4:, :4].transpose(0, 2, 1)).max() < 1e-12 * scale.max()
    for blk in (M[:, :4, :4], M[:, 4:, 4:]):
        assert np.abs(blk - blk.transpose(0, 2, 1)).max() / np.abs(blk).max() < 1e-11
        ev = np.linalg.eigvalsh(0.5 * (blk + blk.transpose(0, 2, 1)))
        assert (ev.min(axis=1) > -1e-11 * np.abs(ev).max(axis=1)).all()
    assert np.isfinite(P).all()
    pick = np.random.default_rng(7).choice(n, 8, replace=False)
    o = oracle_mod.ned_basis_batch(l, verts[pick], a[pick], 1.0, 0.0, (1.0, 1.0))
    compare(l, M[pick], b[pick], P[pick], o)


def test_ned_bad_cell_reported(oracle_mod):
    print("case ned_bad_cell", flush=True)
    rng = np.random.default_rng(9)
    l, n = 2, 3
    nf = 4 ** l
    verts = boxes(rng, n)
    a = np.exp(rng.standard_normal((n, nf)))
    a[1, 5] = np.nan
    M, b, P, st = run_gpu(verts, l, a, np.ones((n, nf)), np.zeros((n, nf)), np.ones((n, nf, 2)))
    assert st[0] == 0 and st[2] == 0
    assert st[1] == 3  # MSFEC_CELL_NONFINITE
    o = oracle_mod.ned_basis_batch(l, verts[[0, 2]], a[[0, 2]], 1.0, 0.0, (1.0, 1.0))
    compare(l, M[[0, 2]], b[[0, 2]], P[[0, 2]], o)


@pytest.mark.parametrize("l", [2, 3, 4])
def test_ned_mixed_precision(oracle_mod, l):
    """FP32 LDL^T on the matrix cores + refinement with FP64 residuals (the
    precision recipe of the Q_Ned configuration): the oracle's basis to the
    FP64 tolerances of this file, every cell converged, and the FP64 path's
    result to 1e-10."""
    print(f"case ned_mixed l={l}", flush=True)
    from msfec_amd import NedBasisBatch, _lib as L
    rng = np.random.default_rng(150 + l)
    n = 6
    nf = 4 ** l
    verts = boxes(rng, n)
    a = np.exp(rng.standard_normal((n, nf)))
    binv = np.exp(0.5 * rng.standard_normal((n, nf)))
    r = np.exp(rng.standard_normal((n, nf)))
    f = rng.standard_normal((n, nf, 2))
    out = {}
    for solver in (L.SOLVER_BAND_FP64_MFMA, L.SOLVER_BAND_MIXED):
        with NedBasisBatch(verts, n_refine_local=l, solver=solver) as nb:
            nb.set_coefficients(a, binv, r, f)
            nb.run()
            out[solver] = (nb.get_global_element_matrix(), nb.get_global_element_rhs(),
                           nb.get_basis(), nb.get_cell_status())
    M, b, P, st = out[L.SOLVER_BAND_MIXED]
    assert (st == 0).all(), st
    o = oracle_mod.ned_basis_batch(l, verts, a, binv, r, f)
    compare(l, M, b, P, o)
    M64, b64, P64, _ = out[L.SOLVER_BAND_FP64_MFMA]
    for x, y in ((M, M64), (b, b64), (P, P64)):
        assert np.abs(x - y).max() <= 1e-10 * np.abs(y).max()


def test_ned_mixed_refinement_passes(oracle_mod):
    """The error of the mixed Q_Ned solve after 1, 2, ... refinement passes
    (printed): it falls by the contraction of the FP32 factor each pass, and
    the cells still short of the stop rule after the last pass are flagged."""
    print("case ned_mixed_passes l=4", flush=True)
    from msfec_amd import NedBasisBatch, _lib as L
    rng = np.random.default_rng(160)
    n, l = 6, 4
    nf = 4 ** l
    verts = boxes(rng, n)
    a = np.exp(rng.standard_normal((n, nf)))
    o = oracle_mod.ned_basis_batch(l, verts, a, 1.0, 0.0, (1.0, 1.0))
    ns = (2 ** l + 1) ** 2
    errs = []
    for passes in (1, 2, 3, 4, 5):
        with NedBasisBatch(verts, n_refine_local=l, solver=L.SOLVER_BAND_MIXED,
                           max_refine=passes) as nb:
            nb.set_coefficients(a, 1.0, 0.0, (1.0, 1.0))
            nb.run()
            P, st = nb.get_basis(), nb.get_cell_status()
        es = max(np.abs(P[c, :, :ns] - o["Psi"][c, :, :ns]).max() / np.abs(o["Psi"][c, :, :ns]).max()
                 for c in range(n))
        eu = max(np.abs(P[c, :, ns:] - o["Psi"][c, :, ns:]).max() / np.abs(o["Psi"][c, :, ns:]).max()
                 for c in range(n))
        errs.append((passes, es, eu, int((st != 0).sum())))
    print("passes, sigma error, u error, flagged cells:", errs, flush=True)
    assert errs[-1][1] < TOL_SIGMA and errs[-1][2] < TOL_U and errs[-1][3] == 0
    # every cell flagged after a single pass (its predicted error is the first
    # solution's error squared, far above 1e-12)
    assert errs[0][3] == n
