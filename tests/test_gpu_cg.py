"""Parity of the matrix-free CG solver (MSFEC_SOLVER_CG) with the FP64
oracle's direct solve, and the uneven batch that routes each refinement
level to its solver.

CG stops at ||r|| <= 1e-13 ||b||.  With the Jacobi-scaled condition numbers
of these cells (up to ~1e4 with the 1e4 inclusions) that bounds the solution
error by ~1e-9 relative; measured, it is at most 3.3e-12 on every case here
(printed by each test), so Psi is held to the metric's TOL_PSI = 1e-10.  The
element matrix, formed with the energy form of the Schur term, is second
order in it: TOL_M = 1e-10 (relative, max over cells of the norm ratio).
"""
import numpy as np
import pytest

from test_gpu_parity import random_boxes, rel, run_gpu, run_oracle

pytestmark = pytest.mark.gpu

CG = 3
TOL_PSI = 1e-10
TOL_M = 1e-10


@pytest.mark.parametrize("dim,l", [(2, 2), (2, 4), (2, 6), (3, 1), (3, 2), (3, 3)])
def test_cg_lognormal(oracle_mod, dim, l):
    print(f"case cg_lognormal dim={dim} l={l}", flush=True)
    rng = np.random.default_rng(300 + 10 * dim + l)
    n_cells = 5
    verts = random_boxes(rng, dim, n_cells)
    coef = np.exp(1.5 * rng.standard_normal((n_cells, (1 << l) ** dim)))
    g = run_gpu(verts, dim, l, coef, solver=CG)
    o = run_oracle(oracle_mod, verts, dim, l, coef)
    print("errors", {k: rel(g[k], o[k]) for k in ("M", "b", "Psi")}, flush=True)
    assert (g["status"] == 0).all(), g["status"]
    assert rel(g["M"], o["M"]) < TOL_M
    assert rel(g["b"], o["b"]) < TOL_M
    assert rel(g["Psi"], o["Psi"]) < TOL_PSI


def test_cg_reaction_and_cell_load(oracle_mod):
    print("case cg_reaction dim=3 l=2", flush=True)
    rng = np.random.default_rng(17)
    dim, l, n_cells = 3, 2, 4
    verts = random_boxes(rng, dim, n_cells)
    nf = (1 << l) ** dim
    coef = np.exp(rng.standard_normal((n_cells, nf)))
    f = rng.standard_normal((n_cells, nf))
    reac = np.exp(rng.standard_normal((n_cells, nf)))
    kw = dict(f=f, rhs_mode=1, reac=reac, reaction_mode=2)
    g = run_gpu(verts, dim, l, coef, solver=CG, **kw)
    o = run_oracle(oracle_mod, verts, dim, l, coef, **kw)
    print("errors", {k: rel(g[k], o[k]) for k in ("M", "b", "Psi")}, flush=True)
    assert (g["status"] == 0).all()
    assert rel(g["M"], o["M"]) < TOL_M
    assert rel(g["b"], o["b"]) < TOL_M
    assert rel(g["Psi"], o["Psi"]) < TOL_PSI


def test_cg_inclusions_l4(oracle_mod):
    """3-D, 16^3 fine cells per coarse cell (3375 interior unknowns, beyond
    the band window), the two cells of the seeded inclusion field with the
    largest 1e4-inclusion volume."""
    print("case cg_inclusions dim=3 l=4", flush=True)
    from msfec_amd import fields as F
    coef_all = F.inclusion_coefficient(3, 16, 4, 410, 0.02, 1e4, seed=20261202)
    verts_all = F.coarse_vertices(3, 16)
    pick = np.argsort(-(coef_all > 10).mean(axis=1))[:2]
    g = run_gpu(verts_all[pick], 3, 4, coef_all[pick], solver=CG)
    o = run_oracle(oracle_mod, verts_all[pick], 3, 4, coef_all[pick])
    print("errors", {k: rel(g[k], o[k]) for k in ("M", "b", "Psi")}, flush=True)
    assert (g["status"] == 0).all()
    assert rel(g["M"], o["M"]) < TOL_M
    assert rel(g["b"], o["b"]) < TOL_M
    assert rel(g["Psi"], o["Psi"]) < TOL_PSI


def test_cg_matches_band_solver():
    """Same cells through CG and the FP64 band solver (matrix cores)."""
    print("case cg_vs_band dim=3 l=3", flush=True)
    rng = np.random.default_rng(23)
    verts = random_boxes(rng, 3, 16)
    coef = np.exp(1.5 * rng.standard_normal((16, 512)))
    g = run_gpu(verts, 3, 3, coef, solver=CG)
    h = run_gpu(verts, 3, 3, coef, solver=1)
    assert rel(g["M"], h["M"]) < TOL_M
    assert rel(g["Psi"], h["Psi"]) < TOL_PSI


def test_cg_bad_cells_and_iteration_cap():
    print("case cg_bad_cells", flush=True)
    from msfec_amd import QBasisBatch, _lib as L
    rng = np.random.default_rng(29)
    verts = random_boxes(rng, 3, 3)
    coef = np.exp(rng.standard_normal((3, 64)))
    coef[1, :] = -1.0
    coef[2, 7] = np.nan
    with QBasisBatch(verts, dim=3, n_refine_local=2, solver=CG) as qb:
        qb.set_coefficient(coef)
        qb.run()
        st = qb.get_cell_status()
        it = qb.cg_iterations()
    assert st[0] == L.CELL_OK and st[1] == L.CELL_NOT_SPD
    assert st[2] in (L.CELL_NOT_SPD, L.CELL_NONFINITE)
    assert (it[0] > 0).all()
    # two iterations cannot reach 1e-13
    coef[1:] = 1.0 + rng.random((2, 64))
    with QBasisBatch(verts, dim=3, n_refine_local=2, solver=CG, cg_max_iter=2) as qb:
        qb.set_coefficient(coef)
        qb.run()
        st = qb.get_cell_status()
        it = qb.cg_iterations()
    assert (st == L.CELL_NOT_CONVERGED).all()
    assert (it == 2).all()


def random_spd_tensors(rng, n_cells, nf, nq, dim):
    """Random SPD tensors per Gauss point, B B^T + 0.5 I, as the symmetric
    3- or 6-vectors of MSFEC_COEF_TENSOR_QP."""
    B = rng.standard_normal((n_cells, nf, nq, dim, dim))
    A = np.einsum("...ij,...kj->...ik", B, B) + 0.5 * np.eye(dim)
    if dim == 2:
        return np.stack([A[..., 0, 0], A[..., 0, 1], A[..., 1, 1]], -1)
    return np.stack([A[..., 0, 0], A[..., 0, 1], A[..., 0, 2], A[..., 1, 1], A[..., 1, 2],
                     A[..., 2, 2]], -1)


@pytest.mark.parametrize("dim,l", [(2, 3), (2, 5), (3, 2), (3, 3)])
def test_cg_tensor_qp_and_qp_load(oracle_mod, dim, l):
    """CG on the assembled stencil: anisotropic tensors and loads at the
    Gauss points, with and without a reaction (all bases solved, or the last
    one from the partition of unity), against the oracle's direct solve and
    the band solver of the same assembled system."""
    print(f"case cg_tensor dim={dim} l={l}", flush=True)
    rng = np.random.default_rng(400 + 10 * dim + l)
    n_cells, N, nq = 4, 1 << l, 1 << dim
    verts = random_boxes(rng, dim, n_cells)
    T = random_spd_tensors(rng, n_cells, N ** dim, nq, dim)
    f = rng.standard_normal((n_cells, N ** dim, nq))
    for reaction in (0.0, 2.0):
        kw = dict(coef_mode=1, f=f, rhs_mode=2, reaction_mode=1 if reaction else 0,
                  reaction_const=reaction)
        g = run_gpu(verts, dim, l, T, solver=CG, **kw)
        o = run_oracle(oracle_mod, verts, dim, l, T, **kw)
        print("reaction", reaction, "errors", {k: rel(g[k], o[k]) for k in ("M", "b", "Psi")},
              flush=True)
        assert (g["status"] == 0).all(), g["status"]
        assert rel(g["M"], o["M"]) < TOL_M, reaction
        assert rel(g["b"], o["b"]) < TOL_M, reaction
        assert rel(g["Psi"], o["Psi"]) < TOL_PSI, reaction
        if l <= 3:
            gb = run_gpu(verts, dim, l, T, solver=0, **kw)
            assert rel(g["M"], gb["M"]) < TOL_M


def test_cg_reference_example_l4(oracle_mod):
    """The reference's example (example_parameters/prm_q_test-01.prm): coarse
    mesh of 4^3 cells (global refinements 2), local refinements 4 (16^3 fine
    cells, 3375 interior unknowns), the rotated oscillatory Diffusion_A
    (frequencies 17, 16, 15, alpha 0.9, scale 1, rotate) and f = sin(2 pi x)
    sin(2 pi y) sin(2 pi z), both at the QGauss(2) points.  Two cells against
    the oracle, which evaluates Diffusion_A inline (q_basis.cc:196-220)."""
    print("case cg_reference_example dim=3 l=4", flush=True)
    from msfec_amd import fields as F
    dim, l = 3, 4
    verts = F.coarse_vertices(dim, 4)[[5, 42]]
    pts = F.qp_coordinates(verts, l)
    T = F.diffusion_A_tensor(pts, freq=(17, 16, 15), scale=(1.0, 1.0, 1.0),
                             alpha=(0.9, 0.9, 0.9), rotate=True)
    f = np.prod(np.sin(2 * np.pi * pts), axis=-1)
    g = run_gpu(verts, dim, l, T, coef_mode=1, f=f, rhs_mode=2, solver=CG)
    o = oracle_mod.basis_batch(
        oracle_mod.make_problem(dim, l, coef_mode=oracle_mod.COEF_ANALYTIC, rhs_mode=2,
                                freq=(17, 16, 15), scale=(1.0, 1.0, 1.0), alpha=(0.9, 0.9, 0.9),
                                rotate=True),
        verts, f=f, solver=oracle_mod.SOLVER_DIRECT, n_threads=2)
    assert (g["status"] == 0).all(), g["status"]
    e = {k: rel(g[k], o[k]) for k in ("M", "b", "Psi")}
    print("reference example errors", e, flush=True)
    assert e["M"] < TOL_M and e["b"] < TOL_M and e["Psi"] < TOL_PSI, e
