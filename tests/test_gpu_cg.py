"""Parity of the matrix-free CG solver (MSFEC_SOLVER_CG) with the FP64
oracle's direct solve, and the uneven batch that routes each refinement
level to its solver.

CG stops at ||r|| <= 1e-13 ||b||.  With the Jacobi-scaled condition numbers
of these cells (up to ~1e4 with the 1e4 inclusions) the solution error is at
most ~1e-9 relative, and the element matrix, formed with the energy form of
the Schur term, is second order in it: TOL_PSI = 1e-9, TOL_M = 1e-10
(relative, max over cells of the norm ratio).
"""
import numpy as np
import pytest

from test_gpu_parity import random_boxes, rel, run_gpu, run_oracle

pytestmark = pytest.mark.gpu

CG = 3
TOL_PSI = 1e-9
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
