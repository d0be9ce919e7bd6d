"""An uneven batch: every coarse cell at its own refinement level, each level
on its solver (band where the window holds the band, CG beyond), against the
FP64 oracle's direct solve.  Tolerances as for the CG solver (1e-9 on Psi,
1e-10 on M and b)."""
import numpy as np
import pytest

from test_gpu_parity import rel

pytestmark = pytest.mark.gpu


def test_uneven_batch_levels_and_parity(oracle_mod):
    print("case uneven nc=4 levels 2/3/4", flush=True)
    from msfec_amd import UnevenBatch, fields as F
    nc = 4
    cen = F.inclusion_centres(3, 2, seed=41)
    lv = F.uneven_levels(3, nc, cen, seed=42)
    blocks = F.inclusion_blocks(3, nc, lv, cen, 0.08, 1e4)
    verts = F.coarse_vertices(3, nc)
    with UnevenBatch(verts, lv, dim=3, coef_blocks=blocks, sample=16) as ub:
        ub.run()
        M, b, st = ub.get_global_element_matrix(), ub.get_global_element_rhs(), ub.get_cell_status()
        summ = ub.summary()
        psi = {l: ub.get_basis(l) for l in summ}
    assert (st == 0).all()
    assert summ[4]["solver"] == "cg"
    assert set(summ) == {2, 3, 4}
    for l in summ:
        P, idx = psi[l]
        o = oracle_mod.basis_batch(oracle_mod.make_problem(3, l), verts[idx],
                                   coef=np.stack([blocks[c] for c in idx]),
                                   solver=oracle_mod.SOLVER_DIRECT)
        assert rel(M[idx], o["M"]) < 1e-10, l
        assert rel(b[idx], o["b"]) < 1e-10, l
        assert rel(P, o["Psi"]) < 1e-9, l
