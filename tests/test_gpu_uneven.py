"""Uneven batches through msfec_qu_* (one handle, every coarse cell at its own
refinement level, each level on the solver chosen in C++: band where the
window holds the band, CG beyond), against the FP64 oracle's direct solve.

Tolerances: M and b 1e-10 relative (the energy form of the Schur term is
second order in the solve error); Psi 1e-10 on every level.  On the CG
level Psi carries the CG stopping error (relative residual 1e-13 times the
condition of the Jacobi-scaled system, which the 1e4 inclusions raise):
measured 0.9e-12 to 5.2e-12 on these cells (printed by the tests)."""
import ctypes

import numpy as np
import pytest

from test_gpu_parity import rel

pytestmark = pytest.mark.gpu

TOL = 1e-10
TOL_PSI_CG = 1e-10


def psi_blocks(ub, P, idx, level):
    nn = ((1 << level) + 1) ** ub.dim
    return np.stack([P[int(ub.basis_offsets[c]):int(ub.basis_offsets[c]) + ub.nb * nn]
                     for c in idx]).reshape(len(idx), ub.nb, nn)


def oracle_check(oracle_mod, ub, verts, blocks, M, b, P, samples):
    errs = {}
    for lv, idx in samples.items():
        o = oracle_mod.basis_batch(oracle_mod.make_problem(3, lv), verts[idx],
                                   coef=np.stack([blocks[c] for c in idx]),
                                   solver=oracle_mod.SOLVER_DIRECT)
        e = {"M": rel(M[idx], o["M"]), "b": rel(b[idx], o["b"]),
             "Psi": rel(psi_blocks(ub, P, idx, lv), o["Psi"])}
        print(f"level {lv} ({len(idx)} cells): {e}", flush=True)
        assert e["M"] < TOL and e["b"] < TOL, (lv, e)
        assert e["Psi"] < (TOL_PSI_CG if lv == 4 else TOL), (lv, e)
        errs[lv] = e
    return errs


def test_uneven_batch_levels_and_parity(oracle_mod):
    print("case uneven nc=4 levels 2/3/4", flush=True)
    from msfec_amd import UnevenBatch, fields as F
    nc = 4
    cen = F.inclusion_centres(3, 2, seed=41)
    lv = F.uneven_levels(3, nc, cen, seed=42)
    blocks = F.inclusion_blocks(3, nc, lv, cen, 0.08, 1e4)
    verts = F.coarse_vertices(3, nc)
    with UnevenBatch(verts, lv, dim=3, coef_blocks=blocks) as ub:
        ub.run()
        M, b, st = ub.get_global_element_matrix(), ub.get_global_element_rhs(), ub.get_cell_status()
        P = ub.get_basis_flat()
        summ = ub.summary()
    assert (st == 0).all()
    assert summ[4]["solver"] == "cg"
    assert set(summ) == {2, 3, 4}
    for l in (2, 3):  # the band levels were timed against CG, with every window that fits
        assert summ[l]["tuned_ms"]["band_mfma_w64"] is not None
        assert summ[l]["tuned_ms"]["cg"] is not None
    assert summ[2]["tuned_ms"]["band_mfma_w32"] is not None  # 3-D l = 2: half-bandwidth 13
    assert summ[2]["tuned_ms"]["band_mfma_w16"] is None      # 13 + 4 > 16
    assert summ[3]["tuned_ms"]["band_mfma_w32"] is None      # 3-D l = 3: 57
    samples = {l: np.nonzero(lv == l)[0] for l in (2, 3, 4)}
    oracle_check(oracle_mod, ub, verts, blocks, M, b, P, samples)


def test_uneven_full_configuration(oracle_mod):
    """The uneven configuration at its size: 16^3 cells, 410 seeded inclusions
    (A = 1e4, radius 0.02) refined 4 levels, the others 2 or 3 by a seeded
    coin.  Invariants on all 4096 cells; the oracle on 4 cells of levels 2
    and 3 and one of level 4."""
    print("case uneven full 16^3, 410 inclusions", flush=True)
    from msfec_amd import UnevenBatch, fields as F
    nc, seed = 16, 20261202
    cen = F.inclusion_centres(3, 410, seed)
    lv = F.uneven_levels(3, nc, cen, seed + 1)
    blocks = F.inclusion_blocks(3, nc, lv, cen, 0.02, 1e4)
    verts = F.coarse_vertices(3, nc)
    with UnevenBatch(verts, lv, dim=3, coef_blocks=blocks) as ub:
        ub.run()
        M, b, st = ub.get_global_element_matrix(), ub.get_global_element_rhs(), ub.get_cell_status()
        P = ub.get_basis_flat()
        it = ub.level(4).cg_iterations()
        summ = ub.summary()
        offs = ub.basis_offsets.copy()
    print(summ, "CG iterations mean", it.mean(), "max", it.max(), flush=True)
    assert (st == 0).all()
    assert summ[4]["solver"] == "cg" and summ[4]["cells"] == (lv == 4).sum()
    scale = np.abs(M).max(axis=(1, 2))
    assert np.max(np.abs(M - M.transpose(0, 2, 1)).max(axis=(1, 2)) / scale) < 1e-12
    assert np.max(np.abs(M.sum(axis=2)).max(axis=1) / scale) < TOL
    vol = (1.0 / nc) ** 3
    assert np.abs(b.sum(axis=1) - vol).max() / vol < TOL
    # partition of unity of every cell's bases
    worst = 0.0
    for c in range(nc ** 3):
        nn = ((1 << int(lv[c])) + 1) ** 3
        blk = P[int(offs[c]):int(offs[c]) + 8 * nn].reshape(8, nn)
        worst = max(worst, float(np.abs(blk.sum(axis=0) - 1.0).max()))
    assert worst < TOL, worst
    rng = np.random.default_rng(5)
    samples = {2: rng.choice(np.nonzero(lv == 2)[0], 4, replace=False),
               3: rng.choice(np.nonzero(lv == 3)[0], 4, replace=False),
               4: rng.choice(np.nonzero(lv == 4)[0], 1, replace=False)}
    oracle_check(oracle_mod, ub, verts, blocks, M, b, P, samples)


def test_uneven_fixed_solver_and_coarse_solve(oracle_mod):
    """A fixed solver (no timing: the matrix-core band where it fits, CG
    beyond) gives the AUTO handle's basis to rounding, and the coarse solve of
    the uneven batch equals the oracle's coarse solve of the same M and b."""
    print("case uneven fixed solver + coarse solve", flush=True)
    from msfec_amd import UnevenBatch, _lib as L, fields as F
    nc = 4
    cen = F.inclusion_centres(3, 3, seed=51)
    lv = F.uneven_levels(3, nc, cen, seed=52)
    blocks = F.inclusion_blocks(3, nc, lv, cen, 0.06, 1e3)
    verts = F.coarse_vertices(3, nc)
    with UnevenBatch(verts, lv, dim=3, coef_blocks=blocks, solver=L.SOLVER_BAND_FP64_MFMA) as ub:
        ub.run()
        M1, b1 = ub.get_global_element_matrix(), ub.get_global_element_rhs()
        summ = ub.summary()
        u, w, it, res = ub.coarse_solve(nc, abs_tol=0.0, rel_tol=1e-13)
    assert summ[2]["solver"] == "band-fp64-mfma" and summ[4]["solver"] == "cg"
    assert summ[2]["tuned_ms"]["band_mfma_w32"] is None  # not timed
    assert summ[2]["window"] == 32  # the smallest window holding the band
    with UnevenBatch(verts, lv, dim=3, coef_blocks=blocks) as ua:
        ua.run()
        M2, b2 = ua.get_global_element_matrix(), ua.get_global_element_rhs()
    assert rel(M1, M2) < TOL and rel(b1, b2) < TOL
    uo, _ = oracle_mod.coarse_solve(3, nc, M1, b1, rtol=1e-14)
    assert it > 0 and np.abs(u - uo).max() <= 1e-10 * np.abs(uo).max()
    assert w.shape == (nc ** 3, 8)


def test_uneven_device_buffers_match_host():
    """Coefficients from device memory and results into device memory equal
    the host-pointer path bit for bit (the gather and scatter kernels)."""
    print("case uneven device buffers", flush=True)
    from msfec_amd import UnevenBatch, _lib as L, fields as F
    from test_gpu_streams import DevBuf, Hip
    hip = Hip()
    nc = 3
    cen = F.inclusion_centres(3, 2, seed=61)
    lv = F.uneven_levels(3, nc, cen, seed=62, coarse=(1, 2, 3))
    blocks = F.inclusion_blocks(3, nc, lv, cen, 0.1, 50.0)
    verts = F.coarse_vertices(3, nc)
    flat = np.concatenate([np.ravel(x) for x in blocks])
    s = ctypes.c_void_p(0)
    with UnevenBatch(verts, lv, dim=3, coef_blocks=blocks, solver=L.SOLVER_BAND_FP64_MFMA) as ub:
        ub.run()
        M, b, P = ub.get_global_element_matrix(), ub.get_global_element_rhs(), ub.get_basis_flat()
        dc = DevBuf(hip, flat, s)
        dM, db = DevBuf(hip, np.zeros_like(M), s), DevBuf(hip, np.zeros_like(b), s)
        dP = DevBuf(hip, np.zeros_like(P), s)
        try:
            L.check(ub._lib.msfec_qu_set_coefficient(ub._h, dc.ptr, flat.size, L.DEVICE))
            ub.run()
            L.check(ub._lib.msfec_qu_get_element(ub._h, dM.ptr, db.ptr, L.DEVICE))
            L.check(ub._lib.msfec_qu_get_basis(ub._h, dP.ptr, L.DEVICE))
            got = []
            for buf, ref in ((dM, M), (db, b), (dP, P)):
                out = np.empty_like(ref)
                hip.ok(hip.L.hipMemcpyAsync(ctypes.c_void_p(out.ctypes.data), buf.ptr, out.nbytes, 2, s),
                       "hipMemcpyAsync")
                hip.ok(hip.L.hipStreamSynchronize(s), "hipStreamSynchronize")
                got.append(out)
        finally:
            for x in (dc, dM, db, dP):
                x.free()
    for x, y, name in zip(got, (M, b, P), ("M", "b", "Psi")):
        np.testing.assert_array_equal(x, y, err_msg=name)
