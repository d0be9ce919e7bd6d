"""Parity of the HIP RT_DQ basis (libmsfec_hip.so, msfec_rt_* entry points)
with the FP64 oracle (oracle/rt_oracle.c) on seeded inputs.

The oracle solves the reference's saddle-point system (flux and cell value)
by LU with partial pivoting; the GPU solves the reduced stream-function
problem by a band L D L^T on the matrix cores.  Both are FP64 direct solves of
conditioning ~max(a)/min(a) (lognormal a); the tolerances are TOL_S = 1e-11
for the fluxes (relative to the largest flux of the cell), TOL_M = 1e-11 for
the flux block of the element matrix (relative to its largest entry) and
1e-12 absolute for the divergence entries, which are -+1.
"""
import numpy as np
import pytest

import test_rt_oracle as R

pytestmark = pytest.mark.gpu

TOL_S = 1e-11
TOL_M = 1e-11


def boxes(rng, n):
    v = np.empty((n, 4, 2))
    for c in range(n):
        x0, y0 = rng.uniform(-1, 1, 2)
        hx, hy = rng.uniform(0.02, 0.5, 2)
        v[c] = [[x0, y0], [x0 + hx, y0], [x0, y0 + hy], [x0 + hx, y0 + hy]]
    return v


def run_gpu(verts, l, a, f=None, tensor=False):
    from msfec_amd import RTDQBasisBatch, _lib as L
    mode = L.COEF_TENSOR_QP if tensor else L.COEF_SCALAR_CELL
    with RTDQBasisBatch(verts, n_refine_local=l, coef_mode=mode) as rb:
        rb.set_coefficients(a, f)
        rb.run()
        return (rb.get_global_element_matrix(), rb.get_global_element_rhs(), rb.get_basis(),
                rb.get_cell_status())


def compare(M, b, S, o, cells=None):
    cells = range(M.shape[0]) if cells is None else cells
    for k, c in enumerate(cells):
        Mo, bo, So = o["M"][k], o["b"][k], o["Sigma"][k]
        assert np.abs(S[c] - So).max() <= TOL_S * np.abs(So).max(), f"cell {c} flux"
        blk = np.abs(Mo[:4, :4]).max()
        assert np.abs(M[c, :4, :4] - Mo[:4, :4]).max() <= TOL_M * blk, f"cell {c} M"
        assert np.abs(M[c, :, 4] - Mo[:, 4]).max() <= 1e-12, f"cell {c} M column 4"
        assert np.abs(M[c, 4, :] - Mo[4, :]).max() <= 1e-12, f"cell {c} M row 4"
        np.testing.assert_allclose(b[c], bo, rtol=1e-13, atol=0)


@pytest.mark.parametrize("l", [1, 2, 3, 4])
def test_rt_matches_oracle_random(oracle_mod, l):
    print(f"case rt random l={l}", flush=True)
    rng = np.random.default_rng(300 + l)
    n = 6
    verts = boxes(rng, n)
    nf = (1 << l) ** 2
    a = np.exp(1.5 * rng.normal(size=(n, nf)))
    f = rng.uniform(0.5, 2.0, size=(n, nf))
    M, b, S, st = run_gpu(verts, l, a, f)
    assert np.all(st == 0), st
    o = oracle_mod.rt_basis_batch(l, verts, a, f)
    assert np.all(o["status"] == 0)
    compare(M, b, S, o)


def test_rt_constant_coefficient_closed_form():
    print("case rt constant coefficient l=3", flush=True)
    l = 3
    verts = R.VERTS[None]
    M, b, S, st = run_gpu(verts, l, 2.5, 3.0)
    assert st[0] == 0
    np.testing.assert_allclose(S[0], R.coarse_fluxes(l), rtol=0, atol=1e-14)
    HX, HY, a = R.HX, R.HY, 2.5
    ref = np.array([[HX / (3 * HY), HX / (6 * HY), 0, 0], [HX / (6 * HY), HX / (3 * HY), 0, 0],
                    [0, 0, HY / (3 * HX), HY / (6 * HX)], [0, 0, HY / (6 * HX), HY / (3 * HX)]]) / a
    np.testing.assert_allclose(M[0, :4, :4], ref, rtol=0, atol=1e-14)
    np.testing.assert_allclose(M[0, :4, 4], [1, -1, 1, -1], atol=1e-14)
    np.testing.assert_allclose(M[0, 4, :4], [-1, 1, -1, 1], atol=1e-14)
    np.testing.assert_allclose(b[0], [0, 0, 0, 0, 3.0 * HX * HY], rtol=1e-14)


def test_rt_finest_level_invariants():
    """l = 5 (the largest band the solver takes): flux balance per fine cell,
    boundary fluxes, symmetry and the stream-function solve of the test module."""
    print("case rt l=5 invariants", flush=True)
    l = 5
    m = 1 << l
    rng = np.random.default_rng(55)
    n = 3
    verts = np.broadcast_to(R.VERTS, (n, 4, 2)).copy()
    a = np.exp(1.5 * rng.normal(size=(n, m * m)))
    M, b, S, st = run_gpu(verts, l, a)
    assert np.all(st == 0), st
    nv = m * (m + 1)
    sgn = np.array([-1.0, 1.0, -1.0, 1.0])
    F = R.coarse_fluxes(l)
    bnd = [q * (m + 1) + p for q in range(m) for p in (0, m)] + \
          [nv + q * m + p for q in (0, m) for p in range(m)]
    for c in range(n):
        np.testing.assert_allclose(S[c][:, bnd], F[:, bnd], atol=1e-15)
        Sc = S[c].reshape(4, -1)
        out = np.zeros((4, m, m))
        v = Sc[:, :nv].reshape(4, m, m + 1)
        h = Sc[:, nv:].reshape(4, m + 1, m)
        out = (v[:, :, 1:] - v[:, :, :-1]) + (h[:, 1:, :] - h[:, :-1, :])
        np.testing.assert_allclose(out, np.broadcast_to(sgn[:, None, None] / m ** 2, out.shape),
                                   atol=1e-13)
        np.testing.assert_allclose(M[c, :4, :4], M[c, :4, :4].T, rtol=1e-13, atol=1e-16)
        Sref = R.stream_function_fluxes(l, a[c])
        assert np.abs(S[c] - Sref).max() <= TOL_S * np.abs(Sref).max()
        Mref = R.rt_mass(l, a[c], Sref)
        assert np.abs(M[c, :4, :4] - Mref).max() <= TOL_M * np.abs(Mref).max()


def test_rt_batch_32x32_l4(oracle_mod):
    """The 2-D mixed-basis workload shape: 32 x 32 unit-square cells, l = 4,
    a = exp(g), g iid N(0, 1) per fine cell; 16 cells checked against the oracle."""
    print("case rt batch 32x32 l=4", flush=True)
    nc, l = 32, 4
    m = 1 << l
    H = 1.0 / nc
    verts = np.array([[[i * H, j * H], [(i + 1) * H, j * H], [i * H, (j + 1) * H],
                       [(i + 1) * H, (j + 1) * H]] for j in range(nc) for i in range(nc)])
    rng = np.random.default_rng(20261202)
    a = np.exp(rng.normal(size=(nc * nc, m * m)))
    M, b, S, st = run_gpu(verts, l, a)
    assert np.all(st == 0)
    pick = rng.choice(nc * nc, 16, replace=False)
    o = oracle_mod.rt_basis_batch(l, verts[pick], a[pick])
    compare(M, b, S, o, cells=pick)


def test_rt_reconstruct():
    print("case rt reconstruct l=3", flush=True)
    from msfec_amd import RTDQBasisBatch
    l = 3
    rng = np.random.default_rng(9)
    verts = boxes(rng, 5)
    a = np.exp(rng.normal(size=(5, 64)))
    w = rng.normal(size=(5, 5))
    with RTDQBasisBatch(verts, n_refine_local=l) as rb:
        rb.set_coefficients(a)
        rb.run()
        S = rb.get_basis()
        sig, u = rb.set_global_weights(w)
    np.testing.assert_allclose(sig, np.einsum("ck,cke->ce", w[:, :4], S), rtol=1e-14, atol=1e-16)
    np.testing.assert_array_equal(u, np.repeat(w[:, 4:5], 64, axis=1))


def test_rt_nonfinite_coefficient_reported():
    print("case rt non-finite coefficient", flush=True)
    rng = np.random.default_rng(3)
    verts = boxes(rng, 3)
    a = np.exp(rng.normal(size=(3, 16)))
    a[1, 5] = np.nan
    M, b, S, st = run_gpu(verts, 2, a)
    assert st[0] == 0 and st[2] == 0 and st[1] != 0


@pytest.mark.parametrize("l", [1, 2, 3, 4])
def test_rt_tensor_matches_oracle(oracle_mod, l):
    """A rotated anisotropic tensor at every Gauss point (lognormal eigenvalues)."""
    print(f"case rt tensor l={l}", flush=True)
    rng = np.random.default_rng(400 + l)
    n = 5
    verts = boxes(rng, n)
    t = np.stack([R.random_tensors(l, 500 + 10 * l + c) for c in range(n)])
    f = rng.uniform(0.5, 2.0, size=(n, (1 << l) ** 2))
    M, b, S, st = run_gpu(verts, l, t, f, tensor=True)
    assert np.all(st == 0), st
    o = oracle_mod.rt_basis_batch(l, verts, t, f, tensor=True)
    assert np.all(o["status"] == 0)
    compare(M, b, S, o)


def test_rt_tensor_isotropic_equals_scalar():
    print("case rt tensor isotropic l=3", flush=True)
    l = 3
    rng = np.random.default_rng(77)
    verts = boxes(rng, 4)
    a = np.exp(1.5 * rng.normal(size=(4, 64)))
    t = np.zeros((4, 64, 4, 3))
    t[..., 0] = a[..., None]
    t[..., 2] = a[..., None]
    Ms, bs, Ss, sts = run_gpu(verts, l, a)
    Mt, bt, St, stt = run_gpu(verts, l, t, tensor=True)
    assert np.all(sts == 0) and np.all(stt == 0)
    np.testing.assert_allclose(St, Ss, rtol=0, atol=1e-13 * np.abs(Ss).max())
    np.testing.assert_allclose(Mt, Ms, rtol=1e-12, atol=1e-14)
    np.testing.assert_array_equal(bt, bs)
