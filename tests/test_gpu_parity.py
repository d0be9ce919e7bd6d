"""Parity of the HIP basis construction (libmsfec_hip.so through the C-ABI)
with the FP64 CPU oracle, on seeded inputs.

Tolerances: the FP64 solvers and the oracle compute in FP64 with direct
solves, so M, b and Psi must agree to a few hundred ulps of the condition
number: TOL = 1e-11 relative (max over cells of ||X_gpu - X_oracle|| /
||X_oracle||).  The mixed solver (FP32 factor, FP64-residual refinement) stops
refining a cell once the predicted error of its solution is below 1e-11;
its element matrix, in energy form, carries about the square of the first
FP32 solution's error times the cancellation in P - S (up to ~3e-11 on the
4096-cell lognormal batch), so its bound is the acceptance bound of the basis
metric, TOL_MIXED = 1e-10.
"""
import numpy as np
import pytest

from conftest import box_vertices

pytestmark = pytest.mark.gpu

TOL = 1e-11
TOL_MIXED = 1e-10


def tol(solver):
    return TOL_MIXED if solver == 2 else TOL


def _lib():
    from msfec_amd import _lib as L
    return L


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    num = np.linalg.norm((a - b).reshape(a.shape[0], -1), axis=1)
    den = np.linalg.norm(b.reshape(b.shape[0], -1), axis=1)
    return float(np.max(num / np.maximum(den, 1e-300)))


# MSFEC_SOLVER_BAND_FP64 (vector ALU), MSFEC_SOLVER_BAND_FP64_MFMA, MSFEC_SOLVER_BAND_MIXED
SOLVERS = (0, 1, 2)
# solvers for tensor coefficients / quadrature-point loads (the mixed solver
# takes a scalar coefficient per fine cell only)
SOLVERS_GENERAL = (0, 1)


def run_gpu(verts, dim, l, coef, coef_mode=0, f=None, rhs_mode=0, rhs_const=1.0,
            reac=None, reaction_mode=0, reaction_const=0.0, solver=1):
    from msfec_amd import QBasisBatch
    with QBasisBatch(verts, dim=dim, n_refine_local=l, coef_mode=coef_mode, rhs_mode=rhs_mode,
                     rhs_const=rhs_const, reaction_mode=reaction_mode,
                     reaction_const=reaction_const, solver=solver) as qb:
        qb.set_coefficient(coef)
        if f is not None:
            qb.set_rhs(f)
        if reac is not None:
            qb.set_reaction(reac)
        qb.run()
        return dict(M=qb.get_global_element_matrix(), b=qb.get_global_element_rhs(),
                    Psi=qb.get_basis(), status=qb.get_cell_status())


def run_oracle(O, verts, dim, l, coef, coef_mode=0, f=None, rhs_mode=0, rhs_const=1.0,
               reac=None, reaction_mode=0, reaction_const=0.0, **kw):
    p = O.make_problem(dim, l, coef_mode=coef_mode, rhs_mode=rhs_mode, rhs_const=rhs_const,
                       reaction_mode=reaction_mode, reaction_const=reaction_const, **kw)
    return O.basis_batch(p, verts, coef=coef, f=f, reaction=reac, solver=O.SOLVER_DIRECT)


def random_boxes(rng, dim, n_cells):
    out = []
    for _ in range(n_cells):
        lo = rng.random(dim)
        hi = lo + 0.02 + 0.1 * rng.random(dim)
        out.append(box_vertices(dim, lo, hi))
    return np.array(out)


@pytest.mark.parametrize("solver", SOLVERS)
@pytest.mark.parametrize("dim,l", [(2, 1), (2, 2), (2, 3), (2, 4), (2, 5), (3, 1), (3, 2), (3, 3)])
def test_scalar_lognormal(oracle_mod, dim, l, solver):
    print(f"case scalar_lognormal dim={dim} l={l} solver={solver}", flush=True)
    rng = np.random.default_rng(100 + 10 * dim + l)
    n_cells = 7
    verts = random_boxes(rng, dim, n_cells)
    N = 1 << l
    coef = np.exp(1.5 * rng.standard_normal((n_cells, N ** dim)))
    g = run_gpu(verts, dim, l, coef, solver=solver)
    o = run_oracle(oracle_mod, verts, dim, l, coef)
    assert (g["status"] == 0).all() and (o["status"] == 0).all()
    assert rel(g["M"], o["M"]) < tol(solver)
    assert rel(g["b"], o["b"]) < tol(solver)
    assert rel(g["Psi"], o["Psi"]) < tol(solver)


@pytest.mark.parametrize("solver", SOLVERS_GENERAL)
@pytest.mark.parametrize("dim,l", [(2, 3), (3, 2)])
def test_tensor_qp_and_rhs_modes(oracle_mod, dim, l, solver):
    print(f"case tensor_qp dim={dim} l={l} solver={solver}", flush=True)
    rng = np.random.default_rng(7 + dim)
    n_cells, N, nq = 5, 1 << l, 1 << dim
    verts = random_boxes(rng, dim, n_cells)
    ncomp = 3 if dim == 2 else 6
    # random SPD tensors per quadrature point: B B^T + 0.5 I
    B = rng.standard_normal((n_cells, N ** dim, nq, dim, dim))
    A = np.einsum("...ij,...kj->...ik", B, B) + 0.5 * np.eye(dim)
    if dim == 2:
        T = np.stack([A[..., 0, 0], A[..., 0, 1], A[..., 1, 1]], -1)
    else:
        T = np.stack([A[..., 0, 0], A[..., 0, 1], A[..., 0, 2], A[..., 1, 1], A[..., 1, 2],
                      A[..., 2, 2]], -1)
    assert T.shape[-1] == ncomp
    for rhs_mode in (1, 2):
        f = rng.standard_normal((n_cells, N ** dim) if rhs_mode == 1 else (n_cells, N ** dim, nq))
        g = run_gpu(verts, dim, l, T, coef_mode=1, f=f, rhs_mode=rhs_mode, solver=solver)
        o = run_oracle(oracle_mod, verts, dim, l, T, coef_mode=1, f=f, rhs_mode=rhs_mode)
        assert rel(g["M"], o["M"]) < tol(solver)
        assert rel(g["b"], o["b"]) < tol(solver)
        assert rel(g["Psi"], o["Psi"]) < tol(solver)


def test_analytic_diffusion_A(oracle_mod):
    """The reference's rotated oscillatory tensor (eqn_coeff_A.cc), evaluated
    at the quadrature points on the host, vs the oracle evaluating it inline."""
    print("case analytic_diffusion_A", flush=True)
    from msfec_amd import fields as F
    dim, l, nc = 3, 2, 4
    verts = F.coarse_vertices(dim, nc)[:16]
    pts = F.qp_coordinates(verts, l)
    T = F.diffusion_A_tensor(pts)
    g = run_gpu(verts, dim, l, T, coef_mode=1)
    o = run_oracle(oracle_mod, verts, dim, l, None, coef_mode=2, freq=(17, 16, 15),
                   alpha=(0.9, 0.9, 0.9), rotate=True)
    assert rel(g["M"], o["M"]) < 1e-10
    assert rel(g["Psi"], o["Psi"]) < 1e-10


@pytest.mark.parametrize("solver", SOLVERS)
@pytest.mark.parametrize("mode", [1, 2])
def test_reaction(oracle_mod, mode, solver):
    print(f"case reaction mode={mode} solver={solver}", flush=True)
    rng = np.random.default_rng(31 + mode)
    dim, l, n_cells = 3, 2, 4
    N = 1 << l
    verts = random_boxes(rng, dim, n_cells)
    coef = np.exp(rng.standard_normal((n_cells, N ** dim)))
    reac = rng.random((n_cells, N ** dim)) if mode == 2 else None
    g = run_gpu(verts, dim, l, coef, reac=reac, reaction_mode=mode, reaction_const=3.0,
                solver=solver)
    o = run_oracle(oracle_mod, verts, dim, l, coef, reac=reac, reaction_mode=mode,
                   reaction_const=3.0)
    assert rel(g["M"], o["M"]) < tol(solver)
    assert rel(g["Psi"], o["Psi"]) < tol(solver)


@pytest.mark.parametrize("solver", SOLVERS)
def test_bad_cells_reported_not_fatal(solver):
    """A cell with a non-positive coefficient is flagged; the others stay valid."""
    print(f"case bad_cells solver={solver}", flush=True)
    rng = np.random.default_rng(5)
    dim, l, n_cells = 3, 2, 3
    N = 1 << l
    verts = random_boxes(rng, dim, n_cells)
    coef = np.exp(rng.standard_normal((n_cells, N ** dim)))
    coef[1, :] = -1.0
    coef[2, 3] = np.nan
    L = _lib()
    g = run_gpu(verts, dim, l, coef, solver=solver)
    assert g["status"][0] == L.CELL_OK
    assert g["status"][1] == L.CELL_NOT_SPD
    assert g["status"][2] in (L.CELL_NOT_SPD, L.CELL_NONFINITE)
    assert np.isfinite(g["M"][0]).all()


@pytest.mark.parametrize("solver", SOLVERS)
def test_headline_batch_invariants(oracle_mod, solver):
    """4096 cells, 3-D, 8^3 fine cells each, lognormal sigma=1.5 field with
    correlation length 1/64: structural invariants on every cell and oracle
    parity on a seeded sample."""
    print(f"case headline_batch solver={solver}", flush=True)
    from msfec_amd import fields as F
    dim, l, nc = 3, 3, 16
    verts = F.coarse_vertices(dim, nc)
    coef = F.lognormal_coefficient(dim, nc, l, 1.5, 1.0 / 64, seed=20261202)
    g = run_gpu(verts, dim, l, coef, solver=solver)
    assert (g["status"] == 0).all()
    M, b, Psi = g["M"], g["b"], g["Psi"]
    scale = np.abs(M).max(axis=(1, 2))
    assert np.max(np.abs(M - M.transpose(0, 2, 1)).max(axis=(1, 2)) / scale) < 1e-12
    # constants are A-harmonic: M 1 = 0 and sum_i psi_i = 1
    assert np.max(np.abs(M.sum(axis=2)).max(axis=1) / scale) < tol(solver)
    assert np.abs(Psi.sum(axis=1) - 1.0).max() < tol(solver)
    # sum_i b_i = int_K f = |K| for f = 1
    vol = (1.0 / nc) ** 3
    assert np.abs(b.sum(axis=1) - vol).max() / vol < tol(solver)
    # diagonal positive, eigenvalues >= 0
    ev = np.linalg.eigvalsh(M)
    assert (ev[:, 0] > -1e-12 * scale).all()
    sample = np.random.default_rng(1).choice(verts.shape[0], 24, replace=False)
    o = run_oracle(oracle_mod, verts[sample], dim, l, coef[sample])
    assert rel(M[sample], o["M"]) < tol(solver)
    assert rel(b[sample], o["b"]) < tol(solver)
    assert rel(Psi[sample], o["Psi"]) < tol(solver)


def test_batch_beyond_32bit_indices(oracle_mod):
    """64^3 = 262 144 cells at l = 3 (~85 GB of device buffers): the panel
    store holds 5.9e9 doubles and the row table 2.2e9, beyond 32-bit element
    indices.  a = exp(g), g iid N(0, 1) per fine cell.  Status, symmetry,
    M 1 = 0 and sum b = |K| on every cell; the last cells and one in the
    middle against the oracle."""
    print("case batch 64^3 cells l=3", flush=True)
    from msfec_amd import QBasisBatch, fields as F
    dim, l, nc = 3, 3, 64
    verts = F.coarse_vertices(dim, nc)
    n = verts.shape[0]
    coef = np.exp(np.random.default_rng(64).standard_normal((n, 512)))
    with QBasisBatch(verts, dim=dim, n_refine_local=l, solver=1) as qb:
        qb.set_coefficient(coef)
        qb.run()
        M, b, st = qb.get_global_element_matrix(), qb.get_global_element_rhs(), qb.get_cell_status()
    assert (st == 0).all()
    scale = np.abs(M).max(axis=(1, 2))
    assert np.max(np.abs(M - M.transpose(0, 2, 1)).max(axis=(1, 2)) / scale) < 1e-12
    assert np.max(np.abs(M.sum(axis=2)).max(axis=1) / scale) < TOL
    vol = (1.0 / nc) ** 3
    assert np.abs(b.sum(axis=1) - vol).max() / vol < TOL
    pick = np.array([n // 2, n - 2, n - 1])
    o = run_oracle(oracle_mod, verts[pick], dim, l, coef[pick])
    assert rel(M[pick], o["M"]) < TOL
    assert rel(b[pick], o["b"]) < TOL


@pytest.mark.parametrize("reaction", [0.0, 1.0])
@pytest.mark.parametrize("dim,nc,l", [(2, 8, 4), (3, 4, 3)])
def test_coarse_solve_and_reconstruct(oracle_mod, dim, nc, l, reaction):
    """The whole method on a structured coarse grid: 2-D 8 x 8 cells refined 4
    levels, a = exp(g) with g iid N(0, 1) (the plain-method configuration,
    with the reaction 1 of -div(A grad u) + u = f and with the reference's
    reaction 0), and 3-D 4^3 cells refined 3 levels.  M, b and Psi of every
    cell against the oracle's direct solve; the coarse solve against the
    oracle's coarse solve of the same M and b, and against that of the
    oracle's own M and b; the reconstruction."""
    print(f"case coarse dim={dim} nc={nc} l={l} reaction={reaction}", flush=True)
    from msfec_amd import QBasisBatch, _lib as L, fields as F
    verts = F.coarse_vertices(dim, nc)
    coef = F.lognormal_coefficient(dim, nc, l, 1.0, 0.0, seed=3 if dim == 3 else 20261202)
    rmode = L.REACTION_CONST if reaction else L.REACTION_NONE
    with QBasisBatch(verts, dim=dim, n_refine_local=l, reaction_mode=rmode,
                     reaction_const=reaction) as qb:
        qb.set_coefficient(coef)
        qb.run()
        M, b, Psi = qb.get_global_element_matrix(), qb.get_global_element_rhs(), qb.get_basis()
        u, w, it, res = qb.coarse_solve(nc, abs_tol=0.0, rel_tol=1e-13)
        uf = qb.set_global_weights(w)
    o = run_oracle(oracle_mod, verts, dim, l, coef, reaction_mode=rmode, reaction_const=reaction)
    assert rel(M, o["M"]) < TOL
    assert rel(b, o["b"]) < TOL
    assert rel(Psi, o["Psi"]) < TOL
    uoo, _ = oracle_mod.coarse_solve(dim, nc, o["M"], o["b"], rtol=1e-14)
    assert np.abs(u - uoo).max() <= 1e-10 * np.abs(uoo).max()
    uo, ito = oracle_mod.coarse_solve(dim, nc, M, b, rtol=1e-14)
    assert it > 0 and ito > 0
    assert np.abs(u - uo).max() <= 1e-10 * np.abs(uo).max()
    ufo = oracle_mod.reconstruct(dim, nc, l, uo, Psi)
    assert np.abs(uf - ufo).max() <= 1e-10 * np.abs(ufo).max()


@pytest.mark.parametrize("dim,nc,l", [(3, 16, 1), (3, 24, 1), (2, 128, 1)])
def test_coarse_solve_grid_sizes(oracle_mod, dim, nc, l):
    """The coarse Jacobi-PCG on grids whose search direction fits the
    workgroup's LDS (3-D 16^3: the adjoint and uneven configurations' coarse
    grid) and on grids beyond it (3-D 24^3, 2-D 128^2: the general kernel),
    against the oracle's coarse solve of the same M and b; a = exp(g), g iid
    N(0, 1.5^2) per fine cell."""
    print(f"case coarse sizes dim={dim} nc={nc} l={l}", flush=True)
    from msfec_amd import QBasisBatch, fields as F
    verts = F.coarse_vertices(dim, nc)
    coef = F.lognormal_coefficient(dim, nc, l, 1.5, 0.0, seed=31 + nc)
    with QBasisBatch(verts, dim=dim, n_refine_local=l) as qb:
        qb.set_coefficient(coef)
        qb.run()
        M, b = qb.get_global_element_matrix(), qb.get_global_element_rhs()
        u, w, it, res = qb.coarse_solve(nc, abs_tol=0.0, rel_tol=1e-13)
    uo, ito = oracle_mod.coarse_solve(dim, nc, M, b, rtol=1e-14)
    print("iterations", it, "oracle", ito, "difference", np.abs(u - uo).max() / np.abs(uo).max(),
          flush=True)
    assert it > 0 and ito > 0
    assert np.abs(u - uo).max() <= 1e-10 * np.abs(uo).max()
    # the weights are the coarse values at each cell's vertices
    assert w.shape == (nc ** dim, 1 << dim) and np.isfinite(w).all()


@pytest.mark.parametrize("solver", SOLVERS)
@pytest.mark.parametrize("dim,l", [(2, 4), (3, 3)])
def test_fused_path_rhs_cell_and_reaction(oracle_mod, dim, l, solver):
    """Scalar coefficient with a per-cell load and a per-cell reaction runs
    through the fused kernel; checked against the oracle."""
    print(f"case fused rhs_cell dim={dim} l={l}", flush=True)
    rng = np.random.default_rng(40 + dim)
    n_cells, N = 6, 1 << l
    verts = random_boxes(rng, dim, n_cells)
    coef = np.exp(rng.standard_normal((n_cells, N ** dim)))
    f = rng.standard_normal((n_cells, N ** dim))
    reac = rng.random((n_cells, N ** dim))
    g = run_gpu(verts, dim, l, coef, f=f, rhs_mode=1, reac=reac, reaction_mode=2, solver=solver)
    o = run_oracle(oracle_mod, verts, dim, l, coef, f=f, rhs_mode=1, reac=reac, reaction_mode=2)
    assert rel(g["M"], o["M"]) < tol(solver)
    assert rel(g["b"], o["b"]) < tol(solver)
    assert rel(g["Psi"], o["Psi"]) < tol(solver)


@pytest.mark.parametrize("dim,l", [(2, 3), (3, 3)])
def test_fused_equals_general_path(dim, l):
    """A I given as a scalar (fused kernel) and as a tensor at the quadrature
    points (assembly + band + element kernels) give the same basis."""
    print(f"case fused_vs_general dim={dim} l={l}", flush=True)
    rng = np.random.default_rng(50 + dim)
    n_cells, N, nq = 5, 1 << l, 1 << dim
    verts = random_boxes(rng, dim, n_cells)
    a = np.exp(1.5 * rng.standard_normal((n_cells, N ** dim)))
    T = np.zeros((n_cells, N ** dim, nq, 3 if dim == 2 else 6))
    for c in ([0, 2] if dim == 2 else [0, 3, 5]):
        T[..., c] = a[:, :, None]
    g1 = run_gpu(verts, dim, l, a)
    g2 = run_gpu(verts, dim, l, T, coef_mode=1)
    assert rel(g1["M"], g2["M"]) < TOL
    assert rel(g1["b"], g2["b"]) < TOL
    assert rel(g1["Psi"], g2["Psi"]) < TOL


@pytest.mark.parametrize("dim,l", [(2, 5), (3, 3)])
def test_solvers_agree(dim, l):
    """Vector-ALU and matrix-core variants of the band factorisation agree to
    rounding (different summation orders, same algorithm)."""
    print(f"case solvers_agree dim={dim} l={l}", flush=True)
    rng = np.random.default_rng(60 + dim)
    n_cells, N = 9, 1 << l
    verts = random_boxes(rng, dim, n_cells)
    coef = np.exp(2.0 * rng.standard_normal((n_cells, N ** dim)))
    g0 = run_gpu(verts, dim, l, coef, solver=0)
    for solver in (1, 2):
        g1 = run_gpu(verts, dim, l, coef, solver=solver)
        assert rel(g0["M"], g1["M"]) < tol(solver)
        assert rel(g0["b"], g1["b"]) < tol(solver)
        assert rel(g0["Psi"], g1["Psi"]) < tol(solver)



@pytest.mark.parametrize("dim,l,window", [(2, 1, 16), (2, 2, 16), (2, 3, 16), (2, 3, 32),
                                          (2, 4, 32), (2, 4, 64), (2, 5, 64), (3, 1, 16),
                                          (3, 1, 32), (3, 2, 32), (3, 2, 64)])
def test_band_windows(oracle_mod, dim, l, window):
    """The matrix-core band factorisation with a 16-, 32- or 64-slot window
    (1, 2 or 4 tile rows of 16; the window must hold the half-bandwidth + 4):
    every window that holds the band gives the oracle's basis, and a window
    too small for it is refused."""
    print(f"case band_window dim={dim} l={l} window={window}", flush=True)
    from msfec_amd import QBasisBatch, MsfecError, _lib as L
    rng = np.random.default_rng(70 + 10 * dim + l)
    n_cells, N = 7, 1 << l
    verts = random_boxes(rng, dim, n_cells)
    coef = np.exp(1.5 * rng.standard_normal((n_cells, N ** dim)))
    with QBasisBatch(verts, dim=dim, n_refine_local=l, solver=L.SOLVER_BAND_FP64_MFMA,
                     band_window=window) as qb:
        assert qb.info.band_window == window
        qb.set_coefficient(coef)
        qb.run()
        g = {"M": qb.get_global_element_matrix(), "b": qb.get_global_element_rhs(),
             "Psi": qb.get_basis(), "status": qb.get_cell_status()}
    o = run_oracle(oracle_mod, verts, dim, l, coef)
    assert (g["status"] == 0).all()
    # b and Psi against the oracle: with this contrast (exp(1.5 N(0,1)) iid
    # per fine cell) the oracle's own dense solve leaves ~1e-12 absolute on
    # entries that are exactly zero here, 1.1e-11 relative on Psi and 1.8e-11
    # on b (3-D l = 2, for every window alike): the metric's 1e-10
    for k, t in (("M", TOL), ("b", TOL_MIXED), ("Psi", TOL_MIXED)):
        assert rel(g[k], o[k]) < t, k
    if window > 16:
        ni = N - 1
        bw = ni + 1 if dim == 2 else ni * ni + ni + 1
        if bw + 4 > window // 2:
            with pytest.raises(MsfecError):
                QBasisBatch(verts, dim=dim, n_refine_local=l, solver=L.SOLVER_BAND_FP64_MFMA,
                            band_window=window // 2)


@pytest.mark.parametrize("dim,l,w1,w2", [(2, 4, 32, 64), (3, 2, 32, 64), (2, 3, 16, 32)])
def test_band_windows_agree(dim, l, w1, w2):
    """Two windows that both hold the band do the same eliminations (an
    entry's updates arrive in the same order; rows outside the band only add
    exact zeros), so the bases agree to a few ulps."""
    print(f"case band_windows_agree dim={dim} l={l} {w1} vs {w2}", flush=True)
    from msfec_amd import QBasisBatch, _lib as L
    rng = np.random.default_rng(90 + dim + l)
    n_cells, N = 9, 1 << l
    verts = random_boxes(rng, dim, n_cells)
    coef = np.exp(2.0 * rng.standard_normal((n_cells, N ** dim)))
    out = []
    for w in (w1, w2):
        with QBasisBatch(verts, dim=dim, n_refine_local=l, solver=L.SOLVER_BAND_FP64_MFMA,
                         band_window=w) as qb:
            qb.set_coefficient(coef)
            qb.run()
            out.append((qb.get_global_element_matrix(), qb.get_basis()))
    assert rel(out[0][0], out[1][0]) < 1e-13
    assert rel(out[0][1], out[1][1]) < 1e-13
