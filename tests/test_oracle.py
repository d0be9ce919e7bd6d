"""Pins the FP64 CPU oracle (oracle/msfec_oracle.c) against facts that do not
come from the oracle itself: closed forms, an independent quadrature,
invariants of the continuous problem, and golden vectors made by a separate
dense numpy/scipy implementation (tests/golden/make_golden.py).  CPU only.
"""
import itertools
import os

import numpy as np
import pytest

from conftest import box_vertices

HERE = os.path.dirname(os.path.abspath(__file__))


def q1_stiffness_closed_form(dim, a, h):
    """Q1 element stiffness on a cube: 3-D (a h / 12) S with S_ii = 4, 0 for
    vertices differing in one coordinate, -1 for two or three; 2-D (a / 6) S
    with S_ii = 4, -1 (one coordinate), -2 (both)."""
    nv = 1 << dim
    S = np.zeros((nv, nv))
    for i in range(nv):
        for j in range(nv):
            diff = bin(i ^ j).count("1")
            S[i, j] = ({0: 4, 1: 0, 2: -1, 3: -1} if dim == 3 else {0: 4, 1: -1, 2: -2})[diff]
    return (a * h / 12.0) * S if dim == 3 else (a / 6.0) * S


def box_tensor_stiffness(h, A):
    """Element stiffness of a box with constant tensor A from 1-D integrals
    (exact): K_ij = sum_{d,e} A_de int d_d phi_i d_e phi_j, each integral a
    product of 1-D stiffness/mass/mixed factors."""
    dim = len(h)
    nv = 1 << dim
    # 1-D factors on [0, h]: mass m, stiffness s, mixed c = int phi_a' phi_b
    m = lambda a, b, hh: hh * (1 / 3 if a == b else 1 / 6)
    s = lambda a, b, hh: (1 if a == b else -1) / hh
    c = lambda a, b, hh: (-1 if a == 0 else 1) * 0.5   # int phi_a' phi_b = d_a * 1/2
    K = np.zeros((nv, nv))
    for i in range(nv):
        bi = [(i >> d) & 1 for d in range(dim)]
        for j in range(nv):
            bj = [(j >> d) & 1 for d in range(dim)]
            tot = 0.0
            for d in range(dim):
                for e in range(dim):
                    prod = 1.0
                    for g in range(dim):
                        if d == e == g:
                            prod *= s(bi[g], bj[g], h[g])
                        elif g == d and d != e:
                            prod *= c(bi[g], bj[g], h[g])
                        elif g == e and d != e:
                            prod *= c(bj[g], bi[g], h[g])
                        else:
                            prod *= m(bi[g], bj[g], h[g])
                    tot += A[d, e] * prod
            K[i, j] = tot
    return K


# ---------------------------------------------------------------- BasisQ1 ---
@pytest.mark.parametrize("dim", [2, 3])
def test_basis_q1_delta_and_partition_of_unity(oracle_mod, dim):
    rng = np.random.default_rng(dim)
    nv = 1 << dim
    # a general (non-box) cell: perturbed box vertices
    v = box_vertices(dim, [0.0] * dim, [1.0] * dim) + 0.1 * rng.standard_normal((nv, dim))
    C = oracle_mod.q1_coeffs(dim, v)
    for i in range(nv):
        for k in range(nv):
            assert abs(oracle_mod.q1_value(dim, C, i, v[k]) - (i == k)) < 1e-12
    for x in rng.random((20, dim)):
        assert abs(sum(oracle_mod.q1_value(dim, C, i, x) for i in range(nv)) - 1.0) < 1e-12


@pytest.mark.parametrize("dim", [2, 3])
def test_basis_q1_box_is_tensor_lagrange(oracle_mod, dim):
    rng = np.random.default_rng(10 + dim)
    lo = rng.random(dim)
    hi = lo + 0.1 + rng.random(dim)
    v = box_vertices(dim, lo, hi)
    C = oracle_mod.q1_coeffs(dim, v)
    for x in lo + (hi - lo) * rng.random((20, dim)):
        xi = (x - lo) / (hi - lo)
        for i in range(1 << dim):
            want = np.prod([xi[d] if (i >> d) & 1 else 1 - xi[d] for d in range(dim)])
            assert abs(oracle_mod.q1_value(dim, C, i, x) - want) < 1e-12


# ---------------------------------------------------------- Diffusion_A ---
def test_diffusion_A_rotation_invariants(oracle_mod):
    p = oracle_mod.make_problem(3, 1, coef_mode=oracle_mod.COEF_ANALYTIC, freq=(17, 16, 15),
                                scale=(1.0, 2.0, 3.0), alpha=(0.9, 0.5, 0.1), rotate=True)
    q = oracle_mod.make_problem(3, 1, coef_mode=oracle_mod.COEF_ANALYTIC, freq=(17, 16, 15),
                                scale=(1.0, 2.0, 3.0), alpha=(0.9, 0.5, 0.1), rotate=False)
    rng = np.random.default_rng(3)
    for x in rng.random((10, 3)):
        A = oracle_mod.diffusion_A(p, x)
        D = oracle_mod.diffusion_A(q, x)
        assert np.allclose(D, np.diag(np.diag(D)), atol=0)
        want = [s * (1 - a * np.sin(2 * np.pi * k * xx)) for s, a, k, xx in
                zip((1.0, 2.0, 3.0), (0.9, 0.5, 0.1), (17, 16, 15), x)]
        assert np.allclose(np.diag(D), want, rtol=1e-14)
        assert np.allclose(A, A.T, atol=1e-14)
        # orthogonal similarity: same spectrum as the diagonal
        assert np.allclose(np.sort(np.linalg.eigvalsh(A)), np.sort(want), rtol=1e-12)


def test_diffusion_A_is_the_zxz_euler_rotation(oracle_mod):
    """eqn_coeff_A.cc:25-39 builds rot from the fixed angles alpha = pi/3,
    beta = pi/6, gamma = pi/4 as R_z(alpha) R_x(beta) R_z(gamma) (intrinsic
    z-x'-z'' Euler angles) and :185-190 sets value = rot * D * transpose(rot).
    scipy's Rotation is an independent construction of that matrix; the
    transposed similarity rot^T D rot (a plausible slip) must fail."""
    from scipy.spatial.transform import Rotation
    R = Rotation.from_euler("ZXZ", [np.pi / 3, np.pi / 6, np.pi / 4]).as_matrix()
    # the first row as eqn_coeff_A.cc:27-30 writes it
    a, b, g = np.pi / 3, np.pi / 6, np.pi / 4
    assert np.allclose(R[0], [np.cos(a) * np.cos(g) - np.sin(a) * np.cos(b) * np.sin(g),
                              -np.cos(a) * np.sin(g) - np.sin(a) * np.cos(b) * np.cos(g),
                              np.sin(a) * np.sin(b)], atol=1e-15)
    scale, alpha, freq = (1.0, 2.0, 3.0), (0.9, 0.5, 0.1), (17, 16, 15)
    p = oracle_mod.make_problem(3, 1, coef_mode=oracle_mod.COEF_ANALYTIC, freq=freq,
                                scale=scale, alpha=alpha, rotate=True)
    rng = np.random.default_rng(21)
    worst_transposed = 0.0
    for x in np.vstack([[0.3, 0.7, 0.1], rng.random((40, 3))]):
        D = np.diag([s * (1 - al * np.sin(2 * np.pi * k * xx))
                     for s, al, k, xx in zip(scale, alpha, freq, x)])
        A = oracle_mod.diffusion_A(p, x)
        assert np.abs(A - R @ D @ R.T).max() < 1e-14 * np.abs(D).max()
        worst_transposed = max(worst_transposed, np.abs(A - R.T @ D @ R).max())
    assert worst_transposed > 0.5


def test_diffusion_A_matches_field_generator(oracle_mod):
    """The host-side quadrature-point tensor used for the GPU path agrees with
    the oracle's inline evaluation."""
    from msfec_amd import fields as F
    p = oracle_mod.make_problem(3, 1, coef_mode=oracle_mod.COEF_ANALYTIC, freq=(17, 16, 15),
                                alpha=(0.9, 0.9, 0.9), rotate=True)
    pts = np.random.default_rng(4).random((7, 3))
    T = F.diffusion_A_tensor(pts)
    for x, t in zip(pts, T):
        A = oracle_mod.diffusion_A(p, x)
        assert np.allclose(t, [A[0, 0], A[0, 1], A[0, 2], A[1, 1], A[1, 2], A[2, 2]], atol=1e-14)


# ------------------------------------------------------ element matrices ---
@pytest.mark.parametrize("dim", [2, 3])
def test_element_stiffness_closed_form(oracle_mod, dim):
    h = 0.37
    a = 2.5
    nq, nc = 1 << dim, (3 if dim == 2 else 6)
    Aq = np.zeros((nq, nc))
    for c in ([0, 2] if dim == 2 else [0, 3, 5]):
        Aq[:, c] = a
    Ke, Fe = oracle_mod.element_matrix(dim, [h] * dim, Aq, np.full(nq, 3.0))
    assert np.allclose(Ke, q1_stiffness_closed_form(dim, a, h), rtol=1e-13, atol=1e-15)
    assert np.allclose(Fe, 3.0 * h ** dim / (1 << dim), rtol=1e-14)


@pytest.mark.parametrize("dim", [2, 3])
def test_element_tensor_box_exact(oracle_mod, dim):
    """Anisotropic box, full constant tensor: QGauss(2) equals the exact
    integral assembled from 1-D factors."""
    rng = np.random.default_rng(20 + dim)
    h = 0.1 + rng.random(dim)
    B = rng.standard_normal((dim, dim))
    A = B @ B.T + np.eye(dim)
    nq = 1 << dim
    sym = [A[0, 0], A[0, 1], A[1, 1]] if dim == 2 else [A[0, 0], A[0, 1], A[0, 2], A[1, 1],
                                                        A[1, 2], A[2, 2]]
    Ke, _ = oracle_mod.element_matrix(dim, h, np.tile(sym, (nq, 1)), np.zeros(nq))
    assert np.allclose(Ke, box_tensor_stiffness(h, A), rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("dim", [2, 3])
def test_element_mass_closed_form(oracle_mod, dim):
    h = [0.3, 0.5, 0.7][:dim]
    nq, nc = 1 << dim, (3 if dim == 2 else 6)
    Ke, _ = oracle_mod.element_matrix(dim, h, np.zeros((nq, nc)), np.zeros(nq), rq=2.0)
    nv = 1 << dim
    want = np.array([[2.0 * np.prod([h[d] * (1 / 3 if ((i ^ j) >> d) & 1 == 0 else 1 / 6)
                                     for d in range(dim)]) for j in range(nv)]
                     for i in range(nv)])
    assert np.allclose(Ke, want, rtol=1e-13)


# ------------------------------------------------------------ local basis ---
@pytest.mark.parametrize("dim,l", [(2, 3), (3, 2)])
def test_constant_coefficient_gives_coarse_q1(oracle_mod, dim, l):
    """A = aI: the trilinear phi_i is A-harmonic and lies in the fine space,
    so psi_i = phi_i and M is the coarse Q1 stiffness."""
    H = 0.25
    lo = np.full(dim, 0.5)
    v = box_vertices(dim, lo, lo + H)[None]
    N = 1 << l
    a = 1.7
    p = oracle_mod.make_problem(dim, l, rhs_const=2.0)
    r = oracle_mod.basis_batch(p, v, coef=np.full((1, N ** dim), a))
    assert r["status"][0] == 0
    assert np.allclose(r["M"][0], q1_stiffness_closed_form(dim, a, H), rtol=1e-12, atol=1e-15)
    S = N + 1
    xi = np.array([[(i // S ** d) % S for d in range(dim)] for i in range(S ** dim)]) / N
    for k in range(1 << dim):
        phi = np.prod([xi[:, d] if (k >> d) & 1 else 1 - xi[:, d] for d in range(dim)], axis=0)
        assert np.abs(r["Psi"][0, k] - phi).max() < 1e-13
    assert np.allclose(r["b"][0], 2.0 * H ** dim / (1 << dim), rtol=1e-12)


@pytest.mark.parametrize("dim,l", [(2, 4), (3, 3)])
def test_invariants_rough_coefficient(oracle_mod, dim, l):
    rng = np.random.default_rng(30 + dim)
    N = 1 << l
    lo = rng.random(dim)
    hi = lo + 0.1 + 0.1 * rng.random(dim)
    v = box_vertices(dim, lo, hi)[None]
    coef = np.exp(2.0 * rng.standard_normal((1, N ** dim)))
    p = oracle_mod.make_problem(dim, l)
    r = oracle_mod.basis_batch(p, v, coef=coef)
    M, b, Psi = r["M"][0], r["b"][0], r["Psi"][0]
    scale = np.abs(M).max()
    assert np.abs(M - M.T).max() < 1e-13 * scale
    assert np.abs(M.sum(axis=1)).max() < 1e-12 * scale                 # M 1 = 0
    assert np.abs(Psi.sum(axis=0) - 1).max() < 1e-12                   # partition of unity
    assert np.linalg.eigvalsh(M)[0] > -1e-12 * scale                    # PSD
    assert abs(b.sum() - np.prod(hi - lo)) < 1e-14                       # sum b = int f
    # energy minimisation: M_ii <= phi_i^T K phi_i = energy of the trilinear
    # extension, evaluated with the oracle's element matrices
    S = N + 1
    h = (hi - lo) / N
    xi = np.array([[(i // S ** d) % S for d in range(dim)] for i in range(S ** dim)]) / N
    nq, nc = 1 << dim, (3 if dim == 2 else 6)
    for k in range(1 << dim):
        phi = np.prod([xi[:, d] if (k >> d) & 1 else 1 - xi[:, d] for d in range(dim)], axis=0)
        e = 0.0
        for z in range(N ** dim):
            cz = [(z // N ** d) % N for d in range(dim)]
            Aq = np.zeros((nq, nc))
            for c in ([0, 2] if dim == 2 else [0, 3, 5]):
                Aq[:, c] = coef[0, z]
            Ke, _ = oracle_mod.element_matrix(dim, h, Aq, np.zeros(nq))
            ids = [sum((cz[d] + ((a >> d) & 1)) * S ** d for d in range(dim)) for a in range(nq)]
            e += phi[ids] @ Ke @ phi[ids]
        assert M[k, k] <= e * (1 + 1e-12)


def test_single_interior_node_closed_form(oracle_mod):
    """l = 1 in 3-D: one interior node, x = -(K_IB g)/K_II by hand."""
    rng = np.random.default_rng(7)
    coef = np.exp(rng.standard_normal((1, 8)))
    v = box_vertices(3, [0, 0, 0], [1, 1, 1])[None]
    p = oracle_mod.make_problem(3, 1)
    r = oracle_mod.basis_batch(p, v, coef=coef)
    # centre node couples to all 26 others through the 8 cells; with A = a_z I
    # on a cube of side 1/2 the couplings are (a h/12) S entries of each cell
    h = 0.5
    S3 = q1_stiffness_closed_form(3, 1.0, h)
    diag = sum(coef[0, z] * S3[7 - z, 7 - z] for z in range(8))
    # node of cell z at local corner (7 - z) is the centre; the other corners
    # carry the trilinear boundary values of basis k at xi in {0, 1/2, 1}
    for k in range(8):
        rhs = 0.0
        for z in range(8):
            cz = [(z >> d) & 1 for d in range(3)]
            a_loc = 7 - z
            for b_loc in range(8):
                if b_loc == a_loc:
                    continue
                xi = [(cz[d] + ((b_loc >> d) & 1)) / 2 for d in range(3)]
                g = np.prod([xi[d] if (k >> d) & 1 else 1 - xi[d] for d in range(3)])
                rhs -= coef[0, z] * S3[a_loc, b_loc] * g
        assert abs(r["Psi"][0, k, 13] - rhs / diag) < 1e-14


def test_cg_ssor_agrees_with_direct(oracle_mod):
    from msfec_amd import fields as F
    verts = F.coarse_vertices(3, 4)[:6]
    coef = F.lognormal_coefficient(3, 4, 3, 1.5, 1 / 16, seed=5)[:6]
    p = oracle_mod.make_problem(3, 3)
    d = oracle_mod.basis_batch(p, verts, coef=coef, solver=oracle_mod.SOLVER_DIRECT)
    c = oracle_mod.basis_batch(p, verts, coef=coef, solver=oracle_mod.SOLVER_CG_SSOR, n_threads=2)
    assert (c["status"] == 0).all() and (c["iters"] > 0).all()
    assert np.abs(c["M"] - d["M"]).max() < 1e-9 * np.abs(d["M"]).max()
    # Psi=None path (M and b only) gives the same numbers
    c2 = oracle_mod.basis_batch(p, verts, coef=coef, solver=oracle_mod.SOLVER_CG_SSOR,
                                want_psi=False)
    assert np.array_equal(c2["M"], c["M"])


def test_cg_nonconvergence_reported(oracle_mod):
    from msfec_amd import fields as F
    verts = F.coarse_vertices(3, 2)[:2]
    coef = F.lognormal_coefficient(3, 2, 3, 1.5, 0.0, seed=6)[:2]
    p = oracle_mod.make_problem(3, 3, cg_maxit=1)
    r = oracle_mod.basis_batch(p, verts, coef=coef, solver=oracle_mod.SOLVER_CG_SSOR)
    assert (r["status"] == 1).all()


def test_golden_vectors(oracle_mod):
    """Oracle vs the independent dense numpy/scipy implementation."""
    z = np.load(os.path.join(HERE, "golden", "q_basis_golden.npz"), allow_pickle=False)
    for i in range(int(z["n_cases"])):
        g = {k.split("_", 1)[1]: z[k] for k in z.files if k.startswith(f"c{i}_")}
        dim, l = int(g["dim"]), int(g["l"])
        v = box_vertices(dim, g["lo"], g["hi"])[None]
        kw = {}
        if g["f"].size:
            kw.update(rhs_mode=oracle_mod.RHS_CELL, reaction_mode=oracle_mod.REACTION_CELL)
        p = oracle_mod.make_problem(dim, l, **kw)
        r = oracle_mod.basis_batch(p, v, coef=g["coef"][None],
                                   f=g["f"][None] if g["f"].size else None,
                                   reaction=g["reac"][None] if g["reac"].size else None)
        assert r["status"][0] == 0
        for name in ("M", "b", "Psi"):
            ref = g[name]
            err = np.abs(r[name][0] - ref).max() / np.abs(ref).max()
            assert err < 1e-12, (i, name, err)


# ---------------------------------------------------------------- coarse ---
def test_coarse_constant_coefficient_equals_standard_q1(oracle_mod):
    """A = I: the multiscale coarse problem is the standard coarse Q1 FEM;
    checked against a dense numpy solve of the closed-form coarse system."""
    dim, nc, l = 2, 6, 2
    from msfec_amd import fields as F
    verts = F.coarse_vertices(dim, nc)
    p = oracle_mod.make_problem(dim, l)
    r = oracle_mod.basis_batch(p, verts, coef=np.ones((nc ** dim, (1 << l) ** dim)))
    u, it = oracle_mod.coarse_solve(dim, nc, r["M"], r["b"])
    assert it > 0
    H = 1.0 / nc
    Ke = q1_stiffness_closed_form(dim, 1.0, H)
    nn = (nc + 1) ** dim
    K = np.zeros((nn, nn))
    F_ = np.zeros(nn)
    for c in range(nc ** dim):
        cc = [(c // nc ** d) % nc for d in range(dim)]
        ids = [sum((cc[d] + ((v >> d) & 1)) * (nc + 1) ** d for d in range(dim))
               for v in range(1 << dim)]
        K[np.ix_(ids, ids)] += Ke
        F_[ids] += H ** dim / (1 << dim)
    coords = np.array([[(i // (nc + 1) ** d) % (nc + 1) for d in range(dim)] for i in range(nn)])
    I = np.where(np.all((coords > 0) & (coords < nc), axis=1))[0]
    uref = np.zeros(nn)
    uref[I] = np.linalg.solve(K[np.ix_(I, I)], F_[I])
    assert np.abs(u - uref).max() < 1e-12 * np.abs(uref).max()
    uf = oracle_mod.reconstruct(dim, nc, l, u, r["Psi"])
    # reconstruction is linear in the weights and reproduces nodal values
    S = (1 << l) + 1
    for c in range(nc ** dim):
        cc = [(c // nc ** d) % nc for d in range(dim)]
        for v in range(1 << dim):
            node_f = sum((((v >> d) & 1) * (S - 1)) * S ** d for d in range(dim))
            node_c = sum((cc[d] + ((v >> d) & 1)) * (nc + 1) ** d for d in range(dim))
            assert abs(uf[c, node_f] - u[node_c]) < 1e-14


@pytest.mark.parametrize("dim,nc,l", [(2, 3, 2), (3, 2, 2)])
def test_adjoint_gradient_matches_finite_differences(oracle_mod, dim, nc, l):
    """The oracle's adjoint gradient of J = 1/2 |u_H - u_target|^2 with respect
    to the log-coefficient, against central differences of its own misfit
    along random directions (truncation O(eps^2) ~ 1e-10)."""
    from msfec_amd import fields as F
    O = oracle_mod
    verts = F.coarse_vertices(dim, nc)
    N = 1 << l
    rng = np.random.default_rng(21 + dim)
    g = 0.8 * rng.standard_normal((nc ** dim, N ** dim))
    p = O.make_problem(dim, l)
    _, ut = O.misfit(p, nc, verts, np.exp(0.5 * rng.standard_normal(g.shape)),
                     np.zeros((nc + 1) ** dim))
    J, grad = O.adjoint_gradient(p, nc, verts, np.exp(g), ut)
    J2, _ = O.misfit(p, nc, verts, np.exp(g), ut)
    assert abs(J - J2) <= 1e-14 * abs(J)
    eps = 1e-5
    for _ in range(3):
        dg = rng.standard_normal(g.shape)
        Jp, _ = O.misfit(p, nc, verts, np.exp(g + eps * dg), ut)
        Jm, _ = O.misfit(p, nc, verts, np.exp(g - eps * dg), ut)
        fd = (Jp - Jm) / (2 * eps)
        ad = float((grad * dg).sum())
        # relative to |grad| |dg|, the largest |<grad, dg>| can be: a random
        # direction can make the derivative itself nearly vanish
        scale = np.linalg.norm(grad) * np.linalg.norm(dg)
        assert abs(fd - ad) <= 1e-8 * scale


def test_coarse_solve_rhs_equals_assembled(oracle_mod):
    """The nodal right-hand-side variant reproduces the assembled solve."""
    O = oracle_mod
    dim, nc = 2, 4
    rng = np.random.default_rng(3)
    # a valid coarse system: element matrices of A = I on unit squares
    Ke, _ = O.element_matrix(2, np.array([0.25, 0.25]), np.tile([1.0, 0.0, 1.0], (4, 1)),
                             np.zeros(4), 0.0)
    M = np.tile(Ke, (nc * nc, 1, 1))
    b = rng.random((nc * nc, 4))
    u1, it1 = O.coarse_solve(dim, nc, M, b)
    rhs = np.zeros((nc + 1) ** dim)
    for c in range(nc * nc):
        cx, cy = c % nc, c // nc
        for v in range(4):
            node = (cx + (v & 1)) + (nc + 1) * (cy + (v >> 1))
            rhs[node] += b[c, v]
    u2, it2 = O.coarse_solve_rhs(dim, nc, M, rhs)
    assert it1 > 0 and it2 > 0
    assert np.abs(u1 - u2).max() <= 1e-13 * np.abs(u1).max()
