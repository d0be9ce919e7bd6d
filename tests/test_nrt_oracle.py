"""Pins the 3-D Nedelec x Raviart-Thomas oracle (oracle/nrt_oracle.c) without the GPU.

The oracle restates NedRTBasis (source/Ned_RT/ned_rt_basis.cc): per coarse
box, a mixed curl/div system over the fine edges (sigma) and faces (u), one
solve per edge basis (12) and face basis (6), and an 18 x 18 element matrix.
What fixes its results independently of its own code:
  - the discrete de Rham sequence: the edge values of the gradient of a nodal
    Q1 field are annihilated by the coupling block (curl grad = 0); the face
    circulations of an edge field are annihilated by the divergence block
    (div curl = 0), and the coupling block maps an edge field to the face mass
    matrix times its circulations (Stokes on every fine face);
  - closed forms: the edge mass matrix integrates c . G c of constant fields
    for a full constant tensor G = A^-1, the face mass |c|^2, the divergence
    block the integral of B for the field x e_x, the load (v, f) of a
    constant f is f_d h_d per interior face;
  - the loads (v, curl N_n) equal the face mass times the circulations of the
    coarse Nedelec function's edge values;
  - a manufactured smooth solution of the mixed problem converges at first
    order when the assembled system is solved on 2^3, 4^3 and 8^3 fine cells;
  - with A^-1 = a I and constant B and no reaction the bases are the coarse
    functions themselves and M is the standard lowest-order element matrix
    (closed form by exact 1-D integrals here);
  - any coefficients: the bases satisfy the assembled interior equations with
    the loads above and carry the coarse traces on the boundary.
"""
import numpy as np
import pytest
import sympy as sp

from oracle import oracle as O

LO = np.array([0.2, 0.1, 0.3])
HI = np.array([0.7, 0.5, 1.0])
VERTS = O.box_vertices_3d(LO, HI)
H = HI - LO
GAUSS = np.array([0.5 - 0.5 / np.sqrt(3), 0.5 + 0.5 / np.sqrt(3)])
# deal.II line order: axis, then the values of the two other coordinates
LINES = [(1, 0, 0), (1, 1, 0), (0, 0, 0), (0, 1, 0), (1, 0, 1), (1, 1, 1),
         (0, 0, 1), (0, 1, 1), (2, 0, 0), (2, 1, 0), (2, 0, 1), (2, 1, 1)]


def others(d):
    return [a for a in range(3) if a != d]


class Dofs:
    """The fine dof numbering documented in nrt_oracle.h."""

    def __init__(self, l):
        self.l = l
        self.m, self.nx, self.ne, self.nf1, self.nf, self.N = O.nrt_sizes(l)

    def edge(self, t, i, j, k):
        m = self.m
        if t == 0:
            return (k * (m + 1) + j) * m + i
        if t == 1:
            return self.nx + (k * m + j) * (m + 1) + i
        return 2 * self.nx + (k * (m + 1) + j) * (m + 1) + i

    def face(self, t, i, j, k):
        m = self.m
        if t == 0:
            return self.ne + (k * m + j) * (m + 1) + i
        if t == 1:
            return self.ne + self.nf1 + (k * (m + 1) + j) * m + i
        return self.ne + 2 * self.nf1 + (k * m + j) * m + i

    def edges(self):
        """(dof, axis, node index (i, j, k) of its start)"""
        m = self.m
        for t in range(3):
            ext = [m + 1] * 3
            ext[t] = m
            for k in range(ext[2]):
                for j in range(ext[1]):
                    for i in range(ext[0]):
                        yield self.edge(t, i, j, k), t, (i, j, k)

    def faces(self):
        """(dof, normal axis, (i, j, k): the face's level along the axis, the
        cell indices along the others)"""
        m = self.m
        for t in range(3):
            ext = [m] * 3
            ext[t] = m + 1
            for k in range(ext[2]):
                for j in range(ext[1]):
                    for i in range(ext[0]):
                        yield self.face(t, i, j, k), t, (i, j, k)

    def edge_on_boundary(self, t, p):
        return any(p[d] in (0, self.m) for d in others(t))

    def face_on_boundary(self, t, p):
        return p[t] in (0, self.m)


def rand_coefs(l, seed, reaction=True):
    rng = np.random.default_rng(seed)
    nc = (1 << l) ** 3
    Lm = rng.normal(size=(nc, 8, 3, 3)) * 0.4 + np.eye(3)
    G = Lm @ np.swapaxes(Lm, -1, -2) + 0.2 * np.eye(3)
    ainv = np.stack([G[..., 0, 0], G[..., 0, 1], G[..., 0, 2], G[..., 1, 1], G[..., 1, 2],
                     G[..., 2, 2]], axis=-1)
    bq = np.exp(rng.normal(size=(nc, 8)))
    rq = np.exp(rng.normal(size=(nc, 8))) if reaction else None
    fq = rng.normal(size=(nc, 8, 3))
    return ainv, bq, rq, fq


def grad_edges(D, q):
    """edge values of the gradient of the nodal Q1 field q[k, j, i]"""
    w = np.zeros(D.ne)
    for e, t, (i, j, k) in D.edges():
        p1 = [i, j, k]
        p1[t] += 1
        w[e] = q[p1[2], p1[1], p1[0]] - q[k, j, i]
    return w


def circulation_values(D, w):
    """face values of curl of the edge field w: the circulation around each
    fine face, counter-clockwise seen from the +normal side (Stokes)"""
    E = D.edge
    c = np.zeros(D.nf)
    for f, t, (i, j, k) in D.faces():
        if t == 0:
            v = w[E(1, i, j, k)] + w[E(2, i, j + 1, k)] - w[E(1, i, j, k + 1)] - w[E(2, i, j, k)]
        elif t == 1:
            v = w[E(2, i, j, k)] + w[E(0, i, j, k + 1)] - w[E(2, i + 1, j, k)] - w[E(0, i, j, k)]
        else:
            v = w[E(0, i, j, k)] + w[E(1, i + 1, j, k)] - w[E(0, i, j + 1, k)] - w[E(1, i, j, k)]
        c[f - D.ne] = v
    return c


def coarse_edge_values(D):
    """fine edge values of the 12 coarse Nedelec functions (their line
    integrals: (1/m) l_f1 l_f2 at the edge's fixed node coordinates)"""
    m = D.m
    lag = lambda s, t: t if s else 1.0 - t
    I = np.zeros((12, D.ne))
    for e, t, p in D.edges():
        for n, (ax, f1, f2) in enumerate(LINES):
            if t == ax:
                o1, o2 = others(ax)
                I[n, e] = lag(f1, p[o1] / m) * lag(f2, p[o2] / m) / m
    return I


def coarse_face_values(D):
    """fine face fluxes of the 6 coarse Raviart-Thomas functions"""
    m = D.m
    lag = lambda s, t: t if s else 1.0 - t
    I = np.zeros((6, D.nf))
    for f, t, p in D.faces():
        for s in range(2):
            I[2 * t + s, f - D.ne] = lag(s, p[t] / m) / (m * m)
    return I


@pytest.mark.parametrize("l", [1, 2])
def test_coupling_block_is_antisymmetric_and_coefficient_free(l):
    D = Dofs(l)
    A1, _ = O.nrt_assemble(l, VERTS, *rand_coefs(l, 1))
    A2, _ = O.nrt_assemble(l, VERTS, *rand_coefs(l, 2))
    ne = D.ne
    C = A1[ne:, :ne]
    assert np.abs(C).max() > 0
    np.testing.assert_allclose(A1[:ne, ne:], -C.T, rtol=0, atol=1e-13 * np.abs(C).max())
    np.testing.assert_allclose(A2[ne:, :ne], C, rtol=0, atol=1e-13 * np.abs(C).max())
    np.testing.assert_allclose(A1[:ne, :ne], A1[:ne, :ne].T, rtol=0,
                               atol=1e-13 * np.abs(A1[:ne, :ne]).max())


@pytest.mark.parametrize("l", [1, 2])
def test_curl_of_gradient_vanishes(l):
    D = Dofs(l)
    A, _ = O.nrt_assemble(l, VERTS, *rand_coefs(l, 3))
    q = np.random.default_rng(4).normal(size=(D.m + 1,) * 3)
    g = grad_edges(D, q)
    C = A[D.ne:, :D.ne]
    assert np.abs(C @ g).max() < 1e-12 * np.abs(C).max() * np.abs(g).max()
    # a wrong edge orientation or line order breaks it
    g_bad = g.copy()
    g_bad[D.edge(1, 1, 0, 1)] *= -1
    assert np.abs(C @ g_bad).max() > 1e-3 * np.abs(C).max() * np.abs(g).max()


@pytest.mark.parametrize("l", [1, 2])
def test_coupling_equals_face_mass_times_circulation(l):
    D = Dofs(l)
    rng = np.random.default_rng(5)
    w = rng.normal(size=D.ne)
    c = circulation_values(D, w)
    nc = D.m ** 3
    ainv, _, _, _ = rand_coefs(l, 6)
    # B random, no reaction: the face block is B div div, which annihilates curls
    A, _ = O.nrt_assemble(l, VERTS, ainv, np.exp(rng.normal(size=(nc, 8))), None)
    Duu = A[D.ne:, D.ne:]
    assert np.abs(Duu @ c).max() < 1e-12 * np.abs(Duu).max() * np.abs(c).max()
    # B = 0, R = 1: the face block is the face mass matrix
    A0, _ = O.nrt_assemble(l, VERTS, ainv, 0.0, 1.0)
    Mf = A0[D.ne:, D.ne:]
    lhs = A[D.ne:, :D.ne] @ w
    np.testing.assert_allclose(lhs, Mf @ c, rtol=0, atol=1e-12 * np.abs(lhs).max())


@pytest.mark.parametrize("l", [1, 2])
def test_mass_matrices_closed_forms(l):
    D = Dofs(l)
    h = H / D.m
    vol = np.prod(H)
    rng = np.random.default_rng(7)
    Lm = rng.normal(size=(3, 3))
    G = Lm @ Lm.T + np.eye(3)
    ainv = np.array([G[0, 0], G[0, 1], G[0, 2], G[1, 1], G[1, 2], G[2, 2]])
    A, _ = O.nrt_assemble(l, VERTS, ainv, 0.0, 1.0)
    for _ in range(3):
        c = rng.normal(size=3)
        w = np.zeros(D.ne)
        for e, t, _p in D.edges():
            w[e] = c[t] * h[t]
        assert w @ A[:D.ne, :D.ne] @ w == pytest.approx(c @ G @ c * vol, rel=1e-13)
        u = np.zeros(D.nf)
        for f, t, _p in D.faces():
            u[f - D.ne] = c[t] * h[others(t)[0]] * h[others(t)[1]]
        assert u @ A[D.ne:, D.ne:] @ u == pytest.approx(c @ c * vol, rel=1e-13)
    # div div with B at the Gauss points, field x e_x (divergence 1)
    nc = D.m ** 3
    bq = np.exp(rng.normal(size=(nc, 8)))
    Ab, _ = O.nrt_assemble(l, VERTS, ainv, bq, None)
    u = np.zeros(D.nf)
    for f, t, p in D.faces():
        if t == 0:
            u[f - D.ne] = (LO[0] + p[0] * h[0]) * h[1] * h[2]
    assert u @ Ab[D.ne:, D.ne:] @ u == pytest.approx(bq.sum() * vol / nc / 8, rel=1e-13)


@pytest.mark.parametrize("l", [1, 2])
def test_load_of_constant_field(l):
    D = Dofs(l)
    h = H / D.m
    f = np.array([0.7, -1.3, 2.1])
    _, F = O.nrt_assemble(l, VERTS, [1, 0, 0, 1, 0, 1], 1.0, None, f)
    assert np.abs(F[:D.ne]).max() == 0
    for dof, t, p in D.faces():
        want = f[t] * h[t] * (0.5 if D.face_on_boundary(t, p) else 1.0)
        assert F[dof] == pytest.approx(want, rel=1e-13)


@pytest.mark.parametrize("l", [1, 2])
def test_curl_loads_are_face_mass_times_coarse_circulations(l):
    D = Dofs(l)
    Fc = O.nrt_curl_loads(l, VERTS)
    A0, _ = O.nrt_assemble(l, VERTS, [1, 0, 0, 1, 0, 1], 0.0, 1.0)
    Mf = A0[D.ne:, D.ne:]
    I = coarse_edge_values(D)
    assert np.abs(Fc[:, :D.ne]).max() == 0
    for n in range(12):
        want = Mf @ circulation_values(D, I[n])
        np.testing.assert_allclose(Fc[n, D.ne:], want, rtol=0, atol=1e-13 * np.abs(want).max())


def _manufactured():
    x, y, z = sp.symbols("x y z")
    X = (x, y, z)
    u = sp.Matrix([sp.sin(y) + x * z, sp.cos(z) + x * y ** 2, sp.sin(x) + y * z])
    G = sp.Matrix([[1.5, 0.3, 0.0], [0.3, 1.0, -0.2], [0.0, -0.2, 0.8]])  # A^-1
    Bf = 2 + x
    Rf = sp.Integer(1)

    def curl(v):
        return sp.Matrix([sp.diff(v[2], y) - sp.diff(v[1], z), sp.diff(v[0], z) - sp.diff(v[2], x),
                          sp.diff(v[1], x) - sp.diff(v[0], y)])

    div = sum(sp.diff(u[d], X[d]) for d in range(3))
    sig = G.inv() * curl(u)
    f = curl(sig) - sp.Matrix([sp.diff(Bf * div, X[d]) for d in range(3)]) + Rf * u
    lam = lambda e: sp.lambdify(X, e, "numpy")
    return lam(list(sig)), lam(list(u)), lam(list(f)), lam(Bf), np.array(G, dtype=float)


def _edge_integrals(D, fn, lo, h):
    """line integrals of the tangential component of fn along every fine edge"""
    gx, gw = np.polynomial.legendre.leggauss(5)
    gx, gw = 0.5 * (gx + 1), 0.5 * gw
    out = np.zeros(D.ne)
    for e, t, p in D.edges():
        pts = np.array([lo + (np.array(p) + (np.arange(3) == t) * s) * h for s in gx])
        vals = np.array(fn(pts[:, 0], pts[:, 1], pts[:, 2]), dtype=float)
        out[e] = (vals[t] * gw).sum() * h[t]
    return out


def _face_fluxes(D, fn, lo, h):
    gx, gw = np.polynomial.legendre.leggauss(5)
    gx, gw = 0.5 * (gx + 1), 0.5 * gw
    out = np.zeros(D.nf)
    for f, t, p in D.faces():
        o1, o2 = others(t)
        acc = 0.0
        for a, wa in zip(gx, gw):
            for b, wb in zip(gx, gw):
                pt = lo + np.array(p, dtype=float) * h
                pt[o1] += a * h[o1]
                pt[o2] += b * h[o2]
                acc += wa * wb * float(fn(*pt)[t])
        out[f - D.ne] = acc * h[o1] * h[o2]
    return out


def test_manufactured_solution_converges():
    sig, u, f, Bf, G = _manufactured()
    errs = []
    for l in (1, 2, 3):
        D = Dofs(l)
        m = D.m
        h = H / m
        nc = m ** 3
        pts = np.zeros((nc, 8, 3))
        for c in range(nc):
            a, b, cc = c % m, (c // m) % m, c // (m * m)
            for q in range(8):
                pts[c, q] = LO + (np.array([a, b, cc]) +
                                  np.array([GAUSS[q & 1], GAUSS[(q >> 1) & 1], GAUSS[(q >> 2) & 1]])) * h
        P = pts.reshape(-1, 3)
        fq = np.stack(f(P[:, 0], P[:, 1], P[:, 2]), axis=-1).reshape(nc, 8, 3)
        bq = np.broadcast_to(Bf(P[:, 0], P[:, 1], P[:, 2]), (nc * 8,)).reshape(nc, 8)
        ainv = np.array([G[0, 0], G[0, 1], G[0, 2], G[1, 1], G[1, 2], G[2, 2]])
        A, F = O.nrt_assemble(l, VERTS, ainv, bq, 1.0, fq)
        xs = np.concatenate([_edge_integrals(D, sig, LO, h), _face_fluxes(D, u, LO, h)])
        bnd = np.zeros(D.N, dtype=bool)
        for e, t, p in D.edges():
            bnd[e] = D.edge_on_boundary(t, p)
        for dof, t, p in D.faces():
            bnd[dof] = D.face_on_boundary(t, p)
        I = ~bnd
        x = xs.copy()
        x[I] = np.linalg.solve(A[np.ix_(I, I)], F[I] - A[np.ix_(I, bnd)] @ xs[bnd])
        scale = np.concatenate([np.full(D.ne, 1.0 / h.mean()), np.full(D.nf, 1.0 / h.mean() ** 2)])
        errs.append(np.abs((x - xs) * scale)[I].max())
    assert errs[0] / errs[1] > 1.6 and errs[1] / errs[2] > 1.6, errs
    assert errs[2] < 0.05


def _coarse_element_matrix(a, beta, fvec):
    """The standard lowest-order Nedelec x Raviart-Thomas element matrix of the
    box with A^-1 = a I, B = beta, no reaction, by exact 1-D integrals:
    int l_s l_t = 1/3 (s = t) or 1/6, int l_s = 1/2, l_s' = -1 or +1."""
    ll = lambda s, t: (1 / 3 if s == t else 1 / 6)
    dl = lambda s: 1.0 if s else -1.0
    vol = np.prod(H)
    M = np.zeros((18, 18))
    for i, (ai, fi1, fi2) in enumerate(LINES):
        for j, (aj, fj1, fj2) in enumerate(LINES):
            if ai == aj:
                o1, o2 = others(ai)
                M[i, j] = a * ll(fi1, fj1) * ll(fi2, fj2) * vol / H[ai] ** 2
    for i, (ax, f1, f2) in enumerate(LINES):
        o1, o2 = others(ax)
        e_ax = np.eye(3)[ax]
        for j in range(6):
            d, s = j // 2, j % 2
            if d == ax:
                continue
            # curl(g e_ax) = grad g x e_ax with g = l_f1(X_o1) l_f2(X_o2) / H_ax;
            # its component along d (one of o1, o2) comes from the derivative
            # along the other one, oo; psi_j = l_s(X_d) e_d / (H_ax H_oo)
            oo, f_oo, f_d = (o2, f2, f1) if d == o1 else (o1, f1, f2)
            cross = np.cross(np.eye(3)[oo], e_ax)[d]
            val = cross * dl(f_oo) / (H[ax] * H[oo]) * ll(f_d, s) * vol / (H[ax] * H[oo])
            M[i, 12 + j] = -val
            M[12 + j, i] = val
    s = np.array([-1.0, 1.0, -1.0, 1.0, -1.0, 1.0])
    M[12:, 12:] = beta * np.outer(s, s) / vol
    b = np.zeros(18)
    for j in range(6):
        b[12 + j] = fvec[j // 2] * H[j // 2] / 2
    return M, b


@pytest.mark.parametrize("l", [1, 2])
def test_constant_coefficients_reproduce_the_coarse_element(l):
    D = Dofs(l)
    a, beta, fvec = 1.7, 0.6, np.array([0.3, -1.1, 0.8])
    M, b, X, st = O.nrt_basis_cell(l, VERTS, [a, 0, 0, a, 0, a], beta, None, fvec)
    assert st == 0
    Mw, bw = _coarse_element_matrix(a, beta, fvec)
    np.testing.assert_allclose(M, Mw, rtol=0, atol=1e-12 * np.abs(Mw).max())
    np.testing.assert_allclose(b, bw, rtol=0, atol=1e-12 * np.abs(bw).max())
    # the bases are the coarse functions: sigma of the edge bases, u of the face bases
    np.testing.assert_allclose(X[:12, :D.ne], coarse_edge_values(D), rtol=0, atol=1e-12)
    np.testing.assert_allclose(X[12:, D.ne:], coarse_face_values(D), rtol=0, atol=1e-12)
    assert np.abs(X[12:, :D.ne]).max() < 1e-11


@pytest.mark.parametrize("l", [1, 2])
def test_bases_solve_the_assembled_system(l):
    D = Dofs(l)
    coefs = rand_coefs(l, 11)
    M, b, X, st = O.nrt_basis_cell(l, VERTS, *coefs)
    assert st == 0
    A, F = O.nrt_assemble(l, VERTS, *coefs)
    Fc = O.nrt_curl_loads(l, VERTS)
    bnd = np.zeros(D.N, dtype=bool)
    for e, t, p in D.edges():
        bnd[e] = D.edge_on_boundary(t, p)
    for dof, t, p in D.faces():
        bnd[dof] = D.face_on_boundary(t, p)
    I = ~bnd
    for n in range(18):
        rhs = Fc[n] if n < 12 else np.zeros(D.N)
        r = (A @ X[n] - rhs)[I]
        assert np.abs(r).max() < 1e-12 * np.abs(A).max() * np.abs(X[n]).max()
    # boundary traces of the coarse functions
    Ie, If = coarse_edge_values(D), coarse_face_values(D)
    for n in range(12):
        np.testing.assert_allclose(X[n, :D.ne][bnd[:D.ne]], Ie[n][bnd[:D.ne]], atol=1e-15)
        assert np.abs(X[n, D.ne:][bnd[D.ne:]]).max() == 0
    for n in range(6):
        np.testing.assert_allclose(X[12 + n, D.ne:][bnd[D.ne:]], If[n][bnd[D.ne:]], atol=1e-15)
        assert np.abs(X[12 + n, :D.ne][bnd[:D.ne]]).max() == 0
    # the element matrix pairs the blocks of the assembled matrix
    # (ned_rt_basis.cc:850-948): edge bases by their edge part, face bases by
    # their face part; the load pairs the face part of face bases with (v, f)
    V = np.zeros((18, D.N))
    V[:12, :D.ne] = X[:12, :D.ne]
    V[12:, D.ne:] = X[12:, D.ne:]
    np.testing.assert_allclose(M, V @ A @ V.T, rtol=0, atol=1e-12 * np.abs(M).max())
    np.testing.assert_allclose(b[12:], V[12:] @ F, rtol=0, atol=1e-12 * np.abs(b).max())
    assert np.abs(b[:12]).max() == 0
    # the mixed blocks pair through the curl: M[face j, edge i] = (curl sigma_i, u_j)
    A0, _ = O.nrt_assemble(l, VERTS, [1, 0, 0, 1, 0, 1], 0.0, 1.0)
    Mf = A0[D.ne:, D.ne:]
    for i in range(12):
        cs = circulation_values(D, X[i, :D.ne])
        for j in range(6):
            want = X[12 + j, D.ne:] @ Mf @ cs
            assert M[12 + j, i] == pytest.approx(want, rel=1e-10, abs=1e-12 * np.abs(M).max())
            assert M[i, 12 + j] == pytest.approx(-want, rel=1e-10, abs=1e-12 * np.abs(M).max())


@pytest.mark.parametrize("l", [1, 2, 3])
def test_sparse_direct_solve_matches_dense(l):
    """nrt_basis_cell_sparse (SuperLU on the triplet assembly, the path that
    reaches l = 4) against the dense LU of nrt_basis_cell."""
    ainv, bq, rq, fq = rand_coefs(l, 60 + l)
    M, b, X = O.nrt_basis_cell_sparse(l, VERTS, ainv, bq, rq, fq)
    Mo, bo, Xo, st = O.nrt_basis_cell(l, VERTS, ainv, bq, rq, fq)
    assert st == 0
    np.testing.assert_allclose(M, Mo, rtol=0, atol=1e-12 * np.abs(Mo).max())
    np.testing.assert_allclose(b, bo, rtol=0, atol=1e-12 * np.abs(bo).max())
    np.testing.assert_allclose(X, Xo, rtol=0, atol=1e-10 * np.abs(Xo).max())


def test_triplet_assembly_equals_dense_assembly():
    ainv, bq, rq, fq = rand_coefs(2, 71)
    A, F = O.nrt_assemble(2, VERTS, ainv, bq, rq, fq)
    As, Fs = O.nrt_assemble_sparse(2, VERTS, ainv, bq, rq, fq)
    np.testing.assert_allclose(As.toarray(), A, rtol=0, atol=1e-14 * np.abs(A).max())
    np.testing.assert_array_equal(Fs, F)


def test_boundary_values_are_the_coarse_traces():
    for l in (1, 3):
        D = Dofs(l)
        G, bnd = O.nrt_boundary(l, VERTS)
        for e, t, p in D.edges():
            assert bnd[e] == D.edge_on_boundary(t, p)
        for f, t, p in D.faces():
            assert bnd[f] == D.face_on_boundary(t, p)
        E, Fv = coarse_edge_values(D), coarse_face_values(D)
        np.testing.assert_allclose(G[:12, :D.ne], np.where(bnd[:D.ne], E, 0.0), rtol=0, atol=1e-15)
        np.testing.assert_allclose(G[12:, D.ne:], np.where(bnd[D.ne:], Fv, 0.0), rtol=0, atol=1e-15)
        assert not G[:12, D.ne:].any() and not G[12:, :D.ne].any()


def test_rejects_bad_cells():
    v = VERTS.copy()
    v[7, 0] += 0.01
    ainv, bq, rq, fq = rand_coefs(1, 1)
    assert O.nrt_basis_cell(1, v, ainv, bq, rq, fq)[3] == 1
    assert O.nrt_basis_cell(4, VERTS, [1, 0, 0, 1, 0, 1], 1.0)[3] == 1
