"""The 2-D Q1 x Nedelec oracle (oracle/ned_oracle.c) against what the
mathematics fixes: discrete de Rham identities of the assembled blocks, closed
forms for constant coefficients, the element-matrix structure and the residual
of the interior system."""
import numpy as np
import pytest

from oracle import oracle as O


def box(x0, y0, hx, hy):
    return np.array([[x0, y0], [x0 + hx, y0], [x0, y0 + hy], [x0 + hx, y0 + hy]])


def incidence(l):
    """Node-to-edge incidence D: the line integral of grad p along each edge
    (tangent +x / +y) is p(end) - p(start)."""
    m = 1 << l
    ns, nh, nv, N = O.ned_sizes(l)
    D = np.zeros((nh + nv, ns))
    for j in range(m + 1):
        for i in range(m):
            e = j * m + i
            D[e, j * (m + 1) + i] -= 1
            D[e, j * (m + 1) + i + 1] += 1
    for j in range(m):
        for i in range(m + 1):
            e = nh + j * (m + 1) + i
            D[e, j * (m + 1) + i] -= 1
            D[e, (j + 1) * (m + 1) + i] += 1
    return D


def node_xy(l, v):
    m = 1 << l
    hx, hy = (v[1, 0] - v[0, 0]) / m, (v[2, 1] - v[0, 1]) / m
    i, j = np.meshgrid(np.arange(m + 1), np.arange(m + 1))
    return v[0, 0] + i.ravel() * hx, v[0, 1] + j.ravel() * hy


def edge_mid(l, v):
    """Midpoints and lengths of the horizontal then vertical edges."""
    m = 1 << l
    hx, hy = (v[1, 0] - v[0, 0]) / m, (v[2, 1] - v[0, 1]) / m
    i, j = np.meshgrid(np.arange(m), np.arange(m + 1))
    hxm, hym = v[0, 0] + (i.ravel() + 0.5) * hx, v[0, 1] + j.ravel() * hy
    i, j = np.meshgrid(np.arange(m + 1), np.arange(m))
    vxm, vym = v[0, 0] + i.ravel() * hx, v[0, 1] + (j.ravel() + 0.5) * hy
    return (hxm, hym, hx), (vxm, vym, hy)


def coarse_phi(n, x, y, v):
    X = (x - v[0, 0]) / (v[1, 0] - v[0, 0])
    Y = (y - v[0, 1]) / (v[2, 1] - v[0, 1])
    px = (1 - X) if (n & 1) == 0 else X
    py = (1 - Y) if (n >> 1) == 0 else Y
    return px * py


def interior_mask(l):
    m = 1 << l
    ns, nh, nv, N = O.ned_sizes(l)
    bnd = np.zeros(N, dtype=bool)
    i, j = np.meshgrid(np.arange(m + 1), np.arange(m + 1))
    bnd[:ns] = ((i == 0) | (j == 0) | (i == m) | (j == m)).ravel()
    i, j = np.meshgrid(np.arange(m), np.arange(m + 1))
    bnd[ns:ns + nh] = ((j == 0) | (j == m)).ravel()
    i, j = np.meshgrid(np.arange(m + 1), np.arange(m))
    bnd[ns + nh:] = ((i == 0) | (i == m)).ravel()
    return ~bnd


def rand_coefs(rng, l):
    nf = (1 << l) ** 2
    return (np.exp(rng.standard_normal(nf)), np.exp(0.5 * rng.standard_normal(nf)),
            np.exp(rng.standard_normal(nf)), rng.standard_normal((nf, 2)))


@pytest.mark.parametrize("l", [1, 2, 3])
def test_assembly_de_rham_identities(l):
    rng = np.random.default_rng(10 + l)
    v = box(0.3, -0.2, 0.7, 0.45)
    a, binv, r, f = rand_coefs(rng, l)
    ns = O.ned_sizes(l)[0]
    D = incidence(l)
    A0, _ = O.ned_assemble(l, v, a, binv, 0.0, f)
    Amass, _ = O.ned_assemble(l, v, 0.0, binv, r, f)
    Mned = Amass[ns:, ns:]
    # curl-curl kills discrete gradients
    assert np.abs(A0[ns:, ns:] @ D).max() < 1e-12 * np.abs(A0[ns:, ns:]).max()
    # (v, grad sigma) = (v, sum_e (D sigma)_e N_e): the Nedelec mass times D,
    # here with the reaction weight r as the mass weight
    Amr, _ = O.ned_assemble(l, v, 0.0, binv, 1.0, f)
    G = A0[ns:, :ns]
    assert np.abs(G - Amr[ns:, ns:] @ D).max() < 1e-13 * np.abs(G).max()
    # the (0,1) block is minus the transpose of the (1,0) block
    assert np.abs(A0[:ns, ns:] + G.T).max() < 1e-15 * np.abs(G).max() + 1e-16
    # the weighted Q1 mass matrix sums to the integral of B^-1
    m = 1 << l
    cell_area = (0.7 / m) * (0.45 / m)
    assert A0[:ns, :ns].sum() == pytest.approx(binv.sum() * cell_area, rel=1e-13)
    # reaction block is the weighted Nedelec mass: positive definite
    assert np.linalg.eigvalsh(0.5 * (Mned + Mned.T)).min() > 0


@pytest.mark.parametrize("l", [1, 2, 3])
def test_assembly_against_fields(l):
    """(v, f) against the interpolant of a constant field, and the curl-curl
    energy of a field with unit curl."""
    rng = np.random.default_rng(20 + l)
    v = box(-1.0, 2.0, 0.5, 0.8)
    a, binv, r, f = rand_coefs(rng, l)
    m = 1 << l
    ns, nh, nv, N = O.ned_sizes(l)
    A, F = O.ned_assemble(l, v, a, binv, 0.0, f)
    (hxm, hym, hx), (vxm, vym, hy) = edge_mid(l, v)
    area = hx * hy
    c = np.array([0.3, -1.7])
    dofs = np.concatenate([np.full(nh, c[0] * hx), np.full(nv, c[1] * hy)])
    assert F[ns:] @ dofs == pytest.approx(area * (f @ c).sum(), rel=1e-13)
    # w = (-y/2, x/2): curl w = 1; line integrals along the edges
    w = np.concatenate([-hym / 2 * hx, vxm / 2 * hy])
    assert w @ A[ns:, ns:] @ w == pytest.approx(area * a.sum(), rel=1e-12)


@pytest.mark.parametrize("l", [2, 3])
def test_edge_bases_are_coarse_nedelec_for_constant_a(l):
    """Constant a and no reaction: sigma = 0 and u = N_E solve the local
    problem exactly (div N_E = 0, curl N_E constant), for any B and f."""
    rng = np.random.default_rng(30 + l)
    v = box(0.1, 0.2, 0.6, 0.35)
    nf = (1 << l) ** 2
    binv = np.exp(rng.standard_normal(nf))
    f = rng.standard_normal((nf, 2))
    M, b, Psi, st = O.ned_basis_cell(l, v, 2.5, binv, 0.0, f)
    assert st == 0
    ns, nh, nv, N = O.ned_sizes(l)
    (hxm, hym, hx), (vxm, vym, hy) = edge_mid(l, v)
    Hx, Hy = 0.6, 0.35
    X = (vxm - 0.1) / Hx
    Y = (hym - 0.2) / Hy
    expect = np.zeros((4, N))
    expect[0, ns + nh:] = (1 - X) / Hy * hy   # N_0 = (0, (1-X)/Hy)
    expect[1, ns + nh:] = X / Hy * hy         # N_1 = (0, X/Hy)
    expect[2, ns:ns + nh] = (1 - Y) / Hx * hx  # N_2 = ((1-Y)/Hx, 0)
    expect[3, ns:ns + nh] = Y / Hx * hx        # N_3 = (Y/Hx, 0)
    assert np.abs(Psi[4:] - expect).max() < 1e-12
    # edge block: a |K| curl N_E curl N_F with curl N_E = s_E / |K|
    s = np.array([-1.0, 1.0, 1.0, -1.0])
    assert np.abs(M[4:, 4:] - 2.5 / (Hx * Hy) * np.outer(s, s)).max() < 1e-10


@pytest.mark.parametrize("l", [2, 3, 4])
def test_element_matrix_structure_and_residual(l):
    rng = np.random.default_rng(40 + l)
    v = box(0.25, 0.5, 0.25, 0.25)
    a, binv, r, f = rand_coefs(rng, l)
    M, b, Psi, st = O.ned_basis_cell(l, v, a, binv, r, f)
    assert st == 0
    assert np.abs(M[:4, 4:] + M[4:, :4].T).max() < 1e-12 * np.abs(M).max()
    for blk in (M[:4, :4], M[4:, 4:]):
        assert np.abs(blk - blk.T).max() < 1e-12 * np.abs(blk).max()
        assert np.linalg.eigvalsh(0.5 * (blk + blk.T)).min() > -1e-12 * np.abs(blk).max()
    assert np.abs(b[:4]).max() == 0.0
    # the interior residual: A psi = source on the interior dofs
    ns, nh, nv, N = O.ned_sizes(l)
    A, F = O.ned_assemble(l, v, a, binv, r, f)
    Amr, _ = O.ned_assemble(l, v, 0.0, binv, 1.0, f)
    D = incidence(l)
    x, y = node_xy(l, v)
    inner = interior_mask(l)
    for n in range(8):
        src = np.zeros(N)
        if n < 4:  # (v, grad phi_n): grad phi_n is D phi_n in the Nedelec space
            src[ns:] = Amr[ns:, ns:] @ (D @ coarse_phi(n, x, y, v))
        res = (A @ Psi[n] - src)[inner]
        assert np.abs(res).max() < 1e-11 * (np.abs(A).max() * np.abs(Psi[n]).max() + 1.0)
        # the boundary values
        if n < 4:
            assert np.abs(Psi[n, :ns][~inner[:ns]] - coarse_phi(n, x, y, v)[~inner[:ns]]).max() < 1e-15
    # b: (psi_u, F) for the edge bases
    assert np.abs(b[4:] - Psi[4:, ns:] @ F[ns:]).max() < 1e-13 * (np.abs(b).max() + 1)


def test_no_full_rhs_vertex_bases_see_only_the_lifting():
    """Without the source, the vertex bases sum to the lifting of phi = 1:
    sigma = 1 on the boundary, u = 0 there; interior from the solve."""
    l = 2
    v = box(0.0, 0.0, 1.0, 1.0)
    M, b, Psi, st = O.ned_basis_cell(l, v, 1.0, 1.0, 0.0, (0.0, 0.0), full_rhs=False)
    assert st == 0
    ns = O.ned_sizes(l)[0]
    inner = interior_mask(l)
    tot = Psi[:4].sum(axis=0)
    assert np.abs(tot[:ns][~inner[:ns]] - 1.0).max() < 1e-15
    assert np.abs(tot[ns:][~inner[ns:]]).max() < 1e-15


@pytest.mark.parametrize("l", [1, 2, 3, 4])
def test_boundary_traces_of_all_bases(l):
    """The lifting data of q_ned_basis.cc:223-357 on every boundary dof, for
    rough coefficients (the traces do not depend on them): a vertex basis n
    carries sigma = phi_n at the boundary nodes and, on every boundary edge,
    the line integral of grad phi_n, i.e. phi_n(end) - phi_n(start) along the
    edge's +x / +y tangent; an edge basis E carries sigma = 0 and the line
    integral of the coarse Nedelec function N_E (deal.II line order
    x = x0, x = x1, y = y0, y = y1)."""
    rng = np.random.default_rng(60 + l)
    x0, y0, Hx, Hy = -0.3, 0.45, 0.7, 0.4
    v = box(x0, y0, Hx, Hy)
    a, binv, r, f = rand_coefs(rng, l)
    M, b, Psi, st = O.ned_basis_cell(l, v, a, binv, r, f)
    assert st == 0
    m = 1 << l
    ns, nh, nv, N = O.ned_sizes(l)
    bnd = ~interior_mask(l)
    x, y = node_xy(l, v)
    hx, hy = Hx / m, Hy / m
    for n in range(4):
        phi = coarse_phi(n, x, y, v)
        want = np.zeros(N)
        want[:ns] = phi
        for j in range(m + 1):          # horizontal edges (i,j)-(i+1,j)
            for i in range(m):
                want[ns + j * m + i] = phi[j * (m + 1) + i + 1] - phi[j * (m + 1) + i]
        for j in range(m):              # vertical edges (i,j)-(i,j+1)
            for i in range(m + 1):
                want[ns + nh + j * (m + 1) + i] = phi[(j + 1) * (m + 1) + i] - phi[j * (m + 1) + i]
        assert np.abs(Psi[n][bnd] - want[bnd]).max() < 1e-15, n
        # the nodal part of the trace is grad phi_n's line integral, not its value:
        # a sign or end swap breaks this
        assert np.abs(want[ns:][bnd[ns:]]).max() > 0.1 / m
    # edge bases: sigma = 0, and the tangential line integrals of N_E
    ii, jj = np.meshgrid(np.arange(m), np.arange(m + 1))          # horizontal edges
    Yh = (jj.ravel() * hy) / Hy
    ii, jj = np.meshgrid(np.arange(m + 1), np.arange(m))          # vertical edges
    Xv = (ii.ravel() * hx) / Hx
    want = np.zeros((4, N))
    want[0, ns + nh:] = (1 - Xv) / Hy * hy
    want[1, ns + nh:] = Xv / Hy * hy
    want[2, ns:ns + nh] = (1 - Yh) / Hx * hx
    want[3, ns:ns + nh] = Yh / Hx * hx
    for e in range(4):
        assert np.abs(Psi[4 + e][bnd] - want[e][bnd]).max() < 1e-15, e
        # the trace is non-zero on exactly one side of the coarse box
        assert np.count_nonzero(np.abs(want[e][bnd]) > 1e-14) == m
