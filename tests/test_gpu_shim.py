"""The C++ QBasis shim in a QMultiscale-shaped driver (map of cells filled by
copy-insertion, run() per cell, coarse solve, set_global_weights), checked
against the oracle.  Runs the compiled driver as a child process."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(REPO, "mpi-msfec_amd", "cpp", "q_multiscale_driver")


@pytest.mark.parametrize("dim,l,nc", [(3, 2, 4), (2, 3, 8)])
def test_multiscale_driver(oracle_mod, dim, l, nc):
    print(f"case shim_driver dim={dim} l={l} nc={nc}", flush=True)
    from msfec_amd import fields as F
    assert os.path.exists(DRIVER), "driver not built"
    coef = F.lognormal_coefficient(dim, nc, l, 1.0, 0.0, seed=9)
    ncell, nv, nn = nc ** dim, 1 << dim, ((1 << l) + 1) ** dim
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as fh:
            fh.write(np.array([dim, l, nc], dtype=np.int32).tobytes())
            fh.write(np.ascontiguousarray(coef, dtype=np.float64).tobytes())
        r = subprocess.run([DRIVER, fin, fout], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        raw = open(fout, "rb").read()
    d = np.frombuffer(raw[:-4], dtype=np.float64)
    o = 0
    M = d[o:o + ncell * nv * nv].reshape(ncell, nv, nv); o += ncell * nv * nv
    b = d[o:o + ncell * nv].reshape(ncell, nv); o += ncell * nv
    w = d[o:o + ncell * nv].reshape(ncell, nv); o += ncell * nv
    uf = d[o:o + ncell * nn].reshape(ncell, nn)
    verts = F.coarse_vertices(dim, nc)
    ref = oracle_mod.basis_batch(oracle_mod.make_problem(dim, l), verts, coef=coef)
    assert np.abs(M - ref["M"]).max() < 1e-11 * np.abs(ref["M"]).max()
    assert np.abs(b - ref["b"]).max() < 1e-11 * np.abs(ref["b"]).max()
    u, _ = oracle_mod.coarse_solve(dim, nc, ref["M"], ref["b"], rtol=1e-14)
    ufo = oracle_mod.reconstruct(dim, nc, l, u, ref["Psi"])
    assert np.abs(uf - ufo).max() < 1e-9 * np.abs(ufo).max()


def read_vtu(path):
    """(points (n, 3), connectivity (cells, k), types, {name: values}) of an
    ASCII VTU file as the shim writes it."""
    import xml.etree.ElementTree as ET
    piece = ET.parse(path).getroot().find("UnstructuredGrid").find("Piece")
    arr = lambda e, t: np.array(e.text.split(), dtype=t)
    pts = arr(piece.find("Points").find("DataArray"), np.float64).reshape(-1, 3)
    cells = {e.get("Name"): e for e in piece.find("Cells").findall("DataArray")}
    types = arr(cells["types"], np.int64)
    conn = arr(cells["connectivity"], np.int64).reshape(len(types), -1)
    assert int(piece.get("NumberOfPoints")) == len(pts)
    assert int(piece.get("NumberOfCells")) == len(types)
    fields = {e.get("Name"): arr(e, np.float64) for e in piece.find("PointData").findall("DataArray")}
    return pts, conn, types, fields


@pytest.mark.parametrize("dim,l,nc", [(2, 2, 3), (3, 1, 2)])
def test_multiscale_driver_vtu_output(oracle_mod, dim, l, nc):
    """output_basis (first cell) and output_global_solution_in_cell (every
    cell): file names, fine grid and the written values."""
    print(f"case shim_driver_vtu dim={dim} l={l} nc={nc}", flush=True)
    from msfec_amd import fields as F
    coef = F.lognormal_coefficient(dim, nc, l, 1.0, 0.0, seed=21)
    ncell, nv, S = nc ** dim, 1 << dim, (1 << l) + 1
    nn = S ** dim
    with tempfile.TemporaryDirectory() as td:
        fin, fout, od = os.path.join(td, "in.bin"), os.path.join(td, "out.bin"), os.path.join(td, "vtu")
        os.mkdir(od)
        with open(fin, "wb") as fh:
            fh.write(np.array([dim, l, nc], dtype=np.int32).tobytes())
            fh.write(np.ascontiguousarray(coef, dtype=np.float64).tobytes())
        r = subprocess.run([DRIVER, fin, fout, od], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        d = np.frombuffer(open(fout, "rb").read()[:-4], dtype=np.float64)
        uf = d[ncell * (nv * nv + 2 * nv):].reshape(ncell, nn)
        names = sorted(os.listdir(od))
        expect = sorted([f"basis_q.00000.cell-0.index-{i:02d}.vtu" for i in range(nv)] +
                        [f"solution.00000.cell-{c}.vtu" for c in range(ncell)])
        assert names == expect
        verts = F.coarse_vertices(dim, nc)
        ref = oracle_mod.basis_batch(oracle_mod.make_problem(dim, l), verts, coef=coef)
        for i in range(nv):
            pts, conn, types, fl = read_vtu(os.path.join(od, f"basis_q.00000.cell-0.index-{i:02d}.vtu"))
            assert len(pts) == nn and np.all(types == (9 if dim == 2 else 12))
            assert np.abs(fl["u"] - ref["Psi"][0, i]).max() < 1e-11
            # basis i is 1 at coarse vertex i (delta property)
            corner = verts[0, i]
            k = np.argmin(np.abs(pts[:, :dim] - corner).sum(axis=1))
            assert abs(fl["u"][k] - 1.0) < 1e-14
        c = ncell - 1
        pts, conn, types, fl = read_vtu(os.path.join(od, f"solution.00000.cell-{c}.vtu"))
        np.testing.assert_allclose(fl["u"], uf[c], rtol=0, atol=1e-15 * np.abs(uf[c]).max())
        lo, hi = verts[c].min(axis=0), verts[c].max(axis=0)
        assert np.allclose(pts[:, :dim].min(axis=0), lo) and np.allclose(pts[:, :dim].max(axis=0), hi)
        # every cell of the connectivity is an axis-aligned fine box of the right size
        h = (hi - lo) / (S - 1)
        ext = pts[conn][:, :, :dim].max(axis=1) - pts[conn][:, :, :dim].min(axis=1)
        assert np.allclose(ext, h)
        # VTK vertex order: the bottom face counter-clockwise (shoelace area > 0)
        q = pts[conn[:, :4]][:, :, :2]
        area = 0.5 * (q[:, :, 0] * np.roll(q[:, :, 1], -1, axis=1)
                      - np.roll(q[:, :, 0], -1, axis=1) * q[:, :, 1]).sum(axis=1)
        assert np.allclose(area, h[0] * h[1])


NED_DRIVER = os.path.join(REPO, "mpi-msfec_amd", "cpp", "q_ned_driver")


def test_ned_driver(oracle_mod):
    """The C++ QNedBasis shim: map of cells, run(), getters and
    set_global_weights (sigma block from the vertex bases, u block from the
    edge bases, q_ned_basis.cc:1158-1182) against the oracle."""
    print("case ned_shim_driver l=3 nc=4", flush=True)
    from msfec_amd import fields as F
    assert os.path.exists(NED_DRIVER), "driver not built"
    l, nc = 3, 4
    a = F.lognormal_coefficient(2, nc, l, 1.0, 0.0, seed=13)
    ncell = nc * nc
    ns, nh, nv, N = oracle_mod.ned_sizes(l)
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as fh:
            fh.write(np.array([l, nc], dtype=np.int32).tobytes())
            fh.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
        r = subprocess.run([NED_DRIVER, fin, fout], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        d = np.frombuffer(open(fout, "rb").read(), dtype=np.float64)
    o = 0
    M = d[o:o + ncell * 64].reshape(ncell, 8, 8); o += ncell * 64
    b = d[o:o + ncell * 8].reshape(ncell, 8); o += ncell * 8
    P = d[o:o + ncell * 8 * N].reshape(ncell, 8, N); o += ncell * 8 * N
    sol = d[o:o + ncell * N].reshape(ncell, N)
    verts = F.coarse_vertices(2, nc)
    ref = oracle_mod.ned_basis_batch(l, verts, a, 1.0, 0.0, (1.0, 1.0))
    for i in (0, 4):
        for j in (0, 4):
            blk = ref["M"][:, i:i + 4, j:j + 4]
            assert np.abs(M[:, i:i + 4, j:j + 4] - blk).max() < 1e-10 * np.abs(blk).max()
    assert np.abs(b - ref["b"]).max() < 1e-10 * np.abs(ref["b"]).max()
    assert np.abs(P[:, :, ns:] - ref["Psi"][:, :, ns:]).max() < 1e-11 * np.abs(ref["Psi"]).max()
    w = np.arange(1.0, 9.0)
    expect = np.concatenate([np.einsum("i,cid->cd", w[:4], P[:, :4, :ns]),
                             np.einsum("i,cid->cd", w[4:], P[:, 4:, ns:])], axis=1)
    assert np.abs(sol - expect).max() < 1e-13 * np.abs(expect).max()


RT_DRIVER = os.path.join(REPO, "mpi-msfec_amd", "cpp", "rt_dq_driver")


def test_rt_dq_driver(oracle_mod):
    """The C++ RTDQBasis shim: map of cells, run(), getters and
    set_global_weights (fluxes sum_i w_i sigma_i, cell values w_4,
    rt_dq_basis.cc:1011-1036) against the oracle."""
    print("case rt_dq_shim_driver l=3 nc=4", flush=True)
    from msfec_amd import fields as F
    assert os.path.exists(RT_DRIVER), "driver not built"
    l, nc = 3, 4
    a = F.lognormal_coefficient(2, nc, l, 1.0, 0.0, seed=17)
    ncell = nc * nc
    nv, ne, nfc = oracle_mod.rt_sizes(l)
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as fh:
            fh.write(np.array([l, nc], dtype=np.int32).tobytes())
            fh.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
        r = subprocess.run([RT_DRIVER, fin, fout], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        d = np.frombuffer(open(fout, "rb").read(), dtype=np.float64)
    o = 0
    M = d[o:o + ncell * 25].reshape(ncell, 5, 5); o += ncell * 25
    b = d[o:o + ncell * 5].reshape(ncell, 5); o += ncell * 5
    S = d[o:o + ncell * 4 * ne].reshape(ncell, 4, ne); o += ncell * 4 * ne
    sol = d[o:o + ncell * (ne + nfc)].reshape(ncell, ne + nfc)
    verts = F.coarse_vertices(2, nc)
    ref = oracle_mod.rt_basis_batch(l, verts, a)
    assert np.abs(M - ref["M"]).max() < 1e-11 * np.abs(ref["M"]).max()
    assert np.abs(b - ref["b"]).max() < 1e-13 * np.abs(ref["b"]).max()
    assert np.abs(S - ref["Sigma"]).max() < 1e-11 * np.abs(ref["Sigma"]).max()
    w = np.arange(1.0, 6.0)
    expect = np.concatenate([np.einsum("i,cie->ce", w[:4], S), np.full((ncell, nfc), w[4])],
                            axis=1)
    assert np.abs(sol - expect).max() < 1e-13 * np.abs(expect).max()


def read_vtu_cells(path):
    """{name: (cells, ncomp) values} of the CellData of an ASCII VTU file."""
    import xml.etree.ElementTree as ET
    piece = ET.parse(path).getroot().find("UnstructuredGrid").find("Piece")
    out = {}
    for e in piece.find("CellData").findall("DataArray"):
        k = int(e.get("NumberOfComponents", "1"))
        out[e.get("Name")] = np.array(e.text.split(), dtype=np.float64).reshape(-1, k)
    return out


def test_rt_dq_driver_vtu_output(oracle_mod):
    """RTDQBasis output_basis (first cell) and output_global_solution_in_cell:
    file names, the flux field at the fine cell centres and u."""
    print("case rt_dq_shim_driver_vtu l=2 nc=3", flush=True)
    from msfec_amd import fields as F
    l, nc = 2, 3
    m = 1 << l
    a = F.lognormal_coefficient(2, nc, l, 1.0, 0.0, seed=23)
    ncell = nc * nc
    nv, ne, nfc = oracle_mod.rt_sizes(l)
    with tempfile.TemporaryDirectory() as td:
        fin, fout, od = os.path.join(td, "in.bin"), os.path.join(td, "out.bin"), os.path.join(td, "vtu")
        os.mkdir(od)
        with open(fin, "wb") as fh:
            fh.write(np.array([l, nc], dtype=np.int32).tobytes())
            fh.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
        r = subprocess.run([RT_DRIVER, fin, fout, od], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        names = sorted(os.listdir(od))
        expect = sorted([f"basis_rt-dq.div.00000.cell-0.index-{i:02d}.vtu" for i in range(4)] +
                        [f"solution.00000.cell-{c}.vtu" for c in range(ncell)])
        assert names == expect
        verts = F.coarse_vertices(2, nc)
        ref = oracle_mod.rt_basis_batch(l, verts, a)
        h = 1.0 / nc / m
        for i in range(4):
            cf = read_vtu_cells(os.path.join(od, f"basis_rt-dq.div.00000.cell-0.index-{i:02d}.vtu"))
            S = ref["Sigma"][0, i]
            sx = 0.5 * (S[:nv].reshape(m, m + 1)[:, :-1] + S[:nv].reshape(m, m + 1)[:, 1:]) / h
            sy = 0.5 * (S[nv:].reshape(m + 1, m)[:-1] + S[nv:].reshape(m + 1, m)[1:]) / h
            np.testing.assert_allclose(cf["sigma"][:, 0], sx.ravel(), rtol=1e-11, atol=1e-11)
            np.testing.assert_allclose(cf["sigma"][:, 1], sy.ravel(), rtol=1e-11, atol=1e-11)
            assert np.all(cf["sigma"][:, 2] == 0.0) and np.all(cf["u"] == 1.0)
        cf = read_vtu_cells(os.path.join(od, f"solution.00000.cell-{ncell - 1}.vtu"))
        np.testing.assert_array_equal(cf["u"][:, 0], np.full(nfc, 5.0))  # w_4 = 5


def test_ned_driver_vtu_output(oracle_mod):
    """QNedBasis output_basis (first cell: 4 .h1 and 4 .curl files) and
    output_global_solution_in_cell: sigma at the fine nodes, u at the fine cell
    centres."""
    print("case ned_shim_driver_vtu l=2 nc=2", flush=True)
    from msfec_amd import fields as F
    l, nc = 2, 2
    m = 1 << l
    a = F.lognormal_coefficient(2, nc, l, 1.0, 0.0, seed=29)
    ncell = nc * nc
    ns, nh, nv, N = oracle_mod.ned_sizes(l)
    with tempfile.TemporaryDirectory() as td:
        fin, fout, od = os.path.join(td, "in.bin"), os.path.join(td, "out.bin"), os.path.join(td, "vtu")
        os.mkdir(od)
        with open(fin, "wb") as fh:
            fh.write(np.array([l, nc], dtype=np.int32).tobytes())
            fh.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
        r = subprocess.run([NED_DRIVER, fin, fout, od], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        d = np.frombuffer(open(fout, "rb").read(), dtype=np.float64)
        P = d[ncell * 72:ncell * 72 + ncell * 8 * N].reshape(ncell, 8, N)
        names = sorted(os.listdir(od))
        expect = sorted([f"basis_q-ned.{k}.00000.cell-0.index-{i:02d}.vtu"
                         for k in ("h1", "curl") for i in range(4)] +
                        [f"solution.00000.cell-{c}.vtu" for c in range(ncell)])
        assert names == expect
        h = 1.0 / nc / m
        for b in range(8):
            kind, i = ("h1", b) if b < 4 else ("curl", b - 4)
            path = os.path.join(od, f"basis_q-ned.{kind}.00000.cell-0.index-{i:02d}.vtu")
            pts, conn, types, fl = read_vtu(path)
            np.testing.assert_array_equal(fl["sigma"], P[0, b, :ns])
            u = read_vtu_cells(path)["u"]
            uh = P[0, b, ns:ns + nh].reshape(m + 1, m)
            uv = P[0, b, ns + nh:].reshape(m, m + 1)
            np.testing.assert_allclose(u[:, 0], (0.5 * (uh[:-1] + uh[1:]) / h).ravel(), rtol=1e-14, atol=1e-15)
            np.testing.assert_allclose(u[:, 1], (0.5 * (uv[:, :-1] + uv[:, 1:]) / h).ravel(), rtol=1e-14, atol=1e-15)


def test_multiscale_driver_uneven_levels(oracle_mod):
    """The C++ map loop with cells of mixed refinement (QBasis(..., level),
    QBatch on msfec_qu_*): element data per level against the oracle, and the
    fine solution of every cell against the oracle's coarse solve and
    reconstruction."""
    print("case shim_driver uneven", flush=True)
    from msfec_amd import fields as F
    assert os.path.exists(DRIVER), "driver not built"
    dim, nc = 3, 4
    cen = F.inclusion_centres(3, 2, seed=81)
    lv = F.uneven_levels(3, nc, cen, seed=82, coarse=(2, 3))
    blocks = F.inclusion_blocks(3, nc, lv, cen, 0.08, 1e3)
    ncell, nv = nc ** dim, 1 << dim
    nns = [((1 << int(l)) + 1) ** dim for l in lv]
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as fh:
            fh.write(np.array([dim, 0, nc], dtype=np.int32).tobytes())
            fh.write(np.asarray(lv, dtype=np.int32).tobytes())
            fh.write(np.concatenate([np.ravel(x) for x in blocks]).astype(np.float64).tobytes())
        r = subprocess.run([DRIVER, fin, fout], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        raw = open(fout, "rb").read()
    d = np.frombuffer(raw[:-4], dtype=np.float64)
    o = 0
    M = d[o:o + ncell * nv * nv].reshape(ncell, nv, nv); o += ncell * nv * nv
    b = d[o:o + ncell * nv].reshape(ncell, nv); o += ncell * nv
    w = d[o:o + ncell * nv].reshape(ncell, nv); o += ncell * nv
    uf = []
    for nn in nns:
        uf.append(d[o:o + nn]); o += nn
    assert o == d.size
    verts = F.coarse_vertices(dim, nc)
    Mo, bo, Po = np.zeros_like(M), np.zeros_like(b), [None] * ncell
    for l in sorted(set(lv.tolist())):
        idx = np.nonzero(lv == l)[0]
        ref = oracle_mod.basis_batch(oracle_mod.make_problem(dim, int(l)), verts[idx],
                                     coef=np.stack([blocks[c] for c in idx]),
                                     solver=oracle_mod.SOLVER_DIRECT)
        Mo[idx], bo[idx] = ref["M"], ref["b"]
        for k, c in enumerate(idx):
            Po[c] = ref["Psi"][k]
    assert np.abs(M - Mo).max() < 1e-10 * np.abs(Mo).max()
    assert np.abs(b - bo).max() < 1e-10 * np.abs(bo).max()
    u, _ = oracle_mod.coarse_solve(dim, nc, Mo, bo, rtol=1e-14)
    for c in range(ncell):
        ix, iy, iz = c % nc, (c // nc) % nc, c // (nc * nc)
        wc = np.array([u[(ix + (k & 1)) + (nc + 1) * ((iy + ((k >> 1) & 1)) + (nc + 1) * (iz + (k >> 2)))]
                       for k in range(nv)])
        assert np.abs(w[c] - wc).max() < 1e-9 * np.abs(u).max()
        ufo = wc @ Po[c]
        assert np.abs(uf[c] - ufo).max() < 1e-9 * np.abs(ufo).max(), c


NRT_DRIVER = os.path.join(REPO, "mpi-msfec_amd", "cpp", "ned_rt_driver")


def nrt_driver_inputs(rng, l, ncell):
    """Seeded SPD A^-1 (6 entries), B and f at the 8 Gauss points of every fine cell."""
    nq = (1 << l) ** 3
    Lm = rng.normal(size=(ncell, nq, 8, 3, 3)) * 0.3 + np.eye(3)
    G = Lm @ np.swapaxes(Lm, -1, -2) + 0.1 * np.eye(3)
    ainv = np.stack([G[..., 0, 0], G[..., 0, 1], G[..., 0, 2], G[..., 1, 1], G[..., 1, 2],
                     G[..., 2, 2]], -1)
    bq = np.exp(rng.normal(size=(ncell, nq, 8)))
    fq = rng.normal(size=(ncell, nq, 8, 3))
    return ainv, bq, fq


def run_nrt_driver(l, nc, ainv, bq, fq, td, outdir=None):
    fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
    with open(fin, "wb") as fh:
        fh.write(np.array([l, nc], dtype=np.int32).tobytes())
        for x in (ainv, bq, fq):
            fh.write(np.ascontiguousarray(x, dtype=np.float64).tobytes())
    args = [NRT_DRIVER, fin, fout] + ([outdir] if outdir else [])
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return np.frombuffer(open(fout, "rb").read(), dtype=np.float64)


def unit_cube_boxes(oracle_mod, nc):
    h = 1.0 / nc
    return np.array([oracle_mod.box_vertices_3d([cx * h, cy * h, cz * h],
                                                [(cx + 1) * h, (cy + 1) * h, (cz + 1) * h])
                     for cz in range(nc) for cy in range(nc) for cx in range(nc)])


def test_ned_rt_driver(oracle_mod):
    """The C++ NedRTBasis shim: map of cells, run(), getters and
    set_global_weights (sigma = sum_{i<12} w_i sigma_i on the edges, u =
    sum_{i>=12} w_i u_i on the faces, ned_rt_basis.cc:1156-1181) against the
    oracle."""
    print("case ned_rt_shim_driver l=2 nc=2", flush=True)
    assert os.path.exists(NRT_DRIVER), "driver not built"
    l, nc = 2, 2
    ncell = nc ** 3
    m, nx, ne, nf1, nf, N = oracle_mod.nrt_sizes(l)
    ainv, bq, fq = nrt_driver_inputs(np.random.default_rng(31), l, ncell)
    with tempfile.TemporaryDirectory() as td:
        d = run_nrt_driver(l, nc, ainv, bq, fq, td)
    o = 0
    M = d[o:o + ncell * 324].reshape(ncell, 18, 18); o += ncell * 324
    b = d[o:o + ncell * 18].reshape(ncell, 18); o += ncell * 18
    X = d[o:o + ncell * 18 * N].reshape(ncell, 18, N); o += ncell * 18 * N
    sol = d[o:o + ncell * N].reshape(ncell, N); o += ncell * N
    assert o == d.size
    verts = unit_cube_boxes(oracle_mod, nc)
    ref = oracle_mod.nrt_basis_batch(l, verts, ainv, bq, None, fq, n_threads=8)
    assert (ref["status"] == 0).all()
    assert np.abs(M - ref["M"]).max() < 1e-11 * np.abs(ref["M"]).max()
    assert np.abs(b - ref["b"]).max() < 1e-11 * np.abs(ref["b"]).max()
    assert np.abs(X - ref["X"]).max() < 1e-11 * np.abs(ref["X"]).max()
    w = np.arange(1.0, 19.0)
    expect = np.concatenate([np.einsum("i,cie->ce", w[:12], X[:, :12, :ne]),
                             np.einsum("i,cie->ce", w[12:], X[:, 12:, ne:])], axis=1)
    assert np.abs(sol - expect).max() <= 1e-13 * np.abs(expect).max()


def test_ned_rt_driver_minres_path(oracle_mod):
    """The NedRTBasis shim at l = 3, where the library solves by MINRES,
    against the oracle's sparse direct solve."""
    print("case ned_rt_shim_driver l=3 nc=2", flush=True)
    assert os.path.exists(NRT_DRIVER), "driver not built"
    l, nc = 3, 2
    ncell = nc ** 3
    m, nx, ne, nf1, nf, N = oracle_mod.nrt_sizes(l)
    ainv, bq, fq = nrt_driver_inputs(np.random.default_rng(43), l, ncell)
    with tempfile.TemporaryDirectory() as td:
        d = run_nrt_driver(l, nc, ainv, bq, fq, td)
    M = d[:ncell * 324].reshape(ncell, 18, 18)
    X = d[ncell * 342:ncell * 342 + ncell * 18 * N].reshape(ncell, 18, N)
    verts = unit_cube_boxes(oracle_mod, nc)
    for c in range(ncell):
        Ms, bs, Xs = oracle_mod.nrt_basis_cell_sparse(l, verts[c], ainv[c], bq[c], None, fq[c])
        assert np.abs(M[c] - Ms).max() <= 1e-8 * np.abs(Ms).max()
        assert np.abs(X[c] - Xs).max() <= 1e-8 * np.abs(Xs).max()


def nrt_cell_centre_fields(v, l, h):
    """Nedelec field (edge line integrals averaged over the 4 parallel edges, over
    the edge length) and Raviart-Thomas field (face fluxes averaged over the 2
    opposite faces, over the face area) at the fine cell centres of a cube of
    fine width h, (m^3, 3) each, cells x fastest."""
    m = 1 << l
    nx, nf1 = m * (m + 1) ** 2, m * m * (m + 1)
    ex = v[:nx].reshape(m + 1, m + 1, m)             # [k, j, i]
    ey = v[nx:2 * nx].reshape(m + 1, m, m + 1)
    ez = v[2 * nx:3 * nx].reshape(m, m + 1, m + 1)
    f = v[3 * nx:]
    fx = f[:nf1].reshape(m, m, m + 1)
    fy = f[nf1:2 * nf1].reshape(m, m + 1, m)
    fz = f[2 * nf1:].reshape(m + 1, m, m)
    sx = (ex[:-1, :-1] + ex[1:, :-1] + ex[:-1, 1:] + ex[1:, 1:]) / (4 * h)
    sy = (ey[:-1, :, :-1] + ey[1:, :, :-1] + ey[:-1, :, 1:] + ey[1:, :, 1:]) / (4 * h)
    sz = (ez[:, :-1, :-1] + ez[:, 1:, :-1] + ez[:, :-1, 1:] + ez[:, 1:, 1:]) / (4 * h)
    ux = (fx[:, :, :-1] + fx[:, :, 1:]) / (2 * h * h)
    uy = (fy[:, :-1] + fy[:, 1:]) / (2 * h * h)
    uz = (fz[:-1] + fz[1:]) / (2 * h * h)
    sig = np.stack([sx.ravel(), sy.ravel(), sz.ravel()], 1)
    u = np.stack([ux.ravel(), uy.ravel(), uz.ravel()], 1)
    return sig, u


def test_ned_rt_driver_vtu_output(oracle_mod):
    """NedRTBasis output_basis (first cell: 12 .curl and 6 .div files) and
    output_global_solution_in_cell: file names, the Nedelec and Raviart-Thomas
    fields at the fine cell centres."""
    print("case ned_rt_shim_driver_vtu l=1 nc=2", flush=True)
    l, nc = 1, 2
    ncell = nc ** 3
    m, nx, ne, nf1, nf, N = oracle_mod.nrt_sizes(l)
    ainv, bq, fq = nrt_driver_inputs(np.random.default_rng(37), l, ncell)
    with tempfile.TemporaryDirectory() as td:
        od = os.path.join(td, "vtu")
        os.mkdir(od)
        d = run_nrt_driver(l, nc, ainv, bq, fq, td, od)
        names = sorted(os.listdir(od))
        expect = sorted([f"basis_ned-rt.curl.00000.cell-0.index-{i:02d}.vtu" for i in range(12)] +
                        [f"basis_ned-rt.div.00000.cell-0.index-{i:02d}.vtu" for i in range(6)] +
                        [f"solution.00000.cell-{c}.vtu" for c in range(ncell)])
        assert names == expect
        verts = unit_cube_boxes(oracle_mod, nc)
        ref = oracle_mod.nrt_basis_batch(l, verts[:1], ainv[:1], bq[:1], None, fq[:1])
        h = 1.0 / nc / m
        for i in range(18):
            kind, j = ("curl", i) if i < 12 else ("div", i - 12)
            cf = read_vtu_cells(os.path.join(od, f"basis_ned-rt.{kind}.00000.cell-0.index-{j:02d}.vtu"))
            sig, u = nrt_cell_centre_fields(ref["X"][0, i], l, h)
            scale = max(np.abs(sig).max(), np.abs(u).max())
            np.testing.assert_allclose(cf["sigma"], sig, rtol=0, atol=1e-10 * scale)
            np.testing.assert_allclose(cf["u"], u, rtol=0, atol=1e-10 * scale)
        sol = d[-ncell * N:].reshape(ncell, N)
        c = ncell - 1
        cf = read_vtu_cells(os.path.join(od, f"solution.00000.cell-{c}.vtu"))
        sig, u = nrt_cell_centre_fields(sol[c], l, h)
        scale = max(np.abs(sig).max(), np.abs(u).max())
        np.testing.assert_allclose(cf["sigma"], sig, rtol=0, atol=1e-10 * scale)
        np.testing.assert_allclose(cf["u"], u, rtol=0, atol=1e-10 * scale)


RT3_DRIVER = os.path.join(REPO, "mpi-msfec_amd", "cpp", "rt_dq3_driver")


def run_rt3_driver(l, nc, ainv, fq, td, outdir=None):
    fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
    with open(fin, "wb") as fh:
        fh.write(np.array([l, nc], dtype=np.int32).tobytes())
        for x in (ainv, fq):
            fh.write(np.ascontiguousarray(x, dtype=np.float64).tobytes())
    args = [RT3_DRIVER, fin, fout] + ([outdir] if outdir else [])
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return np.frombuffer(open(fout, "rb").read(), dtype=np.float64)


def test_rt_dq3_driver(oracle_mod):
    """The C++ RTDQ3Basis shim (the reference's 3-D RTDQBasis surface): map of
    cells, run(), getters, set_global_weights (fluxes sum_{i<6} w_i sigma_i,
    u = w_6, rt_dq_basis.cc:1011-1036) against the oracle, and the VTU files
    of output_basis (6 .div files of the first cell) and
    output_global_solution_in_cell (flux at the fine cell centres, u)."""
    print("case rt_dq3_shim_driver l=2 nc=2", flush=True)
    assert os.path.exists(RT3_DRIVER), "driver not built"
    l, nc = 2, 2
    ncell = nc ** 3
    m, nf, nfc, N = oracle_mod.rt3_sizes(l)
    ainv, _, _ = nrt_driver_inputs(np.random.default_rng(41), l, ncell)
    fq = np.random.default_rng(42).normal(size=(ncell, nfc, 8))
    with tempfile.TemporaryDirectory() as td:
        od = os.path.join(td, "vtu")
        os.mkdir(od)
        d = run_rt3_driver(l, nc, ainv, fq, td, od)
        names = sorted(os.listdir(od))
        expect = sorted([f"basis_rt-dq.div.00000.cell-0.index-{i:02d}.vtu" for i in range(6)] +
                        [f"solution.00000.cell-{c}.vtu" for c in range(ncell)])
        assert names == expect
        o = 0
        M = d[o:o + ncell * 49].reshape(ncell, 7, 7); o += ncell * 49
        b = d[o:o + ncell * 7].reshape(ncell, 7); o += ncell * 7
        S = d[o:o + ncell * 6 * nf].reshape(ncell, 6, nf); o += ncell * 6 * nf
        sol = d[o:o + ncell * N].reshape(ncell, N); o += ncell * N
        assert o == d.size
        verts = unit_cube_boxes(oracle_mod, nc)
        for c in range(ncell):
            Mo, bo, So, st = oracle_mod.rt3_basis_cell(l, verts[c], ainv[c], fq[c])
            assert st == 0
            assert np.abs(M[c] - Mo).max() < 1e-9 * np.abs(Mo).max()
            assert np.abs(b[c] - bo).max() < 1e-12 * np.abs(bo).max()
            assert np.abs(S[c] - So).max() < 1e-9 * np.abs(So).max()
        w = np.arange(1.0, 8.0)
        np.testing.assert_allclose(sol[:, :nf], np.einsum("i,cif->cf", w[:6], S), rtol=0,
                                   atol=1e-13 * np.abs(S).max() * w.sum())
        assert (sol[:, nf:] == 7.0).all()
        # the VTU cell fields: the flux at the fine cell centres
        h = 1.0 / nc / m
        nf1 = m * m * (m + 1)

        def centre(s):
            fx = s[:nf1].reshape(m, m, m + 1)
            fy = s[nf1:2 * nf1].reshape(m, m + 1, m)
            fz = s[2 * nf1:].reshape(m + 1, m, m)
            return np.stack([0.5 * (fx[..., 1:] + fx[..., :-1]) / h ** 2,
                             0.5 * (fy[:, 1:] + fy[:, :-1]) / h ** 2,
                             0.5 * (fz[1:] + fz[:-1]) / h ** 2], -1).reshape(-1, 3)

        for i in range(6):
            cf = read_vtu_cells(os.path.join(od, f"basis_rt-dq.div.00000.cell-0.index-{i:02d}.vtu"))
            sig = centre(S[0, i])
            np.testing.assert_allclose(cf["sigma"], sig, rtol=0, atol=1e-10 * np.abs(sig).max())
            np.testing.assert_allclose(cf["u"], 1.0)
        c = ncell - 1
        cf = read_vtu_cells(os.path.join(od, f"solution.00000.cell-{c}.vtu"))
        sig = centre(sol[c, :nf])
        np.testing.assert_allclose(cf["sigma"], sig, rtol=0, atol=1e-10 * np.abs(sig).max())
        np.testing.assert_allclose(cf["u"], 7.0)
