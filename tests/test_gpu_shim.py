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
