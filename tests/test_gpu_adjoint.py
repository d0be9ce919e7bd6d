"""Adjoint gradient of the coarse misfit on the GPU (msfec_q_adjoint_gradient)
against the oracle's adjoint (itself pinned by finite differences in
tests/test_oracle.py) and against central differences of the GPU forward map.

J(g) = 1/2 sum over free coarse nodes (u_H(g) - u_target)^2, a = exp(g) per
fine cell, u_target the coarse solution of a second seeded field.
Tolerances: the FP64 path reproduces the oracle to 1e-9 relative; the mixed
path solves w = K_II^-1 F_I with its FP32 panels, so its gradient carries
~1e-7 relative error and is held to 1e-5.  Finite differences use
eps = 1e-4 along random directions (truncation ~1e-8); the bound is 1e-6
relative to |grad| |dg|.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def forward(verts, dim, l, nc, g, solver):
    from msfec_amd import QBasisBatch
    qb = QBasisBatch(verts, dim=dim, n_refine_local=l, solver=solver)
    qb.set_coefficient(np.exp(g))
    qb.run()
    return qb


def misfit_gpu(verts, dim, l, nc, g, ut, solver):
    with forward(verts, dim, l, nc, g, solver) as qb:
        u, _, _, _ = qb.coarse_solve(nc, abs_tol=0.0, rel_tol=1e-14)
    free = np.ones_like(u, dtype=bool)
    idx = np.indices((nc + 1,) * dim).reshape(dim, -1)[::-1]
    for d in range(dim):
        free &= (idx[d] > 0) & (idx[d] < nc)
    return 0.5 * float(((u - ut)[free] ** 2).sum())


def target(verts, dim, l, nc, seed):
    from msfec_amd import fields as F
    g2 = np.log(F.lognormal_coefficient(dim, nc, l, 1.0, 0.0, seed=seed))
    with forward(verts, dim, l, nc, g2, 1) as qb:
        u, _, _, _ = qb.coarse_solve(nc, abs_tol=0.0, rel_tol=1e-14)
    return u


@pytest.mark.parametrize("solver", [1, 2])
@pytest.mark.parametrize("dim,nc,l", [(2, 4, 3), (3, 2, 2), (3, 3, 3)])
def test_adjoint_matches_oracle(oracle_mod, dim, nc, l, solver):
    print(f"case adjoint_vs_oracle dim={dim} nc={nc} l={l} solver={solver}", flush=True)
    from msfec_amd import fields as F
    verts = F.coarse_vertices(dim, nc)
    g = np.log(F.lognormal_coefficient(dim, nc, l, 1.0, 0.0, seed=11))
    ut = target(verts, dim, l, nc, 12)
    with forward(verts, dim, l, nc, g, solver) as qb:
        J, grad = qb.adjoint_gradient(nc, ut)
    p = oracle_mod.make_problem(dim, l)
    Jo, go = oracle_mod.adjoint_gradient(p, nc, verts, np.exp(g), ut)
    tol = 1e-9 if solver == 1 else 1e-5
    assert abs(J - Jo) <= tol * abs(Jo)
    assert np.linalg.norm(grad - go) <= tol * np.linalg.norm(go)


def test_adjoint_gradient_into_device_memory():
    """The gradient written into a DeviceBuffer (MSFEC_DEVICE) equals the
    host-landing one bit for bit, and a wrongly sized buffer is refused."""
    print("case adjoint_device_out", flush=True)
    from msfec_amd import DeviceBuffer, fields as F
    dim, nc, l = 3, 3, 3
    verts = F.coarse_vertices(dim, nc)
    g = np.log(F.lognormal_coefficient(dim, nc, l, 1.0, 0.0, seed=13))
    ut = target(verts, dim, l, nc, 14)
    with forward(verts, dim, l, nc, g, 1) as qb:
        J, grad = qb.adjoint_gradient(nc, ut)
        with DeviceBuffer(grad.shape) as d:
            Jd, _ = qb.adjoint_gradient(nc, ut, out=d)
            gd = d.numpy()
        with DeviceBuffer(grad.size - 1) as bad:
            with pytest.raises(ValueError):
                qb.adjoint_gradient(nc, ut, out=bad)
    assert Jd == J
    np.testing.assert_array_equal(gd, grad)


@pytest.mark.parametrize("solver", [1, 2])
def test_adjoint_matches_finite_differences(solver):
    print(f"case adjoint_fd solver={solver}", flush=True)
    from msfec_amd import fields as F
    dim, nc, l = 3, 4, 3
    verts = F.coarse_vertices(dim, nc)
    g = np.log(F.lognormal_coefficient(dim, nc, l, 1.5, 1.0 / 64, seed=20261202))
    ut = target(verts, dim, l, nc, 77)
    with forward(verts, dim, l, nc, g, solver) as qb:
        J, grad = qb.adjoint_gradient(nc, ut)
    rng = np.random.default_rng(9)
    eps = 1e-4
    for _ in range(2):
        dg = rng.standard_normal(g.shape)
        Jp = misfit_gpu(verts, dim, l, nc, g + eps * dg, ut, solver)
        Jm = misfit_gpu(verts, dim, l, nc, g - eps * dg, ut, solver)
        fd = (Jp - Jm) / (2 * eps)
        ad = float((grad * dg).sum())
        scale = np.linalg.norm(grad) * np.linalg.norm(dg)
        assert abs(fd - ad) <= 1e-6 * scale, (fd, ad, scale)


def test_adjoint_headline_batch():
    """16^3 coarse x 8^3 fine (2.1M log-coefficients): the gradient against
    central differences along one random direction."""
    print("case adjoint_headline", flush=True)
    from msfec_amd import fields as F
    dim, nc, l = 3, 16, 3
    verts = F.coarse_vertices(dim, nc)
    g = np.log(F.lognormal_coefficient(dim, nc, l, 1.5, 1.0 / 64, seed=20261202))
    ut = target(verts, dim, l, nc, 20261203)
    with forward(verts, dim, l, nc, g, 1) as qb:
        J, grad = qb.adjoint_gradient(nc, ut)
    assert grad.shape == (nc ** 3, 512) and np.isfinite(grad).all()
    dg = np.random.default_rng(4).standard_normal(g.shape)
    eps = 1e-4
    Jp = misfit_gpu(verts, dim, l, nc, g + eps * dg, ut, 1)
    Jm = misfit_gpu(verts, dim, l, nc, g - eps * dg, ut, 1)
    fd = (Jp - Jm) / (2 * eps)
    ad = float((grad * dg).sum())
    assert abs(fd - ad) <= 1e-6 * np.linalg.norm(grad) * np.linalg.norm(dg)
    assert abs(fd - ad) <= 1e-5 * abs(fd)
