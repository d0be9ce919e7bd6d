"""The torch.autograd wrapper: backward() equals the adjoint gradient and
agrees with central differences of the forward value.  torch is imported
before the library, the way a user of the wrapper would."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_autograd_backward_matches_finite_differences():
    print("case autograd_fd", flush=True)
    import torch
    from msfec_amd import fields as F
    from msfec_amd.autograd import coarse_misfit
    from msfec_amd import QBasisBatch
    dim, nc, l = 3, 3, 2
    verts = F.coarse_vertices(dim, nc)
    g0 = np.log(F.lognormal_coefficient(dim, nc, l, 1.0, 0.0, seed=31))
    with QBasisBatch(verts, dim=dim, n_refine_local=l, solver=1) as qb:
        qb.set_coefficient(F.lognormal_coefficient(dim, nc, l, 1.0, 0.0, seed=32))
        qb.run()
        ut, _, _, _ = qb.coarse_solve(nc, abs_tol=0.0, rel_tol=1e-14)
    g = torch.tensor(g0, dtype=torch.float64, requires_grad=True)
    J = coarse_misfit(g, verts, dim, l, nc, ut)
    J.backward()
    grad = g.grad.numpy()
    dg = np.random.default_rng(3).standard_normal(g0.shape)
    eps = 1e-4
    with torch.no_grad():
        Jp = float(coarse_misfit(torch.tensor(g0 + eps * dg), verts, dim, l, nc, ut))
        Jm = float(coarse_misfit(torch.tensor(g0 - eps * dg), verts, dim, l, nc, ut))
    fd = (Jp - Jm) / (2 * eps)
    ad = float((grad * dg).sum())
    assert abs(fd - ad) <= 1e-6 * np.linalg.norm(grad) * np.linalg.norm(dg)
