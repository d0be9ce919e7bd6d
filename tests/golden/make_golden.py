"""Generate golden vectors for the Q1 multiscale basis with an implementation
that shares no code with the oracle (oracle/msfec_oracle.c):

- fine element matrices by 3-point Gauss-Legendre quadrature (exact for the
  Q1 integrands, but a different rule from the QGauss(2) the reference uses),
  written in numpy with shape functions built from 1-D Lagrange polynomials;
- dense assembly of the whole fine matrix and the load vector;
- the Dirichlet data of basis i: the tensor-product linear Lagrange function
  of coarse vertex i (what BasisQ1 reduces to on a box, basis_q1.tpp:43-67);
- the interior solve K_II x = -K_IB g by scipy.linalg.cho_solve;
- M = Psi K Psi^T and b = Psi F (q_basis.cc:262-300).

The fixtures are small (l <= 2 in 3-D, l <= 3 in 2-D) and committed as
tests/golden/q_basis_golden.npz.  Run:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import itertools
import os

import numpy as np
import scipy.linalg

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "q_basis_golden.npz")

G3 = (np.array([-np.sqrt(0.6), 0.0, np.sqrt(0.6)]) + 1.0) / 2.0
W3 = np.array([5.0, 8.0, 5.0]) / 18.0


def lagrange1(a, x):
    return x if a else 1.0 - x


def dlagrange1(a):
    return 1.0 if a else -1.0


def element(dim, h, Aq_fun, f_fun, r):
    """Element stiffness (+ reaction mass) and load by 3^dim-point Gauss."""
    nv = 1 << dim
    K = np.zeros((nv, nv))
    F = np.zeros(nv)
    for idx in itertools.product(range(3), repeat=dim):
        xi = np.array([G3[i] for i in idx])
        w = np.prod([W3[i] for i in idx]) * np.prod(h)
        A = Aq_fun(xi)
        f = f_fun(xi)
        val = np.zeros(nv)
        grad = np.zeros((nv, dim))
        for a in range(nv):
            bits = [(a >> d) & 1 for d in range(dim)]
            val[a] = np.prod([lagrange1(bits[d], xi[d]) for d in range(dim)])
            for d in range(dim):
                g = dlagrange1(bits[d]) / h[d]
                for e in range(dim):
                    if e != d:
                        g *= lagrange1(bits[e], xi[e])
                grad[a, d] = g
        K += w * (grad @ A @ grad.T + r * np.outer(val, val))
        F += w * f * val
    return K, F


def basis_dense(dim, l, lo, hi, coef, f=None, reac=None):
    N = 1 << l
    S = N + 1
    nv = 1 << dim
    h = (np.asarray(hi) - np.asarray(lo)) / N
    nn = S ** dim
    K = np.zeros((nn, nn))
    F = np.zeros(nn)
    for z in range(N ** dim):
        cz = [(z // N ** d) % N for d in range(dim)]
        a = coef[z]
        fz = 1.0 if f is None else f[z]
        rz = 0.0 if reac is None else reac[z]
        Ke, Fe = element(dim, h, lambda xi: a * np.eye(dim), lambda xi: fz, rz)
        ids = [sum((cz[d] + ((v >> d) & 1)) * S ** d for d in range(dim)) for v in range(nv)]
        K[np.ix_(ids, ids)] += Ke
        F[ids] += Fe
    coords = np.array([[(i // S ** d) % S for d in range(dim)] for i in range(nn)])
    inter = np.all((coords > 0) & (coords < N), axis=1)
    xi = coords / N
    G = np.zeros((nv, nn))
    for k in range(nv):
        G[k] = np.prod([xi[:, d] if (k >> d) & 1 else 1.0 - xi[:, d] for d in range(dim)], axis=0)
    Psi = G.copy()
    I, B = np.where(inter)[0], np.where(~inter)[0]
    if len(I):
        c = scipy.linalg.cho_factor(K[np.ix_(I, I)])
        X = scipy.linalg.cho_solve(c, -K[np.ix_(I, B)] @ G[:, B].T)
        Psi[:, I] = X.T
    M = Psi @ K @ Psi.T
    b = Psi @ F
    return M, b, Psi


def main():
    rng = np.random.default_rng(20261202)
    cases = []
    for dim, l, kind in [(2, 1, "lognormal"), (2, 2, "lognormal"), (2, 3, "lognormal"),
                         (3, 1, "lognormal"), (3, 2, "lognormal"), (3, 2, "reaction_load")]:
        N = 1 << l
        lo = rng.random(dim)
        hi = lo + 0.05 + 0.2 * rng.random(dim)
        coef = np.exp(1.5 * rng.standard_normal(N ** dim))
        f = reac = None
        if kind == "reaction_load":
            f = rng.standard_normal(N ** dim)
            reac = rng.random(N ** dim)
        M, b, Psi = basis_dense(dim, l, lo, hi, coef, f, reac)
        cases.append(dict(dim=dim, l=l, lo=lo, hi=hi, coef=coef,
                          f=np.zeros(0) if f is None else f,
                          reac=np.zeros(0) if reac is None else reac, M=M, b=b, Psi=Psi))
    arrays = {}
    for i, c in enumerate(cases):
        for k, v in c.items():
            arrays[f"c{i}_{k}"] = np.asarray(v)
    arrays["n_cases"] = np.array(len(cases))
    np.savez_compressed(OUT, **arrays)
    print(f"wrote {OUT} with {len(cases)} cases")


if __name__ == "__main__":
    main()
