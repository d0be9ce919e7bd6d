"""ctypes front end of the FP64 CPU oracle (oracle/msfec_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the cpu_baseline leg of bench.py, always as the results to compare with (or, for the CPU
baseline, as the thing timed beside the GPU path).  The product library never
imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

COEF_SCALAR_CELL, COEF_TENSOR_QP, COEF_ANALYTIC = 0, 1, 2
RHS_CONST, RHS_CELL, RHS_QP = 0, 1, 2
REACTION_NONE, REACTION_CONST, REACTION_CELL = 0, 1, 2
SOLVER_DIRECT, SOLVER_CG_SSOR = 0, 1


class Problem(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int),
        ("n_refine", ctypes.c_int),
        ("coef_mode", ctypes.c_int),
        ("rhs_mode", ctypes.c_int),
        ("rhs_const", ctypes.c_double),
        ("reaction_mode", ctypes.c_int),
        ("reaction_const", ctypes.c_double),
        ("freq", ctypes.c_int * 3),
        ("scale", ctypes.c_double * 3),
        ("alpha", ctypes.c_double * 3),
        ("rotate", ctypes.c_int),
        ("cg_maxit", ctypes.c_int),
        ("cg_tol", ctypes.c_double),
        ("ssor_omega", ctypes.c_double),
    ]


def build(force: bool = False) -> str:
    """Compile liboracle.so with the Makefile next to this file."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "all"], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        pp = ctypes.POINTER(Problem)
        L.oq_basis_batch.argtypes = [pp, ctypes.c_int, dp, dp, dp, dp, ctypes.c_int,
                                     ctypes.c_int, dp, dp, dp, ip, ip]
        L.oq_basis_batch.restype = ctypes.c_int
        L.oq_basis_q1_coeffs.argtypes = [ctypes.c_int, dp, dp]
        L.oq_basis_q1_coeffs.restype = ctypes.c_int
        L.oq_basis_q1_value.argtypes = [ctypes.c_int, dp, ctypes.c_int, dp]
        L.oq_basis_q1_value.restype = ctypes.c_double
        L.oq_diffusion_A.argtypes = [pp, dp, dp]
        L.oq_diffusion_A.restype = None
        L.oq_element_matrix.argtypes = [ctypes.c_int, dp, dp, dp, ctypes.c_double, dp, dp]
        L.oq_element_matrix.restype = None
        L.oq_coarse_solve.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, ctypes.c_double,
                                      ctypes.c_int, dp]
        L.oq_coarse_solve.restype = ctypes.c_int
        L.oq_reconstruct.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp, dp]
        L.oq_reconstruct.restype = None
        L.oq_coarse_solve_rhs.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, ctypes.c_double,
                                          ctypes.c_int, dp]
        L.oq_coarse_solve_rhs.restype = ctypes.c_int
        L.oq_misfit.argtypes = [pp, ctypes.c_int, dp, dp, dp, dp, ctypes.c_double, dp, dp]
        L.oq_misfit.restype = ctypes.c_int
        L.oq_adjoint_gradient.argtypes = [pp, ctypes.c_int, dp, dp, dp, dp, ctypes.c_double,
                                          dp, dp]
        L.oq_adjoint_gradient.restype = ctypes.c_int
        L.oq_ned_basis_cell.argtypes = [ctypes.c_int, dp, dp, dp, dp, dp, ctypes.c_int, dp, dp,
                                        dp]
        L.oq_ned_basis_cell.restype = ctypes.c_int
        L.oq_ned_assemble.argtypes = [ctypes.c_int, dp, dp, dp, dp, dp, dp, dp]
        L.oq_ned_assemble.restype = ctypes.c_int
        L.oq_rt_basis_cell.argtypes = [ctypes.c_int, dp, dp, dp, dp, dp, dp, dp]
        L.oq_rt_basis_cell.restype = ctypes.c_int
        L.oq_rt_assemble.argtypes = [ctypes.c_int, dp, dp, dp, dp, dp]
        L.oq_rt_assemble.restype = ctypes.c_int
        L.oq_rt_basis_cell_qp.argtypes = [ctypes.c_int, dp, dp, dp, dp, dp, dp, dp]
        L.oq_rt_basis_cell_qp.restype = ctypes.c_int
        L.oq_rt_assemble_qp.argtypes = [ctypes.c_int, dp, dp, dp, dp, dp]
        L.oq_rt_assemble_qp.restype = ctypes.c_int
        L.oq_nrt_basis_cell.argtypes = [ctypes.c_int, dp, dp, dp, dp, dp, dp, dp, dp]
        L.oq_nrt_basis_cell.restype = ctypes.c_int
        L.oq_nrt_assemble.argtypes = [ctypes.c_int, dp, dp, dp, dp, dp, dp, dp]
        L.oq_nrt_assemble.restype = ctypes.c_int
        L.oq_nrt_curl_loads.argtypes = [ctypes.c_int, dp, dp]
        L.oq_nrt_curl_loads.restype = ctypes.c_int
        ip = ctypes.POINTER(ctypes.c_int)
        L.oq_nrt_assemble_coo.argtypes = [ctypes.c_int, dp, dp, dp, dp, dp, ip, ip, dp, dp]
        L.oq_nrt_assemble_coo.restype = ctypes.c_int
        L.oq_nrt_boundary.argtypes = [ctypes.c_int, dp, dp, ip]
        L.oq_nrt_boundary.restype = ctypes.c_int
        L.oq_rt3_assemble.argtypes = [ctypes.c_int, dp, dp, dp, dp, dp]
        L.oq_rt3_assemble.restype = ctypes.c_int
        L.oq_rt3_basis_cell.argtypes = [ctypes.c_int, dp, dp, dp, dp, dp, dp]
        L.oq_rt3_basis_cell.restype = ctypes.c_int
        L.oq_rt3_assemble_coo.argtypes = [ctypes.c_int, dp, dp, dp, ip, ip, dp, dp]
        L.oq_rt3_assemble_coo.restype = ctypes.c_int
        _lib = L
    return _lib


def _dp(a):
    if a is None:
        return ctypes.POINTER(ctypes.c_double)()
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def make_problem(dim, n_refine, coef_mode=COEF_SCALAR_CELL, rhs_mode=RHS_CONST,
                 rhs_const=1.0, reaction_mode=REACTION_NONE, reaction_const=0.0,
                 freq=(0, 0, 0), scale=(1.0, 1.0, 1.0), alpha=(1.0, 1.0, 1.0),
                 rotate=False, cg_maxit=1000, cg_tol=1e-12, ssor_omega=1.6):
    p = Problem()
    p.dim, p.n_refine, p.coef_mode, p.rhs_mode = dim, n_refine, coef_mode, rhs_mode
    p.rhs_const, p.reaction_mode, p.reaction_const = rhs_const, reaction_mode, reaction_const
    p.freq[:] = list(freq)
    p.scale[:] = list(scale)
    p.alpha[:] = list(alpha)
    p.rotate = int(bool(rotate))
    p.cg_maxit, p.cg_tol, p.ssor_omega = cg_maxit, cg_tol, ssor_omega
    return p


def basis_batch(prob: Problem, verts, coef=None, f=None, reaction=None,
                solver=SOLVER_DIRECT, n_threads=0, want_psi=True):
    """Run the oracle on a batch of cells. Returns dict(M, b, Psi, iters, status)."""
    dim, nref = prob.dim, prob.n_refine
    nv = 1 << dim
    nn = ((1 << nref) + 1) ** dim
    verts = np.ascontiguousarray(verts, dtype=np.float64)
    n_cells = verts.shape[0]
    coef = None if coef is None else np.ascontiguousarray(coef, dtype=np.float64)
    f = None if f is None else np.ascontiguousarray(f, dtype=np.float64)
    reaction = None if reaction is None else np.ascontiguousarray(reaction, dtype=np.float64)
    M = np.zeros((n_cells, nv, nv))
    b = np.zeros((n_cells, nv))
    Psi = np.zeros((n_cells, nv, nn)) if want_psi else None
    iters = np.zeros(n_cells, dtype=np.int32)
    status = np.zeros(n_cells, dtype=np.int32)
    lib().oq_basis_batch(ctypes.byref(prob), n_cells, _dp(verts), _dp(coef), _dp(f),
                         _dp(reaction), solver, n_threads, _dp(M), _dp(b), _dp(Psi),
                         _ip(iters), _ip(status))
    return {"M": M, "b": b, "Psi": Psi, "iters": iters, "status": status}


def q1_coeffs(dim, verts):
    verts = np.ascontiguousarray(verts, dtype=np.float64)
    nv = 1 << dim
    C = np.zeros((nv, nv))
    rc = lib().oq_basis_q1_coeffs(dim, _dp(verts), _dp(C))
    if rc != 0:
        raise ValueError("singular vertex-monomial matrix")
    return C


def q1_value(dim, C, index, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return lib().oq_basis_q1_value(dim, _dp(np.ascontiguousarray(C)), index, _dp(x))


def diffusion_A(prob: Problem, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    A = np.zeros(9)
    lib().oq_diffusion_A(ctypes.byref(prob), _dp(x), _dp(A))
    return A.reshape(3, 3)


def element_matrix(dim, h, Aq, fq, rq=0.0):
    nv = 1 << dim
    Ke = np.zeros((nv, nv))
    Fe = np.zeros(nv)
    lib().oq_element_matrix(dim, _dp(np.ascontiguousarray(h, dtype=np.float64)),
                            _dp(np.ascontiguousarray(Aq, dtype=np.float64)),
                            _dp(np.ascontiguousarray(fq, dtype=np.float64)), float(rq),
                            _dp(Ke), _dp(Fe))
    return Ke, Fe


def coarse_solve(dim, nc, M, b, rtol=1e-14, maxit=100000):
    u = np.zeros((nc + 1) ** dim)
    it = lib().oq_coarse_solve(dim, nc, _dp(np.ascontiguousarray(M, dtype=np.float64)),
                               _dp(np.ascontiguousarray(b, dtype=np.float64)), rtol, maxit,
                               _dp(u))
    return u, it


def reconstruct(dim, nc, n_refine, u_coarse, Psi):
    nn = ((1 << n_refine) + 1) ** dim
    ncell = nc ** dim
    out = np.zeros((ncell, nn))
    lib().oq_reconstruct(dim, nc, n_refine, _dp(np.ascontiguousarray(u_coarse, dtype=np.float64)),
                         _dp(np.ascontiguousarray(Psi, dtype=np.float64)), _dp(out))
    return out


def coarse_solve_rhs(dim, nc, M, rhs_nodes, rtol=1e-14, maxit=100000):
    """The coarse system of coarse_solve with a nodal right-hand side."""
    u = np.zeros((nc + 1) ** dim)
    it = lib().oq_coarse_solve_rhs(dim, nc, _dp(np.ascontiguousarray(M, dtype=np.float64)),
                                   _dp(np.ascontiguousarray(rhs_nodes, dtype=np.float64)), rtol,
                                   maxit, _dp(u))
    return u, it


def misfit(prob, nc, verts, coef, u_target, f=None, rtol=1e-14):
    """J = 1/2 sum over free coarse nodes (u_H - u_target)^2; returns (J, u_H)."""
    dim = prob.dim
    u = np.zeros((nc + 1) ** dim)
    J = ctypes.c_double(0.0)
    st = lib().oq_misfit(ctypes.byref(prob), nc, _dp(np.ascontiguousarray(verts, dtype=np.float64)),
                         _dp(np.ascontiguousarray(coef, dtype=np.float64)),
                         _dp(None if f is None else np.ascontiguousarray(f, dtype=np.float64)),
                         _dp(np.ascontiguousarray(u_target, dtype=np.float64)), rtol,
                         ctypes.byref(J), _dp(u))
    if st != 0:
        raise RuntimeError(f"oq_misfit failed ({st})")
    return J.value, u


def adjoint_gradient(prob, nc, verts, coef, u_target, f=None, rtol=1e-14):
    """(J, dJ/dg) for a = exp(g) per fine cell, dJ/dg in the layout of coef."""
    grad = np.zeros(np.asarray(coef).shape)
    J = ctypes.c_double(0.0)
    st = lib().oq_adjoint_gradient(ctypes.byref(prob), nc,
                                   _dp(np.ascontiguousarray(verts, dtype=np.float64)),
                                   _dp(np.ascontiguousarray(coef, dtype=np.float64)),
                                   _dp(None if f is None else np.ascontiguousarray(f, dtype=np.float64)),
                                   _dp(np.ascontiguousarray(u_target, dtype=np.float64)), rtol,
                                   ctypes.byref(J), _dp(grad))
    if st != 0:
        raise RuntimeError(f"oq_adjoint_gradient failed ({st})")
    return J.value, grad


# ---------------------------------------------------------------------------
# 2-D Q1 x Nedelec basis (ned_oracle.c)
# ---------------------------------------------------------------------------
def ned_sizes(l):
    """(ns, nh, nv, N): nodes, horizontal edges, vertical edges, all dofs."""
    m = 1 << l
    ns, nh, nv = (m + 1) ** 2, m * (m + 1), (m + 1) * m
    return ns, nh, nv, ns + nh + nv


def _ned_coefs(l, a, binv, r, f):
    m = 1 << l
    nf = m * m
    a = np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.float64), (nf,)))
    binv = np.ascontiguousarray(np.broadcast_to(np.asarray(binv, dtype=np.float64), (nf,)))
    r = np.ascontiguousarray(np.broadcast_to(np.asarray(r, dtype=np.float64), (nf,)))
    f = np.ascontiguousarray(np.broadcast_to(np.asarray(f, dtype=np.float64), (nf, 2)))
    return a, binv, r, f


def ned_basis_cell(l, verts, a, binv=1.0, r=0.0, f=(1.0, 1.0), full_rhs=True):
    """One coarse box cell: returns (M 8x8, b 8, Psi 8 x N, status)."""
    a, binv, r, f = _ned_coefs(l, a, binv, r, f)
    v = np.ascontiguousarray(verts, dtype=np.float64).reshape(4, 2)
    N = ned_sizes(l)[3]
    M = np.zeros((8, 8))
    b = np.zeros(8)
    Psi = np.zeros((8, N))
    st = lib().oq_ned_basis_cell(l, _dp(v), _dp(a), _dp(binv), _dp(r), _dp(f), int(full_rhs),
                                 _dp(M), _dp(b), _dp(Psi))
    return M, b, Psi, st


def _map_cells(fn, n, n_threads):
    """fn(c) for every cell, on n_threads host threads (the C calls release
    the interpreter lock; the oracle keeps no state between calls)."""
    if n_threads <= 1:
        return [fn(c) for c in range(n)]
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(n_threads) as ex:
        return list(ex.map(fn, range(n)))


def ned_basis_batch(l, verts, a, binv=1.0, r=0.0, f=(1.0, 1.0), full_rhs=True, n_threads=1):
    """Cells of verts (n, 4, 2) with per-cell coefficient blocks a (n, m*m) etc."""
    verts = np.asarray(verts, dtype=np.float64)
    n = verts.shape[0]
    m = 1 << l

    def per_cell(x, shape):
        x = np.asarray(x, dtype=np.float64)
        return np.broadcast_to(x, (n,) + shape) if x.ndim <= len(shape) else x.reshape((n,) + shape)

    a_, b_, r_ = per_cell(a, (m * m,)), per_cell(binv, (m * m,)), per_cell(r, (m * m,))
    f_ = per_cell(f, (m * m, 2)) if np.asarray(f).ndim >= 2 else np.broadcast_to(
        np.asarray(f, dtype=np.float64), (n, m * m, 2))
    res = _map_cells(lambda c: ned_basis_cell(l, verts[c], a_[c], b_[c], r_[c], f_[c], full_rhs),
                     n, n_threads)
    return {k: np.array([x[i] for x in res]) for i, k in enumerate(("M", "b", "Psi", "status"))}


def ned_assemble(l, verts, a, binv=1.0, r=0.0, f=(1.0, 1.0)):
    """The assembled full system (row = test function) and the load (v, f)."""
    a, binv, r, f = _ned_coefs(l, a, binv, r, f)
    v = np.ascontiguousarray(verts, dtype=np.float64).reshape(4, 2)
    N = ned_sizes(l)[3]
    A = np.zeros((N, N))
    F = np.zeros(N)
    st = lib().oq_ned_assemble(l, _dp(v), _dp(a), _dp(binv), _dp(r), _dp(f), _dp(A), _dp(F))
    assert st == 0
    return A, F


# ---------------------------------------------------------------------------
# 2-D Raviart-Thomas x DQ0 basis (rt_oracle.c)
# ---------------------------------------------------------------------------
def rt_sizes(l):
    """(nv, ne, nc): vertical edges, all edges, fine cells."""
    m = 1 << l
    nv = m * (m + 1)
    return nv, 2 * nv, m * m


def _rt_coefs(l, a, f, tensor=False):
    nf = (1 << l) ** 2
    shp = (nf, 4, 3) if tensor else (nf,)
    a = np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.float64), shp))
    f = np.ascontiguousarray(np.broadcast_to(np.asarray(f, dtype=np.float64), (nf,)))
    return a, f


def rt_basis_cell(l, verts, a, f=1.0, tensor=False):
    """One coarse box cell: returns (M 5x5, b 5, Sigma 4 x ne, U 4 x nc, status).
    a: one value per fine cell, or with tensor=True (xx, xy, yy) at the 4
    Gauss points of every fine cell (m*m x 4 x 3)."""
    a, f = _rt_coefs(l, a, f, tensor)
    v = np.ascontiguousarray(verts, dtype=np.float64).reshape(4, 2)
    _, ne, nc = rt_sizes(l)
    M = np.zeros((5, 5))
    b = np.zeros(5)
    Sig = np.zeros((4, ne))
    U = np.zeros((4, nc))
    fn = lib().oq_rt_basis_cell_qp if tensor else lib().oq_rt_basis_cell
    st = fn(l, _dp(v), _dp(a), _dp(f), _dp(M), _dp(b), _dp(Sig), _dp(U))
    return M, b, Sig, U, st


def rt_basis_batch(l, verts, a, f=1.0, tensor=False, n_threads=1):
    """Cells of verts (n, 4, 2) with per-cell coefficient blocks a (n, m*m, or
    n, m*m, 4, 3 with tensor=True) and f (n, m*m)."""
    verts = np.asarray(verts, dtype=np.float64)
    n = verts.shape[0]
    nf = (1 << l) ** 2

    def per_cell(x, shp):
        x = np.asarray(x, dtype=np.float64)
        return np.broadcast_to(x, (n,) + shp) if x.ndim <= len(shp) else x.reshape((n,) + shp)

    a_ = per_cell(a, (nf, 4, 3) if tensor else (nf,))
    f_ = per_cell(f, (nf,))
    res = _map_cells(lambda c: rt_basis_cell(l, verts[c], a_[c], f_[c], tensor), n, n_threads)
    return {k: np.array([x[i] for x in res])
            for i, k in enumerate(("M", "b", "Sigma", "U", "status"))}


def rt_assemble(l, verts, a, f=1.0, tensor=False):
    """The assembled mixed system over all fine dofs (row = test) and the load."""
    a, f = _rt_coefs(l, a, f, tensor)
    v = np.ascontiguousarray(verts, dtype=np.float64).reshape(4, 2)
    _, ne, nc = rt_sizes(l)
    N = ne + nc
    A = np.zeros((N, N))
    F = np.zeros(N)
    fn = lib().oq_rt_assemble_qp if tensor else lib().oq_rt_assemble
    st = fn(l, _dp(v), _dp(a), _dp(f), _dp(A), _dp(F))
    assert st == 0
    return A, F


# ---------------------------------------------------------------------------
# 3-D Nedelec x Raviart-Thomas basis (nrt_oracle.c)
# ---------------------------------------------------------------------------
def nrt_sizes(l):
    """(m, nx, ne, nf1, nf, N): fine cells per direction, x-edges, edges,
    x-faces, faces, all fine dofs."""
    m = 1 << l
    nx = m * (m + 1) ** 2
    nf1 = m * m * (m + 1)
    return m, nx, 3 * nx, nf1, 3 * nf1, 3 * nx + 3 * nf1


def _nrt_coefs(l, ainv, bq, rq, fq):
    nc = (1 << l) ** 3
    ainv = np.ascontiguousarray(np.broadcast_to(np.asarray(ainv, dtype=np.float64), (nc, 8, 6)))
    bq = np.ascontiguousarray(np.broadcast_to(np.asarray(bq, dtype=np.float64), (nc, 8)))
    rq = None if rq is None else np.ascontiguousarray(
        np.broadcast_to(np.asarray(rq, dtype=np.float64), (nc, 8)))
    fq = None if fq is None else np.ascontiguousarray(
        np.broadcast_to(np.asarray(fq, dtype=np.float64), (nc, 8, 3)))
    return ainv, bq, rq, fq


def nrt_basis_cell(l, verts, ainv, bq=1.0, rq=None, fq=None):
    """One coarse box cell (verts 8 x 3, deal.II vertex order): returns
    (M 18x18, b 18, X 18 x N, status).  ainv: (xx, xy, xz, yy, yz, zz) of
    A^-1 at the 8 Gauss points of every fine cell (m^3 x 8 x 6, or
    broadcastable); bq, rq: m^3 x 8; fq: m^3 x 8 x 3."""
    ainv, bq, rq, fq = _nrt_coefs(l, ainv, bq, rq, fq)
    v = np.ascontiguousarray(verts, dtype=np.float64).reshape(8, 3)
    N = nrt_sizes(l)[5]
    M = np.zeros((18, 18))
    b = np.zeros(18)
    X = np.zeros((18, N))
    st = lib().oq_nrt_basis_cell(l, _dp(v), _dp(ainv), _dp(bq), _dp(rq), _dp(fq), _dp(M),
                                 _dp(b), _dp(X))
    return M, b, X, st


def nrt_basis_batch(l, verts, ainv, bq=1.0, rq=None, fq=None, n_threads=1):
    """Cells of verts (n, 8, 3) with per-cell coefficient blocks (n, m^3, 8, ...)."""
    verts = np.asarray(verts, dtype=np.float64)
    n = verts.shape[0]
    nc = (1 << l) ** 3

    def per_cell(x, shp):
        if x is None:
            return [None] * n
        x = np.asarray(x, dtype=np.float64)
        return np.broadcast_to(x, (n,) + shp) if x.ndim <= len(shp) else x.reshape((n,) + shp)

    a_, b_ = per_cell(ainv, (nc, 8, 6)), per_cell(bq, (nc, 8))
    r_, f_ = per_cell(rq, (nc, 8)), per_cell(fq, (nc, 8, 3))
    res = _map_cells(lambda c: nrt_basis_cell(l, verts[c], a_[c], b_[c], r_[c], f_[c]), n,
                     n_threads)
    return {k: np.array([x[i] for x in res]) for i, k in enumerate(("M", "b", "X", "status"))}


def nrt_assemble(l, verts, ainv, bq=1.0, rq=None, fq=None):
    """The assembled mixed system over all fine dofs (row = test) and the load (v, f)."""
    ainv, bq, rq, fq = _nrt_coefs(l, ainv, bq, rq, fq)
    v = np.ascontiguousarray(verts, dtype=np.float64).reshape(8, 3)
    N = nrt_sizes(l)[5]
    A = np.zeros((N, N))
    F = np.zeros(N)
    st = lib().oq_nrt_assemble(l, _dp(v), _dp(ainv), _dp(bq), _dp(rq), _dp(fq), _dp(A), _dp(F))
    assert st == 0
    return A, F


def nrt_curl_loads(l, verts):
    """(v, curl N_n) of the 12 edge bases over all fine dofs (12 x N)."""
    v = np.ascontiguousarray(verts, dtype=np.float64).reshape(8, 3)
    Fc = np.zeros((12, nrt_sizes(l)[5]))
    st = lib().oq_nrt_curl_loads(l, _dp(v), _dp(Fc))
    assert st == 0
    return Fc


def nrt_assemble_sparse(l, verts, ainv, bq=1.0, rq=None, fq=None):
    """The assembled mixed system over all fine dofs as a scipy CSR matrix (row
    = test) and the load (v, f); l up to 4."""
    import scipy.sparse as sp
    ainv, bq, rq, fq = _nrt_coefs(l, ainv, bq, rq, fq)
    v = np.ascontiguousarray(verts, dtype=np.float64).reshape(8, 3)
    N = nrt_sizes(l)[5]
    nt = (1 << (3 * l)) * 324
    rows = np.empty(nt, dtype=np.int32)
    cols = np.empty(nt, dtype=np.int32)
    vals = np.empty(nt)
    F = np.zeros(N)
    ip = ctypes.POINTER(ctypes.c_int)
    st = lib().oq_nrt_assemble_coo(l, _dp(v), _dp(ainv), _dp(bq), _dp(rq), _dp(fq),
                                   rows.ctypes.data_as(ip), cols.ctypes.data_as(ip), _dp(vals),
                                   _dp(F))
    assert st == 0
    return sp.coo_matrix((vals, (rows, cols)), shape=(N, N)).tocsr(), F


def nrt_boundary(l, verts):
    """(G 18 x N, the fixed values of the bases on the boundary dofs, zero
    inside; on_boundary, N bools)."""
    v = np.ascontiguousarray(verts, dtype=np.float64).reshape(8, 3)
    N = nrt_sizes(l)[5]
    G = np.zeros((18, N))
    b = np.zeros(N, dtype=np.int32)
    st = lib().oq_nrt_boundary(l, _dp(v), _dp(G), b.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    assert st == 0
    return G, b.astype(bool)


def nrt_basis_cell_sparse(l, verts, ainv, bq=1.0, rq=None, fq=None):
    """One cell by a sparse direct solve (SuperLU, scipy.sparse.linalg.splu)
    of the assembled interior system with the 18 right-hand sides: the
    reference's solve_direct (UMFPACK, ned_rt_basis.cc:579-634) restated for
    l up to 4, where the dense LU of nrt_basis_cell does not reach, with two
    steps of iterative refinement.  Returns
    (M 18x18, b 18, X 18 x N); M and b by the block rule of
    assemble_global_element_matrix (:850-948)."""
    import scipy.sparse.linalg as spla
    A, F = nrt_assemble_sparse(l, verts, ainv, bq, rq, fq)
    Fc = nrt_curl_loads(l, verts)
    G, bnd = nrt_boundary(l, verts)
    ne, N = nrt_sizes(l)[2], nrt_sizes(l)[5]
    I = np.flatnonzero(~bnd)
    B = np.flatnonzero(bnd)
    AI = A[I]
    rhs = np.concatenate([Fc[:, I].T, np.zeros((len(I), 6))], axis=1)
    rhs -= AI[:, B] @ G[:, B].T
    AII = AI[:, I].tocsc()
    lu = spla.splu(AII)
    x = lu.solve(rhs)
    # two steps of iterative refinement, as nrt_solve does for the dense LU:
    # the interior system is badly scaled (edge rows ~ h, face rows ~ 1/h^3)
    for _ in range(2):
        x += lu.solve(rhs - AII @ x)
    X = G.copy()
    X[:, I] = x.T
    Ye = A[:, :ne] @ X[:, :ne].T
    Yf = A[:, ne:] @ X[:, ne:].T
    M = np.zeros((18, 18))
    for i in range(18):
        ri = slice(0, ne) if i < 12 else slice(ne, N)
        for j in range(18):
            M[i, j] = X[i, ri] @ (Ye if j < 12 else Yf)[ri, j]
    b = np.array([0.0] * 12 + [X[i, ne:] @ F[ne:] for i in range(12, 18)])
    return M, b, X


def rt3_sizes(l):
    """(m, nf, nc, N): fine cells per direction, fine faces, fine cells, all
    dofs of the 3-D RT_DQ system."""
    m = 1 << l
    nf = 3 * m * m * (m + 1)
    return m, nf, m ** 3, nf + m ** 3


def rt3_basis_cell(l, verts, ainv, fq=None):
    """3-D RT_DQ basis of one box cell: (M 7x7, b 7, Sigma 6 x nf, status)."""
    nc = (1 << l) ** 3
    ainv = np.ascontiguousarray(np.broadcast_to(np.asarray(ainv, dtype=np.float64), (nc, 8, 6)))
    fq = None if fq is None else np.ascontiguousarray(
        np.broadcast_to(np.asarray(fq, dtype=np.float64), (nc, 8)))
    v = np.ascontiguousarray(verts, dtype=np.float64).reshape(8, 3)
    nf = rt3_sizes(l)[1]
    M = np.zeros((7, 7))
    b = np.zeros(7)
    S = np.zeros((6, nf))
    st = lib().oq_rt3_basis_cell(l, _dp(v), _dp(ainv), _dp(fq), _dp(M), _dp(b), _dp(S))
    return M, b, S, st


def rt3_assemble(l, verts, ainv, fq=None):
    nc = (1 << l) ** 3
    ainv = np.ascontiguousarray(np.broadcast_to(np.asarray(ainv, dtype=np.float64), (nc, 8, 6)))
    fq = None if fq is None else np.ascontiguousarray(
        np.broadcast_to(np.asarray(fq, dtype=np.float64), (nc, 8)))
    v = np.ascontiguousarray(verts, dtype=np.float64).reshape(8, 3)
    N = rt3_sizes(l)[3]
    A = np.zeros((N, N))
    F = np.zeros(N)
    assert lib().oq_rt3_assemble(l, _dp(v), _dp(ainv), _dp(fq), _dp(A), _dp(F)) == 0
    return A, F


def rt3_assemble_sparse(l, verts, ainv, fq=None):
    """The assembled 3-D RT_DQ system (fine faces, then fine cells) as a scipy
    CSR matrix (row = test) and the load; l up to 4."""
    import scipy.sparse as sp
    nc = (1 << l) ** 3
    ainv = np.ascontiguousarray(np.broadcast_to(np.asarray(ainv, dtype=np.float64), (nc, 8, 6)))
    fq = None if fq is None else np.ascontiguousarray(
        np.broadcast_to(np.asarray(fq, dtype=np.float64), (nc, 8)))
    v = np.ascontiguousarray(verts, dtype=np.float64).reshape(8, 3)
    N = rt3_sizes(l)[3]
    rows = np.empty(nc * 49, dtype=np.int32)
    cols = np.empty(nc * 49, dtype=np.int32)
    vals = np.empty(nc * 49)
    F = np.zeros(N)
    ip = ctypes.POINTER(ctypes.c_int)
    assert lib().oq_rt3_assemble_coo(l, _dp(v), _dp(ainv), _dp(fq), rows.ctypes.data_as(ip),
                                     cols.ctypes.data_as(ip), _dp(vals), _dp(F)) == 0
    return sp.coo_matrix((vals, (rows, cols)), shape=(N, N)).tocsr(), F


def rt3_basis_cell_sparse(l, verts, ainv, fq=None):
    """oq_rt3_basis_cell by a sparse direct solve (SuperLU) with two steps of
    iterative refinement, for l up to 4: the same boundary fluxes (the coarse
    traces), u pinned in the last fine cell, the loads -+|c|/|K| on the cell
    rows (rt_dq_basis.cc:516-538), and M, b by the block products with u = 1
    (:822-914).  Returns (M 7x7, b 7, Sigma 6 x nf)."""
    import scipy.sparse.linalg as spla
    A, F = rt3_assemble_sparse(l, verts, ainv, fq)
    m, nf, nc, N = rt3_sizes(l)
    G, bnd = nrt_boundary(l, verts)
    ne = nrt_sizes(l)[2]
    Gf, fb = G[12:, ne:], bnd[ne:]
    fix = np.zeros((6, N))
    fix[:, :nf] = np.where(fb, Gf, 0.0)
    free = np.ones(N, dtype=bool)
    free[:nf] = ~fb
    free[N - 1] = False
    I = np.flatnonzero(free)
    rhs = -(A[I] @ fix.T)
    sgn = np.array([-1.0, 1, -1, 1, -1, 1])
    rhs[I >= nf] += sgn / m ** 3
    AII = A[I][:, I].tocsc()
    lu = spla.splu(AII)
    x = lu.solve(rhs)
    for _ in range(2):
        x += lu.solve(rhs - AII @ x)
    X = fix.copy()
    X[:, I] = x.T
    S = X[:, :nf]
    Aff, Afc, Acf = A[:nf, :nf], A[:nf, nf:], A[nf:, :nf]
    one = np.ones(nc)
    M = np.zeros((7, 7))
    M[:6, :6] = S @ (Aff @ S.T)
    M[:6, 6] = S @ (Afc @ one)
    M[6, :6] = one @ (Acf @ S.T)
    M[6, 6] = one @ (A[nf:, nf:] @ one)
    b = np.zeros(7)
    b[6] = F[nf:].sum()
    return M, b, S


def box_vertices_3d(x0, x1):
    """The 8 vertices of the box [x0, x1] (3-vectors) in deal.II order."""
    x0 = np.asarray(x0, dtype=np.float64)
    x1 = np.asarray(x1, dtype=np.float64)
    return np.array([[(x1 if (n >> d) & 1 else x0)[d] for d in range(3)] for n in range(8)])
