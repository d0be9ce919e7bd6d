"""Batched multiscale Q1 basis: the Python face of the C-ABI.

``QBasisBatch`` stands for the whole ``std::map<CellId, QBasis>`` of
``QMultiscale`` (source/Q/q_global.cc:38-87) at once: construction from the
coarse cells plus parameters, ``run()`` (QBasis::run, source/Q/q_basis.cc:
518-603), ``get_global_element_matrix()`` / ``get_global_element_rhs()``
(:361-373), the basis coefficients (``solution_vector``) and
``set_global_weights()`` (:394-416).  Arrays may be numpy (host) or torch
tensors on the GPU (device pointers are passed through, no copy to the host).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L


def _ptr(a):
    """(pointer, loc, keepalive) for a numpy array, a torch tensor or a
    DeviceBuffer."""
    from .device import DeviceBuffer
    if isinstance(a, DeviceBuffer):
        return a.ptr, L.DEVICE, a
    if hasattr(a, "data_ptr") and hasattr(a, "is_cuda"):
        if a.is_cuda:
            if not a.is_contiguous():
                raise ValueError("device tensors must be contiguous")
            return ctypes.c_void_p(a.data_ptr()), L.DEVICE, a
        a = a.numpy()
    arr = np.ascontiguousarray(a, dtype=np.float64)
    return ctypes.c_void_p(arr.ctypes.data), L.HOST, arr


def _out_ptr(a, shape):
    """(pointer, loc, keepalive) for an output buffer the library writes into.
    No copy is ever made: a host array must be a C-contiguous float64 numpy
    array (or CPU torch tensor) of the expected size, a device tensor a
    contiguous float64 one; anything else is an error, not a silent copy."""
    n = int(np.prod(shape))
    from .device import DeviceBuffer
    if isinstance(a, DeviceBuffer):
        if a.nbytes != 8 * n:
            raise ValueError(f"out must hold {n} float64 values")
        return a.ptr, L.DEVICE, a
    if hasattr(a, "data_ptr") and hasattr(a, "is_cuda"):
        import torch
        if a.dtype != torch.float64 or not a.is_contiguous() or a.numel() != n:
            raise ValueError(f"out must be a contiguous float64 tensor of {n} elements")
        return ctypes.c_void_p(a.data_ptr()), (L.DEVICE if a.is_cuda else L.HOST), a
    if (not isinstance(a, np.ndarray) or a.dtype != np.float64 or not a.flags.c_contiguous
            or not a.flags.writeable or a.size != n):
        raise ValueError(f"out must be a writeable C-contiguous float64 array of {n} elements")
    return ctypes.c_void_p(a.ctypes.data), L.HOST, a


class QBasisBatch:
    def __init__(self, vertices, dim=3, n_refine_local=3, coef_mode=L.COEF_SCALAR_CELL,
                 rhs_mode=L.RHS_CONST, rhs_const=1.0, reaction_mode=L.REACTION_NONE,
                 reaction_const=0.0, solver=L.SOLVER_BAND_FP64, timing=False, max_refine=0,
                 cg_tol=0.0, cg_max_iter=0, band_window=0):
        self._lib = L.load()
        self._h = ctypes.c_void_p()
        p = L.Params()
        p.dim, p.n_refine_local, p.coef_mode, p.rhs_mode = dim, n_refine_local, coef_mode, rhs_mode
        p.rhs_const, p.reaction_mode, p.reaction_const = rhs_const, reaction_mode, reaction_const
        p.solver, p.flags = solver, (L.FLAG_TIMING if timing else 0)
        p.max_refine = max_refine
        p.cg_tol, p.cg_max_iter = cg_tol, cg_max_iter
        p.band_window = band_window
        vptr, vloc, keep = _ptr(vertices)
        nv = 1 << dim
        n_cells = int(np.prod(vertices.shape)) // (nv * dim)
        L.check(self._lib.msfec_q_create(ctypes.byref(p), n_cells, vptr, vloc, ctypes.byref(self._h)))
        info = L.Info()
        L.check(self._lib.msfec_q_get_info(self._h, ctypes.byref(info)))
        self.info = info
        self.n_cells, self.dim, self.nb, self.nn = info.n_cells, dim, info.n_basis, info.n_fine_nodes

    # -- data ---------------------------------------------------------------
    def set_coefficient(self, coef):
        ptr, loc, keep = _ptr(coef)
        L.check(self._lib.msfec_q_set_coefficient(self._h, ptr, self.info.coef_count, loc))

    def set_rhs(self, f):
        ptr, loc, keep = _ptr(f)
        L.check(self._lib.msfec_q_set_rhs(self._h, ptr, self.info.rhs_count, loc))

    def set_reaction(self, r):
        ptr, loc, keep = _ptr(r)
        L.check(self._lib.msfec_q_set_reaction(self._h, ptr, self.info.reaction_count, loc))

    # -- QBasis::run ----------------------------------------------------------
    def run(self, stream=None):
        """Enqueue the basis construction of all cells on ``stream`` (a HIP
        stream handle as int / torch.cuda.Stream / None for the default)."""
        if stream is not None and hasattr(stream, "cuda_stream"):
            stream = stream.cuda_stream
        L.check(self._lib.msfec_q_compute(self._h, ctypes.c_void_p(stream or 0)))

    compute = run

    def synchronize(self):
        L.check(self._lib.msfec_q_synchronize(self._h))

    # -- getters --------------------------------------------------------------
    def get_global_element_matrix(self, out=None):
        out = np.empty((self.n_cells, self.nb, self.nb)) if out is None else out
        ptr, loc, keep = _out_ptr(out, (self.n_cells, self.nb, self.nb))
        L.check(self._lib.msfec_q_get_element(self._h, ptr, None, loc))
        return out

    def get_global_element_rhs(self, out=None):
        out = np.empty((self.n_cells, self.nb)) if out is None else out
        ptr, loc, keep = _out_ptr(out, (self.n_cells, self.nb))
        L.check(self._lib.msfec_q_get_element(self._h, None, ptr, loc))
        return out

    def get_basis(self, out=None):
        out = np.empty((self.n_cells, self.nb, self.nn)) if out is None else out
        ptr, loc, keep = _out_ptr(out, (self.n_cells, self.nb, self.nn))
        L.check(self._lib.msfec_q_get_basis(self._h, ptr, loc))
        return out

    def get_cell_status(self):
        st = np.zeros(self.n_cells, dtype=np.int32)
        L.check(self._lib.msfec_q_get_cell_status(self._h, st.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return st

    def set_global_weights(self, weights, out=None):
        """Fine solution u|_K = sum_k w_k psi_k for every cell."""
        out = np.empty((self.n_cells, self.nn)) if out is None else out
        wp, wl, wk = _ptr(weights)
        op, ol, ok = _out_ptr(out, (self.n_cells, self.nn))
        L.check(self._lib.msfec_q_reconstruct(self._h, wp, wl, op, ol))
        return out

    def coarse_solve(self, nc, abs_tol=1e-12, rel_tol=1e-12, max_iter=100000):
        """Assemble and solve the coarse problem (nc^dim cells of [0,1]^dim in
        lexicographic order).  Returns (u_nodes, weights, iterations, residual)."""
        nnode = (nc + 1) ** self.dim
        u = np.empty(nnode)
        w = np.empty((self.n_cells, self.nb))
        it = ctypes.c_int(0)
        res = ctypes.c_double(0.0)
        L.check(self._lib.msfec_coarse_solve(self._h, nc, abs_tol, rel_tol, max_iter,
                                             ctypes.c_void_p(u.ctypes.data), L.HOST,
                                             ctypes.c_void_p(w.ctypes.data), L.HOST,
                                             ctypes.byref(it), ctypes.byref(res)))
        return u, w, it.value, res.value

    def adjoint_gradient(self, nc, u_target, rel_tol=1e-14, max_iter=100000, out=None):
        """(J, dJ/dg) for J = 1/2 |u_H - u_target|^2 over the free coarse nodes
        and the log-coefficient g of every fine cell (a = exp(g)); the batch
        must have been computed (run()) with solver BAND_FP64_MFMA or
        BAND_MIXED on the nc^dim structured grid of coarse_solve.  ``out``
        (n_cells x fine cells float64: a host array, a DeviceBuffer or a
        device tensor) receives the gradient in place."""
        ut = np.ascontiguousarray(u_target, dtype=np.float64)
        shape = (self.n_cells, self.info.n_fine_cells)
        grad = np.empty(shape) if out is None else out
        gptr, gloc, keep = _out_ptr(grad, shape)
        J = ctypes.c_double(0.0)
        L.check(self._lib.msfec_q_adjoint_gradient(
            self._h, nc, ctypes.c_void_p(ut.ctypes.data), L.HOST, rel_tol, max_iter,
            ctypes.byref(J), gptr, gloc))
        return J.value, grad

    def cg_iterations(self):
        """(n_cells, nb) CG iterations of the last compute (SOLVER_CG)."""
        it = np.empty((self.n_cells, self.nb), dtype=np.int32)
        L.check(self._lib.msfec_q_get_cg_iterations(
            self._h, it.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return it

    def kernel_times(self):
        ms = (ctypes.c_float * 8)()
        n = ctypes.c_int(0)
        L.check(self._lib.msfec_q_kernel_times(self._h, ms, 8, ctypes.byref(n)))
        return [ms[i] for i in range(n.value)]

    def close(self):
        if self._h:
            self._lib.msfec_q_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
