"""Batched 2-D Raviart-Thomas x DQ0 (RT_DQ) multiscale basis: the Python face
of the ``msfec_rt_*`` entry points.

``RTDQBasisBatch`` stands for the ``std::map<CellId, RTDQBasis>`` of the
reference's RT_DQ driver at once (source/RT_DQ/rt_dq_basis.cc, whose class is
three-dimensional; this is its two-dimensional form): construction from the
coarse cells, ``run()`` (RTDQBasis::run, :1114-1170),
``get_global_element_matrix()`` / ``get_global_element_rhs()`` (:1096-1106),
the fine fluxes of the bases and ``set_global_weights`` (:1011-1036).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L
from .qbasis import _ptr


def rt_sizes(n_refine_local):
    """(nv, ne, nc): vertical edges, all fine edges, fine cells."""
    m = 1 << n_refine_local
    nv = m * (m + 1)
    return nv, 2 * nv, m * m


class RTDQBasisBatch:
    def __init__(self, vertices, n_refine_local=4, timing=False, coef_mode=L.COEF_SCALAR_CELL):
        """coef_mode: COEF_SCALAR_CELL (a per fine cell) or COEF_TENSOR_QP
        ((xx, xy, yy) of A at the 4 Gauss points of every fine cell)."""
        self._lib = L.load()
        self._h = ctypes.c_void_p()
        p = L.RtParams()
        p.n_refine_local = n_refine_local
        p.flags = L.FLAG_TIMING if timing else 0
        p.coef_mode = coef_mode
        self.coef_mode = coef_mode
        vptr, vloc, keep = _ptr(vertices)
        n_cells = int(np.prod(np.shape(vertices))) // 8
        L.check(self._lib.msfec_rt_create(ctypes.byref(p), n_cells, vptr, vloc,
                                          ctypes.byref(self._h)))
        self.n_cells, self.l = n_cells, n_refine_local
        self.m = 1 << n_refine_local
        _, self.ne, self.nc = rt_sizes(n_refine_local)

    def set_coefficients(self, a, f=None):
        """Diffusion a (n_cells x 4^l, or n_cells x 4^l x 4 x 3 in tensor mode)
        and load f per fine cell (n_cells x 4^l); scalars broadcast, f None
        means 1."""
        shp = (self.n_cells, self.m * self.m)
        ashp = shp + ((4, 3) if self.coef_mode == L.COEF_TENSOR_QP else ())

        def full(x, s):
            if hasattr(x, "is_cuda") and x.is_cuda:
                return x
            return np.ascontiguousarray(np.broadcast_to(np.asarray(x, dtype=np.float64), s))

        aptr, aloc, akeep = _ptr(full(a, ashp))
        if f is None:
            L.check(self._lib.msfec_rt_set_coefficients(self._h, aptr, None, aloc))
            return
        fptr, floc, fkeep = _ptr(full(f, shp))
        if aloc != floc:
            raise ValueError("a and f must live on the same side")
        L.check(self._lib.msfec_rt_set_coefficients(self._h, aptr, fptr, aloc))

    def run(self, stream=None):
        if stream is not None and hasattr(stream, "cuda_stream"):
            stream = stream.cuda_stream
        L.check(self._lib.msfec_rt_compute(self._h, ctypes.c_void_p(stream or 0)))

    compute = run

    def synchronize(self):
        L.check(self._lib.msfec_rt_synchronize(self._h))

    def get_global_element_matrix(self):
        M = np.empty((self.n_cells, 5, 5))
        L.check(self._lib.msfec_rt_get_element(self._h, ctypes.c_void_p(M.ctypes.data), None,
                                               L.HOST))
        return M

    def get_global_element_rhs(self):
        b = np.empty((self.n_cells, 5))
        L.check(self._lib.msfec_rt_get_element(self._h, None, ctypes.c_void_p(b.ctypes.data),
                                               L.HOST))
        return b

    def get_basis(self):
        """Fine fluxes, n_cells x 4 x ne (the u part of every basis is 1)."""
        S = np.empty((self.n_cells, 4, self.ne))
        L.check(self._lib.msfec_rt_get_basis(self._h, ctypes.c_void_p(S.ctypes.data), L.HOST))
        return S

    def set_global_weights(self, w):
        """(sigma n_cells x ne, u n_cells x 4^l) of the weights w (n_cells x 5)."""
        w = np.ascontiguousarray(np.asarray(w, dtype=np.float64).reshape(self.n_cells, 5))
        sig = np.empty((self.n_cells, self.ne))
        u = np.empty((self.n_cells, self.nc))
        L.check(self._lib.msfec_rt_reconstruct(self._h, ctypes.c_void_p(w.ctypes.data), L.HOST,
                                               ctypes.c_void_p(sig.ctypes.data),
                                               ctypes.c_void_p(u.ctypes.data), L.HOST))
        return sig, u

    def get_cell_status(self):
        st = np.empty(self.n_cells, dtype=np.int32)
        L.check(self._lib.msfec_rt_get_cell_status(
            self._h, st.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return st

    def kernel_times(self):
        ms = (ctypes.c_float * 4)()
        n = ctypes.c_int(0)
        L.check(self._lib.msfec_rt_kernel_times(self._h, ms, 4, ctypes.byref(n)))
        return [ms[i] for i in range(n.value)]

    def close(self):
        if self._h:
            self._lib.msfec_rt_destroy(self._h)
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
