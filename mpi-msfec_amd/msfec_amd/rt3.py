"""Batched 3-D Raviart-Thomas x DQ0 (RT_DQ) multiscale basis: the Python face
of the ``msfec_rt3_*`` entry points.

``RTDQ3BasisBatch`` stands for the ``std::map<CellId, RTDQBasis>`` of the
reference's RT_DQ driver (source/RT_DQ/rt_dq_basis.cc, three-dimensional):
construction from the coarse box cells, ``run()`` (RTDQBasis::run,
:1114-1170), ``get_global_element_matrix()`` / ``get_global_element_rhs()``
(:1096-1106), the fine fluxes of the 6 face bases and ``set_global_weights``
(:1011-1036).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L
from .qbasis import _ptr


def rt3_sizes(n_refine_local):
    """(m, nf, nc): fine cells per direction, fine faces, fine cells."""
    m = 1 << n_refine_local
    return m, 3 * m * m * (m + 1), m ** 3


class RTDQ3BasisBatch:
    def __init__(self, vertices, n_refine_local=2, timing=False, max_iter=0, rtol=0.0):
        """vertices: n_cells x 8 x 3, axis-aligned boxes in deal.II vertex order.
        max_iter 0: 20000 CG iterations; rtol 0: 1e-13."""
        self._lib = L.load()
        self._h = ctypes.c_void_p()
        p = L.Rt3Params()
        p.n_refine_local = n_refine_local
        p.flags = L.FLAG_TIMING if timing else 0
        p.max_iter = max_iter
        p.rtol = rtol
        vptr, vloc, keep = _ptr(vertices)
        n_cells = int(np.prod(getattr(vertices, "shape", np.shape(vertices)))) // 24
        L.check(self._lib.msfec_rt3_create(ctypes.byref(p), n_cells, vptr, vloc,
                                           ctypes.byref(self._h)))
        self.n_cells, self.l = n_cells, n_refine_local
        self.m, self.nf, self.nc = rt3_sizes(n_refine_local)

    def set_coefficients(self, ainv, f=None):
        """ainv: (xx, xy, xz, yy, yz, zz) of A^-1 at the 8 Gauss points of every
        fine cell (n_cells x 8^l x 8 x 6, or broadcastable); f: the load at the
        same points (n_cells x 8^l x 8), None for zero."""
        pts = (self.n_cells, self.m ** 3, 8)
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(ainv, dtype=np.float64), pts + (6,)))
        fp = None
        if f is not None:
            fp = np.ascontiguousarray(np.broadcast_to(np.asarray(f, dtype=np.float64), pts))
        L.check(self._lib.msfec_rt3_set_coefficients(
            self._h, ctypes.c_void_p(a.ctypes.data),
            None if fp is None else ctypes.c_void_p(fp.ctypes.data), L.HOST))

    def run(self, stream=None):
        if stream is not None and hasattr(stream, "cuda_stream"):
            stream = stream.cuda_stream
        L.check(self._lib.msfec_rt3_compute(self._h, ctypes.c_void_p(stream or 0)))

    compute = run

    def synchronize(self):
        L.check(self._lib.msfec_rt3_synchronize(self._h))

    def get_global_element_matrix(self):
        M = np.empty((self.n_cells, 7, 7))
        L.check(self._lib.msfec_rt3_get_element(self._h, ctypes.c_void_p(M.ctypes.data), None,
                                                L.HOST))
        return M

    def get_global_element_rhs(self):
        b = np.empty((self.n_cells, 7))
        L.check(self._lib.msfec_rt3_get_element(self._h, None, ctypes.c_void_p(b.ctypes.data),
                                                L.HOST))
        return b

    def get_basis(self):
        """The fine fluxes of the 6 face bases, n_cells x 6 x nf (their u is 1)."""
        S = np.empty((self.n_cells, 6, self.nf))
        L.check(self._lib.msfec_rt3_get_basis(self._h, ctypes.c_void_p(S.ctypes.data), L.HOST))
        return S

    def set_global_weights(self, w):
        """(fluxes sum_{i<6} w_i sigma_i, n_cells x nf; cell values w_6, n_cells x nc)."""
        w = np.asarray(w, dtype=np.float64).reshape(self.n_cells, 7)
        S = self.get_basis()
        return np.einsum("ci,cif->cf", w[:, :6], S), np.repeat(w[:, 6:], self.nc, axis=1)

    def get_cell_status(self):
        st = np.empty(self.n_cells, dtype=np.int32)
        L.check(self._lib.msfec_rt3_get_cell_status(
            self._h, st.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return st

    def get_iterations(self):
        its = np.empty((self.n_cells, 6), dtype=np.int32)
        L.check(self._lib.msfec_rt3_get_iterations(
            self._h, its.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return its

    def kernel_times(self):
        """Mean ms of the timed computes: [whole, face masses + preconditioner,
        CG, element matrix]."""
        ms = (ctypes.c_float * 4)()
        n = ctypes.c_int(0)
        L.check(self._lib.msfec_rt3_kernel_times(self._h, ms, 4, ctypes.byref(n)))
        return [ms[i] for i in range(n.value)]

    def close(self):
        if self._h:
            self._lib.msfec_rt3_destroy(self._h)
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
