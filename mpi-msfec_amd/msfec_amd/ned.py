"""Batched 2-D Q1 x Nedelec (Q_Ned) multiscale basis: the Python face of the
``msfec_ned_*`` entry points.

``NedBasisBatch`` stands for the ``std::map<CellId, QNedBasis>`` of the
reference's Q_Ned driver at once (source/Q_Ned/q_ned_basis.cc, whose class is
three-dimensional; this is its two-dimensional form): construction from the
coarse cells, ``run()`` (QNedBasis::run, :1281-1429),
``get_global_element_matrix()`` / ``get_global_element_rhs()`` and the fine
basis coefficients.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L
from .qbasis import _ptr


def ned_sizes(n_refine_local):
    """(ns, nh, nv, N): nodes, horizontal edges, vertical edges, all fine dofs."""
    m = 1 << n_refine_local
    ns, nh, nv = (m + 1) ** 2, m * (m + 1), (m + 1) * m
    return ns, nh, nv, ns + nh + nv


class NedBasisBatch:
    def __init__(self, vertices, n_refine_local=4, full_rhs=True, timing=False,
                 solver=L.SOLVER_BAND_FP64_MFMA, max_refine=0):
        """solver: SOLVER_BAND_FP64_MFMA (FP64 LDL^T) or SOLVER_BAND_MIXED
        (FP32 LDL^T + FP64 refinement, at most max_refine passes, 0 = 3)."""
        self._lib = L.load()
        self._h = ctypes.c_void_p()
        p = L.NedParams()
        p.n_refine_local, p.full_rhs = n_refine_local, int(bool(full_rhs))
        p.flags = L.FLAG_TIMING if timing else 0
        p.solver, p.max_refine = solver, max_refine
        vptr, vloc, keep = _ptr(vertices)
        n_cells = int(np.prod(np.shape(vertices))) // 8
        L.check(self._lib.msfec_ned_create(ctypes.byref(p), n_cells, vptr, vloc,
                                           ctypes.byref(self._h)))
        self.n_cells, self.l = n_cells, n_refine_local
        self.m = 1 << n_refine_local
        self.N = ned_sizes(n_refine_local)[3]

    def set_coefficients(self, a, binv=1.0, r=0.0, f=(1.0, 1.0)):
        """Per fine cell (n_cells x 4^l; f with a trailing axis of 2); scalars
        and short vectors broadcast."""
        nf = self.m * self.m
        shp = (self.n_cells, nf)

        def full(x, shape):
            if hasattr(x, "is_cuda") and x.is_cuda:
                return x
            return np.ascontiguousarray(np.broadcast_to(np.asarray(x, dtype=np.float64),
                                                        shape))

        args = [_ptr(full(x, s)) for x, s in ((a, shp), (binv, shp), (r, shp),
                                               (f, shp + (2,)))]
        locs = {loc for _, loc, _ in args}
        if len(locs) != 1:
            raise ValueError("all coefficient arrays must live on the same side")
        L.check(self._lib.msfec_ned_set_coefficients(self._h, *[ptr for ptr, _, _ in args],
                                                     locs.pop()))

    def run(self, stream=None):
        if stream is not None and hasattr(stream, "cuda_stream"):
            stream = stream.cuda_stream
        L.check(self._lib.msfec_ned_compute(self._h, ctypes.c_void_p(stream or 0)))

    compute = run

    def synchronize(self):
        L.check(self._lib.msfec_ned_synchronize(self._h))

    def get_global_element_matrix(self):
        M = np.empty((self.n_cells, 8, 8))
        L.check(self._lib.msfec_ned_get_element(self._h, ctypes.c_void_p(M.ctypes.data), None,
                                                L.HOST))
        return M

    def get_global_element_rhs(self):
        b = np.empty((self.n_cells, 8))
        L.check(self._lib.msfec_ned_get_element(self._h, None, ctypes.c_void_p(b.ctypes.data),
                                                L.HOST))
        return b

    def get_basis(self):
        P = np.empty((self.n_cells, 8, self.N))
        L.check(self._lib.msfec_ned_get_basis(self._h, ctypes.c_void_p(P.ctypes.data), L.HOST))
        return P

    def get_cell_status(self):
        st = np.empty(self.n_cells, dtype=np.int32)
        L.check(self._lib.msfec_ned_get_cell_status(
            self._h, st.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return st

    def kernel_times(self):
        ms = (ctypes.c_float * 4)()
        n = ctypes.c_int(0)
        L.check(self._lib.msfec_ned_kernel_times(self._h, ms, 4, ctypes.byref(n)))
        return [ms[i] for i in range(n.value)]

    def close(self):
        if self._h:
            self._lib.msfec_ned_destroy(self._h)
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
