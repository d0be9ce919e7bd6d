"""Batched 3-D Nedelec x Raviart-Thomas (Ned_RT) multiscale basis: the Python
face of the ``msfec_nrt_*`` entry points.

``NedRTBasisBatch`` stands for the ``std::map<CellId, NedRTBasis>`` of the
reference's Ned_RT driver at once (source/Ned_RT/ned_rt_basis.cc): construction
from the coarse cells, ``run()`` (NedRTBasis::run, :1280-1352),
``get_global_element_matrix()`` / ``get_global_element_rhs()`` (:1262-1272),
the fine vectors of the bases and ``set_global_weights`` (:1156-1181).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L
from .qbasis import _ptr


def nrt_sizes(n_refine_local):
    """(m, ne, nf, N): fine cells per direction, fine edges, fine faces, all."""
    m = 1 << n_refine_local
    ne = 3 * m * (m + 1) ** 2
    nf = 3 * m * m * (m + 1)
    return m, ne, nf, ne + nf


class NedRTBasisBatch:
    def __init__(self, vertices, n_refine_local=2, timing=False, solver=0, max_iter=0, rtol=0.0):
        """vertices: n_cells x 8 x 3, axis-aligned boxes in deal.II vertex order.
        solver 0: the dense direct solve for l <= 2, MINRES above;
        SOLVER_CG: MINRES at every l (max_iter 0: 20000, rtol 0: 1e-14)."""
        self._lib = L.load()
        self._h = ctypes.c_void_p()
        p = L.NrtParams()
        p.n_refine_local = n_refine_local
        p.flags = L.FLAG_TIMING if timing else 0
        p.solver = solver
        p.max_iter = max_iter
        p.rtol = rtol
        vptr, vloc, keep = _ptr(vertices)
        n_cells = int(np.prod(getattr(vertices, "shape", np.shape(vertices)))) // 24
        L.check(self._lib.msfec_nrt_create(ctypes.byref(p), n_cells, vptr, vloc,
                                           ctypes.byref(self._h)))
        self.n_cells, self.l = n_cells, n_refine_local
        self.m, self.ne, self.nf, self.N = nrt_sizes(n_refine_local)

    def set_coefficients(self, ainv, b, r=None, f=None):
        """Values at the 8 Gauss points of every fine cell: ainv (n_cells x
        8^l x 8 x 6, the symmetric A^-1), b (n_cells x 8^l x 8), r (same, None
        for no reaction), f (n_cells x 8^l x 8 x 3, None for no load); numpy
        arrays broadcast, DeviceBuffers and device tensors of the full shape
        are passed as they are."""
        pts = (self.n_cells, self.m ** 3, 8)

        from .device import DeviceBuffer

        def full(x, s):
            if isinstance(x, DeviceBuffer) or (hasattr(x, "is_cuda") and x.is_cuda):
                return x
            return np.ascontiguousarray(np.broadcast_to(np.asarray(x, dtype=np.float64), s))

        args, keep, locs = [], [], set()
        for x, s in ((ainv, pts + (6,)), (b, pts), (r, pts), (f, pts + (3,))):
            if x is None:
                args.append(None)
                continue
            ptr, loc, k = _ptr(full(x, s))
            args.append(ptr)
            keep.append(k)
            locs.add(loc)
        if len(locs) != 1:
            raise ValueError("all coefficients must live on the same side")
        L.check(self._lib.msfec_nrt_set_coefficients(self._h, *args, locs.pop()))

    def run(self, stream=None):
        if stream is not None and hasattr(stream, "cuda_stream"):
            stream = stream.cuda_stream
        L.check(self._lib.msfec_nrt_compute(self._h, ctypes.c_void_p(stream or 0)))

    compute = run

    def synchronize(self):
        L.check(self._lib.msfec_nrt_synchronize(self._h))

    def get_global_element_matrix(self):
        M = np.empty((self.n_cells, 18, 18))
        L.check(self._lib.msfec_nrt_get_element(self._h, ctypes.c_void_p(M.ctypes.data), None,
                                                L.HOST))
        return M

    def get_global_element_rhs(self):
        b = np.empty((self.n_cells, 18))
        L.check(self._lib.msfec_nrt_get_element(self._h, None, ctypes.c_void_p(b.ctypes.data),
                                                L.HOST))
        return b

    def get_basis(self):
        """The full fine vectors of the 18 bases, n_cells x 18 x N."""
        X = np.empty((self.n_cells, 18, self.N))
        L.check(self._lib.msfec_nrt_get_basis(self._h, ctypes.c_void_p(X.ctypes.data), L.HOST))
        return X

    def set_global_weights(self, w):
        """(sigma n_cells x ne, u n_cells x nf) of the weights w (n_cells x 18)."""
        w = np.ascontiguousarray(np.asarray(w, dtype=np.float64).reshape(self.n_cells, 18))
        sig = np.empty((self.n_cells, self.ne))
        u = np.empty((self.n_cells, self.nf))
        L.check(self._lib.msfec_nrt_reconstruct(self._h, ctypes.c_void_p(w.ctypes.data), L.HOST,
                                                ctypes.c_void_p(sig.ctypes.data),
                                                ctypes.c_void_p(u.ctypes.data), L.HOST))
        return sig, u

    def get_cell_status(self):
        st = np.empty(self.n_cells, dtype=np.int32)
        L.check(self._lib.msfec_nrt_get_cell_status(
            self._h, st.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return st

    def get_iterations(self):
        """MINRES iterations of every basis solve, n_cells x 18 (MINRES path only)."""
        its = np.empty((self.n_cells, 18), dtype=np.int32)
        L.check(self._lib.msfec_nrt_get_iterations(
            self._h, its.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return its

    def kernel_times(self):
        """Mean ms of the timed computes: [whole compute] on the dense path,
        [whole, local matrices + preconditioner, MINRES, element matrix] on the
        MINRES path."""
        ms = (ctypes.c_float * 4)()
        n = ctypes.c_int(0)
        L.check(self._lib.msfec_nrt_kernel_times(self._h, ms, 4, ctypes.byref(n)))
        return [ms[i] for i in range(n.value)]

    def close(self):
        if self._h:
            self._lib.msfec_nrt_destroy(self._h)
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
