"""Uneven batches: coarse cells with their own refinement levels.

A binding of the msfec_qu_* entry points of the C-ABI.  The library keeps one
batch per level present and picks each level's local solver in C++: with
``solver=SOLVER_AUTO`` the levels the band solvers can take (3-D l <= 3, 2-D
l <= 5) are timed on up to 256 of their own cells with each FP64 solver and
tile shape (matrix-core band with a 16-, 32- or 64-slot window, vector-ALU
band, CG; candidates with a failed cell are rejected) and the fastest is kept
for the handle until ``tune()`` is called again; levels beyond the band window
run matrix-free CG.  A compute forks the levels onto streams of their own, so
they overlap.

Replaces, for cells of mixed refinement, the std::map<CellId, QBasis> loop of
QMultiscale::initialize_and_compute_basis (source/Q/q_global.cc:38-87), whose
cells all share one n_refine_local (source/Q/q_parameters.cc:229).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L
from .qbasis import _ptr

SOLVER_NAMES = {L.SOLVER_BAND_FP64: "band-fp64-valu", L.SOLVER_BAND_FP64_MFMA: "band-fp64-mfma",
                L.SOLVER_BAND_MIXED: "band-mixed", L.SOLVER_CG: "cg"}
#: the timed candidates of a level (solver and band window), in the order
#: msfec_qu_level_info reports them
CANDIDATES = ("band_mfma_w16", "band_mfma_w32", "band_mfma_w64", "band_valu", "cg")


def band_fits(dim: int, level: int) -> bool:
    """Whether the band solvers take this level: half-bandwidth + 4 <= 64."""
    ni = (1 << level) - 1
    bw = ni + 1 if dim == 2 else ni * ni + ni + 1
    return bw + 4 <= 64


class _LevelView:
    """Read-only view of one level's batch (kernel times, CG iterations)."""

    def __init__(self, lib, h):
        self._lib, self._h = lib, h

    def kernel_times(self):
        ms = (ctypes.c_float * 8)()
        n = ctypes.c_int(0)
        L.check(self._lib.msfec_q_kernel_times(self._h, ms, 8, ctypes.byref(n)))
        return [ms[i] for i in range(n.value)]

    def cg_iterations(self):
        info = L.Info()
        L.check(self._lib.msfec_q_get_info(self._h, ctypes.byref(info)))
        it = np.empty((info.n_cells, info.n_basis), dtype=np.int32)
        L.check(self._lib.msfec_q_get_cg_iterations(
            self._h, it.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return it


class UnevenBatch:
    def __init__(self, vertices, levels, dim=3, autotune=True, coef_blocks=None, timing=False,
                 solver=None, rhs_mode=L.RHS_CONST, rhs_const=1.0,
                 reaction_mode=L.REACTION_NONE, reaction_const=0.0, cg_tol=0.0, cg_max_iter=0):
        self._lib = L.load()
        self.dim = dim
        self.vertices = np.ascontiguousarray(vertices, dtype=np.float64)
        self.levels = np.ascontiguousarray(levels, dtype=np.int32)
        self.n_cells = int(self.levels.shape[0])
        self.nb = 1 << dim
        if solver is None:
            solver = L.SOLVER_AUTO if autotune else L.SOLVER_BAND_FP64_MFMA
        p = L.Params()
        p.dim, p.n_refine_local, p.coef_mode, p.rhs_mode = dim, 0, L.COEF_SCALAR_CELL, rhs_mode
        p.rhs_const, p.reaction_mode, p.reaction_const = rhs_const, reaction_mode, reaction_const
        p.solver, p.flags = solver, (L.FLAG_TIMING if timing else 0)
        p.cg_tol, p.cg_max_iter = cg_tol, cg_max_iter
        self._h = ctypes.c_void_p()
        vptr, vloc, keep = _ptr(self.vertices)
        L.check(self._lib.msfec_qu_create(
            ctypes.byref(p), self.n_cells, vptr, vloc,
            self.levels.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), ctypes.byref(self._h)))
        self.coef_offsets = np.zeros(self.n_cells + 1, dtype=np.uint64)
        self.basis_offsets = np.zeros(self.n_cells + 1, dtype=np.uint64)
        sp = ctypes.POINTER(ctypes.c_size_t)
        L.check(self._lib.msfec_qu_offsets(self._h, self.coef_offsets.ctypes.data_as(sp),
                                           self.basis_offsets.ctypes.data_as(sp)))
        self.level_set = sorted(set(self.levels.tolist()))
        self.groups = {lv: {"cells": np.nonzero(self.levels == lv)[0]} for lv in self.level_set}
        if coef_blocks is not None:
            self.set_coefficients(coef_blocks)

    # -- data -------------------------------------------------------------------
    def _flat(self, blocks):
        """Per-cell blocks (a list of arrays, or one (n_cells, m) array for a
        single level) -> the concatenated batch-order layout."""
        if hasattr(blocks, "data_ptr"):
            return blocks
        if isinstance(blocks, np.ndarray) and blocks.ndim == 1:
            flat = blocks
        else:
            flat = np.concatenate([np.asarray(b, dtype=np.float64).ravel() for b in blocks])
        flat = np.ascontiguousarray(flat, dtype=np.float64)
        if flat.size != int(self.coef_offsets[-1]):
            raise ValueError(f"expected {int(self.coef_offsets[-1])} values, got {flat.size}")
        return flat

    def set_coefficients(self, coef_blocks):
        ptr, loc, keep = _ptr(self._flat(coef_blocks))
        L.check(self._lib.msfec_qu_set_coefficient(self._h, ptr, int(self.coef_offsets[-1]), loc))

    set_coefficient = set_coefficients

    def set_rhs(self, f_blocks):
        ptr, loc, keep = _ptr(self._flat(f_blocks))
        L.check(self._lib.msfec_qu_set_rhs(self._h, ptr, int(self.coef_offsets[-1]), loc))

    def set_reaction(self, r_blocks):
        ptr, loc, keep = _ptr(self._flat(r_blocks))
        L.check(self._lib.msfec_qu_set_reaction(self._h, ptr, int(self.coef_offsets[-1]), loc))

    def tune(self):
        L.check(self._lib.msfec_qu_tune(self._h))

    # -- QBasis::run for every cell ---------------------------------------------
    def run(self, stream=None):
        if stream is not None and hasattr(stream, "cuda_stream"):
            stream = stream.cuda_stream
        L.check(self._lib.msfec_qu_compute(self._h, ctypes.c_void_p(stream or 0)))

    compute = run

    def synchronize(self):
        L.check(self._lib.msfec_qu_synchronize(self._h))

    # -- results ------------------------------------------------------------------
    def get_global_element_matrix(self):
        M = np.empty((self.n_cells, self.nb, self.nb))
        L.check(self._lib.msfec_qu_get_element(self._h, ctypes.c_void_p(M.ctypes.data), None, L.HOST))
        return M

    def get_global_element_rhs(self):
        b = np.empty((self.n_cells, self.nb))
        L.check(self._lib.msfec_qu_get_element(self._h, None, ctypes.c_void_p(b.ctypes.data), L.HOST))
        return b

    def get_cell_status(self):
        st = np.empty(self.n_cells, dtype=np.int32)
        L.check(self._lib.msfec_qu_get_cell_status(
            self._h, st.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return st

    def get_basis_flat(self):
        """Psi of every cell in the ragged batch-order layout (basis_offsets)."""
        P = np.empty(int(self.basis_offsets[-1]))
        L.check(self._lib.msfec_qu_get_basis(self._h, ctypes.c_void_p(P.ctypes.data), L.HOST))
        return P

    def get_basis(self, level):
        """(cells of that level, nb, nodes) and their indices in the batch."""
        P = self.get_basis_flat()
        idx = self.groups[level]["cells"]
        nn = ((1 << level) + 1) ** self.dim
        out = np.stack([P[int(self.basis_offsets[c]):int(self.basis_offsets[c]) + self.nb * nn]
                        for c in idx]).reshape(len(idx), self.nb, nn)
        return out, idx

    def coarse_solve(self, nc, abs_tol=1e-12, rel_tol=1e-12, max_iter=100000):
        """As QBasisBatch.coarse_solve, the cells being the nc^dim coarse cells
        in lexicographic order."""
        nnode = (nc + 1) ** self.dim
        u = np.empty(nnode)
        w = np.empty((self.n_cells, self.nb))
        it = ctypes.c_int(0)
        res = ctypes.c_double(0.0)
        L.check(self._lib.msfec_qu_coarse_solve(self._h, nc, abs_tol, rel_tol, max_iter,
                                                ctypes.c_void_p(u.ctypes.data), L.HOST,
                                                ctypes.c_void_p(w.ctypes.data), L.HOST,
                                                ctypes.byref(it), ctypes.byref(res)))
        return u, w, it.value, res.value

    # -- per-level views ------------------------------------------------------------
    def level_info(self, level):
        solver, window, n = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
        ms = (ctypes.c_float * len(CANDIDATES))()
        L.check(self._lib.msfec_qu_level_info(self._h, level, ctypes.byref(solver),
                                              ctypes.byref(window), ctypes.byref(n), ms))
        return {"solver": SOLVER_NAMES[solver.value], "window": window.value, "cells": n.value,
                "tuned_ms": {k: (round(ms[i], 5) if ms[i] >= 0 else None)
                             for i, k in enumerate(CANDIDATES)}}

    def level(self, level):
        h = ctypes.c_void_p()
        L.check(self._lib.msfec_qu_level_handle(self._h, level, ctypes.byref(h)))
        return _LevelView(self._lib, h)

    def summary(self):
        return {lv: self.level_info(lv) for lv in self.level_set}

    def close(self):
        if self._h:
            self._lib.msfec_qu_destroy(self._h)
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
