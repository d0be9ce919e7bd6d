"""Uneven batches: coarse cells with their own refinement levels.

One handle per refinement level present.  The local solver of a level is the
FP64 band solver on the matrix cores where the 64-row window holds the
interior band (3-D l <= 3, 2-D l <= 5) and matrix-free CG beyond it.  With
``autotune=True`` the levels the band solvers can take are timed once on a
sample of their cells with each FP64 solver (matrix-core band, vector-ALU
band, CG) and the fastest is kept (cached per dimension and level for the
process).  All level handles run on the same stream.
"""
from __future__ import annotations

import time

import numpy as np

from . import _lib as L
from .qbasis import QBasisBatch

_TUNED: dict = {}


def band_fits(dim: int, level: int) -> bool:
    ni = (1 << level) - 1
    bw = ni + 1 if dim == 2 else ni * ni + ni + 1
    return bw + 1 <= 64


def _time_solver(verts, blocks, dim, level, solver, reps=3):
    with QBasisBatch(verts, dim=dim, n_refine_local=level, solver=solver) as qb:
        qb.set_coefficient(blocks)
        qb.run()
        qb.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            qb.run()
        qb.synchronize()
        return (time.perf_counter() - t0) / reps


#: the FP64 local solvers a level may use, in the order ties are broken
CANDIDATES = (("band_mfma", L.SOLVER_BAND_FP64_MFMA), ("band_valu", L.SOLVER_BAND_FP64),
              ("cg", L.SOLVER_CG))


def choose_solver(dim, level, verts=None, blocks=None, autotune=True):
    """The solver for one level: the matrix-core band solver where the band
    fits, CG otherwise; when the band fits and autotune is on, the fastest of
    the FP64 solvers (matrix-core band, vector-ALU band, CG) on the given
    cells."""
    if not band_fits(dim, level):
        return L.SOLVER_CG
    if not autotune or verts is None:
        return L.SOLVER_BAND_FP64_MFMA
    key = (dim, level)
    if key not in _TUNED:
        times = {name: _time_solver(verts, blocks, dim, level, sv) for name, sv in CANDIDATES}
        best = min(CANDIDATES, key=lambda c: times[c[0]])
        _TUNED[key] = {"solver": best[1], "sample_cells": int(verts.shape[0]),
                       **{f"{name}_s": t for name, t in times.items()}}
    return _TUNED[key]["solver"]


def tuning_table():
    return dict(_TUNED)


class UnevenBatch:
    def __init__(self, vertices, levels, dim=3, autotune=True, coef_blocks=None,
                 sample=256, timing=False):
        self.dim = dim
        self.vertices = np.ascontiguousarray(vertices, dtype=np.float64)
        self.levels = np.asarray(levels, dtype=int)
        self.n_cells = self.levels.shape[0]
        self.groups = {}
        for lv in sorted(set(self.levels.tolist())):
            idx = np.nonzero(self.levels == lv)[0]
            self.groups[lv] = {"cells": idx}
        if coef_blocks is not None:
            self._plan(coef_blocks, autotune, sample, timing)

    def _plan(self, coef_blocks, autotune, sample, timing):
        for lv, g in self.groups.items():
            idx = g["cells"]
            blocks = np.stack([np.asarray(coef_blocks[c], dtype=np.float64) for c in idx])
            n_s = min(sample, idx.shape[0])
            solver = choose_solver(self.dim, lv, self.vertices[idx[:n_s]], blocks[:n_s],
                                   autotune)
            g["solver"] = solver
            g["blocks"] = blocks
            g["qb"] = QBasisBatch(self.vertices[idx], dim=self.dim, n_refine_local=lv,
                                  solver=solver, timing=timing)
            g["qb"].set_coefficient(blocks)

    def set_coefficients(self, coef_blocks, autotune=True, sample=256, timing=False):
        self._plan(coef_blocks, autotune, sample, timing)

    def run(self, stream=None):
        for g in self.groups.values():
            g["qb"].run(stream)

    def synchronize(self):
        for g in self.groups.values():
            g["qb"].synchronize()

    def get_global_element_matrix(self):
        nb = 1 << self.dim
        M = np.empty((self.n_cells, nb, nb))
        for g in self.groups.values():
            M[g["cells"]] = g["qb"].get_global_element_matrix()
        return M

    def get_global_element_rhs(self):
        nb = 1 << self.dim
        b = np.empty((self.n_cells, nb))
        for g in self.groups.values():
            b[g["cells"]] = g["qb"].get_global_element_rhs()
        return b

    def get_cell_status(self):
        st = np.empty(self.n_cells, dtype=np.int32)
        for g in self.groups.values():
            st[g["cells"]] = g["qb"].get_cell_status()
        return st

    def get_basis(self, level):
        """(cells of that level, nb, nodes) and their indices in the batch."""
        g = self.groups[level]
        return g["qb"].get_basis(), g["cells"]

    def summary(self):
        return {int(lv): {"cells": int(g["cells"].shape[0]),
                          "solver": {0: "band-fp64-valu", 1: "band-fp64-mfma", 2: "band-mixed",
                                     3: "cg"}[g["solver"]]}
                for lv, g in self.groups.items()}

    def close(self):
        for g in self.groups.values():
            if "qb" in g:
                g["qb"].close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
