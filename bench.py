#!/usr/bin/env python3
"""Benchmark: multiscale Q1 basis construction on one MI355X.

Workload (the metric's configuration): 3-D Q1 pairing, 16^3 = 4096 coarse
cells of the unit cube, each refined 3 levels (8^3 fine cells, 343 interior
unknowns, 8 basis right-hand sides), A = exp(g) piecewise constant per fine
cell with g a seeded stationary Gaussian field (sigma 1.5, correlation length
1/64) on the 128^3 fine grid, f = 1.  One step = the whole batch: fine
assembly + lifting, the 8 local solves per cell, Psi, the 8x8 element matrices
and element right-hand sides (QBasis::run of every cell).  Inputs are resident
in HBM before the timed region.

Prints one JSON line (see the repository's benchmark contract).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "mpi-msfec_amd"))

import numpy as np  # noqa: E402

DIM, NREF, NC = 3, 3, 16
SIGMA, CORR, SEED = 1.5, 1.0 / 64, 20261202
FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 (vector and matrix)
FP32_MFMA_PEAK_TFLOPS = 157.3  # MI355X FP32 matrix cores, dense
# local solvers: name -> (solver id, dominant kernel, its arithmetic).  The
# value line is HEADLINE, the fastest path that meets the 1e-10 bound on this
# workload; the others, the mixed-precision one among them, are timed beside
# it and reported with their ratio to it.
SOLVERS = {"band-fp64-mfma": (1, "q_blk_factor_kernel<double>", "f64"),
           "band-mixed": (2, "q_blk_factor_kernel<float>", "f32"),
           "band-fp64-valu": (0, "q_band_kernel", "f64")}
SOLVER_TEXT = {"band-mixed": "banded Cholesky FP32 on the matrix cores + iterative refinement "
                             "with FP64 residuals until (|dx|/|x|)^2 <= 1e-11 per cell",
               "band-fp64-valu": "banded Cholesky FP64 on the vector ALUs",
               "band-fp64-mfma": "banded Cholesky FP64 on the matrix cores"}
HEADLINE = "band-fp64-mfma"
TRAFFIC_FILE = os.path.join(REPO, "profiles", "hbm_traffic.json")
HBM_PEAK_GBS = 8000.0     # MI355X HBM3E


def band_flops_per_cell(n: int, p: int, nrhs: int) -> float:
    """Algorithmic FP64 flops of the banded LDL^T (rank-1 updates inside the
    band), forward and backward solves with nrhs right-hand sides."""
    f = 0.0
    for j in range(n):
        b = min(p, n - 1 - j)
        f += b * (b + 1)          # symmetric rank-1 update (b(b+1)/2 FMAs)
        f += b                    # column scaling
        f += 2.0 * 2 * b * nrhs   # forward + backward substitution
    return f


def dense_flops_per_cell(n: int, nrhs: int) -> float:
    """Dense Cholesky + solves: n^3/3 + 2 r n^2."""
    return n ** 3 / 3.0 + 2.0 * nrhs * n * n


def band_bytes_per_cell(n: int, p: int, nlow: int, nrhs: int = 8) -> float:
    """Compulsory HBM bytes of the FP64 band kernel per cell: stencil rows and
    lifting right-hand sides in, the band factor L written once and read once
    by the backward solve, forward results written and read, solution out."""
    return 8.0 * (n * nlow + n * nrhs + 2 * n * (p + 1) + 2 * n * nrhs + n * nrhs)


def mixed_factor_bytes_per_cell(n: int, p: int, nlow: int, nrhs: int = 8) -> float:
    """The same for the FP32 factor kernel of the mixed solver: FP64 rows and
    right-hand sides in, the FP32 factor and forward results written and read
    back, the FP64 first solution out."""
    return 8.0 * (n * nlow + n * nrhs) + 4.0 * (2 * n * (p + 1) + 2 * n * nrhs) + 8.0 * n * nrhs


def compulsory_bytes_per_cell(dim: int, nref: int, nb: int = 0) -> float:
    """HBM bytes a basis construction must move per coarse cell: the fine-cell
    coefficients and the vertices in, Psi on all fine nodes, the element
    matrix and right-hand side out (51 520 B for 3-D, l = 3)."""
    N = 1 << nref
    nb = nb or (1 << dim)
    return 8.0 * (N ** dim + nb * dim + nb * (N + 1) ** dim + nb * nb + nb)


def measured_traffic(kernel: str, n_cells: int):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE and WRITE_SIZE runs, each scaled by its reading for
    a copy of known size), or None when no summary matches this kernel and
    batch size."""
    try:
        with open(TRAFFIC_FILE) as fh:
            t = json.load(fh)
        k = t["kernels"][kernel]
        if t["n_cells"] != n_cells:
            return None
        return round((k["fetch_kib"] * t["fetch_scale"] + k["write_kib"] * t["write_scale"]) * 1024)
    except (OSError, KeyError, ValueError):
        return None


def host_cores() -> int:
    """The host cores this process may run on (its affinity mask), every one
    of them used by the CPU baseline like mpirun -n nproc -- capped only by
    OMP_NUM_THREADS where the machine sets it, which is the share of the host
    granted to a job on one GPU."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    cap = os.environ.get("OMP_NUM_THREADS", "")
    return min(n, int(cap)) if cap.isdigit() and int(cap) > 0 else n


def host_cores_text(used: int) -> str:
    try:
        mask = len(os.sched_getaffinity(0))
    except AttributeError:
        mask = os.cpu_count() or 1
    return (f"{used} threads ({mask} cores in the affinity mask, "
            f"OMP_NUM_THREADS={os.environ.get('OMP_NUM_THREADS', 'unset')})")


def rel_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    num = np.linalg.norm((a - b).reshape(a.shape[0], -1), axis=1)
    den = np.linalg.norm(b.reshape(b.shape[0], -1), axis=1)
    return float(np.max(num / np.maximum(den, 1e-300)))


def timed_reps(fn, reps: int):
    """Walls of `reps` calls of fn after one untimed warm-up call."""
    fn()
    walls = []
    for _ in range(reps):
        t0 = time.perf_counter()
        out = fn()
        walls.append(time.perf_counter() - t0)
    return walls, out


def cpu_baseline(verts, coef, budget_s: float = 20.0):
    """The reference algorithm on the host cores: the oracle's CG + SSOR(1.6)
    path (abs tol 1e-12, maxit 1000 -- QBasis::solve_iterative, the default of
    every example parameter file), OpenMP over cells like mpirun ranks; one
    warm-up, then the best and the median of 5 runs over a seeded sample of
    about budget_s / 2 CPU-seconds.  Beside it, the direct variant (dense FP64
    Cholesky of the interior system, the 'use direct solver basis' analogue)
    on a smaller sample, best of 3."""
    from oracle import oracle as O
    O.build()
    cores = host_cores()
    p = O.make_problem(DIM, NREF)
    rng = np.random.default_rng(7)
    order = rng.permutation(verts.shape[0])

    def sized(solver, budget):
        # calibrate on a few cells, then size the sample to ~budget CPU-seconds
        cal = order[:cores]
        t0 = time.perf_counter()
        O.basis_batch(p, verts[cal], coef=coef[cal], solver=solver, n_threads=cores,
                      want_psi=False)
        per_cell_cpu = (time.perf_counter() - t0) * cores / len(cal)
        n = int(min(verts.shape[0], max(cores, budget / max(per_cell_cpu, 1e-6))))
        return order[:max(cores, (n // cores) * cores)]

    samp = sized(O.SOLVER_CG_SSOR, budget_s / 2)
    walls, r = timed_reps(lambda: O.basis_batch(p, verts[samp], coef=coef[samp],
                                                solver=O.SOLVER_CG_SSOR, n_threads=cores,
                                                want_psi=False), 5)
    dsamp = sized(O.SOLVER_DIRECT, budget_s / 6)
    dwalls, _ = timed_reps(lambda: O.basis_batch(p, verts[dsamp], coef=coef[dsamp],
                                                 solver=O.SOLVER_DIRECT, n_threads=cores,
                                                 want_psi=False), 3)
    n_sample = len(samp)
    return {
        "value": n_sample / min(walls),
        "unit": "cells/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{n_sample} of the 4096 cells (seeded), CG+SSOR(1.6) abs tol 1e-12 per "
                  f"basis, OpenMP on {host_cores_text(cores)}, max CG iterations "
                  f"{int(r['iters'].max())}; best of 5 runs after a warm-up, walls "
                  + ", ".join(f"{w:.2f}" for w in walls) + " s",
        "median_value": n_sample / float(np.median(walls)),
        "direct_value": len(dsamp) / min(dwalls),
        "direct_sample": f"{len(dsamp)} cells, dense FP64 Cholesky of K_II (the direct "
                         f"solver), best of 3 after a warm-up",
    }


NED_NREF, NED_NC, NED_SIGMA = 4, 32, 1.0


def ned_config(steps: int, warmup: int, check: bool, cpu: bool):
    """The 2-D Q1 x Nedelec configuration, timed beside the headline: 32 x 32
    coarse cells of the unit square, 16 x 16 fine cells each (289 nodes + 544
    edges, 705 interior unknowns, 4 + 4 bases), a = exp(g) with g iid
    N(0, 1) per fine cell (seed 20261202), B = 1, R = 0, f = (1, 1)."""
    from msfec_amd import NedBasisBatch, _lib as L, fields as F
    verts = F.coarse_vertices(2, NED_NC)
    a = F.lognormal_coefficient(2, NED_NC, NED_NREF, NED_SIGMA, 0.0, SEED)
    n_cells = verts.shape[0]
    # the configuration's precision recipe: FP32 LDL^T + FP64 refinement,
    # timed beside the FP64 LDL^T
    mixed = {}
    mb = NedBasisBatch(verts, n_refine_local=NED_NREF, timing=True, solver=L.SOLVER_BAND_MIXED)
    mb.set_coefficients(a, 1.0, 0.0, (1.0, 1.0))
    for _ in range(max(warmup, 1)):
        mb.run()
    mb.synchronize()
    mb.kernel_times()
    t0 = time.perf_counter()
    for _ in range(steps):
        mb.run()
    mb.synchronize()
    wm = time.perf_counter() - t0
    km = mb.kernel_times()
    mixed = {"solver": "FP32 LDL^T on the matrix cores + refinement with FP64 residuals",
             "cells_per_s": round(n_cells / (wm / steps), 1),
             "ms_per_step": round(wm * 1e3 / steps, 4),
             "kernel_ms": {"rows": round(km[0], 4), "factor_refine": round(km[1], 4),
                           "psi_element": round(km[2], 4)},
             "failed_cells": int((mb.get_cell_status() != 0).sum())}
    mixed_out = (mb.get_global_element_matrix(), mb.get_global_element_rhs(), mb.get_basis())
    mb.close()
    nb = NedBasisBatch(verts, n_refine_local=NED_NREF, timing=True)
    nb.set_coefficients(a, 1.0, 0.0, (1.0, 1.0))
    for _ in range(warmup):
        nb.run()
    nb.synchronize()
    if warmup > 0:
        nb.kernel_times()
    t0 = time.perf_counter()
    for _ in range(steps):
        nb.run()
    nb.synchronize()
    wall = time.perf_counter() - t0
    kt = nb.kernel_times()
    n_int, p = 705, 51
    flops_cell = band_flops_per_cell(n_int, p, 8)
    tflops = flops_cell * n_cells / (kt[1] * 1e-3) / 1e12
    out = {"workload": "2D Q1 x Nedelec basis, 32x32=1024 coarse cells x 16x16 fine cells "
                       "(833 DoFs, 705 interior, 4+4 RHS), symmetric indefinite band LDL^T FP64 "
                       "on the matrix cores (4x4 pivot blocks), equilibrated, no pivoting "
                       "(edge-after-node order)",
           "cells_per_s": round(n_cells / (wall / steps), 1),
           "ms_per_step": round(wall * 1e3 / steps, 4),
           "kernel_ms": {"rows": round(kt[0], 4), "factor_solve": round(kt[1], 4),
                         "psi_element": round(kt[2], 4)},
           "failed_cells": int((nb.get_cell_status() != 0).sum()),
           "roofline": {"bound": "mfma", "kernel": "q_blk_factor_kernel<double, 2, 1, true>",
                        "flops_per_cell": flops_cell, "achieved_tflops": round(tflops, 3),
                        "peak_tflops": FP64_PEAK_TFLOPS,
                        "frac": round(tflops / FP64_PEAK_TFLOPS, 4)}}
    cores = host_cores()
    samp = np.random.default_rng(11).choice(n_cells, 16 * cores if cpu else 4, replace=False)
    if check or cpu:
        from oracle import oracle as O
        O.build()
        t0 = time.perf_counter()
        o = O.ned_basis_batch(NED_NREF, verts[samp], a[samp], 1.0, 0.0, (1.0, 1.0),
                              n_threads=cores if cpu else 1)
        t_cpu = time.perf_counter() - t0
        if check:
            ns = (2 ** NED_NREF + 1) ** 2
            for dst, (M, b, P) in ((out, (nb.get_global_element_matrix(),
                                          nb.get_global_element_rhs(), nb.get_basis())),
                                   (mixed, mixed_out)):
                dst["rel_err_vs_fp64_oracle"] = {
                    "M": rel_err(M[samp], o["M"]), "b": rel_err(b[samp, 4:], o["b"][:, 4:]),
                    "Psi_sigma": rel_err(P[samp, :, :ns], o["Psi"][:, :, :ns]),
                    "Psi_u": rel_err(P[samp, :, ns:], o["Psi"][:, :, ns:]),
                    "cells_checked": len(samp)}
        if cpu:
            out["cpu_baseline"] = {"value": round(len(samp) / t_cpu, 3), "unit": "cells/s",
                                   "cores": cores, "kind": "port",
                                   "sample": f"{len(samp)} of the 1024 cells (seeded): dense "
                                             "assembly, equilibrated LU with partial pivoting "
                                             "+ 3 refinement steps (the UMFPACK direct solve "
                                             f"restated), one cell per thread on "
                                             f"{host_cores_text(cores)}, wall {t_cpu:.2f} s"}
    nb.close()
    out["mixed"] = mixed
    out["mixed_vs_fp64"] = round(mixed["cells_per_s"] / out["cells_per_s"], 3)
    return out


RT_NREF, RT_NC, RT_SIGMA = 4, 32, 1.0


def rt_config(steps: int, warmup: int, check: bool, cpu: bool):
    """The 2-D Raviart-Thomas x DQ0 configuration, timed beside the headline:
    32 x 32 coarse cells of the unit square, 16 x 16 fine cells each (544 edge
    fluxes + 256 cell values; the reduced stream-function system has 225
    unknowns and half-bandwidth 16), 4 face bases, a = exp(g) with g iid
    N(0, 1) per fine cell (seed 20261202), f = 1."""
    from msfec_amd import RTDQBasisBatch, fields as F
    verts = F.coarse_vertices(2, RT_NC)
    a = F.lognormal_coefficient(2, RT_NC, RT_NREF, RT_SIGMA, 0.0, SEED)
    n_cells = verts.shape[0]
    rb = RTDQBasisBatch(verts, n_refine_local=RT_NREF, timing=True)
    rb.set_coefficients(a)
    for _ in range(warmup):
        rb.run()
    rb.synchronize()
    if warmup > 0:
        rb.kernel_times()
    t0 = time.perf_counter()
    for _ in range(steps):
        rb.run()
    rb.synchronize()
    wall = time.perf_counter() - t0
    kt = rb.kernel_times()
    m = 1 << RT_NREF
    n_int, p, ne = (m - 1) ** 2, m, 2 * m * (m + 1)
    flops_cell = band_flops_per_cell(n_int, p, 4)
    tflops = flops_cell * n_cells / (kt[1] * 1e-3) / 1e12
    # element kernel: psi in (n x 8 doubles), a and f in, fluxes and M, b out
    elem_bytes = 8.0 * (n_int * 8 + 2 * m * m + 4 * ne + 30)
    elem_gbs = elem_bytes * n_cells / (kt[2] * 1e-3) / 1e9
    out = {"workload": "2D Raviart-Thomas x DQ0 basis, 32x32=1024 coarse cells x 16x16 fine "
                       "cells (544 fluxes + 256 cell values per basis, 4 face bases); the flux "
                       "is the coarse function plus the curl of a Q1 stream function: 225 "
                       "unknowns, half-bandwidth 16, band LDL^T FP64 on the matrix cores",
           "cells_per_s": round(n_cells / (wall / steps), 1),
           "ms_per_step": round(wall * 1e3 / steps, 4),
           "kernel_ms": {"rows": round(kt[0], 4), "factor_solve": round(kt[1], 4),
                         "flux_element": round(kt[2], 4)},
           "failed_cells": int((rb.get_cell_status() != 0).sum()),
           "roofline": {"bound": "mfma", "kernel": "q_blk_factor_kernel<double, 2, 1, true>",
                        "flops_per_cell": flops_cell, "achieved_tflops": round(tflops, 4),
                        "peak_tflops": FP64_PEAK_TFLOPS,
                        "frac": round(tflops / FP64_PEAK_TFLOPS, 5)},
           "element_roofline": {"bound": "hbm", "kernel": "rt_element_kernel",
                                "bytes_per_cell": elem_bytes, "achieved_gbs": round(elem_gbs, 1),
                                "peak_gbs": HBM_PEAK_GBS,
                                "frac": round(elem_gbs / HBM_PEAK_GBS, 4)}}
    if check or cpu:
        from oracle import oracle as O
        O.build()
        cores = host_cores()
        samp = np.random.default_rng(11).choice(n_cells, min(n_cells, 64 * cores) if cpu else 8,
                                                replace=False)
        t0 = time.perf_counter()
        o = O.rt_basis_batch(RT_NREF, verts[samp], a[samp], n_threads=cores if cpu else 1)
        t_cpu = time.perf_counter() - t0
        if check:
            M, b, S = rb.get_global_element_matrix(), rb.get_global_element_rhs(), rb.get_basis()
            out["rel_err_vs_fp64_oracle"] = {
                "M": rel_err(M[samp], o["M"]), "b": rel_err(b[samp, 4:], o["b"][:, 4:]),
                "Sigma": rel_err(S[samp], o["Sigma"]), "cells_checked": len(samp)}
        if cpu:
            out["cpu_baseline"] = {"value": round(len(samp) / t_cpu, 3), "unit": "cells/s",
                                   "cores": cores, "kind": "port",
                                   "sample": f"{len(samp)} of the 1024 cells (seeded): dense "
                                             "assembly of the mixed system, LU with partial "
                                             "pivoting of the saddle-point system (the UMFPACK "
                                             f"direct solve restated), one cell per thread on "
                                             f"{host_cores_text(cores)}, wall {t_cpu:.2f} s"}
    rb.close()
    return out


NRT_NREF, NRT_NC = 2, 16


def nrt_flops_per_cell(n: int, nrhs: int) -> float:
    """Algorithmic FP64 flops of the dense symmetric indefinite L D L^T of an
    n x n system (n^3 / 3) and its forward and backward sweeps for nrhs
    right-hand sides (2 n^2 each)."""
    return n ** 3 / 3.0 + 2.0 * n * n * nrhs


def nrt_config(steps: int, warmup: int, check: bool, cpu: bool):
    """The 3-D Nedelec x Raviart-Thomas configuration: 16^3 coarse cells of the
    unit cube, 4^3 fine cells each (300 edges + 240 faces; 108 + 144 = 252
    interior unknowns), 12 edge and 6 face bases, the coefficients of
    example_parameters/prm_ned_rt_test-01.prm at the Gauss points (rotated
    DiffusionInverse_A with frequencies 17/16/15 and contrast 0.9, Diffusion_B
    with frequency 14, the gradient right-hand side, no reaction)."""
    from msfec_amd import NedRTBasisBatch, fields as F
    verts = F.box_grid_3d(NRT_NC)
    pts = F.gauss_points_3d(verts, NRT_NREF)
    ainv = F.diffusion_inverse_A_tensor(pts)
    bq = F.diffusion_B(pts)
    fq = F.ned_rt_rhs(pts)
    n_cells = verts.shape[0]
    nb = NedRTBasisBatch(verts, n_refine_local=NRT_NREF, timing=True)
    nb.set_coefficients(ainv, bq, None, fq)
    for _ in range(warmup):
        nb.run()
    nb.synchronize()
    if warmup > 0:
        nb.kernel_times()
    t0 = time.perf_counter()
    for _ in range(steps):
        nb.run()
    nb.synchronize()
    wall = time.perf_counter() - t0
    kt = nb.kernel_times()
    m = 1 << NRT_NREF
    n_int = 3 * m * (m - 1) ** 2 + 3 * (m - 1) * m * m
    flops_cell = nrt_flops_per_cell(n_int, 18)
    tflops = flops_cell * n_cells / (kt[0] * 1e-3) / 1e12
    # compulsory HBM bytes per cell: the Gauss-point coefficients in (6 + 1 + 3
    # values at 8 points of 64 fine cells), the 18 fine vectors, M and b out
    hbm_cell = 8.0 * (m ** 3 * 8 * 10 + 18 * nb.N + 18 * 18 + 18)
    out = {"workload": "3D Nedelec x Raviart-Thomas basis, 16^3=4096 coarse cells x 4^3 fine "
                       "cells (300 edges + 240 faces, 252 interior unknowns), 12 edge + 6 face "
                       "bases, coefficients of prm_ned_rt_test-01 at the Gauss points; one "
                       "workgroup per cell: dense symmetric indefinite L D L^T FP64 on the "
                       "matrix cores with the 18 right-hand sides, backward sweep, element matrix",
           "cells_per_s": round(n_cells / (wall / steps), 1),
           "ms_per_step": round(wall * 1e3 / steps, 4),
           "kernel_ms": {"basis": round(kt[0], 4)},
           "failed_cells": int((nb.get_cell_status() != 0).sum()),
           "roofline": {"bound": "mfma", "kernel": "nrt_basis_kernel<2>",
                        "flops_per_cell": flops_cell, "achieved_tflops": round(tflops, 4),
                        "peak_tflops": FP64_PEAK_TFLOPS,
                        "frac": round(tflops / FP64_PEAK_TFLOPS, 5),
                        "hbm_bytes_per_cell": hbm_cell,
                        "hbm_frac": round(hbm_cell * n_cells / (kt[0] * 1e-3) / 1e9
                                          / HBM_PEAK_GBS, 5)}}
    if check or cpu:
        from oracle import oracle as O
        O.build()
        cores = host_cores()
        samp = np.random.default_rng(13).choice(n_cells, min(n_cells, 128 * cores) if cpu else 8,
                                                replace=False)
        t0 = time.perf_counter()
        o = O.nrt_basis_batch(NRT_NREF, verts[samp], ainv[samp], bq[samp], None, fq[samp],
                              n_threads=cores if cpu else 1)
        t_cpu = time.perf_counter() - t0
        if check:
            M, b, X = nb.get_global_element_matrix(), nb.get_global_element_rhs(), nb.get_basis()
            out["rel_err_vs_fp64_oracle"] = {
                "M": rel_err(M[samp], o["M"]), "b": rel_err(b[samp], o["b"]),
                "X": rel_err(X[samp], o["X"]), "cells_checked": len(samp)}
        if cpu:
            out["cpu_baseline"] = {"value": round(len(samp) / t_cpu, 3), "unit": "cells/s",
                                   "cores": cores, "kind": "port",
                                   "sample": f"{len(samp)} of the 4096 cells (seeded): dense "
                                             "assembly of the mixed system over all fine dofs, "
                                             "LU with partial pivoting and iterative refinement "
                                             "of the interior system (the UMFPACK direct solve "
                                             "restated), 18 right-hand sides, one cell per "
                                             f"thread on {host_cores_text(cores)}, wall "
                                             f"{t_cpu:.2f} s"}
    nb.close()
    return out


NRT4_NREF, NRT4_NC = 4, 4


def nrt_minres_bytes_per_iteration(n: int, ns: int, nu: int, nfc: int, nc0: int,
                                   edge_schwarz: bool) -> float:
    """Algorithmic bytes of one MINRES iteration of one basis: 16 passes over
    the n interior values (the product reads v once and reads and writes the
    first residual slot; the second pass reads both residuals and the
    preconditioner and writes one; the third reads v, w, w2, x, the new
    residual and the preconditioner and writes w, x and v); a share of the
    cell's local matrices (99 values per fine cell, read by the 18 bases);
    the coarse correction (the nu face values restricted, the nc0 x nc0
    inverse read); and, with the edge additive Schwarz, 5 passes over the ns
    interior edges (z cleared, y gathered, z read and written, z read for
    the next v) and a share of the 78 inverse entries per fine cell."""
    b = 16 * n + 99 * nfc / 18.0 + nu + nc0 * nc0
    if edge_schwarz:
        b += 5 * ns + 78 * nfc / 18.0
    return 8.0 * b


def nrt4_config(steps: int, warmup: int, check: bool, cpu: bool):
    """The reference's own Ned_RT example: prm_ned_rt_test-01 (global
    refinements 2: 4^3 = 64 coarse cells of the unit cube; local refinements
    4: 16^3 fine cells per coarse cell, 13 872 edges + 13 056 faces, 22 320
    interior unknowns), its coefficients at the Gauss points, 12 edge + 6
    face bases, solved by matrix-free MINRES."""
    from msfec_amd import NedRTBasisBatch, fields as F
    verts = F.box_grid_3d(NRT4_NC)
    pts = F.gauss_points_3d(verts, NRT4_NREF)
    ainv = F.diffusion_inverse_A_tensor(pts)
    bq = F.diffusion_B(pts)
    fq = F.ned_rt_rhs(pts)
    n_cells = verts.shape[0]
    nb = NedRTBasisBatch(verts, n_refine_local=NRT4_NREF, timing=True)
    nb.set_coefficients(ainv, bq, None, fq)
    for _ in range(warmup):
        nb.run()
    nb.synchronize()
    if warmup > 0:
        nb.kernel_times()
    t0 = time.perf_counter()
    for _ in range(steps):
        nb.run()
    nb.synchronize()
    wall = time.perf_counter() - t0
    kt = nb.kernel_times()
    its = nb.get_iterations()
    m = 1 << NRT4_NREF
    ns, nu = 3 * m * (m - 1) ** 2, 3 * (m - 1) * m * m
    nblk = m // 2 if m <= 4 else 4  # aggregates per direction of the coarse correction
    bpi = nrt_minres_bytes_per_iteration(ns + nu, ns, nu, m ** 3, 3 * nblk ** 3, NRT4_NREF >= 4)
    gbs = float(its.sum()) * bpi / (kt[2] * 1e-3) / 1e9
    out = {"workload": "3D Nedelec x Raviart-Thomas basis, the reference's prm_ned_rt_test-01: "
                       "4^3=64 coarse cells x 16^3 fine cells (13872 edges + 13056 faces, "
                       "22320 interior unknowns), 12 edge + 6 face bases, its coefficients at "
                       "the Gauss points; matrix-free MINRES (relative tolerance 1e-14) "
                       "preconditioned by an element additive Schwarz on the edges and a "
                       "diagonal plus 192-dof coarse correction on the faces, one workgroup "
                       "per basis",
           "cells_per_s": round(n_cells / (wall / steps), 2),
           "ms_per_step": round(wall * 1e3 / steps, 3),
           "kernel_ms": {"total": round(kt[0], 3), "local_and_preconditioner": round(kt[1], 3),
                         "minres": round(kt[2], 3), "element": round(kt[3], 3)},
           "minres_iterations": {"mean": round(float(its.mean()), 1), "min": int(its.min()),
                                 "max": int(its.max())},
           "failed_cells": int((nb.get_cell_status() != 0).sum()),
           "roofline": {"bound": "hbm", "kernel": "nrt_minres_kernel<4>",
                        "bytes_per_iteration_basis": bpi, "achieved_gbs": round(gbs, 1),
                        "peak_gbs": HBM_PEAK_GBS, "frac": round(gbs / HBM_PEAK_GBS, 4)}}
    if check or cpu:
        from oracle import oracle as O
        O.build()
        c = 21
        t0 = time.perf_counter()
        Mo, bo, Xo = O.nrt_basis_cell_sparse(NRT4_NREF, verts[c], ainv[c], bq[c], None, fq[c])
        t_cpu = time.perf_counter() - t0
        if check:
            M, b, X = nb.get_global_element_matrix(), nb.get_global_element_rhs(), nb.get_basis()
            out["rel_err_vs_fp64_oracle"] = {
                "M": rel_err(M[c:c + 1], Mo[None]), "b": rel_err(b[c:c + 1], bo[None]),
                "X": rel_err(X[c:c + 1], Xo[None]), "cells_checked": 1,
                "oracle": "sparse direct solve (SuperLU)"}
        if cpu:
            out["cpu_baseline"] = {"value": round(1.0 / t_cpu, 4), "unit": "cells/s", "cores": 1,
                                   "kind": "port",
                                   "sample": "1 of the 64 cells (cell 21): triplet assembly of "
                                             "the mixed system, sparse LU (SuperLU, scipy) of "
                                             "the interior system with the 18 right-hand sides "
                                             "(the UMFPACK direct solve restated), element "
                                             f"matrix; one thread, wall {t_cpu:.2f} s"}
    nb.close()
    return out


RT3_NREF, RT3_NC = 4, 4


def rt3_cg_bytes_per_iteration(nfc: int, ns: int) -> float:
    """Algorithmic HBM bytes of one Jacobi-PCG iteration of one 3-D RT_DQ
    basis (vectors over the ns interior edges): the 21 face-mass entries of
    every fine cell, p read and q read and written in the element-by-element
    product, then per edge x and r read and written, p, q and the Jacobi
    diagonal read in the update (56 B), and the diagonal, r, p read, p and q
    written in the direction update (40 B)."""
    return 8.0 * 21 * nfc + 24.0 * ns + 96.0 * ns


def rt3_config(steps: int, warmup: int, check: bool, cpu: bool):
    """The reference's own 3-D RT_DQ example, prm_rt_dq_test-01 (global
    refinements 2: 4^3 = 64 coarse cells of the unit cube; local refinements
    4: 16^3 fine cells, 13 056 faces + 4096 cell values per basis), its
    DiffusionInverse_A and right-hand side at the Gauss points, 6 face bases."""
    from msfec_amd import RTDQ3BasisBatch, fields as F
    verts = F.box_grid_3d(RT3_NC)
    pts = F.gauss_points_3d(verts, RT3_NREF)
    ainv = F.diffusion_inverse_A_tensor(pts)
    fq = F.rt_dq_rhs_3d(pts)
    n_cells = verts.shape[0]
    rb = RTDQ3BasisBatch(verts, n_refine_local=RT3_NREF, timing=True)
    rb.set_coefficients(ainv, fq)
    for _ in range(warmup):
        rb.run()
    rb.synchronize()
    if warmup > 0:
        rb.kernel_times()
    t0 = time.perf_counter()
    for _ in range(steps):
        rb.run()
    rb.synchronize()
    wall = time.perf_counter() - t0
    kt = rb.kernel_times()
    its = rb.get_iterations()
    m = 1 << RT3_NREF
    ns = 3 * m * (m - 1) ** 2
    bpi = rt3_cg_bytes_per_iteration(m ** 3, ns)
    gbs = float(its.sum()) * bpi / (kt[2] * 1e-3) / 1e9
    out = {"workload": "3D Raviart-Thomas x DQ0 basis, the reference's prm_rt_dq_test-01: "
                       "4^3=64 coarse cells x 16^3 fine cells (13056 faces + 4096 cell values), "
                       "6 face bases, its coefficients and right-hand side at the Gauss points; "
                       "the flux is the coarse function plus the curl of an interior Nedelec "
                       f"field ({ns} interior edges), curl-curl system by Jacobi-PCG "
                       "(relative tolerance 1e-13), one workgroup per basis",
           "cells_per_s": round(n_cells / (wall / steps), 2),
           "ms_per_step": round(wall * 1e3 / steps, 3),
           "kernel_ms": {"total": round(kt[0], 3), "local_and_diagonal": round(kt[1], 3),
                         "cg": round(kt[2], 3), "element": round(kt[3], 3)},
           "cg_iterations": {"mean": round(float(its.mean()), 1), "min": int(its.min()),
                             "max": int(its.max())},
           "failed_cells": int((rb.get_cell_status() != 0).sum()),
           "roofline": {"bound": "hbm", "kernel": "rt3_cg_kernel<4>",
                        "bytes_per_iteration_basis": bpi, "achieved_gbs": round(gbs, 1),
                        "peak_gbs": HBM_PEAK_GBS, "frac": round(gbs / HBM_PEAK_GBS, 4)}}
    if check or cpu:
        from oracle import oracle as O
        O.build()
        c = 21
        t0 = time.perf_counter()
        Mo, bo, So = O.rt3_basis_cell_sparse(RT3_NREF, verts[c], ainv[c], fq[c])
        t_cpu = time.perf_counter() - t0
        if check:
            M, b, S = rb.get_global_element_matrix(), rb.get_global_element_rhs(), rb.get_basis()
            out["rel_err_vs_fp64_oracle"] = {
                "M": rel_err(M[c:c + 1], Mo[None]), "b": rel_err(b[c:c + 1], bo[None]),
                "Sigma": rel_err(S[c:c + 1], So[None]), "cells_checked": 1,
                "oracle": "sparse direct solve of the saddle-point system (SuperLU)"}
        if cpu:
            out["cpu_baseline"] = {"value": round(1.0 / t_cpu, 4), "unit": "cells/s", "cores": 1,
                                   "kind": "port",
                                   "sample": "1 of the 64 cells (cell 21): triplet assembly of "
                                             "the mixed system, sparse LU (SuperLU, scipy) of "
                                             "the saddle-point system with the 6 right-hand "
                                             "sides (the UMFPACK direct solve restated), element "
                                             f"matrix; one thread, wall {t_cpu:.2f} s"}
    rb.close()
    return out


def cg_flops_per_iteration(dim: int, level: int) -> float:
    """Algorithmic FP64 flops of one Jacobi-PCG iteration of one basis on one
    cell: the matrix-free 3^dim-point product over the interior rows (a
    multiply-add per coupling) and the 3 dot products, 3 updates and the
    preconditioner (2 flops each per row)."""
    n = ((1 << level) - 1) ** dim
    return n * (2.0 * 3 ** dim + 2.0 * 7)


def uneven_config(steps: int, warmup: int, check: bool, cpu: bool):
    """The uneven 3-D configuration, timed beside the headline: 16^3 coarse
    cells; 410 seeded inclusions (A = 1e4, radius 0.02) in a background A = 1;
    the cells holding an inclusion centre refined 4 levels (3375 interior
    unknowns, matrix-free CG), the others 2 or 3 levels by a seeded coin
    (band solver or CG, whichever the library's timing on up to 256 cells of
    the level picks); f = 1.  One handle (msfec_qu_*), levels overlapped on
    streams of their own."""
    from msfec_amd import UnevenBatch, fields as F
    nc = 16
    cen = F.inclusion_centres(3, 410, SEED)
    lv = F.uneven_levels(3, nc, cen, SEED + 1)
    blocks = F.inclusion_blocks(3, nc, lv, cen, 0.02, 1e4)
    verts = F.coarse_vertices(3, nc)
    ub = UnevenBatch(verts, lv, dim=3, coef_blocks=blocks, timing=True)
    t0 = time.perf_counter()
    ub.tune()
    t_tune = time.perf_counter() - t0
    for _ in range(warmup):
        ub.run()
    ub.synchronize()
    for l in ub.level_set:
        if warmup > 0:
            ub.level(l).kernel_times()
    t0 = time.perf_counter()
    for _ in range(steps):
        ub.run()
    ub.synchronize()
    wall = time.perf_counter() - t0
    summ = ub.summary()
    per_level = {}
    for l in ub.level_set:
        kt = ub.level(l).kernel_times()
        per_level[str(l)] = {"cells": summ[l]["cells"], "solver": summ[l]["solver"],
                             "band_window": summ[l]["window"] or None,
                             "tuned_ms": summ[l]["tuned_ms"],
                             "kernel_ms": [round(t, 4) for t in kt[:4]]}
    out = {"workload": "uneven 3D Q1 batch: 16^3 cells, levels 2/3 (band or CG, timed choice) "
                       "and 4 (3375 interior DoFs, matrix-free Jacobi-PCG FP64) for the cells "
                       "holding one of 410 inclusions (A=1e4, r=0.02); one msfec_qu handle, "
                       "levels forked onto their own streams",
           "cells_per_s": round(ub.n_cells / (wall / steps), 1),
           "ms_per_step": round(wall * 1e3 / steps, 4),
           "levels": per_level,
           "tune_s": round(t_tune, 3),
           "failed_cells": int((ub.get_cell_status() != 0).sum())}
    it = ub.level(4).cg_iterations()
    out["cg_iterations_l4"] = {"mean": round(float(it.mean()), 1), "max": int(it.max())}
    cg_ms = per_level["4"]["kernel_ms"][1]
    if summ[4]["solver"] == "cg" and cg_ms > 0:
        fl = float(it[:, :7].sum()) * cg_flops_per_iteration(3, 4)
        tf = fl / (cg_ms * 1e-3) / 1e12
        out["cg_roofline"] = {"bound": "fp64 vector ALU (operands LDS-resident)",
                              "kernel": "q_cg_kernel<3, false>", "kernel_ms": cg_ms,
                              "flops_per_iteration_basis": cg_flops_per_iteration(3, 4),
                              "iterations_total": int(it[:, :7].sum()),
                              "achieved_tflops": round(tf, 3), "peak_tflops": FP64_PEAK_TFLOPS,
                              "frac": round(tf / FP64_PEAK_TFLOPS, 4)}
    if check or cpu:
        from oracle import oracle as O
        O.build()
        rng = np.random.default_rng(11)
        cells = {l: np.nonzero(lv == l)[0] for l in (2, 3, 4)}
        samp = {2: rng.choice(cells[2], 8, replace=False), 3: rng.choice(cells[3], 8, replace=False),
                4: rng.choice(cells[4], 1, replace=False)}
        if check:
            M, b = ub.get_global_element_matrix(), ub.get_global_element_rhs()
            P = ub.get_basis_flat()
            errs = {}
            for l, idx in samp.items():
                o = O.basis_batch(O.make_problem(3, l), verts[idx],
                                  coef=np.stack([blocks[c] for c in idx]), solver=O.SOLVER_DIRECT)
                nn = ((1 << l) + 1) ** 3
                Pl = np.stack([P[int(ub.basis_offsets[c]):int(ub.basis_offsets[c]) + 8 * nn]
                               for c in idx]).reshape(len(idx), 8, nn)
                errs[f"l{l}"] = {"M": rel_err(M[idx], o["M"]), "b": rel_err(b[idx], o["b"]),
                                 "Psi": rel_err(Pl, o["Psi"]), "cells": len(idx)}
            out["rel_err_vs_fp64_oracle"] = errs
        if cpu:
            # the reference's default CG + SSOR on a sample of each level,
            # scaled to the batch's level mix
            cores = host_cores()
            t_cell = {}
            for l in (2, 3, 4):
                idx = rng.choice(cells[l], cores, replace=False)
                t0 = time.perf_counter()
                O.basis_batch(O.make_problem(3, l), verts[idx],
                              coef=np.stack([blocks[c] for c in idx]), solver=O.SOLVER_CG_SSOR,
                              n_threads=cores, want_psi=False)
                t_cell[l] = (time.perf_counter() - t0) / len(idx)
            total = sum(t_cell[l] * len(cells[l]) for l in (2, 3, 4))
            out["cpu_baseline"] = {"value": round(ub.n_cells / total, 1), "unit": "cells/s",
                                   "cores": cores, "kind": "port",
                                   "sample": f"{cores} cells per level, CG+SSOR(1.6) abs tol 1e-12, "
                                             f"OpenMP on {host_cores_text(cores)}, per-level "
                                             "times scaled to the batch's level counts"}
    ub.close()
    return out


def plain_config(steps: int, warmup: int, check: bool, cpu: bool):
    """The plain-method configuration: 2-D -div(A grad u) + u = f on the unit
    square, 8 x 8 coarse cells refined 4 levels (16 x 16 fine cells, 289
    nodes, 4 bases each), A = exp(g) with g iid N(0, 1) per fine cell (seed
    20261202), reaction 1, f = 1.  Timed: the basis construction of the 64
    cells, and the whole method (bases, coarse assembly, coarse CG to 1e-12,
    weights).  Checked: M, b and Psi of all 64 cells and the coarse solution
    against the oracle."""
    from msfec_amd import QBasisBatch, _lib as L, fields as F
    dim, nc, l = 2, 8, 4
    verts = F.coarse_vertices(dim, nc)
    coef = F.lognormal_coefficient(dim, nc, l, 1.0, 0.0, SEED)
    # kernel times from a handle with HIP events between the kernels, the
    # rates from one without them
    qt = QBasisBatch(verts, dim=dim, n_refine_local=l, solver=L.SOLVER_BAND_FP64_MFMA,
                     reaction_mode=L.REACTION_CONST, reaction_const=1.0, timing=True)
    qt.set_coefficient(coef)
    for _ in range(warmup):
        qt.run()
    qt.synchronize()
    if warmup > 0:
        qt.kernel_times()
    for _ in range(steps):
        qt.run()
    qt.synchronize()
    kt = qt.kernel_times()
    qt.close()
    qb = QBasisBatch(verts, dim=dim, n_refine_local=l, solver=L.SOLVER_BAND_FP64_MFMA,
                     reaction_mode=L.REACTION_CONST, reaction_const=1.0)
    qb.set_coefficient(coef)
    for _ in range(warmup):
        qb.run()
    qb.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        qb.run()
    qb.synchronize()
    wall = time.perf_counter() - t0
    t0 = time.perf_counter()
    for _ in range(steps):
        qb.run()
        u, w, it, res = qb.coarse_solve(nc, abs_tol=1e-12, rel_tol=1e-12)
    wall_all = time.perf_counter() - t0
    out = {"workload": "2D Q1, -div(A grad u) + u = f, 8x8 coarse cells x 16x16 fine cells "
                       "(225 interior DoFs, 4 RHS), A=exp(g) g iid N(0,1), reaction 1, f=1; "
                       "banded Cholesky FP64 on the matrix cores + coarse Jacobi-PCG",
           "cells_per_s": round(64 / (wall / steps), 1),
           "ms_per_step": round(wall * 1e3 / steps, 4),
           "kernel_ms": {"rows": round(kt[0], 4), "factor_solve": round(kt[1], 4)},
           "method_ms_per_step": round(wall_all * 1e3 / steps, 4),
           "coarse_cg_iterations": it,
           "failed_cells": int((qb.get_cell_status() != 0).sum())}
    if check:
        from oracle import oracle as O
        O.build()
        o = O.basis_batch(O.make_problem(dim, l, reaction_mode=O.REACTION_CONST, reaction_const=1.0),
                          verts, coef=coef, solver=O.SOLVER_DIRECT)
        M, b, P = qb.get_global_element_matrix(), qb.get_global_element_rhs(), qb.get_basis()
        uo, _ = O.coarse_solve(dim, nc, o["M"], o["b"], rtol=1e-14)
        out["rel_err_vs_fp64_oracle"] = {"M": rel_err(M, o["M"]), "b": rel_err(b, o["b"]),
                                         "Psi": rel_err(P, o["Psi"]),
                                         "u_coarse": float(np.abs(u - uo).max() / np.abs(uo).max()),
                                         "cells_checked": 64}
    if cpu:
        from oracle import oracle as O
        O.build()
        cores = host_cores()
        pp = O.make_problem(dim, l, reaction_mode=O.REACTION_CONST, reaction_const=1.0)
        walls, r = timed_reps(lambda: O.basis_batch(pp, verts, coef=coef, solver=O.SOLVER_CG_SSOR,
                                                    n_threads=cores, want_psi=False), 5)
        out["cpu_baseline"] = {
            "value": round(64 / min(walls), 1), "unit": "cells/s", "cores": cores, "kind": "port",
            "sample": f"all 64 cells, CG+SSOR(1.6) abs tol 1e-12 per basis, OpenMP on "
                      f"{host_cores_text(cores)}, best of 5 runs after a warm-up (median "
                      f"{64 / float(np.median(walls)):.0f} cells/s)"}
    qb.close()
    return out


def adjoint_config(steps: int, check: bool):
    """The adjoint configuration: the headline batch (16^3 x 8^3, 2.1 M fine
    log-coefficients g, a = exp(g)), J = 1/2 |u_H - u_target|^2 over the free
    coarse nodes with u_target the coarse solution of a second seeded field
    (sigma 1.0, iid).  One step = the forward basis construction with FP32
    factors + FP64 refinement and the backward pass (coarse solve, adjoint
    coarse solve, w = K_II^-1 F_I from the stored panels, the per-fine-cell
    gradient) with the gradient left in device memory, as an optimiser on
    the GPU would take it; also timed with the gradient downloaded to host
    memory each step.  Checked against a central difference of the forward
    map along a seeded direction."""
    from msfec_amd import DeviceBuffer, QBasisBatch, _lib as L, fields as F
    nc = NC
    verts = F.coarse_vertices(DIM, nc)
    g = np.log(F.lognormal_coefficient(DIM, nc, NREF, SIGMA, CORR, SEED))
    g2 = np.log(F.lognormal_coefficient(DIM, nc, NREF, 1.0, 0.0, SEED + 1))

    def coarse(gv, solver):
        q = QBasisBatch(verts, dim=DIM, n_refine_local=NREF, solver=solver)
        q.set_coefficient(np.exp(gv))
        q.run()
        u, _, _, _ = q.coarse_solve(nc, abs_tol=0.0, rel_tol=1e-14)
        q.close()
        return u

    ut = coarse(g2, L.SOLVER_BAND_FP64_MFMA)
    qb = QBasisBatch(verts, dim=DIM, n_refine_local=NREF, solver=L.SOLVER_BAND_MIXED)
    qb.set_coefficient(np.exp(g))
    qb.run()
    J, grad = qb.adjoint_gradient(nc, ut)
    dgrad = DeviceBuffer(grad.shape)
    qb.adjoint_gradient(nc, ut, out=dgrad)
    t0 = time.perf_counter()
    for _ in range(steps):
        qb.run()
        J, _ = qb.adjoint_gradient(nc, ut, out=dgrad)
    wall = time.perf_counter() - t0
    t0 = time.perf_counter()
    for _ in range(steps):
        qb.run()
        J, grad = qb.adjoint_gradient(nc, ut)
    wall_host = time.perf_counter() - t0
    same = bool(np.array_equal(dgrad.numpy(), grad))
    dgrad.free()
    out = {"workload": "adjoint of J = 1/2 |u_H - u_target|^2 through the basis solves and the "
                       "coarse CG, 16^3 coarse x 8^3 fine cells (2,097,152 log-coefficients), "
                       "FP32 factors + FP64 refinement; gradient in device memory",
           "ms_per_step": round(wall * 1e3 / steps, 4),
           "gradients_per_s": round(steps / wall, 2),
           "ms_per_step_gradient_to_host": round(wall_host * 1e3 / steps, 4),
           "device_gradient_equals_host": same,
           "J": J}
    if check:
        free = np.ones((nc + 1,) * DIM, dtype=bool)
        free[0] = free[-1] = False
        free[:, 0] = free[:, -1] = False
        free[:, :, 0] = free[:, :, -1] = False
        free = free.ravel()

        def misfit(gv):
            u = coarse(gv, L.SOLVER_BAND_MIXED)
            return 0.5 * float(((u - ut)[free] ** 2).sum())

        dg = np.random.default_rng(4).standard_normal(g.shape)
        eps = 1e-4
        fd = (misfit(g + eps * dg) - misfit(g - eps * dg)) / (2 * eps)
        ad = float((grad * dg).sum())
        out["fd_check"] = {"directional_fd": fd, "directional_adjoint": ad,
                           "abs_diff_over_grad_norm_dg_norm":
                               abs(fd - ad) / (np.linalg.norm(grad) * np.linalg.norm(dg)),
                           "rel_diff": abs(fd - ad) / abs(fd), "eps": eps}
    qb.close()
    return out


def copy_probe(mib: int, reps: int = 3):
    """Device-to-device copies of a known size on the library's HIP runtime.
    Profiled with the FETCH_SIZE / WRITE_SIZE counters, the copy kernel's
    reading against its known bytes gives the scale factor applied to the
    same counters of the basis kernels."""
    import ctypes
    from msfec_amd import _lib as L
    L.load()  # the runtime the library links (libamdhip64.so.7)
    hip = ctypes.CDLL("libamdhip64.so.7")
    nbytes = mib << 20
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    if hip.hipMemset(src, 1, ctypes.c_size_t(nbytes)) != 0:
        raise RuntimeError("hipMemset failed")
    for _ in range(reps):
        if hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copy_probe_bytes": nbytes, "copy_probe_kib": nbytes // 1024, "reps": reps}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-ned", action="store_true",
                    help="skip the 2-D Q1 x Nedelec configuration timed beside the headline")
    ap.add_argument("--no-rt", action="store_true",
                    help="skip the 2-D and 3-D Raviart-Thomas x DQ0 configurations timed beside the headline")
    ap.add_argument("--no-nrt", action="store_true",
                    help="skip the 3-D Nedelec x Raviart-Thomas configuration timed beside the "
                         "headline")
    ap.add_argument("--no-uneven", action="store_true",
                    help="skip the uneven 3-D configuration timed beside the headline")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the 2-D 8x8 plain-method configuration timed beside the headline")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the adjoint configuration timed beside the headline")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run MIB-sized device copies (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only: the per-cell basis "
                                   "construction would shard over GPUs by coarse cells, but "
                                   "no multi-GPU path is built"}))
        sys.exit(2)

    # The library owns its device buffers and stream work; torch is not
    # imported so that exactly one HIP runtime (the one the library was built
    # against) lives in the process.
    from msfec_amd import QBasisBatch, _lib as L, fields as F

    verts = F.coarse_vertices(DIM, NC)
    coef = F.lognormal_coefficient(DIM, NC, NREF, SIGMA, CORR, SEED)

    def measure(solver, steps, warmup, timing=True):
        qb = QBasisBatch(verts, dim=DIM, n_refine_local=NREF, solver=solver, timing=timing)
        qb.set_coefficient(coef)          # host -> HBM copy, before the timed region
        for _ in range(warmup):
            qb.run()
        qb.synchronize()
        if warmup > 0 and timing:
            qb.kernel_times()  # drop the warm-up events
        qb.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            qb.run()
        qb.synchronize()
        wall = time.perf_counter() - t0
        # HIP events on the launch stream, averaged over the timed steps
        kt = qb.kernel_times() if timing else None
        return qb, wall, kt

    # the oracle's dense FP64 direct solve on a seeded sample of cells: every
    # solver's M, b and Psi are compared with it
    oracle_ref = None
    samp = np.random.default_rng(11).choice(verts.shape[0], 32, replace=False)
    if not args.no_check:
        from oracle import oracle as O
        O.build()
        oracle_ref = O.basis_batch(O.make_problem(DIM, NREF), verts[samp], coef=coef[samp],
                                   solver=O.SOLVER_DIRECT)

    def errors(q):
        if oracle_ref is None:
            return None
        M, b, Psi = q.get_global_element_matrix(), q.get_global_element_rhs(), q.get_basis()
        return {"M": rel_err(M[samp], oracle_ref["M"]), "b": rel_err(b[samp], oracle_ref["b"]),
                "Psi": rel_err(Psi[samp], oracle_ref["Psi"]), "cells_checked": len(samp)}

    # the other local solvers, timed beside the headline on fewer steps
    others = {}
    for name, (solver, kname, _) in SOLVERS.items():
        if name == HEADLINE:
            continue
        k = max(2, args.steps // 4)
        q, w, t = measure(solver, k, min(args.warmup, 1))
        others[name] = {"cells_per_s": round(q.info.n_cells / (w / k), 1),
                        "ms_per_step": round(w * 1e3 / k, 4),
                        "kernel": kname, "kernel_ms": [round(x, 4) for x in t[:5]],
                        "failed_cells": int((q.get_cell_status() != 0).sum())}
        e = errors(q)
        if e is not None:
            others[name]["rel_err_vs_fp64_oracle"] = e
        q.close()
    solver, kname, arith = SOLVERS[HEADLINE]
    # kernel times from a run with HIP events between the kernels; the value
    # from a run without them (the events are markers in the stream, not work)
    qt, _, kt = measure(solver, args.steps, args.warmup)
    qt.close()
    qb, wall, _ = measure(solver, args.steps, args.warmup, timing=False)
    info = qb.info
    ms_step = wall * 1e3 / args.steps
    status = qb.get_cell_status()
    cells_per_s = info.n_cells / (ms_step * 1e-3)

    n, p = info.n_interior, info.bandwidth
    nlow = 14
    flops_cell = band_flops_per_cell(n, p, 8)
    if arith == "f32":
        bytes_cell = mixed_factor_bytes_per_cell(n, p, nlow)
        peak_tflops = FP32_MFMA_PEAK_TFLOPS
    else:
        bytes_cell = band_bytes_per_cell(n, p, nlow)
        peak_tflops = FP64_PEAK_TFLOPS
    band_ms = kt[1]  # the dominant kernel: factor + forward + backward sweep
    tflops = flops_cell * info.n_cells / (band_ms * 1e-3) / 1e12
    f_frac = tflops / peak_tflops
    # the kernel's own round trips (band factor written and read back, ...)
    gbs = bytes_cell * info.n_cells / (band_ms * 1e-3) / 1e9
    # the step's compulsory HBM bytes: coefficients and vertices in, Psi, M, b out
    comp_cell = compulsory_bytes_per_cell(DIM, NREF)
    comp_gbs = comp_cell * info.n_cells / (ms_step * 1e-3) / 1e9
    traffic = measured_traffic(kname, info.n_cells)
    rows_traffic = measured_traffic("q_rows_kernel", info.n_cells)
    roof = {"bound": "mfma", "achieved": round(tflops, 3), "peak": peak_tflops,
            "unit": "TFLOP/s", "frac": round(f_frac, 4), "traffic": traffic,
            "kernel": kname, "kernel_ms": round(band_ms, 4),
            "flops_per_cell": flops_cell,
            "flops_note": "band Cholesky inside the half-bandwidth + forward and backward "
                          "sweeps of the 8 lifting right-hand sides (the work this kernel does; "
                          "the dense n^3/3 + 2rn^2 figure would exceed the peak)",
            "kernel_bytes_per_cell": bytes_cell, "kernel_gbs": round(gbs, 1),
            "kernel_hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
            "compulsory_bytes_per_cell": comp_cell,
            "step_compulsory_gbs": round(comp_gbs, 1),
            "step_compulsory_hbm_frac": round(comp_gbs / HBM_PEAK_GBS, 4),
            "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                            "copy-calibrated, profiles/hbm_traffic.json)",
            "dense_equiv_tflops": round(dense_flops_per_cell(n, 8) * info.n_cells /
                                        (band_ms * 1e-3) / 1e12, 3)}
    if traffic is not None and rows_traffic is not None:
        step_traffic = traffic + rows_traffic
        roof["step_traffic_per_cell"] = round(step_traffic / info.n_cells)
        roof["step_traffic_vs_compulsory"] = round(step_traffic / (comp_cell * info.n_cells), 2)
    # mixed precision against the fastest FP64 path (the headline or a path beside it)
    rates = {k: v["cells_per_s"] for k, v in others.items()}
    rates[HEADLINE] = cells_per_s
    fp64_best = max((v for k, v in rates.items() if "fp64" in k), default=None)
    mixed_vs_fp64 = (round(rates["band-mixed"] / fp64_best, 3)
                     if fp64_best and "band-mixed" in rates else None)

    out = {
        "metric": "basis cells/sec, 3D Q 8^3 fine/cell: mixed vs FP64 GPU at ≤1e-10 rel err",
        "value": round(cells_per_s, 1),
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if arith == "f32" else "f64",
        "data": "synthetic: seeded lognormal A=exp(g), g Gaussian sigma 1.5, corr. length 1/64 "
                "on the 128^3 fine grid (seed 20261202), f=1",
        "config": {"workload": "3D Q1 basis construction, 16^3=4096 coarse cells x 8^3 fine "
                               "cells (343 interior DoFs, 8 RHS), " + SOLVER_TEXT[HEADLINE],
                   "n_cells": info.n_cells, "n_refine_local": NREF,
                   "solver": HEADLINE, "parallelism": "single GPU"},
        "other_solvers": others,
        "mixed_vs_fp64": mixed_vs_fp64,
        # the FP64 path launches two kernels (the element matrix is the factor
        # kernel's epilogue); the mixed path's refinement kernels are listed
        # under other_solvers
        "kernel_ms": ({"rows": round(kt[0], 4), "factor_solve": round(kt[1], 4)}
                      if arith == "f64" else
                      {"rows": round(kt[0], 4), "factor_solve": round(kt[1], 4),
                       "refine_residual": round(kt[2], 4), "refine_sweep": round(kt[3], 4)}),
        "wall_s_timed": round(wall, 3),
        "failed_cells": int((status != 0).sum()),
        "roofline": roof,
    }
    e = errors(qb)
    if e is not None:
        out["rel_err_vs_fp64_oracle"] = e
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(verts, coef)
    qb.close()
    others_cfg = {}
    if not args.no_ned:
        others_cfg["q_ned_2d_32x32_l4"] = ned_config(
            max(2, args.steps // 2), min(args.warmup, 2), not args.no_check,
            not args.no_cpu_baseline)
    if not args.no_rt:
        others_cfg["rt_dq_2d_32x32_l4"] = rt_config(
            max(2, args.steps // 2), min(args.warmup, 2), not args.no_check,
            not args.no_cpu_baseline)
    if not args.no_nrt:
        others_cfg["ned_rt_3d_16x16x16_l2"] = nrt_config(
            max(2, args.steps // 2), min(args.warmup, 2), not args.no_check,
            not args.no_cpu_baseline)
    if not args.no_nrt:
        others_cfg["ned_rt_3d_4x4x4_l4"] = nrt4_config(
            max(2, args.steps // 4), min(args.warmup, 1), not args.no_check,
            not args.no_cpu_baseline)
    if not args.no_rt:
        others_cfg["rt_dq_3d_4x4x4_l4"] = rt3_config(
            max(2, args.steps // 4), min(args.warmup, 1), not args.no_check,
            not args.no_cpu_baseline)
    if not args.no_uneven:
        others_cfg["q_uneven_3d_16x16x16"] = uneven_config(
            max(2, args.steps // 4), min(args.warmup, 1), not args.no_check,
            not args.no_cpu_baseline)
    if not args.no_plain:
        others_cfg["q_plain_2d_8x8_l4"] = plain_config(max(2, args.steps // 2),
                                                       min(args.warmup, 2), not args.no_check,
                                                       not args.no_cpu_baseline)
    if not args.no_adjoint:
        others_cfg["q_adjoint_3d_16x16x16"] = adjoint_config(max(2, args.steps // 10),
                                                             not args.no_check)
    if others_cfg:
        out["other_configs"] = others_cfg
    print(json.dumps(out))


if __name__ == "__main__":
    main()
