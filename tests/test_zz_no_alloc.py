"""msfec_q_compute_noalloc (the entry point that allocates nothing, scratch
in a static device array) against msfec_q_compute with the FP64 matrix-core
solver, whose kernels it shares: bit-identical M, b, Psi and status, on two
calls in a row through the C-ABI, the second with a smaller input (the
scratch keeps the first call's contents)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class Hip:
    def __init__(self):
        from msfec_amd import _lib
        _lib.load()
        self.L = ctypes.CDLL("libamdhip64.so.7")  # the runtime the library already uses
        vp = ctypes.c_void_p
        for name, args in (("hipMalloc", [ctypes.POINTER(vp), ctypes.c_size_t]),
                           ("hipFree", [vp]),
                           ("hipMemcpy", [vp, vp, ctypes.c_size_t, ctypes.c_int]),
                           ("hipDeviceSynchronize", [])):
            getattr(self.L, name).argtypes = args
            getattr(self.L, name).restype = ctypes.c_int

    def ok(self, st, what):
        assert st == 0, f"{what} returned {st}"


class Dev:
    def __init__(self, hip, nbytes):
        self.hip, self.nbytes = hip, nbytes
        self.ptr = ctypes.c_void_p()
        hip.ok(hip.L.hipMalloc(ctypes.byref(self.ptr), nbytes), "hipMalloc")

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes == self.nbytes
        self.hip.ok(self.hip.L.hipMemcpy(self.ptr, ctypes.c_void_p(arr.ctypes.data), arr.nbytes, 1),
                    "hipMemcpy H2D")

    def get(self, shape, dtype):
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes == self.nbytes
        self.hip.ok(self.hip.L.hipMemcpy(ctypes.c_void_p(out.ctypes.data), self.ptr, out.nbytes, 2),
                    "hipMemcpy D2H")
        return out

    def free(self):
        self.hip.L.hipFree(self.ptr)


def reference(verts, coef, dim, l):
    from msfec_amd import QBasisBatch, _lib as L
    with QBasisBatch(verts, dim=dim, n_refine_local=l, solver=L.SOLVER_BAND_FP64_MFMA) as qb:
        qb.set_coefficient(coef)
        qb.run()
        return (qb.get_global_element_matrix(), qb.get_global_element_rhs(), qb.get_basis(),
                qb.get_cell_status())


def noalloc(hip, verts, coef, dim, l):
    from msfec_amd import _lib as L
    n = verts.shape[0]
    nb, nn = 1 << dim, ((1 << l) + 1) ** dim
    p = L.Params()
    p.dim, p.n_refine_local, p.coef_mode, p.rhs_mode, p.rhs_const = dim, l, L.COEF_SCALAR_CELL, 0, 1.0
    p.reaction_mode, p.reaction_const, p.solver = 0, 0.0, L.SOLVER_BAND_FP64_MFMA
    bufs = {"v": Dev(hip, verts.nbytes), "c": Dev(hip, coef.nbytes),
            "P": Dev(hip, n * nb * nn * 8), "M": Dev(hip, n * nb * nb * 8), "b": Dev(hip, n * nb * 8),
            "s": Dev(hip, n * 4)}
    try:
        bufs["v"].put(verts)
        bufs["c"].put(coef)
        lib = L.load()
        L.check(lib.msfec_q_compute_noalloc(ctypes.byref(p), n, bufs["v"].ptr, bufs["c"].ptr,
                                            bufs["P"].ptr, bufs["M"].ptr, bufs["b"].ptr,
                                            bufs["s"].ptr, ctypes.c_void_p(0)))
        hip.ok(hip.L.hipDeviceSynchronize(), "hipDeviceSynchronize")
        return (bufs["M"].get((n, nb, nb), np.float64), bufs["b"].get((n, nb), np.float64),
                bufs["P"].get((n, nb, nn), np.float64), bufs["s"].get((n,), np.int32))
    finally:
        for b in bufs.values():
            b.free()


def test_noalloc_matches_compute_on_two_calls():
    from msfec_amd import fields as F
    hip = Hip()
    dim, l = 3, 3
    # first call: 8 x 8 x 11 = 704 cells, two scratch chunks (498 + 206)
    nc1 = (8, 8, 11)
    verts1 = F.coarse_vertices(dim, 8)
    verts1 = np.concatenate([verts1, verts1[:192] + np.array([0.0, 0.0, 1.0])])
    coef1 = F.lognormal_coefficient(dim, 8, l, 1.5, 1.0 / 64, seed=41)
    coef1 = np.concatenate([coef1, np.exp(np.random.default_rng(42).standard_normal((192, 512)))])
    assert verts1.shape[0] == int(np.prod(nc1))
    # second call: 27 cells with another field, one chunk, scratch holds the first call's data
    verts2 = F.coarse_vertices(dim, 3)
    coef2 = F.lognormal_coefficient(dim, 3, l, 1.0, 0.0, seed=43)
    for k, (v, c) in enumerate(((verts1, coef1), (verts2, coef2))):
        print(f"case noalloc call {k}: {v.shape[0]} cells, 3-D l=3", flush=True)
        want = reference(v, c, dim, l)
        got = noalloc(hip, v, c, dim, l)
        assert (want[3] == 0).all()
        for name, x, y in zip(("M", "b", "Psi", "status"), got, want):
            np.testing.assert_array_equal(x, y, err_msg=f"call {k}: {name}")
