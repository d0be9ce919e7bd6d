"""Device-resident inputs and outputs across the C-ABI (loc = MSFEC_DEVICE):
vertices, coefficients and weights uploaded to device buffers, element
matrices, bases and reconstructions written to device buffers.  Each result
must equal, bit for bit, the host-pointer path on the same data.  The device
buffers come from the HIP runtime the library itself links (hipMalloc through
ctypes), each sized for what the entry point writes."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def hip_runtime():
    """The HIP runtime libmsfec_hip.so itself uses: the library is loaded
    first, then its runtime is opened by soname, which returns the instance
    already in the process (torch, when imported earlier, brings its own copy
    of the runtime, and a second instance opened by another name sees no
    device)."""
    from msfec_amd import _lib
    _lib.load()
    return ctypes.CDLL("libamdhip64.so.7")

DEV = 1


class DevBuf:
    """A hipMalloc'd buffer of n doubles."""
    hip = None

    def __init__(self, n):
        if DevBuf.hip is None:
            h = hip_runtime()
            h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
            h.hipFree.argtypes = [ctypes.c_void_p]
            h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
            h.hipDeviceSynchronize.argtypes = []
            DevBuf.hip = h
        self.n = int(n)
        self.ptr = ctypes.c_void_p()
        assert DevBuf.hip.hipMalloc(ctypes.byref(self.ptr), self.n * 8) == 0

    @classmethod
    def of(cls, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = cls(a.size)
        assert cls.hip.hipMemcpy(b.ptr, a.ctypes.data, b.n * 8, 1) == 0  # host to device
        return b

    def get(self, shape):
        out = np.empty(self.n)
        assert DevBuf.hip.hipMemcpy(out.ctypes.data, self.ptr, self.n * 8, 2) == 0  # device to host
        return out.reshape(shape)

    def free(self):
        if self.ptr:
            DevBuf.hip.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


def check(st):
    from msfec_amd import _lib as L
    L.check(st)


def test_q_device_pointers():
    print("case device pointers q", flush=True)
    from msfec_amd import QBasisBatch, fields as F
    dim, l, nc = 3, 2, 3
    verts = F.coarse_vertices(dim, nc)
    coef = F.lognormal_coefficient(dim, nc, l, 1.0, 0.0, seed=41)
    n, nb, nn = nc ** dim, 1 << dim, ((1 << l) + 1) ** dim
    w = np.random.default_rng(1).normal(size=(n, nb))
    with QBasisBatch(verts, dim=dim, n_refine_local=l, solver=1) as qb:
        qb.set_coefficient(coef)
        qb.run()
        ref = (qb.get_global_element_matrix(), qb.get_global_element_rhs(), qb.get_basis(),
               qb.set_global_weights(w))
    bufs = {k: DevBuf.of(v) for k, v in (("v", verts), ("c", coef), ("w", w))}
    outs = {"M": DevBuf(n * nb * nb), "b": DevBuf(n * nb), "P": DevBuf(n * nb * nn),
            "u": DevBuf(n * nn)}
    try:
        from msfec_amd import _lib as LL
        L = LL.load()
        prm = LL.Params()
        prm.dim, prm.n_refine_local, prm.solver = dim, l, 1
        prm.coef_mode, prm.rhs_mode, prm.rhs_const = 0, 0, 1.0
        h = ctypes.c_void_p()
        check(L.msfec_q_create(ctypes.byref(prm), n, bufs["v"].ptr, DEV, ctypes.byref(h)))
        try:
            check(L.msfec_q_set_coefficient(h, bufs["c"].ptr, coef.size, DEV))
            check(L.msfec_q_compute(h, None))
            check(L.msfec_q_get_element(h, outs["M"].ptr, outs["b"].ptr, DEV))
            check(L.msfec_q_get_basis(h, outs["P"].ptr, DEV))
            check(L.msfec_q_reconstruct(h, bufs["w"].ptr, DEV, outs["u"].ptr, DEV))
        finally:
            L.msfec_q_destroy(h)
        got = (outs["M"].get((n, nb, nb)), outs["b"].get((n, nb)), outs["P"].get((n, nb, nn)),
               outs["u"].get((n, nn)))
        for x, y in zip(got, ref):
            np.testing.assert_array_equal(x, y)
    finally:
        for b in list(bufs.values()) + list(outs.values()):
            b.free()


def test_ned_device_pointers():
    print("case device pointers ned", flush=True)
    from msfec_amd import NedBasisBatch, fields as F
    from msfec_amd import _lib as LL
    l, nc = 3, 3
    verts = F.coarse_vertices(2, nc)
    a = F.lognormal_coefficient(2, nc, l, 1.0, 0.0, seed=43)
    n, nf = nc * nc, (1 << l) ** 2
    from msfec_amd.ned import ned_sizes
    N = ned_sizes(l)[3]
    binv, r, f = np.ones((n, nf)), np.zeros((n, nf)), np.ones((n, nf, 2))
    with NedBasisBatch(verts, n_refine_local=l) as nb:
        nb.set_coefficients(a, binv, r, f)
        nb.run()
        ref = (nb.get_global_element_matrix(), nb.get_global_element_rhs(), nb.get_basis())
    bufs = {k: DevBuf.of(v) for k, v in (("v", verts), ("a", a), ("bi", binv), ("r", r), ("f", f))}
    outs = {"M": DevBuf(n * 64), "b": DevBuf(n * 8), "P": DevBuf(n * 8 * N)}
    try:
        L = LL.load()
        prm = LL.NedParams()
        prm.n_refine_local, prm.full_rhs, prm.flags = l, 1, 0
        h = ctypes.c_void_p()
        check(L.msfec_ned_create(ctypes.byref(prm), n, bufs["v"].ptr, DEV, ctypes.byref(h)))
        try:
            check(L.msfec_ned_set_coefficients(h, bufs["a"].ptr, bufs["bi"].ptr, bufs["r"].ptr,
                                               bufs["f"].ptr, DEV))
            check(L.msfec_ned_compute(h, None))
            check(L.msfec_ned_get_element(h, outs["M"].ptr, outs["b"].ptr, DEV))
            check(L.msfec_ned_get_basis(h, outs["P"].ptr, DEV))
        finally:
            L.msfec_ned_destroy(h)
        got = (outs["M"].get((n, 8, 8)), outs["b"].get((n, 8)), outs["P"].get((n, 8, N)))
        for x, y in zip(got, ref):
            np.testing.assert_array_equal(x, y)
    finally:
        for b in list(bufs.values()) + list(outs.values()):
            b.free()


def test_rt_device_pointers():
    print("case device pointers rt", flush=True)
    from msfec_amd import RTDQBasisBatch, fields as F
    from msfec_amd import _lib as LL
    l, nc = 3, 3
    verts = F.coarse_vertices(2, nc)
    a = F.lognormal_coefficient(2, nc, l, 1.0, 0.0, seed=47)
    n, m = nc * nc, 1 << l
    nf, ne = m * m, 2 * m * (m + 1)
    f = np.random.default_rng(2).uniform(0.5, 2.0, size=(n, nf))
    w = np.random.default_rng(3).normal(size=(n, 5))
    with RTDQBasisBatch(verts, n_refine_local=l) as rb:
        rb.set_coefficients(a, f)
        rb.run()
        ref = (rb.get_global_element_matrix(), rb.get_global_element_rhs(), rb.get_basis(),
               *rb.set_global_weights(w))
    bufs = {k: DevBuf.of(v) for k, v in (("v", verts), ("a", a), ("f", f), ("w", w))}
    outs = {"M": DevBuf(n * 25), "b": DevBuf(n * 5), "S": DevBuf(n * 4 * ne),
            "s": DevBuf(n * ne), "u": DevBuf(n * nf)}
    try:
        L = LL.load()
        prm = LL.RtParams()
        prm.n_refine_local, prm.flags = l, 0
        h = ctypes.c_void_p()
        check(L.msfec_rt_create(ctypes.byref(prm), n, bufs["v"].ptr, DEV, ctypes.byref(h)))
        try:
            check(L.msfec_rt_set_coefficients(h, bufs["a"].ptr, bufs["f"].ptr, DEV))
            check(L.msfec_rt_compute(h, None))
            check(L.msfec_rt_get_element(h, outs["M"].ptr, outs["b"].ptr, DEV))
            check(L.msfec_rt_get_basis(h, outs["S"].ptr, DEV))
            check(L.msfec_rt_reconstruct(h, bufs["w"].ptr, DEV, outs["s"].ptr, outs["u"].ptr, DEV))
        finally:
            L.msfec_rt_destroy(h)
        got = (outs["M"].get((n, 5, 5)), outs["b"].get((n, 5)), outs["S"].get((n, 4, ne)),
               outs["s"].get((n, ne)), outs["u"].get((n, nf)))
        for x, y in zip(got, ref):
            np.testing.assert_array_equal(x, y)
    finally:
        for b in list(bufs.values()) + list(outs.values()):
            b.free()
