"""Input uploads are ordered on the handle's stream.  On a non-blocking
stream (one that does not wait for the default stream) compute ->
set_coefficient -> compute runs with no host synchronisation in between: the
first compute must finish with the first input and the second must see the
second, for host-resident and for device-resident inputs, in each of the
three pairings.  Every result is compared bit for bit with an eager run."""
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
        for name, args in (("hipStreamCreateWithFlags", [ctypes.POINTER(vp), ctypes.c_uint]),
                           ("hipStreamDestroy", [vp]),
                           ("hipStreamSynchronize", [vp]),
                           ("hipMalloc", [ctypes.POINTER(vp), ctypes.c_size_t]),
                           ("hipFree", [vp]),
                           ("hipMemcpyAsync", [vp, vp, ctypes.c_size_t, ctypes.c_int, vp])):
            getattr(self.L, name).argtypes = args
            getattr(self.L, name).restype = ctypes.c_int

    def ok(self, st, what):
        assert st == 0, f"{what} returned {st}"


class DevBuf:
    """A device copy of a float64 array, written on `stream` (so its producer
    is on the handle's stream, as msfec.h requires of device inputs)."""
    def __init__(self, hip, arr, stream):
        self.hip, self.arr = hip, np.ascontiguousarray(arr, dtype=np.float64)
        self.ptr = ctypes.c_void_p()
        hip.ok(hip.L.hipMalloc(ctypes.byref(self.ptr), self.arr.nbytes), "hipMalloc")
        hip.ok(hip.L.hipMemcpyAsync(self.ptr, ctypes.c_void_p(self.arr.ctypes.data), self.arr.nbytes,
                                    1, stream), "hipMemcpyAsync")
        hip.ok(hip.L.hipStreamSynchronize(stream), "hipStreamSynchronize")

    def free(self):
        self.hip.L.hipFree(self.ptr)


@pytest.fixture()
def nb_stream():
    hip = Hip()
    s = ctypes.c_void_p()
    hip.ok(hip.L.hipStreamCreateWithFlags(ctypes.byref(s), 1), "hipStreamCreateWithFlags")  # non-blocking
    yield hip, s
    hip.L.hipStreamSynchronize(s)
    hip.L.hipStreamDestroy(s)


def run_overlapped(batch, stream, set_data, results, d0, d1):
    """eager(d0), eager(d1); then compute(d0) -> set(d1) -> read (must be d0's),
    compute(d1) -> read (must be d1's), with no host sync between the calls."""
    eager = []
    for d in (d0, d1):
        set_data(d)
        batch.run(stream.value)
        eager.append(results())
    set_data(d0)
    batch.run(stream.value)
    set_data(d1)                 # must wait for the compute in flight
    got0 = results()
    batch.run(stream.value)
    got1 = results()
    for k, got in ((0, got0), (1, got1)):
        for x, y in zip(got, eager[k]):
            np.testing.assert_array_equal(x, y, err_msg=f"compute with input {k}")


@pytest.mark.parametrize("loc", ["host", "device"])
def test_q_uploads_ordered_on_nonblocking_stream(nb_stream, loc):
    print(f"case streams q loc={loc}", flush=True)
    from msfec_amd import QBasisBatch, _lib as L, fields as F
    hip, s = nb_stream
    dim, l, nc = 3, 3, 8            # 512 cells: long enough for a race to show
    verts = F.coarse_vertices(dim, nc)
    data = [F.lognormal_coefficient(dim, nc, l, 1.5, 1.0 / 64, seed=k) for k in (31, 32)]
    with QBasisBatch(verts, dim=dim, n_refine_local=l, solver=L.SOLVER_BAND_FP64_MFMA) as qb:
        if loc == "host":
            setter = qb.set_coefficient
            d0, d1 = data
            bufs = []
        else:
            bufs = [DevBuf(hip, d, s) for d in data]
            d0, d1 = bufs

            def setter(b):
                L.check(qb._lib.msfec_q_set_coefficient(qb._h, b.ptr, qb.info.coef_count, L.DEVICE))
        try:
            run_overlapped(qb, s, setter, lambda: (qb.get_global_element_matrix(), qb.get_basis()),
                           d0, d1)
        finally:
            for b in bufs:
                b.free()


def test_ned_uploads_ordered_on_nonblocking_stream(nb_stream):
    print("case streams ned", flush=True)
    from msfec_amd import NedBasisBatch, fields as F
    hip, s = nb_stream
    verts = F.coarse_vertices(2, 16)
    data = [F.lognormal_coefficient(2, 16, 4, 1.0, 0.0, seed=k) for k in (33, 34)]
    with NedBasisBatch(verts, n_refine_local=4) as nb:
        run_overlapped(nb, s, lambda a: nb.set_coefficients(a, 1.0, 0.0, (1.0, 1.0)),
                       lambda: (nb.get_global_element_matrix(), nb.get_basis()), *data)


def test_rt_uploads_ordered_on_nonblocking_stream(nb_stream):
    print("case streams rt", flush=True)
    from msfec_amd import RTDQBasisBatch, fields as F
    hip, s = nb_stream
    verts = F.coarse_vertices(2, 16)
    data = [F.lognormal_coefficient(2, 16, 4, 1.0, 0.0, seed=k) for k in (35, 36)]
    with RTDQBasisBatch(verts, n_refine_local=4) as rb:
        run_overlapped(rb, s, rb.set_coefficients,
                       lambda: (rb.get_global_element_matrix(), rb.get_basis()), *data)


@pytest.mark.parametrize("l,n", [(2, 1024), (3, 8)])
def test_nrt_uploads_ordered_on_nonblocking_stream(nb_stream, l, n):
    print(f"case streams nrt l={l} n={n}", flush=True)
    from msfec_amd import NedRTBasisBatch
    from test_gpu_nrt import boxes, rand_coefs
    hip, s = nb_stream
    rng = np.random.default_rng(37 + l)
    verts = boxes(rng, n)
    d0, d1 = rand_coefs(rng, n, l), rand_coefs(rng, n, l)
    with NedRTBasisBatch(verts, n_refine_local=l) as nb:
        run_overlapped(nb, s, lambda d: nb.set_coefficients(*d),
                       lambda: (nb.get_global_element_matrix(), nb.get_basis()), d0, d1)


def test_rt3_uploads_ordered_on_nonblocking_stream(nb_stream):
    print("case streams rt3 l=3 n=8", flush=True)
    from msfec_amd import RTDQ3BasisBatch
    from test_gpu_nrt import boxes, rand_coefs
    hip, s = nb_stream
    rng = np.random.default_rng(61)
    n, l = 8, 3
    verts = boxes(rng, n)
    d0, d1 = [(rand_coefs(rng, n, l, reaction=False, load=False)[0],
               rng.normal(size=(n, (1 << l) ** 3, 8))) for _ in range(2)]
    with RTDQ3BasisBatch(verts, n_refine_local=l) as rb:
        run_overlapped(rb, s, lambda d: rb.set_coefficients(*d),
                       lambda: (rb.get_global_element_matrix(), rb.get_basis()), d0, d1)
