"""The compute entry points under HIP graph capture: msfec_q_compute,
msfec_ned_compute, msfec_rt_compute and msfec_nrt_compute enqueue kernels only (no memsets, no
host synchronisation), so a stream capture of one call replays the whole
batch.  Each graph is replayed three times with new coefficients uploaded in
between (same device buffers), with eager work and synchronisations between
the replays, and every replay is checked against an eager compute of the
same data (bit for bit: the same kernels on the same inputs)."""
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


class Hip:
    def __init__(self):
        self.L = hip_runtime()
        vp = ctypes.c_void_p
        for name, args in (("hipStreamCreate", [ctypes.POINTER(vp)]),
                           ("hipStreamDestroy", [vp]),
                           ("hipStreamSynchronize", [vp]),
                           ("hipStreamBeginCapture", [vp, ctypes.c_int]),
                           ("hipStreamEndCapture", [vp, ctypes.POINTER(vp)]),
                           ("hipGraphInstantiate", [ctypes.POINTER(vp), vp, vp, vp, ctypes.c_size_t]),
                           ("hipGraphLaunch", [vp, vp]),
                           ("hipGraphExecDestroy", [vp]),
                           ("hipGraphDestroy", [vp])):
            fn = getattr(self.L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int

    def ok(self, st, what):
        assert st == 0, f"{what} returned {st}"

    def capture(self, stream, enqueue):
        """Graph exec of what enqueue() puts on stream."""
        self.ok(self.L.hipStreamBeginCapture(stream, 1), "hipStreamBeginCapture")  # thread-local mode
        enqueue()
        g = ctypes.c_void_p()
        self.ok(self.L.hipStreamEndCapture(stream, ctypes.byref(g)), "hipStreamEndCapture")
        ex = ctypes.c_void_p()
        self.ok(self.L.hipGraphInstantiate(ctypes.byref(ex), g, None, None, 0), "hipGraphInstantiate")
        return g, ex


def replay_check(batch, stream, set_data, results, datasets):
    """Eager results per dataset, then a captured compute replayed per dataset."""
    hip = Hip()
    eager = []
    for d in datasets:
        set_data(d)
        batch.run(stream.value)
        eager.append(results())
    set_data(datasets[0])
    g, ex = hip.capture(stream, lambda: batch.run(stream.value))
    try:
        for rep, k in enumerate([0, 1, 0]):
            set_data(datasets[k])
            hip.ok(hip.L.hipGraphLaunch(ex, stream), "hipGraphLaunch")
            hip.ok(hip.L.hipStreamSynchronize(stream), "hipStreamSynchronize")
            got = results()
            for x, y in zip(got, eager[k]):
                np.testing.assert_array_equal(x, y, err_msg=f"replay {rep} (data {k})")
            batch.run(stream.value)  # eager work between the replays
            hip.ok(hip.L.hipStreamSynchronize(stream), "hipStreamSynchronize")
    finally:
        hip.L.hipGraphExecDestroy(ex)
        hip.L.hipGraphDestroy(g)


@pytest.fixture()
def stream():
    hip = Hip()
    s = ctypes.c_void_p()
    hip.ok(hip.L.hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
    yield s
    hip.L.hipStreamSynchronize(s)
    hip.L.hipStreamDestroy(s)


@pytest.mark.parametrize("solver", [0, 1, 2, 3])
def test_q_compute_graph_replay(stream, solver):
    print(f"case graph q solver={solver}", flush=True)
    from msfec_amd import QBasisBatch, fields as F
    dim, l, nc = 3, 2, 3
    verts = F.coarse_vertices(dim, nc)
    data = [F.lognormal_coefficient(dim, nc, l, 1.0, 0.0, seed=s) for s in (5, 6)]
    with QBasisBatch(verts, dim=dim, n_refine_local=l, solver=solver) as qb:
        replay_check(qb, stream, qb.set_coefficient,
                     lambda: (qb.get_global_element_matrix(), qb.get_global_element_rhs(),
                              qb.get_basis(), qb.get_cell_status()), data)


def test_ned_compute_graph_replay(stream):
    print("case graph ned", flush=True)
    from msfec_amd import NedBasisBatch, fields as F
    verts = F.coarse_vertices(2, 4)
    data = [F.lognormal_coefficient(2, 4, 3, 1.0, 0.0, seed=s) for s in (7, 8)]
    with NedBasisBatch(verts, n_refine_local=3) as nb:
        replay_check(nb, stream, lambda a: nb.set_coefficients(a, 1.0, 0.0, (1.0, 1.0)),
                     lambda: (nb.get_global_element_matrix(), nb.get_global_element_rhs(),
                              nb.get_basis(), nb.get_cell_status()), data)


def test_rt_compute_graph_replay(stream):
    print("case graph rt", flush=True)
    from msfec_amd import RTDQBasisBatch, fields as F
    verts = F.coarse_vertices(2, 4)
    data = [F.lognormal_coefficient(2, 4, 3, 1.0, 0.0, seed=s) for s in (9, 10)]
    with RTDQBasisBatch(verts, n_refine_local=3) as rb:
        replay_check(rb, stream, rb.set_coefficients,
                     lambda: (rb.get_global_element_matrix(), rb.get_global_element_rhs(),
                              rb.get_basis(), rb.get_cell_status()), data)


def test_uneven_compute_graph_replay(stream):
    """msfec_qu_compute forks the levels onto streams of their own and joins
    them back with event waits: captured from the caller's stream, the graph
    holds every level's kernels.  The first eager compute does the timing."""
    print("case graph uneven", flush=True)
    from msfec_amd import UnevenBatch, fields as F
    nc = 3
    cen = F.inclusion_centres(3, 2, seed=71)
    lv = F.uneven_levels(3, nc, cen, seed=72)
    verts = F.coarse_vertices(3, nc)
    data = [F.inclusion_blocks(3, nc, lv, cen, r, a) for r, a in ((0.1, 1e3), (0.15, 30.0))]
    with UnevenBatch(verts, lv, dim=3) as ub:
        replay_check(ub, stream, ub.set_coefficients,
                     lambda: (ub.get_global_element_matrix(), ub.get_global_element_rhs(),
                              ub.get_basis_flat(), ub.get_cell_status()), data)


@pytest.mark.parametrize("l,solver", [(2, 0), (2, 3), (3, 0)])
def test_nrt_compute_graph_replay(stream, l, solver):
    """msfec_nrt_compute on both paths (the dense kernel; MINRES with its
    five launches) replayed from a graph, bit for bit the eager results."""
    print(f"case graph nrt l={l} solver={solver}", flush=True)
    from msfec_amd import NedRTBasisBatch
    from test_gpu_nrt import boxes, rand_coefs
    rng = np.random.default_rng(51 + l)
    n = 8
    verts = boxes(rng, n)
    data = [rand_coefs(rng, n, l) for _ in range(2)]
    with NedRTBasisBatch(verts, n_refine_local=l, solver=solver) as nb:
        replay_check(nb, stream, lambda d: nb.set_coefficients(*d),
                     lambda: (nb.get_global_element_matrix(), nb.get_global_element_rhs(),
                              nb.get_basis(), nb.get_cell_status()), data)


@pytest.mark.parametrize("l", [2, 3])
def test_rt3_compute_graph_replay(stream, l):
    """msfec_rt3_compute (five kernel launches, no memset) replayed from a
    graph, bit for bit the eager results."""
    print(f"case graph rt3 l={l}", flush=True)
    from msfec_amd import RTDQ3BasisBatch
    from test_gpu_nrt import boxes, rand_coefs
    rng = np.random.default_rng(71 + l)
    n = 8
    verts = boxes(rng, n)
    data = [(rand_coefs(rng, n, l, reaction=False, load=False)[0],
             rng.normal(size=(n, (1 << l) ** 3, 8))) for _ in range(2)]
    with RTDQ3BasisBatch(verts, n_refine_local=l) as rb:
        replay_check(rb, stream, lambda d: rb.set_coefficients(*d),
                     lambda: (rb.get_global_element_matrix(), rb.get_global_element_rhs(),
                              rb.get_basis(), rb.get_cell_status()), data)
