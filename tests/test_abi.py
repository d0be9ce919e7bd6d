"""The C-ABI library loads, exports every symbol include/msfec.h declares, and
rejects invalid requests before touching the GPU.  CPU only: no compute call."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import box_vertices

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(REPO, "include", "msfec.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(msfec_[a-z_0-9]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    from msfec_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libmsfec_hip.so is not built (run __graft_entry__.build())")
    return _lib


def test_exports_every_declared_symbol(lib):
    L = lib.load()
    names = declared_functions()
    assert "msfec_q_create" in names and len(names) >= 15
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(lib.EXPORTED)


def test_version_string(lib):
    assert b"gfx950" in lib.load().msfec_version()


def _create(lib, dim=3, l=2, verts=None, **kw):
    L = lib.load()
    p = lib.Params()
    p.dim, p.n_refine_local, p.coef_mode, p.rhs_mode = dim, l, kw.get("coef_mode", 0), 0
    p.rhs_const, p.reaction_mode, p.reaction_const, p.solver = 1.0, 0, 0.0, kw.get("solver", 0)
    if verts is None:
        verts = box_vertices(dim, [0.0] * dim, [1.0] * dim)[None]
    verts = np.ascontiguousarray(verts, dtype=np.float64)
    h = ctypes.c_void_p()
    st = L.msfec_q_create(ctypes.byref(p), verts.shape[0], ctypes.c_void_p(verts.ctypes.data),
                          lib.HOST, ctypes.byref(h))
    return st, L.msfec_last_error().decode()


def test_rejects_bad_dimension(lib):
    st, msg = _create(lib, dim=4)
    assert st == lib.ERR_INVALID and "dim" in msg


def test_rejects_too_wide_band(lib):
    st, msg = _create(lib, dim=3, l=4)
    assert st == lib.ERR_UNSUPPORTED and "bandwidth" in msg


def test_rejects_non_box_cell(lib):
    v = box_vertices(3, [0, 0, 0], [1, 1, 1])
    v[5, 1] += 0.1
    st, msg = _create(lib, verts=v[None])
    assert st == lib.ERR_UNSUPPORTED and "axis-aligned" in msg


def test_rejects_degenerate_cell(lib):
    v = box_vertices(2, [0, 0], [0, 1])
    st, msg = _create(lib, dim=2, verts=v[None])
    assert st == lib.ERR_INVALID


def test_rejects_unknown_modes(lib):
    st, _ = _create(lib, coef_mode=7)
    assert st == lib.ERR_INVALID
    st, _ = _create(lib, solver=9)
    assert st == lib.ERR_INVALID


def test_mixed_solver_needs_scalar_coefficient(lib):
    """The mixed-precision solver works on the scalar-coefficient row table;
    a tensor coefficient is refused before any device work."""
    st, msg = _create(lib, coef_mode=1, solver=lib.SOLVER_BAND_MIXED)
    assert st == lib.ERR_UNSUPPORTED and "mixed" in msg


def test_null_arguments(lib):
    L = lib.load()
    assert L.msfec_q_create(None, 1, None, 0, None) == lib.ERR_INVALID
    assert L.msfec_q_compute(None, None) == lib.ERR_INVALID
    assert L.msfec_q_get_element(None, None, None, 0) == lib.ERR_INVALID
    L.msfec_q_destroy(None)


def test_product_path_has_no_fallback():
    """The package refuses to run without the HIP library (no CPU path)."""
    from msfec_amd import _lib
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        saved = _lib._lib
        try:
            _lib._lib = None
            _lib.load("/nonexistent/libmsfec_hip.so")
        finally:
            _lib._lib = saved


def _ned_create(lib, l=2, verts=None):
    L = lib.load()
    p = lib.NedParams()
    p.n_refine_local, p.full_rhs, p.flags = l, 1, 0
    if verts is None:
        verts = box_vertices(2, [0.0, 0.0], [1.0, 1.0])[None]
    verts = np.ascontiguousarray(verts, dtype=np.float64)
    h = ctypes.c_void_p()
    st = L.msfec_ned_create(ctypes.byref(p), verts.shape[0], ctypes.c_void_p(verts.ctypes.data),
                            lib.HOST, ctypes.byref(h))
    return st, L.msfec_last_error().decode()


def test_ned_rejects_non_box_and_bad_level(lib):
    v = box_vertices(2, [0, 0], [1, 1])
    v[3, 0] += 0.1
    st, msg = _ned_create(lib, verts=v[None])
    assert st == lib.ERR_INVALID and "axis-aligned" in msg
    st, msg = _ned_create(lib, l=0)
    assert st == lib.ERR_INVALID
    st, msg = _ned_create(lib, l=5)
    assert st == lib.ERR_UNSUPPORTED and "bandwidth" in msg
    L = lib.load()
    assert L.msfec_ned_create(None, 1, None, 0, None) == lib.ERR_INVALID
    assert L.msfec_ned_compute(None, None) == lib.ERR_INVALID


def test_cg_solver_limits(lib):
    """CG lifts the band limit (3-D l = 4 accepted up to the allocation, with
    scalar or tensor coefficients) but refuses cells beyond its workgroup."""
    st, msg = _create(lib, dim=3, l=5, solver=lib.SOLVER_CG)
    assert st == lib.ERR_UNSUPPORTED and "CG" in msg


def _rt_create(lib, l=2, verts=None, coef_mode=0):
    L = lib.load()
    p = lib.RtParams()
    p.n_refine_local, p.flags, p.coef_mode = l, 0, coef_mode
    if verts is None:
        verts = box_vertices(2, [0.0, 0.0], [1.0, 1.0])[None]
    verts = np.ascontiguousarray(verts, dtype=np.float64)
    h = ctypes.c_void_p()
    st = L.msfec_rt_create(ctypes.byref(p), verts.shape[0], ctypes.c_void_p(verts.ctypes.data),
                           lib.HOST, ctypes.byref(h))
    return st, L.msfec_last_error().decode()


def test_rt_rejects_non_box_and_bad_level(lib):
    v = box_vertices(2, [0, 0], [1, 1])
    v[2, 0] -= 0.1
    st, msg = _rt_create(lib, verts=v[None])
    assert st == lib.ERR_INVALID and "axis-aligned" in msg
    st, msg = _rt_create(lib, l=0)
    assert st == lib.ERR_INVALID
    st, msg = _rt_create(lib, l=6)
    assert st == lib.ERR_UNSUPPORTED and "band" in msg
    st, msg = _rt_create(lib, coef_mode=7)
    assert st == lib.ERR_INVALID and "coef_mode" in msg
    L = lib.load()
    assert L.msfec_rt_create(None, 1, None, 0, None) == lib.ERR_INVALID
    assert L.msfec_rt_compute(None, None) == lib.ERR_INVALID
    assert L.msfec_rt_reconstruct(None, None, 0, None, None, 0) == lib.ERR_INVALID


def test_output_buffers_are_never_silently_copied():
    """A getter's out= buffer that the library could not write in place (wrong
    dtype, not C-contiguous, wrong size, read-only) is rejected up front
    instead of being filled through a temporary copy."""
    from msfec_amd.qbasis import _out_ptr
    shape = (3, 8, 8)
    ok = np.empty(shape)
    ptr, loc, keep = _out_ptr(ok, shape)
    assert ptr.value == ok.ctypes.data and keep is ok
    bad = [np.empty(shape, dtype=np.float32), np.empty((8, 8, 3)).transpose(2, 0, 1),
           np.empty((3, 8, 7)), np.empty(shape).tolist()]
    ro = np.empty(shape)
    ro.flags.writeable = False
    bad.append(ro)
    for b in bad:
        with pytest.raises(ValueError):
            _out_ptr(b, shape)


def test_noalloc_rejects_before_touching_the_gpu(lib):
    """The allocation-free entry point validates its request on the host:
    null pointers and unsupported solvers / modes / levels return an error
    before any device call."""
    L = lib.load()
    p = lib.Params()
    p.dim, p.n_refine_local, p.coef_mode, p.rhs_mode, p.rhs_const = 3, 3, 0, 0, 1.0
    p.reaction_mode, p.reaction_const, p.solver = 0, 0.0, lib.SOLVER_BAND_FP64_MFMA
    fake = ctypes.c_void_p(0x1000)  # never dereferenced: every case fails validation first
    args = [fake] * 7 + [ctypes.c_void_p(0)]
    assert L.msfec_q_compute_noalloc(ctypes.byref(p), 4, None, *args[1:]) == lib.ERR_INVALID
    for field, value in (("solver", lib.SOLVER_BAND_FP64), ("solver", lib.SOLVER_CG),
                         ("coef_mode", 1), ("rhs_mode", 1), ("reaction_mode", 2),
                         ("n_refine_local", 4)):
        q = lib.Params()
        ctypes.pointer(q)[0] = p
        setattr(q, field, value)
        st = L.msfec_q_compute_noalloc(ctypes.byref(q), 4, *args)
        assert st == lib.ERR_UNSUPPORTED, (field, value, st)
        assert "noalloc" in L.msfec_last_error().decode() or "band" in L.msfec_last_error().decode()
    q = lib.Params()
    ctypes.pointer(q)[0] = p
    q.dim = 4
    assert L.msfec_q_compute_noalloc(ctypes.byref(q), 4, *args) == lib.ERR_INVALID
    assert L.msfec_q_compute_noalloc(ctypes.byref(p), 0, *args) == lib.ERR_INVALID


def test_band_window_validated_on_the_host(lib):
    """band_window: 0, 16, 32 or 64 slots, and large enough for the half-
    bandwidth + 4; checked before any device call."""
    L = lib.load()
    verts = np.ascontiguousarray(box_vertices(3, [0.0] * 3, [1.0] * 3)[None])
    for window, l, want in ((48, 2, lib.ERR_INVALID), (16, 2, lib.ERR_UNSUPPORTED),
                            (32, 3, lib.ERR_UNSUPPORTED), (16, 3, lib.ERR_UNSUPPORTED)):
        p = lib.Params()
        p.dim, p.n_refine_local, p.rhs_const, p.solver = 3, l, 1.0, lib.SOLVER_BAND_FP64_MFMA
        p.band_window = window
        h = ctypes.c_void_p()
        st = L.msfec_q_create(ctypes.byref(p), 1, ctypes.c_void_p(verts.ctypes.data), lib.HOST,
                              ctypes.byref(h))
        assert st == want, (window, l, st, L.msfec_last_error())
        assert "window" in L.msfec_last_error().decode()


def test_uneven_create_rejects_on_the_host(lib):
    """msfec_qu_create: levels outside the supported range, tensor
    coefficients and unknown solvers are refused before any device call."""
    L = lib.load()
    verts = np.ascontiguousarray(np.stack([box_vertices(3, [0.0] * 3, [1.0] * 3)] * 2))
    cases = (({"solver": lib.SOLVER_AUTO}, [2, 5], lib.ERR_UNSUPPORTED),
             ({"solver": lib.SOLVER_AUTO}, [0, 2], lib.ERR_UNSUPPORTED),
             ({"solver": 9}, [2, 3], lib.ERR_INVALID),
             ({"solver": lib.SOLVER_AUTO, "coef_mode": 1}, [2, 3], lib.ERR_UNSUPPORTED),
             ({"solver": lib.SOLVER_AUTO, "rhs_mode": 2}, [2, 3], lib.ERR_UNSUPPORTED))
    for fields, levels, want in cases:
        p = lib.Params()
        p.dim, p.rhs_const = 3, 1.0
        for k, v in fields.items():
            setattr(p, k, v)
        lv = np.array(levels, dtype=np.int32)
        h = ctypes.c_void_p()
        st = L.msfec_qu_create(ctypes.byref(p), 2, ctypes.c_void_p(verts.ctypes.data), lib.HOST,
                               lv.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), ctypes.byref(h))
        assert st == want, (fields, levels, st, L.msfec_last_error())


def _nrt_create(lib, l=2, verts=None, solver=0, max_iter=0, rtol=0.0):
    L = lib.load()
    p = lib.NrtParams()
    p.n_refine_local, p.flags = l, 0
    p.solver, p.max_iter, p.rtol = solver, max_iter, rtol
    if verts is None:
        verts = box_vertices(3, [0.0] * 3, [1.0] * 3)[None]
    verts = np.ascontiguousarray(verts, dtype=np.float64)
    h = ctypes.c_void_p()
    st = L.msfec_nrt_create(ctypes.byref(p), verts.shape[0], ctypes.c_void_p(verts.ctypes.data),
                            lib.HOST, ctypes.byref(h))
    return st, L.msfec_last_error().decode()


def test_nrt_rejects_non_box_and_bad_level(lib):
    v = box_vertices(3, [0, 0, 0], [1, 1, 1])
    v[6, 2] += 0.05
    st, msg = _nrt_create(lib, verts=v[None])
    assert st == lib.ERR_INVALID and "axis-aligned" in msg
    st, msg = _nrt_create(lib, l=0)
    assert st == lib.ERR_INVALID
    st, msg = _nrt_create(lib, l=5)
    assert st == lib.ERR_UNSUPPORTED and "above 4" in msg
    st, msg = _nrt_create(lib, solver=lib.SOLVER_BAND_MIXED)
    assert st == lib.ERR_INVALID and "solver" in msg
    st, msg = _nrt_create(lib, max_iter=-1)
    assert st == lib.ERR_INVALID
    st, msg = _nrt_create(lib, rtol=float("nan"))
    assert st == lib.ERR_INVALID
    L = lib.load()
    assert L.msfec_nrt_create(None, 1, None, 0, None) == lib.ERR_INVALID
    assert L.msfec_nrt_destroy(None) == lib.OK


def _rt3_create(lib, l=2, verts=None, max_iter=0, rtol=0.0, n=None):
    L = lib.load()
    p = lib.Rt3Params()
    p.n_refine_local, p.flags, p.max_iter, p.rtol = l, 0, max_iter, rtol
    if verts is None:
        verts = box_vertices(3, [0.0] * 3, [1.0] * 3)[None]
    verts = np.ascontiguousarray(verts, dtype=np.float64)
    h = ctypes.c_void_p()
    st = L.msfec_rt3_create(ctypes.byref(p), verts.shape[0] if n is None else n,
                            ctypes.c_void_p(verts.ctypes.data), lib.HOST, ctypes.byref(h))
    return st, L.msfec_last_error().decode()


def test_rt3_rejects_non_box_bad_level_and_arguments(lib):
    """The 3-D RT_DQ batch checks its arguments before any device call."""
    v = box_vertices(3, [0, 0, 0], [1, 1, 1])
    v[3, 0] += 0.05
    st, msg = _rt3_create(lib, verts=v[None])
    assert st == lib.ERR_INVALID and "axis-aligned" in msg
    assert _rt3_create(lib, l=0)[0] == lib.ERR_INVALID
    st, msg = _rt3_create(lib, l=5)
    assert st == lib.ERR_UNSUPPORTED and "above 4" in msg
    assert _rt3_create(lib, max_iter=-1)[0] == lib.ERR_INVALID
    assert _rt3_create(lib, rtol=float("nan"))[0] == lib.ERR_INVALID
    assert _rt3_create(lib, n=0)[0] == lib.ERR_INVALID
    L = lib.load()
    assert L.msfec_rt3_create(None, 1, None, 0, None) == lib.ERR_INVALID
    assert L.msfec_rt3_destroy(None) == lib.OK
    assert L.msfec_rt3_compute(None, None) == lib.ERR_INVALID
    assert L.msfec_rt3_get_element(None, None, None, 0) == lib.ERR_INVALID
