"""The MATLAB gateway multem_amd/mex/ilc_multem_hip.cpp, compiled against
the emulated matrix API of tests/mex_host and driven with the MATLAB struct
that multem_input.parameters builds (field names of parameters.m:6-239).

CPU tests: argument handling and the errors the engine reports before any
device work.  GPU tests: every output kind of set_struct_multislice
(ilc_multem.cu:381-481) equals, bit for bit, what the C ABI returns to the
Python binding for the same input.
"""
import ctypes as C
import os

import numpy as np
import pytest

from multem_amd import configs, engine

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "mex_host", "libmexhost.so")

_dp = C.POINTER(C.c_double)


def _lib():
    if not os.path.exists(LIB):
        pytest.skip("tests/mex_host not built (make -C tests/mex_host)")
    L = C.CDLL(LIB)
    L.mh_double.restype = C.c_void_p
    L.mh_double.argtypes = [C.c_size_t, C.c_size_t, _dp, _dp]
    L.mh_struct.restype = C.c_void_p
    L.mh_struct.argtypes = [C.c_size_t]
    L.mh_set.argtypes = [C.c_void_p, C.c_size_t, C.c_char_p, C.c_void_p]
    L.mh_get.restype = C.c_void_p
    L.mh_get.argtypes = [C.c_void_p, C.c_size_t, C.c_char_p]
    L.mh_shape.argtypes = [C.c_void_p] + [C.POINTER(C.c_size_t)] * 2 + [C.POINTER(C.c_int)] * 2
    L.mh_data.argtypes = [C.c_void_p, _dp, _dp]
    L.mh_free.argtypes = [C.c_void_p]
    L.mh_call.argtypes = [C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_char_p, C.c_size_t]
    return L


def _mat(L, a):
    """numpy array -> MATLAB double matrix (column-major); 1-D arrays are rows."""
    a = np.asarray(a)
    if a.ndim == 0:
        a = a.reshape(1, 1)
    elif a.ndim == 1:
        a = a.reshape(1, -1)
    m, n = a.shape
    re = np.ascontiguousarray(np.real(a).T, dtype=np.float64).ravel()
    im = np.ascontiguousarray(np.imag(a).T, dtype=np.float64).ravel() if np.iscomplexobj(a) else None
    return L.mh_double(m, n, re.ctypes.data_as(_dp), im.ctypes.data_as(_dp) if im is not None else None)


def to_mx(L, p):
    """multem_amd Parameters -> (system_conf, input_multislice) MATLAB structs."""
    sc = L.mh_struct(1)
    for k, v in p.system_conf.__dict__.items():
        L.mh_set(sc, 0, k.encode(), _mat(L, float(v)))
    s = L.mh_struct(1)
    for k, v in p.__dict__.items():
        if k in ("system_conf", "detector_cir", "detector_type", "detector_fR") or v is None:
            continue
        L.mh_set(s, 0, k.encode(), _mat(L, v if isinstance(v, np.ndarray) else float(v)))
    det = L.mh_struct(1)
    L.mh_set(det, 0, b"type", _mat(L, float(p.detector_type)))
    cir = L.mh_struct(len(p.detector_cir))
    for i, (a, b) in enumerate(p.detector_cir):
        L.mh_set(cir, i, b"inner_ang", _mat(L, float(a)))
        L.mh_set(cir, i, b"outer_ang", _mat(L, float(b)))
    L.mh_set(det, 0, b"cir", cir)
    if p.detector_fR:
        mat = L.mh_struct(len(p.detector_fR))
        for i, f in enumerate(p.detector_fR):
            L.mh_set(mat, i, b"fR", _mat(L, np.asarray(f, dtype=np.float64)))
        L.mh_set(det, 0, b"matrix", mat)
    L.mh_set(s, 0, b"detector", det)
    return sc, s


def get(L, a, *path):
    """Follows (name, index) steps; returns the numpy value of a numeric leaf."""
    for name, idx in path:
        a = L.mh_get(a, idx, name.encode())
        assert a, name
    m, n, cx, st = C.c_size_t(), C.c_size_t(), C.c_int(), C.c_int()
    L.mh_shape(a, C.byref(m), C.byref(n), C.byref(cx), C.byref(st))
    if st.value:
        return (m.value, n.value)
    re, im = np.zeros(m.value * n.value), np.zeros(m.value * n.value)
    L.mh_data(a, re.ctypes.data_as(_dp), im.ctypes.data_as(_dp))
    v = re + 1j * im if cx.value else re
    return v.reshape(n.value, m.value)     # column-major m x n -> [col][row]: MATLAB (iy, ix) -> [ix, iy]


def call(L, *args):
    arr = (C.c_void_p * len(args))(*args)
    out = C.c_void_p()
    err = C.create_string_buffer(1024)
    rc = L.mh_call(len(args), arr, C.byref(out), err, 1024)
    return rc, out.value, err.value.decode()


def test_gateway_reports_engine_errors():
    L = _lib()
    p = configs.si_ewrs_256(nc=2)
    p.simulation_type = 0                       # not a simulation type
    sc, s = to_mx(L, p)
    rc, _, err = call(L, sc, s)
    assert rc == 1 and "simulation_type" in err, err
    p = configs.si_ewrs_256(nc=2)
    p.system_conf.device = 1                    # CPU: refused before any device work
    sc, s = to_mx(L, p)
    rc, _, err = call(L, sc, s)
    assert rc == 1 and "CPU" in err, err
    rc, _, err = call(L)                        # no arguments
    assert rc == 1 and "usage" in err
    for a in (sc, s):
        L.mh_free(a)


def _compare(L, p):
    sc, s = to_mx(L, p)
    rc, o, err = call(L, sc, s)
    assert rc == 0, err
    ref = engine.run_multislice(p)
    assert get(L, o, ("dx", 0))[0, 0] == ref.dx and get(L, o, ("dy", 0))[0, 0] == ref.dy
    assert np.array_equal(get(L, o, ("thick", 0)).ravel(), ref.thick[:ref.n_thk])
    assert np.array_equal(get(L, o, ("x", 0)).ravel(), ref.x[:ref.nx])
    assert get(L, o, ("data", 0)) == (1, ref.n_thk)
    for k in range(ref.n_thk):
        if ref.image_tot is not None:
            assert get(L, o, ("data", 0), ("image_tot", k)) == (1, ref.n_det)
            for d in range(ref.n_det):
                assert np.array_equal(get(L, o, ("data", 0), ("image_tot", k), ("image", d)), ref.image_tot[k, d])
                if ref.image_coh is not None:
                    assert np.array_equal(get(L, o, ("data", 0), ("image_coh", k), ("image", d)), ref.image_coh[k, d])
        if ref.m2psi_tot is not None:
            assert np.array_equal(get(L, o, ("data", 0), ("m2psi_tot", k)), ref.m2psi_tot[k])
        if ref.m2psi_coh is not None:
            assert np.array_equal(get(L, o, ("data", 0), ("m2psi_coh", k)), ref.m2psi_coh[k])
        if ref.psi_coh is not None:
            assert np.array_equal(get(L, o, ("data", 0), ("psi_coh", k)), ref.psi_coh[k])
    for a in (sc, s, o):
        L.mh_free(a)


@pytest.mark.gpu
def test_gateway_stem_two_detectors_coherent():
    L = _lib()
    p = configs.au_stem_512(ns=3, nconf=2, n=256, ncells=4, nc=3)
    p.detector_cir = [(20.0, 60.0), (80.0, 200.0)]
    p.pn_coh_contrib = 1
    _compare(L, p)


@pytest.mark.gpu
def test_gateway_cbed_hrtem_ewrs():
    L = _lib()
    _compare(L, configs.pt_cbed_1024(nconf=2, n=256, na=4, nb=6, nc=8, thick=[10.0, 20.0]))
    p = configs.si_ewrs_256(nc=3)
    _compare(L, p)                              # psi_coh (complex) and m2psi_tot
    p.simulation_type = 32
    p.obj_lens_c_10, p.obj_lens_outer_aper_ang = -40.0, 25.0
    _compare(L, p)
    p = configs.si_ewrs_256(nc=3)
    p.system_conf.precision = 2                 # the double path through the same gateway
    _compare(L, p)


@pytest.mark.gpu
def test_gateway_matrix_detectors():
    L = _lib()
    p = configs.au_stem_512(ns=3, nconf=1, n=256, ncells=4, nc=3)
    p.detector_type = 3
    rng = np.random.default_rng(5)
    p.detector_fR = [rng.random((p.ny, p.nx)) for _ in range(3)]
    _compare(L, p)


@pytest.mark.gpu
def test_gateway_precession_and_hollow_cone():
    """ped_theta / ped_nrot and hci_theta / hci_nrot reach the engine through
    the gateway (ilc_multem.cu:338-347)."""
    L = _lib()
    p = configs.si_ewrs_256(nc=3, n=128)
    p.simulation_type = 41
    p.ped_theta, p.ped_nrot = 1.0, 5
    _compare(L, p)
    p.simulation_type = 42
    p.hci_theta, p.hci_nrot = 0.8, 3
    p.obj_lens_c_10, p.obj_lens_outer_aper_ang = -40.0, 25.0
    _compare(L, p)
