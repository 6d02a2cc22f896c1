"""Python loader for the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use
this module; it is the checker the HIP engine is compared against.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from multem_amd import capi

_HERE = os.path.dirname(os.path.abspath(__file__))
_libs = {}


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib(precision="f64"):
    if precision not in _libs:
        path = os.path.join(_HERE, "liboracle_%s.so" % precision)
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        P = C.POINTER
        sysp, inp, outp = P(capi.SystemConf), P(capi.InputMultislice), P(capi.OutputMultislice)
        dp = P(C.c_double)
        L.orc_last_error.restype = C.c_char_p
        L.orc_run.argtypes = [sysp, inp, outp, C.c_int]
        L.orc_output_dims.argtypes = [sysp, inp, outp]
        L.orc_spec_slicing.argtypes = [sysp, inp, C.c_int, P(capi.Slice), C.c_int, P(C.c_int)]
        L.orc_projected_potential.argtypes = [sysp, inp, C.c_int, C.c_int, dp]
        L.orc_transmission_function.argtypes = [sysp, inp, C.c_int, C.c_int, dp]
        L.orc_incident_wave.argtypes = [sysp, inp, dp]
        L.orc_propagate.argtypes = [sysp, inp, dp, C.c_double, C.c_int, dp]
        L.orc_displaced_atoms.argtypes = [inp, C.c_int, dp, C.c_int, P(C.c_int)]
        L.orc_atom_table.argtypes = [C.c_int, C.c_double] + [dp] * 7
        L.orc_lobato_VR.argtypes = [C.c_int, C.c_int, dp, dp, dp, dp]
        L.orc_subslice_rule.argtypes = [dp, dp]
        L.orc_subslice_table.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double] + [dp] * 4
        L.orc_bessel.argtypes = [C.c_int, C.c_double]
        L.orc_bessel.restype = C.c_double
        L.orc_lambda.argtypes = [C.c_double]
        L.orc_lambda.restype = C.c_double
        L.orc_vr_factor.argtypes = [C.c_double, C.c_double]
        L.orc_vr_factor.restype = C.c_double
        L.orc_fft2.argtypes = [C.c_int, C.c_int, dp, C.c_int]
        L.orc_mt19937_64.argtypes = [C.c_uint64, C.c_int, P(C.c_uint64)]
        L.orc_uniform_int.argtypes = [C.c_uint64, C.c_int, P(C.c_int)]
        L.orc_normal.argtypes = [C.c_uint64, C.c_int, dp]
        L.orc_gauss_rule.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, dp, dp]
        L.orc_lens_quadrature.argtypes = [inp, C.c_int, C.c_int, dp, dp, dp]
        _libs[precision] = L
    return _libs[precision]


class OracleError(RuntimeError):
    pass


def _check(L, st):
    if st != 0:
        raise OracleError("%s: %s" % (capi.STATUS_NAMES.get(st, st), L.orc_last_error().decode()))


def default_threads():
    """Host threads for the oracle: the cores this process may use, at most 16
    (the CPU share of one GPU's host)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return max(1, min(n, 16))


def run(params, precision="f64", nthreads=None):
    L = lib(precision)
    s, keep = params.to_struct()
    sysc = params.system_conf.as_struct()
    d = capi.OutputMultislice()
    _check(L, L.orc_output_dims(C.byref(sysc), C.byref(s), C.byref(d)))
    out = capi.Output(d)
    o = out.as_struct()
    nt = nthreads if nthreads else default_threads()
    _check(L, L.orc_run(C.byref(sysc), C.byref(s), C.byref(o), int(nt)))
    return out


def spec_slicing(params, iconf=0, precision="f64"):
    L = lib(precision)
    s, keep = params.to_struct()
    sysc = params.system_conf.as_struct()
    n = C.c_int(0)
    _check(L, L.orc_spec_slicing(C.byref(sysc), C.byref(s), iconf, None, 0, C.byref(n)))
    arr = (capi.Slice * max(n.value, 1))()
    _check(L, L.orc_spec_slicing(C.byref(sysc), C.byref(s), iconf, arr, n.value, C.byref(n)))
    return [(a.z_0, a.z_e, a.z_int_0, a.z_int_e, a.iatom_0, a.iatom_e, a.ithk) for a in arr[:n.value]]


def projected_potential(params, islice, iconf=1, precision="f64"):
    L = lib(precision)
    s, keep = params.to_struct()
    sysc = params.system_conf.as_struct()
    V = np.zeros(params.nx * params.ny)
    _check(L, L.orc_projected_potential(C.byref(sysc), C.byref(s), iconf, islice, capi.dptr(V)))
    return V.reshape(params.nx, params.ny)


def transmission_function(params, islice, iconf=1, precision="f64"):
    L = lib(precision)
    s, keep = params.to_struct()
    sysc = params.system_conf.as_struct()
    T = np.zeros(2 * params.nx * params.ny)
    _check(L, L.orc_transmission_function(C.byref(sysc), C.byref(s), iconf, islice, capi.dptr(T)))
    return T.view(np.complex128).reshape(params.nx, params.ny)


def incident_wave(params, precision="f64"):
    L = lib(precision)
    s, keep = params.to_struct()
    sysc = params.system_conf.as_struct()
    P = np.zeros(2 * params.nx * params.ny)
    _check(L, L.orc_incident_wave(C.byref(sysc), C.byref(s), capi.dptr(P)))
    return P.view(np.complex128).reshape(params.nx, params.ny)


def propagate(params, psi, z, space=1, precision="f64"):
    L = lib(precision)
    s, keep = params.to_struct()
    sysc = params.system_conf.as_struct()
    pin = np.ascontiguousarray(psi, dtype=np.complex128).view(np.float64).ravel()
    pout = np.zeros_like(pin)
    _check(L, L.orc_propagate(C.byref(sysc), C.byref(s), capi.dptr(pin), float(z), int(space), capi.dptr(pout)))
    return pout.view(np.complex128).reshape(params.nx, params.ny)


def displaced_atoms(params, iconf, precision="f64"):
    L = lib(precision)
    s, keep = params.to_struct()
    n = C.c_int(0)
    _check(L, L.orc_displaced_atoms(C.byref(s), iconf, None, 0, C.byref(n)))
    buf = np.zeros(4 * max(n.value, 1))
    _check(L, L.orc_displaced_atoms(C.byref(s), iconf, capi.dptr(buf), n.value, C.byref(n)))
    return buf[:4 * n.value].reshape(-1, 4)


def atom_table(Z, dR_min, precision="f64"):
    L = lib(precision)
    arrs = [np.zeros(128) for _ in range(6)]
    scal = np.zeros(4)
    _check(L, L.orc_atom_table(Z, float(dR_min), *[capi.dptr(a) for a in arrs], capi.dptr(scal)))
    R, R2, c0, c1, c2, c3 = arrs
    return dict(R=R, R2=R2, c0=c0, c1=c1, c2=c2, c3=c3, R_min=scal[0], R_max=scal[1], R2_tap=scal[2], tap_cf=scal[3])


def fft2(a, inverse=False, precision="f64"):
    L = lib(precision)
    nx, ny = a.shape
    buf = np.ascontiguousarray(a, dtype=np.complex128).view(np.float64).ravel().copy()
    _check(L, L.orc_fft2(nx, ny, capi.dptr(buf), int(bool(inverse))))
    return buf.view(np.complex128).reshape(nx, ny)


def gauss_rule(kind, n, alpha=0.0, a=0.0, b=1.0, precision="f64"):
    """Gauss rule of kind 5 (generalized Laguerre) or 6 (generalized Hermite)."""
    L = lib(precision)
    x, w = np.zeros(max(n, 1)), np.zeros(max(n, 1))
    _check(L, L.orc_gauss_rule(int(kind), int(n), float(alpha), float(a), float(b), capi.dptr(x), capi.dptr(w)))
    return x, w


def lens_quadrature(params, which, precision="f64"):
    """which 0: condenser defocus points (x, w); 1: condenser source points
    (x, y, w); 2: objective defocus points (x, w)."""
    L = lib(precision)
    s, keep = params.to_struct()
    nmax = 4096
    x, y, w = np.zeros(nmax), np.zeros(nmax), np.zeros(nmax)
    n = L.orc_lens_quadrature(C.byref(s), int(which), nmax, capi.dptr(x), capi.dptr(y), capi.dptr(w))
    if n < 0:
        raise RuntimeError(L.orc_last_error().decode())
    return (x[:n], y[:n], w[:n]) if which == 1 else (x[:n], w[:n])


def subslice_rule(precision="f64"):
    """The 128-node rule on [-1, 1] the sub-slice integrals use."""
    L = lib(precision)
    x, w = np.zeros(128), np.zeros(128)
    L.orc_subslice_rule(capi.dptr(x), capi.dptr(w))
    return x, w


def subslice_table(Z, dR_min, z_0, z_e, za, precision="f64"):
    """c0..c3 (128 each) of the sub-slice spline of element Z over [z_0, z_e)
    for an atom at height za, on the knots of atom_table(Z, dR_min)."""
    L = lib(precision)
    arrs = [np.zeros(128) for _ in range(4)]
    _check(L, L.orc_subslice_table(int(Z), float(dR_min), float(z_0), float(z_e), float(za),
                                   *[capi.dptr(a) for a in arrs]))
    return arrs
