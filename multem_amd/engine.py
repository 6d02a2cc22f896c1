"""ctypes binding of libmultem_hip.so (the HIP engine).

The library is built in-tree by `python -m multem_amd.build`; importing this
module never falls back to any other implementation -- a missing library is
an error.
"""
import ctypes as C
import os

import numpy as np

from . import capi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libmultem_hip.so")
_lib = None


class EngineError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s: %s" % (capi.STATUS_NAMES.get(status, status), msg))
        self.status = status


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError("HIP engine not built: %s (run python -m multem_amd.build)" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        sysp, inp, outp = P(capi.SystemConf), P(capi.InputMultislice), P(capi.OutputMultislice)
        dp = P(C.c_double)
        L.mt_hip_abi_version.restype = C.c_int
        L.mt_hip_last_error.restype = C.c_char_p
        L.mt_hip_output_dims.argtypes = [sysp, inp, outp]
        L.mt_hip_run_multislice.argtypes = [sysp, inp, outp]
        L.mt_hip_run_multislice_f32.argtypes = [sysp, inp, outp]
        L.mt_hip_run_multislice_f64.argtypes = [sysp, inp, outp]
        L.mt_hip_run_multislice_group.argtypes = [sysp, P(inp), C.c_int, outp]
        L.mt_hip_spec_slicing.argtypes = [sysp, inp, C.c_int, P(capi.Slice), C.c_int, P(C.c_int)]
        L.mt_hip_projected_potential.argtypes = [sysp, inp, C.c_int, C.c_int, dp]
        L.mt_hip_transmission_function.argtypes = [sysp, inp, C.c_int, C.c_int, dp]
        L.mt_hip_projected_potential_sum.argtypes = [sysp, inp, C.c_int, dp]
        L.mt_hip_incident_wave.argtypes = [sysp, inp, dp]
        L.mt_hip_propagate.argtypes = [sysp, inp, dp, C.c_double, C.c_int, dp]
        L.mt_hip_propagate_noalloc.argtypes = [sysp, inp, dp, C.c_double, C.c_int, dp]
        L.mt_hip_request_stop.argtypes = []
        L.mt_hip_progress.argtypes = [P(C.c_int), P(C.c_int)]
        L.mt_hip_set_memory_budget.argtypes = [C.c_size_t]
        L.mt_hip_last_plan.argtypes = [C.c_char_p, C.c_size_t]
        L.mt_hip_last_timing.argtypes = [dp, dp, dp]
        L.mt_hip_set_batch.argtypes = [C.c_int, C.c_int]
        L.mt_hip_set_wave_pad.argtypes = [C.c_int]
        L.mt_hip_set_batch_chains.argtypes = [C.c_int]
        L.mt_hip_kernel_timing.argtypes = [dp, P(C.c_int), dp, P(C.c_int), dp]
        L.mt_hip_synchronize.argtypes = [C.c_int]
        L.mt_hip_kept_device_bytes.argtypes = [C.c_int, C.POINTER(C.c_size_t)]
        L.mt_hip_release_device_memory.argtypes = [C.c_int]
        L.mt_hip_atom_table.argtypes = [C.c_int, C.c_double] + [dp] * 6
        L.mt_hip_set_autotune.argtypes = [C.c_int]
        L.mt_hip_stream_copy_bandwidth.argtypes = [C.c_int, C.c_size_t, C.c_int, P(C.c_double)]
        L.mt_hip_set_pass_variants.argtypes = [C.c_int, C.c_int, C.c_int]
        L.mt_hip_tune_log.argtypes = [C.c_char_p, C.c_size_t]
        if L.mt_hip_abi_version() != capi.MT_ABI_VERSION:
            raise RuntimeError("libmultem_hip ABI mismatch")
        _lib = L
    return _lib


def _check(st):
    if st != 0:
        raise EngineError(st, lib().mt_hip_last_error().decode())


def output_dims(params):
    s, keep = params.to_struct()
    sysc = params.system_conf.as_struct()
    d = capi.OutputMultislice()
    _check(lib().mt_hip_output_dims(C.byref(sysc), C.byref(s), C.byref(d)))
    return d


def run_multislice(params, precision=None):
    """ilc_multem on the GPU; returns a capi.Output.  The wave path runs in
    params.system_conf.precision (1 float, 2 double) unless `precision`
    forces one through the _f32 / _f64 entry."""
    L = lib()
    s, keep = params.to_struct()
    sysc = params.system_conf.as_struct()
    if precision is not None:
        sysc.precision = int(precision)     # output sizes of the forced precision's model
    d = capi.OutputMultislice()
    _check(L.mt_hip_output_dims(C.byref(sysc), C.byref(s), C.byref(d)))
    out = capi.Output(d)
    o = out.as_struct()
    entry = {None: L.mt_hip_run_multislice, 1: L.mt_hip_run_multislice_f32, 2: L.mt_hip_run_multislice_f64}[precision]
    _check(entry(C.byref(sysc), C.byref(s), C.byref(o)))
    return out


def run_multislice_group(params_list):
    """Several STEM inputs (grid sizes may differ) in one run; returns one
    capi.Output per input, as run_multislice would for each alone."""
    L = lib()
    n = len(params_list)
    structs, keeps, outs = [], [], []
    sysc = params_list[0].system_conf.as_struct()
    for p in params_list:
        st, keep = p.to_struct()
        structs.append(st)
        keeps.append(keep)
        d = capi.OutputMultislice()
        _check(L.mt_hip_output_dims(C.byref(sysc), C.byref(st), C.byref(d)))
        outs.append(capi.Output(d))
    ins = (C.POINTER(capi.InputMultislice) * n)(*[C.pointer(st) for st in structs])
    os_ = (capi.OutputMultislice * n)(*[o.as_struct() for o in outs])
    _check(L.mt_hip_run_multislice_group(C.byref(sysc), ins, n, os_))
    return outs


def spec_slicing(params, iconf=0):
    L = lib()
    s, keep = params.to_struct()
    sysc = params.system_conf.as_struct()
    n = C.c_int(0)
    _check(L.mt_hip_spec_slicing(C.byref(sysc), C.byref(s), iconf, None, 0, C.byref(n)))
    arr = (capi.Slice * max(n.value, 1))()
    _check(L.mt_hip_spec_slicing(C.byref(sysc), C.byref(s), iconf, arr, n.value, C.byref(n)))
    return [(a.z_0, a.z_e, a.z_int_0, a.z_int_e, a.iatom_0, a.iatom_e, a.ithk) for a in arr[:n.value]]


def projected_potential(params, islice, iconf=1):
    s, keep = params.to_struct()
    sysc = params.system_conf.as_struct()
    V = np.zeros(params.nx * params.ny)
    _check(lib().mt_hip_projected_potential(C.byref(sysc), C.byref(s), iconf, islice, capi.dptr(V)))
    return V.reshape(params.nx, params.ny)


def projected_potential_sum(params, iconf=1):
    """Sum over all slices of the projected potential, [nx, ny] (natural layout)."""
    s, keep = params.to_struct()
    sysc = params.system_conf.as_struct()
    V = np.zeros(params.nx * params.ny)
    _check(lib().mt_hip_projected_potential_sum(C.byref(sysc), C.byref(s), iconf, capi.dptr(V)))
    return V.reshape(params.nx, params.ny)


def transmission_function(params, islice, iconf=1):
    s, keep = params.to_struct()
    sysc = params.system_conf.as_struct()
    T = np.zeros(2 * params.nx * params.ny)
    _check(lib().mt_hip_transmission_function(C.byref(sysc), C.byref(s), iconf, islice, capi.dptr(T)))
    return T.view(np.complex128).reshape(params.nx, params.ny)


def incident_wave(params):
    s, keep = params.to_struct()
    sysc = params.system_conf.as_struct()
    P = np.zeros(2 * params.nx * params.ny)
    _check(lib().mt_hip_incident_wave(C.byref(sysc), C.byref(s), capi.dptr(P)))
    return P.view(np.complex128).reshape(params.nx, params.ny)


def propagate(params, psi, z, space=1):
    s, keep = params.to_struct()
    sysc = params.system_conf.as_struct()
    pin = np.ascontiguousarray(psi, dtype=np.complex128).view(np.float64).ravel()
    pout = np.zeros_like(pin)
    _check(lib().mt_hip_propagate(C.byref(sysc), C.byref(s), capi.dptr(pin), float(z), int(space), capi.dptr(pout)))
    return pout.view(np.complex128).reshape(params.nx, params.ny)


def propagate_noalloc(params, psi, z, space=1):
    """propagate() through mt_hip_propagate_noalloc: no device allocation,
    nx * ny at most NOALLOC_MAX_NXY."""
    s, keep = params.to_struct()
    sysc = params.system_conf.as_struct()
    pin = np.ascontiguousarray(psi, dtype=np.complex128).view(np.float64).ravel()
    pout = np.zeros_like(pin)
    _check(lib().mt_hip_propagate_noalloc(C.byref(sysc), C.byref(s), capi.dptr(pin), float(z), int(space),
                                          capi.dptr(pout)))
    return pout.view(np.complex128).reshape(params.nx, params.ny)


NOALLOC_MAX_NXY = 1024 * 1024


def synchronize(device=0):
    _check(lib().mt_hip_synchronize(int(device)))


def kept_device_bytes(device=0):
    """Bytes of device memory the library keeps on `device` for later runs."""
    b = C.c_size_t(0)
    _check(lib().mt_hip_kept_device_bytes(int(device), C.byref(b)))
    return b.value


def release_device_memory(device=0):
    """Returns the device blocks kept between runs to the HIP runtime."""
    _check(lib().mt_hip_release_device_memory(int(device)))


def atom_table(Z, dR_min):
    arrs = [np.zeros(128) for _ in range(5)]
    scal = np.zeros(4)
    _check(lib().mt_hip_atom_table(int(Z), float(dR_min), *[capi.dptr(a) for a in arrs], capi.dptr(scal)))
    R2, c0, c1, c2, c3 = arrs
    return dict(R2=R2, c0=c0, c1=c1, c2=c2, c3=c3, R_min=scal[0], R_max=scal[1], R_tap=scal[2], tap_cf=scal[3])


def set_batch(batch=0, use_graph=True):
    lib().mt_hip_set_batch(int(batch), int(bool(use_graph)))


def set_chains(chains=0):
    """Concurrent chains a STEM batch is split into (0: automatic, 1: one;
    at most 4).  Results do not depend on it."""
    lib().mt_hip_set_batch_chains(int(chains))


def set_wave_pad(pad=-1):
    """Row pitch of the STEM wave batches: ny + pad elements (pad < 0: the
    default, 0).  Results do not depend on it."""
    lib().mt_hip_set_wave_pad(int(pad))


PASS_VARIANTS = 11


def stream_copy_bandwidth(nbytes=1 << 30, reps=10, device=0):
    """Read + write GB/s of the library's streaming copy kernel."""
    g = C.c_double(0)
    _check(lib().mt_hip_stream_copy_bandwidth(int(device), int(nbytes), int(reps), C.byref(g)))
    return g.value


def set_autotune(mode=1):
    """0: fixed launch variants; 1: tune once per shape (default); 2: tune every run."""
    lib().mt_hip_set_autotune(int(mode))


def set_pass_variants(row=-1, col=-1, col_det=-1):
    """Force slice-step launch variants (-1: tuner's choice)."""
    lib().mt_hip_set_pass_variants(int(row), int(col), int(col_det))


def tune_log():
    buf = C.create_string_buffer(2048)
    _check(lib().mt_hip_tune_log(buf, 2048))
    return buf.value.decode()


def set_memory_budget(nbytes):
    lib().mt_hip_set_memory_budget(int(nbytes))


def request_stop():
    lib().mt_hip_request_stop()


def progress():
    a, b = C.c_int(0), C.c_int(0)
    lib().mt_hip_progress(C.byref(a), C.byref(b))
    return a.value, b.value


def last_plan():
    buf = C.create_string_buffer(512)
    _check(lib().mt_hip_last_plan(buf, 512))
    return buf.value.decode()


def last_timing():
    a, b, c = C.c_double(0), C.c_double(0), C.c_double(0)
    lib().mt_hip_last_timing(C.byref(a), C.byref(b), C.byref(c))
    r, nr, cc, nc, tf = C.c_double(0), C.c_int(0), C.c_double(0), C.c_int(0), C.c_double(0)
    lib().mt_hip_kernel_timing(C.byref(r), C.byref(nr), C.byref(cc), C.byref(nc), C.byref(tf))
    return dict(ms_slice=a.value, ms_total=b.value, probe_slices=c.value, ms_row=r.value, n_row=nr.value,
                ms_col=cc.value, n_col=nc.value, ms_tf=tf.value)
