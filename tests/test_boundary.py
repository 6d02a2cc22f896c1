"""The C-ABI library on the CPU: it loads, exports every symbol the header
declares, and its host-side model (input reading, slicing, output sizes,
error codes) agrees with the oracle.  No device work."""
import os
import re
import subprocess

import numpy as np
import pytest

from multem_amd import capi, configs, engine
from oracle import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_exports_every_declared_symbol(engine_lib):
    inc = os.path.join(ROOT, "include")
    declared = set()
    for h in sorted(os.listdir(inc)):
        if h.endswith(".h"):
            declared |= set(re.findall(r"\b(mt_hip_\w+)\s*\(", open(os.path.join(inc, h)).read()))
    assert len(declared) >= 30
    nm = subprocess.run(["nm", "-D", "--defined-only", engine.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (mt_hip_\w+)", nm))
    assert declared <= exported, declared - exported
    for name in declared:
        getattr(engine_lib, name)


def test_abi_struct_sizes():
    import ctypes as C
    # 31 doubles of lens parameters, then 4 ints and 4 more doubles
    assert C.sizeof(capi.Lens) == 8 * 31 + 8 + 8 * 3 + 8 + 8 + 8
    assert C.sizeof(capi.DetectorCir) == 16


@pytest.mark.parametrize("maker", [
    configs.si_ewrs_256,
    lambda: configs.au_stem_512(ns=4, nconf=2),
    lambda: configs.pt_cbed_1024(nconf=2, nc=30),
    lambda: configs.sto_stem(n=512, ncells=12, nc=25),
])
def test_slicing_matches_oracle(maker):
    p = maker()
    a = engine.spec_slicing(p)
    b = oracle.spec_slicing(p, precision="f32")
    assert len(a) == len(b)
    for x, y in zip(a, b):
        assert x[4:] == y[4:]
        assert abs(x[0] - y[0]) < 1e-6 and abs(x[1] - y[1]) < 1e-6


def test_fp32_slicing_quirk_srtio3():
    """SrTiO3 [001], 24x24x50 cells: atomic planes at z = 78.1 A and 156.2 A
    (40 and 80 times a/2) sit on edges of the plane finder's 0.1 A histogram
    (Identify_Planes, src/types.cuh:1166-1372).  The reference's FP32 path
    bins z in float and places the slice boundaries around those planes
    differently from its FP64 path; the engine's FP32 slicing is the FP32
    oracle's, and differs from the FP64 slicing at exactly slices 40-42 and
    80-82.  The 6x6x10-cell specimen slices alike in both precisions."""
    p = configs.sto_stem(n=1024, ncells=24, nc=50, ns=2)
    s_eng = engine.spec_slicing(p)
    s32, s64 = oracle.spec_slicing(p, precision="f32"), oracle.spec_slicing(p, precision="f64")
    assert len(s_eng) == len(s32) == len(s64) == 101
    for a, b in zip(s_eng, s32):
        assert a[4:] == b[4:] and abs(a[0] - b[0]) < 1e-4 and abs(a[1] - b[1]) < 1e-4
    differ = [k for k, (a, b) in enumerate(zip(s32, s64)) if any(abs(x - y) > 1e-4 for x, y in zip(a, b))]
    assert differ == [40, 41, 42, 80, 81, 82]
    # the FP32 slice 40 is 0.39 A thinner and slice 42 as much thicker
    assert abs((s32[40][1] - s32[40][0]) - (s64[40][1] - s64[40][0]) + 0.39) < 0.01
    q = configs.sto_stem(n=256, ncells=6, nc=10, ns=2)
    for a, b in zip(oracle.spec_slicing(q, precision="f32"), oracle.spec_slicing(q, precision="f64")):
        assert a[4:] == b[4:] and abs(a[0] - b[0]) < 1e-4


@pytest.mark.parametrize("maker", [
    configs.si_ewrs_256,
    lambda: configs.au_stem_512(ns=8, nconf=2),
    lambda: configs.pt_cbed_1024(nconf=2, nc=30),
])
def test_output_dims_match_oracle(maker):
    import ctypes as C
    p = maker()
    d = engine.output_dims(p)
    s, keep = p.to_struct()
    sysc = p.system_conf.as_struct()
    o = capi.OutputMultislice()
    assert oracle.lib("f32").orc_output_dims(C.byref(sysc), C.byref(s), C.byref(o)) == 0
    for f in ("output_type", "nx", "ny", "n_thk", "n_det"):
        assert getattr(d, f) == getattr(o, f), f
    assert abs(d.dx - o.dx) < 1e-6 and abs(d.dy - o.dy) < 1e-6


def test_scan_grid_square_region():
    p = configs.au_stem_512(ns=64, nconf=1)
    d = engine.output_dims(p)
    assert (d.nx, d.ny) == (64, 64)
    assert abs(d.dx - 4.078 / 64) < 1e-6


@pytest.mark.parametrize("field,value,status", [
    ("simulation_type", 61, 2),       # EELS: not provided
    ("simulation_type", 7, 2),        # not a simulation type
    ("potential_type", 4, 2),         # Kirkland tables: not provided
    ("nx", 300, 2),                   # 3 x 100: not 2^k, 3 x 2^k or 5 x 2^k
    ("nx", 112, 2),                   # 7 x 16: prime factor 7
    ("ny", 4096, 2),                  # beyond 2048
    ("nx", 0, 1),
    ("E_0", 0.0, 1),
    ("iw_type", 3, 1),                # user-defined wave without iw_psi
    ("interaction_model", 4, 1),
    ("potential_slicing", 4, 2),      # auto: no implementation in the reference
    ("theta", 95.0, 1),               # tilt beyond 90 degrees
])
def test_errors_are_reported(field, value, status):
    p = configs.si_ewrs_256()
    setattr(p, field, value)
    with pytest.raises(engine.EngineError) as e:
        engine.output_dims(p)
    assert e.value.status == status
    assert str(e.value)


def test_cpu_device_is_refused():
    p = configs.si_ewrs_256()
    p.system_conf.device = 1
    with pytest.raises(engine.EngineError) as e:
        engine.output_dims(p)
    assert e.value.status == 2


def test_empty_specimen_is_refused():
    p = configs.si_ewrs_256()
    p.spec_atoms = np.zeros((0, 8))
    with pytest.raises(engine.EngineError) as e:
        engine.output_dims(p)
    assert e.value.status == 1
