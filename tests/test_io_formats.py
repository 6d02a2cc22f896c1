"""File formats of the MATLAB layer (multem_amd/io.py).

The output layout is pinned by the one output file the reference ships,
mex_examples_multem/stem_sim_Au_def_0.00_nconf_2.mat (a STEM run: 16 x 16
scan, 12 annular detectors, one thickness; its inputs are not stored, so
it pins the struct layout, not values): our reader reads it, and our
writer, given its images, writes a file with the same struct tree, field
names, shapes and values.  Inputs round-trip through JSON and .mat, and a
struct saved from plain MATLAB-named fields reads back to the same engine
input.
"""
import os

import numpy as np
import pytest
import scipy.io as sio

from multem_amd import capi, configs, engine
from multem_amd import io as mio

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "stem_sim_Au_def_0.00_nconf_2.mat")


def _tree(path, var):
    """(field path, shape) of every leaf of a MATLAB struct, in order."""
    s = sio.loadmat(path, squeeze_me=False, struct_as_record=True)[var]
    leaves = []

    def walk(a, pre):
        if a.dtype.names:
            for i in range(a.size):
                for f in a.dtype.names:
                    walk(a.flat[i][f], "%s(%d).%s" % (pre, i, f))
        else:
            leaves.append((pre, a.shape, a))
    walk(s, var)
    return leaves


def test_reads_the_reference_output_file():
    r = mio.load_output(GOLD)
    assert r["dx"] == pytest.approx(5.1690239906311035) and len(r["data"]) == 1
    imgs = r["data"][0]["image_tot"]
    assert len(imgs) == 12 and all(im.shape == (16, 16) for im in imgs)
    v = np.array(imgs)
    assert v.min() > 0 and v.max() < 0.1
    assert r["thick"][0] == pytest.approx(10.174078941345215)


def test_writer_reproduces_the_reference_struct(tmp_path):
    r = mio.load_output(GOLD)
    d = capi.OutputMultislice()
    d.output_type, d.nx, d.ny, d.n_thk, d.n_det = capi.OT_IMAGE_TOT, 16, 16, 1, 12
    d.dx = d.dy = r["dx"]
    out = capi.Output(d)
    # the reference file's x and y have 256 entries; our writer stores the scan axes
    out.x[:], out.y[:] = r["x"][:16], r["y"][:16]
    out.thick[:] = r["thick"]
    for k, im in enumerate(r["data"][0]["image_tot"]):
        out.image_tot[0, k] = im.T                   # MATLAB (iy, ix) -> [ix, iy]
    path = str(tmp_path / "o.mat")
    mio.save_output(out, path, name="output_multem")
    ref = [(p, s, a) for p, s, a in _tree(GOLD, "output_multem") if not p.endswith((".x", ".y"))]
    ours = [(p, s, a) for p, s, a in _tree(path, "output_multem") if not p.endswith((".x", ".y"))]
    assert [(p, s) for p, s, _ in ours] == [(p, s) for p, s, _ in ref]
    for (p, _, a), (_, _, b) in zip(ours, ref):
        assert np.array_equal(a, b), p


def test_image_orientation_is_ny_by_nx(tmp_path):
    d = capi.OutputMultislice()
    d.output_type, d.nx, d.ny, d.n_thk, d.n_det = capi.OT_M2PSI_TOT, 4, 6, 2, 0
    out = capi.Output(d)
    out.m2psi_tot[:] = np.arange(out.m2psi_tot.size).reshape(out.m2psi_tot.shape)
    path = str(tmp_path / "o.mat")
    mio.save_output(out, path)
    back = mio.load_output(path)
    for k in range(2):
        im = back["data"][k]["m2psi_tot"]
        assert im.shape == (6, 4)                     # ny x nx
        assert np.array_equal(im, out.m2psi_tot[k].T)


def _cases():
    p = configs.au_stem_512(ns=4, nconf=2, n=256, ncells=4, nc=3)
    p.detector_cir = [(20.0, 60.0), (80.0, 200.0)]
    q = configs.pt_cbed_1024(nconf=2, n=256, na=4, nb=6, nc=8, thick=[10.0, 20.0])
    q.illumination_model = 4
    h = configs.si_ewrs_256(nc=3)
    h.simulation_type = 32
    h.system_conf.precision = 2
    m = configs.au_stem_512(ns=4, nconf=1, n=64, ncells=2, nc=3)
    m.detector_type = 3
    m.detector_fR = [np.random.default_rng(i).random((m.ny, m.nx)) for i in range(2)]
    w = configs.si_ewrs_256(nc=2)
    w.iw_type = 3
    w.iw_psi = np.exp(1j * np.linspace(0, 1, w.nx * w.ny)).reshape(w.ny, w.nx)
    return [p, q, h, w, m]


def _same(a, b):
    assert a.system_conf.__dict__ == b.system_conf.__dict__
    assert sorted(a.__dict__) == sorted(b.__dict__)
    for k, v in a.__dict__.items():
        u = b.__dict__[k]
        if k == "system_conf":
            continue
        if v is None or u is None:
            assert v is None and u is None, k
        elif isinstance(v, np.ndarray) or isinstance(u, np.ndarray):
            assert np.array_equal(np.asarray(v), np.asarray(u)), k
        elif k == "detector_fR":
            assert len(v) == len(u) and all(np.array_equal(x, y) for x, y in zip(v, u)), k
        elif k == "detector_cir":
            assert [tuple(map(float, x)) for x in v] == [tuple(map(float, x)) for x in u], k
        else:
            assert float(v) == float(u), (k, v, u)
            assert isinstance(u, int) == isinstance(a.__class__().__dict__.get(k, u), int), k


@pytest.mark.parametrize("ext", [".json", ".mat"])
def test_input_round_trip(tmp_path, ext):
    for i, p in enumerate(_cases()):
        path = str(tmp_path / ("in%d%s" % (i, ext)))
        mio.save_input(p, path)
        _same(p, mio.load_input(path))


def test_matlab_named_struct_reads_to_the_same_input(tmp_path):
    """A struct saved from MATLAB field names (as `save('in.mat',
    'input_multislice')` would hold it) gives the engine the same input:
    same output sizes from the C ABI (no device work)."""
    p = configs.au_stem_512(ns=4, nconf=2, n=256, ncells=4, nc=3)
    fields = {k: v for k, v in p.__dict__.items() if k not in ("system_conf", "detector_cir", "detector_type",
                                                               "detector_fR") and v is not None}
    cir = np.empty((1, 1), dtype=object)
    cir[0, 0] = {"inner_ang": p.detector_cir[0][0], "outer_ang": p.detector_cir[0][1]}
    fields["detector"] = {"type": 1, "cir": cir}
    fields["spec_cryst_na"] = 4.0                    # builder-only fields are ignored
    path = str(tmp_path / "in.mat")
    sio.savemat(path, {"input_multislice": fields})
    q = mio.load_input(path)
    a, b = engine.output_dims(p), engine.output_dims(q)
    for f in ("output_type", "nx", "ny", "n_thk", "n_det", "dx", "dy"):
        assert getattr(a, f) == getattr(b, f), f
