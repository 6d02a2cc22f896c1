"""File formats of MULTEM's MATLAB layer.

Inputs: the `multem_input.parameters` struct (parameters.m:6-239) with its
field names, as a MATLAB .mat file (a struct variable, e.g. saved with
`save('in.mat', 'input_multislice')`, and optionally a `system_conf`
struct) or as JSON with the same names (`{"system_conf": {...},
"input_multislice": {...}}`; arrays as nested lists, `spec_atoms` one row
per atom, `detector` as `{"type": 1, "cir": [{"inner_ang": ..,
"outer_ang": ..}, ...]}`).

Outputs: the struct `ilc_multem` returns (set_struct_multislice,
mex_files_multem/ilc_multem.cu:381-481) -- dx, dy, x, y, thick and
data(ithk) with image_tot(idet).image / image_coh (STEM), m2psi_tot /
psi_coh (EWFS, EWRS) or m2psi_tot / m2psi_coh (CBED, HRTEM), images
ny x nx -- written to and read from .mat files.

MATLAB files are read with scipy.io.loadmat (MAT v5 data only).
"""
import json

import numpy as np
import scipy.io as sio

from .parameters import Parameters, SystemConf

_SC_FIELDS = ("precision", "device", "cpu_nthread", "gpu_device")


def _scalar_or_array(v):
    a = np.asarray(v, dtype=np.float64)
    return float(a.ravel()[0]) if a.size == 1 and a.ndim <= 2 and not isinstance(v, (list, tuple)) else a


def _set_field(p, name, v):
    if name == "detector":
        _set_detector(p, v)
        return
    if isinstance(v, dict) and set(v) == {"re", "im"}:     # complex array in JSON
        v = np.asarray(v["re"], dtype=np.float64) + 1j * np.asarray(v["im"], dtype=np.float64)
    cur = getattr(p, name, None)
    if isinstance(cur, np.ndarray) or name in ("spec_atoms", "thick", "iw_x", "iw_y", "iw_psi"):
        a = np.asarray(v)
        if name == "spec_atoms":
            a = np.atleast_2d(np.asarray(a, dtype=np.float64))
        elif name == "iw_psi":
            a = np.asarray(a, dtype=np.complex128) if np.iscomplexobj(a) else np.asarray(a, dtype=np.float64)
        else:
            a = np.atleast_1d(np.asarray(a, dtype=np.float64)).ravel()
        setattr(p, name, a)
    elif isinstance(cur, int) and not isinstance(cur, bool):
        setattr(p, name, int(round(float(np.asarray(v).ravel()[0]))))
    else:
        setattr(p, name, float(np.asarray(v).ravel()[0]))


def _elements(v):
    if v is None:
        return []
    if isinstance(v, (list, tuple)):
        return list(v)
    return list(np.atleast_1d(np.asarray(v, dtype=object)).ravel())


def _set_detector(p, det):
    if isinstance(det, dict):
        typ, cir, mat = det.get("type", 1), det.get("cir", []), det.get("matrix", [])
    else:                                   # mat_struct
        typ, cir, mat = getattr(det, "type", 1), getattr(det, "cir", []), getattr(det, "matrix", [])
    p.detector_fR = [np.asarray(m["fR"] if isinstance(m, dict) else m.fR, dtype=np.float64)
                     for m in _elements(mat)]
    p.detector_type = int(np.asarray(typ).ravel()[0])
    if not isinstance(cir, (list, tuple, np.ndarray)):
        cir = [cir]
    pairs = []
    for c in np.atleast_1d(np.asarray(cir, dtype=object)).ravel():
        if isinstance(c, dict):
            a, b = c["inner_ang"], c["outer_ang"]
        else:
            a, b = c.inner_ang, c.outer_ang
        pairs.append((float(a), float(b)))
    p.detector_cir = pairs


def _fields_of(obj):
    if isinstance(obj, dict):
        return obj.items()
    return ((f, getattr(obj, f)) for f in obj._fieldnames)


def _params_from(inp, sc=None):
    p = Parameters()
    for k, v in _fields_of(inp):
        if k == "system_conf":
            sc = v
            continue
        if k.startswith("spec_cryst_") or k.startswith("eels_") or k.startswith("eftem_"):
            continue     # read by the reference for other simulation types / builders; no effect here
        _set_field(p, k, v)
    if sc is not None:
        p.system_conf = SystemConf(**{k: int(np.asarray(v).ravel()[0]) for k, v in _fields_of(sc) if k in _SC_FIELDS})
    return p


def load_input(path, name=None):
    """Parameters from a .mat (struct variable `name`, or the only struct
    with a simulation_type field) or a .json file."""
    if str(path).endswith(".json"):
        with open(path) as f:
            d = json.load(f)
        inp = d.get("input_multislice", d)
        return _params_from(inp, d.get("system_conf"))
    d = sio.loadmat(path, squeeze_me=True, struct_as_record=False, chars_as_strings=True)
    structs = {k: v for k, v in d.items() if not k.startswith("__") and hasattr(v, "_fieldnames")}
    if name is None:
        cands = [k for k, v in structs.items() if "simulation_type" in v._fieldnames]
        if len(cands) != 1:
            raise ValueError("%s: name the input struct (found %s)" % (path, sorted(structs)))
        name = cands[0]
    sc = next((v for v in structs.values() if "precision" in v._fieldnames), None)
    return _params_from(structs[name], sc)


def _input_dict(p):
    d = {}
    for k, v in p.__dict__.items():
        if k in ("system_conf", "detector_cir", "detector_type", "detector_fR"):
            continue
        if v is None:
            continue
        if isinstance(v, np.ndarray):
            if np.iscomplexobj(v):
                d[k] = {"re": v.real.tolist(), "im": v.imag.tolist()}
            else:
                d[k] = v.tolist()
        else:
            d[k] = v
    d["detector"] = {"type": int(p.detector_type),
                     "cir": [{"inner_ang": a, "outer_ang": b} for a, b in p.detector_cir],
                     "matrix": [{"fR": np.asarray(f, dtype=np.float64).tolist()} for f in p.detector_fR]}
    return d


def save_input(p, path):
    """Writes Parameters as .json or .mat (variables input_multislice and system_conf)."""
    if str(path).endswith(".json"):
        with open(path, "w") as f:
            json.dump({"system_conf": dict(p.system_conf.__dict__), "input_multislice": _input_dict(p)}, f, indent=1)
        return
    d = _input_dict(p)
    if "iw_psi" in d and isinstance(d["iw_psi"], dict):
        d["iw_psi"] = np.asarray(d["iw_psi"]["re"]) + 1j * np.asarray(d["iw_psi"]["im"])
    d["spec_atoms"] = np.atleast_2d(np.asarray(p.spec_atoms, dtype=np.float64))
    cir = _struct_array(len(p.detector_cir), ["inner_ang", "outer_ang"])
    for i, (a, b) in enumerate(p.detector_cir):
        cir[0, i]["inner_ang"], cir[0, i]["outer_ang"] = a, b
    mat = _struct_array(len(p.detector_fR), ["fR"])
    for i, f in enumerate(p.detector_fR):
        mat[0, i]["fR"] = np.asarray(f, dtype=np.float64)
    d["detector"] = {"type": int(p.detector_type), "cir": cir, "matrix": mat}
    sio.savemat(path, {"input_multislice": d, "system_conf": dict(p.system_conf.__dict__)}, oned_as="row")


# ---------------------------------------------------------------------------
def _struct_array(n, names):
    """1 x n MATLAB struct array (numpy record array) with the given fields."""
    return np.zeros((1, n), dtype=[(f, object) for f in names])


def output_struct(out):
    """A capi.Output as the struct of set_struct_multislice (images ny x nx)."""
    o = {"dx": out.dx, "dy": out.dy, "x": out.x[:out.nx], "y": out.y[:out.ny], "thick": out.thick[:out.n_thk]}
    if out.image_tot is not None:
        names = ["image_tot"] + (["image_coh"] if out.image_coh is not None else [])
    else:
        names = [f for f in ("m2psi_tot", "m2psi_coh", "psi_coh") if getattr(out, f) is not None]
    data = _struct_array(out.n_thk, names)
    for k in range(out.n_thk):
        for name in names:
            arr = getattr(out, name)
            if name.startswith("image"):
                dets = _struct_array(out.n_det, ["image"])
                for d in range(out.n_det):
                    dets[0, d]["image"] = arr[k, d].T
                data[0, k][name] = dets
            else:
                data[0, k][name] = arr[k].T
    o["data"] = data
    return o


def save_output(out, path, name="output_multislice"):
    """Writes a capi.Output as the MATLAB struct ilc_multem returns."""
    sio.savemat(path, {name: output_struct(out)}, oned_as="row")


def load_output(path, name=None):
    """Reads an ilc_multem output struct from a .mat file: dict with dx, dy,
    x, y, thick and data, a list over thicknesses of dicts whose image_tot /
    image_coh are lists over detectors of ny x nx images, or whose
    m2psi_tot / m2psi_coh / psi_coh are ny x nx arrays."""
    d = sio.loadmat(path, squeeze_me=True, struct_as_record=False)
    structs = {k: v for k, v in d.items() if not k.startswith("__") and hasattr(v, "_fieldnames")}
    if name is None:
        cands = [k for k, v in structs.items() if "data" in v._fieldnames]
        if len(cands) != 1:
            raise ValueError("%s: name the output struct (found %s)" % (path, sorted(structs)))
        name = cands[0]
    s = structs[name]
    res = {f: np.atleast_1d(np.asarray(getattr(s, f), dtype=np.float64)) for f in ("x", "y", "thick")}
    res["dx"], res["dy"] = float(s.dx), float(s.dy)
    data = []
    for e in np.atleast_1d(np.asarray(s.data, dtype=object)).ravel():
        t = {}
        for f in e._fieldnames:
            v = getattr(e, f)
            if f in ("image_tot", "image_coh"):
                t[f] = [np.asarray(x.image) for x in np.atleast_1d(np.asarray(v, dtype=object)).ravel()]
            else:
                t[f] = np.asarray(v)
        data.append(t)
    res["data"] = data
    return res
