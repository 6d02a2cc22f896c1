"""Input parameters with the field names, defaults and units of MULTEM's
MATLAB input class `multem_input.parameters` (matlab_functions/+multem_input/
parameters.m:6-239), which `ilc_multem` reads (ilc_multem.cu:37-379).

    p = Parameters()
    p.simulation_type = 11
    p.spec_atoms, p.spec_lx, p.spec_ly, p.spec_lz, a, b, c, p.spec_dz = fcc001_xtl(79, 4.078, 8, 8, 20, 2, 0.085)
    ...
    out = p.ilc_multem()            # HIP engine on the GPU
"""
import ctypes as C

import numpy as np

from . import capi


class SystemConf:
    """system_conf (matlab_functions/+multem_input/system_conf.m)."""

    def __init__(self, precision=1, device=2, cpu_nthread=1, gpu_device=0):
        self.precision = precision      # 1: float, 2: double
        self.device = device            # 1: CPU, 2: GPU
        self.cpu_nthread = cpu_nthread
        self.gpu_device = gpu_device

    def as_struct(self):
        return capi.SystemConf(int(self.precision), int(self.device), int(self.cpu_nthread), int(self.gpu_device))


_LENS_DEFAULTS = dict(
    m=0.0, c_10=14.0312, c_12=0.0, phi_12=0.0, c_21=0.0, phi_21=0.0, c_23=0.0, phi_23=0.0,
    c_30=1e-3, c_32=0.0, phi_32=0.0, c_34=0.0, phi_34=0.0,
    c_41=0.0, phi_41=0.0, c_43=0.0, phi_43=0.0, c_45=0.0, phi_45=0.0,
    c_50=0.0, c_52=0.0, phi_52=0.0, c_54=0.0, phi_54=0.0, c_56=0.0, phi_56=0.0,
    inner_aper_ang=0.0, outer_aper_ang=21.0,
    si_a=1.0, si_sigma=0.0072, si_beta=0.0, si_rad_npts=4, si_azm_npts=4,
    ti_a=1.0, ti_sigma=32.0, ti_beta=0.0, ti_npts=10,
    zero_defocus_type=1, zero_defocus_plane=0.0,
)


class Parameters:
    """multem_input.parameters; attribute names are the MATLAB field names."""

    def __init__(self):
        self.system_conf = SystemConf()
        self.interaction_model = 1
        self.potential_type = 6
        self.operation_mode = 1
        self.memory_size = 0
        self.reverse_multislice = 0
        self.pn_model = 1
        self.pn_coh_contrib = 0
        self.pn_single_conf = 0
        self.pn_nconf = 1
        self.pn_dim = 110
        self.pn_seed = 300183
        self.spec_atoms = np.zeros((0, 8))
        self.spec_dz = 0.25
        self.spec_lx = 10.0
        self.spec_ly = 10.0
        self.spec_lz = 10.0
        self.spec_rot_theta = 0.0
        self.spec_rot_u0 = np.array([0.0, 0.0, 1.0])
        self.spec_rot_center_type = 1
        self.spec_rot_center_p = np.array([0.0, 0.0, 0.0])
        self.thick_type = 1
        self.thick = np.zeros(1)
        self.potential_slicing = 1
        self.nx = 256
        self.ny = 256
        self.bwl = 0
        self.simulation_type = 52
        self.iw_type = 4
        self.iw_psi = None
        self.iw_x = np.zeros(1)
        self.iw_y = np.zeros(1)
        self.E_0 = 300.0
        self.theta = 0.0
        self.phi = 0.0
        self.ped_nrot = 360                   # precession (simulation_type 41): azimuths, cone angle (deg)
        self.ped_theta = 3.0
        self.hci_nrot = 360                   # hollow-cone illumination (42)
        self.hci_theta = 3.0
        self.illumination_model = 3
        self.temporal_spatial_incoh = 1
        for k, v in _LENS_DEFAULTS.items():
            setattr(self, "cond_lens_" + k, v)
            setattr(self, "obj_lens_" + k, v)
        self.obj_lens_outer_aper_ang = 24.0
        self.detector_type = 1
        self.detector_cir = [(40.0, 160.0)]   # (inner_ang, outer_ang) in mrad
        self.detector_fR = []                 # detector_type 3: ny x nx sensitivity maps, centred (g = 0 at nx/2, ny/2)
        self.scanning_type = 2
        self.scanning_periodic = 0
        self.scanning_square_pxs = 1
        self.scanning_ns = 10
        self.scanning_x0 = 0.0
        self.scanning_y0 = 0.0
        self.scanning_xe = 4.078
        self.scanning_ye = 4.078
        self.output_area_ix_0 = 0
        self.output_area_iy_0 = 0
        self.output_area_ix_e = 0
        self.output_area_iy_e = 0

    # ------------------------------------------------------------------
    def copy(self):
        q = Parameters.__new__(Parameters)
        q.__dict__ = {k: (v.copy() if isinstance(v, np.ndarray) else (list(v) if isinstance(v, list) else v))
                      for k, v in self.__dict__.items()}
        q.system_conf = SystemConf(**self.system_conf.__dict__)
        return q

    def to_struct(self):
        """Returns (ctypes struct, keepalive list)."""
        keep = []

        def arr(a):
            a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
            keep.append(a)
            return a

        s = capi.InputMultislice()
        for f in ("simulation_type", "interaction_model", "potential_type", "operation_mode", "reverse_multislice",
                  "pn_model", "pn_coh_contrib", "pn_single_conf", "pn_nconf", "pn_dim", "pn_seed",
                  "thick_type", "potential_slicing", "nx", "ny", "bwl", "iw_type",
                  "illumination_model", "temporal_spatial_incoh",
                  "scanning_type", "scanning_periodic", "scanning_square_pxs", "scanning_ns", "detector_type",
                  "output_area_ix_0", "output_area_iy_0", "output_area_ix_e", "output_area_iy_e",
                  "spec_rot_center_type", "ped_nrot", "hci_nrot"):
            setattr(s, f, int(getattr(self, f)))
        for f in ("spec_rot_u0", "spec_rot_center_p"):
            v = np.asarray(getattr(self, f), dtype=np.float64).ravel()
            getattr(s, f)[:] = [float(v[i]) if i < v.size else 0.0 for i in range(3)]
        for f in ("spec_lx", "spec_ly", "spec_lz", "spec_dz", "spec_rot_theta", "E_0", "theta", "phi",
                  "ped_theta", "hci_theta",
                  "scanning_x0", "scanning_y0", "scanning_xe", "scanning_ye"):
            setattr(s, f, float(getattr(self, f)))
        atoms = np.asarray(self.spec_atoms, dtype=np.float64)
        if atoms.ndim != 2:
            raise ValueError("spec_atoms must be an n x (4..8) matrix")
        a_cm = arr(atoms.T.ravel())          # column-major, as MATLAB passes it
        s.spec_atoms = capi.dptr(a_cm)
        s.spec_atoms_rows, s.spec_atoms_cols = atoms.shape
        th = arr(np.atleast_1d(self.thick))
        s.thick, s.n_thick = capi.dptr(th), th.size
        ix = arr(np.atleast_1d(self.iw_x))
        iy = arr(np.atleast_1d(self.iw_y))
        n = min(ix.size, iy.size)
        s.iw_x, s.iw_y, s.n_iw_xy = capi.dptr(ix), capi.dptr(iy), n
        if self.iw_psi is not None:
            p = np.ascontiguousarray(np.asarray(self.iw_psi, dtype=np.complex128).T).view(np.float64).ravel()
            keep.append(p)
            s.iw_psi = capi.dptr(p)
        for lens_name in ("cond_lens", "obj_lens"):
            L = getattr(s, lens_name)
            for k in _LENS_DEFAULTS:
                v = getattr(self, lens_name + "_" + k)
                if k in ("ti_npts", "si_rad_npts", "si_azm_npts", "zero_defocus_type"):
                    setattr(L, k, int(v))
                else:
                    setattr(L, k, float(v))
        dets = (capi.DetectorCir * max(len(self.detector_cir), 1))()
        for i, (a, b) in enumerate(self.detector_cir):
            dets[i].inner_ang, dets[i].outer_ang = float(a), float(b)
        keep.append(dets)
        s.detector_cir = C.cast(dets, C.POINTER(capi.DetectorCir))
        s.n_detector = len(self.detector_cir)
        if int(self.detector_type) == 3:
            fr = arr(np.concatenate([np.asarray(f, dtype=np.float64).T.ravel() for f in self.detector_fR])
                     if self.detector_fR else np.zeros(1))
            s.detector_fR = capi.dptr(fr)
            s.n_detector = len(self.detector_fR)
        return s, keep

    # ------------------------------------------------------------------
    def ilc_multem(self):
        """Runs the simulation on the HIP engine (ilc_multem)."""
        from . import engine
        return engine.run_multislice(self)

    def ilc_projected_potential(self, islice=0, iconf=1):
        from . import engine
        return engine.projected_potential(self, islice, iconf)

    def ilc_transmission_function(self, islice=0, iconf=1):
        from . import engine
        return engine.transmission_function(self, islice, iconf)

    def ilc_incident_wave(self):
        from . import engine
        return engine.incident_wave(self)

    def ilc_propagate(self, psi, z, space=1):
        from . import engine
        return engine.propagate(self, psi, z, space)

    def ilc_spec_slicing(self, iconf=0):
        from . import engine
        return engine.spec_slicing(self, iconf)
