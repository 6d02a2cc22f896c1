"""ctypes mirror of include/multem_hip.h.

The structures here must match the C declarations field for field; the
library checks MT_ABI_VERSION at load time.
"""
import ctypes as C

import numpy as np

MT_ABI_VERSION = 4

MT_OK = 0
STATUS_NAMES = {
    0: "MT_OK", 1: "MT_ERR_INVALID_ARGUMENT", 2: "MT_ERR_UNSUPPORTED",
    3: "MT_ERR_OUT_OF_MEMORY", 4: "MT_ERR_DEVICE", 5: "MT_ERR_STOPPED", 6: "MT_ERR_NO_DEVICE",
}

# output kinds (eTEM_Output_Type)
OT_IMAGE_TOT_COH, OT_IMAGE_TOT, OT_M2PSI_TOT_COH, OT_M2PSI_TOT = 1, 2, 3, 4
OT_M2PSI_TOT_PSI_COH, OT_PSI_COH, OT_PSI_0, OT_V, OT_TRANS = 5, 6, 7, 8, 9

_dp = C.POINTER(C.c_double)


class SystemConf(C.Structure):
    _fields_ = [("precision", C.c_int), ("device", C.c_int),
                ("cpu_nthread", C.c_int), ("gpu_device", C.c_int)]


LENS_FIELDS = [
    "m", "c_10", "c_12", "phi_12", "c_21", "phi_21", "c_23", "phi_23",
    "c_30", "c_32", "phi_32", "c_34", "phi_34",
    "c_41", "phi_41", "c_43", "phi_43", "c_45", "phi_45",
    "c_50", "c_52", "phi_52", "c_54", "phi_54", "c_56", "phi_56",
    "inner_aper_ang", "outer_aper_ang", "ti_a", "ti_sigma", "ti_beta",
]


class Lens(C.Structure):
    _fields_ = ([(f, C.c_double) for f in LENS_FIELDS]
                + [("ti_npts", C.c_int), ("si_a", C.c_double), ("si_sigma", C.c_double),
                   ("si_beta", C.c_double), ("si_rad_npts", C.c_int), ("si_azm_npts", C.c_int),
                   ("zero_defocus_type", C.c_int), ("zero_defocus_plane", C.c_double)])


class DetectorCir(C.Structure):
    _fields_ = [("inner_ang", C.c_double), ("outer_ang", C.c_double)]


class InputMultislice(C.Structure):
    _fields_ = [
        ("simulation_type", C.c_int), ("interaction_model", C.c_int), ("potential_type", C.c_int),
        ("operation_mode", C.c_int), ("reverse_multislice", C.c_int),
        ("pn_model", C.c_int), ("pn_coh_contrib", C.c_int), ("pn_single_conf", C.c_int),
        ("pn_nconf", C.c_int), ("pn_dim", C.c_int), ("pn_seed", C.c_int),
        ("spec_atoms", _dp), ("spec_atoms_rows", C.c_int), ("spec_atoms_cols", C.c_int),
        ("spec_lx", C.c_double), ("spec_ly", C.c_double), ("spec_lz", C.c_double), ("spec_dz", C.c_double),
        ("spec_rot_theta", C.c_double),
        ("thick_type", C.c_int), ("thick", _dp), ("n_thick", C.c_int),
        ("potential_slicing", C.c_int),
        ("nx", C.c_int), ("ny", C.c_int), ("bwl", C.c_int),
        ("iw_type", C.c_int), ("iw_psi", _dp), ("iw_x", _dp), ("iw_y", _dp), ("n_iw_xy", C.c_int),
        ("E_0", C.c_double), ("theta", C.c_double), ("phi", C.c_double),
        ("illumination_model", C.c_int), ("temporal_spatial_incoh", C.c_int),
        ("cond_lens", Lens), ("obj_lens", Lens),
        ("scanning_type", C.c_int), ("scanning_periodic", C.c_int), ("scanning_square_pxs", C.c_int),
        ("scanning_ns", C.c_int),
        ("scanning_x0", C.c_double), ("scanning_y0", C.c_double),
        ("scanning_xe", C.c_double), ("scanning_ye", C.c_double),
        ("detector_type", C.c_int), ("n_detector", C.c_int), ("detector_cir", C.POINTER(DetectorCir)),
        ("output_area_ix_0", C.c_int), ("output_area_iy_0", C.c_int),
        ("output_area_ix_e", C.c_int), ("output_area_iy_e", C.c_int),
        ("spec_rot_u0", C.c_double * 3), ("spec_rot_center_type", C.c_int), ("spec_rot_center_p", C.c_double * 3),
        ("detector_fR", _dp),
        ("ped_theta", C.c_double), ("ped_nrot", C.c_int), ("hci_theta", C.c_double), ("hci_nrot", C.c_int),
    ]


class OutputMultislice(C.Structure):
    _fields_ = [
        ("output_type", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("n_thk", C.c_int), ("n_det", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double),
        ("x", _dp), ("y", _dp), ("thick", _dp),
        ("image_tot", _dp), ("image_coh", _dp), ("m2psi_tot", _dp), ("m2psi_coh", _dp), ("psi_coh", _dp),
    ]


class Slice(C.Structure):
    _fields_ = [("z_0", C.c_double), ("z_e", C.c_double), ("z_int_0", C.c_double),
                ("z_int_e", C.c_double), ("iatom_0", C.c_int), ("iatom_e", C.c_int), ("ithk", C.c_int)]


def dptr(a):
    """Pointer to a C-contiguous float64 numpy array (or None)."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


class Output:
    """Host-side result with the fields of MATLAB's output_multislice.

    image_tot / image_coh: (n_thk, n_det, nx, ny) -- pixel (ix, iy) of the scan;
    m2psi_tot / m2psi_coh: (n_thk, nx, ny); psi_coh: complex (n_thk, nx, ny).
    Arrays are indexed [.., ix, iy]; transpose the last two axes for MATLAB's ny x nx view.
    """

    def __init__(self, dims):
        self.output_type = dims.output_type
        self.nx, self.ny, self.n_thk, self.n_det = dims.nx, dims.ny, dims.n_thk, dims.n_det
        self.dx, self.dy = dims.dx, dims.dy
        n = self.nx * self.ny
        self.x = np.zeros(max(self.nx, 1))
        self.y = np.zeros(max(self.ny, 1))
        self.thick = np.zeros(max(self.n_thk, 1))
        ot = self.output_type
        self._image_tot = np.zeros(self.n_thk * max(self.n_det, 1) * n) if ot in (OT_IMAGE_TOT, OT_IMAGE_TOT_COH) else None
        self._image_coh = np.zeros(self.n_thk * max(self.n_det, 1) * n) if ot == OT_IMAGE_TOT_COH else None
        self._m2psi_tot = np.zeros(self.n_thk * n) if ot in (OT_M2PSI_TOT, OT_M2PSI_TOT_COH, OT_M2PSI_TOT_PSI_COH) else None
        self._m2psi_coh = np.zeros(self.n_thk * n) if ot == OT_M2PSI_TOT_COH else None
        self._psi_coh = np.zeros(2 * self.n_thk * n) if ot in (OT_PSI_COH, OT_M2PSI_TOT_PSI_COH) else None

    def as_struct(self):
        o = OutputMultislice()
        o.output_type, o.nx, o.ny, o.n_thk, o.n_det = self.output_type, self.nx, self.ny, self.n_thk, self.n_det
        o.dx, o.dy = self.dx, self.dy
        o.x, o.y, o.thick = dptr(self.x), dptr(self.y), dptr(self.thick)
        o.image_tot, o.image_coh = dptr(self._image_tot), dptr(self._image_coh)
        o.m2psi_tot, o.m2psi_coh, o.psi_coh = dptr(self._m2psi_tot), dptr(self._m2psi_coh), dptr(self._psi_coh)
        return o

    @property
    def image_tot(self):
        return None if self._image_tot is None else self._image_tot.reshape(self.n_thk, self.n_det, self.nx, self.ny)

    @property
    def image_coh(self):
        return None if self._image_coh is None else self._image_coh.reshape(self.n_thk, self.n_det, self.nx, self.ny)

    @property
    def m2psi_tot(self):
        return None if self._m2psi_tot is None else self._m2psi_tot.reshape(self.n_thk, self.nx, self.ny)

    @property
    def m2psi_coh(self):
        return None if self._m2psi_coh is None else self._m2psi_coh.reshape(self.n_thk, self.nx, self.ny)

    @property
    def psi_coh(self):
        if self._psi_coh is None:
            return None
        return self._psi_coh.view(np.complex128).reshape(self.n_thk, self.nx, self.ny)
