"""The benchmark workloads as ilc_multem inputs.

si_ewrs_256     Si [001] 4x4x10 cells, 256^2, 300 kV plane wave, exit wave (EWRS)
au_stem_512     Au [001] 8x8x20 cells, 512^2, STEM ADF 80-200 mrad, 64x64 scan, 16 phonon configs
sto_stem        SrTiO3 [001] STEM at 256^2 / 512^2 / 1024^2 (6/12/24 cells, 10/25/50 thick)
sto_group       the three sto_stem sizes as one uneven-size group
pt_cbed_1024    Pt [110] CBED, 1024^2, 200 kV, 10 mrad, -100 A defocus, 20 phonon configs
"""
import numpy as np

from .crystal import fcc001_xtl, fcc110_xtl, si001_xtl, srtio3001_xtl
from .parameters import Parameters


def _base():
    p = Parameters()
    p.system_conf.precision = 1
    p.system_conf.device = 2
    p.interaction_model = 1
    p.potential_type = 6
    p.potential_slicing = 1
    p.illumination_model = 1
    p.temporal_spatial_incoh = 1
    p.bwl = 0
    p.E_0 = 300.0
    p.theta = 0.0
    p.phi = 0.0
    return p


def si_ewrs_256(nc=10, n=256):
    """Si [001], a=5.431 A, 4x4 cells x nc thick, plane wave, exit wave in real space."""
    p = _base()
    p.simulation_type = 52
    p.pn_model = 1
    a = 5.431
    p.spec_atoms, p.spec_lx, p.spec_ly, p.spec_lz, _, _, _, p.spec_dz = si001_xtl(4, 4, nc, 4, 0.085, a=a)
    p.nx = p.ny = n
    p.iw_type = 1
    p.iw_x = np.array([0.5 * p.spec_lx])
    p.iw_y = np.array([0.5 * p.spec_ly])
    p.thick_type = 1
    p.obj_lens_zero_defocus_type = 3
    return p


def au_stem_512(ns=64, nconf=16, n=512, ncells=8, nc=20):
    """Au [001] STEM: 21.4 mrad probe, zero defocus, Cs=0, ADF 80-200 mrad, scan over one cell."""
    p = _base()
    p.simulation_type = 11
    a = 4.078
    p.spec_atoms, p.spec_lx, p.spec_ly, p.spec_lz, _, _, _, p.spec_dz = fcc001_xtl(79, a, ncells, ncells, nc, 2, 0.085)
    p.nx = p.ny = n
    p.pn_model = 3 if nconf > 0 else 1
    p.pn_nconf = max(nconf, 1)
    p.pn_dim = 110
    p.pn_seed = 300183
    p.pn_coh_contrib = 0
    p.pn_single_conf = 0
    p.iw_type = 2
    p.cond_lens_c_10 = 0.0
    p.cond_lens_c_30 = 0.0
    p.cond_lens_c_50 = 0.0
    p.cond_lens_inner_aper_ang = 0.0
    p.cond_lens_outer_aper_ang = 21.4
    p.cond_lens_zero_defocus_type = 1
    p.cond_lens_zero_defocus_plane = 0.0
    p.thick_type = 1
    p.scanning_type = 2
    p.scanning_periodic = 1
    p.scanning_square_pxs = 1
    p.scanning_ns = ns
    c0 = ncells // 2
    p.scanning_x0 = c0 * a
    p.scanning_y0 = c0 * a
    p.scanning_xe = (c0 + 1) * a
    p.scanning_ye = (c0 + 1) * a
    p.detector_type = 1
    p.detector_cir = [(80.0, 200.0)]
    return p


def sto_stem(n=256, ncells=6, nc=10, ns=16, nconf=1):
    """SrTiO3 [001] STEM, 25 mrad probe, 80-200 mrad ring, phonon rms 0.07 (Sr, Ti) / 0.09 (O)."""
    p = _base()
    p.simulation_type = 11
    p.spec_atoms, p.spec_lx, p.spec_ly, p.spec_lz, a, _, _, p.spec_dz = srtio3001_xtl(ncells, ncells, nc, 2, 0.07, rmsd_O=0.09)
    p.nx = p.ny = n
    p.pn_model = 3
    p.pn_nconf = nconf
    p.pn_dim = 110
    p.pn_seed = 300183
    p.iw_type = 2
    p.cond_lens_c_10 = 0.0
    p.cond_lens_c_30 = 0.0
    p.cond_lens_outer_aper_ang = 25.0
    p.cond_lens_zero_defocus_type = 1
    p.thick_type = 1
    p.scanning_type = 2
    p.scanning_periodic = 1
    p.scanning_ns = ns
    c0 = ncells // 2
    p.scanning_x0 = p.scanning_y0 = c0 * a
    p.scanning_xe = p.scanning_ye = (c0 + 1) * a
    p.detector_cir = [(80.0, 200.0)]
    return p


def sto_group(ns=16, nconf=4):
    """SrTiO3 [001] at 256^2 (6x6 cells, 10 thick), 512^2 (12x12, 25) and
    1024^2 (24x24, 50): 21, 51 and 101 slices, ns x ns probes each."""
    return [sto_stem(n=256, ncells=6, nc=10, ns=ns, nconf=nconf),
            sto_stem(n=512, ncells=12, nc=25, ns=ns, nconf=nconf),
            sto_stem(n=1024, ncells=24, nc=50, ns=ns, nconf=nconf)]


def pt_cbed_1024(nconf=20, n=1024, na=10, nb=14, nc=None, thick=None):
    """Pt [110] CBED at 200 kV, 10 mrad, -100 A defocus; thicknesses every ~50 A."""
    p = _base()
    p.simulation_type = 21
    a = 3.924
    if nc is None:
        nc = int(round(500.0 / (a / np.sqrt(2))))
    p.spec_atoms, p.spec_lx, p.spec_ly, p.spec_lz, _, _, c, p.spec_dz = fcc110_xtl(78, a, na, nb, nc, 2, 0.075)
    p.nx = p.ny = n
    p.E_0 = 200.0
    p.pn_model = 3
    p.pn_nconf = nconf
    p.pn_dim = 110
    p.pn_seed = 300183
    p.iw_type = 2
    p.iw_x = np.array([0.5 * p.spec_lx])
    p.iw_y = np.array([0.5 * p.spec_ly])
    p.cond_lens_c_10 = -100.0
    p.cond_lens_c_30 = 0.0
    p.cond_lens_outer_aper_ang = 10.0
    p.cond_lens_zero_defocus_type = 1
    p.thick_type = 2
    if thick is None:
        thick = np.arange(50.0, nc * c + 1e-3, 50.0)
    p.thick = np.asarray(thick, dtype=np.float64)
    return p
