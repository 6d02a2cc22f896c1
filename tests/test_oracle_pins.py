"""Pins of the CPU oracle against what fixes it independently of its own code:
scipy's exact Bessel functions, numpy's FFT, libstdc++'s random engines, the
reference's own parameter table, closed forms and physical invariants.
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest
from scipy import integrate, special

from multem_amd import configs
from oracle import oracle
from tests.conftest import rel_l2

REF = "/root/reference/src/atomic_data.hpp"


def test_bessel_against_scipy():
    L = oracle.lib("f64")
    x = np.concatenate([np.linspace(0.01, 2.0, 60), np.linspace(2.0, 40.0, 80)])
    for which, f, tol in [(0, special.i0, 2e-7), (1, special.i1, 2e-7), (2, special.k0, 2e-7), (3, special.k1, 3e-7)]:
        got = np.array([L.orc_bessel(which, float(v)) for v in x])
        ref = f(x)
        err = np.max(np.abs(got - ref) / np.abs(ref))
        assert err < tol, (which, err)


@pytest.mark.parametrize("shape", [(64, 64), (128, 32), (256, 512), (96, 80), (1536, 64), (160, 1280), (384, 640)])
def test_fft_against_numpy(shape):
    rng = np.random.default_rng(3)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    f = oracle.fft2(a, inverse=False)
    b = oracle.fft2(a, inverse=True)
    assert rel_l2(f, np.fft.fft2(a)) < 1e-13
    assert rel_l2(b, np.fft.ifft2(a) * a.size) < 1e-13
    assert rel_l2(oracle.fft2(a, precision="f32"), np.fft.fft2(a)) < 2e-6


def test_fft_known_answers():
    n = 64
    d = np.zeros((n, n), complex)
    d[0, 0] = 1
    assert np.allclose(oracle.fft2(d), np.ones((n, n)))
    ix, iy = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    plane = np.exp(2j * np.pi * (3 * ix + 5 * iy) / n)
    F = oracle.fft2(plane)
    expect = np.zeros((n, n), complex)
    expect[3, 5] = n * n
    assert np.allclose(F, expect, atol=1e-9)


@pytest.fixture(scope="module")
def rng_ref(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("rng")
    exe = str(d / "rng_ref")
    src = os.path.join(os.path.dirname(__file__), "golden", "rng_ref.cpp")
    subprocess.run(["g++", "-O2", "-std=c++17", src, "-o", exe], check=True)
    return exe


@pytest.mark.parametrize("seed", [300183, 1, 2147483647])
def test_rng_against_libstdcxx(rng_ref, seed):
    n = 2000
    out = subprocess.run([rng_ref, str(seed), str(n)], capture_output=True, text=True, check=True).stdout.split()
    mt = np.array([int(v) for v in out[:n]], dtype=np.uint64)
    ui = np.array([int(v) for v in out[n:2 * n]])
    nf = np.array([float(v) for v in out[2 * n:3 * n]])
    nd = np.array([float(v) for v in out[3 * n:]])
    import ctypes as C
    L32, L64 = oracle.lib("f32"), oracle.lib("f64")
    got_mt = (C.c_uint64 * n)()
    L64.orc_mt19937_64(seed, n, got_mt)
    assert np.array_equal(np.array(got_mt[:], dtype=np.uint64), mt)
    got_ui = (C.c_int * n)()
    L64.orc_uniform_int(seed, n, got_ui)
    assert np.array_equal(np.array(got_ui[:]), ui)
    g32 = np.zeros(n)
    L32.orc_normal(seed, n, g32.ctypes.data_as(C.POINTER(C.c_double)))
    assert np.array_equal(g32.astype(np.float32), nf.astype(np.float32))
    g64 = np.zeros(n)
    L64.orc_normal(seed, n, g64.ctypes.data_as(C.POINTER(C.c_double)))
    assert np.array_equal(g64, nd)


def test_wavelength_closed_form():
    L = oracle.lib("f64")
    # lambda = hc / sqrt(E (2 m c^2 + E)), E in keV
    for E, lam in [(300.0, 0.0196875), (200.0, 0.0250793)]:
        assert abs(L.orc_lambda(E) - lam) < 1e-7


@pytest.mark.skipif(not os.path.exists(REF), reason="reference sources not present")
def test_lobato_table_matches_reference_file():
    src = open(REF).read()
    body = src[src.index("void Load_feg_Lobato_neutral_0_12"):]
    body = body[:body.index("\n\t\t\t}")]
    ref = {}
    for m in re.finditer(r"data_table\[(\d+)\]\.feg\[0\]\.(cl|cnl)\[(\d)\] = ([0-9.eE+-]+);", body):
        ref[(int(m.group(1)), m.group(2), int(m.group(3)))] = float(m.group(4))
    inc = open(os.path.join(os.path.dirname(__file__), "..", "multem_amd", "csrc", "lobato_table.inc")).read()
    rows = re.findall(r"^\s*\{([-0-9.eE+, ]+)\},\s*$", inc[inc.index("mt_lobato_coef"):], re.M)
    assert len(rows) == 103
    for z, row in enumerate(rows):
        vals = [float(v) for v in row.split(",")]
        for j in range(5):
            assert vals[2 * j] == ref[(z, "cl", j)]
            assert vals[2 * j + 1] == ref[(z, "cnl", j)]
    A = {int(m.group(1)): int(m.group(2)) for m in re.finditer(r"data_table\[(\d+)\]\.Z = \d+;\s*data_table\[\d+\]\.A = (\d+);", src)}
    elem = re.findall(r"\{\s*(\d+),\s*(\d+),", inc[inc.index("mt_element_table"):inc.index("mt_lobato_coef")])
    assert all(int(a) == A[int(z) - 1] for z, a in elem)


@pytest.mark.parametrize("Z", [6, 8, 14, 22, 38, 78, 79])
def test_projected_potential_is_z_integral_of_3d_potential(Z):
    """VR(R) = integral over z of Vr(sqrt(R^2 + z^2)): pins the K0/K1 closed form
    and its coefficients against the 3D Lobato potential."""
    L = oracle.lib("f64")
    import ctypes as C
    R = np.array([0.05, 0.2, 0.5, 1.0, 2.0])
    VR = np.zeros_like(R)
    dp = C.POINTER(C.c_double)
    L.orc_lobato_VR(Z, len(R), R.ctypes.data_as(dp), VR.ctypes.data_as(dp), None, None)

    def vr3(r):
        rr = np.array([r])
        out = np.zeros(1)
        L.orc_lobato_VR(Z, 1, rr.ctypes.data_as(dp), None, None, out.ctypes.data_as(dp))
        return out[0]

    for Ri, VRi in zip(R, VR):
        num = 2 * integrate.quad(lambda z: vr3(np.hypot(Ri, z)), 0, np.inf, limit=400, epsabs=0, epsrel=1e-10)[0]
        assert abs(num - VRi) / abs(VRi) < 2e-6, (Z, Ri, num, VRi)


@pytest.mark.parametrize("Z", [8, 14, 79])
def test_spline_table_invariants(Z):
    t = oracle.atom_table(Z, 0.05)
    R2, c0, c1, c2, c3 = t["R2"], t["c0"], t["c1"], t["c2"], t["c3"]
    # Hermite: each cubic hits the next knot's value and slope (in R^2)
    h = R2[1:] - R2[:-1]
    v_end = ((c3[:-1] * h + c2[:-1]) * h + c1[:-1]) * h + c0[:-1]
    d_end = (3 * c3[:-1] * h + 2 * c2[:-1]) * h + c1[:-1]
    assert np.max(np.abs(v_end - c0[1:]) / np.max(np.abs(c0))) < 1e-12
    assert np.max(np.abs(d_end - c1[1:]) / np.max(np.abs(c1))) < 1e-9
    # the taper brings the last knot to zero; knots below R_tap are the exact VR
    assert abs(c0[-1]) < 1e-12 * abs(c0[0])
    R = np.sqrt(R2)
    below = R < np.sqrt(t["R2_tap"])
    VR = np.zeros_like(R)
    import ctypes as C
    dp = C.POINTER(C.c_double)
    Rc = np.ascontiguousarray(R)
    oracle.lib("f64").orc_lobato_VR(Z, len(R), Rc.ctypes.data_as(dp), VR.ctypes.data_as(dp), None, None)
    assert np.allclose(c0[below], VR[below], rtol=1e-13)
    # knots are log-spaced between R_min and R_max
    assert np.allclose(np.diff(np.log(R)), np.log(t["R_max"] / t["R_min"]) / 127, rtol=1e-9)


def test_single_atom_potential_is_projected_potential():
    """One atom at the box centre: V(r) follows VR(|r - r_a|) up to the spline
    interpolation error inside R_tap and vanishes beyond R_max."""
    p = configs.si_ewrs_256()
    a = 10.0
    p.spec_atoms = np.array([[79, a / 2, a / 2, 0.0, 0.085, 1, 0, 0], [79, a / 2, a / 2, 3.0, 0.085, 1, 0, 0]])
    p.spec_lx = p.spec_ly = a
    p.spec_lz = 3.0
    p.nx = p.ny = 128
    V = oracle.projected_potential(p, 0, precision="f64")
    d = a / 128
    ix = np.arange(128)
    X, Y = np.meshgrid(ix * d - a / 2, ix * d - a / 2, indexing="ij")
    R = np.hypot(X, Y)
    t = oracle.atom_table(79, d)
    inside = (R > t["R_min"]) & (R < np.sqrt(t["R2_tap"]))
    import ctypes as C
    dp = C.POINTER(C.c_double)
    Rc = np.ascontiguousarray(R[inside])
    VR = np.zeros_like(Rc)
    oracle.lib("f64").orc_lobato_VR(79, len(Rc), Rc.ctypes.data_as(dp), VR.ctypes.data_as(dp), None, None)
    assert np.max(np.abs(V[inside] - VR) / VR) < 2e-3
    assert np.all(V[R > t["R_max"] + 1e-9] == 0)


@pytest.mark.parametrize("maker,count", [
    (lambda: configs.si_ewrs_256(nc=10), 41),
    (lambda: configs.au_stem_512(ns=2, nconf=1, nc=20), 41),
    (lambda: configs.sto_stem(nc=10), 21),
])
def test_slice_counts(maker, count):
    p = maker()
    sl = oracle.spec_slicing(p)
    assert len(sl) == count
    # slices tile the specimen without gaps, each plane's atoms in one slice
    for a, b in zip(sl[:-1], sl[1:]):
        assert abs(a[1] - b[0]) < 1e-12
        assert a[5] + 1 == b[4]


def test_flux_conservation_and_fp32_path():
    p = configs.si_ewrs_256()
    r64 = oracle.run(p, "f64")
    r32 = oracle.run(p, "f32")
    psi = r64.psi_coh[0]
    assert abs(np.mean(np.abs(psi) ** 2) - 1.0) < 1e-12
    # the reference's FP32 path stays within a few 1e-6 of its FP64 path
    assert rel_l2(r32.psi_coh, r64.psi_coh) < 1e-5


def test_propagator_composition():
    p = configs.si_ewrs_256()
    rng = np.random.default_rng(5)
    psi = rng.standard_normal((256, 256)) + 1j * rng.standard_normal((256, 256))
    a = oracle.propagate(p, oracle.propagate(p, psi, 3.0), 4.5)
    b = oracle.propagate(p, psi, 7.5)
    assert rel_l2(a, b) < 1e-12
    # unitary without a bandwidth limit
    assert abs(np.linalg.norm(b) / np.linalg.norm(psi) - 1) < 1e-12


def test_probe_normalisation_and_position():
    p = configs.au_stem_512(ns=2, nconf=1, n=256, ncells=4, nc=2)
    p.simulation_type = 21
    p.iw_x = np.array([0.25 * p.spec_lx])
    p.iw_y = np.array([0.75 * p.spec_ly])
    psi = oracle.incident_wave(p)
    assert abs(np.sum(np.abs(psi) ** 2) / psi.size - 1) < 1e-12
    ix, iy = np.unravel_index(np.argmax(np.abs(psi)), psi.shape)
    assert abs(ix - 64) <= 1 and abs(iy - 192) <= 1


def test_phonon_displacements():
    p = configs.au_stem_512(ns=2, nconf=3, n=256, ncells=4, nc=10)
    u = oracle.displaced_atoms(p, 0)
    for iconf in (1, 2, 3):
        d = oracle.displaced_atoms(p, iconf)
        dx, dy, dz = (d[:, k] - u[:, k] for k in range(3))
        assert np.all(dz == 0)                   # pn_dim = 110
        assert abs(np.std(dx) - 0.085) < 0.01 and abs(np.std(dy) - 0.085) < 0.01
        assert abs(np.mean(dx)) < 0.01
    assert not np.array_equal(oracle.displaced_atoms(p, 1), oracle.displaced_atoms(p, 2))


def test_threaded_cbed_equals_sequential():
    # configurations run in parallel but are accumulated in configuration order
    p = configs.pt_cbed_1024(nconf=5, n=128, na=3, nb=4, nc=6, thick=[4.0, 8.0])
    a = oracle.run(p, "f64", nthreads=1).m2psi_tot.copy()
    b = oracle.run(p, "f64", nthreads=4).m2psi_tot.copy()
    assert np.array_equal(a, b)


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_tilted_vacuum_gives_one_spot(precision):
    # Through an empty specimen a tilted plane wave keeps unit modulus; the
    # output ramp exp(-2 pi i g0.r) then puts all intensity in one Fourier
    # pixel, -g0 / dg, where sin(theta) is quantised to n lambda dg.
    p = configs.si_ewrs_256(nc=2, n=128)
    p.spec_atoms = p.spec_atoms.copy()
    p.spec_atoms[:, 5] = 0.0                      # occupancy 0: no potential
    p.simulation_type = 51                        # EWFS
    p.obj_lens_zero_defocus_type = 1
    lam = 0.0196875                               # 300 kV
    dg = 1.0 / p.spec_lx
    n_tilt = 3
    p.theta = np.degrees(np.arcsin(n_tilt * lam * dg * 1.2))   # quantised to n = round(3.6) = 4
    p.phi = 0.0
    out = oracle.run(p, precision)
    I = np.abs(out.psi_coh[0]) ** 2              # [nx, ny], centred (fftshift) layout
    k = np.unravel_index(np.argmax(I), I.shape)
    assert I.sum() == pytest.approx(1.0, rel=1e-5)
    assert I[k] == pytest.approx(1.0, rel=1e-5)
    assert (k[0] - p.nx // 2, k[1] - p.ny // 2) == (-4, 0)


def test_stem_coherent_image_bounds():
    # one configuration: the coherent image is the total image; several:
    # |sum w psi|^2 <= sum w |psi|^2 (sum w = 1) at every pixel
    p = configs.au_stem_512(ns=3, nconf=1, n=128, ncells=2, nc=3)
    p.pn_coh_contrib = 1
    p.pn_model = 1
    o = oracle.run(p, "f64")
    assert np.allclose(o.image_coh, o.image_tot, rtol=1e-12, atol=0)
    p = configs.au_stem_512(ns=3, nconf=4, n=128, ncells=2, nc=3)
    p.pn_coh_contrib = 1
    o = oracle.run(p, "f64")
    assert np.all(o.image_coh <= o.image_tot * (1 + 1e-12))
    assert np.all(o.image_coh < o.image_tot)     # phonons remove coherent intensity from the ADF ring


def _hrtem_case(**kw):
    p = configs.si_ewrs_256(nc=3, n=128)
    p.simulation_type = 32                  # HRTEM
    p.obj_lens_zero_defocus_type = 3
    p.obj_lens_c_10, p.obj_lens_c_30 = 0.0, 0.0
    p.obj_lens_outer_aper_ang = 20.0
    p.illumination_model = 1
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def test_hrtem_coherent_is_aperture_filtered_exit_wave():
    # aberration-free objective: image = |IFFT(aperture * O(g))|^2 with O the
    # reciprocal-space exit wave (EWFS output, centred, scaled by 1/N^2)
    p = _hrtem_case()
    img = oracle.run(p, "f64").m2psi_tot[0]
    q = p.copy()
    q.simulation_type = 51
    O = oracle.run(q, "f64").psi_coh[0]                 # [nx, ny], centred k-space
    nx, ny = p.nx, p.ny
    lam = 0.0196875
    gx = (np.arange(nx) - nx // 2) / p.spec_lx
    gy = (np.arange(ny) - ny // 2) / p.spec_ly
    g2 = gx[:, None] ** 2 + gy[None, :] ** 2
    mask = g2 < (np.sin(20e-3) / lam) ** 2
    psi = np.fft.ifft2(np.fft.ifftshift(O * mask)) * (nx * ny)
    want = np.fft.fftshift(np.abs(psi) ** 2)
    assert rel_l2(img, want) < 1e-9


def test_hrtem_partial_coherence_without_spread_equals_coherent():
    kw = dict(obj_lens_c_10=-60.0, obj_lens_c_30=1.2, obj_lens_ti_sigma=0.0, cond_lens_si_sigma=0.0)
    a = oracle.run(_hrtem_case(illumination_model=1, **kw), "f64").m2psi_tot[0]
    b = oracle.run(_hrtem_case(illumination_model=2, **kw), "f64").m2psi_tot[0]
    assert rel_l2(a, b) < 1e-12
    # a finite focal spread damps the high frequencies: less image contrast
    c = oracle.run(_hrtem_case(illumination_model=2, **dict(kw, obj_lens_ti_sigma=40.0)), "f64").m2psi_tot[0]
    assert np.std(c) < np.std(b)


@pytest.mark.parametrize("illum", [1, 2, 4])
def test_hrtem_coherent_image_of_one_configuration(illum):
    """One configuration: the coherent image |IFFT(lens psi_coh)|^2 is the
    total image (set_m2psi_coh, src/wave_function.cuh:206-215)."""
    p = configs.si_ewrs_256(nc=3)
    p.simulation_type = 32
    p.pn_model = 1
    p.pn_coh_contrib = 1
    p.illumination_model = illum
    p.temporal_spatial_incoh = 2
    p.obj_lens_c_10, p.obj_lens_c_30, p.obj_lens_outer_aper_ang = -40.0, 0.8, 25.0
    p.obj_lens_ti_sigma, p.obj_lens_ti_npts = 20.0, 5
    r = oracle.run(p, "f64")
    assert r.m2psi_coh is not None
    assert np.array_equal(r.m2psi_coh, r.m2psi_tot)


# ---------------------------------------------------------------------------
# Closed forms from physical constants (CODATA, scipy.constants), independent
# of the oracle's own constants: wavelength, interaction constant, Fresnel
# propagator, aberration function, detector normalisation, partially coherent
# envelope.  Each check fails under a flipped sign or a wrong scale.

def _lambda_codata(E_kev):
    """Relativistic electron wavelength (A): h c / sqrt(E (2 m c^2 + E))."""
    from scipy import constants as K
    E = E_kev * 1e3 * K.e
    mc2 = K.m_e * K.c ** 2
    return K.h * K.c / np.sqrt(E * (2.0 * mc2 + E)) * 1e10


def _sigma_codata(E_kev):
    """Interaction constant (rad / (V A)): 2 pi / (lambda V) (m c^2 + e V) / (2 m c^2 + e V)."""
    from scipy import constants as K
    V = E_kev * 1e3
    mc2 = K.m_e * K.c ** 2 / K.e
    return 2.0 * np.pi / (_lambda_codata(E_kev) * V) * (mc2 + V) / (2.0 * mc2 + V)


def test_interaction_constant_closed_form():
    # Vr_factor = gamma lambda / c_Potf / cos(theta) (src/cgpu_fcns.cuh:137-144) is the
    # interaction constant sigma of the projected potential in V A
    L = oracle.lib("f64")
    for E, published in [(300.0, 6.5265e-4), (200.0, 7.2884e-4), (100.0, 9.2441e-4)]:
        s = _sigma_codata(E)
        assert abs(s / published - 1) < 1e-4            # the textbook values
        assert abs(L.orc_vr_factor(E, 0.0) / s - 1) < 2e-6
        th = np.radians(0.7)
        assert abs(L.orc_vr_factor(E, th) * np.cos(th) / s - 1) < 2e-6
        assert abs(L.orc_lambda(E) / _lambda_codata(E) - 1) < 2e-6


def _vacuum(p):
    p.spec_atoms = p.spec_atoms.copy()
    p.spec_atoms[:, 5] = 0.0                               # occupancy 0: no potential
    return p


@pytest.mark.parametrize("reverse", [0, 1])
@pytest.mark.parametrize("tilt", [None, (2, 0.0), (3, 90.0)])
def test_propagator_plane_wave_closed_form(reverse, tilt):
    """A plane wave of frequency g through the Fresnel propagator of
    distance z picks up exp(-i s pi lambda z |g - g0|^2), s = +1 (-1 for
    reverse multislice), with |psi(g)| kept (propagate,
    src/cgpu_fcns.cuh:1835-1845; w = -mul_sign pi lambda z,
    src/input_multislice.cuh:908-912); the reciprocal output carries the
    1/(nx ny) of the unnormalised FFT (src/propagator.cuh:69-86)."""
    p = configs.si_ewrs_256(nc=2, n=128)
    p.reverse_multislice = reverse
    nx, ny = p.nx, p.ny
    lam = _lambda_codata(p.E_0)
    g0 = np.zeros(2)
    if tilt:
        # sin(theta) on the tilt grid n lambda dg: g0 = n dg (cos phi, sin phi)
        n, phi = tilt
        p.theta = np.degrees(np.arcsin(n * lam / p.spec_lx))
        p.phi = phi
        g0 = n / p.spec_lx * np.array([np.cos(np.radians(phi)), np.sin(np.radians(phi))])
    ix, iy = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    sgn = -1.0 if reverse else 1.0
    for kx, ky in [(3, 5), (-7, 2), (40, -33), (0, 0)]:
        psi = np.exp(2j * np.pi * (kx * ix / nx + ky * iy / ny))
        g2 = (kx / p.spec_lx - g0[0]) ** 2 + (ky / p.spec_ly - g0[1]) ** 2
        for z in (2.0, -13.5, 120.0):
            out = oracle.propagate(p, psi, z)
            phase = -sgn * np.pi * lam * z * g2
            # the tolerance follows the phase: the reference's constants give
            # lambda to ~1e-7 of CODATA's
            assert np.allclose(np.abs(out), 1.0, rtol=0, atol=1e-12)
            d = np.angle(out / (np.exp(1j * phase) * psi))
            assert np.max(np.abs(d)) < 2e-6 * abs(phase) + 1e-12, (kx, ky, z)
        f = oracle.propagate(p, psi, 7.0, space=2)     # centred reciprocal layout
        k = ((kx + nx // 2) % nx, (ky + ny // 2) % ny)
        assert abs(abs(f[k]) - 1.0) < 1e-12
        f[k] = 0
        assert np.max(np.abs(f)) < 1e-12


def _centre_probe(n=128, **lens):
    p = configs.au_stem_512(ns=2, nconf=1, n=n, ncells=4, nc=2)
    p.simulation_type = 21
    p.iw_x = np.array([0.5 * p.spec_lx])
    p.iw_y = np.array([0.5 * p.spec_ly])
    p.cond_lens_outer_aper_ang = 30.0
    for k, v in lens.items():
        setattr(p, "cond_lens_" + k, v)
    return p


def _probe_spectrum(p):
    """fPsi(g) (natural k order) of the oracle's probe: the probe is
    returned in natural real space, the shifted layout puts the box centre
    at index 0, and the inverse FFT is unnormalised."""
    psi = oracle.incident_wave(p)
    return np.fft.fft2(np.fft.fftshift(psi)) / psi.size


def _k_grid(p):
    gx = np.fft.fftfreq(p.nx, d=p.spec_lx / p.nx)
    gy = np.fft.fftfreq(p.ny, d=p.spec_ly / p.ny)
    GX, GY = np.meshgrid(gx, gy, indexing="ij")
    return GX, GY


def test_probe_aberration_phases_closed_form():
    """Phase of the probe spectrum, Lens::set_input_data / eval_* and
    exp_i_chi (src/types.cuh:2924-2976,3180-3300; src/cgpu_fcns.cuh:1949-1979):
    chi = -pi lambda c_10 g^2 - pi Cs lambda^3 g^4 / 2 - pi lambda c_12 g^2
    sin(2 (phi_g - phi_12)), Cs in mm (x 1e7 A), angles in degrees.  Ratios
    of two probes leave only the changed term."""
    lam = _lambda_codata(300.0)
    a = _centre_probe(c_10=0.0, c_30=0.0)
    b = _centre_probe(c_10=85.0, c_30=0.0)
    c = _centre_probe(c_10=85.0, c_30=1.3)
    d = _centre_probe(c_10=85.0, c_30=1.3, c_12=40.0, phi_12=25.0)
    Fa, Fb, Fc, Fd = (_probe_spectrum(q) for q in (a, b, c, d))
    GX, GY = _k_grid(a)
    g2 = GX ** 2 + GY ** 2
    inside = np.abs(Fa) > 0.5 * np.abs(Fa).max()
    assert inside.sum() > 100
    # equal moduli: the aperture and the normalisation do not depend on chi
    assert np.allclose(np.abs(Fb)[inside], np.abs(Fa)[inside], rtol=1e-12)
    def check(num, den, phase):
        # the tolerance follows the phase: the reference's constants give
        # lambda to ~1e-7 of CODATA's (lambda^3 for Cs)
        d = np.angle(num[inside] / den[inside] * np.exp(-1j * phase))
        assert np.max(np.abs(d)) < 1e-6 * np.max(np.abs(phase)) + 1e-9

    g2i = g2[inside]
    check(Fb, Fa, -np.pi * lam * 85.0 * g2i)
    check(Fc, Fb, -np.pi * 1.3e7 * lam ** 3 * g2i ** 2 / 2)
    phig = np.arctan2(GY, GX)[inside]
    check(Fd, Fc, -np.pi * lam * 40.0 * g2i * np.sin(2 * (phig - np.radians(25.0))))
    # the aperture is sin(alpha)/lambda, alpha in mrad
    gmax = np.sin(30e-3) / lam
    assert np.all(g2[inside] < gmax ** 2)
    assert np.all(np.abs(Fa)[g2 > gmax ** 2 * 1.0001] < 1e-15)     # FFT round-off only


def test_stem_detector_normalisation_in_vacuum():
    """Through vacuum |psi(g)|^2 is the probe's normalised spectrum: a ring
    covering the whole grid and a ring just outside the aperture give 1, a
    ring outside the aperture gives 0 (integrated_intensity_over_det,
    src/wave_function.cuh:107-133; probe normalised to sum |fPsi|^2 = 1,
    src/gpu_fcns.cuh:1612-1625); with atoms the full ring still gives 1
    (unitary slices without a bandwidth limit)."""
    p = configs.au_stem_512(ns=3, nconf=1, n=128, ncells=2, nc=3)
    p.detector_cir = [(0.0, 1000.0), (0.0, 22.0), (25.0, 200.0)]
    q = _vacuum(p.copy())
    img = oracle.run(q, "f64").image_tot          # [n_thk][n_det][nscan]
    assert np.allclose(img[0, 0], 1.0, rtol=0, atol=1e-12)
    assert np.allclose(img[0, 1], 1.0, rtol=0, atol=1e-12)
    assert np.all(np.abs(img[0, 2]) < 1e-20)
    img = oracle.run(p, "f64").image_tot
    assert np.allclose(img[0, 0], 1.0, rtol=0, atol=1e-12)
    assert np.all(img[0, 2] > 1e-4)                  # the specimen scatters into the ADF ring


def _two_beam(illum, **lens):
    """HRTEM of the user wave 1 + eps exp(2 pi i g.r) through vacuum: the
    image's Fourier coefficient at g is eps E(g) exp(i chi(g)) times the
    vacuum propagation phase (the 0 beam is unchanged)."""
    p = _vacuum(configs.si_ewrs_256(nc=2, n=128))
    p.simulation_type = 32
    p.iw_type = 3
    p.iw_x, p.iw_y = np.array([0.5 * p.spec_lx]), np.array([0.5 * p.spec_ly])
    p.illumination_model = illum
    p.temporal_spatial_incoh = 1
    p.obj_lens_outer_aper_ang = 60.0
    for k, v in lens.items():
        setattr(p, k, v)
    ix, iy = np.meshgrid(np.arange(p.nx), np.arange(p.ny), indexing="ij")
    out = {}
    for kx, ky in [(5, 0), (9, 4), (-6, 11), (17, -13)]:
        psi = 1 + 0.05 * np.exp(2j * np.pi * (kx * ix / p.nx + ky * iy / p.ny))
        p.iw_psi = psi.T.copy()
        img = oracle.run(p, "f64").m2psi_tot[0]
        F = np.fft.fft2(img) / img.size
        out[(kx, ky)] = F[kx % p.nx, ky % p.ny]
    return p, out


def test_objective_lens_ctf_and_pctf_envelope_closed_form():
    """Coherent CTF exp(i chi) and the partially coherent envelope of
    apply_PCTF (src/cgpu_fcns.cuh:2010-2040; src/microscope_effects.cuh:57-145):
    E(g) = exp(-(spa + temp) / u) / sqrt(u), u = 1 + (pi th_c D)^2 g^2,
    temp = (pi D lambda g^2)^2 / 4, spa = (pi th_c (Cs lambda^2 g^2 + df))^2 g^2,
    D = sqrt2 ti_sigma, th_c = asin(sqrt2 si_sigma lambda)
    (Lens::set_ti_sigma / set_si_sigma, src/types.cuh:2978-2989)."""
    lam = _lambda_codata(300.0)
    df, cs_mm, ti, si = -60.0, 1.2, 40.0, 0.01
    lens = dict(obj_lens_c_10=df, obj_lens_c_30=cs_mm, obj_lens_ti_sigma=ti, cond_lens_si_sigma=si)
    p, coh = _two_beam(1, **lens)
    _, pcoh = _two_beam(2, **lens)
    _, coh2 = _two_beam(1, **dict(lens, obj_lens_c_10=df + 30.0))
    D, thc, cs = np.sqrt(2) * ti, np.arcsin(np.sqrt(2) * si * lam), cs_mm * 1e7
    for (kx, ky), a in coh.items():
        g2 = (kx / p.spec_lx) ** 2 + (ky / p.spec_ly) ** 2
        u = 1 + (np.pi * thc * D) ** 2 * g2
        temp = (np.pi * D * lam * g2) ** 2 / 4
        spa = (np.pi * thc * (cs * lam ** 2 * g2 + df)) ** 2 * g2
        E = np.exp(-(spa + temp) / u) / np.sqrt(u)
        r = pcoh[(kx, ky)] / a
        # compared in the exponent: the reference's constants differ from
        # CODATA's by ~3e-7, which moves ln E by a few 1e-6 of itself
        assert abs(r.imag) < 1e-9 * E + 1e-14 and r.real > 0     # same chi, real envelope
        assert abs(np.log(r.real) - np.log(E)) < 5e-6 * abs(np.log(E)) + 1e-9, ((kx, ky), r, E)
        assert abs(abs(a) - 0.05) < 1e-12           # coherent: |exp(i chi)| = 1
        # a 30 A defocus change multiplies the coefficient by exp(-i pi lambda 30 g^2)
        assert abs(coh2[(kx, ky)] / a - np.exp(-1j * np.pi * lam * 30.0 * g2)) < 1e-6
