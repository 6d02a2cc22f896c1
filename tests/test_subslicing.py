"""Sub-slicing (potential_slicing = 3, dz_Sub): the z-integrated potential of
each (slice, atom) pair, z displacements, and the slicing, on the CPU.

The reference's intended path is set_atom_Vp's sub-slice branch with
linear_Vz / cubic_poly_coef (src/projected_potential.cuh:253-264,
src/gpu_fcns.cuh:175-285); in the reference snapshot the call computing the
coefficients is commented out (src/projected_potential.cuh:118,148), so its
output cannot pin this one.  The oracle's restatement is pinned by the
mathematics instead:
  - the 128-node tanh-sinh rule integrates known functions on its range;
  - the sub-slices of one atom covering its whole extent add up to the
    projected potential of the element table (closed form in K0/K1), knot by
    knot, values and R^2-derivatives;
  - a thin slab through the atom gives thickness x Vr(R) (the 3D potential);
  - integrals over adjacent ranges add; an atom at the centre of a slab gives
    twice the half slab.
The engine's dz_Sub slicing (boundaries, atom ranges per configuration) is
compared with the oracle's through the C ABI.
"""
import numpy as np
import pytest

from multem_amd import configs, engine
from oracle import oracle


def _sub_params(nconf=2, pn_dim=111, dz=1.0, n=256, ncells=4, nc=3):
    p = configs.au_stem_512(ns=2, nconf=nconf, n=n, ncells=ncells, nc=nc)
    p.potential_slicing = 3
    p.spec_dz = dz
    p.pn_dim = pn_dim
    return p


def test_rule_integrates_known_functions():
    """The reference's rule is a tanh-sinh trapezoid in t over the whole
    truncated range [-1 + eps, 1 - eps], end weights not halved: it
    integrates smooth functions on that range to about the end weights
    (6e-7) times the function's size."""
    x, w = oracle.subslice_rule()
    eps = 10 * np.finfo(np.float32).eps
    Lr = 1 - eps
    assert np.all(np.diff(x) > 0) and -Lr <= x[0] and x[-1] <= Lr
    assert abs(x[0] + x[-1]) < 1e-12 and np.allclose(w, w[::-1], rtol=1e-12)
    h = 2 * np.arcsinh(np.arctanh(Lr) / (np.pi / 2)) / 127   # step in t
    assert w[0] < 1e-6 and abs(w[64] / (h * np.pi / 2) - 1) < 1e-3   # tiny ends, h pi/2 in the middle
    cases = [(lambda u: 1 + 0 * u, 2 * Lr),
             (lambda u: u ** 2, 2 * Lr ** 3 / 3),
             (lambda u: 1 / (1 + u ** 2), 2 * np.arctan(Lr)),
             (lambda u: np.exp(3 * u), (np.exp(3 * Lr) - np.exp(-3 * Lr)) / 3),
             (lambda u: np.log1p(u + 1.0), _int_log(Lr))]   # log(2 + x)
    for f, exact in cases:
        fmax = np.max(np.abs(f(x)))
        assert abs(np.sum(w * f(x)) - exact) < 1.5e-6 * fmax


def _int_log(L):
    # integral of log(2 + x) over [-L, L]
    f = lambda u: (u) * np.log(u) - u
    return f(2 + L) - f(2 - L)


@pytest.mark.parametrize("Z", [8, 38, 79])
def test_subslices_add_up_to_projected_potential(Z):
    dR = 0.05
    tab = oracle.atom_table(Z, dR)
    Rm = tab["R_max"]
    edges = [-3 * Rm, -0.7 * Rm, -0.11, 0.37, 0.9 * Rm, 3 * Rm]
    c0 = np.zeros(128)
    c1 = np.zeros(128)
    for a, b in zip(edges[:-1], edges[1:]):
        s0, s1, _, _ = oracle.subslice_table(Z, dR, a, b, 0.0)
        c0 += s0
        c1 += s1
    # agreement to the rule's accuracy (about 1e-6 per slab; the z tails
    # beyond 3 R_max are below 1e-7 of the value at the first knot); a wrong
    # factor, sign or term would miss by 1e-2 or more
    scale = abs(tab["c0"][0])
    assert np.max(np.abs(c0 - tab["c0"])) < 5e-6 * scale, np.max(np.abs(c0 - tab["c0"])) / scale
    sd = np.max(np.abs(tab["c1"]))
    assert np.max(np.abs(c1 - tab["c1"])) < 5e-6 * sd


def test_thin_slab_is_thickness_times_3d_potential():
    Z, dR, d = 79, 0.05, 1e-3
    tab = oracle.atom_table(Z, dR)
    c0, _, _, _ = oracle.subslice_table(Z, dR, -d / 2, d / 2, 0.0)
    R = np.sqrt(tab["R2"])
    Vr3 = np.zeros(128)
    L = oracle.lib("f64")
    from multem_amd import capi
    L.orc_lobato_VR(Z, 128, capi.dptr(np.ascontiguousarray(R)), None, None, capi.dptr(Vr3))
    # below the taper radius and at R >> d, where the next term of the
    # expansion, (d/R)^2/24 times a factor of order one, is below 1e-6
    m = (R > 300 * d) & (tab["R2"] < tab["R2_tap"])
    assert np.count_nonzero(m) > 20
    assert np.max(np.abs(c0[m] / (d * Vr3[m]) - 1)) < 3e-6


def test_ranges_add_and_centre_split_is_symmetric():
    Z, dR = 38, 0.04
    a = oracle.subslice_table(Z, dR, -1.3, 0.4, 0.2)
    b = oracle.subslice_table(Z, dR, 0.4, 2.9, 0.2)
    ab = oracle.subslice_table(Z, dR, -1.3, 2.9, 0.2)
    # to the rule's accuracy: its end-node error does not cancel between splits
    for k in range(2):
        assert np.max(np.abs(a[k] + b[k] - ab[k])) < 2e-5 * np.max(np.abs(ab[k]))
    # the two halves of a slab centred on the atom map onto each other node for node
    full = oracle.subslice_table(Z, dR, -0.8, 0.8, 0.0)
    half = oracle.subslice_table(Z, dR, 0.0, 0.8, 0.0)
    for k in range(2):
        assert np.max(np.abs(full[k] - 2 * half[k])) < 1e-12 * np.max(np.abs(full[k]))


def test_fp32_table_follows_fp64():
    c64 = oracle.subslice_table(79, 0.05, -1.0, 1.2, 0.1)
    c32 = oracle.subslice_table(79, 0.05, -1.0, 1.2, 0.1, precision="f32")
    assert np.max(np.abs(c32[0] - c64[0])) < 1e-5 * np.max(np.abs(c64[0]))


@pytest.mark.parametrize("precision", [1, 2])
def test_engine_subslicing_matches_oracle(precision):
    p = _sub_params()
    p.system_conf.precision = precision
    o = "f32" if precision == 1 else "f64"
    base = engine.spec_slicing(p, 0)
    ref = oracle.spec_slicing(p, 0, precision=o)
    assert len(base) == len(ref) > 8
    for a, b in zip(base, ref):
        assert np.allclose(a[:4], b[:4], rtol=0, atol=1e-4) and a[4:] == b[4:], (a, b)
    # contiguous boundaries from z_int_min to z_int_max, wider atom ranges
    zs = [s[0] for s in base] + [base[-1][1]]
    assert np.all(np.diff(zs) > 0)
    assert all(s[2] <= s[0] and s[3] >= s[1] for s in base)
    assert base[-1][6] == 0 and all(s[6] == -1 for s in base[:-1])   # whole specimen at the last slice
    for iconf in (1, 2):
        e, r = engine.spec_slicing(p, iconf), oracle.spec_slicing(p, iconf, precision=o)
        assert [s[4:6] for s in e] == [s[4:6] for s in r], iconf


@pytest.mark.parametrize("slicing, pn_dim, moved", [(3, 111, True), (3, 110, False), (2, 111, False)])
def test_z_displacements_only_with_subslicing(slicing, pn_dim, moved):
    """pn_dim's z digit displaces z only with sub-slicing (the oracle's
    configuration atoms, src/input_multislice.cuh:702-705); the z draws use
    the configuration's third seed, rms = the atom's sigma."""
    p = _sub_params(pn_dim=pn_dim)
    p.potential_slicing = slicing
    z0 = np.sort(oracle.displaced_atoms(p, 0)[:, 2])
    z1 = np.sort(oracle.displaced_atoms(p, 1)[:, 2])
    dz = z1 - z0
    if moved:
        assert 0.05 < np.std(dz) < 0.12   # sigma 0.085
    else:
        assert np.array_equal(z0, z1)
