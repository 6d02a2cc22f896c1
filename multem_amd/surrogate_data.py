"""Training and evaluation pairs for the frozen-phonon surrogate, made by the engine.

One pair is a seeded random specimen with
  input  channel 0: the 2-configuration frozen-phonon ADF image,
         channel 1: the still-atom projected potential summed over all
                    slices along the (possibly tilted) beam path, sampled
                    at the scan points,
  target          : the 64-configuration ADF image of the same specimen
(configurations 1..2 of the 64-configuration run are the 2-configuration
run's, since configuration i is seeded by (pn_seed, i)).

Specimens: fcc, diamond-cubic or perovskite [001], a in [3.5, 5.5] A, Z in
6..82 (perovskite X site O), rms displacement 0.05-0.12 A, thickness
20-200 A; 300 kV, 20-25 mrad probe, zero defocus, Cs = 0, 60-200 mrad ring;
ns x ns periodic scan over ns/32 unit cells per side (one cell at the
training size ns = 32, so every scan samples a cell by 32 points; a scan
larger than the box repeats it, the specimen being periodic); the lateral box is the largest
number of cells that keeps the 256^2 grid's Nyquist frequency beyond the
200 mrad ring.  Illumination tilt theta is uniform in [0, tilt_max] degrees
with a uniform azimuth (tilt_max = 0 by default: untilted).

    python -m multem_amd.surrogate_data --n 200 --seed 1 --out pairs.npz
"""
import argparse
import sys
import time

import numpy as np

from . import engine
from .configs import _base
from .crystal import diamond001_xtl, fcc001_xtl, perovskite001_xtl

GRID = 256
BOX_MAX = 0.5 * GRID / (0.200 / 0.0196875)   # Angstrom: Nyquist 1/(2 dR) at 200 mrad, 300 kV


def random_specimen(rng, ns=32, nconf=64, tilt_max=0.0):
    """Parameters of one random specimen (STEM, `nconf` frozen-phonon configurations)."""
    kind = int(rng.integers(0, 3))
    a = float(rng.uniform(3.5, 5.5))
    sigma = float(rng.uniform(0.05, 0.12))
    thick = float(rng.uniform(20.0, 200.0))
    nl = max(2, int(BOX_MAX // a))
    nc = max(1, int(round(thick / a)))
    if kind == 0:
        Z = int(rng.integers(6, 83))
        xtl = fcc001_xtl(Z, a, nl, nl, nc, 2, sigma)
        desc = "fcc Z=%d" % Z
    elif kind == 1:
        Z = int(rng.integers(6, 83))
        xtl = diamond001_xtl(Z, a, nl, nl, nc, 4, sigma)
        desc = "diamond Z=%d" % Z
    else:
        ZA, ZB = (int(z) for z in rng.choice(np.arange(6, 83), size=2, replace=False))
        xtl = perovskite001_xtl(ZA, ZB, 8, a, nl, nl, nc, 2, sigma)
        desc = "perovskite A=%d B=%d" % (ZA, ZB)
    p = _base()
    p.simulation_type = 11
    p.spec_atoms, p.spec_lx, p.spec_ly, p.spec_lz, _, _, _, p.spec_dz = xtl
    p.nx = p.ny = GRID
    p.pn_model = 3
    p.pn_nconf = nconf
    p.pn_dim = 110
    p.pn_seed = int(rng.integers(1, 2 ** 31 - 1))
    p.iw_type = 2
    p.cond_lens_outer_aper_ang = float(rng.uniform(20.0, 25.0))
    p.cond_lens_zero_defocus_type = 1
    p.thick_type = 1
    p.scanning_type = 2
    p.scanning_periodic = 1
    p.scanning_square_pxs = 1
    p.scanning_ns = ns
    c0 = nl // 2
    k = max(1, ns // 32)   # cells scanned per side: 32 scan points per cell
    p.scanning_x0 = p.scanning_y0 = c0 * a
    p.scanning_xe = p.scanning_ye = (c0 + k) * a
    p.detector_type = 1
    p.detector_cir = [(60.0, 200.0)]
    if tilt_max > 0:
        p.theta = float(rng.uniform(0.0, tilt_max))
        p.phi = float(rng.uniform(0.0, 360.0))
    meta = dict(kind=kind, a=a, sigma=sigma, thick=nc * a, cells=nl, desc=desc, theta=p.theta, phi=p.phi)
    return p, meta


def potential_map(p):
    """Still-atom projected potential summed over all slices, sampled (bilinear,
    periodic) at the scan points: [ns, ns] with [ix, iy] as the ADF image.

    Under an illumination tilt (theta, phi) the wave drifts laterally by
    -z sin(theta) (cos phi, sin phi) at depth z below the entrance surface
    (the tilted propagator, |g - g0|^2), so the probe entering at r meets
    the atoms near r - z sin(theta) u at depth z.  The map is then the
    projection along the tilted path: every atom is moved by
    +z sin(theta) u (periodically wrapped) before projecting."""
    q = p.copy()
    q.pn_model = 1
    q.pn_nconf = 1
    if p.theta:
        at = np.array(q.spec_atoms, dtype=np.float64, copy=True)
        t, f = np.radians(p.theta), np.radians(p.phi)
        zr = at[:, 3] - at[:, 3].min()
        at[:, 1] = np.mod(at[:, 1] + zr * np.sin(t) * np.cos(f), p.spec_lx)
        at[:, 2] = np.mod(at[:, 2] + zr * np.sin(t) * np.sin(f), p.spec_ly)
        q.spec_atoms = at
        q.theta = 0.0
    V = engine.projected_potential_sum(q, 1)   # [nx, ny], pixel ix at x = ix*dx
    n = p.scanning_ns
    xs = p.scanning_x0 + (p.scanning_xe - p.scanning_x0) * np.arange(n) / n
    ys = p.scanning_y0 + (p.scanning_ye - p.scanning_y0) * np.arange(n) / n
    fx, fy = xs / (p.spec_lx / p.nx), ys / (p.spec_ly / p.ny)
    ix0, iy0 = np.floor(fx).astype(int), np.floor(fy).astype(int)
    tx, ty = (fx - ix0)[:, None], (fy - iy0)[None, :]
    i0, i1 = ix0 % p.nx, (ix0 + 1) % p.nx
    j0, j1 = iy0 % p.ny, (iy0 + 1) % p.ny
    return ((1 - tx) * (1 - ty) * V[np.ix_(i0, j0)] + tx * (1 - ty) * V[np.ix_(i1, j0)]
            + (1 - tx) * ty * V[np.ix_(i0, j1)] + tx * ty * V[np.ix_(i1, j1)])


def adf(p, nconf):
    q = p.copy()
    q.pn_nconf = nconf
    return engine.run_multislice(q).image_tot[0, 0].copy()


def make_pair(p, n_small=2, n_big=64):
    t0 = time.perf_counter()
    small = adf(p, n_small)
    t1 = time.perf_counter()
    vmap = potential_map(p)
    t2 = time.perf_counter()
    big = adf(p, n_big)
    t3 = time.perf_counter()
    return small, vmap, big, dict(t_small=t1 - t0, t_vmap=t2 - t1, t_big=t3 - t2)


def generate(n, seed, ns=32, t_max=None, tilt_max=0.0, log=sys.stderr):
    rng = np.random.default_rng(seed)
    S, Vm, T, metas = [], [], [], []
    t_start = time.perf_counter()
    for k in range(n):
        p, meta = random_specimen(rng, ns, tilt_max=tilt_max)
        small, vmap, big, tm = make_pair(p)
        meta.update(tm)
        S.append(small)
        Vm.append(vmap)
        T.append(big)
        metas.append(meta)
        rr = np.sqrt(np.mean((small - big) ** 2) / np.mean(big ** 2))
        print("pair %d/%d %s a=%.3f t=%.0f A: 2-conf rel rms %.4f, %.2f s" %
              (k + 1, n, meta["desc"], meta["a"], meta["thick"], rr, tm["t_small"] + tm["t_vmap"] + tm["t_big"]),
              file=log, flush=True)
        if t_max and time.perf_counter() - t_start > t_max:
            break
    return np.array(S), np.array(Vm), np.array(T), metas


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=100)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--ns", type=int, default=32)
    ap.add_argument("--t-max", type=float, default=None, help="stop after this many seconds")
    ap.add_argument("--tilt-max", type=float, default=0.0, help="largest illumination tilt (degrees)")
    ap.add_argument("--out", required=True)
    a = ap.parse_args()
    S, Vm, T, metas = generate(a.n, a.seed, a.ns, a.t_max, a.tilt_max)
    np.savez_compressed(a.out, adf_small=S.astype(np.float32), vmap=Vm.astype(np.float32),
                        adf_big=T.astype(np.float32), a=np.array([m["a"] for m in metas]),
                        thick=np.array([m["thick"] for m in metas]), kind=np.array([m["kind"] for m in metas]),
                        theta=np.array([m["theta"] for m in metas]), phi=np.array([m["phi"] for m in metas]),
                        seed=a.seed, tilt_max=a.tilt_max)
    print("wrote %d pairs to %s" % (len(S), a.out), file=sys.stderr)


if __name__ == "__main__":
    main()
