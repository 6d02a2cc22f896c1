"""Synthetic crystalline specimens.

`crystal_by_lays` follows Crystal_Spec (src/xtl_build.hpp:39-133): unit-cell
layers are tiled na x nb (boundary positions included, x <= na*a), stacked nc
times along z, and the first layer is repeated on top as a closing layer.
The *_xtl builders follow crystalline_materials/{fcc001,fcc110,Si001,
SrTiO3001}_xtl.m.  Rows are Z x y z sigma occ region charge.
"""
import numpy as np

_EPS_REL_F = 1e-4   # Epsilon<float>::rel; the builder's bound check tolerance


def crystal_by_lays(na, nb, nc, a, b, c, ulayers):
    """ulayers: list of (k x 8) arrays in fractional coordinates."""
    xmin, xmax = 0.0 - _EPS_REL_F, na * a + _EPS_REL_F
    ymin, ymax = 0.0 - _EPS_REL_F, nb * b + _EPS_REL_F
    lays = []
    for ul in ulayers:
        ul = np.asarray(ul, dtype=np.float64)
        rows = []
        for j in range(nb + 1):
            for i in range(na + 1):
                for k in range(ul.shape[0]):
                    x = (i + ul[k, 1]) * a
                    y = (j + ul[k, 2]) * b
                    if xmin <= x <= xmax and ymin <= y <= ymax:
                        r = ul[k].copy()
                        r[1], r[2], r[3] = x, y, c * ul[k, 3]
                        rows.append(r)
        lays.append(np.array(rows).reshape(-1, 8))
    out = []
    for k in range(nc):
        for lay in lays:
            L = lay.copy()
            L[:, 3] += c * k
            out.append(L)
    top = lays[0].copy()
    top[:, 3] += c * nc
    out.append(top)
    return np.vstack(out)


def _finish(atoms, na, nb, nc, a, b, c, ncu):
    return atoms, na * a, nb * b, nc * c, a, b, c, c / ncu


def fcc001_xtl(Z, l_c, na, nb, nc, ncu, rmsd_3d):
    l1 = [[Z, 0.0, 0.0, 0.0, rmsd_3d, 1, 0, 0], [Z, 0.5, 0.5, 0.0, rmsd_3d, 1, 0, 0]]
    l2 = [[Z, 0.0, 0.5, 0.5, rmsd_3d, 1, 0, 0], [Z, 0.5, 0.0, 0.5, rmsd_3d, 1, 0, 0]]
    return _finish(crystal_by_lays(na, nb, nc, l_c, l_c, l_c, [l1, l2]), na, nb, nc, l_c, l_c, l_c, ncu)


def fcc110_xtl(Z, l_c, na, nb, nc, ncu, rmsd_3d):
    a, b, c = l_c / np.sqrt(2), l_c, l_c / np.sqrt(2)
    l1 = [[Z, 0.0, 0.0, 0.0, rmsd_3d, 1, 0, 0]]
    l2 = [[Z, 0.5, 0.5, 0.5, rmsd_3d, 1, 0, 0]]
    return _finish(crystal_by_lays(na, nb, nc, a, b, c, [l1, l2]), na, nb, nc, a, b, c, ncu)


def si001_xtl(na, nb, nc, ncu, rmsd_3d, a=5.4307):
    L = [[[14, 0.00, 0.00, 0.00], [14, 0.50, 0.50, 0.00]],
         [[14, 0.25, 0.25, 0.25], [14, 0.75, 0.75, 0.25]],
         [[14, 0.00, 0.50, 0.50], [14, 0.50, 0.00, 0.50]],
         [[14, 0.25, 0.75, 0.75], [14, 0.75, 0.25, 0.75]]]
    ul = [[r + [rmsd_3d, 1, 0, 0] for r in lay] for lay in L]
    return _finish(crystal_by_lays(na, nb, nc, a, a, a, ul), na, nb, nc, a, a, a, ncu)


def diamond001_xtl(Z, l_c, na, nb, nc, ncu, rmsd_3d):
    """Diamond-cubic [001] of element Z with lattice constant l_c (Si001_xtl's layers)."""
    L = [[[Z, 0.00, 0.00, 0.00], [Z, 0.50, 0.50, 0.00]],
         [[Z, 0.25, 0.25, 0.25], [Z, 0.75, 0.75, 0.25]],
         [[Z, 0.00, 0.50, 0.50], [Z, 0.50, 0.00, 0.50]],
         [[Z, 0.25, 0.75, 0.75], [Z, 0.75, 0.25, 0.75]]]
    ul = [[r + [rmsd_3d, 1, 0, 0] for r in lay] for lay in L]
    return _finish(crystal_by_lays(na, nb, nc, l_c, l_c, l_c, ul), na, nb, nc, l_c, l_c, l_c, ncu)


def srtio3001_xtl(na, nb, nc, ncu, rmsd_3d, a=3.9050, rmsd_O=None):
    """SrTiO3 [001]; rmsd_O (if given) is the O rms displacement, rmsd_3d that of Sr and Ti."""
    so = rmsd_3d if rmsd_O is None else rmsd_O
    l1 = [[38, 0.0, 0.0, 0.0, rmsd_3d, 1, 0, 0], [8, 0.5, 0.5, 0.0, so, 1, 0, 0]]
    l2 = [[8, 0.0, 0.5, 0.5, so, 1, 0, 0], [8, 0.5, 0.0, 0.5, so, 1, 0, 0],
          [22, 0.5, 0.5, 0.5, rmsd_3d, 1, 0, 0]]
    return _finish(crystal_by_lays(na, nb, nc, a, a, a, [l1, l2]), na, nb, nc, a, a, a, ncu)


def perovskite001_xtl(ZA, ZB, ZX, l_c, na, nb, nc, ncu, rmsd_3d):
    l1 = [[ZA, 0.0, 0.0, 0.0, rmsd_3d, 1, 0, 0], [ZX, 0.5, 0.5, 0.0, rmsd_3d, 1, 0, 0]]
    l2 = [[ZX, 0.0, 0.5, 0.5, rmsd_3d, 1, 0, 0], [ZX, 0.5, 0.0, 0.5, rmsd_3d, 1, 0, 0],
          [ZB, 0.5, 0.5, 0.5, rmsd_3d, 1, 0, 0]]
    return _finish(crystal_by_lays(na, nb, nc, l_c, l_c, l_c, [l1, l2]), na, nb, nc, l_c, l_c, l_c, ncu)
