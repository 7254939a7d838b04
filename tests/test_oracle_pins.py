"""Pins of the oracle's findGdotR, voxel sum and image assembly by checks
that do not restate them.  CPU only.

  findGdotR (Lib_DynamicalTwoBeamImaging.f90:1160-1265)
      brute force: every periodic image of every atom over 5x5x5 supercell
      replicas, rotated and offset, kept where it falls inside the link-cell
      box; the oracle's points (27 replicas, link-cell order) are the same
      set, with the same cell coordinates and g.r.
  voxel sum (:1877-1892)
      a perfect bcc crystal and an allowed reflection: every g.r is the same
      modulo 2 pi, so every voxel's normalised phase factor is that constant.
  image (:1870-1951)
      a perfect crystal images every column as the perfect-crystal column
      (x = 1, rho = 1), whose RK4 march approaches intensityAtDepthZ
      (Lib_DiffractionConditions.f90:1209-1258) at fourth order.
"""
import itertools
import math

import numpy as np
import pytest

from oracle import oracle as orc
from simulated_tem_amd import synth
from simulated_tem_amd.abi import mat_from_cm


def _condition(cells, sg=0.0, g=(1, 1, 0), k=(0, 0, 1), d=None, perfect=False, n_tomo=1, tilt=0,
               zpad=False):
    cfg = synth.scaled("c0", cells=cells, d=d)
    cfg.update(sg=[sg], g=[g], k=k, zpad=zpad)
    if perfect:
        cfg.update(noise=0.0, n_centres=0, n_voids=0)
    x, box = synth.make_atoms(cfg)
    if n_tomo > 1:
        cfg.update(n_tomo=n_tomo, tomo_angle=20.0)
    x, conds = synth.conditions(cfg, x, box)
    return x, conds[tilt][0]


def _cm(a):
    return mat_from_cm(a)


def _brute_points(rec, plan, x, span=2):
    """Cell coordinates (in link-cell units) and g.r of every replica image
    inside the link-cell box, over (2 span + 1)^3 supercell replicas."""
    RU = _cm(plan.RU)
    ia = _cm(plan.ia_col)
    A = _cm(rec.super_a) * np.array(list(rec.super_n))[None, :]
    delta = np.array(list(rec.delta))
    lcd = np.array(list(plan.lc_delta))
    gg = np.array(list(plan.gg))
    E = np.array([plan.Ex, plan.Ey, plan.Ez])
    pos, gd, margin = [], [], np.inf
    for t in itertools.product(range(-span, span + 1), repeat=3):
        y = (x + A @ np.array(t, float)) @ RU.T + delta
        c = (y - lcd) @ ia.T
        inside = np.all((c >= 0) & (c < E), axis=1)
        # distance of every image to the nearest box face (ambiguity check)
        margin = min(margin, float(np.min(np.abs(np.concatenate([c, c - E], axis=1)))))
        pos.append(c[inside])
        gd.append(y[inside] @ gg)
    return np.concatenate(pos), np.concatenate(gd), margin


def _oracle_points(oc):
    cs, ids, frac = oc.cells()
    p = oc.plan
    ncell = p.Ex * p.Ey * p.Ez
    cell = np.repeat(np.arange(ncell), np.diff(cs))
    cz, rem = np.divmod(cell, p.Ex * p.Ey)
    cy, cx = np.divmod(rem, p.Ex)
    pos = frac + np.stack([cx, cy, cz], axis=1)
    gd = oc.gdotr()[ids - 1]
    return pos, gd


def _sorted(pos, gd):
    o = np.lexsort(np.round(pos[:, ::-1], 6).T)
    return pos[o], gd[o]


@pytest.mark.parametrize("kind", ["untilted", "sg_tilt", "tomo_tilt", "zone_111"])
def test_find_gdotr_brute_force_replicas(kind):
    if kind == "untilted":
        x, im = _condition((5, 4, 6))
    elif kind == "sg_tilt":
        x, im = _condition((5, 4, 6), sg=0.02, zpad=True)
    elif kind == "tomo_tilt":
        x, im = _condition((5, 6, 4), sg=0.01, g=(0, 1, 1), n_tomo=3, tilt=0, zpad=True)
    else:
        x, im = _condition((6, 5, 5), g=(1, -1, 0), k=(1, 1, 1), d=(14.0, 15.0, 11.0))
    oc = orc.Condition(im.rec, x)
    bpos, bgd, margin = _brute_points(im.rec, oc.plan, x)
    # no image within rounding of a box face: the set is unambiguous
    assert margin > 1e-9
    opos, ogd = _oracle_points(oc)
    assert oc.n_points == len(bpos) > 0
    bpos, bgd = _sorted(bpos, bgd)
    opos, ogd = _sorted(opos, ogd)
    assert np.max(np.abs(bpos - opos)) < 1e-9
    assert np.max(np.abs(bgd - ogd)) < 1e-9 * max(1.0, np.max(np.abs(bgd)))


@pytest.mark.parametrize("sg,g", [(0.0, (1, 1, 0)), (0.015, (2, 0, 0))])
def test_voxel_sum_perfect_crystal_constant_phase(sg, g):
    """bcc with h + k + l even: g.R is a multiple of 2 pi for every lattice
    site R, so g.r = g.delta' (mod 2 pi) for every point and S / |S| is one
    unit phase factor on every voxel."""
    x, im = _condition((6, 6, 6), sg=sg, g=g, perfect=True)
    oc = orc.Condition(im.rec, x)
    gd = oc.gdotr()
    ref = np.exp(1j * gd[0])
    assert np.max(np.abs(np.exp(1j * gd) - ref)) < 1e-10
    ph, rho, cnt, r2m = oc.field(window=(3, 9, 2, 7), nthreads=4)
    assert np.all(cnt > 0)
    assert np.max(np.abs(np.abs(ph) - 1.0)) < 1e-12
    assert np.max(np.abs(ph - ref)) < 1e-10
    # a lattice site within 3 sigma of a voxel at a lattice site: r2min = 0
    # there, and every voxel has ~28 neighbours (2 (4 pi / 3) 1.5^3 on average)
    assert 20 <= cnt.mean() <= 36


@pytest.mark.parametrize("atoms", [[[9.0, 9.5, 10.0]], [[0.3, 18.7, 1.1]],
                                   [[9.0, 9.5, 10.0], [10.1, 9.9, 12.0], [0.2, 0.4, 18.8]]])
def test_voxel_sum_few_atoms_min_image_brute_force(atoms):
    """One to three atoms: the oracle's field equals a minimum-image search
    over the link-cell box (counts exactly, phase factors to 1e-13); a
    voxel within 3 sigma of one image only has that image's e^{i g.r}."""
    from tests.brute import min_image_field
    x, im = _condition((6, 6, 6))
    pts = np.array(atoms, float)
    oc = orc.Condition(im.rec, pts)
    ph, rho, cnt, r2m = oc.field(nthreads=4)
    bx, bcnt, gap = min_image_field(oc, im.rec)
    assert gap > 1e-9
    assert np.array_equal(cnt, bcnt) and cnt.max() >= 1
    assert np.max(np.abs(ph - bx)) < 1e-13
    assert np.all(ph[cnt == 0] == 0)


@pytest.mark.parametrize("sg,n_prec", [(0.0, 1), (0.01, 1), (0.01, 4)])
def test_image_perfect_crystal_vs_intensity_at_depth(sg, n_prec):
    x, im = _condition((8, 8, 8), sg=sg, perfect=True)
    oc = orc.Condition(im.rec, x, 0.005, n_prec)
    p = oc.plan
    img = oc.image(window=(0, 6, 0, 5), nthreads=4)
    nz = p.Mz
    col = np.mean([orc.integrate_column(p.aa[q], p.b0, p.bg, p.dz, np.ones(nz + 1, complex), np.ones(nz + 1))
                   for q in range(n_prec)])
    # every column is the perfect-crystal column (a constant phase factor of
    # c_g leaves |phi_g|^2 unchanged)
    assert np.max(np.abs(img - col)) < 1e-12
    # and that column is intensityAtDepthZ up to the RK4 error at this step
    z = p.dz * (nz - nz % 2)
    want = np.mean([orc.intensity_at_depth(im.rec.xi0, im.rec.xig, p.aa[q] / (2 * math.pi), z)
                    for q in range(n_prec)])
    assert abs(col - want) < 2e-6
    assert np.max(np.abs(img - want)) < 2e-6
