"""Brute-force restatements used as independent checks by the tests (not
collected as tests): the voxel phase field of a handful of points by a
minimum-image search over the link-cell box instead of link cells."""
import numpy as np

from simulated_tem_amd.abi import mat_from_cm


def point_positions(oc):
    """Positions (A, link-cell frame) and g.r of an oracle condition's points."""
    cs, ids, frac = oc.cells()
    p = oc.plan
    ncell = p.Ex * p.Ey * p.Ez
    cell = np.repeat(np.arange(ncell), np.diff(cs))
    cz, rem = np.divmod(cell, p.Ex * p.Ey)
    cy, cx = np.divmod(rem, p.Ex)
    a = np.diag(mat_from_cm(p.a_col))
    pos = (frac + np.stack([cx, cy, cz], axis=1)) * a
    return pos, oc.gdotr()[ids - 1]


def min_image_field(oc, rec):
    """(x [Mz, My, Mx] complex, count, smallest distance to the cutoff) of the
    voxel sum S = sum exp(-d^2 / 2 sigma^2 + i g.r) over points within
    3 sigma, d the minimum-image displacement over the link-cell box; x =
    S / max(|S|, 1e-16).  Cuboid link cells wider than 3 sigma (then the
    27-cell search sees exactly the minimum images within range)."""
    p = oc.plan
    pos, gd = point_positions(oc)
    L = np.array([p.Ex, p.Ey, p.Ez]) * np.diag(mat_from_cm(p.a_col))
    va = np.diag(mat_from_cm(rec.gk_a))
    lcd = np.array(list(p.lc_delta))
    iz, iy, ix = np.meshgrid(np.arange(p.Mz), np.arange(p.My), np.arange(p.Mx), indexing="ij")
    v = np.stack([ix, iy, iz], axis=-1) * va - lcd
    S = np.zeros(v.shape[:3], complex)
    cnt = np.zeros(v.shape[:3], np.int32)
    gap = np.inf
    for q, g in zip(pos, gd):
        d = q - v
        d -= L * np.round(d / L)
        r2 = np.sum(d * d, axis=-1)
        gap = min(gap, float(np.min(np.abs(r2 - p.rc2))))
        inr = r2 <= p.rc2
        S += np.where(inr, np.exp(-r2 * p.i2s2 + 1j * g), 0.0)
        cnt += inr
    x = S / np.maximum(np.abs(S), 1e-16)
    return x, cnt, gap
