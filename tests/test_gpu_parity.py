"""HIP path vs the FP64 oracle, through the C-ABI.

Bars (per the imaging contract):
  link cells, point ids, in-cell offsets, g.r   bit-exact
  voxel membership (neighbour count), r2min     bit-exact in FP64 mode,
                                                identical counts in FAST mode
  FP64 image                                    max |dI| <= 1e-10
  FAST image                                    max |dI| <= 1e-4
  blur                                          max |d| <= 1e-15
"""
import ctypes as C
import json
import math
import os

import numpy as np
import pytest

from oracle import oracle as orc
from simulated_tem_amd import abi, imaging, synth
from simulated_tem_amd.imaging import DynamicalTwoBeamImaging, input_supercell, lattice_a0
from tests.brute import min_image_field

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
FAST_TOL = 1e-4
NTH = min(16, os.cpu_count() or 1)


def _cond(name, cells=None, sg=None, density=None, g=None, d=None, k=None):
    cfg = synth.scaled(name, cells, d)
    if sg is not None:
        cfg["sg"] = [sg]
    if density is not None:
        cfg["density"] = density
    if g is not None:
        cfg["g"] = [g]
    if k is not None:
        cfg["k"] = k
    cfg["n_tomo"] = 1
    x, box = synth.make_atoms(cfg)
    x, conds = synth.conditions(cfg, x, box)
    return x, conds[0][0]


CASES = {
    "c0_sg0": lambda: _cond("c0"),
    "c0_sg0.01_density": lambda: _cond("c0", sg=0.01, density=True),
    "c1_small_zpad": lambda: _cond("c1", cells=(22, 18, 12)),
    "c2_small_g200_density": lambda: _cond("c2", cells=(14, 12, 30), sg=-0.02),
    "c0_k111_g1m10": lambda: _cond("c0", cells=(9, 11, 7), g=(1, -1, 0), k=(1, 1, 1),
                                   d=(25.0, 27.0, 19.0)),
    # a 33 x 31 x 29 voxel grid: the last tile along every axis is partial
    # (4-voxel tiles in x and y, 4- and 8-voxel tiles in z)
    "c0_k111_ragged": lambda: _cond("c0", cells=(10, 9, 8), g=(1, -1, 0), k=(1, 1, 1),
                                    d=(26.1, 24.7, 22.9)),
    # link cells thinner than 3 sigma along z (4.725 A, Ez = 8), and a foil
    # of three cells (Ez = 3, 2.63 A): points two cells away in z can lie
    # within 3 sigma and neighbourList1b leaves them out
    "c0_thin_z": lambda: _cond("c0", cells=(12, 12, 12), d=(37.9824, 37.9824, 37.8)),
    "c0_foil_ez3": lambda: _cond("c0", cells=(12, 12, 3), d=(37.9824, 37.9824, 7.9)),
}


@pytest.fixture(scope="module", params=sorted(CASES))
def case(request):
    print("\n[case]", request.param, flush=True)
    x, im = CASES[request.param]()
    oc = orc.Condition(im.rec, x)
    return request.param, x, im, oc


def test_linkcells_bitexact(case):
    name, x, im, oc = case
    cs, ids, frac, gd = im.linkcells(x)
    ocs, oids, ofrac = oc.cells()
    assert len(ids) == oc.n_points
    assert np.array_equal(cs, ocs)
    assert np.array_equal(ids, oids)
    assert np.array_equal(frac, ofrac)
    assert np.array_equal(gd, oc.gdotr())
    # every point in exactly one cell, offsets in [0,1)
    assert np.array_equal(np.sort(ids), np.arange(1, len(ids) + 1))
    assert np.all(frac >= 0) and np.all(frac < 1)


def test_binning_after_smaller_and_larger_conditions():
    """The workspace's point and slot arrays grow for a larger condition and
    are reused, larger than needed, by a smaller one.  Small, large, small,
    large in one process: the link cells and the FAST and FP64 images of each
    equal the oracle's (cells bit for bit) every time."""
    small = _cond("c0")
    large = _cond("c1", cells=(22, 18, 12))
    for x, im in (small, large, small, large):
        oc = orc.Condition(im.rec, x)
        cs, ids, frac, gd = im.linkcells(x)
        ocs, oids, ofrac = oc.cells()
        assert np.array_equal(cs, ocs) and np.array_equal(ids, oids) and np.array_equal(frac, ofrac)
        ref = orc.Condition(im.rec, x, 0.005, 2).image(nthreads=NTH)
        assert np.max(np.abs(im.computeD2BI_img(x, 0.005, 2, mode="fp64") - ref)) <= FP64_TOL
        assert np.max(np.abs(im.computeD2BI_img(x, 0.005, 2, mode="fast") - ref)) <= FAST_TOL


def test_field_fp64_membership_bitexact(case):
    name, x, im, oc = case
    ph, rho, cnt, r2m = im.field(x, "fp64")
    oph, orho, ocnt, or2m = oc.field(nthreads=NTH)
    assert np.array_equal(cnt, ocnt)
    assert np.array_equal(r2m, or2m)
    assert np.max(np.abs(ph - oph)) < 1e-12
    assert np.max(np.abs(rho - orho)) < 1e-14


@pytest.fixture(params=["0", "1", "2", "3"])
def fast_variant(request, monkeypatch):
    """Every FAST voxel kernel: a lane per voxel (0), a lane per row of four
    voxels and quarter of the candidates of a 4 x 4 x 4 tile (1), a lane per
    row and half of the candidates of a 4 x 4 x 8 tile (2), a lane per row of
    eight voxels and quarter of the candidates of an 8 x 4 x 4 tile (3)."""
    monkeypatch.setenv("D2BI_FAST_VARIANT", request.param)
    return request.param


def test_field_fast_imaging_kernel(case, fast_variant):
    """The imaging kernel's own field (no diagnostics) against the oracle.
    Leaving out or adding one atom at the 3 sigma cut moves x by its weight
    exp(-4.5) over |S| (~2e-3 for bcc), far above the FP32 error (~3e-7), so
    the bound below also checks that membership is exact."""
    name, x, im, oc = case
    ph, rho, _, _ = im.field(x, "fast", diag=False)
    oph, orho, ocnt, or2m = oc.field(nthreads=NTH)
    err = np.abs(ph - oph)
    print("[fast field, imaging kernel] %s v%s max|dx| %.3e mean %.3e max|drho| %.3e"
          % (name, fast_variant, err.max(), err.mean(), np.max(np.abs(rho - orho))), flush=True)
    assert err.max() < 1e-5
    assert np.max(np.abs(rho - orho)) < 1e-5


def test_field_fast_membership_exact(case):
    """The FAST kernel's diagnostic form (a lane per voxel, neighbour counts
    and nearest ranges): counts identical to the oracle's."""
    name, x, im, oc = case
    ph, rho, cnt, r2m = im.field(x, "fast")
    oph, orho, ocnt, or2m = oc.field(nthreads=NTH)
    assert np.array_equal(cnt, ocnt)                      # guard band keeps membership exact
    occ = ocnt > 0
    assert np.max(np.abs(r2m[occ] - or2m[occ]) / np.maximum(or2m[occ], 1.0)) < 1e-5
    err = np.abs(ph - oph)
    print("[fast field] %s max|dx| %.3e mean %.3e" % (name, err.max(), err.mean()), flush=True)
    assert np.max(np.abs(rho - orho)) < 1e-5


@pytest.mark.parametrize("name", ["c0_sg0.01_density", "c2_small_g200_density", "c1_small_zpad"])
@pytest.mark.parametrize("n_prec", [8, 19, 40])
def test_fast_integrators_agree(name, n_prec, monkeypatch):
    """8 or more precession points per pass go through the shared RK4 step
    matrix (passes of 32, 16, 8, then the stage-by-stage march for the rest:
    19 = 16 + 2 + 1, 40 = 32 + 8).  It is the same map as the stage-by-stage
    march: the two FAST images differ by FP32 rounding, and both meet the
    FAST bar against the oracle's FP64 image."""
    print("\n[integrators]", name, n_prec, flush=True)
    x, im = CASES[name]()
    ref = orc.Condition(im.rec, x, 0.005, n_prec).image(nthreads=NTH)
    poly = im.computeD2BI_img(x, 0.005, n_prec, mode="fast")
    monkeypatch.setenv("D2BI_FAST_INTEGRATOR", "stage")
    stage = im.computeD2BI_img(x, 0.005, n_prec, mode="fast")
    d = np.max(np.abs(poly - stage))
    print("[integrators] %s P=%d |poly - stage| %.3e  poly vs fp64 %.3e  stage vs fp64 %.3e"
          % (name, n_prec, d, np.max(np.abs(poly - ref)), np.max(np.abs(stage - ref))), flush=True)
    assert d <= 2e-6
    assert np.max(np.abs(poly - ref)) <= FAST_TOL
    assert np.max(np.abs(stage - ref)) <= FAST_TOL


@pytest.mark.parametrize("n_prec", [1, 5])
def test_image_fp64_and_fast(case, n_prec, fast_variant):
    name, x, im, oc = case
    ref = orc.Condition(im.rec, x, 0.005, n_prec).image(nthreads=NTH)
    g64 = im.computeD2BI_img(x, 0.005, n_prec, mode="fp64")
    gf = im.computeD2BI_img(x, 0.005, n_prec, mode="fast")
    e64 = np.max(np.abs(g64 - ref))
    ef = np.abs(gf - ref)
    print("[image] %s P=%d v%s fp64 max %.3e  fast max %.3e mean %.3e  I in [%.4f, %.4f]"
          % (name, n_prec, fast_variant, e64, ef.max(), ef.mean(), ref.min(), ref.max()), flush=True)
    assert e64 <= FP64_TOL
    assert ef.max() <= FAST_TOL


@pytest.mark.parametrize("n_app,ap_angle,density", [(1, 1e-4, False), (8, 1e-3, True), (30, 3e-3, False)])
def test_image_aperture(n_app, ap_angle, density, fast_variant):
    """computeD2BI_img with apertureAngle / nAperturePoints against the
    oracle's joint voxel loop (one phase sum per aperture point, integrations
    in (prec, app) order); nApp = 1 is the plain image bit for bit.  n = 8
    takes the reference's table of points, n = 30 the sunflower placement."""
    x, im = _cond("c0", cells=(9, 11, 7), g=(1, -1, 0), k=(1, 1, 1), d=(25.0, 27.0, 19.0), sg=0.01,
                  density=density)
    ref = orc.compute_img_aperture(im.rec, x, 0.005, 3, ap_angle, n_app, nthreads=NTH)
    g64 = im.computeD2BI_img(x, 0.005, 3, mode="fp64", apertureAngle=ap_angle, nAperturePoints=n_app)
    gf = im.computeD2BI_img(x, 0.005, 3, mode="fast", apertureAngle=ap_angle, nAperturePoints=n_app)
    e64 = np.max(np.abs(g64 - ref))
    ef = np.abs(gf - ref)
    print("[aperture] n=%d angle=%g v%s fp64 max %.3e  fast max %.3e mean %.3e"
          % (n_app, ap_angle, fast_variant, e64, ef.max(), ef.mean()), flush=True)
    assert e64 <= FP64_TOL
    assert ef.max() <= FAST_TOL
    if n_app == 1:
        assert np.array_equal(g64, im.computeD2BI_img(x, 0.005, 3, mode="fp64"))
        assert np.array_equal(gf, im.computeD2BI_img(x, 0.005, 3, mode="fast"))


@pytest.mark.parametrize("n_app,density,geom", [(7, True, "k111"), (8, False, "k111"), (6, False, "thin_z")])
def test_image_aperture_shared_pass(n_app, density, geom, fast_variant, monkeypatch):
    """The aperture sum's shared pass (one binning, the phases of groups of
    four and two points per slot, one neighbour walk per group) against a
    binning and voxel pass per point (D2BI_APERTURE_SHARED=0): the same image
    to the bit in both modes, and the same statistics.  n = 7 is a group of
    four, one of two and a single point; thin z cells take the row kernel with
    the z-offset test.  Variants 0 and 3 have no shared kernel and walk once
    per point over the one binning."""
    if geom == "k111":
        x, im = _cond("c0", cells=(9, 11, 7), g=(1, -1, 0), k=(1, 1, 1), d=(25.0, 27.0, 19.0), sg=0.01,
                      density=density)
    else:
        x, im = _cond("c0", cells=(12, 12, 12), d=(37.9824, 37.9824, 37.8), sg=0.01, density=density)
    out = {}
    for shared in ("1", "0"):
        monkeypatch.setenv("D2BI_APERTURE_SHARED", shared)
        for mode in ("fp64", "fast"):
            st = abi.D2biStats()
            img = im.computeD2BI_img(x, 0.005, 3, mode=mode, apertureAngle=2e-3, nAperturePoints=n_app, stats=st)
            out[shared, mode] = (img, st.n_points, st.n_guard, st.max_per_cell, st.n_cells)
    print("[aperture shared] n=%d %s v%s guard fast %d" % (n_app, geom, fast_variant, out["1", "fast"][2]),
          flush=True)
    for mode in ("fp64", "fast"):
        a, b = out["1", mode], out["0", mode]
        assert np.array_equal(a[0], b[0]), mode
        assert a[1:] == b[1:], mode


def test_blur_matches_oracle():
    rng = np.random.default_rng(7)
    # windows cut by both edges, one row or column, a height that is no
    # multiple of the y pass's rows per lane, and the widest kernel (nk = 256)
    for (ny, nx, t) in [(94, 94, 3.1446), (800, 800, 3.16), (33, 517, 0.4), (300, 7, 12.3),
                        (13, 90, 5.0), (1, 40, 2.0), (40, 1, 2.0), (600, 20, 51.0)]:
        img = rng.uniform(size=(ny, nx))
        got = imaging.gaussianBlurImage(img, t)
        want = orc.blur(img, t)
        assert np.max(np.abs(got - want)) <= 1e-15, (ny, nx, t)


def _kat():
    return json.load(open(os.path.join(os.path.dirname(__file__), "golden", "linkcell_kat.json")))


def test_linkcell_kat_gpu():
    """testLib_LinkCell3D.f90: atom 100's neighbours within r = 1, in order."""
    k = _kat()
    pos = np.array(k["positions"])
    lc = imaging.LinkCell3D(np.eye(3), *k["N"])
    lc.add(pos)
    olc = orc.LinkCells(np.eye(3), *k["N"], orc.estimate_max_count(120, 200))
    for i, p in enumerate(pos):
        olc.add(i + 1, p)
    cs, ids, frac = olc.flatten(200)
    slot = int(np.where(ids == 100)[0][0])
    c = int(np.searchsorted(cs, slot, side="right") - 1)
    cz, rem = divmod(c, 24)
    cy, cx = divmod(rem, 6)
    y100 = frac[slot] + np.array([cx, cy, cz])
    (gid, gr2), = lc.neighbourList([y100], 1.0)
    assert list(gid) == [e["id"] for e in k["neigh_r1"]]
    for e, rr in zip(k["neigh_r1"], gr2):
        assert abs(math.sqrt(rr) - e["d"]) <= 5.1e-9


@pytest.mark.parametrize("seed", [11, 12])
def test_linkcell_query_matches_oracle(seed):
    rng = np.random.default_rng(seed)
    N = rng.integers(3, 8, size=3)
    a = np.diag(rng.uniform(0.9, 1.6, size=3))
    a[0, 1] = 0.1                                        # a sheared cell, general 3x3 path
    rc = 0.8
    box = a @ np.diag(N)
    pts = rng.uniform(0, 1, size=(2000, 3)) @ box.T
    q = rng.uniform(0, 1, size=(300, 3)) @ box.T
    lc = imaging.LinkCell3D(a, *N)
    lc.add(pts)
    got = lc.neighbourList(q, rc)
    olc = orc.LinkCells(a, *N, 8)
    for i, p in enumerate(pts):
        olc.add(i + 1, p)
    for qq, (gid, gr2) in zip(q, got):
        oid, or2 = olc.neighbours(qq, rc)
        assert np.array_equal(gid, oid)
        assert np.array_equal(gr2, or2)


def test_empty_and_single_atom():
    cfg = synth.scaled("c0", cells=(6, 6, 6))
    x, box = synth.make_atoms(cfg)
    _, conds = synth.conditions(cfg, x, box)
    im = conds[0][0]
    img = im.computeD2BI_img(np.zeros((0, 3)), mode="fp64")
    assert np.all(img == 0.0)
    img = im.computeD2BI_img(np.zeros((0, 3)), mode="fast")
    assert np.all(img == 0.0)
    # one atom: its periodic images in the link-cell box are points with
    # their own g.r; every voxel within 3 sigma of one image has the phase
    # factor e^{i g.r} of that image (a minimum-image brute force over the
    # link-cell box, not the link cells), every other voxel x = 0
    one = np.array([[9.0, 9.5, 10.0]])
    ph, rho, cnt, r2m = im.field(one, "fp64")
    oc = orc.Condition(im.rec, one)
    assert oc.n_points >= 1
    bx, bcnt, gap = min_image_field(oc, im.rec)
    assert gap > 1e-9
    hit = cnt > 0
    assert np.any(hit)
    assert np.array_equal(cnt, bcnt)
    assert np.max(np.abs(ph - bx)) < 1e-13
    assert np.all(np.abs(np.abs(ph[hit]) - 1.0) < 1e-13)
    assert np.all(ph[~hit] == 0)
    img64 = im.computeD2BI_img(one, mode="fp64")
    ref = oc.image()
    assert np.max(np.abs(img64 - ref)) <= FP64_TOL


@pytest.mark.parametrize("mode", ["fast", "fp64"])
def test_image_device_output(mode):
    """d2bi_gpu_image_dev_out leaves the image in HBM (a DeviceBuffer):
    bit-identical to the host-output entry point, density on and off; a
    host buffer or a wrong shape is refused."""
    import ctypes as C
    from simulated_tem_amd import abi
    for density in (False, True):
        x, im = _cond("c0", cells=(10, 10, 10), sg=0.01, density=density)
        at = imaging.Atoms(x)
        p = im.plan(len(x))
        host = im.computeD2BI_img(at, 0.005, 3, mode=mode)
        out = imaging.DeviceBuffer((p.My, p.Mx))
        got = im.computeD2BI_img(at, 0.005, 3, mode=mode, out=out)
        assert got is out
        assert np.array_equal(out.numpy(), host)
    with pytest.raises(ValueError):
        im.computeD2BI_img(at, 0.0, 1, out=imaging.DeviceBuffer((p.My, p.Mx + 1)))
    hostbuf = np.zeros((p.My, p.Mx))
    rc = abi.load().d2bi_gpu_image_dev_out(C.byref(im.rec), at.h, 0.0, 1, 1,
                                           C.c_void_p(hostbuf.ctypes.data), None)
    assert rc == -1
    # the library still works after the refused call
    assert np.array_equal(im.computeD2BI_img(at, 0.005, 3, mode=mode), host)


@pytest.mark.parametrize("mode", ["fast", "fp64"])
def test_batch_graph_replays_equal_eager(mode):
    """d2bi_gpu_image_batch captures the whole batch into one HIP graph on the
    first call and replays it afterwards.  Every replay, with eager work and
    synchronisations in between, must equal the eager single-condition images
    bit for bit (same kernels, deterministic sums)."""
    x, im1 = _cond("c0", cells=(10, 10, 10))
    _, im2 = _cond("c0", cells=(10, 10, 10), sg=0.015, density=True)
    x3, im3 = _cond("c1", cells=(12, 9, 7))
    at = imaging.Atoms(x)
    at3 = imaging.Atoms(x3)
    conds = [(im1, at, 0.005, 3), (im2, at, 0.005, 1), (im3, at3, 0.004, 6)]
    eager = [c[0].computeD2BI_img(c[1], c[2], c[3], mode=mode) for c in conds]
    for rep in range(4):
        out = imaging.computeD2BI_img_batch(conds, mode=mode)
        for a, b in zip(out, eager):
            assert np.array_equal(a, b), rep
        # eager work between replays (a different condition through the same workspace)
        im2.computeD2BI_img(at3 if rep % 2 else at, 0.005, 2, mode="fast")
    ref = orc.Condition(im3.rec, x3, 0.004, 6).image()
    tol = FP64_TOL if mode == "fp64" else FAST_TOL
    assert np.max(np.abs(out[2] - ref)) <= tol
    # the device-output form: the same images, left in HBM one after another
    dev = imaging.DeviceBuffer((imaging.batch_image_size(conds),))
    for rep in range(2):
        assert imaging.computeD2BI_img_batch(conds, mode=mode, out=dev) is dev
        assert np.array_equal(dev.numpy(), np.concatenate([e.ravel() for e in eager])), rep
    with pytest.raises(ValueError):
        imaging.computeD2BI_img_batch(conds, mode=mode, out=imaging.DeviceBuffer((7,)))
    host = np.zeros(imaging.batch_image_size(conds))
    recs = (imaging.D2biRecord * 3)(*[c[0].rec for c in conds])
    hs = (C.c_void_p * 3)(*[c[1].h for c in conds])
    pa = np.array([c[2] for c in conds])
    npr = np.array([c[3] for c in conds], dtype=np.int32)
    rc = abi.load().d2bi_gpu_image_batch_dev_out(recs, hs, 3, pa.ctypes.data_as(C.POINTER(C.c_double)),
                                                 npr.ctypes.data_as(C.POINTER(C.c_int32)), imaging.MODES[mode],
                                                 C.c_void_p(host.ctypes.data), None)
    assert rc != 0                                   # host memory is refused


@pytest.mark.parametrize("mode", ["fast", "fp64"])
def test_batch_after_reupload_same_count(mode):
    """The next frame of a series: one atom set freed and another of the same
    count uploaded (which may get the freed set's addresses back).  The batch
    must not replay the graph of the first set, whose point counts it holds:
    the new frame's images equal its eager images."""
    x, im = _cond("c0", cells=(10, 10, 10))
    at = imaging.Atoms(x)
    st1 = imaging.D2biStats()
    first = imaging.computeD2BI_img_batch([(im, at, 0.005, 3)], mode=mode, stats=st1)[0]
    assert np.array_equal(first, im.computeD2BI_img(at, 0.005, 3, mode=mode))
    del at
    c = x.mean(axis=0)
    x2 = c + 0.97 * (x - c)                          # same count, fewer points in the cell box
    at2 = imaging.Atoms(x2)
    st2 = imaging.D2biStats()
    second = imaging.computeD2BI_img_batch([(im, at2, 0.005, 3)], mode=mode, stats=st2)[0]
    assert st2.n_points != st1.n_points
    assert np.array_equal(second, im.computeD2BI_img(at2, 0.005, 3, mode=mode))
    assert not np.array_equal(second, first)


@pytest.mark.parametrize("mode", ["fast", "fp64"])
def test_batch_empty_and_sparse_atom_sets(mode):
    """A batch whose atom sets include an empty one, one atom far outside the
    supercell (no point in any link cell) and a single atom: every image
    equals its eager image, and the first two image to zero."""
    x, im = _cond("c0", cells=(6, 6, 6))
    empty = imaging.Atoms(np.zeros((0, 3)))
    far = imaging.Atoms(np.array([[1.0e6, 1.0e6, 1.0e6]]))
    one = imaging.Atoms(np.array([[9.0, 9.5, 10.0]]))
    full = imaging.Atoms(x)
    conds = [(im, empty, 0.005, 3), (im, far, 0.0, 1), (im, one, 0.0, 1), (im, full, 0.005, 2),
             (im, empty, 0.0, 9)]
    eager = [c[0].computeD2BI_img(c[1], c[2], c[3], mode=mode) for c in conds]
    for rep in range(2):
        out = imaging.computeD2BI_img_batch(conds, mode=mode)
        for a, b in zip(out, eager):
            assert np.array_equal(a, b), rep
    for i in (0, 1, 4):
        assert np.all(out[i] == 0.0)
    assert np.any(out[2] != 0.0) and np.any(out[3] != 0.0)


def test_batch_branches_fit_free_memory(monkeypatch):
    """Each concurrent branch of a batch holds a workspace sized for the
    largest condition (branch_bytes); extra branches run only where what the
    workspaces add, branch 0's own included, fits in 85 % of the free device
    memory.  From an empty library (no workspace kept): no free memory gives
    one branch, free memory for 1.5 workspaces one branch, for 2.5 two, and
    plenty four; the images are the same bit for bit on any count."""
    x, im1 = _cond("c0", cells=(10, 10, 10))
    _, im2 = _cond("c0", cells=(10, 10, 10), sg=0.015, density=True)
    x3, im3 = _cond("c1", cells=(12, 9, 7))

    monkeypatch.setenv("D2BI_TUNE", "0")             # every branch that fits runs (no timing)

    def run(mem):
        imaging.finalize()                           # drop every workspace
        if mem is None:
            monkeypatch.delenv("D2BI_BATCH_MEM", raising=False)
        else:
            monkeypatch.setenv("D2BI_BATCH_MEM", str(int(mem)))
        at, at3 = imaging.Atoms(x), imaging.Atoms(x3)
        conds = [(im1, at, 0.005, 3), (im2, at, 0.005, 1), (im3, at3, 0.004, 6), (im1, at3, 0.0, 1)]
        st = imaging.D2biStats()
        out = imaging.computeD2BI_img_batch(conds, mode="fast", stats=st)
        return out, st.n_branches, st.branch_bytes

    try:
        ref, nb, W = run(1e15)
        assert nb == 4 and W > 0
        for mem, want in [(0, 1), (1.5 * W / 0.85, 1), (2.5 * W / 0.85, 2), (1e15, 4)]:
            out, nb, W2 = run(mem)
            print("[branches] free %.3e -> %d branches (workspace %d B)" % (mem, nb, W2), flush=True)
            assert W2 == W
            assert nb == want, (mem, nb)
            for a, b in zip(out, ref):
                assert np.array_equal(a, b), mem
    finally:
        monkeypatch.delenv("D2BI_BATCH_MEM", raising=False)
        monkeypatch.delenv("D2BI_TUNE", raising=False)
        imaging.finalize()


def test_uneven_batch_c4_style():
    """A small uneven batch like c4: mixed M (3, 4, 6), zone axes 001 and
    01-1, g in {110, 1-10, 200, 211}, 1-8 precession points, density on and
    off, all in one graph; each image against the oracle."""
    cfgs = [c for c in synth.c4_configs(seed=77, n=40, max_sites=3.0e4)][:6]
    data = synth.c4_conditions(cfgs)
    batch, refs = [], []
    for x, conds in data:
        at = imaging.Atoms(x)
        for im, pa, npr in conds:
            batch.append((im, at, pa, npr))
            refs.append(orc.Condition(im.rec, x, pa, npr).image(nthreads=NTH))
    assert len(set(c["M"] for c in cfgs)) >= 2
    out = imaging.computeD2BI_img_batch(batch, mode="fast")
    out2 = imaging.computeD2BI_img_batch(batch, mode="fast")          # graph replay
    for a, b, r in zip(out, out2, refs):
        assert np.array_equal(a, b)
        assert np.max(np.abs(a - r)) <= FAST_TOL
    out64 = imaging.computeD2BI_img_batch(batch, mode="fp64")
    for a, r in zip(out64, refs):
        assert np.max(np.abs(a - r)) <= FP64_TOL


def test_golden_w_loop_image(golden_dir):
    """Data/test/...ng1.00.dat via -U <header RU> -R I -d <header extents>: the
    header holds RU to 6 decimals and the extents to 3, so the reproduction
    is approximate; testD2BI.sh's min/max/avg gate is +-1/256."""
    w = np.load(os.path.join(golden_dir, "w_loop_golden.npz"))
    a0 = lattice_a0(3.1652)
    box = w["box"][:, 1] - w["box"][:, 0]
    sa, sn = input_supercell(np.diag(box), 4, a0)
    im = DynamicalTwoBeamImaging(sa, sn, a0, (0, 0, 1), (2, 0, 0), 200.0, 4, 4)
    im.setConventionalUnitCell(a0 * np.eye(3))
    im.setVoxelSupercellSize(w["space"])
    im.setSupercellTilt(w["rotation_cm"].reshape(3, 3).T, np.eye(3))
    im.setExtinctionDistances(*w["xi"])
    im.setUseDensityField(False)
    x = w["x"]
    g64 = im.computeD2BI_img(x, mode="fp64")
    gf = im.computeD2BI_img(x, mode="fast")
    ref = orc.Condition(im.rec, x).image(nthreads=NTH)
    assert np.max(np.abs(g64 - ref)) <= FP64_TOL
    assert np.max(np.abs(gf - ref)) <= FAST_TOL
    gold = w["image"]
    d = np.abs(g64 - gold)
    print("[golden] max|d| %.3e mean|d| %.3e  min/max/avg %.6f %.6f %.6f"
          % (d.max(), d.mean(), g64.min(), g64.max(), g64.mean()), flush=True)
    for got, want in zip((g64.min(), g64.max(), g64.mean()), (0.000028, 0.230127, 0.081501)):
        assert abs(got - want) <= 0.0039


SHEAR = np.array([[1.0, 0.04, 0.02], [0.0, 1.0, 0.03], [0.0, 0.0, 1.0]])


def test_triclinic_supercell_w_loop(golden_dir):
    """A sheared (triclinic) periodic supercell: the W loop's atoms and box
    under x -> S x, S upper unitriangular (LAMMPS tilt factors xy, xz, yz =
    0.04 ly, 0.02 lz, 0.03 lz).  findGdotR's 27 replicas are translations
    along the sheared vectors (getSuperVector, :904-918); FP64 and FAST
    against the oracle."""
    w = np.load(os.path.join(golden_dir, "w_loop_golden.npz"))
    a0 = lattice_a0(3.1652)
    box = w["box"][:, 1] - w["box"][:, 0]
    sa, sn = input_supercell(SHEAR @ np.diag(box), 4, a0)
    assert sa[0, 1] != 0 and sa[0, 2] != 0 and sa[1, 2] != 0
    im = DynamicalTwoBeamImaging(sa, sn, a0, (0, 0, 1), (2, 0, 0), 200.0, 4, 4)
    im.setConventionalUnitCell(a0 * np.eye(3))
    im.setVoxelSupercellSize(w["space"])
    im.setSupercellTilt(w["rotation_cm"].reshape(3, 3).T, np.eye(3))
    im.setExtinctionDistances(*w["xi"])
    im.setUseDensityField(False)
    x = w["x"] @ SHEAR.T
    g64 = im.computeD2BI_img(x, mode="fp64")
    gf = im.computeD2BI_img(x, mode="fast")
    ref = orc.Condition(im.rec, x).image(nthreads=NTH)
    e64, ef = np.max(np.abs(g64 - ref)), np.max(np.abs(gf - ref))
    print("[triclinic] fp64 %.3e fast %.3e  I in [%.4f, %.4f]" % (e64, ef, ref.min(), ref.max()), flush=True)
    assert e64 <= FP64_TOL and ef <= FAST_TOL
    # the shear changes the image (the replicas move)
    plain = DynamicalTwoBeamImaging(*input_supercell(np.diag(box), 4, a0), a0, (0, 0, 1), (2, 0, 0), 200.0, 4, 4)
    plain.setConventionalUnitCell(a0 * np.eye(3))
    plain.setVoxelSupercellSize(w["space"])
    plain.setSupercellTilt(w["rotation_cm"].reshape(3, 3).T, np.eye(3))
    plain.setExtinctionDistances(*w["xi"])
    plain.setUseDensityField(False)
    assert np.max(np.abs(plain.computeD2BI_img(w["x"], mode="fp64") - g64)) > 1e-3


def test_c2_full_size_density_parity():
    """c2 at full size: 640x640x1600 voxels, 20.5e6 atoms, g = 200, density
    field on, crystal coordinates up to ~1.4e3 A (g.r up to ~5e3 rad).  All
    three s_g conditions on the GPU in FP64 and FAST: FAST vs FP64 over the
    whole image; FP64 and FAST against oracle windows (whole crystal binned)
    for s_g = -0.02."""
    print("\n[c2 full] generating 20.5M atoms", flush=True)
    cfg = dict(synth.CONFIGS["c2"])
    x, box = synth.make_atoms(cfg)
    x, conds = synth.conditions(cfg, x, box)
    assert len(conds) == 3 and conds[0][0].rec.use_density == 1
    at = imaging.Atoms(x)
    for k, (im, pa, npr) in enumerate(conds):
        g64 = im.computeD2BI_img(at, pa, npr, mode="fp64")
        gf = im.computeD2BI_img(at, pa, npr, mode="fast")
        assert g64.shape == (640, 640)
        e = np.abs(gf - g64)
        print("[c2 full] s_g %+.3f fast vs fp64: max %.3e mean %.3e" % (im.getDeviationParameter(), e.max(),
                                                                        e.mean()), flush=True)
        assert e.max() <= FAST_TOL
        assert np.all(g64 >= 0) and np.all(g64 <= 1)
        if k == 0:
            print("[c2 full] oracle binning", flush=True)
            oc = orc.Condition(im.rec, x, pa, npr)
            for (ix0, iy0) in [(0, 0), (316, 320), (632, 630)]:
                ref = oc.image(window=(ix0, ix0 + 8, iy0, iy0 + 10), nthreads=NTH)
                e64 = np.max(np.abs(g64[iy0:iy0 + 10, ix0:ix0 + 8] - ref))
                ef = np.max(np.abs(gf[iy0:iy0 + 10, ix0:ix0 + 8] - ref))
                print("[c2 full] window (%d,%d) fp64 %.3e fast %.3e" % (ix0, iy0, e64, ef), flush=True)
                assert e64 <= FP64_TOL
                assert ef <= FAST_TOL
    del at


@pytest.fixture(scope="module")
def c1_full():
    print("\n[c1 full] generating 8M atoms", flush=True)
    cfg = dict(synth.CONFIGS["c1"])
    x, box = synth.make_atoms(cfg)
    x, conds = synth.conditions(cfg, x, box)
    im = conds[0][0]
    at = imaging.Atoms(x)
    return x, im, at


def test_c1_full_size_window_parity(c1_full):
    """Full 800x800x400 image on the GPU; the oracle recomputes a window of
    columns (with the whole crystal binned) for the 1e-10 / 1e-4 bars."""
    x, im, at = c1_full
    g64 = im.computeD2BI_img(at, mode="fp64")
    gf = im.computeD2BI_img(at, mode="fast")
    assert g64.shape == (800, 800)
    print("[c1 full] oracle binning", flush=True)
    oc = orc.Condition(im.rec, x)
    for (ix0, iy0) in [(0, 0), (392, 397), (792, 790)]:
        ref = oc.image(window=(ix0, ix0 + 8, iy0, iy0 + 10), nthreads=NTH)
        e64 = np.max(np.abs(g64[iy0:iy0 + 10, ix0:ix0 + 8] - ref))
        ef = np.max(np.abs(gf[iy0:iy0 + 10, ix0:ix0 + 8] - ref))
        print("[c1 full] window (%d,%d) fp64 %.3e fast %.3e" % (ix0, iy0, e64, ef), flush=True)
        assert e64 <= FP64_TOL
        assert ef <= FAST_TOL
    # full-image error gate FAST vs FP64 (size-independent property)
    e = np.abs(gf - g64)
    print("[c1 full] fast vs fp64: max %.3e mean %.3e" % (e.max(), e.mean()), flush=True)
    assert e.max() <= FAST_TOL
    assert np.all(g64 >= 0) and np.all(g64 <= 1)


_IMAGE_IN_CHILD = r"""
import sys
import numpy as np
sys.path.insert(0, sys.argv[1])
from simulated_tem_amd import imaging, synth
imaging.init(0)
cfg = synth.scaled("c0", (10, 9, 8), (26.1, 24.7, 22.9))
cfg.update(sg=[0.01], density=True, g=[(1, -1, 0)], k=(1, 1, 1), n_tomo=1)
x, box = synth.make_atoms(cfg)
x, conds = synth.conditions(cfg, x, box)
im = conds[0][0]
np.save(sys.argv[2], im.computeD2BI_img(x, 0.005, 9, mode="fast"))
"""


def test_fast_images_identical_across_processes(tmp_path):
    """A FAST image does not depend on the process that computes it: two
    fresh processes image the same condition bit for bit alike (the voxel
    kernel variant and the integrator are fixed choices, the sums have a
    fixed order)."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    outs = []
    for i in range(2):
        out = tmp_path / ("img%d.npy" % i)
        env = dict(os.environ)
        env.pop("D2BI_FAST_VARIANT", None)
        env.pop("D2BI_FAST_TUNE", None)
        print("[process] %d" % i, flush=True)
        r = subprocess.run([sys.executable, "-c", _IMAGE_IN_CHILD, root, str(out)], env=env,
                           capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout + r.stderr
        outs.append(np.load(out))
    assert outs[0].size > 0 and np.all(np.isfinite(outs[0]))
    assert np.array_equal(outs[0], outs[1])
