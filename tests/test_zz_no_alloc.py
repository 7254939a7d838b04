"""d2bi_gpu_image_noalloc (the entry point that allocates no device memory:
its workspace is a static block per device) against d2bi_gpu_image_dev_out,
the entry point it is a variant of, through the C-ABI: images equal bit for
bit on two calls in a row, the second with a smaller condition; device
memory unchanged across the calls; a condition beyond the block refused
with D2BI_ERR_CAPACITY, the library working afterwards."""
import ctypes as C

import numpy as np
import pytest

from simulated_tem_amd import abi, imaging, synth

pytestmark = pytest.mark.gpu


def _cond(cells, sg=0.01, density=False):
    cfg = synth.scaled("c0", cells=cells)
    cfg.update(sg=[sg], density=density)
    x, box = synth.make_atoms(cfg)
    x, conds = synth.conditions(cfg, x, box)
    return x, conds[0][0]


def _free_bytes():
    hip = imaging._hip_runtime()
    f, t = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(f), C.byref(t)) == 0
    return f.value


def _noalloc(im, at, pa, npr, mode, out):
    st = imaging.D2biStats()
    rc = abi.load().d2bi_gpu_image_noalloc(C.byref(im.rec), at.h, float(pa), int(npr), imaging.MODES[mode],
                                           C.c_void_p(out.ptr), C.byref(st))
    return rc, st


@pytest.mark.parametrize("mode", ["fast", "fp64"])
def test_noalloc_equals_dev_out_two_calls(mode):
    # c0's size (80^3 voxels, 16,000 sites) with density, then a smaller one
    cases = [(_cond((20, 20, 20), density=True), 0.005, 3), (_cond((8, 7, 6), sg=0.0), 0.0, 1)]
    runs = []
    for (x, im), pa, npr in cases:
        at = imaging.Atoms(x)
        p = im.plan(len(x))
        ref = imaging.DeviceBuffer((p.My, p.Mx))
        im.computeD2BI_img(at, pa, npr, mode=mode, out=ref)
        runs.append((im, at, pa, npr, ref.numpy(), imaging.DeviceBuffer((p.My, p.Mx))))
    free0 = _free_bytes()
    for rep in range(2):
        for im, at, pa, npr, want, out in runs:
            print("[noalloc] %s rep %d %dx%d P=%d" % (mode, rep, want.shape[1], want.shape[0], npr), flush=True)
            rc, st = _noalloc(im, at, pa, npr, mode, out)
            assert rc == 0, abi.load().d2bi_last_error()
            got = out.numpy()
            assert np.array_equal(got, want)
            assert np.all(np.isfinite(got)) and np.any(got != 0)
            assert st.n_points > 0
    free1 = _free_bytes()
    print("[noalloc] free device bytes before %d after %d" % (free0, free1), flush=True)
    assert free1 == free0


def test_noalloc_refuses_beyond_scratch_then_works():
    x, im = _cond((48, 48, 48))                      # 192^3 voxels: 113 MB of FAST field
    at = imaging.Atoms(x)
    p = im.plan(len(x))
    out = imaging.DeviceBuffer((p.My, p.Mx))
    rc, _ = _noalloc(im, at, 0.0, 1, "fast", out)
    assert rc == -3                                  # D2BI_ERR_CAPACITY
    assert b"scratch" in abi.load().d2bi_last_error()
    xs, ims = _cond((8, 7, 6))
    ats = imaging.Atoms(xs)
    ps = ims.plan(len(xs))
    ref = imaging.DeviceBuffer((ps.My, ps.Mx))
    ims.computeD2BI_img(ats, 0.0, 1, mode="fast", out=ref)
    outs = imaging.DeviceBuffer((ps.My, ps.Mx))
    rc, _ = _noalloc(ims, ats, 0.0, 1, "fast", outs)
    assert rc == 0, abi.load().d2bi_last_error()
    assert np.array_equal(outs.numpy(), ref.numpy())
    # a host pointer for the image is refused
    host = np.zeros((ps.My, ps.Mx))
    st = imaging.D2biStats()
    rc = abi.load().d2bi_gpu_image_noalloc(C.byref(ims.rec), ats.h, 0.0, 1, 1, C.c_void_p(host.ctypes.data),
                                           C.byref(st))
    assert rc == -1
