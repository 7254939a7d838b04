"""The launch tuner leaves every image bit-identical.

The tuner times launch choices that no thread's arithmetic depends on: the
FAST integrators' workgroup size (64, 128, 256; a lane integrates one pixel
column alone), the FAST voxel kernel's tile order over the workgroups (per-XCD
ranges, linear, or rows dealt to the XCDs in turn; a wave sums one tile alone)
and a batch graph's concurrent
branch count (1..4; a condition's kernels do not depend on its branch).  Every
forced choice, the tuned choice, the untuned defaults and a choice read back
from a tuning cache file must give the same images bit for bit, eager and
batched.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from simulated_tem_amd import imaging, synth

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TUNE_VARS = ("D2BI_TUNE", "D2BI_TUNE_CACHE", "D2BI_INT_BLOCK", "D2BI_TILE_ORDER", "D2BI_BATCH_BRANCHES",
             "D2BI_FAST_VARIANT", "D2BI_FAST_TUNE")


def _conds():
    """Three FAST conditions: passes of 32 + 8 and of 8 + 1 precession
    points, a density field, and a ragged 33 x 31 x 29 voxel grid."""
    out = []
    for name, cells, d, upd, pa, npr in [
            ("c1", (22, 18, 12), None, {}, 0.005, 40),
            ("c0", (10, 10, 10), None, {"sg": [0.015], "density": True}, 0.004, 9),
            ("c0", (10, 9, 8), (26.1, 24.7, 22.9), {"g": [(1, -1, 0)], "k": (1, 1, 1)}, 0.005, 3)]:
        cfg = synth.scaled(name, cells, d)
        cfg.update(upd)
        cfg["n_tomo"] = 1
        x, box = synth.make_atoms(cfg)
        x, cs = synth.conditions(cfg, x, box)
        out.append((cs[0][0], imaging.Atoms(x), pa, npr))
    return out


def _images(conds):
    eager = [im.computeD2BI_img(at, pa, npr, mode="fast") for (im, at, pa, npr) in conds]
    batch = imaging.computeD2BI_img_batch(conds, mode="fast")
    return eager, batch


@pytest.fixture
def clean_env(monkeypatch):
    for v in TUNE_VARS:
        monkeypatch.delenv(v, raising=False)
    return monkeypatch


def test_forced_launch_choices_bit_identical(clean_env):
    mp = clean_env
    conds = _conds()
    mp.setenv("D2BI_TUNE", "0")
    ref_e, ref_b = _images(conds)
    for a, b in zip(ref_e, ref_b):
        assert np.array_equal(a, b)
    assert all(np.all(np.isfinite(a)) and np.any(a != 0) for a in ref_e)
    for block in ("64", "128", "256"):
        for order in ("0", "1", "2"):
            mp.setenv("D2BI_INT_BLOCK", block)
            mp.setenv("D2BI_TILE_ORDER", order)
            print("\n[forced] block %s order %s" % (block, order), flush=True)
            e, b = _images(conds)
            for i in range(len(conds)):
                assert np.array_equal(e[i], ref_e[i]), (block, order, i)
                assert np.array_equal(b[i], ref_e[i]), (block, order, i)
    mp.delenv("D2BI_INT_BLOCK")
    mp.delenv("D2BI_TILE_ORDER")
    for nb in ("1", "2", "3", "4"):
        mp.setenv("D2BI_BATCH_BRANCHES", nb)
        st = imaging.D2biStats()
        b = imaging.computeD2BI_img_batch(conds, mode="fast", stats=st)
        print("[forced] %s branches -> %d" % (nb, st.n_branches), flush=True)
        assert st.n_branches == min(int(nb), len(conds))
        for i in range(len(conds)):
            assert np.array_equal(b[i], ref_e[i]), (nb, i)
    # tuning on (this process may already hold tuned choices for these shapes)
    mp.delenv("D2BI_BATCH_BRANCHES")
    mp.delenv("D2BI_TUNE")
    e, b = _images(conds)
    for i in range(len(conds)):
        assert np.array_equal(e[i], ref_e[i]) and np.array_equal(b[i], ref_e[i]), i


_CHILD = r"""
import sys
import numpy as np
sys.path.insert(0, sys.argv[1])
from simulated_tem_amd import imaging
from tests.test_gpu_tune import _conds, _images
imaging.init(0)
e, b = _images(_conds())
np.savez(sys.argv[2], *(e + b))
"""


def _child(out, env):
    r = subprocess.run([sys.executable, "-c", _CHILD, ROOT, str(out)], env=env, cwd=ROOT, capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    with np.load(out) as z:
        return [z["arr_%d" % i] for i in range(len(z.files))]


def _cache_lines(path):
    with open(path) as f:
        return [tuple(int(v) for v in line.split()) for line in f if line.strip()]


def test_tuning_cache_file_and_fresh_processes(tmp_path):
    """A fresh process tunes every shape it meets and appends the choices
    to D2BI_TUNE_CACHE; a second process reads them back (adding nothing);
    a third runs untuned.  All three image alike, bit for bit."""
    cache = tmp_path / "tune.txt"
    env = {k: v for k, v in os.environ.items() if k not in TUNE_VARS}
    env["D2BI_TUNE_CACHE"] = str(cache)
    print("\n[process] tuning", flush=True)
    tuned = _child(tmp_path / "a.npz", env)
    lines = _cache_lines(cache)
    kinds = sorted(set(ln[0] for ln in lines))
    print("[process] %d cache lines, kinds %s" % (len(lines), kinds), flush=True)
    assert all(len(ln) == 10 for ln in lines)
    assert kinds == [1, 2, 3]                    # workgroup size, tile order, branches
    for ln in lines:
        if ln[0] == 1:
            assert ln[9] in (64, 128, 256)
        elif ln[0] == 2:
            assert ln[9] in (0, 1, 2)
        else:
            assert 1 <= ln[9] <= 3               # three conditions
    print("[process] from the cache", flush=True)
    cached = _child(tmp_path / "b.npz", env)
    assert _cache_lines(cache) == lines          # nothing timed again
    env.pop("D2BI_TUNE_CACHE")
    env["D2BI_TUNE"] = "0"
    print("[process] untuned", flush=True)
    plain = _child(tmp_path / "c.npz", env)
    assert len(tuned) == 6
    for a, b, c in zip(tuned, cached, plain):
        assert np.array_equal(a, c) and np.array_equal(b, c)
    for i in range(3):
        assert np.array_equal(tuned[i], tuned[3 + i])      # eager == batch
