"""Imaging throughput benchmark: imaged voxels/s (voxelise + integrate),
batched FP32 against plain per-condition FP64.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c3]

Workload c3 (the default, the metric's batched configuration): the c1
crystal (bcc, 200 x 200 x 100 cells, 8.0e6 atoms, Eshelby centres, voids,
-zpad) imaged as a tomography series of 21 tilts (+-20 deg) x 2 g-vectors
(110, 011) on a [001] zone = 42 conditions, each re-binned and re-voxelised
(800 x 800 x 400 voxels) and integrated at 32 precession points (0.005 rad):
1.344e3 integrations, 3.44e11 imaged voxels per batch.  A step is one whole
batch through d2bi_gpu_image_batch_dev_out (one captured HIP graph; atoms
resident in HBM, the 42 images left in HBM), the window of the reference's
T_D2BI timer (d2bi.f90:1392-1410) summed over the series.  The first call
plans, counts and captures the graph; it is timed apart (first_call_s).

Beside the timed batch:
  fp64_value       the plain per-condition FP64 path (d2bi_gpu_image_dev_out,
                   FP64 mode, one condition per call) on conditions spread
                   over the series; fast_over_fp64 = value / fp64_value
  errors_vs_fp64   max and mean |I_FAST - I_FP64| per checked image
  roofline         k_voxel_tile (the largest kernel of the batch), FP32 VALU:
                   algorithmic work 388 flop per voxel (28.27 in-sphere atoms
                   x 13 flop + 20) over its device time, measured with HIP
                   events around its launches on the library stream while
                   imaging conditions of the series one at a time (the same
                   kernels and launch shapes as in the graph)
  roofline_integrate
                   k_integrate_poly<32> by its real operation count: per RK4
                   double step and column, 399 flop to form the step matrix
                   and 76 flop per precession point (Horner at the point's
                   alpha, one 2x2 complex product); the 78.5 flop per voxel
                   and condition of the stage-by-stage march would overstate
                   it
  cpu_baseline     the oracle (FP64 C restatement, OpenMP over pixel rows) on
                   one condition of the series: the whole crystal binned and a
                   window of columns integrated at all 32 precession points,
                   extrapolated to the full image
  single           one c1 condition at P = 1 (the unbatched path)

Single-condition workloads (c0, c1, c2) time one condition per step instead.
Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import math
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "imaged voxels/s (voxelise+integrate), batched FP32 vs plain FP64, err<=1e-4"
PEAK_FP32_TFLOPS = 157.3             # MI355X published FP32 vector/matrix peak
PEAK_TRANS_PER_S = 157.3e12 / 16     # transcendentals at a quarter of the wave64 VALU rate
FLOPS_PER_VOXEL = 388.0              # voxelisation, per voxel and field
TRANS_PER_VOXEL = 30.0
POLY_FORM_FLOPS = 399.0              # k_integrate_poly: step matrix, per double step and column
POLY_POINT_FLOPS = 76.0              # k_integrate_poly: per precession point, double step and column
STAGE_POINT_FLOPS = 140.0            # k_integrate_fast: per precession point, double step and column
# HBM bytes per launch of k_voxel_tile from rocprofv3 FETCH_SIZE / WRITE_SIZE
# passes, scaled by a 1 GiB copy's counts, with the library's default launch
# choices (profiles/k_voxel_tile_c1_traffic.md, profiles/c3_batch_pmc.md); a
# profiler cannot run inside the timed process, so the figure is recorded per
# workload
TRAFFIC_BYTES = {"c1": 4.557e9, "c3": 4.104e9}
SINGLE = ("c0", "c1", "c2")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def integrate_flops(p, n_prec):
    """Real flops of the FAST integration of one condition: passes of 32, 16
    or 8 points through the shared step matrix, fewer stage by stage."""
    cols = p.Mx * p.My
    ds = p.Mz // 2
    left, f = n_prec, 0.0
    while left > 0:
        np_ = 32 if left >= 32 else 16 if left >= 16 else 8 if left >= 8 else 4 if left >= 4 else \
            2 if left >= 2 else 1
        f += (POLY_FORM_FLOPS + POLY_POINT_FLOPS * np_) if np_ >= 8 else STAGE_POINT_FLOPS * np_
        left -= np_
    return cols * ds * f


def host_info():
    model = platform.processor() or ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    return model, os.cpu_count() or 1, affinity


def cpu_threads():
    """Threads for the CPU baseline: every core this process may run on,
    capped by OMP_NUM_THREADS when the environment sets one (a GPU host
    grants one GPU's share of its cores)."""
    _, nproc, affinity = host_info()
    n = affinity
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def cpu_baseline(im, x, prec_angle, n_prec, sample_cols, nthreads):
    from oracle import oracle as orc
    t0 = time.perf_counter()
    oc = orc.Condition(im.rec, x, prec_angle, n_prec)
    t_bin = time.perf_counter() - t0
    p = oc.plan
    w = sample_cols
    ix0, iy0 = p.Mx // 2 - w // 2, p.My // 2 - w // 2
    t0 = time.perf_counter()
    oc.image(window=(ix0, ix0 + w, iy0, iy0 + w), nthreads=nthreads)
    t_win = time.perf_counter() - t0
    v_full = p.Mx * p.My * p.Mz
    t_full = t_bin + t_win * (p.Mx * p.My) / (w * w)
    model, nproc, affinity = host_info()
    return {"value": v_full * n_prec / t_full, "unit": "imaged voxels/s", "cores": nthreads, "kind": "port",
            "nproc": nproc, "affinity_cores": affinity, "cpu_model": model,
            "sample": "oracle FP64, one condition at P=%d: whole crystal binned (%.1f s) + %dx%d columns x %d "
                      "slices (%.2f s, %d threads), extrapolated to %dx%dx%d" %
                      (n_prec, t_bin, w, w, p.Mz, t_win, nthreads, p.Mx, p.My, p.Mz)}


def kernel_times(imaging, flat, idx, reps=2):
    """Device time of the voxel kernel and of the integration kernels per
    condition (HIP events on the library stream), imaging conditions idx one
    at a time in FAST mode; one untimed pass first."""
    from simulated_tem_amd.abi import D2biStats
    out = {}
    for rep in range(reps + 1):
        for i in idx:
            im, at, pa, npr = flat[i][:4]
            p = im.plan(flat[i][4], pa, npr)
            d = imaging.DeviceBuffer((p.My, p.Mx))
            st = D2biStats()
            im.computeD2BI_img(at, pa, npr, mode="fast", stats=st, out=d)
            if rep > 0:
                out.setdefault(i, []).append((st.ms_voxel, st.ms_integrate, st.ms_bin, st.ms_total))
    return out


def load_workload(name, imaging, synth):
    """[(imaging, Atoms, precAngle, nPrec, nAtoms)] and, per entry, its host atoms."""
    if name == "c4":
        data = synth.c4_conditions(synth.c4_configs())
        flat, xs = [], []
        for x, conds in data:
            at = imaging.Atoms(x)
            flat.extend((im, at, pa, npr, len(x)) for (im, pa, npr) in conds)
            xs.extend(x for _ in conds)
        return flat, xs
    cfg = dict(synth.CONFIGS[name])
    x, box = synth.make_atoms(cfg)
    x, conds = synth.conditions(cfg, x, box)
    at = imaging.Atoms(x)
    return [(im, at, pa, npr, len(x)) for (im, pa, npr) in conds], [x] * len(conds)


def batch_main(args, imaging, synth, np):
    log("[bench] generating workload", args.workload)
    flat, xs = load_workload(args.workload, imaging, synth)
    batch = [f[:4] for f in flat]
    imaged = 0
    nvox_total = 0
    for im, at, pa, npr, n in flat:
        p = im.plan(n, pa, npr)
        imaged += p.Mx * p.My * p.Mz * npr
        nvox_total += p.Mx * p.My * p.Mz
    log("[bench] %d conditions, %.3e imaged voxels per batch" % (len(batch), imaged))
    dev = imaging.DeviceBuffer((imaging.batch_image_size(batch),))
    st = imaging.D2biStats()
    t0 = time.perf_counter()
    imaging.computeD2BI_img_batch(batch, mode="fast", out=dev, stats=st)    # plans, counts, captures
    first = time.perf_counter() - t0
    log("[bench] first call (plan, count, capture, run) %.3f s, %d branches" % (first, st.n_branches))
    for _ in range(args.warmup):
        imaging.computeD2BI_img_batch(batch, mode="fast", out=dev)
    log("[bench] warmup done")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        imaging.computeD2BI_img_batch(batch, mode="fast", out=dev)
    t1 = time.perf_counter()
    s_per_batch = (t1 - t0) / args.steps
    value = imaged / s_per_batch
    log("[bench] batch fast: %.4f s/batch, %.3e imaged voxels/s" % (s_per_batch, value))
    fast_all = dev.numpy()
    # the same batch with the images copied to host memory (PCIe included)
    t0 = time.perf_counter()
    imgs = imaging.computeD2BI_img_batch(batch, mode="fast")
    s_host = time.perf_counter() - t0
    assert np.array_equal(fast_all, np.concatenate([i.ravel() for i in imgs]))
    log("[bench] batch fast, images copied to host: %.4f s/batch" % s_host)

    # plain per-condition FP64 on conditions spread over the series
    nchk = max(1, min(args.fp64_conditions, len(flat)))
    idx = sorted(set(np.linspace(0, len(flat) - 1, nchk).round().astype(int).tolist()))
    errs, imaged64, t64 = [], 0, 0.0
    for i in idx:
        im, at, pa, npr, n = flat[i]
        p = im.plan(n, pa, npr)
        d64 = imaging.DeviceBuffer((p.My, p.Mx))
        im.computeD2BI_img(at, pa, npr, mode="fp64", out=d64)            # untimed first pass
        t0 = time.perf_counter()
        im.computeD2BI_img(at, pa, npr, mode="fp64", out=d64)
        t64 += time.perf_counter() - t0
        imaged64 += p.Mx * p.My * p.Mz * npr
        e = np.abs(imgs[i] - d64.numpy())
        errs.append({"condition": int(i), "g": [int(v) for v in im.rec.g], "max": float(e.max()),
                     "mean": float(e.mean())})
        log("[bench] condition %d g=%s fast vs fp64: max %.3e mean %.3e" % (i, errs[-1]["g"], e.max(), e.mean()))
    fp64_value = imaged64 / t64
    log("[bench] plain fp64: %.3e imaged voxels/s (%d conditions), fast/fp64 %.2f"
        % (fp64_value, len(idx), value / fp64_value))

    # live kernel times of the batch's kernels (eager, same launches)
    kt = kernel_times(imaging, flat, idx[:3])
    ms_vox = float(np.mean([v[0] for vs in kt.values() for v in vs]))
    ms_int = float(np.mean([v[1] for vs in kt.values() for v in vs]))
    ms_bin = float(np.mean([v[2] for vs in kt.values() for v in vs]))
    im0, _, pa0, npr0, n0 = flat[idx[0]]
    p0 = im0.plan(n0, pa0, npr0)
    vox0 = p0.Mx * p0.My * p0.Mz
    ach = vox0 * FLOPS_PER_VOXEL / (ms_vox * 1e-3) / 1e12
    roof = {"bound": "valu", "pipe": "valu_fp32", "kernel": "k_voxel_tile<false,false,1,8,4,1>",
            "achieved": ach, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": ach / PEAK_FP32_TFLOPS,
            "traffic": TRAFFIC_BYTES.get(args.workload),
            "traffic_unit": "B per launch (rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, default launch choices)",
            "ms_per_launch": ms_vox, "work": "%d voxels x %g flop" % (vox0, FLOPS_PER_VOXEL),
            "transcendental_frac": vox0 * TRANS_PER_VOXEL / (ms_vox * 1e-3) / PEAK_TRANS_PER_S}
    fi = integrate_flops(p0, npr0)
    ach_i = fi / (ms_int * 1e-3) / 1e12
    roof_i = {"bound": "valu", "pipe": "valu_fp32", "kernel": "k_integrate_poly<%d,false>" % min(32, npr0),
              "achieved": ach_i, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": ach_i / PEAK_FP32_TFLOPS,
              "ms_per_launch": ms_int,
              "work": "%d columns x %d double steps x (%g + %g x %d) flop" %
                      (p0.Mx * p0.My, p0.Mz // 2, POLY_FORM_FLOPS, POLY_POINT_FLOPS, min(32, npr0))}
    log("[bench] per condition: voxel %.3f ms (%.3f of FP32 peak), integrate %.3f ms (%.3f), binning %.3f ms"
        % (ms_vox, roof["frac"], ms_int, roof_i["frac"], ms_bin))
    emax = max(e["max"] for e in errs)
    out = {
        "metric": METRIC, "value": value, "unit": "imaged voxels/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": s_per_batch * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded bcc crystal, Eshelby centres, voids)",
        "config": {"workload": args.workload, "conditions": len(batch), "atoms": int(n0),
                   "voxels_per_condition": [p0.Mx, p0.My, p0.Mz], "n_prec": int(npr0),
                   "imaged_voxels_per_batch": int(imaged)},
        "first_call_s": first, "branches": int(st.n_branches),
        "value_image_to_host": imaged / s_host,
        "fp64_value": fp64_value, "fp64_conditions": [e["condition"] for e in errs],
        "fast_over_fp64": value / fp64_value,
        "max_abs_err_vs_fp64": emax, "mean_abs_err_vs_fp64": float(np.mean([e["mean"] for e in errs])),
        "errors_vs_fp64": errs,
        "ms_bin": ms_bin, "ms_voxel": ms_vox, "ms_integrate": ms_int,
        "roofline": roof, "roofline_integrate": roof_i,
    }
    return out, flat, xs, idx


def single_main(name, steps, warmup, imaging, synth, np, fp64_steps=3):
    """One condition of a single-condition workload per step (c0, c1, c2)."""
    from simulated_tem_amd.abi import D2biStats
    cfg = dict(synth.CONFIGS[name])
    x, box = synth.make_atoms(cfg)
    x, conds = synth.conditions(cfg, x, box)
    im, pa, npr = conds[0]
    p = im.plan(len(x), pa, npr)
    vox = p.Mx * p.My * p.Mz
    imaged = vox * npr
    at = imaging.Atoms(x)
    d = imaging.DeviceBuffer((p.My, p.Mx))
    for _ in range(warmup):
        im.computeD2BI_img(at, pa, npr, mode="fast", out=d)
    ms_vox, ms_int, ms_bin = [], [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        st = D2biStats()
        im.computeD2BI_img(at, pa, npr, mode="fast", stats=st, out=d)
        ms_vox.append(st.ms_voxel)
        ms_int.append(st.ms_integrate)
        ms_bin.append(st.ms_bin)
    t1 = time.perf_counter()
    img = d.numpy()
    value = imaged * steps / (t1 - t0)
    log("[bench] %s fast: %.3f ms/condition, %.3e imaged voxels/s" % (name, (t1 - t0) * 1e3 / steps, value))
    out = {"workload": name, "atoms": len(x), "voxels": [p.Mx, p.My, p.Mz], "n_prec": npr,
           "value": value, "ms_per_condition": (t1 - t0) * 1e3 / steps,
           "ms_bin": float(np.mean(ms_bin)), "ms_voxel": float(np.mean(ms_vox)),
           "ms_integrate": float(np.mean(ms_int))}
    if fp64_steps > 0:
        d64 = imaging.DeviceBuffer((p.My, p.Mx))
        im.computeD2BI_img(at, pa, npr, mode="fp64", out=d64)
        t0 = time.perf_counter()
        for _ in range(fp64_steps):
            im.computeD2BI_img(at, pa, npr, mode="fp64", out=d64)
        t1 = time.perf_counter()
        e = np.abs(img - d64.numpy())
        out.update(fp64_value=imaged * fp64_steps / (t1 - t0), max_abs_err_vs_fp64=float(e.max()),
                   mean_abs_err_vs_fp64=float(e.mean()))
        out["fast_over_fp64"] = value / out["fp64_value"]
        log("[bench] %s fp64: %.3e imaged voxels/s, max err %.3e" % (name, out["fp64_value"], e.max()))
    t_vox = out["ms_voxel"] * 1e-3
    ach = vox * FLOPS_PER_VOXEL / t_vox / 1e12
    out["roofline"] = {"bound": "valu", "pipe": "valu_fp32", "kernel": "k_voxel_tile<false,false,1,8,4,1>",
                       "achieved": ach, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                       "frac": ach / PEAK_FP32_TFLOPS, "traffic": TRAFFIC_BYTES.get(name),
                       "ms_per_launch": out["ms_voxel"]}
    return out, im, x, at


def calibrate_copy(mib, reps=3):
    """A plain device-to-device copy of known size, so that a profiler's
    FETCH_SIZE / WRITE_SIZE per dispatch can be scaled to bytes."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "reps": reps,
                      "GB_per_s": 8 * n / dt / 1e9}), flush=True)


def blur_only(n, reps=5, t=3.16):
    """Blur a seeded n x n image reps times (d2bi_gpu_blur: host image in,
    host image out) for a kernel trace of k_blur_x / k_blur_y; t = 3.16 px is
    c1's blur (2.5 A over 0.79 A pixels)."""
    import numpy as np
    from simulated_tem_amd import imaging
    imaging.init(0)
    f = np.random.default_rng(5).random((n, n))
    imaging.gaussianBlurImage(f, t)
    t0 = time.perf_counter()
    for _ in range(reps):
        imaging.gaussianBlurImage(f, t)
    dt = (time.perf_counter() - t0) / reps
    print(json.dumps({"blur_pixels": n * n, "t": t, "reps": reps, "ms_per_call_host_io": 1e3 * dt,
                      "kernel_bytes_per_pass": 16 * n * n}), flush=True)


def aperture_only(n_app, name="c1", angle=2e-3):
    """The aperture sum of workload name's condition over n_app points
    (computeD2BI_img with apertureAngle, nAperturePoints; host atoms in, host
    image out), timed with the shared voxel pass and with a binning and voxel
    pass per point (D2BI_APERTURE_SHARED=0), FAST and FP64; the two images of
    a mode must agree to the bit."""
    import numpy as np
    from simulated_tem_amd import imaging, synth
    imaging.init(0)
    cfg = dict(synth.CONFIGS[name])
    cfg["n_tomo"] = 1
    x, box = synth.make_atoms(cfg)
    x, conds = synth.conditions(cfg, x, box)
    im, pa, npr = conds[0]
    out = {"aperture_points": n_app, "workload": name, "atoms": len(x), "angle": angle}
    for mode, reps in (("fast", 3), ("fp64", 2)):
        imgs = {}
        for shared in ("1", "0"):
            os.environ["D2BI_APERTURE_SHARED"] = shared
            im.computeD2BI_img(x, pa, npr, mode=mode, apertureAngle=angle, nAperturePoints=n_app)
            t0 = time.perf_counter()
            for _ in range(reps):
                imgs[shared] = im.computeD2BI_img(x, pa, npr, mode=mode, apertureAngle=angle,
                                                  nAperturePoints=n_app)
            out["%s_%s_ms" % (mode, "shared" if shared == "1" else "per_point")] = \
                1e3 * (time.perf_counter() - t0) / reps
        os.environ.pop("D2BI_APERTURE_SHARED", None)
        out["%s_identical" % mode] = bool(np.array_equal(imgs["1"], imgs["0"]))
        log("[bench] aperture %s: shared %.1f ms, per point %.1f ms" %
            (mode, out[mode + "_shared_ms"], out[mode + "_per_point_ms"]))
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed steps (batches, or conditions for c0-c2)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c3", help="c3 (tilt series, default), c4 (uneven batch), c0, c1, c2")
    ap.add_argument("--fp64-conditions", type=int, default=3,
                    help="batch conditions timed in plain FP64 and checked per image")
    ap.add_argument("--cpu-cols", type=int, default=224,
                    help="side of the column window the CPU baseline images")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--single", default="c1", help="extra single-condition workload ('' to skip)")
    ap.add_argument("--blur", type=int, default=0, metavar="N",
                    help="only blur a seeded N x N image a few times (for a kernel trace of the blur)")
    ap.add_argument("--aperture", type=int, default=0, metavar="N",
                    help="only time c1's aperture sum over N points, shared pass vs a pass per point")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build images on one GPU; multi-GPU runs are not part of it "
              "(conditions or image tiles would shard across GPUs with no reduction).",
              file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.blur:
        blur_only(args.blur)
        return
    if args.aperture:
        aperture_only(args.aperture)
        return

    import numpy as np
    from simulated_tem_amd import imaging, synth
    imaging.init(0)

    if args.workload in SINGLE:
        s, im, x, at = single_main(args.workload, args.steps, args.warmup, imaging, synth, np)
        out = {"metric": METRIC, "value": s["value"], "unit": "imaged voxels/s", "n_gpus": 1,
               "steps": args.steps, "warmup": args.warmup, "ms_per_step": s["ms_per_condition"],
               "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
               "data": "synthetic (seeded bcc crystal, Eshelby centres, voids)",
               "config": {"workload": args.workload, "atoms": s["atoms"], "voxels": s["voxels"],
                          "n_prec": s["n_prec"]}}
        out.update({k: v for k, v in s.items() if k not in ("workload", "value", "atoms", "voxels", "n_prec")})
        if not args.no_cpu:
            nth = cpu_threads()
            log("[bench] cpu baseline (%d threads)" % nth)
            out["cpu_baseline"] = cpu_baseline(im, x, 0.0, 1, 2 * args.cpu_cols, nth)
        print(json.dumps(out), flush=True)
        return

    out, flat, xs, idx = batch_main(args, imaging, synth, np)
    if args.single:
        log("[bench] single-condition workload", args.single)
        out["single"] = single_main(args.single, 5, 2, imaging, synth, np, fp64_steps=1)[0]
    if not args.no_cpu:
        nth = cpu_threads()
        im, at, pa, npr, n = flat[idx[0]]
        log("[bench] cpu baseline (%d threads)" % nth)
        out["cpu_baseline"] = cpu_baseline(im, xs[idx[0]], pa, npr, args.cpu_cols, nth)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
