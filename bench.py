"""Imaging throughput benchmark: imaged voxels/s (voxelise + integrate).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c1]

A step is one imaging condition of the workload's first condition, from
device-resident atoms (uploaded before timing) to the image in HBM:
link-cell binning, voxelisation and RK4 column integration, the window of
the reference's T_D2BI timer (d2bi.f90:1392-1410).  value_image_to_host is
the same with the image copied to host memory (PCIe included).  Imaged voxels per step =
Mx*My*Mz*nPrecessionAngle.  The timed steps run the FAST (FP32) mode; the
plain per-condition FP64 mode is timed beside it (fp64_value) with the FAST
vs FP64 intensity error of the same image.

Prints ONE JSON line.  roofline: the dominant kernel (k_voxel_tile), timed
live with HIP events around its launch on the library's stream; algorithmic
work per voxel = 388 FP32 flops + 30 transcendentals (28.27 in-sphere atoms
x 13 flops + 1 exp, plus normalisation and density).  cpu_baseline: the
oracle's C restatement (FP64, OpenMP over pixel rows) on a window of the
same workload, with the whole crystal binned.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3          # MI355X published FP32 vector/matrix peak
PEAK_TRANS_PER_S = 157.3e12 / 16     # transcendentals: quarter rate of one wave64 instruction per clock per CU
FLOPS_PER_VOXEL = 388.0
TRANS_PER_VOXEL = 30.0
# HBM bytes per k_voxel_tile launch from rocprofv3 FETCH_SIZE / WRITE_SIZE
# passes scaled by a 1 GiB copy's counts (profiles/k_voxel_tile_c1_traffic.md);
# a profiler cannot run inside the timed process, so the figure is recorded
# per workload
TRAFFIC_BYTES = {"c1": 2.644e9}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(im, x, sample_cols, nthreads):
    from oracle import oracle as orc
    t0 = time.perf_counter()
    oc = orc.Condition(im.rec, x)
    t_bin = time.perf_counter() - t0
    p = oc.plan
    w = sample_cols
    ix0, iy0 = p.Mx // 2 - w // 2, p.My // 2 - w // 2
    t0 = time.perf_counter()
    oc.image(window=(ix0, ix0 + w, iy0, iy0 + w), nthreads=nthreads)
    t_win = time.perf_counter() - t0
    v_full = p.Mx * p.My * p.Mz
    v_win = w * w * p.Mz
    t_full = t_bin + t_win * v_full / v_win
    return {"value": v_full / t_full, "unit": "imaged voxels/s", "cores": nthreads, "kind": "port",
            "sample": "oracle FP64: whole crystal binned (%.1f s) + %dx%d columns x %d slices "
                      "(%.2f s, %d threads), extrapolated to %dx%dx%d" %
                      (t_bin, w, w, p.Mz, t_win, nthreads, p.Mx, p.My, p.Mz)}


def batch_bench(args, imaging, synth):
    """Batched conditions through the graph-captured batch API: every
    condition re-binned and re-voxelised, P precession points integrated per
    voxel field, the images left in HBM (value) or copied to host memory
    (value_image_to_host, PCIe included).  Plain FP64 = eager per-condition
    calls on a subset."""
    import numpy as np
    log("[bench] batch workload", args.batch)
    if args.batch == "c4":
        # the uneven batch: 64 seeded configurations, each under its own
        # g-vectors, voxel step and precession points
        data = synth.c4_conditions(synth.c4_configs())
        flat = []
        for x, conds in data:
            at = imaging.Atoms(x)
            flat.extend((im, at, pa, npr, len(x)) for (im, pa, npr) in conds)
    else:
        cfg = dict(synth.CONFIGS[args.batch])
        x, box = synth.make_atoms(cfg)
        x, conds = synth.conditions(cfg, x, box)
        at = imaging.Atoms(x)
        flat = [(im, at, pa, npr, len(x)) for (im, pa, npr) in conds]
    conds = [(im, pa, npr) for (im, at, pa, npr, n) in flat]
    batch = [(im, at, pa, npr) for (im, at, pa, npr, n) in flat]
    imaged = 0
    for im, at, pa, npr, n in flat:
        p = im.plan(n, pa, npr)
        imaged += p.Mx * p.My * p.Mz * npr
    dev = imaging.DeviceBuffer((imaging.batch_image_size(batch),))
    t0 = time.perf_counter()
    imaging.computeD2BI_img_batch(batch, mode="fast", out=dev)       # plans, captures
    log("[bench] batch first call (capture) %.2f s" % (time.perf_counter() - t0))
    t0 = time.perf_counter()
    for _ in range(args.batch_steps):
        imaging.computeD2BI_img_batch(batch, mode="fast", out=dev)
    t1 = time.perf_counter()
    s_per_batch = (t1 - t0) / args.batch_steps
    value = imaged / s_per_batch
    log("[bench] batch fast: %.3f s/batch, %.3e imaged voxels/s" % (s_per_batch, value))
    t0 = time.perf_counter()
    imgs = imaging.computeD2BI_img_batch(batch, mode="fast")
    s_host = time.perf_counter() - t0
    assert np.array_equal(dev.numpy(), np.concatenate([i.ravel() for i in imgs]))
    log("[bench] batch fast, images copied to host: %.3f s/batch" % s_host)
    idx = np.linspace(0, len(conds) - 1, max(1, args.batch_fp64)).round().astype(int)
    errs, imaged64 = [], 0
    t64 = 0.0
    for i in idx:
        im, at, pa, npr, n = flat[i]
        p = im.plan(n, pa, npr)
        t0 = time.perf_counter()
        g64 = im.computeD2BI_img(at, pa, npr, mode="fp64")
        t64 += time.perf_counter() - t0
        imaged64 += p.Mx * p.My * p.Mz * npr
        errs.append(float(np.max(np.abs(g64 - imgs[i]))))
    fp64_value = imaged64 / t64
    log("[bench] batch fp64 (plain, %d conditions): %.3e imaged voxels/s, max err %.3e"
        % (len(idx), fp64_value, max(errs)))
    return {"workload": args.batch, "conditions": len(conds), "n_prec": max(c[2] for c in conds),
            "imaged_voxels_per_batch": imaged, "value": value, "unit": "imaged voxels/s",
            "s_per_batch": s_per_batch, "steps": args.batch_steps,
            "value_image_to_host": imaged / s_host,
            "fp64_value": fp64_value, "fp64_conditions": len(idx),
            "fast_over_fp64": value / fp64_value, "max_abs_err_vs_fp64": max(errs)}


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="c1")
    ap.add_argument("--fp64-steps", type=int, default=3)
    ap.add_argument("--cpu-cols", type=int, default=384,
                    help="side of the column window the CPU baseline images (~10 s on 16 threads for c1)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch", default="c3", help="batched workload: c3 (tilt series), c4 (uneven), '' to skip")
    ap.add_argument("--batch-steps", type=int, default=2)
    ap.add_argument("--batch-fp64", type=int, default=3, help="conditions timed in plain FP64")
    ap.add_argument("--blur", type=int, default=0, metavar="N",
                    help="only blur a seeded N x N image a few times (for a kernel trace of the blur)")
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

    import numpy as np
    from simulated_tem_amd import imaging, synth
    from simulated_tem_amd.abi import D2biStats

    imaging.init(0)
    cfg = dict(synth.CONFIGS[args.workload])
    log("[bench] generating workload", args.workload)
    x, box = synth.make_atoms(cfg)
    x, conds = synth.conditions(cfg, x, box)
    im, prec_angle, n_prec = conds[0]
    p = im.plan(len(x), prec_angle, n_prec)
    vox = p.Mx * p.My * p.Mz
    imaged = vox * n_prec
    atoms = imaging.Atoms(x)
    log("[bench] %d atoms, %dx%dx%d voxels, P=%d" % (len(x), p.Mx, p.My, p.Mz, n_prec))

    # the image stays in HBM inside the timed steps (inputs and outputs
    # resident); the rate with the image copied to host memory is timed after
    d_img = imaging.DeviceBuffer((p.My, p.Mx))

    def step(mode, st, out=None):
        return im.computeD2BI_img(atoms, prec_angle, n_prec, mode=mode, stats=st, out=out)

    st = D2biStats()
    for _ in range(args.warmup):
        step("fast", st, d_img)
    log("[bench] warmup done")
    ms_vox, ms_int, ms_bin = [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = D2biStats()
        step("fast", st, d_img)
        ms_vox.append(st.ms_voxel)
        ms_int.append(st.ms_integrate)
        ms_bin.append(st.ms_bin)
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / args.steps
    value = imaged * args.steps / (t1 - t0)
    img_fast = d_img.numpy()
    log("[bench] fast: %.3f ms/step, %.3e imaged voxels/s" % (ms_step, value))
    n_h = max(1, min(5, args.steps))
    t0 = time.perf_counter()
    for _ in range(n_h):
        img_host = step("fast", D2biStats())
    value_host = imaged * n_h / (time.perf_counter() - t0)
    assert np.array_equal(img_host, img_fast)
    log("[bench] fast with the image copied to host memory: %.3e imaged voxels/s" % value_host)

    # plain per-condition FP64 path, same condition
    fp64_value, err, ms64 = None, None, None
    if args.fp64_steps > 0:
        st64 = D2biStats()
        img64 = step("fp64", st64)
        ms64 = []
        t0 = time.perf_counter()
        for _ in range(args.fp64_steps):
            st64 = D2biStats()
            img64 = step("fp64", st64)
            ms64.append((st64.ms_total, st64.ms_voxel))
        t1 = time.perf_counter()
        fp64_value = imaged * args.fp64_steps / (t1 - t0)
        err = np.abs(img_fast - img64)
        log("[bench] fp64: %.3e imaged voxels/s, fast vs fp64 max %.3e" % (fp64_value, err.max()))

    t_vox = float(np.mean(ms_vox)) * 1e-3
    achieved = vox * FLOPS_PER_VOXEL / t_vox / 1e12
    roof = {"bound": "mfma", "pipe": "valu_fp32", "kernel": "k_voxel_tile",
            "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS, "traffic": TRAFFIC_BYTES.get(args.workload),
            "traffic_unit": "B per launch (rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0)",
            "ms_per_launch": t_vox * 1e3,
            "transcendental_frac": vox * TRANS_PER_VOXEL / t_vox / PEAK_TRANS_PER_S}
    out = {
        "metric": "imaged voxels/s (voxelise+integrate), batched FP32 vs plain FP64, err<=1e-4",
        "value": value, "unit": "imaged voxels/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded bcc crystal, Eshelby centres, voids)",
        "config": {"workload": args.workload, "atoms": len(x), "voxels": [p.Mx, p.My, p.Mz],
                   "n_prec": n_prec, "g": list(cfg["g"][0]), "sg": cfg["sg"][0]},
        "value_image_to_host": value_host,
        "fp64_value": fp64_value, "fast_over_fp64": value / fp64_value if fp64_value else None,
        "max_abs_err_vs_fp64": float(err.max()) if err is not None else None,
        "mean_abs_err_vs_fp64": float(err.mean()) if err is not None else None,
        "ms_bin": float(np.mean(ms_bin)), "ms_voxel": float(np.mean(ms_vox)),
        "ms_integrate": float(np.mean(ms_int)),
        "ms_fp64_total": float(np.mean([m[0] for m in ms64])) if ms64 else None,
        "ms_fp64_voxel": float(np.mean([m[1] for m in ms64])) if ms64 else None,
        "n_guard": int(st.n_guard),
        "roofline": roof,
    }
    if args.batch:
        out["batch"] = batch_bench(args, imaging, synth)
    if not args.no_cpu:
        nth = min(16, os.cpu_count() or 1)
        log("[bench] cpu baseline (%d threads)" % nth)
        out["cpu_baseline"] = cpu_baseline(im, x, args.cpu_cols, nth)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
