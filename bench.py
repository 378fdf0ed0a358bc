"""Trainer-step benchmark on one MI355X.

Workload (the benchmark's single-view configuration): 1M Gaussians (means
U(-5,5)^3, log-scales N(-4,0.5), random unit quaternions, opacity logits
N(0,1), DC U(0,1) as logits, SH degree 2 rest N(0,0.05)), a ring of 64 views
at radius 8 looking at the origin, 1600x900, f=1100. Targets are renders of a
second seeded 1M-Gaussian set. A training iteration = projection, tile
binning + per-tile depth sort, raster forward + Charbonnier loss, raster
backward and chain for one view picked by the seeded camera RNG; Adam
(position lr 1.6e-4) after every view in the single-view step, once per
batch of 8 views in the batched step. All three SH bands are active
(iteration counter starts at 2000, past the one-band-per-1000 ramp).

Runs, all from the same start on the same workload:
  f32          the plain FP32 single-view step (the reference's arithmetic and storage)
  bf16         single-view, SH coefficients and raster-record colours stored in bf16,
               Adam moments and all arithmetic f32 (the metric's mixed precision)
  mixed        single-view, SH coefficients and all Adam moments in bf16 (beyond the
               metric's rule that moments stay f32; reported, not the headline)
  bf16 x8      the batched step: 8 views projected and binned in one launch per
               stage, rasterized forward and backward in one persistent launch
               each, one Adam step, replayed as a hipGraph (x8o: each view's
               raster on a second stream beside the binning of the next views
               instead)
  bf16 x16     the same with 16 views per step (reported beside the headline)
Before timing, each run's trainer picks its raster shape (tile side 8, 16 or
32 pixels x LDS batch of 64, 128 or 256 keys) with its autotuner on views
0.. of the ring (--no-autotune keeps 16 x 256); config.runs records the
choice and the sweep's pass times. --sweep runs the autotuner on config 4
instead (each of its 8 view sizes alone, then one 8-view batch).
`value` is the batched bf16 step's training iterations per second: views
trained per second (the trainer's iteration counter advances by one per view,
as in the reference's stepOnce). `config.runs` holds every run's rate, and
`config.psnr` the PSNR drift of each 16-bit storage mode against FP32 after
2000 iterations under the same schedule.

Prints one JSON line with the roofline of the headline run's dominant kernel
(HIP events on the trainer's stream), a measured HBM ceiling (a plain 1 GiB
device-to-device copy), a CPU baseline (the C restatement under oracle/ on
the host's cores, one full step of the same workload) and the error of one
step of the headline storage mode against the FP64 restatement on a
4096-Gaussian crop of the scene.
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "splat.js_amd", "py"))

import numpy as np  # noqa: E402

import gsabi  # noqa: E402
import scenes  # noqa: E402

METRIC = "train iters/s @1M splats 1600x900 (fwd+bwd+sort), mixed vs FP32, ≤0.05dB PSNR"
HBM_PEAK_GBS = 8000.0
PRECS = {"f32": gsabi.GS_PREC_F32, "mixed": gsabi.GS_PREC_MIXED, "bf16": gsabi.GS_PREC_BF16,
         "fp16": gsabi.GS_PREC_FP16}
BATCH = 8


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_workload(n, views, W, H, f, seed=0, gt_seed=1):
    p, sh, cams = scenes.config2(n=n, seed=seed, views=views, W=W, H=H, f=f)
    gt_p, gt_sh, _ = scenes.config2(n=n, seed=gt_seed, views=views, W=W, H=H, f=f)
    t0 = time.time()
    gt = gsabi.GSTrainer(shDeg=2, maxSplats=n)
    blank = [{"tw": W, "th": H, "rgb": np.zeros(W * H * 3, np.float32)}]
    gt.setup({"data": gt_p, "n": n, "sh": gt_sh}, [dict(cams[0], imgIdx=0)], blank, W, H, 8.0)
    images = []
    for c in cams:
        img = gt.renderView(c)
        images.append({"tw": W, "th": H, "rgb": np.clip(img[..., :3], 0, 1).reshape(-1).copy()})
    gt.destroy()
    log(f"rendered {views} targets in {time.time() - t0:.1f}s")
    return p, sh, cams, images


def algorithmic_bytes(name, n, vis, keys, npx, shC, prec, views=1):
    """HBM bytes each kernel has to move per launch. SH coefficients take sb
    bytes (4 f32, 2 in every 16-bit mode), Adam moments mb (4, or 2 in mode
    mixed); the raster record is 48 B of f32 geometry and colour read per key
    (32 B in the 16-bit-colour modes: geometry and packed colours).
    project   n*(56 params + 16 key/rect/count) + vis*(sb*shC SH + 48 record write)
    project_views (one thread per Gaussian for all views) n*(56 params + sb*shC SH)
              + views*(n*16 key/rect/count + vis*48 record write)
    raster fwd  keys*(4 id + rec) + px*(16 img + 4 T + 4 end + 16 dL/dC + 4 target)
    raster bwd  keys*(4 id + rec + 16) + px*(16 dL/dC + 4 T + 4 end) + vis*40 gradient sums
    raster fwd+bwd (one launch, each tile's backward right after its forward)
                keys*(4 id + rec + 16) + px*44 + vis*40: the records and ids once, the
                forward's per-pixel traffic, the gradient sums
    update    n*(48 sums + 4 count + 2*56 params + 2*2*14*mb m/v) + n*(2*sb + 2*2*mb)*shC SH and its m/v
              (the batched update reads every view's sums and counts)
    bin_sort  n*(16 count pass + 40 scatter pass) + keys*(8 scattered key + 8 sort read + 4 sorted id)
    keys, vis and npx are per view; views > 1 for the batched launches (*_views:
    one launch covers every view of the step)."""
    sb = 4.0 if prec == "f32" else 2.0
    mb = 2.0 if prec == "mixed" else 4.0
    rec = 32.0 if prec in ("bf16", "fp16") else 48.0
    if name == "project":
        return views * (n * 72.0 + vis * (sb * shC + 48.0))
    if name == "project_views":
        return n * (56.0 + sb * shC) + views * (n * 16.0 + vis * 48.0)
    if name in ("raster_fwd_loss", "raster_fwd_views"):
        return views * (keys * (4.0 + rec) + npx * 44.0)
    if name in ("raster_bwd", "raster_bwd_views"):
        return views * (keys * (20.0 + rec) + npx * 24.0 + vis * 40.0)
    if name in ("raster_fwd_bwd", "raster_fwd_bwd_views"):
        return views * (keys * (20.0 + rec) + npx * 44.0 + vis * 40.0)
    if name in ("update", "update_views"):
        return n * (52.0 * views + 112.0 + 4.0 * 14.0 * mb) + n * (2.0 * sb + 4.0 * mb) * shC
    if name in ("bin_sort", "bin_sort_views"):
        return views * (n * 56.0 + keys * 20.0)
    return None


def _trainer(p, sh, cams, images, W, H, n, prec, views=1, graph=1, launch=gsabi.GS_VIEWS_PERSISTENT):
    t = gsabi.GSTrainer(shDeg=2, maxSplats=n, posLr=1.6e-4, seed=7, precision=PRECS[prec], viewsPerStep=views,
                        graph=graph, viewsLaunch=launch)
    t.setup({"data": p, "n": n, "sh": sh}, cams, images, W, H, 8.0)
    t.iter = 2000
    return t


def autotune(t, views, args):
    """The trainer's raster shape from its autotuner (gs_autotune) on
    `views`, or the default 16 x 256 with --no-autotune."""
    if args.no_autotune:
        return {"tile": 16, "batch": 256, "autotuned": False}
    res = t.autotune(views, reps=3)
    info = t.info()
    ms = {f"{e['tile']}x{e['batch']}": round(e["ms"], 3) for e in res if e["status"] == gsabi.GS_TUNE_MEASURED}
    log(f"autotune over views {views}: {ms} -> tile {info.tile}, batch {info.batch}")
    return {"tile": info.tile, "batch": info.batch, "autotuned": True, "pass_ms": ms}


def parse_run(name):
    """'bf16' -> ('bf16', 0, False); 'bf16x8' -> ('bf16', 8, False); 'bf16x8o'
    (overlapped launches) -> ('bf16', 8, True)."""
    m = re.fullmatch(r"(f32|bf16|fp16|mixed)(?:x(\d+)(o?))?", name)
    if not m:
        raise SystemExit(f"bench.py: unknown run {name!r}")
    return m.group(1), int(m.group(2) or 0), m.group(3) == "o"


def run_single(prec, p, sh, cams, images, W, H, args):
    """Warm up, time args.steps single-view steps with per-stage HIP events
    on the trainer's stream, then sample key and visible counts over 16 more
    steps (they vary by view)."""
    t = _trainer(p, sh, cams, images, W, H, args.n, prec)
    shape = autotune(t, [0], args)
    for _ in range(args.warmup):
        t.stepOnce()
    t.sync()
    t.profile(True)
    t.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t.stepOnce()
    t.sync()
    dt = time.perf_counter() - t0
    prof = t.profile_read()
    t.profile(False)
    keys, vis = [], []
    for _ in range(16):
        t.stepOnce()
        s = t.step_stats()
        keys.append(s.lastKeys)
        vis.append(s.lastVisible)
    t.destroy()
    log(f"{prec}: {args.steps / dt:.1f} iters/s")
    return {"its": args.steps / dt, "ms_per_iter": 1e3 * dt / args.steps, "prof": prof,
            "keys": float(np.mean(keys)), "vis": float(np.mean(vis)), "shape": shape}


def run_batched(prec, p, sh, cams, images, W, H, args, launch=gsabi.GS_VIEWS_PERSISTENT, batch=BATCH):
    """Batched steps of `batch` views: warm up (the second step captures the
    graph), time args.steps batched steps replayed from the graph, then a
    profiled pass (eager persistent launches with per-stage HIP events, one
    stream) and key counts."""
    t = _trainer(p, sh, cams, images, W, H, args.n, prec, views=batch, launch=launch)
    shape = autotune(t, list(range(batch)), args)
    picks = [-1] * batch
    for _ in range(max(args.warmup, 2)):
        t.stepViews(picks)
    t.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t.stepViews(picks)
    t.sync()
    dt = time.perf_counter() - t0
    t.profile(True)
    for _ in range(16):
        t.stepViews(picks)
    t.sync()
    prof = t.profile_read()
    t.profile(False)
    keys, vis = [], []
    for _ in range(4):
        t.stepViews(picks)
        s = t.step_stats()
        keys.append(s.lastKeys / batch)
        vis.append(s.lastVisible / batch)
    t.destroy()
    its = args.steps * batch / dt
    form = "overlapped" if launch == gsabi.GS_VIEWS_OVERLAP else "persistent"
    log(f"{prec} x{batch} {form}: {its:.1f} iters/s ({args.steps / dt:.1f} batched steps/s)")
    return {"its": its, "ms_per_iter": 1e3 * dt / (args.steps * batch), "steps_per_s": args.steps / dt, "views": batch,
            "prof": prof, "keys": float(np.mean(keys)), "vis": float(np.mean(vis)), "shape": shape}


def psnr_after(prec, p, sh, cams, images, W, H, n, iters, batched):
    """Mean PSNR over all views after `iters` training iterations from the
    same start (single-view steps, or batched steps of BATCH views)."""
    t = _trainer(p, sh, cams, images, W, H, n, prec, views=BATCH if batched else 1)
    if batched:
        for _ in range(iters // BATCH):
            t.stepViews([-1] * BATCH)
    else:
        for _ in range(iters):
            t.stepOnce()
    v = float(np.mean([t.evalCamPsnr(ci) for ci in range(len(cams))]))
    t.destroy()
    return v


def cpu_leg(p, sh, cam, image, W, H, radius, n, threads, headline_prec):
    """The CPU side (runs on the host only): one full f32 step of the C
    restatement on the same inputs and view, timed over `threads` threads;
    then the error of one GPU step of the headline storage mode against the
    FP64 restatement on a 4096-Gaussian crop (the Gaussians inside [-1, 1]^3)
    of the same scene, with the same view."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle as O  # CPU restatement: the baseline and the FP64 checker, never the product
    from helpers import OracleTrainer
    O.set_threads(threads)
    ref = OracleTrainer(p, sh, [dict(cam, imgIdx=0)], [image], radius, opts={"posLr": 1.6e-4}, shDeg=2)
    ref.iter = 2000
    t0 = time.perf_counter()
    ref.step(0)
    sec = time.perf_counter() - t0
    # FP64 error on a crop
    inside = np.all(np.abs(p[:, :3]) <= 1.0, axis=1)
    idx = np.nonzero(inside)[0][:4096]
    cp, csh = np.ascontiguousarray(p[idx]), np.ascontiguousarray(sh[idx])
    st = O.Stage(O.make_opts(shDeg=2), "f64")
    ocam = O.make_cam(cam, shActiveDeg=2)
    tgt = O.pack_targets(np.asarray(image["rgb"], np.float32).reshape(-1, 3))
    proj, rect, cnt = st.project(ocam, cp, csh)
    ids, ts, _ = st.bin_sort(ocam, proj, rect, cnt)
    img64, _, _ = st.render_fwd(ocam, proj, ids, ts)
    g64 = O.analytic_grads(st, ocam, cp, csh, tgt, fixed=False)[0]
    O.set_threads(1)
    t = gsabi.GSTrainer(shDeg=2, maxSplats=cp.shape[0], precision=PRECS[headline_prec])
    t.setup({"data": cp, "n": cp.shape[0], "sh": csh}, [dict(cam, imgIdx=0)], [image], W, H, radius)
    t.iter = 2000
    g = t.step_debug(0, want=("img", "gradF"))
    t.destroy()
    err = np.abs(g["img"][:, :3].astype(np.float64) - img64[:, :3]).reshape(-1)
    meds = []
    for s in range(14):
        a, b = g["gradF"][:, s].astype(np.float64), g64[:, s]
        big = np.abs(b) > 1e-2 * max(np.abs(b).max(), 1e-30)
        meds.append(float(np.median(np.abs(a[big] - b[big]) / np.abs(b[big]))))
    fp64 = {"storage": headline_prec, "sample": f"one step on a {cp.shape[0]}-Gaussian crop of the scene, "
                                                 f"view 0, {W}x{H}",
            "image_abs_mean": float(f"{err.mean():.3g}"), "image_abs_p999": float(f"{np.quantile(err, 0.999):.3g}"),
            "image_abs_max": float(f"{err.max():.3g}"),
            "grad_median_rel_worst_slot": float(f"{max(meds):.3g}")}
    return sec, fp64


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200, help="timed steps per run (batched: steps of 8 views)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--views", type=int, default=64)
    ap.add_argument("--width", type=int, default=1600)
    ap.add_argument("--height", type=int, default=900)
    ap.add_argument("--runs", default="f32,bf16,mixed,bf16x8,bf16x16",
                    help="comma list of f32, bf16, fp16, mixed (single-view), <prec>x<V> (batched steps of V "
                         "views, persistent launches; bf16x8 is the headline) and <prec>x<V>o (overlapped)")
    ap.add_argument("--psnr-iters", type=int, default=2000, help="0 skips the PSNR drift runs")
    ap.add_argument("--threads", type=int, default=0, help="CPU baseline threads (0: all, at most 16)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--config5", action="store_true",
                    help="instead: the stress configuration (3M Gaussians, SH 3, 1920x1080, 4-view batches)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default raster shape (16-pixel tiles, 256-key LDS batches)")
    ap.add_argument("--sweep", action="store_true",
                    help="instead: autotuner sweeps of config 4 (each view size alone, then one 8-view batch)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one plain 1 GiB device-to-device copy (scales HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this trainer runs on one GPU; sharding across GPUs is a design only",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    if args.config5:
        run_config5(args)
        return
    if args.sweep:
        run_sweep(args)
        return
    W, H, f = args.width, args.height, 1100.0 * args.width / 1600.0
    p, sh, cams, images = build_workload(args.n, args.views, W, H, f)
    runs = {}
    for r in args.runs.split(","):
        prec, nv, over = parse_run(r)
        if nv:
            runs[r] = run_batched(prec, p, sh, cams, images, W, H, args, batch=nv,
                                  launch=gsabi.GS_VIEWS_OVERLAP if over else gsabi.GS_VIEWS_PERSISTENT)
        else:
            runs[r] = run_single(r, p, sh, cams, images, W, H, args)
    head_name = "bf16x8" if "bf16x8" in runs else list(runs)[-1]
    head = runs[head_name]
    head_prec = parse_run(head_name)[0]
    views = head.get("views", 1)

    npx = W * H
    shC = 24
    kernels = {}
    for name, v in head["prof"].items():
        avg_ms = v["ms"] / max(v["launches"], 1)
        nv = views if name.endswith("_views") else 1
        b = algorithmic_bytes(name, args.n, head["vis"], head["keys"], npx, shC, head_prec, nv)
        kernels[name] = {"avg_ms": round(avg_ms, 4), "launches_per_step": round(v["launches"] / 16, 2),
                         "GBps": None if b is None else round(b / (avg_ms * 1e-3) / 1e9, 1)}
    single = {k: v for k, v in kernels.items() if v["GBps"] is not None}
    dom = max(single, key=lambda k: single[k]["avg_ms"] * head["prof"][k]["launches"])
    nv = views if dom.endswith("_views") else 1
    dom_b = algorithmic_bytes(dom, args.n, head["vis"], head["keys"], npx, shC, head_prec, nv)
    achieved = dom_b / (single[dom]["avg_ms"] * 1e-3) / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", f"traffic_{dom}_1M_1600x900.json")
    if os.path.exists(tf):
        traffic = json.load(open(tf)).get("bytes_per_launch")

    copy_gbps = round(copy_ceiling(), 1)
    log(f"device-to-device copy: {copy_gbps} GB/s (read + write)")

    psnr = {}
    if args.psnr_iters > 0:
        for batched in (False, True):
            base = psnr_after("f32", p, sh, cams, images, W, H, args.n, args.psnr_iters, batched)
            tag = f"x{BATCH}" if batched else "single"
            psnr[tag] = {"f32_db": round(base, 4)}
            for prec in ("bf16", "fp16", "mixed"):
                v = psnr_after(prec, p, sh, cams, images, W, H, args.n, args.psnr_iters, batched)
                psnr[tag][f"{prec}_delta_db"] = round(v - base, 4)
            log(f"psnr after {args.psnr_iters} iterations ({tag}): {psnr[tag]}")

    cpu, fp64 = None, None
    if not args.no_cpu_baseline:
        threads = args.threads or min(16, os.cpu_count() or 1)
        log(f"cpu leg: one f32 step of the C restatement on {threads} threads; FP64 crop check")
        sec, fp64 = cpu_leg(p, sh, cams[0], images[0], W, H, 8.0, args.n, threads, head_prec)
        cpu = {"value": round(1.0 / sec, 5), "unit": "iters/s", "cores": threads, "kind": "port",
               "sample": f"1 full training step (project, sort, fwd, loss, bwd, chain, Adam) of the same "
                         f"{args.n}-Gaussian {W}x{H} workload on view 0, f32 C restatement with OpenMP, "
                         f"{sec:.1f}s"}

    storage = {"f32": "f32", "bf16": "bf16 SH coefficients and raster-record colours, f32 Adam moments",
               "fp16": "fp16 SH coefficients and raster-record colours, f32 Adam moments",
               "mixed": "bf16 SH coefficients and bf16 Adam moments"}
    config = {"workload": f"{'batched ' + str(BATCH) + '-view' if views > 1 else 'single-view'} trainer step, "
                          f"{args.n} Gaussians SH deg 2, {args.views}-view ring r=8, {W}x{H} f={f:g}, "
                          f"{head_prec} storage",
              "gaussians": args.n, "width": W, "height": H, "views": args.views, "sh_degree": 2,
              "storage": storage[head_prec], "arithmetic": "f32 (f64 nowhere on the GPU)",
              "iteration": "one view's projection, binning + sort, raster fwd + loss, raster bwd and chain; "
                           "Adam after each view (single) or once per batch (batched)",
              "views_per_step": views,
              "keys_per_view": round(head["keys"]), "visible_per_view": round(head["vis"]),
              "sorted_keys_per_s": round(head["keys"] * head["its"]),   # over the whole step
              "kernels": kernels, "hbm_copy_ceiling_GBps": copy_gbps,
              "runs": {k: {"iters_per_s": round(v["its"], 2), "ms_per_iter": round(v["ms_per_iter"], 4),
                           **({"batched_steps_per_s": round(v["steps_per_s"], 2)} if "steps_per_s" in v else {}),
                           "storage": storage[parse_run(k)[0]], "raster_shape": v["shape"],
                           **({"views_per_step": parse_run(k)[1],
                               "launch": "overlapped" if parse_run(k)[2] else "persistent"} if parse_run(k)[1] else {})}
                       for k, v in runs.items()},
              "psnr": psnr, "psnr_protocol": f"mean PSNR over all {args.views} views after {args.psnr_iters} "
                                            f"iterations from the same start and seed; each mode against f32 "
                                            f"with the same schedule (single-view, or batches of {BATCH})",
              "fp64_error": fp64}
    # keys per second of the binning launch itself (K / t_sort): the keys one
    # launch bins over its event-timed duration
    kb = kernels.get("bin_sort_views" if views > 1 else "bin_sort")
    if kb and kb.get("avg_ms"):
        config["binning_keys_per_s"] = round(head["keys"] * views / (kb["avg_ms"] * 1e-3))
    if "f32" in runs:
        config["ratio_vs_f32_single"] = round(head["its"] / runs["f32"]["its"], 3)
    line = {
        "metric": METRIC, "value": round(head["its"], 2), "unit": "iters/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(head["ms_per_iter"] * views, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": config,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    if args.copy_calibration:
        copy_calibration()


def run_sweep(args):
    """Autotuner sweeps on config 4 (1M Gaussians, 8 views from 800x450 to
    1920x1080, bf16 storage): each view size alone (single-view passes),
    then the 8 views as one batched pass. One JSON line per sweep."""
    p, sh, cams = scenes.config4(n=args.n, seed=0)
    images = [{"tw": c["w"], "th": c["h"], "rgb": np.full(c["w"] * c["h"] * 3, 0.4, np.float32)} for c in cams]
    maxW, maxH = max(c["w"] for c in cams), max(c["h"] for c in cams)
    t = gsabi.GSTrainer(shDeg=2, maxSplats=args.n, precision=gsabi.GS_PREC_BF16, viewsPerStep=BATCH)
    t.setup({"data": p, "n": args.n, "sh": sh}, cams, images, maxW, maxH, 8.0)
    t.iter = 2500
    status = {gsabi.GS_TUNE_MEASURED: "measured", gsabi.GS_TUNE_TOO_MANY_TILES: "too many tiles",
              gsabi.GS_TUNE_MORE_DROPS: "more budget drops"}
    for views in [[k] for k in range(len(cams))] + [list(range(len(cams)))]:
        res = t.autotune(views, reps=5)
        info = t.info()
        label = (f"{cams[views[0]]['w']}x{cams[views[0]]['h']} single view" if len(views) == 1
                 else f"{len(views)}-view batch")
        print(json.dumps({"sweep": label, "best": f"{info.tile}x{info.batch}",
                          "shapes": {f"{e['tile']}x{e['batch']}": {"ms": round(e["ms"], 4), "keys": e["keys"],
                                                                    "status": status[e["status"]]}
                                     for e in res}}), flush=True)
    t.destroy()


def device_used_bytes():
    """Bytes of device memory in use on the current GPU (hipMemGetInfo)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return total.value - free.value


def run_config5(args):
    """The stress configuration: 3M Gaussians, SH degree 3, log-scales
    N(-3, 1), 1920x1080, batches of 4 views. The key budget holds every key
    (the reference's default 12M would drop most tiles of these views).
    Reports device memory in use (peak over the run, minus what was in use
    before the trainer), keys per view, sorted keys/s and iterations (views)/s."""
    n, W, H, V = 3_000_000, 1920, 1080, 4
    p, sh, cams = scenes.config5(n=n, seed=0, views=V, W=W, H=H)
    gp, gsh, _ = scenes.config5(n=n, seed=1, views=V, W=W, H=H)
    base = device_used_bytes()
    gt = gsabi.GSTrainer(shDeg=3, maxSplats=n, entriesCap=400_000_000)
    blank = [{"tw": W, "th": H, "rgb": np.zeros(W * H * 3, np.float32)}]
    gt.setup({"data": gp, "n": n, "sh": gsh}, [dict(cams[0], imgIdx=0)], blank, W, H, 8.0)
    images = [{"tw": W, "th": H, "rgb": np.clip(gt.renderView(c)[..., :3], 0, 1).reshape(-1).copy()} for c in cams]
    gt.destroy()
    out = {}
    for prec in ("f32", "bf16"):
        t = gsabi.GSTrainer(shDeg=3, maxSplats=n, entriesCap=400_000_000, posLr=1.6e-4, seed=7,
                            precision=PRECS[prec], viewsPerStep=V)
        t.setup({"data": p, "n": n, "sh": sh}, cams, images, W, H, 8.0)
        t.iter = 5000
        peak = device_used_bytes()
        for _ in range(2):
            t.stepViews(list(range(V)))
        t.sync()
        peak = max(peak, device_used_bytes())
        steps = max(2, args.steps // 20)
        t0 = time.perf_counter()
        for _ in range(steps):
            t.stepViews(list(range(V)))
        t.sync()
        dt = time.perf_counter() - t0
        t.profile(True)
        t.stepViews(list(range(V)))
        t.sync()
        prof = t.profile_read()
        t.profile(False)
        st = t.step_stats()
        peak = max(peak, device_used_bytes())
        t.destroy()
        keys = st.lastKeys / V
        if "bin_sort_views" in prof:   # every view's binning in one launch per stage
            bin_ms = prof["bin_sort_views"]["ms"] / max(prof["bin_sort_views"]["launches"], 1) / V
        else:
            bin_ms = prof["bin_sort"]["ms"] / max(prof["bin_sort"]["launches"], 1)
        out[prec] = {"iters_per_s": round(steps * V / dt, 2), "batched_steps_per_s": round(steps / dt, 3),
                     "keys_per_view": int(keys), "visible_per_view": int(st.lastVisible / V),
                     "sorted_keys_per_s": round(keys / (bin_ms * 1e-3)), "bin_sort_ms": round(bin_ms, 3),
                     "peak_device_GB": round((peak - base) / 1e9, 2),
                     "stage_ms": {k: round(v["ms"] / max(v["launches"], 1), 3) for k, v in prof.items()}}
        log(f"config 5 {prec}: {out[prec]}")
    print(json.dumps({"config5": {"workload": f"{n} Gaussians SH deg 3, log-scales N(-3,1), {W}x{H}, "
                                              f"{V}-view batches, key budget 400M", "runs": out}}), flush=True)


def copy_ceiling(nbytes=1 << 30, reps=5):
    """Measured HBM ceiling: a plain device-to-device copy of nbytes (one
    read and one write of every byte), timed with HIP events, best of reps.
    Returns GB/s of read + write traffic."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    e0, e1 = C.c_void_p(), C.c_void_p()
    assert hip.hipEventCreate(C.byref(e0)) == 0 and hip.hipEventCreate(C.byref(e1)) == 0
    best = None
    for _ in range(reps + 1):      # the first copy warms up
        assert hip.hipEventRecord(e0, None) == 0
        assert hip.hipMemcpyAsync(b, a, C.c_size_t(nbytes), 3, None) == 0  # hipMemcpyDeviceToDevice
        assert hip.hipEventRecord(e1, None) == 0
        assert hip.hipEventSynchronize(e1) == 0
        ms = C.c_float()
        assert hip.hipEventElapsedTime(C.byref(ms), e0, e1) == 0
        best = ms.value if best is None else min(best, ms.value)
    hip.hipEventDestroy(e0)
    hip.hipEventDestroy(e1)
    hip.hipFree(a)
    hip.hipFree(b)
    return 2.0 * nbytes / (best * 1e-3) / 1e9


def copy_calibration(nbytes=1 << 30):
    """One hipMemcpy device-to-device of nbytes: a kernel whose HBM traffic is
    known (nbytes read + nbytes written), to scale FETCH_SIZE/WRITE_SIZE."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy calibration: {nbytes} B device-to-device")


if __name__ == "__main__":
    main()
