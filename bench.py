"""Benchmark of the volume-filter hot path on one MI355X.

Workload (one "step"): the circle-fitting forward of runCircleFittingMethod
(tube-segmentation.cpp:212-939) on a seeded 512^3 thin-vessel MRA-like volume
(120 tubes, radius 0.5-3, Rician noise 0.08): small-scale branch (blur 0.5 ->
vector field -> TDF r = 0.5..3.0) and large-scale branch (blur 0.5 -> vector
field -> GVF Euler, mu = 0.05, 400 iterations, field stored in FP16 with
float arithmetic -> TDF r = 2.5..4.0, step 0.5) -> combine.

Headline value: GVF throughput in Gvoxel-iter/s (voxels x iterations / GVF
time) of the FP16-stored field.  Also reported: forward ms per step, the same
GVF in plain FP32 storage (speed-up), the GVF kernel's achieved HBM bandwidth
against the 8 TB/s peak, the oracle's FP32 GVF on the host CPU, and one
forward + adjoint step of the differentiable FP32 chain at the same size
(blur 0.5 -> vector field -> GVF 400 iterations checkpointed every 40 ->
TDF r = 0.5..4.0, backward to dL/dvolume, dL/dmu, dL/dsigma) with its
device memory at the backward's peak.

--fit-steps K runs K gradient-descent steps fitting mu and sigma of that
chain on the noisy volume to the TDF of the noise-free rendering of the
same tree (mu 0.05, sigma 0.5), starting from mu 0.03, sigma 0.8, and
prints one JSON line for the fit.

All inputs are resident in device memory before the timed region; timings use
HIP events recorded on the stream the kernels run on.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tube-segmentation-framework_amd"))

import numpy as np  # noqa: E402

PEAK_HBM = 8.0e12  # bytes/s, MI355X HBM3E


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--iters", type=int, default=400)
    ap.add_argument("--mu", type=float, default=0.05)
    ap.add_argument("--storage", choices=["f16", "bf16", "sn16", "f32"], default="f16",
                    help="vector-field storage; sn16 is the reference's own 16-bit format")
    ap.add_argument("--fuse", type=int, default=2, choices=[1, 2, 4], help="GVF steps per launch")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-f32-compare", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-downstream", action="store_true",
                    help="skip timing the stages around the filters (ingest, centerline graph, segmentation)")
    ap.add_argument("--checkpoint-every", type=int, default=40)
    ap.add_argument("--fit-steps", type=int, default=0,
                    help="run this many mu/sigma fitting steps instead of the forward benchmark")
    ap.add_argument("--pmc-workload", action="store_true",
                    help="run a short fixed workload for hardware-counter passes (no timing output)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the volume-filter path runs on a single device "
              "(multi-GPU z-slab sharding is designed, not built); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import tsf
    from tsf import synth

    if tsf.device_count() < 1:
        raise RuntimeError("no GPU visible")
    S = {"f32": tsf.STORAGE_F32, "f16": tsf.STORAGE_F16, "bf16": tsf.STORAGE_BF16,
         "sn16": tsf.STORAGE_SN16}[args.storage]
    scale = 2.0 ** 14 if S == tsf.STORAGE_F16 else 1.0
    n = args.size
    shape = (n, n, n)
    N = n ** 3

    t0 = time.time()
    vol, _ = synth.config5(shape=shape)
    log(f"synthetic {shape} volume in {time.time() - t0:.1f}s")

    stream = tsf.Stream()
    d_vol = tsf.DeviceArray.from_host(vol, stream=stream)
    if args.fit_steps > 0:
        fit(args, vol, d_vol, shape, stream)
        return
    d_blur = tsf.DeviceArray(shape, np.float32)
    F_small = tsf.Field(shape, S, scale)
    F = tsf.Field(shape, S, scale)
    V = tsf.Field(shape, S, scale)
    ws = tsf.DeviceArray((tsf.gvf_euler_workspace_size(shape, S),), np.uint8)
    t_s, r_s = tsf.DeviceArray(shape, np.float32), tsf.DeviceArray(shape, np.float32)
    t_l, r_l = tsf.DeviceArray(shape, np.float32), tsf.DeviceArray(shape, np.float32)
    tws = tsf.DeviceArray((tsf.tdf_circle_workspace_size(shape, S),), np.uint8)
    ev = [tsf.Event() for _ in range(4)]

    if args.pmc_workload:
        # Fixed launches for counter passes: k_gvf_copy_finish on F32 planes is
        # the calibration copy (reads 3 x 4 B, writes 4 x 4 B per voxel), then
        # GVF steps in the benchmark storage (and FP32 with one, two and four
        # steps per launch) and both TDF passes.
        F32c, V32c = tsf.Field(shape), tsf.Field(shape)
        tsf.blur_gaussian(d_vol, d_blur, 0.5, stream)
        tsf.create_vector_field(d_blur, F32c, 0.2, 1, stream)
        tsf.gvf_euler(F32c, V32c, args.mu, 0, 1, None, stream)
        tsf.create_vector_field(d_blur, F, 0.2, 1, stream)
        tsf.gvf_euler(F, V, args.mu, 4, 1, ws, stream)  # single steps
        tsf.gvf_euler(F, V, args.mu, 8, 2, ws, stream)  # two steps per launch
        tsf.gvf_euler(F32c, V32c, args.mu, 2, 1, None, stream)
        tsf.gvf_euler(F32c, V32c, args.mu, 4, 2, None, stream)
        tsf.gvf_euler(F32c, V32c, args.mu, 8, 4, None, stream)  # four steps per launch
        tsf.tdf_circle(F, 0.5, 3.0, 0.5, t_s, r_s, tws, stream)
        tsf.tdf_circle(V, 2.5, 4.0, 0.5, t_l, r_l, tws, stream)
        stream.synchronize()
        log("pmc workload done")
        return

    def forward(Fs, Fl, Vl, gws, tws, fuse):
        ev[0].record(stream)
        # small-radius branch (radius-min < 2.5)
        tsf.blur_gaussian(d_vol, d_blur, 0.5, stream)
        tsf.create_vector_field(d_blur, Fs, 0.2, 1, stream)
        tsf.tdf_circle(Fs, 0.5, 3.0, 0.5, t_s, r_s, tws, stream)
        # large-radius branch
        tsf.blur_gaussian(d_vol, d_blur, 0.5, stream)
        tsf.create_vector_field(d_blur, Fl, 0.2, 1, stream)
        ev[1].record(stream)
        tsf.gvf_euler(Fl, Vl, args.mu, args.iters, fuse, gws, stream)
        ev[2].record(stream)
        tsf.tdf_circle(Vl, 2.5, 4.0, 0.5, t_l, r_l, tws, stream)
        tsf.combine(t_s, r_s, t_l, r_l, stream)
        ev[3].record(stream)

    def step():
        forward(F_small, F, V, ws, tws, args.fuse)

    for i in range(args.warmup):
        step()
        stream.synchronize()
        log(f"warmup {i} done")
    fwd, gvf = [], []
    stream.synchronize()
    wall0 = time.time()
    for i in range(args.steps):
        step()
        stream.synchronize()
        fwd.append(ev[0].elapsed_ms(ev[3]))
        gvf.append(ev[1].elapsed_ms(ev[2]))
        log(f"step {i}: forward {fwd[-1]:.1f} ms, GVF {gvf[-1]:.1f} ms")
    stream.synchronize()
    wall = time.time() - wall0

    gvf_ms = float(np.median(gvf))
    fwd_ms = float(np.median(fwd))
    gvox = N * args.iters / (gvf_ms * 1e-3) / 1e9
    # launches: fours, then at most one two-step and one single (as tsf_gvf_euler schedules them)
    spl = args.fuse
    quads = args.iters // 4 if spl == 4 else 0
    pairs = (args.iters - 4 * quads) // 2 if spl >= 2 else 0
    launches = max(1, quads + pairs + (args.iters - 4 * quads - 2 * pairs))
    bytes_per_voxel = 9 * (4 if S == tsf.STORAGE_F32 else 2)  # read v, f; write v (3 components each)
    launch_ms = gvf_ms / launches
    achieved = bytes_per_voxel * N / (launch_ms * 1e-3)

    # the stages around the filters, on this forward's TDF, radius and field
    # (before the FP32 comparison below reuses the TDF and radius buffers)
    downstream = None
    if not args.no_downstream:
        downstream = downstream_stages(vol, shape, t_s, r_s, V, stream)

    # the same forward with float storage, for the speed-up of the
    # reduced-precision form: two steps per launch, and the plain FP32 path
    # (one step per launch)
    f32 = None
    if not args.no_f32_compare and S != tsf.STORAGE_F32:
        Fs32, F32, V32 = tsf.Field(shape), tsf.Field(shape), tsf.Field(shape)
        ws32 = tsf.DeviceArray((tsf.gvf_euler_workspace_size(shape, 0),), np.uint8)
        tws32 = tsf.DeviceArray((tsf.tdf_circle_workspace_size(shape, tsf.STORAGE_F32),), np.uint8)
        res = {}
        f32_fuse = 4  # four steps per launch: the fastest FP32 form
        for fz in (f32_fuse, 1):
            forward(Fs32, F32, V32, ws32, tws32, fz)  # warm
            stream.synchronize()
            tf, tg = [], []
            for _ in range(max(1, args.steps)):
                forward(Fs32, F32, V32, ws32, tws32, fz)
                stream.synchronize()
                tf.append(ev[0].elapsed_ms(ev[3]))
                tg.append(ev[1].elapsed_ms(ev[2]))
            res[fz] = (float(np.median(tf)), float(np.median(tg)))
        (f32_fwd, f32_ms), (f32_fwd1, f32_1) = res[f32_fuse], res[1]
        f32 = {"steps_per_launch": f32_fuse, "gvf_ms": round(f32_ms, 3),
               "gvoxel_iter_per_s": round(N * args.iters / (f32_ms * 1e-3) / 1e9, 2),
               "forward_ms": round(f32_fwd, 3),
               "speedup_reduced_vs_f32": round(f32_ms / gvf_ms, 3),
               "hbm_frac_per_launch": round(36 * N / (f32_ms / (args.iters // f32_fuse) * 1e-3) / PEAK_HBM, 4),
               "single_step_gvf_ms": round(f32_1, 3),
               "single_step_gvoxel_iter_per_s": round(N * args.iters / (f32_1 * 1e-3) / 1e9, 2),
               "single_step_forward_ms": round(f32_fwd1, 3),
               "speedup_reduced_vs_f32_single_step": round(f32_1 / gvf_ms, 3),
               "forward_speedup_reduced_vs_plain_f32": round(f32_fwd1 / fwd_ms, 3)}
        del Fs32, F32, V32, ws32, tws32

    adjoint = None
    if not args.no_adjoint:
        del F_small, F, V, ws, tws, t_s, r_s, t_l, r_l  # the forward's buffers
        adjoint = adjoint_step(args, d_vol, shape, stream)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(vol)

    # HBM bytes per launch of the dominant kernel, measured by the two
    # rocprofv3 counter passes recorded in profiles/ (null when not measured
    # for this storage type)
    traffic = None
    gk = {4: "k_gvf_fused4<{}, false>", 2: "k_gvf_fused2<{}, false>"}.get(args.fuse, "k_gvf_march<{}, 2, false>")
    kname = {"f16": gk.format("F16"), "f32": gk.format("F32"), "sn16": gk.format("SN16"),
             "bf16": gk.format("BF16")}.get(args.storage)
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic_circle_forward_512cubed.json")
    if kname and n == 512 and os.path.exists(tpath):
        with open(tpath) as fh:
            rec = json.load(fh)["kernels"].get(kname)
        if rec:
            traffic = round(rec["traffic_bytes"] / 1e9, 4)

    out = {
        "metric": f"GVF Gvoxel-iter/s ({args.storage.upper()}-stored field, {n}^3, mu {args.mu}, {args.iters} iterations)",
        "value": round(gvox, 3),
        "unit": "Gvoxel-iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(fwd_ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f16": "f16-storage/f32-arith", "bf16": "bf16-storage/f32-arith", "sn16": "snorm16-storage/f32-arith",
                  "f32": "f32"}[args.storage],
        "data": "synthetic (seeded 512^3 thin-vessel MRA-like tree, Rician noise)",
        "config": {"workload": f"circle-fitting forward at {n}^3: blur, vector field, GVF Euler "
                               f"{args.iters} it mu {args.mu} ({args.storage} storage), TDF small 0.5-3 + large "
                               f"2.5-4, combine", "voxels": N, "gvf_iterations": args.iters,
                   "fuse": args.fuse, "parallelism": "single-gpu"},
        "gvf_ms": round(gvf_ms, 3),
        "forward_ms": round(fwd_ms, 3),
        "wall_s_timed": round(wall, 3),
        "f32_gvf": f32,
        "adjoint": adjoint,
        "downstream": downstream,
        "roofline": {"bound": "hbm", "achieved": round(achieved / 1e9, 1), "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM, 4), "traffic": traffic, "traffic_unit": "GB/launch",
                     "algorithmic_gb_per_launch": round(bytes_per_voxel * N / 1e9, 4),
                     "kernel": kname or "k_gvf", "steps_per_launch": args.fuse,
                     "bytes_per_voxel_per_launch": bytes_per_voxel,
                     "launch_ms": round(launch_ms, 4)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def downstream_stages(vol, shape, tdf, radius, field, stream):
    """Times the stages either side of the filters on the same volume, each
    from device-resident inputs: ingest of the raw volume as 16-bit MRA
    intensities (limits, threshold-cropping scan lines, toFloat of the
    region), the spline TDF on the GVF field, the centerline graph from the forward's TDF, radius and GVF
    field (candidates, compaction, suppression, vertices, linking, components,
    rasterisation; the Neuro-Vessels-MRA preset's thresholds) and the
    inverse-gradient segmentation of that centerline."""
    import tsf
    N = int(np.prod(shape))
    ev = [tsf.Event() for _ in range(2)]

    def timed(fn, reps=3):
        fn()
        stream.synchronize()
        ts = []
        for _ in range(reps):
            ev[0].record(stream)
            r = fn()
            ev[1].record(stream)
            stream.synchronize()
            ts.append(ev[0].elapsed_ms(ev[1]))
        return float(np.median(ts)), r

    out = {}
    # ingest: 16-bit intensities 100..300 (the MRA preset's window)
    raw = tsf.DeviceArray.from_host(np.round(vol * 200.0 + 100.0).astype(np.uint16), stream=stream)
    ms_lim, lim = timed(lambda: tsf.volume_limits(raw, stream))
    ms_scan, lines = timed(lambda: tsf.crop_scan_lines(raw, tsf.CROP_THRESHOLD, 150.0, stream))
    off, size = tsf.crop_region(*lines, 10, False)
    dst = tsf.DeviceArray((size[2], size[1], size[0]), np.float32)
    ms_tf, _ = timed(lambda: tsf.to_float_region(raw, off, dst, 100.0, 300.0, stream))
    Nc = size[0] * size[1] * size[2]
    out["ingest"] = {
        "raw": "uint16", "limits": [lim[0], lim[1]], "crop_offset": off, "crop_size": size,
        "limits_ms": round(ms_lim, 3), "limits_gb_s": round(2 * N / (ms_lim * 1e-3) / 1e9, 1),
        "scan_lines_ms": round(ms_scan, 3), "scan_lines_gb_s": round(2 * N / (ms_scan * 1e-3) / 1e9, 1),
        "to_float_region_ms": round(ms_tf, 3), "to_float_region_gb_s": round(6 * Nc / (ms_tf * 1e-3) / 1e9, 1),
        "bytes": "limits 2 B/voxel read; threshold scan lines 2 B/voxel read (z-lines and x-lines in one pass); "
                 "toFloat 2 B read + 4 B written per cropped voxel; times include the host synchronisation"}
    del raw, dst

    # the spline TDF in place of the large circle TDF (use-spline-tdf: radii
    # from 1.5), on the forward's GVF field
    st, sr = tsf.DeviceArray(shape, np.float32), tsf.DeviceArray(shape, np.float32)
    sws = tsf.DeviceArray((tsf.tdf_circle_workspace_size(shape, field.storage),), np.uint8)
    ms_spl, _ = timed(lambda: tsf.tdf_spline(field, 1.5, 4.0, 0.5, st, sr, 0.1, sws, stream), reps=2)
    out["spline_tdf"] = {"ms": round(ms_spl, 3), "radii": "1.5-4.0 step 0.5, 12 arms",
                         "voxels_tdf_positive": int((st.download(stream) > 0).sum())}
    del st, sr, sws

    # centerline graph
    d = {}
    mask = tsf.DeviceArray(shape, np.uint8)
    hws = tsf.DeviceArray((tsf.hp_compact_workspace_size(shape),), np.uint8)

    def candidates():
        tsf.candidates(tdf, mask, 0.5, stream)
        c = tsf.hp_compact(mask, None, hws, stream)
        pos = tsf.DeviceArray((max(c, 1), 3), np.int32)
        tsf.hp_compact(mask, pos, hws, stream)
        cp = tsf.DeviceArray(shape, np.uint8)
        cp.zero(stream)
        tsf.candidate_nms(tdf, radius, field, pos, c, 8, cp, stream)
        cells = tsf.DeviceArray(shape, np.uint8)
        cells.zero(stream)
        tsf.cube_max(tdf, cp, 5, cells, stream)
        nv = tsf.hp_compact(cells, None, hws, stream)
        verts = tsf.DeviceArray((max(nv, 1), 3), np.int32)
        tsf.hp_compact(cells, verts, hws, stream)
        return c, nv, verts

    ms_cand, (ncand, nv, verts) = timed(candidates, reps=2)
    lws = tsf.DeviceArray((tsf.link_centerpoints_workspace_size(nv),), np.uint8)
    cl = tsf.DeviceArray(shape, np.uint8)

    def graph():
        ne = tsf.link_centerpoints(verts, nv, tdf, 10.0, 0.5, None, lws, stream)
        edges = tsf.DeviceArray((max(ne, 1), 2), np.int32)
        tsf.link_centerpoints(verts, nv, tdf, 10.0, 0.5, edges, lws, stream)
        labels, sizes = tsf.DeviceArray((nv,), np.int32), tsf.DeviceArray((nv,), np.int32)
        tsf.graph_components(edges, ne, nv, labels, sizes, stream)
        cl.zero(stream)
        tsf.rasterize_tree_edges(edges, ne, verts, labels, sizes, 5, cl, stream)
        return ne

    ms_graph, ne = timed(graph, reps=2)
    d.update({"candidates": ncand, "vertices": nv, "edges": ne, "candidates_to_vertices_ms": round(ms_cand, 3),
              "link_components_rasterise_ms": round(ms_graph, 3),
              "centerline_voxels": int(cl.download(stream).sum())})
    del mask, hws, lws
    out["centerline"] = d

    # segmentation of that centerline
    seg = tsf.DeviceArray(shape, np.uint8)
    sws = tsf.DeviceArray((tsf.segment_inverse_gradient_workspace_size(shape),), np.uint8)
    ms_seg, passes = timed(lambda: tsf.segment_inverse_gradient(cl, radius, field, seg, sws, stream), reps=2)
    out["segmentation"] = {"ms": round(ms_seg, 3), "growing_passes": passes,
                           "ms_per_pass": round(ms_seg / max(passes, 1), 3),
                           "voxels": int(seg.download(stream).sum())}
    log(f"downstream: {out}")
    return out


CHAIN = dict(sigma=0.5, rmin=0.5, rmax=4.0, rstep=0.5)


def adjoint_step(args, d_vol, shape, stream):
    """One forward + backward of the FP32 chain, timed with events on the
    chain's stream; a first step warms up.  dL/dT = 1 where T > 0 (every
    responding voxel carries gradient)."""
    import tsf
    from tsf.chain import Ctx, TubeFilterFunction as TF
    free0, total = tsf.mem_info()
    ev = [tsf.Event() for _ in range(3)]
    kw = dict(mu=args.mu, sigma=CHAIN["sigma"], iterations=args.iters, rmin=CHAIN["rmin"], rmax=CHAIN["rmax"],
              rstep=CHAIN["rstep"], checkpoint_every=args.checkpoint_every, stream=stream)
    res = []
    for rep in range(2):
        ctx = Ctx()
        ev[0].record(stream)
        T, R = TF.forward(ctx, d_vol, **kw)
        ev[1].record(stream)
        stream.synchronize()
        if rep == 0:
            gT = tsf.DeviceArray.from_host((T.download() > 0).astype(np.float32))
        ev[1].record(stream)
        tb = time.time()
        _, dmu, dsigma = TF.backward(ctx, gT, stream)
        ev[2].record(stream)
        stream.synchronize()
        res.append((ev[0].elapsed_ms(ev[1]), ev[1].elapsed_ms(ev[2]), (time.time() - tb) * 1e3,
                    ctx.backward_used_bytes))
        log(f"adjoint step {rep}: forward {res[-1][0]:.1f} ms, backward {res[-1][1]:.1f} ms, "
            f"dmu {dmu:.5g}, dsigma {dsigma:.5g}")
        del ctx, T, R
    fwd_ms, bwd_ms, bwd_wall, used = res[-1]
    return {"workload": f"FP32 chain at {shape[0]}^3: blur {CHAIN['sigma']}, vector field, GVF {args.iters} it "
                        f"mu {args.mu} checkpointed every {args.checkpoint_every}, TDF {CHAIN['rmin']}-"
                        f"{CHAIN['rmax']} step {CHAIN['rstep']}; backward to dvolume, dmu, dsigma",
            "forward_ms": round(fwd_ms, 2), "backward_ms": round(bwd_ms, 2),
            "step_ms": round(fwd_ms + bwd_ms, 2), "backward_wall_ms": round(bwd_wall, 2),
            "hbm_peak_gb": round(used / 1e9, 2), "hbm_chain_gb": round((used - (total - free0)) / 1e9, 2),
            "hbm_total_gb": round(total / 1e9, 1)}


def fit(args, vol, d_vol, shape, stream):
    import tsf
    from tsf import synth
    from tsf.chain import Ctx, TubeFilterFunction as TF, fit_mu_sigma
    clean, _ = synth.mra_thin(shape, 120, 505, scale=0.0)
    d_clean = tsf.DeviceArray.from_host(clean, stream=stream)
    ctx = Ctx()
    target, _ = TF.forward(ctx, d_clean, 0.05, 0.5, args.iters, CHAIN["rmin"], CHAIN["rmax"], CHAIN["rstep"],
                           checkpoint_every=args.iters, stream=stream)
    stream.synchronize()
    del ctx, d_clean
    t0 = time.time()
    hist, mu, sigma = fit_mu_sigma(d_vol, target, 0.03, 0.8, args.fit_steps, args.iters, CHAIN["rmin"],
                                   CHAIN["rmax"], CHAIN["rstep"], checkpoint_every=args.checkpoint_every,
                                   stream=stream, log=log)
    wall = time.time() - t0
    ms = [h["ms"] for h in hist]
    print(json.dumps({
        "metric": "mu/sigma fit: ms per gradient-descent step (FP32 chain forward + adjoint)",
        "value": round(float(np.median(ms)), 1), "unit": "ms/step", "higher_is_better": False,
        "steps": args.fit_steps, "wall_s": round(wall, 1),
        "config": {"workload": f"fit mu, sigma at {shape[0]}^3 to the noise-free TDF (mu 0.05, sigma 0.5); "
                               f"GVF {args.iters} it checkpointed every {args.checkpoint_every}, TDF "
                               f"{CHAIN['rmin']}-{CHAIN['rmax']}", "start": {"mu": 0.03, "sigma": 0.8}},
        "loss_first": hist[0]["loss"], "loss_last": hist[-1]["loss"], "mu": mu, "sigma": sigma,
        "hbm_peak_gb": round(max(h["hbm_bytes"] or 0 for h in hist) / 1e9, 2),
        "history": [{k: (round(v, 6) if isinstance(v, float) else v) for k, v in h.items() if k != "hbm_bytes"}
                    for h in hist]}), flush=True)


def cpu_baseline(vol):
    """Oracle FP32 GVF (OpenMP, all host threads) on a bounded sample: the
    central 256^3 block of the same volume after the same blur and vector
    field, timed over enough iterations for ~12 s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    c = min(256, vol.shape[0])
    o = (vol.shape[0] - c) // 2
    sub = np.ascontiguousarray(vol[o:o + c, o:o + c, o:o + c])
    F = oracle.vector_field(oracle.blur(sub, 0.5, "f32"), 0.2, 1, "f32")
    oracle.gvf_euler(F, 0.05, 4, "f32")  # thread pool and page warm-up
    # the cost per iteration from the difference of 48 and 16 iterations (a
    # call also has a fixed cost: copies in and out), then ~12 s timed
    t = []
    for n in (16, 48):
        t0 = time.time()
        oracle.gvf_euler(F, 0.05, n, "f32")
        t.append(time.time() - t0)
    iters = max(16, int(12.0 / max((t[1] - t[0]) / 32, 1e-3)))
    t0 = time.time()
    oracle.gvf_euler(F, 0.05, iters, "f32")
    dt = time.time() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(sub.size * iters / dt / 1e9, 4), "unit": "Gvoxel-iter/s", "cores": threads,
            "kind": "port", "sample": f"oracle FP32 GVF Euler, {c}^3 centre block, {iters} iterations, {dt:.1f}s"}


if __name__ == "__main__":
    main()
