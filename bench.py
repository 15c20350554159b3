"""Benchmark of the volume-filter hot path on one MI355X.

Workload (one "step"): the circle-fitting forward of runCircleFittingMethod
(tube-segmentation.cpp:212-939) on a seeded 512^3 thin-vessel MRA-like volume
(120 tubes, radius 0.5-3, Rician noise 0.08): small-scale branch (blur 0.5 ->
vector field -> TDF r = 0.5..3.0) and large-scale branch (blur 0.5 -> vector
field -> GVF Euler, mu = 0.05, 400 iterations -> TDF r = 2.5..4.0, step 0.5)
-> combine.

The default is the precision form: the large branch's vector field stored
as SNORM16 (the reference's 16-bit vector format), the GVF iterate streamed
as F16 with stochastic rounding, two steps per launch (tsf_gvf_euler_sr),
the last --tail iterations in float with the final iterate F32; both TDFs
take the Hessian from F32 planes and the circle samples from F16 texels
(tsf_tdf_circle_mixed); the small branch's vector field is F32.  --form
plain runs the fields in one storage type throughout (tsf_gvf_euler,
round-to-nearest 16-bit stores).

Headline value: GVF throughput in Gvoxel-iter/s (voxels x iterations / GVF
time, float tail included).  Also reported: forward ms per step; the same
forward in FP32 with four GVF steps per launch (the fastest FP32 form) and
with one (the plain FP32 path), and the speed-ups; the roofline of the
dominant kernel (the two-step 16-bit launch: its average duration from the
difference of two GVF runs that differ only in the number of such launches,
timed with HIP events on its stream; its algorithmic bytes, 18 per voxel: v
and f read, v written, 3 x 2 bytes each); the errors against FP64 measured
by tests/test_gpu_precision.py (recorded in profiles/); the oracle's FP32
GVF on the host CPU; one forward + adjoint step of the differentiable FP32
chain at the same size (blur 0.5 -> vector field -> GVF 400 iterations
checkpointed every 40 -> TDF r = 0.5..4.0, backward to dL/dvolume, dL/dmu,
dL/dsigma) with its device memory at the backward's peak.

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
SR_SEED = 7


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
    ap.add_argument("--form", choices=["precision", "plain"], default="precision",
                    help="precision: 16-bit GVF iterate with stochastic rounding, float tail, mixed TDF; "
                         "plain: every field in --storage")
    ap.add_argument("--storage", choices=["sn16", "f16", "bf16", "f32"], default=None,
                    help="precision form: storage of the GVF's f (sn16, the default: iterate F16; f16; bf16); "
                         "plain form: storage of every field (default f16; sn16 is the reference's own format)")
    ap.add_argument("--tail", type=int, default=32, help="precision form: float iterations at the end of the GVF")
    ap.add_argument("--fuse", type=int, default=2, choices=[1, 2, 4], help="plain form: GVF steps per launch")
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
    precision = args.form == "precision"
    if args.storage is None:
        args.storage = "sn16" if precision else "f16"
    if precision and args.storage == "f32":
        ap.error("the precision form's f is sn16, f16 or bf16")

    import tsf
    from tsf import synth

    if tsf.device_count() < 1:
        raise RuntimeError("no GPU visible")
    S = {"f32": tsf.STORAGE_F32, "f16": tsf.STORAGE_F16, "bf16": tsf.STORAGE_BF16,
         "sn16": tsf.STORAGE_SN16}[args.storage]
    scale = 2.0 ** 14 if S == tsf.STORAGE_F16 else 1.0
    # the precision form's 16-bit texels of the TDF samples
    TS, TSC = (tsf.STORAGE_BF16, 1.0) if args.storage == "bf16" else (tsf.STORAGE_F16, 2.0 ** 14)
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
    F = tsf.Field(shape, S, scale)  # the large branch's vector field (the GVF's f)
    t_s, r_s = tsf.DeviceArray(shape, np.float32), tsf.DeviceArray(shape, np.float32)
    t_l, r_l = tsf.DeviceArray(shape, np.float32), tsf.DeviceArray(shape, np.float32)
    ev = [tsf.Event() for _ in range(4)]
    if precision:
        F_small = tsf.Field(shape)
        V = tsf.Field(shape)
        ws = tsf.DeviceArray((tsf.gvf_euler_sr_workspace_size(shape, args.tail),), np.uint8)
        tws = tsf.DeviceArray((tsf.tdf_circle_mixed_workspace_size(shape),), np.uint8)
    else:
        F_small = tsf.Field(shape, S, scale)
        V = tsf.Field(shape, S, scale)
        ws = tsf.DeviceArray((tsf.gvf_euler_workspace_size(shape, S),), np.uint8)
        tws = tsf.DeviceArray((tsf.tdf_circle_workspace_size(shape, S),), np.uint8)

    if args.pmc_workload:
        # Fixed launches for counter passes: k_sr_finish (reads 3 x 4 B, writes
        # 4 B per voxel) calibrates the memory-side byte counters; then two
        # GVF runs of the precision form (SNORM16 f: a first launch reading
        # v0 from f, two body launches, a last one writing F32), the plain
        # FP32 GVF with four steps per launch, and both mixed TDF passes.
        Fsn = tsf.Field(shape, tsf.STORAGE_SN16)
        F32c, V32c = tsf.Field(shape), tsf.Field(shape)
        tsf.blur_gaussian(d_vol, d_blur, 0.5, stream)
        tsf.create_vector_field(d_blur, F32c, 0.2, 1, stream)
        tsf.create_vector_field(d_blur, Fsn, 0.2, 1, stream)
        sws = tsf.DeviceArray((tsf.gvf_euler_sr_workspace_size(shape, 4),), np.uint8)
        tsf.gvf_euler_sr(Fsn, V32c, args.mu, 8, SR_SEED, 4, sws, stream)  # body, then 4 float steps (k_sr_finish)
        tsf.gvf_euler_sr(Fsn, V32c, args.mu, 8, SR_SEED, 0, sws, stream)
        tsf.gvf_euler(F32c, V32c, args.mu, 8, 4, None, stream)
        tsf.tdf_circle_mixed(F32c, 0.5, 3.0, 0.5, t_s, r_s, tsf.STORAGE_F16, 2.0 ** 14, None, stream)
        tsf.tdf_circle_mixed(V32c, 2.5, 4.0, 0.5, t_l, r_l, tsf.STORAGE_F16, 2.0 ** 14, None, stream)
        stream.synchronize()
        log("pmc workload done")
        return

    def gvf(Fl, Vl, gws, iters, tail=None, fuse=None):
        if precision and fuse is None:
            tsf.gvf_euler_sr(Fl, Vl, args.mu, iters, SR_SEED, args.tail if tail is None else tail, gws, stream)
        else:
            tsf.gvf_euler(Fl, Vl, args.mu, iters, fuse or args.fuse, gws, stream)

    def forward(Fs, Fl, Vl, gws, tws, fuse=None):
        mixed = precision and fuse is None
        ev[0].record(stream)
        # small-radius branch (radius-min < 2.5)
        tsf.blur_gaussian(d_vol, d_blur, 0.5, stream)
        tsf.create_vector_field(d_blur, Fs, 0.2, 1, stream)
        if mixed:
            tsf.tdf_circle_mixed(Fs, 0.5, 3.0, 0.5, t_s, r_s, TS, TSC, tws, stream)
        else:
            tsf.tdf_circle(Fs, 0.5, 3.0, 0.5, t_s, r_s, tws, stream)
        # large-radius branch
        tsf.blur_gaussian(d_vol, d_blur, 0.5, stream)
        tsf.create_vector_field(d_blur, Fl, 0.2, 1, stream)
        ev[1].record(stream)
        gvf(Fl, Vl, gws, args.iters, fuse=fuse)
        ev[2].record(stream)
        if mixed:
            tsf.tdf_circle_mixed(Vl, 2.5, 4.0, 0.5, t_l, r_l, TS, TSC, tws, stream)
        else:
            tsf.tdf_circle(Vl, 2.5, 4.0, 0.5, t_l, r_l, tws, stream)
        tsf.combine(t_s, r_s, t_l, r_l, stream)
        ev[3].record(stream)

    for i in range(args.warmup):
        forward(F_small, F, V, ws, tws)
        stream.synchronize()
        log(f"warmup {i} done")
    fwd, gv = [], []
    stream.synchronize()
    wall0 = time.time()
    for i in range(args.steps):
        forward(F_small, F, V, ws, tws)
        stream.synchronize()
        fwd.append(ev[0].elapsed_ms(ev[3]))
        gv.append(ev[1].elapsed_ms(ev[2]))
        log(f"step {i}: forward {fwd[-1]:.1f} ms, GVF {gv[-1]:.1f} ms")
    stream.synchronize()
    wall = time.time() - wall0
    gvf_ms = float(np.median(gv))
    fwd_ms = float(np.median(fwd))
    gvox = N * args.iters / (gvf_ms * 1e-3) / 1e9

    # the dominant kernel's average launch: two GVF runs that differ only in
    # K launches of it (each followed by its small face-copy kernel)
    def gvf_time(iters, reps=3):
        ts = []
        for _ in range(reps):
            ev[1].record(stream)
            gvf(F, V, ws, iters, tail=0)
            ev[2].record(stream)
            stream.synchronize()
            ts.append(ev[1].elapsed_ms(ev[2]))
        return float(np.median(ts))

    if precision:
        K = 50
        spl = 2
        launch_ms = (gvf_time(2 * K + 4) - gvf_time(4)) / K
        kname = {"sn16": "k_gvf_sr2<F16, F16, SN16, false>", "f16": "k_gvf_sr2<F16, F16, F16, false>",
                 "bf16": "k_gvf_sr2<BF16, BF16, BF16, false>"}[args.storage]
        bytes_per_voxel = 18  # v and f read, v written: 3 components x 2 bytes each
    else:
        spl = args.fuse
        K = 48 // spl
        launch_ms = (gvf_time(spl * (K + 2)) - gvf_time(2 * spl)) / K
        gk = {4: "k_gvf_fused4<{}, false>", 2: "k_gvf_fused2<{}, false>"}.get(spl, "k_gvf_march<{}, 2, false>")
        kname = gk.format(args.storage.upper())
        bytes_per_voxel = 9 * (4 if S == tsf.STORAGE_F32 else 2)
    achieved = bytes_per_voxel * N / (launch_ms * 1e-3)

    # the stages around the filters, on this forward's TDF, radius and field
    # (before the FP32 comparison below reuses the TDF and radius buffers)
    downstream = None
    if not args.no_downstream:
        downstream = downstream_stages(vol, shape, t_s, r_s, V, stream)

    # the same forward in FP32: four GVF steps per launch (the fastest FP32
    # form) and one (the plain FP32 path)
    f32 = None
    if not args.no_f32_compare and not (S == tsf.STORAGE_F32 and not precision):
        Fs32, F32, V32 = tsf.Field(shape), tsf.Field(shape), tsf.Field(shape)
        ws32 = tsf.DeviceArray((tsf.gvf_euler_workspace_size(shape, 0),), np.uint8)
        tws32 = tsf.DeviceArray((tsf.tdf_circle_workspace_size(shape, tsf.STORAGE_F32),), np.uint8)
        res = {}
        for fz in (4, 1):
            forward(Fs32, F32, V32, ws32, tws32, fz)  # warm
            stream.synchronize()
            tf, tg = [], []
            for _ in range(max(1, args.steps)):
                forward(Fs32, F32, V32, ws32, tws32, fz)
                stream.synchronize()
                tf.append(ev[0].elapsed_ms(ev[3]))
                tg.append(ev[1].elapsed_ms(ev[2]))
            res[fz] = (float(np.median(tf)), float(np.median(tg)))
            log(f"FP32 forward, {fz} GVF steps per launch: {res[fz][0]:.1f} ms, GVF {res[fz][1]:.1f} ms")
        (f4_fwd, f4_gvf), (f1_fwd, f1_gvf) = res[4], res[1]
        f32 = {"fastest": {"gvf_steps_per_launch": 4, "gvf_ms": round(f4_gvf, 3), "forward_ms": round(f4_fwd, 3),
                           "gvoxel_iter_per_s": round(N * args.iters / (f4_gvf * 1e-3) / 1e9, 2),
                           "hbm_frac_per_launch": round(36 * N / (f4_gvf / (args.iters // 4) * 1e-3) / PEAK_HBM, 4)},
               "plain": {"gvf_steps_per_launch": 1, "gvf_ms": round(f1_gvf, 3), "forward_ms": round(f1_fwd, 3),
                         "gvoxel_iter_per_s": round(N * args.iters / (f1_gvf * 1e-3) / 1e9, 2)},
               "speedup_gvf_vs_fastest_f32": round(f4_gvf / gvf_ms, 3),
               "speedup_forward_vs_fastest_f32": round(f4_fwd / fwd_ms, 3),
               "speedup_gvf_vs_plain_f32": round(f1_gvf / gvf_ms, 3),
               "speedup_forward_vs_plain_f32": round(f1_fwd / fwd_ms, 3)}
        del Fs32, F32, V32, ws32, tws32

    adjoint = None
    if not args.no_adjoint:
        del F_small, F, V, ws, tws, t_s, r_s, t_l, r_l  # the forward's buffers
        adjoint = adjoint_step(args, d_vol, shape, stream)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(vol)

    # HBM bytes per launch of the dominant kernel: memory-side read and write
    # counters of the committed kernels, recorded in profiles/ with the
    # command that produced them (null when not measured for this kernel)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic_512cubed.json")
    if n == 512 and os.path.exists(tpath):
        with open(tpath) as fh:
            rec = json.load(fh)["kernels"].get(kname)
        if rec:
            traffic = round(rec["bytes_per_dispatch"] / 1e9, 4)
    # errors against FP64 (tests/test_gpu_precision.py, C5 and C3 crops)
    errors = None
    epath = os.path.join(ROOT, "profiles", "precision_vs_fp64.json")
    if os.path.exists(epath):
        with open(epath) as fh:
            errors = json.load(fh)["forms"].get(f"precision/{args.storage}/tail{args.tail}" if precision else
                                                f"plain/{args.storage}")

    form = (f"precision form: f {args.storage.upper()}, iterate {'BF16' if args.storage == 'bf16' else 'F16'} "
            f"stochastic rounding, last {args.tail} iterations F32, mixed TDF" if precision else
            f"{args.storage} storage, round to nearest, {args.fuse} steps per launch")
    out = {
        "metric": f"GVF Gvoxel-iter/s ({n}^3, mu {args.mu}, {args.iters} iterations; {form})",
        "value": round(gvox, 3),
        "unit": "Gvoxel-iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(fwd_ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f32-arith; GVF iterate f16 (bf16) stochastic rounding, f " + args.storage + ", TDF samples f16"
                  if precision else {"f16": "f16-storage/f32-arith", "bf16": "bf16-storage/f32-arith",
                                     "sn16": "snorm16-storage/f32-arith", "f32": "f32"}[args.storage]),
        "data": "synthetic (seeded 512^3 thin-vessel MRA-like tree, Rician noise)",
        "config": {"workload": f"circle-fitting forward at {n}^3: blur, vector field, GVF Euler "
                               f"{args.iters} it mu {args.mu}, TDF small 0.5-3 + large 2.5-4, combine ({form})",
                   "voxels": N, "gvf_iterations": args.iters, "form": args.form, "storage": args.storage,
                   "f32_tail": args.tail if precision else None, "parallelism": "single-gpu"},
        "gvf_ms": round(gvf_ms, 3),
        "forward_ms": round(fwd_ms, 3),
        "wall_s_timed": round(wall, 3),
        "f32": f32,
        "errors_vs_fp64": errors,
        "adjoint": adjoint,
        "downstream": downstream,
        "roofline": {"bound": "hbm", "achieved": round(achieved / 1e9, 1), "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM, 4), "traffic": traffic, "traffic_unit": "GB/launch",
                     "algorithmic_gb_per_launch": round(bytes_per_voxel * N / 1e9, 4),
                     "kernel": kname, "steps_per_launch": spl, "bytes_per_voxel_per_launch": bytes_per_voxel,
                     "launch_ms": round(launch_ms, 4),
                     "launch_ms_method": f"difference of two GVF runs {K} launches apart, HIP events"},
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
        "history": [{k: (float(f"{v:.6g}") if isinstance(v, float) else v) for k, v in h.items() if k != "hbm_bytes"}
                    for h in hist]}), flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_stages(oracle, vol):
    """The forward's stages in the oracle's FP32 restatement (all threads) on
    the central 128^3 block: median of 3 after one warm-up pass, ms."""
    c = min(128, vol.shape[0])
    o = (vol.shape[0] - c) // 2
    sub = np.ascontiguousarray(vol[o:o + c, o:o + c, o:o + c])
    names = ["blur", "create vector field", "TDF small (r 0.5-3)", "GVF (mu 0.05, 400 iterations)",
             "TDF large (r 2.5-4)", "combine", "candidates + HP compaction"]
    runs = []
    for rep in range(4):
        t = []
        def timed(f):
            t0 = time.perf_counter()
            r = f()
            t.append((time.perf_counter() - t0) * 1e3)
            return r
        b = timed(lambda: oracle.blur(sub, 0.5, "f32"))
        F = timed(lambda: oracle.vector_field(b, 0.2, 1, "f32"))
        ts, rs = timed(lambda: oracle.tdf_circle(F, 0.5, 3.0, 0.5, "f32"))
        V = timed(lambda: oracle.gvf_euler(F, 0.05, 400, "f32"))
        tl, rl = timed(lambda: oracle.tdf_circle(V, 2.5, 4.0, 0.5, "f32"))
        tc, _ = timed(lambda: oracle.combine(ts, rs, tl, rl, "f32"))
        timed(lambda: oracle.hp_compact(oracle.candidates(tc, 0.5, "f32"), "pyramid"))
        if rep:
            runs.append(t)
    return {"sample": f"{c}^3 centre block, median of 3 after a warm-up",
            "ms": {n: round(float(np.median([r[i] for r in runs])), 1) for i, n in enumerate(names)}}


def cpu_baseline(vol):
    """Oracle FP32 GVF (OpenMP, all host threads) on a bounded sample: the
    central 256^3 block of the same volume after the same blur and vector
    field, timed over enough iterations for ~12 s.  Also: the same GVF on
    one thread (the 128^3 block, ~3 s), the forward's stages on the 128^3
    block, and the CPU model."""
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
    threads = oracle.set_threads()
    out = {"value": round(sub.size * iters / dt / 1e9, 4), "unit": "Gvoxel-iter/s", "cores": threads,
           "kind": "port", "sample": f"oracle FP32 GVF Euler, {c}^3 centre block, {iters} iterations, {dt:.1f}s",
           "cpu_model": cpu_model()}
    out["stages"] = cpu_stages(oracle, vol)
    # one thread, on the 128^3 block
    c1 = min(128, vol.shape[0])
    o1 = (vol.shape[0] - c1) // 2
    F1 = np.ascontiguousarray(F[:, o1 - o:o1 - o + c1, o1 - o:o1 - o + c1, o1 - o:o1 - o + c1]) if c1 < c else F
    oracle.set_threads(1)
    try:
        oracle.gvf_euler(F1, 0.05, 2, "f32")
        t0 = time.time()
        n1 = 8
        oracle.gvf_euler(F1, 0.05, n1, "f32")
        n1 = max(8, int(3.0 / max((time.time() - t0) / n1, 1e-4)))
        t0 = time.time()
        oracle.gvf_euler(F1, 0.05, n1, "f32")
        d1 = time.time() - t0
    finally:
        oracle.set_threads(threads)
    out["one_thread"] = {"value": round(c1 ** 3 * n1 / d1 / 1e9, 4), "unit": "Gvoxel-iter/s",
                         "sample": f"oracle FP32 GVF Euler, {c1}^3 centre block, {n1} iterations, {d1:.1f}s"}
    return out


if __name__ == "__main__":
    main()
