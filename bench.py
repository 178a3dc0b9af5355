"""Benchmark: path vertices per second of the plain path tracer at 1080p.

Workload: procedural scene of 1,000,000 triangles (centres uniform in
[-10, 10]^3, LogNormal(-2, 0.5) edge offsets; 60% diffuse, 30% GGX metal,
10% dielectric), 1920x1080, 1 sample per pixel, 5 bounces, Russian roulette
from bounce 3, Medium quality, guiding off.  One step = one lmb_trace of one
frame (seed_offset = step index) with the scene resident in HBM.

value = path vertices (closest-hit queries) of the K timed frames / wall time;
the timed frames record no per-kernel events.
roofline: the traversal ("extend") kernel, algorithmic bytes per path vertex
184 + 80 * L = 664 (48 B ray in, 48 B ray / hit out, one 36 B triangle, 4 B
material index, 32 B material, 16 B accumulate, one 80 B node per BVH level,
L = ceil(log8(N/4)) = 6) over its launch durations, measured with HIP events
on the tracer's stream in a second pass over the same K frames (same seeds,
so the same paths and vertex count; events_pass_ms_per_step is that pass's
frame time); the traversal alone moves 84 + 80 * L = 564 of them
(traversal_only_frac).
cpu_baseline: the CPU oracle (C++/OpenMP restatement of the reference
shaders, "port") on a bounded band of rows of the same frame.

spectral_restir: the spectral + ReSTIR GI frame loop (guided mode, four
wavelengths, Cauchy dispersion, temporal + spatial reuse) over a 60-frame
camera orbit, in ms per frame.
batched_uneven: 64 uneven scenes in one batch of persistent launches
(hipGraph, autotuned launch shape) against one lmb_trace per image.
cornell_cfg1: the Cornell box at 256x256, 16 spp, depth 8, GPU and CPU.
guided_medium: the reference's own path (Markov-chain guiding, world grid,
per-pixel reservoir, no RR) on the 1080p scene.
time_to_relmse: the neural-cache tracer against the plain tracer (both 8
bounces) on the same scene: GPU time for each to reach the relMSE that the
plain tracer has at 16, 64 and 256 spp, judged against a 4096-spp plain
reference accumulated in fp64; plus the matrix-core rate of the cache's
inference and training kernels against the bf16 dense peak (with the spread
rule, and as configured: every path ends at its second vertex; with the first
scatter stratified, and from the reference's hash draws), and the CPU
oracle's time to the 256-spp relMSE (a band of rows, scaled to the frame).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--triangles N]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(spec, budget_s: float = 12.0) -> dict:
    import oracle
    cam = spec.camera().to_bytes()
    t0 = time.perf_counter()
    osc = oracle.Scene(spec)
    log(f"cpu baseline: oracle BVH built in {time.perf_counter() - t0:.1f} s")
    kw = dict(quality=1, samples=1, bounces=5, rr_start=3, seed_offset=0)
    mid = spec.height // 2
    t0 = time.perf_counter()
    _, _, _, v = osc.trace_plain(cam, spec.width, spec.height, row0=mid, rows=4, **kw)
    dt = time.perf_counter() - t0
    rows = max(4, min(spec.height - mid, int(4 * budget_s / max(dt, 1e-3))))
    t0 = time.perf_counter()
    _, _, _, v = osc.trace_plain(cam, spec.width, spec.height, row0=mid - rows // 2, rows=rows, **kw)
    dt = time.perf_counter() - t0
    return {"value": v / dt, "unit": "path-vertices/s", "cores": oracle.num_threads(),
            "kind": "port",
            "sample": f"{rows} rows x {spec.width} px around the image centre, same scene/camera/"
                      f"seeds, 1 spp, 5 bounces, RR from 3 ({v} vertices in {dt:.1f} s); CPU "
                      f"restatement, not lavapipe"}


MFMA_BF16_PEAK_TFLOPS = 2500.0
# Algorithmic FLOPs of the cache network per query (312 encoded inputs):
# 2 * (312*64 + 4*64*64 + 64*3); a training sample costs about 3x that.
NRC_FLOP_FWD = 2 * (312 * 64 + 4 * 64 * 64 + 64 * 3)
NRC_FLOP_TRAIN = 3 * NRC_FLOP_FWD


def rel_mse(img, ref) -> float:
    import numpy as np
    return float(np.mean((img - ref) ** 2 / (ref ** 2 + 0.01)))


class _Hip:
    """The HIP runtime that the library itself is linked against (the same
    loaded libamdhip64), through ctypes: events on the null stream, which the
    library's calls use when given no stream, device allocations and copies.
    (torch brings a HIP runtime of its own, which cannot start in a process
    where the library's already has.)"""

    def __init__(self):
        import ctypes
        self.ct = ctypes
        self.lib = ctypes.CDLL("libamdhip64.so")
        self.lib.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.lib.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.lib.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p,
                                                 ctypes.c_void_p]
        self.lib.hipEventDestroy.argtypes = [ctypes.c_void_p]
        self.lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.lib.hipFree.argtypes = [ctypes.c_void_p]
        self.lib.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                            ctypes.c_int, ctypes.c_void_p]
        self.lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                       ctypes.c_int]

    def _ok(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed with HIP error {rc}")

    def event(self):
        e = self.ct.c_void_p()
        self._ok(self.lib.hipEventCreate(self.ct.byref(e)), "hipEventCreate")
        return e

    def record(self, e):
        self._ok(self.lib.hipEventRecord(e, None), "hipEventRecord")

    def elapsed_s(self, a, b) -> float:
        ms = self.ct.c_float()
        self._ok(self.lib.hipEventElapsedTime(self.ct.byref(ms), a, b), "hipEventElapsedTime")
        return ms.value * 1e-3

    def destroy(self, e):
        self.lib.hipEventDestroy(e)

    def sync(self):
        self._ok(self.lib.hipDeviceSynchronize(), "hipDeviceSynchronize")

    def malloc(self, n: int):
        p = self.ct.c_void_p()
        self._ok(self.lib.hipMalloc(self.ct.byref(p), n), "hipMalloc")
        return p

    def free(self, p):
        self.lib.hipFree(p)

    def copy_d2d_async(self, dst, src: int, n: int):
        self._ok(self.lib.hipMemcpyAsync(dst, self.ct.c_void_p(src), n, 3, None), "hipMemcpyAsync")

    def read(self, src, out):
        self._ok(self.lib.hipMemcpy(out.ctypes.data_as(self.ct.c_void_p), src, out.nbytes, 2),
                 "hipMemcpy")


def cache_reference(api, scene, spec, cam, ref_spp: int):
    """The plain tracer's 4096-spp fp64 reference and its 16 / 64 / 256-spp
    points (8 bounces, RR from 3) for the time-to-relMSE legs."""
    import numpy as np
    W, H = spec.width, spec.height
    popts = dict(bounces=8, rr_start=3, guiding=api.Guiding.OFF,
                 restir_temporal=api.RestirTemporalMode.Disabled)

    hip = _Hip()
    ev0, ev1 = hip.event(), hip.event()

    def plain(spp, seed, timed=False):
        pipe = api.RayTracer(api.Medium).create_pipeline(api.PathTracerOpts(samples=spp, **popts))
        b = api.DataBuffers(W, H, 1, api.BufferType(0))
        if timed:  # buffers allocated and the kernels loaded outside the timed call
            pipe.dispatch(scene, cam, b, seed + 1_000_000)
        hip.sync()
        hip.record(ev0)
        pipe.dispatch(scene, cam, b, seed)
        hip.record(ev1)
        hip.sync()
        dt = hip.elapsed_s(ev0, ev1)
        return b.output(api.Output.COLOR)[..., :3].astype(np.float64), dt

    t0 = time.perf_counter()
    chunk = 64
    ref = np.zeros((H, W, 3))
    for i in range(ref_spp // chunk):
        ref += plain(chunk, 10_000_000 + i)[0]
    ref /= ref_spp // chunk
    log(f"cache leg: reference {ref_spp} spp in {time.perf_counter() - t0:.1f} s")
    plain_pts = []
    for n_spp in (16, 64, 256):
        img, dt = plain(n_spp, 20_000 + n_spp, timed=True)
        plain_pts.append({"spp": n_spp, "seconds": dt, "relmse": rel_mse(img, ref)})
        log(f"cache leg: plain {n_spp} spp {dt * 1e3:.1f} ms relMSE {plain_pts[-1]['relmse']:.5f}")
    return ref, plain_pts, f"plain tracer {ref_spp} spp, fp64 accumulation of {chunk}-spp dispatches"


def cache_leg(api, scene, spec, cam, reference, frames: int, warm: int,
              lr: float = 1e-3, spp: int = 8, steps: int = 16, overlap: bool = True,
              warm_paths: int = 524288, warm_steps: int = 256, ema: float = 0.0,
              defer_specular: int = 2, stratify: bool = True) -> dict:
    """Time-to-relMSE of the cached tracer against the plain tracer, both with
    8-bounce paths (RR from bounce 3), judged against the plain reference of
    cache_reference (`reference`).  The cache starts untrained: `warm`
    training frames (lmb_trace_nrc_train: about warm_paths paths at random
    pixels train it with warm_steps Adam steps, no image; timed), then
    `frames` frames of `spp` samples (the camera rays are traced once per
    frame), `steps` Adam steps per frame over the frame's 65536 training
    records (overlap: a training pass of its own and the steps run on the
    cache's own stream beside the frame's samples, which are traced eight at
    a time; otherwise the training paths come from the first sample).  The
    frames run back to back on one stream as in a progressive render, each
    image pushed into a frame ring on the GPU; times are stream times
    (events) to the end of each checkpoint frame, and the checkpoint averages
    are scored after the run.  The plain tracer's points are timed the same way
    (one dispatch each).  defer_specular: how many specular scatters may
    move a path's cache vertex one vertex on (0: always the second vertex).
    stratify: the first scatter draws from each pixel's scrambled Sobol
    sequence (False: the reference's hash draws, as the plain tracer's)."""
    import numpy as np
    W, H = spec.width, spec.height
    popts = dict(bounces=8, rr_start=3, guiding=api.Guiding.OFF,
                 restir_temporal=api.RestirTemporalMode.Disabled)
    ref, plain_pts, ref_text = reference
    frame_opts = api.NeuralCacheOpts(learning_rate=lr, train_steps=steps, overlap_training=overlap,
                                     ema_decay=ema, defer_specular=defer_specular,
                                     stratify=stratify)
    cache = api.NeuralCache(api.NeuralCacheOpts(**{**frame_opts.__dict__, "train_paths": warm_paths,
                                                   "train_steps": warm_steps}))
    warm_pipe = api.RayTracer(api.Medium).create_pipeline(api.PathTracerOpts(samples=1, **popts))
    pipe = api.RayTracer(api.Medium).create_pipeline(api.PathTracerOpts(samples=spp, **popts))
    b = api.DataBuffers(W, H, 1, api.BufferType(0))
    # A progressive render: the warm-up and the frames back to back on one
    # stream, each frame's image pushed into a frame ring (the example
    # application's accumulation, one slot per frame); at a checkpoint the
    # ring's average is copied into a snapshot on the device.  Times are
    # stream times (events) to the end of each checkpoint frame, less the
    # averages and copies of the checkpoints before it; the snapshots are
    # read and scored after the run.
    hip = _Hip()
    ring = api.FrameRing(W, H, frames)
    rgb_ptr, rgb_bytes = ring.output_device_ptr()
    checks = sorted(set(range(1, 65, 2)) | set(range(64, frames + 1, 8)) | {frames})
    snap_mem = {n: hip.malloc(rgb_bytes) for n in checks}
    ev_start, ev_warm, ev_frames = hip.event(), hip.event(), hip.event()
    hip.sync()
    hip.record(ev_start)
    for f in range(warm):
        warm_pipe.train_cached(scene, cam, b, cache, f)
    hip.record(ev_warm)
    hip.sync()  # the options change below is a host call
    cache.set_opts(frame_opts)
    hip.record(ev_frames)
    snaps = []
    for n in range(1, frames + 1):
        pipe.dispatch_cached(scene, cam, b, cache, warm + n)
        ring.push(b)
        if n in snap_mem:
            ev_a, ev_b = hip.event(), hip.event()
            hip.record(ev_a)
            ring.average_on_device()
            hip.copy_d2d_async(snap_mem[n], rgb_ptr, rgb_bytes)
            hip.record(ev_b)
            snaps.append((n, ev_a, ev_b))
    hip.sync()
    t_warm = hip.elapsed_s(ev_start, ev_warm)
    pts = []  # the host gap at the options change is not counted
    excluded = 0.0
    img = np.empty((H, W, 3), np.float32)
    for n, ev_a, ev_b in snaps:
        t_frames = hip.elapsed_s(ev_frames, ev_a) - excluded
        excluded += hip.elapsed_s(ev_a, ev_b)
        hip.read(snap_mem[n], img)
        pts.append({"frames": n, "samples": n * spp, "seconds": t_warm + t_frames,
                    "relmse": rel_mse(img.astype(np.float64), ref)})
        hip.destroy(ev_a)
        hip.destroy(ev_b)
    for p in snap_mem.values():
        hip.free(p)
    for e in (ev_start, ev_warm, ev_frames):
        hip.destroy(e)
    ring.close()
    t = pts[-1]["seconds"]
    log(f"cache leg (spread bound {defer_specular}, stratify {stratify}): {warm} + {frames} frames in {t:.2f} s, "
        f"final relMSE {pts[-1]['relmse']:.5f}")
    # Matrix-core figures from a few more frames with per-launch events (kept
    # out of the timed run: the events add about a quarter to a frame): the
    # training rate from frames that train, the inference rate from frames
    # with the training paused, so that no training kernel shares the GPU
    # with the inference launch.
    b.set_profiling(True)
    b.kernel_stats()
    prof_frames = 4
    for f in range(prof_frames):
        pipe.dispatch_cached(scene, cam, b, cache, warm + frames + 1 + f)
    ks = b.kernel_stats()
    st = cache.stats()
    cache.set_opts(api.NeuralCacheOpts(**{**frame_opts.__dict__, "train_paths": 0}))
    for f in range(prof_frames):
        pipe.dispatch_cached(scene, cam, b, cache, warm + frames + 1 + prof_frames + f)
    ks_inf = b.kernel_stats()
    st_inf = cache.stats()
    b.set_profiling(False)
    targets = []
    for p in plain_pts:
        hit = next((q for q in pts if q["relmse"] <= p["relmse"]), None)
        targets.append({"plain_spp": p["spp"], "relmse": p["relmse"], "plain_seconds": p["seconds"],
                        "cache_seconds": hit["seconds"] if hit else None,
                        "speedup": p["seconds"] / hit["seconds"] if hit else None})
    infer_launches = ks_inf["cache_infer_launches"]
    infer_ms = ks_inf["cache_infer_ms"]
    train_ms = ks["cache_train_ms"]
    infer_flops = st_inf["queries"] * NRC_FLOP_FWD * infer_launches
    train_flops = st["records"] * NRC_FLOP_TRAIN * prof_frames
    infer_tf = infer_flops / (infer_ms * 1e-3) / 1e12 if infer_ms else 0.0
    train_tf = train_flops / (train_ms * 1e-3) / 1e12 if train_ms else 0.0
    return {
        "workload": f"procedural {W}x{H}, 8 bounces, RR from 3; cache: paths end at the second "
                    + ("vertex (a specular scatter moves it one vertex on, at most "
                       f"{defer_specular} times)" if defer_specular else "vertex, always") +
                    f", 65536 training paths "
                    f"and {steps} Adam steps (lr {lr:g}) per frame"
                    f"{' on a second stream beside the samples' if overlap else ''}; {warm} "
                    f"training frames of {warm_paths} paths and {warm_steps} steps (no image, "
                    f"timed), then frames of {spp} spp; first scatter from "
                    + ("each pixel's Owen-scrambled Sobol sequence" if stratify
                       else "the reference's hash draws"),
        "reference": ref_text,
        "plain": plain_pts,
        "cache_points": pts,
        "warm_ms_per_frame": t_warm / max(warm, 1) * 1e3,
        "cache_ms_per_frame": (t - t_warm) / frames * 1e3,
        "targets": targets,
        "headline": targets[-1],
        "mfma": {
            "infer_tflops": infer_tf, "infer_frac": infer_tf / MFMA_BF16_PEAK_TFLOPS,
            "infer_ms_per_launch": infer_ms / max(infer_launches, 1),
            "queries_per_launch": st_inf["queries"],
            "infer_measured_on": "frames with the training paused (no training kernel beside it)",
            "train_tflops": train_tf, "train_frac": train_tf / MFMA_BF16_PEAK_TFLOPS,
            "train_ms_per_frame": train_ms / prof_frames,
            "records_per_frame": st["records"],
            "peak_tflops": MFMA_BF16_PEAK_TFLOPS, "flop_per_query": NRC_FLOP_FWD,
        },
    }


def spectral_leg(api, scenes, triangles: int, width: int, height: int, frames: int) -> dict:
    """The spectral + ReSTIR GI frame loop: the procedural scene with 20% of
    its dielectrics dispersive (Cauchy A = 1.5, B = 0.004 um^2), four
    wavelengths per path, Markov-chain guiding with a 1.5M-entry world grid, a
    camera orbit of radius 25 (6 degrees per frame), 1 spp per frame, per-pixel
    reservoirs, temporal reuse, 5 spatial neighbours in a 61-px window and a
    confidence cap of 20.  One frame = trace + temporal + spatial +
    advance_frame; the timed region holds `frames` frames back to back."""
    spec = scenes.procedural(triangles, seed=2, width=width, height=height, dispersive=0.2)
    scene = spec.build()
    opts = api.PathTracerOpts(samples=1, bounces=5, rr_start=3, guiding=api.Guiding.MARKOV_CHAIN,
                              restir_temporal=api.RestirTemporalMode.PerPixel)
    pipe = api.RayTracer(api.Spectral).create_pipeline(opts)
    bufs = api.DataBuffers(width, height, 1_500_000, api.BufferType.all())
    proj = api.perspective(45.0, width / height, 0.1, 200.0)
    cams = []
    for f in range(frames + 2):
        a = math.radians(6.0 * f)
        eye = (-25.0 * math.sin(a), 0.0, -25.0 * math.cos(a))
        d = (-eye[0] / 25.0, 0.0, -eye[2] / 25.0)
        cams.append(api.Camera.from_proj_view(proj, api.look_to_rh(eye, d, (0.0, 1.0, 0.0))))
    cap = api.AdvanceOptions(20)

    def frame(f):
        pipe.dispatch(scene, cams[f], bufs, f)
        bufs.temporal_resample(f)
        bufs.spatial_resample(5, 61, f)
        bufs.advance_frame(api.BufferType.all(), cap)

    for f in range(2):  # warm-up: first-use allocations of the guided state
        frame(f)
    api.sync()
    bufs.set_profiling(True)
    bufs.kernel_stats()
    t0 = time.perf_counter()
    for f in range(2, frames + 2):
        frame(f)
    api.sync()
    dt = time.perf_counter() - t0
    ks = bufs.kernel_stats()
    img = bufs.output(api.Output.COLOR)
    log(f"spectral leg: {frames} frames in {dt:.3f} s, {ks['path_vertices']} path vertices")
    return {
        "workload": f"procedural {triangles} triangles, 20% of dielectrics dispersive (Cauchy "
                    f"A=1.5, B=0.004 um^2), {width}x{height}, 4 wavelengths/path (380-780 nm), "
                    f"1 spp, 5 bounces, RR from 3, Markov-chain guiding (world grid 1.5M), "
                    f"per-pixel reservoir, temporal + 5 spatial neighbours (61 px), M cap 20, "
                    f"{frames}-frame orbit of radius 25",
        "frames": frames,
        "seconds": dt,
        "ms_per_frame": dt / frames * 1e3,
        "path_vertices_per_s": ks["path_vertices"] / dt,
        "extend_ms_per_frame": ks["extend_ms"] / frames,
        "shade_ms_per_frame": ks["shade_ms"] / frames,
        "image_finite": bool(math.isfinite(float(img.sum()))),
    }


def cornell_leg(api, scenes, frames: int, cpu: bool) -> dict:
    """Config 1: the Cornell box built in code, 256x256, 16 spp, depth 8
    (no RR), Medium, plain; one frame = one dispatch of 16 spp.  The CPU row
    is the oracle (C++/OpenMP restatement, "port") on the same frame."""
    spec = scenes.cornell_box(256, 256)
    scene = spec.build()
    opts = api.PathTracerOpts(samples=16, bounces=8, guiding=api.Guiding.OFF,
                              restir_temporal=api.RestirTemporalMode.Disabled)
    pipe = api.RayTracer(api.Medium).create_pipeline(opts)
    bufs = api.DataBuffers(256, 256, 1, api.BufferType(0))
    pipe.dispatch(scene, spec.camera(), bufs, 0)
    api.sync()
    bufs.set_profiling(True)
    bufs.kernel_stats()
    t0 = time.perf_counter()
    for f in range(frames):
        pipe.dispatch(scene, spec.camera(), bufs, 1 + f)
    api.sync()
    dt = time.perf_counter() - t0
    verts = bufs.kernel_stats()["path_vertices"]
    per_sample = verts + frames * 15 * 256 * 256   # camera rays counted once per sample
    out = {
        "workload": "Cornell box (36 triangles), 256x256, 16 spp, depth 8, no RR, Medium, plain",
        "frames": frames, "ms_per_frame": dt / frames * 1e3,
        "path_vertices_per_s": verts / dt,
        "path_vertices_per_s_counting_camera_rays_per_sample": per_sample / dt,
        "note": "the GPU traces each pixel's camera ray once for all 16 samples (no jitter); "
                "the second rate counts it once per sample, as the CPU row does",
    }
    if cpu:
        import oracle
        osc = oracle.Scene(spec)
        t0 = time.perf_counter()
        _, _, _, v = osc.trace_plain(spec.camera().to_bytes(), 256, 256, quality=1, samples=16,
                                     bounces=8, seed_offset=1)
        cdt = time.perf_counter() - t0
        out["cpu"] = {"value": v / cdt, "unit": "path-vertices/s", "seconds_per_frame": cdt,
                      "cores": oracle.num_threads(), "kind": "port",
                      "sample": "the whole 256x256 16-spp frame; CPU restatement, not lavapipe"}
    return out


def cpu_time_to_relmse(spec, reference, rows: int = 8) -> dict:
    """The CPU oracle's time to reach the plain tracer's 256-spp relMSE on the
    1080p scene: it renders 256 spp (8 bounces, RR from 3) on a band of rows
    in the image centre; the full-frame time is that band's time scaled by
    height / rows (the band's relMSE against the same reference is reported
    beside the GPU's full-frame figure)."""
    import numpy as np
    import oracle
    ref = reference[0]
    osc = oracle.Scene(spec)
    row0 = spec.height // 2 - rows // 2
    t0 = time.perf_counter()
    c, _, _, v = osc.trace_plain(spec.camera().to_bytes(), spec.width, spec.height, row0=row0,
                                 rows=rows, quality=1, samples=256, bounces=8, rr_start=3,
                                 seed_offset=20_256)
    dt = time.perf_counter() - t0
    band = c[..., :3].astype(np.float64)
    return {"seconds_full_frame": dt * spec.height / rows, "band_seconds": dt,
            "band_relmse": rel_mse(band, ref[row0:row0 + rows]),
            "band_path_vertices_per_s": v / dt, "cores": oracle.num_threads(), "kind": "port",
            "sample": f"{rows} rows x {spec.width} px at 256 spp, 8 bounces, RR from 3, scaled to "
                      f"{spec.height} rows; CPU restatement, not lavapipe"}


def guided_leg(api, scene, spec, frames: int) -> dict:
    """The reference's own path (rt_main): Markov-chain guiding with the
    world grid (1.5M cells), per-pixel ReSTIR reservoir, Medium, one
    wavelength, 1 spp, 5 bounces, no Russian roulette, on the 1M-triangle
    1080p scene; one frame = trace + advance_frame (static camera)."""
    W, H = spec.width, spec.height
    opts = api.PathTracerOpts(samples=1, bounces=5, guiding=api.Guiding.MARKOV_CHAIN,
                              restir_temporal=api.RestirTemporalMode.PerPixel)
    pipe = api.RayTracer(api.Medium).create_pipeline(opts)
    bufs = api.DataBuffers(W, H, 1_500_000, api.BufferType.all())
    cam = spec.camera()
    cap = api.AdvanceOptions(20)
    for f in range(3):
        pipe.dispatch(scene, cam, bufs, f)
        bufs.advance_frame(api.BufferType.all(), cap)
    api.sync()
    bufs.set_profiling(True)
    bufs.kernel_stats()
    t0 = time.perf_counter()
    for f in range(frames):
        pipe.dispatch(scene, cam, bufs, 3 + f)
        bufs.advance_frame(api.BufferType.all(), cap)
    api.sync()
    dt = time.perf_counter() - t0
    ks = bufs.kernel_stats()
    verts = ks["path_vertices"]
    L = math.ceil(math.log(1_000_000 / 4, 8))
    bpv = 184 + 80 * L + 768
    log(f"guided leg: {frames} frames in {dt:.3f} s, {verts} path vertices")
    return {
        "workload": f"reference path (rt_main): procedural 1000000 triangles, {W}x{H}, Medium, "
                    "Markov-chain guiding (world grid 1.5M), per-pixel reservoir, 1 spp, 5 "
                    "bounces, no RR; frame = trace + advance_frame",
        "frames": frames, "ms_per_frame": dt / frames * 1e3,
        "path_vertices_per_s": verts / dt,
        "extend_ms_per_frame": ks["extend_ms"] / frames,
        "shade_and_guide_ms_per_frame": ks["shade_ms"] / frames,
        "roofline": {"bound": "hbm", "scope": "whole frame", "bytes_per_vertex": bpv,
                     "achieved": verts * bpv / dt / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": verts * bpv / dt / 1e9 / HBM_PEAK_GBS},
    }


def batch_leg(api, scenes, frames: int, eager_only: bool = False) -> dict:
    """64 uneven procedural scenes (1k-500k triangles log-uniform, 128^2 to
    1024^2 images) rendered as one batch: 4 spp, 6 bounces (no RR), Medium,
    plain.  Compared with rendering the same 64 images one lmb_trace
    at a time; the batch's launch shape comes from an autotuner sweep."""
    t0 = time.perf_counter()
    specs = scenes.batch_workload()
    items = [(sp.build(), sp.camera(), sp.width, sp.height) for sp in specs]
    tris = sum(len(sp.vertices) // 3 for sp in specs)
    pixels = sum(sp.width * sp.height for sp in specs)
    log(f"batch leg: {len(items)} scenes, {tris} triangles, {pixels} pixels built in "
        f"{time.perf_counter() - t0:.1f} s")
    opts = api.PathTracerOpts(samples=4, bounces=6, guiding=api.Guiding.OFF,
                              restir_temporal=api.RestirTemporalMode.Disabled)
    # one image at a time
    pipe = api.RayTracer(api.Medium).create_pipeline(opts)
    bufs = [api.DataBuffers(w, h, 1, api.BufferType(0)) for _, _, w, h in items]
    for (scene, cam, _, _), b in zip(items, bufs):
        pipe.dispatch(scene, cam, b, 0)
    api.sync()
    t0 = time.perf_counter()
    for f in range(frames):
        for (scene, cam, _, _), b in zip(items, bufs):
            pipe.dispatch(scene, cam, b, 1 + f)
    api.sync()
    seq_ms = (time.perf_counter() - t0) / frames * 1e3
    seq_verts = sum(b.path_vertices() for b in bufs)
    del bufs

    batch = api.Batch(items, opts, api.Medium)

    def timed(cfg):
        batch.set_config(cfg)
        batch.trace(0)
        api.sync()
        t0 = time.perf_counter()
        for f in range(frames):
            batch.trace(1 + f)
        api.sync()
        return (time.perf_counter() - t0) / frames * 1e3

    eager_ms = timed(api.BatchConfig(use_graph=False))
    if eager_only:   # the default launch shape without graphs or the sweep (kernel profiles)
        log(f"batch leg (eager only): one-by-one {seq_ms:.1f} ms, batched {eager_ms:.1f} ms")
        return {"one_by_one_ms_per_frame": seq_ms, "batched_eager_ms_per_frame": eager_ms}
    graph_ms = timed(api.BatchConfig())
    sweep = [api.BatchConfig(tile_w=tw, tile_h=th, blocks_per_cu=k, refill=16, pass_pixels=pp)
             for tw, th in ((8, 8), (64, 1)) for k in (4, 8, 16) for pp in (0, 1 << 22)]
    t0 = time.perf_counter()
    best, ms = batch.autotune(sweep, frames=1)
    tune_s = time.perf_counter() - t0
    best_ms = timed(best)
    verts = batch.path_vertices()
    log(f"batch leg: one-by-one {seq_ms:.1f} ms, batched {eager_ms:.1f} ms eager / "
        f"{graph_ms:.1f} ms graph, autotuned {best_ms:.1f} ms ({tune_s:.1f} s sweep)")
    return {
        "workload": f"{len(items)} procedural scenes, {tris} triangles (1k-500k log-uniform), "
                    f"{pixels} pixels (128^2-1024^2), 4 spp, 6 bounces, no RR, Medium, "
                    f"2/3 diffuse 1/3 GGX",
        "one_by_one_ms_per_frame": seq_ms,
        "batched_eager_ms_per_frame": eager_ms,
        "batched_graph_ms_per_frame": graph_ms,
        "autotuned": {"config": best.__dict__, "ms_per_frame": best_ms, "sweep_seconds": tune_s,
                      "sweep": [{"config": c.__dict__, "ms": m} for c, m in zip(sweep, ms)]},
        "path_vertices_per_frame": verts,
        "path_vertices_per_s": verts / (best_ms * 1e-3),
        "speedup_vs_one_by_one": seq_ms / best_ms,
        "vertices_match_one_by_one": verts == seq_verts,
    }


def calibrate_copy(gib: int = 1, launches: int = 3) -> int:
    """A plain device-to-device copy of known size (torch's copy kernel) for
    scaling rocprofv3 FETCH_SIZE / WRITE_SIZE readings to bytes."""
    import torch
    n = gib << 30
    src = torch.ones(n, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(launches):
        dst.copy_(src)
    torch.cuda.synchronize()
    log(f"copied {gib} GiB {launches} times")
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--triangles", type=int, default=1_000_000)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-cache-leg", action="store_true")
    ap.add_argument("--no-spectral-leg", action="store_true")
    ap.add_argument("--spectral-frames", type=int, default=60)
    ap.add_argument("--no-batch-leg", action="store_true")
    ap.add_argument("--no-cornell-leg", action="store_true")
    ap.add_argument("--cornell-frames", type=int, default=10)
    ap.add_argument("--no-guided-leg", action="store_true")
    ap.add_argument("--guided-frames", type=int, default=10)
    ap.add_argument("--batch-frames", type=int, default=3)
    ap.add_argument("--batch-eager-only", action="store_true",
                    help="batch leg with the default launch shape, no graph, no autotune")
    ap.add_argument("--ref-spp", type=int, default=4096)
    ap.add_argument("--cache-frames", type=int, default=96)
    ap.add_argument("--cache-warmup", type=int, default=4)
    ap.add_argument("--cache-lr", type=float, default=1e-3)
    ap.add_argument("--cache-spp", type=int, default=8)
    ap.add_argument("--cache-steps", type=int, default=16)
    ap.add_argument("--cache-ema", type=float, default=0.0,
                    help="inference from an exponential moving average of the weights")
    ap.add_argument("--cache-warm-paths", type=int, default=524288)
    ap.add_argument("--cache-warm-steps", type=int, default=256)
    ap.add_argument("--cache-no-overlap", action="store_true",
                    help="cache training on the frame's stream after its samples")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 1 GiB device copy (rocprofv3 counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        return calibrate_copy()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(image rows would shard across GPUs, one band of rows per rank). "
              "Run with --gpus 1.",
              file=sys.stderr)
        return 2

    from lambent_amd import api, scenes

    t0 = time.perf_counter()
    spec = scenes.procedural(args.triangles, seed=2, width=args.width, height=args.height)
    log(f"scene generated in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    scene = spec.build()
    st = scene.bvh_stats()
    log(f"scene uploaded in {time.perf_counter() - t0:.1f} s (BVH {st.num_nodes} nodes, depth "
        f"{st.max_depth}, build {st.build_ms:.0f} ms, {st.device_bytes / 1e6:.0f} MB)")
    cam = spec.camera()
    opts = api.PathTracerOpts(samples=1, bounces=5, rr_start=3, guiding=api.Guiding.OFF,
                              restir_temporal=api.RestirTemporalMode.Disabled)
    pipe = api.RayTracer(api.Medium).create_pipeline(opts)
    bufs = api.DataBuffers(spec.width, spec.height, 1, api.BufferType(0))

    for w in range(args.warmup):
        pipe.dispatch(scene, cam, bufs, 1_000_000 + w)
    api.sync()
    log(f"warmup done ({args.warmup} frames)")

    # The timed frames, with no per-kernel events: a timed HIP event is a
    # marker on the stream (about 2 us per record, 20 per frame).
    t0 = time.perf_counter()
    for k in range(args.steps):
        pipe.dispatch(scene, cam, bufs, k)
    api.sync()
    elapsed = time.perf_counter() - t0
    # The same frames (same seeds, so the same paths and vertex count) once
    # more with HIP events around every traversal and shading launch on the
    # tracer's stream: the kernels' durations for the roofline.
    bufs.set_profiling(True)
    bufs.kernel_stats()  # reset
    api.sync()
    t1 = time.perf_counter()
    for k in range(args.steps):
        pipe.dispatch(scene, cam, bufs, k)
    api.sync()
    elapsed_events = time.perf_counter() - t1
    ks = bufs.kernel_stats()
    bufs.set_profiling(False)
    verts = ks["path_vertices"]
    log(f"timed {args.steps} frames in {elapsed:.3f} s ({elapsed_events:.3f} s with per-kernel "
        f"events), {verts} path vertices")

    value = verts / elapsed
    L = math.ceil(math.log(max(args.triangles, 4) / 4, 8))
    # algorithmic bytes per path vertex: 48 ray in + 48 ray / hit out + 36
    # triangle + 4 material index + 32 material + 16 accumulate + one 80-B
    # node per BVH level; of these the traversal itself moves 84 + 80 L
    bytes_per_vertex = 184 + 80 * L
    traversal_bytes = 84 + 80 * L
    achieved = verts * bytes_per_vertex / (ks["extend_ms"] * 1e-3) / 1e9 if ks["extend_ms"] else 0.0
    out = {
        "metric": "path-vertices/s @1080p",
        "value": value,
        "unit": "path-vertices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"plain path tracer, procedural {args.triangles} triangles, "
                        f"{spec.width}x{spec.height}, 1 spp, 5 bounces, RR from bounce 3, "
                        f"Medium quality, guiding off",
            "triangles": args.triangles, "width": spec.width, "height": spec.height, "spp": 1,
            "bounces": 5, "rr_start": 3, "parallelism": "single GPU",
            "bvh_nodes": st.num_nodes, "bvh_depth": st.max_depth,
            "vertices_per_frame": verts / args.steps,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_extend (BVH8 traversal)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "bytes_per_vertex": bytes_per_vertex,
            "traversal_only_bytes_per_vertex": traversal_bytes,
            "traversal_only_frac": achieved * traversal_bytes / bytes_per_vertex / HBM_PEAK_GBS,
            "extend_ms_per_step": ks["extend_ms"] / args.steps,
            "shade_ms_per_step": ks["shade_ms"] / args.steps,
            "extend_launches": ks["extend_launches"],
            "events_pass_ms_per_step": elapsed_events / args.steps * 1e3,
        },
        "cpu_baseline": None,
    }
    # HBM-side bytes per k_extend launch from the last calibrated rocprofv3
    # --pmc FETCH_SIZE / WRITE_SIZE runs of this workload (profiles/).
    tfile = ROOT / "profiles" / "k_extend_traffic.json"
    if tfile.exists() and args.triangles == 1_000_000 and (args.width, args.height) == (1920, 1080):
        t = json.loads(tfile.read_text())
        out["roofline"]["traffic"] = t["traffic_bytes_per_launch"]
        out["roofline"]["traffic_source"] = t["source"]
        out["roofline"]["algorithmic_bytes_per_launch"] = (verts / ks["extend_launches"]
                                                           * bytes_per_vertex)
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(spec)
        except Exception as exc:  # the GPU figure stands on its own
            log(f"cpu baseline failed: {exc!r}")
    if not args.no_cornell_leg:
        out["cornell_cfg1"] = cornell_leg(api, scenes, args.cornell_frames,
                                          not args.no_cpu_baseline)
    if not args.no_guided_leg:
        out["guided_medium"] = guided_leg(api, scene, spec, args.guided_frames)
    if not args.no_cache_leg:
        reference = cache_reference(api, scene, spec, cam, args.ref_spp)
        leg = dict(frames=args.cache_frames, warm=args.cache_warmup, lr=args.cache_lr,
                   spp=args.cache_spp, steps=args.cache_steps, overlap=not args.cache_no_overlap,
                   warm_paths=args.cache_warm_paths, warm_steps=args.cache_warm_steps,
                   ema=args.cache_ema)
        out["time_to_relmse"] = cache_leg(api, scene, spec, cam, reference, **leg)
        # the config as written: every path ends at its second vertex
        out["time_to_relmse_vertex2"] = cache_leg(api, scene, spec, cam, reference,
                                                  defer_specular=0, **leg)
        # the cache alone: the first scatter from the reference's hash draws
        out["time_to_relmse_hash_draws"] = cache_leg(api, scene, spec, cam, reference,
                                                     stratify=False, **leg)
        if not args.no_cpu_baseline:
            try:
                out["time_to_relmse"]["cpu_plain_256spp"] = cpu_time_to_relmse(spec, reference)
            except Exception as exc:
                log(f"cpu time-to-relMSE failed: {exc!r}")
    if not args.no_spectral_leg:
        del scene
        out["spectral_restir"] = spectral_leg(api, scenes, args.triangles, args.width,
                                              args.height, args.spectral_frames)
    if not args.no_batch_leg:
        out["batched_uneven"] = batch_leg(api, scenes, args.batch_frames, args.batch_eager_only)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
