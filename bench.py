"""Benchmark: path vertices per second of the plain path tracer at 1080p.

Workload: procedural scene of 1,000,000 triangles (centres uniform in
[-10, 10]^3, LogNormal(-2, 0.5) edge offsets; 60% diffuse, 30% GGX metal,
10% dielectric), 1920x1080, 1 sample per pixel, 5 bounces, Russian roulette
from bounce 3, Medium quality, guiding off.  One step = one lmb_trace of one
frame (seed_offset = step index) with the scene resident in HBM.

value = path vertices (closest-hit queries) of the K timed frames / wall time.
roofline: the traversal ("extend") kernel, algorithmic bytes per vertex
84 + 80 * L (ray 32 B in, hit 16 B out, one 36 B triangle, one 80 B node per
BVH level, L = ceil(log8(N/4)) = 6) over its HIP-event launch durations.
cpu_baseline: the CPU oracle (C++/OpenMP restatement of the reference
shaders, "port") on a bounded band of rows of the same frame.

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


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--triangles", type=int, default=1_000_000)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(image rows would shard across GPUs; see DESIGN.md). Run with --gpus 1.",
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

    bufs.set_profiling(True)
    bufs.kernel_stats()  # reset
    api.sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        pipe.dispatch(scene, cam, bufs, k)
    api.sync()
    elapsed = time.perf_counter() - t0
    ks = bufs.kernel_stats()
    verts = ks["path_vertices"]
    log(f"timed {args.steps} frames in {elapsed:.3f} s, {verts} path vertices")

    value = verts / elapsed
    L = math.ceil(math.log(max(args.triangles, 4) / 4, 8))
    bytes_per_vertex = 84 + 80 * L
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
            "extend_ms_per_step": ks["extend_ms"] / args.steps,
            "shade_ms_per_step": ks["shade_ms"] / args.steps,
            "extend_launches": ks["extend_launches"],
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
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
