#!/usr/bin/env python3
"""Throughput of the TressFX 4.1 strand simulation + SDF collision path on one
MI355X, in vertex-substeps/s.

Workload (the one-GPU configuration the metric is quoted on): 20,000 guide
strands x 32 vertices with 4 follow strands per guide (3.2 M vertices), a
128^3 SDF (2.5 mm cells) rebuilt every frame from a 20,480-triangle noisy
icosphere, wind with a 0.3 Hz gust, 2 substeps per frame, length iterations 2.
A step is one frame: SDF rebuild + 2 x (fused substep incl. SDF collision).
Inputs (asset, per-frame collision meshes) are resident in HBM before the timed
region; the per-substep parameter block and bone matrix are uploaded by the
library, as the reference uploads its constant buffer.

Printed line: one JSON object with value = vertex-substeps/s, the roofline of
the dominant kernel (the fused substep, HBM-bound), the plain per-instance
dispatch for comparison and the CPU transcription's throughput.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes_per_substep(G, F, V):
    """Minimal streams of one substep: per guide vertex read Pos, Prev, RestPos
    (16 B each) and RestLength (4 B), write Pos, Prev, Tangent (16 B each) =
    100 B; per follow vertex write Pos, Tangent = 32 B; per guide strand bone
    skinning record 32 B + vertex-1 PrevPrev read/write 32 B = 64 B; per follow
    strand its root offset 16 B."""
    return G * V * 100 + G * F * V * 32 + G * 64 + G * F * 16


def read_traffic(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def cpu_baseline(wl, frames, nthreads, budget_s):
    """The FP32 transcription (oracle) on the host, multithreaded over strands
    (and triangles for the SDF build), same workload and frame order."""
    import numpy as np
    import oracle
    import tressfx_amd as tfx
    from tressfx_amd import synth
    a = wl.asset
    h = oracle.OracleHair(a, np.float32)
    p = tfx.SimulationParams()
    bones = np.tile(np.eye(4, dtype=np.float32).reshape(1, 16), (tfx.TFX_MAX_NUM_BONES, 1))
    origin, cell, dims = wl.sdf_grid
    origin = np.asarray(origin, np.float32)
    margin = np.float32(np.float32(wl.collision_margin) * np.float32(cell))
    dt = 1.0 / synth.FRAME_RATE
    t0 = time.perf_counter()
    done_substeps = 0
    frame = 0
    for frame in range(frames):
        t = frame * dt
        verts = wl.mesh.vertices(t, wl.bone_fn(t))
        grid = oracle.sdf_build(verts, wl.mesh.indices, origin, np.float32(cell), dims, np.float32, nthreads)
        for s in range(wl.substeps):
            ts = (frame + (s + 1) / wl.substeps) * dt
            bones[0] = wl.bone_fn(ts).reshape(16)
            wl.settings.wind_magnitude = wl.wind_fn(ts)
            oracle.pack_sim_params(wl.settings, np.float32(dt / wl.substeps), frame * wl.substeps + s,
                                   a.num_vertices_per_strand, a.num_follow_strands_per_guide, into=p)
            h.simulate(p, bones, nthreads)
            p.reset_positions = 0.0
            oracle.sdf_collide(h.pos, h.prev, a.num_vertices_per_strand, grid, origin, np.float32(cell), dims,
                               margin, nthreads)
            done_substeps += 1
        if time.perf_counter() - t0 > budget_s:
            break
    el = time.perf_counter() - t0
    return a.num_total_vertices * done_substeps / el, done_substeps, frame + 1, el


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--plain-steps", type=int, default=20)
    ap.add_argument("--cpu-frames", type=int, default=600)
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--guides", type=int, default=20000)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this path runs on one GPU per process; multi-GPU runs are not part of this build "
              "(instances would shard across GPUs with no per-substep exchange).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import tressfx_amd as tfx
    from tressfx_amd import harness, synth

    tfx.init(0)
    torch.cuda.init()
    stream = torch.cuda.Stream()
    sh = ctypes_stream = stream.cuda_stream

    wl = synth.config2(num_guides=args.guides)
    a = wl.asset
    G, F, V = a.num_guide_strands, a.num_follow_strands_per_guide, a.num_vertices_per_strand
    N = a.num_total_vertices
    W, K = args.warmup, args.steps
    frames = list(range(W + K))
    drv = harness.FrameDriver(torch, wl, frames, fused=True)
    print(f"[bench] config2 G={G} F={F} V={V} N={N} substeps={wl.substeps} warmup={W} steps={K}",
          file=sys.stderr, flush=True)

    with torch.cuda.stream(stream):
        for f in range(W):
            drv.frame(f, sh)
        stream.synchronize()
        tfx.profiling_enable(True)
        for lbl in ("FusedSubstep", "InitializeSignedDistanceField", "ConstructSignedDistanceField",
                    "FinalizeSignedDistanceField"):
            tfx.profiling_read(lbl)
        evs = [torch.cuda.Event(enable_timing=True) for _ in range(K + 1)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        evs[0].record(stream)
        for k in range(K):
            drv.frame(W + k, sh)
            evs[k + 1].record(stream)
        stream.synchronize()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        tfx.profiling_enable(False)
    elapsed = t1 - t0
    fused_ms, fused_n = tfx.profiling_read("FusedSubstep")
    sdf_ms = sum(tfx.profiling_read(l)[0] for l in ("InitializeSignedDistanceField", "ConstructSignedDistanceField",
                                                     "FinalizeSignedDistanceField"))
    frame_ms = np.array([evs[i].elapsed_time(evs[i + 1]) for i in range(K)])
    value = N * wl.substeps * K / elapsed
    st = drv.hair.download_state()
    assert np.isfinite(st["pos"]).all(), "non-finite positions"

    bytes_sub = algorithmic_bytes_per_substep(G, F, V)
    avg_s = fused_ms / max(fused_n, 1) / 1e3
    achieved = bytes_sub / avg_s / 1e9
    traffic = read_traffic(os.path.join(REPO, "profiles", "fused_substep_config2_traffic.json"))

    # plain per-instance dispatch (reference launch structure), same workload
    plain = None
    if args.plain_steps > 0:
        pframes = list(range(W + args.plain_steps))
        pdrv = harness.FrameDriver(torch, wl, pframes, fused=False)
        with torch.cuda.stream(stream):
            for f in range(W):
                pdrv.frame(f, sh)
            stream.synchronize()
            p0 = time.perf_counter()
            for k in range(args.plain_steps):
                pdrv.frame(W + k, sh)
            stream.synchronize()
            p1 = time.perf_counter()
        pv = N * wl.substeps * args.plain_steps / (p1 - p0)
        plain = {"value": pv, "ms_per_step": (p1 - p0) * 1e3 / args.plain_steps, "steps": args.plain_steps}
        del pdrv

    cpu = None
    if not args.no_cpu:
        nthreads = max(1, min(16, os.cpu_count() or 1))
        cwl = synth.config2(num_guides=args.guides)
        cv, csub, cfr, cel = cpu_baseline(cwl, args.cpu_frames, nthreads, args.cpu_budget)
        cpu = {"value": cv, "unit": "vertex-substeps/s", "cores": nthreads, "kind": "port",
               "sample": f"FP32 transcription on the host, {cfr} frames = {csub} substeps of the same workload "
                         f"(SDF build + simulate + collide), {cel:.1f} s wall"}

    out = {
        "metric": "vertex-substeps/s",
        "value": value,
        "unit": "vertex-substeps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": elapsed * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"config2: {G} guide strands x {V} vertices, {F} follow strands per guide "
                        f"({N} vertices), 128^3 SDF at 2.5 mm rebuilt per frame from a "
                        f"{wl.mesh.num_triangles}-triangle noisy icosphere, wind 3 m/s +-20% gust, "
                        f"{wl.substeps} substeps/frame, length iterations 2; step = one frame",
            "guide_strands": G, "follow_per_guide": F, "vertices_per_strand": V, "substeps": wl.substeps,
            "path": "fused substep kernel (one persistent launch per substep)",
        },
        "latency_ms": {"p50": float(np.percentile(frame_ms, 50)), "p99": float(np.percentile(frame_ms, 99)),
                       "unit": "GPU ms per frame"},
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic.get("bytes_per_launch") if traffic else None,
            "kernel": "k_fused_substep<32,1024>",
            "algorithmic_bytes_per_launch": bytes_sub,
            "avg_launch_us": avg_s * 1e6,
            "launches": fused_n,
        },
        "sdf_build_ms_per_frame": sdf_ms / K,
        "plain": plain,
        "speedup_vs_plain": (value / plain["value"]) if plain else None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
