#!/usr/bin/env python3
"""Throughput of the TressFX 4.1 strand simulation + SDF collision path on one
MI355X, in vertex-substeps/s.

Workload (the one-GPU configuration the metric is quoted on): 20,000 guide
strands x 32 vertices with 4 follow strands per guide (3.2 M vertices), a
128^3 SDF (2.5 mm cells) rebuilt every frame from a 20,480-triangle noisy
icosphere, wind with a 0.3 Hz gust, 2 substeps per frame, length iterations 2.
A step is one frame: SDF rebuild + 2 x (substep incl. SDF collision; the
batch runs this workload's collision as a pass after each fused launch).
Inputs (asset, per-frame collision meshes) are resident in HBM before the timed
region; the per-substep parameter block and bone matrix are uploaded by the
library, as the reference uploads its constant buffer.

The timed region replays the frame as a hipGraph through the batched form
(one member).  Printed line: one JSON object with value = vertex-substeps/s,
the roofline of the dominant kernel (the fused substep; its duration from HIP
events around each launch in a second, directly launched leg; `substep_frac`
counts the collide pass too), the plain
per-instance dispatch for comparison, the 256-instance batched workload
against its plain dispatch, the served form (continuous batching over arriving
instances and step requests; and its capacity, every instance's frames queued
at once and drained) and the CPU transcription's throughput.  The headline and
every leg use correctly rounded arithmetic (bit-exact with the FP32
transcription); `fast_arithmetic` entries repeat the config-2 and config-3
legs with TFX_ARITH_FAST (hardware division and square root, within the FP64
tolerance).
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
    """The PMC traffic record of the fused kernel, if it was measured on the
    sources that are built now (else None)."""
    from tressfx_amd.profile_summary import fused_source_digest
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    return rec if rec.get("source_digest") == fused_source_digest() else None


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


def _timed_frames(torch, stream, drv, frames, **kw):
    """Runs `frames` on `stream` between two device synchronisations; returns
    (wall seconds, GPU ms per frame from events recorded between frames)."""
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(len(frames) + 1)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    evs[0].record(stream)
    for k, f in enumerate(frames):
        drv.frame(f, stream.cuda_stream, **kw)
        evs[k + 1].record(stream)
    stream.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    import numpy as np
    return t1 - t0, np.array([evs[i].elapsed_time(evs[i + 1]) for i in range(len(frames))])


def _log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def sdf_rebuild_bytes(cells, mesh_vertices, triangles):
    """Init write + finalize read/write of every cell (12 B), the mesh's
    vertices (32 B) and indices (12 B per triangle)."""
    return 12 * cells + 32 * mesh_vertices + 12 * triangles


def fur_leg(torch, tfx, harness, stream, wl5, wls3, args):
    """Config 5 alone -- the reference dispatch, the fused substep (pass
    timers around every launch: the V = 8 fused kernel's duration) and the
    batched form's graph-replayed frame over the same frames -- then, with
    the config-4 mix, the fur as one heavy tenant stepped at 60 frames/s."""
    import numpy as np
    from tressfx_amd import synth
    a = wl5.asset
    G, F, V = a.num_guide_strands, a.num_follow_strands_per_guide, a.num_vertices_per_strand
    n5, subs = a.num_total_vertices, wl5.substeps
    cells = int(np.prod(wl5.sdf_grid[2]))
    W5, K5, P5 = 2, args.fur_steps, args.fur_plain_steps
    legs = {}
    timers = None
    for mode, k in (("graph", K5), ("fused", K5), ("plain", P5)):
        if k <= 0:
            continue
        f5 = list(range(W5 + k))
        d5 = harness.SkinnedFrameDriver(torch, wl5, f5, mode=mode)
        with torch.cuda.stream(stream):
            for f in range(W5):
                d5.frame(f, stream.cuda_stream)
            stream.synchronize()
            if mode == "fused":
                tfx.profiling_enable(True)
                tfx.profiling_read("FusedSubstep")
            el5, ms5 = _timed_frames(torch, stream, d5, f5[W5:])
            if mode == "fused":
                tfx.profiling_enable(False)
                timers = tfx.profiling_read("FusedSubstep")
        legs[mode] = {"value": n5 * subs * k / el5, "ms_per_step": el5 * 1e3 / k, "steps": k,
                      "latency_ms": {"p50": float(np.percentile(ms5, 50)), "p99": float(np.percentile(ms5, 99))}}
        _log(f"fur {mode} leg {el5 * 1e3 / k:.3f} ms/frame")
        del d5
    bytes_sub = algorithmic_bytes_per_substep(G, F, V)
    bytes_frame = subs * bytes_sub + sdf_rebuild_bytes(cells, wl5.mesh.positions.shape[0], wl5.mesh.num_triangles)
    out = {
        "workload": f"config5: {G} guide strands x {V} vertices, {F} follow strands per guide ({n5} vertices), "
                    f"roots on a {wl5.mesh.num_triangles}-triangle noisy ellipsoid (0.5 x 0.3 x 0.3 m) skinned "
                    f"by 4 bones every frame, {wl5.sdf_grid[2][0]}^3 SDF rebuilt per frame, {subs} substeps, "
                    f"length iterations 2, VSP 0.7",
        "unit": "vertex-substeps/s",
        "algorithmic_bytes_per_frame": bytes_frame,
        "legs": legs,
    }
    if "graph" in legs:
        g = legs["graph"]
        out["value"] = g["value"]
        out["frame_hbm_frac"] = bytes_frame / (g["ms_per_step"] * 1e-3) / (HBM_PEAK_GBS * 1e9)
        if "plain" in legs:
            out["speedup_vs_plain"] = g["value"] / legs["plain"]["value"]
    if timers and timers[1]:
        avg_s = timers[0] / timers[1] / 1e3
        out["roofline"] = {"bound": "hbm", "kernel": f"k_fused_substep<{V},*>", "avg_launch_us": avg_s * 1e6,
                           "launches": timers[1], "algorithmic_bytes_per_launch": bytes_sub,
                           "achieved": bytes_sub / avg_s / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": bytes_sub / avg_s / 1e9 / HBM_PEAK_GBS}
    if wls3 is not None and args.fur_served_seconds > 0:
        srv = tfx.HairServer(hbm_budget_bytes=200 << 30, max_instances=1000, substeps_per_frame=2,
                             substep_dt=float(np.float32(1.0 / synth.FRAME_RATE / 2)), max_inflight=2)
        run = harness.ServedRun(srv, wls3, initial=args.served_initial, duration_s=args.fur_served_seconds,
                                inst_rate=20.0, lifetime_s=30.0, req_rate=5000.0, ks=(1, 2, 4), tenants=[wl5],
                                tenant_hz=60.0)
        _log(f"served window with the fur tenant, {args.fur_served_seconds:.0f} s")
        res = run.run(log=_log)
        srv.close()
        out["in_mix"] = {
            "workload": "the config-4 request mix (as the served leg) with the fur resident as one more instance "
                        "(3 substeps per frame), its client requesting one frame every 1/60 s",
            "value": res["vertex_substeps"] / res["elapsed_s"], "unit": "vertex-substeps/s",
            "latency_ms": {"p50": res["latency_p50_ms"], "p99": res["latency_p99_ms"], "max": res["latency_max_ms"]},
            "tenant": res["tenants"][0], "passes": res["passes"],
            "mean_instances_per_pass": res["mean_instances_per_pass"], "window_s": res["elapsed_s"],
        }
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=600, help="timed frames of the config-2 workload")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--plain-steps", type=int, default=20)
    ap.add_argument("--batched-steps", type=int, default=20, help="timed frames of the 256-instance workload")
    ap.add_argument("--batched-plain-steps", type=int, default=3)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-fast", action="store_true", help="skip the TFX_ARITH_FAST legs")
    ap.add_argument("--no-autotune", action="store_true",
                    help="batched legs keep the occupancy heuristic's launch shapes instead of the autotuner's")
    ap.add_argument("--served-seconds", type=float, default=12.0,
                    help="length of the served-form request window (0: skip)")
    ap.add_argument("--served-initial", type=int, default=256, help="instances resident when the window opens")
    ap.add_argument("--served-burst-frames", type=int, default=24,
                    help="served capacity: frames queued at once on every config-3 instance (0: skip)")
    ap.add_argument("--fur-steps", type=int, default=10, help="timed frames of the config-5 fur workload")
    ap.add_argument("--fur-plain-steps", type=int, default=3)
    ap.add_argument("--fur-served-seconds", type=float, default=8.0,
                    help="served window with the fur as a heavy tenant (0: skip)")
    ap.add_argument("--no-fur", action="store_true")
    ap.add_argument("--cpu-frames", type=int, default=600)
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--guides", type=int, default=20000)
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="after the legs, copy a device buffer of this many bytes 3 times (a known-size "
                         "kernel that calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this path runs on one GPU per process; multi-GPU runs are not part of this build "
              "(instances would shard across GPUs with no per-substep exchange).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from tressfx_amd import synth

    host_workers = max(1, min(16, os.cpu_count() or 1))
    # Workloads are generated before the process touches the GPU (the
    # batched one by forked workers).
    wl = synth.config2(num_guides=args.guides)
    wls3 = None
    if not args.no_batched:
        _log(f"generating the 256-instance workload with {host_workers} processes")
        wls3 = synth.config3(workers=host_workers)
    wl5 = None
    if not args.no_fur:
        _log("generating the fur workload (1,000,000 x 8 on a 200,000-triangle body)")
        wl5 = synth.config5()

    import torch
    import tressfx_amd as tfx
    from tressfx_amd import harness

    tfx.init(0)
    torch.cuda.init()
    stream = torch.cuda.Stream()

    a = wl.asset
    G, F, V = a.num_guide_strands, a.num_follow_strands_per_guide, a.num_vertices_per_strand
    N = a.num_total_vertices
    W, K = args.warmup, args.steps
    frames = list(range(W + K))
    drv = harness.BatchDriver(torch, [wl], frames)
    _log(f"config2 G={G} F={F} V={V} N={N} substeps={wl.substeps} warmup={W} steps={K}")

    with torch.cuda.stream(stream):
        for f in range(W):
            drv.frame(f, stream.cuda_stream, use_graph=True)
        stream.synchronize()
        elapsed, frame_ms = _timed_frames(torch, stream, drv, frames[W:], use_graph=True)
        _log(f"graph leg {elapsed * 1e3 / K:.3f} ms/frame")
    st = drv.hairs[0].download_state()
    assert np.isfinite(st["pos"]).all(), "non-finite positions"
    value = N * wl.substeps * K / elapsed

    # Kernel leg: a second object runs the same frames from the same initial
    # state, launched directly with a HIP event pair around every pass on the
    # launch stream, so the fused kernel's average duration covers the same
    # launches (same states, same per-launch work) as the timed region.
    labels = ("FusedSubstep", "InitializeSignedDistanceField", "ConstructSignedDistanceField",
              "FinalizeSignedDistanceField", "BatchFrame", "CollideHairVerticesWithSdf")
    kdrv = harness.BatchDriver(torch, [wl], frames, meshes=drv.meshes)
    with torch.cuda.stream(stream):
        for f in range(W):
            kdrv.frame(f, stream.cuda_stream, use_graph=False)
        stream.synchronize()
        tfx.profiling_enable(True)
        for lbl in labels:
            tfx.profiling_read(lbl)
        _timed_frames(torch, stream, kdrv, frames[W:], use_graph=False)
        tfx.profiling_enable(False)
        timers = {lbl: tfx.profiling_read(lbl) for lbl in labels}
    kst = kdrv.hairs[0].download_state()
    graph_equals_direct = bool(all(np.array_equal(st[k].view(np.uint32), kst[k].view(np.uint32))
                                   for k in ("pos", "prev", "tangent")))
    del drv, kdrv

    fused_ms, fused_n = timers["FusedSubstep"]
    sdf_ms = sum(timers[l][0] for l in labels[1:4])
    bytes_sub = algorithmic_bytes_per_substep(G, F, V)
    avg_s = fused_ms / max(fused_n, 1) / 1e3
    achieved = bytes_sub / avg_s / 1e9
    # The batch defers this workload's SDF collision to a collide pass after
    # each fused launch (its 625 strand blocks fit in one round of resident
    # blocks): the substep is the two launches.
    coll_ms, coll_n = timers["CollideHairVerticesWithSdf"]
    coll_s = coll_ms / max(coll_n, 1) / 1e3 if coll_n else 0.0
    sub_s = avg_s + coll_s
    traffic = read_traffic(os.path.join(REPO, "profiles", "fused_substep_config2_traffic.json"))

    # plain per-instance dispatch (reference launch structure), same workload
    plain = None
    if args.plain_steps > 0:
        pframes = list(range(W + args.plain_steps))
        pdrv = harness.FrameDriver(torch, wl, pframes, fused=False)
        with torch.cuda.stream(stream):
            for f in range(W):
                pdrv.frame(f, stream.cuda_stream)
            pel, _ = _timed_frames(torch, stream, pdrv, pframes[W:])
        plain = {"value": N * wl.substeps * args.plain_steps / pel, "ms_per_step": pel * 1e3 / args.plain_steps,
                 "steps": args.plain_steps}
        _log(f"plain leg {plain['ms_per_step']:.3f} ms/frame")
        del pdrv

    # The same legs with TFX_ARITH_FAST (hardware division and square root in
    # the simulation and collision kernels, within the FP64 tolerance).
    fast = None
    if not args.no_fast:
        tfx.set_arithmetic(tfx.TFX_ARITH_FAST)
        try:
            fdrv = harness.BatchDriver(torch, [wl], frames)
            with torch.cuda.stream(stream):
                for f in range(W):
                    fdrv.frame(f, stream.cuda_stream, use_graph=True)
                fel, _ = _timed_frames(torch, stream, fdrv, frames[W:], use_graph=True)
            del fdrv
            fast = {"value": N * wl.substeps * K / fel, "ms_per_step": fel * 1e3 / K, "steps": K}
            if args.plain_steps > 0:
                pdrv = harness.FrameDriver(torch, wl, pframes, fused=False)
                with torch.cuda.stream(stream):
                    for f in range(W):
                        pdrv.frame(f, stream.cuda_stream)
                    pel, _ = _timed_frames(torch, stream, pdrv, pframes[W:])
                del pdrv
                fast["plain"] = {"value": N * wl.substeps * args.plain_steps / pel,
                                 "ms_per_step": pel * 1e3 / args.plain_steps}
                fast["speedup_vs_plain"] = fast["value"] / plain["value"]
            _log(f"fast-arithmetic graph leg {fast['ms_per_step']:.3f} ms/frame")
        finally:
            tfx.set_arithmetic(tfx.TFX_ARITH_IEEE)

    batched = None
    if wls3 is not None:
        n3 = sum(w.asset.num_total_vertices for w in wls3)
        W3, K3, P3 = 3, args.batched_steps, args.batched_plain_steps
        f3 = list(range(W3 + K3))
        bdrv = harness.BatchDriver(torch, wls3, f3)
        tune = 2 if not args.no_autotune else 0
        with torch.cuda.stream(stream):
            for f in range(W3):
                bdrv.frame(f, stream.cuda_stream, use_graph=True, autotune_repeats=tune if f == W3 - 1 else 0)
            bel, bms = _timed_frames(torch, stream, bdrv, f3[W3:], use_graph=True)
        _log(f"batched leg {bel * 1e3 / K3:.3f} ms/frame")
        bval = n3 * 2 * K3 / bel
        autotune = getattr(bdrv, "autotune_records", None)
        del bdrv
        pval = None
        if P3 > 0:
            pf3 = list(range(1 + P3))
            pdrv = harness.PlainMultiDriver(torch, wls3, pf3)
            with torch.cuda.stream(stream):
                pdrv.frame(0, stream.cuda_stream)
                pel, _ = _timed_frames(torch, stream, pdrv, pf3[1:])
            pval = n3 * 2 * P3 / pel
            _log(f"batched-plain leg {pel * 1e3 / P3:.3f} ms/frame")
            del pdrv
        lengths = {v: sum(w.asset.num_guide_strands for w in wls3 if w.asset.num_vertices_per_strand == v)
                   for v in (16, 32, 64)}
        batched = {
            "workload": f"config3: {len(wls3)} instances, {sum(w.asset.num_guide_strands for w in wls3)} strands "
                        f"(log-uniform 500..50000 per instance; by vertices per strand {lengths}), {n3} vertices, "
                        f"per-instance rigid trajectories and 64^3..128^3 SDFs rebuilt per frame from "
                        f"5120-triangle noisy icospheres, 2 substeps; one graph replay per frame",
            "value": bval, "unit": "vertex-substeps/s", "ms_per_step": bel * 1e3 / K3, "steps": K3,
            "latency_ms": {"p50": float(np.percentile(bms, 50)), "p99": float(np.percentile(bms, 99))},
            "plain": {"value": pval, "steps": P3} if pval else None,
            "speedup_vs_plain": (bval / pval) if pval else None,
            "autotune": autotune,
        }
        if not args.no_fast:
            tfx.set_arithmetic(tfx.TFX_ARITH_FAST)
            try:
                bdrv = harness.BatchDriver(torch, wls3, f3)
                with torch.cuda.stream(stream):
                    for f in range(W3):
                        bdrv.frame(f, stream.cuda_stream, use_graph=True, autotune_repeats=tune if f == W3 - 1 else 0)
                    fel, fms = _timed_frames(torch, stream, bdrv, f3[W3:], use_graph=True)
                fast_tune = getattr(bdrv, "autotune_records", None)
                del bdrv
                fb = {"value": n3 * 2 * K3 / fel, "ms_per_step": fel * 1e3 / K3, "steps": K3,
                      "latency_ms": {"p50": float(np.percentile(fms, 50)), "p99": float(np.percentile(fms, 99))},
                      "speedup_vs_plain": (n3 * 2 * K3 / fel / pval) if pval else None, "autotune": fast_tune}
                if P3 > 0:
                    pdrv = harness.PlainMultiDriver(torch, wls3, pf3)
                    with torch.cuda.stream(stream):
                        pdrv.frame(0, stream.cuda_stream)
                        pel, _ = _timed_frames(torch, stream, pdrv, pf3[1:])
                    del pdrv
                    fb["plain"] = {"value": n3 * 2 * P3 / pel, "steps": P3}
                    fb["speedup_vs_fast_plain"] = fb["value"] / fb["plain"]["value"]
                batched["fast_arithmetic"] = fb
                _log(f"fast-arithmetic batched leg {fb['ms_per_step']:.3f} ms/frame")
            finally:
                tfx.set_arithmetic(tfx.TFX_ARITH_IEEE)

    served = None
    if wls3 is not None and args.served_seconds > 0:
        srv = tfx.HairServer(hbm_budget_bytes=200 << 30, max_instances=1000, substeps_per_frame=2,
                             substep_dt=float(np.float32(1.0 / synth.FRAME_RATE / 2)), max_inflight=2)
        run = harness.ServedRun(srv, wls3, initial=args.served_initial, duration_s=args.served_seconds,
                                inst_rate=20.0, lifetime_s=30.0, req_rate=5000.0, ks=(1, 2, 4))
        _log(f"served window {args.served_seconds:.0f} s, {args.served_initial} instances resident at start")
        res = run.run(log=_log)
        srv.close()
        sv = res["vertex_substeps"] / res["elapsed_s"]
        served = {
            "workload": "config4: instances drawn from the 256 config-3 instances, Poisson arrivals 20/s, "
                        "lifetimes Exp(30 s), step requests Poisson 5000/s of k in {1,2,4} frames on random "
                        f"resident instances, 200 GB admission budget, {args.served_initial} resident at start, "
                        f"{args.served_seconds:.0f} s window",
            "value": sv, "unit": "vertex-substeps/s",
            "latency_ms": {"p50": res["latency_p50_ms"], "p99": res["latency_p99_ms"], "mean": res["latency_mean_ms"],
                           "max": res["latency_max_ms"]},
            "requests_completed": res["completed_requests"], "frames_completed": res["completed_frames"],
            "offered_request_rate": res["offered_request_rate"], "passes": res["passes"],
            "mean_instances_per_pass": res["mean_instances_per_pass"], "window_s": res["elapsed_s"],
            "instances_retired": res["instances_retired"], "admissions_refused": res["admissions_refused"],
            "resident_at_end": res["resident_at_end"],
            "speedup_vs_plain": (sv / batched["plain"]["value"]) if batched and batched.get("plain") else None,
        }
        if args.served_burst_frames > 0:
            srv = tfx.HairServer(hbm_budget_bytes=200 << 30, max_instances=1000, substeps_per_frame=2,
                                 substep_dt=float(np.float32(1.0 / synth.FRAME_RATE / 2)), max_inflight=2)
            brun = harness.ServedRun(srv, wls3, initial=0, duration_s=0.0, inst_rate=0.0, req_rate=0.0)
            _log(f"served capacity: {len(wls3)} instances x {args.served_burst_frames} frames queued at once")
            bres = brun.burst(args.served_burst_frames)
            bv = bres["vertex_substeps"] / bres["window_s"]
            served["capacity"] = {
                "workload": f"the {len(wls3)} config-3 instances resident, {args.served_burst_frames} frames of each "
                            "submitted at once and drained (requests outrun the GPU: every pass coalesces one "
                            "frame of every instance)",
                "value": bv, "unit": "vertex-substeps/s", "window_s": bres["window_s"], "passes": bres["passes"],
                "mean_instances_per_pass": bres["mean_instances_per_pass"],
                "speedup_vs_plain": (bv / batched["plain"]["value"]) if batched and batched.get("plain") else None,
            }
            if not args.no_fast:
                tfx.set_arithmetic(tfx.TFX_ARITH_FAST)
                try:
                    fres = brun.burst(args.served_burst_frames)
                finally:
                    tfx.set_arithmetic(tfx.TFX_ARITH_IEEE)
                fv = fres["vertex_substeps"] / fres["window_s"]
                served["capacity"]["fast_arithmetic"] = {
                    "value": fv, "window_s": fres["window_s"], "passes": fres["passes"],
                    "speedup_vs_plain": (fv / batched["plain"]["value"]) if batched and batched.get("plain") else None,
                }
            srv.close()

    fur = None
    if wl5 is not None:
        fur = fur_leg(torch, tfx, harness, stream, wl5, wls3, args)

    cpu = None
    if not args.no_cpu:
        cwl = synth.config2(num_guides=args.guides)
        cv, csub, cfr, cel = cpu_baseline(cwl, args.cpu_frames, host_workers, args.cpu_budget)
        cpu = {"value": cv, "unit": "vertex-substeps/s", "cores": host_workers, "kind": "port",
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
            "path": "batched form, one member: per frame one hipGraph replay = parameter upload, "
                    "SDF build (4 launches) beside the first substep's fused launch, per substep a fused launch "
                    "and a collide pass",
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
            "collide_pass_us": coll_s * 1e6,
            "collide_passes": coll_n,
            "substep_achieved": bytes_sub / sub_s / 1e9,
            "substep_frac": bytes_sub / sub_s / 1e9 / HBM_PEAK_GBS,
        },
        "sdf_build_ms_per_frame": sdf_ms / K,
        "graph_equals_direct": graph_equals_direct,
        "plain": plain,
        "fast_arithmetic": fast,
        "speedup_vs_plain": (value / plain["value"]) if plain else None,
        "batched": batched,
        "served": served,
        "fur": fur,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)

    if args.copy_bytes > 0:
        n = args.copy_bytes // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst


if __name__ == "__main__":
    main()
