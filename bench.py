#!/usr/bin/env python3
"""Throughput of the TressFX 4.1 strand simulation + SDF collision path on one
MI355X, in vertex-substeps/s.

Headline workload (the batched form the metric is quoted on): 256 hair
instances of uneven size -- strand counts log-uniform in [500, 50,000],
16/32/64 vertices per strand (2.92 M strands, 76.4 M vertices) -- each with its
own rigid head trajectory and its own 64^3-128^3 SDF rebuilt every frame from
a 5,120-triangle noisy icosphere, wind, 2 substeps per frame.  A step is one
frame: one hipGraph replay = parameter upload, the batched SDF build, and per
substep one persistent fused launch per strand length.  Inputs (assets,
per-frame collision meshes) are resident in HBM before the timed region.  The
headline runs the fast arithmetic (hardware reciprocal and square root, fused
multiply-adds), which the test suite holds to 1e-5 relative per substep
against the FP64 transcription at this size; `bit_exact_arithmetic` repeats
the workload in the correctly rounded arithmetic that equals the FP32
transcription bit for bit.

Printed line: one JSON object -- value, per-frame latency, the roofline of
the dominant kernel (the fused launch of the strand length that takes the
most time; its duration from HIP events stamped by each launch, in a second,
directly launched leg over the same frames from the same initial state), the
plain per-instance dispatch in the same arithmetic (the denominator of the
batched-vs-plain ratio), the correctly rounded repeat, the single large
instance of config 2, the served form (continuous batching at the Poisson
steady state of resident instances, and its capacity) and the config-5 fur,
and the CPU transcription's throughput on a bounded sample.
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


def read_traffic(path, kernel=None):
    """The PMC traffic record of the fused kernel, if it was measured on the
    sources and build flags that are built now, for `kernel` (else None)."""
    from tressfx_amd.profile_summary import fused_source_digest
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    if rec.get("source_digest") != fused_source_digest():
        return None
    return rec if kernel is None or rec.get("kernel") == kernel else None


def cpu_threads():
    """Host threads for the CPU transcription: the process's CPU share (the
    GPU box gives one GPU's share of its cores), and the machine's own count."""
    share = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cap = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return (min(share, cap) if cap > 0 else share), share, os.cpu_count() or 1


def lscpu_info():
    """Model name and core counts of the host, from lscpu."""
    import subprocess
    out = {}
    try:
        text = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
    except (OSError, subprocess.SubprocessError):
        return out
    keys = {"Model name": "model", "CPU(s)": "cpus", "Thread(s) per core": "threads_per_core",
            "Core(s) per socket": "cores_per_socket", "Socket(s)": "sockets"}
    for line in text.splitlines():
        k, _, v = line.partition(":")
        if k.strip() in keys and keys[k.strip()] not in out:
            out[keys[k.strip()]] = v.strip()
    try:
        out["physical_cores"] = int(out["cores_per_socket"]) * int(out["sockets"])
    except (KeyError, ValueError):
        pass
    return out


def cpu_frames(wls, frames, nthreads):
    """The FP32 transcription (oracle) on the host, multithreaded over strands
    (and triangles for the SDF build): for each workload, from its rest state,
    `frames` frames of SDF build + per substep simulate and collide, in the
    GPU's frame order.  Returns (vertex-substeps, seconds); only the
    transcription's calls are timed (not the input preparation)."""
    import numpy as np
    import oracle
    import tressfx_amd as tfx
    from tressfx_amd import synth
    work, el = 0, 0.0
    bones = np.tile(np.eye(4, dtype=np.float32).reshape(1, 16), (tfx.TFX_MAX_NUM_BONES, 1))
    for wl in wls:
        a = wl.asset
        h = oracle.OracleHair(a, np.float32)
        p = tfx.SimulationParams()
        origin, cell, dims = wl.sdf_grid
        origin = np.asarray(origin, np.float32)
        margin = np.float32(np.float32(wl.collision_margin) * np.float32(cell))
        dt = 1.0 / synth.FRAME_RATE
        settings = tfx.SimulationSettings.from_buffer_copy(wl.settings)
        for frame in range(frames):
            t = frame * dt
            verts = wl.mesh.vertices(t, wl.bone_fn(t))
            t0 = time.perf_counter()
            grid = oracle.sdf_build(verts, wl.mesh.indices, origin, np.float32(cell), dims, np.float32, nthreads)
            el += time.perf_counter() - t0
            for s in range(wl.substeps):
                ts = (frame + (s + 1) / wl.substeps) * dt
                bones[0] = np.asarray(wl.bone_fn(ts), np.float32).reshape(16)
                if wl.wind_fn is not None:
                    settings.wind_magnitude = wl.wind_fn(ts)
                oracle.pack_sim_params(settings, np.float32(dt / wl.substeps), frame * wl.substeps + s,
                                       a.num_vertices_per_strand, a.num_follow_strands_per_guide, into=p)
                t0 = time.perf_counter()
                h.simulate(p, bones, nthreads)
                oracle.sdf_collide(h.pos, h.prev, a.num_vertices_per_strand, grid, origin, np.float32(cell), dims,
                                   margin, nthreads)
                el += time.perf_counter() - t0
                p.reset_positions = 0.0
                work += a.num_total_vertices
    return work, el


def cpu_leg(wls3, wl2, runs=3):
    """The CPU baseline: the transcription on a bounded sample of the headline
    workload (every 4th config-3 instance, 2 frames each) with every thread of
    the process's CPU share and on one thread (every 32nd instance, 1 frame),
    and config 2 (the single large instance) the same two ways; each the
    median of `runs` runs."""
    import statistics
    threads, share, hw = cpu_threads()
    host = lscpu_info()
    host.update({"hardware_concurrency": hw, "process_cpu_share": share})

    def med(fn):
        vals = []
        for _ in range(runs):
            w, el = fn()
            vals.append(w / el)
        return statistics.median(vals), vals

    s_multi = [wls3[i] for i in range(0, len(wls3), 4)]
    s_single = [wls3[i] for i in range(0, len(wls3), 32)]
    v_multi, r_multi = med(lambda: cpu_frames(s_multi, 2, threads))
    _log(f"cpu config3 sample, {threads} threads: {v_multi / 1e6:.1f} M vertex-substeps/s")
    v_single, r_single = med(lambda: cpu_frames(s_single, 1, 1))
    _log(f"cpu config3 sample, 1 thread: {v_single / 1e6:.2f} M vertex-substeps/s")
    c2_multi, c2_rm = med(lambda: cpu_frames([wl2], 4, threads))
    c2_single, c2_rs = med(lambda: cpu_frames([wl2], 1, 1))
    _log(f"cpu config2: {c2_multi / 1e6:.1f} M ({threads} threads), {c2_single / 1e6:.2f} M (1 thread)")
    nv_multi = sum(w.asset.num_total_vertices for w in s_multi)
    nv_single = sum(w.asset.num_total_vertices for w in s_single)
    return {
        "value": v_multi, "unit": "vertex-substeps/s", "cores": threads, "kind": "port",
        "sample": f"FP32 transcription (oracle) on the host: {len(s_multi)} of the 256 config-3 instances "
                  f"(every 4th, {nv_multi} vertices), 2 frames each (SDF build + 2 x (simulate + collide)), "
                  f"{threads} threads; median of {runs} runs",
        "runs": r_multi,
        "single_thread": {"value": v_single, "cores": 1, "runs": r_single,
                          "sample": f"{len(s_single)} config-3 instances (every 32nd, {nv_single} vertices), "
                                    f"1 frame each, one thread; median of {runs} runs"},
        "config2": {"value": c2_multi, "runs": c2_rm, "single_thread": c2_single, "single_thread_runs": c2_rs,
                    "sample": f"config 2 (20,000 x 32 + 4 follow, 128^3 SDF), 4 frames on {threads} threads, "
                              f"1 frame on one thread; median of {runs} runs each"},
        "host": host,
        "threads_note": "threads = the process's CPU share on the GPU box (the cores one GPU's runs may use), "
                        "capped by OMP_NUM_THREADS when set",
    }


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
    timers, vb5 = None, None
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
                per_vb = {vb: tfx.profiling_read(f"FusedSubstep/V{V}/B{vb}") for vb in (512, 1024, 2048)}
                vb5 = max(per_vb, key=lambda k: per_vb[k][0])
                timers = (sum(t[0] for t in per_vb.values()), sum(t[1] for t in per_vb.values()))
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
        out["roofline"] = {"bound": "hbm", "kernel": fused_kernel_name(V, vb5, "fast" if tfx.get_arithmetic() ==
                                                                  tfx.TFX_ARITH_FAST else "ieee"),
                           "avg_launch_us": avg_s * 1e6,
                           "launches": timers[1], "algorithmic_bytes_per_launch": bytes_sub,
                           "achieved": bytes_sub / avg_s / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": bytes_sub / avg_s / 1e9 / HBM_PEAK_GBS}
    if wls3 is not None and args.fur_served_seconds > 0:
        srv = tfx.HairServer(hbm_budget_bytes=200 << 30, max_instances=1000, substeps_per_frame=2,
                             substep_dt=float(np.float32(1.0 / synth.FRAME_RATE / 2)), max_inflight=2)
        initial, _ = served_initial(args)
        run = harness.ServedRun(srv, wls3, initial=initial, duration_s=args.fur_served_seconds,
                                inst_rate=20.0, lifetime_s=30.0, req_rate=5000.0, ks=(1, 2, 4), tenants=[wl5],
                                tenant_hz=60.0)
        _log(f"served window with the fur tenant, {args.fur_served_seconds:.0f} s, {initial} instances resident "
             "at start")
        res = run.run(log=_log)
        srv.close()
        out["in_mix"] = {
            "workload": "the config-4 request mix (as the served leg, opened with "
                        f"{initial} instances resident) with the fur resident as one more instance "
                        "(3 substeps per frame), its client requesting one frame every 1/60 s",
            "resident_at_start": initial, "resident_at_end": res["resident_at_end"],
            "value": res["vertex_substeps"] / res["elapsed_s"], "unit": "vertex-substeps/s",
            "latency_ms": {"p50": res["latency_p50_ms"], "p99": res["latency_p99_ms"], "max": res["latency_max_ms"]},
            "tenant": res["tenants"][0], "passes": res["passes"],
            "mean_instances_per_pass": res["mean_instances_per_pass"], "window_s": res["elapsed_s"],
        }
    return out


def class_bytes(wls):
    """Algorithmic bytes of one substep per strand length (guide and follow
    streams of every member of that length)."""
    out = {}
    for w in wls:
        a = w.asset
        V = a.num_vertices_per_strand
        out[V] = out.get(V, 0) + algorithmic_bytes_per_substep(a.num_guide_strands, a.num_follow_strands_per_guide, V)
    return out


def frame_sdf_bytes(wls):
    import numpy as np
    return sum(sdf_rebuild_bytes(int(np.prod(w.sdf_grid[2])), w.mesh.vertices(0.0).shape[0], w.mesh.num_triangles)
               for w in wls if w.mesh is not None)


def fused_kernel_name(V, variant_vertices, arith):
    ns = "fast" if arith == "fast" else "exact"
    return f"tfx::{ns}::k_fused_substep<{V}, {variant_vertices}>"


def batched_leg(torch, tfx, harness, stream, wls, arith, W, K, plain_steps, tune=2, kernel_leg=True, label=""):
    """The workloads through the batched form: W warm-up frames (the
    autotuner runs on the last), K timed graph-replayed frames; then a second
    set of objects runs the same frames from the same initial state with direct
    launches and pass timers (per strand length: each launch stamped by its
    own begin and end), ending bit-identical; then the plain per-instance
    dispatch of the same workloads.  `arith`: "fast" or "ieee"."""
    import numpy as np
    mode = tfx.TFX_ARITH_FAST if arith == "fast" else tfx.TFX_ARITH_IEEE
    tfx.set_arithmetic(mode)
    try:
        frames = list(range(W + K))
        n = sum(w.asset.num_total_vertices for w in wls)
        subs = wls[0].substeps
        bdrv = harness.BatchDriver(torch, wls, frames)
        with torch.cuda.stream(stream):
            for f in range(W):
                bdrv.frame(f, stream.cuda_stream, use_graph=True, autotune_repeats=tune if f == W - 1 else 0)
            stream.synchronize()
            el, ms = _timed_frames(torch, stream, bdrv, frames[W:], use_graph=True)
        _log(f"{label} batched graph leg ({arith}) {el * 1e3 / K:.3f} ms/frame (GPU p50 {np.percentile(ms, 50):.3f} ms)")
        records = getattr(bdrv, "autotune_records", None)
        for r in records or []:
            _log(f"{label} autotune V={r['vertices_per_strand']}: variant {r['variant']} x {r['blocks_per_cu']}/CU "
                 f"{r['chosen_us']:.0f} us (heuristic {r['default_us']:.0f} us), deferred {r['sdf_deferred']}")
        out = {"value": n * subs * K / el, "unit": "vertex-substeps/s", "ms_per_step": el * 1e3 / K, "steps": K,
               "warmup": W, "latency_ms": {"p50": float(np.percentile(ms, 50)), "p99": float(np.percentile(ms, 99)),
                                            "unit": "GPU ms per frame"},
               "arithmetic": arith, "autotune": records}
        if kernel_leg:
            kdrv = harness.BatchDriver(torch, wls, frames, meshes=bdrv.meshes)
            for r in records or []:
                kdrv.batch.set_launch_shape(r["vertices_per_strand"], r["variant"], r["blocks_per_cu"])
            lengths = sorted(class_bytes(wls))
            # fused launches are timed per strand length and strand-block size
            fused_labels = {(V, vb): f"FusedSubstep/V{V}/B{vb}" for V in lengths for vb in (512, 1024, 2048)}
            labels = list(fused_labels.values()) + [
                "InitializeSignedDistanceField", "ConstructSignedDistanceField", "FinalizeSignedDistanceField",
                "CollideHairVerticesWithSdf"]
            with torch.cuda.stream(stream):
                for f in range(W):
                    kdrv.frame(f, stream.cuda_stream, use_graph=False)
                stream.synchronize()
                tfx.profiling_enable(True)
                for lbl in labels:
                    tfx.profiling_read(lbl)
                kel, kms = _timed_frames(torch, stream, kdrv, frames[W:], use_graph=False)
                tfx.profiling_enable(False)
                timers = {lbl: tfx.profiling_read(lbl) for lbl in labels}
            out["direct_launch_ms_per_step"] = kel * 1e3 / K
            same = True
            for hb, hk in zip(bdrv.hairs, kdrv.hairs):
                a = hb.download_state(keys=("pos", "prev"))
                b = hk.download_state(keys=("pos", "prev"))
                same = same and all(np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)) for k in a)
                if not same:
                    break
            out["graph_equals_direct"] = bool(same)
            del kdrv
            cb = class_bytes(wls)
            classes = []
            for V in lengths:
                by_vb = {vb: timers[lbl] for (v, vb), lbl in fused_labels.items() if v == V and timers[lbl][1]}
                if not by_vb:
                    continue
                ms_tot = sum(t[0] for t in by_vb.values())
                launches = sum(t[1] for t in by_vb.values())
                avg_s = ms_tot / launches / 1e3
                vb = max(by_vb, key=lambda k: by_vb[k][0])      # the block size that took the time
                classes.append({"vertices_per_strand": V, "kernel": fused_kernel_name(V, vb, arith),
                                "algorithmic_bytes_per_launch": cb[V], "avg_launch_us": avg_s * 1e6,
                                "launches": launches, "achieved": cb[V] / avg_s / 1e9,
                                "frac": cb[V] / avg_s / 1e9 / HBM_PEAK_GBS,
                                "ms_per_frame": ms_tot / K})
            out["fused_classes"] = classes
            out["sdf_build_ms_per_frame"] = {k: timers[k][0] / K for k in labels if "SignedDistance" in k}
            out["collide_pass_ms_per_frame"] = timers["CollideHairVerticesWithSdf"][0] / K
            fused_ms = sum(c["ms_per_frame"] for c in classes)
            out["fused_ms_per_frame"] = fused_ms
            _log(f"{label} kernel leg ({arith}): {kel * 1e3 / K:.3f} ms/frame direct; fused {fused_ms:.3f} ms/frame "
                 + ", ".join(f"V{c['vertices_per_strand']} {c['avg_launch_us']:.0f} us ({c['frac']:.3f})" for c in classes)
                 + f"; SDF build {sum(out['sdf_build_ms_per_frame'].values()):.3f} ms/frame")
            if fused_ms > 0:
                out["fused_all_classes_frac"] = subs * sum(cb.values()) / (fused_ms * 1e-3) / 1e9 / HBM_PEAK_GBS
            out["frame_hbm_frac"] = (subs * sum(cb.values()) + frame_sdf_bytes(wls)) / (el / K) / 1e9 / HBM_PEAK_GBS
        del bdrv
        if plain_steps > 0:
            pf = list(range(1 + plain_steps))
            pdrv = harness.PlainMultiDriver(torch, wls, pf) if len(wls) > 1 else harness.FrameDriver(
                torch, wls[0], pf, fused=False)
            with torch.cuda.stream(stream):
                pdrv.frame(0, stream.cuda_stream)
                pel, _ = _timed_frames(torch, stream, pdrv, pf[1:])
            del pdrv
            out["plain"] = {"value": n * subs * plain_steps / pel, "ms_per_step": pel * 1e3 / plain_steps,
                            "steps": plain_steps, "arithmetic": arith,
                            "path": "per-instance dispatch with the reference's launch structure: per frame each "
                                    "SDF Update (3 launches), per substep the six simulation stages of every "
                                    "object (local shape repeated per iteration) and CollideWithHair per object"}
            out["speedup_vs_plain"] = out["value"] / out["plain"]["value"]
            _log(f"{label} plain leg ({arith}) {pel * 1e3 / plain_steps:.3f} ms/frame")
        return out
    finally:
        tfx.set_arithmetic(tfx.TFX_ARITH_IEEE)


def roofline_of(leg, traffic):
    """The roofline object of a batched leg: its dominant kernel (the strand
    length whose fused launches take the most time)."""
    cls = leg.get("fused_classes") or []
    if not cls:
        return None
    dom = max(cls, key=lambda c: c["ms_per_frame"])
    return {
        "bound": "hbm", "achieved": dom["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": dom["frac"],
        "traffic": traffic.get("bytes_per_launch") if traffic else None,
        "kernel": dom["kernel"], "algorithmic_bytes_per_launch": dom["algorithmic_bytes_per_launch"],
        "avg_launch_us": dom["avg_launch_us"], "launches": dom["launches"],
        "classes": [{k: c[k] for k in ("vertices_per_strand", "kernel", "algorithmic_bytes_per_launch",
                                        "avg_launch_us", "achieved", "frac")} for c in cls],
        "all_classes_frac": leg.get("fused_all_classes_frac"),
        "frame_frac": leg.get("frame_hbm_frac"),
    }


def served_initial(args):
    """Instances resident when a served window opens: --served-initial, or the
    Poisson steady state of the arrival process (20/s x 30 s mean lifetime)."""
    steady = int(round(20.0 * 30.0))
    return (args.served_initial if args.served_initial >= 0 else steady), steady


def served_leg(tfx, harness, wls3, args, plain_value):
    """The served form in the fast arithmetic: a request window opened at the
    Poisson steady state of resident instances (arrival rate x mean lifetime),
    then its capacity (every instance's frames queued at once, drained) in both
    arithmetics."""
    import numpy as np
    from tressfx_amd import synth
    initial, steady = served_initial(args)
    out = {}
    tfx.set_arithmetic(tfx.TFX_ARITH_FAST)
    try:
        if args.served_seconds > 0:
            srv = tfx.HairServer(hbm_budget_bytes=200 << 30, max_instances=1000, substeps_per_frame=2,
                                 substep_dt=float(np.float32(1.0 / synth.FRAME_RATE / 2)), max_inflight=2)
            run = harness.ServedRun(srv, wls3, initial=initial, duration_s=args.served_seconds, inst_rate=20.0,
                                    lifetime_s=30.0, req_rate=5000.0, ks=(1, 2, 4))
            _log(f"served window {args.served_seconds:.0f} s, {initial} instances resident at start "
                 f"(steady state {steady})")
            res = run.run(log=_log)
            srv.close()
            sv = res["vertex_substeps"] / res["elapsed_s"]
            out.update({
                "workload": "config4: instances drawn from the 256 config-3 instances, Poisson arrivals 20/s, "
                            "lifetimes Exp(30 s), step requests Poisson 5000/s of k in {1,2,4} frames on random "
                            f"resident instances, 200 GB admission budget, {initial} resident when the window opens "
                            f"(the steady state 20/s x 30 s = {steady}), {args.served_seconds:.0f} s window, "
                            "fast arithmetic",
                "value": sv, "unit": "vertex-substeps/s",
                "latency_ms": {"p50": res["latency_p50_ms"], "p99": res["latency_p99_ms"],
                               "mean": res["latency_mean_ms"], "max": res["latency_max_ms"]},
                "requests_completed": res["completed_requests"], "frames_completed": res["completed_frames"],
                "offered_request_rate": res["offered_request_rate"], "passes": res["passes"],
                "mean_instances_per_pass": res["mean_instances_per_pass"], "window_s": res["elapsed_s"],
                "instances_retired": res["instances_retired"], "admissions_refused": res["admissions_refused"],
                "resident_at_start": initial, "resident_at_end": res["resident_at_end"], "steady_state": steady,
                "resident_bytes_at_end": res["resident_bytes"],
                "speedup_vs_plain": sv / plain_value["fast"] if plain_value.get("fast") else None,
                "note": "throughput at offered load: the request process, not the GPU, sets it",
            })
        if args.served_burst_frames > 0:
            srv = tfx.HairServer(hbm_budget_bytes=200 << 30, max_instances=1000, substeps_per_frame=2,
                                 substep_dt=float(np.float32(1.0 / synth.FRAME_RATE / 2)), max_inflight=2)
            brun = harness.ServedRun(srv, wls3, initial=0, duration_s=0.0, inst_rate=0.0, req_rate=0.0)
            _log(f"served capacity: {len(wls3)} instances x {args.served_burst_frames} frames queued at once")
            cap = {}
            for arith, mode in (("fast", tfx.TFX_ARITH_FAST), ("ieee", tfx.TFX_ARITH_IEEE)):
                tfx.set_arithmetic(mode)
                r = brun.burst(args.served_burst_frames)
                v = r["vertex_substeps"] / r["window_s"]
                cap[arith] = {"value": v, "window_s": r["window_s"], "passes": r["passes"],
                              "mean_instances_per_pass": r["mean_instances_per_pass"],
                              "speedup_vs_plain": v / plain_value[arith] if plain_value.get(arith) else None}
            srv.close()
            out["capacity"] = {
                "workload": f"the {len(wls3)} config-3 instances resident, {args.served_burst_frames} frames of each "
                            "submitted at once and drained (requests outrun the GPU: every pass coalesces one "
                            "frame of every instance); speedup against the plain dispatch in the same arithmetic",
                "unit": "vertex-substeps/s", **cap}
    finally:
        tfx.set_arithmetic(tfx.TFX_ARITH_IEEE)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40, help="timed frames of the headline (config-3) workload")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--plain-steps", type=int, default=3, help="timed frames of the plain per-instance dispatch")
    ap.add_argument("--no-ieee", action="store_true", help="skip the correctly rounded repeat of the headline")
    ap.add_argument("--no-config2", action="store_true")
    ap.add_argument("--config2-steps", type=int, default=200)
    ap.add_argument("--no-autotune", action="store_true",
                    help="batched legs keep the occupancy heuristic's launch shapes instead of the autotuner's")
    ap.add_argument("--served-seconds", type=float, default=12.0,
                    help="length of the served-form request window (0: skip)")
    ap.add_argument("--served-initial", type=int, default=-1,
                    help="instances resident when the window opens (-1: the Poisson steady state, 600)")
    ap.add_argument("--served-burst-frames", type=int, default=24,
                    help="served capacity: frames queued at once on every config-3 instance (0: skip)")
    ap.add_argument("--fur-steps", type=int, default=10, help="timed frames of the config-5 fur workload")
    ap.add_argument("--fur-plain-steps", type=int, default=3)
    ap.add_argument("--fur-served-seconds", type=float, default=8.0,
                    help="served window with the fur as a heavy tenant (0: skip)")
    ap.add_argument("--no-fur", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-runs", type=int, default=3)
    ap.add_argument("--host-workers", type=int, default=0,
                    help="processes generating the 256-instance workload (0: the host's CPU share, at most 16; "
                         "1: in this process, for runs under a profiler whose preload every worker would load)")
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

    host_workers = args.host_workers if args.host_workers > 0 else max(1, min(16, cpu_threads()[0]))
    # Workloads are generated before the process touches the GPU, the batched
    # one by freshly started worker processes: a profiler that collects
    # counters initialises the GPU before this script runs, and a process
    # forked from one that has is not safe to run.
    _log(f"generating the 256-instance workload with {host_workers} processes")
    wls3 = synth.config3(workers=host_workers, start_method="spawn")
    wl2 = synth.config2()
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
    tune = 0 if args.no_autotune else 2
    W, K = args.warmup, args.steps

    # headline: config 3 in the fast arithmetic
    head = batched_leg(torch, tfx, harness, stream, wls3, "fast", W, K, args.plain_steps, tune=tune, label="config3")
    dom = max(head.get("fused_classes") or [{}], key=lambda c: c.get("ms_per_frame", 0.0))
    traffic = read_traffic(os.path.join(REPO, "profiles", "fused_substep_config3_traffic.json"), dom.get("kernel"))
    ieee = None
    if not args.no_ieee:
        ieee = batched_leg(torch, tfx, harness, stream, wls3, "ieee", W, K, args.plain_steps, tune=tune,
                           label="config3")
    n3 = sum(w.asset.num_total_vertices for w in wls3)
    lengths = {v: sum(w.asset.num_guide_strands for w in wls3 if w.asset.num_vertices_per_strand == v)
               for v in (16, 32, 64)}

    config2 = None
    if not args.no_config2:
        config2 = {"workload": f"config2: {wl2.asset.num_guide_strands} guide strands x "
                               f"{wl2.asset.num_vertices_per_strand} vertices, "
                               f"{wl2.asset.num_follow_strands_per_guide} follow strands per guide "
                               f"({wl2.asset.num_total_vertices} vertices), 128^3 SDF at 2.5 mm rebuilt per frame "
                               f"from a {wl2.mesh.num_triangles}-triangle noisy icosphere, wind 3 m/s +-20% gust, "
                               "2 substeps/frame, length iterations 2; the batched form with one member"}
        for arith in ("fast", "ieee"):
            leg = batched_leg(torch, tfx, harness, stream, [wl2], arith, 10, args.config2_steps, args.plain_steps * 4,
                              tune=tune, label="config2")
            cls2 = leg.get("fused_classes") or [{}]
            leg["roofline"] = roofline_of(leg, read_traffic(
                os.path.join(REPO, "profiles", "fused_substep_config2_traffic.json"), cls2[0].get("kernel")))
            config2[arith] = leg

    plain_value = {"fast": head.get("plain", {}).get("value"),
                   "ieee": (ieee or {}).get("plain", {}).get("value")}
    served = served_leg(tfx, harness, wls3, args, plain_value) if (args.served_seconds > 0 or
                                                                   args.served_burst_frames > 0) else None
    if served and "value" in served:
        served["speedup_vs_plain"] = served["value"] / plain_value["fast"] if plain_value["fast"] else None

    fur = None
    if wl5 is not None:
        tfx.set_arithmetic(tfx.TFX_ARITH_FAST)
        try:
            fur = fur_leg(torch, tfx, harness, stream, wl5, wls3, args)
            fur["arithmetic"] = "fast"
        finally:
            tfx.set_arithmetic(tfx.TFX_ARITH_IEEE)

    cpu = None if args.no_cpu else cpu_leg(wls3, wl2, runs=args.cpu_runs)

    out = {
        "metric": "vertex-substeps/s",
        "value": head["value"],
        "unit": "vertex-substeps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"config3: {len(wls3)} hair instances per launch, {sum(lengths.values())} guide strands "
                        f"(log-uniform 500..50000 per instance; by vertices per strand {lengths}), {n3} vertices, "
                        "per-instance rigid yaw/pitch trajectories and 64^3..128^3 SDFs rebuilt per frame from "
                        "5120-triangle noisy icospheres, wind, 2 substeps/frame; step = one frame",
            "path": "batched form: per frame one hipGraph replay = parameter upload, the SDFs' build (band mark, "
                    "initialise, construct, finalise, pack: one launch per phase for all 256 grids), per substep "
                    "one persistent fused launch per strand length with the SDF collision inside; launch shapes "
                    "from the autotuner",
            "arithmetic": "fast: hardware reciprocal and square root, fused multiply-adds (held to 1e-5 relative "
                          "per substep against the FP64 transcription at full size by the GPU tests); "
                          "bit_exact_arithmetic repeats the workload bit-exact with the FP32 transcription",
        },
        "latency_ms": head["latency_ms"],
        "roofline": roofline_of(head, traffic),
        "cpu_baseline": cpu,
        "speedup_vs_plain": head.get("speedup_vs_plain"),
        "plain": head.get("plain"),
        "graph_equals_direct": head.get("graph_equals_direct"),
        "fused_ms_per_frame": head.get("fused_ms_per_frame"),
        "sdf_build_ms_per_frame": head.get("sdf_build_ms_per_frame"),
        "autotune": head.get("autotune"),
        "bit_exact_arithmetic": ieee and {k: ieee.get(k) for k in (
            "value", "ms_per_step", "latency_ms", "speedup_vs_plain", "plain", "graph_equals_direct",
            "fused_ms_per_frame", "sdf_build_ms_per_frame", "fused_classes", "frame_hbm_frac")},
        "config2": config2,
        "served": served,
        "fur": fur,
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
