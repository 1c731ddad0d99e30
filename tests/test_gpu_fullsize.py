"""Parity at the workloads' full sizes.

- config 1 (1,000 x 16 strands, sphere collider through the capsule path,
  1 substep): all 120 frames stepwise, plain dispatch and fused substep, bit-exact
  vs the FP32 transcription and within 1e-5 relative of FP64 at every vertex;
- config 2 (20,000 x 32 + 4 follow strands per guide, 128^3 SDF rebuilt per
  frame, wind, 2 substeps) through the bench's path -- the batched form with
  the frame replayed as a hipGraph -- frame by frame from the GPU's own state vs
  the FP32 transcription (SDF build, two substeps of simulation + collision);
- 256 heterogeneous instances (the config-3 size distribution, strand counts
  scaled down) through the graph-replayed batch vs the per-object dispatch.
"""
import numpy as np
import pytest

import oracle
import tressfx_amd as tfx
from tressfx_amd import harness, synth

from tfx_testutil import check_fp64, first_mismatch, fp64_reference

pytestmark = pytest.mark.gpu


def _compare(got, ref, keys, what):
    for k in keys:
        mm = first_mismatch(got[k], getattr(ref, k) if not isinstance(ref, dict) else ref[k])
        assert mm is None, f"{what}: {k} differs at {mm[0]} gpu={mm[1]!r} ref={mm[2]!r} ({mm[3]} words)"


@pytest.mark.parametrize("fused", [False, True])
def test_config1_full_120_frames_stepwise(gpu, fused):
    print(f"[config1] 1000x16, 120 frames, fused={fused}", flush=True)
    wl = synth.config1()
    hair = tfx.TressFXHairObject(wl.asset)
    hair.SetCollisionCapsules(*wl.capsules)
    sim = tfx.TressFXSimulation()
    o32 = oracle.OracleHair(wl.asset, np.float32)
    dt = np.float32(1.0 / synth.FRAME_RATE)
    worst = 0.0
    keys = ("pos", "prev", "tangent") if fused else ("pos", "prev", "prevprev", "tangent")
    for frame in range(120):
        t = (frame + 1) / synth.FRAME_RATE
        hair.UpdateBoneMatrices(wl.bone_fn(t))
        hair.UpdateSimulationParameters(wl.settings, dt)
        params, bones = hair.simulation_params()
        st0 = hair.download_state()
        o32.set_state(st0)
        check64 = frame % 10 == 9
        if check64:
            ref, sens = fp64_reference(wl.asset, st0, lambda h: h.simulate(params, bones), seed=frame)
        if fused:
            sim.SimulateFused([hair])
        else:
            sim.Simulate([hair])
        o32.simulate(params, bones)
        st = hair.download_state()
        _compare(st, o32, keys, f"frame {frame}")
        if check64:
            worst = max(worst, check_fp64(st, ref, sens, f"frame {frame}", 16))
    print(f"worst relative error vs FP64: {worst:.3e} (bound 1e-5 at every vertex)")
    moved = np.abs(st["pos"][:, :3] - wl.asset.positions[:, :3]).max()
    assert moved > 1e-3, "the hair never moved"


def test_config2_full_graph_frames_vs_oracle(gpu):
    torch = gpu
    print("[config2] 20000x32 F=4, graph-replayed frames vs FP32 transcription", flush=True)
    wl = synth.config2()
    a = wl.asset
    V = a.num_vertices_per_strand
    frames = list(range(48))
    check = {0, 1, 2, 3, 20, 46, 47}
    drv = harness.BatchDriver(torch, [wl], frames)
    hair, sdf, b = drv.hairs[0], drv.sdfs[0], drv.batch
    stream = torch.cuda.Stream()
    origin, cell, dims = wl.sdf_grid
    origin32 = np.asarray(origin, np.float32)
    margin32 = np.float32(np.float32(wl.collision_margin) * np.float32(cell))
    o32 = oracle.OracleHair(a, np.float32)
    hits = 0
    for f in frames:
        st0 = hair.download_state() if f in check else None
        sdf.SetMesh(drv.descs[f][0])
        b.begin_frame()
        staged = []
        for s in range(drv.substeps):
            counter = hair.info().simulation_frame
            bones, st = drv.inputs[(f, s)]
            b.stage_substep(bones, st, drv.dt_sub)
            p, bn = hair.simulation_params()
            if counter < 2:
                p.reset_positions = 1.0   # the flag IncreaseSimulationFrame cleared after staging
            staged.append((p, bn))
        b.run_frame(stream.cuda_stream, use_graph=True)
        torch.cuda.synchronize()
        if f not in check:
            continue
        t = f / synth.FRAME_RATE
        verts = wl.mesh.vertices(t, wl.bone_fn(t))
        grid = oracle.sdf_build(verts, wl.mesh.indices, origin32, np.float32(cell), dims, np.float32)
        o32.set_state(st0)
        for p, bn in staged:
            o32.simulate(p, bn)
            before = o32.pos.copy()
            oracle.sdf_collide(o32.pos, o32.prev, V, grid, origin32, np.float32(cell), dims, margin32)
            hits += int((before != o32.pos).any(axis=1).sum())
        got = hair.download_state()
        _compare(got, o32, ("pos", "prev", "tangent"), f"frame {f}")
        assert np.array_equal(sdf.download().view(np.uint32), grid.view(np.uint32)), f"frame {f}: SDF differs"
    assert hits > 1000, "too few SDF collisions to exercise the collision path"


def test_256_instances_graph_equals_per_object_dispatch(gpu):
    torch = gpu
    print("[256 instances] graph-replayed batch vs per-object dispatch", flush=True)
    spec = synth.config3_spec(256, seed=7, min_strands=100, max_strands=4000)
    wls = [synth.config3_instance(i, G, V, n, seed=7, subdivisions=3) for i, (G, V, n) in enumerate(spec)]
    frames = list(range(4))
    bdrv = harness.BatchDriver(torch, wls, frames)
    pdrv = harness.PlainMultiDriver(torch, wls, frames, meshes=bdrv.meshes)
    stream = torch.cuda.Stream()
    for f in frames:
        bdrv.frame(f, stream.cuda_stream, use_graph=True)
        pdrv.frame(f, stream.cuda_stream)
    torch.cuda.synchronize()
    for i, (hb, hp) in enumerate(zip(bdrv.hairs, pdrv.hairs)):
        _compare(hb.download_state(), hp.download_state(), ("pos", "prev", "tangent"), f"instance {i} {spec[i]}")
    for i, (sb, sp) in enumerate(zip(bdrv.sdfs, pdrv.sdfs)):
        assert np.array_equal(sb.download().view(np.uint32), sp.download().view(np.uint32)), f"SDF {i} differs"
