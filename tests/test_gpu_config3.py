"""The batched workload at its full size: 256 instances with strand counts
log-uniform in [500, 50,000] and 16, 32 or 64 vertices per strand (2.92 M
strands, 76.4 M vertices), each with its own rigid trajectory, 64^3-128^3 SDF
rebuilt per frame from a noisy icosphere, wind, 2 substeps.

- Default arithmetic: the batch, every frame one replayed hipGraph, with the
  automatic collision placement (inside the fused launches here) and with
  every class's collision deferred to a collide pass (TFX_SDF_DEFER_ALWAYS,
  the SDF build beside the first substep), against the reference's
  per-object dispatch (SDF Update, then per substep Simulate over all objects
  and CollideWithHair per object) -- every instance's positions, previous
  positions and tangents and every SDF grid bit for bit after every frame.  The per-object dispatch of one instance per strand
  length and of the largest instance is checked substep by substep against
  the FP32 transcription (bit for bit, SDF grid included) and the FP64 one
  (1e-5 relative at every vertex), from the GPU's own state.
- TFX_ARITH_FAST: the batch runs one substep per replayed frame graph (the
  same inputs; the grid is rebuilt from the same mesh, so the substeps are
  those of a two-substep frame), and the same instances are checked substep by
  substep against the FP64 transcription at 1e-5 relative at every vertex.

Each check stops at its first mismatch.
"""
import numpy as np
import pytest

import oracle
import tressfx_amd as tfx
from tressfx_amd import harness, synth

from tfx_testutil import check_fp64, first_mismatch, fp64_reference

pytestmark = pytest.mark.gpu

FRAMES = list(range(4))
CHECK_FRAMES = (2, 3)       # frames 0 and 1 reset to the skinned rest pose


@pytest.fixture
def workloads(config3_workloads):
    return config3_workloads


def _checked_instances(wls):
    """The smallest instance of each strand length and the largest overall."""
    pick = set()
    for V in (16, 32, 64):
        ids = [i for i, w in enumerate(wls) if w.asset.num_vertices_per_strand == V]
        pick.add(min(ids, key=lambda i: wls[i].asset.num_guide_strands))
    pick.add(max(range(len(wls)), key=lambda i: wls[i].asset.num_total_vertices))
    return sorted(pick)


def _oracle_substep(wl, hair, sdf, params, bones, grid32, st0, dtype):
    cp = sdf.collision_params(hair)
    dims = (cp.num_cells_x, cp.num_cells_y, cp.num_cells_z)
    V = wl.asset.num_vertices_per_strand

    def advance(h):
        h.simulate(params, bones)
        g = grid32 if dtype == np.float32 else grid32.astype(np.float64)
        oracle.sdf_collide(h.pos, h.prev, V, g, np.array(cp.origin[:3], dtype), dtype(cp.cell_size), dims,
                           dtype(cp.collision_margin))
    return advance


@pytest.mark.timeout(900)
def test_config3_full_size_graph_batch_equals_per_object_and_oracles(gpu, workloads):
    torch = gpu
    wls = workloads
    checked = _checked_instances(wls)
    print(f"[config3] batch (graph) vs per-object dispatch, {len(FRAMES)} frames; oracle instances "
          f"{[(i, wls[i].asset.num_guide_strands, wls[i].asset.num_vertices_per_strand) for i in checked]}",
          flush=True)
    bdrv = harness.BatchDriver(torch, wls, FRAMES)
    # a second batch with every class's collision deferred to a collide pass
    # (and the SDF build beside the first substep's launches)
    ddrv = harness.BatchDriver(torch, wls, FRAMES, meshes=bdrv.meshes)
    ddrv.batch.set_sdf_deferral(tfx.TFX_SDF_DEFER_ALWAYS)
    pdrv = harness.PlainMultiDriver(torch, wls, FRAMES, meshes=bdrv.meshes)
    stream = torch.cuda.Stream()
    worst = 0.0
    for f in FRAMES:
        bdrv.frame(f, stream.cuda_stream, use_graph=True)
        ddrv.frame(f, stream.cuda_stream, use_graph=True)
        # the per-object dispatch, with the checked instances stepped against the oracles
        k = f - pdrv.first
        for sdf, d in zip(pdrv.sdfs, pdrv.descs[k]):
            sdf.SetMesh(d)
            sdf.Update(pdrv.system, stream.cuda_stream)
        check = f in CHECK_FRAMES
        grids = {}
        if check:
            torch.cuda.synchronize()
            for i in checked:
                wl = wls[i]
                origin, cell, dims = wl.sdf_grid
                t = f / synth.FRAME_RATE
                grids[i] = pdrv.sdfs[i].download()
                want = oracle.sdf_build(wl.mesh.vertices(t, wl.bone_fn(t)), wl.mesh.indices,
                                        np.asarray(origin, np.float32), np.float32(cell), dims, np.float32)
                assert np.array_equal(grids[i].view(np.uint32), want.view(np.uint32)), f"frame {f} instance {i}: SDF"
        for s in range(pdrv.substeps):
            for hair, wl, (bones, wind) in zip(pdrv.hairs, pdrv.wls, pdrv.inputs[(f, s)]):
                hair.UpdateBoneMatrices(bones)
                if wind is not None:
                    wl.settings.wind_magnitude = wind
                hair.UpdateSimulationParameters(wl.settings, pdrv.dt_sub)
            before = {}
            if check:
                torch.cuda.synchronize()
                before = {i: (pdrv.hairs[i].download_state(), pdrv.hairs[i].simulation_params()) for i in checked}
            pdrv.sim.Simulate(pdrv.hairs, stream.cuda_stream)
            for hair, sdf in zip(pdrv.hairs, pdrv.sdfs):
                sdf.CollideWithHair(pdrv.system, hair, stream.cuda_stream)
            if check:
                torch.cuda.synchronize()
                for i in checked:
                    wl, hair = wls[i], pdrv.hairs[i]
                    st0, (params, bones) = before[i]
                    got = hair.download_state()
                    o32 = oracle.OracleHair(wl.asset, np.float32)
                    o32.set_state(st0)
                    _oracle_substep(wl, hair, pdrv.sdfs[i], params, bones, grids[i], st0, np.float32)(o32)
                    for key in ("pos", "prev", "prevprev", "tangent"):
                        mm = first_mismatch(got[key], getattr(o32, key))
                        assert mm is None, f"frame {f} sub {s} instance {i}: {key} vs FP32 transcription {mm}"
                    ref, sens = fp64_reference(wl.asset, st0, _oracle_substep(wl, hair, pdrv.sdfs[i], params, bones,
                                                                            grids[i], st0, np.float64), seed=f)
                    worst = max(worst, check_fp64(got, ref, sens, f"frame {f} sub {s} instance {i}",
                                                  wl.asset.num_vertices_per_strand))
        torch.cuda.synchronize()
        keys = ("pos", "prev", "tangent")
        for i, (hb, hd, hp) in enumerate(zip(bdrv.hairs, ddrv.hairs, pdrv.hairs)):
            b = hp.download_state(keys=keys)
            for name, h in (("batch", hb), ("deferred batch", hd)):
                a = h.download_state(keys=keys)
                for key in keys:
                    mm = first_mismatch(a[key], b[key])
                    assert mm is None, (f"frame {f} {name} instance {i} (G={wls[i].asset.num_guide_strands}): "
                                        f"{key} {mm}")
        for i, (sb, sd, sp) in enumerate(zip(bdrv.sdfs, ddrv.sdfs, pdrv.sdfs)):
            want = sp.download().view(np.uint32)
            assert np.array_equal(sb.download().view(np.uint32), want), f"frame {f} SDF {i}"
            assert np.array_equal(sd.download().view(np.uint32), want), f"frame {f} deferred batch SDF {i}"
        print(f"[config3] frame {f}: 256 instances bit-identical (batch, deferred batch, per-object)", flush=True)
    print(f"[config3] worst relative error vs FP64: {worst:.3e} (bound 1e-5 at every vertex)")
    moved = np.abs(pdrv.hairs[checked[-1]].download_state()["pos"][:, :3]
                   - wls[checked[-1]].asset.positions[:, :3]).max()
    assert moved > 1e-3, "the hair never moved"


@pytest.mark.timeout(900)
def test_config3_full_size_fast_arithmetic_vs_fp64(gpu, workloads):
    torch = gpu
    wls = workloads
    checked = _checked_instances(wls)
    print(f"[config3 fast] batch one substep per graph replay vs FP64, instances {checked}", flush=True)
    bdrv = harness.BatchDriver(torch, wls, FRAMES)
    stream = torch.cuda.Stream()
    b = bdrv.batch
    worst, differs = 0.0, 0
    tfx.set_arithmetic(tfx.TFX_ARITH_FAST)
    try:
        for f in FRAMES:
            k = f - bdrv.first
            for s in range(bdrv.substeps):
                for sdf, d in zip(bdrv.sdfs, bdrv.descs[k]):
                    sdf.SetMesh(d)
                check = f in CHECK_FRAMES
                if check:
                    torch.cuda.synchronize()
                    before = {i: bdrv.hairs[i].download_state() for i in checked}
                b.begin_frame()
                bones, st = bdrv.inputs[(f, s)]
                b.stage_substep(bones, st, bdrv.dt_sub)
                staged = {}
                if check:
                    for i in checked:
                        p, bn = bdrv.hairs[i].simulation_params()
                        staged[i] = (p, bn)
                b.run_frame(stream.cuda_stream, use_graph=True)
                if not check:
                    continue
                torch.cuda.synchronize()
                for i in checked:
                    wl, hair, sdf = wls[i], bdrv.hairs[i], bdrv.sdfs[i]
                    params, bn = staged[i]
                    grid = sdf.download()
                    got = hair.download_state()
                    adv32 = _oracle_substep(wl, hair, sdf, params, bn, grid, before[i], np.float32)
                    o32 = oracle.OracleHair(wl.asset, np.float32)
                    o32.set_state(before[i])
                    adv32(o32)
                    differs += int((got["pos"] != o32.pos).any(axis=1).sum())
                    ref, sens = fp64_reference(wl.asset, before[i],
                                               _oracle_substep(wl, hair, sdf, params, bn, grid, before[i],
                                                               np.float64), seed=f)
                    worst = max(worst, check_fp64(got, ref, sens, f"fast frame {f} sub {s} instance {i}",
                                                  wl.asset.num_vertices_per_strand))
            print(f"[config3 fast] frame {f} done", flush=True)
    finally:
        tfx.set_arithmetic(tfx.TFX_ARITH_IEEE)
    print(f"[config3 fast] worst relative error vs FP64: {worst:.3e} (bound 1e-5 at every vertex); "
          f"vertices differing from the FP32 transcription: {differs}")
    assert differs > 0, "the fast kernels did not run"
