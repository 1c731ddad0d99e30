"""Config 5, fur stress, at full size: 1,000,000 guide strands x 8 vertices
with 2 follow strands each (24 M vertices) rooted on a 200,000-triangle noisy
ellipsoid that 4 bones skin every frame, a 256^3 SDF rebuilt from the skinned
mesh every frame, 3 substeps, length iterations 2, VSP 0.7.

- Frame by frame from the GPU's own state: the skinned mesh
  (TressFXBoneSkinning) and the SDF built from it bit-exact against the FP32
  transcription, then every substep of the reference dispatch (Simulate +
  CollideWithHair) bit-exact against it, with the collision actually hitting.
- The fused substep and the graph-replayed batched frame equal the reference
  dispatch bit for bit over the same frames.
"""
import numpy as np
import pytest

import oracle
from tressfx_amd import harness, synth

from tfx_testutil import first_mismatch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fur():
    print("[config5] generating 1,000,000 x 8 fur on a 200,000-triangle body", flush=True)
    return synth.config5()


def _same(a, b, keys, what):
    for k in keys:
        mm = first_mismatch(a[k], b[k] if isinstance(b, dict) else getattr(b, k))
        assert mm is None, f"{what}: {k} differs at {mm[0]} gpu={mm[1]!r} ref={mm[2]!r} ({mm[3]} words)"


def test_config5_full_reference_dispatch_vs_oracle(gpu, fur):
    torch = gpu
    wl = fur
    a = wl.asset
    assert (a.num_guide_strands, a.num_vertices_per_strand, a.num_follow_strands_per_guide) == (1000000, 8, 2)
    assert wl.mesh.num_triangles == 200000 and wl.substeps == 3
    frames = [0, 1]
    print("[config5] reference dispatch vs FP32 transcription, frames 0-1", flush=True)
    drv = harness.SkinnedFrameDriver(torch, wl, frames, mode="plain")
    body = wl.mesh
    rest = np.zeros((body.positions.shape[0], 8), np.float32)
    rest[:, :3], rest[:, 4:7] = body.positions, body.normals
    origin, cell, dims = wl.sdf_grid
    origin32 = np.asarray(origin, np.float32)
    o32 = oracle.OracleHair(a, np.float32)
    V = a.num_vertices_per_strand
    hits = 0
    for k, f in enumerate(frames):
        drv.skin.Update(drv.mb[k])
        drv.sdf.SetMesh(drv.skin.mesh())
        drv.sdf.Update(drv.system)
        verts = oracle.mesh_skin(rest, body.skinning, drv.mb[k])
        mm = first_mismatch(drv.skin.download(), verts)
        assert mm is None, f"frame {f}: skinned body differs at {mm}"
        cells = drv.sdf.download()
        ref = oracle.sdf_build(verts, body.indices, origin32, np.float32(cell), dims, np.float32)
        mm = first_mismatch(cells, ref)
        assert mm is None, f"frame {f}: SDF differs at {mm}"
        for s in range(wl.substeps):
            drv.hair.UpdateBoneMatrices(drv.hb[k, s].reshape(-1, 16))
            drv.hair.UpdateSimulationParameters(wl.settings, drv.dt_sub)
            params, bones = drv.hair.simulation_params()
            o32.set_state(drv.hair.download_state())
            drv.sim.Simulate([drv.hair])
            drv.sdf.CollideWithHair(drv.system, drv.hair)
            o32.simulate(params, bones)
            cp = drv.sdf.collision_params(drv.hair)
            before = o32.pos.copy()
            oracle.sdf_collide(o32.pos, o32.prev, V, cells, np.array(cp.origin[:3], np.float32),
                               np.float32(cp.cell_size), (cp.num_cells_x, cp.num_cells_y, cp.num_cells_z),
                               np.float32(cp.collision_margin))
            hits += int((before != o32.pos).any(axis=1).sum())
            _same(drv.hair.download_state(), o32, ("pos", "prev", "prevprev", "tangent"), f"frame {f} substep {s}")
    print(f"collision hits over {len(frames) * wl.substeps} substeps: {hits}")
    assert hits > 1000, "the fur never touched the body"


def test_config5_full_fused_and_graph_equal_reference_dispatch(gpu, fur):
    torch = gpu
    wl = fur
    frames = list(range(4))
    states = {}
    stream = torch.cuda.Stream()    # graph replay needs a non-default stream
    for mode in ("plain", "fused", "graph"):
        print(f"[config5] {mode}, frames 0-3", flush=True)
        drv = harness.SkinnedFrameDriver(torch, wl, frames, mode=mode)
        with torch.cuda.stream(stream):
            for f in frames:
                drv.frame(f, stream.cuda_stream)
        stream.synchronize()
        states[mode] = drv.hair.download_state()
        del drv
    for mode in ("fused", "graph"):
        _same(states[mode], states["plain"], ("pos", "prev", "tangent"), f"{mode} vs reference dispatch")
    moved = np.abs(states["plain"]["pos"][:, :3] - wl.asset.positions[:, :3]).max()
    assert np.isfinite(states["plain"]["pos"]).all() and moved > 1e-4
