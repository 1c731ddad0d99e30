"""tfx_simulation_simulate_fused_noalloc (the fused substep whose tables live
in a static device array, for callers that cannot allocate) against
tfx_simulation_simulate_fused, bit for bit, through the C-ABI: two calls in a
row, the first with three objects (16, 32 and 64 vertices per strand, follow
strands, SDF collision, a multi-bone skinning record), the second with one
smaller object, each over a few frames; and its size limits."""
import numpy as np
import pytest

import tressfx_amd as tfx
from tressfx_amd import synth

from tfx_testutil import first_mismatch, set_substep_inputs
from test_gpu_parity import _mesh_setup, make_hair, small_workload

pytestmark = pytest.mark.gpu


def _run(torch, groups, frames):
    """groups: [(workload, sdf or None)]; every object exists twice, one
    stepped by each entry point; compares after every substep."""
    sim = tfx.TressFXSimulation()
    a = [make_hair(wl) for wl, _ in groups]
    b = [make_hair(wl) for wl, _ in groups]
    sdfs = [sdf for _, sdf in groups]
    for frame in range(frames):
        for sub in range(2):
            for (wl, _), ha, hb in zip(groups, a, b):
                set_substep_inputs(ha, wl, frame, sub)
                set_substep_inputs(hb, wl, frame, sub)
            sim.SimulateFused(a, sdfs)
            sim.SimulateFusedNoAlloc(b, sdfs)
            torch.cuda.synchronize()
            for i, (ha, hb) in enumerate(zip(a, b)):
                sa, sb = ha.download_state(), hb.download_state()
                for k in ("pos", "prev", "tangent"):
                    mm = first_mismatch(sb[k], sa[k])
                    assert mm is None, f"frame {frame} sub {sub} object {i}: {k} differs at {mm}"
                V = ha.info().num_vertices_per_strand
                v01 = np.arange(sa["pos"].shape[0]) % V < 2
                assert np.array_equal(sa["prevprev"][v01].view(np.uint32), sb["prevprev"][v01].view(np.uint32))
                assert ha.info().simulation_frame == hb.info().simulation_frame
    return a, b


def test_noalloc_equals_fused_two_calls(gpu):
    torch = gpu
    print("[noalloc] call 1: three objects with SDFs; call 2: one smaller object", flush=True)
    wl_m, dm = _mesh_setup(torch, G=300, V=32, F=2)
    origin, cell, dims = wl_m.sdf_grid
    system = tfx.TressFXSDFCollisionSystem()
    sdf = tfx.TressFXSDFCollision(dm.update(0.0), collision_margin=wl_m.collision_margin,
                                  explicit=(origin, cell, dims))
    sdf.Update(system)
    wl16 = small_workload(200, 16, 0, seed=31)
    wl64 = small_workload(70, 64, 1, seed=32, capsules=False)
    # a skinning record that reads bones 0..2 with three weights
    skin = np.asarray(wl16.asset.bone_skinning_data).copy()
    skin[:, 0:4] = (0.0, 1.0, 2.0, 0.0)
    skin[:, 4:8] = (0.5, 0.3, 0.2, 0.0)
    wl16.asset.bone_skinning_data[:] = skin
    base_bone = wl16.bone_fn
    wl16.bone_fn = lambda t: np.stack([np.asarray(base_bone(t), np.float32).reshape(16),
                                       synth.yaw_pitch_matrix(0.2 * t, 0.1).reshape(16),
                                       synth.yaw_pitch_matrix(-0.3 * t, 0.05).reshape(16)])
    first = [(wl16, sdf), (wl_m, sdf), (wl64, None)]
    a1, b1 = _run(torch, first, frames=3)
    moved = np.abs(b1[1].download_state()["pos"][:, :3] - wl_m.asset.positions[:, :3]).max()
    assert moved > 1e-4, "the hair never moved"
    # the second call is smaller: one object, fewer strands, other length
    wl8 = small_workload(90, 8, 0, seed=33)
    _run(torch, [(wl8, sdf)], frames=2)


def test_noalloc_limits(gpu):
    wl = small_workload(8, 16, 0, seed=34)
    hairs = [make_hair(wl) for _ in range(65)]
    sim = tfx.TressFXSimulation()
    with pytest.raises(tfx.TFXError) as e:
        sim.SimulateFusedNoAlloc(hairs)
    assert "at most 64" in str(e.value)
    sim.SimulateFusedNoAlloc(hairs[:64])   # at the limit
