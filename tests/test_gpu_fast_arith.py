"""TFX_ARITH_FAST: the simulation stages, the fused substep and the collision
lookups with the hardware reciprocal and square root, against the FP64 oracle
stepwise.  The bound is the one the default arithmetic meets: relative 1e-5
per substep at every vertex.  The
SDF build keeps correctly rounded arithmetic (its grid is bit-exact in both
modes), and switching back to TFX_ARITH_IEEE restores bit-exactness with the
FP32 transcription, also for a batch whose frame graph was captured in the
fast mode.
"""
import numpy as np
import pytest

import oracle
import tressfx_amd as tfx
from tressfx_amd import synth

from tfx_testutil import check_fp64, first_mismatch, fp64_reference, set_substep_inputs
from test_gpu_parity import _mesh_setup, make_hair, small_workload

pytestmark = pytest.mark.gpu



@pytest.fixture
def fast(gpu):
    tfx.set_arithmetic(tfx.TFX_ARITH_FAST)
    assert tfx.get_arithmetic() == tfx.TFX_ARITH_FAST
    try:
        yield gpu
    finally:
        tfx.set_arithmetic(tfx.TFX_ARITH_IEEE)


def test_arithmetic_mode_errors(gpu):
    with pytest.raises(tfx.TFXError):
        tfx.set_arithmetic(2)
    assert tfx.get_arithmetic() == tfx.TFX_ARITH_IEEE


@pytest.mark.parametrize("V,F,wind", [(16, 0, True), (32, 2, True), (64, 1, False), (8, 4, True)])
def test_fast_substeps_vs_fp64(fast, V, F, wind):
    """Simulate() (reference dispatch structure) in the fast mode, stepwise vs
    FP64; the fast kernels do run (results differ from the FP32 transcription)."""
    print(f"[fast substeps] V={V} F={F} wind={wind}", flush=True)
    wl = small_workload(128, V, F, wind=wind)
    hair = make_hair(wl)
    sim = tfx.TressFXSimulation()
    o32 = oracle.OracleHair(wl.asset, np.float32)
    worst, differs = 0.0, 0
    for frame in range(8):
        for sub in range(wl.substeps):
            set_substep_inputs(hair, wl, frame, sub)
            params, bones = hair.simulation_params()
            st0 = hair.download_state()
            o32.set_state(st0)
            ref, sens = fp64_reference(wl.asset, st0, lambda h: h.simulate(params, bones), seed=frame)
            sim.Simulate([hair])
            o32.simulate(params, bones)
            st = hair.download_state()
            differs += int((st["pos"] != o32.pos).any(axis=1).sum())
            worst = max(worst, check_fp64(st, ref, sens, f"frame {frame} sub {sub}", V))
    print(f"worst relative error vs FP64: {worst:.3e} (bound 1e-5 at every vertex); "
          f"vertex-substeps differing from the FP32 transcription: {differs}")
    assert differs > 0, "the fast kernels did not run"


def test_fast_collide_and_fused_vs_fp64(fast):
    """Simulate + CollideWithHair and the fused substep in the fast mode, with
    SDF, wind and follow strands, stepwise vs FP64; the SDF grid stays
    bit-exact with the FP32 transcription."""
    torch = fast
    wl, dm = _mesh_setup(torch)
    origin, cell, dims = wl.sdf_grid
    system = tfx.TressFXSDFCollisionSystem()
    sim = tfx.TressFXSimulation()
    hair_a = tfx.TressFXHairObject(wl.asset)
    hair_b = tfx.TressFXHairObject(wl.asset)
    desc = dm.update(0.0)
    sdf = tfx.TressFXSDFCollision(desc, collision_margin=wl.collision_margin, explicit=(origin, cell, dims))
    V = wl.asset.num_vertices_per_strand
    hits = 0
    for frame in range(6):
        desc = dm.update(frame / synth.FRAME_RATE)
        sdf.SetMesh(desc)
        sdf.Update(system)
        grid32 = sdf.download()
        ref32 = oracle.sdf_build(dm.verts_h, dm.mesh.indices, origin, np.float32(cell), dims, np.float32)
        mm = first_mismatch(grid32, ref32)
        assert mm is None, f"frame {frame}: SDF differs from the FP32 transcription at {mm}"
        for sub in range(wl.substeps):
            for h in (hair_a, hair_b):
                set_substep_inputs(h, wl, frame, sub)
            params, bones = hair_a.simulation_params()
            st0 = hair_a.download_state()
            hair_b.upload_state(st0)
            cp = sdf.collision_params(hair_a)
            g64 = grid32.astype(np.float64)

            def advance64(h64):
                h64.simulate(params, bones)
                oracle.sdf_collide(h64.pos, h64.prev, V, g64, np.array(cp.origin[:3], np.float64),
                                   float(cp.cell_size), dims, float(cp.collision_margin))

            ref, sens = fp64_reference(wl.asset, st0, advance64, seed=frame)
            before = ref.pos.copy()
            sim.Simulate([hair_a])
            sdf.CollideWithHair(system, hair_a)
            sim.SimulateFused([hair_b], [sdf])
            check_fp64(hair_a.download_state(), ref, sens, f"plain frame {frame} sub {sub}", V)
            check_fp64(hair_b.download_state(), ref, sens, f"fused frame {frame} sub {sub}", V)
            h64 = oracle.OracleHair(wl.asset, np.float64)
            h64.set_state(st0)
            h64.simulate(params, bones)
            hits += int((h64.pos != before).any(axis=1).sum())
    assert hits > 0, "the SDF never moved a vertex"


def test_fast_batch_graph_then_ieee(fast):
    """A batch replays its frame graph in the fast mode (each frame vs the
    FP64 oracle over its two substeps), then in the default mode (the graph is
    re-captured) bit-exact with the per-object dispatch of the FP32
    transcription's contract."""
    torch = fast
    wl, dm = _mesh_setup(torch, G=192, V=32, F=1)
    origin, cell, dims = wl.sdf_grid
    system = tfx.TressFXSDFCollisionSystem()
    desc = dm.update(0.0)
    hair = tfx.TressFXHairObject(wl.asset)
    sdf = tfx.TressFXSDFCollision(desc, collision_margin=wl.collision_margin, explicit=(origin, cell, dims))
    ref_hair = tfx.TressFXHairObject(wl.asset)
    ref_sdf = tfx.TressFXSDFCollision(desc, collision_margin=wl.collision_margin, explicit=(origin, cell, dims))
    sim = tfx.TressFXSimulation()
    batch = tfx.HairBatch()
    batch.add(hair, sdf)
    side = torch.cuda.Stream()
    V = wl.asset.num_vertices_per_strand
    dt = np.float32(1.0 / synth.FRAME_RATE / wl.substeps)
    for frame in range(6):
        mode = tfx.TFX_ARITH_FAST if frame < 3 else tfx.TFX_ARITH_IEEE
        tfx.set_arithmetic(mode)
        desc = dm.update(frame / synth.FRAME_RATE)
        sdf.SetMesh(desc)
        ref_sdf.SetMesh(desc)
        torch.cuda.synchronize()
        st0 = hair.download_state()
        ref_hair.upload_state(st0)
        ins = []
        for sub in range(wl.substeps):
            t = (frame + (sub + 1) / wl.substeps) / synth.FRAME_RATE
            bones = np.asarray(wl.bone_fn(t), np.float32).reshape(1, 16)
            st = tfx.SimulationSettings.from_buffer_copy(wl.settings)
            st.wind_magnitude = wl.wind_fn(t)
            ins.append((bones, st))
        batch.begin_frame()
        for bones, st in ins:
            batch.stage_substep(bones, st, dt)
        batch.run_frame(side.cuda_stream, use_graph=True)
        torch.cuda.synchronize()
        got = hair.download_state()
        # the same frame through the per-object dispatch, in the same mode
        ref_sdf.Update(system)
        steps = []
        for bones, st in ins:
            ref_hair.UpdateBoneMatrices(bones.reshape(-1, 16))
            ref_hair.UpdateSimulationParameters(st, dt)
            steps.append(ref_hair.simulation_params())
            sim.Simulate([ref_hair])
            ref_sdf.CollideWithHair(system, ref_hair)
        want = ref_hair.download_state()
        if mode == tfx.TFX_ARITH_IEEE:
            for k in ("pos", "prev", "tangent"):
                mm = first_mismatch(got[k], want[k])
                assert mm is None, f"frame {frame} (default mode after fast frames) {k}: {mm}"
        else:
            g64 = ref_sdf.download().astype(np.float64)
            cp = sdf.collision_params(hair)

            def advance64(h64):
                for params, bones in steps:
                    h64.simulate(params, bones)
                    oracle.sdf_collide(h64.pos, h64.prev, V, g64, np.array(cp.origin[:3], np.float64),
                                       float(cp.cell_size), dims, float(cp.collision_margin))

            ref, sens = fp64_reference(wl.asset, st0, advance64, seed=frame)
            check_fp64(got, ref, sens, f"batched frame {frame} (fast)", V)
