"""The simulation-to-renderer hand-off (TressFXHairObject.cpp:559-583,
TressFXStrands.hlsl:33-35, 82-89): a barrier from the simulation stream
orders a reader on another stream after every enqueued substep, and the
loaded asset's triangle index buffer addresses the simulated Pos/Tangent
buffers in place -- index / 2 is a vertex, each triangle pair spans one
segment of one strand, and the draw at a LOD density ends on a whole group of
a guide and its follow strands."""
import numpy as np
import pytest

import tressfx_amd as tfx
from tressfx_amd import synth

from test_loaders import TFX, TFXBONE, read_tfxbone
from tfx_testutil import first_mismatch, set_substep_inputs

pytestmark = pytest.mark.gpu


def test_transition_orders_the_reader_after_the_simulation(gpu):
    torch = gpu
    print("[render] stream hand-off", flush=True)
    wl = synth.config2(num_guides=4000)
    hair = tfx.TressFXHairObject(wl.asset)
    sim = tfx.TressFXSimulation()
    a, b = torch.cuda.Stream(), torch.cuda.Stream()
    for frame in range(3):
        for sub in range(2):
            set_substep_inputs(hair, wl, frame, sub)
            sim.Simulate([hair], a.cuda_stream)
    hair.TransitionSimToRendering(a.cuda_stream, b.cuda_stream)
    seen = hair.download_state(b.cuda_stream)       # enqueued on b, after the barrier
    a.synchronize()
    final = hair.download_state()
    for k in ("pos", "tangent"):
        mm = first_mismatch(seen[k], final[k])
        assert mm is None, f"the reader on the other stream saw {k} before the simulation ended: {mm}"
    hair.TransitionRenderingToSim(b.cuda_stream, a.cuda_stream)
    hair.TransitionSimToRendering(a.cuda_stream, a.cuda_stream)   # same stream: nothing to do


def test_render_index_buffer_reads_the_strands_in_place(gpu):
    print("[render] index buffer layout", flush=True)
    names, _ = read_tfxbone(TFXBONE)
    asset = tfx.TressFXAsset()
    asset.LoadHairData(TFX)
    asset.GenerateFollowHairs(2, 0.3, 0.004, seed=9)
    asset.ProcessAsset()
    asset.LoadBoneData(TFXBONE, names)
    ha = asset.hair_asset()
    V, F, S = ha.num_vertices_per_strand, ha.num_follow_strands_per_guide, ha.num_total_strands
    _, _, tri = asset.render_data()
    assert tri.size == 6 * S * (V - 1)
    hair = tfx.TressFXHairObject(ha)
    sim = tfx.TressFXSimulation()
    st = tfx.SimulationSettings.defaults()
    bones = np.tile(np.eye(4, dtype=np.float32).reshape(1, 16), (len(names), 1))
    for _ in range(4):
        hair.UpdateBoneMatrices(bones)
        hair.UpdateSimulationParameters(st, np.float32(1.0 / 120.0))
        sim.Simulate([hair])
    state = hair.download_state()
    vid = (tri // 2).reshape(-1, 6)                          # the vertex each index reads
    seg = vid[:, 0]
    assert np.array_equal(vid, np.stack([seg, seg, seg + 1, seg + 1, seg, seg + 1], 1))
    assert np.all(seg // V == (seg + 1) // V), "a segment crosses a strand boundary"
    assert np.array_equal(np.unique(seg), np.arange(S * V).reshape(S, V)[:, :-1].reshape(-1))
    # the positions the shader reads are the simulated strands: every segment
    # spans two neighbouring vertices of one strand (length near its rest
    # length; a vertex of another strand would be a jump across the head)
    p = state["pos"][:, :3].astype(np.float64)
    length = np.linalg.norm(p[seg + 1] - p[seg], axis=1)
    rest = ha.rest_lengths[seg].astype(np.float64)
    assert np.all(length <= 1.5 * rest + 1e-6)
    assert np.mean(np.abs(length - rest) / rest) < 0.02
    # the draw at density 0.5 ends on a whole guide group
    n = tfx.num_indices_to_render(tri.size, V, F, 0.5)
    strands = n // (6 * (V - 1))
    assert n % (6 * (V - 1)) == 0 and strands % (F + 1) == 0 and 0 < strands < S
