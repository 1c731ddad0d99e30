"""The batched form (tfx_batch_*): objects of mixed strand length, some with
an SDF collider, advanced together -- launched directly and replayed from a
hipGraph -- must equal the per-object reference dispatch (Simulate, then
CollideWithHair per object, src/Simulation.cpp:50-131) bit for bit, and the
per-object path is itself pinned to the oracle by test_gpu_parity."""
import numpy as np
import pytest

import tressfx_amd as tfx
from tressfx_amd import synth

from tfx_testutil import DeviceMesh, first_mismatch, set_substep_inputs
from test_gpu_parity import make_hair, small_workload

pytestmark = pytest.mark.gpu

# (guides, V, follow per guide, SDF: None | ("explicit", n) | ("cells_x", n))
MEMBERS = [(50, 16, 0, ("explicit", 64)), (300, 32, 1, ("cells_x", 40)), (70, 64, 0, None),
           (33, 8, 2, ("explicit", 48)), (129, 4, 0, None), (211, 32, 0, ("explicit", 64))]
CELL = 0.005


class Member:
    def __init__(self, torch, i, G, V, F, sdf_kind, copies):
        self.wl = small_workload(G, V, F, seed=200 + i)
        self.hairs = [make_hair(self.wl) for _ in range(copies)]
        self.dm = None
        self.sdfs = [None] * copies
        if sdf_kind is not None:
            rng = np.random.Generator(np.random.PCG64(300 + i))
            mesh = synth.NoisySphereMesh(rng, radius=0.1, subdivisions=2 + (i % 2), amp=0.01)
            self.dm = DeviceMesh(torch, mesh, synth.head_yaw)
            desc = self.dm.update(0.0)
            kind, n = sdf_kind
            if kind == "explicit":
                ex = ((-0.5 * n * CELL,) * 3, CELL, (n,) * 3)
                self.sdfs = [tfx.TressFXSDFCollision(desc, collision_margin=3.0, explicit=ex) for _ in range(copies)]
            else:
                self.sdfs = [tfx.TressFXSDFCollision(desc, num_cells_in_x=n, collision_margin=1.0)
                             for _ in range(copies)]

    def substep_inputs(self, frame, sub):
        wl = self.wl
        t = (frame + (sub + 1) / wl.substeps) / synth.FRAME_RATE
        bones = np.asarray(wl.bone_fn(t), np.float32).reshape(-1)
        st = tfx.SimulationSettings.from_buffer_copy(wl.settings)
        if wl.wind_fn is not None:
            st.wind_magnitude = wl.wind_fn(t)
        return bones, st


def _run(torch, members, frames, substeps_of, plain_copy, batch_copies, stream_of, graph_of, batches=None):
    """Advances copy `plain_copy` of every member through the per-object
    reference dispatch and each batch in `batch_copies` through the batch
    API (new batches unless given); compares after every frame."""
    system = tfx.TressFXSDFCollisionSystem()
    sim = tfx.TressFXSimulation()
    if batches is None:
        batches = {}
        for c in batch_copies:
            b = tfx.HairBatch()
            for m in members:
                b.add(m.hairs[c], m.sdfs[c])
            assert len(b) == len(members)
            batches[c] = b
    for frame in frames:
        nsub = substeps_of(frame)
        for m in members:
            if m.dm is not None:
                desc = m.dm.update(frame / synth.FRAME_RATE)
                for s in m.sdfs:
                    s.SetMesh(desc)
        torch.cuda.synchronize()
        # per-object reference order
        for m in members:
            if m.sdfs[plain_copy] is not None:
                m.sdfs[plain_copy].Update(system)
        for sub in range(nsub):
            for m in members:
                m.wl.substeps = nsub
                set_substep_inputs(m.hairs[plain_copy], m.wl, frame, sub)
            sim.Simulate([m.hairs[plain_copy] for m in members])
            for m in members:
                if m.sdfs[plain_copy] is not None:
                    m.sdfs[plain_copy].CollideWithHair(system, m.hairs[plain_copy])
        # batches
        for c, b in batches.items():
            b.begin_frame()
            for sub in range(nsub):
                ins = [m.substep_inputs(frame, sub) for m in members]
                bones = np.stack([x[0] for x in ins])
                settings = (tfx.SimulationSettings * len(members))(*[x[1] for x in ins])
                b.stage_substep(bones, settings, np.float32(1.0 / synth.FRAME_RATE / nsub))
            b.run_frame(stream_of(c), use_graph=graph_of(c))
        torch.cuda.synchronize()
        for m in members:
            ref = m.hairs[plain_copy].download_state()
            for c in batches:
                got = m.hairs[c].download_state()
                for k in ("pos", "prev", "tangent"):
                    mm = first_mismatch(got[k], ref[k])
                    assert mm is None, (f"frame {frame} copy {c} V={m.wl.asset.num_vertices_per_strand} "
                                        f"G={m.wl.asset.num_guide_strands} {k}: {mm}")
                assert m.hairs[c].info().simulation_frame == m.hairs[plain_copy].info().simulation_frame
    return batches


def test_batch_direct_and_graph_equal_per_object_dispatch(gpu):
    torch = gpu
    print("[batch] direct + graph vs per-object, 6 members, 5 frames", flush=True)
    members = [Member(torch, i, *spec, copies=3) for i, spec in enumerate(MEMBERS)]
    side = torch.cuda.Stream()
    _run(torch, members, range(5), lambda f: 2, plain_copy=0, batch_copies=(1, 2),
         stream_of=lambda c: None if c == 1 else side.cuda_stream, graph_of=lambda c: c == 2)


def test_batch_sdf_deferral_modes_equal_per_object_dispatch(gpu):
    """The SDF collision inside the fused launches (TFX_SDF_DEFER_NEVER) and in
    a collide pass after each substep's fused launches (_ALWAYS; _AUTO picks
    it for these small classes, which fit in one round of resident blocks)
    give the per-object dispatch's results bit for bit, directly and from a
    graph, the deferral changed between frames of one batch included."""
    torch = gpu
    print("[batch] SDF deferral never/always/auto vs per-object, 6 members, 4 frames", flush=True)
    members = [Member(torch, i, *spec, copies=4) for i, spec in enumerate(MEMBERS)]
    side = torch.cuda.Stream()
    modes = {1: tfx.TFX_SDF_DEFER_NEVER, 2: tfx.TFX_SDF_DEFER_ALWAYS, 3: tfx.TFX_SDF_DEFER_AUTO}
    batches = {}
    for c, mode in modes.items():
        b = tfx.HairBatch()
        for m in members:
            b.add(m.hairs[c], m.sdfs[c])
        b.set_sdf_deferral(mode)
        batches[c] = b
    _run(torch, members, range(2), lambda f: 2, plain_copy=0, batch_copies=tuple(modes), batches=batches,
         stream_of=lambda c: side.cuda_stream, graph_of=lambda c: c != 1)
    batches[1].set_sdf_deferral(tfx.TFX_SDF_DEFER_ALWAYS)
    batches[2].set_sdf_deferral(tfx.TFX_SDF_DEFER_NEVER)
    _run(torch, members, range(2, 4), lambda f: 2, plain_copy=0, batch_copies=tuple(modes), batches=batches,
         stream_of=lambda c: side.cuda_stream, graph_of=lambda c: c != 1)
    with pytest.raises(tfx.TFXError):
        batches[1].set_sdf_deferral(2)


def test_batch_zero_substep_frame_joins_the_sdf_build(gpu):
    """With a deferring class the batch builds the SDFs on a second stream that
    the first substep joins; a frame with no substeps must join it too (a
    graph capture would otherwise end with unjoined work, and a direct frame
    could let the next frame's upload race the build).  Frames of 2, 0, 2
    substeps, direct and graph-replayed, automatic deferral (which these
    small classes take): hair state and every SDF equal the per-object
    dispatch after each frame."""
    torch = gpu
    print("[batch] zero-substep frame with the overlapped SDF build", flush=True)
    members = [Member(torch, i, *spec, copies=3) for i, spec in enumerate(MEMBERS)]
    side = torch.cuda.Stream()
    counts = [2, 0, 2, 0]
    batches = None
    for f in range(len(counts)):
        batches = _run(torch, members, [f], lambda fr: counts[fr], plain_copy=0, batch_copies=(1, 2),
                       batches=batches, stream_of=lambda c: side.cuda_stream, graph_of=lambda c: c == 2)
        torch.cuda.synchronize()
        for i, m in enumerate(members):
            if m.sdfs[0] is None:
                continue
            want = m.sdfs[0].download().view(np.uint32)
            for c in (1, 2):
                assert np.array_equal(m.sdfs[c].download().view(np.uint32), want), f"frame {f} member {i} copy {c}"


def test_batch_graph_recaptures_on_substep_change(gpu):
    """A frame with another substep count re-captures its slot's graph; both
    slots alternate between frames."""
    torch = gpu
    print("[batch] graph with 2,2,1,3,2,2 substeps", flush=True)
    members = [Member(torch, i, *spec, copies=2) for i, spec in enumerate(MEMBERS[:4])]
    side = torch.cuda.Stream()
    counts = [2, 2, 1, 3, 2, 2]
    _run(torch, members, range(len(counts)), lambda f: counts[f], plain_copy=0, batch_copies=(1,),
         stream_of=lambda c: side.cuda_stream, graph_of=lambda c: True)


def test_batch_sparse_sdf_build_and_shared_sdf(gpu):
    """The batched SDF build initialises and finalises only the bricks that the
    triangles' node boxes overlap (resetting the previous build's others); its
    grid must equal the per-object build of the same mesh after every frame,
    with the mesh rotating and deforming, the grid origin following its
    bounding box, graph and direct frames alternating, and a per-object build
    of the batched SDF in between (which drops the band history).  Two hair
    objects share that SDF in the batch (one build per frame) and must equal
    the per-object dispatch colliding both with one SDF; a third has its own
    128^3 grid (with it the build takes four lanes per brick)."""
    torch = gpu
    print("[batch] sparse SDF build, shared SDF, 8 frames", flush=True)
    specs = [(120, 16, 0), (90, 32, 1), (64, 16, 0)]
    wls = [small_workload(G, V, F, seed=400 + i) for i, (G, V, F) in enumerate(specs)]
    rng = np.random.Generator(np.random.PCG64(77))
    mesh = synth.NoisySphereMesh(rng, radius=0.1, subdivisions=3, amp=0.015, anim_hz=2.0)
    dm = DeviceMesh(torch, mesh, synth.head_yaw)
    desc = dm.update(0.0)
    sdf_ref, sdf_b = (tfx.TressFXSDFCollision(desc, num_cells_in_x=40, collision_margin=1.0) for _ in range(2))
    big = ((-0.16, -0.16, -0.16), 0.0025, (128, 128, 128))
    big_ref, big_b = (tfx.TressFXSDFCollision(desc, collision_margin=1.0, explicit=big) for _ in range(2))
    colliders_ref, colliders_b = [sdf_ref, sdf_ref, big_ref], [sdf_b, sdf_b, big_b]
    hairs_ref = [make_hair(w) for w in wls]
    hairs_b = [make_hair(w) for w in wls]
    batch = tfx.HairBatch()
    for h, c in zip(hairs_b, colliders_b):
        batch.add(h, c)
    system = tfx.TressFXSDFCollisionSystem()
    sim = tfx.TressFXSimulation()
    side = torch.cuda.Stream()
    members = [Member.__new__(Member) for _ in wls]
    for m, w in zip(members, wls):
        m.wl = w
    prev, changed = None, 0
    for frame in range(8):
        desc = dm.update(0.1 * frame)        # 0.1 s apart: the head turns and the band moves
        for c in (sdf_ref, sdf_b, big_ref, big_b):
            c.SetMesh(desc)
        torch.cuda.synchronize()
        if frame == 5:
            sdf_b.Update(system)
        sdf_ref.Update(system)
        big_ref.Update(system)
        for sub in range(2):
            for h, w in zip(hairs_ref, wls):
                w.substeps = 2
                set_substep_inputs(h, w, frame, sub)
            sim.Simulate(hairs_ref)
            for h, c in zip(hairs_ref, colliders_ref):
                c.CollideWithHair(system, h)
        batch.begin_frame()
        for sub in range(2):
            ins = [m.substep_inputs(frame, sub) for m in members]
            bones = np.stack([x[0] for x in ins])
            settings = (tfx.SimulationSettings * len(members))(*[x[1] for x in ins])
            batch.stage_substep(bones, settings, np.float32(1.0 / synth.FRAME_RATE / 2))
        batch.run_frame(side.cuda_stream, use_graph=frame % 2 == 1)
        torch.cuda.synchronize()
        ref = sdf_ref.download()
        for name, got, want in (("shared", sdf_b.download(), ref), ("128^3", big_b.download(), big_ref.download())):
            mm = first_mismatch(got, want)
            assert mm is None, f"frame {frame}: batched {name} SDF differs at cell {mm[0]}: {mm[1]!r} vs {mm[2]!r}"
        if prev is not None:
            changed += int((prev != ref).sum())
        prev = ref
        for i, (h, hr) in enumerate(zip(hairs_b, hairs_ref)):
            got, want = h.download_state(), hr.download_state()
            for k in ("pos", "prev", "tangent"):
                mm = first_mismatch(got[k], want[k])
                assert mm is None, f"frame {frame} hair {i} {k}: {mm}"
    assert changed > 10000, "the mesh barely moved: the band history is not exercised"


def test_batch_membership_change(gpu):
    """Removing and adding members re-lays out the batch; the remaining
    members keep stepping exactly."""
    torch = gpu
    print("[batch] membership change", flush=True)
    members = [Member(torch, i, *spec, copies=2) for i, spec in enumerate(MEMBERS[:4])]
    side = torch.cuda.Stream()
    batches = _run(torch, members, range(2), lambda f: 2, plain_copy=0, batch_copies=(1,),
                   stream_of=lambda c: side.cuda_stream, graph_of=lambda c: True)
    b = batches[1]
    b.remove(members[1].hairs[1])
    assert len(b) == 3
    with pytest.raises(tfx.TFXError):
        b.remove(members[1].hairs[1])
    with pytest.raises(tfx.TFXError):
        b.add(members[0].hairs[1], members[0].sdfs[1])
    rest = [members[0], members[2], members[3]]
    _run(torch, rest, range(2, 4), lambda f: 2, plain_copy=0, batch_copies=(1,),
         stream_of=lambda c: side.cuda_stream, graph_of=lambda c: True, batches=batches)
    # and back in, with another launch shape
    b.add(members[1].hairs[1], members[1].sdfs[1])
    b.set_launch_shape(32, 0)
    # neither copy of the removed object advanced meanwhile (the wind depends
    # on the frame counter, so the counters must agree as well as the states)
    assert members[1].hairs[1].info().simulation_frame == members[1].hairs[0].info().simulation_frame
    _run(torch, rest + [members[1]], range(4, 6), lambda f: 2, plain_copy=0, batch_copies=(1,),
         stream_of=lambda c: side.cuda_stream, graph_of=lambda c: True, batches=batches)


def test_batch_launch_shapes_per_length_and_pass_timer_labels(gpu):
    """Launch shapes pinned per strand length stay with their length (each
    length's blocks-per-CU cap is its own), results equal the per-object
    dispatch, and the pass timers record each fused launch under
    "FusedSubstep/V<V>/B<block vertices>": reading a length or a length and
    block size takes exactly its launches, "FusedSubstep" takes the rest."""
    torch = gpu
    print("[batch] pinned launch shapes per length and fused pass-timer labels", flush=True)
    members = [Member(torch, i, *spec, copies=2) for i, spec in enumerate(MEMBERS)]
    b = tfx.HairBatch()
    for m in members:
        b.add(m.hairs[1], m.sdfs[1])
    b.set_launch_shape(16, 0, 4)      # 512-vertex blocks
    b.set_launch_shape(32, 1, 1)      # 1024-vertex blocks, one per CU
    b.set_launch_shape(64, 1, 3)      # 1024-vertex blocks
    tfx.profiling_enable(True)
    try:
        tfx.profiling_read("FusedSubstep")
        _run(torch, members, range(2), lambda f: 2, plain_copy=0, batch_copies=(1,), batches={1: b},
             stream_of=lambda c: None, graph_of=lambda c: False)
        torch.cuda.synchronize()
        assert tfx.profiling_read("FusedSubstep/V16/B1024")[1] == 0
        assert tfx.profiling_read("FusedSubstep/V16/B512")[1] == 4
        assert tfx.profiling_read("FusedSubstep/V32")[1] == 4        # both 32-vertex members, one launch each
        assert tfx.profiling_read("FusedSubstep/V32/B1024")[1] == 0  # already read
        assert tfx.profiling_read("FusedSubstep/V64/B1024")[1] == 4
        # the 4- and 8-vertex classes: one launch per substep each
        ms, n = tfx.profiling_read("FusedSubstep")
        assert n == 8 and ms > 0.0
    finally:
        tfx.profiling_enable(False)


def test_batch_errors(gpu):
    b = tfx.HairBatch()
    st = tfx.SimulationSettings.defaults()
    with pytest.raises(tfx.TFXError):
        b.stage_substep(None, st, 0.01)          # no frame begun
    with pytest.raises(tfx.TFXError):
        b.run_frame()
    b.begin_frame()
    b.run_frame()                                # empty batch: nothing to do
    with pytest.raises(tfx.TFXError):
        b.set_launch_shape(12, 0)


def test_batch_autotune_keeps_results(gpu):
    """The autotuner times every launch shape of each strand length on the
    staged frame (from a snapshot of the state, restored after every run) and
    pins the fastest: the frame it was called in, and later graph-replayed
    frames with the pinned shapes, equal the per-object reference dispatch bit
    for bit."""
    torch = gpu
    print("[batch] autotune, 4 members, 4 frames", flush=True)
    members = [Member(torch, i, *spec, copies=2) for i, spec in enumerate(MEMBERS[:4])]
    b = tfx.HairBatch()
    for m in members:
        b.add(m.hairs[1], m.sdfs[1])
    system = tfx.TressFXSDFCollisionSystem()
    sim = tfx.TressFXSimulation()
    side = torch.cuda.Stream()
    records = None
    for frame in range(4):
        for m in members:
            if m.dm is not None:
                desc = m.dm.update(frame / synth.FRAME_RATE)
                for s in m.sdfs:
                    s.SetMesh(desc)
        torch.cuda.synchronize()
        for m in members:
            if m.sdfs[0] is not None:
                m.sdfs[0].Update(system)
        for sub in range(2):
            for m in members:
                m.wl.substeps = 2
                set_substep_inputs(m.hairs[0], m.wl, frame, sub)
            sim.Simulate([m.hairs[0] for m in members])
            for m in members:
                if m.sdfs[0] is not None:
                    m.sdfs[0].CollideWithHair(system, m.hairs[0])
        b.begin_frame()
        for sub in range(2):
            ins = [m.substep_inputs(frame, sub) for m in members]
            b.stage_substep(np.stack([x[0] for x in ins]), (tfx.SimulationSettings * len(members))(*[x[1] for x in ins]),
                            np.float32(1.0 / synth.FRAME_RATE / 2))
        if frame == 1:
            records = b.autotune(side.cuda_stream, repeats=2)
        b.run_frame(side.cuda_stream, use_graph=frame >= 1)
        torch.cuda.synchronize()
        for m in members:
            ref, got = m.hairs[0].download_state(), m.hairs[1].download_state()
            for k in ("pos", "prev", "tangent"):
                mm = first_mismatch(got[k], ref[k])
                assert mm is None, f"frame {frame} V={m.wl.asset.num_vertices_per_strand} {k}: {mm}"
    lengths = sorted({m.wl.asset.num_vertices_per_strand for m in members})
    assert sorted(r["vertices_per_strand"] for r in records) == lengths
    for r in records:
        print(r)
        assert r["variant"] in (0, 1, 2) and r["blocks_per_cu"] >= 1 and r["candidates"] >= 1
        assert 0 < r["chosen_us"] and (r["default_us"] == 0 or r["chosen_us"] <= r["default_us"])
