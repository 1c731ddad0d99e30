"""GPU parity of the HIP path against the CPU restatement (oracle).

From identical state, every kernel must reproduce the FP32 transcription bit
for bit (both evaluate the reference's expressions in the same order with IEEE
division/sqrt and no contraction), and every substep must be within 1e-5
relative per vertex position of the FP64 oracle.  Multi-frame runs are
compared stepwise: before every substep the GPU state is loaded into the
oracle.
"""
import numpy as np
import pytest

import oracle
import tressfx_amd as tfx
from tressfx_amd import synth
from tfx_testutil import (REL_TOL, DeviceMesh, bits_equal, check_fp64, first_mismatch, fp64_reference,
                          oracle_pair, rel_err, set_substep_inputs)

pytestmark = pytest.mark.gpu

STAGE_NAMES = ["IntegrationAndGlobalShapeConstraints", "CalculateStrandLevelData", "VelocityShockPropagation",
               "LocalShapeConstraints", "LengthConstriantsWindAndCollision", "UpdateFollowHairVertices"]


def small_workload(G, V, F, wind=True, capsules=True, seed=7):
    rng = np.random.Generator(np.random.PCG64(seed))
    roots, nrm = synth.hemisphere_roots(G, 0.1, rng)
    strands = synth.grow_strands(roots, nrm, V, 0.005, 0.15, rng)
    asset = synth.build_asset(strands, F, rng, follow_sigma=0.002, tip_separation=0.5 if F else 0.0)
    s = synth.settings_config2()
    s.tip_separation = 0.5
    if not wind:
        s.wind_magnitude = 0.0
    caps = None
    if capsules:
        caps = (np.array([[0.0, 0.0, 0.0, 0.1], [0.02, 0.1, 0.0, 0.03]], np.float32),
                np.array([[0.0, 1e-3, 0.0, 0.1], [0.05, 0.14, 0.0, 0.02]], np.float32))
    wl = synth.Workload(asset, s, 2, lambda t: synth.yaw_pitch_matrix(0.6 * np.sin(2 * np.pi * t),
                                                                       0.3 * np.sin(2 * np.pi * 0.7 * t)),
                        capsules=caps, wind_fn=(lambda t: synth.wind_gust(t)) if wind else None)
    return wl


def make_hair(wl):
    hair = tfx.TressFXHairObject(wl.asset)
    if wl.capsules is not None:
        hair.SetCollisionCapsules(*wl.capsules)
    return hair


def compare_state_fp32(gpu_st, orc, what, keys=("pos", "prev", "prevprev", "tangent")):
    for k in keys:
        mm = first_mismatch(gpu_st[k], getattr(orc, k))
        assert mm is None, f"{what}: {k} differs from the FP32 oracle at {mm[0]} gpu={mm[1]!r} oracle={mm[2]!r} ({mm[3]} words)"


@pytest.mark.parametrize("V,F", [(16, 0), (32, 2), (4, 1), (8, 0), (64, 1)])
def test_stages_bit_exact_vs_fp32_oracle(gpu, V, F):
    """Each of the six kernels, one at a time, from the GPU's own state."""
    print(f"[stages] V={V} F={F}", flush=True)
    wl = small_workload(96, V, F)
    hair = make_hair(wl)
    sim = tfx.TressFXSimulation()
    o32 = oracle.OracleHair(wl.asset, np.float32)
    for frame in range(3):
        for sub in range(wl.substeps):
            set_substep_inputs(hair, wl, frame, sub)
            params, bones = hair.simulation_params()
            o32.set_state(hair.download_state())
            for stage in range(6):
                reps = params.cpu_local_shape_iterations if stage == 3 else 1
                for _ in range(reps):
                    sim.RunStage(hair, stage)
                    o32.run_stage(params, bones, stage)
                st = hair.download_state()
                compare_state_fp32(st, o32, f"frame {frame} sub {sub} stage {STAGE_NAMES[stage]}")
                if stage in (1,):
                    G = wl.asset.num_guide_strands
                    gi = np.arange(G) * (F + 1)
                    mm = first_mismatch(st["strand_level"][gi], o32.strand_level[gi])
                    assert mm is None, f"strand level data differs {mm}"
            hair.IncreaseSimulationFrame()


@pytest.mark.parametrize("V,F,wind", [(16, 0, False), (32, 4, True), (64, 0, True), (8, 2, True)])
def test_substeps_vs_oracles(gpu, V, F, wind):
    """Simulate() (reference dispatch structure) stepwise vs FP32 (bit-exact)
    and FP64 (<= 1e-5 relative) over several frames."""
    print(f"[substeps] V={V} F={F} wind={wind}", flush=True)
    wl = small_workload(128, V, F, wind=wind)
    hair = make_hair(wl)
    sim = tfx.TressFXSimulation()
    o32 = oracle.OracleHair(wl.asset, np.float32)
    worst = 0.0
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
            compare_state_fp32(st, o32, f"frame {frame} sub {sub}")
            worst = max(worst, check_fp64(st, ref, sens, f"frame {frame} sub {sub}", V))
    print(f"worst relative error vs FP64: {worst:.3e} (bound 1e-5 at every vertex)")


def test_reset_and_frame_counter(gpu):
    wl = small_workload(64, 16, 0)
    hair = make_hair(wl)
    sim = tfx.TressFXSimulation()
    for frame in range(3):
        set_substep_inputs(hair, wl, frame, 0)
        p, _ = hair.simulation_params()
        assert p.reset_positions == (1.0 if hair.info().simulation_frame < 2 else 0.0)
        sim.Simulate([hair])
    assert hair.info().simulation_frame == 3
    hair.ResetPositions()
    p, _ = hair.simulation_params()
    assert p.reset_positions == 1.0


def _mesh_setup(torch, G=256, V=32, F=2, grid_n=64, cell=0.005, subdiv=3, seed=11):
    rng = np.random.Generator(np.random.PCG64(seed))
    roots, nrm = synth.hemisphere_roots(G, 0.1, rng)
    strands = synth.grow_strands(roots, nrm, V, 0.005, 0.15, rng)
    asset = synth.build_asset(strands, F, rng, follow_sigma=0.002)
    mesh = synth.NoisySphereMesh(rng, radius=0.1, subdivisions=subdiv, amp=0.01)
    wl = synth.Workload(asset, synth.settings_config2(), 2, synth.head_yaw, mesh=mesh,
                        sdf_grid=((-0.5 * grid_n * cell,) * 3, cell, (grid_n,) * 3), wind_fn=synth.wind_gust,
                        collision_margin=3.0)   # 15 mm: vertices near the roots lie inside it
    dm = DeviceMesh(torch, mesh, synth.head_yaw)
    return wl, dm


def test_sdf_build_vs_oracle(gpu):
    torch = gpu
    wl, dm = _mesh_setup(torch)
    desc = dm.update(0.25)
    origin, cell, dims = wl.sdf_grid
    sdf = tfx.TressFXSDFCollision(desc, collision_margin=0.5, explicit=(origin, cell, dims))
    system = tfx.TressFXSDFCollisionSystem()
    sdf.Update(system)
    g = sdf.download()
    ref32 = oracle.sdf_build(dm.verts_h, dm.mesh.indices, origin, np.float32(cell), dims, np.float32)
    mm = first_mismatch(g, ref32)
    assert mm is None, f"SDF differs from the FP32 oracle at cell {mm[0]}: gpu={mm[1]} oracle={mm[2]} ({mm[3]} cells)"
    ref64 = oracle.sdf_build(dm.verts_h.astype(np.float64), dm.mesh.indices, np.asarray(origin, np.float64),
                             float(np.float32(cell)), dims, np.float64)
    band = (g != np.float32(1e10)) & (ref64 != float(np.float32(1e10)))
    assert np.array_equal(g == np.float32(1e10), ref64 == float(np.float32(1e10)))
    assert np.abs(g[band] - ref64[band]).max() <= 1e-6
    assert band.sum() > 1000


def test_sdf_reference_grid_sizing(gpu):
    """The numCellsInX constructor (TressFXSDFCollision.cpp:38-86) vs its transcription."""
    torch = gpu
    wl, dm = _mesh_setup(torch)
    desc = dm.update(0.0)
    sdf = tfx.TressFXSDFCollision(desc, num_cells_in_x=40, collision_margin=0.5)
    gi = sdf.grid()
    lib = oracle.load()
    import ctypes
    o = (ctypes.c_float * 3)()
    c = ctypes.c_float()
    d = (ctypes.c_int * 3)()
    tot = ctypes.c_int()
    mn = (ctypes.c_float * 3)(*desc.initial_bbox_min)
    mx = (ctypes.c_float * 3)(*desc.initial_bbox_max)
    cur = (ctypes.c_float * 3)(*desc.bbox_min)
    lib.orc_sdf_grid_from_bbox(ctypes.addressof(mn), ctypes.addressof(mx), ctypes.addressof(cur), 40,
                               ctypes.addressof(o), ctypes.addressof(c), ctypes.addressof(d), ctypes.addressof(tot))
    assert (gi.num_cells_x, gi.num_cells_y, gi.num_cells_z) == tuple(d)
    assert gi.num_total_cells == tot.value
    assert gi.cell_size == c.value
    assert tuple(gi.origin) == tuple(o)
    system = tfx.TressFXSDFCollisionSystem()
    desc2 = dm.update(0.5)
    sdf.SetMesh(desc2)
    sdf.Update(system)
    g = sdf.download()
    gi = sdf.grid()
    ref = oracle.sdf_build(dm.verts_h, dm.mesh.indices, np.array(gi.origin, np.float32), np.float32(gi.cell_size),
                           (gi.num_cells_x, gi.num_cells_y, gi.num_cells_z), np.float32)
    assert bits_equal(g, ref)


@pytest.mark.parametrize("deferral,follow_w", [("auto", "guide"), ("never", "guide"), ("auto", "own")])
def test_collide_and_fused_vs_oracle(gpu, deferral, follow_w):
    """Plain Simulate + CollideWithHair and the fused substep, stepwise against
    the FP32 oracle (bit-exact) and the FP64 oracle (1e-5), with SDF, wind and
    follow strands; the fused substep's SDF collision as a collide pass after
    its launch (auto: this object fits in one round of resident blocks) or
    inside it (never).  follow_w: the follow vertices carry their guide
    vertex's w, as GenerateFollowHairs makes them (the fused kernel writes
    them as whole records), or a w of their own, which UpdateFollowHairVertices
    keeps (it writes x, y and z)."""
    torch = gpu
    wl, dm = _mesh_setup(torch)
    if follow_w == "own":
        a = wl.asset
        G, F, V = a.num_guide_strands, a.num_follow_strands_per_guide, a.num_vertices_per_strand
        pos = a.positions.reshape(G, F + 1, V, 4)
        pos[:, 1:, 2:, 3] = 0.5          # movable, but not the guide's 1.0
    origin, cell, dims = wl.sdf_grid
    system = tfx.TressFXSDFCollisionSystem()
    sim = tfx.TressFXSimulation()
    sim.set_sdf_deferral(tfx.TFX_SDF_DEFER_AUTO if deferral == "auto" else tfx.TFX_SDF_DEFER_NEVER)
    hair_a = tfx.TressFXHairObject(wl.asset)   # plain dispatch
    hair_b = tfx.TressFXHairObject(wl.asset)   # fused
    desc = dm.update(0.0)
    sdf = tfx.TressFXSDFCollision(desc, collision_margin=wl.collision_margin, explicit=(origin, cell, dims))
    o32 = oracle.OracleHair(wl.asset, np.float32)
    G, F, V = wl.asset.num_guide_strands, wl.asset.num_follow_strands_per_guide, wl.asset.num_vertices_per_strand
    guide_v01 = ((np.arange(G) * (F + 1))[:, None] * V + np.arange(2)[None, :]).reshape(-1)
    hits = 0
    for frame in range(6):
        desc = dm.update(frame / synth.FRAME_RATE)
        sdf.SetMesh(desc)
        sdf.Update(system)
        grid32 = sdf.download()
        for sub in range(wl.substeps):
            for h in (hair_a, hair_b):
                set_substep_inputs(h, wl, frame, sub)
            params, bones = hair_a.simulation_params()
            st0 = hair_a.download_state()
            hair_b.upload_state(st0)
            o32.set_state(st0)
            cp = sdf.collision_params(hair_a)
            g64 = grid32.astype(np.float64)

            def advance64(h64):
                h64.simulate(params, bones)
                oracle.sdf_collide(h64.pos, h64.prev, V, g64, np.array(cp.origin[:3], np.float64),
                                   float(cp.cell_size), dims, float(cp.collision_margin))

            ref, sens = fp64_reference(wl.asset, st0, advance64, seed=frame)
            sim.Simulate([hair_a])
            sdf.CollideWithHair(system, hair_a)
            sim.SimulateFused([hair_b], [sdf])
            o32.simulate(params, bones)
            before = o32.pos.copy()
            oracle.sdf_collide(o32.pos, o32.prev, V, grid32, np.array(cp.origin[:3], np.float32), np.float32(cp.cell_size),
                               dims, np.float32(cp.collision_margin))
            hits += int((before != o32.pos).any(axis=1).sum())
            sa = hair_a.download_state()
            sb = hair_b.download_state()
            compare_state_fp32(sa, o32, f"plain frame {frame} sub {sub}")
            compare_state_fp32(sb, o32, f"fused frame {frame} sub {sub}", keys=("pos", "prev", "tangent"))
            assert bits_equal(sb["prevprev"][guide_v01], o32.prevprev[guide_v01])
            check_fp64(sa, ref, sens, f"frame {frame} sub {sub}", V)
    assert hits > 0, "the SDF never moved a vertex: the test does not exercise the collision"


def test_collide_gradient_lookup_cases_vs_oracle(gpu):
    """CollideWithHair on placed vertices, bit-exact against the FP32 oracle.
    The forward-difference lookups at p + offset * h share the centre's cell
    on the unmoved axes and step at most one cell on the moved one; the
    vertices here hit every branch of that: steps inside the cell and into the
    next, steps out of the grid's lookup range (the far face of the last
    usable cell), coordinates exactly on cell boundaries or one step below
    them, -0 coordinates (where p + 0 * h changes the bits, taking the lookups
    as written) and plain near-surface vertices."""
    torch = gpu
    rng = np.random.default_rng(3)
    mesh = synth.NoisySphereMesh(rng, radius=0.1, subdivisions=4, amp=0.0, anim_hz=0.0)
    dm = DeviceMesh(torch, mesh)
    desc = dm.update(0.0)
    n, cell = 44, np.float32(0.005)
    origin = np.full(3, -0.11, np.float32)
    sdf = tfx.TressFXSDFCollision(desc, collision_margin=1.0, explicit=(tuple(origin), float(cell), (n, n, n)))
    system = tfx.TressFXSDFCollisionSystem()
    sdf.Update(system)
    grid32 = sdf.download()
    h = np.float32(0.1) * cell
    G, V = 320, 16
    roots, nrm = synth.hemisphere_roots(G, 0.1, rng)
    asset = synth.build_asset(synth.grow_strands(roots, nrm, V, 0.005, 0.15, rng), 0, rng)
    hair = tfx.TressFXHairObject(asset)
    st = hair.download_state()
    pts = []
    # the last usable cell is n - 3; its far face is where the +h step leaves the range
    edge = origin[0] + np.float32(n - 2) * cell
    for k in range(3):
        for x in np.linspace(edge - np.float32(2.0) * h, edge - np.float32(1e-6), 24, dtype=np.float32):
            q = np.zeros(3, np.float32)
            q[k] = x
            pts.append(q)
    for k in range(3):                                        # cell boundaries and one step below them
        for j in range(38, 42):
            b = origin[k] + np.float32(j) * cell
            for x in (b, b - h, np.nextafter(b, np.float32(0.0)), np.nextafter(b - h, np.float32(1.0))):
                q = np.zeros(3, np.float32)
                q[k] = x
                q[(k + 1) % 3] = np.float32(0.002)
                pts.append(q)
    for k in range(3):                                        # -0 on the other axes
        for r in np.linspace(0.096, 0.1005, 16, dtype=np.float32):
            for sgn in (1.0, -1.0):
                q = np.full(3, -0.0, np.float32)
                q[k] = np.float32(sgn) * r
                pts.append(q)
    d = rng.normal(size=(4000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    pts += list((d * rng.uniform(0.096, 0.104, (4000, 1))).astype(np.float32))
    pts = np.asarray(pts, np.float32)
    usable = np.arange(G * V).reshape(G, V)[:, 2:].reshape(-1)
    assert len(pts) <= len(usable)
    pos = st["pos"].copy()
    pos[usable[:len(pts)], :3] = pts
    prev = pos.copy()
    prev[:, :3] += np.float32(1e-4)
    st["pos"], st["prev"] = pos, prev
    hair.upload_state(st)
    sdf.CollideWithHair(system, hair)
    got = hair.download_state()
    cp = sdf.collision_params(hair)
    rp, rq = pos.copy(), prev.copy()
    oracle.sdf_collide(rp, rq, V, grid32, origin, cell, (n, n, n), np.float32(cp.collision_margin))
    for k, ref in (("pos", rp), ("prev", rq)):
        mm = first_mismatch(got[k], ref)
        assert mm is None, f"{k} differs at {mm[0]} gpu={mm[1]!r} ref={mm[2]!r} ({mm[3]} words)"
    moved = (rp != pos).any(axis=1)
    assert moved.sum() > 2000
    negz = np.signbit(pts) & (pts == 0)
    assert moved[usable[:len(pts)]][negz.any(axis=1)].sum() > 20, "no -0 vertex reached the gradient lookups"


def test_collide_mixed_difference_variant_vs_oracle(gpu):
    """CollideWithHair with the mixed forward/backward-difference kernel
    (TressFXSDFCollision.hlsl:597-715, every vertex including the roots),
    stepwise against the FP32 transcription, bit for bit; the fused and
    batched paths refuse a mixed SDF (TFX_ERR_UNSUPPORTED)."""
    torch = gpu
    print("[collide mixed] plain vs oracle", flush=True)
    wl, dm = _mesh_setup(torch, seed=23)
    origin, cell, dims = wl.sdf_grid
    system = tfx.TressFXSDFCollisionSystem()
    sim = tfx.TressFXSimulation()
    hair = tfx.TressFXHairObject(wl.asset)
    sdf = tfx.TressFXSDFCollision(dm.update(0.0), collision_margin=wl.collision_margin, explicit=(origin, cell, dims))
    sdf.SetCollideVariant(tfx.TFX_SDF_COLLIDE_MIXED)
    o32 = oracle.OracleHair(wl.asset, np.float32)
    V = wl.asset.num_vertices_per_strand
    hits = roots = 0
    for frame in range(4):
        sdf.SetMesh(dm.update(frame / synth.FRAME_RATE))
        sdf.Update(system)
        grid32 = sdf.download()
        for sub in range(wl.substeps):
            set_substep_inputs(hair, wl, frame, sub)
            params, bones = hair.simulation_params()
            o32.set_state(hair.download_state())
            cp = sdf.collision_params(hair)
            sim.Simulate([hair])
            sdf.CollideWithHair(system, hair)
            o32.simulate(params, bones)
            before = o32.pos.copy()
            oracle.sdf_collide(o32.pos, o32.prev, V, grid32, np.array(cp.origin[:3], np.float32),
                               np.float32(cp.cell_size), dims, np.float32(cp.collision_margin), mixed=True)
            moved = (before != o32.pos).any(axis=1)
            hits += int(moved.sum())
            roots += int(moved.reshape(-1, V)[:, :2].sum())
            compare_state_fp32(hair.download_state(), o32, f"mixed frame {frame} sub {sub}")
    print(f"mixed collide hits {hits} (root vertices {roots})")
    assert hits > 0
    with pytest.raises(tfx.TFXError) as e:
        sim.SimulateFused([hair], [sdf])
    assert e.value.code == 6
    b = tfx.HairBatch()
    b.add(tfx.TressFXHairObject(wl.asset), sdf)
    with pytest.raises(tfx.TFXError) as e:
        b.begin_frame()
    assert e.value.code == 6
    sdf.SetCollideVariant(tfx.TFX_SDF_COLLIDE_FORWARD)
    sim.SimulateFused([hair], [sdf])
    with pytest.raises(tfx.TFXError):
        sdf.SetCollideVariant(7)


@pytest.mark.parametrize("variant", [-1, 0, 1, 2])
def test_fused_multi_object_mixed_lengths(gpu, variant):
    """Several objects of different strand lengths in one fused call (one
    persistent launch per length class) equal their plain results."""
    sim_p = tfx.TressFXSimulation()
    sim_f = tfx.TressFXSimulation()
    objs = []
    for i, (G, V, F) in enumerate([(50, 16, 0), (300, 32, 1), (70, 64, 0), (33, 8, 2), (129, 4, 0)]):
        wl = small_workload(G, V, F, seed=100 + i)
        objs.append((wl, make_hair(wl), make_hair(wl)))
    for V in (4, 8, 16, 32, 64):
        sim_f.SetLaunchShape(V, variant, 2 if variant == 1 else 0)
    for frame in range(3):
        for sub in range(2):
            for wl, a, b in objs:
                set_substep_inputs(a, wl, frame, sub)
                set_substep_inputs(b, wl, frame, sub)
            sim_p.Simulate([a for _, a, _ in objs])
            sim_f.SimulateFused([b for _, _, b in objs])
            for wl, a, b in objs:
                sa, sb = a.download_state(), b.download_state()
                for k in ("pos", "prev", "tangent"):
                    mm = first_mismatch(sb[k], sa[k])
                    assert mm is None, f"V={wl.asset.num_vertices_per_strand} {k}: {mm}"


@pytest.mark.parametrize("iters", [1, 3, 5, 7])
@pytest.mark.parametrize("V", [16, 32])
def test_fused_local_shape_iteration_counts(gpu, V, iters):
    """The pipelined local-shape sweep (up to four passes per group, one lane
    each) with pass counts that leave idle lanes or need several groups, on
    enough strands that they straddle wave boundaries: fused equals plain
    (the reference re-dispatch) every substep, and plain equals the FP32
    transcription on the first."""
    print(f"[fused lsc] V={V} iterations={iters}", flush=True)
    wl = small_workload(300, V, 1, seed=60 + iters)
    wl.settings.local_constraints_iterations = iters
    a, b = make_hair(wl), make_hair(wl)
    sim_p, sim_f = tfx.TressFXSimulation(), tfx.TressFXSimulation()
    o32 = oracle.OracleHair(wl.asset, np.float32)
    for frame in range(3):
        for sub in range(2):
            set_substep_inputs(a, wl, frame, sub)
            set_substep_inputs(b, wl, frame, sub)
            if frame == 0 and sub == 0:
                params, bones = a.simulation_params()
                assert params.cpu_local_shape_iterations == iters
                o32.set_state(a.download_state())
                o32.simulate(params, bones)
            sim_p.Simulate([a])
            sim_f.SimulateFused([b])
            sa, sb = a.download_state(), b.download_state()
            if frame == 0 and sub == 0:
                compare_state_fp32(sa, o32, "plain vs FP32 transcription")
            for k in ("pos", "prev", "tangent"):
                mm = first_mismatch(sb[k], sa[k])
                assert mm is None, f"frame {frame} substep {sub}: fused {k} differs from plain at {mm}"


@pytest.mark.parametrize("variant", [0, 1, 2])
def test_fused_sdf_launch_variants(gpu, variant):
    """Fused simulation + SDF collision at every block size equals plain
    Simulate + CollideWithHair, with follow strands (staged and not)."""
    torch = gpu
    print(f"[fused sdf variants] variant={variant}", flush=True)
    system = tfx.TressFXSDFCollisionSystem()
    sim_p = tfx.TressFXSimulation()
    sim_f = tfx.TressFXSimulation()
    objs = []
    for i, (G, V, F) in enumerate([(200, 32, 1), (90, 16, 5), (40, 64, 2)]):
        wl, dm = _mesh_setup(torch, G=G, V=V, F=F, seed=40 + i)
        desc = dm.update(0.0)
        origin, cell, dims = wl.sdf_grid
        sdfs = [tfx.TressFXSDFCollision(desc, collision_margin=wl.collision_margin, explicit=(origin, cell, dims))
                for _ in range(2)]
        objs.append((wl, dm, sdfs, tfx.TressFXHairObject(wl.asset), tfx.TressFXHairObject(wl.asset)))
        sim_f.SetLaunchShape(V, variant)
    for frame in range(3):
        for wl, dm, sdfs, a, b in objs:
            desc = dm.update(frame / synth.FRAME_RATE)
            for s in sdfs:
                s.SetMesh(desc)
                s.Update(system)
        for sub in range(2):
            for wl, dm, sdfs, a, b in objs:
                set_substep_inputs(a, wl, frame, sub)
                set_substep_inputs(b, wl, frame, sub)
            sim_p.Simulate([o[3] for o in objs])
            for wl, dm, sdfs, a, b in objs:
                sdfs[0].CollideWithHair(system, a)
            sim_f.SimulateFused([o[4] for o in objs], [o[2][1] for o in objs])
            for wl, dm, sdfs, a, b in objs:
                sa, sb = a.download_state(), b.download_state()
                for k in ("pos", "prev", "tangent"):
                    mm = first_mismatch(sb[k], sa[k])
                    assert mm is None, f"V={wl.asset.num_vertices_per_strand} frame {frame} sub {sub} {k}: {mm}"
