"""The served form (tfx_server_*): continuous batching over resident
instances.  Each instance's result is the same alone, in a mix of other
instances with interleaved requests, and through the per-object reference
dispatch (bone skinning -> SDF Update -> per substep Simulate + CollideWithHair,
src/Simulation.cpp:50-131), bit for bit; admission against the HBM budget;
retirement; error paths."""
import numpy as np
import pytest

import tressfx_amd as tfx
from tressfx_amd import harness, synth

from tfx_testutil import first_mismatch

pytestmark = pytest.mark.gpu

SUBSTEPS = 2
DT = np.float32(1.0 / synth.FRAME_RATE / SUBSTEPS)


def make_instance(i, G, V, grid_n=64, seed=3):
    wl = synth.config3_instance(i, G, V, grid_n, seed=seed, subdivisions=3)
    verts = wl.mesh.vertices(0.0)   # static noise; moved by the head bone every frame
    skin = np.zeros((verts.shape[0], 8), np.float32)
    skin[:, 4] = 1.0
    mesh = dict(positions=verts[:, :3], normals=verts[:, 4:7], skinning=skin, indices=wl.mesh.indices, follow_bone=0)
    return wl, mesh


def inputs(wl, f0, frames):
    hb = np.zeros((frames, SUBSTEPS, 16), np.float32)
    mb = np.zeros((frames, 16), np.float32)
    wind = np.zeros((frames, SUBSTEPS), np.float32)
    for i, f in enumerate(range(f0, f0 + frames)):
        mb[i] = wl.bone_fn(f / synth.FRAME_RATE).reshape(16)
        for s in range(SUBSTEPS):
            t = (f + (s + 1) / SUBSTEPS) / synth.FRAME_RATE
            hb[i, s] = wl.bone_fn(t).reshape(16)
            wind[i, s] = wl.wind_fn(t)
    return hb, mb, wind


def admit(server, wl, mesh):
    return server.admit(wl.asset, wl.settings, mesh=mesh, sdf_grid=wl.sdf_grid, collision_margin=wl.collision_margin)


def per_object(wl, mesh, frames):
    """The same frames through the per-object reference dispatch."""
    hair = tfx.TressFXHairObject(wl.asset)
    sk = tfx.TressFXBoneSkinning(mesh["positions"], mesh["skinning"], mesh["indices"], normals=mesh["normals"])
    origin, cell, dims = wl.sdf_grid
    sdf = tfx.TressFXSDFCollision(sk.mesh(), collision_margin=wl.collision_margin, explicit=(origin, cell, dims))
    system = tfx.TressFXSDFCollisionSystem()
    sim = tfx.TressFXSimulation()
    hb, mb, wind = inputs(wl, 0, frames)
    for i in range(frames):
        sk.Update(mb[i])
        sdf.SetMesh(sk.mesh())
        sdf.Update(system)
        for s in range(SUBSTEPS):
            hair.UpdateBoneMatrices(hb[i, s])
            st = tfx.SimulationSettings.from_buffer_copy(wl.settings)
            st.wind_magnitude = float(wind[i, s])
            hair.UpdateSimulationParameters(st, DT)
            sim.Simulate([hair])
            sdf.CollideWithHair(system, hair)
    return hair.download_state()


def compare(a, b, what):
    for k in ("pos", "prev", "tangent"):
        mm = first_mismatch(a[k], b[k])
        assert mm is None, f"{what}: {k} differs at {mm}"


@pytest.mark.parametrize("mix_deferral", ["never", "auto"])
def test_alone_in_mix_and_per_object_are_identical(gpu, mix_deferral):
    """The alone runs keep the SDF collision inside the fused launches (the
    server's default); the mix runs it there too (never) or, chosen per pass
    (auto: these passes fit in one round of resident blocks), as a collide
    pass after each substep's fused launches."""
    print(f"[server] alone vs mix ({mix_deferral}) vs per-object", flush=True)
    x = make_instance(0, 900, 32)
    y = make_instance(1, 400, 16)
    others = [make_instance(2 + k, G, V) for k, (G, V) in enumerate([(1500, 16), (300, 64), (700, 32), (2000, 16)])]
    frames = 5

    alone = {}
    for name, (wl, mesh) in (("x", x), ("y", y)):
        srv = tfx.HairServer(substeps_per_frame=SUBSTEPS, substep_dt=float(DT))
        iid = admit(srv, wl, mesh)
        hb, mb, wind = inputs(wl, 0, frames)
        srv.wait(srv.submit(iid, hb, mb, wind))
        alone[name] = srv.hair(iid).download_state()
        srv.close()

    srv = tfx.HairServer(substeps_per_frame=SUBSTEPS, substep_dt=float(DT), max_inflight=2)
    srv.set_sdf_deferral(tfx.TFX_SDF_DEFER_NEVER if mix_deferral == "never" else tfx.TFX_SDF_DEFER_AUTO)
    ids = [admit(srv, wl, mesh) for wl, mesh in others]
    xid, yid = admit(srv, *x), admit(srv, *y)
    pos = {i: 0 for i in ids}
    tickets = []

    def push(i, wl, k):
        hb, mb, wind = inputs(wl, pos[i], k)
        pos[i] += k
        tickets.append(srv.submit(i, hb, mb, wind))

    for i, (wl, _) in zip(ids, others):
        push(i, wl, 3)
    hx, mx, wx = inputs(x[0], 0, frames)
    tickets.append(srv.submit(xid, hx[:2], mx[:2], wx[:2]))      # x joins while others run
    for i, (wl, _) in zip(ids[:2], others[:2]):
        push(i, wl, 4)
    hy, my, wy = inputs(y[0], 0, frames)
    tickets.append(srv.submit(yid, hy, my, wy))
    tickets.append(srv.submit(xid, hx[2:], mx[2:], wx[2:]))
    for i, (wl, _) in zip(ids[2:], others[2:]):
        push(i, wl, 2)
    for t in tickets:
        srv.wait(t)
    st = srv.stats()
    print(st)
    assert st["completed_requests"] == len(tickets)
    assert st["mean_instances_per_pass"] > 1.5, "requests were not batched"
    assert st["passes"] < st["completed_frames"]
    compare(srv.hair(xid).download_state(), alone["x"], "x in the mix vs alone")
    compare(srv.hair(yid).download_state(), alone["y"], "y in the mix vs alone")
    compare(alone["x"], per_object(*x, frames), "x served vs per-object dispatch")
    compare(alone["y"], per_object(*y, frames), "y served vs per-object dispatch")
    srv.close()


def test_own_substeps_and_four_bone_body_in_a_mix(gpu):
    """A reduced config-5 fur instance (3 substeps in a server whose frame is 2
    substeps; a body skinned by 4 bones, the hair by the same 4) gives the same
    result alone, in a mix with 2-substep instances, and through the per-object
    dispatch at dt = frame / 3; the 2-substep instances it ran beside match
    their own per-object dispatch."""
    print("[server] own substeps + 4-bone body", flush=True)
    fur = synth.config5(num_guides=3000, freq=20, grid_n=64, cell=0.01)
    fmesh = fur.mesh.skinned_mesh()
    frames = 4
    hb, mb, _ = harness.frame_inputs(fur, frames)
    assert hb.shape == (frames, 3, 4, 16) and mb.shape == (frames, 4, 16)
    dt3 = np.float32(np.float64(DT) * SUBSTEPS / 3)

    def admit_fur(srv):
        return srv.admit(fur.asset, fur.settings, mesh=fmesh, sdf_grid=fur.sdf_grid,
                         collision_margin=fur.collision_margin, substeps=3)

    srv = tfx.HairServer(substeps_per_frame=SUBSTEPS, substep_dt=float(DT))
    fid = admit_fur(srv)
    srv.wait(srv.submit(fid, hb, mb))
    alone = srv.hair(fid).download_state()
    srv.close()

    others = [make_instance(20 + k, G, V) for k, (G, V) in enumerate([(800, 16), (300, 32), (200, 64)])]
    srv = tfx.HairServer(substeps_per_frame=SUBSTEPS, substep_dt=float(DT), max_inflight=2)
    ids = [admit(srv, wl, mesh) for wl, mesh in others]
    fid = admit_fur(srv)
    tickets = [srv.submit(i, *inputs(wl, 0, 3)) for i, (wl, _) in zip(ids, others)]
    tickets.append(srv.submit(fid, hb[:2], mb[:2]))
    tickets += [srv.submit(i, *inputs(wl, 3, 2)) for i, (wl, _) in zip(ids, others)]
    tickets.append(srv.submit(fid, hb[2:], mb[2:]))
    for t in tickets:
        srv.wait(t)
    st = srv.stats()
    assert st["mean_instances_per_pass"] > 1.5, "requests were not batched"
    compare(srv.hair(fid).download_state(), alone, "fur in the mix vs alone")
    compare(srv.hair(ids[0]).download_state(), per_object(*others[0], 5), "2-substep neighbour vs per-object")
    compare(srv.hair(ids[2]).download_state(), per_object(*others[2], 5), "V=64 neighbour vs per-object")
    srv.close()

    hair = tfx.TressFXHairObject(fur.asset)
    sk = tfx.TressFXBoneSkinning(fmesh["positions"], fmesh["skinning"], fmesh["indices"], normals=fmesh["normals"])
    origin, cell, dims = fur.sdf_grid
    sdf = tfx.TressFXSDFCollision(sk.mesh(), collision_margin=fur.collision_margin, explicit=(origin, cell, dims))
    system = tfx.TressFXSDFCollisionSystem()
    sim = tfx.TressFXSimulation()
    for i in range(frames):
        sk.Update(mb[i])
        sdf.SetMesh(sk.mesh())
        sdf.Update(system)
        for s in range(3):
            hair.UpdateBoneMatrices(hb[i, s])
            hair.UpdateSimulationParameters(fur.settings, dt3)
            sim.Simulate([hair])
            sdf.CollideWithHair(system, hair)
    compare(alone, hair.download_state(), "fur served vs per-object dispatch at dt = frame / 3")


def test_admission_budget_and_retire(gpu):
    print("[server] admission", flush=True)
    wl, mesh = make_instance(7, 500, 16)
    probe = tfx.HairServer(substeps_per_frame=SUBSTEPS, substep_dt=float(DT))
    i0 = admit(probe, wl, mesh)
    need = probe.stats()["resident_bytes"]
    probe.close()
    assert need > 0
    srv = tfx.HairServer(hbm_budget_bytes=int(need * 1.5), substeps_per_frame=SUBSTEPS, substep_dt=float(DT))
    a = admit(srv, wl, mesh)
    with pytest.raises(tfx.TFXError) as e:
        admit(srv, wl, mesh)
    assert e.value.code == 4
    hb, mb, wind = inputs(wl, 0, 2)
    t = srv.submit(a, hb, mb, wind)
    srv.retire(a)                      # frees once its queued frames are done
    srv.wait(t)
    srv.drain()
    with pytest.raises(tfx.TFXError):
        srv.submit(a, hb, mb, wind)
    b = admit(srv, wl, mesh)           # the budget is free again
    srv.wait(srv.submit(b, hb, mb, wind))
    assert srv.stats()["resident_instances"] == 1
    srv.close()


def test_drain_waits_for_retired_instances_to_be_freed(gpu):
    """Many instances retired in a row (the worker frees the first while the
    next are retired) and drained: drain returns with every one freed and the
    budget back at zero, and the server keeps serving."""
    print("[server] retire many, drain", flush=True)
    srv = tfx.HairServer(substeps_per_frame=SUBSTEPS, substep_dt=float(DT))
    made = [make_instance(20 + i, 60 + 10 * i, 8 if i % 2 else 16) for i in range(3)]
    for rnd in range(2):
        ids = [admit(srv, *made[i % 3]) for i in range(16)]
        for i, iid in enumerate(ids):
            srv.submit(iid, *inputs(made[i % 3][0], 0, 1), track=False)
        srv.drain()
        for iid in ids:
            srv.retire(iid)
        srv.drain()
        st = srv.stats()
        assert st["resident_instances"] == 0 and st["resident_bytes"] == 0, (rnd, st)
    srv.close()


def test_server_limits_and_errors(gpu):
    srv = tfx.HairServer(max_instances=1, substeps_per_frame=SUBSTEPS, substep_dt=float(DT))
    wl, mesh = make_instance(8, 100, 8)
    a = admit(srv, wl, mesh)
    with pytest.raises(tfx.TFXError) as e:
        admit(srv, wl, mesh)
    assert e.value.code == 4
    with pytest.raises(tfx.TFXError):
        srv.retire(12345)
    with pytest.raises(tfx.TFXError):
        srv.poll(999999)
    hb, mb, wind = inputs(wl, 0, 1)
    with pytest.raises(tfx.TFXError):
        srv.submit(a + 100, hb, mb, wind)
    srv.close()
    srv = tfx.HairServer(substeps_per_frame=SUBSTEPS, substep_dt=float(DT))
    with pytest.raises(tfx.TFXError) as e:
        srv.admit(wl.asset, wl.settings, mesh=mesh, sdf_grid=wl.sdf_grid, substeps=17)
    assert e.value.code == 1
    srv.close()
