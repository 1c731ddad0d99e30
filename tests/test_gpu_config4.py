"""The served form in the request loop's regime: the config-3 size
distribution (strand counts log-uniform in [500, 50,000], 16/32/64 vertices
per strand), 300 instances resident at the start, new instances arriving as a
Poisson process at 20/s with exponential lifetimes (mean 30 s) and step
requests of k in {1, 2, 4} frames arriving as a Poisson process at 5,000/s on
random resident instances, with the SDF collision deferral chosen per pass.
After the window drains, instances still resident -- the one with the most
frames of each strand length -- must equal, bit for bit, the per-object
reference dispatch of the same frames (collision-mesh skinning, SDF Update,
per substep UpdateBoneMatrices, UpdateSimulationParameters, Simulate,
CollideWithHair; src/Simulation.cpp:50-131).
"""
import numpy as np
import pytest

import tressfx_amd as tfx
from tressfx_amd import harness, synth

from tfx_testutil import first_mismatch

pytestmark = pytest.mark.gpu

DT = np.float32(1.0 / synth.FRAME_RATE / 2)


def per_object(wl, mesh, hb, mb, wind, frames):
    hair = tfx.TressFXHairObject(wl.asset)
    sk = tfx.TressFXBoneSkinning(mesh["positions"], mesh["skinning"], mesh["indices"], normals=mesh["normals"])
    origin, cell, dims = wl.sdf_grid
    sdf = tfx.TressFXSDFCollision(sk.mesh(), collision_margin=wl.collision_margin, explicit=(origin, cell, dims))
    system = tfx.TressFXSDFCollisionSystem()
    sim = tfx.TressFXSimulation()
    for f in range(frames):
        sk.Update(mb[f])
        sdf.SetMesh(sk.mesh())
        sdf.Update(system)
        for s in range(wl.substeps):
            hair.UpdateBoneMatrices(hb[f, s])
            st = tfx.SimulationSettings.from_buffer_copy(wl.settings)
            st.wind_magnitude = float(wind[f, s])
            hair.UpdateSimulationParameters(st, DT)
            sim.Simulate([hair])
            sdf.CollideWithHair(system, hair)
    return hair.download_state(keys=("pos", "prev", "tangent"))


@pytest.mark.timeout(900)
def test_config4_served_window_samples_equal_per_object_dispatch(gpu, config3_workloads):
    wls = config3_workloads
    srv = tfx.HairServer(substeps_per_frame=2, substep_dt=float(DT), max_inflight=2)
    srv.set_sdf_deferral(tfx.TFX_SDF_DEFER_AUTO)
    run = harness.ServedRun(srv, wls, initial=300, duration_s=4.0, inst_rate=20.0, lifetime_s=30.0,
                            req_rate=5000.0, seed=11)
    print("[config4] served window: 300 resident, Poisson arrivals 20/s, requests 5,000/s, 4 s", flush=True)
    out = run.run(log=lambda m: print(m, flush=True))
    print({k: out[k] for k in ("resident_at_end", "instances_retired", "requests_submitted", "frames_submitted",
                               "passes", "mean_instances_per_pass", "latency_p50_ms", "latency_p99_ms")}, flush=True)
    assert out["resident_at_end"] >= 256
    assert out["instances_retired"] > 0, "no instance retired during the window"
    assert out["mean_instances_per_pass"] > 2, "requests were not coalesced"
    assert out["completed_requests"] == out["requests_submitted"]
    # the instance with the most frames of each strand length (frames within the input horizon)
    best = {}
    for iid, (w, n) in run.resident.items():
        V = wls[w].asset.num_vertices_per_strand
        if 0 < n <= run.horizon and n > best.get(V, (None, None, 0))[2]:
            best[V] = (iid, w, n)
    assert set(best) == {16, 32, 64}
    for V, (iid, w, n) in sorted(best.items()):
        print(f"[config4] instance {iid}: workload {w} (G={wls[w].asset.num_guide_strands}, V={V}), {n} frames",
              flush=True)
        got = srv.hair(iid).download_state(keys=("pos", "prev", "tangent"))
        hb, mb, wind = run.inputs[w]
        want = per_object(wls[w], run.meshes[w], hb, mb, wind, n)
        for k in ("pos", "prev", "tangent"):
            mm = first_mismatch(got[k], want[k])
            assert mm is None, f"instance {iid} (workload {w}, {n} frames): {k} differs from per-object at {mm}"
    srv.close()
