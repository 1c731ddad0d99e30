"""Host logic of the harness without a GPU: the served-form load schedule."""
import numpy as np

from tressfx_amd import harness, synth


class _FakeServer:
    substeps = 2


def test_served_schedule_rates_and_inputs():
    wls = [synth.config3_instance(i, 60, V, 64, seed=9, subdivisions=2) for i, V in enumerate((16, 32))]
    run = harness.ServedRun(_FakeServer(), wls, initial=2, duration_s=20.0, inst_rate=20.0, req_rate=500.0,
                            ks=(1, 2, 4), seed=1, horizon=16)
    # Poisson counts within 5 standard deviations of their means
    assert abs(len(run.arrivals) - 400) < 5 * 20
    assert abs(len(run.requests) - 10000) < 5 * 100
    assert np.all(np.diff(run.arrivals) >= 0) and np.all(np.diff(run.requests) >= 0)
    assert set(np.unique(run.req_k)) == {1, 2, 4}
    hb, mb, wind = run.inputs[0]
    assert hb.shape == (16, 2, 16) and mb.shape == (16, 16) and wind.shape == (16, 2)
    # frame f, substep s is the trajectory at (f + (s + 1) / substeps) / 60
    np.testing.assert_array_equal(hb[3, 1], wls[0].bone_fn((3 + 1.0) / synth.FRAME_RATE).reshape(16))
    np.testing.assert_array_equal(mb[5], wls[0].bone_fn(5 / synth.FRAME_RATE).reshape(16))
    m = run.meshes[1]
    assert m["skinning"].shape == (m["positions"].shape[0], 8) and np.all(m["skinning"][:, 4] == 1.0)
