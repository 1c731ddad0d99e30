"""The library's host parameter packing against the oracle's.

TressFXHairObject keeps two constant blocks, m_SimCB[2], indexed by the
simulation frame's parity (TressFXHairObject.h:104, 129, 160):
UpdateBoneMatrices and UpdateSimulationParameters (with SetWind) write
m_SimCB[frame % 2] (TressFXHairObject.cpp:366-437, 191-235), ResetPositions
sets its reset flag and IncreaseSimulationFrame clears it and advances the
frame.  The test drives a hair object through the C-ABI and, beside it, a
model of those two blocks whose packing is the oracle's pinned transcription
(oracle.pack_sim_params), and requires the block the library hands the next
Simulate (tfx_hair_get_simulation_params) to equal the model's current block
bit for bit, for callers that update the parameters every frame: frames 0-3
and 150-157 (sin^2 of the wind gust near 1 at 157), wind directions off every
axis, V < 64 and V = 64 (the LSC iteration split of :402-411), iteration
counts 1, 2 and 5, ResetPositions between frames, and capsules set once.
"""
import ctypes

import numpy as np
import pytest

import oracle
import tressfx_amd as tfx
from tressfx_amd import synth

pytestmark = pytest.mark.gpu

# Fields of the reference's TressFXSimulationParams (TressFXConstantBuffers.h:41-110)
# plus m_CPULocalShapeIterations; num_bones, simulation_frame and the pads are
# this library's bookkeeping.
REFERENCE_FIELDS = ("wind", "wind1", "wind2", "wind3", "shape", "grav_time_tip", "sim_ints", "counts", "vsp",
                    "reset_positions", "clamp_position_delta", "center_and_radius0", "center_and_radius1",
                    "num_collision_capsules", "cpu_local_shape_iterations")


def _bits(p, name):
    v = getattr(p, name)
    return bytes(v) if isinstance(v, ctypes.Array) else bytes(dict(p._fields_)[name](v))


def _show(p, name):
    v = getattr(p, name)
    return np.ctypeslib.as_array(v).tolist() if isinstance(v, ctypes.Array) else v


class DoubleBufferedBlocks:
    """m_SimCB[2] of the reference, packed by the oracle's transcription."""

    def __init__(self, V, F, capsules):
        self.V, self.F = V, F
        self.cb = [tfx.SimulationParams(), tfx.SimulationParams()]
        self.bones = [np.tile(np.eye(4, dtype=np.float32).reshape(1, 16), (512, 1)) for _ in range(2)]
        for p in self.cb:
            c0, c1 = capsules
            p.num_collision_capsules[0] = len(c0)
            for i in range(len(c0)):
                for k in range(4):
                    p.center_and_radius0[i][k] = float(c0[i][k])
                    p.center_and_radius1[i][k] = float(c1[i][k])
        self.frame = 0
        self.reset_positions()          # the constructor's ResetPositions (TressFXHairObject.cpp:157)

    def current(self):
        return self.cb[self.frame % 2]

    def update_bones(self, m):
        m = np.asarray(m, np.float32).reshape(-1, 16)
        self.bones[self.frame % 2][:m.shape[0]] = m

    def update(self, settings, dt):
        oracle.pack_sim_params(settings, dt, self.frame, self.V, self.F, into=self.current())

    def reset_positions(self):
        self.current().reset_positions = 1.0

    def increase_frame(self):
        self.current().reset_positions = 0.0
        self.frame += 1


def _settings(rng, iters):
    s = synth.settings_config2()
    d = rng.normal(size=3)
    s.wind_direction[0], s.wind_direction[1], s.wind_direction[2] = (float(x) for x in d)
    s.wind_magnitude = float(rng.uniform(0.5, 6.0))
    s.local_constraints_iterations = iters
    s.length_constraints_iterations = int(rng.integers(1, 4))
    s.vsp_coeff = float(rng.uniform(0.5, 1.0))
    s.damping = float(rng.uniform(0.0, 0.2))
    s.tip_separation = float(rng.uniform(0.0, 1.0))
    s.global_constraint_stiffness = float(rng.uniform(0.0, 0.3))
    s.global_constraints_range = float(rng.uniform(0.0, 1.0))
    return s


@pytest.mark.parametrize("V,F", [(16, 0), (32, 2), (64, 1), (8, 0)])
@pytest.mark.parametrize("iters", [1, 2, 5])
def test_host_packing_equals_oracle_double_buffered_blocks(gpu, V, F, iters):
    print(f"[host params] V={V} F={F} local iterations={iters}", flush=True)
    rng = np.random.default_rng(1000 * V + 10 * F + iters)
    G = 8
    r2 = np.random.Generator(np.random.PCG64(5))
    roots, nrm = synth.hemisphere_roots(G, 0.1, r2)
    asset = synth.build_asset(synth.grow_strands(roots, nrm, V, 0.005, 0.15, r2), F, r2, follow_sigma=0.002)
    caps = (np.array([[0.0, 0.0, 0.0, 0.1], [0.02, 0.1, 0.0, 0.03]], np.float32),
            np.array([[0.0, 1e-3, 0.0, 0.1], [0.05, 0.14, 0.0, 0.02]], np.float32))
    hair = tfx.TressFXHairObject(asset)
    hair.SetCollisionCapsules(*caps)
    model = DoubleBufferedBlocks(V, F, caps)
    frames_checked = 0
    for frame in range(158):
        s = _settings(rng, iters)
        dt = np.float32(rng.uniform(1.0 / 240, 1.0 / 30))
        bones = np.stack([synth.yaw_pitch_matrix(float(rng.uniform(-1, 1)), float(rng.uniform(-1, 1)),
                                                 tuple(rng.uniform(-0.1, 0.1, 3))).reshape(16) for _ in range(3)])
        if frame in (2, 151, 155):             # ResetPositions between frames, before the update
            hair.ResetPositions()
            model.reset_positions()
        hair.UpdateBoneMatrices(bones)
        model.update_bones(bones)
        hair.UpdateSimulationParameters(s, dt)
        model.update(s, dt)
        if frame in (3, 156):                  # ... and after it
            hair.ResetPositions()
            model.reset_positions()
        if frame <= 3 or frame >= 150:
            got, got_bones = hair.simulation_params()
            want = model.current()
            for name in REFERENCE_FIELDS:
                assert _bits(got, name) == _bits(want, name), (
                    f"frame {frame}: {name} library {_show(got, name)} oracle {_show(want, name)}")
            assert np.array_equal(got_bones[:3].view(np.uint32), model.bones[frame % 2][:3].view(np.uint32)), frame
            assert got.simulation_frame == frame
            frames_checked += 1
        hair.IncreaseSimulationFrame()
        model.increase_frame()
    assert frames_checked == 12
    # the wind pyramid is non-trivial: four distinct corners off every axis
    w = np.array([list(got.wind)[:3], list(got.wind1)[:3], list(got.wind2)[:3], list(got.wind3)[:3]])
    assert np.linalg.matrix_rank(w) == 3 and np.all(np.abs(w) > 0)
