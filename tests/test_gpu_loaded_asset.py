"""A loaded asset through the whole path: the Ratboy mohawk excerpt (70 strands
padded to 128 guides, 32 vertices, two follow strands per guide, multi-bone
skinning from its .tfxbone) and the left-hand collision-mesh excerpt skinned by
its own bones into an SDF.  Per substep the plain dispatch + CollideWithHair
is bit-exact against the FP32 transcription, and the fused substep equals it."""
import numpy as np
import pytest

import oracle
import tressfx_amd as tfx

from tfx_testutil import first_mismatch
from test_loaders import TFX, TFXBONE, TFXMESH, read_tfxbone

pytestmark = pytest.mark.gpu


def _small_bone(rng, angle=0.15, shift=0.002):
    """A rigid bone near the identity: rotation by up to `angle` radians about a
    random axis, translation up to `shift` (row-major row-vector float4x4)."""
    ax = rng.normal(size=3)
    ax /= np.linalg.norm(ax)
    a = rng.uniform(-angle, angle)
    k = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
    r = np.eye(3) + np.sin(a) * k + (1 - np.cos(a)) * (k @ k)
    m = np.eye(4)
    m[:3, :3] = r.T
    m[3, :3] = rng.uniform(-shift, shift, 3)
    return m


def test_ratboy_excerpt_plain_fused_and_oracle(gpu):
    print("[loaded asset] ratboy excerpt", flush=True)
    names, _ = read_tfxbone(TFXBONE)
    asset = tfx.TressFXAsset()
    asset.LoadHairData(TFX)
    asset.GenerateFollowHairs(2, 0.3, 0.004, seed=5)
    asset.ProcessAsset()
    asset.LoadBoneData(TFXBONE, names)
    ha = asset.hair_asset()
    V = ha.num_vertices_per_strand

    lines = [l for l in open(TFXMESH).read().splitlines() if l and not l.startswith("#")]
    mesh_names = [l.split()[1] for l in lines[1:1 + int(lines[0].split()[1])]]
    mesh = tfx.load_collision_mesh(TFXMESH, mesh_names, follow_bone=mesh_names[0])
    sk = tfx.TressFXBoneSkinning(mesh["positions"], mesh["skinning"], mesh["indices"], normals=mesh["normals"],
                                 follow_bone=mesh["follow_bone"])
    # the hand sits away from the head: move its mesh onto the hair so they touch
    p = ha.positions[:, :3]
    shift = p[: ha.num_total_vertices].mean(axis=0) - mesh["positions"].mean(axis=0)
    rest = np.zeros((mesh["positions"].shape[0], 8), np.float32)
    rest[:, :3], rest[:, 4:7] = mesh["positions"], mesh["normals"]

    plain = tfx.TressFXHairObject(ha)
    fused = tfx.TressFXHairObject(ha)
    system = tfx.TressFXSDFCollisionSystem()
    sdf = tfx.TressFXSDFCollision(sk.mesh(), num_cells_in_x=48, collision_margin=2.0)
    sim = tfx.TressFXSimulation()
    st = tfx.SimulationSettings.defaults()
    st.gravity_magnitude = 9.8
    st.wind_magnitude = 2.0
    rng = np.random.Generator(np.random.PCG64(17))
    o32 = oracle.OracleHair(ha, np.float32)
    dt = np.float32(1.0 / 120.0)
    hits = 0
    for frame in range(4):
        mb = np.stack([_small_bone(rng).reshape(16) for _ in mesh_names]).astype(np.float32)
        mb[:, 12:15] += shift.astype(np.float32)
        sk.Update(mb)
        sdf.SetMesh(sk.mesh())
        sdf.Update(system)
        grid = sdf.grid()
        cells = sdf.download()
        verts = oracle.mesh_skin(rest, mesh["skinning"], mb)
        ref_cells = oracle.sdf_build(verts, mesh["indices"], np.array(grid.origin, np.float32),
                                     np.float32(grid.cell_size), (grid.num_cells_x, grid.num_cells_y,
                                                                  grid.num_cells_z), np.float32)
        assert first_mismatch(cells, ref_cells) is None, f"frame {frame}: SDF of the skinned mesh differs"
        hb = np.stack([_small_bone(rng, 0.1, 0.001).reshape(16) for _ in names]).astype(np.float32)
        for sub in range(2):
            for h in (plain, fused):
                h.UpdateBoneMatrices(hb)
                h.UpdateSimulationParameters(st, dt)
            params, bones = plain.simulation_params()
            o32.set_state(plain.download_state())
            fused.upload_state(plain.download_state())
            sim.Simulate([plain])
            sdf.CollideWithHair(system, plain)
            sim.SimulateFused([fused], [sdf])
            o32.simulate(params, bones)
            cp = sdf.collision_params(plain)
            before = o32.pos.copy()
            oracle.sdf_collide(o32.pos, o32.prev, V, cells, np.array(cp.origin[:3], np.float32),
                               np.float32(cp.cell_size), (cp.num_cells_x, cp.num_cells_y, cp.num_cells_z),
                               np.float32(cp.collision_margin))
            hits += int((before != o32.pos).any(axis=1).sum())
            sp, sf = plain.download_state(), fused.download_state()
            for k in ("pos", "prev", "tangent"):
                mm = first_mismatch(sp[k], getattr(o32, k))
                assert mm is None, f"frame {frame} sub {sub}: plain {k} vs FP32 transcription {mm}"
                mm = first_mismatch(sf[k], sp[k])
                assert mm is None, f"frame {frame} sub {sub}: fused {k} vs plain {mm}"
    assert np.isfinite(sp["pos"]).all()
    print(f"collision hits: {hits}")
