"""Collision-mesh skinning (TressFXBoneSkinning) through the C-ABI: the skinned
vertex buffer bit-exact vs the FP32 transcription for 4-bone blends, the
follow-bone bounding box, and the skinned mesh feeding the SDF build (bit-exact
vs the transcription's SDF of the transcription's skinned mesh)."""
import numpy as np
import pytest

import oracle
import tressfx_amd as tfx
from tressfx_amd import synth

from tfx_testutil import first_mismatch
from test_oracle import _rand_bone

pytestmark = pytest.mark.gpu


def _skinned_sphere(rng, subdivisions=3, nbones=6):
    mesh = synth.NoisySphereMesh(rng, radius=0.1, subdivisions=subdivisions, amp=0.01, anim_hz=0.0)
    v = mesh.vertices(0.0)
    n = v.shape[0]
    skin = np.zeros((n, 8), np.float32)
    skin[:, :4] = rng.integers(0, nbones, (n, 4)).astype(np.float32)
    w = rng.uniform(0.0, 1.0, (n, 4)).astype(np.float32)
    w[:, 2:] *= rng.uniform(0, 1, (n, 2)) > 0.5   # some zero-weight slots
    w[:, 0] += 0.1
    skin[:, 4:] = w
    return mesh, v, skin


def test_mesh_skinning_vs_oracle_and_bbox(gpu):
    torch = gpu
    rng = np.random.Generator(np.random.PCG64(21))
    mesh, v, skin = _skinned_sphere(rng)
    sk = tfx.TressFXBoneSkinning(v[:, :3], skin, mesh.indices, normals=v[:, 4:7], follow_bone=2)
    d0 = sk.mesh()
    rest = np.zeros_like(v)
    rest[:, :3], rest[:, 4:7] = v[:, :3], v[:, 4:7]
    assert np.array_equal(sk.download().view(np.uint32), rest.view(np.uint32)), "the buffer starts as the rest mesh"
    # TressFXBoneSkinning.cpp:377-390: centroid +- largest distance
    c = np.zeros(3, np.float32)
    for p in v[:, :3]:
        c = (c + p).astype(np.float32)
    c = (c / np.float32(v.shape[0])).astype(np.float32)
    rad = np.sqrt(((v[:, :3] - c) ** 2).sum(axis=1)).max()
    assert np.allclose(np.array(d0.initial_bbox_min), c - rad, atol=1e-6)
    for frame in range(3):
        bones = np.tile(np.eye(4, dtype=np.float32).reshape(1, 16), (6, 1))
        for b in range(6):
            bones[b] = _rand_bone(rng, 0.05).reshape(16)
        sk.Update(bones)
        got = sk.download()
        ref = oracle.mesh_skin(rest, skin, bones)
        mm = first_mismatch(got, ref)
        assert mm is None, f"frame {frame}: skinned vertex differs {mm}"
        d = sk.mesh()
        cen = (np.array(d.initial_bbox_min) + np.array(d.initial_bbox_max)) * 0.5
        m = bones[2].astype(np.float64).reshape(4, 4)
        shift = cen @ m[:3, :3] + m[3, :3] - cen
        assert np.allclose(np.array(d.bbox_min), np.array(d.initial_bbox_min) + shift, atol=1e-6)


def test_skinned_mesh_feeds_the_sdf(gpu):
    """Skin -> SDF Update each frame (the reference order, src/Simulation.cpp:72-82),
    bit-exact vs the transcription."""
    torch = gpu
    rng = np.random.Generator(np.random.PCG64(22))
    mesh, v, skin = _skinned_sphere(rng, subdivisions=3, nbones=3)
    sk = tfx.TressFXBoneSkinning(v[:, :3], skin, mesh.indices, normals=v[:, 4:7], follow_bone=0)
    rest = np.zeros_like(v)
    rest[:, :3], rest[:, 4:7] = v[:, :3], v[:, 4:7]
    system = tfx.TressFXSDFCollisionSystem()
    sdf = tfx.TressFXSDFCollision(sk.mesh(), num_cells_in_x=40, collision_margin=0.5)
    for frame in range(3):
        bones = np.stack([_rand_bone(rng, 0.02).reshape(16) for _ in range(3)]).astype(np.float32)
        sk.Update(bones)
        sdf.SetMesh(sk.mesh())
        sdf.Update(system)
        g = sdf.grid()
        got = sdf.download()
        verts = oracle.mesh_skin(rest, skin, bones)
        ref = oracle.sdf_build(verts, mesh.indices, np.array(g.origin, np.float32), np.float32(g.cell_size),
                               (g.num_cells_x, g.num_cells_y, g.num_cells_z), np.float32)
        mm = first_mismatch(got, ref)
        assert mm is None, f"frame {frame}: SDF differs {mm}"
        assert (got < 1e9).sum() > 1000
