"""SDF build on adversarial meshes, bit-exact against the FP32 transcription.

The construct pass skips an atomic whose value is no smaller than the node's
current one and spreads each triangle's box over a team of lanes; neither may
change the grid.  The meshes mix a closed icosphere with triangles that stress
SignedDistancePointToTriangle and the atomicMin ordering: duplicates with
reversed winding (|d| ties between +d and -d), repeated and collinear
vertices, point triangles, slivers, triangles far smaller than a cell, one
triangle larger than the grid, and triangles partly or wholly outside it.
Both build paths (per object and batched) are compared, and a second build
over the first one's grid must give the same result.
"""
import numpy as np
import pytest

import oracle
import tressfx_amd as tfx
from tressfx_amd import synth

from tfx_testutil import DeviceMesh, first_mismatch

pytestmark = pytest.mark.gpu


class StaticMesh:
    def __init__(self, pos, indices):
        self.v = np.zeros((pos.shape[0], 8), np.float32)
        self.v[:, :3] = pos
        self.v[:, 3] = 1.0
        self.indices = np.ascontiguousarray(indices, np.uint32).reshape(-1, 3)

    @property
    def num_triangles(self):
        return self.indices.shape[0]

    def vertices(self, t, bone=None):
        return self.v


def adversarial_mesh(seed):
    rng = np.random.default_rng(seed)
    unit, faces = synth.icosphere(3)
    pos = [np.asarray(unit, np.float64) * 0.08]
    tris = [np.asarray(faces, np.int64)]
    base = len(pos[0])

    def add(points):
        nonlocal base
        pts = np.asarray(points, np.float64).reshape(-1, 3)
        tris.append(base + np.arange(len(pts)).reshape(-1, 3))
        pos.append(pts)
        base += len(pts)

    sphere = pos[0]
    f = tris[0]
    pick = rng.choice(len(f), 64, replace=False)
    add(sphere[f[pick][:, ::-1]])                        # reversed-winding duplicates
    add(sphere[f[pick[:32]]] * 1.0)                      # same-winding duplicates
    c = rng.normal(size=(200, 3))
    c = 0.09 * c / np.linalg.norm(c, axis=1, keepdims=True)
    e = rng.normal(size=(200, 3)) * 0.004
    add(np.stack([c, c, c + e], 1))                      # repeated vertex
    add(np.stack([c, c, c], 1))                          # point triangles
    add(np.stack([c, c + e, c + 2.0 * e], 1))            # collinear
    add(np.stack([c, c + e, c + 0.5 * e + 1e-7 * rng.normal(size=(200, 3))], 1))   # slivers
    add(np.stack([c, c + 1e-6 * rng.normal(size=(200, 3)), c + 1e-6 * rng.normal(size=(200, 3))], 1))  # tiny
    add([[-1.0, -1.0, 0.031], [2.0, -1.0, 0.029], [-1.0, 2.0, 0.03]])          # larger than the grid
    add([[0.1, 0.1, 0.1], [0.2, 0.1, 0.1], [0.1, 0.2, 0.15]])                  # partly outside
    add([[1.0, 1.0, 1.0], [1.1, 1.0, 1.0], [1.0, 1.1, 1.0]])                   # wholly outside
    return StaticMesh(np.concatenate(pos).astype(np.float32), np.concatenate(tris))


@pytest.mark.parametrize("seed", [0, 1])
def test_adversarial_mesh_sdf_bit_exact(gpu, seed):
    torch = gpu
    mesh = adversarial_mesh(seed)
    dm = DeviceMesh(torch, mesh)
    desc = dm.update(0.0)
    n, cell = 48, np.float32(0.005)
    origin = (-0.12, -0.12, -0.12)
    print(f"[sdf adversarial] seed {seed}: {mesh.num_triangles} triangles, {n}^3 grid", flush=True)
    ref = oracle.sdf_build(dm.verts_h, mesh.indices, np.asarray(origin, np.float32), cell, (n, n, n), np.float32)
    assert (ref != np.float32(1e10)).sum() > 5000

    sdf = tfx.TressFXSDFCollision(desc, collision_margin=0.5, explicit=(origin, float(cell), (n, n, n)))
    system = tfx.TressFXSDFCollisionSystem()
    for _ in range(2):    # the second build starts from the first one's grid
        sdf.Update(system)
        mm = first_mismatch(sdf.download(), ref)
        assert mm is None, f"per-object build differs at cell {mm[0]}: gpu={mm[1]!r} ref={mm[2]!r} ({mm[3]} cells)"

    batch = tfx.HairBatch()
    G, V = 64, 16
    rng = np.random.Generator(np.random.PCG64(seed))
    roots, nrm = synth.hemisphere_roots(G, 0.1, rng)
    asset = synth.build_asset(synth.grow_strands(roots, nrm, V, 0.005, 0.15, rng), 0, rng)
    hair = tfx.TressFXHairObject(asset)
    sdf2 = tfx.TressFXSDFCollision(desc, collision_margin=0.5, explicit=(origin, float(cell), (n, n, n)))
    batch.add(hair, sdf2)
    bones = np.tile(np.eye(4, dtype=np.float32).reshape(1, 16), (1, 1))
    batch.begin_frame()
    batch.stage_substep(bones, synth.settings_config2(), 1.0 / 60.0)
    batch.run_frame(use_graph=False)
    torch.cuda.synchronize()
    mm = first_mismatch(sdf2.download(), ref)
    assert mm is None, f"batched build differs at cell {mm[0]}: gpu={mm[1]!r} ref={mm[2]!r} ({mm[3]} cells)"
