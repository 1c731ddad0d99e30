"""GenerateTransforms (TressFXSimulation.hlsl:1035-1142, the mesh-based strand
transform the reference builds under USE_MESH_BASED_HAIR_TRANSFORM) on the
GPU against the FP32 transcription, bit for bit; records of follow strands
and of strands whose triangle lies past the mesh are not written."""
import ctypes

import numpy as np
import pytest

import oracle
import tressfx_amd as tfx
from tressfx_amd import synth

from tfx_testutil import first_mismatch

pytestmark = pytest.mark.gpu


def test_generate_transforms_vs_oracle(gpu):
    torch = gpu
    print("[transforms] GenerateTransforms vs FP32 transcription", flush=True)
    rng = np.random.Generator(np.random.PCG64(43))
    G, F, V, T = 3000, 2, 16, 500
    roots, nrm = synth.hemisphere_roots(G, 0.1, rng)
    strands = synth.grow_strands(roots, nrm, V, 0.005, 0.15, rng)
    asset = synth.build_asset(strands, F, rng)
    S = G * (F + 1)
    mesh = np.ones((3 * T, 4), np.float32)
    mesh[:, :3] = rng.normal(size=(3 * T, 3)) * 0.1
    th = 0.4
    R = np.array([[np.cos(th), 0, -np.sin(th)], [0, 1, 0], [np.sin(th), 0, np.cos(th)]])
    skinned = mesh.copy()
    skinned[:, :3] = mesh[:, :3] @ R + np.array([0.01, 0.02, -0.03]) + rng.normal(size=(3 * T, 3)) * 1e-3
    mapping = (tfx.HairToTriangleMapping * S)()
    for g in range(G):
        s = g * (F + 1)
        mapping[s].triangle_index = int(rng.integers(0, T))
        b = rng.dirichlet([1.0, 1.0, 1.0])
        mapping[s].barycentric[0], mapping[s].barycentric[1], mapping[s].barycentric[2] = (float(x) for x in b)
    mapping[(G - 1) * (F + 1)].triangle_index = T + 5          # past the mesh: skipped
    hair = tfx.TressFXHairObject(asset)
    map_bytes = np.frombuffer(bytes(mapping), np.uint8).copy()
    d_map = torch.from_numpy(map_bytes).cuda()
    d_mesh = torch.from_numpy(mesh).cuda()
    d_skin = torch.from_numpy(skinned.astype(np.float32)).cuda()
    sentinel = np.full((S, 20), -7.0, np.float32)
    d_out = torch.from_numpy(sentinel).cuda()
    hair.GenerateTransforms(d_map.data_ptr(), d_mesh.data_ptr(), d_skin.data_ptr(), 3 * T, d_out.data_ptr())
    torch.cuda.synchronize()
    got = d_out.cpu().numpy()
    ref = oracle.generate_transforms(asset.positions, G, F, V, mapping, mesh, skinned.astype(np.float32),
                                     np.float32)
    guides = np.arange(G - 1) * (F + 1)
    mm = first_mismatch(got[guides], ref[guides])
    assert mm is None, f"transform differs from the FP32 transcription at {mm}"
    others = np.setdiff1d(np.arange(S), guides)
    assert np.all(got[others] == -7.0), "a follow strand's or a skipped strand's record was written"
    # the transforms take each root to its point on the moved triangle
    m = got[guides, :16].reshape(-1, 4, 4).astype(np.float64)
    root = np.concatenate([asset.positions[guides * V, :3], np.ones((len(guides), 1))], 1)
    tri = np.array([mapping[s].triangle_index for s in guides])
    bary = np.array([[mapping[s].barycentric[k] for k in range(3)] for s in guides])
    target = np.einsum("nk,nkj->nj", bary, skinned[(3 * tri[:, None] + np.arange(3)[None, :]), :3].astype(np.float64))
    assert np.allclose(np.einsum("ni,nij->nj", root, m)[:, :3], target, atol=1e-5)
