"""The asset loaders against independent reads of the same files: excerpts of
the reference's shipped Ratboy assets (tests/golden/, written by
make_ratboy_fixtures.py) and, where the reference is mounted, the full files.
LoadHairData padding, GenerateFollowHairs geometry, ProcessAsset arrays
(tangents, rest lengths, thickness, triangle indices), LoadBoneData name
mapping and padding, and the .tfxmesh collision-mesh loader."""
import os
import struct

import numpy as np
import pytest

import tressfx_amd as tfx

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TFX = os.path.join(GOLDEN, "ratboy_mohawk_70.tfx")
TFXBONE = os.path.join(GOLDEN, "ratboy_mohawk_70.tfxbone")
TFXMESH = os.path.join(GOLDEN, "ratboy_left_hand_400.tfxmesh")
FULL = "/root/reference/bin/Objects/HairAsset/Ratboy"


def read_tfx(path):
    b = open(path, "rb").read()
    version, n, V, off_pos, off_uv = struct.unpack_from("<fIIII", b, 0)
    pos = np.frombuffer(b, np.float32, n * V * 4, off_pos).reshape(n * V, 4)
    uv = np.frombuffer(b, np.float32, n * 2, off_uv).reshape(n, 2) if off_uv else np.zeros((n, 2), np.float32)
    return version, n, V, pos, uv


def read_tfxbone(path):
    b = open(path, "rb").read()
    p = 0
    (nb,) = struct.unpack_from("<i", b, p)
    p += 4
    names = []
    for _ in range(nb):
        _, ln = struct.unpack_from("<ii", b, p)
        p += 8
        names.append(b[p:p + ln].split(b"\0")[0].decode())
        p += ln
    (ns,) = struct.unpack_from("<i", b, p)
    p += 4
    recs = []
    for _ in range(ns):
        vals = struct.unpack_from("<i" + "if" * 4, b, p)
        p += 36
        recs.append(([vals[1 + 2 * j] for j in range(4)], [vals[2 + 2 * j] for j in range(4)]))
    return names, recs


def norm(v):
    d = np.sqrt((v[..., 0] * v[..., 0] + v[..., 1] * v[..., 1]) + v[..., 2] * v[..., 2]).astype(np.float32)
    return np.where((d != 0)[..., None], v / np.where(d != 0, d, 1)[..., None], v).astype(np.float32)


def load(path=TFX, follow=0, tip=0.0, radius=0.0, seed=0, bones=None, engine=None):
    a = tfx.TressFXAsset()
    a.LoadHairData(path)
    if follow:
        a.GenerateFollowHairs(follow, tip, radius, seed)
    a.ProcessAsset()
    if bones:
        a.LoadBoneData(bones, engine)
    return a


def check_hair_data(path):
    version, n, V, pos, uv = read_tfx(path)
    a = load(path)
    info = a.info()
    G = (n - n % 64) + 64                        # TressFXAsset.cpp:111
    assert info == dict(file_strands=n, num_guide_strands=G, num_follow_strands_per_guide=0, num_vertices_per_strand=V)
    h = a.hair_asset()
    assert np.array_equal(h.positions[:n * V].view(np.uint32), pos.view(np.uint32))
    last = pos[(n - 1) * V:n * V]
    for s in range(n, G):                        # made-up strands copy the last one (:139-149)
        assert np.array_equal(h.positions[s * V:(s + 1) * V], last)
    suv, thick, tri = a.render_data()
    assert np.array_equal(suv[:n], uv) and np.all(suv[n:] == uv[-1])
    return a, h, V


def test_tfx_loader_padding_and_arrays():
    a, h, V = check_hair_data(TFX)
    S = h.num_total_strands
    p = h.positions.reshape(S, V, 4)[..., :3]
    # ComputeRestLengths (:377-399), float32 in the reference's order
    d = p[:, :-1] - p[:, 1:]
    rl = np.zeros((S, V), np.float32)
    rl[:, :-1] = np.sqrt((d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2])
    assert np.array_equal(h.rest_lengths.reshape(S, V), rl)
    # ComputeStrandTangent (:304-342); the last vertex keeps (0, 0, 0)
    t = np.zeros((S, V, 3), np.float32)
    t[:, 0] = norm(p[:, 1] - p[:, 0])
    t[:, 1:-1] = norm((norm(p[:, 1:-1] - p[:, :-2]) + norm(p[:, 2:] - p[:, 1:-1])).astype(np.float32))
    assert np.array_equal(h.tangents.reshape(S, V, 4)[..., :3], t)
    assert np.all(h.tangents[:, 3] == 0.0)
    # FillTriangleIndexArray (:276-302)
    _, thick, tri = a.render_data()
    ids = np.arange(S * V).reshape(S, V)[:, :-1].reshape(-1)
    exp = np.stack([2 * ids, 2 * ids + 1, 2 * ids + 2, 2 * ids + 2, 2 * ids + 1, 2 * ids + 3], 1).reshape(-1)
    assert np.array_equal(tri, exp)
    # ComputeThicknessCoeffs (:344-375) with its repeated division
    seg = np.sqrt((d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]).astype(np.float32)
    for s in (0, 5, S - 1):
        tv, ln = np.float32(0), np.float32(0)
        for x in seg[s]:
            tv, ln = np.float32(tv + x), np.float32(ln + x)
        for i in range(V):
            tv = np.float32(tv / ln)
            with np.errstate(invalid="ignore"):
                e = np.float32(np.sqrt(np.float32(1) - tv * tv))   # NaN once tv > 1, as in the reference
            got = thick[s * V + i]
            assert got == e or (np.isnan(got) and np.isnan(e))
    # follow root offsets are zero with w unused before GenerateFollowHairs
    assert np.all(h.follow_root_offsets[:, :3] == 0)


@pytest.mark.skipif(not os.path.exists(FULL), reason="reference assets not mounted")
def test_full_ratboy_mohawk():
    a, h, V = check_hair_data(os.path.join(FULL, "Ratboy_mohawk.tfx"))
    assert (h.num_guide_strands, V) == (2240, 32)    # 2228 strands padded to 2240


def test_generate_follow_hairs():
    r, tip = 0.012, 0.5
    a = load(follow=3, tip=tip, radius=r, seed=11)
    h = a.hair_asset()
    G, F, V = h.num_guide_strands, h.num_follow_strands_per_guide, h.num_vertices_per_strand
    assert F == 3 and h.num_total_strands == G * 4
    p = h.positions.reshape(G, F + 1, V, 4)
    off = h.follow_root_offsets.reshape(G, F + 1, 4)
    assert np.all(off[:, :, 3] == (np.arange(G) * (F + 1))[:, None])      # w = guide strand index
    assert np.all(off[:, 0, :3] == 0)
    factor = (np.float32(tip) * (np.arange(V, dtype=np.float32) / np.float32(V)) + np.float32(1)).astype(np.float32)
    exp = (p[:, :1, :, :3] + (off[:, 1:, None, :3] * factor[None, None, :, None]).astype(np.float32)).astype(np.float32)
    assert np.array_equal(p[:, 1:, :, :3], exp)
    assert np.array_equal(p[:, 1:, :, 3], np.broadcast_to(p[:, :1, :, 3], p[:, 1:, :, 3].shape))
    # offsets lie in the plane orthogonal to the root segment, within the square of side 2r
    v01 = norm(p[:, 0, 1, :3] - p[:, 0, 0, :3])
    o = off[:, 1:, :3].astype(np.float64)
    assert np.abs((o * v01[:, None, :]).sum(-1)).max() < 1e-8
    assert np.linalg.norm(o, axis=-1).max() <= r * np.sqrt(2) * (1 + 1e-6)
    assert np.linalg.norm(o, axis=-1).std() > r / 10
    # seeded: same seed same offsets, another seed other offsets
    assert np.array_equal(load(follow=3, tip=tip, radius=r, seed=11).hair_asset().follow_root_offsets,
                          h.follow_root_offsets)
    assert not np.array_equal(load(follow=3, tip=tip, radius=r, seed=12).hair_asset().follow_root_offsets,
                              h.follow_root_offsets)


def test_load_bone_data_mapping_and_padding():
    names, recs = read_tfxbone(TFXBONE)
    # an engine skeleton that lists the file's bones in reverse, plus one it lacks
    engine = list(reversed(names[1:])) + ["not_in_file"]
    a = load(follow=2, tip=0.0, radius=0.01, bones=TFXBONE, engine=engine)
    h = a.hair_asset()
    G, F = h.num_guide_strands, h.num_follow_strands_per_guide
    skin = h.bone_skinning_data.reshape(G, F + 1, 8)[:, 0]
    eid = {n: (engine.index(n) if n in engine else -1) for n in names}
    for i, (bi, w) in enumerate(recs):
        ids = [float(eid[names[b]]) for b in bi]
        ids = [0.0 if x == -1.0 else x for x in ids]    # :463-468
        assert list(skin[i, :4]) == ids, i
        assert np.array_equal(skin[i, 4:], np.array(w, np.float32)), i
    for i in range(len(recs), G):                        # padding guides copy the last record
        assert np.array_equal(skin[i], skin[len(recs) - 1])
    # a .tfxbone with more strands than the asset's guides is refused (:438-439)
    small = os.path.join(GOLDEN, "ratboy_mohawk_70.tfx")
    with pytest.raises(tfx.TFXError):
        b = tfx.TressFXAsset()
        b.LoadHairData(small)
        b.ProcessAsset()
        b.LoadBoneData(os.path.join(FULL, "Ratboy_mohawk.tfxbone") if os.path.exists(FULL) else "/nonexistent", names)


def test_tfxmesh_loader():
    lines = [l for l in open(TFXMESH).read().splitlines() if l and not l.startswith("#")]
    nb = int(lines[0].split()[1])
    names = [l.split()[1] for l in lines[1:1 + nb]]
    nv = int(lines[1 + nb].split()[1])
    vl = [l.split() for l in lines[2 + nb:2 + nb + nv]]
    nt = int(lines[2 + nb + nv].split()[1])
    tl = [l.split() for l in lines[3 + nb + nv:3 + nb + nv + nt]]
    engine = names[::-1]
    m = tfx.load_collision_mesh(TFXMESH, engine, follow_bone=names[2])
    pos = np.array([[float(x) for x in v[1:4]] for v in vl], np.float32)
    nrm = np.array([[float(x) for x in v[4:7]] for v in vl], np.float32)
    assert np.array_equal(m["positions"], pos) and np.array_equal(m["normals"], nrm)
    ids = np.array([[engine.index(names[int(b)]) for b in v[7:11]] for v in vl], np.float32)
    w = np.array([[float(x) for x in v[11:15]] for v in vl], np.float32)
    assert np.array_equal(m["skinning"][:, :4], ids) and np.array_equal(m["skinning"][:, 4:], w)
    assert np.array_equal(m["indices"], np.array([[int(x) for x in t[1:4]] for t in tl], np.uint32))
    assert m["follow_bone"] == engine.index(names[2])


def test_loader_errors():
    with pytest.raises(tfx.TFXError):
        tfx.TressFXAsset().LoadHairData("/nonexistent.tfx")
    with pytest.raises(tfx.TFXError):
        tfx.load_collision_mesh("/nonexistent.tfxmesh", [])
