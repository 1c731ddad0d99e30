"""Seeded synthetic workloads (hair assets, collision meshes, trajectories).

Seeds follow `seed = 1000 * config + instance`; the generator is numpy's
PCG64.  Conventions of the reference's assets are kept:
  - positions are float4 with w = inverse mass: 0 for vertices 0 and 1, 1 after
    (tool/Maya/TressFX_Exporter.py:1585-1592);
  - rest lengths are |p_j - p_{j+1}| with a trailing 0 per strand
    (TressFXAsset.cpp:377-399);
  - tangents follow ComputeStrandTangent (TressFXAsset.cpp:304-342); their w,
    which the reference leaves uninitialised, is 0 here;
  - follow strands are interleaved after their guide, offset in the root's
    tangent plane, w copied from the guide, offset.w = guide strand index
    (TressFXAsset.cpp:177-248);
  - bone matrices are row-major row-vector 4x4 (p' = [p,1] M).
"""
import math

import numpy as np

from . import HairAsset, SimulationSettings

FRAME_RATE = 60.0


def rng_for(config, instance=0):
    return np.random.Generator(np.random.PCG64(1000 * config + instance))


# ---------------------------------------------------------------------------
# Strands
# ---------------------------------------------------------------------------
def _perp_basis(d):
    """Two unit vectors orthogonal to each row of d (GetTangentVectors-like)."""
    a = np.where(np.abs(d[:, :1]) > 0.9, np.array([[0.0, 1.0, 0.0]]), np.array([[1.0, 0.0, 0.0]]))
    t0 = np.cross(d, a)
    t0 /= np.linalg.norm(t0, axis=1, keepdims=True)
    t1 = np.cross(d, t0)
    return t0, t1


def hemisphere_roots(n, radius, rng):
    """Area-uniform roots on the upper (+y) hemisphere of a sphere at the origin."""
    y = rng.uniform(0.0, 1.0, n)
    phi = rng.uniform(0.0, 2.0 * math.pi, n)
    r = np.sqrt(np.maximum(0.0, 1.0 - y * y))
    nrm = np.stack([r * np.cos(phi), y, r * np.sin(phi)], axis=1)
    return radius * nrm, nrm


def grow_strands(roots, normals, V, segment, sigma, rng):
    """Random-walk strands: segment k's direction is segment k-1's rotated by an
    angle ~ N(0, sigma) about a random axis perpendicular to it."""
    G = roots.shape[0]
    pos = np.zeros((G, V, 3))
    pos[:, 0] = roots
    d = normals.copy()
    for k in range(1, V):
        t0, t1 = _perp_basis(d)
        ang = rng.normal(0.0, sigma, G)
        phi = rng.uniform(0.0, 2.0 * math.pi, G)
        axis = np.cos(phi)[:, None] * t0 + np.sin(phi)[:, None] * t1
        # Rodrigues rotation of d about axis (axis is perpendicular to d)
        d = d * np.cos(ang)[:, None] + np.cross(axis, d) * np.sin(ang)[:, None]
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        pos[:, k] = pos[:, k - 1] + segment * d
    return pos


def strand_tangents(p32):
    """ComputeStrandTangent on float32 strands [S, V, 3] (Vector3::Normalize
    leaves zero vectors unchanged); the last vertex keeps a zero tangent."""
    S, V, _ = p32.shape

    def nz(v):
        d = np.sqrt((v[..., 0] * v[..., 0] + v[..., 1] * v[..., 1]) + v[..., 2] * v[..., 2])
        return v / np.where(d != 0, d, np.float32(1.0))[..., None]

    t = np.zeros_like(p32)
    t[:, 0] = nz(p32[:, 1] - p32[:, 0])
    if V > 2:
        pre = nz(p32[:, 1:V - 1] - p32[:, 0:V - 2])
        nxt = nz(p32[:, 2:V] - p32[:, 1:V - 1])
        t[:, 1:V - 1] = nz((pre + nxt).astype(np.float32))
    return t


def build_asset(guide_pos, F, rng, follow_sigma=0.002, tip_separation=0.0, bone_index=0, skin=None):
    """A HairAsset from guide strands [G, V, 3] (float64), with F follow strands
    per guide (offsets ~ N(0, follow_sigma) in the root tangent plane)."""
    G, V, _ = guide_pos.shape
    S = G * (F + 1)
    pos = np.zeros((G, F + 1, V, 4), np.float32)
    gp32 = guide_pos.astype(np.float32)
    w = np.ones(V, np.float32)
    w[:2] = 0.0
    pos[:, 0, :, :3] = gp32
    pos[:, 0, :, 3] = w
    offsets = np.zeros((G, F + 1, 4), np.float32)
    offsets[:, :, 3] = (np.arange(G) * (F + 1)).astype(np.float32)[:, None]
    if F > 0:
        v01 = gp32[:, 1].astype(np.float64) - gp32[:, 0]
        v01 /= np.linalg.norm(v01, axis=1, keepdims=True)
        t0, t1 = _perp_basis(v01)
        a = rng.normal(0.0, follow_sigma, (G, F))
        b = rng.normal(0.0, follow_sigma, (G, F))
        off = (a[..., None] * t0[:, None, :] + b[..., None] * t1[:, None, :]).astype(np.float32)
        offsets[:, 1:, :3] = off
        factor = (np.float32(tip_separation) * (np.arange(V, dtype=np.float32) / np.float32(V)) + np.float32(1.0))
        pos[:, 1:, :, :3] = gp32[:, None, :, :] + off[:, :, None, :] * factor[None, None, :, None]
        pos[:, 1:, :, 3] = w
    pos = pos.reshape(S, V, 4)
    # rest lengths (float32, TressFXAsset.cpp:377-399)
    d = pos[:, :-1, :3] - pos[:, 1:, :3]
    rl = np.zeros((S, V), np.float32)
    rl[:, :-1] = np.sqrt((d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2])
    tan = np.zeros((S, V, 4), np.float32)
    tan[..., :3] = strand_tangents(pos[..., :3])
    if skin is None:
        skin = np.zeros((S, 8), np.float32)
        skin[:, 0] = bone_index
        skin[:, 4] = 1.0
    return HairAsset(G, F, V, pos.reshape(-1, 4), tan.reshape(-1, 4), rl.reshape(-1), offsets.reshape(-1, 4),
                     skin)


# ---------------------------------------------------------------------------
# Rigid trajectories (bone matrices, row-major row-vector)
# ---------------------------------------------------------------------------
def yaw_pitch_matrix(yaw, pitch=0.0, translation=(0.0, 0.0, 0.0)):
    cy, sy = math.cos(yaw), math.sin(yaw)
    cp, sp = math.cos(pitch), math.sin(pitch)
    ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])      # column-vector R_y
    rx = np.array([[1, 0, 0], [0, cp, -sp], [0, sp, cp]])      # column-vector R_x
    r = ry @ rx
    m = np.eye(4)
    m[:3, :3] = r.T                                            # row-vector convention
    m[3, :3] = translation
    return m.astype(np.float32)


def head_yaw(t, amplitude_deg=30.0, freq_hz=1.0):
    """Config-1/2 head: rigid yaw of +-30 degrees at 1 Hz."""
    return yaw_pitch_matrix(math.radians(amplitude_deg) * math.sin(2.0 * math.pi * freq_hz * t))


# ---------------------------------------------------------------------------
# Collision meshes
# ---------------------------------------------------------------------------
def _icosahedron():
    t = (1.0 + 5 ** 0.5) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    return [np.array(p, float) / np.linalg.norm(p) for p in verts], faces


def _outward(V, Fc):
    """Flip faces so that each winds counter-clockwise seen from outside."""
    n = np.cross(V[Fc[:, 1]] - V[Fc[:, 0]], V[Fc[:, 2]] - V[Fc[:, 0]])
    flip = np.einsum("ij,ij->i", n, V[Fc[:, 0]]) < 0
    Fc[flip] = Fc[flip][:, [0, 2, 1]]
    return Fc


def icosphere(subdivisions):
    """Unit icosphere, outward counter-clockwise winding. 20*4^k triangles."""
    v, faces = _icosahedron()
    for _ in range(subdivisions):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = v[a] + v[b]
                v.append(m / np.linalg.norm(m))
                cache[key] = len(v) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    V = np.array(v)
    return V, _outward(V, np.array(faces, dtype=np.uint32))


def geodesic_sphere(freq):
    """Unit geodesic sphere of frequency `freq`: every icosahedron face split
    into freq^2 triangles on a barycentric lattice and projected onto the
    sphere (20 freq^2 triangles, 10 freq^2 + 2 vertices), outward
    counter-clockwise winding.  A lattice point is identified by its integer
    barycentric weights on the icosahedron's vertex ids, so points shared by
    neighbouring faces are one vertex with bit-identical coordinates."""
    base, faces = _icosahedron()
    base = np.array(base)
    k = int(freq)
    ii, jj = np.meshgrid(np.arange(k + 1), np.arange(k + 1), indexing="ij")
    keep = ii + jj <= k
    ii, jj = ii[keep], jj[keep]
    lid = -np.ones((k + 1, k + 1), np.int64)          # lattice (i, j) -> point within a face
    lid[ii, jj] = np.arange(ii.size)
    keys = []
    for a, b, c in faces:
        ids = np.array([a, b, c])
        w = np.stack([k - ii - jj, ii, jj], 1)
        vid = np.where(w > 0, ids[None, :], -1)        # a zero weight does not name its vertex
        w = np.where(w > 0, w, 0)
        order = np.argsort(vid, axis=1, kind="stable")
        keys.append(np.concatenate([np.take_along_axis(vid, order, 1), np.take_along_axis(w, order, 1)], 1))
    keys = np.concatenate(keys)
    uniq, inv = np.unique(keys, axis=0, return_inverse=True)
    inv = inv.reshape(-1)
    pts = np.zeros((uniq.shape[0], 3))
    for q in range(3):
        vq, wq = uniq[:, q], uniq[:, 3 + q]
        pts += np.where((vq >= 0)[:, None], base[np.maximum(vq, 0)] * wq[:, None], 0.0)
    pts /= np.linalg.norm(pts, axis=1, keepdims=True)
    # triangles of one face's lattice: up (i,j),(i+1,j),(i,j+1); down (i+1,j),(i+1,j+1),(i,j+1)
    ui, uj = np.nonzero(np.add.outer(np.arange(k), np.arange(k)) < k)
    up = np.stack([lid[ui, uj], lid[ui + 1, uj], lid[ui, uj + 1]], 1)
    di, dj = np.nonzero(np.add.outer(np.arange(k), np.arange(k)) < k - 1)
    down = np.stack([lid[di + 1, dj], lid[di + 1, dj + 1], lid[di, dj + 1]], 1)
    local = np.concatenate([up, down])
    npf = ii.size
    tri = np.concatenate([inv[f * npf + local] for f in range(len(faces))]).astype(np.uint32)
    return pts, _outward(pts, tri)


class ValueNoise3:
    """3-D lattice value noise with smoothstep trilinear interpolation."""

    def __init__(self, rng, size=64):
        self.size = size
        self.table = rng.uniform(-1.0, 1.0, (size, size, size))

    def __call__(self, p):
        s = self.size
        pf = np.floor(p)
        f = p - pf
        i = pf.astype(np.int64) % s
        u = f * f * (3.0 - 2.0 * f)
        out = 0.0
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    w = ((u[:, 0] if dx else 1 - u[:, 0]) * (u[:, 1] if dy else 1 - u[:, 1]) *
                         (u[:, 2] if dz else 1 - u[:, 2]))
                    out = out + w * self.table[(i[:, 0] + dx) % s, (i[:, 1] + dy) % s, (i[:, 2] + dz) % s]
        return out


class NoisySphereMesh:
    """Icosphere of radius r displaced along the normal by a 3-octave value
    noise of amplitude `amp`, animated with period 1/anim_hz, then moved by a
    rigid bone matrix.  Produces StandardVertex float32[M, 8] arrays."""

    def __init__(self, rng, radius=0.1, subdivisions=5, amp=0.01, anim_hz=0.5, freq=6.0, scale=(1.0, 1.0, 1.0)):
        self.unit, self.indices = icosphere(subdivisions)
        self.radius = radius
        self.amp = amp
        self.anim_hz = anim_hz
        self.freq = freq
        self.scale = np.asarray(scale, float)
        self.noise = ValueNoise3(rng)

    @property
    def num_triangles(self):
        return self.indices.shape[0]

    def _displaced(self, t):
        ph = 2.0 * math.pi * self.anim_hz * t
        shift = 0.5 * np.array([math.cos(ph), math.sin(ph), 0.3 * math.cos(ph)])
        n = self.unit
        disp = np.zeros(n.shape[0])
        norm = 0.0
        for o in range(3):
            a = 0.5 ** o
            disp += a * self.noise(n * self.freq * (2 ** o) + shift * (2 ** o) + 17.0 * o)
            norm += a
        r = self.radius + self.amp * disp / norm
        return n * self.scale * r[:, None]

    def vertices(self, t, bone=None):
        if self.anim_hz == 0.0:
            # static displacement: evaluated once, then only moved rigidly
            if getattr(self, "_static", None) is None:
                self._static = self._displaced(0.0)
            p = self._static
        else:
            p = self._displaced(t)
        n = self.unit
        if bone is not None:
            ph = np.concatenate([p, np.ones((p.shape[0], 1))], axis=1) @ bone.astype(np.float64)
            p = ph[:, :3]
        out = np.zeros((p.shape[0], 8), np.float32)
        out[:, :3] = p
        out[:, 3] = 1.0
        out[:, 4:7] = n
        return out


# ---------------------------------------------------------------------------
# Configurations
# ---------------------------------------------------------------------------
def settings_config1():
    """dt 1/60, gravity 9.8, damping 0.035, global 0.1 over 0.3, local 0.8,
    length iterations 1, VSP 0.9, wind off."""
    s = SimulationSettings.defaults()
    s.gravity_magnitude = 9.8
    s.damping = 0.035
    s.global_constraint_stiffness = 0.1
    s.global_constraints_range = 0.3
    s.local_constraint_stiffness = 0.8
    s.length_constraints_iterations = 1
    s.vsp_coeff = 0.9
    s.wind_magnitude = 0.0
    return s


def settings_config2():
    s = settings_config1()
    s.length_constraints_iterations = 2
    s.wind_magnitude = 3.0
    s.wind_direction[0], s.wind_direction[1], s.wind_direction[2] = 1.0, 0.0, 0.5
    return s


def wind_gust(t, base=3.0):
    """3 m/s with a +-20 % sinusoidal gust at 0.3 Hz."""
    return base * (1.0 + 0.2 * math.sin(2.0 * math.pi * 0.3 * t))


class Workload:
    """One hair asset with its trajectory, collider and settings."""

    def __init__(self, asset, settings, substeps, bone_fn, mesh=None, sdf_grid=None, capsules=None, wind_fn=None,
                 collision_margin=0.5):
        self.asset = asset
        self.settings = settings
        self.substeps = substeps
        self.bone_fn = bone_fn
        self.mesh = mesh
        self.sdf_grid = sdf_grid          # (origin[3], cell, (nx, ny, nz))
        self.capsules = capsules          # (c0[n,4], c1[n,4])
        self.wind_fn = wind_fn
        self.collision_margin = collision_margin


def config1(num_guides=1000, V=16, seed_instance=0):
    rng = rng_for(1, seed_instance)
    roots, nrm = hemisphere_roots(num_guides, 0.1, rng)
    strands = grow_strands(roots, nrm, V, 0.005, 0.15, rng)
    asset = build_asset(strands, 0, rng)
    eps = 1e-3
    caps = (np.array([[0.0, 0.0, 0.0, 0.1]], np.float32), np.array([[0.0, eps, 0.0, 0.1]], np.float32))
    return Workload(asset, settings_config1(), 1, head_yaw, capsules=caps)


def config2(num_guides=20000, V=32, F=4, grid_n=128, cell=0.0025, subdivisions=5, seed_instance=0):
    rng = rng_for(2, seed_instance)
    roots, nrm = hemisphere_roots(num_guides, 0.1, rng)
    strands = grow_strands(roots, nrm, V, 0.005, 0.15, rng)
    asset = build_asset(strands, F, rng, follow_sigma=0.002)
    mesh = NoisySphereMesh(rng, radius=0.1, subdivisions=subdivisions, amp=0.01, anim_hz=0.5)
    half = 0.5 * grid_n * cell
    grid = ((-half, -half, -half), cell, (grid_n, grid_n, grid_n))
    return Workload(asset, settings_config2(), 2, head_yaw, mesh=mesh, sdf_grid=grid, wind_fn=wind_gust)


def rigid_trajectory(rng):
    """Independent yaw and pitch sinusoids: amplitude U(10, 45) degrees,
    frequency U(0.3, 2) Hz, random phase."""
    amp = np.radians(rng.uniform(10.0, 45.0, 2))
    hz = rng.uniform(0.3, 2.0, 2)
    ph = rng.uniform(0.0, 2.0 * math.pi, 2)

    def bone(t):
        return yaw_pitch_matrix(amp[0] * math.sin(2.0 * math.pi * hz[0] * t + ph[0]),
                                amp[1] * math.sin(2.0 * math.pi * hz[1] * t + ph[1]))
    bone.params = (amp, hz, ph)
    return bone


def yaw_pitch_matrices(yaw, pitch):
    """yaw_pitch_matrix over arrays of angles: float32[T, 16]."""
    yaw, pitch = np.asarray(yaw, np.float64), np.asarray(pitch, np.float64)
    cy, sy, cp, sp = np.cos(yaw), np.sin(yaw), np.cos(pitch), np.sin(pitch)
    z = np.zeros_like(cy)
    # r = R_y R_x (column-vector); the row-vector matrix holds r transposed
    r = np.stack([np.stack([cy, sy * sp, sy * cp], -1), np.stack([z, cp, -sp], -1),
                  np.stack([-sy, cy * sp, cy * cp], -1)], -2)
    m = np.zeros(yaw.shape + (4, 4))
    m[..., :3, :3] = np.swapaxes(r, -1, -2)
    m[..., 3, 3] = 1.0
    return m.reshape(yaw.shape + (16,)).astype(np.float32)


def trajectory_bones(bone_fn, times):
    """bone_fn at every time of `times`: float32[..., 16] for one bone,
    float32[..., nb, 16] for several; vectorised for rigid_trajectory and
    FurBody."""
    times = np.asarray(times, np.float64)
    params = getattr(bone_fn, "params", None)
    if params is not None:
        amp, hz, ph = params
        return yaw_pitch_matrices(amp[0] * np.sin(2.0 * np.pi * hz[0] * times + ph[0]),
                                  amp[1] * np.sin(2.0 * np.pi * hz[1] * times + ph[1]))
    vec = getattr(getattr(bone_fn, "__self__", None), "bone_matrices", None)
    if vec is not None:
        return vec(times)
    flat = np.stack([np.asarray(bone_fn(float(t)), np.float32).reshape(-1, 16) for t in times.reshape(-1)])
    if flat.shape[1] == 1:
        return flat.reshape(times.shape + (16,))
    return flat.reshape(times.shape + flat.shape[1:])


CONFIG3_LENGTHS = (16, 32, 64)
CONFIG3_LENGTH_P = (0.5, 0.35, 0.15)
CONFIG3_GRIDS = (64, 80, 96, 112, 128)
CONFIG3_EXTENT = 0.32     # SDF grid edge in metres (128 cells at 2.5 mm)


def config3_spec(num_instances=256, seed=0, min_strands=500, max_strands=50000):
    """Per-instance (guide strands, vertices per strand, SDF cells per edge):
    strands log-uniform in [min_strands, max_strands], lengths from
    {16, 32, 64} with p = {0.5, 0.35, 0.15}, SDF 64^3..128^3."""
    rng = rng_for(3, 100000 + seed)
    G = np.exp(rng.uniform(math.log(min_strands), math.log(max_strands), num_instances)).round().astype(int)
    V = rng.choice(CONFIG3_LENGTHS, size=num_instances, p=CONFIG3_LENGTH_P)
    n = rng.choice(CONFIG3_GRIDS, size=num_instances)
    return [(int(g), int(v), int(c)) for g, v, c in zip(G, V, n)]


def _c3_rng(i, seed, stream):
    return np.random.Generator(np.random.PCG64([3, seed, i, stream]))


def config3_asset(i, G, V, seed=0):
    """Strands of instance i of the batched workload (no follow strands)."""
    rng = _c3_rng(i, seed, 0)
    roots, nrm = hemisphere_roots(G, 0.1, rng)
    strands = grow_strands(roots, nrm, V, 0.005, 0.15, rng)
    return build_asset(strands, 0, rng)


def config3_instance(i, G, V, grid_n, seed=0, subdivisions=4, asset=None):
    """One instance of the batched workload: its own strands, rigid head
    trajectory and noisy-icosphere collider (static noise, moved rigidly);
    config-2 solver settings and wind, 2 substeps, no follow strands."""
    if asset is None:
        asset = config3_asset(i, G, V, seed)
    rng = _c3_rng(i, seed, 1)
    mesh = NoisySphereMesh(rng, radius=0.1, subdivisions=subdivisions, amp=0.01, anim_hz=0.0)
    cell = CONFIG3_EXTENT / grid_n
    half = 0.5 * CONFIG3_EXTENT
    grid = ((-half, -half, -half), cell, (grid_n, grid_n, grid_n))
    return Workload(asset, settings_config2(), 2, rigid_trajectory(rng), mesh=mesh, sdf_grid=grid,
                    wind_fn=wind_gust)


def _c3_asset_job(args):
    return config3_asset(*args)


def config3(num_instances=256, seed=0, workers=1, start_method="fork", **kw):
    """The batched workload; assets are generated by `workers` processes
    started with `start_method` ("fork" only before the process touches the
    GPU; "spawn" afterwards)."""
    spec = config3_spec(num_instances, seed, **kw)
    jobs = [(i, G, V, seed) for i, (G, V, n) in enumerate(spec)]
    if workers > 1:
        import multiprocessing as mp
        with mp.get_context(start_method).Pool(workers) as pool:
            assets = pool.map(_c3_asset_job, jobs, chunksize=1)
    else:
        assets = [config3_asset(*j) for j in jobs]
    return [config3_instance(i, G, V, n, seed, asset=a) for (i, (G, V, n)), a in zip(enumerate(spec), assets)]


# ---------------------------------------------------------------------------
# Config 5: fur on a skinned body
# ---------------------------------------------------------------------------
FUR_SEMI_AXES = (0.25, 0.15, 0.15)                  # a 0.5 x 0.3 x 0.3 m body
FUR_BONE_X = (-0.1875, -0.0625, 0.0625, 0.1875)     # bone pivots along the body's long axis
FUR_BONE_SPACING = 0.125


def fur_skin_weights(x):
    """Skinning of a point at body coordinate x to the 4 bones: hat functions
    centred on the bone pivots (a partition of unity, x clamped to the end
    pivots).  float32[n, 8]: bone ids as float, then weights."""
    xc = np.clip(np.asarray(x, np.float64), FUR_BONE_X[0], FUR_BONE_X[-1])
    w = np.maximum(0.0, 1.0 - np.abs(xc[:, None] - np.asarray(FUR_BONE_X)[None, :]) / FUR_BONE_SPACING)
    out = np.zeros((xc.shape[0], 8), np.float32)
    out[:, :4] = np.arange(4, dtype=np.float32)
    out[:, 4:] = w
    return out


class FurBody:
    """The config-5 collision body: a closed geodesic sphere of 20 freq^2
    triangles (freq 100: 200,000) scaled to semi-axes 0.25 x 0.15 x 0.15 m and
    displaced radially by a static 3-octave value noise of amplitude `amp`,
    skinned to 4 bones spaced along x.  Bone b bends the body about its pivot
    (FUR_BONE_X[b], 0, 0) with independent yaw and pitch sinusoids (amplitude
    U(5, 20) degrees, frequency U(0.3, 2) Hz, random phase)."""

    def __init__(self, rng, freq=100, amp=0.01, noise_freq=6.0):
        unit, self.indices = geodesic_sphere(freq)
        noise = ValueNoise3(rng)
        disp = np.zeros(unit.shape[0])
        norm = 0.0
        for o in range(3):
            a = 0.5 ** o
            disp += a * noise(unit * noise_freq * (2 ** o) + 17.0 * o)
            norm += a
        p = unit * np.asarray(FUR_SEMI_AXES)[None, :] + (amp * disp / norm)[:, None] * unit
        tri = self.indices.astype(np.int64)
        fn = np.cross(p[tri[:, 1]] - p[tri[:, 0]], p[tri[:, 2]] - p[tri[:, 0]])
        vn = np.zeros_like(p)
        for q in range(3):
            np.add.at(vn, tri[:, q], fn)
        vn /= np.linalg.norm(vn, axis=1, keepdims=True)
        self.positions = p.astype(np.float32)
        self.normals = vn.astype(np.float32)
        self.face_normals = fn / np.linalg.norm(fn, axis=1, keepdims=True)
        self.face_areas = 0.5 * np.linalg.norm(fn, axis=1)
        self.skinning = fur_skin_weights(p[:, 0])
        self.amp = np.radians(rng.uniform(5.0, 20.0, (4, 2)))
        self.hz = rng.uniform(0.3, 2.0, (4, 2))
        self.ph = rng.uniform(0.0, 2.0 * math.pi, (4, 2))

    @property
    def num_triangles(self):
        return self.indices.shape[0]

    def bone_matrices(self, times):
        """float32[..., 4, 16] at every time of `times`: bone b is the
        rotation R_y(yaw) R_x(pitch) about its pivot (row-vector convention,
        p' = (p - c) R + c)."""
        t = np.asarray(times, np.float64)[..., None]
        yaw = self.amp[:, 0] * np.sin(2.0 * np.pi * self.hz[:, 0] * t + self.ph[:, 0])
        pitch = self.amp[:, 1] * np.sin(2.0 * np.pi * self.hz[:, 1] * t + self.ph[:, 1])
        m = yaw_pitch_matrices(yaw, pitch).astype(np.float64).reshape(t.shape[:-1] + (4, 4, 4))
        c = np.zeros((4, 3))
        c[:, 0] = FUR_BONE_X
        r = m[..., :3, :3]
        m[..., 3, :3] = c - np.einsum("bi,...bij->...bj", c, r)
        return m.reshape(t.shape[:-1] + (4, 16)).astype(np.float32)

    def bone_fn(self, t):
        return self.bone_matrices(t)

    def skinned_mesh(self):
        """The rest mesh with its skinning, as TressFXBoneSkinning takes it."""
        return dict(positions=self.positions, normals=self.normals, skinning=self.skinning, indices=self.indices,
                    follow_bone=0)

    def sample_surface(self, n, rng):
        """n area-uniform points on the rest surface with their face normals."""
        tri = rng.choice(self.indices.shape[0], size=n, p=self.face_areas / self.face_areas.sum())
        r1, r2 = rng.uniform(0.0, 1.0, n), rng.uniform(0.0, 1.0, n)
        s1 = np.sqrt(r1)
        w = np.stack([1.0 - s1, s1 * (1.0 - r2), s1 * r2], 1)
        p = self.positions.astype(np.float64)[self.indices[tri].astype(np.int64)]
        return np.einsum("nk,nkj->nj", w, p), self.face_normals[tri]


def settings_config5():
    """Config-1 solver settings with length iterations 2 and VSP 0.7."""
    s = settings_config1()
    s.length_constraints_iterations = 2
    s.vsp_coeff = 0.7
    return s


def config5(num_guides=1_000_000, V=8, F=2, freq=100, grid_n=256, cell=0.0025, follow_sigma=0.0005,
            seed_instance=0, body=None):
    """Fur stress: num_guides x V strands (0.002 m segments, the same curl
    generation) rooted area-uniformly on the 4-bone skinned FurBody, F follow
    strands per guide (offsets N(0, follow_sigma)), each guide skinned with the
    body's weights at its root; 256^3 SDF of 2.5 mm cells centred on the body
    rebuilt every frame from the skinned mesh; 3 substeps."""
    rng = rng_for(5, seed_instance)
    if body is None:
        body = FurBody(rng, freq)
    roots, nrm = body.sample_surface(num_guides, rng)
    strands = grow_strands(roots, nrm, V, 0.002, 0.15, rng)
    skin = np.repeat(fur_skin_weights(roots[:, 0]), F + 1, axis=0)
    asset = build_asset(strands, F, rng, follow_sigma=follow_sigma, skin=skin)
    half = 0.5 * grid_n * cell
    grid = ((-half, -half, -half), cell, (grid_n, grid_n, grid_n))
    return Workload(asset, settings_config5(), 3, body.bone_fn, mesh=body, sdf_grid=grid)
