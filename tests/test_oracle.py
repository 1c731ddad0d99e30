"""Pins of the CPU restatement (oracle) against closed forms, invariants and
independent re-derivations.  The reference has no tests, golden vectors or
fixtures for this path (it is HLSL with no CPU build), so each oracle function
is checked against what the mathematics fixes, chosen so that a dropped term,
a wrong sign or index, or a transposed operand fails."""
import ctypes
import math

import numpy as np
import pytest

import oracle
import tressfx_amd as tfx
from tressfx_amd import synth

rng0 = np.random.default_rng(1234)


# ---------------------------------------------------------------------------
# helpers
# ---------------------------------------------------------------------------
def make_params(dt=1.0 / 60, gravity=0.0, damping=0.0, k_local=0.0, k_global=0.0, g_range=0.0, length_iters=0,
                vsp=0.0, vsp_thr=1e9, reset=0.0, clamp=20.0, wind=None, capsules=None, V=4, F=0, lsc=1):
    p = tfx.SimulationParams()
    p.shape[0], p.shape[1], p.shape[2], p.shape[3] = damping, k_local, k_global, g_range
    p.grav_time_tip[0], p.grav_time_tip[1] = gravity, dt
    p.sim_ints[0] = length_iters
    p.counts[0], p.counts[1], p.counts[2] = 64 // V, F, V
    p.vsp[0], p.vsp[1] = vsp, vsp_thr
    p.reset_positions = reset
    p.clamp_position_delta = clamp
    p.cpu_local_shape_iterations = lsc
    if wind is not None:
        for k, w in enumerate((p.wind, p.wind1, p.wind2, p.wind3)):
            for c in range(3):
                w[c] = wind[k][c]
    if capsules is not None:
        c0, c1 = capsules
        p.num_collision_capsules[0] = len(c0)
        for i in range(len(c0)):
            for c in range(4):
                p.center_and_radius0[i][c] = c0[i][c]
                p.center_and_radius1[i][c] = c1[i][c]
    return p


def identity_bones():
    return np.tile(np.eye(4, dtype=np.float32).reshape(1, 16), (512, 1))


def make_asset(strands, F=0, w=None, offsets=None, skin=None):
    """strands: [G, V, 3] guide positions (follow strands are copies)."""
    strands = np.asarray(strands, np.float64)
    G, V, _ = strands.shape
    S = G * (F + 1)
    pos = np.zeros((G, F + 1, V, 4), np.float32)
    pos[:, :, :, :3] = strands[:, None].astype(np.float32)
    ww = np.ones(V, np.float32) if w is None else np.asarray(w, np.float32)
    pos[..., 3] = ww
    pos = pos.reshape(S * V, 4)
    d = np.diff(pos.reshape(S, V, 4)[..., :3], axis=1)
    rl = np.zeros((S, V), np.float32)
    rl[:, :-1] = np.linalg.norm(d, axis=2)
    off = np.zeros((S, 4), np.float32) if offsets is None else np.asarray(offsets, np.float32)
    if skin is None:
        skin = np.zeros((S, 8), np.float32)
        skin[:, 4] = 1.0
    return tfx.HairAsset(G, F, V, pos, np.zeros_like(pos), rl.reshape(-1), off, skin)


def rot_matrix_rowvec(R, t=(0, 0, 0), dtype=np.float32):
    m = np.eye(4)
    m[:3, :3] = np.asarray(R).T
    m[3, :3] = t
    return m.astype(dtype)


def random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def quat_to_matrix(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def straight_strands(G=2, V=4, seg=0.01, direction=(0, 1, 0), origin=(0, 0, 0)):
    d = np.asarray(direction, float)
    d /= np.linalg.norm(d)
    s = np.zeros((G, V, 3))
    for g in range(G):
        for v in range(V):
            s[g, v] = np.asarray(origin) + np.array([0.1 * g, 0, 0]) + v * seg * d
    return s


# ---------------------------------------------------------------------------
# quaternion helpers
# ---------------------------------------------------------------------------
def test_make_quaternion_matches_rotation_all_branches():
    lib = oracle.load()
    rots = [random_rotation(rng0) for _ in range(200)]
    # trace <= 0 branches: rotations by ~180 degrees about each axis
    for ax in np.eye(3):
        c, s = math.cos(math.pi - 0.01), math.sin(math.pi - 0.01)
        K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
        rots.append(np.eye(3) + s * K + (1 - c) * K @ K)
    for R in rots:
        m = rot_matrix_rowvec(R, dtype=np.float64)
        q = np.zeros(4)
        lib.orc64_make_quaternion(m.ctypes.data, q.ctypes.data)
        assert abs(np.linalg.norm(q) - 1) < 1e-9
        assert np.allclose(quat_to_matrix(q), R, atol=1e-9)


def test_quat_from_two_unit_vectors_rotates_u_onto_v():
    lib = oracle.load()
    for _ in range(200):
        u = rng0.normal(size=3); u /= np.linalg.norm(u)
        v = rng0.normal(size=3); v /= np.linalg.norm(v)
        q = np.zeros(4)
        lib.orc64_quat_from_two_unit_vectors(u.ctypes.data, v.ctypes.data, q.ctypes.data)
        out = np.zeros(3)
        lib.orc64_mult_quat_vec(q.ctypes.data, u.ctypes.data, out.ctypes.data)
        assert np.allclose(out, v, atol=1e-12)
        # and it is the shortest arc: the axis is perpendicular to both
        assert abs(np.dot(q[:3], u)) < 1e-12 and abs(np.dot(q[:3], v)) < 1e-12
    # antiparallel branch: a half turn about an axis perpendicular to u
    for u in (np.array([1.0, 0, 0]), np.array([0, 0, 1.0]), np.array([0.6, 0.8, 0])):
        q = np.zeros(4)
        lib.orc64_quat_from_two_unit_vectors(u.ctypes.data, (-u).copy().ctypes.data, q.ctypes.data)
        assert abs(q[3]) < 1e-12 and abs(np.dot(q[:3], u)) < 1e-12
        out = np.zeros(3)
        lib.orc64_mult_quat_vec(q.ctypes.data, u.ctypes.data, out.ctypes.data)
        assert np.allclose(out, -u, atol=1e-12)


def test_mult_quat_vec_is_rotation():
    lib = oracle.load()
    for _ in range(50):
        ax = rng0.normal(size=3); ax /= np.linalg.norm(ax)
        ang = rng0.uniform(-3, 3)
        q = np.r_[math.sin(ang / 2) * ax, math.cos(ang / 2)]
        v = rng0.normal(size=3)
        out = np.zeros(3)
        lib.orc64_mult_quat_vec(q.ctypes.data, v.ctypes.data, out.ctypes.data)
        K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
        R = np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * K @ K
        assert np.allclose(out, R @ v, atol=1e-12)


# ---------------------------------------------------------------------------
# IntegrationAndGlobalShapeConstraints
# ---------------------------------------------------------------------------
def test_integrate_free_fall_closed_form():
    """decay 1, gravity g, start at rest: x_n = x_0 - g dt^2 n(n+1)/2."""
    a = make_asset(straight_strands(3, 8), w=[0, 0, 1, 1, 1, 1, 1, 1])
    h = oracle.OracleHair(a, np.float64)
    g, dt = 9.8, 1.0 / 60
    p = make_params(dt=dt, gravity=g, V=8)
    x0 = h.pos.copy()
    for n in range(1, 6):
        h.run_stage(p, identity_bones(), 0)
        exp = x0.copy()
        mov = x0[:, 3] > 0
        exp[mov, 1] -= g * dt * dt * n * (n + 1) / 2
        assert np.allclose(h.pos[:, :3], exp[:, :3], atol=1e-12), n
        # buffer rotation: Prev is the position this step started from
    assert np.allclose(h.prevprev[:, :3], h.prev[:, :3] + (h.prev[:, :3] - h.pos[:, :3]) * 0 + np.array([0, g * dt * dt * 4, 0]) * (x0[:, 3:4] > 0), atol=1e-12)


def test_integrate_damping_geometric():
    """Velocity decays by exp(-60 k dt) per step without gravity."""
    a = make_asset(straight_strands(2, 4), w=[0, 0, 1, 1])
    h = oracle.OracleHair(a, np.float64)
    k, dt = 0.3, 1.0 / 60
    vel = np.array([0.001, -0.002, 0.0005])
    st = h.state()
    st["prev"] = st["pos"].copy()
    st["prev"][:, :3] -= vel * (st["pos"][:, 3:4] > 0)
    h.set_state(st)
    x0 = h.pos[:, :3].copy()
    p = make_params(dt=dt, damping=k, V=4)
    h.run_stage(p, identity_bones(), 0)
    decay = math.exp(-k * dt * 60)
    exp = x0 + decay * vel * (h.pos[:, 3:4] > 0)
    assert np.allclose(h.pos[:, :3], exp, atol=1e-13)


def test_skinning_rigid_transform_and_pinned_vertices():
    """One bone, weight 1: immovable vertices equal R p + t exactly; reset puts
    every vertex on the skinned rest pose."""
    a = make_asset(straight_strands(3, 8), w=[0, 0, 1, 1, 1, 1, 1, 1])
    h = oracle.OracleHair(a, np.float64)
    R = random_rotation(rng0)
    t = np.array([0.1, -0.2, 0.3])
    bones = identity_bones()
    bones[0] = rot_matrix_rowvec(R, t).reshape(16)
    p = make_params(gravity=9.8, V=8, reset=1.0)
    h.run_stage(p, bones, 0)
    rest = a.positions[:, :3].astype(np.float64)
    skinned = rest @ R.T + t
    Rf = bones[0].reshape(4, 4).astype(np.float64)
    skinned_f = rest @ Rf[:3, :3] + Rf[3, :3]
    pinned = a.positions[:, 3] == 0
    assert np.allclose(h.pos[pinned, :3], skinned[pinned], atol=1e-6)
    assert np.allclose(h.pos[pinned, :3], skinned_f[pinned], atol=1e-14)
    # reset: Prev = skinned rest for all vertices
    assert np.allclose(h.prev[:, :3], skinned_f, atol=1e-14)


def test_skinning_weight_blend_four_bones():
    """Weights blend matrices element-wise and are renormalised by their sum."""
    a = make_asset(straight_strands(1, 4), w=[0, 0, 0, 0],
                   skin=np.array([[3, 7, 11, 2, 0.5, 1.0, 0.5, 0.0]], np.float32))
    h = oracle.OracleHair(a, np.float64)
    bones = identity_bones()
    mats = {}
    for b in (3, 7, 11, 2):
        mats[b] = rot_matrix_rowvec(random_rotation(rng0), rng0.normal(size=3))
        bones[b] = mats[b].reshape(16)
    p = make_params(V=4, reset=1.0)
    h.run_stage(p, bones, 0)
    M = (0.5 * mats[3].astype(np.float64) + 1.0 * mats[7] + 0.5 * mats[11]) / 2.0  # weight 0 bone skipped
    rest = np.c_[a.positions[:, :3].astype(np.float64), np.ones(4)]
    assert np.allclose(h.pos[:, :3], (rest @ M)[:, :3], atol=1e-12)


def test_global_shape_snap_and_range_cutoff():
    a = make_asset(straight_strands(2, 16), w=[0, 0] + [1] * 14)
    h = oracle.OracleHair(a, np.float64)
    st = h.state()
    st["pos"][:, :3] += 0.01 * rng0.normal(size=(st["pos"].shape[0], 3)) * (st["pos"][:, 3:4] > 0)
    st["prev"] = st["pos"].copy()
    h.set_state(st)
    before = h.pos.copy()
    p = make_params(V=16, k_global=1.0, g_range=0.5)
    h.run_stage(p, identity_bones(), 0)
    v = np.tile(np.arange(16), 2)
    inside = v < 0.5 * 16
    rest = a.positions[:, :3]
    assert np.allclose(h.pos[inside, :3], rest[inside], atol=1e-15)
    outside = ~inside
    assert np.allclose(h.pos[outside, :3], before[outside, :3], atol=1e-15)


# ---------------------------------------------------------------------------
# CalculateStrandLevelData + VelocityShockPropagation
# ---------------------------------------------------------------------------
def test_vsp_full_coefficient_is_rigid_motion():
    """With coefficient 1 every vertex v >= 2 receives the rigid transform that
    maps old vertices 0,1 onto new ones (VSP :757-758)."""
    a = make_asset(straight_strands(2, 8, direction=(0.2, 1, 0.1)), w=[0, 0] + [1] * 6)
    h = oracle.OracleHair(a, np.float64)
    R = random_rotation(rng0)
    t = np.array([0.01, 0.02, -0.03])
    bones = identity_bones()
    bones[0] = rot_matrix_rowvec(R, t).reshape(16)
    p = make_params(V=8, vsp=1.0)
    h.run_stage(p, bones, 0)   # vertices 0,1 jump to the skinned pose, others integrate (at rest)
    old = h.prev.copy()
    mid = h.pos.copy()
    h.run_stage(p, bones, 1)
    h.run_stage(p, bones, 2)
    Rf = bones[0].reshape(4, 4).astype(np.float64)
    for g in range(2):
        s = slice(g * 8 + 2, g * 8 + 8)
        # old 0,1 -> new 0,1 is the bone transform composed with the strand's
        # prior pose: the rest pose (old) maps by R; check the vertices moved rigidly
        exp = mid[s, :3] @ Rf[:3, :3] + Rf[3, :3]
        assert np.allclose(h.pos[s, :3], exp, atol=1e-6)
    # coefficient 0 leaves positions untouched
    h2 = oracle.OracleHair(a, np.float64)
    p0 = make_params(V=8, vsp=0.0)
    h2.run_stage(p0, bones, 0)
    b2 = h2.pos.copy()
    h2.run_stage(p0, bones, 1)
    h2.run_stage(p0, bones, 2)
    assert np.array_equal(h2.pos, b2)


def test_vsp_accel_threshold_forces_coefficient_one():
    a = make_asset(straight_strands(1, 4), w=[0, 0, 1, 1])
    h = oracle.OracleHair(a, np.float64)
    st = h.state()
    st["pos"][1, :3] += np.array([0.0, 0.0, 2.0])   # pseudo-acceleration |x1 - 2 x1' + x1''| = 2
    h.set_state(st)
    h.run_stage(make_params(V=4, vsp=0.3, vsp_thr=1.5), identity_bones(), 1)
    assert h.strand_level[0, 11] == 1.0
    h.run_stage(make_params(V=4, vsp=0.3, vsp_thr=2.5), identity_bones(), 1)
    assert abs(h.strand_level[0, 11] - np.float64(np.float32(0.3))) < 1e-12


# ---------------------------------------------------------------------------
# LocalShapeConstraints
# ---------------------------------------------------------------------------
def test_lsc_fixed_point_and_zero_stiffness():
    strands = synth.grow_strands(*synth.hemisphere_roots(4, 0.1, rng0), 16, 0.005, 0.3, rng0)
    a = make_asset(strands, w=[0, 0] + [1] * 14)
    R = random_rotation(rng0)
    bones = identity_bones()
    bones[0] = rot_matrix_rowvec(R, (0.1, 0, 0)).reshape(16)
    h = oracle.OracleHair(a, np.float64)
    p = make_params(V=16, k_local=0.9, reset=1.0)
    h.run_stage(p, bones, 0)   # everything on the skinned rest pose
    h.run_stage(p, bones, 1)
    before = h.pos.copy()
    h.run_stage(p, bones, 3)
    assert np.allclose(h.pos, before, atol=1e-14)
    # stiffness 0: a perturbed strand does not move
    st = h.state()
    st["pos"][:, :3] += 0.002 * rng0.normal(size=(st["pos"].shape[0], 3)) * (st["pos"][:, 3:4] > 0)
    h.set_state(st)
    b2 = h.pos.copy()
    h.run_stage(make_params(V=16, k_local=0.0), bones, 3)
    assert np.array_equal(h.pos, b2)


def _shortest_arc(a, b):
    """Rodrigues rotation matrix taking unit a onto unit b (not antiparallel)."""
    v = np.cross(a, b)
    c = np.dot(a, b)
    K = np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])
    return np.eye(3) + K + K @ K / (1 + c)


def test_lsc_gauss_seidel_by_hand_V4():
    """V = 4: two sequential steps, re-derived with rotation matrices."""
    b = np.array([[0, 0, 0], [0, 0.01, 0], [0.002, 0.02, 0], [0.006, 0.029, 0.001]], float)
    a = make_asset(b[None], w=[0, 1, 1, 1])
    h = oracle.OracleHair(a, np.float64)
    p_ = b + np.array([[0, 0, 0], [0.001, 0, 0.0005], [-0.002, 0.001, 0.001], [0.003, -0.002, 0.0]])
    st = h.state()
    st["pos"][:, :3] = p_
    h.set_state(st)
    h.strand_level[0, :4] = [0, 0, 0, 1]
    k = 0.8
    h.run_stage(make_params(V=4, k_local=k), identity_bones(), 3)
    s = 0.5 * min(np.float32(k), np.float32(0.95))
    P = p_.copy()
    for i in (1, 2):
        last = P[i] - P[i - 1]
        lastb = b[i] - b[i - 1]
        Rg = _shortest_arc(lastb / np.linalg.norm(lastb), last / np.linalg.norm(last))
        target = Rg @ (b[i + 1] - b[i]) + P[i]
        d = s * (target - P[i + 1])
        P[i] -= d      # both movable
        P[i + 1] += d
    assert np.allclose(h.pos[:, :3], P, atol=1e-13)


# ---------------------------------------------------------------------------
# LengthConstriantsWindAndCollision
# ---------------------------------------------------------------------------
def test_length_single_pair_pinned_end_projects_exactly():
    a = make_asset(np.array([[[0, 0, 0], [0, 0.01, 0], [0, 0.02, 0], [0, 0.03, 0]]]), w=[0, 1, 1, 1])
    h = oracle.OracleHair(a, np.float64)
    st = h.state()
    st["pos"][1, :3] = [0.003, 0.02, -0.001]
    h.set_state(st)
    h.run_stage(make_params(V=4, length_iters=1), identity_bones(), 4)
    # pass A moves only vertex 1 (vertex 0 pinned) onto |x1 - x0| = L
    L = np.float64(np.float32(0.01))
    # then pass B (1,2) may move it again; check pass A alone via the closed form on a V=2-like setup
    d = np.array([0.003, 0.02, -0.001])
    x1 = d / np.linalg.norm(d) * L
    # pass B pair (1,2): both movable, each moves half of the stretch
    x2 = np.array([0, 0.02, 0])
    delta = x2 - x1
    dist = np.linalg.norm(delta)
    corr = (1 - np.float64(np.float32(0.01)) / dist) * delta
    exp1 = x1 + 0.5 * corr
    exp2 = x2 - 0.5 * corr
    assert np.allclose(h.pos[1, :3], exp1, atol=1e-14)
    assert np.allclose(h.pos[2, :3], exp2, atol=1e-14)


def test_length_converges_to_rest_lengths():
    strands = straight_strands(2, 16, seg=0.005)
    a = make_asset(strands, w=[0, 0] + [1] * 14)
    h = oracle.OracleHair(a, np.float64)
    st = h.state()
    st["pos"][:, :3] += 0.003 * rng0.normal(size=(32, 3)) * (st["pos"][:, 3:4] > 0)
    h.set_state(st)
    h.run_stage(make_params(V=16, length_iters=4000), identity_bones(), 4)
    seg = np.linalg.norm(np.diff(h.pos.reshape(2, 16, 4)[..., :3], axis=1), axis=2)
    assert np.allclose(seg, np.float64(np.float32(0.005)), atol=1e-9)


def test_wind_closed_form_and_strand_blend():
    """force = -(d x w) x d = -|d|^2 w_perp, blend a = (strand % 20)/20."""
    G, V = 3, 6
    strands = np.zeros((G, V, 3))
    strands[:, :, 0] = np.arange(V) * 0.01  # along x
    a = make_asset(strands, w=[0, 0] + [1] * 4)
    h = oracle.OracleHair(a, np.float64)
    W = [np.array([0, 1.0, 0]), np.array([0, 0, 2.0]), np.array([1.0, 1.0, 0]), np.array([0, 0, 0.0])]
    dt = 0.1
    h.run_stage(make_params(V=V, dt=dt, wind=W), identity_bones(), 4)
    for g in range(G):
        aa = (g % 20) / 20.0
        w = aa * W[0] + (1 - aa) * W[1] + aa * W[2] + (1 - aa) * W[3]
        for v in range(V):
            x = strands[g, v].copy()
            if 2 <= v <= V - 2:
                d = np.array([-0.01, 0, 0])
                wperp = w - np.dot(w, d) / np.dot(d, d) * d
                x = x - np.dot(d, d) * wperp * dt * dt
            assert np.allclose(h.pos[g * V + v, :3], x, atol=1e-15), (g, v)


def test_tangent_unit_and_clamp_formula():
    a = make_asset(straight_strands(1, 8, direction=(1, 2, 3)), w=[0, 0] + [1] * 6)
    h = oracle.OracleHair(a, np.float64)
    st = h.state()
    st["prev"][5, :3] = st["pos"][5, :3] - np.array([30.0, 0, 40.0])  # |delta| = 50 > C = 20
    h.set_state(st)
    h.run_stage(make_params(V=8, clamp=20.0), identity_bones(), 4)
    t = h.tangent[:, :3]
    assert np.allclose(np.linalg.norm(t, axis=1), 1.0, atol=1e-15)
    d = np.array([1, 2, 3]) / math.sqrt(14)
    assert np.allclose(t, d, atol=1e-12)
    # clamp: delta scaled by C^2/|delta|^2 (not C/|delta|)
    delta = np.array([30.0, 0, 40.0]) * (400.0 / 2500.0)
    assert np.allclose(h.prev[5, :3], h.pos[5, :3] - delta, atol=1e-12)


def test_capsule_sphere_projection_and_prev_reset():
    lib = oracle.load()
    c0 = np.array([0.0, 0.0, 0.0, 0.1])
    c1 = np.array([0.0, 1e-3, 0.0, 0.1])
    for _ in range(100):
        d = rng0.normal(size=3)
        d /= np.linalg.norm(d)
        if d[1] > 0:
            d[1] = -d[1]  # below the first cap: sphere-end region
        r = rng0.uniform(0.01, 0.2)
        cur = np.r_[r * d, 1.0]
        out = np.zeros(3)
        hit = lib.orc64_capsule_collision(cur.ctypes.data, cur.ctypes.data, c0.ctypes.data, c1.ctypes.data,
                                          out.ctypes.data)
        if r < 0.1:
            assert hit == 1 and np.allclose(out, 0.1 * d, atol=1e-14)
        else:
            assert hit == 0 and np.allclose(out, cur[:3])
    # pinned vertices are never moved
    cur = np.array([0.01, 0, 0, 0.0])
    out = np.zeros(3)
    assert lib.orc64_capsule_collision(cur.ctypes.data, cur.ctypes.data, c0.ctypes.data, c1.ctypes.data,
                                       out.ctypes.data) == 0
    # a hit sets Prev = Pos in the stage
    a = make_asset(straight_strands(1, 4, seg=0.01, origin=(0, -0.05, 0), direction=(1, 0, 0)), w=[0, 0, 1, 1])
    h = oracle.OracleHair(a, np.float64)
    h.run_stage(make_params(V=4, capsules=([c0], [c1])), identity_bones(), 4)
    assert np.array_equal(h.prev[2:], h.pos[2:])
    assert np.allclose(np.linalg.norm(h.pos[2:, :3], axis=1), 0.1, atol=1e-14)


def test_capsule_second_cap_and_cylinder_friction_closed_forms():
    """CapsuleCollision (TressFXSimulation.hlsl:462-522) on a tapered capsule
    (r0 = 0.05 at p0, r1 = 0.02 at p1) placed by a random rigid motion, in its
    own frame (axis = local y from 0 to L = 0.2):
    - beyond p1 (dist1 < 0, :492-502): inside r1 the point goes to
      p1 + r1 * (p - p1) / |p - p1|; outside r1 (but inside r0) nothing moves;
    - between the caps (:504-519): with x = the axis point at the same height
      y and r(y) = r0 + (r1 - r0) y / L, a point inside r(y) ends at radius
      r(y) in its own radial direction, at height old_y + 0.4 (cur_y - old_y)
      (the axial motion since the previous position kept at friction 0.4, the
      radial motion replaced by the projection); outside r(y) nothing moves."""
    lib = oracle.load()
    rng = np.random.default_rng(77)
    L, r0, r1 = 0.2, 0.05, 0.02
    for trial in range(40):
        R = random_rotation(rng)
        t = rng.uniform(-0.5, 0.5, 3)
        to_world = lambda q: R @ np.asarray(q, float) + t
        c0 = np.r_[to_world([0, 0, 0]), r0]
        c1 = np.r_[to_world([0, L, 0]), r1]

        def collide(cur_local, old_local):
            cur = np.r_[to_world(cur_local), 1.0]
            old = np.r_[to_world(old_local), 1.0]
            out = np.zeros(3)
            hit = lib.orc64_capsule_collision(cur.ctypes.data, old.ctypes.data, c0.ctypes.data, c1.ctypes.data,
                                              out.ctypes.data)
            return hit, R.T @ (out - t)

        # second cap
        d = rng.normal(size=3)
        d[1] = abs(d[1]) + 0.2
        d /= np.linalg.norm(d)
        r = rng.uniform(0.001, 0.019)
        hit, out = collide(np.array([0, L, 0]) + r * d, [0, 0.1, 0])
        assert hit == 1 and np.allclose(out, np.array([0, L, 0]) + r1 * d, atol=1e-12), trial
        hit, out = collide(np.array([0, L, 0]) + 0.03 * d, [0, 0.1, 0])
        assert hit == 0 and np.allclose(out, np.array([0, L, 0]) + 0.03 * d, atol=1e-12), trial
        # middle cylinder with friction
        y = rng.uniform(0.01, L - 0.01)
        ry = r0 + (r1 - r0) * y / L
        phi = rng.uniform(0, 2 * np.pi)
        radial = np.array([np.cos(phi), 0.0, np.sin(phi)])
        cur = np.array([0, y, 0]) + rng.uniform(0.05, 0.95) * ry * radial
        old = cur + rng.normal(scale=0.01, size=3)
        hit, out = collide(cur, old)
        want = np.array([ry * radial[0], old[1] + 0.4 * (cur[1] - old[1]), ry * radial[2]])
        assert hit == 1 and np.allclose(out, want, atol=1e-12), (trial, out, want)
        far = np.array([0, y, 0]) + 1.05 * ry * radial
        hit, out = collide(far, old)
        assert hit == 0 and np.allclose(out, far, atol=1e-12), trial


def test_follow_closed_form():
    G, F, V = 2, 3, 8
    strands = straight_strands(G, V)
    offs = np.zeros((G * (F + 1), 4), np.float32)
    offs[:, :3] = rng0.normal(size=(G * (F + 1), 3)) * 0.01
    a = make_asset(strands, F=F, offsets=offs)
    h = oracle.OracleHair(a, np.float64)
    st = h.state()
    st["pos"][:, :3] += rng0.normal(size=st["pos"][:, :3].shape) * 0.001
    st["tangent"][:, :3] = rng0.normal(size=st["pos"][:, :3].shape)
    h.set_state(st)
    tip = 0.75
    p = make_params(V=V, F=F)
    p.grav_time_tip[2] = tip
    h.run_stage(p, identity_bones(), 5)
    pos = h.pos.reshape(G, F + 1, V, 4)
    tan = h.tangent.reshape(G, F + 1, V, 4)
    for g in range(G):
        for j in range(1, F + 1):
            for v in range(V):
                f = np.float64(np.float32(tip)) * (v / V) + 1.0
                exp = pos[g, 0, v, :3] + f * offs[g * (F + 1) + j, :3].astype(np.float64)
                assert np.allclose(pos[g, j, v, :3], exp, atol=1e-15)
                assert np.array_equal(tan[g, j, v], tan[g, 0, v])


# ---------------------------------------------------------------------------
# SDF
# ---------------------------------------------------------------------------
def test_float_flip3_order_and_bijection():
    lib = oracle.load()
    xs = np.r_[rng0.normal(size=2000) * 10.0 ** rng0.integers(-30, 30, 2000), 0.0, -0.0, 1e10, -1e10, 5e-45, -5e-45]
    xs = xs.astype(np.float32)
    f = np.array([lib.orc_float_flip3(float(x)) for x in xs], np.uint32)
    back = np.array([lib.orc_ifloat_flip3(int(u)) for u in f], np.uint32).view(np.float32)
    assert np.array_equal(back.view(np.uint32), xs.view(np.uint32))
    # order: |x| ascending, then +x before -x
    key = sorted(range(len(xs)), key=lambda i: (abs(float(xs[i])), 0 if np.signbit(xs[i]) == 0 else 1))
    assert np.all(np.diff(f[key].astype(np.int64)) >= 0)


def _ericson_closest(p, a, b, c):
    """Real-Time Collision Detection, 5.1.5 (FP64), independent of the oracle."""
    ab, ac, ap = b - a, c - a, p - a
    d1, d2 = ab @ ap, ac @ ap
    if d1 <= 0 and d2 <= 0:
        return a
    bp = p - b
    d3, d4 = ab @ bp, ac @ bp
    if d3 >= 0 and d4 <= d3:
        return b
    vc = d1 * d4 - d3 * d2
    if vc <= 0 and d1 >= 0 and d3 <= 0:
        return a + d1 / (d1 - d3) * ab
    cp = p - c
    d5, d6 = ab @ cp, ac @ cp
    if d6 >= 0 and d5 <= d6:
        return c
    vb = d5 * d2 - d1 * d6
    if vb <= 0 and d2 >= 0 and d6 <= 0:
        return a + d2 / (d2 - d6) * ac
    va = d3 * d6 - d5 * d4
    if va <= 0 and (d4 - d3) >= 0 and (d5 - d6) >= 0:
        return b + (d4 - d3) / ((d4 - d3) + (d5 - d6)) * (c - b)
    denom = 1.0 / (va + vb + vc)
    return a + ab * (vb * denom) + ac * (vc * denom)


def test_signed_distance_point_triangle_vs_ericson():
    lib = oracle.load()
    for _ in range(3000):
        a, b, c = rng0.normal(size=(3, 3))
        p = rng0.normal(size=3) * 2
        q = _ericson_closest(p, a, b, c)
        n = np.cross(b - a, c - a)
        exp = np.linalg.norm(p - q) * (-1 if np.dot(p - a, n) < 0 else 1)
        got = lib.orc64_signed_distance_point_triangle(p.ctypes.data, a.ctypes.data, b.ctypes.data, c.ctypes.data)
        assert abs(got - exp) < 1e-9, (got, exp)


def test_sdf_build_order_independence_and_sphere():
    mesh = synth.NoisySphereMesh(np.random.default_rng(5), radius=0.1, subdivisions=3, amp=0.0)
    verts = mesh.vertices(0.0)
    idx = mesh.indices
    n, cell = 48, 0.006
    origin = np.full(3, -0.5 * n * cell, np.float32)
    g1 = oracle.sdf_build(verts, idx, origin, np.float32(cell), (n, n, n), np.float32, nthreads=4)
    perm = np.random.default_rng(9).permutation(idx.shape[0])
    g2 = oracle.sdf_build(verts, idx[perm], origin, np.float32(cell), (n, n, n), np.float32, nthreads=3)
    assert np.array_equal(g1.view(np.uint32), g2.view(np.uint32))
    # narrow band: far cells keep 1e10; near cells approximate |p| - r (faceted sphere)
    ii = np.indices((n, n, n)).reshape(3, -1)[::-1].T  # x fastest
    pts = ii * np.float32(cell) + origin
    r = np.linalg.norm(pts, axis=1)
    g = g1
    assert np.all(g[np.abs(r - 0.1) > 0.05] == np.float32(1e10))
    band = np.abs(r - 0.1) < 0.01
    assert np.all(g[band] != np.float32(1e10))
    assert np.allclose(g[band], r[band] - 0.1, atol=0.0015)


def test_collide_plane_sdf_exact():
    """A linear field is reproduced exactly by trilinear interpolation and its
    forward differences give the exact normal: a vertex at depth d below the
    plane moves to +margin along the normal; others stay put."""
    n, cell = 16, 0.01
    origin = np.zeros(3)
    ii = np.indices((n, n, n)).reshape(3, -1)[::-1].T
    pts = ii * cell + origin
    nrm = np.array([0.0, 0.6, 0.8])
    z0 = 0.07
    sdf = pts @ nrm - z0
    V = 4
    pos = np.zeros((8, 4))
    pos[:, :3] = [[0.05, 0.05, 0.05]] * 4 + [[0.06, 0.06, 0.07]] * 4
    pos[:, 3] = 1
    prev = pos.copy()
    before = pos.copy()
    margin = 0.005
    oracle.sdf_collide(pos, prev, V, sdf, origin, cell, (n, n, n), margin)
    for i in range(8):
        d = before[i, :3] @ nrm - z0
        if i % V < 2 or d >= margin:
            assert np.array_equal(pos[i], before[i]), i
        else:
            exp = before[i, :3] + nrm * (margin - d)
            assert np.allclose(pos[i, :3], exp, atol=1e-12), i
            assert np.array_equal(prev[i], pos[i])
    # outside the grid (or within 2 cells of the far faces) nothing happens
    pos2 = np.array([[0, 0, 0, 1.0]] * 2 + [[0.155, 0.155, 0.155, 1.0], [-0.5, 0.02, 0.02, 1.0]])
    p2 = pos2.copy()
    oracle.sdf_collide(pos2, pos2.copy(), 4, sdf, origin, cell, (n, n, n), 10.0)
    assert np.array_equal(pos2, p2)


def test_collide_mixed_differences_plane_exact():
    """The mixed forward/backward-difference variant (TressFXSDFCollision.hlsl
    :597-715) on a linear field: every vertex -- strand roots included, as
    that kernel has no root test -- at depth d below the plane moves to
    +margin along the normal, whether each axis takes the forward or, near the
    cell's far face, the backward difference (a sign slip in either would turn
    the normal); the step h = 0.1 * cell is added to cell-unit weights, which
    only scales the gradient.  A cube holding a 1e10 node and vertices outside
    the grid stay put."""
    n, cell = 16, 0.01
    origin = np.zeros(3)
    ii = np.indices((n, n, n)).reshape(3, -1)[::-1].T
    pts = ii * cell + origin
    nrm = np.array([0.0, 0.6, 0.8])
    z0 = 0.07
    sdf = pts @ nrm - z0
    # interpolation weights: x = 0.5 (forward), x = 0.9995 (backward), y/z likewise
    pos = np.array([[0.055, 0.055, 0.045, 0.0], [0.059995, 0.059995, 0.04, 1.0], [0.0455, 0.04, 0.059995, 1.0],
                    [0.06, 0.06, 0.07, 1.0], [0.155, 0.155, 0.155, 1.0], [-0.5, 0.02, 0.02, 1.0]])
    h = 0.1 * cell
    frac = (pos[:, :3] - origin) / cell - np.floor((pos[:, :3] - origin) / cell)
    assert np.all(frac[0] + h < 1.0)                                     # forward in every axis
    assert frac[1, 0] + h >= 1.0 and frac[1, 1] + h >= 1.0 and frac[1, 2] + h < 1.0   # backward in x and y
    assert frac[2, 2] + h >= 1.0                                         # backward in z
    prev = pos.copy()
    before = pos.copy()
    margin = 0.005
    oracle.sdf_collide(pos, prev, 4, sdf, origin, cell, (n, n, n), margin, mixed=True)
    for i in range(4):
        d = before[i, :3] @ nrm - z0
        if d >= margin:
            assert np.array_equal(pos[i], before[i]), i
        else:
            exp = before[i, :3] + nrm * (margin - d)
            assert np.allclose(pos[i, :3], exp, atol=1e-12), i
            assert np.array_equal(prev[i, :3], pos[i, :3]) and prev[i, 3] == before[i, 3], i
    moved = [i for i in range(4) if not np.array_equal(pos[i], before[i])]
    assert 0 in moved and 1 in moved and 2 in moved                       # a root (index 0) moves too
    assert np.array_equal(pos[4:], before[4:])                            # beyond n - 2 cells, outside
    # the forward variant leaves that root alone
    f = before.copy()
    oracle.sdf_collide(f, f.copy(), 4, sdf, origin, cell, (n, n, n), margin)
    assert np.array_equal(f[0], before[0]) and not np.array_equal(f[2], before[2])
    # a cube with a 1e10 node is skipped
    sdf2 = sdf.copy()
    c = (np.floor((before[1, :3] - origin) / cell)).astype(int)
    sdf2[(c[2] * n + c[1]) * n + c[0] + 1] = 1e10
    p2 = before.copy()
    oracle.sdf_collide(p2, p2.copy(), 4, sdf2, origin, cell, (n, n, n), margin, mixed=True)
    assert np.array_equal(p2[1], before[1])


def _transform_case(rng, G, F, V, T):
    """A non-indexed mesh of T triangles, a mapping of every guide to a
    random triangle and barycentric point, and rest strands rooted there."""
    mesh = np.ones((3 * T, 4))
    mesh[:, :3] = rng.normal(size=(3 * T, 3))
    S = G * (F + 1)
    mapping = (tfx.HairToTriangleMapping * S)()
    rest = np.zeros((S * V, 4))
    rest[:, 3] = 1.0
    for g in range(G):
        s = g * (F + 1)
        t = int(rng.integers(0, T))
        b = rng.dirichlet([1.0, 1.0, 1.0]).astype(np.float32)
        mapping[s].triangle_index = t
        mapping[s].barycentric[0], mapping[s].barycentric[1], mapping[s].barycentric[2] = (float(x) for x in b)
        tri = mesh[3 * t:3 * t + 3, :3]
        rest[s * V, :3] = b.astype(np.float64) @ tri
    return mesh, mapping, rest


def test_generate_transforms_closed_forms():
    """GenerateTransforms (TressFXSimulation.hlsl:1035-1142) in FP64: an
    unmoved mesh gives the identity transform and quaternion (0, 0, 0, 1);
    the mesh turned by theta about z (row-vector p R) and moved by T gives R
    as the rotation block, maps each root to its point on the moved triangle,
    and the quaternion (0, 0, -sin(theta/2), cos(theta/2)) of the diagonal
    formula; follow strands and triangles past the mesh keep zero records."""
    rng = np.random.Generator(np.random.PCG64(41))
    G, F, V, T = 20, 1, 4, 7
    mesh, mapping, rest = _transform_case(rng, G, F, V, T)
    out = oracle.generate_transforms(rest, G, F, V, mapping, mesh, mesh, np.float64)
    guides = np.arange(G) * (F + 1)
    for s in guides:
        m = out[s, :16].reshape(4, 4)
        assert np.allclose(m, np.eye(4), atol=1e-12), s
        assert np.allclose(out[s, 16:], [0, 0, 0, 1], atol=1e-12)
    assert np.all(out[np.setdiff1d(np.arange(G * (F + 1)), guides)] == 0)
    th = 0.7
    R = np.array([[np.cos(th), np.sin(th), 0], [-np.sin(th), np.cos(th), 0], [0, 0, 1]])
    Tr = np.array([0.3, -1.2, 0.5])
    moved = mesh.copy()
    moved[:, :3] = mesh[:, :3] @ R + Tr
    out = oracle.generate_transforms(rest, G, F, V, mapping, mesh, moved, np.float64)
    for s in guides:
        m = out[s, :16].reshape(4, 4)
        assert np.allclose(m[:3, :3], R, atol=1e-12), s
        root = np.append(rest[s * V, :3], 1.0)
        assert np.allclose((root @ m)[:3], rest[s * V, :3] @ R + Tr, atol=1e-12), s
        assert np.allclose(out[s, 16:], [0, 0, -np.sin(th / 2), np.cos(th / 2)], atol=1e-12), s
    mapping[0].triangle_index = T          # past the mesh
    out = oracle.generate_transforms(rest, G, F, V, mapping, mesh, moved, np.float64)
    assert np.all(out[0] == 0)


# ---------------------------------------------------------------------------
# Host parameter packing (UpdateSimulationParameters / SetWind)
# ---------------------------------------------------------------------------
def test_set_wind_corners_closed_form():
    s = tfx.SimulationSettings.defaults()
    s.wind_magnitude = 2.0
    s.wind_direction[0], s.wind_direction[1], s.wind_direction[2] = 1.0, 0.0, 0.0
    for frame in (0, 157):
        p = oracle.pack_sim_params(s, np.float32(1 / 60), frame, 16, 0)
        wM = 2.0 * (math.sin(frame * 0.01) ** 2 + 0.5)
        c, sn = math.cos(math.radians(40)), math.sin(math.radians(40))
        exp = [(c, 0, -sn), (c, 0, sn), (c, sn, 0), (c, -sn, 0)]   # X rotated by 40 deg about +Y, -Y, +Z, -Z
        for w, e in zip((p.wind, p.wind1, p.wind2, p.wind3), exp):
            assert np.allclose([w[0], w[1], w[2]], np.array(e) * wM, atol=1e-5)
            assert w[3] == 0.0
    # fields and the LSC split (TressFXHairObject.cpp:402-411)
    p16 = oracle.pack_sim_params(s, np.float32(0.01), 5, 16, 2)
    assert p16.sim_ints[1] == 1 and p16.cpu_local_shape_iterations == s.local_constraints_iterations
    p64 = oracle.pack_sim_params(s, np.float32(0.01), 5, 64, 0)
    assert p64.sim_ints[1] == s.local_constraints_iterations and p64.cpu_local_shape_iterations == 1
    assert p16.clamp_position_delta == 20.0 and p16.counts[0] == 4 and p16.counts[1] == 2 and p16.counts[2] == 16
    assert p16.reset_positions == 0.0
    assert oracle.pack_sim_params(s, np.float32(0.01), 1, 16, 0).reset_positions == 1.0


def test_sdf_grid_sizing_comma_padding():
    """numCellsInX = 10 over a 1 x 2 x 3 box: cell 0.1, padding 8 cells on every
    axis (the comma expression assigns x's value to all three)."""
    lib = oracle.load()
    mn = (ctypes.c_float * 3)(0, 0, 0)
    mx = (ctypes.c_float * 3)(1, 2, 3)
    cur = (ctypes.c_float * 3)(0.5, 0.5, 0.5)
    o = (ctypes.c_float * 3)()
    c = ctypes.c_float()
    d = (ctypes.c_int * 3)()
    tot = ctypes.c_int()
    lib.orc_sdf_grid_from_bbox(ctypes.addressof(mn), ctypes.addressof(mx), ctypes.addressof(cur), 10,
                               ctypes.addressof(o), ctypes.addressof(c), ctypes.addressof(d), ctypes.addressof(tot))
    assert abs(c.value - 0.1) < 1e-7
    assert tuple(d) in {(26, 36, 46), (25, 35, 45), (26, 35, 45), (25, 36, 46)}
    assert np.allclose(list(o), [0.5 - 0.8] * 3, atol=1e-6)
    assert tot.value == int(np.float32(1.4) * d[0] * d[1] * d[2]) or abs(tot.value - 1.4 * d[0] * d[1] * d[2]) <= 2


def test_fp32_and_fp64_transcriptions_agree_on_a_substep():
    """Stepwise, the FP32 instantiation stays within 1e-5 of the FP64 one at
    every vertex."""
    from tfx_testutil import check_fp64, fp64_reference
    wl = synth.config1(num_guides=64)
    o32 = oracle.OracleHair(wl.asset, np.float32)
    p = oracle.pack_sim_params(wl.settings, np.float32(1 / 60), 0, 16, 0)
    p.num_collision_capsules[0] = 1
    for k in range(4):
        p.center_and_radius0[0][k] = float(wl.capsules[0][0][k])
        p.center_and_radius1[0][k] = float(wl.capsules[1][0][k])
    bones = identity_bones()
    for f in range(12):
        bones[0] = wl.bone_fn((f + 1) / 60).reshape(16)
        st0 = o32.state()
        ref, sens = fp64_reference(wl.asset, st0, lambda h: h.simulate(p, bones), seed=f)
        o32.simulate(p, bones)
        check_fp64({"pos": o32.pos, "prev": o32.prev}, ref, sens, f"frame {f}", 16)
        p.reset_positions = 0.0 if f >= 1 else 1.0


def _rand_bone(rng, scale=1.0):
    """A random rigid bone as a row-major row-vector float4x4 (rotation rows, translation row 3)."""
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    r = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    m = np.eye(4)
    m[:3, :3] = r.T
    m[3, :3] = rng.normal(scale=scale, size=3)
    return m


def test_mesh_skin_identity_rigid_blend_and_ignored_slots():
    """BoneSkinning (TressFXBoneSkinning.hlsl:57-94): identity bones leave the
    mesh bit-identical; one bone of weight 1 is the rigid transform p·R + t with
    normals rotated only; unnormalised weights blend the matrices by w/sum(w);
    slots 1..3 with weight <= 0 are ignored whatever bone they name."""
    rng = np.random.Generator(np.random.PCG64(5))
    n = 500
    rest = np.zeros((n, 8), np.float32)
    rest[:, :3] = rng.uniform(-0.2, 0.2, (n, 3))
    rest[:, 4:7] = rng.uniform(0.1, 1.0, (n, 3))
    skin = np.zeros((n, 8), np.float32)
    skin[:, 4] = 1.0
    eye = np.tile(np.eye(4, dtype=np.float32).reshape(1, 16), (512, 1))
    out = oracle.mesh_skin(rest, skin, eye)
    assert np.array_equal(out.view(np.uint32), rest.view(np.uint32))

    bones = np.tile(np.eye(4).reshape(1, 16), (512, 1))
    b1, b2, b3 = _rand_bone(rng), _rand_bone(rng), _rand_bone(rng, 5.0)
    bones[1], bones[2], bones[3] = b1.reshape(16), b2.reshape(16), b3.reshape(16)
    bones32 = bones.astype(np.float32)
    m1, m2 = bones32[1].astype(np.float64).reshape(4, 4), bones32[2].astype(np.float64).reshape(4, 4)
    p, nn = rest[:, :3].astype(np.float64), rest[:, 4:7].astype(np.float64)
    # rigid: bone 1, weight 1
    skin = np.zeros((n, 8), np.float32)
    skin[:, 0] = 1
    skin[:, 4] = 1.0
    o64 = oracle.mesh_skin(rest, skin, bones32, np.float64)
    assert np.abs(o64[:, :3] - (p @ m1[:3, :3] + m1[3, :3])).max() < 1e-12
    assert np.abs(o64[:, 4:7] - nn @ m1[:3, :3]).max() < 1e-12
    assert np.array_equal(o64[:, 3], rest[:, 3]) and np.array_equal(o64[:, 7], rest[:, 7])
    # blend: bones 1 and 2 with weights 2 and 6 (normalised by their sum); slot 2
    # names bone 3 with weight 0 and slot 3 names bone 3 with a negative weight
    skin = np.zeros((n, 8), np.float32)
    skin[:, 0], skin[:, 1], skin[:, 2], skin[:, 3] = 1, 2, 3, 3
    skin[:, 4], skin[:, 5], skin[:, 6], skin[:, 7] = 2.0, 6.0, 0.0, -1.0
    o64 = oracle.mesh_skin(rest, skin, bones32, np.float64)
    mb = (2.0 * m1 + 6.0 * m2) / 8.0
    assert np.abs(o64[:, :3] - (p @ mb[:3, :3] + mb[3, :3])).max() < 1e-12
    assert np.abs(o64[:, 4:7] - nn @ mb[:3, :3]).max() < 1e-12
    # the FP32 transcription agrees with FP64 to rounding
    o32 = oracle.mesh_skin(rest, skin, bones32, np.float32)
    assert np.abs(o32.astype(np.float64) - o64).max() < 1e-5
