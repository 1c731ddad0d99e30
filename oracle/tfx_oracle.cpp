// tfx_oracle.cpp — CPU restatement of the TressFX 4.1 strand simulation and
// SDF collision path.  TEST INFRASTRUCTURE ONLY: it is the checker that the
// HIP path is compared against; nothing in the product links or calls it.
//
// A scalar transcription of the reference's dispatch sequence, templated on
// the arithmetic type (R = float for the FP32 transcription, R = double for the
// FP64 reference).  Each function cites the reference file:line it follows.
// The per-thread structure of the HLSL kernels is replaced by the equivalent
// serial order:
//   - LocalShapeConstraints is Gauss–Seidel along the strand, in memory
//     (TressFXSimulation.hlsl:794-825);
//   - wind is Jacobi: every vertex reads its neighbour's pre-wind position
//     (:909-911, the read happens before any lane of the wave stores);
//   - length constraints are red–black pairs, pass A then pass B (:920-938).
// Precision conventions shared with the GPU path (HLSL leaves them open):
//   normalize(v) = v / sqrt(dot(v,v));  length(v) = sqrt(dot(v,v));
//   dot = (x*x' + y*y') + z*z';  min/max = IEEE minNum/maxNum (fmin/fmax);
//   decay = exp(-damping * dt * 60) evaluated once per substep on the host;
//   float -> int is a saturating truncation (NaN -> 0) as D3D's ftoi.
// Build with -ffp-contract=off so no product is fused into an add.
//
// Parity pins (tests/test_oracle.py): closed forms for integration, skinning,
// global shape, VSP, LSC fixed points, length projection, wind, follow hair,
// FloatFlip3 ordering, point–triangle distance (Ericson's closest point in
// FP64), SDF order independence and the plane-SDF collision response.
// Host parameter packing (SetWind) is pinned against closed-form rotations;
// collision-mesh skinning against rigid transforms and matrix blends.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../include/tressfx_amd.h"

namespace {

template <class R> struct V3 { R x, y, z; };
template <class R> struct V4 { R x, y, z, w; };

template <class R> inline V3<R> v3(R x, R y, R z) { return V3<R>{x, y, z}; }
template <class R> inline V3<R> xyz(const V4<R>& a) { return V3<R>{a.x, a.y, a.z}; }
template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return V3<R>{a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return V3<R>{a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> operator*(R s, V3<R> a) { return V3<R>{s * a.x, s * a.y, s * a.z}; }
template <class R> inline V3<R> operator*(V3<R> a, R s) { return V3<R>{a.x * s, a.y * s, a.z * s}; }
template <class R> inline V3<R> operator/(V3<R> a, R s) { return V3<R>{a.x / s, a.y / s, a.z / s}; }
template <class R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline V3<R> cross(V3<R> a, V3<R> b) {
    return V3<R>{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R> inline R length(V3<R> a) { return std::sqrt(dot(a, a)); }
template <class R> inline V3<R> normalize(V3<R> a) { return a / std::sqrt(dot(a, a)); }
template <class R> inline R fmn(R a, R b) { return std::fmin(a, b); }
template <class R> inline R fmx(R a, R b) { return std::fmax(a, b); }

// D3D ftoi: truncation toward zero, NaN -> 0, saturating at the int range.
template <class R> inline int ftoi(R x) {
    if (!(x == x)) return 0;
    if (x >= R(2147483648.0)) return 2147483647;
    if (x <= R(-2147483648.0)) return (-2147483647 - 1);
    return (int)x;
}

template <class R> inline V4<R> load4(const R* p, size_t i) { return V4<R>{p[4 * i], p[4 * i + 1], p[4 * i + 2], p[4 * i + 3]}; }
template <class R> inline void store4(R* p, size_t i, V4<R> v) { p[4 * i] = v.x; p[4 * i + 1] = v.y; p[4 * i + 2] = v.z; p[4 * i + 3] = v.w; }
template <class R> inline void store3(R* p, size_t i, V3<R> v) { p[4 * i] = v.x; p[4 * i + 1] = v.y; p[4 * i + 2] = v.z; }

// ---------------------------------------------------------------------------
// Threading: strands are independent within every stage.
// ---------------------------------------------------------------------------
template <class F> void parallel_for(int n, int nthreads, F&& f) {
    if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    if (nthreads > n) nthreads = std::max(1, n);
    if (nthreads == 1) { for (int i = 0; i < n; ++i) f(i); return; }
    std::vector<std::thread> th;
    std::atomic<int> next(0);
    const int chunk = std::max(1, n / (nthreads * 8));
    for (int t = 0; t < nthreads; ++t)
        th.emplace_back([&]() {
            for (;;) {
                int b = next.fetch_add(chunk);
                if (b >= n) break;
                int e = std::min(n, b + chunk);
                for (int i = b; i < e; ++i) f(i);
            }
        });
    for (auto& t : th) t.join();
}

// ---------------------------------------------------------------------------
// Quaternion and matrix helpers, TressFXSimulation.hlsl:207-363
// ---------------------------------------------------------------------------
// IsMovable :207-212
template <class R> inline bool is_movable(const V4<R>& p) { return p.w > R(0); }

// ConstraintMultiplier :214-230
template <class R> inline void constraint_multiplier(const V4<R>& a, const V4<R>& b, R& m0, R& m1) {
    if (is_movable(a)) {
        if (is_movable(b)) { m0 = R(0.5); m1 = R(0.5); }
        else { m0 = R(1); m1 = R(0); }
    } else {
        if (is_movable(b)) { m0 = R(0); m1 = R(1); }
        else { m0 = R(0); m1 = R(0); }
    }
}

// MakeQuaternion(float4x4) :247-288; m[i][j] is the logical (row i, col j) element.
template <class R> V4<R> make_quaternion(const R m[4][4]) {
    V4<R> q;
    R trace = m[0][0] + m[1][1] + m[2][2];
    if (trace > R(0)) {
        R r = std::sqrt(trace + R(1));
        q.w = R(0.5) * r;
        r = R(0.5) / r;
        q.x = (m[1][2] - m[2][1]) * r;
        q.y = (m[2][0] - m[0][2]) * r;
        q.z = (m[0][1] - m[1][0]) * r;
    } else {
        int i = 0, j = 1, k = 2;
        if (m[1][1] > m[0][0]) { i = 1; j = 2; k = 0; }
        if (m[2][2] > m[i][i]) { i = 2; j = 0; k = 1; }
        R r = std::sqrt(m[i][i] - m[j][j] - m[k][k] + R(1));
        R qq[4];
        qq[i] = R(0.5) * r;
        r = R(0.5) / r;
        q.w = (m[j][k] - m[k][j]) * r;
        qq[j] = (m[j][i] + m[i][j]) * r;
        qq[k] = (m[k][i] + m[i][k]) * r;
        q.x = qq[0]; q.y = qq[1]; q.z = qq[2];
    }
    return q;
}

// MultQuaternionAndVector :305-315
template <class R> inline V3<R> mult_quat_vec(const V4<R>& q, V3<R> v) {
    V3<R> qv{q.x, q.y, q.z};
    V3<R> uv = cross(qv, v);
    V3<R> uuv = cross(qv, uv);
    uv = uv * (R(2) * q.w);
    uuv = uuv * R(2);
    return (v + uv) + uuv;
}

// NormalizeQuaternion :329-342
template <class R> inline V4<R> normalize_quat(V4<R> q) {
    V4<R> qq = q;
    R n = q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w;
    if (n < R(1e-10f)) { qq.w = R(1); return qq; }
    R s = R(1) / std::sqrt(n);
    qq.x *= s; qq.y *= s; qq.z *= s; qq.w *= s;
    return qq;
}

// QuatFromTwoUnitVectors :345-363
template <class R> inline V4<R> quat_from_two_unit_vectors(V3<R> u, V3<R> v) {
    R r = R(1) + dot(u, v);
    V3<R> n;
    if (r < R(1e-7f)) {
        r = R(0);
        n = std::fabs(u.x) > std::fabs(u.z) ? V3<R>{-u.y, u.x, R(0)} : V3<R>{R(0), -u.z, u.y};
    } else {
        n = cross(u, v);
    }
    return normalize_quat(V4<R>{n.x, n.y, n.z, r});
}

// ApplyVertexBoneSkinning (matrix path) :524-572. Returns the skinned
// position; the blended matrix's quaternion goes to `quat`.
template <class R> struct BlendedBone { R m[4][4]; V4<R> quat; };

// `bones` is the object's full 512-entry table.  Indices are validated to lie
// in [0, 511] when the object is created; the clamp only keeps reads in bounds.
template <class R> BlendedBone<R> blend_bones(const float* skin /* 8 floats */, const float* bones) {
    auto bone = [&](float fidx) -> const float* {
        int idx = ftoi<float>(fidx);
        idx = std::max(0, std::min(idx, TFX_MAX_NUM_BONES - 1));
        return bones + 16 * idx;
    };
    BlendedBone<R> b;
    const float* m0 = bone(skin[0]);
    R w0 = (R)skin[4];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) b.m[r][c] = (R)m0[4 * r + c] * w0;
    R wsum = w0;
    for (int i = 1; i < 4; ++i) {
        R wi = (R)skin[4 + i];
        if (wi > R(0)) {
            const float* mi = bone(skin[i]);
            for (int r = 0; r < 4; ++r)
                for (int c = 0; c < 4; ++c) b.m[r][c] = b.m[r][c] + (R)mi[4 * r + c] * wi;
            wsum = wsum + wi;
        }
    }
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) b.m[r][c] = b.m[r][c] / wsum;
    b.quat = make_quaternion<R>(b.m);
    return b;
}

// mul(float4(p, 1), M).xyz  (row vector times row-major matrix)
template <class R> inline V3<R> transform_point(const BlendedBone<R>& b, V3<R> p) {
    const auto& m = b.m;
    return V3<R>{((p.x * m[0][0] + p.y * m[1][0]) + p.z * m[2][0]) + m[3][0],
                 ((p.x * m[0][1] + p.y * m[1][1]) + p.z * m[2][1]) + m[3][1],
                 ((p.x * m[0][2] + p.y * m[1][2]) + p.z * m[2][2]) + m[3][2]};
}

// ApplyDistanceConstraint :377-387 (stiffness 1)
template <class R> inline void apply_distance_constraint(V4<R>& p0, V4<R>& p1, R target) {
    V3<R> delta = xyz(p1) - xyz(p0);
    R distance = fmx(length(delta), R(1e-7f));
    R stretching = R(1) - target / distance;
    delta = stretching * delta;
    R m0, m1;
    constraint_multiplier(p0, p1, m0, m1);
    V3<R> d0 = (m0 * delta) * R(1);
    V3<R> d1 = (m1 * delta) * R(1);
    p0.x += d0.x; p0.y += d0.y; p0.z += d0.z;
    p1.x -= d1.x; p1.y -= d1.y; p1.z -= d1.z;
}

// CapsuleCollision :462-522 (friction 0.4)
template <class R> bool capsule_collision(const V4<R>& cur, const V4<R>& old, V3<R>& newp,
                                          const V4<R>& c0, const V4<R>& c1, R friction) {
    const R radius0 = c0.w, radius1 = c1.w;
    newp = xyz(cur);
    if (!is_movable(cur)) return false;
    V3<R> segment = xyz(c1) - xyz(c0);
    V3<R> delta0 = xyz(cur) - xyz(c0);
    V3<R> delta1 = xyz(c1) - xyz(cur);
    R dist0 = dot(delta0, segment);
    R dist1 = dot(delta1, segment);
    if (dist0 < R(0)) {
        if (dot(delta0, delta0) < radius0 * radius0) {
            V3<R> n = normalize(delta0);
            newp = radius0 * n + xyz(c0);
            return true;
        }
        return false;
    }
    if (dist1 < R(0)) {
        if (dot(delta1, delta1) < radius1 * radius1) {
            V3<R> nd = V3<R>{-delta1.x, -delta1.y, -delta1.z};
            V3<R> n = normalize(nd);
            newp = radius1 * n + xyz(c1);
            return true;
        }
        return false;
    }
    V3<R> x = (dist0 * xyz(c1) + dist1 * xyz(c0)) / (dist0 + dist1);
    V3<R> delta = xyz(cur) - x;
    R radius_at_x = (dist0 * radius1 + dist1 * radius0) / (dist0 + dist1);
    if (dot(delta, delta) < radius_at_x * radius_at_x) {
        V3<R> n = normalize(delta);
        V3<R> vec = xyz(cur) - xyz(old);
        V3<R> segN = normalize(segment);
        V3<R> vecTangent = dot(vec, segN) * segN;
        V3<R> vecNormal = vec - vecTangent;
        newp = xyz(old) + friction * vecTangent + ((vecNormal + radius_at_x * n) - delta);
        return true;
    }
    return false;
}

// ---------------------------------------------------------------------------
// One hair object's arrays (float4 streams, strand-major, follows interleaved)
// ---------------------------------------------------------------------------
template <class R> struct Hair {
    int G, F, V;
    R* pos; R* prev; R* prevprev; R* tangent; R* sld;  // sld: 12 per strand
    const R* rest; const R* restlen; const R* follow_off;
    const float* skin;  // 8 floats per strand (indices as float, weights)
};

struct Params {
    const tfx_simulation_params* p;
    const float* bones;
};

// IntegrationAndGlobalShapeConstraints, TressFXSimulation.hlsl:609-666
// (helpers: CalcIndicesInVertexLevelMaster :402-414, Integrate :440-446,
// UpdateFinalVertexPositions :581-586).  One guide strand per task.
template <class R> void stage_integrate(Hair<R>& h, const Params& P, int nthreads) {
    const tfx_simulation_params& p = *P.p;
    const int V = h.V, F = h.F;
    const R damping = (R)p.shape[0];
    const R gstiff = (R)p.shape[2], grange = (R)p.shape[3];
    const R g = (R)p.grav_time_tip[0], dt = (R)p.grav_time_tip[1];
    // decay = exp(-dampingCoeff * g_TimeStep * 60) (:443), once per substep
    const R decay = std::exp((-damping * dt) * R(60));
    const bool reset = p.reset_positions != 0.0f;
    parallel_for(h.G, nthreads, [&](int gi) {
        const size_t strand = (size_t)gi * (F + 1);
        BlendedBone<R> bb = blend_bones<R>(h.skin + 8 * strand, P.bones);
        for (int v = 0; v < V; ++v) {
            const size_t i = strand * V + v;
            V4<R> initialPos = load4(h.rest, i);
            V3<R> sk = transform_point(bb, xyz(initialPos));
            initialPos.x = sk.x; initialPos.y = sk.y; initialPos.z = sk.z;
            V4<R> shared = load4(h.pos, i);
            V4<R> currentPos = shared;
            V4<R> oldPos = load4(h.prev, i);
            if (reset) { currentPos = initialPos; oldPos = initialPos; }
            if (is_movable(currentPos)) {
                // Integrate :440-446: cur + decay*(cur-old) + force*dt*dt
                V3<R> force{g * R(0), g * R(-1), g * R(0)};
                V3<R> c = xyz(currentPos), o = xyz(oldPos);
                V3<R> n = (c + decay * (c - o)) + (force * dt) * dt;
                shared.x = n.x; shared.y = n.y; shared.z = n.z;
            } else {
                shared = initialPos;
            }
            if (gstiff > R(0) && grange != R(0)) {
                if (is_movable(shared)) {
                    if ((R)(float)v < grange * (R)(float)V) {
                        V3<R> del = gstiff * (xyz(initialPos) - xyz(shared));
                        shared.x += del.x; shared.y += del.y; shared.z += del.z;
                    }
                }
            }
            // UpdateFinalVertexPositions
            store4(h.prevprev, i, load4(h.prev, i));
            store4(h.prev, i, currentPos);
            store4(h.pos, i, shared);
        }
    });
}

// CalculateStrandLevelData :677-728 (CalcIndicesInStrandLevelMaster :424-431)
template <class R> void stage_strand_level(Hair<R>& h, const Params& P, int nthreads) {
    const tfx_simulation_params& p = *P.p;
    const int V = h.V, F = h.F;
    parallel_for(h.G, nthreads, [&](int gi) {
        const size_t strand = (size_t)gi * (F + 1);
        const size_t root = strand * V;
        V4<R> oo1 = load4(h.prevprev, root + 1);
        V4<R> o0 = load4(h.prev, root), o1 = load4(h.prev, root + 1);
        V4<R> n0 = load4(h.pos, root), n1 = load4(h.pos, root + 1);
        V3<R> u = normalize(xyz(o1) - xyz(o0));
        V3<R> v = normalize(xyz(n1) - xyz(n0));
        V4<R> rot = quat_from_two_unit_vectors(u, v);
        V3<R> trans = xyz(n0) - mult_quat_vec(rot, xyz(o0));
        R vspCoeff = (R)p.vsp[0];
        R thr = (R)p.vsp[1];
        // length(pos_new[1] - 2.0*pos_old[1] + pos_old_old[1]) over float4
        R ax = (n1.x - R(2) * o1.x) + oo1.x;
        R ay = (n1.y - R(2) * o1.y) + oo1.y;
        R az = (n1.z - R(2) * o1.z) + oo1.z;
        R aw = (n1.w - R(2) * o1.w) + oo1.w;
        R accel = std::sqrt(((ax * ax + ay * ay) + az * az) + aw * aw);
        if (accel > thr) vspCoeff = R(1);
        BlendedBone<R> bb = blend_bones<R>(h.skin + 8 * strand, P.bones);
        R* s = h.sld + 12 * strand;
        s[0] = bb.quat.x; s[1] = bb.quat.y; s[2] = bb.quat.z; s[3] = bb.quat.w;
        s[4] = rot.x; s[5] = rot.y; s[6] = rot.z; s[7] = rot.w;
        s[8] = trans.x; s[9] = trans.y; s[10] = trans.z; s[11] = vspCoeff;
    });
}

// VelocityShockPropagation :739-762
template <class R> void stage_vsp(Hair<R>& h, const Params&, int nthreads) {
    const int V = h.V, F = h.F;
    parallel_for(h.G, nthreads, [&](int gi) {
        const size_t strand = (size_t)gi * (F + 1);
        const R* s = h.sld + 12 * strand;
        V4<R> q{s[4], s[5], s[6], s[7]};
        V3<R> t{s[8], s[9], s[10]};
        R c = s[11];
        for (int v = 2; v < V; ++v) {
            const size_t i = strand * V + v;
            V3<R> pn = xyz(load4(h.pos, i)), po = xyz(load4(h.prev, i));
            V3<R> rn = (R(1) - c) * pn + c * (mult_quat_vec(q, pn) + t);
            V3<R> ro = (R(1) - c) * po + c * (mult_quat_vec(q, po) + t);
            store3(h.pos, i, rn);
            store3(h.prev, i, ro);
        }
    });
}

// LocalShapeConstraints :773-829 — one pass (the host repeats it).
template <class R> void stage_lsc(Hair<R>& h, const Params& P, int nthreads) {
    const tfx_simulation_params& p = *P.p;
    const int V = h.V, F = h.F;
    const R stiff = R(0.5) * fmn((R)p.shape[1], R(0.95f));
    parallel_for(h.G, nthreads, [&](int gi) {
        const size_t strand = (size_t)gi * (F + 1);
        const size_t root = strand * V;
        const R* s = h.sld + 12 * strand;
        V4<R> boneQuat{s[0], s[1], s[2], s[3]};
        for (int lv = 1; lv < V - 1; ++lv) {
            const size_t i = root + lv;
            V4<R> pos = load4(h.pos, i);
            V4<R> pos_plus_one = load4(h.pos, i + 1);
            V4<R> pos_minus_one = load4(h.pos, i - 1);
            V3<R> bindPos = mult_quat_vec(boneQuat, xyz(load4(h.rest, i)));
            V3<R> bindPos_plus_one = mult_quat_vec(boneQuat, xyz(load4(h.rest, i + 1)));
            V3<R> bindPos_minus_one = mult_quat_vec(boneQuat, xyz(load4(h.rest, i - 1)));
            V3<R> lastVec = xyz(pos) - xyz(pos_minus_one);
            // InverseQuaternion(boneQuat) is computed at :808 and unused.
            V3<R> vecBindPose = bindPos_plus_one - bindPos;
            V3<R> lastVecBindPose = bindPos - bindPos_minus_one;
            V4<R> rotGlobal = quat_from_two_unit_vectors(normalize(lastVecBindPose), normalize(lastVec));
            V3<R> orgPos = mult_quat_vec(rotGlobal, vecBindPose) + xyz(pos);
            V3<R> del = stiff * (orgPos - xyz(pos_plus_one));
            if (is_movable(pos)) { pos.x -= del.x; pos.y -= del.y; pos.z -= del.z; }
            if (is_movable(pos_plus_one)) { pos_plus_one.x += del.x; pos_plus_one.y += del.y; pos_plus_one.z += del.z; }
            store3(h.pos, i, xyz(pos));
            store3(h.pos, i + 1, xyz(pos_plus_one));
        }
    });
}

// LengthConstriantsWindAndCollision :875-975 (ResolveCapsuleCollisions :832-863)
template <class R> void stage_length(Hair<R>& h, const Params& P, int nthreads) {
    const tfx_simulation_params& p = *P.p;
    const int V = h.V, F = h.F;
    const R dt = (R)p.grav_time_tip[1];
    const bool wind_on = p.wind[0] != 0.0f || p.wind[1] != 0.0f || p.wind[2] != 0.0f;
    const int iters = p.sim_ints[0];
    const int ncaps = std::min(std::max(p.num_collision_capsules[0], 0), TFX_MAX_NUM_COLLISION_CAPSULES);
    V3<R> W0{(R)p.wind[0], (R)p.wind[1], (R)p.wind[2]};
    V3<R> W1{(R)p.wind1[0], (R)p.wind1[1], (R)p.wind1[2]};
    V3<R> W2{(R)p.wind2[0], (R)p.wind2[1], (R)p.wind2[2]};
    V3<R> W3{(R)p.wind3[0], (R)p.wind3[1], (R)p.wind3[2]};
    parallel_for(h.G, nthreads, [&](int gi) {
        const size_t strand = (size_t)gi * (F + 1);
        const size_t root = strand * V;
        std::vector<V4<R>> sp(V);
        std::vector<R> sl(V);
        for (int v = 0; v < V; ++v) { sp[v] = load4(h.pos, root + v); sl[v] = h.restlen[root + v]; }
        if (wind_on) {
            // Jacobi: all reads precede all writes
            std::vector<V3<R>> force(V, V3<R>{R(0), R(0), R(0)});
            R a = (R)(float)(strand % 20) / R(20);
            V3<R> w = ((a * W0 + (R(1) - a) * W1) + a * W2) + (R(1) - a) * W3;
            for (int v = 2; v < V - 1; ++v) {
                V3<R> d = xyz(sp[v]) - xyz(sp[v + 1]);
                V3<R> c = cross(cross(d, w), d);
                force[v] = V3<R>{-c.x, -c.y, -c.z};
            }
            for (int v = 2; v < V - 1; ++v) {
                V3<R> f = (force[v] * dt) * dt;
                sp[v].x += f.x; sp[v].y += f.y; sp[v].z += f.z;
            }
        }
        const int a = V / 2, b = (V - 1) / 2;
        for (int it = 0; it < iters; ++it) {
            for (int k = 0; k < a; ++k) apply_distance_constraint(sp[2 * k], sp[2 * k + 1], sl[2 * k]);
            for (int k = 0; k < b; ++k) apply_distance_constraint(sp[2 * k + 1], sp[2 * k + 2], sl[2 * k + 1]);
        }
        std::vector<V4<R>> oldp(V);
        std::vector<char> hit(V, 0);
        for (int v = 0; v < V; ++v) {
            oldp[v] = load4(h.prev, root + v);
            for (int c = 0; c < ncaps; ++c) {
                V4<R> c0{(R)p.center_and_radius0[c][0], (R)p.center_and_radius0[c][1], (R)p.center_and_radius0[c][2], (R)p.center_and_radius0[c][3]};
                V4<R> c1{(R)p.center_and_radius1[c][0], (R)p.center_and_radius1[c][1], (R)p.center_and_radius1[c][2], (R)p.center_and_radius1[c][3]};
                V3<R> np;
                bool col = capsule_collision(sp[v], oldp[v], np, c0, c1, R(0.4f));
                if (col) { sp[v].x = np.x; sp[v].y = np.y; sp[v].z = np.z; hit[v] = 1; }
            }
        }
        const R C = (R)p.clamp_position_delta;
        for (int v = 0; v < V; ++v) {
            const size_t i = root + v;
            int ti = (v == V - 1) ? v - 1 : v;
            V3<R> tangent = xyz(sp[ti + 1]) - xyz(sp[ti]);
            store3(h.tangent, i, normalize(tangent));
            V3<R> positionDelta = xyz(sp[v]) - xyz(oldp[v]);
            R speedSqr = dot(positionDelta, positionDelta);
            if (speedSqr > C * C) {
                R sc = (C * C) / speedSqr;
                positionDelta = positionDelta * sc;
                store3(h.prev, i, xyz(sp[v]) - positionDelta);
            }
            store4(h.pos, i, sp[v]);
            if (hit[v]) store4(h.prev, i, sp[v]);
        }
    });
}

// UpdateFollowHairVertices :978-1001
template <class R> void stage_follow(Hair<R>& h, const Params& P, int nthreads) {
    const tfx_simulation_params& p = *P.p;
    const int V = h.V, F = h.F;
    const R tip = (R)p.grav_time_tip[2];
    if (F <= 0) return;
    parallel_for(h.G, nthreads, [&](int gi) {
        const size_t strand = (size_t)gi * (F + 1);
        for (int v = 0; v < V; ++v) {
            const size_t gv = strand * V + v;
            V4<R> gp = load4(h.pos, gv);
            V4<R> gt = load4(h.tangent, gv);
            for (int j = 0; j < F; ++j) {
                const size_t fv = gv + (size_t)V * (j + 1);
                const size_t fs = strand + j + 1;
                R factor = tip * ((R)(float)v / (R)(float)V) + R(1);
                V3<R> off = xyz(load4(h.follow_off, fs));
                store3(h.pos, fv, xyz(gp) + factor * off);
                store4(h.tangent, fv, gt);
            }
        }
    });
}

template <class R> void run_stage(Hair<R>& h, const Params& P, int stage, int nthreads) {
    switch (stage) {
        case 0: stage_integrate(h, P, nthreads); break;
        case 1: stage_strand_level(h, P, nthreads); break;
        case 2: stage_vsp(h, P, nthreads); break;
        case 3: stage_lsc(h, P, nthreads); break;
        case 4: stage_length(h, P, nthreads); break;
        case 5: stage_follow(h, P, nthreads); break;
        default: break;
    }
}

// TressFXSimulation::Simulate, TressFXSimulation.cpp:77-112 for one object:
// the LSC pass is re-dispatched GetCPULocalShapeIterations() times (:64-72).
template <class R> void simulate(Hair<R>& h, const Params& P, int nthreads) {
    run_stage(h, P, 0, nthreads);
    run_stage(h, P, 1, nthreads);
    run_stage(h, P, 2, nthreads);
    for (int i = 0; i < P.p->cpu_local_shape_iterations; ++i) run_stage(h, P, 3, nthreads);
    run_stage(h, P, 4, nthreads);
    run_stage(h, P, 5, nthreads);
}

// ---------------------------------------------------------------------------
// SDF, TressFXSDFCollision.hlsl
// ---------------------------------------------------------------------------
// FloatFlip3 / IFloatFlip3 :92-100 (and the 64-bit analogue for the FP64 grid:
// rotate the sign bit to the least significant position).
inline uint32_t float_flip3(float f) { uint32_t u; std::memcpy(&u, &f, 4); return (u << 1) | (u >> 31); }
inline uint32_t ifloat_flip3(uint32_t f) { return (f >> 1) | (f << 31); }
inline uint64_t double_flip3(double f) { uint64_t u; std::memcpy(&u, &f, 8); return (u << 1) | (u >> 63); }
inline uint64_t idouble_flip3(uint64_t f) { return (f >> 1) | (f << 63); }

template <class R> struct Grid {
    V3<R> origin; R cell; int nx, ny, nz;
};

// GetSdfCoordinates :103-113
template <class R> inline void sdf_coords(const Grid<R>& g, V3<R> p, int c[3]) {
    V3<R> s = (p - g.origin) / g.cell;
    c[0] = ftoi(s.x); c[1] = ftoi(s.y); c[2] = ftoi(s.z);
}
// GetSdfCellPosition :115-121
template <class R> inline V3<R> sdf_cell_pos(const Grid<R>& g, int x, int y, int z) {
    V3<R> c = V3<R>{(R)(float)x, (R)(float)y, (R)(float)z} * g.cell;
    return c + g.origin;
}
// GetSdfCellIndex :123-129 (wrapping int arithmetic as on the GPU)
template <class R> inline int sdf_index(const Grid<R>& g, int x, int y, int z) {
    uint32_t perLine = (uint32_t)g.nx, perPlane = (uint32_t)g.nx * (uint32_t)g.ny;
    return (int)(perPlane * (uint32_t)z + perLine * (uint32_t)y + (uint32_t)x);
}

// DistancePointToEdge :147-159
template <class R> inline R distance_point_to_edge(V3<R> p, V3<R> x0, V3<R> x1) {
    V3<R> x10 = x1 - x0;
    R t = dot(x1 - p, x10) / dot(x10, x10);
    t = fmx(R(0), fmn(t, R(1)));
    V3<R> a = p - (t * x0 + (R(1) - t) * x1);
    return length(a);
}

// SignedDistancePointToTriangle :163-215 (the normal it also forms is unused)
template <class R> R signed_distance_point_triangle(V3<R> p, V3<R> x0, V3<R> x1, V3<R> x2) {
    R d = R(0);
    V3<R> x02 = x0 - x2;
    R l0 = length(x02) + R(1e-30f);
    x02 = x02 / l0;
    V3<R> x12 = x1 - x2;
    R l1 = dot(x12, x02);
    x12 = x12 - l1 * x02;
    R l2 = length(x12) + R(1e-30f);
    x12 = x12 / l2;
    V3<R> px2 = p - x2;
    R b = dot(x12, px2) / l2;
    R a = (dot(x02, px2) - l1 * b) / l0;
    R c = (R(1) - a) - b;
    V3<R> nTri = cross(x1 - x0, x2 - x0);
    const R tol = R(1e-8f);
    if (a >= -tol && b >= -tol && c >= -tol) {
        V3<R> n = p - ((a * x0 + b * x1) + c * x2);
        d = length(n);
    } else {
        d = distance_point_to_edge(p, x0, x1);
        R d12 = distance_point_to_edge(p, x1, x2);
        R d02 = distance_point_to_edge(p, x0, x2);
        d = fmn(d, d12);
        d = fmn(d, d02);
    }
    d = (dot(p - x0, nTri) < R(0)) ? -d : d;
    return d;
}

template <class R> struct FlipTraits;
template <> struct FlipTraits<float> {
    using U = uint32_t;
    static U flip(float f) { return float_flip3(f); }
    static float unflip(U u) { float f; uint32_t b = ifloat_flip3(u); std::memcpy(&f, &b, 4); return f; }
};
template <> struct FlipTraits<double> {
    using U = uint64_t;
    static U flip(double f) { return double_flip3(f); }
    static double unflip(U u) { double f; uint64_t b = idouble_flip3(u); std::memcpy(&f, &b, 8); return f; }
};

// InitializeSignedDistanceField :301-312, ConstructSignedDistanceField
// :329-397, FinalizeSignedDistanceField :400-412.  The construct step is
// multithreaded over triangles with a CAS atomic-min, so the grid is the same
// for any thread count or triangle order.
template <class R> void sdf_build(const R* verts /* 8 per vertex */, const uint32_t* idx, int ntri,
                                  const Grid<R>& g, R* out, int nthreads) {
    using T = FlipTraits<R>;
    using U = typename T::U;
    const size_t ncell = (size_t)g.nx * g.ny * g.nz;
    std::vector<std::atomic<U>> cells(ncell);
    const U init = T::flip(R(1e10f));
    for (size_t i = 0; i < ncell; ++i) cells[i].store(init, std::memory_order_relaxed);
    const R margin = g.cell;
    parallel_for(ntri, nthreads, [&](int t) {
        V3<R> tri[3];
        for (int k = 0; k < 3; ++k) {
            uint32_t vi = idx[3 * t + k];
            tri[k] = V3<R>{verts[8 * (size_t)vi], verts[8 * (size_t)vi + 1], verts[8 * (size_t)vi + 2]};
        }
        V3<R> mn{fmn(tri[0].x, fmn(tri[1].x, tri[2].x)), fmn(tri[0].y, fmn(tri[1].y, tri[2].y)), fmn(tri[0].z, fmn(tri[1].z, tri[2].z))};
        V3<R> mx{fmx(tri[0].x, fmx(tri[1].x, tri[2].x)), fmx(tri[0].y, fmx(tri[1].y, tri[2].y)), fmx(tri[0].z, fmx(tri[1].z, tri[2].z))};
        mn = mn - V3<R>{margin, margin, margin};
        mx = mx + V3<R>{margin, margin, margin};
        int gmin[3], gmax[3];
        sdf_coords(g, mn, gmin);
        sdf_coords(g, mx, gmax);
        const int dims[3] = {g.nx, g.ny, g.nz};
        for (int k = 0; k < 3; ++k) {
            gmin[k] = (int)((uint32_t)gmin[k] - 1u);
            gmax[k] = (int)((uint32_t)gmax[k] + 1u);
            gmin[k] = std::max(0, std::min(gmin[k], dims[k] - 1));
            gmax[k] = std::max(0, std::min(gmax[k], dims[k] - 1));
        }
        for (int z = gmin[2]; z <= gmax[2]; ++z)
            for (int y = gmin[1]; y <= gmax[1]; ++y)
                for (int x = gmin[0]; x <= gmax[0]; ++x) {
                    int ci = sdf_index(g, x, y, z);
                    V3<R> cp = sdf_cell_pos(g, x, y, z);
                    R dist = signed_distance_point_triangle(cp, tri[0], tri[1], tri[2]);
                    U du = T::flip(dist);
                    std::atomic<U>& cell = cells[(size_t)ci];
                    U cur = cell.load(std::memory_order_relaxed);
                    while (du < cur && !cell.compare_exchange_weak(cur, du, std::memory_order_relaxed)) {}
                }
    });
    for (size_t i = 0; i < ncell; ++i) out[i] = T::unflip(cells[i].load(std::memory_order_relaxed));
}

// LinearInterpolate :414-417, BilinearInterpolate :435-438, TrilinearInterpolate :466-469
template <class R> inline R lerp_(R a, R b, R t) { return a * (R(1) - t) + b * t; }
template <class R> inline R bilerp(R a, R b, R c, R d, R p, R q) { return lerp_(lerp_(a, b, p), lerp_(c, d, p), q); }

// GetSignedDistance :472-527
template <class R> R get_signed_distance(const Grid<R>& g, const R* sdf, V3<R> p) {
    const R INITIAL = R(1e10f);
    int c[3];
    sdf_coords(g, p, c);
    if (!(0 <= c[0] && c[0] < g.nx - 2) || !(0 <= c[1] && c[1] < g.ny - 2) || !(0 <= c[2] && c[2] < g.nz - 2))
        return INITIAL;
    int base = sdf_index(g, c[0], c[1], c[2]);
    int X = 1, Y = g.nx, Z = g.ny * g.nx;
    int ids[8] = {base, base + X, base + Y, base + Y + X, base + Z, base + Z + X, base + Z + Y, base + Z + Y + X};
    R d[8];
    for (int j = 0; j < 8; ++j) {
        R v = sdf[ids[j]];
        if (v == INITIAL) return INITIAL;
        d[j] = v;
    }
    V3<R> cp = sdf_cell_pos(g, c[0], c[1], c[2]);
    V3<R> t = (p - cp) / g.cell;
    return lerp_(bilerp(d[0], d[1], d[2], d[3], t.x, t.y), bilerp(d[4], d[5], d[6], d[7], t.x, t.y), t.z);
}

// CollideHairVerticesWithSdf_forward :531-595, over all N vertices.
template <class R> void sdf_collide(R* pos, R* prev, int N, int V, const R* sdf, const Grid<R>& g, R margin,
                                    int nthreads) {
    const int S = (V > 0) ? N / V : 0;
    parallel_for(S, nthreads, [&](int s) {
        for (int lv = 2; lv < V; ++lv) {
            const size_t i = (size_t)s * V + lv;
            V4<R> hv = load4(pos, i);
            V3<R> p = xyz(hv);
            R distance = get_signed_distance(g, sdf, p);
            if (distance > margin) continue;
            R hstep = R(0.1f) * g.cell;
            R fx = get_signed_distance(g, sdf, V3<R>{p.x + R(1) * hstep, p.y + R(0) * hstep, p.z + R(0) * hstep});
            R fy = get_signed_distance(g, sdf, V3<R>{p.x + R(0) * hstep, p.y + R(1) * hstep, p.z + R(0) * hstep});
            R fz = get_signed_distance(g, sdf, V3<R>{p.x + R(0) * hstep, p.y + R(0) * hstep, p.z + R(1) * hstep});
            V3<R> grad = V3<R>{fx - distance, fy - distance, fz - distance} / hstep;
            V3<R> n = normalize(grad);
            if (distance < margin) {
                V3<R> proj = p + n * (margin - distance);
                store3(pos, i, proj);
                store3(prev, i, proj);
            }
        }
    });
}

// CollideHairVerticesWithSdf :597-715, the variant mixing forward and
// backward differences (compiled by the reference, not dispatched): every
// vertex, the gradient from the same eight nodes as the distance, h = 0.1 *
// cell added to the cell-unit interpolation weights as written.
template <class R> void sdf_collide_mixed(R* pos, R* prev, int N, const R* sdf, const Grid<R>& g, R margin,
                                          int nthreads) {
    const R INITIAL = R(1e10f);
    parallel_for(N, nthreads, [&](int i) {
        V4<R> hv = load4(pos, (size_t)i);
        V3<R> p = xyz(hv);
        int c[3];
        sdf_coords(g, p, c);
        if (!(0 <= c[0] && c[0] < g.nx - 2) || !(0 <= c[1] && c[1] < g.ny - 2) || !(0 <= c[2] && c[2] < g.nz - 2))
            return;
        int base = sdf_index(g, c[0], c[1], c[2]);
        int X = 1, Y = g.nx, Z = g.ny * g.nx;
        int ids[8] = {base, base + X, base + Y, base + Y + X, base + Z, base + Z + X, base + Z + Y, base + Z + Y + X};
        R d[8];
        for (int j = 0; j < 8; ++j) {
            d[j] = sdf[ids[j]];
            if (d[j] == INITIAL) return;
        }
        auto tri = [&](R x, R y, R z) {
            return lerp_(bilerp(d[0], d[1], d[2], d[3], x, y), bilerp(d[4], d[5], d[6], d[7], x, y), z);
        };
        V3<R> cp = sdf_cell_pos(g, c[0], c[1], c[2]);
        V3<R> t = (p - cp) / g.cell;
        R dv = tri(t.x, t.y, t.z);
        R h = R(0.1f) * g.cell;
        V3<R> dir{(t.x + h < R(1)) ? R(1) : R(-1), (t.y + h < R(1)) ? R(1) : R(-1), (t.z + h < R(1)) ? R(1) : R(-1)};
        V3<R> nb{tri(t.x + h * dir.x, t.y, t.z), tri(t.x, t.y + h * dir.y, t.z), tri(t.x, t.y, t.z + h * dir.z)};
        V3<R> grad = V3<R>{dir.x * (nb.x - dv), dir.y * (nb.y - dv), dir.z * (nb.z - dv)} / h;
        V3<R> n = normalize(grad);
        if (dv < margin) {
            V3<R> proj = p + n * (margin - dv);
            store3(pos, (size_t)i, proj);
            store3(prev, (size_t)i, proj);
        }
    });
}

// GenerateTransforms, TressFXSimulation.hlsl:1035-1142.  Row-major 4x4
// products summed left to right, ((a0 b0 + a1 b1) + a2 b2) + a3 b3.
template <class R> struct M4 { R m[4][4]; };
template <class R> M4<R> m4_mul(const M4<R>& a, const M4<R>& b) {
    M4<R> c;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            c.m[i][j] = ((a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j]) + a.m[i][2] * b.m[2][j]) + a.m[i][3] * b.m[3][j];
    return c;
}
template <class R> M4<R> m4_frame(V3<R> t, V3<R> n, V3<R> b) {
    M4<R> m{};
    m.m[0][0] = t.x; m.m[0][1] = t.y; m.m[0][2] = t.z;
    m.m[1][0] = n.x; m.m[1][1] = n.y; m.m[1][2] = n.z;
    m.m[2][0] = b.x; m.m[2][1] = b.y; m.m[2][2] = b.z;
    m.m[3][3] = R(1);
    return m;
}
template <class R> M4<R> m4_triangle_frame(V3<R> v1, V3<R> v2, V3<R> v3) {
    V3<R> tangent = normalize(v1 - v3);
    V3<R> tangent2 = v2 - v3;
    V3<R> normal = normalize(cross(tangent, tangent2));
    V3<R> binormal = normalize(cross(normal, tangent));
    return m4_frame(tangent, normal, binormal);
}
template <class R> void generate_transforms(const R* rest, int G, int F, int V, const tfx_hair_to_triangle_mapping* map,
                                            const R* mesh, const R* skinned, int nverts, R* out, int nthreads) {
    parallel_for(G, nthreads, [&](int g) {
        const size_t s = (size_t)g * (F + 1);
        const size_t tri = (size_t)map[s].triangle_index * 3;
        if ((long long)tri + 2 >= (long long)nverts) return;  // a triangle past the mesh: record not written
        const R a = map[s].barycentric[0], b = map[s].barycentric[1], c = map[s].barycentric[2];
        V3<R> v1 = xyz(load4(mesh, tri)), v2 = xyz(load4(mesh, tri + 1)), v3 = xyz(load4(mesh, tri + 2));
        V3<R> s1 = xyz(load4(skinned, tri)), s2 = xyz(load4(skinned, tri + 1)), s3 = xyz(load4(skinned, tri + 2));
        V3<R> tfmPos = (s1 * a + s2 * b) + s3 * c;
        V3<R> initialPos = xyz(load4(rest, s * V));
        M4<R> triM = m4_triangle_frame(v1, v2, v3);
        M4<R> tfmTriM = m4_triangle_frame(s1, s2, s3);
        M4<R> triT{};
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) triT.m[i][j] = triM.m[j][i];
        M4<R> rot = m4_mul(triT, tfmTriM);
        M4<R> tr{};
        for (int i = 0; i < 4; ++i) tr.m[i][i] = R(1);
        tr.m[3][0] = -initialPos.x; tr.m[3][1] = -initialPos.y; tr.m[3][2] = -initialPos.z;
        rot = m4_mul(tr, rot);
        tr.m[3][0] = tfmPos.x; tr.m[3][1] = tfmPos.y; tr.m[3][2] = tfmPos.z;
        M4<R> tfm = m4_mul(rot, tr);
        R q[4];
        q[3] = std::sqrt(((R(1) + tfm.m[0][0]) + tfm.m[1][1]) + tfm.m[2][2]) / R(2);
        q[0] = (tfm.m[2][1] - tfm.m[1][2]) / (R(4) * q[3]);
        q[1] = (tfm.m[0][2] - tfm.m[2][0]) / (R(4) * q[3]);
        q[2] = (tfm.m[1][0] - tfm.m[0][1]) / (R(4) * q[3]);
        R* o = out + s * 20;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) o[4 * i + j] = tfm.m[i][j];
        for (int k = 0; k < 4; ++k) o[16 + k] = q[k];
    });
}

// ---------------------------------------------------------------------------
// Host parameter packing, TressFXHairObject.cpp:172-235 and :380-437
// (Quaternion.cpp :79-89 SetRotation(axis, angle), :262-284 products).
// ---------------------------------------------------------------------------
struct HQ { float x, y, z, w; };
struct HV { float x, y, z; };
inline HQ hq_axis_angle(HV axis, float angle) {
    float half = 0.5f * angle;
    float sh = std::sin(half);
    return HQ{sh * axis.x, sh * axis.y, sh * axis.z, std::cos(half)};
}
inline HQ hq_mul(HQ a, HQ b) {
    HQ q;
    q.w = a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z;
    q.x = a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y;
    q.y = a.w * b.y + a.y * b.w + a.z * b.x - a.x * b.z;
    q.z = a.w * b.z + a.z * b.w + a.x * b.y - a.y * b.x;
    return q;
}
inline HV hv_cross(HV a, HV b) { return HV{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline HV hq_rot(HQ q, HV v) {
    HV qv{q.x, q.y, q.z};
    HV uv = hv_cross(qv, v);
    HV uuv = hv_cross(qv, uv);
    float s = 2.0f * q.w;
    uv = HV{uv.x * s, uv.y * s, uv.z * s};
    uuv = HV{uuv.x * 2.0f, uuv.y * 2.0f, uuv.z * 2.0f};
    return HV{(v.x + uv.x) + uuv.x, (v.y + uv.y) + uuv.y, (v.z + uv.z) + uuv.z};
}
// Vector3::Normalize, Vector3D.cpp:48-62 (leaves a zero vector unchanged)
inline HV hv_normalize(HV v) {
    float d = std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    if (d == 0) return v;
    return HV{v.x / d, v.y / d, v.z / d};
}

void set_wind(tfx_simulation_params& cb, const float dir[3], float mag, int frame) {
    float s = std::sin((float)frame * 0.01f);
    float wM = mag * (std::pow(s, 2.0f) + 0.5f);
    HV wd = hv_normalize(HV{dir[0], dir[1], dir[2]});
    HV X{1.0f, 0.0f, 0.0f};
    HV xw = hv_cross(X, wd);
    HQ rotX{0, 0, 0, 1};
    float angle = std::asin(std::sqrt(xw.x * xw.x + xw.y * xw.y + xw.z * xw.z));
    if (angle > 0.001) rotX = hq_axis_angle(hv_normalize(xw), angle);
    const float cone = 40.f * 3.14159265359f / 180;
    HV axes[4] = {{0, 1.0f, 0}, {0, -1.0f, 0}, {0, 0, 1.0f}, {0, 0, -1.0f}};
    float* outs[4] = {cb.wind, cb.wind1, cb.wind2, cb.wind3};
    for (int k = 0; k < 4; ++k) {
        HQ rot = hq_axis_angle(axes[k], cone);
        HV nd = hq_rot(hq_mul(rotX, rot), X);
        outs[k][0] = nd.x * wM; outs[k][1] = nd.y * wM; outs[k][2] = nd.z * wM; outs[k][3] = 0;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// C entry points (called from tests/ and bench.py's CPU-baseline leg only)
// ---------------------------------------------------------------------------
typedef struct orc_hair {
    int num_guides, num_follow, num_vertices, pad;
    void* pos; void* prev; void* prevprev; void* tangent; void* strand_level;
    const void* rest; const void* rest_lengths; const void* follow_offsets;
    const float* skin;
} orc_hair;

template <class R> static Hair<R> to_hair(const orc_hair* o) {
    Hair<R> h;
    h.G = o->num_guides; h.F = o->num_follow; h.V = o->num_vertices;
    h.pos = (R*)o->pos; h.prev = (R*)o->prev; h.prevprev = (R*)o->prevprev; h.tangent = (R*)o->tangent;
    h.sld = (R*)o->strand_level; h.rest = (const R*)o->rest; h.restlen = (const R*)o->rest_lengths;
    h.follow_off = (const R*)o->follow_offsets; h.skin = o->skin;
    return h;
}

// BoneSkinning, TressFXBoneSkinning.hlsl:57-94: every collision-mesh vertex
// (StandardVertex: float4 position, float4 normal) by the weight-normalised
// bone blend; position [p,1]·M and normal [n,0]·M, w components kept.
template <class R> void mesh_skin(const float* rest, const float* skin, int n, const float* bones, R* out,
                                  int nthreads) {
    parallel_for(n, nthreads, [&](int i) {
        const BlendedBone<R> b = blend_bones<R>(skin + 8 * i, bones);
        const float* v = rest + 8 * i;
        R* o = out + 8 * i;
        for (int c = 0; c < 3; ++c) {
            o[c] = (((R)v[0] * b.m[0][c] + (R)v[1] * b.m[1][c]) + (R)v[2] * b.m[2][c]) + R(1) * b.m[3][c];
            o[4 + c] = (((R)v[4] * b.m[0][c] + (R)v[5] * b.m[1][c]) + (R)v[6] * b.m[2][c]) + R(0) * b.m[3][c];
        }
        o[3] = (R)v[3];
        o[7] = (R)v[7];
    });
}

extern "C" {



int orc32_run_stage(const orc_hair* h, const tfx_simulation_params* p, const float* bones, int stage, int nthreads) {
    Hair<float> hh = to_hair<float>(h); Params P{p, bones}; run_stage(hh, P, stage, nthreads); return 0;
}
int orc64_run_stage(const orc_hair* h, const tfx_simulation_params* p, const float* bones, int stage, int nthreads) {
    Hair<double> hh = to_hair<double>(h); Params P{p, bones}; run_stage(hh, P, stage, nthreads); return 0;
}
int orc32_simulate(const orc_hair* h, const tfx_simulation_params* p, const float* bones, int nthreads) {
    Hair<float> hh = to_hair<float>(h); Params P{p, bones}; simulate(hh, P, nthreads); return 0;
}
int orc64_simulate(const orc_hair* h, const tfx_simulation_params* p, const float* bones, int nthreads) {
    Hair<double> hh = to_hair<double>(h); Params P{p, bones}; simulate(hh, P, nthreads); return 0;
}

int orc32_sdf_build(const float* verts, const uint32_t* idx, int ntri, const float* origin, float cell,
                    int nx, int ny, int nz, float* out, int nthreads) {
    Grid<float> g{V3<float>{origin[0], origin[1], origin[2]}, cell, nx, ny, nz};
    sdf_build<float>(verts, idx, ntri, g, out, nthreads); return 0;
}
int orc64_sdf_build(const double* verts, const uint32_t* idx, int ntri, const double* origin, double cell,
                    int nx, int ny, int nz, double* out, int nthreads) {
    Grid<double> g{V3<double>{origin[0], origin[1], origin[2]}, cell, nx, ny, nz};
    sdf_build<double>(verts, idx, ntri, g, out, nthreads); return 0;
}
int orc32_sdf_collide(float* pos, float* prev, int N, int V, const float* sdf, const float* origin, float cell,
                      int nx, int ny, int nz, float margin, int nthreads) {
    Grid<float> g{V3<float>{origin[0], origin[1], origin[2]}, cell, nx, ny, nz};
    sdf_collide<float>(pos, prev, N, V, sdf, g, margin, nthreads); return 0;
}
int orc64_sdf_collide(double* pos, double* prev, int N, int V, const double* sdf, const double* origin, double cell,
                      int nx, int ny, int nz, double margin, int nthreads) {
    Grid<double> g{V3<double>{origin[0], origin[1], origin[2]}, cell, nx, ny, nz};
    sdf_collide<double>(pos, prev, N, V, sdf, g, margin, nthreads); return 0;
}
int orc32_sdf_collide_mixed(float* pos, float* prev, int N, const float* sdf, const float* origin, float cell,
                            int nx, int ny, int nz, float margin, int nthreads) {
    Grid<float> g{V3<float>{origin[0], origin[1], origin[2]}, cell, nx, ny, nz};
    sdf_collide_mixed<float>(pos, prev, N, sdf, g, margin, nthreads); return 0;
}
int orc64_sdf_collide_mixed(double* pos, double* prev, int N, const double* sdf, const double* origin, double cell,
                            int nx, int ny, int nz, double margin, int nthreads) {
    Grid<double> g{V3<double>{origin[0], origin[1], origin[2]}, cell, nx, ny, nz};
    sdf_collide_mixed<double>(pos, prev, N, sdf, g, margin, nthreads); return 0;
}
int orc32_generate_transforms(const float* rest, int G, int F, int V, const tfx_hair_to_triangle_mapping* map,
                              const float* mesh, const float* skinned, int nverts, float* out, int nthreads) {
    generate_transforms<float>(rest, G, F, V, map, mesh, skinned, nverts, out, nthreads); return 0;
}
int orc64_generate_transforms(const double* rest, int G, int F, int V, const tfx_hair_to_triangle_mapping* map,
                              const double* mesh, const double* skinned, int nverts, double* out, int nthreads) {
    generate_transforms<double>(rest, G, F, V, map, mesh, skinned, nverts, out, nthreads); return 0;
}
float orc32_get_signed_distance(const float* sdf, const float* origin, float cell, int nx, int ny, int nz, const float* p) {
    Grid<float> g{V3<float>{origin[0], origin[1], origin[2]}, cell, nx, ny, nz};
    return get_signed_distance<float>(g, sdf, V3<float>{p[0], p[1], p[2]});
}

uint32_t orc_float_flip3(float f) { return float_flip3(f); }
uint32_t orc_ifloat_flip3(uint32_t u) { return ifloat_flip3(u); }

float orc32_signed_distance_point_triangle(const float* p, const float* a, const float* b, const float* c) {
    return signed_distance_point_triangle<float>(V3<float>{p[0], p[1], p[2]}, V3<float>{a[0], a[1], a[2]},
                                                 V3<float>{b[0], b[1], b[2]}, V3<float>{c[0], c[1], c[2]});
}
double orc64_signed_distance_point_triangle(const double* p, const double* a, const double* b, const double* c) {
    return signed_distance_point_triangle<double>(V3<double>{p[0], p[1], p[2]}, V3<double>{a[0], a[1], a[2]},
                                                  V3<double>{b[0], b[1], b[2]}, V3<double>{c[0], c[1], c[2]});
}
void orc64_make_quaternion(const double* m16, double* q) {
    double m[4][4];
    for (int r = 0; r < 4; ++r) for (int c = 0; c < 4; ++c) m[r][c] = m16[4 * r + c];
    V4<double> x = make_quaternion<double>(m);
    q[0] = x.x; q[1] = x.y; q[2] = x.z; q[3] = x.w;
}
void orc64_quat_from_two_unit_vectors(const double* u, const double* v, double* q) {
    V4<double> x = quat_from_two_unit_vectors<double>(V3<double>{u[0], u[1], u[2]}, V3<double>{v[0], v[1], v[2]});
    q[0] = x.x; q[1] = x.y; q[2] = x.z; q[3] = x.w;
}
void orc64_mult_quat_vec(const double* q, const double* v, double* out) {
    V3<double> r = mult_quat_vec<double>(V4<double>{q[0], q[1], q[2], q[3]}, V3<double>{v[0], v[1], v[2]});
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
int orc64_capsule_collision(const double* cur, const double* old, const double* c0, const double* c1, double* out) {
    V3<double> np;
    bool hit = capsule_collision<double>(V4<double>{cur[0], cur[1], cur[2], cur[3]}, V4<double>{old[0], old[1], old[2], old[3]},
                                         np, V4<double>{c0[0], c0[1], c0[2], c0[3]}, V4<double>{c1[0], c1[1], c1[2], c1[3]}, 0.4);
    out[0] = np.x; out[1] = np.y; out[2] = np.z;
    return hit ? 1 : 0;
}

// UpdateSimulationParameters (TressFXHairObject.cpp:380-437) for an object
// with V vertices per strand and F follow strands per guide, at `frame`.
void orc_pack_sim_params(const tfx_simulation_settings* s, float dt, int frame, int V, int F,
                         tfx_simulation_params* cb) {
    cb->vsp[0] = s->vsp_coeff;
    cb->vsp[1] = s->vsp_accel_threshold;
    cb->shape[0] = s->damping;
    cb->shape[1] = s->local_constraint_stiffness;
    cb->shape[2] = s->global_constraint_stiffness;
    cb->shape[3] = s->global_constraints_range;
    cb->grav_time_tip[0] = s->gravity_magnitude;
    cb->grav_time_tip[1] = dt;
    cb->grav_time_tip[2] = s->tip_separation;
    cb->sim_ints[2] = 0;
    cb->counts[0] = TFX_SIM_THREAD_GROUP_SIZE / V;
    cb->counts[2] = V;
    cb->counts[1] = F;
    cb->clamp_position_delta = 20.0f;
    if (V >= TFX_MIN_VERTS_PER_STRAND_FOR_GPU_ITERATION) {
        cb->sim_ints[1] = s->local_constraints_iterations;
        cb->cpu_local_shape_iterations = 1;
    } else {
        cb->sim_ints[1] = 1;
        cb->cpu_local_shape_iterations = s->local_constraints_iterations;
    }
    cb->sim_ints[0] = s->length_constraints_iterations;
    set_wind(*cb, s->wind_direction, s->wind_magnitude, frame);
    if (frame < 2) cb->reset_positions = 1.0f;
}

// TressFXSDFCollision constructor grid sizing (TressFXSDFCollision.cpp:38-71,
// with the comma-expression padding of :58-59) and UpdateSDFGrid (:88-93).
void orc_sdf_grid_from_bbox(const float* init_min, const float* init_max, const float* cur_min, int numCellsInX,
                            float* origin, float* cell, int* dims, int* total_cells) {
    float c = (init_max[0] - init_min[0]) / numCellsInX;
    int extra = (int)(0.8f * (float)numCellsInX);
    float pad = extra * c;
    float bmin[3], bmax[3];
    for (int k = 0; k < 3; ++k) { bmin[k] = init_min[k] - pad; bmax[k] = init_max[k] + pad; }
    for (int k = 0; k < 3; ++k) dims[k] = (int)((bmax[k] - bmin[k]) / c);
    *cell = c;
    int tot = (int)(1.4f * dims[0] * dims[1] * dims[2]);
    *total_cells = tot;
    for (int k = 0; k < 3; ++k) origin[k] = cur_min[k] - pad;
}

int orc32_mesh_skin(const float* rest, const float* skin, int n, const float* bones, float* out, int nthreads) {
    mesh_skin<float>(rest, skin, n, bones, out, nthreads); return 0;
}
int orc64_mesh_skin(const float* rest, const float* skin, int n, const float* bones, double* out, int nthreads) {
    mesh_skin<double>(rest, skin, n, bones, out, nthreads); return 0;
}

}  // extern "C"
