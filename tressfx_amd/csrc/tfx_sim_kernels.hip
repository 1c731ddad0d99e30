// tfx_sim_kernels.hip — the six simulation stages as separate gfx950 kernels,
// launched with the reference's structure (one launch per stage per object,
// TressFXSimulation.cpp:51-112).  Lanes map to guide vertices strand-major
// (lane = strand * V + v inside a 256-thread block, 256/V strands per block),
// so each strand's float4 streams are read as contiguous 16*V-byte runs.
// Grids cover guide strands only (the reference over-dispatches by (F+1) and
// relies on D3D12's dropped out-of-bounds writes, TressFXSimulation.hlsl:409).
#include "tfx_device.h"

namespace tfx {
inline namespace TFX_KERNEL_NS {

constexpr int kBlock = 256;

// IntegrationAndGlobalShapeConstraints, TressFXSimulation.hlsl:609-666
__global__ void __launch_bounds__(kBlock) k_integrate(DevHair h) {
    const int V = h.V;
    const int gid = blockIdx.x * kBlock + threadIdx.x;
    const int g = gid / V, v = gid - g * V;
    if (g >= h.G) return;
    const DevParams& P = *h.params;
    const size_t strand = (size_t)g * (h.F + 1);
    const size_t i = strand * V + v;
    BlendedBone bb;
    blend_bones(h.skin[strand], h.bones, bb);
    float4 initialPos = h.rest[i];
    f3 sk = bone_transform(bb.m, bb.t, xyz(initialPos));
    initialPos.x = sk.x; initialPos.y = sk.y; initialPos.z = sk.z;
    float4 shared = h.pos[i];
    float4 currentPos = shared;
    const float4 prevOld = h.prev[i];
    float4 oldPos = prevOld;
    if (P.reset != 0.0f) { currentPos = initialPos; oldPos = initialPos; }
    if (movable(currentPos)) {
        const float g_ = P.gravity, dt = P.dt;
        f3 force{g_ * 0.0f, g_ * -1.0f, g_ * 0.0f};
        f3 c = xyz(currentPos), o = xyz(oldPos);
        f3 n = add(add(c, mul(P.decay, sub(c, o))), mul(mul(force, dt), dt));
        shared.x = n.x; shared.y = n.y; shared.z = n.z;
    } else {
        shared = initialPos;
    }
    if (P.global_stiffness > 0.0f && P.global_range != 0.0f) {
        if (movable(shared)) {
            if ((float)v < P.global_range * (float)V) {
                f3 del = mul(P.global_stiffness, sub(xyz(initialPos), xyz(shared)));
                shared.x += del.x; shared.y += del.y; shared.z += del.z;
            }
        }
    }
    h.prevprev[i] = prevOld;
    h.prev[i] = currentPos;
    h.pos[i] = shared;
}

// CalculateStrandLevelData :677-728 — one lane per guide strand.
__global__ void __launch_bounds__(kBlock) k_strand_level(DevHair h) {
    const int g = blockIdx.x * kBlock + threadIdx.x;
    if (g >= h.G) return;
    const DevParams& P = *h.params;
    const size_t strand = (size_t)g * (h.F + 1);
    const size_t root = strand * h.V;
    const float4 oo1 = h.prevprev[root + 1];
    const float4 o0 = h.prev[root], o1 = h.prev[root + 1];
    const float4 n0 = h.pos[root], n1 = h.pos[root + 1];
    f3 u = normalize3(sub(xyz(o1), xyz(o0)));
    f3 v = normalize3(sub(xyz(n1), xyz(n0)));
    float4 rot = quat_two_vectors(u, v);
    f3 trans = sub(xyz(n0), qrot(rot, xyz(o0)));
    float ax = (n1.x - 2.0f * o1.x) + oo1.x;
    float ay = (n1.y - 2.0f * o1.y) + oo1.y;
    float az = (n1.z - 2.0f * o1.z) + oo1.z;
    float aw = (n1.w - 2.0f * o1.w) + oo1.w;
    float accel = sqrtf(((ax * ax + ay * ay) + az * az) + aw * aw);
    float vspCoeff = P.vsp_coeff;
    if (accel > P.vsp_accel_threshold) vspCoeff = 1.0f;
    BlendedBone bb;
    blend_bones(h.skin[strand], h.bones, bb);
    float4 sq = quat_from_matrix3(bb.m);
    tfx_strand_level_data& d = h.sld[strand];
    d.skinning_quat[0] = sq.x; d.skinning_quat[1] = sq.y; d.skinning_quat[2] = sq.z; d.skinning_quat[3] = sq.w;
    d.vsp_quat[0] = rot.x; d.vsp_quat[1] = rot.y; d.vsp_quat[2] = rot.z; d.vsp_quat[3] = rot.w;
    d.vsp_translation[0] = trans.x; d.vsp_translation[1] = trans.y; d.vsp_translation[2] = trans.z;
    d.vsp_translation[3] = vspCoeff;
}

// VelocityShockPropagation :739-762
__global__ void __launch_bounds__(kBlock) k_vsp(DevHair h) {
    const int V = h.V;
    const int gid = blockIdx.x * kBlock + threadIdx.x;
    const int g = gid / V, v = gid - g * V;
    if (g >= h.G || v < 2) return;
    const size_t strand = (size_t)g * (h.F + 1);
    const size_t i = strand * V + v;
    const tfx_strand_level_data& d = h.sld[strand];
    const float4 q = make_float4(d.vsp_quat[0], d.vsp_quat[1], d.vsp_quat[2], d.vsp_quat[3]);
    const f3 t{d.vsp_translation[0], d.vsp_translation[1], d.vsp_translation[2]};
    const float c = d.vsp_translation[3];
    f3 pn = xyz(h.pos[i]), po = xyz(h.prev[i]);
    f3 rn = add(mul(1.0f - c, pn), mul(c, add(qrot(q, pn), t)));
    f3 ro = add(mul(1.0f - c, po), mul(c, add(qrot(q, po), t)));
    store_xyz(&h.pos[i], rn);
    store_xyz(&h.prev[i], ro);
}

// LocalShapeConstraints :773-829 — one lane per guide strand, Gauss–Seidel
// along the strand in global memory (one pass; the host repeats it).
__global__ void __launch_bounds__(kBlock) k_lsc(DevHair h) {
    const int g = blockIdx.x * kBlock + threadIdx.x;
    if (g >= h.G) return;
    const int V = h.V;
    const DevParams& P = *h.params;
    const size_t strand = (size_t)g * (h.F + 1);
    const size_t root = strand * V;
    const float stiff = 0.5f * fmn(P.local_stiffness, 0.95f);
    const tfx_strand_level_data& d = h.sld[strand];
    const float4 bq = make_float4(d.skinning_quat[0], d.skinning_quat[1], d.skinning_quat[2], d.skinning_quat[3]);
    // The window (i-1, i, i+1) slides along the strand; the values step i
    // stores are exactly those step i+1 would read back.
    float4 pm = h.pos[root];
    float4 pos = h.pos[root + 1];
    f3 bindM = qrot(bq, xyz(h.rest[root]));
    f3 bind = qrot(bq, xyz(h.rest[root + 1]));
    for (int lv = 1; lv < V - 1; ++lv) {
        const size_t i = root + lv;
        float4 pp = h.pos[i + 1];
        f3 bindP = qrot(bq, xyz(h.rest[i + 1]));
        f3 lastVec = sub(xyz(pos), xyz(pm));
        f3 vecBindPose = sub(bindP, bind);
        f3 lastVecBindPose = sub(bind, bindM);
        f3 org = add(rotate_from_to(normalize3(lastVecBindPose), normalize3(lastVec), vecBindPose), xyz(pos));
        f3 del = mul(stiff, sub(org, xyz(pp)));
        if (movable(pos)) { pos.x -= del.x; pos.y -= del.y; pos.z -= del.z; }
        if (movable(pp)) { pp.x += del.x; pp.y += del.y; pp.z += del.z; }
        store_xyz(&h.pos[i], xyz(pos));
        store_xyz(&h.pos[i + 1], xyz(pp));
        pm = pos; pos = pp;
        bindM = bind; bind = bindP;
    }
}

// LengthConstriantsWindAndCollision :875-975.  256/V strands per block staged
// in LDS; wind reads the neighbour before any write (Jacobi), length
// constraints are red–black with a barrier per pass.
__global__ void __launch_bounds__(kBlock) k_length(DevHair h) {
    __shared__ float4 sPos[kBlock];
    __shared__ float sLen[kBlock];
    const int V = h.V;
    const int t = threadIdx.x;
    const int gid = blockIdx.x * kBlock + t;
    const int g = gid / V, v = gid - g * V;
    const bool valid = g < h.G;
    const DevParams& P = *h.params;
    const size_t strand = (size_t)g * (h.F + 1);
    const size_t i = strand * V + v;
    if (valid) { sPos[t] = h.pos[i]; sLen[t] = h.restlen[i]; }
    __syncthreads();
    if (P.wind_on) {
        f3 force{0.0f, 0.0f, 0.0f};
        const bool act = valid && v >= 2 && v < V - 1;
        if (act) {
            float a = (float)(uint32_t)(strand % 20) / 20.0f;
            f3 w = add(add(add(mul(a, xyz(P.wind0)), mul(1.0f - a, xyz(P.wind1))), mul(a, xyz(P.wind2))),
                       mul(1.0f - a, xyz(P.wind3)));
            f3 d = sub(xyz(sPos[t]), xyz(sPos[t + 1]));
            f3 c = cross3(cross3(d, w), d);
            force = f3{-c.x, -c.y, -c.z};
        }
        __syncthreads();
        if (act) {
            f3 f = mul(mul(force, P.dt), P.dt);
            sPos[t].x += f.x; sPos[t].y += f.y; sPos[t].z += f.z;
        }
    }
    __syncthreads();
    const int a = V / 2, b = (V - 1) / 2;
    for (int it = 0; it < P.length_iters; ++it) {
        // pass A: pair (2k, 2k+1) owned by the lane of vertex 2k
        if (valid && (v & 1) == 0 && (v >> 1) < a) {
            float4 p0 = sPos[t], p1 = sPos[t + 1];
            distance_constraint(p0, p1, sLen[t]);
            sPos[t] = p0; sPos[t + 1] = p1;
        }
        __syncthreads();
        // pass B: pair (2k+1, 2k+2) owned by the lane of vertex 2k+1
        if (valid && (v & 1) == 1 && (v >> 1) < b) {
            float4 p0 = sPos[t], p1 = sPos[t + 1];
            distance_constraint(p0, p1, sLen[t]);
            sPos[t] = p0; sPos[t + 1] = p1;
        }
        __syncthreads();
    }
    float4 oldPos = make_float4(0.f, 0.f, 0.f, 0.f);
    bool hit = false;
    if (valid) {
        oldPos = h.prev[i];
        float4 cur = sPos[t];
        for (int c = 0; c < P.num_capsules; ++c) {
            f3 np;
            if (capsule_collision(cur, oldPos, np, P.cap0[c], P.cap1[c])) {
                cur.x = np.x; cur.y = np.y; cur.z = np.z;
                hit = true;
            }
        }
        sPos[t] = cur;
    }
    __syncthreads();
    if (!valid) return;
    const int ti = (v == V - 1) ? t - 1 : t;
    f3 tangent = sub(xyz(sPos[ti + 1]), xyz(sPos[ti]));
    store_xyz(&h.tangent[i], normalize3(tangent));
    const float4 sp = sPos[t];
    f3 delta = sub(xyz(sp), xyz(oldPos));
    float speedSqr = dot3(delta, delta);
    const float C = P.clamp;
    if (speedSqr > C * C) {
        float sc = (C * C) / speedSqr;
        delta = mul(delta, sc);
        store_xyz(&h.prev[i], sub(xyz(sp), delta));
    }
    h.pos[i] = sp;
    if (hit) h.prev[i] = sp;
}

// UpdateFollowHairVertices :978-1001
__global__ void __launch_bounds__(kBlock) k_follow(DevHair h) {
    const int V = h.V, F = h.F;
    const int gid = blockIdx.x * kBlock + threadIdx.x;
    const int g = gid / V, v = gid - g * V;
    if (g >= h.G || F <= 0) return;
    const DevParams& P = *h.params;
    const size_t strand = (size_t)g * (F + 1);
    const size_t i = strand * V + v;
    const float4 gp = h.pos[i];
    const float4 gt = h.tangent[i];
    const float factor = P.tip_sep * ((float)v / (float)V) + 1.0f;
    for (int j = 0; j < F; ++j) {
        const size_t fv = i + (size_t)V * (j + 1);
        const f3 off = xyz(h.follow_off[strand + j + 1]);
        store_xyz(&h.pos[fv], add(xyz(gp), mul(factor, off)));
        h.tangent[fv] = gt;
    }
}

// GenerateTransforms, TressFXSimulation.hlsl:1035-1142
// (USE_MESH_BASED_HAIR_TRANSFORM): one lane per guide strand.  Row-major 4x4
// products summed left to right.
struct M4 { float m[4][4]; };

__device__ __forceinline__ M4 m4_mul(const M4& a, const M4& b) {
    M4 c;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
            c.m[i][j] = ((a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j]) + a.m[i][2] * b.m[2][j]) + a.m[i][3] * b.m[3][j];
    return c;
}

__device__ __forceinline__ M4 m4_triangle_frame(f3 v1, f3 v2, f3 v3) {
    const f3 tangent = normalize3(sub(v1, v3));
    const f3 tangent2 = sub(v2, v3);
    const f3 normal = normalize3(cross3(tangent, tangent2));
    const f3 binormal = normalize3(cross3(normal, tangent));
    M4 m;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) m.m[i][j] = 0.0f;
    m.m[0][0] = tangent.x; m.m[0][1] = tangent.y; m.m[0][2] = tangent.z;
    m.m[1][0] = normal.x; m.m[1][1] = normal.y; m.m[1][2] = normal.z;
    m.m[2][0] = binormal.x; m.m[2][1] = binormal.y; m.m[2][2] = binormal.z;
    m.m[3][3] = 1.0f;
    return m;
}

__global__ void __launch_bounds__(kBlock) k_generate_transforms(const float4* __restrict__ rest, int G, int F, int V,
                                                                const tfx_hair_to_triangle_mapping* __restrict__ map,
                                                                const float4* __restrict__ mesh,
                                                                const float4* __restrict__ skinned, int nverts,
                                                                tfx_strand_transform* __restrict__ out) {
    const int g = blockIdx.x * kBlock + threadIdx.x;
    if (g >= G) return;
    const size_t s = (size_t)g * (F + 1);
    const tfx_hair_to_triangle_mapping mp = map[s];
    const long long tri = (long long)mp.triangle_index * 3;
    if (tri + 2 >= (long long)nverts) return;
    const float a = mp.barycentric[0], b = mp.barycentric[1], c = mp.barycentric[2];
    const f3 v1 = xyz(mesh[tri]), v2 = xyz(mesh[tri + 1]), v3 = xyz(mesh[tri + 2]);
    const f3 s1 = xyz(skinned[tri]), s2 = xyz(skinned[tri + 1]), s3 = xyz(skinned[tri + 2]);
    const f3 tfmPos = add(add(mul(s1, a), mul(s2, b)), mul(s3, c));
    const f3 initialPos = xyz(rest[s * V]);
    const M4 triM = m4_triangle_frame(v1, v2, v3);
    const M4 tfmTriM = m4_triangle_frame(s1, s2, s3);
    M4 triT;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) triT.m[i][j] = triM.m[j][i];
    M4 rot = m4_mul(triT, tfmTriM);
    M4 tr;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) tr.m[i][j] = (i == j) ? 1.0f : 0.0f;
    tr.m[3][0] = -initialPos.x; tr.m[3][1] = -initialPos.y; tr.m[3][2] = -initialPos.z;
    rot = m4_mul(tr, rot);
    tr.m[3][0] = tfmPos.x; tr.m[3][1] = tfmPos.y; tr.m[3][2] = tfmPos.z;
    const M4 tfm = m4_mul(rot, tr);
    float q[4];
    q[3] = sqrtf(((1.0f + tfm.m[0][0]) + tfm.m[1][1]) + tfm.m[2][2]) / 2.0f;
    q[0] = (tfm.m[2][1] - tfm.m[1][2]) / (4.0f * q[3]);
    q[1] = (tfm.m[0][2] - tfm.m[2][0]) / (4.0f * q[3]);
    q[2] = (tfm.m[1][0] - tfm.m[0][1]) / (4.0f * q[3]);
    tfx_strand_transform o;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) o.tfm[4 * i + j] = tfm.m[i][j];
#pragma unroll
    for (int k = 0; k < 4; ++k) o.quat[k] = q[k];
    out[s] = o;
}

}  // namespace TFX_KERNEL_NS
}  // namespace tfx

// Launchers used by the C-ABI layer.
extern "C" hipError_t TFX_KSYM(tfx_launch_stage)(int stage, const tfx::DevHair* h, hipStream_t stream) {
    using namespace tfx;
    const int nv = h->G * h->V;
    const int vblocks = (nv + kBlock - 1) / kBlock;
    const int sblocks = (h->G + kBlock - 1) / kBlock;
    if (h->G <= 0) return hipSuccess;
    switch (stage) {
        case 0: hipLaunchKernelGGL(k_integrate, dim3(vblocks), dim3(kBlock), 0, stream, *h); break;
        case 1: hipLaunchKernelGGL(k_strand_level, dim3(sblocks), dim3(kBlock), 0, stream, *h); break;
        case 2: hipLaunchKernelGGL(k_vsp, dim3(vblocks), dim3(kBlock), 0, stream, *h); break;
        case 3: hipLaunchKernelGGL(k_lsc, dim3(sblocks), dim3(kBlock), 0, stream, *h); break;
        case 4: hipLaunchKernelGGL(k_length, dim3(vblocks), dim3(kBlock), 0, stream, *h); break;
        case 5: if (h->F > 0) hipLaunchKernelGGL(k_follow, dim3(vblocks), dim3(kBlock), 0, stream, *h); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

extern "C" hipError_t TFX_KSYM(tfx_launch_generate_transforms)(const float4* rest, int G, int F, int V,
                                                     const tfx_hair_to_triangle_mapping* map, const float4* mesh,
                                                     const float4* skinned, int nverts, tfx_strand_transform* out,
                                                     hipStream_t stream) {
    using namespace tfx;
    if (G <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_generate_transforms, dim3((G + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, rest, G, F, V,
                       map, mesh, skinned, nverts, out);
    return hipGetLastError();
}
