// tfx_device.h — device-side records and FP32 math for the gfx950 kernels.
//
// The arithmetic helpers fix the operation order that HLSL leaves open, so
// that a kernel evaluates each reference expression in one defined way:
//   dot(a,b) = (a.x*b.x + a.y*b.y) + a.z*b.z
//   normalize(v) = v / sqrt(dot(v,v)) (IEEE division and square root)
//   min/max = IEEE minNum/maxNum (v_min_f32 / v_max_f32)
//   float -> int = truncation, NaN -> 0, saturating (D3D ftoi)
// All HIP sources are compiled with -ffp-contract=off, so no product is
// fused into an add behind the source's back.
//
// The kernel sources are built twice.  The default build divides and takes
// square roots correctly rounded; the second (TFX_FAST_BUILD, with hipcc's
// -fno-hip-fp32-correctly-rounded-divide-sqrt) uses the hardware reciprocal
// and square root.  That build's kernels live in tfx::fast and its C entry
// points carry a _fast suffix; the host chooses a build per launch.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/tressfx_amd.h"

#ifdef TFX_FAST_BUILD
#define TFX_KERNEL_NS fast
#define TFX_KSYM(name) name##_fast
#else
#define TFX_KERNEL_NS exact
#define TFX_KSYM(name) name
#endif

namespace tfx {

struct f3 { float x, y, z; };

// Address-space qualifiers apply to the device pass; the host pass, which
// only type-checks device functions, sees plain pointers.
#if defined(__HIP_DEVICE_COMPILE__)
#define TFX_ADDRSPACE(n) __attribute__((address_space(n)))
#else
#define TFX_ADDRSPACE(n)
#endif

// Loads and stores through the global address space: pointers that come out
// of the object tables are generic, and generic (flat) accesses also count
// against the LDS wait counter, so every wait for one of them waits for LDS
// traffic too.
// A 16-byte record (float4, uint4) is moved as a native 4-dword vector: the
// copy constructor of HIP's vector types would take the record's address as
// a generic pointer again and turn the access into a flat one.
template <typename T>
__device__ __forceinline__ T ldg(const T* p) {
    if constexpr (sizeof(T) == 16 && alignof(T) >= 16) {
        typedef unsigned int u4 __attribute__((ext_vector_type(4)));
        const u4 v = *(const TFX_ADDRSPACE(1) u4*)p;
        T out;
        __builtin_memcpy(&out, &v, 16);
        return out;
    } else if constexpr (sizeof(T) == 8 && alignof(T) >= 8) {
        typedef unsigned int u2 __attribute__((ext_vector_type(2)));
        const u2 v = *(const TFX_ADDRSPACE(1) u2*)p;
        T out;
        __builtin_memcpy(&out, &v, 8);
        return out;
    } else if constexpr (sizeof(T) > 16 && sizeof(T) % 4 == 0) {
        // larger records (the 32-byte skinning record) word by word
        const TFX_ADDRSPACE(1) unsigned int* w = (const TFX_ADDRSPACE(1) unsigned int*)p;
        unsigned int words[sizeof(T) / 4];
#pragma unroll
        for (int i = 0; i < (int)(sizeof(T) / 4); ++i) words[i] = w[i];
        T out;
        __builtin_memcpy(&out, words, sizeof(T));
        return out;
    } else {
        return *(const TFX_ADDRSPACE(1) T*)p;
    }
}
template <typename T>
__device__ __forceinline__ void stg(T* p, const T& v) {
    if constexpr (sizeof(T) == 16 && alignof(T) >= 16) {
        typedef unsigned int u4 __attribute__((ext_vector_type(4)));
        u4 w;
        __builtin_memcpy(&w, &v, 16);
        *(TFX_ADDRSPACE(1) u4*)p = w;
    } else {
        *(TFX_ADDRSPACE(1) T*)p = v;
    }
}
// Host-written records a kernel only reads (object tables, per-substep
// parameters): read through the constant address space from a wave-uniform
// address, they become scalar loads that the compiler may hoist and reuse.
template <typename T>
using const_as = const TFX_ADDRSPACE(4) T;
template <typename T>
__device__ __forceinline__ const_as<T>* uniform_const(const T* p) {
    const uint64_t a = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
    return (const_as<T>*)(((uint64_t)hi << 32) | lo);
}
// A whole record copied out of such memory, word by word.
template <typename T>
__device__ __forceinline__ T load_uniform(const T* p) {
    static_assert(sizeof(T) % 4 == 0, "word-sized record");
    const_as<uint32_t>* src = (const_as<uint32_t>*)uniform_const(p);
    T out;
    uint32_t* dst = reinterpret_cast<uint32_t*>(&out);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) dst[i] = src[i];
    return out;
}
__device__ __forceinline__ float4 ld4(const_as<float4>& v) { return make_float4(v.x, v.y, v.z, v.w); }

__device__ __forceinline__ f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 xyz(const float4& a) { return f3{a.x, a.y, a.z}; }
__device__ __forceinline__ f3 add(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 sub(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 mul(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ f3 mul(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 divs(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross3(f3 a, f3 b) {
    return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float len3(f3 a) { return sqrtf(dot3(a, a)); }
#ifndef TFX_FAST_BUILD
__device__ __forceinline__ f3 normalize3(f3 a) { return divs(a, sqrtf(dot3(a, a))); }
#else
// the hardware reciprocal square root: one transcendental instead of a square
// root and a reciprocal
__device__ __forceinline__ f3 normalize3(f3 a) { return mul(a, __builtin_amdgcn_rsqf(dot3(a, a))); }
#endif
// a / b, bit-identical to the IEEE quotient, for a divisor b shared by many
// divisions, given y = recip_fixed(b): q = a*y, then twice r = a - b*q (exact
// in an FMA) and q += r*y.  With y = 1/b correctly rounded, |a| >= 2^-96 (the
// residual stays exactly representable) and the result in [2^-100, FLT_MAX]
// this is the correctly rounded quotient (Markstein's correction); every
// other case (zeros, underflow, overflow, NaNs, divisors outside
// [2^-100, 2^100]) is recomputed by IEEE division.
__host__ __device__ __forceinline__ float recip_fixed(float b) {
    const float m = b < 0.0f ? -b : b;
    return (m >= 0x1p-100f && m <= 0x1p100f) ? 1.0f / b : __builtin_nanf("");
}
__device__ __forceinline__ float div_fixed(float a, float b, float y) {
    float q = a * y;
    float r = __builtin_fmaf(-q, b, a);
    q = __builtin_fmaf(r, y, q);
    r = __builtin_fmaf(-q, b, a);
    q = __builtin_fmaf(r, y, q);
    if (!(__builtin_fabsf(a) >= 0x1p-96f && __builtin_fabsf(q) >= 0x1p-100f &&
          __builtin_fabsf(q) <= 0x1.fffffep127f)) {
        // An empty volatile statement cannot be speculated: it keeps this
        // rarely taken path a branch instead of a division on every lane.
        float aa = a;
        asm volatile("" : "+v"(aa));
        q = aa / b;
    }
    return q;
}
__device__ __forceinline__ float fmn(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ float fmx(float a, float b) { return fmaxf(a, b); }

__device__ __forceinline__ int ftoi(float x) {
    if (!(x == x)) return 0;
    if (x >= 2147483648.0f) return 2147483647;
    if (x <= -2147483648.0f) return (-2147483647 - 1);
    return (int)x;
}

__device__ __forceinline__ bool movable(const float4& p) { return p.w > 0.0f; }

// MakeQuaternion(float4x4), TressFXSimulation.hlsl:247-288, on the blended
// matrix's upper 3x3 (m[r*3+c] = logical element (r, c)).
__device__ __forceinline__ float4 quat_from_matrix3(const float m[9]) {
    float4 q;
    float trace = m[0] + m[4] + m[8];
    if (trace > 0.0f) {
        float r = sqrtf(trace + 1.0f);
        q.w = 0.5f * r;
        r = 0.5f / r;
        q.x = (m[5] - m[7]) * r;
        q.y = (m[6] - m[2]) * r;
        q.z = (m[1] - m[3]) * r;
    } else {
        // (i, j, k) is one of the cyclic triples (0,1,2), (1,2,0), (2,0,1);
        // each is written out so no array is indexed at run time.
        int i = 0;
        if (m[4] > m[0]) i = 1;
        if (m[8] > m[i * 4]) i = 2;
        if (i == 0) {
            float r = sqrtf(m[0] - m[4] - m[8] + 1.0f);
            q.x = 0.5f * r;
            r = 0.5f / r;
            q.w = (m[5] - m[7]) * r;
            q.y = (m[3] + m[1]) * r;
            q.z = (m[6] + m[2]) * r;
        } else if (i == 1) {
            float r = sqrtf(m[4] - m[8] - m[0] + 1.0f);
            q.y = 0.5f * r;
            r = 0.5f / r;
            q.w = (m[6] - m[2]) * r;
            q.z = (m[7] + m[5]) * r;
            q.x = (m[1] + m[3]) * r;
        } else {
            float r = sqrtf(m[8] - m[0] - m[4] + 1.0f);
            q.z = 0.5f * r;
            r = 0.5f / r;
            q.w = (m[1] - m[3]) * r;
            q.x = (m[2] + m[6]) * r;
            q.y = (m[5] + m[7]) * r;
        }
    }
    return q;
}

// MultQuaternionAndVector :305-315
__device__ __forceinline__ f3 qrot(const float4& q, f3 v) {
    f3 qv{q.x, q.y, q.z};
    f3 uv = cross3(qv, v);
    f3 uuv = cross3(qv, uv);
    uv = mul(uv, 2.0f * q.w);
    uuv = mul(uuv, 2.0f);
    return add(add(v, uv), uuv);
}

// NormalizeQuaternion :329-342
__device__ __forceinline__ float4 normalize_quat(float4 q) {
    float n = q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w;
    if (n < 1e-10f) { q.w = 1.0f; return q; }
    float s = 1.0f / sqrtf(n);
    return make_float4(q.x * s, q.y * s, q.z * s, q.w * s);
}

// QuatFromTwoUnitVectors :345-363
__device__ __forceinline__ float4 quat_two_vectors(f3 u, f3 v) {
    float r = 1.0f + dot3(u, v);
    f3 n;
    if (r < 1e-7f) {
        r = 0.0f;
        n = fabsf(u.x) > fabsf(u.z) ? f3{-u.y, u.x, 0.0f} : f3{0.0f, -u.z, u.y};
    } else {
        n = cross3(u, v);
    }
    return normalize_quat(make_float4(n.x, n.y, n.z, r));
}

// MultQuaternionAndVector(QuatFromTwoUnitVectors(u, v), x): x rotated by the
// rotation taking u onto v, the step of LocalShapeConstraints :808-814.  The
// fast build applies that rotation by Rodrigues' formula with the axis and
// sine folded into c = u x v and the cosine d = u.v:
//   R x = d x + c x x + ((c.x) / (1 + d)) c
// (no quaternion, one division); for antiparallel u and v (the branch of
// QuatFromTwoUnitVectors :349-355) the half turn about an axis a
// perpendicular to u, R x = 2 ((a.x) / (a.a)) a - x.
__device__ __forceinline__ f3 rotate_from_to(f3 u, f3 v, f3 x) {
#ifndef TFX_FAST_BUILD
    return qrot(quat_two_vectors(u, v), x);
#else
    const float d = dot3(u, v);
    const float r = 1.0f + d;
    if (r < 1e-7f) {
        const f3 a = fabsf(u.x) > fabsf(u.z) ? f3{-u.y, u.x, 0.0f} : f3{0.0f, -u.z, u.y};
        return sub(mul(2.0f * dot3(a, x) / dot3(a, a), a), x);
    }
    const f3 c = cross3(u, v);
    return add(add(mul(d, x), cross3(c, x)), mul(dot3(c, x) / r, c));
#endif
}

#ifdef TFX_FAST_BUILD
// The fast build's goal position of one LocalShapeConstraints step
// (:799-829): p + R x, R the rotation taking the unit bind direction u onto
// the direction of lastVec = p - pm, x the bind-pose edge.  With v = lastVec
// / |lastVec| and s = u + v,
//   R x = x - ((s.x) / (1 + u.v)) s + 2 (u.x) v,
// evaluated from lastVec's three dot products and one reciprocal square root,
// so that the step's dependent chain is short; antiparallel u and v take the
// half turn of rotate_from_to.
__device__ __forceinline__ f3 lsc_target(f3 u, f3 x, f3 pm, f3 p) {
    const f3 lv = sub(p, pm);
    const float ll = dot3(lv, lv), ul = dot3(u, lv), xl = dot3(x, lv), ux = dot3(u, x);
    const float rs = __builtin_amdgcn_rsqf(ll);
    const float r = 1.0f + ul * rs;
    const f3 v = mul(lv, rs);
    const float k = (ux + xl * rs) / r;
    const float w = 2.0f * ux;
    const f3 base = add(p, x);
    f3 out{base.x + w * v.x - k * (u.x + v.x), base.y + w * v.y - k * (u.y + v.y), base.z + w * v.z - k * (u.z + v.z)};
    const bool anti = r < 1e-7f;
    if (__builtin_expect(__any(anti), 0)) {
        if (anti) {
            const f3 a = fabsf(u.x) > fabsf(u.z) ? f3{-u.y, u.x, 0.0f} : f3{0.0f, -u.z, u.y};
            out = add(sub(mul(2.0f * dot3(a, x) / dot3(a, a), a), x), p);
        }
    }
    return out;
}
#endif

// Blended bone of one strand: ApplyVertexBoneSkinning's matrix path
// (:549-567) — 3x3 rotation/scale part plus translation row.
struct BlendedBone {
    float m[9];   // rows 0..2, cols 0..2
    float t[3];   // row 3, cols 0..2
};

__device__ __forceinline__ const float* bone_ptr(const float* bones, float fidx) {
    int idx = ftoi(fidx);
    idx = max(0, min(idx, TFX_MAX_NUM_BONES - 1));
    return bones + 16 * idx;
}

// Blends the 4x4 matrices element-wise (the unused fourth column is skipped:
// no result of the path reads it).  All four matrices are read up front.  A
// slot of weight <= 0 is not blended, and it reads slot 0's matrix instead of
// its own: a packed bone table may hold only the bones that slot 0 and the
// weighted slots name.
__device__ __forceinline__ void blend_bones(const tfx_bone_skinning_data& sd, const float* bones, BlendedBone& b) {
    float mat[4][12];
    const float* m0 = bone_ptr(bones, sd.bone_index[0]);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float* mi = (i == 0 || !(sd.bone_weight[i] > 0.0f)) ? m0 : bone_ptr(bones, sd.bone_index[i]);
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) mat[i][r * 3 + c] = ldg(&mi[4 * r + c]);
    }
    const float w0 = sd.bone_weight[0];
    float acc[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) acc[k] = mat[0][k] * w0;
    float wsum = w0;
#pragma unroll
    for (int i = 1; i < 4; ++i) {
        const float wi = sd.bone_weight[i];
        if (wi > 0.0f) {
#pragma unroll
            for (int k = 0; k < 12; ++k) acc[k] = acc[k] + mat[i][k] * wi;
            wsum = wsum + wi;
        }
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) b.m[k] = acc[k] / wsum;
#pragma unroll
    for (int k = 0; k < 3; ++k) b.t[k] = acc[9 + k] / wsum;
}

// blend_bones with the matrices read one after another into the running sum
// (same operations in the same order, a quarter of the registers): for
// callers with other state live.
__device__ __forceinline__ void blend_bones_lean(const tfx_bone_skinning_data& sd, const float* bones, BlendedBone& b) {
    const float* m0 = bone_ptr(bones, sd.bone_index[0]);
    const float w0 = sd.bone_weight[0];
    float acc[12];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[r * 3 + c] = ldg(&m0[4 * r + c]) * w0;
    float wsum = w0;
#pragma unroll 1
    for (int i = 1; i < 4; ++i) {
        const float wi = sd.bone_weight[i];
        if (wi > 0.0f) {
            const float* mi = bone_ptr(bones, sd.bone_index[i]);
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) acc[r * 3 + c] = acc[r * 3 + c] + ldg(&mi[4 * r + c]) * wi;
            wsum = wsum + wi;
        }
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) b.m[k] = acc[k] / wsum;
#pragma unroll
    for (int k = 0; k < 3; ++k) b.t[k] = acc[9 + k] / wsum;
}

// mul(float4(p, 1), M).xyz
__device__ __forceinline__ f3 bone_transform(const float m[9], const float t[3], f3 p) {
    return f3{((p.x * m[0] + p.y * m[3]) + p.z * m[6]) + t[0],
              ((p.x * m[1] + p.y * m[4]) + p.z * m[7]) + t[1],
              ((p.x * m[2] + p.y * m[5]) + p.z * m[8]) + t[2]};
}

// ApplyDistanceConstraint :377-387 with ConstraintMultiplier :214-230.
__device__ __forceinline__ void distance_constraint(float4& p0, float4& p1, float target) {
    f3 delta = sub(xyz(p1), xyz(p0));
    float distance = fmx(len3(delta), 1e-7f);
    float stretching = 1.0f - target / distance;
    delta = mul(stretching, delta);
    const bool a = movable(p0), b = movable(p1);
    const float m0 = a ? (b ? 0.5f : 1.0f) : 0.0f;
    const float m1 = a ? (b ? 0.5f : 0.0f) : (b ? 1.0f : 0.0f);
    f3 d0 = mul(m0, delta);
    f3 d1 = mul(m1, delta);
    p0.x += d0.x; p0.y += d0.y; p0.z += d0.z;
    p1.x -= d1.x; p1.y -= d1.y; p1.z -= d1.z;
}

// CapsuleCollision :462-522 (friction 0.4).
__device__ __forceinline__ bool capsule_collision(const float4& cur, const float4& old, f3& newp,
                                                  const float4& c0, const float4& c1) {
    const float friction = 0.4f;
    const float radius0 = c0.w, radius1 = c1.w;
    newp = xyz(cur);
    if (!movable(cur)) return false;
    f3 segment = sub(xyz(c1), xyz(c0));
    f3 delta0 = sub(xyz(cur), xyz(c0));
    f3 delta1 = sub(xyz(c1), xyz(cur));
    float dist0 = dot3(delta0, segment);
    float dist1 = dot3(delta1, segment);
    if (dist0 < 0.0f) {
        if (dot3(delta0, delta0) < radius0 * radius0) {
            f3 n = normalize3(delta0);
            newp = add(mul(radius0, n), xyz(c0));
            return true;
        }
        return false;
    }
    if (dist1 < 0.0f) {
        if (dot3(delta1, delta1) < radius1 * radius1) {
            f3 n = normalize3(f3{-delta1.x, -delta1.y, -delta1.z});
            newp = add(mul(radius1, n), xyz(c1));
            return true;
        }
        return false;
    }
    f3 x = divs(add(mul(dist0, xyz(c1)), mul(dist1, xyz(c0))), dist0 + dist1);
    f3 delta = sub(xyz(cur), x);
    float radius_at_x = (dist0 * radius1 + dist1 * radius0) / (dist0 + dist1);
    if (dot3(delta, delta) < radius_at_x * radius_at_x) {
        f3 n = normalize3(delta);
        f3 vec = sub(xyz(cur), xyz(old));
        f3 segN = normalize3(segment);
        f3 vecTangent = mul(dot3(vec, segN), segN);
        f3 vecNormal = sub(vec, vecTangent);
        newp = add(add(xyz(old), mul(friction, vecTangent)), sub(add(vecNormal, mul(radius_at_x, n)), delta));
        return true;
    }
    return false;
}

// Per-object, per-substep parameter block as the kernels read it: the
// tfx_simulation_params fields the path uses, the host-evaluated decay
// exp(-damping*dt*60) and the object's SDF collision constants.
struct alignas(16) DevParams {
    float4 wind0, wind1, wind2, wind3;
    float damping, local_stiffness, global_stiffness, global_range;
    float gravity, dt, tip_sep, decay;
    int length_iters, lsc_passes, wind_on, num_capsules;
    float vsp_coeff, vsp_accel_threshold, reset, clamp;
    float4 cap0[TFX_MAX_NUM_COLLISION_CAPSULES];
    float4 cap1[TFX_MAX_NUM_COLLISION_CAPSULES];
    // SDF collision for the fused substep (CollideHairVerticesWithSdf_forward)
    float sdf_ox, sdf_oy, sdf_oz, sdf_cell;
    int sdf_nx, sdf_ny, sdf_nz, has_sdf;
    float sdf_margin;
    int tangent_w_uniform;  // all guide tangent w equal -> tangent_w
    float tangent_w;
    float sdf_inv_cell;
    int sdf_bx, sdf_by, sdf_bz, pad2;
};

// Device pointers and counts of one hair object.
struct DevHair {
    float4* pos;
    float4* prev;
    float4* prevprev;
    float4* tangent;
    tfx_strand_level_data* sld;
    const float4* rest;
    const float* restlen;
    const float4* follow_off;
    const tfx_bone_skinning_data* skin;
    const float* bones;          // 512 x 16
    const DevParams* params;
    const float* sdf;            // float cells of the hair's SDF (fused path)
    const uint8_t* sdf_bricks;   // its dilated brick occupancy
    const uint32_t* sdf_brick_bits;  // the same, one bit per brick
    // The fused kernel's processing order of the guide strands: work item
    // (hair, g0) takes strands order[g0], order[g0 + 1], ... (nullptr: g0,
    // g0 + 1, ...) when the launch collides with the SDF.  Strands are
    // independent within a substep, so the order changes no result; a
    // spatially coherent one makes the strands of a block share SDF cells.
    const int* order;
    int G, F, V;
    // 1 when every follow vertex's w equals its guide vertex's w (as
    // GenerateFollowHairs makes them, TressFXAsset.cpp:234): the fused
    // kernel then writes a follow position as a whole record with the
    // guide's w instead of x, y and z one by one.
    int follow_w_guide;
};

// One entry of the fused kernel's work list: a block of guide strands of one
// object.
struct WorkItem {
    int hair;   // index into the DevHair table
    int g0;     // first guide strand of the block
};

// Device words a caller of the fused launches keeps for their work queues:
// two per strand-length class (V = 4, 8, 16, 32, 64), zero when a launch
// starts, and left at zero by every launch that completes.  Launches that
// share the words must not overlap in time (one stream does that).
constexpr int kFusedQueueWords = 16;
__host__ __device__ constexpr int fused_queue_class(int V) {
    return V == 4 ? 0 : V == 8 ? 1 : V == 16 ? 2 : V == 32 ? 3 : 4;
}

// SDF grid constants (TressFXSDFCollisionParams subset).
struct SdfGrid {
    float ox, oy, oz, cell;
    int nx, ny, nz, pad;
    // Occupancy of 4^3-cell bricks, dilated by kBrickHalo cells: brick b is
    // non-zero if any cell within kBrickHalo cells of it holds a distance
    // other than 1e10.  Null = no map (every lookup is exact).
    const uint8_t* bricks;
    const uint32_t* brick_bits; // the same occupancy, one bit per brick
    float inv_cell;             // 1 / cell, for the conservative brick test only
    int bx, by, bz;             // bricks per axis
};

constexpr int kBrickShift = 2;  // 4^3 cells per brick
constexpr int kBrickHalo = 2;

// One SDF of a batched build: its cells, mesh and grid.
struct SdfJob {
    uint32_t* cells;
    const tfx_standard_vertex* verts;
    const uint32_t* idx;
    int ntri;
    int ncell;
    SdfGrid g;
    // Sparse bookkeeping, per 4^3-cell brick (the occupancy map's bricks):
    // band_cur receives the bricks this build's construct pass can touch;
    // band_prev holds those of the grid's previous build (null: unknown, every
    // other brick is reset).  Null band_cur: the whole grid is initialised
    // and finalised.
    uint8_t* band_cur;
    const uint8_t* band_prev;
    uint8_t* band_next;   // the band the next build writes: cleared at the end of this one
};

// Whether a batched build of njobs grids of at most max_cells nodes keeps the
// band bookkeeping: it saves writing and un-flipping whole grids, at the cost
// of one more launch, which pays once the grids hold 2^16 bricks in all.
__host__ __device__ inline bool sdf_multi_sparse(int max_cells, int njobs) {
    return (long long)(max_cells / 64) * njobs >= (1 << 16);
}

// One mesh of a batched skinning pass (TressFXBoneSkinning.hlsl:57-94).
struct MeshSkinJob {
    tfx_standard_vertex* out;
    const tfx_standard_vertex* rest;
    const tfx_bone_skinning_data* skin;
    const float* bones;  // row-major row-vector float4x4 per bone
    int n;
    int pad;
};

// GetSdfCoordinates :103-113
__device__ __forceinline__ void sdf_coords(const SdfGrid& g, f3 p, int& cx, int& cy, int& cz) {
    cx = ftoi((p.x - g.ox) / g.cell);
    cy = ftoi((p.y - g.oy) / g.cell);
    cz = ftoi((p.z - g.oz) / g.cell);
}
// GetSdfCellIndex :123-129 (wrapping int arithmetic)
__device__ __forceinline__ int sdf_index(const SdfGrid& g, int x, int y, int z) {
    uint32_t perLine = (uint32_t)g.nx, perPlane = (uint32_t)g.nx * (uint32_t)g.ny;
    return (int)(perPlane * (uint32_t)z + perLine * (uint32_t)y + (uint32_t)x);
}
// GetSdfCellPosition :115-121
__device__ __forceinline__ f3 sdf_cell_pos(const SdfGrid& g, int x, int y, int z) {
    return f3{(float)x * g.cell + g.ox, (float)y * g.cell + g.oy, (float)z * g.cell + g.oz};
}

__device__ __forceinline__ float lerp_(float a, float b, float t) { return a * (1.0f - t) + b * t; }

// GetSignedDistance :472-527 (1e10 = outside the band or the grid).
// All eight corners are in the grid once the range test passed, so the loads
// may be issued before the reference's "any corner is 1e10" exit is applied
// (same result).  kAllCorners: issue all eight at once (one memory round
// trip, for latency-bound callers); otherwise corner 0 first, which settles
// the common far-from-the-surface case with one load.
// Conservative test: false only if the exact lookup at p reads no cell other
// than 1e10 (or leaves the grid).  The approximate cell index (p - o) * 1/cell
// is within one cell of the exact ftoi((p - o) / cell); the lookup reads cells
// cx..cx+1 per axis, all within the brick halo of the approximate cell's brick.
__device__ __forceinline__ bool sdf_brick_of(const SdfGrid& g, f3 p, int& b) {
    const float ax = floorf((p.x - g.ox) * g.inv_cell);
    const float ay = floorf((p.y - g.oy) * g.inv_cell);
    const float az = floorf((p.z - g.oz) * g.inv_cell);
    // Outside [-2, n] the exact index is outside [0, n-3]: GetSignedDistance returns 1e10.
    if (!(ax >= -2.0f && ax <= (float)g.nx && ay >= -2.0f && ay <= (float)g.ny && az >= -2.0f && az <= (float)g.nz))
        return false;
    const int bxi = min(max((int)ax, 0) >> kBrickShift, g.bx - 1);
    const int byi = min(max((int)ay, 0) >> kBrickShift, g.by - 1);
    const int bzi = min(max((int)az, 0) >> kBrickShift, g.bz - 1);
    b = (bzi * g.by + byi) * g.bx + bxi;
    return true;
}

__device__ __forceinline__ bool sdf_maybe_near(const SdfGrid& g, f3 p) {
    if (!g.bricks) return true;
    int b;
    return sdf_brick_of(g, p, b) && ldg(&g.bricks[b]) != 0;
}

template <bool kAllCorners>
__device__ __forceinline__ float sdf_distance_t(const SdfGrid& g, const float* __restrict__ sdf, f3 p) {
    const float INITIAL = 1e10f;
    int cx, cy, cz;
    sdf_coords(g, p, cx, cy, cz);
    if (!(0 <= cx && cx < g.nx - 2) || !(0 <= cy && cy < g.ny - 2) || !(0 <= cz && cz < g.nz - 2)) return INITIAL;
    const int base = sdf_index(g, cx, cy, cz);
    const int X = 1, Y = g.nx, Z = g.ny * g.nx;
    const int ids[8] = {base, base + X, base + Y, base + Y + X, base + Z, base + Z + X, base + Z + Y, base + Z + Y + X};
    float d[8];
    if (!kAllCorners) {
        d[0] = ldg(&sdf[ids[0]]);
        if (d[0] == INITIAL) return INITIAL;
    }
#pragma unroll
    for (int j = kAllCorners ? 0 : 1; j < 8; ++j) d[j] = ldg(&sdf[ids[j]]);
    bool outside = false;
#pragma unroll
    for (int j = 0; j < 8; ++j) outside |= (d[j] == INITIAL);
    if (outside) return INITIAL;
    f3 cp = sdf_cell_pos(g, cx, cy, cz);
    f3 t = divs(sub(p, cp), g.cell);
    float a = lerp_(lerp_(d[0], d[1], t.x), lerp_(d[2], d[3], t.x), t.y);
    float b = lerp_(lerp_(d[4], d[5], t.x), lerp_(d[6], d[7], t.x), t.y);
    return lerp_(a, b, t.z);
}

// One axis of a GetSignedDistance lookup (:472-527): the cell coordinate of
// GetSdfCoordinates and the weight (p - GetSdfCellPosition) / cell, both
// functions of that axis's coordinate alone.
struct SdfAxis {
    int c;
    float t;
};

__device__ __forceinline__ SdfAxis sdf_axis(float v, float o, float cell) {
    const int c = ftoi((v - o) / cell);
    return SdfAxis{c, (v - ((float)c * cell + o)) / cell};
}

// GetSignedDistance's blend of the eight nodes of a cube (1e10 if any is 1e10).
__device__ __forceinline__ float sdf_blend(float d0, float d1, float d2, float d3, float d4, float d5, float d6,
                                           float d7, float tx, float ty, float tz) {
    const float INITIAL = 1e10f;
    if (d0 == INITIAL || d1 == INITIAL || d2 == INITIAL || d3 == INITIAL || d4 == INITIAL || d5 == INITIAL ||
        d6 == INITIAL || d7 == INITIAL)
        return INITIAL;
    float a = lerp_(lerp_(d0, d1, tx), lerp_(d2, d3, tx), ty);
    float b = lerp_(lerp_(d4, d5, tx), lerp_(d6, d7, tx), ty);
    return lerp_(a, b, tz);
}

// The forward-difference gradient step of CollideHairVerticesWithSdf_forward
// :559-593 once the centre distance is known: GetSignedDistance at
// p + offset[i] * h.  Each of those points moves one coordinate by h and adds
// 0 * h to the other two, which leaves them bit-identical unless they are -0;
// an unchanged coordinate keeps the centre's cell and weight, and the moved
// one stays in the centre's cell or steps into the next (the quotient is
// monotone and h is a tenth of a cell).  So the three lookups read the
// centre's eight nodes plus, per stepping axis, the four nodes of the next
// face: one round of at most 20 loads and 12 divisions instead of three
// rounds of 8 loads and 18 divisions, with the same values in the same
// arithmetic.  Anything else (a -0 coordinate, a centre cube outside the grid,
// non-finite values) takes the three lookups as written.
template <bool kAllCorners>
__device__ __forceinline__ bool sdf_project(const SdfGrid& g, const float* __restrict__ sdf, float margin, f3 p,
                                            float distance, f3& out) {
    const float h = 0.1f * g.cell;
    const f3 qx{p.x + 1.0f * h, p.y + 0.0f * h, p.z + 0.0f * h};
    const f3 qy{p.x + 0.0f * h, p.y + 1.0f * h, p.z + 0.0f * h};
    const f3 qz{p.x + 0.0f * h, p.y + 0.0f * h, p.z + 1.0f * h};
    const SdfAxis cx = sdf_axis(p.x, g.ox, g.cell), cy = sdf_axis(p.y, g.oy, g.cell), cz = sdf_axis(p.z, g.oz, g.cell);
    const SdfAxis sx = sdf_axis(qx.x, g.ox, g.cell), sy = sdf_axis(qy.y, g.oy, g.cell), sz = sdf_axis(qz.z, g.oz, g.cell);
    const bool same_rest = __float_as_uint(qy.x) == __float_as_uint(p.x) &&
                           __float_as_uint(qx.y) == __float_as_uint(p.y) &&
                           __float_as_uint(qx.z) == __float_as_uint(p.z);
    const bool centre_in = (0 <= cx.c && cx.c < g.nx - 2) && (0 <= cy.c && cy.c < g.ny - 2) &&
                           (0 <= cz.c && cz.c < g.nz - 2);
    const bool stx = sx.c == cx.c + 1, sty = sy.c == cy.c + 1, stz = sz.c == cz.c + 1;
    const bool simple = same_rest && centre_in && (stx || sx.c == cx.c) && (sty || sy.c == cy.c) &&
                        (stz || sz.c == cz.c);
    f3 f;
    if (simple) {
        const int X = 1, Y = g.nx, Z = g.ny * g.nx;
        const int base = sdf_index(g, cx.c, cy.c, cz.c);
        // a stepped lookup whose cube leaves the grid's range returns 1e10
        const bool inx = !stx || sx.c < g.nx - 2, iny = !sty || sy.c < g.ny - 2, inz = !stz || sz.c < g.nz - 2;
        const float d0 = ldg(&sdf[base]), d1 = ldg(&sdf[base + X]), d2 = ldg(&sdf[base + Y]),
                    d3 = ldg(&sdf[base + Y + X]), d4 = ldg(&sdf[base + Z]), d5 = ldg(&sdf[base + Z + X]),
                    d6 = ldg(&sdf[base + Z + Y]), d7 = ldg(&sdf[base + Z + Y + X]);
        float ex0 = 0.f, ex1 = 0.f, ex2 = 0.f, ex3 = 0.f;
        float ey0 = 0.f, ey1 = 0.f, ey2 = 0.f, ey3 = 0.f;
        float ez0 = 0.f, ez1 = 0.f, ez2 = 0.f, ez3 = 0.f;
        if (stx && inx) {
            const int b = base + 2 * X;
            ex0 = ldg(&sdf[b]); ex1 = ldg(&sdf[b + Y]); ex2 = ldg(&sdf[b + Z]); ex3 = ldg(&sdf[b + Z + Y]);
        }
        if (sty && iny) {
            const int b = base + 2 * Y;
            ey0 = ldg(&sdf[b]); ey1 = ldg(&sdf[b + X]); ey2 = ldg(&sdf[b + Z]); ey3 = ldg(&sdf[b + Z + X]);
        }
        if (stz && inz) {
            const int b = base + 2 * Z;
            ez0 = ldg(&sdf[b]); ez1 = ldg(&sdf[b + X]); ez2 = ldg(&sdf[b + Y]); ez3 = ldg(&sdf[b + Y + X]);
        }
        const float INITIAL = 1e10f;
        f.x = !inx ? INITIAL
                   : (stx ? sdf_blend(d1, ex0, d3, ex1, d5, ex2, d7, ex3, sx.t, cy.t, cz.t)
                          : sdf_blend(d0, d1, d2, d3, d4, d5, d6, d7, sx.t, cy.t, cz.t));
        f.y = !iny ? INITIAL
                   : (sty ? sdf_blend(d2, d3, ey0, ey1, d6, d7, ey2, ey3, cx.t, sy.t, cz.t)
                          : sdf_blend(d0, d1, d2, d3, d4, d5, d6, d7, cx.t, sy.t, cz.t));
        f.z = !inz ? INITIAL
                   : (stz ? sdf_blend(d4, d5, d6, d7, ez0, ez1, ez2, ez3, cx.t, cy.t, sz.t)
                          : sdf_blend(d0, d1, d2, d3, d4, d5, d6, d7, cx.t, cy.t, sz.t));
    } else {
        f.x = sdf_distance_t<kAllCorners>(g, sdf, qx);
        f.y = sdf_distance_t<kAllCorners>(g, sdf, qy);
        f.z = sdf_distance_t<kAllCorners>(g, sdf, qz);
    }
    f3 grad = divs(f3{f.x - distance, f.y - distance, f.z - distance}, h);
    f3 n = normalize3(grad);
    if (distance < margin) {
        out = add(p, mul(n, margin - distance));
        return true;
    }
    return false;
}

// CollideHairVerticesWithSdf_forward :546-595 for one vertex position.
// Returns true (and the projected position) when the vertex is moved.
template <bool kAllCorners = false>
__device__ __forceinline__ bool sdf_collide_vertex(const SdfGrid& g, const float* __restrict__ sdf, float margin,
                                                   f3 p, f3& out) {
    if (!sdf_maybe_near(g, p)) return false;  // the exact distance is 1e10
    float distance = sdf_distance_t<kAllCorners>(g, sdf, p);
    if (distance > margin) return false;
    return sdf_project<kAllCorners>(g, sdf, margin, p, distance, out);
}

// U independent vertices at once: every centre lookup is issued before any
// result is used, then the gradient lookups of the vertices inside the margin.
template <int U>
__device__ __forceinline__ void sdf_collide_batch(const SdfGrid& g, const float* __restrict__ sdf, float margin,
                                                  const f3 (&p)[U], const bool (&act)[U], f3 (&out)[U],
                                                  bool (&hit)[U]) {
    float d[U];
#pragma unroll
    for (int u = 0; u < U; ++u) d[u] = (act[u] && sdf_maybe_near(g, p[u])) ? sdf_distance_t<true>(g, sdf, p[u]) : 1e10f;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        hit[u] = false;
        out[u] = p[u];
        if (act[u] && !(d[u] > margin)) hit[u] = sdf_project<true>(g, sdf, margin, p[u], d[u], out[u]);
    }
}

// CollideHairVerticesWithSdf :597-715, the variant mixing forward and
// backward differences (the reference builds it but dispatches the forward
// one): the gradient comes from the eight nodes of the vertex's own cube, per
// axis forward unless weight + h reaches the cube's far face; h = 0.1 * cell
// is added to the cell-unit weights as written.  No root-vertex test.
__device__ __forceinline__ bool sdf_collide_mixed_vertex(const SdfGrid& g, const float* __restrict__ sdf,
                                                         float margin, f3 p, f3& out) {
    const float INITIAL = 1e10f;
    int cx, cy, cz;
    sdf_coords(g, p, cx, cy, cz);
    if (!(0 <= cx && cx < g.nx - 2) || !(0 <= cy && cy < g.ny - 2) || !(0 <= cz && cz < g.nz - 2)) return false;
    const int base = sdf_index(g, cx, cy, cz);
    const int X = 1, Y = g.nx, Z = g.ny * g.nx;
    const int ids[8] = {base, base + X, base + Y, base + Y + X, base + Z, base + Z + X, base + Z + Y, base + Z + Y + X};
    float d[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) d[j] = ldg(&sdf[ids[j]]);
    bool outside = false;
#pragma unroll
    for (int j = 0; j < 8; ++j) outside |= (d[j] == INITIAL);
    if (outside) return false;
    auto tri = [&](float x, float y, float z) {
        const float a = lerp_(lerp_(d[0], d[1], x), lerp_(d[2], d[3], x), y);
        const float b = lerp_(lerp_(d[4], d[5], x), lerp_(d[6], d[7], x), y);
        return lerp_(a, b, z);
    };
    const f3 t = divs(sub(p, sdf_cell_pos(g, cx, cy, cz)), g.cell);
    const float dv = tri(t.x, t.y, t.z);
    const float h = 0.1f * g.cell;
    const f3 dir{(t.x + h < 1.0f) ? 1.0f : -1.0f, (t.y + h < 1.0f) ? 1.0f : -1.0f, (t.z + h < 1.0f) ? 1.0f : -1.0f};
    const f3 nb{tri(t.x + h * dir.x, t.y, t.z), tri(t.x, t.y + h * dir.y, t.z), tri(t.x, t.y, t.z + h * dir.z)};
    const f3 grad = divs(f3{dir.x * (nb.x - dv), dir.y * (nb.y - dv), dir.z * (nb.z - dv)}, h);
    const f3 n = normalize3(grad);
    if (dv < margin) {
        out = add(p, mul(n, margin - dv));
        return true;
    }
    return false;
}

__device__ __forceinline__ void store_xyz(float4* p, f3 v) {
    float* f = reinterpret_cast<float*>(p);
    stg(&f[0], v.x);
    stg(&f[1], v.y);
    stg(&f[2], v.z);
}

}  // namespace tfx
