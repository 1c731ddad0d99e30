// tfx_fused_kernel.hip — one simulation substep (the six stages of
// TressFXSimulation::Simulate, TressFXSimulation.cpp:77-112, followed by
// TressFXSDFCollision::CollideWithHair, TressFXSDFCollision.cpp:171-221) as a
// single persistent kernel over a work list of (object, strand block) items.
//
// Each 256-thread block owns B = VB/V guide strands of one object at a time:
//   P0 lane-per-strand   blend the strand's bone matrices (once per strand)
//   P1 vertex-parallel   skinning + Verlet + global shape   (Integration...)
//   P2 lane-per-strand   VSP rigid transform from vertices 0,1 (CalculateStrandLevelData)
//   P3 vertex-parallel   apply VSP                          (VelocityShockPropagation)
//   P4 lane-per-strand   Gauss–Seidel bend/twist sweep      (LocalShapeConstraints)
//   P5 vertex-parallel   wind (Jacobi), red–black length pairs, capsules
//   P6 vertex-parallel   tangent, clamp, follow strands, SDF projection, store
// Positions live in LDS, strand-major with one float4 of padding per strand
// (index = strand*(V+1) + v), so the vertex-parallel phases read consecutive
// float4s and the lane-per-strand phases (stride V+1 float4) spread over all
// banks.  Vertex-parallel lanes take consecutive vertices of a strand, so every
// global stream is read and written in contiguous runs.  Pos, Prev and
// RestPos are read once and Pos, Prev, Tangent written once per guide vertex;
// follow strands are written once.  PrevPrev is written for vertices 0 and 1
// only (the only ones any stage reads) and StrandLevelData stays on chip.
#include "tfx_device.h"

namespace tfx {

constexpr int kFusedBlock = 256;

template <int V, int VB>
struct FusedShape {
    static constexpr int B = VB / V;                 // guide strands per block
    static constexpr int NPT = VB / kFusedBlock;     // vertices per thread in vertex-parallel phases
    static constexpr int STRIDE = V + 1;             // LDS float4 per strand (padded)
    static_assert(VB % kFusedBlock == 0, "VB must be a multiple of the block size");
    static_assert(B >= 1 && B * V == VB, "VB must be a multiple of V");
};

template <int V, int VB>
__global__ void __launch_bounds__(kFusedBlock) k_fused_substep(const DevHair* __restrict__ hairs,
                                                               const WorkItem* __restrict__ items, int nitems) {
    using S = FusedShape<V, VB>;
    constexpr int B = S::B, NPT = S::NPT, STRIDE = S::STRIDE;
    __shared__ float4 sPos[B * STRIDE];
    __shared__ float4 sBind[B * STRIDE];  // P1->P4 rest pose rotated by the skinning quaternion; P6 guide tangents
    __shared__ float sMat[B][12];      // P0->P1 blended bone; P2->P3 vsp quat (0..3), translation+coeff (4..7)
    __shared__ float4 sSkinQ[B];       // P0->P4 skinning quaternion
    __shared__ float4 sV01[B][3];      // P1->P2 prev(v0), prev(v1), prevprev(v1)

    const int t = threadIdx.x;
    // Lane-per-strand phases use threads tt = t rotated by a block-dependent
    // wave offset, so that blocks with B <= 64 do not all load SIMD 0.
    const int tt = (t + 64 * (blockIdx.x & 3)) & (kFusedBlock - 1);

    for (int w = blockIdx.x; w < nitems; w += gridDim.x) {
        const WorkItem item = items[w];
        const int hi = __builtin_amdgcn_readfirstlane(item.hair);
        const int g0 = __builtin_amdgcn_readfirstlane(item.g0);
        const DevHair h = hairs[hi];
        const DevParams& P = *h.params;
        const int F = h.F;
        const int nb = min(B, h.G - g0);

        // ---- P0: blended bone per strand ---------------------------------
        for (int s = tt; s < nb; s += kFusedBlock) {
            const size_t strand = (size_t)(g0 + s) * (F + 1);
            BlendedBone bb;
            blend_bones(h.skin[strand], h.bones, bb);
#pragma unroll
            for (int k = 0; k < 9; ++k) sMat[s][k] = bb.m[k];
#pragma unroll
            for (int k = 0; k < 3; ++k) sMat[s][9 + k] = bb.t[k];
            sSkinQ[s] = quat_from_matrix3(bb.m);
        }
        __syncthreads();

        // ---- P1: integration and global shape ------------------------------
        float4 prevReg[NPT];
        const bool reset = P.reset != 0.0f;
        const float grav = P.gravity, dt = P.dt, decay = P.decay;
        const bool gshape = P.global_stiffness > 0.0f && P.global_range != 0.0f;
#pragma unroll
        for (int k = 0; k < NPT; ++k) {
            const int lv = k * kFusedBlock + t;
            const int ls = lv / V, v = lv % V;
            prevReg[k] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (ls < nb) {
                const size_t i = ((size_t)(g0 + ls) * (F + 1)) * V + v;
                float4 initialPos = h.rest[i];
                float4 shared = h.pos[i];
                const float4 prevOld = h.prev[i];
                float m[9], tr[3];
#pragma unroll
                for (int q = 0; q < 9; ++q) m[q] = sMat[ls][q];
#pragma unroll
                for (int q = 0; q < 3; ++q) tr[q] = sMat[ls][9 + q];
                // bind pose for LocalShapeConstraints: MultQuaternionAndVector(boneQuat, rest)
                const f3 bind = qrot(sSkinQ[ls], xyz(initialPos));
                sBind[ls * STRIDE + v] = make_float4(bind.x, bind.y, bind.z, 0.0f);
                f3 sk = bone_transform(m, tr, xyz(initialPos));
                initialPos.x = sk.x; initialPos.y = sk.y; initialPos.z = sk.z;
                float4 currentPos = shared;
                float4 oldPos = prevOld;
                if (reset) { currentPos = initialPos; oldPos = initialPos; }
                if (movable(currentPos)) {
                    f3 force{grav * 0.0f, grav * -1.0f, grav * 0.0f};
                    f3 c = xyz(currentPos), o = xyz(oldPos);
                    f3 n = add(add(c, mul(decay, sub(c, o))), mul(mul(force, dt), dt));
                    shared.x = n.x; shared.y = n.y; shared.z = n.z;
                } else {
                    shared = initialPos;
                }
                if (gshape && movable(shared) && (float)v < P.global_range * (float)V) {
                    f3 del = mul(P.global_stiffness, sub(xyz(initialPos), xyz(shared)));
                    shared.x += del.x; shared.y += del.y; shared.z += del.z;
                }
                sPos[ls * STRIDE + v] = shared;
                prevReg[k] = currentPos;
                if (v < 2) {
                    h.prevprev[i] = prevOld;
                    sV01[ls][v] = currentPos;
                    if (v == 1) sV01[ls][2] = prevOld;
                }
            }
        }
        __syncthreads();

        // ---- P2: strand-level VSP transform --------------------------------
        for (int s = tt; s < nb; s += kFusedBlock) {
            const float4 n0 = sPos[s * STRIDE], n1 = sPos[s * STRIDE + 1];
            const float4 o0 = sV01[s][0], o1 = sV01[s][1], oo1 = sV01[s][2];
            f3 u = normalize3(sub(xyz(o1), xyz(o0)));
            f3 vv = normalize3(sub(xyz(n1), xyz(n0)));
            float4 rot = quat_two_vectors(u, vv);
            f3 trans = sub(xyz(n0), qrot(rot, xyz(o0)));
            float ax = (n1.x - 2.0f * o1.x) + oo1.x;
            float ay = (n1.y - 2.0f * o1.y) + oo1.y;
            float az = (n1.z - 2.0f * o1.z) + oo1.z;
            float aw = (n1.w - 2.0f * o1.w) + oo1.w;
            float accel = sqrtf(((ax * ax + ay * ay) + az * az) + aw * aw);
            float coeff = P.vsp_coeff;
            if (accel > P.vsp_accel_threshold) coeff = 1.0f;
            sMat[s][0] = rot.x; sMat[s][1] = rot.y; sMat[s][2] = rot.z; sMat[s][3] = rot.w;
            sMat[s][4] = trans.x; sMat[s][5] = trans.y; sMat[s][6] = trans.z; sMat[s][7] = coeff;
        }
        __syncthreads();

        // ---- P3: velocity shock propagation --------------------------------
#pragma unroll
        for (int k = 0; k < NPT; ++k) {
            const int lv = k * kFusedBlock + t;
            const int ls = lv / V, v = lv % V;
            if (ls < nb && v >= 2) {
                const float4 q = make_float4(sMat[ls][0], sMat[ls][1], sMat[ls][2], sMat[ls][3]);
                const f3 tv{sMat[ls][4], sMat[ls][5], sMat[ls][6]};
                const float c = sMat[ls][7];
                float4& sp = sPos[ls * STRIDE + v];
                f3 pn = xyz(sp), po = xyz(prevReg[k]);
                f3 rn = add(mul(1.0f - c, pn), mul(c, add(qrot(q, pn), tv)));
                f3 ro = add(mul(1.0f - c, po), mul(c, add(qrot(q, po), tv)));
                sp.x = rn.x; sp.y = rn.y; sp.z = rn.z;
                prevReg[k].x = ro.x; prevReg[k].y = ro.y; prevReg[k].z = ro.z;
            }
        }
        __syncthreads();

        // ---- P4: local shape constraints (Gauss–Seidel sweeps) -------------
        if (P.lsc_passes > 0) {
            const float stiff = 0.5f * fmn(P.local_stiffness, 0.95f);
            for (int s = tt; s < nb; s += kFusedBlock) {
                float4* sp = &sPos[s * STRIDE];
                const float4* sb = &sBind[s * STRIDE];
                for (int pass = 0; pass < P.lsc_passes; ++pass) {
                    float4 pm = sp[0];
                    float4 p = sp[1];
                    f3 bindM = xyz(sb[0]);
                    f3 bind = xyz(sb[1]);
#pragma unroll 2
                    for (int lv = 1; lv < V - 1; ++lv) {
                        float4 pp = sp[lv + 1];
                        f3 bindP = xyz(sb[lv + 1]);
                        f3 lastVec = sub(xyz(p), xyz(pm));
                        f3 vecBindPose = sub(bindP, bind);
                        f3 lastVecBindPose = sub(bind, bindM);
                        float4 rg = quat_two_vectors(normalize3(lastVecBindPose), normalize3(lastVec));
                        f3 org = add(qrot(rg, vecBindPose), xyz(p));
                        f3 del = mul(stiff, sub(org, xyz(pp)));
                        if (movable(p)) { p.x -= del.x; p.y -= del.y; p.z -= del.z; }
                        if (movable(pp)) { pp.x += del.x; pp.y += del.y; pp.z += del.z; }
                        sp[lv] = p;
                        pm = p; p = pp;
                        bindM = bind; bind = bindP;
                    }
                    sp[V - 1] = p;
                }
            }
            __syncthreads();
        }

        // ---- P5: wind, length constraints, capsules ------------------------
        if (P.wind_on) {
            f3 force[NPT];
#pragma unroll
            for (int k = 0; k < NPT; ++k) {
                const int lv = k * kFusedBlock + t;
                const int ls = lv / V, v = lv % V;
                force[k] = f3{0.f, 0.f, 0.f};
                if (ls < nb && v >= 2 && v < V - 1) {
                    const int strand = (g0 + ls) * (F + 1);
                    float a = (float)(uint32_t)(strand % 20) / 20.0f;
                    f3 wv = add(add(add(mul(a, xyz(P.wind0)), mul(1.0f - a, xyz(P.wind1))), mul(a, xyz(P.wind2))),
                                mul(1.0f - a, xyz(P.wind3)));
                    f3 d = sub(xyz(sPos[ls * STRIDE + v]), xyz(sPos[ls * STRIDE + v + 1]));
                    f3 c = cross3(cross3(d, wv), d);
                    force[k] = f3{-c.x, -c.y, -c.z};
                }
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < NPT; ++k) {
                const int lv = k * kFusedBlock + t;
                const int ls = lv / V, v = lv % V;
                if (ls < nb && v >= 2 && v < V - 1) {
                    f3 f = mul(mul(force[k], dt), dt);
                    float4& sp = sPos[ls * STRIDE + v];
                    sp.x += f.x; sp.y += f.y; sp.z += f.z;
                }
            }
            __syncthreads();
        }
        float restLen[NPT];
#pragma unroll
        for (int k = 0; k < NPT; ++k) {
            const int lv = k * kFusedBlock + t;
            const int ls = lv / V, v = lv % V;
            restLen[k] = 0.0f;
            if (ls < nb) restLen[k] = h.restlen[((size_t)(g0 + ls) * (F + 1)) * V + v];
        }
        for (int it = 0; it < P.length_iters; ++it) {
#pragma unroll
            for (int k = 0; k < NPT; ++k) {
                const int lv = k * kFusedBlock + t;
                const int ls = lv / V, v = lv % V;
                if (ls < nb && (v & 1) == 0 && (v >> 1) < V / 2) {
                    float4 a = sPos[ls * STRIDE + v], b = sPos[ls * STRIDE + v + 1];
                    distance_constraint(a, b, restLen[k]);
                    sPos[ls * STRIDE + v] = a; sPos[ls * STRIDE + v + 1] = b;
                }
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < NPT; ++k) {
                const int lv = k * kFusedBlock + t;
                const int ls = lv / V, v = lv % V;
                if (ls < nb && (v & 1) == 1 && (v >> 1) < (V - 1) / 2) {
                    float4 a = sPos[ls * STRIDE + v], b = sPos[ls * STRIDE + v + 1];
                    distance_constraint(a, b, restLen[k]);
                    sPos[ls * STRIDE + v] = a; sPos[ls * STRIDE + v + 1] = b;
                }
            }
            __syncthreads();
        }
        uint32_t hitMask = 0;
        if (P.num_capsules > 0) {
#pragma unroll
            for (int k = 0; k < NPT; ++k) {
                const int lv = k * kFusedBlock + t;
                const int ls = lv / V, v = lv % V;
                if (ls < nb) {
                    float4 cur = sPos[ls * STRIDE + v];
                    for (int c = 0; c < P.num_capsules; ++c) {
                        f3 np;
                        if (capsule_collision(cur, prevReg[k], np, P.cap0[c], P.cap1[c])) {
                            cur.x = np.x; cur.y = np.y; cur.z = np.z;
                            hitMask |= 1u << k;
                        }
                    }
                    sPos[ls * STRIDE + v] = cur;
                }
            }
            __syncthreads();
        }

        // ---- P6a: guide tangent, clamp, SDF collision, stores ---------------
        // Vertices are taken in groups of four so that all their SDF lookups
        // are in flight together.
        const bool has_sdf = P.has_sdf != 0;
        SdfGrid grid;
        grid.ox = P.sdf_ox; grid.oy = P.sdf_oy; grid.oz = P.sdf_oz; grid.cell = P.sdf_cell;
        grid.nx = P.sdf_nx; grid.ny = P.sdf_ny; grid.nz = P.sdf_nz; grid.pad = 0;
        const float margin = P.sdf_margin;
        const float C = P.clamp;
        const bool tw_uniform = P.tangent_w_uniform != 0;
        const float* __restrict__ sdfp = h.sdf;
        constexpr int U = NPT < 4 ? NPT : 4;
#pragma unroll
        for (int k0 = 0; k0 < NPT; k0 += U) {
            f3 cp[U], outp[U];
            bool act[U], hit[U];
            float4 pvs[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = k0 + u;
                const int lv = k * kFusedBlock + t;
                const int ls = lv / V, v = lv % V;
                act[u] = false;
                cp[u] = f3{0.f, 0.f, 0.f};
                pvs[u] = prevReg[k];
                if (ls < nb) {
                    const float4 sp = sPos[ls * STRIDE + v];
                    float4 pv = prevReg[k];
                    f3 delta = sub(xyz(sp), xyz(pv));
                    const float speedSqr = dot3(delta, delta);
                    if (speedSqr > C * C) {
                        const float sc = (C * C) / speedSqr;
                        delta = mul(delta, sc);
                        f3 np = sub(xyz(sp), delta);
                        pv.x = np.x; pv.y = np.y; pv.z = np.z;
                    }
                    if (hitMask & (1u << k)) pv = sp;
                    pvs[u] = pv;
                    cp[u] = xyz(sp);
                    act[u] = has_sdf && v >= 2;
                }
            }
            sdf_collide_batch<U>(grid, sdfp, margin, cp, act, outp, hit);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = k0 + u;
                const int lv = k * kFusedBlock + t;
                const int ls = lv / V, v = lv % V;
                if (ls < nb) {
                    const size_t i = ((size_t)(g0 + ls) * (F + 1)) * V + v;
                    const int base = ls * STRIDE;
                    const int ti = (v == V - 1) ? v - 1 : v;
                    const f3 tan = normalize3(sub(xyz(sPos[base + ti + 1]), xyz(sPos[base + ti])));
                    float4 pout = sPos[base + v];
                    float4 pv = pvs[u];
                    if (hit[u]) {
                        pout.x = outp[u].x; pout.y = outp[u].y; pout.z = outp[u].z;
                        pv.x = outp[u].x; pv.y = outp[u].y; pv.z = outp[u].z;
                    }
                    h.pos[i] = pout;
                    h.prev[i] = pv;
                    float tw;
                    if (tw_uniform) {
                        tw = P.tangent_w;
                        h.tangent[i] = make_float4(tan.x, tan.y, tan.z, tw);
                    } else {
                        tw = h.tangent[i].w;
                        store_xyz(&h.tangent[i], tan);
                    }
                    // sBind is dead after the LSC sweep: it now carries the
                    // guide tangent (with the w the follow strands copy) to P6b.
                    sBind[base + v] = make_float4(tan.x, tan.y, tan.z, tw);
                }
            }
        }

        // ---- P6b: follow strands (UpdateFollowHairVertices + collision) -----
        // Follow vertex (j, lv) of the block; consecutive lanes take
        // consecutive vertices of one follow strand; four per lane at a time.
        if (F > 0) {
            __syncthreads();
            const int per = nb * V;
            const int nf = F * per;
            const float tip = P.tip_sep;
            for (int idx0 = t; idx0 < nf; idx0 += 4 * kFusedBlock) {
                f3 fp[4], outp[4];
                bool act[4], hit[4], ok[4];
                size_t fi[4];
                float4 ft[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int idx = idx0 + u * kFusedBlock;
                    ok[u] = idx < nf;
                    act[u] = false;
                    fp[u] = f3{0.f, 0.f, 0.f};
                    fi[u] = 0;
                    ft[u] = make_float4(0.f, 0.f, 0.f, 0.f);
                    if (ok[u]) {
                        const int j = idx / per + 1;
                        const int lv = idx - (j - 1) * per;
                        const int ls = lv / V, v = lv % V;
                        const size_t strand = (size_t)(g0 + ls) * (F + 1);
                        fi[u] = (strand + j) * V + v;
                        const float4 sp = sPos[ls * STRIDE + v];
                        ft[u] = sBind[ls * STRIDE + v];
                        const float factor = tip * ((float)v / (float)V) + 1.0f;
                        const f3 off = xyz(h.follow_off[strand + j]);
                        fp[u] = add(xyz(sp), mul(factor, off));
                        act[u] = has_sdf && v >= 2;
                    }
                }
                sdf_collide_batch<4>(grid, sdfp, margin, fp, act, outp, hit);
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (!ok[u]) continue;
                    if (hit[u]) store_xyz(&h.prev[fi[u]], outp[u]);
                    store_xyz(&h.pos[fi[u]], hit[u] ? outp[u] : fp[u]);
                    h.tangent[fi[u]] = ft[u];
                }
            }
        }
        __syncthreads();  // LDS is reused by the next work item
    }
}

template <int V, int VB>
static hipError_t launch_fused(const DevHair* hairs, const WorkItem* items, int nitems, int grid, hipStream_t stream) {
    if (nitems <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_fused_substep<V, VB>), dim3(grid), dim3(kFusedBlock), 0, stream, hairs, items, nitems);
    return hipGetLastError();
}

}  // namespace tfx

// Vertices per block for each strand length and launch-shape variant
// (variant 0: 1024 vertices, no register spills; variant 1: 2048 vertices).
extern "C" int tfx_fused_block_vertices(int V, int variant) {
    if (V != 4 && V != 8 && V != 16 && V != 32 && V != 64) return 0;
    const int vb = variant == 1 ? 2048 : 1024;
    return V == 4 ? vb / 2 : vb;
}

extern "C" hipError_t tfx_launch_fused(int V, int variant, const tfx::DevHair* hairs, const tfx::WorkItem* items,
                                       int nitems, int grid, hipStream_t stream) {
    using namespace tfx;
    if (variant == 1) {
        switch (V) {
            case 4: return launch_fused<4, 1024>(hairs, items, nitems, grid, stream);
            case 8: return launch_fused<8, 2048>(hairs, items, nitems, grid, stream);
            case 16: return launch_fused<16, 2048>(hairs, items, nitems, grid, stream);
            case 32: return launch_fused<32, 2048>(hairs, items, nitems, grid, stream);
            case 64: return launch_fused<64, 2048>(hairs, items, nitems, grid, stream);
            default: return hipErrorInvalidValue;
        }
    }
    switch (V) {
        case 4: return launch_fused<4, 512>(hairs, items, nitems, grid, stream);
        case 8: return launch_fused<8, 1024>(hairs, items, nitems, grid, stream);
        case 16: return launch_fused<16, 1024>(hairs, items, nitems, grid, stream);
        case 32: return launch_fused<32, 1024>(hairs, items, nitems, grid, stream);
        case 64: return launch_fused<64, 1024>(hairs, items, nitems, grid, stream);
        default: return hipErrorInvalidValue;
    }
}
