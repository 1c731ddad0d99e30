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
#include <hip/hip_ext.h>

#include <type_traits>

#include "tfx_device.h"

namespace tfx {
inline namespace TFX_KERNEL_NS {

constexpr int kFusedBlock = 256;
constexpr int kMaxStagedFollow = 4;  // follow strands per guide whose offsets are staged in LDS
constexpr int kMaxBrickWords = 1024; // SDF grids of up to 32768 bricks keep their occupancy in LDS

// Orders a wave's LDS accesses before and after this point (LDS executes a
// wave's accesses in order; the fences keep the compiler from moving them).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// sdf_collide_batch for the U vertices of each lane of a wave, with the
// projections of the vertices inside the margin packed onto consecutive
// lanes: the distance lookups of all U slots go out together, then the lanes
// of the wave take the wave's hits one each (64 at a time), so every hit's
// gradient lookup is in flight at once and each lane projects at most one
// vertex per round, instead of one gradient round trip per slot.  `wl`: 64
// float4 of LDS private to the calling wave.  Every lane of the wave calls
// it.  Per vertex the same lookups in the same arithmetic as
// sdf_collide_batch.
template <int U>
__device__ __forceinline__ void sdf_collide_wave(const SdfGrid& g, const float* __restrict__ sdf, float margin,
                                                 const f3 (&p)[U], const bool (&act)[U], f3 (&out)[U],
                                                 bool (&hit)[U], float4* wl) {
    float d[U];
#pragma unroll
    for (int u = 0; u < U; ++u) d[u] = (act[u] && sdf_maybe_near(g, p[u])) ? sdf_distance_t<true>(g, sdf, p[u]) : 1e10f;
    bool want[U];
    int pos[U];
    int total = 0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        hit[u] = false;
        out[u] = p[u];
        want[u] = act[u] && !(d[u] > margin);
        const unsigned long long m = __ballot(want[u]);
        pos[u] = total + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
        total += __popcll(m);
    }
    const int lane = __lane_id();
    for (int base = 0; base < total; base += 64) {
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (want[u] && pos[u] >= base && pos[u] < base + 64) wl[pos[u] - base] = make_float4(p[u].x, p[u].y, p[u].z, d[u]);
        wave_sync();
        if (base + lane < total) {
            const float4 e = wl[lane];
            f3 o;
            const bool h = sdf_project<true>(g, sdf, margin, xyz(e), e.w, o);
            wl[lane] = make_float4(o.x, o.y, o.z, h ? 1.0f : 0.0f);
        }
        wave_sync();
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (want[u] && pos[u] >= base && pos[u] < base + 64) {
                const float4 r = wl[pos[u] - base];
                hit[u] = r.w != 0.0f;
                if (hit[u]) out[u] = xyz(r);
            }
        wave_sync();
    }
}

// x of lane + 1 and of lane - 1 of the wave (DPP wave_shl:1 / wave_shr:1;
// the last and the first lane get 0).
__device__ __forceinline__ float from_next_lane(float x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x130, 0xf, 0xf, false));
}
__device__ __forceinline__ float from_prev_lane(float x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x138, 0xf, 0xf, false));
}

template <int V, int VB>
struct FusedShape {
    static constexpr int B = VB / V;                 // guide strands per block
    static constexpr int NPT = VB / kFusedBlock;     // vertices per thread in vertex-parallel phases
    static constexpr int STRIDE = V + 1;             // LDS float4 per strand (padded)
    static_assert(VB % kFusedBlock == 0, "VB must be a multiple of the block size");
    static_assert(B >= 1 && B * V == VB, "VB must be a multiple of V");
};

// Occupancy target: 4 blocks per CU for 512-vertex blocks, 3 for larger ones
// (the register budget of 128 and 168 VGPRs per lane).
template <int V, int VB>
__global__ void __launch_bounds__(kFusedBlock, (VB <= 512 ? 4 : 3)) k_fused_substep(const DevHair* __restrict__ hairs,
                                                               const WorkItem* __restrict__ items, int nitems,
                                                               int defer_sdf, uint32_t* __restrict__ queue) {
    using S = FusedShape<V, VB>;
    constexpr int B = S::B, NPT = S::NPT, STRIDE = S::STRIDE;
    __shared__ float4 sPos[B * STRIDE];
    // P1->P4: bind-pose vectors of LocalShapeConstraints (:799-807) per vertex i,
    //   vecBindPose = bind[i+1] - bind[i] in sVB[i].xyz, n = normalize(bind[i] - bind[i-1])
    //   in (sVB[i].w, sNB[i].x, sNB[i].y), bind = MultQuaternionAndVector(skinningQuat, rest).
    //   P6: sVB carries the guide tangents.
    __shared__ float4 sVB[B * STRIDE];
    __shared__ __align__(16) float2 sNB[B * STRIDE];
    __shared__ float sMat[B][12];      // P0->P1 blended bone; P2->P3 vsp quat (0..3), translation+coeff (4..7)
    __shared__ float4 sSkinQ[B];       // P0->P1 skinning quaternion
    __shared__ float4 sV01[B][3];      // P1->P2 prev(v0), prev(v1), prevprev(v1)
    __shared__ float4 sOff[B][kMaxStagedFollow];  // P0->P6b follow root offsets (F <= kMaxStagedFollow)
    __shared__ uint32_t sBits[kMaxBrickWords];    // SDF brick occupancy, one bit per brick
    __shared__ int sIdx[B];            // P1->P6b the item's guide strands (the object's processing order)
    __shared__ int sQueueCount;
    __shared__ uint32_t sNext;         // the block's next work item

    const int t = threadIdx.x;
    // Lane-per-strand phases use threads tt = t rotated by a block-dependent
    // wave offset, so that blocks with B <= 64 do not all load SIMD 0.
    const int tt = (t + 64 * (blockIdx.x & 3)) & (kFusedBlock - 1);

    // P0 of a work item: the blended bone and skinning quaternion of each of
    // its strands, threads first, first + count, ... taking one strand each.
    // `lean`: with the item's vertex registers live (during P4).
    // The object's processing order, used when the launch makes its SDF
    // lookups (their cells are what neighbouring strands share); without
    // them index order keeps each block's streams contiguous.
    auto order_of = [&](const DevHair& hh) -> const int* { return hh.sdf != nullptr && !defer_sdf ? hh.order : nullptr; };
    auto blend_strands = [&](const DevHair& hh, int gg0, int nnb, int first, int count, auto lean) {
        const int* ord = order_of(hh);
        for (int s = first; s < nnb; s += count) {
            const int gs = ord ? ldg(&ord[gg0 + s]) : gg0 + s;
            const size_t strand = (size_t)gs * (hh.F + 1);
            BlendedBone bb;
            if constexpr (decltype(lean)::value) blend_bones_lean(ldg(&hh.skin[strand]), hh.bones, bb);
            else blend_bones(ldg(&hh.skin[strand]), hh.bones, bb);
#pragma unroll
            for (int k = 0; k < 9; ++k) sMat[s][k] = bb.m[k];
#pragma unroll
            for (int k = 0; k < 3; ++k) sMat[s][9 + k] = bb.t[k];
            sSkinQ[s] = quat_from_matrix3(bb.m);
        }
    };
    // Whether the current item's P0 was done ahead, during the previous
    // item's LocalShapeConstraints phase.
    bool blended = false;
    // The occupancy sBits holds (null: none).
    const uint32_t* bits_in_lds = nullptr;

    // Block b's first work item is item b; the items past the first round
    // are taken from the launch's queue (queue[0] counts the takes, item
    // gridDim.x + count), one at a time as each block gets to its next item,
    // so that blocks whose items cost more take fewer of them and the blocks
    // of a launch finish together.  A block takes its next item in P1, after
    // the phase's stream loads (so that waiting for them does not wait for
    // the atomic, which all blocks issue at once in their first item), and
    // publishes it at the end of P3 for P4, where the idle waves start on it.
    // A launch whose first round covers the list takes nothing.
    const bool queued = (uint32_t)nitems > gridDim.x;
    uint32_t w = blockIdx.x;
    while (w < (uint32_t)nitems) {
        const WorkItem item = items[w];
        const int hi = __builtin_amdgcn_readfirstlane(item.hair);
        const int g0 = __builtin_amdgcn_readfirstlane(item.g0);
        const DevHair h = load_uniform(hairs + hi);
        const_as<DevParams>& P = *uniform_const(h.params);
        const int F = h.F;
        const int nb = min(B, h.G - g0);

        // SDF brick occupancy into LDS (read by P6's conservative early-out).
        // defer_sdf: the SDF collision is left to k_sdf_collide_items, run
        // after the substep (the reference's own order: Simulate, then
        // CollideWithHair; the tangents and follow strands come from the
        // positions before the collision either way).
        const bool has_sdf = P.has_sdf != 0 && !defer_sdf;
        const int brick_words = (P.sdf_bx * P.sdf_by * P.sdf_bz + 31) >> 5;
        const bool bits_lds = has_sdf && h.sdf_brick_bits != nullptr && brick_words <= kMaxBrickWords;
        // (kept from the block's previous item when that item read the same
        // occupancy: a block's successive items mostly belong to one object,
        // and no launch writes the occupancy it reads)
        if (bits_lds && h.sdf_brick_bits != bits_in_lds)
            for (int k = t; k < brick_words; k += kFusedBlock) sBits[k] = ldg(&h.sdf_brick_bits[k]);
        bits_in_lds = bits_lds ? h.sdf_brick_bits : nullptr;

        // ---- P0: blended bone per strand ---------------------------------
        const int* ord = order_of(h);
        for (int s = t; s < nb; s += kFusedBlock) sIdx[s] = ord ? ldg(&ord[g0 + s]) : g0 + s;
        if (!blended) blend_strands(h, g0, nb, tt, kFusedBlock, std::false_type{});
        if (F > 0 && F <= kMaxStagedFollow)
            for (int s = tt; s < nb; s += kFusedBlock) {
                const size_t strand = (size_t)(ord ? ldg(&ord[g0 + s]) : g0 + s) * (F + 1);
                for (int j = 0; j < F; ++j) sOff[s][j] = ldg(&h.follow_off[strand + 1 + j]);
            }
        __syncthreads();

        // ---- P1: integration and global shape ------------------------------
        // Every global load of the phase is issued before any of its results
        // is used (lanes past the block's last strand read the first strand's
        // vertices and discard them).
        float4 prevReg[NPT];
        float4 restIn[NPT], posIn[NPT], prevIn[NPT];
#pragma unroll
        for (int k = 0; k < NPT; ++k) {
            const int lv = k * kFusedBlock + t;
            const int ls = lv / V, v = lv % V;
            const size_t i = ((size_t)sIdx[ls < nb ? ls : 0] * (F + 1)) * V + v;
            restIn[k] = ldg(&h.rest[i]);
            posIn[k] = ldg(&h.pos[i]);
            prevIn[k] = ldg(&h.prev[i]);
        }
        // The length constraints take one lane per vertex pair (P5): pair
        // slot m of this thread starts at block vertex pairLv[m] (even phase;
        // the odd phase's pair starts one vertex later).  Lanes 0-31 of a wave
        // take the pairs of the 64 vertices the wave holds for k = 2m, lanes
        // 32-63 those for k = 2m + 1, so a strand's pairs stay in the wave that
        // holds its vertices.  Both rest lengths a slot needs are adjacent.
        constexpr int NPP = NPT / 2;
        static_assert(NPT % 2 == 0, "pairs take two vertex slots");
        int pairLv[NPP];
        float2 restLen[NPP];
#pragma unroll
        for (int m = 0; m < NPP; ++m) {
            pairLv[m] = (2 * m + ((t >> 5) & 1)) * kFusedBlock + (t & ~63) + 2 * (t & 31);
            const int ls = pairLv[m] / V, v = pairLv[m] % V;
            const size_t i = ((size_t)sIdx[ls < nb ? ls : 0] * (F + 1)) * V + v;
            restLen[m] = ldg(reinterpret_cast<const float2*>(&h.restlen[i]));
        }
        uint32_t taken = (uint32_t)nitems;
        if (queued && t == 0) taken = gridDim.x + atomicAdd(&queue[0], 1u);
        const bool reset = P.reset != 0.0f;
        const float grav = P.gravity, dt = P.dt, decay = P.decay;
        const bool gshape = P.global_stiffness > 0.0f && P.global_range != 0.0f;
        const float gstiff = P.global_stiffness, grange = P.global_range;
#pragma unroll
        for (int k = 0; k < NPT; ++k) {
            const int lv = k * kFusedBlock + t;
            const int ls = lv / V, v = lv % V;
            prevReg[k] = make_float4(0.f, 0.f, 0.f, 0.f);
            const bool valid = ls < nb;
            const size_t i = ((size_t)sIdx[valid ? ls : 0] * (F + 1)) * V + v;
            float4 initialPos = valid ? restIn[k] : make_float4(0.f, 0.f, 0.f, 0.f);
            // Bind pose of LocalShapeConstraints, MultQuaternionAndVector(boneQuat,
            // rest) (:799-801); a strand's vertices sit in consecutive lanes of one
            // wave (64 % V == 0), so the neighbours come from lane +-1 (a
            // wave-wide DPP shift; the values the first and last lanes of the
            // wave get are not used).
            const f3 bind = qrot(sSkinQ[valid ? ls : 0], xyz(initialPos));
            const f3 bindP{from_next_lane(bind.x), from_next_lane(bind.y), from_next_lane(bind.z)};
            const f3 bindM{from_prev_lane(bind.x), from_prev_lane(bind.y), from_prev_lane(bind.z)};
            if (valid) {
                const f3 vb = sub(bindP, bind);
                f3 nbp{0.f, 0.f, 0.f};
                if (v >= 1 && v < V - 1) nbp = normalize3(sub(bind, bindM));
                sVB[ls * STRIDE + v] = make_float4(vb.x, vb.y, vb.z, nbp.x);
                sNB[ls * STRIDE + v] = make_float2(nbp.y, nbp.z);
                float4 shared = posIn[k];
                const float4 prevOld = prevIn[k];
                float m[9], tr[3];
#pragma unroll
                for (int q = 0; q < 9; ++q) m[q] = sMat[ls][q];
#pragma unroll
                for (int q = 0; q < 3; ++q) tr[q] = sMat[ls][9 + q];
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
                if (gshape && movable(shared) && (float)v < grange * (float)V) {
                    f3 del = mul(gstiff, sub(xyz(initialPos), xyz(shared)));
                    shared.x += del.x; shared.y += del.y; shared.z += del.z;
                }
                sPos[ls * STRIDE + v] = shared;
                prevReg[k] = currentPos;
                if (v < 2) {
                    stg(&h.prevprev[i], prevOld);
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
        if (t == 0) sNext = taken;
        __syncthreads();

        // ---- P4: local shape constraints (Gauss–Seidel sweeps) -------------
        // Step i of a pass reads vertex i+1 from LDS (vertices i-1 and i come
        // from its own previous step), stores vertex i, which is then final
        // for the pass, and stores vertex i+1 as the step left it (step i+1
        // replaces it; after the last step it is vertex V-1's final value).
        // So up to kLscPipe passes run together, pass k on its own lane
        // kLscLag = 3 steps behind pass k-1.  Step i's reads (vertex i+1 and
        // the bind pose of i) are issued one step early, during step i-1:
        // pass k-1 stored vertex i+1 final at its step i+1, one step before
        // pass k's step i-1, and from then on stores only vertices above i+1;
        // and pass k-1 never reads a vertex that pass k has stored.  Every
        // step reads exactly the values the sequential order gives it, and
        // its LDS reads are in flight while the step before it computes.  The
        // lanes of one strand are adjacent in one wave; the wave runs the
        // steps in lock step and orders its LDS accesses by a wave-scope fence
        // per step.
        // The waves the sweep leaves idle compute P0 of the block's next work
        // item meanwhile (sMat and sSkinQ are dead from P4 on: P1 last read
        // them, and P2/P3 their VSP reuse).
        constexpr int kLscPipe = 4, kLscLag = 3;
        const int npass = P.lsc_passes > 0 ? P.lsc_passes : 0;
        // 1, 2 or 4 lanes per strand: a count that divides 64 keeps the
        // lanes of a strand in one wave (with 3 or more passes the extra
        // lanes of a group are idle).
        const int lanes = npass >= 3 ? kLscPipe : npass;
        const int lsc_threads = min(((nb * lanes + 63) >> 6) << 6, kFusedBlock);
        const uint32_t wn = sNext;
        const bool ahead = wn < (uint32_t)nitems && lsc_threads < kFusedBlock;
        if (ahead && tt >= lsc_threads) {
            const WorkItem nx = items[wn];
            const DevHair hn = load_uniform(hairs + __builtin_amdgcn_readfirstlane(nx.hair));
            const int gn = __builtin_amdgcn_readfirstlane(nx.g0);
            blend_strands(hn, gn, min(B, hn.G - gn), tt - lsc_threads, kFusedBlock - lsc_threads, std::true_type{});
        }
        if (npass > 0) {
            const float stiff = 0.5f * fmn(P.local_stiffness, 0.95f);
            for (int item = tt; item < nb * lanes; item += kFusedBlock) {
                const int s = item / lanes, j = item - s * lanes;
                float4* sp = &sPos[s * STRIDE];
                const float4* svb = &sVB[s * STRIDE];
                const float2* snb = &sNB[s * STRIDE];
                // The sweep is one dependent chain per strand: raise this
                // wave's issue priority while it runs.
                __builtin_amdgcn_s_setprio(3);
                for (int group = 0; group < npass; group += lanes) {
                    const bool live = group + j < npass;
                    const int lag = kLscLag * j;
                    const int steps = (V - 2) + kLscLag * (min(lanes, npass - group) - 1);
                    // Every lane runs every iteration, so the loop body
                    // branches only to start a pass: a lane outside its pass's
                    // steps works on clamped indices and stores into the
                    // strand's padding slot (index V), which no step uses.
                    // Reads for step lv+1 are issued at step lv (for the first
                    // step, before the loop); vertex 0 never moves.
                    const float4 root = sp[0];
                    int ln = min(max(1 - lag, 1), V - 2);
                    float4 ppN = sp[ln + 1], vbwN = svb[ln];
                    float2 nyzN = snb[ln];
                    float4 pm = root, p = root;
                    auto step = [&](const int it) {
                        const int lv = it - lag;
                        const bool act = live && lv >= 1 && lv <= V - 2;
                        if (lv == 1) { pm = root; p = sp[1]; }
                        float4 pp = ppN;
                        const float4 vbw = vbwN;
                        const float2 nyz = nyzN;
                        ln = min(max(lv + 1, 1), V - 2);
                        ppN = sp[ln + 1];
                        vbwN = svb[ln];
                        nyzN = snb[ln];
                        const f3 vecBindPose = xyz(vbw);
                        const f3 nLastBind{vbw.w, nyz.x, nyz.y};
#ifndef TFX_FAST_BUILD
                        f3 lastVec = sub(xyz(p), xyz(pm));
                        f3 org = add(rotate_from_to(nLastBind, normalize3(lastVec), vecBindPose), xyz(p));
                        f3 del = mul(stiff, sub(org, xyz(pp)));
                        if (movable(p)) { p.x -= del.x; p.y -= del.y; p.z -= del.z; }
                        if (movable(pp)) { pp.x += del.x; pp.y += del.y; pp.z += del.z; }
#else
                        const f3 org = lsc_target(nLastBind, vecBindPose, xyz(pm), xyz(p));
                        // the stiffness of an immovable vertex is zero
                        const f3 dif = sub(org, xyz(pp));
                        const float sp0 = movable(p) ? stiff : 0.0f, sp1 = movable(pp) ? stiff : 0.0f;
                        p.x -= sp0 * dif.x; p.y -= sp0 * dif.y; p.z -= sp0 * dif.z;
                        pp.x += sp1 * dif.x; pp.y += sp1 * dif.y; pp.z += sp1 * dif.z;
#endif
                        // vertex lv is final for this pass; lv+1 is final
                        // once the next step has moved it (or, after the
                        // last step, now: vertex V-1)
                        sp[act ? lv : V] = p;
                        sp[act ? lv + 1 : V] = pp;
                        pm = p; p = pp;
                        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                        __builtin_amdgcn_wave_barrier();
                        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    };
                    // four steps per trip: the compiler renames the values a
                    // step hands to the next instead of copying them
                    int it = 1;
                    for (; it + 3 <= steps; it += 4) {
                        step(it);
                        step(it + 1);
                        step(it + 2);
                        step(it + 3);
                    }
                    for (; it <= steps; ++it) step(it);
                }
                __builtin_amdgcn_s_setprio(0);
            }
        }
        blended = ahead;
        __syncthreads();

        // ---- P5: wind, length constraints, capsules ------------------------
        // In the vertex-parallel mapping a strand's vertices lie in
        // consecutive lanes of one wave (64 % V == 0), and P5 reads only
        // vertices of the strand it writes, so its steps (and P6a's reads of
        // them) are ordered within each wave instead of across the block.
        if (P.wind_on) {
            f3 force[NPT];
#pragma unroll
            for (int k = 0; k < NPT; ++k) {
                const int lv = k * kFusedBlock + t;
                const int ls = lv / V, v = lv % V;
                force[k] = f3{0.f, 0.f, 0.f};
                if (ls < nb && v >= 2 && v < V - 1) {
                    const int strand = sIdx[ls] * (F + 1);
                    float a = (float)(uint32_t)(strand % 20) / 20.0f;
                    f3 wv = add(add(add(mul(a, xyz(ld4(P.wind0))), mul(1.0f - a, xyz(ld4(P.wind1)))), mul(a, xyz(ld4(P.wind2)))),
                                mul(1.0f - a, xyz(ld4(P.wind3))));
                    f3 d = sub(xyz(sPos[ls * STRIDE + v]), xyz(sPos[ls * STRIDE + v + 1]));
                    f3 c = cross3(cross3(d, wv), d);
                    force[k] = f3{-c.x, -c.y, -c.z};
                }
            }
            wave_sync();
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
            wave_sync();
        }
        // Red-black length constraints: the even pairs (v, v+1), then the
        // odd ones, each pair on its own lane.
        for (int it = 0; it < P.length_iters; ++it) {
#pragma unroll
            for (int m = 0; m < NPP; ++m) {
                const int ls = pairLv[m] / V, v = pairLv[m] % V;
                if (ls < nb) {
                    float4 a = sPos[ls * STRIDE + v], b = sPos[ls * STRIDE + v + 1];
                    distance_constraint(a, b, restLen[m].x);
                    sPos[ls * STRIDE + v] = a; sPos[ls * STRIDE + v + 1] = b;
                }
            }
            wave_sync();
#pragma unroll
            for (int m = 0; m < NPP; ++m) {
                const int ls = pairLv[m] / V, v = pairLv[m] % V + 1;
                if (ls < nb && (v >> 1) < (V - 1) / 2) {
                    float4 a = sPos[ls * STRIDE + v], b = sPos[ls * STRIDE + v + 1];
                    distance_constraint(a, b, restLen[m].y);
                    sPos[ls * STRIDE + v] = a; sPos[ls * STRIDE + v + 1] = b;
                }
            }
            wave_sync();
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
                        if (capsule_collision(cur, prevReg[k], np, ld4(P.cap0[c]), ld4(P.cap1[c]))) {
                            cur.x = np.x; cur.y = np.y; cur.z = np.z;
                            hitMask |= 1u << k;
                        }
                    }
                    sPos[ls * STRIDE + v] = cur;
                }
            }
            wave_sync();
        }

        // ---- P6a: guide tangent, clamp, SDF collision, stores ---------------
        // Vertices are taken in groups of four so that all their SDF lookups
        // are in flight together.  A vertex whose brick neighbourhood holds no
        // distance other than 1e10 skips the lookup (it would return 1e10).
        SdfGrid grid;
        grid.ox = P.sdf_ox; grid.oy = P.sdf_oy; grid.oz = P.sdf_oz; grid.cell = P.sdf_cell;
        grid.nx = P.sdf_nx; grid.ny = P.sdf_ny; grid.nz = P.sdf_nz; grid.pad = 0;
        grid.bricks = bits_lds ? nullptr : h.sdf_bricks;
        grid.brick_bits = nullptr;
        grid.inv_cell = P.sdf_inv_cell;
        grid.bx = P.sdf_bx; grid.by = P.sdf_by; grid.bz = P.sdf_bz;
        auto maybe_near = [&](f3 p) -> bool {
            if (!bits_lds) return true;
            int b;
            return sdf_brick_of(grid, p, b) && ((sBits[b >> 5] >> (b & 31)) & 1u);
        };
        const float margin = P.sdf_margin;
        const float C = P.clamp;
        const bool tw_uniform = P.tangent_w_uniform != 0;
        const float* __restrict__ sdfp = h.sdf;
        constexpr int U = 2;
#pragma unroll
        for (int k0 = 0; k0 < NPT; k0 += U) {
            f3 cp[U], outp[U];
            bool act[U], hit[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = k0 + u;
                const int lv = k * kFusedBlock + t;
                const int ls = lv / V, v = lv % V;
                act[u] = false;
                cp[u] = f3{0.f, 0.f, 0.f};
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
                    prevReg[k] = pv;  // the previous position the vertex keeps
                    cp[u] = xyz(sp);
                    act[u] = has_sdf && v >= 2 && maybe_near(cp[u]);
                }
            }
            // sNB is dead from P4 on: each wave packs its projections in a
            // 64-float4 part of it (B * STRIDE >= VB + B >= 512 slots)
            sdf_collide_wave<U>(grid, sdfp, margin, cp, act, outp, hit, reinterpret_cast<float4*>(sNB) + (t >> 6) * 64);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = k0 + u;
                const int lv = k * kFusedBlock + t;
                const int ls = lv / V, v = lv % V;
                if (ls < nb) {
                    const size_t i = ((size_t)sIdx[ls] * (F + 1)) * V + v;
                    const int base = ls * STRIDE;
                    const int ti = (v == V - 1) ? v - 1 : v;
                    const f3 tan = normalize3(sub(xyz(sPos[base + ti + 1]), xyz(sPos[base + ti])));
                    float4 pout = sPos[base + v];
                    float4 pv = prevReg[k];
                    if (hit[u]) {
                        pout.x = outp[u].x; pout.y = outp[u].y; pout.z = outp[u].z;
                        pv.x = outp[u].x; pv.y = outp[u].y; pv.z = outp[u].z;
                    }
                    stg(&h.pos[i], pout);
                    stg(&h.prev[i], pv);
                    float tw;
                    if (tw_uniform) {
                        tw = P.tangent_w;
                        stg(&h.tangent[i], make_float4(tan.x, tan.y, tan.z, tw));
                    } else {
                        tw = ldg(&h.tangent[i].w);
                        store_xyz(&h.tangent[i], tan);
                    }
                    // sVB is dead after the LSC sweep: it now carries the
                    // guide tangent (with the w the follow strands copy) to P6b.
                    sVB[base + v] = make_float4(tan.x, tan.y, tan.z, tw);
                }
            }
        }

        // ---- P6b: follow strands (UpdateFollowHairVertices + collision) -----
        // Follow vertex idx = (j - 1) * nb * V + lv of the block; consecutive
        // lanes take consecutive vertices of one follow strand.  Pass 1 stores
        // every vertex that provably stays put and queues the others (in the
        // dead sNB storage); pass 2 runs the SDF projection densely over the
        // queue, four vertices per lane at a time.
        if (F > 0) {
            if (t == 0) sQueueCount = 0;
            __syncthreads();
            const int per = nb * V;
            const int nf = F * per;
            const float tip = P.tip_sep;
            const bool staged = F <= kMaxStagedFollow;
            uint16_t* queue = reinterpret_cast<uint16_t*>(sNB);   // 4 * B * STRIDE >= F * VB entries when staged
            // follow strand j (1..F) of the item's guide vertex lv
            auto follow_vertex_at = [&](int j, int lv, size_t& fi, f3& fp, float4& ft, int& v, float& gw) {
                const int ls = lv / V;
                v = lv % V;
                const size_t strand = (size_t)sIdx[ls] * (F + 1);
                fi = (strand + j) * V + v;
                const float4 sp = sPos[ls * STRIDE + v];
                gw = sp.w;
                ft = sVB[ls * STRIDE + v];
                const float factor = tip * ((float)v / (float)V) + 1.0f;
                const f3 off = xyz(staged ? sOff[ls][j - 1] : ldg(&h.follow_off[strand + j]));
                fp = add(xyz(sp), mul(factor, off));
            };
            auto follow_vertex = [&](int idx, size_t& fi, f3& fp, float4& ft, int& v, float& gw) {
                const int j = idx / per + 1;
                follow_vertex_at(j, idx - (j - 1) * per, fi, fp, ft, v, gw);
            };
            if (staged) {
                // strand by strand, so that no lane divides by the item's size
                for (int j = 1; j <= F; ++j) {
                    for (int lv = t; lv < per; lv += kFusedBlock) {
                        const int idx = (j - 1) * per + lv;
                        size_t fi;
                        f3 fp;
                        float4 ft;
                        int v;
                        float gw;
                        follow_vertex_at(j, lv, fi, fp, ft, v, gw);
                        if (has_sdf && v >= 2 && maybe_near(fp)) {
                            queue[atomicAdd(&sQueueCount, 1)] = (uint16_t)idx;
                        } else {
                            if (h.follow_w_guide) stg(&h.pos[fi], make_float4(fp.x, fp.y, fp.z, gw));
                            else store_xyz(&h.pos[fi], fp);
                            stg(&h.tangent[fi], ft);
                        }
                    }
                }
                __syncthreads();
            }
            const int nq = staged ? sQueueCount : nf;
            // Store index and tangent of follow vertex idx, recomputed after the
            // projection rather than held across it.
            auto follow_slot = [&](int idx, size_t& fi, float4& ft) {
                const int j = idx / per + 1;
                const int lv = idx - (j - 1) * per;
                const int ls = lv / V, v = lv % V;
                fi = ((size_t)sIdx[ls] * (F + 1) + j) * V + v;
                ft = sVB[ls * STRIDE + v];
            };
            for (int q0 = t; q0 < nq; q0 += 4 * kFusedBlock) {
                f3 fp[4], outp[4];
                bool act[4], hit[4], ok[4];
                int qi[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int q = q0 + u * kFusedBlock;
                    ok[u] = q < nq;
                    act[u] = false;
                    fp[u] = f3{0.f, 0.f, 0.f};
                    qi[u] = 0;
                    if (ok[u]) {
                        int v;
                        size_t fi;
                        float4 ft;
                        float gw;
                        qi[u] = staged ? (int)queue[q] : q;
                        follow_vertex(qi[u], fi, fp[u], ft, v, gw);
                        act[u] = has_sdf && v >= 2 && (staged || maybe_near(fp[u]));
                    }
                }
                sdf_collide_batch<4>(grid, sdfp, margin, fp, act, outp, hit);
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (!ok[u]) continue;
                    size_t fi;
                    float4 ft;
                    follow_slot(qi[u], fi, ft);
                    if (hit[u]) store_xyz(&h.prev[fi], outp[u]);
                    store_xyz(&h.pos[fi], hit[u] ? outp[u] : fp[u]);
                    stg(&h.tangent[fi], ft);
                }
            }
        }
        __syncthreads();  // LDS is reused by the next work item
        w = wn;
    }
    // The queue is left at zero for the next launch: each block counts
    // itself out (queue[1]) after its last take, so the last one to do so
    // knows that every take of the launch has been made.
    if (queued && t == 0 && atomicAdd(&queue[1], 1u) == gridDim.x - 1) {
        __hip_atomic_store(&queue[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&queue[1], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

template <int V, int VB>
static hipError_t launch_fused(const DevHair* hairs, const WorkItem* items, int nitems, int grid, int defer_sdf,
                               uint32_t* queue, hipStream_t stream, hipEvent_t t0, hipEvent_t t1) {
    if (nitems <= 0) return hipSuccess;
    if (t0)
        hipExtLaunchKernelGGL((k_fused_substep<V, VB>), dim3(grid), dim3(kFusedBlock), 0, stream, t0, t1, 0, hairs,
                              items, nitems, defer_sdf, queue);
    else
        hipLaunchKernelGGL((k_fused_substep<V, VB>), dim3(grid), dim3(kFusedBlock), 0, stream, hairs, items, nitems,
                           defer_sdf, queue);
    return hipGetLastError();
}

// CollideHairVerticesWithSdf_forward :531-595 for the objects whose fused
// launch deferred it: one 64-lane block per work item (hair, first vertex),
// one vertex per lane, local vertices 0 and 1 skipped.  A lane's lookups form
// one dependent chain (brick occupancy, then the first corner, then the other
// seven, then the gradient); many small blocks keep many chains in flight,
// where the fused launch takes them one group after another inside each
// strand block.
constexpr int kCollideBlock = 64;

__global__ void __launch_bounds__(kCollideBlock) k_sdf_collide_items(const DevHair* __restrict__ hairs,
                                                                     const WorkItem* __restrict__ items) {
    const WorkItem item = items[blockIdx.x];
    const int hi = __builtin_amdgcn_readfirstlane(item.hair);
    const int v0 = __builtin_amdgcn_readfirstlane(item.g0);
    const DevHair h = load_uniform(hairs + hi);
    const_as<DevParams>& P = *uniform_const(h.params);
    const int i = v0 + (int)threadIdx.x;
    const int V = h.V, n = h.G * (h.F + 1) * V;
    if (P.has_sdf == 0 || i >= n || i % V < 2) return;
    SdfGrid grid;
    grid.ox = P.sdf_ox; grid.oy = P.sdf_oy; grid.oz = P.sdf_oz; grid.cell = P.sdf_cell;
    grid.nx = P.sdf_nx; grid.ny = P.sdf_ny; grid.nz = P.sdf_nz; grid.pad = 0;
    grid.bricks = h.sdf_bricks;
    grid.brick_bits = nullptr;
    grid.inv_cell = P.sdf_inv_cell;
    grid.bx = P.sdf_bx; grid.by = P.sdf_by; grid.bz = P.sdf_bz;
    f3 out;
    if (sdf_collide_vertex<false>(grid, h.sdf, P.sdf_margin, xyz(ldg(&h.pos[i])), out)) {
        store_xyz(&h.pos[i], out);
        store_xyz(&h.prev[i], out);
    }
}

}  // namespace TFX_KERNEL_NS
}  // namespace tfx

// Vertices per strand block for each strand length and launch-shape variant:
// variant 0 = 512, 1 = 1024, 2 = 2048 guide vertices (at most 128 strands).
extern "C" int TFX_KSYM(tfx_fused_block_vertices)(int V, int variant) {
    if (V != 4 && V != 8 && V != 16 && V != 32 && V != 64) return 0;
    if (variant < 0 || variant > 2) return 0;
    const int vb = 512 << variant;
    return vb / V > 128 ? 128 * V : vb;
}

// t0, t1: events stamped by the launch itself (kernel begin and end), or null.
extern "C" hipError_t TFX_KSYM(tfx_launch_collide_items)(const tfx::DevHair* hairs, const tfx::WorkItem* items,
                                                         int nitems, hipStream_t stream, hipEvent_t t0,
                                                         hipEvent_t t1) {
    using namespace tfx;
    if (nitems <= 0) return hipSuccess;
    if (t0)
        hipExtLaunchKernelGGL(k_sdf_collide_items, dim3(nitems), dim3(kCollideBlock), 0, stream, t0, t1, 0, hairs,
                              items);
    else
        hipLaunchKernelGGL(k_sdf_collide_items, dim3(nitems), dim3(kCollideBlock), 0, stream, hairs, items);
    return hipGetLastError();
}

// Vertices per work item of tfx_launch_collide_items.
extern "C" int TFX_KSYM(tfx_collide_chunk)() { return tfx::kCollideBlock; }

extern "C" hipError_t TFX_KSYM(tfx_launch_fused)(int V, int variant, const tfx::DevHair* hairs, const tfx::WorkItem* items,
                                       int nitems, int grid, int defer_sdf, uint32_t* queue,
                                       hipStream_t stream, hipEvent_t t0, hipEvent_t t1) {
    using namespace tfx;
    const int vb = TFX_KSYM(tfx_fused_block_vertices)(V, variant);
    if (queue == nullptr) return hipErrorInvalidValue;
    uint32_t* q = queue + 2 * fused_queue_class(V);
    switch (V * 10000 + vb) {
        case 4 * 10000 + 512: return launch_fused<4, 512>(hairs, items, nitems, grid, defer_sdf, q, stream, t0, t1);
        case 8 * 10000 + 512: return launch_fused<8, 512>(hairs, items, nitems, grid, defer_sdf, q, stream, t0, t1);
        case 8 * 10000 + 1024: return launch_fused<8, 1024>(hairs, items, nitems, grid, defer_sdf, q, stream, t0, t1);
        case 16 * 10000 + 512: return launch_fused<16, 512>(hairs, items, nitems, grid, defer_sdf, q, stream, t0, t1);
        case 16 * 10000 + 1024: return launch_fused<16, 1024>(hairs, items, nitems, grid, defer_sdf, q, stream, t0, t1);
        case 16 * 10000 + 2048: return launch_fused<16, 2048>(hairs, items, nitems, grid, defer_sdf, q, stream, t0, t1);
        case 32 * 10000 + 512: return launch_fused<32, 512>(hairs, items, nitems, grid, defer_sdf, q, stream, t0, t1);
        case 32 * 10000 + 1024: return launch_fused<32, 1024>(hairs, items, nitems, grid, defer_sdf, q, stream, t0, t1);
        case 32 * 10000 + 2048: return launch_fused<32, 2048>(hairs, items, nitems, grid, defer_sdf, q, stream, t0, t1);
        case 64 * 10000 + 512: return launch_fused<64, 512>(hairs, items, nitems, grid, defer_sdf, q, stream, t0, t1);
        case 64 * 10000 + 1024: return launch_fused<64, 1024>(hairs, items, nitems, grid, defer_sdf, q, stream, t0, t1);
        case 64 * 10000 + 2048: return launch_fused<64, 2048>(hairs, items, nitems, grid, defer_sdf, q, stream, t0, t1);
        default: return hipErrorInvalidValue;
    }
}

// Resident blocks per CU of one launch shape (0 if the shape does not exist).
extern "C" int TFX_KSYM(tfx_fused_occupancy)(int V, int variant) {
    using namespace tfx;
    const int vb = TFX_KSYM(tfx_fused_block_vertices)(V, variant);
    int n = 0;
    hipError_t e = hipErrorInvalidValue;
#define TFX_OCC(Vv, VBv) \
    if (V == Vv && vb == VBv) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fused_substep<Vv, VBv>, kFusedBlock, 0);
    TFX_OCC(4, 512) TFX_OCC(8, 512) TFX_OCC(8, 1024) TFX_OCC(16, 512) TFX_OCC(16, 1024) TFX_OCC(16, 2048)
    TFX_OCC(32, 512) TFX_OCC(32, 1024) TFX_OCC(32, 2048) TFX_OCC(64, 512) TFX_OCC(64, 1024) TFX_OCC(64, 2048)
#undef TFX_OCC
    return e == hipSuccess ? n : 0;
}
