// tfx_sdf_kernels.hip — signed distance field build from a triangle mesh and
// hair-vs-SDF collision, TressFXSDFCollision.hlsl.
//
// Build: Initialize (fill FloatFlip3(1e10)) -> Construct (per triangle, an
// atomicMin of FloatFlip3(signed distance) into every node of the triangle's
// padded cell box) -> Finalize (IFloatFlip3 in place).  atomicMin on the
// rotated-sign bit pattern orders by |d| with ties going to +d, so the grid is
// bit-identical for any triangle order or lane mapping.
//
// Construct maps a team of lanes to one triangle and spreads the z-columns of
// the triangle's cell box over the team (instead of one thread looping over
// the whole box), so large and small triangles share a wave without one lane
// serialising it; a node's atomic is skipped when its current value is
// already no larger.
#include "tfx_device.h"

namespace tfx {
inline namespace TFX_KERNEL_NS {

constexpr int kSdfBlock = 256;
constexpr int kConstructBlock = 256;  // threads of a construct block (one team per triangle)

// FloatFlip3 / IFloatFlip3 :92-100
__device__ __forceinline__ uint32_t float_flip3(float f) {
    uint32_t u = __float_as_uint(f);
    return (u << 1) | (u >> 31);
}
__device__ __forceinline__ uint32_t ifloat_flip3(uint32_t u) { return (u >> 1) | (u << 31); }

// Per-triangle constants of SignedDistancePointToTriangle :163-215 that do
// not depend on the query point.
struct TriPrep {
    f3 x0, x1, x2;
    f3 x02n;   // normalised x0 - x2
    f3 x12n;   // x1 - x2 made orthogonal to x02n, normalised
    float l0, l1, l2;
    f3 nTri;
    float rl0, rl2;          // 1 / l0, 1 / l2 (correctly rounded)
    f3 e10, e21, e20;        // edge vectors x1 - x0, x2 - x1, x2 - x0
    float q10, q21, q20;     // their squared lengths
    float rq10, rq21, rq20;  // and the reciprocals of those
};

__device__ __forceinline__ TriPrep tri_prep(f3 x0, f3 x1, f3 x2) {
    TriPrep t;
    t.x0 = x0; t.x1 = x1; t.x2 = x2;
    f3 x02 = sub(x0, x2);
    t.l0 = len3(x02) + 1e-30f;
    x02 = divs(x02, t.l0);
    f3 x12 = sub(x1, x2);
    t.l1 = dot3(x12, x02);
    x12 = sub(x12, mul(t.l1, x02));
    t.l2 = len3(x12) + 1e-30f;
    x12 = divs(x12, t.l2);
    t.x02n = x02;
    t.x12n = x12;
    t.nTri = cross3(sub(x1, x0), sub(x2, x0));
    t.rl0 = recip_fixed(t.l0);
    t.rl2 = recip_fixed(t.l2);
    t.e10 = sub(x1, x0);
    t.e21 = sub(x2, x1);
    t.e20 = sub(x2, x0);
    t.q10 = dot3(t.e10, t.e10);
    t.q21 = dot3(t.e21, t.e21);
    t.q20 = dot3(t.e20, t.e20);
    t.rq10 = recip_fixed(t.q10);
    t.rq21 = recip_fixed(t.q21);
    t.rq20 = recip_fixed(t.q20);
    return t;
}

// The six dot products SignedDistancePointToTriangle (:163-215) takes of the
// query point p: with px2 = p - x2, x12n.px2 and x02n.px2 (the barycentric
// frame), (x1-p).e10, (x2-p).e21 and (x2-p).e20 (DistancePointToEdge's
// parameter along each edge, before the division), and (p-x0).nTri (the
// side).  Each is (X + Y) + Z of one per-axis term per coordinate, a term
// depending on that coordinate of p alone.
struct TriDots {
    float b, a, e10, e21, e20, side;
};
__device__ __forceinline__ void tri_dot_terms(const TriPrep& t, float pv, int k, float (&o)[6]) {
    const float c0 = k == 0 ? t.x0.x : k == 1 ? t.x0.y : t.x0.z;
    const float c1 = k == 0 ? t.x1.x : k == 1 ? t.x1.y : t.x1.z;
    const float c2 = k == 0 ? t.x2.x : k == 1 ? t.x2.y : t.x2.z;
    const float px2 = pv - c2;
    o[0] = (k == 0 ? t.x12n.x : k == 1 ? t.x12n.y : t.x12n.z) * px2;
    o[1] = (k == 0 ? t.x02n.x : k == 1 ? t.x02n.y : t.x02n.z) * px2;
    o[2] = (c1 - pv) * (k == 0 ? t.e10.x : k == 1 ? t.e10.y : t.e10.z);
    o[3] = (c2 - pv) * (k == 0 ? t.e21.x : k == 1 ? t.e21.y : t.e21.z);
    o[4] = (c2 - pv) * (k == 0 ? t.e20.x : k == 1 ? t.e20.y : t.e20.z);
    o[5] = (pv - c0) * (k == 0 ? t.nTri.x : k == 1 ? t.nTri.y : t.nTri.z);
}

// The squared length DistancePointToEdge takes the square root of, given
// the edge parameter's numerator u = (x1 - p).x10.
__device__ __forceinline__ float sq_distance_point_to_edge_u(f3 p, f3 x0, f3 x1, float u, float q, float rq) {
    float t = div_fixed(u, q, rq);
    t = fmx(0.0f, fmn(t, 1.0f));
    f3 a = sub(p, add(mul(t, x0), mul(1.0f - t, x1)));
    return dot3(a, a);
}

// SignedDistancePointToTriangle :163-215 from the point and its six dot
// products.  Outside the face region the reference takes min(sqrt(e01),
// sqrt(e12), sqrt(e02)) of the three squared edge distances; the correctly
// rounded square root is monotone, so sqrt(min(e01, e12, e02)) is the same
// value (also for infinities and NaNs, which fmin drops) with one square root
// instead of three; the face region's length is the square root of its
// squared length too, so one square root serves both.
__device__ __forceinline__ float signed_distance(const TriPrep& t, f3 p, const TriDots& dt) {
    float b = div_fixed(dt.b, t.l2, t.rl2);
    float a = div_fixed(dt.a - t.l1 * b, t.l0, t.rl0);
    float c = (1.0f - a) - b;
    const float tol = 1e-8f;
    float e;
    if (a >= -tol && b >= -tol && c >= -tol) {
        f3 n = sub(p, add(add(mul(a, t.x0), mul(b, t.x1)), mul(c, t.x2)));
        e = dot3(n, n);
    } else {
        e = sq_distance_point_to_edge_u(p, t.x0, t.x1, dt.e10, t.q10, t.rq10);
        e = fmn(e, sq_distance_point_to_edge_u(p, t.x1, t.x2, dt.e21, t.q21, t.rq21));
        e = fmn(e, sq_distance_point_to_edge_u(p, t.x0, t.x2, dt.e20, t.q20, t.rq20));
    }
    const float d = sqrtf(e);
    return (dt.side < 0.0f) ? -d : d;
}

// i / n and i % n for 0 <= i < 2^24, n >= 1: the float estimate is within one
// of the quotient, and one correction step makes it exact.
__device__ __forceinline__ void divmod24(int i, int n, float inv_n, int& q, int& r) {
    q = (int)((float)i * inv_n);
    r = i - q * n;
    if (r < 0) {
        q -= 1;
        r += n;
    } else if (r >= n) {
        q += 1;
        r -= n;
    }
}

// Finalize of one node: IFloatFlip3, and a node closer than 1e10 marks every
// brick whose kBrickHalo-dilated extent contains it.
__device__ __forceinline__ void finalize_cell(uint32_t* __restrict__ cells, int i, const SdfGrid& g, float inv_nx,
                                              float inv_ny, bool fast_index) {
    const uint32_t u = cells[i];
    cells[i] = ifloat_flip3(u);
    if (u == float_flip3(1e10f) || !g.bricks) return;
    int x, y, z, r;
    if (fast_index) {
        divmod24(i, g.nx, inv_nx, r, x);
        divmod24(r, g.ny, inv_ny, z, y);
    } else {
        x = i % g.nx;
        r = i / g.nx;
        y = r % g.ny;
        z = r / g.ny;
    }
    const int x0 = max(x - kBrickHalo, 0) >> kBrickShift, x1 = min((x + kBrickHalo) >> kBrickShift, g.bx - 1);
    const int y0 = max(y - kBrickHalo, 0) >> kBrickShift, y1 = min((y + kBrickHalo) >> kBrickShift, g.by - 1);
    const int z0 = max(z - kBrickHalo, 0) >> kBrickShift, z1 = min((z + kBrickHalo) >> kBrickShift, g.bz - 1);
    uint8_t* bricks = const_cast<uint8_t*>(g.bricks);
    for (int bz = z0; bz <= z1; ++bz)
        for (int by = y0; by <= y1; ++by)
            for (int bx = x0; bx <= x1; ++bx) bricks[(bz * g.by + by) * g.bx + bx] = 1;
}

// Finalize of the four nodes 4q..4q+3, one quad of an x-row (nx % 4 == 0):
// the same result as finalize_cell on each.  Each touched node x marks the
// bricks (x-2)>>2 .. (x+2)>>2 along x, and every such range holds the quad's
// own brick, so their union is the range of the first and last touched node.
__device__ __forceinline__ void finalize_quad(uint32_t* __restrict__ cells, int q, const SdfGrid& g, float inv_nq,
                                              float inv_ny, bool fast_index) {
    uint4* p = reinterpret_cast<uint4*>(cells) + q;
    const uint4 u = ldg(p);
    stg(p, make_uint4(ifloat_flip3(u.x), ifloat_flip3(u.y), ifloat_flip3(u.z), ifloat_flip3(u.w)));
    const uint32_t far = float_flip3(1e10f);
    const int m = (u.x != far ? 1 : 0) | (u.y != far ? 2 : 0) | (u.z != far ? 4 : 0) | (u.w != far ? 8 : 0);
    if (m == 0 || !g.bricks) return;
    const int nq = g.nx >> 2;
    int xq, y, z, r;
    if (fast_index) {
        divmod24(q, nq, inv_nq, r, xq);
        divmod24(r, g.ny, inv_ny, z, y);
    } else {
        xq = q % nq;
        r = q / nq;
        y = r % g.ny;
        z = r / g.ny;
    }
    const int xlo = 4 * xq + __builtin_ctz(m), xhi = 4 * xq + 31 - __builtin_clz(m);
    const int x0 = max(xlo - kBrickHalo, 0) >> kBrickShift, x1 = min((xhi + kBrickHalo) >> kBrickShift, g.bx - 1);
    const int y0 = max(y - kBrickHalo, 0) >> kBrickShift, y1 = min((y + kBrickHalo) >> kBrickShift, g.by - 1);
    const int z0 = max(z - kBrickHalo, 0) >> kBrickShift, z1 = min((z + kBrickHalo) >> kBrickShift, g.bz - 1);
    uint8_t* bricks = const_cast<uint8_t*>(g.bricks);
    for (int bz = z0; bz <= z1; ++bz)
        for (int by = y0; by <= y1; ++by)
            for (int bx = x0; bx <= x1; ++bx) bricks[(bz * g.by + by) * g.bx + bx] = 1;
}

// Whole-grid passes over `ncell` nodes: four nodes per lane and iteration
// when the rows split into quads, else one.
__device__ __forceinline__ void init_cells(uint32_t* __restrict__ sdf, int ncell, int lane, int stride) {
    const uint32_t v = float_flip3(1e10f);
    if ((ncell & 3) == 0) {
        uint4* p = reinterpret_cast<uint4*>(sdf);
        for (int q = lane; q < (ncell >> 2); q += stride) stg(&p[q], make_uint4(v, v, v, v));
    } else {
        for (int i = lane; i < ncell; i += stride) stg(&sdf[i], v);
    }
}

__device__ __forceinline__ void finalize_cells(uint32_t* __restrict__ sdf, int ncell, const SdfGrid& g, int lane,
                                               int stride) {
    const bool fast = ncell <= (1 << 24);
    if ((g.nx & 3) == 0) {
        const float inv_nq = 1.0f / (float)(g.nx >> 2), inv_ny = 1.0f / (float)g.ny;
        for (int q = lane; q < (ncell >> 2); q += stride) finalize_quad(sdf, q, g, inv_nq, inv_ny, fast);
    } else {
        const float inv_nx = 1.0f / (float)g.nx, inv_ny = 1.0f / (float)g.ny;
        for (int i = lane; i < ncell; i += stride) finalize_cell(sdf, i, g, inv_nx, inv_ny, fast);
    }
}

// InitializeSignedDistanceField :301-312, plus clearing the brick occupancy.
__global__ void __launch_bounds__(kSdfBlock) k_sdf_init(uint32_t* __restrict__ sdf, int ncell,
                                                        uint8_t* __restrict__ bricks, int nbrick) {
    init_cells(sdf, ncell, blockIdx.x * kSdfBlock + threadIdx.x, gridDim.x * kSdfBlock);
    if (bricks)
        for (int i = blockIdx.x * kSdfBlock + threadIdx.x; i < nbrick; i += gridDim.x * kSdfBlock) bricks[i] = 0;
}

// FinalizeSignedDistanceField :400-412.  A cell that ends up closer than 1e10
// also marks every brick whose kBrickHalo-dilated extent contains it.
__global__ void __launch_bounds__(kSdfBlock) k_sdf_finalize(uint32_t* __restrict__ sdf, int ncell, SdfGrid g) {
    finalize_cells(sdf, ncell, g, blockIdx.x * kSdfBlock + threadIdx.x, gridDim.x * kSdfBlock);
}

// Word w of the one-bit-per-brick occupancy (bricks 32w..32w+31): a full
// word's 32 bytes are read as two 16-byte loads.
__device__ __forceinline__ uint32_t pack_brick_word(const uint8_t* __restrict__ bricks, int nbrick, int w) {
    uint32_t m = 0;
    if (w * 32 + 32 <= nbrick) {
        const uint4* p = reinterpret_cast<const uint4*>(bricks + w * 32);
        const uint4 a = ldg(p), b = ldg(p + 1);
        const uint32_t v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
        for (int k = 0; k < 8; ++k)
#pragma unroll
            for (int j = 0; j < 4; ++j) m |= (((v[k] >> (8 * j)) & 0xffu) != 0 ? 1u : 0u) << (4 * k + j);
    } else {
        for (int k = 0; w * 32 + k < nbrick; ++k) m |= (bricks[w * 32 + k] != 0 ? 1u : 0u) << k;
    }
    return m;
}

// Brick occupancy packed to one bit per brick (word w covers bricks 32w..32w+31).
__global__ void __launch_bounds__(kSdfBlock) k_sdf_pack_bricks(const uint8_t* __restrict__ bricks, int nbrick,
                                                               uint32_t* __restrict__ bits) {
    const int w = blockIdx.x * kSdfBlock + threadIdx.x;
    if (w * 32 >= nbrick) return;
    bits[w] = pack_brick_word(bricks, nbrick, w);
}

// Triangle `tri`'s vertices and its padded node box (ConstructSignedDistanceField
// :329-397: AABB +- one cell, GetSdfCoordinates, -1/+1, clamped to the grid).
__device__ __forceinline__ void triangle_node_box(const tfx_standard_vertex* verts, const uint32_t* idx, int tri,
                                                  const SdfGrid& g, f3& x0, f3& x1, f3& x2, int (&lo)[3],
                                                  int (&hi)[3]) {
    const uint32_t i0 = ldg(&idx[3 * tri + 0]), i1 = ldg(&idx[3 * tri + 1]), i2 = ldg(&idx[3 * tri + 2]);
    auto vertex = [&](uint32_t i) {
        const float* q = verts[i].position;
        return f3{ldg(&q[0]), ldg(&q[1]), ldg(&q[2])};
    };
    x0 = vertex(i0);
    x1 = vertex(i1);
    x2 = vertex(i2);
    const float m = g.cell;  // MARGIN
    f3 mn{fmn(x0.x, fmn(x1.x, x2.x)) - m, fmn(x0.y, fmn(x1.y, x2.y)) - m, fmn(x0.z, fmn(x1.z, x2.z)) - m};
    f3 mx{fmx(x0.x, fmx(x1.x, x2.x)) + m, fmx(x0.y, fmx(x1.y, x2.y)) + m, fmx(x0.z, fmx(x1.z, x2.z)) + m};
    sdf_coords(g, mn, lo[0], lo[1], lo[2]);
    sdf_coords(g, mx, hi[0], hi[1], hi[2]);
    const int dims[3] = {g.nx, g.ny, g.nz};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        lo[k] = (int)((uint32_t)lo[k] - 1u);
        hi[k] = (int)((uint32_t)hi[k] + 1u);
        lo[k] = max(0, min(lo[k], dims[k] - 1));
        hi[k] = max(0, min(hi[k], dims[k] - 1));
    }
}

// The nodes of one triangle's padded box (ConstructSignedDistanceField
// :329-397), spread over the kTeam lanes of a team: each lane walks whole
// z-columns of the box, with the x and y terms of the six dot products (their
// partial sums come first in the reference's order) formed once per column
// and the z terms once per node.  `sink(ci, local, f)` takes each node's
// FloatFlip3(signed distance): ci its grid index (GetSdfCellIndex :123-129,
// wrapping int arithmetic), local its index in the box `ub` (x fastest, extents
// uex, uey), which must contain this box.
template <int kTeam, typename Sink>
__device__ __forceinline__ void construct_nodes(const TriPrep& tp, const int (&lo)[3], const int (&hi)[3],
                                                const SdfGrid& g, int lane, const int (&ub)[3], int uex, int uey,
                                                Sink&& sink) {
    // An axis whose wrapped lower bound passes the upper one (a vertex more
    // than 2^31 cells below the origin) has no nodes, as in the reference's
    // nested loops.
    const int ex = max(hi[0] - lo[0] + 1, 0), ey = max(hi[1] - lo[1] + 1, 0);
    const int ncol = ex * ey;
    // Exact small-integer division through the float reciprocal: for
    // 0 <= c < 2^16, (c + 0.5) / e stays more than 0.5/e - 2^-8/e from any
    // integer, beyond the rounding of the two float operations.
    const bool fast = ncol < 65536;
    const float inv_x = 1.0f / (float)ex;
    const uint32_t plane = (uint32_t)g.nx * (uint32_t)g.ny;
    const int uplane = uex * uey;
    for (int c = lane; c < ncol; c += kTeam) {
        int xx, yy;
        if (fast) {
            yy = (int)(((float)c + 0.5f) * inv_x);
            xx = c - yy * ex;
        } else {
            xx = c % ex;
            yy = c / ex;
        }
        const int x = lo[0] + xx, y = lo[1] + yy;
        const float px = (float)x * g.cell + g.ox, py = (float)y * g.cell + g.oy;
        float tx[6], ty[6], xy[6];
        tri_dot_terms(tp, px, 0, tx);
        tri_dot_terms(tp, py, 1, ty);
#pragma unroll
        for (int k = 0; k < 6; ++k) xy[k] = tx[k] + ty[k];
        uint32_t ci = plane * (uint32_t)lo[2] + (uint32_t)g.nx * (uint32_t)y + (uint32_t)x;
        int local = ((lo[2] - ub[2]) * uey + (y - ub[1])) * uex + (x - ub[0]);
        for (int z = lo[2]; z <= hi[2]; ++z, ci += plane, local += uplane) {
            const float pz = (float)z * g.cell + g.oz;
            float tz[6];
            tri_dot_terms(tp, pz, 2, tz);
            const TriDots dt{xy[0] + tz[0], xy[1] + tz[1], xy[2] + tz[2], xy[3] + tz[3], xy[4] + tz[4], xy[5] + tz[5]};
            sink(ci, local, float_flip3(signed_distance(tp, f3{px, py, pz}, dt)));
        }
    }
}

// atomicMin of FloatFlip3 bits into a grid node, issued only if the node's
// current value (agent-coherent load) is larger: the field only decreases, so
// a value read at any time bounds the final one from above and skipping is
// exact.
__device__ __forceinline__ void grid_min(uint32_t* __restrict__ cells, uint32_t ci, uint32_t f) {
    TFX_ADDRSPACE(1) uint32_t* cell = (TFX_ADDRSPACE(1) uint32_t*)&cells[(int)ci];
    if (f < __hip_atomic_load(cell, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        __hip_atomic_fetch_min(cell, f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Nodes (4 bytes each) of the block's LDS staging box.
constexpr int kStageNodes = 6144;

// One round of ConstructSignedDistanceField :329-397 for the block's
// kConstructBlock / kTeam consecutive triangles tri0, tri0 + 1, ... (a team each).
// Neighbouring triangles of a mesh usually follow each other in its index
// buffer, so their padded boxes overlap: when the union of the round's boxes
// holds at most kStageNodes nodes, every node's minimum over the round is
// first taken in LDS (atomicMin on the same FloatFlip3 bits) and each touched
// node then goes to the grid once; otherwise every node goes to the grid
// directly.  Either way each grid node ends as the minimum over all
// triangles, so the grid is the same.  Called by every thread of the block.
template <int kTeam>
__device__ __forceinline__ void construct_round(uint32_t* __restrict__ cells, const tfx_standard_vertex* verts,
                                                const uint32_t* idx, int tri0, int ntri, const SdfGrid& g,
                                                uint32_t* sStage, int* sBox) {
    const int t = threadIdx.x, lane = t % kTeam, tri = tri0 + t / kTeam;
    if (t < 3) {
        sBox[t] = 0x7fffffff;
        sBox[3 + t] = -0x7fffffff - 1;
    }
    __syncthreads();
    f3 x0, x1, x2;
    int lo[3] = {0, 0, 0}, hi[3] = {-1, -1, -1};
    const bool valid = tri < ntri;
    if (valid) triangle_node_box(verts, idx, tri, g, x0, x1, x2, lo, hi);
    const bool nodes = valid && lo[0] <= hi[0] && lo[1] <= hi[1] && lo[2] <= hi[2];
    if (nodes && lane == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            atomicMin(&sBox[k], lo[k]);
            atomicMax(&sBox[3 + k], hi[k]);
        }
    }
    __syncthreads();
    // the union box (extents clamped, so that the volume cannot overflow);
    // staged only if no team is left out of it
    int ub[3], ue[3];
    long long vol = 1;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        ub[k] = sBox[k];
        const long long e = (long long)sBox[3 + k] - (long long)ub[k] + 1;
        ue[k] = (int)min(max(e, 0LL), (long long)kStageNodes + 1);
        vol *= ue[k];
    }
    const bool staged = vol > 0 && vol <= kStageNodes;
    if (staged) {
        for (int n = t; n < (int)vol; n += kConstructBlock) sStage[n] = 0xffffffffu;
        __syncthreads();
    }
    if (nodes) {
        const TriPrep tp = tri_prep(x0, x1, x2);
        if (staged)
            construct_nodes<kTeam>(tp, lo, hi, g, lane, ub, ue[0], ue[1],
                                   [&](uint32_t, int local, uint32_t f) { atomicMin(&sStage[local], f); });
        else
            construct_nodes<kTeam>(tp, lo, hi, g, lane, lo, 0, 0,
                                   [&](uint32_t ci, int, uint32_t f) { grid_min(cells, ci, f); });
    }
    if (staged) {
        __syncthreads();
        const uint32_t plane = (uint32_t)g.nx * (uint32_t)g.ny;
        const int uplane = ue[0] * ue[1];
        const float inv_x = 1.0f / (float)ue[0], inv_p = 1.0f / (float)uplane;
        for (int n = t; n < (int)vol; n += kConstructBlock) {
            const uint32_t f = sStage[n];
            if (f == 0xffffffffu) continue;
            // n < kStageNodes < 2^16: exact float-reciprocal division
            const int zz = (int)(((float)n + 0.5f) * inv_p), r = n - zz * uplane;
            const int yy = (int)(((float)r + 0.5f) * inv_x), xx = r - yy * ue[0];
            const uint32_t ci = plane * (uint32_t)(ub[2] + zz) + (uint32_t)g.nx * (uint32_t)(ub[1] + yy) +
                                (uint32_t)(ub[0] + xx);
            grid_min(cells, ci, f);
        }
    }
    __syncthreads();  // the next round resets sBox and reuses sStage
}

// ConstructSignedDistanceField :329-397: kTeam lanes per triangle, one round
// of consecutive triangles per block.
template <int kTeam>
__global__ void __launch_bounds__(kConstructBlock) k_sdf_construct(uint32_t* __restrict__ sdf,
                                                             const tfx_standard_vertex* __restrict__ verts,
                                                             const uint32_t* __restrict__ idx, int ntri, SdfGrid g) {
    __shared__ uint32_t sStage[kStageNodes];
    __shared__ int sBox[6];
    construct_round<kTeam>(sdf, verts, idx, blockIdx.x * (kConstructBlock / kTeam), ntri, g, sStage, sBox);
}

// CollideHairVerticesWithSdf_forward :531-595 — one lane per vertex, guide
// and follow strands alike; local vertices 0 and 1 are skipped.
__global__ void __launch_bounds__(kSdfBlock) k_sdf_collide(float4* __restrict__ pos, float4* __restrict__ prev,
                                                           int n, int V, const float* __restrict__ sdf, SdfGrid g,
                                                           float margin) {
    const int i = blockIdx.x * kSdfBlock + threadIdx.x;
    if (i >= n) return;
    const int lv = i % V;
    if (lv == 0 || lv == 1) return;
    const float4 hv = pos[i];
    f3 out;
    if (sdf_collide_vertex<false>(g, sdf, margin, xyz(hv), out)) {
        store_xyz(&pos[i], out);
        store_xyz(&prev[i], out);
    }
}

// CollideHairVerticesWithSdf :597-715 — one lane per vertex, every vertex.
__global__ void __launch_bounds__(kSdfBlock) k_sdf_collide_mixed(float4* __restrict__ pos, float4* __restrict__ prev,
                                                                 int n, const float* __restrict__ sdf, SdfGrid g,
                                                                 float margin) {
    const int i = blockIdx.x * kSdfBlock + threadIdx.x;
    if (i >= n) return;
    const float4 hv = pos[i];
    f3 out;
    if (sdf_collide_mixed_vertex(g, sdf, margin, xyz(hv), out)) {
        store_xyz(&pos[i], out);
        store_xyz(&prev[i], out);
    }
}

}  // namespace TFX_KERNEL_NS
}  // namespace tfx

// phase 0 = Initialize, 1 = Construct, 2 = Finalize, 3 = pack the brick occupancy to bits
extern "C" hipError_t TFX_KSYM(tfx_launch_sdf_phase)(int phase, uint32_t* sdf, int ncell, const tfx_standard_vertex* verts,
                                           const uint32_t* idx, int ntri, tfx::SdfGrid g, hipStream_t stream) {
    using namespace tfx;
    // init and finalize take four nodes per lane where the grid allows it
    const int cblocks = std::max(1, std::min((ncell / 4 + kSdfBlock - 1) / kSdfBlock, 4096));
    uint8_t* bricks = const_cast<uint8_t*>(g.bricks);
    const int nbrick = g.bx * g.by * g.bz;
    if (phase == 0) {
        if (ncell > 0)
            hipLaunchKernelGGL(k_sdf_init, dim3(cblocks), dim3(kSdfBlock), 0, stream, sdf, ncell, bricks, nbrick);
    } else if (phase == 1) {
        if (ntri > 0) {
            // one grid per launch: 16 lanes per triangle keep enough blocks
            // in flight for a single mesh (the batched build takes 8)
            constexpr int kTeam = 16;
            const long long lanes = (long long)ntri * kTeam;
            const int tblocks = (int)((lanes + kConstructBlock - 1) / kConstructBlock);
            hipLaunchKernelGGL(k_sdf_construct<kTeam>, dim3(tblocks), dim3(kConstructBlock), 0, stream, sdf, verts, idx, ntri,
                               g);
        }
    } else if (phase == 2) {
        if (ncell > 0) hipLaunchKernelGGL(k_sdf_finalize, dim3(cblocks), dim3(kSdfBlock), 0, stream, sdf, ncell, g);
    } else if (phase == 3) {
        if (bricks && g.brick_bits) {
            const int words = (nbrick + 31) / 32;
            hipLaunchKernelGGL(k_sdf_pack_bricks, dim3((words + kSdfBlock - 1) / kSdfBlock), dim3(kSdfBlock), 0, stream,
                               g.bricks, nbrick, const_cast<uint32_t*>(g.brick_bits));
        }
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

extern "C" hipError_t TFX_KSYM(tfx_launch_sdf_collide_mixed)(float4* pos, float4* prev, int n, const float* sdf, tfx::SdfGrid g,
                                                   float margin, hipStream_t stream) {
    using namespace tfx;
    if (n <= 0) return hipSuccess;
    const int blocks = (n + kSdfBlock - 1) / kSdfBlock;
    hipLaunchKernelGGL(k_sdf_collide_mixed, dim3(blocks), dim3(kSdfBlock), 0, stream, pos, prev, n, sdf, g, margin);
    return hipGetLastError();
}

extern "C" hipError_t TFX_KSYM(tfx_launch_sdf_collide)(float4* pos, float4* prev, int n, int V, const float* sdf,
                                             tfx::SdfGrid g, float margin, hipStream_t stream) {
    using namespace tfx;
    if (n <= 0) return hipSuccess;
    const int blocks = (n + kSdfBlock - 1) / kSdfBlock;
    hipLaunchKernelGGL(k_sdf_collide, dim3(blocks), dim3(kSdfBlock), 0, stream, pos, prev, n, V, sdf, g, margin);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Batched build: one launch per phase for many SDFs.  blockIdx.y selects the
// job; the job table lives in device memory so that a captured graph replays
// with new meshes and grid origins.
// ---------------------------------------------------------------------------
namespace tfx {
inline namespace TFX_KERNEL_NS {

// Brick of node i of a whole-grid pass.
__device__ __forceinline__ int node_brick(int i, const SdfGrid& g, float inv_nx, float inv_ny, bool fast) {
    int x, y, z, r;
    if (fast) {
        divmod24(i, g.nx, inv_nx, r, x);
        divmod24(r, g.ny, inv_ny, z, y);
    } else {
        x = i % g.nx;
        r = i / g.nx;
        y = r % g.ny;
        z = r / g.ny;
    }
    return ((z >> kBrickShift) * g.by + (y >> kBrickShift)) * g.bx + (x >> kBrickShift);
}

// The batched build's sparse form (SdfJob::band_cur).  Only bricks that a
// triangle's node box overlaps can leave Construct with a value other than
// FloatFlip3(1e10), so after Finalize every other node holds 1e10 — the value
// it already holds if it was outside the previous build's band as well.
// Initialize therefore writes FloatFlip3(1e10) into this build's band, resets
// the previous band's other bricks to 1e10 and leaves the rest; Finalize
// un-flips this build's band only.  The grid equals the full build's.  Marks
// beyond the touched bricks cost work, never results, so a band only has to
// cover them.

// Step 1: every brick a triangle's node box overlaps (the box Construct walks).
// The band was cleared at the end of the build before (a band that was not,
// after a build without bookkeeping, only holds extra marks: those bricks are
// initialised and finalised to 1e10, which they hold anyway).
__global__ void __launch_bounds__(kSdfBlock) k_sdf_band_mark_multi(const SdfJob* __restrict__ jobs) {
    const SdfJob& j = jobs[blockIdx.y];
    if (!j.band_cur) return;
    const SdfGrid g = j.g;
    for (int tri = blockIdx.x * kSdfBlock + threadIdx.x; tri < j.ntri; tri += gridDim.x * kSdfBlock) {
        f3 x0, x1, x2;
        int lo[3], hi[3];
        triangle_node_box(j.verts, j.idx, tri, g, x0, x1, x2, lo, hi);
        for (int bz = lo[2] >> kBrickShift; bz <= hi[2] >> kBrickShift; ++bz)
            for (int by = lo[1] >> kBrickShift; by <= hi[1] >> kBrickShift; ++by)
                for (int bx = lo[0] >> kBrickShift; bx <= hi[0] >> kBrickShift; ++bx)
                    j.band_cur[(bz * g.by + by) * g.bx + bx] = 1;
    }
}

// Step 3, InitializeSignedDistanceField :301-312 over the band (or the grid).
// With quads (nx % 4 == 0) kL lanes take one brick, each 16 / kL of its quads
// (4 nodes along x, one per y and z of the brick): neighbouring lanes write
// neighbouring quads and clean bricks cost one byte read.  kL trades lanes
// for work per lane: many grids fill the GPU with one lane per brick, one
// grid needs more.
template <int kL>
__global__ void __launch_bounds__(kSdfBlock) k_sdf_init_multi(const SdfJob* __restrict__ jobs) {
    const SdfJob& j = jobs[blockIdx.y];
    const int lane = blockIdx.x * kSdfBlock + threadIdx.x, stride = gridDim.x * kSdfBlock;
    const SdfGrid& g = j.g;
    const int nbrick = g.bx * g.by * g.bz;
    uint8_t* bricks = const_cast<uint8_t*>(g.bricks);
    if (!j.band_cur || (g.nx & 3) != 0) {
        // the occupancy map is rebuilt by Finalize
        if (bricks)
            for (int i = lane; i < nbrick; i += stride) bricks[i] = 0;
    }
    if (!j.band_cur) {
        init_cells(j.cells, j.ncell, lane, stride);
        return;
    }
    const uint32_t key = float_flip3(1e10f), flat = __float_as_uint(1e10f);
    if ((g.nx & 3) == 0) {
        const int nq = g.nx >> 2;
        uint4* p = reinterpret_cast<uint4*>(j.cells);
        for (int w = lane; w < nbrick * kL; w += stride) {
            const int b = w / kL, part = w % kL;
            if (bricks && part == 0) bricks[b] = 0;
            const bool cur = j.band_cur[b] != 0;
            if (!cur && j.band_prev && !j.band_prev[b]) continue;
            const uint32_t v = cur ? key : flat;
            const int bx = b % g.bx, r = b / g.bx, by = r % g.by, bz = r / g.by;
            for (int k = part; k < 16; k += kL) {
                const int y = (by << kBrickShift) + (k & 3), z = (bz << kBrickShift) + (k >> 2);
                if (y < g.ny && z < g.nz) stg(&p[(z * g.ny + y) * nq + bx], make_uint4(v, v, v, v));
            }
        }
    } else {
        const bool fast = j.ncell <= (1 << 24);
        const float inv_nx = 1.0f / (float)g.nx, inv_ny = 1.0f / (float)g.ny;
        for (int i = lane; i < j.ncell; i += stride) {
            const int b = node_brick(i, g, inv_nx, inv_ny, fast);
            if (j.band_cur[b]) stg(&j.cells[i], key);
            else if (!j.band_prev || j.band_prev[b]) stg(&j.cells[i], flat);
        }
    }
}

template <int kTeam>
__global__ void __launch_bounds__(kConstructBlock) k_sdf_construct_multi(const SdfJob* __restrict__ jobs) {
    __shared__ uint32_t sStage[kStageNodes];
    __shared__ int sBox[6];
    const SdfJob& j = jobs[blockIdx.y];
    const SdfGrid g = j.g;
    constexpr int kPerRound = kConstructBlock / kTeam;
    for (int tri0 = blockIdx.x * kPerRound; tri0 < j.ntri; tri0 += gridDim.x * kPerRound)
        construct_round<kTeam>(j.cells, j.verts, j.idx, tri0, j.ntri, g, sStage, sBox);
}

// Finalize of quads part, part + kL, ... of brick (bx, by, bz) (the brick's
// quad k has y = 4 by + k % 4, z = 4 bz + k / 4), nx % 4 == 0: the same
// un-flip and the same occupancy marks as finalize_quad on each quad.  A
// touched node marks the bricks its kBrickHalo-dilated extent meets, all within
// one brick of its own, so the marks are gathered as a 3x3x3 mask (bit
// 9 (dz+1) + 3 (dy+1) + (dx+1)) and each marked brick is stored once per lane.
template <int kL>
__device__ __forceinline__ void finalize_brick(uint32_t* __restrict__ cells, const SdfGrid& g, int bx, int by, int bz,
                                               int part) {
    const int nq = g.nx >> 2;
    uint4* p = reinterpret_cast<uint4*>(cells);
    const uint32_t far = float_flip3(1e10f);
    uint32_t marks = 0;
    auto quad = [&](int y, int z, const uint4& u) {
        const int q = (z * g.ny + y) * nq + bx;
        stg(&p[q], make_uint4(ifloat_flip3(u.x), ifloat_flip3(u.y), ifloat_flip3(u.z), ifloat_flip3(u.w)));
        const int m = (u.x != far ? 1 : 0) | (u.y != far ? 2 : 0) | (u.z != far ? 4 : 0) | (u.w != far ? 8 : 0);
        if (m == 0) return;
        const int xlo = 4 * bx + __builtin_ctz(m), xhi = 4 * bx + 31 - __builtin_clz(m);
        const int x0 = max(xlo - kBrickHalo, 0) >> kBrickShift, x1 = min((xhi + kBrickHalo) >> kBrickShift, g.bx - 1);
        const int y0 = max(y - kBrickHalo, 0) >> kBrickShift, y1 = min((y + kBrickHalo) >> kBrickShift, g.by - 1);
        const int z0 = max(z - kBrickHalo, 0) >> kBrickShift, z1 = min((z + kBrickHalo) >> kBrickShift, g.bz - 1);
        // bits lo..hi of a 3-bit row, relative to the own brick at bit 1
        auto row = [](int lo, int hi, int own) { return ((2u << (hi - own + 1)) - 1u) & ~((1u << (lo - own + 1)) - 1u); };
        const uint32_t xm = row(x0, x1, bx), ym = row(y0, y1, by), zm = row(z0, z1, bz);
#pragma unroll
        for (int dz = 0; dz < 3; ++dz)
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
                if (((zm >> dz) & 1u) && ((ym >> dy) & 1u)) marks |= xm << (9 * dz + 3 * dy);
    };
    const int ybase = by << kBrickShift, zbase = bz << kBrickShift;
    constexpr int kQ = 16 / kL;           // quads per lane
    constexpr int kRound = kQ < 8 ? kQ : 8; // loads in flight together
    if (ybase + 4 <= g.ny && zbase + 4 <= g.nz) {
        for (int r0 = 0; r0 < kQ; r0 += kRound) {
            uint4 u[kRound];
#pragma unroll
            for (int i = 0; i < kRound; ++i) {
                const int k = part + kL * (r0 + i);
                u[i] = ldg(&p[((zbase + (k >> 2)) * g.ny + ybase + (k & 3)) * nq + bx]);
            }
#pragma unroll
            for (int i = 0; i < kRound; ++i) {
                const int k = part + kL * (r0 + i);
                quad(ybase + (k & 3), zbase + (k >> 2), u[i]);
            }
        }
    } else {
        for (int k = part; k < 16; k += kL) {
            const int y = ybase + (k & 3), z = zbase + (k >> 2);
            if (y < g.ny && z < g.nz) quad(y, z, ldg(&p[(z * g.ny + y) * nq + bx]));
        }
    }
    if (!marks || !g.bricks) return;
    uint8_t* bricks = const_cast<uint8_t*>(g.bricks);
    for (int k = 0; k < 27; ++k)
        if ((marks >> k) & 1u)
            bricks[((bz + k / 9 - 1) * g.by + (by + (k / 3) % 3 - 1)) * g.bx + (bx + k % 3 - 1)] = 1;
}

// FinalizeSignedDistanceField :400-412 over the band (or the grid), kL lanes
// per brick as in the initialisation.
template <int kL>
__global__ void __launch_bounds__(kSdfBlock) k_sdf_finalize_multi(const SdfJob* __restrict__ jobs) {
    const SdfJob& j = jobs[blockIdx.y];
    const int lane = blockIdx.x * kSdfBlock + threadIdx.x, stride = gridDim.x * kSdfBlock;
    if (!j.band_cur) {
        finalize_cells(j.cells, j.ncell, j.g, lane, stride);
        return;
    }
    const SdfGrid& g = j.g;
    if ((g.nx & 3) == 0) {
        const int nbrick = g.bx * g.by * g.bz;
        for (int w = lane; w < nbrick * kL; w += stride) {
            const int b = w / kL;
            if (!j.band_cur[b]) continue;
            const int bx = b % g.bx, r = b / g.bx;
            finalize_brick<kL>(j.cells, g, bx, r % g.by, r / g.by, w % kL);
        }
    } else {
        const bool fast = j.ncell <= (1 << 24);
        const float inv_nx = 1.0f / (float)g.nx, inv_ny = 1.0f / (float)g.ny;
        for (int i = lane; i < j.ncell; i += stride)
            if (j.band_cur[node_brick(i, g, inv_nx, inv_ny, fast)]) finalize_cell(j.cells, i, g, inv_nx, inv_ny, fast);
    }
}

// Occupancy bits, and the next build's band cleared (this build has read it
// as its previous band).
__global__ void __launch_bounds__(kSdfBlock) k_sdf_pack_multi(const SdfJob* __restrict__ jobs) {
    const SdfJob& j = jobs[blockIdx.y];
    const int nbrick = j.g.bx * j.g.by * j.g.bz;
    uint32_t* bits = const_cast<uint32_t*>(j.g.brick_bits);
    for (int w = blockIdx.x * kSdfBlock + threadIdx.x; w * 32 < nbrick; w += gridDim.x * kSdfBlock) {
        if (bits && j.g.bricks) bits[w] = pack_brick_word(j.g.bricks, nbrick, w);
        if (j.band_next)
            for (int k = w * 32; k < min(w * 32 + 32, nbrick); ++k) j.band_next[k] = 0;
    }
}

}  // namespace TFX_KERNEL_NS
}  // namespace tfx

// Batched build of `njobs` SDFs (job table in device memory).  `max_cells` /
// `max_tris` bound the per-job work and size the grid's x extent.
extern "C" hipError_t TFX_KSYM(tfx_launch_sdf_multi)(int phase, const tfx::SdfJob* jobs, int njobs, int max_cells, int max_tris,
                                           hipStream_t stream) {
    using namespace tfx;
    if (njobs <= 0) return hipSuccess;
    const int cx = std::max(1, std::min((max_cells / 4 + kSdfBlock - 1) / kSdfBlock, 1024));
    // With band bookkeeping (sdf_multi_sparse, the same choice the host made
    // when it staged the jobs): a brick per lane for the marks, and kL lanes
    // per brick for init / finalize, one if the grids hold 2^18 bricks in all,
    // else four.
    const bool sparse = sdf_multi_sparse(max_cells, njobs);
    const int max_bricks = std::max(1, max_cells / 64);
    const int lpb = (long long)max_bricks * njobs >= (1 << 18) ? 1 : 4;
    const int lxg = std::max(1, std::min((max_bricks * lpb + kSdfBlock - 1) / kSdfBlock, 1024));
    auto band_pass = [&](void (*k1)(const SdfJob*), void (*k4)(const SdfJob*)) {
        hipLaunchKernelGGL(lpb == 1 ? k1 : k4, dim3(lxg, njobs), dim3(kSdfBlock), 0, stream, jobs);
    };
    switch (phase) {
        case 0:
            if (sparse) {
                const int txg = std::max(1, std::min((max_tris + kSdfBlock - 1) / kSdfBlock, 1024));
                hipLaunchKernelGGL(k_sdf_band_mark_multi, dim3(txg, njobs), dim3(kSdfBlock), 0, stream, jobs);
                band_pass(k_sdf_init_multi<1>, k_sdf_init_multi<4>);
            } else {
                hipLaunchKernelGGL(k_sdf_init_multi<1>, dim3(cx, njobs), dim3(kSdfBlock), 0, stream, jobs);
            }
            break;
        case 1: {
            // 8 lanes per triangle (32 triangles per staged round) over many
            // grids; a single grid takes 16 (smaller rounds, more blocks)
            auto launch = [&](auto team) {
                constexpr int kTeam = decltype(team)::value;
                const int tx = std::max(1, std::min((max_tris * kTeam + kConstructBlock - 1) / kConstructBlock, 4096));
                hipLaunchKernelGGL(k_sdf_construct_multi<kTeam>, dim3(tx, njobs), dim3(kConstructBlock), 0, stream,
                                   jobs);
            };
            if (njobs == 1) launch(std::integral_constant<int, 16>{});
            else launch(std::integral_constant<int, 8>{});
            break;
        }
        case 2:
            if (sparse) band_pass(k_sdf_finalize_multi<1>, k_sdf_finalize_multi<4>);
            else hipLaunchKernelGGL(k_sdf_finalize_multi<1>, dim3(cx, njobs), dim3(kSdfBlock), 0, stream, jobs);
            break;
        case 3: hipLaunchKernelGGL(k_sdf_pack_multi, dim3(std::max(1, cx / 32), njobs), dim3(kSdfBlock), 0, stream, jobs);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
