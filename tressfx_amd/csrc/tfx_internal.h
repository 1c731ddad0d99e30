// tfx_internal.h — host-side object definitions shared by the C-ABI
// translation units (tfx_capi.hip, tfx_batch.hip): error reporting, per-pass
// GPU timing, pinned staging, the hair / simulation / SDF objects and the
// packing of their device parameter blocks.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "tfx_device.h"

using tfx::DevHair;
using tfx::DevParams;
using tfx::SdfGrid;

// The kernel sources are built twice (tfx_device.h): tfx_x is the correctly
// rounded build's entry point, tfx_x_fast the hardware-division build's.
// g_arith (TFX_ARITH_*) picks the build for every launch issued after it is
// set; kl_x forwards to that build.  The SDF build always runs the correctly
// rounded kernels, so the grid is the same in both modes (hardware division
// could move a triangle's node box by a node and with it the band).
inline std::atomic<int> g_arith{TFX_ARITH_IEEE};
#define TFX_DUAL(ret, name, params, args)                                           \
    extern "C" ret tfx_##name params;                                                \
    extern "C" ret tfx_##name##_fast params;                                         \
    inline ret kl_##name params { return g_arith == TFX_ARITH_FAST ? tfx_##name##_fast args : tfx_##name args; }
TFX_DUAL(hipError_t, launch_stage, (int stage, const tfx::DevHair* h, hipStream_t stream), (stage, h, stream))
extern "C" hipError_t tfx_launch_sdf_phase(int phase, uint32_t* sdf, int ncell, const tfx_standard_vertex* verts,
                                           const uint32_t* idx, int ntri, tfx::SdfGrid g, hipStream_t stream);
inline hipError_t kl_launch_sdf_phase(int phase, uint32_t* sdf, int ncell, const tfx_standard_vertex* verts,
                                      const uint32_t* idx, int ntri, tfx::SdfGrid g, hipStream_t stream) {
    return tfx_launch_sdf_phase(phase, sdf, ncell, verts, idx, ntri, g, stream);
}
TFX_DUAL(hipError_t, launch_sdf_collide,
         (float4* pos, float4* prev, int n, int V, const float* sdf, tfx::SdfGrid g, float margin, hipStream_t stream),
         (pos, prev, n, V, sdf, g, margin, stream))
TFX_DUAL(hipError_t, launch_generate_transforms,
         (const float4* rest, int G, int F, int V, const tfx_hair_to_triangle_mapping* map, const float4* mesh,
          const float4* skinned, int nverts, tfx_strand_transform* out, hipStream_t stream),
         (rest, G, F, V, map, mesh, skinned, nverts, out, stream))
TFX_DUAL(hipError_t, launch_sdf_collide_mixed,
         (float4* pos, float4* prev, int n, const float* sdf, tfx::SdfGrid g, float margin, hipStream_t stream),
         (pos, prev, n, sdf, g, margin, stream))
TFX_DUAL(hipError_t, launch_fused,
         (int V, int variant, const tfx::DevHair* hairs, const tfx::WorkItem* items, int nitems, int grid,
          int defer_sdf, uint32_t* queue, hipStream_t stream, hipEvent_t t0, hipEvent_t t1),
         (V, variant, hairs, items, nitems, grid, defer_sdf, queue, stream, t0, t1))
TFX_DUAL(hipError_t, launch_collide_items,
         (const tfx::DevHair* hairs, const tfx::WorkItem* items, int nitems, hipStream_t stream, hipEvent_t t0,
          hipEvent_t t1),
         (hairs, items, nitems, stream, t0, t1))
TFX_DUAL(int, collide_chunk, (), ())
TFX_DUAL(int, fused_block_vertices, (int V, int variant), (V, variant))
TFX_DUAL(int, fused_occupancy, (int V, int variant), (V, variant))
extern "C" hipError_t tfx_launch_sdf_multi(int phase, const tfx::SdfJob* jobs, int njobs, int max_cells, int max_tris,
                                           hipStream_t stream);
inline hipError_t kl_launch_sdf_multi(int phase, const tfx::SdfJob* jobs, int njobs, int max_cells, int max_tris,
                                      hipStream_t stream) {
    return tfx_launch_sdf_multi(phase, jobs, njobs, max_cells, max_tris, stream);
}
#undef TFX_DUAL

// ---------------------------------------------------------------------------
// Errors
// ---------------------------------------------------------------------------
inline thread_local std::string g_last_error;
inline thread_local int g_device = 0;

inline int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define HIP_TRY(expr)                                                                         \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) return fail(TFX_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

#define CHECK_ARG(cond, msg)                                            \
    do {                                                                \
        if (!(cond)) return fail(TFX_ERR_INVALID_ARGUMENT, "%s", msg);  \
    } while (0)

// ---------------------------------------------------------------------------
// Per-pass GPU timing (the role of GetDevice()->GetTimeStamp(name) after each
// dispatch, TressFXSimulation.cpp:87-106, TressFXSDFCollision.cpp:130-164):
// when enabled, a pair of HIP events brackets every kernel launch on its
// stream and the elapsed times are summed per pass name.
// ---------------------------------------------------------------------------
namespace tfx_timing {
struct TimerRec { hipEvent_t a, b; };
inline std::mutex g_tmu;
inline bool g_timing = false;
inline std::map<std::string, std::vector<TimerRec>> g_timers;
inline std::vector<hipEvent_t> g_event_pool;

inline hipEvent_t pool_event() {
    if (!g_event_pool.empty()) {
        hipEvent_t e = g_event_pool.back();
        g_event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

struct PassTimer {
    const char* label;
    hipStream_t s;
    hipEvent_t a = nullptr, b = nullptr;
    bool dispatch;
    bool stamped = false;
    // in_dispatch: the pass is one kernel launched with start() and stop()
    // (hipExtLaunchKernel), whose own begin and end stamp the pair: the
    // kernel's duration without the gap before its dispatch; the launch's
    // status goes through launched(), and a pair whose launch failed (never
    // stamped) is returned to the pool instead of being recorded.  Otherwise
    // the pair is recorded on the stream around the pass.
    PassTimer(const char* l, hipStream_t st, bool in_dispatch = false) : label(l), s(st), dispatch(in_dispatch) {
        if (!label) return;
        std::lock_guard<std::mutex> g(g_tmu);
        if (!g_timing) return;
        a = pool_event();
        if (!a) return;
        if (dispatch) {
            b = pool_event();
            if (!b) { g_event_pool.push_back(a); a = nullptr; }
            return;
        }
        (void)hipEventRecord(a, s);
    }
    hipEvent_t start() const { return dispatch ? a : nullptr; }
    hipEvent_t stop() const { return dispatch ? b : nullptr; }
    hipError_t launched(hipError_t e) {
        stamped = e == hipSuccess;
        return e;
    }
    ~PassTimer() {
        if (!a) return;
        std::lock_guard<std::mutex> g(g_tmu);
        if (dispatch && !stamped) {
            g_event_pool.push_back(a);
            if (b) g_event_pool.push_back(b);
            return;
        }
        if (!dispatch) {
            b = pool_event();
            if (!b) return;
            (void)hipEventRecord(b, s);
        }
        g_timers[label].push_back(TimerRec{a, b});
    }
};
}  // namespace tfx_timing
using tfx_timing::PassTimer;

// Pass-timer label of the fused launches of strand length V in strand blocks
// of VB guide vertices: "FusedSubstep/V<V>/B<VB>" (the launch's kernel is
// k_fused_substep<V, VB>).
inline const char* fused_label(int V, int VB) {
    static const char* const names[5][3] = {
        {"FusedSubstep/V4/B512", "FusedSubstep/V4/B1024", "FusedSubstep/V4/B2048"},
        {"FusedSubstep/V8/B512", "FusedSubstep/V8/B1024", "FusedSubstep/V8/B2048"},
        {"FusedSubstep/V16/B512", "FusedSubstep/V16/B1024", "FusedSubstep/V16/B2048"},
        {"FusedSubstep/V32/B512", "FusedSubstep/V32/B1024", "FusedSubstep/V32/B2048"},
        {"FusedSubstep/V64/B512", "FusedSubstep/V64/B1024", "FusedSubstep/V64/B2048"}};
    const int vi = V == 4 ? 0 : V == 8 ? 1 : V == 16 ? 2 : V == 32 ? 3 : V == 64 ? 4 : -1;
    const int bi = VB == 512 ? 0 : VB == 1024 ? 1 : VB == 2048 ? 2 : -1;
    return vi < 0 || bi < 0 ? "FusedSubstep" : names[vi][bi];
}

template <class T> inline hipError_t dalloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) return hipSuccess;
    return hipMalloc((void**)p, count * sizeof(T));
}

// ---------------------------------------------------------------------------
// Pinned staging ring: host writes slot k while slot k-1's copy may still be
// in flight (the role of the reference's double-buffered m_SimCB[2]).
// ---------------------------------------------------------------------------
struct Staging {
    void* host[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    size_t bytes = 0;
    int next = 0;

    int ensure(size_t n) {
        if (n <= bytes) return TFX_OK;
        release();
        size_t cap = std::max<size_t>(n, 4096);
        for (int k = 0; k < 2; ++k) {
            HIP_TRY(hipHostMalloc(&host[k], cap, hipHostMallocDefault));
            HIP_TRY(hipEventCreateWithFlags(&ev[k], hipEventDisableTiming));
        }
        bytes = cap;
        return TFX_OK;
    }
    // Returns a host slot that is free to overwrite.
    int acquire(void** out, int* slot) {
        int k = next;
        next ^= 1;
        HIP_TRY(hipEventSynchronize(ev[k]));
        *out = host[k];
        *slot = k;
        return TFX_OK;
    }
    int release_slot(int k, hipStream_t s) {
        HIP_TRY(hipEventRecord(ev[k], s));
        return TFX_OK;
    }
    void release() {
        for (int k = 0; k < 2; ++k) {
            if (ev[k]) { (void)hipEventSynchronize(ev[k]); (void)hipEventDestroy(ev[k]); ev[k] = nullptr; }
            if (host[k]) { (void)hipHostFree(host[k]); host[k] = nullptr; }
        }
        bytes = 0;
    }
    ~Staging() { release(); }
};

// ---------------------------------------------------------------------------
// Objects
// ---------------------------------------------------------------------------
struct tfx_hair {
    int G = 0, F = 0, V = 0, S = 0, N = 0;
    int device = 0;
    float4* pos = nullptr;
    float4* prev = nullptr;
    float4* prevprev = nullptr;
    float4* tangent = nullptr;
    tfx_strand_level_data* sld = nullptr;
    float4* rest = nullptr;
    float* restlen = nullptr;
    float4* follow_off = nullptr;
    tfx_bone_skinning_data* skin = nullptr;
    float* bones_d = nullptr;
    DevParams* params_d = nullptr;
    int* order = nullptr;                 // fused kernel's guide-strand order (nullptr: index order)

    tfx_simulation_params params{};       // m_SimCB (single block: kernel arguments are copied at launch)
    std::vector<float> bones_h;           // 512 x 16
    int bones_dirty = 0;                  // leading matrices not yet on the device
    int frame = 0;                        // m_SimulationFrame
    int bones_used = 1;                   // 1 + the largest bone index the skinning data references
    int tangent_w_uniform = 1;
    int follow_w_guide = 0;               // every follow vertex's w equals its guide vertex's w
    float tangent_w = 0.0f;
    Staging staging;
    hipEvent_t handoff = nullptr;         // sim <-> rendering stream barrier (created on first use)
};

struct tfx_simulation {
    int device = 0;
    Staging staging;
    void* table_d = nullptr;   // DevHair[] followed by WorkItem[]
    size_t table_bytes = 0;
    uint32_t* queue_d = nullptr;  // the fused launches' work queues (kFusedQueueWords)
    int variant[65];           // launch-shape variant per V (-1: chosen per call from occupancy)
    int per_cu[65];            // persistent blocks per CU per V (<= 0: the kernel's occupancy)
    int defer_sdf = TFX_SDF_DEFER_AUTO;  // batched form: SDF collision in the fused launch or a pass after it
};

struct tfx_sdf_collision_system {
    int device = 0;
};

struct tfx_sdf {
    tfx_sdf_mesh_desc mesh{};
    float origin[3] = {0, 0, 0};
    float cell = 0;
    int nx = 0, ny = 0, nz = 0;
    int total_cells = 0;     // allocation
    float margin = 0;        // in cell units
    float pad_boundary = 0;
    int explicit_grid = 0;
    uint32_t* cells = nullptr;
    uint8_t* bricks = nullptr;   // dilated 4^3-brick occupancy (conservative early-out of collision lookups)
    uint32_t* brick_bits = nullptr;
    int bx = 0, by = 0, bz = 0;
    int device = 0;
    int collide_variant = TFX_SDF_COLLIDE_FORWARD;
    // Bricks touched by the batched builds (tfx::SdfJob::band_cur), two per
    // brick: the build in flight writes band[band_parity], the previous one's
    // is the other.  band_valid = 0 once anything else has written the grid.
    uint8_t* band = nullptr;
    int band_parity = 0;
    int band_valid = 0;
};

// The band pointers of the next batched build of s, advancing its history
// (sparse: the build keeps the band bookkeeping, see tfx::sdf_multi_sparse;
// otherwise it rewrites the whole grid and the history is dropped).  Called
// once per build, in build order, right before it is launched.
inline void sdf_band_stage(tfx_sdf* s, tfx::SdfJob& jb, bool sparse) {
    jb.band_cur = nullptr;
    jb.band_prev = nullptr;
    jb.band_next = nullptr;
    if (!s->band || !sparse) {
        s->band_valid = 0;
        return;
    }
    const size_t nbrick = (size_t)s->bx * s->by * s->bz;
    uint8_t* cur = s->band + (s->band_parity ? nbrick : 0);
    uint8_t* other = s->band + (s->band_parity ? 0 : nbrick);
    jb.band_cur = cur;
    jb.band_prev = s->band_valid ? other : nullptr;
    jb.band_next = other;
    s->band_parity ^= 1;
    s->band_valid = 1;
}

struct tfx_mesh_skinning {
    int n = 0, ntri = 0, follow_bone = 0;
    tfx_standard_vertex* rest = nullptr;   // initial vertex positions (StandardVertex, w = 0)
    tfx_standard_vertex* out = nullptr;    // skinned vertices (the SDF's mesh buffer)
    tfx_bone_skinning_data* skin = nullptr;
    uint32_t* idx = nullptr;
    float* bones_d = nullptr;              // 512 float4x4
    std::vector<float> bones_h;
    int bones_used = 1;                    // 1 + the largest bone index the skinning data can read
    float bb_min[3] = {0, 0, 0}, bb_max[3] = {0, 0, 0};   // rest bounding cube
    float cur_min[3] = {0, 0, 0}, cur_max[3] = {0, 0, 0}; // after the last update
    Staging staging;
};

extern "C" hipError_t tfx_launch_mesh_skin_multi(const tfx::MeshSkinJob* jobs, int njobs, int max_vertices,
                                                 hipStream_t stream);

// The bounding box that follows the follow bone (TressFXBoneSkinning::
// GetBoundingBox, TressFXBoneSkinning.cpp:439-452): the rest cube moved by the
// translation the bone matrix gives its centre.
inline void mesh_follow_bbox(tfx_mesh_skinning* m) {
    const float* b = &m->bones_h[16 * m->follow_bone];
    float c[3], nc[3];
    for (int k = 0; k < 3; ++k) c[k] = (m->bb_max[k] + m->bb_min[k]) * 0.5f;
    for (int k = 0; k < 3; ++k) nc[k] = ((c[0] * b[k] + c[1] * b[4 + k]) + c[2] * b[8 + k]) + b[12 + k];
    for (int k = 0; k < 3; ++k) {
        m->cur_min[k] = m->bb_min[k] + (nc[k] - c[k]);
        m->cur_max[k] = m->bb_max[k] + (nc[k] - c[k]);
    }
}

// Bone table entries a skinning record can read: slot 0 always, slots 1..3
// when their weight is positive, index ftoi(bone_index) clamped to [0, 511]
// (bone_ptr).  Returns 1 + the largest.
inline int skin_bones_used(const tfx_bone_skinning_data* sd, long long count, long long stride) {
    int used = 1;
    for (long long i = 0; i < count; i += stride)
        for (int k = 0; k < 4; ++k)
            if (k == 0 || sd[i].bone_weight[k] > 0.0f) {
                const float fi = sd[i].bone_index[k];
                int idx = !(fi == fi) ? 0 : fi >= 2147483648.0f ? 2147483647 : fi <= -2147483648.0f ? (-2147483647 - 1)
                                                                                                     : (int)fi;
                idx = std::max(0, std::min(idx, TFX_MAX_NUM_BONES - 1));
                used = std::max(used, idx + 1);
            }
    return used;
}

inline void set_identity_bones(std::vector<float>& b) {
    b.assign(16 * TFX_MAX_NUM_BONES, 0.0f);
    for (int i = 0; i < TFX_MAX_NUM_BONES; ++i)
        for (int k = 0; k < 4; ++k) b[16 * i + 5 * k] = 1.0f;
}

inline void hair_free(tfx_hair* h) {
    (void)hipFree(h->pos); (void)hipFree(h->prev); (void)hipFree(h->prevprev); (void)hipFree(h->tangent);
    (void)hipFree(h->sld); (void)hipFree(h->rest); (void)hipFree(h->restlen); (void)hipFree(h->follow_off);
    (void)hipFree(h->skin); (void)hipFree(h->bones_d); (void)hipFree(h->params_d); (void)hipFree(h->order);
    if (h->handoff) (void)hipEventDestroy(h->handoff);
}

// Whether every follow vertex of `positions` (the object's layout) carries
// its guide vertex's w, bit for bit.
inline void compute_follow_w(tfx_hair* h, const float* positions) {
    h->follow_w_guide = 0;
    if (h->F == 0 || !positions) return;
    for (int g = 0; g < h->G; ++g) {
        const size_t base = (size_t)g * (h->F + 1) * h->V;
        for (int j = 1; j <= h->F; ++j)
            for (int v = 0; v < h->V; ++v)
                if (std::memcmp(&positions[4 * (base + (size_t)j * h->V + v) + 3], &positions[4 * (base + v) + 3], 4))
                    return;
    }
    h->follow_w_guide = 1;
}

inline void compute_tangent_w(tfx_hair* h, const float* tangents) {
    h->tangent_w_uniform = 1;
    h->tangent_w = 0.0f;
    if (!tangents) return;
    bool first = true;
    uint32_t w0 = 0;
    for (int g = 0; g < h->G; ++g) {
        const size_t base = (size_t)g * (h->F + 1) * h->V;
        for (int v = 0; v < h->V; ++v) {
            uint32_t w;
            std::memcpy(&w, &tangents[4 * (base + v) + 3], 4);
            if (first) { w0 = w; first = false; }
            else if (w != w0) { h->tangent_w_uniform = 0; return; }
        }
    }
    std::memcpy(&h->tangent_w, &w0, 4);
}

// DevParams from the packed reference parameter block.
inline void pack_dev_params(const tfx_hair* h, const tfx_sdf* sdf, DevParams& d) {
    const tfx_simulation_params& p = h->params;
    std::memset(&d, 0, sizeof(d));
    d.wind0 = make_float4(p.wind[0], p.wind[1], p.wind[2], p.wind[3]);
    d.wind1 = make_float4(p.wind1[0], p.wind1[1], p.wind1[2], p.wind1[3]);
    d.wind2 = make_float4(p.wind2[0], p.wind2[1], p.wind2[2], p.wind2[3]);
    d.wind3 = make_float4(p.wind3[0], p.wind3[1], p.wind3[2], p.wind3[3]);
    d.damping = p.shape[0];
    d.local_stiffness = p.shape[1];
    d.global_stiffness = p.shape[2];
    d.global_range = p.shape[3];
    d.gravity = p.grav_time_tip[0];
    d.dt = p.grav_time_tip[1];
    d.tip_sep = p.grav_time_tip[2];
    // Integrate's decay = exp(-dampingCoeff * g_TimeStep * 60) (TressFXSimulation.hlsl:443), once per substep.
    d.decay = std::exp((-p.shape[0] * p.grav_time_tip[1]) * 60.0f);
    d.length_iters = p.sim_ints[0];
    d.lsc_passes = p.cpu_local_shape_iterations;
    d.wind_on = (p.wind[0] != 0.0f || p.wind[1] != 0.0f || p.wind[2] != 0.0f) ? 1 : 0;
    d.num_capsules = std::min(std::max(p.num_collision_capsules[0], 0), TFX_MAX_NUM_COLLISION_CAPSULES);
    d.vsp_coeff = p.vsp[0];
    d.vsp_accel_threshold = p.vsp[1];
    d.reset = p.reset_positions;
    d.clamp = p.clamp_position_delta;
    for (int c = 0; c < TFX_MAX_NUM_COLLISION_CAPSULES; ++c) {
        d.cap0[c] = make_float4(p.center_and_radius0[c][0], p.center_and_radius0[c][1], p.center_and_radius0[c][2],
                                p.center_and_radius0[c][3]);
        d.cap1[c] = make_float4(p.center_and_radius1[c][0], p.center_and_radius1[c][1], p.center_and_radius1[c][2],
                                p.center_and_radius1[c][3]);
    }
    if (sdf) {
        d.sdf_ox = sdf->origin[0]; d.sdf_oy = sdf->origin[1]; d.sdf_oz = sdf->origin[2];
        d.sdf_cell = sdf->cell;
        d.sdf_nx = sdf->nx; d.sdf_ny = sdf->ny; d.sdf_nz = sdf->nz;
        d.has_sdf = 1;
        // m_CollisionMargin * m_CellSize (TressFXSDFCollision.cpp:194)
        d.sdf_margin = sdf->margin * sdf->cell;
        d.sdf_inv_cell = 1.0f / sdf->cell;
        d.sdf_bx = sdf->bx; d.sdf_by = sdf->by; d.sdf_bz = sdf->bz;
    }
    d.tangent_w_uniform = h->tangent_w_uniform;
    d.tangent_w = h->tangent_w;
}

// TressFXHairObject::UpdateConstantBuffer: stream-ordered upload of the
// parameter block (and of the bone matrices changed since the last upload).
inline int upload_params(tfx_hair* h, const tfx_sdf* sdf, hipStream_t stream) {
    const size_t bone_bytes = (size_t)h->bones_dirty * 16 * sizeof(float);
    int rc = h->staging.ensure(sizeof(DevParams) + 16 * TFX_MAX_NUM_BONES * sizeof(float));
    if (rc) return rc;
    void* host;
    int slot;
    if ((rc = h->staging.acquire(&host, &slot))) return rc;
    DevParams* dp = (DevParams*)host;
    pack_dev_params(h, sdf, *dp);
    HIP_TRY(hipMemcpyAsync(h->params_d, dp, sizeof(DevParams), hipMemcpyHostToDevice, stream));
    if (bone_bytes) {
        float* hb = (float*)((char*)host + sizeof(DevParams));
        std::memcpy(hb, h->bones_h.data(), bone_bytes);
        HIP_TRY(hipMemcpyAsync(h->bones_d, hb, bone_bytes, hipMemcpyHostToDevice, stream));
        h->bones_dirty = 0;
    }
    return h->staging.release_slot(slot, stream);
}

inline DevHair dev_hair(const tfx_hair* h, const tfx_sdf* sdf) {
    DevHair d;
    d.pos = h->pos; d.prev = h->prev; d.prevprev = h->prevprev; d.tangent = h->tangent; d.sld = h->sld;
    d.rest = h->rest; d.restlen = h->restlen; d.follow_off = h->follow_off; d.skin = h->skin;
    d.bones = h->bones_d; d.params = h->params_d;
    d.sdf = sdf ? reinterpret_cast<const float*>(sdf->cells) : nullptr;
    d.sdf_bricks = sdf ? sdf->bricks : nullptr;
    d.sdf_brick_bits = sdf ? sdf->brick_bits : nullptr;
    d.order = h->order;
    d.G = h->G; d.F = h->F; d.V = h->V; d.follow_w_guide = h->follow_w_guide;
    return d;
}

inline SdfGrid sdf_grid(const tfx_sdf* s) {
    SdfGrid g;
    g.ox = s->origin[0]; g.oy = s->origin[1]; g.oz = s->origin[2];
    g.cell = s->cell;
    g.nx = s->nx; g.ny = s->ny; g.nz = s->nz; g.pad = 0;
    g.bricks = s->bricks;
    g.brick_bits = s->brick_bits;
    g.inv_cell = 1.0f / s->cell;
    g.bx = s->bx; g.by = s->by; g.bz = s->bz;
    return g;
}

inline int cu_count(int device) {
    static thread_local int cached[64] = {0};
    if (device >= 0 && device < 64 && cached[device]) return cached[device];
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || n <= 0) n = 256;
    if (device >= 0 && device < 64) cached[device] = n;
    return n;
}

// Launch shape of one strand-length class: the variant fixed by
// tfx_simulation_set_launch_shape, or else the one whose resident blocks cover
// the class's strand blocks in the fewest rounds (ties: more resident blocks).
struct Shape { int variant; int resident; };

inline Shape choose_shape(const tfx_simulation* sim, int V, tfx_hair* const* hairs, int n, int cus) {
    auto resident_of = [&](int variant) {
        const int occ = kl_fused_occupancy(V, variant);
        const int cap = sim->per_cu[V];
        const int per_cu = cap > 0 ? std::min(cap, std::max(occ, 1)) : std::max(occ, 1);
        return per_cu * cus;
    };
    if (sim->variant[V] >= 0) return Shape{sim->variant[V], resident_of(sim->variant[V])};
    Shape best{0, resident_of(0)};
    long long best_rounds = -1;
    for (int variant = 0; variant <= 2; ++variant) {
        const int vb = kl_fused_block_vertices(V, variant);
        if (variant > 0 && vb == kl_fused_block_vertices(V, variant - 1)) continue;
        const int B = vb / V;
        long long items = 0;
        for (int i = 0; i < n; ++i)
            if (hairs[i]->V == V) items += (hairs[i]->G + B - 1) / B;
        const int res = resident_of(variant);
        const long long rounds = (items + res - 1) / res;
        if (best_rounds < 0 || rounds < best_rounds || (rounds == best_rounds && res > best.resident)) {
            best = Shape{variant, res};
            best_rounds = rounds;
        }
    }
    return best;
}

