// tfx_capi.hip — the C-ABI of include/tressfx_amd.h: object lifetimes, the
// host-side parameter packing of TressFXHairObject, and the dispatch sequences
// of TressFXSimulation and TressFXSDFCollision over HIP streams.
#include <hip/hip_runtime.h>

#include <algorithm>
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

#include "tfx_internal.h"

using namespace tfx_timing;

extern "C" int tfx_profiling_enable(int enable) {
    std::lock_guard<std::mutex> g(g_tmu);
    g_timing = enable != 0;
    return TFX_OK;
}

extern "C" int tfx_profiling_read(const char* label, double* total_ms, int* launches) {
    if (!label || !total_ms || !launches) return fail(TFX_ERR_INVALID_ARGUMENT, "tfx_profiling_read: null argument");
    // `label` itself and every label that extends it with '/' (the fused
    // launches are recorded per strand length and block size,
    // "FusedSubstep/V16/B512", ...).
    std::vector<TimerRec> recs;
    {
        std::lock_guard<std::mutex> g(g_tmu);
        const size_t n = std::strlen(label);
        for (auto it = g_timers.begin(); it != g_timers.end();) {
            const std::string& k = it->first;
            if (k.compare(0, n, label) == 0 && (k.size() == n || k[n] == '/')) {
                recs.insert(recs.end(), it->second.begin(), it->second.end());
                it = g_timers.erase(it);
            } else {
                ++it;
            }
        }
    }
    double sum = 0.0;
    int rc = TFX_OK;
    for (auto& r : recs) {
        float ms = 0.f;
        if (hipEventSynchronize(r.b) != hipSuccess || hipEventElapsedTime(&ms, r.a, r.b) != hipSuccess)
            rc = fail(TFX_ERR_HIP, "tfx_profiling_read: event query failed");
        sum += ms;
    }
    {
        std::lock_guard<std::mutex> g(g_tmu);
        for (auto& r : recs) { g_event_pool.push_back(r.a); g_event_pool.push_back(r.b); }
    }
    *total_ms = sum;
    *launches = (int)recs.size();
    return rc;
}


// ---------------------------------------------------------------------------
// Host math of TressFXHairObject::SetWind (TressFXHairObject.cpp:172-235,
// Quaternion.cpp:79-89, :262-284; Vector3D.cpp:48-62).
// ---------------------------------------------------------------------------
namespace {
struct Q { float x, y, z, w; };
struct Vec { float x, y, z; };
Q q_axis_angle(Vec a, float ang) {
    float half = 0.5f * ang;
    float s = std::sin(half);
    return Q{s * a.x, s * a.y, s * a.z, std::cos(half)};
}
Q q_mul(Q a, Q b) {
    Q q;
    q.w = a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z;
    q.x = a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y;
    q.y = a.w * b.y + a.y * b.w + a.z * b.x - a.x * b.z;
    q.z = a.w * b.z + a.z * b.w + a.x * b.y - a.y * b.x;
    return q;
}
Vec v_cross(Vec a, Vec b) { return Vec{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
Vec q_rot(Q q, Vec v) {
    Vec qv{q.x, q.y, q.z};
    Vec uv = v_cross(qv, v), uuv = v_cross(qv, uv);
    float s = 2.0f * q.w;
    uv = Vec{uv.x * s, uv.y * s, uv.z * s};
    uuv = Vec{uuv.x * 2.0f, uuv.y * 2.0f, uuv.z * 2.0f};
    return Vec{(v.x + uv.x) + uuv.x, (v.y + uv.y) + uuv.y, (v.z + uv.z) + uuv.z};
}
Vec v_normalize(Vec v) {
    float d = std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    if (d == 0) return v;
    return Vec{v.x / d, v.y / d, v.z / d};
}
void set_wind(tfx_simulation_params& cb, const float dir[3], float mag, int frame) {
    float s = std::sin((float)frame * 0.01f);
    float wM = mag * (std::pow(s, 2.0f) + 0.5f);
    Vec wd = v_normalize(Vec{dir[0], dir[1], dir[2]});
    Vec X{1.0f, 0.0f, 0.0f};
    Vec xw = v_cross(X, wd);
    Q rotX{0, 0, 0, 1};
    float angle = std::asin(std::sqrt(xw.x * xw.x + xw.y * xw.y + xw.z * xw.z));
    if (angle > 0.001) rotX = q_axis_angle(v_normalize(xw), angle);
    const float cone = 40.f * 3.14159265359f / 180;
    const Vec axes[4] = {{0, 1.0f, 0}, {0, -1.0f, 0}, {0, 0, 1.0f}, {0, 0, -1.0f}};
    float* outs[4] = {cb.wind, cb.wind1, cb.wind2, cb.wind3};
    for (int k = 0; k < 4; ++k) {
        Vec nd = q_rot(q_mul(rotX, q_axis_angle(axes[k], cone)), X);
        outs[k][0] = nd.x * wM; outs[k][1] = nd.y * wM; outs[k][2] = nd.z * wM; outs[k][3] = 0;
    }
}
}  // namespace

// ---------------------------------------------------------------------------
// Library
// ---------------------------------------------------------------------------
extern "C" const char* tfx_last_error(void) { return g_last_error.c_str(); }

extern "C" int tfx_get_version(int* major, int* minor, int* patch) {
    if (major) *major = 4;
    if (minor) *minor = 0;
    if (patch) *patch = 0;
    return TFX_OK;
}

static int noalloc_scratch(int device, void** out);

extern "C" int tfx_init(int hip_device) {
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    CHECK_ARG(hip_device >= 0 && hip_device < n, "tfx_init: no such HIP device");
    HIP_TRY(hipSetDevice(hip_device));
    g_device = hip_device;
    // the no-allocation path's scratch array on this device
    void* scratch_ptr = nullptr;
    return noalloc_scratch(hip_device, &scratch_ptr);
}

extern "C" int tfx_set_arithmetic(int mode) {
    CHECK_ARG(mode == TFX_ARITH_IEEE || mode == TFX_ARITH_FAST, "tfx_set_arithmetic: unknown mode");
    g_arith = mode;
    return TFX_OK;
}

extern "C" int tfx_get_arithmetic(int* mode) {
    CHECK_ARG(mode, "tfx_get_arithmetic: null argument");
    *mode = g_arith;
    return TFX_OK;
}

extern "C" void tfx_simulation_settings_default(tfx_simulation_settings* s) {
    if (!s) return;
    s->vsp_coeff = 0.8f;
    s->vsp_accel_threshold = 1.4f;
    s->local_constraint_stiffness = 0.9f;
    s->local_constraints_iterations = 2;
    s->global_constraint_stiffness = 0.0f;
    s->global_constraints_range = 0.0f;
    s->length_constraints_iterations = 2;
    s->damping = 0.08f;
    s->gravity_magnitude = 0.0f;
    s->tip_separation = 0.0f;
    s->wind_magnitude = 0.0f;
    s->wind_direction[0] = 1.0f;
    s->wind_direction[1] = 0.0f;
    s->wind_direction[2] = 0.0f;
    s->wind_angle_radians = 3.1415926f / 180.0f * 40.0f;
    s->clamp_position_delta = 20.0f;
}

// ---------------------------------------------------------------------------
// TressFXHairObject
// ---------------------------------------------------------------------------
// The fused kernel's processing order of the guide strands: by the Morton
// code of each guide's root (vertex 0), 10 bits per axis over the roots'
// bounding box, ties in index order.  A strand block then holds neighbouring
// strands, whose SDF lookups share cells.  Empty when the order is the index
// order already.
static std::vector<int> guide_order(const float* positions, int G, int F, int V) {
    std::vector<int> order;
    if (G < 2) return order;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    auto root = [&](int g, int k) { return positions[4 * ((size_t)g * (F + 1) * V) + k]; };
    for (int g = 0; g < G; ++g)
        for (int k = 0; k < 3; ++k) {
            const float x = root(g, k);
            if (std::isfinite(x)) { lo[k] = std::min(lo[k], x); hi[k] = std::max(hi[k], x); }
        }
    auto spread = [](uint64_t x) {
        x &= 0x3ff;
        x = (x | (x << 16)) & 0x030000ffull;
        x = (x | (x << 8)) & 0x0300f00full;
        x = (x | (x << 4)) & 0x030c30c3ull;
        x = (x | (x << 2)) & 0x09249249ull;
        return x;
    };
    std::vector<uint64_t> key(G);
    for (int g = 0; g < G; ++g) {
        uint64_t code = 0;
        for (int k = 0; k < 3; ++k) {
            const float ext = hi[k] - lo[k];
            const float x = root(g, k);
            int q = 0;
            if (ext > 0.0f && std::isfinite(x)) q = (int)std::min(1023.0f, std::max(0.0f, (x - lo[k]) / ext * 1023.0f));
            code |= spread((uint64_t)q) << k;
        }
        key[g] = (code << 32) | (uint64_t)g;   // unique keys: ties keep index order
    }
    std::vector<int> idx(G);
    for (int g = 0; g < G; ++g) idx[g] = g;
    std::sort(idx.begin(), idx.end(), [&](int a, int b) { return key[a] < key[b]; });
    bool identity = true;
    for (int g = 0; g < G && identity; ++g) identity = idx[g] == g;
    if (!identity) order.swap(idx);
    return order;
}

extern "C" int tfx_hair_create(const tfx_asset_desc* a, tfx_hair** out) {
    CHECK_ARG(a && out, "tfx_hair_create: null argument");
    *out = nullptr;
    const int V = a->num_vertices_per_strand;
    CHECK_ARG(V > 2 && V <= TFX_SIM_THREAD_GROUP_SIZE && TFX_SIM_THREAD_GROUP_SIZE % V == 0,
              "tfx_hair_create: vertices per strand must be 4, 8, 16, 32 or 64 (TressFXAsset.cpp:118)");
    CHECK_ARG(a->num_guide_strands > 0, "tfx_hair_create: no guide strands");
    CHECK_ARG(a->num_follow_strands_per_guide >= 0, "tfx_hair_create: negative follow count");
    CHECK_ARG(a->positions && a->rest_lengths && a->bone_skinning_data, "tfx_hair_create: missing asset array");
    const long long S = (long long)a->num_guide_strands * (a->num_follow_strands_per_guide + 1);
    const long long N = S * V;
    CHECK_ARG(N < (1ll << 31), "tfx_hair_create: more than 2^31 vertices");
    CHECK_ARG(a->num_follow_strands_per_guide == 0 || a->follow_root_offsets,
              "tfx_hair_create: follow strands need follow_root_offsets");
    // Bone indices used by the skinning blend must address the 512-matrix table.
    for (long long s = 0; s < S; s += (a->num_follow_strands_per_guide + 1)) {
        const tfx_bone_skinning_data& sd = a->bone_skinning_data[s];
        for (int k = 0; k < 4; ++k) {
            if (k > 0 && !(sd.bone_weight[k] > 0.0f)) continue;
            float fi = sd.bone_index[k];
            if (!(fi >= 0.0f && fi < (float)TFX_MAX_NUM_BONES))
                return fail(TFX_ERR_INVALID_ARGUMENT, "tfx_hair_create: strand %lld bone index %g outside [0, 512)", s,
                            (double)fi);
        }
    }
    tfx_hair* h = new tfx_hair();
    h->G = a->num_guide_strands;
    h->F = a->num_follow_strands_per_guide;
    h->V = V;
    h->S = (int)S;
    h->N = (int)N;
    h->device = g_device;
    set_identity_bones(h->bones_h);
    h->bones_dirty = TFX_MAX_NUM_BONES;
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = dalloc(&h->pos, N);
    if (e == hipSuccess) e = dalloc(&h->prev, N);
    if (e == hipSuccess) e = dalloc(&h->prevprev, N);
    if (e == hipSuccess) e = dalloc(&h->tangent, N);
    if (e == hipSuccess) e = dalloc(&h->sld, S);
    if (e == hipSuccess) e = dalloc(&h->rest, N);
    if (e == hipSuccess) e = dalloc(&h->restlen, N);
    if (e == hipSuccess) e = dalloc(&h->follow_off, S);
    if (e == hipSuccess) e = dalloc(&h->skin, S);
    if (e == hipSuccess) e = dalloc(&h->bones_d, 16 * TFX_MAX_NUM_BONES);
    if (e == hipSuccess) e = dalloc(&h->params_d, 1);
    if (e != hipSuccess) {
        hair_free(h);
        delete h;
        return fail(e == hipErrorOutOfMemory ? TFX_ERR_OUT_OF_MEMORY : TFX_ERR_HIP, "tfx_hair_create: %s",
                    hipGetErrorString(e));
    }
    // TressFXDynamicState::UploadGPUData: Pos = Prev = PrevPrev = positions.
    const size_t v4 = (size_t)N * sizeof(float4);
    std::vector<float> zeros;
    HIP_TRY(hipMemcpy(h->pos, a->positions, v4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(h->prev, a->positions, v4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(h->prevprev, a->positions, v4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(h->rest, a->positions, v4, hipMemcpyHostToDevice));
    if (a->tangents) HIP_TRY(hipMemcpy(h->tangent, a->tangents, v4, hipMemcpyHostToDevice));
    else HIP_TRY(hipMemset(h->tangent, 0, v4));
    HIP_TRY(hipMemcpy(h->restlen, a->rest_lengths, (size_t)N * sizeof(float), hipMemcpyHostToDevice));
    if (a->follow_root_offsets)
        HIP_TRY(hipMemcpy(h->follow_off, a->follow_root_offsets, (size_t)S * sizeof(float4), hipMemcpyHostToDevice));
    else
        HIP_TRY(hipMemset(h->follow_off, 0, (size_t)S * sizeof(float4)));
    HIP_TRY(hipMemcpy(h->skin, a->bone_skinning_data, (size_t)S * sizeof(tfx_bone_skinning_data),
                      hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(h->sld, 0, (size_t)S * sizeof(tfx_strand_level_data)));
    HIP_TRY(hipMemcpy(h->bones_d, h->bones_h.data(), h->bones_h.size() * sizeof(float), hipMemcpyHostToDevice));
    h->bones_dirty = 0;
    {
        const std::vector<int> order = guide_order(a->positions, h->G, h->F, V);
        if (!order.empty()) {
            hipError_t oe = dalloc(&h->order, order.size());
            if (oe == hipSuccess) oe = hipMemcpy(h->order, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice);
            if (oe != hipSuccess) {
                hair_free(h);
                delete h;
                return fail(oe == hipErrorOutOfMemory ? TFX_ERR_OUT_OF_MEMORY : TFX_ERR_HIP, "tfx_hair_create: %s",
                            hipGetErrorString(oe));
            }
        }
    }
    compute_tangent_w(h, a->tangents);
    compute_follow_w(h, a->positions);
    // The matrices the kernels can read (guide strands' skinning records).
    h->bones_used = skin_bones_used(a->bone_skinning_data, S, h->F + 1);
    // Constant block before the first UpdateSimulationParameters: zero, with
    // the counts of the object and the reset flag of the constructor (:157).
    std::memset(&h->params, 0, sizeof(h->params));
    h->params.counts[0] = TFX_SIM_THREAD_GROUP_SIZE / V;
    h->params.counts[1] = h->F;
    h->params.counts[2] = V;
    h->params.cpu_local_shape_iterations = 0;
    h->params.reset_positions = 1.0f;
    h->params.clamp_position_delta = 0.0f;
    HIP_TRY(hipDeviceSynchronize());
    *out = h;
    return TFX_OK;
}

extern "C" int tfx_hair_destroy(tfx_hair* h) {
    if (!h) return TFX_OK;
    (void)hipDeviceSynchronize();
    hair_free(h);
    delete h;
    return TFX_OK;
}

extern "C" int tfx_hair_update_bone_matrices(tfx_hair* h, const float* m, int n) {
    CHECK_ARG(h && (m || n == 0) && n >= 0, "tfx_hair_update_bone_matrices: bad argument");
    n = std::min(n, TFX_MAX_NUM_BONES);
    std::memcpy(h->bones_h.data(), m, (size_t)n * 16 * sizeof(float));
    h->bones_dirty = std::max(h->bones_dirty, n);
    h->params.num_bones = std::max(h->params.num_bones, n);
    return TFX_OK;
}

extern "C" int tfx_hair_update_simulation_parameters(tfx_hair* h, const tfx_simulation_settings* s, float dt) {
    CHECK_ARG(h && s, "tfx_hair_update_simulation_parameters: null argument");
    tfx_simulation_params& cb = h->params;
    cb.vsp[0] = s->vsp_coeff;
    cb.vsp[1] = s->vsp_accel_threshold;
    cb.shape[0] = s->damping;
    cb.shape[1] = s->local_constraint_stiffness;
    cb.shape[2] = s->global_constraint_stiffness;
    cb.shape[3] = s->global_constraints_range;
    cb.grav_time_tip[0] = s->gravity_magnitude;
    cb.grav_time_tip[1] = dt;
    cb.sim_ints[2] = 0;
    cb.counts[0] = TFX_SIM_THREAD_GROUP_SIZE / h->V;
    cb.counts[2] = h->V;
    cb.counts[1] = h->F;
    cb.grav_time_tip[2] = s->tip_separation;
    cb.clamp_position_delta = 20.0f;  // TressFXHairObject.cpp:397 (settings value ignored)
    if (h->V >= TFX_MIN_VERTS_PER_STRAND_FOR_GPU_ITERATION) {
        cb.sim_ints[1] = s->local_constraints_iterations;
        cb.cpu_local_shape_iterations = 1;
    } else {
        cb.sim_ints[1] = 1;
        cb.cpu_local_shape_iterations = s->local_constraints_iterations;
    }
    cb.sim_ints[0] = s->length_constraints_iterations;
    set_wind(cb, s->wind_direction, s->wind_magnitude, h->frame);
    if (h->frame < 2) cb.reset_positions = 1.0f;
    cb.simulation_frame = h->frame;
    return TFX_OK;
}

extern "C" int tfx_hair_set_collision_capsules(tfx_hair* h, const float* c0, const float* c1, int n) {
    CHECK_ARG(h && n >= 0 && n <= TFX_MAX_NUM_COLLISION_CAPSULES && (n == 0 || (c0 && c1)),
              "tfx_hair_set_collision_capsules: bad argument (at most 8 capsules)");
    std::memset(h->params.center_and_radius0, 0, sizeof(h->params.center_and_radius0));
    std::memset(h->params.center_and_radius1, 0, sizeof(h->params.center_and_radius1));
    for (int c = 0; c < n; ++c)
        for (int k = 0; k < 4; ++k) {
            h->params.center_and_radius0[c][k] = c0[4 * c + k];
            h->params.center_and_radius1[c][k] = c1[4 * c + k];
        }
    h->params.num_collision_capsules[0] = n;
    return TFX_OK;
}

extern "C" int tfx_hair_reset_positions(tfx_hair* h) {
    CHECK_ARG(h, "tfx_hair_reset_positions: null hair");
    h->params.reset_positions = 1.0f;
    return TFX_OK;
}

extern "C" int tfx_hair_increase_simulation_frame(tfx_hair* h) {
    CHECK_ARG(h, "tfx_hair_increase_simulation_frame: null hair");
    h->params.reset_positions = 0.0f;
    h->frame++;
    return TFX_OK;
}

extern "C" int tfx_hair_get_info(const tfx_hair* h, tfx_hair_info* o) {
    CHECK_ARG(h && o, "tfx_hair_get_info: null argument");
    o->num_guide_strands = h->G;
    o->num_follow_strands_per_guide = h->F;
    o->num_vertices_per_strand = h->V;
    o->num_total_strands = h->S;
    o->num_total_vertices = h->N;
    o->simulation_frame = h->frame;
    o->cpu_local_shape_iterations = h->params.cpu_local_shape_iterations;
    o->pad = 0;
    return TFX_OK;
}

extern "C" int tfx_hair_get_simulation_params(const tfx_hair* h, tfx_simulation_params* o, float* bones) {
    CHECK_ARG(h && o, "tfx_hair_get_simulation_params: null argument");
    *o = h->params;
    if (bones) std::memcpy(bones, h->bones_h.data(), h->bones_h.size() * sizeof(float));
    return TFX_OK;
}

extern "C" int tfx_hair_get_dynamic_state(tfx_hair* h, tfx_dynamic_state* o) {
    CHECK_ARG(h && o, "tfx_hair_get_dynamic_state: null argument");
    o->positions = (float*)h->pos;
    o->positions_prev = (float*)h->prev;
    o->positions_prev_prev = (float*)h->prevprev;
    o->tangents = (float*)h->tangent;
    o->strand_level_data = h->sld;
    return TFX_OK;
}

extern "C" int tfx_hair_download_state(tfx_hair* h, void* stream, float* pos, float* prev, float* pp, float* tan,
                                       float* sld) {
    CHECK_ARG(h, "tfx_hair_download_state: null hair");
    hipStream_t s = (hipStream_t)stream;
    const size_t v4 = (size_t)h->N * sizeof(float4);
    if (pos) HIP_TRY(hipMemcpyAsync(pos, h->pos, v4, hipMemcpyDeviceToHost, s));
    if (prev) HIP_TRY(hipMemcpyAsync(prev, h->prev, v4, hipMemcpyDeviceToHost, s));
    if (pp) HIP_TRY(hipMemcpyAsync(pp, h->prevprev, v4, hipMemcpyDeviceToHost, s));
    if (tan) HIP_TRY(hipMemcpyAsync(tan, h->tangent, v4, hipMemcpyDeviceToHost, s));
    if (sld) HIP_TRY(hipMemcpyAsync(sld, h->sld, (size_t)h->S * sizeof(tfx_strand_level_data), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return TFX_OK;
}

extern "C" int tfx_hair_upload_state(tfx_hair* h, void* stream, const float* pos, const float* prev, const float* pp,
                                     const float* tan, const float* sld) {
    CHECK_ARG(h, "tfx_hair_upload_state: null hair");
    hipStream_t s = (hipStream_t)stream;
    const size_t v4 = (size_t)h->N * sizeof(float4);
    if (pos) HIP_TRY(hipMemcpyAsync(h->pos, pos, v4, hipMemcpyHostToDevice, s));
    if (prev) HIP_TRY(hipMemcpyAsync(h->prev, prev, v4, hipMemcpyHostToDevice, s));
    if (pp) HIP_TRY(hipMemcpyAsync(h->prevprev, pp, v4, hipMemcpyHostToDevice, s));
    if (tan) HIP_TRY(hipMemcpyAsync(h->tangent, tan, v4, hipMemcpyHostToDevice, s));
    if (sld) HIP_TRY(hipMemcpyAsync(h->sld, sld, (size_t)h->S * sizeof(tfx_strand_level_data), hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (tan) compute_tangent_w(h, tan);
    if (pos) compute_follow_w(h, pos);
    return TFX_OK;
}

// ---------------------------------------------------------------------------
// TressFXSimulation
// ---------------------------------------------------------------------------
static const char* kStageNames[6] = {"IntegrationAndGlobalShapeConstraints", "CalculateStrandLevelData",
                                     "VelocityShockPropagation", "LocalShapeConstraints",
                                     "LengthConstriantsWindAndCollision", "UpdateFollowHairVertices"};

extern "C" int tfx_simulation_create(tfx_simulation** out) {
    CHECK_ARG(out, "tfx_simulation_create: null argument");
    tfx_simulation* s = new tfx_simulation();
    s->device = g_device;
    for (int v = 0; v <= 64; ++v) s->variant[v] = s->per_cu[v] = -1;
    *out = s;
    return TFX_OK;
}

extern "C" int tfx_simulation_destroy(tfx_simulation* s) {
    if (!s) return TFX_OK;
    (void)hipDeviceSynchronize();
    (void)hipFree(s->table_d);
    (void)hipFree(s->queue_d);
    delete s;
    return TFX_OK;
}

extern "C" int tfx_simulation_run_stage(tfx_simulation* sim, void* stream, tfx_hair* h, int stage) {
    CHECK_ARG(sim && h && stage >= 0 && stage <= 5, "tfx_simulation_run_stage: bad argument");
    hipStream_t s = (hipStream_t)stream;
    int rc = upload_params(h, nullptr, s);
    if (rc) return rc;
    DevHair d = dev_hair(h, nullptr);
    PassTimer tm(kStageNames[stage], s);
    HIP_TRY(kl_launch_stage(stage, &d, s));
    return TFX_OK;
}

// TressFXSimulation::Simulate, TressFXSimulation.cpp:77-112.
extern "C" int tfx_simulation_simulate(tfx_simulation* sim, void* stream, tfx_hair* const* hairs, int n) {
    CHECK_ARG(sim && (hairs || n == 0) && n >= 0, "tfx_simulation_simulate: bad argument");
    hipStream_t s = (hipStream_t)stream;
    for (int i = 0; i < n; ++i) {
        CHECK_ARG(hairs[i], "tfx_simulation_simulate: null hair");
        int rc = upload_params(hairs[i], nullptr, s);
        if (rc) return rc;
    }
    for (int stage = 0; stage < 6; ++stage) {
        for (int i = 0; i < n; ++i) {
            DevHair d = dev_hair(hairs[i], nullptr);
            const int reps = (stage == 3) ? hairs[i]->params.cpu_local_shape_iterations : 1;
            for (int r = 0; r < reps; ++r) {
                PassTimer tm(kStageNames[stage], s);
                HIP_TRY(kl_launch_stage(stage, &d, s));
            }
        }
    }
    for (int i = 0; i < n; ++i) tfx_hair_increase_simulation_frame(hairs[i]);
    return TFX_OK;
}

extern "C" int tfx_simulation_set_sdf_deferral(tfx_simulation* sim, int mode) {
    CHECK_ARG(sim, "tfx_simulation_set_sdf_deferral: null simulation");
    CHECK_ARG(mode >= TFX_SDF_DEFER_AUTO && mode <= TFX_SDF_DEFER_ALWAYS,
              "tfx_simulation_set_sdf_deferral: mode must be TFX_SDF_DEFER_AUTO, _NEVER or _ALWAYS");
    sim->defer_sdf = mode;
    return TFX_OK;
}

extern "C" int tfx_simulation_set_launch_shape(tfx_simulation* sim, int V, int variant, int blocks_per_cu) {
    CHECK_ARG(sim && kl_fused_block_vertices(V, 0) > 0, "tfx_simulation_set_launch_shape: bad vertices per strand");
    CHECK_ARG(variant >= -1 && variant <= 2, "tfx_simulation_set_launch_shape: variant must be -1 (automatic), 0, 1 or 2");
    CHECK_ARG(blocks_per_cu >= 0 && blocks_per_cu <= 64, "tfx_simulation_set_launch_shape: blocks_per_cu out of range");
    sim->variant[V] = variant;
    sim->per_cu[V] = blocks_per_cu;
    return TFX_OK;
}

static int ensure_table(tfx_simulation* sim, size_t bytes) {
    if (!sim->queue_d) {
        HIP_TRY(hipMalloc((void**)&sim->queue_d, tfx::kFusedQueueWords * sizeof(uint32_t)));
        HIP_TRY(hipMemset(sim->queue_d, 0, tfx::kFusedQueueWords * sizeof(uint32_t)));
        HIP_TRY(hipDeviceSynchronize());
    }
    if (bytes <= sim->table_bytes) return TFX_OK;
    if (sim->table_d) { HIP_TRY(hipDeviceSynchronize()); HIP_TRY(hipFree(sim->table_d)); sim->table_d = nullptr; }
    size_t cap = std::max<size_t>(bytes * 2, 1 << 16);
    HIP_TRY(hipMalloc(&sim->table_d, cap));
    sim->table_bytes = cap;
    return TFX_OK;
}


// One fused substep for a set of objects: per V class, one persistent launch
// over the (object, strand block) work list.
extern "C" int tfx_simulation_simulate_fused(tfx_simulation* sim, void* stream, tfx_hair* const* hairs,
                                             tfx_sdf* const* sdfs, int n) {
    CHECK_ARG(sim && (hairs || n == 0) && n >= 0, "tfx_simulation_simulate_fused: bad argument");
    hipStream_t s = (hipStream_t)stream;
    for (int i = 0; i < n; ++i) {
        CHECK_ARG(hairs[i], "tfx_simulation_simulate_fused: null hair");
        if (sdfs && sdfs[i] && sdfs[i]->collide_variant != TFX_SDF_COLLIDE_FORWARD)
            return fail(TFX_ERR_UNSUPPORTED, "tfx_simulation_simulate_fused: the fused substep collides with "
                                             "forward differences only (SDF %d has the mixed variant)", i);
        int rc = upload_params(hairs[i], sdfs ? sdfs[i] : nullptr, s);
        if (rc) return rc;
    }
    // Table layout: DevHair[n] | WorkItem[total]
    std::vector<tfx::WorkItem> items;
    std::vector<int> vclass_begin(65, 0), vclass_count(65, 0), vclass_variant(65, 0), vclass_resident(65, 0);
    const int cus = cu_count(sim->device);
    for (int V = 4; V <= 64; V *= 2) {
        vclass_begin[V] = (int)items.size();
        long long guides = 0;
        for (int i = 0; i < n; ++i)
            if (hairs[i]->V == V) guides += hairs[i]->G;
        if (!guides) continue;
        const Shape sh = choose_shape(sim, V, hairs, n, cus);
        vclass_variant[V] = sh.variant;
        vclass_resident[V] = sh.resident;
        const int B = kl_fused_block_vertices(V, sh.variant) / V;
        for (int i = 0; i < n; ++i)
            if (hairs[i]->V == V)
                for (int g0 = 0; g0 < hairs[i]->G; g0 += B) items.push_back(tfx::WorkItem{i, g0});
        vclass_count[V] = (int)items.size() - vclass_begin[V];
    }
    // SDF collision as a collide pass after the fused launches for the strand
    // lengths whose blocks fit in one round of resident blocks (as the
    // batched form, tfx_simulation_set_sdf_deferral)
    std::vector<int> defer(65, 0);
    for (int V = 4; V <= 64; V *= 2) {
        const int cnt = vclass_count[V], mode = sim->defer_sdf;
        defer[V] = cnt > 0 && (mode == TFX_SDF_DEFER_ALWAYS || (mode == TFX_SDF_DEFER_AUTO && cnt <= vclass_resident[V]));
    }
    const int cbegin = (int)items.size();
    const int chunk = kl_collide_chunk();
    for (int i = 0; i < n; ++i)
        if (defer[hairs[i]->V] && sdfs && sdfs[i])
            for (int v0 = 0; v0 < hairs[i]->N; v0 += chunk) items.push_back(tfx::WorkItem{i, v0});
    const int ccount = (int)items.size() - cbegin;
    const size_t hair_bytes = (size_t)n * sizeof(DevHair);
    const size_t item_off = (hair_bytes + 255) & ~size_t(255);
    const size_t bytes = item_off + items.size() * sizeof(tfx::WorkItem);
    int rc = ensure_table(sim, bytes);
    if (rc) return rc;
    if ((rc = sim->staging.ensure(bytes))) return rc;
    void* host;
    int slot;
    if ((rc = sim->staging.acquire(&host, &slot))) return rc;
    DevHair* dh = (DevHair*)host;
    for (int i = 0; i < n; ++i) dh[i] = dev_hair(hairs[i], sdfs ? sdfs[i] : nullptr);
    std::memcpy((char*)host + item_off, items.data(), items.size() * sizeof(tfx::WorkItem));
    HIP_TRY(hipMemcpyAsync(sim->table_d, host, bytes, hipMemcpyHostToDevice, s));
    if ((rc = sim->staging.release_slot(slot, s))) return rc;
    const DevHair* dtab = (const DevHair*)sim->table_d;
    const tfx::WorkItem* ditems = (const tfx::WorkItem*)((char*)sim->table_d + item_off);
    for (int V = 4; V <= 64; V *= 2) {
        const int cnt = vclass_count[V];
        if (!cnt) continue;
        const int grid = std::min(cnt, vclass_resident[V]);
        PassTimer tm(fused_label(V, kl_fused_block_vertices(V, vclass_variant[V])), s, true);
        HIP_TRY(tm.launched(kl_launch_fused(V, vclass_variant[V], dtab, ditems + vclass_begin[V], cnt, grid,
                                            defer[V], sim->queue_d, s, tm.start(), tm.stop())));
    }
    if (ccount) {
        PassTimer tm("CollideHairVerticesWithSdf", s, true);
        HIP_TRY(tm.launched(kl_launch_collide_items(dtab, ditems + cbegin, ccount, s, tm.start(), tm.stop())));
    }
    for (int i = 0; i < n; ++i) tfx_hair_increase_simulation_frame(hairs[i]);
    return TFX_OK;
}

// ---------------------------------------------------------------------------
// The fused substep without allocations (tfx_simulation_simulate_fused_noalloc)
// ---------------------------------------------------------------------------
// Its tables live in `scratch`, one array per GPU the library runs on, held
// until the process ends: per call at most kNaMaxHairs objects, kNaMaxBones
// bone matrices (the matrices each object's skinning can read) and
// kNaMaxItems strand blocks.  The tables are written by staging launches
// whose records travel as kernel arguments, so no host buffer has to outlive
// the call and nothing is allocated on the host or the device.
namespace {
constexpr size_t na_align(size_t x) { return (x + 255) & ~size_t(255); }
constexpr int kNaMaxHairs = 64;
constexpr int kNaMaxBones = 2048;
constexpr int kNaMaxItems = 65536;
constexpr int kNaStageBones = 24;  // matrices per staging launch (kernel arguments stay under 4 KiB)
constexpr size_t kNaParamsOff = 0;
constexpr size_t kNaHairsOff = na_align(kNaMaxHairs * sizeof(DevParams));
constexpr size_t kNaBonesOff = kNaHairsOff + na_align(kNaMaxHairs * sizeof(DevHair));
constexpr size_t kNaItemsOff = kNaBonesOff + na_align((size_t)kNaMaxBones * 16 * sizeof(float));
constexpr size_t kNaQueueOff = kNaItemsOff + na_align((size_t)kNaMaxItems * sizeof(tfx::WorkItem));
constexpr size_t kNaScratchBytes = kNaQueueOff + tfx::kFusedQueueWords * sizeof(uint32_t);

// One staging launch: the object's parameter block and table row (when
// nitems > 0, its first launch), a chunk of its bone matrices, its work items.
struct NaStage {
    DevParams params;
    DevHair hair;
    int slot, bone_first, nbones, item_first, nitems, block_strands, pad0, pad1;
    float bones[kNaStageBones * 16];
};
static_assert(sizeof(NaStage) <= 4000, "staging record must fit the kernel-argument space");
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNaScratchBytes];

__global__ void k_noalloc_stage(unsigned char* base, NaStage st) {
    DevParams* params = reinterpret_cast<DevParams*>(base + kNaParamsOff);
    DevHair* hairs = reinterpret_cast<DevHair*>(base + kNaHairsOff);
    tfx::WorkItem* items = reinterpret_cast<tfx::WorkItem*>(base + kNaItemsOff);
    float* bones = const_cast<float*>(st.hair.bones);
    const int t = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
    if (st.nitems > 0 && t == 0) {
        params[st.slot] = st.params;
        hairs[st.slot] = st.hair;
    }
    // the fused launches' work queues start at zero (every completed launch
    // also leaves them there)
    if (st.slot == 0 && st.bone_first == 0 && t < tfx::kFusedQueueWords)
        reinterpret_cast<uint32_t*>(base + kNaQueueOff)[t] = 0u;
    for (int k = t; k < st.nbones * 16; k += stride) bones[st.bone_first * 16 + k] = st.bones[k];
    for (int k = t; k < st.nitems; k += stride) items[st.item_first + k] = tfx::WorkItem{st.slot, k * st.block_strands};
}

// The device address of `scratch` on each device (the array's name in host
// code is only a stand-in), looked up with the device current.
static std::mutex g_na_mu;
static void* g_na_scratch[64] = {nullptr};

static int noalloc_scratch(int device, void** out) {
    CHECK_ARG(device >= 0 && device < 64, "no-allocation path: device ordinal out of range");
    std::lock_guard<std::mutex> g(g_na_mu);
    if (!g_na_scratch[device]) {
        int cur = 0;
        HIP_TRY(hipGetDevice(&cur));
        CHECK_ARG(cur == device, "no-allocation path: the object's device is not current (call tfx_init)");
        void* ptr = nullptr;
        HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        g_na_scratch[device] = ptr;
    }
    *out = g_na_scratch[device];
    return TFX_OK;
}

extern "C" int tfx_simulation_simulate_fused_noalloc(tfx_simulation* sim, void* stream, tfx_hair* const* hairs,
                                                     tfx_sdf* const* sdfs, int n) {
    CHECK_ARG(sim && (hairs || n == 0) && n >= 0, "tfx_simulation_simulate_fused_noalloc: bad argument");
    if (n > kNaMaxHairs)
        return fail(TFX_ERR_INVALID_ARGUMENT, "tfx_simulation_simulate_fused_noalloc: %d hair objects (at most %d)", n,
                    kNaMaxHairs);
    if (n == 0) return TFX_OK;
    hipStream_t s = (hipStream_t)stream;
    int bones_total = 0;
    for (int i = 0; i < n; ++i) {
        CHECK_ARG(hairs[i], "tfx_simulation_simulate_fused_noalloc: null hair");
        CHECK_ARG(hairs[i]->device == sim->device, "tfx_simulation_simulate_fused_noalloc: objects on another device");
        if (sdfs && sdfs[i] && sdfs[i]->collide_variant != TFX_SDF_COLLIDE_FORWARD)
            return fail(TFX_ERR_UNSUPPORTED, "tfx_simulation_simulate_fused_noalloc: the fused substep collides with "
                                             "forward differences only (SDF %d has the mixed variant)", i);
        bones_total += hairs[i]->bones_used;
    }
    if (bones_total > kNaMaxBones)
        return fail(TFX_ERR_INVALID_ARGUMENT, "tfx_simulation_simulate_fused_noalloc: %d bone matrices (at most %d)",
                    bones_total, kNaMaxBones);
    void* base = nullptr;
    int rc = noalloc_scratch(sim->device, &base);
    if (rc) return rc;
    unsigned char* b = static_cast<unsigned char*>(base);
    // strand blocks per strand-length class, contiguous per class
    int cls_begin[65] = {0}, cls_count[65] = {0}, cls_variant[65] = {0}, cls_resident[65] = {0}, cls_B[65] = {0};
    int item_first[kNaMaxHairs], nitems[kNaMaxHairs];
    const int cus = cu_count(sim->device);
    int total = 0;
    for (int V = 4; V <= 64; V *= 2) {
        cls_begin[V] = total;
        bool any = false;
        for (int i = 0; i < n; ++i) any = any || hairs[i]->V == V;
        if (!any) continue;
        const Shape sh = choose_shape(sim, V, hairs, n, cus);
        cls_variant[V] = sh.variant;
        cls_resident[V] = sh.resident;
        cls_B[V] = kl_fused_block_vertices(V, sh.variant) / V;
        for (int i = 0; i < n; ++i)
            if (hairs[i]->V == V) {
                item_first[i] = total;
                nitems[i] = (hairs[i]->G + cls_B[V] - 1) / cls_B[V];
                total += nitems[i];
            }
        cls_count[V] = total - cls_begin[V];
    }
    if (total > kNaMaxItems)
        return fail(TFX_ERR_INVALID_ARGUMENT, "tfx_simulation_simulate_fused_noalloc: %d strand blocks (at most %d)",
                    total, kNaMaxItems);
    // staging launches: per object its records, bone matrices and work items
    int bone_at = 0;
    for (int i = 0; i < n; ++i) {
        tfx_hair* h = hairs[i];
        tfx_sdf* sdf = sdfs ? sdfs[i] : nullptr;
        NaStage st;
        std::memset(&st, 0, sizeof(st));
        pack_dev_params(h, sdf, st.params);
        st.hair = dev_hair(h, sdf);
        st.hair.params = reinterpret_cast<const DevParams*>(b + kNaParamsOff) + i;
        st.hair.bones = reinterpret_cast<const float*>(b + kNaBonesOff) + (size_t)16 * bone_at;
        st.slot = i;
        st.block_strands = cls_B[h->V];
        for (int first = 0; first < h->bones_used || first == 0; first += kNaStageBones) {
            st.bone_first = first;
            st.nbones = std::min(kNaStageBones, h->bones_used - first);
            std::memcpy(st.bones, h->bones_h.data() + (size_t)16 * first, (size_t)16 * st.nbones * sizeof(float));
            st.item_first = item_first[i];
            st.nitems = first == 0 ? nitems[i] : 0;
            const int work = std::max(st.nitems, st.nbones * 16);
            const int blocks = std::max(1, std::min((work + 255) / 256, 64));
            hipLaunchKernelGGL(k_noalloc_stage, dim3(blocks), dim3(256), 0, s, b, st);
            HIP_TRY(hipGetLastError());
        }
        bone_at += h->bones_used;
    }
    // the fused launches, the SDF collision inside them
    const DevHair* dtab = reinterpret_cast<const DevHair*>(b + kNaHairsOff);
    const tfx::WorkItem* ditems = reinterpret_cast<const tfx::WorkItem*>(b + kNaItemsOff);
    for (int V = 4; V <= 64; V *= 2) {
        const int cnt = cls_count[V];
        if (!cnt) continue;
        HIP_TRY(kl_launch_fused(V, cls_variant[V], dtab, ditems + cls_begin[V], cnt, std::min(cnt, cls_resident[V]), 0,
                                reinterpret_cast<uint32_t*>(b + kNaQueueOff), s, nullptr, nullptr));
    }
    for (int i = 0; i < n; ++i) tfx_hair_increase_simulation_frame(hairs[i]);
    return TFX_OK;
}

// ---------------------------------------------------------------------------
// TressFXSDFCollisionSystem / TressFXSDFCollision
// ---------------------------------------------------------------------------
extern "C" int tfx_sdf_collision_system_create(tfx_sdf_collision_system** out) {
    CHECK_ARG(out, "tfx_sdf_collision_system_create: null argument");
    *out = new tfx_sdf_collision_system();
    (*out)->device = g_device;
    return TFX_OK;
}

extern "C" int tfx_sdf_collision_system_destroy(tfx_sdf_collision_system* s) {
    delete s;
    return TFX_OK;
}

static int sdf_alloc(tfx_sdf* s) {
    const long long cells = (long long)s->nx * s->ny * s->nz;
    CHECK_ARG(s->nx > 0 && s->ny > 0 && s->nz > 0 && cells < (1ll << 31), "tfx_sdf: grid dimensions out of range");
    CHECK_ARG(s->total_cells >= cells, "tfx_sdf: allocation smaller than the grid");
    hipError_t e = hipMalloc((void**)&s->cells, (size_t)s->total_cells * sizeof(uint32_t));
    if (e != hipSuccess)
        return fail(e == hipErrorOutOfMemory ? TFX_ERR_OUT_OF_MEMORY : TFX_ERR_HIP, "tfx_sdf: %s", hipGetErrorString(e));
    HIP_TRY(hipMemset(s->cells, 0, (size_t)s->total_cells * sizeof(uint32_t)));
    s->bx = (s->nx + (1 << tfx::kBrickShift) - 1) >> tfx::kBrickShift;
    s->by = (s->ny + (1 << tfx::kBrickShift) - 1) >> tfx::kBrickShift;
    s->bz = (s->nz + (1 << tfx::kBrickShift) - 1) >> tfx::kBrickShift;
    e = hipMalloc((void**)&s->bricks, (size_t)s->bx * s->by * s->bz);
    if (e != hipSuccess)
        return fail(e == hipErrorOutOfMemory ? TFX_ERR_OUT_OF_MEMORY : TFX_ERR_HIP, "tfx_sdf: %s", hipGetErrorString(e));
    const size_t nbrick = (size_t)s->bx * s->by * s->bz;
    e = hipMalloc((void**)&s->brick_bits, ((nbrick + 31) / 32) * sizeof(uint32_t));
    if (e != hipSuccess)
        return fail(e == hipErrorOutOfMemory ? TFX_ERR_OUT_OF_MEMORY : TFX_ERR_HIP, "tfx_sdf: %s", hipGetErrorString(e));
    e = hipMalloc((void**)&s->band, 2 * nbrick);
    if (e != hipSuccess)
        return fail(e == hipErrorOutOfMemory ? TFX_ERR_OUT_OF_MEMORY : TFX_ERR_HIP, "tfx_sdf: %s", hipGetErrorString(e));
    HIP_TRY(hipMemset(s->band, 0, 2 * nbrick));
    // Until the first Update every brick is "maybe occupied": lookups stay exact.
    HIP_TRY(hipMemset(s->bricks, 1, nbrick));
    HIP_TRY(hipMemset(s->brick_bits, 0xff, ((nbrick + 31) / 32) * sizeof(uint32_t)));
    return TFX_OK;
}

// TressFXSDFCollision constructor, TressFXSDFCollision.cpp:38-86.
extern "C" int tfx_sdf_create(const tfx_sdf_mesh_desc* mesh, int numCellsInX, float margin, tfx_sdf** out) {
    CHECK_ARG(mesh && out && numCellsInX > 0, "tfx_sdf_create: bad argument");
    *out = nullptr;
    tfx_sdf* s = new tfx_sdf();
    s->mesh = *mesh;
    s->margin = margin;
    s->device = g_device;
    const float* bmin0 = mesh->initial_bbox_min;
    const float* bmax0 = mesh->initial_bbox_max;
    s->cell = (bmax0[0] - bmin0[0]) / numCellsInX;
    const int extra = (int)(0.8f * (float)numCellsInX);
    // m_PaddingBoundary = a, b, c; — a comma expression: every axis gets `a`.
    s->pad_boundary = extra * s->cell;
    float bmin[3], bmax[3];
    for (int k = 0; k < 3; ++k) {
        s->origin[k] = bmin0[k] - s->pad_boundary;  // UpdateSDFGrid(initial bbox)
        bmin[k] = bmin0[k] - s->pad_boundary;
        bmax[k] = bmax0[k] + s->pad_boundary;
    }
    s->nx = (int)((bmax[0] - bmin[0]) / s->cell);
    s->ny = (int)((bmax[1] - bmin[1]) / s->cell);
    s->nz = (int)((bmax[2] - bmin[2]) / s->cell);
    const float alloc = 1.4f * s->nx * s->ny * s->nz;
    s->total_cells = (int)std::min<double>((double)alloc, 2147483647.0);
    int rc = sdf_alloc(s);
    if (rc) { delete s; return rc; }
    *out = s;
    return TFX_OK;
}

extern "C" int tfx_sdf_create_explicit(const tfx_sdf_mesh_desc* mesh, const float origin[3], float cell, int nx,
                                       int ny, int nz, float margin, tfx_sdf** out) {
    CHECK_ARG(mesh && origin && out && cell > 0.0f, "tfx_sdf_create_explicit: bad argument");
    *out = nullptr;
    tfx_sdf* s = new tfx_sdf();
    s->mesh = *mesh;
    s->margin = margin;
    s->device = g_device;
    s->explicit_grid = 1;
    s->cell = cell;
    for (int k = 0; k < 3; ++k) s->origin[k] = origin[k];
    s->nx = nx; s->ny = ny; s->nz = nz;
    s->total_cells = (int)((long long)nx * ny * nz);
    int rc = sdf_alloc(s);
    if (rc) { delete s; return rc; }
    *out = s;
    return TFX_OK;
}

extern "C" int tfx_sdf_destroy(tfx_sdf* s) {
    if (!s) return TFX_OK;
    (void)hipDeviceSynchronize();
    (void)hipFree(s->cells);
    (void)hipFree(s->bricks);
    (void)hipFree(s->brick_bits);
    (void)hipFree(s->band);
    delete s;
    return TFX_OK;
}

extern "C" int tfx_sdf_set_mesh(tfx_sdf* s, const tfx_sdf_mesh_desc* mesh) {
    CHECK_ARG(s && mesh, "tfx_sdf_set_mesh: null argument");
    s->mesh = *mesh;
    return TFX_OK;
}

// TressFXSDFCollision::Update, TressFXSDFCollision.cpp:95-169.
extern "C" int tfx_sdf_update(tfx_sdf* s, tfx_sdf_collision_system* sys, void* stream) {
    CHECK_ARG(s && sys, "tfx_sdf_update: null argument");
    if (!s->mesh.vertices || !s->mesh.indices) return TFX_OK;  // no input mesh: nothing to do (:100)
    if (!s->explicit_grid)
        for (int k = 0; k < 3; ++k) s->origin[k] = s->mesh.bbox_min[k] - s->pad_boundary;  // UpdateSDFGrid
    const int ncell = s->nx * s->ny * s->nz;
    s->band_valid = 0;  // the whole grid is rewritten: a batched build's band history no longer applies
    static const char* kNames[4] = {"InitializeSignedDistanceField", "ConstructSignedDistanceField",
                                    "FinalizeSignedDistanceField", "FinalizeSignedDistanceField"};
    for (int phase = 0; phase < 4; ++phase) {
        PassTimer tm(kNames[phase], (hipStream_t)stream);
        HIP_TRY(kl_launch_sdf_phase(phase, s->cells, ncell, s->mesh.vertices, s->mesh.indices, s->mesh.num_triangles,
                                     sdf_grid(s), (hipStream_t)stream));
    }
    return TFX_OK;
}

// TressFXSDFCollision::CollideWithHair, TressFXSDFCollision.cpp:171-221.
extern "C" int tfx_sdf_collide_with_hair(tfx_sdf* s, tfx_sdf_collision_system* sys, void* stream, tfx_hair* h) {
    CHECK_ARG(s && sys && h, "tfx_sdf_collide_with_hair: null argument");
    if (!s->mesh.vertices || !s->mesh.indices) return TFX_OK;
    PassTimer tm("CollideHairVerticesWithSdf", (hipStream_t)stream);
    const float* cells = reinterpret_cast<const float*>(s->cells);
    if (s->collide_variant == TFX_SDF_COLLIDE_MIXED)
        HIP_TRY(kl_launch_sdf_collide_mixed(h->pos, h->prev, h->N, cells, sdf_grid(s), s->margin * s->cell,
                                             (hipStream_t)stream));
    else
        HIP_TRY(kl_launch_sdf_collide(h->pos, h->prev, h->N, h->V, cells, sdf_grid(s), s->margin * s->cell,
                                       (hipStream_t)stream));
    return TFX_OK;
}

// The compute entry point TressFXSDFCollisionSystem binds for CollideWithHair
// (TressFXSDFCollision.h:49 names CollideHairVerticesWithSdf_forward; the
// shader also holds CollideHairVerticesWithSdf, :597-715).
extern "C" int tfx_sdf_set_collide_variant(tfx_sdf* s, int variant) {
    CHECK_ARG(s, "tfx_sdf_set_collide_variant: null sdf");
    CHECK_ARG(variant == TFX_SDF_COLLIDE_FORWARD || variant == TFX_SDF_COLLIDE_MIXED,
              "tfx_sdf_set_collide_variant: unknown variant");
    s->collide_variant = variant;
    return TFX_OK;
}

extern "C" int tfx_sdf_get_grid(const tfx_sdf* s, tfx_sdf_grid_info* o) {
    CHECK_ARG(s && o, "tfx_sdf_get_grid: null argument");
    for (int k = 0; k < 3; ++k) o->origin[k] = s->origin[k];
    o->cell_size = s->cell;
    o->num_cells_x = s->nx; o->num_cells_y = s->ny; o->num_cells_z = s->nz;
    o->num_total_cells = s->total_cells;
    o->collision_margin = s->margin;
    o->padding_boundary = s->pad_boundary;
    o->explicit_grid = s->explicit_grid;
    o->pad = 0;
    return TFX_OK;
}

extern "C" int tfx_sdf_get_data_buffer(tfx_sdf* s, float** p) {
    CHECK_ARG(s && p, "tfx_sdf_get_data_buffer: null argument");
    *p = reinterpret_cast<float*>(s->cells);
    return TFX_OK;
}

extern "C" int tfx_sdf_download(tfx_sdf* s, void* stream, float* host) {
    CHECK_ARG(s && host, "tfx_sdf_download: null argument");
    hipStream_t st = (hipStream_t)stream;
    HIP_TRY(hipMemcpyAsync(host, s->cells, (size_t)s->nx * s->ny * s->nz * sizeof(float), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return TFX_OK;
}

extern "C" int tfx_sdf_get_collision_params(const tfx_sdf* s, const tfx_hair* h, tfx_sdf_collision_params* o) {
    CHECK_ARG(s && h && o, "tfx_sdf_get_collision_params: null argument");
    std::memset(o, 0, sizeof(*o));
    for (int k = 0; k < 3; ++k) o->origin[k] = s->origin[k];
    o->cell_size = s->cell;
    o->num_cells_x = s->nx; o->num_cells_y = s->ny; o->num_cells_z = s->nz;
    o->collision_margin = s->margin * s->cell;
    o->num_total_hair_vertices = h->N;
    o->num_hair_vertices_per_strand = h->V;
    return TFX_OK;
}

// ---------------------------------------------------------------------------
// Rendering hand-off and level of detail
// ---------------------------------------------------------------------------
namespace {

int stream_barrier(tfx_hair* h, void* from, void* to, const char* what) {
    if (!h) return fail(TFX_ERR_INVALID_ARGUMENT, "%s: null hair", what);
    if (from == to) return TFX_OK;
    HIP_TRY(hipSetDevice(h->device));
    if (!h->handoff) HIP_TRY(hipEventCreateWithFlags(&h->handoff, hipEventDisableTiming));
    HIP_TRY(hipEventRecord(h->handoff, (hipStream_t)from));
    HIP_TRY(hipStreamWaitEvent((hipStream_t)to, h->handoff, 0));
    return TFX_OK;
}

}  // namespace

// TressFXDynamicState::TransitionSimToRendering, TressFXHairObject.cpp:559-569.
extern "C" int tfx_hair_transition_sim_to_rendering(tfx_hair* h, void* from_stream, void* to_stream) {
    return stream_barrier(h, from_stream, to_stream, "tfx_hair_transition_sim_to_rendering");
}

// TressFXDynamicState::TransitionRenderingToSim, TressFXHairObject.cpp:571-583.
extern "C" int tfx_hair_transition_rendering_to_sim(tfx_hair* h, void* from_stream, void* to_stream) {
    return stream_barrier(h, from_stream, to_stream, "tfx_hair_transition_rendering_to_sim");
}

// TressFXRenderingSettings defaults, TressFXSettings.h:95-128.
extern "C" void tfx_rendering_lod_settings_defaults(tfx_rendering_lod_settings* o) {
    if (!o) return;
    o->lod_start_distance = 1.0f;
    o->lod_end_distance = 5.0f;
    o->lod_percent = 0.5f;
    o->lod_width_multiplier = 2.0f;
    o->shadow_lod_start_distance = 1.0f;
    o->shadow_lod_end_distance = 5.0f;
    o->shadow_lod_percent = 0.5f;
    o->shadow_lod_width_multiplier = 2.0f;
    o->fiber_radius = 0.0021f;
    o->enable_hair_lod = 0;
}

// TressFXHairObject::UpdateRenderingParameters, the LOD part, :461-481.
extern "C" int tfx_rendering_lod(const tfx_rendering_lod_settings* s, float distance, int shadow_update,
                                 float* fiber_radius, float* hair_density) {
    CHECK_ARG(s && fiber_radius && hair_density, "tfx_rendering_lod: null argument");
    float density = 1.0f;
    float radius = s->fiber_radius;
    if (s->enable_hair_lod) {
        const float a = shadow_update ? s->shadow_lod_start_distance : s->lod_start_distance;
        const float b = shadow_update ? s->shadow_lod_end_distance : s->lod_end_distance;
        const float minD = std::min(a, b), maxD = std::max(a, b);
        if (distance > minD) {
            const float ratio = std::min((distance - minD) / std::max(maxD - minD, 0.00001f), 1.0f);
            const float maxR = radius * (shadow_update ? s->shadow_lod_width_multiplier : s->lod_width_multiplier);
            radius = radius + ratio * (maxR - radius);
            density = 1.0f + ratio * ((shadow_update ? s->shadow_lod_percent : s->lod_percent) - 1.0f);
        }
    }
    *fiber_radius = radius;
    *hair_density = density;
    return TFX_OK;
}

// The draw size of TressFXHairObject::DrawStrands, :260-277, in its uint32
// arithmetic.
extern "C" int tfx_num_indices_to_render(long long total_indices, int V, int F, float density, long long* out) {
    CHECK_ARG(out && total_indices >= 0 && total_indices <= 0xffffffffll && V >= 2 && F >= 0,
              "tfx_num_indices_to_render: bad argument");
    uint32_t prims = (uint32_t)total_indices / 3u;
    if (density != 1.0f) {
        prims = (uint32_t)((float)prims * density);
        const uint32_t per_strand = (uint32_t)(V - 1) * 2u;
        const uint32_t rem = prims % per_strand;
        prims = rem > 0 ? prims + per_strand - rem : prims;
        prims = prims - (prims % (per_strand * (uint32_t)(F + 1)));
    }
    *out = (long long)prims * 3;
    return TFX_OK;
}

// GenerateTransforms, TressFXSimulation.hlsl:1035-1142.
extern "C" int tfx_hair_generate_transforms(tfx_hair* h, const tfx_hair_to_triangle_mapping* map, const float* mesh,
                                            const float* skinned, int nverts, tfx_strand_transform* out, void* stream) {
    CHECK_ARG(h && map && mesh && skinned && out && nverts >= 0, "tfx_hair_generate_transforms: bad argument");
    HIP_TRY(hipSetDevice(h->device));
    HIP_TRY(kl_launch_generate_transforms(h->rest, h->G, h->F, h->V, map, reinterpret_cast<const float4*>(mesh),
                                           reinterpret_cast<const float4*>(skinned), nverts, out, (hipStream_t)stream));
    return TFX_OK;
}
