// tfx_mesh.hip — collision-mesh bone skinning, the step before the SDF build
// (TressFXBoneSkinning, src/TressFX/TressFXBoneSkinning.cpp:287-504 and
// src/Shaders/TressFXBoneSkinning.hlsl:57-94): every vertex of the rest mesh is
// moved by the weight-normalised blend of up to four bone matrices.  The
// object is the TressFXSDFInputMeshInterface an SDF reads.
#include "tfx_internal.h"

namespace tfx {

constexpr int kSkinBlock = 256;

// One vertex of BoneSkinning (TressFXBoneSkinning.hlsl:57-94): the blended
// row-major row-vector matrix M = sum w_i B[idx_i] / sum w_i (slot 0 always,
// slots 1..3 when their weight is positive); position [p,1]·M, normal [n,0]·M;
// the w components keep the rest buffer's values.
__device__ __forceinline__ void skin_vertex(const tfx_standard_vertex& in, const tfx_bone_skinning_data& sd,
                                            const float* __restrict__ bones, tfx_standard_vertex& out) {
    const float* m0 = bone_ptr(bones, sd.bone_index[0]);
    const float w0 = sd.bone_weight[0];
    float m[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) m[k] = m0[k] * w0;
    float wsum = w0;
#pragma unroll
    for (int i = 1; i < 4; ++i) {
        const float wi = sd.bone_weight[i];
        if (wi > 0.0f) {
            const float* mi = bone_ptr(bones, sd.bone_index[i]);
#pragma unroll
            for (int k = 0; k < 16; ++k) m[k] = m[k] + mi[k] * wi;
            wsum = wsum + wi;
        }
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) m[k] = m[k] / wsum;
    const float px = in.position[0], py = in.position[1], pz = in.position[2];
    const float nx = in.normal[0], ny = in.normal[1], nz = in.normal[2];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        out.position[c] = ((px * m[c] + py * m[4 + c]) + pz * m[8 + c]) + 1.0f * m[12 + c];
        out.normal[c] = ((nx * m[c] + ny * m[4 + c]) + nz * m[8 + c]) + 0.0f * m[12 + c];
    }
    out.position[3] = in.position[3];
    out.normal[3] = in.normal[3];
}

__global__ void __launch_bounds__(kSkinBlock) k_mesh_skin(tfx_standard_vertex* __restrict__ out,
                                                          const tfx_standard_vertex* __restrict__ rest,
                                                          const tfx_bone_skinning_data* __restrict__ skin, int n,
                                                          const float* __restrict__ bones) {
    const int i = blockIdx.x * kSkinBlock + threadIdx.x;
    if (i >= n) return;
    tfx_standard_vertex v;
    skin_vertex(rest[i], skin[i], bones, v);
    out[i] = v;
}

// Batched form: job blockIdx.y skins its mesh with the bones it points to.
__global__ void __launch_bounds__(kSkinBlock) k_mesh_skin_multi(const MeshSkinJob* __restrict__ jobs) {
    const MeshSkinJob& j = jobs[blockIdx.y];
    for (int i = blockIdx.x * kSkinBlock + threadIdx.x; i < j.n; i += gridDim.x * kSkinBlock) {
        tfx_standard_vertex v;
        skin_vertex(j.rest[i], j.skin[i], j.bones, v);
        j.out[i] = v;
    }
}

}  // namespace tfx

extern "C" hipError_t tfx_launch_mesh_skin_multi(const tfx::MeshSkinJob* jobs, int njobs, int max_vertices,
                                                 hipStream_t stream) {
    using namespace tfx;
    if (njobs <= 0) return hipSuccess;
    const int bx = std::max(1, std::min((max_vertices + kSkinBlock - 1) / kSkinBlock, 1024));
    hipLaunchKernelGGL(k_mesh_skin_multi, dim3(bx, njobs), dim3(kSkinBlock), 0, stream, jobs);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------

// TressFXBoneSkinning::Initialize, TressFXBoneSkinning.cpp:287-399: device
// copies of the rest vertices (StandardVertex, w = 0), skinning data and
// indices; the skinned buffer starts as the rest mesh; the bounding cube is the
// vertex centroid +- the largest centroid distance (:377-390).
extern "C" int tfx_mesh_skinning_create(const tfx_skinned_mesh_desc* d, tfx_mesh_skinning** out) {
    CHECK_ARG(d && out, "tfx_mesh_skinning_create: null argument");
    CHECK_ARG(d->num_vertices > 0 && d->num_triangles >= 0 && d->positions && d->skinning &&
                  (d->num_triangles == 0 || d->indices),
              "tfx_mesh_skinning_create: bad mesh description");
    *out = nullptr;
    const int n = d->num_vertices;
    for (long long k = 0; k < 3ll * d->num_triangles; ++k)
        CHECK_ARG(d->indices[k] < (uint32_t)n, "tfx_mesh_skinning_create: triangle index out of range");
    tfx_mesh_skinning* m = new tfx_mesh_skinning();
    m->n = n;
    m->ntri = d->num_triangles;
    m->follow_bone = std::min(std::max(d->follow_bone, 0), TFX_MAX_NUM_BONES - 1);
    m->bones_used = std::max(skin_bones_used(d->skinning, n, 1), m->follow_bone + 1);
    std::vector<tfx_standard_vertex> rest(n);
    for (int i = 0; i < n; ++i) {
        std::memset(&rest[i], 0, sizeof(tfx_standard_vertex));
        for (int k = 0; k < 3; ++k) {
            rest[i].position[k] = d->positions[3 * i + k];
            rest[i].normal[k] = d->normals ? d->normals[3 * i + k] : 0.0f;
        }
    }
    auto cleanup = [&](int rc) { tfx_mesh_skinning_destroy(m); return rc; };
    const size_t vb = (size_t)n * sizeof(tfx_standard_vertex);
    if (hipMalloc((void**)&m->rest, vb) != hipSuccess || hipMalloc((void**)&m->out, vb) != hipSuccess ||
        hipMalloc((void**)&m->skin, (size_t)n * sizeof(tfx_bone_skinning_data)) != hipSuccess ||
        hipMalloc((void**)&m->idx, std::max<size_t>(1, (size_t)3 * m->ntri) * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc((void**)&m->bones_d, 16 * TFX_MAX_NUM_BONES * sizeof(float)) != hipSuccess)
        return cleanup(fail(TFX_ERR_OUT_OF_MEMORY, "tfx_mesh_skinning_create: device allocation failed"));
    if (hipMemcpy(m->rest, rest.data(), vb, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->out, rest.data(), vb, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->skin, d->skinning, (size_t)n * sizeof(tfx_bone_skinning_data), hipMemcpyHostToDevice) !=
            hipSuccess ||
        (m->ntri && hipMemcpy(m->idx, d->indices, (size_t)3 * m->ntri * sizeof(uint32_t), hipMemcpyHostToDevice) !=
                        hipSuccess))
        return cleanup(fail(TFX_ERR_HIP, "tfx_mesh_skinning_create: upload failed"));
    set_identity_bones(m->bones_h);
    if (hipMemcpy(m->bones_d, m->bones_h.data(), m->bones_h.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(fail(TFX_ERR_HIP, "tfx_mesh_skinning_create: upload failed"));
    // bounding cube of the rest mesh (Vector3 float arithmetic, vertex order)
    float c[3] = {0.f, 0.f, 0.f};
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) c[k] = c[k] + d->positions[3 * i + k];
    for (int k = 0; k < 3; ++k) c[k] = c[k] / (float)n;
    float radius = 0.0f;
    for (int i = 0; i < n; ++i) {
        const float dx = d->positions[3 * i] - c[0], dy = d->positions[3 * i + 1] - c[1],
                    dz = d->positions[3 * i + 2] - c[2];
        radius = std::fmax(radius, std::sqrt(dx * dx + dy * dy + dz * dz));
    }
    for (int k = 0; k < 3; ++k) {
        m->bb_min[k] = c[k] - radius;
        m->bb_max[k] = c[k] + radius;
        m->cur_min[k] = m->bb_min[k];
        m->cur_max[k] = m->bb_max[k];
    }
    *out = m;
    return TFX_OK;
}

extern "C" int tfx_mesh_skinning_destroy(tfx_mesh_skinning* m) {
    if (!m) return TFX_OK;
    (void)hipDeviceSynchronize();
    (void)hipFree(m->rest);
    (void)hipFree(m->out);
    (void)hipFree(m->skin);
    (void)hipFree(m->idx);
    (void)hipFree(m->bones_d);
    delete m;
    return TFX_OK;
}

// TressFXBoneSkinning::Update, TressFXBoneSkinning.cpp:460-504.
extern "C" int tfx_mesh_skinning_update(tfx_mesh_skinning* m, const float* bones, int num_bones, void* stream) {
    CHECK_ARG(m && (bones || num_bones == 0) && num_bones >= 0, "tfx_mesh_skinning_update: bad argument");
    hipStream_t st = (hipStream_t)stream;
    const int nb = std::min(num_bones, TFX_MAX_NUM_BONES);
    if (nb) std::memcpy(m->bones_h.data(), bones, (size_t)nb * 16 * sizeof(float));
    mesh_follow_bbox(m);
    int rc = m->staging.ensure(16 * TFX_MAX_NUM_BONES * sizeof(float));
    if (rc) return rc;
    void* host;
    int slot;
    if ((rc = m->staging.acquire(&host, &slot))) return rc;
    const size_t bytes = (size_t)std::max(nb, 1) * 16 * sizeof(float);
    std::memcpy(host, m->bones_h.data(), bytes);
    HIP_TRY(hipMemcpyAsync(m->bones_d, host, bytes, hipMemcpyHostToDevice, st));
    if ((rc = m->staging.release_slot(slot, st))) return rc;
    PassTimer tm("BoneSkinning", st);
    const int blocks = (m->n + tfx::kSkinBlock - 1) / tfx::kSkinBlock;
    hipLaunchKernelGGL(tfx::k_mesh_skin, dim3(blocks), dim3(tfx::kSkinBlock), 0, st, m->out, m->rest, m->skin, m->n,
                       m->bones_d);
    HIP_TRY(hipGetLastError());
    return TFX_OK;
}

// The TressFXSDFInputMeshInterface view (TressFXBoneSkinning.h:70-81).
extern "C" int tfx_mesh_skinning_get_mesh(const tfx_mesh_skinning* m, tfx_sdf_mesh_desc* out) {
    CHECK_ARG(m && out, "tfx_mesh_skinning_get_mesh: null argument");
    std::memset(out, 0, sizeof(*out));
    out->vertices = m->out;
    out->indices = m->idx;
    out->num_vertices = m->n;
    out->num_triangles = m->ntri;
    for (int k = 0; k < 3; ++k) {
        out->bbox_min[k] = m->cur_min[k];
        out->bbox_max[k] = m->cur_max[k];
        out->initial_bbox_min[k] = m->bb_min[k];
        out->initial_bbox_max[k] = m->bb_max[k];
    }
    return TFX_OK;
}

extern "C" int tfx_mesh_skinning_download(tfx_mesh_skinning* m, void* stream, tfx_standard_vertex* host) {
    CHECK_ARG(m && host, "tfx_mesh_skinning_download: null argument");
    hipStream_t st = (hipStream_t)stream;
    HIP_TRY(hipMemcpyAsync(host, m->out, (size_t)m->n * sizeof(tfx_standard_vertex), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return TFX_OK;
}
