// tfx_asset.hip — TressFXAsset's loaders and processing (host code):
//   .tfx       LoadHairData          src/TressFX/TressFXAsset.cpp:94-173
//   follows    GenerateFollowHairs   :177-248
//   processing ProcessAsset          :250-399 (tangents, thickness, rest lengths, triangle indices)
//   .tfxbone   LoadBoneData          :401-480
//   .tfxmesh   TressFXBoneSkinning::LoadTressFXCollisionMeshData, TressFXBoneSkinning.cpp:100-286
// The arithmetic is the reference's Vector3 arithmetic in float (component by
// component, sums left to right, Normalize = v / sqrt(x*x + y*y + z*z) unless
// the length is 0).  Where the reference reads uninitialised memory (the w of a
// Vector3 built from three floats, i.e. tangent w) this code writes 0; where it
// calls rand() (follow-hair offsets) it uses a seeded 32-bit LCG.
#include "tfx_internal.h"

#include <cstdint>
#include <fstream>
#include <sstream>

namespace {

constexpr int kGroup = TFX_SIM_THREAD_GROUP_SIZE;
constexpr int kMaxInfluence = 4;  // TRESSFX_MAX_INFLUENTIAL_BONE_COUNT

struct V4 {
    float x, y, z, w;
};

inline V4 sub3(const V4& a, const V4& b) { return V4{a.x - b.x, a.y - b.y, a.z - b.z, 0.0f}; }
inline V4 add3(const V4& a, const V4& b) { return V4{a.x + b.x, a.y + b.y, a.z + b.z, 0.0f}; }
inline V4 scale3(const V4& a, float s) { return V4{a.x * s, a.y * s, a.z * s, 0.0f}; }
inline float length3(const V4& a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
// Vector3::Normalize (Vector3D.cpp:48-62): unchanged when the length is 0.
inline V4 normalize3(V4 a) {
    const float d = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    if (d == 0.0f) return a;
    a.x = a.x / d;
    a.y = a.y / d;
    a.z = a.z / d;
    return a;
}

// GetTangentVectors, TressFXAsset.cpp:37-63.
void tangent_vectors(const V4& n, V4& t0, V4& t1) {
    if (std::fabs(n.z) > 0.707f) {
        const float a = n.y * n.y + n.z * n.z;
        const float k = 1.0f / std::sqrt(a);
        t0 = V4{0.0f, -n.z * k, n.y * k, 0.0f};
        t1 = V4{a * k, -n.x * t0.z, n.x * t0.y, 0.0f};
    } else {
        const float a = n.x * n.x + n.y * n.y;
        const float k = 1.0f / std::sqrt(a);
        t0 = V4{-n.y * k, n.x * k, 0.0f, 0.0f};
        t1 = V4{-n.z * t0.y, n.z * t0.x, a * k, 0.0f};
    }
}

bool read_file(const char* path, std::vector<unsigned char>& out) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    const long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (n < 0) {
        std::fclose(f);
        return false;
    }
    out.resize((size_t)n);
    const size_t got = n ? std::fread(out.data(), 1, (size_t)n, f) : 0;
    std::fclose(f);
    return got == (size_t)n;
}

int engine_id(const std::vector<std::string>& names, const char* const* engine, int n_engine, const std::string& name) {
    (void)names;
    for (int i = 0; i < n_engine; ++i)
        if (engine[i] && name == engine[i]) return i;
    return -1;  // GetBoneIdByName of a name the skeleton does not have
}

}  // namespace

struct tfx_asset {
    int file_strands = 0;
    int G = 0, F = 0, V = 0, S = 0, N = 0;
    std::vector<V4> positions;             // [N]
    std::vector<float> strand_uv;          // [S][2]
    std::vector<V4> follow_offsets;        // [S]
    std::vector<V4> tangents;              // [N]
    std::vector<float> rest_lengths;       // [N]
    std::vector<float> thickness;          // [N]
    std::vector<int32_t> triangle_indices; // [6 * S * (V - 1)]
    std::vector<tfx_bone_skinning_data> skin;  // [S]
    bool processed = false;
};

struct tfx_collision_mesh {
    int n = 0, ntri = 0, follow_bone = 0;
    std::vector<float> positions, normals;  // [n][3]
    std::vector<tfx_bone_skinning_data> skin;
    std::vector<uint32_t> indices;
    std::vector<std::string> bone_names;
};

// LoadHairData, TressFXAsset.cpp:94-173.
extern "C" int tfx_asset_load_hair_data(const char* path, tfx_asset** out) {
    if (!path || !out) return fail(TFX_ERR_INVALID_ARGUMENT, "tfx_asset_load_hair_data: null argument");
    *out = nullptr;
    std::vector<unsigned char> bytes;
    if (!read_file(path, bytes)) return fail(TFX_ERR_NOT_FOUND, "tfx_asset_load_hair_data: cannot read %s", path);
    // TressFXTFXFileHeader, TressFXFileFormat.h:34-52: float version, 7 uint32, 32 reserved
    constexpr size_t kHeader = 4 + 7 * 4 + 32 * 4;
    if (bytes.size() < kHeader) return fail(TFX_ERR_INVALID_ARGUMENT, "%s: shorter than a .tfx header", path);
    float version;
    uint32_t h[7];
    std::memcpy(&version, bytes.data(), 4);
    std::memcpy(h, bytes.data() + 4, sizeof(h));
    const uint32_t num_strands = h[0], V = h[1], off_pos = h[2], off_uv = h[3];
    if (!(version >= 4.0f))  // the reference refuses versions below its major version (:105)
        return fail(TFX_ERR_UNSUPPORTED, "%s: .tfx version %g below 4", path, version);
    if (!(V > 2 && V <= (uint32_t)kGroup && kGroup % V == 0))
        return fail(TFX_ERR_INVALID_ARGUMENT, "%s: %u vertices per strand (4, 8, 16, 32 or 64)", path, V);
    if (num_strands == 0 || num_strands > (1u << 26))
        return fail(TFX_ERR_INVALID_ARGUMENT, "%s: %u strands", path, num_strands);
    const size_t pos_bytes = (size_t)num_strands * V * 16;
    if ((size_t)off_pos + pos_bytes > bytes.size())
        return fail(TFX_ERR_INVALID_ARGUMENT, "%s: vertex positions past the end of the file", path);
    const bool has_uv = off_uv != 0;
    if (has_uv && (size_t)off_uv + (size_t)num_strands * 8 > bytes.size())
        return fail(TFX_ERR_INVALID_ARGUMENT, "%s: strand UVs past the end of the file", path);
    tfx_asset* a = new tfx_asset();
    a->file_strands = (int)num_strands;
    // guides padded to a multiple of the thread group, always adding 1..64 (:111)
    a->G = (int)(num_strands - num_strands % kGroup) + kGroup;
    a->F = 0;
    a->V = (int)V;
    a->S = a->G;
    a->N = a->S * a->V;
    a->positions.assign(a->N, V4{0, 0, 0, 0});
    std::memcpy(a->positions.data(), bytes.data() + off_pos, pos_bytes);
    for (int i = 0; i < a->G - (int)num_strands; ++i)  // made-up strands copy the last one (:139-149)
        for (int j = 0; j < a->V; ++j)
            a->positions[(size_t)(num_strands + i) * a->V + j] = a->positions[(size_t)(num_strands - 1) * a->V + j];
    a->strand_uv.assign((size_t)2 * a->S, 0.0f);
    if (has_uv) std::memcpy(a->strand_uv.data(), bytes.data() + off_uv, (size_t)num_strands * 8);
    for (int i = (int)num_strands; i < a->G; ++i) {
        a->strand_uv[2 * i] = a->strand_uv[2 * (num_strands - 1)];
        a->strand_uv[2 * i + 1] = a->strand_uv[2 * (num_strands - 1) + 1];
    }
    a->follow_offsets.assign(a->S, V4{0, 0, 0, 0});
    *out = a;
    return TFX_OK;
}

// GenerateFollowHairs, TressFXAsset.cpp:177-248.  The random offsets use a
// 32-bit LCG seeded with `seed` (the reference calls rand()):
// U(-r, r) = (lcg / 2^32) * (r - (-r)) + (-r).
extern "C" int tfx_asset_generate_follow_hairs(tfx_asset* a, int num_follow, float tip_separation, float max_radius,
                                               unsigned int seed) {
    if (!a || num_follow < 0) return fail(TFX_ERR_INVALID_ARGUMENT, "tfx_asset_generate_follow_hairs: bad argument");
    if (a->F != 0 || a->processed)
        return fail(TFX_ERR_INVALID_ARGUMENT, "tfx_asset_generate_follow_hairs: call once, before processing");
    if (num_follow == 0) return TFX_OK;
    a->F = num_follow;
    a->S = a->G * (a->F + 1);
    a->N = a->S * a->V;
    const std::vector<V4> guide = a->positions;
    const std::vector<float> guide_uv = a->strand_uv;
    a->positions.assign(a->N, V4{0, 0, 0, 0});
    a->strand_uv.assign((size_t)2 * a->S, 0.0f);
    a->follow_offsets.assign(a->S, V4{0, 0, 0, 0});
    uint32_t state = seed;
    auto rnd = [&](float lo, float hi) {
        state = state * 1664525u + 1013904223u;
        return ((float)state / 4294967296.0f) * (hi - lo) + lo;
    };
    const int V = a->V;
    for (int i = 0; i < a->G; ++i) {
        const int gs = i * (a->F + 1);
        const size_t root = (size_t)gs * V;
        for (int k = 0; k < V; ++k) a->positions[root + k] = guide[(size_t)i * V + k];
        a->strand_uv[2 * gs] = guide_uv[2 * i];
        a->strand_uv[2 * gs + 1] = guide_uv[2 * i + 1];
        a->follow_offsets[gs] = V4{0.0f, 0.0f, 0.0f, (float)gs};
        const V4 v01 = normalize3(sub3(a->positions[root + 1], a->positions[root]));
        V4 t0, t1;
        tangent_vectors(v01, t0, t1);
        for (int j = 0; j < a->F; ++j) {
            const int fs = gs + j + 1;
            const size_t froot = (size_t)fs * V;
            a->strand_uv[2 * fs] = a->strand_uv[2 * gs];
            a->strand_uv[2 * fs + 1] = a->strand_uv[2 * gs + 1];
            const float r0 = rnd(-max_radius, max_radius);
            const float r1 = rnd(-max_radius, max_radius);
            V4 offset = add3(scale3(t0, r0), scale3(t1, r1));
            offset.w = (float)gs;
            a->follow_offsets[fs] = offset;
            for (int k = 0; k < V; ++k) {
                const V4& g = a->positions[root + k];
                const float factor = tip_separation * ((float)k / (float)V) + 1.0f;
                V4 f = add3(g, scale3(offset, factor));
                f.w = g.w;
                a->positions[froot + k] = f;
            }
        }
    }
    return TFX_OK;
}

// ProcessAsset, TressFXAsset.cpp:250-399.
extern "C" int tfx_asset_process(tfx_asset* a) {
    if (!a) return fail(TFX_ERR_INVALID_ARGUMENT, "tfx_asset_process: null asset");
    const int V = a->V;
    // ComputeStrandTangent (:304-342); the last vertex keeps a zero tangent
    a->tangents.assign(a->N, V4{0, 0, 0, 0});
    for (int s = 0; s < a->S; ++s) {
        const size_t r = (size_t)s * V;
        a->tangents[r] = normalize3(sub3(a->positions[r + 1], a->positions[r]));
        for (int i = 1; i < V - 1; ++i) {
            const V4 pre = normalize3(sub3(a->positions[r + i], a->positions[r + i - 1]));
            const V4 next = normalize3(sub3(a->positions[r + i + 1], a->positions[r + i]));
            a->tangents[r + i] = normalize3(add3(pre, next));
        }
    }
    // ComputeThicknessCoeffs (:344-375), including its repeated division of tVal
    a->thickness.assign(a->N, 0.0f);
    size_t idx = 0;
    for (int s = 0; s < a->S; ++s) {
        const size_t r = (size_t)s * V;
        float len = 0.0f, t = 0.0f;
        for (int i = 1; i < V; ++i) {
            const float d = length3(sub3(a->positions[r + i], a->positions[r + i - 1]));
            t += d;
            len += d;
        }
        for (int i = 0; i < V; ++i) {
            t /= len;
            a->thickness[idx++] = std::sqrt(1.0f - t * t);
        }
    }
    // ComputeRestLengths (:377-399): the last entry of a strand is 0
    a->rest_lengths.assign(a->N, 0.0f);
    idx = 0;
    for (int s = 0; s < a->S; ++s) {
        const size_t r = (size_t)s * V;
        for (int j = 0; j < V - 1; ++j) a->rest_lengths[idx++] = length3(sub3(a->positions[r + j], a->positions[r + j + 1]));
        a->rest_lengths[idx++] = 0.0f;
    }
    // FillTriangleIndexArray (:276-302)
    a->triangle_indices.assign((size_t)6 * a->S * (V - 1), 0);
    int32_t id = 0;
    size_t c = 0;
    for (int s = 0; s < a->S; ++s) {
        for (int j = 0; j < V - 1; ++j) {
            const int32_t t6[6] = {2 * id, 2 * id + 1, 2 * id + 2, 2 * id + 2, 2 * id + 1, 2 * id + 3};
            for (int k = 0; k < 6; ++k) a->triangle_indices[c++] = t6[k];
            ++id;
        }
        ++id;
    }
    // a skinning table sized for the strands (LoadBoneData fills the guides)
    if (a->skin.size() != (size_t)a->S) {
        a->skin.assign(a->S, tfx_bone_skinning_data{{0, 0, 0, 0}, {1, 0, 0, 0}});
    }
    a->processed = true;
    return TFX_OK;
}

// LoadBoneData, TressFXAsset.cpp:401-480: bone names mapped to the engine's
// ids (their index in `engine_bone_names`), skinning records of the file's
// strands at guide slots i*(F+1), padding guides copying the last record.
extern "C" int tfx_asset_load_bone_data(tfx_asset* a, const char* path, const char* const* engine_bone_names,
                                        int num_engine_bones) {
    if (!a || !path || (num_engine_bones > 0 && !engine_bone_names) || num_engine_bones < 0)
        return fail(TFX_ERR_INVALID_ARGUMENT, "tfx_asset_load_bone_data: bad argument");
    std::vector<unsigned char> b;
    if (!read_file(path, b)) return fail(TFX_ERR_NOT_FOUND, "tfx_asset_load_bone_data: cannot read %s", path);
    size_t p = 0;
    auto rd = [&](void* dst, size_t n) {
        if (p + n > b.size()) return false;
        std::memcpy(dst, b.data() + p, n);
        p += n;
        return true;
    };
    int32_t num_bones = 0;
    if (!rd(&num_bones, 4) || num_bones < 0 || num_bones > 100000)
        return fail(TFX_ERR_INVALID_ARGUMENT, "%s: bad bone count", path);
    std::vector<std::string> names(num_bones);
    for (int i = 0; i < num_bones; ++i) {
        int32_t bone_index, len;
        if (!rd(&bone_index, 4) || !rd(&len, 4) || len < 0 || len > 4096)
            return fail(TFX_ERR_INVALID_ARGUMENT, "%s: bad bone name record %d", path, i);
        std::string s(len, '\0');
        if (!rd(&s[0], (size_t)len)) return fail(TFX_ERR_INVALID_ARGUMENT, "%s: truncated bone names", path);
        names[i] = std::string(s.c_str());  // the stored length includes the terminating zero
    }
    std::vector<int> ids(num_bones);
    for (int i = 0; i < num_bones; ++i) ids[i] = engine_id(names, engine_bone_names, num_engine_bones, names[i]);
    int32_t num_strands = 0;
    if (!rd(&num_strands, 4) || num_strands < 0) return fail(TFX_ERR_INVALID_ARGUMENT, "%s: bad strand count", path);
    if (a->G < num_strands)  // :438-439
        return fail(TFX_ERR_INVALID_ARGUMENT, "%s: %d strands but the asset has %d guides", path, num_strands, a->G);
    a->skin.assign(a->S, tfx_bone_skinning_data{{0, 0, 0, 0}, {1, 0, 0, 0}});
    tfx_bone_skinning_data sd{{0, 0, 0, 0}, {1, 0, 0, 0}};
    for (int i = 0; i < num_strands; ++i) {
        int32_t index;
        if (!rd(&index, 4)) return fail(TFX_ERR_INVALID_ARGUMENT, "%s: truncated skinning data", path);
        for (int j = 0; j < kMaxInfluence; ++j) {
            int32_t bi;
            float w;
            if (!rd(&bi, 4) || !rd(&w, 4)) return fail(TFX_ERR_INVALID_ARGUMENT, "%s: truncated skinning data", path);
            if (bi < 0 || bi >= num_bones) return fail(TFX_ERR_INVALID_ARGUMENT, "%s: bone index %d", path, bi);
            sd.bone_index[j] = (float)ids[bi];
            sd.bone_weight[j] = w;
        }
        // no associated bone (-1) becomes bone 0; its weight is zero (:463-468)
        for (int j = 0; j < kMaxInfluence; ++j) sd.bone_index[j] = sd.bone_index[j] == -1.0f ? 0.0f : sd.bone_index[j];
        a->skin[(size_t)i * (a->F + 1)] = sd;
    }
    for (int i = num_strands; i < a->G; ++i) a->skin[(size_t)i * (a->F + 1)] = sd;  // :474-477
    return TFX_OK;
}

extern "C" int tfx_asset_destroy(tfx_asset* a) {
    delete a;
    return TFX_OK;
}

extern "C" int tfx_asset_get_desc(const tfx_asset* a, tfx_asset_desc* d) {
    if (!a || !d) return fail(TFX_ERR_INVALID_ARGUMENT, "tfx_asset_get_desc: null argument");
    if (!a->processed) return fail(TFX_ERR_INVALID_ARGUMENT, "tfx_asset_get_desc: call tfx_asset_process first");
    std::memset(d, 0, sizeof(*d));
    d->num_guide_strands = a->G;
    d->num_follow_strands_per_guide = a->F;
    d->num_vertices_per_strand = a->V;
    d->positions = &a->positions[0].x;
    d->tangents = &a->tangents[0].x;
    d->rest_lengths = a->rest_lengths.data();
    d->follow_root_offsets = &a->follow_offsets[0].x;
    d->bone_skinning_data = a->skin.data();
    return TFX_OK;
}

extern "C" int tfx_asset_get_info(const tfx_asset* a, int* file_strands, int* guide_strands, int* follow_per_guide,
                                  int* vertices_per_strand) {
    if (!a) return fail(TFX_ERR_INVALID_ARGUMENT, "tfx_asset_get_info: null asset");
    if (file_strands) *file_strands = a->file_strands;
    if (guide_strands) *guide_strands = a->G;
    if (follow_per_guide) *follow_per_guide = a->F;
    if (vertices_per_strand) *vertices_per_strand = a->V;
    return TFX_OK;
}

// The render-side arrays of ProcessAsset.
extern "C" int tfx_asset_get_render_data(const tfx_asset* a, const float** strand_uv, const float** thickness,
                                         const int32_t** triangle_indices, int* num_triangle_indices) {
    if (!a || !a->processed) return fail(TFX_ERR_INVALID_ARGUMENT, "tfx_asset_get_render_data: process the asset first");
    if (strand_uv) *strand_uv = a->strand_uv.data();
    if (thickness) *thickness = a->thickness.data();
    if (triangle_indices) *triangle_indices = a->triangle_indices.data();
    if (num_triangle_indices) *num_triangle_indices = (int)a->triangle_indices.size();
    return TFX_OK;
}

// LoadTressFXCollisionMeshData, TressFXBoneSkinning.cpp:100-286: a text file of
// "numOfBones", "numOfVertices" (index x y z nx ny nz b0 b1 b2 b3 w0 w1 w2 w3)
// and "numOfTriangles" (index i0 i1 i2) blocks; '#' lines are comments.
extern "C" int tfx_collision_mesh_load(const char* path, const char* const* engine_bone_names, int num_engine_bones,
                                       const char* follow_bone, tfx_collision_mesh** out) {
    if (!path || !out || num_engine_bones < 0 || (num_engine_bones > 0 && !engine_bone_names))
        return fail(TFX_ERR_INVALID_ARGUMENT, "tfx_collision_mesh_load: bad argument");
    *out = nullptr;
    std::ifstream in(path);
    if (!in.is_open()) return fail(TFX_ERR_NOT_FOUND, "tfx_collision_mesh_load: cannot read %s", path);
    tfx_collision_mesh* m = new tfx_collision_mesh();
    auto bad = [&](const char* what) {
        delete m;
        return fail(TFX_ERR_INVALID_ARGUMENT, "%s: %s", path, what);
    };
    auto next_data_line = [&](std::vector<std::string>& tok) {
        std::string line;
        while (std::getline(in, line)) {
            if (line.empty() || line[0] == '#') continue;
            tok.clear();
            std::istringstream ss(line);
            std::string t;
            while (ss >> t) tok.push_back(t);
            if (tok.empty()) continue;
            return true;
        }
        return false;
    };
    std::vector<std::string> tok;
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream ss(line);
        std::vector<std::string> t;
        std::string w;
        while (ss >> w) t.push_back(w);
        if (t.empty()) continue;
        const std::string& key = t[0];
        if (key.find("numOfBones") != std::string::npos && t.size() > 1) {
            const int nb = std::atoi(t[1].c_str());
            for (int i = 0; i < nb; ++i) {
                if (!next_data_line(tok) || tok.size() < 2) return bad("truncated bone names");
                m->bone_names.push_back(tok[1]);
            }
        } else if (key.find("numOfVertices") != std::string::npos && t.size() > 1) {
            m->n = std::atoi(t[1].c_str());
            if (m->n <= 0) return bad("no vertices");
            m->positions.assign((size_t)3 * m->n, 0.0f);
            m->normals.assign((size_t)3 * m->n, 0.0f);
            m->skin.assign(m->n, tfx_bone_skinning_data{{0, 0, 0, 0}, {0, 0, 0, 0}});
            for (int i = 0; i < m->n; ++i) {
                if (!next_data_line(tok) || tok.size() != 15) return bad("vertex line without 15 fields");
                if (std::atoi(tok[0].c_str()) != i) return bad("vertex lines out of order");
                for (int k = 0; k < 3; ++k) {
                    m->positions[3 * i + k] = (float)std::atof(tok[1 + k].c_str());
                    m->normals[3 * i + k] = (float)std::atof(tok[4 + k].c_str());
                }
                for (int k = 0; k < 4; ++k) {
                    const int bi = std::atoi(tok[7 + k].c_str());
                    if (bi < 0 || bi >= (int)m->bone_names.size()) return bad("bone index out of range");
                    m->skin[i].bone_index[k] =
                        (float)engine_id(m->bone_names, engine_bone_names, num_engine_bones, m->bone_names[bi]);
                    m->skin[i].bone_weight[k] = (float)std::atof(tok[11 + k].c_str());
                }
            }
        } else if (key.find("numOfTriangles") != std::string::npos && t.size() > 1) {
            m->ntri = std::atoi(t[1].c_str());
            if (m->ntri < 0) return bad("bad triangle count");
            m->indices.assign((size_t)3 * m->ntri, 0);
            for (int i = 0; i < m->ntri; ++i) {
                if (!next_data_line(tok) || tok.size() != 4) return bad("triangle line without 4 fields");
                if (std::atoi(tok[0].c_str()) != i) return bad("triangle lines out of order");
                for (int k = 0; k < 3; ++k) {
                    const long v = std::atol(tok[1 + k].c_str());
                    if (v < 0 || v >= m->n) return bad("triangle index out of range");
                    m->indices[3 * i + k] = (uint32_t)v;
                }
            }
        }
    }
    if (m->n == 0) return bad("no numOfVertices block");
    m->follow_bone = follow_bone ? engine_id(m->bone_names, engine_bone_names, num_engine_bones, follow_bone) : 0;
    if (m->follow_bone < 0) m->follow_bone = 0;
    // an engine id of -1 (a bone the skeleton lacks) reads bone 0
    for (auto& s : m->skin)
        for (int k = 0; k < 4; ++k)
            if (s.bone_index[k] < 0.0f) s.bone_index[k] = 0.0f;
    *out = m;
    return TFX_OK;
}

extern "C" int tfx_collision_mesh_get_desc(const tfx_collision_mesh* m, tfx_skinned_mesh_desc* d) {
    if (!m || !d) return fail(TFX_ERR_INVALID_ARGUMENT, "tfx_collision_mesh_get_desc: null argument");
    std::memset(d, 0, sizeof(*d));
    d->num_vertices = m->n;
    d->num_triangles = m->ntri;
    d->positions = m->positions.data();
    d->normals = m->normals.data();
    d->skinning = m->skin.data();
    d->indices = m->indices.data();
    d->follow_bone = m->follow_bone;
    return TFX_OK;
}

extern "C" int tfx_collision_mesh_destroy(tfx_collision_mesh* m) {
    delete m;
    return TFX_OK;
}
