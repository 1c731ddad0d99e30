/* tressfx_amd.hpp — C++ classes with the reference's engine-interface names
 * (src/TressFX/TressFXSimulation.h, TressFXHairObject.h, TressFXSDFCollision.h,
 * TressFXSDFInputMeshInterface.h, TressFXAsset.h, TressFXSettings.h) over the
 * C-ABI of tressfx_amd.h.  Header only; link with libtressfx_amd.so.
 *
 * Differences from the reference signatures, all forced by the platform:
 *   EI_CommandContext&  ->  void* stream (a hipStream_t; nullptr = default stream)
 *   EI_Device*          ->  the device chosen by tfx_init (TressFXSimulation::Initialize(int device))
 *   EI_Resource&        ->  device pointers
 *   errors              ->  tressfx_amd::Error (the reference asserts)
 */
#ifndef TRESSFX_AMD_HPP_
#define TRESSFX_AMD_HPP_

#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "tressfx_amd.h"

namespace tressfx_amd {

class Error : public std::runtime_error {
public:
    Error(int code, const char* what) : std::runtime_error(std::string("tressfx_amd: ") + what), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

inline void check(int rc) {
    if (rc != TFX_OK) throw Error(rc, tfx_last_error());
}

/* TressFXSimulationSettings, TressFXSettings.h:35-87 (defaults of its constructor). */
struct TressFXSimulationSettings : tfx_simulation_settings {
    TressFXSimulationSettings() { tfx_simulation_settings_default(this); }
};

/* TressFXAsset (TressFXAsset.h:48-120): host arrays, strand-major, follow
 * strands after their guide.  Filled either directly or by the loaders, in the
 * reference's order LoadHairData, GenerateFollowHairs, ProcessAsset,
 * LoadBoneData (the engine's bone names give the engine bone ids). */
struct TressFXAsset {
    int m_numGuideStrands = 0;
    int m_numFollowStrandsPerGuide = 0;
    int m_numVerticesPerStrand = 0;
    std::vector<float> m_positions;          /* float4 per vertex, w = inverse mass */
    std::vector<float> m_tangents;           /* float4 per vertex */
    std::vector<float> m_restLengths;        /* float per vertex */
    std::vector<float> m_followRootOffsets;  /* float4 per strand */
    std::vector<tfx_bone_skinning_data> m_boneSkinningData;  /* per strand */

    TressFXAsset() = default;
    TressFXAsset(const TressFXAsset&) = delete;
    TressFXAsset& operator=(const TressFXAsset&) = delete;
    TressFXAsset(TressFXAsset&& o) noexcept { *this = std::move(o); }
    TressFXAsset& operator=(TressFXAsset&& o) noexcept {
        if (this != &o) {
            m_numGuideStrands = o.m_numGuideStrands;
            m_numFollowStrandsPerGuide = o.m_numFollowStrandsPerGuide;
            m_numVerticesPerStrand = o.m_numVerticesPerStrand;
            m_positions = std::move(o.m_positions);
            m_tangents = std::move(o.m_tangents);
            m_restLengths = std::move(o.m_restLengths);
            m_followRootOffsets = std::move(o.m_followRootOffsets);
            m_boneSkinningData = std::move(o.m_boneSkinningData);
            tfx_asset_destroy(loader_);
            loader_ = o.loader_;
            o.loader_ = nullptr;
        }
        return *this;
    }
    ~TressFXAsset() { tfx_asset_destroy(loader_); }

    int NumTotalStrands() const { return m_numGuideStrands * (m_numFollowStrandsPerGuide + 1); }
    int NumTotalVertices() const { return NumTotalStrands() * m_numVerticesPerStrand; }

    bool LoadHairData(const char* tfxPath) {
        tfx_asset_destroy(loader_);
        loader_ = nullptr;
        check(tfx_asset_load_hair_data(tfxPath, &loader_));
        return true;
    }
    bool GenerateFollowHairs(int numFollowHairsPerGuideHair, float tipSeparationFactor, float maxRadiusAroundGuideHair,
                             unsigned int seed = 0) {
        check(tfx_asset_generate_follow_hairs(loader_, numFollowHairsPerGuideHair, tipSeparationFactor,
                                              maxRadiusAroundGuideHair, seed));
        return true;
    }
    bool ProcessAsset() {
        check(tfx_asset_process(loader_));
        copy_from_loader();
        return true;
    }
    bool LoadBoneData(const char* tfxbonePath, const std::vector<std::string>& engineBoneNames) {
        std::vector<const char*> names;
        for (const std::string& n : engineBoneNames) names.push_back(n.c_str());
        check(tfx_asset_load_bone_data(loader_, tfxbonePath, names.data(), (int)names.size()));
        copy_from_loader();
        return true;
    }

    tfx_asset_desc desc() const {
        tfx_asset_desc d{};
        d.num_guide_strands = m_numGuideStrands;
        d.num_follow_strands_per_guide = m_numFollowStrandsPerGuide;
        d.num_vertices_per_strand = m_numVerticesPerStrand;
        d.positions = m_positions.data();
        d.tangents = m_tangents.empty() ? nullptr : m_tangents.data();
        d.rest_lengths = m_restLengths.empty() ? nullptr : m_restLengths.data();
        d.follow_root_offsets = m_followRootOffsets.empty() ? nullptr : m_followRootOffsets.data();
        d.bone_skinning_data = m_boneSkinningData.empty() ? nullptr : m_boneSkinningData.data();
        return d;
    }

private:
    void copy_from_loader() {
        tfx_asset_desc d{};
        check(tfx_asset_get_desc(loader_, &d));
        m_numGuideStrands = d.num_guide_strands;
        m_numFollowStrandsPerGuide = d.num_follow_strands_per_guide;
        m_numVerticesPerStrand = d.num_vertices_per_strand;
        const size_t S = (size_t)NumTotalStrands(), N = (size_t)NumTotalVertices();
        m_positions.assign(d.positions, d.positions + 4 * N);
        m_tangents.assign(d.tangents, d.tangents + 4 * N);
        m_restLengths.assign(d.rest_lengths, d.rest_lengths + N);
        m_followRootOffsets.assign(d.follow_root_offsets, d.follow_root_offsets + 4 * S);
        m_boneSkinningData.assign(d.bone_skinning_data, d.bone_skinning_data + S);
    }
    tfx_asset* loader_ = nullptr;
};

/* TressFXHairObject, TressFXHairObject.h:85-191 (simulation members). */
class TressFXHairObject {
public:
    explicit TressFXHairObject(const TressFXAsset* asset) {
        const tfx_asset_desc d = asset->desc();
        check(tfx_hair_create(&d, &h_));
    }
    ~TressFXHairObject() { tfx_hair_destroy(h_); }
    TressFXHairObject(const TressFXHairObject&) = delete;
    TressFXHairObject& operator=(const TressFXHairObject&) = delete;

    /* row-major, row-vector float4x4 matrices (16 floats each) */
    void UpdateBoneMatrices(const float* pBoneMatrices, int numBoneMatrices) {
        check(tfx_hair_update_bone_matrices(h_, pBoneMatrices, numBoneMatrices));
    }
    void UpdateSimulationParameters(const TressFXSimulationSettings* parameters, float timeStep) {
        check(tfx_hair_update_simulation_parameters(h_, parameters, timeStep));
    }
    void SetCollisionCapsules(const float* centerAndRadius0, const float* centerAndRadius1, int numCapsules) {
        check(tfx_hair_set_collision_capsules(h_, centerAndRadius0, centerAndRadius1, numCapsules));
    }
    void ResetPositions() { check(tfx_hair_reset_positions(h_)); }
    // TressFXDynamicState::TransitionSimToRendering / TransitionRenderingToSim
    // (TressFXHairObject.cpp:559-583): later work on toStream waits for fromStream.
    void TransitionSimToRendering(void* fromStream, void* toStream) {
        check(tfx_hair_transition_sim_to_rendering(h_, fromStream, toStream));
    }
    void TransitionRenderingToSim(void* fromStream, void* toStream) {
        check(tfx_hair_transition_rendering_to_sim(h_, fromStream, toStream));
    }
    void IncreaseSimulationFrame() { check(tfx_hair_increase_simulation_frame(h_)); }
    tfx_dynamic_state GetDynamicState() const {
        tfx_dynamic_state s{};
        check(tfx_hair_get_dynamic_state(h_, &s));
        return s;
    }
    int GetNumTotalHairVertices() const { return info().num_total_vertices; }
    int GetNumTotalHairStrands() const { return info().num_total_strands; }
    int GetNumVerticesPerStrand() const { return info().num_vertices_per_strand; }
    int GetCPULocalShapeIterations() const { return info().cpu_local_shape_iterations; }
    int GetNumFollowHairsPerGuideHair() const { return info().num_follow_strands_per_guide; }

    tfx_hair* handle() const { return h_; }

private:
    tfx_hair_info info() const {
        tfx_hair_info i{};
        check(tfx_hair_get_info(h_, &i));
        return i;
    }
    tfx_hair* h_ = nullptr;
};

/* TressFXSimulation, TressFXSimulation.h:34-48. */
class TressFXSimulation {
public:
    TressFXSimulation() { check(tfx_simulation_create(&s_)); }
    ~TressFXSimulation() { tfx_simulation_destroy(s_); }
    TressFXSimulation(const TressFXSimulation&) = delete;
    TressFXSimulation& operator=(const TressFXSimulation&) = delete;

    static void Initialize(int device) { check(tfx_init(device)); }
    /* Kernel arithmetic (TFX_ARITH_IEEE / TFX_ARITH_FAST), process-wide; no
     * counterpart in the reference. */
    static void SetArithmetic(int mode) { check(tfx_set_arithmetic(mode)); }
    static int GetArithmetic() {
        int m = 0;
        check(tfx_get_arithmetic(&m));
        return m;
    }

    void Simulate(void* stream, std::vector<TressFXHairObject*>& hairObjects) {
        std::vector<tfx_hair*> h;
        for (TressFXHairObject* o : hairObjects) h.push_back(o->handle());
        check(tfx_simulation_simulate(s_, stream, h.data(), (int)h.size()));
    }

    tfx_simulation* handle() const { return s_; }

private:
    tfx_simulation* s_ = nullptr;
};

/* TressFXSDFInputMeshInterface, TressFXSDFInputMeshInterface.h:33-57, with
 * device pointers in place of EI_Resource. */
class TressFXSDFInputMeshInterface {
public:
    virtual ~TressFXSDFInputMeshInterface() {}
    virtual const tfx_standard_vertex* GetMeshBuffer() = 0;          /* device */
    virtual const unsigned int* GetTrimeshVertexIndicesBuffer() = 0;  /* device, 3 per triangle */
    virtual int GetNumMeshVertices() = 0;
    virtual int GetNumMeshTriangle() = 0;
    virtual size_t GetSizeOfMeshElement() { return sizeof(tfx_standard_vertex); }
    virtual void GetBoundingBox(float min[3], float max[3]) = 0;
    virtual void GetInitialBoundingBox(float min[3], float max[3]) = 0;

    tfx_sdf_mesh_desc desc() {
        tfx_sdf_mesh_desc d{};
        d.vertices = GetMeshBuffer();
        d.indices = GetTrimeshVertexIndicesBuffer();
        d.num_vertices = GetNumMeshVertices();
        d.num_triangles = GetNumMeshTriangle();
        GetBoundingBox(d.bbox_min, d.bbox_max);
        GetInitialBoundingBox(d.initial_bbox_min, d.initial_bbox_max);
        return d;
    }
};

/* TressFXBoneSkinning (TressFXBoneSkinning.h:50-131): a collision mesh loaded
 * from a .tfxmesh and skinned on the GPU every Update; it is the mesh
 * interface an SDF reads. */
class TressFXBoneSkinning : public TressFXSDFInputMeshInterface {
public:
    TressFXBoneSkinning() = default;
    ~TressFXBoneSkinning() override { tfx_mesh_skinning_destroy(m_); }
    TressFXBoneSkinning(const TressFXBoneSkinning&) = delete;
    TressFXBoneSkinning& operator=(const TressFXBoneSkinning&) = delete;

    bool LoadTressFXCollisionMeshData(const char* tfxmeshPath, const std::vector<std::string>& engineBoneNames,
                                      const char* followBone) {
        std::vector<const char*> names;
        for (const std::string& n : engineBoneNames) names.push_back(n.c_str());
        tfx_collision_mesh* cm = nullptr;
        check(tfx_collision_mesh_load(tfxmeshPath, names.data(), (int)names.size(), followBone, &cm));
        tfx_skinned_mesh_desc d{};
        int rc = tfx_collision_mesh_get_desc(cm, &d);
        if (!rc) {
            tfx_mesh_skinning_destroy(m_);
            m_ = nullptr;
            rc = tfx_mesh_skinning_create(&d, &m_);  /* Initialize: device buffers */
        }
        tfx_collision_mesh_destroy(cm);
        check(rc);
        return true;
    }
    /* Update(commandContext, time): the world-space bone matrices of the frame. */
    void Update(void* stream, const float* boneMatrices, int numBones) {
        check(tfx_mesh_skinning_update(m_, boneMatrices, numBones, stream));
    }

    const tfx_standard_vertex* GetMeshBuffer() override { return view().vertices; }
    const unsigned int* GetTrimeshVertexIndicesBuffer() override { return view().indices; }
    int GetNumMeshVertices() override { return view().num_vertices; }
    int GetNumMeshTriangle() override { return view().num_triangles; }
    void GetBoundingBox(float min[3], float max[3]) override {
        const tfx_sdf_mesh_desc d = view();
        for (int a = 0; a < 3; ++a) { min[a] = d.bbox_min[a]; max[a] = d.bbox_max[a]; }
    }
    void GetInitialBoundingBox(float min[3], float max[3]) override {
        const tfx_sdf_mesh_desc d = view();
        for (int a = 0; a < 3; ++a) { min[a] = d.initial_bbox_min[a]; max[a] = d.initial_bbox_max[a]; }
    }

private:
    tfx_sdf_mesh_desc view() const {
        tfx_sdf_mesh_desc d{};
        check(tfx_mesh_skinning_get_mesh(m_, &d));
        return d;
    }
    tfx_mesh_skinning* m_ = nullptr;
};

/* TressFXSDFCollisionSystem, TressFXSDFCollision.h:40-56. */
class TressFXSDFCollisionSystem {
public:
    TressFXSDFCollisionSystem() { check(tfx_sdf_collision_system_create(&s_)); }
    ~TressFXSDFCollisionSystem() { tfx_sdf_collision_system_destroy(s_); }
    TressFXSDFCollisionSystem(const TressFXSDFCollisionSystem&) = delete;
    TressFXSDFCollisionSystem& operator=(const TressFXSDFCollisionSystem&) = delete;
    tfx_sdf_collision_system* handle() const { return s_; }

private:
    tfx_sdf_collision_system* s_ = nullptr;
};

/* TressFXSDFCollision, TressFXSDFCollision.h:58-124.  The mesh interface is
 * borrowed (not owned) and read at every Update, as the reference does. */
class TressFXSDFCollision {
public:
    TressFXSDFCollision(TressFXSDFInputMeshInterface* pCollMesh, int numCellsInX, float collisionMargin)
        : mesh_(pCollMesh) {
        const tfx_sdf_mesh_desc d = mesh_->desc();
        check(tfx_sdf_create(&d, numCellsInX, collisionMargin, &s_));
    }
    /* Explicit grid (no reference counterpart). */
    TressFXSDFCollision(TressFXSDFInputMeshInterface* pCollMesh, const float origin[3], float cellSize, int nx,
                        int ny, int nz, float collisionMargin)
        : mesh_(pCollMesh) {
        const tfx_sdf_mesh_desc d = mesh_->desc();
        check(tfx_sdf_create_explicit(&d, origin, cellSize, nx, ny, nz, collisionMargin, &s_));
    }
    ~TressFXSDFCollision() { tfx_sdf_destroy(s_); }
    TressFXSDFCollision(const TressFXSDFCollision&) = delete;
    TressFXSDFCollision& operator=(const TressFXSDFCollision&) = delete;

    void Update(void* stream, TressFXSDFCollisionSystem& system) {
        const tfx_sdf_mesh_desc d = mesh_->desc();
        check(tfx_sdf_set_mesh(s_, &d));
        check(tfx_sdf_update(s_, system.handle(), stream));
    }
    // TFX_SDF_COLLIDE_FORWARD (the reference's dispatch) or TFX_SDF_COLLIDE_MIXED.
    void SetCollideVariant(int variant) { check(tfx_sdf_set_collide_variant(s_, variant)); }
    void CollideWithHair(void* stream, TressFXSDFCollisionSystem& system, TressFXHairObject& hairObject) {
        check(tfx_sdf_collide_with_hair(s_, system.handle(), stream, hairObject.handle()));
    }
    float GetSDFCollisionMargin() const { return grid().collision_margin; }
    const float* GetSDFDataGPUBuffer() const {
        float* p = nullptr;
        check(tfx_sdf_get_data_buffer(s_, &p));
        return p;
    }
    float GetGridCellSize() const { return grid().cell_size; }
    void GetGridOrigin(float origin[3]) const {
        const tfx_sdf_grid_info g = grid();
        for (int a = 0; a < 3; ++a) origin[a] = g.origin[a];
    }
    void GetGridNumCells(int& x, int& y, int& z) const {
        const tfx_sdf_grid_info g = grid();
        x = g.num_cells_x;
        y = g.num_cells_y;
        z = g.num_cells_z;
    }
    int GetGridNumTotalCells() const { return grid().num_total_cells; }

    tfx_sdf* handle() const { return s_; }

private:
    tfx_sdf_grid_info grid() const {
        tfx_sdf_grid_info g{};
        check(tfx_sdf_get_grid(s_, &g));
        return g;
    }
    TressFXSDFInputMeshInterface* mesh_;
    tfx_sdf* s_ = nullptr;
};

}  // namespace tressfx_amd

#endif /* TRESSFX_AMD_HPP_ */
