/*
 * tressfx_amd.h — C-ABI of the MI355X-native TressFX 4.1 strand simulation and
 * SDF collision path.
 *
 * Every entry point returns an int status (TFX_OK == 0); on failure a
 * thread-local message is available from tfx_last_error().  No entry point
 * aborts.  Device pointers are plain `void*`/`float*` into HBM; `stream` is a
 * hipStream_t passed as `void*` (NULL = the default stream).  Calls on one
 * object are stream-ordered and not internally synchronised: use one stream per
 * object at a time.
 *
 * Data layout at the boundary is the reference's (renderer contract):
 *   - per-vertex streams are float4 AoS, strand-major (index = strand*V + v);
 *     follow strands are interleaved after their guide (strand = g*(F+1)+j);
 *   - w of a position is its inverse mass (0 = pinned);
 *   - bone matrices are row-major, row-vector convention (p' = [p,1]*M,
 *     translation in row 3), 16 floats each.
 * Citations are to files of the TressFX 4.1 reference tree.
 */
#ifndef TRESSFX_AMD_H_
#define TRESSFX_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TFX_SIM_THREAD_GROUP_SIZE 64       /* src/TressFX/TressFXCommon.h:34 */
#define TFX_MAX_NUM_BONES 512              /* src/TressFX/AMD_TressFX.h:45-47 */
#define TFX_MAX_NUM_COLLISION_CAPSULES 8   /* src/TressFX/TressFXCommon.h:32 */
#define TFX_MIN_VERTS_PER_STRAND_FOR_GPU_ITERATION 64 /* src/TressFX/TressFXHairObject.cpp:40 */

enum tfx_status {
    TFX_OK = 0,
    TFX_ERR_INVALID_ARGUMENT = 1,
    TFX_ERR_OUT_OF_MEMORY = 2,
    TFX_ERR_HIP = 3,
    TFX_ERR_ADMISSION = 4,
    TFX_ERR_NOT_FOUND = 5,
    TFX_ERR_UNSUPPORTED = 6
};

/* ------------------------------------------------------------------------ */
/* Records and parameter blocks                                              */
/* ------------------------------------------------------------------------ */

/* TressFXSimulationSettings, src/TressFX/TressFXSettings.h:35-87.
 * tfx_simulation_settings_default() fills the reference defaults (:38-56). */
typedef struct tfx_simulation_settings {
    float vsp_coeff;                      /* m_vspCoeff */
    float vsp_accel_threshold;            /* m_vspAccelThreshold */
    float local_constraint_stiffness;     /* m_localConstraintStiffness */
    int   local_constraints_iterations;   /* m_localConstraintsIterations */
    float global_constraint_stiffness;    /* m_globalConstraintStiffness */
    float global_constraints_range;       /* m_globalConstraintsRange */
    int   length_constraints_iterations;  /* m_lengthConstraintsIterations */
    float damping;                        /* m_damping */
    float gravity_magnitude;              /* m_gravityMagnitude */
    float tip_separation;                 /* m_tipSeparation */
    float wind_magnitude;                 /* m_windMagnitude */
    float wind_direction[3];              /* m_windDirection */
    float wind_angle_radians;             /* m_windAngleRadians (unused by the path) */
    float clamp_position_delta;           /* m_clampPositionDelta (ignored: 20 is hard-coded, TressFXHairObject.cpp:397) */
} tfx_simulation_settings;

/* TressFXSimulationParams, src/TressFX/TressFXConstantBuffers.h:41-110, i.e. the
 * HLSL cbuffer tressfxSimParameters (src/Shaders/TressFXSimulation.hlsl:49-79),
 * plus the capsule block of the TRESSFX_COLLISION_CAPSULES build (:72-76) and
 * the per-object host state the dispatch sequence reads
 * (m_CPULocalShapeIterations, TressFXHairObject.h:120).  Bone matrices are kept
 * apart (tfx_hair_update_bone_matrices). */
typedef struct tfx_simulation_params {
    float wind[4];            /* g_Wind  (w unused) */
    float wind1[4];           /* g_Wind1 */
    float wind2[4];           /* g_Wind2 */
    float wind3[4];           /* g_Wind3 */
    float shape[4];           /* damping, local stiffness, global stiffness, global range */
    float grav_time_tip[4];   /* gravity magnitude, time step, tip separation, unused */
    int   sim_ints[4];        /* length iterations, local iterations, collision flag, unused */
    int   counts[4];          /* strands per 64-thread group, follow per guide, vertices per strand, unused */
    float vsp[4];             /* vsp coefficient, vsp accel threshold, unused, unused */
    float reset_positions;    /* g_ResetPositions */
    float clamp_position_delta; /* g_ClampPositionDelta */
    float pad1;
    float pad2;
    float center_and_radius0[TFX_MAX_NUM_COLLISION_CAPSULES][4]; /* g_centerAndRadius0 */
    float center_and_radius1[TFX_MAX_NUM_COLLISION_CAPSULES][4]; /* g_centerAndRadius1 */
    int   num_collision_capsules[4];                             /* g_numCollisionCapsules.x */
    int   cpu_local_shape_iterations; /* LSC re-dispatch count, TressFXHairObject.cpp:402-411 */
    int   num_bones;                  /* matrices supplied by UpdateBoneMatrices */
    int   simulation_frame;           /* m_SimulationFrame at packing time */
    int   pad3;
} tfx_simulation_params;

/* TressFXBoneSkinningData, src/TressFX/TressFXAsset.h:42-46 (indices stored as float). */
typedef struct tfx_bone_skinning_data {
    float bone_index[4];
    float bone_weight[4];
} tfx_bone_skinning_data;

/* TressFXStrandLevelData, src/TressFX/TressFXHairObject.h:78-83. */
typedef struct tfx_strand_level_data {
    float skinning_quat[4];
    float vsp_quat[4];
    float vsp_translation[4]; /* xyz translation, w = VSP coefficient */
} tfx_strand_level_data;

/* StandardVertex, src/Shaders/TressFXSDFCollision.hlsl:35-39. */
typedef struct tfx_standard_vertex {
    float position[4];
    float normal[4];
} tfx_standard_vertex;

/* The host arrays of a processed TressFXAsset (src/TressFX/TressFXAsset.h:50-100):
 * what the TressFXHairObject constructor copies to the GPU
 * (TressFXHairObject.cpp:44-164).  All arrays are copied; the caller may free
 * them after tfx_hair_create returns. */
typedef struct tfx_asset_desc {
    int num_guide_strands;              /* m_numGuideStrands */
    int num_follow_strands_per_guide;   /* m_numFollowStrandsPerGuide */
    int num_vertices_per_strand;        /* m_numVerticesPerStrand: 4, 8, 16, 32 or 64 */
    int pad;
    const float* positions;             /* float4[S*V]  m_positions (also the rest pose) */
    const float* tangents;              /* float4[S*V]  m_tangents (may be NULL: zeros) */
    const float* rest_lengths;          /* float [S*V]  m_restLengths */
    const float* follow_root_offsets;   /* float4[S]    m_followRootOffsets (may be NULL if F == 0) */
    const tfx_bone_skinning_data* bone_skinning_data; /* [S] m_boneSkinningData */
} tfx_asset_desc;

typedef struct tfx_hair_info {
    int num_guide_strands;
    int num_follow_strands_per_guide;
    int num_vertices_per_strand;
    int num_total_strands;     /* GetNumTotalHairStrands */
    int num_total_vertices;    /* GetNumTotalHairVertices */
    int simulation_frame;      /* m_SimulationFrame */
    int cpu_local_shape_iterations; /* GetCPULocalShapeIterations */
    int pad;
} tfx_hair_info;

/* TressFXDynamicState (TressFXHairObject.h:54-76): device pointers. */
typedef struct tfx_dynamic_state {
    float* positions;           /* float4[N] */
    float* positions_prev;      /* float4[N] */
    float* positions_prev_prev; /* float4[N] */
    float* tangents;            /* float4[N] */
    tfx_strand_level_data* strand_level_data; /* [S] */
} tfx_dynamic_state;

/* What TressFXSDFInputMeshInterface (src/TressFX/TressFXSDFInputMeshInterface.h:33-57)
 * hands the SDF builder.  vertices/indices are borrowed device pointers that
 * must stay valid through every tfx_sdf_update that uses them. */
typedef struct tfx_sdf_mesh_desc {
    const tfx_standard_vertex* vertices;  /* GetMeshBuffer(), device */
    const uint32_t* indices;              /* GetTrimeshVertexIndicesBuffer(), device, 3 per triangle */
    int num_vertices;                     /* GetNumMeshVertices() */
    int num_triangles;                    /* GetNumMeshTriangle() */
    float bbox_min[3];                    /* GetBoundingBox() */
    float bbox_max[3];
    float initial_bbox_min[3];            /* GetInitialBoundingBox() */
    float initial_bbox_max[3];
} tfx_sdf_mesh_desc;

/* TressFXSDFCollisionParams, src/TressFX/TressFXConstantBuffers.h:119-134. */
typedef struct tfx_sdf_collision_params {
    float origin[4];
    float cell_size;
    int   num_cells_x;
    int   num_cells_y;
    int   num_cells_z;
    int   max_marching_cubes_vertices;
    float marching_cubes_isolevel;
    float collision_margin;          /* already scaled by the cell size */
    int   num_hair_vertices_per_strand;
    int   num_total_hair_vertices;
    float pad1, pad2, pad3;
} tfx_sdf_collision_params;

typedef struct tfx_sdf_grid_info {
    float origin[3];
    float cell_size;
    int   num_cells_x, num_cells_y, num_cells_z;
    int   num_total_cells;     /* allocation: 1.4 * X*Y*Z for the reference constructor */
    float collision_margin;    /* in cell units, as passed to the constructor */
    float padding_boundary;
    int   explicit_grid;
    int   pad;
} tfx_sdf_grid_info;

typedef struct tfx_hair tfx_hair;
typedef struct tfx_simulation tfx_simulation;
typedef struct tfx_sdf_collision_system tfx_sdf_collision_system;
typedef struct tfx_sdf tfx_sdf;
typedef struct tfx_batch tfx_batch;

/* ------------------------------------------------------------------------ */
/* Library                                                                   */
/* ------------------------------------------------------------------------ */
const char* tfx_last_error(void);
/* Per-pass GPU timing, the counterpart of the reference's
 * GetDevice()->GetTimeStamp(name) after each dispatch (TressFXSimulation.cpp:87-106,
 * TressFXSDFCollision.cpp:130-209).  While enabled, HIP events bracket every
 * kernel launch; tfx_profiling_read waits for and sums the launches recorded
 * under `label` or under any label extending it with '/' (a stage name,
 * "FusedSubstep" -- recorded per strand length and strand-block size as
 * "FusedSubstep/V16/B512" etc., so "FusedSubstep/V16" reads one length --,
 * "InitializeSignedDistanceField", "ConstructSignedDistanceField",
 * "FinalizeSignedDistanceField", "CollideHairVerticesWithSdf") and forgets them. */
int tfx_profiling_enable(int enable);

/* ---- Mesh-based strand transforms ---------------------------------------- */
/* HairToTriangleMapping and Transforms, TressFXSimulation.hlsl:146-158. */
typedef struct tfx_hair_to_triangle_mapping {
    uint32_t mesh_index;
    uint32_t triangle_index;     /* triangle of the (non-indexed) mesh: vertices 3t, 3t+1, 3t+2 */
    float barycentric[3];
    uint32_t reserved;
} tfx_hair_to_triangle_mapping;
typedef struct tfx_strand_transform {
    float tfm[16];               /* row-major float4x4, row-vector convention */
    float quat[4];               /* x, y, z, w */
} tfx_strand_transform;
/* GenerateTransforms (TressFXSimulation.hlsl:1035-1142, built when
 * USE_MESH_BASED_HAIR_TRANSFORM is 1): for every guide strand g, with
 * s = g * (F + 1), the transform taking the strand's root from the rest
 * triangle mapping[s] of `mesh_vertices` to the same triangle of
 * `skinned_vertices` (float4 per triangle corner, device memory), written to
 * transforms[s] (device memory, one record per strand). */
int tfx_hair_generate_transforms(tfx_hair* hair, const tfx_hair_to_triangle_mapping* mapping,
                                 const float* mesh_vertices, const float* skinned_vertices, int num_mesh_vertices,
                                 tfx_strand_transform* transforms, void* stream);
/* (A strand whose triangle lies past num_mesh_vertices is skipped: its record is not written.) */

/* ---- Rendering level of detail (host arithmetic, no device) -------------- */
/* The LOD fields of TressFXRenderingSettings (TressFXSettings.h:93-128). */
typedef struct tfx_rendering_lod_settings {
    float lod_start_distance;           /* m_LODStartDistance (1) */
    float lod_end_distance;             /* m_LODEndDistance (5) */
    float lod_percent;                  /* m_LODPercent (0.5) */
    float lod_width_multiplier;         /* m_LODWidthMultiplier (2) */
    float shadow_lod_start_distance;    /* m_ShadowLODStartDistance (1) */
    float shadow_lod_end_distance;      /* m_ShadowLODEndDistance (5) */
    float shadow_lod_percent;           /* m_ShadowLODPercent (0.5) */
    float shadow_lod_width_multiplier;  /* m_ShadowLODWidthMultiplier (2) */
    float fiber_radius;                 /* m_FiberRadius (0.0021) */
    int enable_hair_lod;                /* m_EnableHairLOD (0); gates the shadow LOD too, as in the reference */
} tfx_rendering_lod_settings;
void tfx_rendering_lod_settings_defaults(tfx_rendering_lod_settings* out);
/* The LOD part of TressFXHairObject::UpdateRenderingParameters
 * (TressFXHairObject.cpp:461-481): the strand fiber radius and the hair
 * density (m_LODHairDensity) at a camera distance, for the colour or the
 * shadow pass. */
int tfx_rendering_lod(const tfx_rendering_lod_settings* settings, float distance, int shadow_update,
                      float* fiber_radius, float* hair_density);
/* The index count of the strand draw at a hair density
 * (TressFXHairObject.cpp:260-277): total_indices / 3 primitives scaled by
 * the density, rounded up to whole strands and down to whole groups of a
 * guide and its follow strands.  total_indices = 6 * strands * (V - 1)
 * (TressFXAsset::GetNumHairTriangleIndices). */
int tfx_num_indices_to_render(long long total_indices, int vertices_per_strand, int follow_per_guide,
                              float hair_density, long long* num_indices);
int tfx_profiling_read(const char* label, double* total_ms, int* launches);
int tfx_get_version(int* major, int* minor, int* patch); /* AMD_TressFX.h:33-35 */
/* Selects the HIP device used by objects created afterwards on this thread. */
int tfx_init(int hip_device);
void tfx_simulation_settings_default(tfx_simulation_settings* out);

/* ------------------------------------------------------------------------ */
/* TressFXHairObject  (src/TressFX/TressFXHairObject.h:85-191)               */
/* ------------------------------------------------------------------------ */
/* TressFXHairObject(TressFXAsset*, ...) — TressFXHairObject.cpp:44-164.
 * Pos = Prev = PrevPrev = positions (TressFXDynamicState::UploadGPUData, :528-558),
 * then ResetPositions() (:157). Synchronises the device before returning. */
int tfx_hair_create(const tfx_asset_desc* asset, tfx_hair** out);
int tfx_hair_destroy(tfx_hair* hair);
/* UpdateBoneMatrices, TressFXHairObject.cpp:366-373 (count clamped to 512). */
int tfx_hair_update_bone_matrices(tfx_hair* hair, const float* bone_matrices_ws, int num_bone_matrices);
/* UpdateSimulationParameters, TressFXHairObject.cpp:380-437 (+ SetWind :191-235). */
int tfx_hair_update_simulation_parameters(tfx_hair* hair, const tfx_simulation_settings* settings, float time_step);
/* The capsule block of the TRESSFX_COLLISION_CAPSULES build (TressFXHairObject.cpp:422-431):
 * center_and_radius0/1 are float4[n] (xyz centre, w radius); n <= 8. */
int tfx_hair_set_collision_capsules(tfx_hair* hair, const float* center_and_radius0,
                                    const float* center_and_radius1, int num_capsules);
/* ResetPositions, TressFXHairObject.h:104. */
int tfx_hair_reset_positions(tfx_hair* hair);
/* IncreaseSimulationFrame, TressFXHairObject.h:129 (tfx_simulation_simulate calls it). */
int tfx_hair_increase_simulation_frame(tfx_hair* hair);
/* GetNumTotalHairVertices / Strands / VerticesPerStrand / FollowHairsPerGuideHair (:114-118). */
int tfx_hair_get_info(const tfx_hair* hair, tfx_hair_info* out);
/* The parameter block the next Simulate will use (m_SimCB[m_SimulationFrame % 2]). */
int tfx_hair_get_simulation_params(const tfx_hair* hair, tfx_simulation_params* out,
                                   float* bone_matrices_out /* 16*num_bones or NULL */);
/* GetDynamicState, TressFXHairObject.h:112. */
int tfx_hair_get_dynamic_state(tfx_hair* hair, tfx_dynamic_state* out);
/* TressFXDynamicState::TransitionSimToRendering / TransitionRenderingToSim
 * (TressFXHairObject.cpp:559-583): the resource barriers between the
 * simulation and the strand renderer.  Here a barrier between two streams:
 * work enqueued later on `to_stream` waits for everything enqueued so far on
 * `from_stream` (nothing to do when they are the same stream).  A renderer
 * then reads the Pos/Tangent buffers of tfx_hair_get_dynamic_state in place:
 * float4 per vertex, strand-major, follow strands after their guide, vertex
 * index = triangle index / 2 of the asset's index buffer (TressFXStrands.hlsl:33-35, 82-89). */
int tfx_hair_transition_sim_to_rendering(tfx_hair* hair, void* from_stream, void* to_stream);
int tfx_hair_transition_rendering_to_sim(tfx_hair* hair, void* from_stream, void* to_stream);
/* Snapshot / restore of the dynamic state (host arrays; NULL entries skipped).
 * Both are stream-ordered and synchronise `stream` before returning. */
int tfx_hair_download_state(tfx_hair* hair, void* stream, float* positions, float* positions_prev,
                            float* positions_prev_prev, float* tangents, float* strand_level_data);
int tfx_hair_upload_state(tfx_hair* hair, void* stream, const float* positions, const float* positions_prev,
                          const float* positions_prev_prev, const float* tangents,
                          const float* strand_level_data);

/* ------------------------------------------------------------------------ */
/* TressFXSimulation  (src/TressFX/TressFXSimulation.h:34-48)                */
/* ------------------------------------------------------------------------ */
/* Initialize, TressFXSimulation.cpp:34-43. */
int tfx_simulation_create(tfx_simulation** out);
int tfx_simulation_destroy(tfx_simulation* sim);
/* Simulate, TressFXSimulation.cpp:77-112: the reference's dispatch structure —
 * for each of the six stages, one launch per hair object (LocalShapeConstraints
 * repeated GetCPULocalShapeIterations() times), then IncreaseSimulationFrame. */
int tfx_simulation_simulate(tfx_simulation* sim, void* stream, tfx_hair* const* hairs, int num_hairs);
/* Runs a single stage (0 Integration+GlobalShape, 1 StrandLevelData, 2 VSP,
 * 3 LocalShapeConstraints (one pass), 4 LengthConstraintsWindAndCollision,
 * 5 UpdateFollowHairVertices) on one hair with its current parameters; for
 * stage-by-stage parity checks.  Does not advance the frame. */
int tfx_simulation_run_stage(tfx_simulation* sim, void* stream, tfx_hair* hair, int stage);
/* One substep of all six stages as a single fused kernel per hair (strand
 * blocks staged in LDS), followed by the hair's SDF collision if `sdf` is
 * non-NULL; then IncreaseSimulationFrame.  Same results as
 * tfx_simulation_simulate + tfx_sdf_collide_with_hair, except that PrevPrev is
 * maintained only for vertices 0 and 1 and StrandLevelData is not written. */
int tfx_simulation_simulate_fused(tfx_simulation* sim, void* stream, tfx_hair* const* hairs,
                                  tfx_sdf* const* sdfs, int num_hairs);
/* tfx_simulation_simulate_fused without any allocation, on the host or the
 * device.  No counterpart in the reference: it is for callers that cannot
 * allocate -- code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc.  Same results as tfx_simulation_simulate_fused, with
 * the SDF collision always inside the fused launches (no collide pass).  Its
 * tables (parameter blocks, object rows, bone matrices, work list) live in a
 * static device array of about 0.7 MiB per GPU, held until the process ends
 * and shared by every call on that GPU: calls that overlap on one GPU race,
 * and the array keeps the last call's contents.  Limits per call: 64 objects,
 * 2048 bone matrices (the matrices each object's skinning records can read),
 * 65,536 strand blocks; larger calls fail with TFX_ERR_INVALID_ARGUMENT.
 * tfx_init resolves the array's device address (else the first call does,
 * with the device current).  No pass timers. */
int tfx_simulation_simulate_fused_noalloc(tfx_simulation* sim, void* stream, tfx_hair* const* hairs,
                                          tfx_sdf* const* sdfs, int num_hairs);
/* Launch shape of the fused kernel for strands of `vertices_per_strand`
 * vertices: `variant` 0, 1, 2 = strand blocks of 512, 1024, 2048 guide
 * vertices (at most 128 strands), -1 = chosen per call (fewest rounds of
 * resident blocks); `blocks_per_cu` caps the persistent grid per CU for this
 * strand length (0 = the kernel's occupancy).  Each strand length keeps its
 * own shape.  Results do not depend on the shape. */
int tfx_simulation_set_launch_shape(tfx_simulation* sim, int vertices_per_strand, int variant, int blocks_per_cu);
/* As tfx_batch_set_sdf_deferral, for tfx_simulation_simulate_fused (values
 * TFX_SDF_DEFER_*, defined with the batched form below; default automatic). */
int tfx_simulation_set_sdf_deferral(tfx_simulation* sim, int mode);

/* ------------------------------------------------------------------------ */
/* TressFXSDFCollisionSystem / TressFXSDFCollision                          */
/* (src/TressFX/TressFXSDFCollision.h:40-124)                               */
/* ------------------------------------------------------------------------ */
int tfx_sdf_collision_system_create(tfx_sdf_collision_system** out); /* Initialize, .h:43-50 */
int tfx_sdf_collision_system_destroy(tfx_sdf_collision_system* sys);
/* TressFXSDFCollision(pDevice, pCollMesh, name, numCellsInX, collisionMargin),
 * TressFXSDFCollision.cpp:38-86: grid sized from the mesh's initial bounding box. */
int tfx_sdf_create(const tfx_sdf_mesh_desc* mesh, int num_cells_in_x, float collision_margin, tfx_sdf** out);
/* Explicit grid (origin fixed, cell size and dimensions given). */
int tfx_sdf_create_explicit(const tfx_sdf_mesh_desc* mesh, const float origin[3], float cell_size,
                            int num_cells_x, int num_cells_y, int num_cells_z,
                            float collision_margin, tfx_sdf** out);
int tfx_sdf_destroy(tfx_sdf* sdf);
/* Replace the mesh source (new device buffers and/or bounding box). */
int tfx_sdf_set_mesh(tfx_sdf* sdf, const tfx_sdf_mesh_desc* mesh);
/* Update, TressFXSDFCollision.cpp:95-169: origin from the current bbox, then
 * Initialize -> Construct -> Finalize. */
int tfx_sdf_update(tfx_sdf* sdf, tfx_sdf_collision_system* sys, void* stream);
/* Which collision kernel CollideWithHair runs (the compute entry point the
 * reference's TressFXSDFCollisionSystem binds, TressFXSDFCollision.h:49):
 * TFX_SDF_COLLIDE_FORWARD, the default, is CollideHairVerticesWithSdf_forward
 * (TressFXSDFCollision.hlsl:531-595); TFX_SDF_COLLIDE_MIXED is
 * CollideHairVerticesWithSdf (:597-715), forward/backward differences inside
 * one cube and no root-vertex test.  The fused and batched paths run the
 * forward kernel only and refuse a mixed SDF with TFX_ERR_UNSUPPORTED. */
#define TFX_SDF_COLLIDE_FORWARD 0
#define TFX_SDF_COLLIDE_MIXED 1
int tfx_sdf_set_collide_variant(tfx_sdf* sdf, int variant);
/* CollideWithHair, TressFXSDFCollision.cpp:171-221. */
int tfx_sdf_collide_with_hair(tfx_sdf* sdf, tfx_sdf_collision_system* sys, void* stream, tfx_hair* hair);
int tfx_sdf_get_grid(const tfx_sdf* sdf, tfx_sdf_grid_info* out);  /* grid getters, .h:76-93 */
/* GetSDFDataGPUBuffer, .h:71-72: device float[X*Y*Z] (bits as written by Finalize).
 * The grid is read-only to the caller: the batched build relies on its
 * previous contents, so a caller that writes it must tfx_sdf_update before the
 * next batched frame. */
int tfx_sdf_get_data_buffer(tfx_sdf* sdf, float** device_ptr);
int tfx_sdf_download(tfx_sdf* sdf, void* stream, float* host_cells /* X*Y*Z */);
/* The constant block CollideWithHair would upload for `hair` (TressFXSDFCollision.cpp:184-198). */
int tfx_sdf_get_collision_params(const tfx_sdf* sdf, const tfx_hair* hair, tfx_sdf_collision_params* out);

/* Arithmetic of the kernels.  No counterpart in the reference, whose HLSL
 * leaves the precision of division and square root to the driver.
 * TFX_ARITH_IEEE (default): correctly rounded float division and square root;
 *   every kernel equals the FP32 transcription bit for bit.
 * TFX_ARITH_FAST: the hardware reciprocal and square root (a few ulp) in the
 *   simulation stages, the fused substep and the collision lookups; results
 *   stay within the stepwise FP64 tolerance (1e-5 relative per substep).  The
 *   SDF build keeps correctly rounded arithmetic: its grid is the same in
 *   both modes.
 * Process-wide; applies to launches issued after the call (a batch re-captures
 * its frame graph). */
#define TFX_ARITH_IEEE 0
#define TFX_ARITH_FAST 1
int tfx_set_arithmetic(int mode);
int tfx_get_arithmetic(int* mode);

/* ------------------------------------------------------------------------ */
/* Asset loading (TressFXAsset, src/TressFX/TressFXAsset.h:48-120)           */
/* The reference's order: LoadHairData, GenerateFollowHairs, ProcessAsset,   */
/* LoadBoneData; then tfx_asset_get_desc feeds tfx_hair_create.              */
/* ------------------------------------------------------------------------ */
typedef struct tfx_asset tfx_asset;
/* LoadHairData, TressFXAsset.cpp:94-173: a .tfx file (version >= 4); guide
 * strands padded to a multiple of 64 by copies of the last strand. */
int tfx_asset_load_hair_data(const char* tfx_path, tfx_asset** out);
/* GenerateFollowHairs, TressFXAsset.cpp:177-248; the offsets' random numbers
 * come from a 32-bit LCG seeded with `seed` (the reference uses rand()). */
int tfx_asset_generate_follow_hairs(tfx_asset* asset, int num_follow_hairs_per_guide, float tip_separation_factor,
                                    float max_radius_around_guide_hair, unsigned int seed);
/* ProcessAsset, TressFXAsset.cpp:250-399: tangents, thickness coefficients,
 * rest lengths, triangle indices. */
int tfx_asset_process(tfx_asset* asset);
/* LoadBoneData, TressFXAsset.cpp:401-480: a .tfxbone file; its bone names are
 * mapped to engine bone ids = their index in engine_bone_names (GetBoneIdByName). */
int tfx_asset_load_bone_data(tfx_asset* asset, const char* tfxbone_path, const char* const* engine_bone_names,
                             int num_engine_bones);
/* The processed arrays (valid while the asset lives), for tfx_hair_create. */
int tfx_asset_get_desc(const tfx_asset* asset, tfx_asset_desc* out);
int tfx_asset_get_info(const tfx_asset* asset, int* file_strands, int* guide_strands, int* follow_per_guide,
                       int* vertices_per_strand);
/* Render-side arrays of ProcessAsset: strand UV float2[S], thickness float[N],
 * triangle indices int32[6*S*(V-1)]. */
int tfx_asset_get_render_data(const tfx_asset* asset, const float** strand_uv, const float** thickness_coeffs,
                              const int32_t** triangle_indices, int* num_triangle_indices);
int tfx_asset_destroy(tfx_asset* asset);

/* ------------------------------------------------------------------------ */
/* Collision-mesh skinning (TressFXBoneSkinning, src/TressFX/TressFXBoneSkinning.h:50-131) */
/* ------------------------------------------------------------------------ */
typedef struct tfx_mesh_skinning tfx_mesh_skinning;
/* The content of a loaded .tfxmesh (LoadTressFXCollisionMeshData,
 * TressFXBoneSkinning.cpp:100-286) with bone indices already mapped to engine
 * bone ids (GetBoneIdByName).  Arrays are copied. */
typedef struct tfx_skinned_mesh_desc {
    int num_vertices;
    int num_triangles;
    const float* positions;                /* float3[num_vertices] rest positions */
    const float* normals;                  /* float3[num_vertices] or NULL */
    const tfx_bone_skinning_data* skinning; /* [num_vertices] engine bone ids (as float), weights */
    const uint32_t* indices;               /* [3 * num_triangles] */
    int follow_bone;                       /* engine bone the bounding box follows (m_followBone) */
    int pad;
} tfx_skinned_mesh_desc;
/* Initialize, TressFXBoneSkinning.cpp:287-399 (device buffers, bounding cube). */
typedef struct tfx_collision_mesh tfx_collision_mesh;
/* LoadTressFXCollisionMeshData, TressFXBoneSkinning.cpp:100-286: a .tfxmesh
 * text file; bone names mapped to engine ids (index in engine_bone_names, a
 * name the skeleton lacks reads bone 0); follow_bone names the bone the
 * bounding box follows. */
int tfx_collision_mesh_load(const char* tfxmesh_path, const char* const* engine_bone_names, int num_engine_bones,
                            const char* follow_bone, tfx_collision_mesh** out);
int tfx_collision_mesh_get_desc(const tfx_collision_mesh* mesh, tfx_skinned_mesh_desc* out);
int tfx_collision_mesh_destroy(tfx_collision_mesh* mesh);
int tfx_mesh_skinning_create(const tfx_skinned_mesh_desc* mesh, tfx_mesh_skinning** out);
int tfx_mesh_skinning_destroy(tfx_mesh_skinning* mesh);
/* Update, TressFXBoneSkinning.cpp:460-504: new world-space bone matrices
 * (row-major row-vector float4x4), then BoneSkinning over every vertex. */
int tfx_mesh_skinning_update(tfx_mesh_skinning* mesh, const float* bone_matrices, int num_bones, void* stream);
/* The TressFXSDFInputMeshInterface view of the skinned mesh (.h:70-81):
 * device buffers, GetBoundingBox (:439-452) and GetInitialBoundingBox (:454-458),
 * for tfx_sdf_set_mesh. */
int tfx_mesh_skinning_get_mesh(const tfx_mesh_skinning* mesh, tfx_sdf_mesh_desc* out);
/* Copies the skinned vertex buffer (StandardVertex[num_vertices]) to the host. */
int tfx_mesh_skinning_download(tfx_mesh_skinning* mesh, void* stream, tfx_standard_vertex* host_vertices);

/* ------------------------------------------------------------------------ */
/* Batched form (no counterpart in the reference, whose Simulate dispatches  */
/* every stage once per object, TressFXSimulation.cpp:51-75): many objects   */
/* of uneven size advance in one persistent launch per substep and strand    */
/* length, each colliding with its own SDF; SDFs rebuilt by one launch per   */
/* build phase; the frame optionally captured into a hipGraph.               */
/* Per frame: tfx_batch_begin_frame, tfx_batch_stage_substep per substep,    */
/* tfx_batch_run_frame.  Objects are not internally synchronised: a member   */
/* must not be used through the per-object API while a frame is in flight.  */
/* ------------------------------------------------------------------------ */
int tfx_batch_create(tfx_batch** out);
int tfx_batch_destroy(tfx_batch* batch);
/* Adds an object (and the SDF it collides with, or NULL). */
int tfx_batch_add(tfx_batch* batch, tfx_hair* hair, tfx_sdf* sdf);
int tfx_batch_remove(tfx_batch* batch, tfx_hair* hair);
int tfx_batch_count(const tfx_batch* batch, int* num_members);
/* As tfx_simulation_set_launch_shape, for the batch's fused launches. */
int tfx_batch_set_launch_shape(tfx_batch* batch, int vertices_per_strand, int variant, int blocks_per_cu);
/* Where the batch runs the SDF collision (CollideHairVerticesWithSdf_forward,
 * TressFXSDFCollision.hlsl:531-595) of a strand-length class: inside its fused
 * launch, or in a collide pass launched after each substep's fused launches
 * (the reference's order: the substep's Simulate, then CollideWithHair).
 * Results are identical either way.  TFX_SDF_DEFER_AUTO (default) defers for
 * a class whose strand blocks fit in one round of resident blocks. */
#define TFX_SDF_DEFER_AUTO (-1)
#define TFX_SDF_DEFER_NEVER 0
#define TFX_SDF_DEFER_ALWAYS 1
int tfx_batch_set_sdf_deferral(tfx_batch* batch, int mode);
/* Autotuner of the batch's launch shapes (no counterpart in the reference).
 * Call with a frame staged (after tfx_batch_stage_substep, before
 * tfx_batch_run_frame), best after frames have run so that states and SDFs are
 * representative.  For every strand-length class, each strand-block size and
 * each persistent-blocks-per-CU count up to the kernel's occupancy runs the
 * class's fused launch of the staged first substep (and its collide pass,
 * for a shape that defers the SDF collision) `repeats` times from the
 * members' state (restored after every run: results are unaffected), and the
 * fastest is pinned for the class until tfx_batch_set_launch_shape changes
 * it.  One record per class (at most max_out written, *num_out = classes). */
typedef struct tfx_autotune_record {
    int vertices_per_strand;
    int variant;         /* chosen: 0, 1, 2 = strand blocks of 512, 1024, 2048 guide vertices */
    int blocks_per_cu;   /* chosen persistent blocks per CU */
    int candidates;      /* shapes timed */
    float chosen_us;     /* the chosen shape's fastest launch */
    float default_us;    /* the shape in use before the call (0 if not among the candidates) */
    int sdf_deferred;    /* the chosen shape defers the SDF collision (tfx_batch_set_sdf_deferral) */
} tfx_autotune_record;
int tfx_batch_autotune(tfx_batch* batch, void* stream, int repeats, tfx_autotune_record* out, int max_out,
                       int* num_out);
/* Takes a free staging slot and moves every member's SDF grid to its mesh's
 * current bounding box (meshes are set with tfx_sdf_set_mesh beforehand). */
int tfx_batch_begin_frame(tfx_batch* batch);
/* Stages one substep for every member, in member order: bone matrices
 * (member m's at bones + m*bones_stride floats, bones_stride/16 matrices; NULL
 * keeps the current ones), UpdateSimulationParameters with settings[m] (or
 * settings[0] when num_settings == 1) and time_step, wind_magnitude[m]
 * overriding the settings' magnitude (NULL: no override), then
 * IncreaseSimulationFrame.  At most 16 substeps per frame. */
int tfx_batch_stage_substep(tfx_batch* batch, const float* bones, int bones_stride,
                            const tfx_simulation_settings* settings, int num_settings, float time_step,
                            const float* wind_magnitude);
/* Runs the staged frame on `stream`: one parameter upload, the SDF build of
 * every member that has one, then the staged substeps (fused simulation +
 * collision).  use_graph != 0: captured once per slot, substep count and
 * membership, then replayed (needs a non-default stream). */
int tfx_batch_run_frame(tfx_batch* batch, void* stream, int use_graph);

/* ------------------------------------------------------------------------ */
/* Served form (no counterpart in the reference): a request loop over        */
/* resident instances with continuous batching.  A worker thread repeatedly   */
/* advances every instance that has queued frames by one frame in a single   */
/* batched pass (collision-mesh skinning, SDF builds, fused substeps);        */
/* requests that arrive while a pass runs join the next one.  Instances are   */
/* admitted against an HBM byte budget.  Instances belong to the server:     */
/* their objects may be read (tfx_server_get_objects) only while they have no */
/* queued or running frames.                                                  */
/* ------------------------------------------------------------------------ */
typedef struct tfx_server tfx_server;
typedef struct tfx_server_config {
    long long hbm_budget_bytes;  /* admission budget for instance state (hair, SDF, mesh buffers) */
    int max_instances;           /* resident instance limit */
    int substeps_per_frame;      /* substeps of every frame (1..16) */
    int device;                  /* HIP device of the worker */
    int max_inflight;            /* batched passes queued on the GPU at once (1..4) */
    float substep_dt;            /* time step of every substep (UpdateSimulationParameters' timeStep) */
    int pad;
} tfx_server_config;
typedef struct tfx_instance_desc {
    const tfx_asset_desc* asset;
    const tfx_skinned_mesh_desc* mesh;   /* collision mesh, skinned every frame; NULL = no SDF collision */
    float sdf_origin[3];                 /* explicit SDF grid when sdf_dims[0] > 0 ... */
    float sdf_cell;
    int sdf_dims[3];
    int sdf_num_cells_in_x;              /* ... else sized from the mesh (TressFXSDFCollision.cpp:38-86) */
    float collision_margin;              /* in cells */
    int substeps;                        /* substeps per frame of this instance, each frame/substeps long
                                            (frame = server substeps_per_frame x substep_dt); 0 = the
                                            server's substeps_per_frame */
    tfx_simulation_settings settings;
    const float* capsules0;              /* float4[num_capsules] (or NULL) */
    const float* capsules1;
    int num_capsules;
    int pad2;
} tfx_instance_desc;
typedef struct tfx_server_stats {
    long long completed_requests;
    long long completed_frames;
    long long vertex_substeps;       /* hair vertices x substeps advanced */
    long long passes;                /* batched passes launched */
    double mean_instances_per_pass;
    double latency_p50_ms;           /* submit -> completion of a request, since the last reset */
    double latency_p99_ms;
    double latency_mean_ms;
    double latency_max_ms;
    double elapsed_s;                /* wall time since the last reset */
    long long resident_bytes;
    int resident_instances;
    int queued_requests;
} tfx_server_stats;
int tfx_server_create(const tfx_server_config* config, tfx_server** out);
/* Drains queued work, then frees every instance. */
int tfx_server_destroy(tfx_server* server);
/* TFX_ERR_ADMISSION if the instance does not fit the budget or the instance
 * limit now (it may be retried once others retire). */
int tfx_server_admit(tfx_server* server, const tfx_instance_desc* desc, int* instance_id);
/* Frees the instance once its queued frames are done; later submits fail. */
int tfx_server_retire(tfx_server* server, int instance_id);
/* Queues `frames` frames for an instance, run in submission order:
 * hair_bones[frames][substeps][num_bones*16] (the hair object's bone
 * matrices for each of the instance's substeps),
 * mesh_bones[frames][num_mesh_bones*16] (the collision mesh's, per frame;
 * NULL keeps the current ones), wind[frames][substeps] (wind magnitude
 * overrides, or NULL).  ticket NULL:
 * nothing to poll; the request only counts in the statistics. */
int tfx_server_submit(tfx_server* server, int instance_id, int frames, const float* hair_bones, int num_bones,
                      const float* mesh_bones, int num_mesh_bones, const float* wind, unsigned long long* ticket);
/* done = 1 once every frame of the request has completed on the GPU;
 * latency_ms = completion - submission. */
int tfx_server_poll(tfx_server* server, unsigned long long ticket, int* done, double* latency_ms);
int tfx_server_wait(tfx_server* server, unsigned long long ticket, double* latency_ms);
/* As tfx_batch_set_sdf_deferral, for the passes the server launches from now
 * on.  The server's default is TFX_SDF_DEFER_NEVER: its passes are small (a
 * few instances), where the collide pass's extra launch costs as much as it
 * saves. */
int tfx_server_set_sdf_deferral(tfx_server* server, int mode);
/* Blocks until nothing is queued or running and the memory of retired
 * instances is released. */
int tfx_server_drain(tfx_server* server);
int tfx_server_stats_read(tfx_server* server, tfx_server_stats* out, int reset);
/* The instance's hair object and SDF (NULL if none), for inspection while it is idle. */
int tfx_server_get_objects(tfx_server* server, int instance_id, tfx_hair** hair, tfx_sdf** sdf);

#ifdef __cplusplus
}
#endif

#endif /* TRESSFX_AMD_H_ */
