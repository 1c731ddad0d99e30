"""MI355X-native TressFX 4.1 strand simulation and SDF collision.

Python mirror of the reference's engine interface over the C-ABI of
``include/tressfx_amd.h`` (the HIP library ``libtressfx_amd.so``):

    TressFXSimulation            src/TressFX/TressFXSimulation.h:34-48
    TressFXHairObject            src/TressFX/TressFXHairObject.h:85-191
    TressFXSDFCollisionSystem    src/TressFX/TressFXSDFCollision.h:40-56
    TressFXSDFCollision          src/TressFX/TressFXSDFCollision.h:58-124
    TressFXSimulationSettings    src/TressFX/TressFXSettings.h:35-87

Every compute call goes to the HIP library; if it is missing the import of the
library raises (there is no CPU fallback).
"""
import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libtressfx_amd.so")

TFX_OK = 0
TFX_MAX_NUM_BONES = 512
TFX_MAX_NUM_COLLISION_CAPSULES = 8

c_float_p = ctypes.POINTER(ctypes.c_float)


class TFXError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"tressfx_amd error {code}: {msg}")
        self.code = code


# ---------------------------------------------------------------------------
# ABI records (layouts of include/tressfx_amd.h)
# ---------------------------------------------------------------------------
class SimulationSettings(ctypes.Structure):
    _fields_ = [
        ("vsp_coeff", ctypes.c_float),
        ("vsp_accel_threshold", ctypes.c_float),
        ("local_constraint_stiffness", ctypes.c_float),
        ("local_constraints_iterations", ctypes.c_int),
        ("global_constraint_stiffness", ctypes.c_float),
        ("global_constraints_range", ctypes.c_float),
        ("length_constraints_iterations", ctypes.c_int),
        ("damping", ctypes.c_float),
        ("gravity_magnitude", ctypes.c_float),
        ("tip_separation", ctypes.c_float),
        ("wind_magnitude", ctypes.c_float),
        ("wind_direction", ctypes.c_float * 3),
        ("wind_angle_radians", ctypes.c_float),
        ("clamp_position_delta", ctypes.c_float),
    ]

    @classmethod
    def defaults(cls):
        """TressFXSimulationSettings() defaults, TressFXSettings.h:38-56."""
        s = cls()
        s.vsp_coeff = 0.8
        s.vsp_accel_threshold = 1.4
        s.local_constraint_stiffness = 0.9
        s.local_constraints_iterations = 2
        s.global_constraint_stiffness = 0.0
        s.global_constraints_range = 0.0
        s.length_constraints_iterations = 2
        s.damping = 0.08
        s.gravity_magnitude = 0.0
        s.tip_separation = 0.0
        s.wind_magnitude = 0.0
        s.wind_direction[0], s.wind_direction[1], s.wind_direction[2] = 1.0, 0.0, 0.0
        s.wind_angle_radians = np.float32(3.1415926 / 180.0 * 40.0)
        s.clamp_position_delta = 20.0
        return s


TressFXSimulationSettings = SimulationSettings


class SimulationParams(ctypes.Structure):
    _fields_ = [
        ("wind", ctypes.c_float * 4),
        ("wind1", ctypes.c_float * 4),
        ("wind2", ctypes.c_float * 4),
        ("wind3", ctypes.c_float * 4),
        ("shape", ctypes.c_float * 4),
        ("grav_time_tip", ctypes.c_float * 4),
        ("sim_ints", ctypes.c_int * 4),
        ("counts", ctypes.c_int * 4),
        ("vsp", ctypes.c_float * 4),
        ("reset_positions", ctypes.c_float),
        ("clamp_position_delta", ctypes.c_float),
        ("pad1", ctypes.c_float),
        ("pad2", ctypes.c_float),
        ("center_and_radius0", (ctypes.c_float * 4) * TFX_MAX_NUM_COLLISION_CAPSULES),
        ("center_and_radius1", (ctypes.c_float * 4) * TFX_MAX_NUM_COLLISION_CAPSULES),
        ("num_collision_capsules", ctypes.c_int * 4),
        ("cpu_local_shape_iterations", ctypes.c_int),
        ("num_bones", ctypes.c_int),
        ("simulation_frame", ctypes.c_int),
        ("pad3", ctypes.c_int),
    ]


class AssetDesc(ctypes.Structure):
    _fields_ = [
        ("num_guide_strands", ctypes.c_int),
        ("num_follow_strands_per_guide", ctypes.c_int),
        ("num_vertices_per_strand", ctypes.c_int),
        ("pad", ctypes.c_int),
        ("positions", c_float_p),
        ("tangents", c_float_p),
        ("rest_lengths", c_float_p),
        ("follow_root_offsets", c_float_p),
        ("bone_skinning_data", ctypes.c_void_p),
    ]


class HairInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "num_guide_strands", "num_follow_strands_per_guide", "num_vertices_per_strand", "num_total_strands",
        "num_total_vertices", "simulation_frame", "cpu_local_shape_iterations", "pad")]


class DynamicState(ctypes.Structure):
    _fields_ = [
        ("positions", ctypes.c_void_p),
        ("positions_prev", ctypes.c_void_p),
        ("positions_prev_prev", ctypes.c_void_p),
        ("tangents", ctypes.c_void_p),
        ("strand_level_data", ctypes.c_void_p),
    ]


class SdfMeshDesc(ctypes.Structure):
    _fields_ = [
        ("vertices", ctypes.c_void_p),
        ("indices", ctypes.c_void_p),
        ("num_vertices", ctypes.c_int),
        ("num_triangles", ctypes.c_int),
        ("bbox_min", ctypes.c_float * 3),
        ("bbox_max", ctypes.c_float * 3),
        ("initial_bbox_min", ctypes.c_float * 3),
        ("initial_bbox_max", ctypes.c_float * 3),
    ]


class SdfGridInfo(ctypes.Structure):
    _fields_ = [
        ("origin", ctypes.c_float * 3),
        ("cell_size", ctypes.c_float),
        ("num_cells_x", ctypes.c_int),
        ("num_cells_y", ctypes.c_int),
        ("num_cells_z", ctypes.c_int),
        ("num_total_cells", ctypes.c_int),
        ("collision_margin", ctypes.c_float),
        ("padding_boundary", ctypes.c_float),
        ("explicit_grid", ctypes.c_int),
        ("pad", ctypes.c_int),
    ]


class SdfCollisionParams(ctypes.Structure):
    _fields_ = [
        ("origin", ctypes.c_float * 4),
        ("cell_size", ctypes.c_float),
        ("num_cells_x", ctypes.c_int),
        ("num_cells_y", ctypes.c_int),
        ("num_cells_z", ctypes.c_int),
        ("max_marching_cubes_vertices", ctypes.c_int),
        ("marching_cubes_isolevel", ctypes.c_float),
        ("collision_margin", ctypes.c_float),
        ("num_hair_vertices_per_strand", ctypes.c_int),
        ("num_total_hair_vertices", ctypes.c_int),
        ("pad1", ctypes.c_float),
        ("pad2", ctypes.c_float),
        ("pad3", ctypes.c_float),
    ]


class SkinnedMeshDesc(ctypes.Structure):
    _fields_ = [
        ("num_vertices", ctypes.c_int),
        ("num_triangles", ctypes.c_int),
        ("positions", c_float_p),
        ("normals", c_float_p),
        ("skinning", ctypes.c_void_p),
        ("indices", ctypes.c_void_p),
        ("follow_bone", ctypes.c_int),
        ("pad", ctypes.c_int),
    ]


class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("hbm_budget_bytes", ctypes.c_longlong),
        ("max_instances", ctypes.c_int),
        ("substeps_per_frame", ctypes.c_int),
        ("device", ctypes.c_int),
        ("max_inflight", ctypes.c_int),
        ("substep_dt", ctypes.c_float),
        ("pad", ctypes.c_int),
    ]


class InstanceDesc(ctypes.Structure):
    _fields_ = [
        ("asset", ctypes.POINTER(AssetDesc)),
        ("mesh", ctypes.POINTER(SkinnedMeshDesc)),
        ("sdf_origin", ctypes.c_float * 3),
        ("sdf_cell", ctypes.c_float),
        ("sdf_dims", ctypes.c_int * 3),
        ("sdf_num_cells_in_x", ctypes.c_int),
        ("collision_margin", ctypes.c_float),
        ("substeps", ctypes.c_int),
        ("settings", SimulationSettings),
        ("capsules0", c_float_p),
        ("capsules1", c_float_p),
        ("num_capsules", ctypes.c_int),
        ("pad2", ctypes.c_int),
    ]


class HairToTriangleMapping(ctypes.Structure):
    """tfx_hair_to_triangle_mapping (HairToTriangleMapping, TressFXSimulation.hlsl:152-158)."""
    _fields_ = [("mesh_index", ctypes.c_uint32), ("triangle_index", ctypes.c_uint32),
                ("barycentric", ctypes.c_float * 3), ("reserved", ctypes.c_uint32)]


class StrandTransform(ctypes.Structure):
    """tfx_strand_transform (Transforms, TressFXSimulation.hlsl:146-150)."""
    _fields_ = [("tfm", ctypes.c_float * 16), ("quat", ctypes.c_float * 4)]


class RenderingLodSettings(ctypes.Structure):
    """tfx_rendering_lod_settings: the LOD fields of TressFXRenderingSettings
    (TressFXSettings.h:93-128)."""
    _fields_ = [
        ("lod_start_distance", ctypes.c_float),
        ("lod_end_distance", ctypes.c_float),
        ("lod_percent", ctypes.c_float),
        ("lod_width_multiplier", ctypes.c_float),
        ("shadow_lod_start_distance", ctypes.c_float),
        ("shadow_lod_end_distance", ctypes.c_float),
        ("shadow_lod_percent", ctypes.c_float),
        ("shadow_lod_width_multiplier", ctypes.c_float),
        ("fiber_radius", ctypes.c_float),
        ("enable_hair_lod", ctypes.c_int),
    ]

    @classmethod
    def defaults(cls):
        s = cls()
        load_library().tfx_rendering_lod_settings_defaults(ctypes.byref(s))
        return s


class AutotuneRecord(ctypes.Structure):
    """tfx_autotune_record: the launch shape the batch autotuner chose for one
    strand length, its fastest launch and that of the shape in use before."""
    _fields_ = [("vertices_per_strand", ctypes.c_int), ("variant", ctypes.c_int), ("blocks_per_cu", ctypes.c_int),
                ("candidates", ctypes.c_int), ("chosen_us", ctypes.c_float), ("default_us", ctypes.c_float),
                ("sdf_deferred", ctypes.c_int)]


class ServerStats(ctypes.Structure):
    _fields_ = [
        ("completed_requests", ctypes.c_longlong),
        ("completed_frames", ctypes.c_longlong),
        ("vertex_substeps", ctypes.c_longlong),
        ("passes", ctypes.c_longlong),
        ("mean_instances_per_pass", ctypes.c_double),
        ("latency_p50_ms", ctypes.c_double),
        ("latency_p99_ms", ctypes.c_double),
        ("latency_mean_ms", ctypes.c_double),
        ("latency_max_ms", ctypes.c_double),
        ("elapsed_s", ctypes.c_double),
        ("resident_bytes", ctypes.c_longlong),
        ("resident_instances", ctypes.c_int),
        ("queued_requests", ctypes.c_int),
    ]


# Symbols of include/tressfx_amd.h: name -> (restype, argtypes)
_VP = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float
TFX_SDF_COLLIDE_FORWARD = 0
TFX_SDF_COLLIDE_MIXED = 1
TFX_ARITH_IEEE = 0
TFX_ARITH_FAST = 1
TFX_SDF_DEFER_AUTO = -1
TFX_SDF_DEFER_NEVER = 0
TFX_SDF_DEFER_ALWAYS = 1

EXPORTS = {
    "tfx_last_error": (ctypes.c_char_p, []),
    "tfx_get_version": (_I, [ctypes.POINTER(_I)] * 3),
    "tfx_profiling_enable": (_I, [_I]),
    "tfx_profiling_read": (_I, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_I)]),
    "tfx_init": (_I, [_I]),
    "tfx_simulation_settings_default": (None, [ctypes.POINTER(SimulationSettings)]),
    "tfx_hair_create": (_I, [ctypes.POINTER(AssetDesc), ctypes.POINTER(_VP)]),
    "tfx_hair_destroy": (_I, [_VP]),
    "tfx_hair_update_bone_matrices": (_I, [_VP, c_float_p, _I]),
    "tfx_hair_update_simulation_parameters": (_I, [_VP, ctypes.POINTER(SimulationSettings), _F]),
    "tfx_hair_set_collision_capsules": (_I, [_VP, c_float_p, c_float_p, _I]),
    "tfx_hair_reset_positions": (_I, [_VP]),
    "tfx_hair_increase_simulation_frame": (_I, [_VP]),
    "tfx_hair_get_info": (_I, [_VP, ctypes.POINTER(HairInfo)]),
    "tfx_hair_get_simulation_params": (_I, [_VP, ctypes.POINTER(SimulationParams), c_float_p]),
    "tfx_hair_get_dynamic_state": (_I, [_VP, ctypes.POINTER(DynamicState)]),
    "tfx_sdf_set_collide_variant": (_I, [_VP, _I]),
    "tfx_set_arithmetic": (_I, [_I]),
    "tfx_get_arithmetic": (_I, [_VP]),
    "tfx_hair_generate_transforms": (_I, [_VP, _VP, _VP, _VP, _I, _VP, _VP]),
    "tfx_hair_transition_sim_to_rendering": (_I, [_VP, _VP, _VP]),
    "tfx_hair_transition_rendering_to_sim": (_I, [_VP, _VP, _VP]),
    "tfx_rendering_lod_settings_defaults": (None, [ctypes.POINTER(RenderingLodSettings)]),
    "tfx_rendering_lod": (_I, [ctypes.POINTER(RenderingLodSettings), ctypes.c_float, _I, ctypes.POINTER(ctypes.c_float),
                               ctypes.POINTER(ctypes.c_float)]),
    "tfx_num_indices_to_render": (_I, [ctypes.c_longlong, _I, _I, ctypes.c_float, ctypes.POINTER(ctypes.c_longlong)]),
    "tfx_hair_download_state": (_I, [_VP, _VP, c_float_p, c_float_p, c_float_p, c_float_p, c_float_p]),
    "tfx_hair_upload_state": (_I, [_VP, _VP, c_float_p, c_float_p, c_float_p, c_float_p, c_float_p]),
    "tfx_simulation_create": (_I, [ctypes.POINTER(_VP)]),
    "tfx_simulation_destroy": (_I, [_VP]),
    "tfx_simulation_simulate": (_I, [_VP, _VP, ctypes.POINTER(_VP), _I]),
    "tfx_simulation_run_stage": (_I, [_VP, _VP, _VP, _I]),
    "tfx_simulation_simulate_fused": (_I, [_VP, _VP, ctypes.POINTER(_VP), ctypes.POINTER(_VP), _I]),
    "tfx_simulation_simulate_fused_noalloc": (_I, [_VP, _VP, ctypes.POINTER(_VP), ctypes.POINTER(_VP), _I]),
    "tfx_simulation_set_launch_shape": (_I, [_VP, _I, _I, _I]),
    "tfx_sdf_collision_system_create": (_I, [ctypes.POINTER(_VP)]),
    "tfx_sdf_collision_system_destroy": (_I, [_VP]),
    "tfx_sdf_create": (_I, [ctypes.POINTER(SdfMeshDesc), _I, _F, ctypes.POINTER(_VP)]),
    "tfx_sdf_create_explicit": (_I, [ctypes.POINTER(SdfMeshDesc), c_float_p, _F, _I, _I, _I, _F, ctypes.POINTER(_VP)]),
    "tfx_sdf_destroy": (_I, [_VP]),
    "tfx_sdf_set_mesh": (_I, [_VP, ctypes.POINTER(SdfMeshDesc)]),
    "tfx_sdf_update": (_I, [_VP, _VP, _VP]),
    "tfx_sdf_collide_with_hair": (_I, [_VP, _VP, _VP, _VP]),
    "tfx_sdf_get_grid": (_I, [_VP, ctypes.POINTER(SdfGridInfo)]),
    "tfx_sdf_get_data_buffer": (_I, [_VP, ctypes.POINTER(_VP)]),
    "tfx_sdf_download": (_I, [_VP, _VP, c_float_p]),
    "tfx_sdf_get_collision_params": (_I, [_VP, _VP, ctypes.POINTER(SdfCollisionParams)]),
    "tfx_asset_load_hair_data": (_I, [ctypes.c_char_p, ctypes.POINTER(_VP)]),
    "tfx_asset_generate_follow_hairs": (_I, [_VP, _I, _F, _F, ctypes.c_uint]),
    "tfx_asset_process": (_I, [_VP]),
    "tfx_asset_load_bone_data": (_I, [_VP, ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p), _I]),
    "tfx_asset_get_desc": (_I, [_VP, ctypes.POINTER(AssetDesc)]),
    "tfx_asset_get_info": (_I, [_VP] + [ctypes.POINTER(_I)] * 4),
    "tfx_asset_get_render_data": (_I, [_VP, ctypes.POINTER(c_float_p), ctypes.POINTER(c_float_p),
                                       ctypes.POINTER(ctypes.POINTER(ctypes.c_int32)), ctypes.POINTER(_I)]),
    "tfx_asset_destroy": (_I, [_VP]),
    "tfx_collision_mesh_load": (_I, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p), _I, ctypes.c_char_p,
                                     ctypes.POINTER(_VP)]),
    "tfx_collision_mesh_get_desc": (_I, [_VP, ctypes.POINTER(SkinnedMeshDesc)]),
    "tfx_collision_mesh_destroy": (_I, [_VP]),
    "tfx_mesh_skinning_create": (_I, [ctypes.POINTER(SkinnedMeshDesc), ctypes.POINTER(_VP)]),
    "tfx_mesh_skinning_destroy": (_I, [_VP]),
    "tfx_mesh_skinning_update": (_I, [_VP, c_float_p, _I, _VP]),
    "tfx_mesh_skinning_get_mesh": (_I, [_VP, ctypes.POINTER(SdfMeshDesc)]),
    "tfx_mesh_skinning_download": (_I, [_VP, _VP, _VP]),
    "tfx_server_create": (_I, [ctypes.POINTER(ServerConfig), ctypes.POINTER(_VP)]),
    "tfx_server_destroy": (_I, [_VP]),
    "tfx_server_admit": (_I, [_VP, ctypes.POINTER(InstanceDesc), ctypes.POINTER(_I)]),
    "tfx_server_retire": (_I, [_VP, _I]),
    "tfx_server_submit": (_I, [_VP, _I, _I, c_float_p, _I, c_float_p, _I, c_float_p,
                               ctypes.POINTER(ctypes.c_ulonglong)]),
    "tfx_server_poll": (_I, [_VP, ctypes.c_ulonglong, ctypes.POINTER(_I), ctypes.POINTER(ctypes.c_double)]),
    "tfx_server_wait": (_I, [_VP, ctypes.c_ulonglong, ctypes.POINTER(ctypes.c_double)]),
    "tfx_server_drain": (_I, [_VP]),
    "tfx_server_set_sdf_deferral": (_I, [_VP, _I]),
    "tfx_server_stats_read": (_I, [_VP, ctypes.POINTER(ServerStats), _I]),
    "tfx_server_get_objects": (_I, [_VP, _I, ctypes.POINTER(_VP), ctypes.POINTER(_VP)]),
    "tfx_batch_create": (_I, [ctypes.POINTER(_VP)]),
    "tfx_batch_destroy": (_I, [_VP]),
    "tfx_batch_add": (_I, [_VP, _VP, _VP]),
    "tfx_batch_remove": (_I, [_VP, _VP]),
    "tfx_batch_count": (_I, [_VP, ctypes.POINTER(_I)]),
    "tfx_batch_set_launch_shape": (_I, [_VP, _I, _I, _I]),
    "tfx_batch_set_sdf_deferral": (_I, [_VP, _I]),
    "tfx_simulation_set_sdf_deferral": (_I, [_VP, _I]),
    "tfx_batch_autotune": (_I, [_VP, _VP, _I, _VP, _I, _VP]),
    "tfx_batch_begin_frame": (_I, [_VP]),
    "tfx_batch_stage_substep": (_I, [_VP, c_float_p, _I, ctypes.POINTER(SimulationSettings), _I, _F, c_float_p]),
    "tfx_batch_run_frame": (_I, [_VP, _VP, _I]),
}

_lib = None


def load_library(path=LIB_PATH):
    """Loads libtressfx_amd.so and binds every exported symbol. Raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built: run `python -m tressfx_amd.build` (hipcc, gfx950)")
    lib = ctypes.CDLL(path)
    for name, (res, args) in EXPORTS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(rc):
    if rc != TFX_OK:
        msg = _lib.tfx_last_error()
        raise TFXError(rc, msg.decode() if msg else "")


def _fptr(a):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(c_float_p)


def init(device=0):
    lib = load_library()
    _check(lib.tfx_init(device))
    return lib


def profiling_enable(on=True):
    _check(load_library().tfx_profiling_enable(1 if on else 0))


def set_arithmetic(mode):
    """TFX_ARITH_IEEE (default: correctly rounded division and square root,
    bit-exact with the FP32 transcription) or TFX_ARITH_FAST (the hardware
    reciprocal and square root, within the FP64 tolerance), for every launch
    issued afterwards."""
    _check(load_library().tfx_set_arithmetic(int(mode)))


def get_arithmetic():
    m = ctypes.c_int()
    _check(load_library().tfx_get_arithmetic(ctypes.byref(m)))
    return m.value


def rendering_lod(settings, distance, shadow_update=False):
    """(fiber radius, hair density) at a camera distance (the LOD part of
    UpdateRenderingParameters, TressFXHairObject.cpp:461-481)."""
    r, d = ctypes.c_float(), ctypes.c_float()
    _check(load_library().tfx_rendering_lod(ctypes.byref(settings), ctypes.c_float(distance), 1 if shadow_update else 0,
                                            ctypes.byref(r), ctypes.byref(d)))
    return r.value, d.value


def num_indices_to_render(total_indices, vertices_per_strand, follow_per_guide, hair_density):
    """Index count of the strand draw at a hair density (TressFXHairObject.cpp:260-277)."""
    out = ctypes.c_longlong()
    _check(load_library().tfx_num_indices_to_render(int(total_indices), int(vertices_per_strand),
                                                    int(follow_per_guide), ctypes.c_float(hair_density),
                                                    ctypes.byref(out)))
    return out.value


def profiling_read(label):
    """(total milliseconds, launches) recorded under `label` since the last read."""
    ms = ctypes.c_double()
    n = ctypes.c_int()
    _check(load_library().tfx_profiling_read(label.encode(), ctypes.byref(ms), ctypes.byref(n)))
    return ms.value, n.value


# ---------------------------------------------------------------------------
# Host asset arrays (TressFXAsset, src/TressFX/TressFXAsset.h:50-100)
# ---------------------------------------------------------------------------
class HairAsset:
    """Processed asset arrays: positions/tangents float32[S*V,4], rest_lengths
    float32[S*V], follow_root_offsets float32[S,4], bone_skinning_data
    float32[S,8] (4 indices as float, 4 weights)."""

    def __init__(self, num_guide_strands, num_follow_strands_per_guide, num_vertices_per_strand, positions,
                 tangents, rest_lengths, follow_root_offsets, bone_skinning_data):
        self.num_guide_strands = int(num_guide_strands)
        self.num_follow_strands_per_guide = int(num_follow_strands_per_guide)
        self.num_vertices_per_strand = int(num_vertices_per_strand)
        self.positions = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 4)
        self.tangents = np.ascontiguousarray(tangents, dtype=np.float32).reshape(-1, 4)
        self.rest_lengths = np.ascontiguousarray(rest_lengths, dtype=np.float32).reshape(-1)
        self.follow_root_offsets = np.ascontiguousarray(follow_root_offsets, dtype=np.float32).reshape(-1, 4)
        self.bone_skinning_data = np.ascontiguousarray(bone_skinning_data, dtype=np.float32).reshape(-1, 8)

    @property
    def num_total_strands(self):
        return self.num_guide_strands * (self.num_follow_strands_per_guide + 1)

    @property
    def num_total_vertices(self):
        return self.num_total_strands * self.num_vertices_per_strand

    def desc(self):
        d = AssetDesc()
        d.num_guide_strands = self.num_guide_strands
        d.num_follow_strands_per_guide = self.num_follow_strands_per_guide
        d.num_vertices_per_strand = self.num_vertices_per_strand
        d.positions = _fptr(self.positions)
        d.tangents = _fptr(self.tangents)
        d.rest_lengths = _fptr(self.rest_lengths)
        d.follow_root_offsets = _fptr(self.follow_root_offsets)
        d.bone_skinning_data = self.bone_skinning_data.ctypes.data
        return d


# ---------------------------------------------------------------------------
# Engine-interface mirror
# ---------------------------------------------------------------------------
class TressFXHairObject:
    """TressFXHairObject(TressFXAsset*, ...), TressFXHairObject.cpp:44-164."""

    def __init__(self, asset: HairAsset):
        self._lib = load_library()
        h = ctypes.c_void_p()
        desc = asset.desc()
        _check(self._lib.tfx_hair_create(ctypes.byref(desc), ctypes.byref(h)))
        self.handle = h
        self.num_guide_strands = asset.num_guide_strands
        self.num_follow = asset.num_follow_strands_per_guide
        self.V = asset.num_vertices_per_strand

    def close(self):
        if self.handle:
            self._lib.tfx_hair_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def UpdateBoneMatrices(self, matrices):
        m = np.ascontiguousarray(matrices, dtype=np.float32).reshape(-1, 16)
        _check(self._lib.tfx_hair_update_bone_matrices(self.handle, _fptr(m), m.shape[0]))

    def UpdateSimulationParameters(self, settings: SimulationSettings, time_step):
        _check(self._lib.tfx_hair_update_simulation_parameters(self.handle, ctypes.byref(settings),
                                                               ctypes.c_float(time_step)))

    def SetCollisionCapsules(self, center_and_radius0, center_and_radius1):
        c0 = np.ascontiguousarray(center_and_radius0, dtype=np.float32).reshape(-1, 4)
        c1 = np.ascontiguousarray(center_and_radius1, dtype=np.float32).reshape(-1, 4)
        assert c0.shape == c1.shape
        _check(self._lib.tfx_hair_set_collision_capsules(self.handle, _fptr(c0), _fptr(c1), c0.shape[0]))

    def ResetPositions(self):
        _check(self._lib.tfx_hair_reset_positions(self.handle))

    def IncreaseSimulationFrame(self):
        _check(self._lib.tfx_hair_increase_simulation_frame(self.handle))

    def info(self):
        o = HairInfo()
        _check(self._lib.tfx_hair_get_info(self.handle, ctypes.byref(o)))
        return o

    def GetNumTotalHairVertices(self):
        return self.info().num_total_vertices

    def GetNumTotalHairStrands(self):
        return self.info().num_total_strands

    def GetNumVerticesPerStrand(self):
        return self.V

    def GetNumFollowHairsPerGuideHair(self):
        return self.num_follow

    def GetCPULocalShapeIterations(self):
        return self.info().cpu_local_shape_iterations

    def simulation_params(self):
        """The packed parameter block and the 512-entry bone table."""
        p = SimulationParams()
        bones = np.zeros((TFX_MAX_NUM_BONES, 16), dtype=np.float32)
        _check(self._lib.tfx_hair_get_simulation_params(self.handle, ctypes.byref(p), _fptr(bones)))
        return p, bones

    def GetDynamicState(self):
        o = DynamicState()
        _check(self._lib.tfx_hair_get_dynamic_state(self.handle, ctypes.byref(o)))
        return o

    def GenerateTransforms(self, mapping_ptr, mesh_vertices_ptr, skinned_vertices_ptr, num_mesh_vertices,
                           transforms_ptr, stream=None):
        """GenerateTransforms (TressFXSimulation.hlsl:1035-1142) on device buffers."""
        _check(self._lib.tfx_hair_generate_transforms(self.handle, mapping_ptr, mesh_vertices_ptr,
                                                       skinned_vertices_ptr, int(num_mesh_vertices),
                                                       transforms_ptr, stream))

    def TransitionSimToRendering(self, from_stream, to_stream):
        """Work enqueued later on to_stream waits for from_stream (TressFXHairObject.cpp:559-569)."""
        _check(self._lib.tfx_hair_transition_sim_to_rendering(self.handle, from_stream, to_stream))

    def TransitionRenderingToSim(self, from_stream, to_stream):
        """The barrier back (TressFXHairObject.cpp:571-583)."""
        _check(self._lib.tfx_hair_transition_rendering_to_sim(self.handle, from_stream, to_stream))

    def download_state(self, stream=None, keys=("pos", "prev", "prevprev", "tangent", "strand_level")):
        """Host copies of the dynamic state arrays named in `keys`."""
        n = self.GetNumTotalHairVertices()
        s = self.GetNumTotalHairStrands()
        st = {k: np.zeros((n, 4), np.float32) for k in ("pos", "prev", "prevprev", "tangent") if k in keys}
        if "strand_level" in keys:
            st["strand_level"] = np.zeros((s, 12), np.float32)
        p = lambda k: _fptr(st[k]) if k in st else None
        _check(self._lib.tfx_hair_download_state(self.handle, stream, p("pos"), p("prev"), p("prevprev"),
                                                 p("tangent"), p("strand_level")))
        return st

    def upload_state(self, state, stream=None):
        keep = {k: np.ascontiguousarray(state[k], np.float32) for k in state if state[k] is not None}
        p = lambda k: _fptr(keep[k]) if k in keep else None
        _check(self._lib.tfx_hair_upload_state(self.handle, stream, p("pos"), p("prev"), p("prevprev"),
                                               p("tangent"), p("strand_level")))


class TressFXSimulation:
    """TressFXSimulation, TressFXSimulation.h:34-48."""

    def __init__(self):
        self._lib = load_library()
        h = ctypes.c_void_p()
        _check(self._lib.tfx_simulation_create(ctypes.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            self._lib.tfx_simulation_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def Simulate(self, hair_objects, stream=None):
        arr = (ctypes.c_void_p * len(hair_objects))(*[h.handle for h in hair_objects])
        _check(self._lib.tfx_simulation_simulate(self.handle, stream, arr, len(hair_objects)))

    def RunStage(self, hair, stage, stream=None):
        _check(self._lib.tfx_simulation_run_stage(self.handle, stream, hair.handle, stage))

    def SetLaunchShape(self, vertices_per_strand, variant, blocks_per_cu=0):
        _check(self._lib.tfx_simulation_set_launch_shape(self.handle, vertices_per_strand, variant, blocks_per_cu))

    def set_sdf_deferral(self, mode):
        """TFX_SDF_DEFER_*: SDF collision inside SimulateFused's fused launches
        or as a collide pass after them (same results)."""
        _check(self._lib.tfx_simulation_set_sdf_deferral(self.handle, int(mode)))

    def SimulateFused(self, hair_objects, sdfs=None, stream=None):
        n = len(hair_objects)
        arr = (ctypes.c_void_p * n)(*[h.handle for h in hair_objects])
        sarr = None
        if sdfs is not None:
            sarr = (ctypes.c_void_p * n)(*[(s.handle if s is not None else None) for s in sdfs])
        _check(self._lib.tfx_simulation_simulate_fused(self.handle, stream, arr, sarr, n))

    def SimulateFusedNoAlloc(self, hair_objects, sdfs=None, stream=None):
        """SimulateFused through tfx_simulation_simulate_fused_noalloc (no allocation; its limits are
        64 objects, 2048 bone matrices and 65,536 strand blocks per call)."""
        n = len(hair_objects)
        arr = (ctypes.c_void_p * n)(*[h.handle for h in hair_objects])
        sarr = None
        if sdfs is not None:
            sarr = (ctypes.c_void_p * n)(*[(s.handle if s is not None else None) for s in sdfs])
        _check(self._lib.tfx_simulation_simulate_fused_noalloc(self.handle, stream, arr, sarr, n))


class TressFXSDFCollisionSystem:
    def __init__(self):
        self._lib = load_library()
        h = ctypes.c_void_p()
        _check(self._lib.tfx_sdf_collision_system_create(ctypes.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            self._lib.tfx_sdf_collision_system_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TressFXSDFCollision:
    """TressFXSDFCollision, TressFXSDFCollision.h:58-124.  `mesh` is an
    SdfMeshDesc whose vertex/index pointers are device buffers."""

    def __init__(self, mesh: SdfMeshDesc, num_cells_in_x=None, collision_margin=0.0, explicit=None):
        self._lib = load_library()
        h = ctypes.c_void_p()
        if explicit is not None:
            origin, cell, dims = explicit
            o = np.asarray(origin, np.float32)
            _check(self._lib.tfx_sdf_create_explicit(ctypes.byref(mesh), _fptr(o), ctypes.c_float(cell),
                                                     int(dims[0]), int(dims[1]), int(dims[2]),
                                                     ctypes.c_float(collision_margin), ctypes.byref(h)))
        else:
            _check(self._lib.tfx_sdf_create(ctypes.byref(mesh), int(num_cells_in_x), ctypes.c_float(collision_margin),
                                            ctypes.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            self._lib.tfx_sdf_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def SetMesh(self, mesh: SdfMeshDesc):
        _check(self._lib.tfx_sdf_set_mesh(self.handle, ctypes.byref(mesh)))

    def Update(self, system, stream=None):
        _check(self._lib.tfx_sdf_update(self.handle, system.handle, stream))

    def SetCollideVariant(self, variant):
        """TFX_SDF_COLLIDE_FORWARD (default) or TFX_SDF_COLLIDE_MIXED."""
        _check(self._lib.tfx_sdf_set_collide_variant(self.handle, int(variant)))

    def CollideWithHair(self, system, hair, stream=None):
        _check(self._lib.tfx_sdf_collide_with_hair(self.handle, system.handle, stream, hair.handle))

    def grid(self):
        g = SdfGridInfo()
        _check(self._lib.tfx_sdf_get_grid(self.handle, ctypes.byref(g)))
        return g

    def collision_params(self, hair):
        p = SdfCollisionParams()
        _check(self._lib.tfx_sdf_get_collision_params(self.handle, hair.handle, ctypes.byref(p)))
        return p

    def download(self, stream=None):
        g = self.grid()
        out = np.zeros(g.num_cells_x * g.num_cells_y * g.num_cells_z, np.float32)
        _check(self._lib.tfx_sdf_download(self.handle, stream, _fptr(out)))
        return out


class TressFXBoneSkinning:
    """TressFXBoneSkinning (src/TressFX/TressFXBoneSkinning.h:50-131): a
    collision mesh skinned on the GPU every Update; `mesh()` is the
    TressFXSDFInputMeshInterface view an SDF reads.
    positions/normals float32[n,3], skinning float32[n,8] (engine bone ids as
    float, then weights), indices uint32[t,3]."""

    def __init__(self, positions, skinning, indices, normals=None, follow_bone=0):
        self._lib = load_library()
        self._pos = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
        self._nrm = None if normals is None else np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
        self._skin = np.ascontiguousarray(skinning, np.float32).reshape(-1, 8)
        self._idx = np.ascontiguousarray(indices, np.uint32).reshape(-1, 3)
        d = SkinnedMeshDesc()
        d.num_vertices = self._pos.shape[0]
        d.num_triangles = self._idx.shape[0]
        d.positions = _fptr(self._pos)
        d.normals = _fptr(self._nrm) if self._nrm is not None else None
        d.skinning = self._skin.ctypes.data
        d.indices = self._idx.ctypes.data
        d.follow_bone = int(follow_bone)
        h = ctypes.c_void_p()
        _check(self._lib.tfx_mesh_skinning_create(ctypes.byref(d), ctypes.byref(h)))
        self.handle = h
        self.num_vertices = d.num_vertices

    def close(self):
        if self.handle:
            self._lib.tfx_mesh_skinning_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def Update(self, bone_matrices, stream=None):
        m = np.ascontiguousarray(bone_matrices, np.float32).reshape(-1, 16)
        _check(self._lib.tfx_mesh_skinning_update(self.handle, _fptr(m), m.shape[0], stream))

    def mesh(self):
        d = SdfMeshDesc()
        _check(self._lib.tfx_mesh_skinning_get_mesh(self.handle, ctypes.byref(d)))
        return d

    def download(self, stream=None):
        """The skinned StandardVertex buffer as float32[n, 8] (synchronous)."""
        out = np.zeros((self.num_vertices, 8), np.float32)
        _check(self._lib.tfx_mesh_skinning_download(self.handle, stream, out.ctypes.data))
        return out


class HairBatch:
    """The batched form over many TressFXHairObject members (tfx_batch_*)."""

    def __init__(self):
        self._lib = load_library()
        h = ctypes.c_void_p()
        _check(self._lib.tfx_batch_create(ctypes.byref(h)))
        self.handle = h
        self.members = []

    def close(self):
        if self.handle:
            self._lib.tfx_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, hair, sdf=None):
        _check(self._lib.tfx_batch_add(self.handle, hair.handle, sdf.handle if sdf is not None else None))
        self.members.append((hair, sdf))

    def remove(self, hair):
        _check(self._lib.tfx_batch_remove(self.handle, hair.handle))
        self.members = [m for m in self.members if m[0] is not hair]

    def __len__(self):
        n = ctypes.c_int()
        _check(self._lib.tfx_batch_count(self.handle, ctypes.byref(n)))
        return n.value

    def set_launch_shape(self, vertices_per_strand, variant, blocks_per_cu=0):
        _check(self._lib.tfx_batch_set_launch_shape(self.handle, vertices_per_strand, variant, blocks_per_cu))

    def set_sdf_deferral(self, mode):
        """TFX_SDF_DEFER_AUTO, _NEVER or _ALWAYS: SDF collision inside the
        fused launches or in a collide pass after them (same results)."""
        _check(self._lib.tfx_batch_set_sdf_deferral(self.handle, int(mode)))

    def autotune(self, stream=None, repeats=2):
        """Times every launch shape of each strand length on the staged frame's
        first substep and pins the fastest (results are unaffected); call
        between stage_substep and run_frame.  Returns one dict per strand length."""
        recs = (AutotuneRecord * 8)()
        n = ctypes.c_int()
        _check(self._lib.tfx_batch_autotune(self.handle, stream, int(repeats), recs, 8, ctypes.byref(n)))
        return [{f: getattr(recs[i], f) for f, _ in AutotuneRecord._fields_} for i in range(min(n.value, 8))]

    def begin_frame(self):
        _check(self._lib.tfx_batch_begin_frame(self.handle))

    def stage_substep(self, bones, settings, time_step, wind_magnitude=None):
        """bones: float32 [members, k*16] (or None); settings: one
        SimulationSettings or a ctypes array of one per member."""
        bp, stride = None, 0
        if bones is not None:
            b = np.ascontiguousarray(bones, np.float32)
            b = b.reshape(b.shape[0], -1)
            bp, stride = _fptr(b), b.shape[1]
            self._keep = b
        if isinstance(settings, SimulationSettings):
            sp, ns = ctypes.pointer(settings), 1
        else:
            sp, ns = ctypes.cast(settings, ctypes.POINTER(SimulationSettings)), len(settings)
        wp = None
        if wind_magnitude is not None:
            w = np.ascontiguousarray(wind_magnitude, np.float32)
            wp = _fptr(w)
            self._keepw = w
        _check(self._lib.tfx_batch_stage_substep(self.handle, bp, stride, sp, ns, ctypes.c_float(time_step), wp))

    def run_frame(self, stream=None, use_graph=False):
        _check(self._lib.tfx_batch_run_frame(self.handle, stream, 1 if use_graph else 0))


class _BorrowedHair(TressFXHairObject):
    """A hair object owned by someone else (the server): same accessors, no destroy."""

    def __init__(self, handle):
        self._lib = load_library()
        self.handle = ctypes.c_void_p(handle)
        info = self.info()
        self.num_guide_strands = info.num_guide_strands
        self.num_follow = info.num_follow_strands_per_guide
        self.V = info.num_vertices_per_strand

    def close(self):
        self.handle = None


class HairServer:
    """The served form (tfx_server_*): continuous batching over resident instances."""

    def __init__(self, hbm_budget_bytes=200 << 30, max_instances=1000, substeps_per_frame=2,
                 substep_dt=1.0 / 120.0, max_inflight=2, device=0):
        self._lib = load_library()
        c = ServerConfig()
        c.hbm_budget_bytes = int(hbm_budget_bytes)
        c.max_instances = int(max_instances)
        c.substeps_per_frame = int(substeps_per_frame)
        c.device = int(device)
        c.max_inflight = int(max_inflight)
        c.substep_dt = float(substep_dt)
        h = ctypes.c_void_p()
        _check(self._lib.tfx_server_create(ctypes.byref(c), ctypes.byref(h)))
        self.handle = h
        self.substeps = int(substeps_per_frame)
        self._substeps = {}    # instance id -> its substeps per frame

    def close(self):
        if self.handle:
            self._lib.tfx_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def admit(self, asset, settings, mesh=None, sdf_grid=None, sdf_num_cells_in_x=0, collision_margin=0.5,
              capsules=None, substeps=0):
        """mesh: dict(positions, skinning, indices, normals=None, follow_bone=0);
        sdf_grid: (origin, cell, dims) explicit grid, else sdf_num_cells_in_x;
        substeps: the instance's substeps per frame (0: the server's).
        Raises TFXError (code TFX_ERR_ADMISSION = 4) when it does not fit now."""
        d = InstanceDesc()
        d.substeps = int(substeps)
        ad = asset.desc()
        d.asset = ctypes.pointer(ad)
        keep = [ad, asset]
        if mesh is not None:
            pos = np.ascontiguousarray(mesh["positions"], np.float32).reshape(-1, 3)
            nrm = mesh.get("normals")
            nrm = None if nrm is None else np.ascontiguousarray(nrm, np.float32).reshape(-1, 3)
            sk = np.ascontiguousarray(mesh["skinning"], np.float32).reshape(-1, 8)
            idx = np.ascontiguousarray(mesh["indices"], np.uint32).reshape(-1, 3)
            md = SkinnedMeshDesc()
            md.num_vertices, md.num_triangles = pos.shape[0], idx.shape[0]
            md.positions = _fptr(pos)
            md.normals = _fptr(nrm) if nrm is not None else None
            md.skinning = sk.ctypes.data
            md.indices = idx.ctypes.data
            md.follow_bone = int(mesh.get("follow_bone", 0))
            d.mesh = ctypes.pointer(md)
            keep += [pos, nrm, sk, idx, md]
            if sdf_grid is not None:
                origin, cell, dims = sdf_grid
                for k in range(3):
                    d.sdf_origin[k] = float(origin[k])
                    d.sdf_dims[k] = int(dims[k])
                d.sdf_cell = float(cell)
            else:
                d.sdf_num_cells_in_x = int(sdf_num_cells_in_x)
            d.collision_margin = float(collision_margin)
        d.settings = settings
        if capsules is not None:
            c0 = np.ascontiguousarray(capsules[0], np.float32).reshape(-1, 4)
            c1 = np.ascontiguousarray(capsules[1], np.float32).reshape(-1, 4)
            d.capsules0, d.capsules1, d.num_capsules = _fptr(c0), _fptr(c1), c0.shape[0]
            keep += [c0, c1]
        out = ctypes.c_int()
        _check(self._lib.tfx_server_admit(self.handle, ctypes.byref(d), ctypes.byref(out)))
        self._substeps[out.value] = int(substeps) or self.substeps
        return out.value

    def retire(self, instance_id):
        _check(self._lib.tfx_server_retire(self.handle, int(instance_id)))
        self._substeps.pop(int(instance_id), None)

    def submit(self, instance_id, hair_bones, mesh_bones=None, wind=None, track=True):
        """hair_bones float32[frames, substeps, nb, 16] (or [frames, substeps, 16]) with the
        instance's substeps; mesh_bones float32[frames, nmb, 16] or None;
        wind float32[frames, substeps] or None."""
        subs = self._substeps.get(int(instance_id), self.substeps)
        hb = np.ascontiguousarray(hair_bones, np.float32)
        frames = hb.shape[0]
        hb = hb.reshape(frames, subs, -1)
        nb = hb.shape[2] // 16
        mb, nmb = None, 0
        if mesh_bones is not None:
            mb = np.ascontiguousarray(mesh_bones, np.float32).reshape(frames, -1)
            nmb = mb.shape[1] // 16
        w = None if wind is None else np.ascontiguousarray(wind, np.float32).reshape(frames, subs)
        t = ctypes.c_ulonglong()
        _check(self._lib.tfx_server_submit(self.handle, int(instance_id), frames, _fptr(hb), nb,
                                           _fptr(mb) if mb is not None else None, nmb,
                                           _fptr(w) if w is not None else None,
                                           ctypes.byref(t) if track else None))
        return t.value if track else None

    def poll(self, ticket):
        done, lat = ctypes.c_int(), ctypes.c_double()
        _check(self._lib.tfx_server_poll(self.handle, ctypes.c_ulonglong(ticket), ctypes.byref(done),
                                         ctypes.byref(lat)))
        return bool(done.value), lat.value

    def wait(self, ticket):
        lat = ctypes.c_double()
        _check(self._lib.tfx_server_wait(self.handle, ctypes.c_ulonglong(ticket), ctypes.byref(lat)))
        return lat.value

    def drain(self):
        _check(self._lib.tfx_server_drain(self.handle))

    def set_sdf_deferral(self, mode):
        """As HairBatch.set_sdf_deferral, for the passes launched from now on."""
        _check(self._lib.tfx_server_set_sdf_deferral(self.handle, int(mode)))

    def stats(self, reset=False):
        st = ServerStats()
        _check(self._lib.tfx_server_stats_read(self.handle, ctypes.byref(st), 1 if reset else 0))
        return {n: getattr(st, n) for n, _ in ServerStats._fields_}

    def hair(self, instance_id):
        """The instance's hair object (borrowed; the instance must be idle)."""
        h, sdf = ctypes.c_void_p(), ctypes.c_void_p()
        _check(self._lib.tfx_server_get_objects(self.handle, int(instance_id), ctypes.byref(h), ctypes.byref(sdf)))
        return _BorrowedHair(h.value)


def _names(names):
    names = [n.encode() if isinstance(n, str) else n for n in (names or [])]
    arr = (ctypes.c_char_p * max(len(names), 1))(*names)
    return arr, len(names)


class TressFXAsset:
    """TressFXAsset (src/TressFX/TressFXAsset.h:48-120) built by the library's
    loaders, in the reference's order: LoadHairData, GenerateFollowHairs,
    ProcessAsset, LoadBoneData; `hair_asset()` is what TressFXHairObject takes."""

    def __init__(self):
        self._lib = load_library()
        self.handle = None

    def close(self):
        if self.handle:
            self._lib.tfx_asset_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def LoadHairData(self, path):
        self.close()
        h = ctypes.c_void_p()
        _check(self._lib.tfx_asset_load_hair_data(os.fsencode(path), ctypes.byref(h)))
        self.handle = h
        return True

    def GenerateFollowHairs(self, num_follow_hairs_per_guide, tip_separation_factor, max_radius, seed=0):
        _check(self._lib.tfx_asset_generate_follow_hairs(self.handle, int(num_follow_hairs_per_guide),
                                                         ctypes.c_float(tip_separation_factor),
                                                         ctypes.c_float(max_radius), ctypes.c_uint(seed)))
        return True

    def ProcessAsset(self):
        _check(self._lib.tfx_asset_process(self.handle))
        return True

    def LoadBoneData(self, path, engine_bone_names):
        arr, n = _names(engine_bone_names)
        _check(self._lib.tfx_asset_load_bone_data(self.handle, os.fsencode(path), arr, n))
        return True

    def info(self):
        v = [ctypes.c_int() for _ in range(4)]
        _check(self._lib.tfx_asset_get_info(self.handle, *[ctypes.byref(x) for x in v]))
        return dict(file_strands=v[0].value, num_guide_strands=v[1].value, num_follow_strands_per_guide=v[2].value,
                    num_vertices_per_strand=v[3].value)

    def hair_asset(self):
        """Copies of the processed arrays as a HairAsset."""
        d = AssetDesc()
        _check(self._lib.tfx_asset_get_desc(self.handle, ctypes.byref(d)))
        G, F, V = d.num_guide_strands, d.num_follow_strands_per_guide, d.num_vertices_per_strand
        S = G * (F + 1)
        N = S * V

        def arr(ptr, n):
            return np.ctypeslib.as_array(ctypes.cast(ptr, c_float_p), shape=(n,)).copy()
        return HairAsset(G, F, V, arr(d.positions, 4 * N).reshape(N, 4), arr(d.tangents, 4 * N).reshape(N, 4),
                         arr(d.rest_lengths, N), arr(d.follow_root_offsets, 4 * S).reshape(S, 4),
                         arr(d.bone_skinning_data, 8 * S).reshape(S, 8))

    def render_data(self):
        uv, th = c_float_p(), c_float_p()
        tri, n = ctypes.POINTER(ctypes.c_int32)(), ctypes.c_int()
        _check(self._lib.tfx_asset_get_render_data(self.handle, ctypes.byref(uv), ctypes.byref(th),
                                                   ctypes.byref(tri), ctypes.byref(n)))
        info = self.info()
        S = info["num_guide_strands"] * (info["num_follow_strands_per_guide"] + 1)
        N = S * info["num_vertices_per_strand"]
        return (np.ctypeslib.as_array(uv, shape=(2 * S,)).reshape(S, 2).copy(),
                np.ctypeslib.as_array(th, shape=(N,)).copy(),
                np.ctypeslib.as_array(tri, shape=(n.value,)).copy())


def load_collision_mesh(path, engine_bone_names, follow_bone=None):
    """A .tfxmesh (TressFXBoneSkinning::LoadTressFXCollisionMeshData) as the
    mesh dict TressFXBoneSkinning and HairServer take."""
    lib = load_library()
    arr, n = _names(engine_bone_names)
    h = ctypes.c_void_p()
    _check(lib.tfx_collision_mesh_load(os.fsencode(path), arr, n,
                                       follow_bone.encode() if follow_bone else None, ctypes.byref(h)))
    try:
        d = SkinnedMeshDesc()
        _check(lib.tfx_collision_mesh_get_desc(h, ctypes.byref(d)))
        nv, nt = d.num_vertices, d.num_triangles
        pos = np.ctypeslib.as_array(d.positions, shape=(3 * nv,)).reshape(nv, 3).copy()
        nrm = np.ctypeslib.as_array(d.normals, shape=(3 * nv,)).reshape(nv, 3).copy()
        skin = np.ctypeslib.as_array(ctypes.cast(d.skinning, c_float_p), shape=(8 * nv,)).reshape(nv, 8).copy()
        idx = np.ctypeslib.as_array(ctypes.cast(d.indices, ctypes.POINTER(ctypes.c_uint32)),
                                    shape=(3 * max(nt, 0),)).reshape(-1, 3).copy()
        return dict(positions=pos, normals=nrm, skinning=skin, indices=idx, follow_bone=d.follow_bone)
    finally:
        lib.tfx_collision_mesh_destroy(h)
