"""ctypes front of the CPU restatement (libtfx_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker.  The product never uses it.
"""
import ctypes
import os
import subprocess

import numpy as np

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(ORACLE_DIR, "libtfx_oracle.so")

_VP = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float
_D = ctypes.c_double


class OrcHair(ctypes.Structure):
    _fields_ = [
        ("num_guides", _I), ("num_follow", _I), ("num_vertices", _I), ("pad", _I),
        ("pos", _VP), ("prev", _VP), ("prevprev", _VP), ("tangent", _VP), ("strand_level", _VP),
        ("rest", _VP), ("rest_lengths", _VP), ("follow_offsets", _VP), ("skin", _VP),
    ]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
    return LIB


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    lib = ctypes.CDLL(LIB)
    for p in ("32", "64"):
        getattr(lib, f"orc{p}_run_stage").argtypes = [ctypes.POINTER(OrcHair), _VP, _VP, _I, _I]
        getattr(lib, f"orc{p}_simulate").argtypes = [ctypes.POINTER(OrcHair), _VP, _VP, _I]
    R = {"32": _F, "64": _D}
    for p in ("32", "64"):
        getattr(lib, f"orc{p}_sdf_build").argtypes = [_VP, _VP, _I, _VP, R[p], _I, _I, _I, _VP, _I]
        getattr(lib, f"orc{p}_sdf_collide").argtypes = [_VP, _VP, _I, _I, _VP, _VP, R[p], _I, _I, _I, R[p], _I]
        getattr(lib, f"orc{p}_sdf_collide_mixed").argtypes = [_VP, _VP, _I, _VP, _VP, R[p], _I, _I, _I, R[p], _I]
        getattr(lib, f"orc{p}_generate_transforms").argtypes = [_VP, _I, _I, _I, _VP, _VP, _VP, _I, _VP, _I]
    for p in ("32", "64"):
        getattr(lib, f"orc{p}_mesh_skin").argtypes = [_VP, _VP, _I, _VP, _VP, _I]
    lib.orc32_get_signed_distance.argtypes = [_VP, _VP, _F, _I, _I, _I, _VP]
    lib.orc32_get_signed_distance.restype = _F
    lib.orc_float_flip3.argtypes = [_F]
    lib.orc_float_flip3.restype = ctypes.c_uint32
    lib.orc_ifloat_flip3.argtypes = [ctypes.c_uint32]
    lib.orc_ifloat_flip3.restype = ctypes.c_uint32
    lib.orc32_signed_distance_point_triangle.argtypes = [_VP] * 4
    lib.orc32_signed_distance_point_triangle.restype = _F
    lib.orc64_signed_distance_point_triangle.argtypes = [_VP] * 4
    lib.orc64_signed_distance_point_triangle.restype = _D
    lib.orc64_make_quaternion.argtypes = [_VP, _VP]
    lib.orc64_quat_from_two_unit_vectors.argtypes = [_VP, _VP, _VP]
    lib.orc64_mult_quat_vec.argtypes = [_VP, _VP, _VP]
    lib.orc64_capsule_collision.argtypes = [_VP] * 5
    lib.orc64_capsule_collision.restype = _I
    lib.orc_pack_sim_params.argtypes = [_VP, _F, _I, _I, _I, _VP]
    lib.orc_sdf_grid_from_bbox.argtypes = [_VP, _VP, _VP, _I, _VP, _VP, _VP, _VP]
    _lib = lib
    return lib


def _p(a):
    return a.ctypes.data if a is not None else None


class OracleHair:
    """Host copy of one hair object's arrays in float32 or float64."""

    def __init__(self, asset, dtype=np.float64):
        self.dtype = np.dtype(dtype)
        dt = self.dtype
        self.G = asset.num_guide_strands
        self.F = asset.num_follow_strands_per_guide
        self.V = asset.num_vertices_per_strand
        self.rest = np.ascontiguousarray(asset.positions, dt)
        self.rest_lengths = np.ascontiguousarray(asset.rest_lengths, dt)
        self.follow_offsets = np.ascontiguousarray(asset.follow_root_offsets, dt)
        self.skin = np.ascontiguousarray(asset.bone_skinning_data, np.float32)
        self.pos = self.rest.copy()
        self.prev = self.rest.copy()
        self.prevprev = self.rest.copy()
        self.tangent = np.ascontiguousarray(asset.tangents, dt)
        self.strand_level = np.zeros((self.G * (self.F + 1), 12), dt)

    def set_state(self, st):
        """Copies of the given arrays (the caller's arrays are never advanced in place)."""
        for k in ("pos", "prev", "prevprev", "tangent", "strand_level"):
            if k in st and st[k] is not None:
                setattr(self, k, np.array(st[k], self.dtype, order="C", copy=True))

    def state(self):
        return {k: getattr(self, k).copy() for k in ("pos", "prev", "prevprev", "tangent", "strand_level")}

    def _struct(self):
        h = OrcHair()
        h.num_guides, h.num_follow, h.num_vertices = self.G, self.F, self.V
        h.pos, h.prev, h.prevprev = _p(self.pos), _p(self.prev), _p(self.prevprev)
        h.tangent, h.strand_level = _p(self.tangent), _p(self.strand_level)
        h.rest, h.rest_lengths, h.follow_offsets = _p(self.rest), _p(self.rest_lengths), _p(self.follow_offsets)
        h.skin = _p(self.skin)
        return h

    def _prefix(self):
        return "orc64" if self.dtype == np.float64 else "orc32"

    def run_stage(self, params, bones, stage, nthreads=0):
        lib = load()
        h = self._struct()
        b = np.ascontiguousarray(bones, np.float32)
        getattr(lib, self._prefix() + "_run_stage")(ctypes.byref(h), ctypes.addressof(params), _p(b), stage, nthreads)

    def simulate(self, params, bones, nthreads=0):
        lib = load()
        h = self._struct()
        b = np.ascontiguousarray(bones, np.float32)
        getattr(lib, self._prefix() + "_simulate")(ctypes.byref(h), ctypes.addressof(params), _p(b), nthreads)


def sdf_build(verts, indices, origin, cell, dims, dtype=np.float64, nthreads=0):
    lib = load()
    dt = np.dtype(dtype)
    v = np.ascontiguousarray(verts, dt)
    idx = np.ascontiguousarray(indices, np.uint32)
    o = np.ascontiguousarray(origin, dt)
    out = np.zeros(int(dims[0]) * int(dims[1]) * int(dims[2]), dt)
    fn = lib.orc64_sdf_build if dt == np.float64 else lib.orc32_sdf_build
    fn(_p(v), _p(idx), idx.size // 3, _p(o), cell, int(dims[0]), int(dims[1]), int(dims[2]), _p(out), nthreads)
    return out


def sdf_collide(pos, prev, V, sdf, origin, cell, dims, margin, nthreads=0, mixed=False):
    """CollideHairVerticesWithSdf_forward (or, with mixed=True, the mixed
    forward/backward-difference CollideHairVerticesWithSdf) in place."""
    lib = load()
    dt = pos.dtype
    o = np.ascontiguousarray(origin, dt)
    n = pos.shape[0]
    if mixed:
        fn = lib.orc64_sdf_collide_mixed if dt == np.float64 else lib.orc32_sdf_collide_mixed
        fn(_p(pos), _p(prev), n, _p(np.ascontiguousarray(sdf, dt)), _p(o), cell, int(dims[0]), int(dims[1]),
           int(dims[2]), margin, nthreads)
        return
    fn = lib.orc64_sdf_collide if dt == np.float64 else lib.orc32_sdf_collide
    fn(_p(pos), _p(prev), n, V, _p(np.ascontiguousarray(sdf, dt)), _p(o), cell, int(dims[0]), int(dims[1]),
       int(dims[2]), margin, nthreads)


def generate_transforms(rest, G, F, V, mapping, mesh_vertices, skinned_vertices, dtype=np.float32, nthreads=0):
    """GenerateTransforms transcription: rest float[S*V, 4] (initial hair
    positions), mapping: a ctypes array or numpy buffer of
    tfx_hair_to_triangle_mapping[S], mesh/skinned float[3*T, 4] ->
    float[S, 20] (row-major 4x4, quaternion); rows of follow strands, and of
    strands whose triangle lies past the mesh, stay zero."""
    lib = load()
    dt = np.dtype(dtype)
    r = np.ascontiguousarray(rest, dt)
    m = np.ascontiguousarray(mesh_vertices, dt)
    k = np.ascontiguousarray(skinned_vertices, dt)
    mp = np.ascontiguousarray(mapping)
    S = G * (F + 1)
    out = np.zeros((S, 20), dt)
    fn = lib.orc64_generate_transforms if dt == np.float64 else lib.orc32_generate_transforms
    fn(_p(r), G, F, V, mp.ctypes.data, _p(m), _p(k), m.shape[0], _p(out), nthreads)
    return out


def mesh_skin(rest, skin, bones, dtype=np.float32, nthreads=0):
    """BoneSkinning transcription: rest StandardVertex float32[n, 8], skin
    float32[n, 8] (bone ids as float, weights), bones float32[512, 16] ->
    skinned float[n, 8] of `dtype`."""
    lib = load()
    r = np.ascontiguousarray(rest, np.float32).reshape(-1, 8)
    sk = np.ascontiguousarray(skin, np.float32).reshape(-1, 8)
    b = np.zeros((512, 16), np.float32)
    bb = np.asarray(bones, np.float32).reshape(-1, 16)
    b[:bb.shape[0]] = bb
    b[bb.shape[0]:] = np.eye(4, dtype=np.float32).reshape(16)
    out = np.zeros(r.shape, np.dtype(dtype))
    fn = lib.orc64_mesh_skin if np.dtype(dtype) == np.float64 else lib.orc32_mesh_skin
    fn(_p(r), _p(sk), r.shape[0], _p(b), _p(out), nthreads)
    return out


def pack_sim_params(settings, dt, frame, V, F, into=None):
    """UpdateSimulationParameters transcription; `into` keeps the fields the
    call does not set (the object's persistent block)."""
    from tressfx_amd import SimulationParams
    lib = load()
    p = into if into is not None else SimulationParams()
    lib.orc_pack_sim_params(ctypes.addressof(settings), dt, frame, V, F, ctypes.addressof(p))
    return p
