"""Shared helpers of the parity tests: workload stepping on the GPU path and
the oracle side by side."""
import ctypes

import numpy as np

import oracle
import tressfx_amd as tfx
from tressfx_amd import synth

# Parity tolerance against the FP64 oracle (relative per vertex position per
# substep): |p_gpu - p_64| / max(|p_64|, 1e-2 m) <= 1e-5.
REL_TOL = 1e-5
REL_FLOOR = 1e-2


def rel_err(a, b):
    """Per-vertex relative position error of float4 arrays (xyz only)."""
    a = np.asarray(a, np.float64)[:, :3]
    b = np.asarray(b, np.float64)[:, :3]
    num = np.linalg.norm(a - b, axis=1)
    den = np.maximum(np.linalg.norm(b, axis=1), REL_FLOOR)
    return num / den


EPS32 = 2.0 ** -24


def perturbed_state(st, seed):
    """The state with every position component rounded-off by up to one float32
    unit in the last place (relative +-2^-24), in FP64."""
    r = np.random.Generator(np.random.PCG64(seed))
    out = {}
    for k, v in st.items():
        a = np.asarray(v, np.float64).copy()
        if k in ("pos", "prev", "prevprev"):
            a[:, :3] *= 1.0 + r.uniform(-EPS32, EPS32, a[:, :3].shape)
        out[k] = a
    return out


def fp64_reference(asset, st0, advance, seed=0):
    """The FP64 oracle's substep from st0, and the conditioning of every vertex.

    `advance(hair64)` runs the substep on an oracle.OracleHair(float64).  It is
    run from st0 and from st0 perturbed by one float32 rounding; the relative
    spread `sens` (per vertex, pos and prev) is what any FP32 evaluation has to
    be allowed at that vertex, because its inputs are themselves rounded.  It
    does not widen the bound (check_fp64 holds every vertex to 1e-5); it names
    the cause when a vertex fails: a spread near 1e-5 marks a branch that flips
    under rounding or a nearly collapsed segment (normalising a vector of
    length << |p| amplifies rounding by |p| / length)."""
    ref = oracle.OracleHair(asset, np.float64)
    ref.set_state(st0)
    advance(ref)
    pert = oracle.OracleHair(asset, np.float64)
    pert.set_state(perturbed_state(st0, seed))
    advance(pert)
    sens = {k: rel_err(getattr(pert, k), getattr(ref, k)) for k in ("pos", "prev")}
    return ref, sens


def check_fp64(gpu_st, ref, sens, what, V=None):
    """Asserts |p_gpu - p_64| / max(|p_64|, 1 cm) <= 1e-5 at every vertex
    (pos and prev); a failure names the vertex (index, and strand and vertex
    within it when V is given), its error and its FP64 conditioning.  Returns
    the worst error."""
    worst = 0.0
    for k in ("pos", "prev"):
        e = rel_err(gpu_st[k], getattr(ref, k))
        if e.size and e.max() > REL_TOL:
            i = int(np.argmax(e))
            where = f"vertex {i}" + (f" (strand {i // V}, vertex {i % V})" if V else "")
            raise AssertionError(
                f"{what}: {k} rel err {e[i]:.3e} at {where} exceeds {REL_TOL:g}; "
                f"{int((e > REL_TOL).sum())} vertices over; FP64 spread under one float32 rounding there "
                f"{sens[k][i]:.3e} (near 1e-5: a branch flipping under rounding or a collapsed segment)")
        worst = max(worst, float(e.max()) if e.size else 0.0)
    return worst


def bits_equal(a, b):
    return np.array_equal(np.ascontiguousarray(a, np.float32).view(np.uint32),
                          np.ascontiguousarray(b, np.float32).view(np.uint32))


def first_mismatch(a, b):
    au = np.ascontiguousarray(a, np.float32).view(np.uint32)
    bu = np.ascontiguousarray(b, np.float32).view(np.uint32)
    idx = np.argwhere(au != bu)
    if idx.size == 0:
        return None
    i = tuple(idx[0])
    return i, np.asarray(a)[i], np.asarray(b)[i], int(idx.shape[0])


class DeviceMesh:
    """Collision mesh whose vertex/index buffers live on the GPU (torch)."""

    def __init__(self, torch, mesh, bone_fn=None):
        self.torch = torch
        self.mesh = mesh
        self.bone_fn = bone_fn
        self.idx_d = torch.from_numpy(mesh.indices.reshape(-1).astype(np.int32)).cuda()
        self.verts_h = None
        self.verts_d = None

    def update(self, t):
        bone = self.bone_fn(t) if self.bone_fn else None
        self.verts_h = self.mesh.vertices(t, bone)
        v = self.torch.from_numpy(self.verts_h)
        if self.verts_d is None or self.verts_d.shape != v.shape:
            self.verts_d = v.cuda()
        else:
            self.verts_d.copy_(v)
        return self.desc()

    def desc(self, initial=None):
        d = tfx.SdfMeshDesc()
        d.vertices = self.verts_d.data_ptr()
        d.indices = self.idx_d.data_ptr()
        d.num_vertices = self.verts_h.shape[0]
        d.num_triangles = self.mesh.indices.shape[0]
        p = self.verts_h[:, :3]
        mn, mx = p.min(axis=0), p.max(axis=0)
        for k in range(3):
            d.bbox_min[k], d.bbox_max[k] = float(mn[k]), float(mx[k])
        ini = initial if initial is not None else (mn, mx)
        for k in range(3):
            d.initial_bbox_min[k], d.initial_bbox_max[k] = float(ini[0][k]), float(ini[1][k])
        return d


def set_substep_inputs(hair, wl, frame, sub):
    """Per-substep host inputs: bones at the substep end time, parameters with
    dt/substeps and the wind schedule (Simulation.cpp:121-132 order)."""
    t = (frame + (sub + 1) / wl.substeps) / synth.FRAME_RATE
    hair.UpdateBoneMatrices(wl.bone_fn(t))
    if wl.wind_fn is not None:
        wl.settings.wind_magnitude = wl.wind_fn(t)
    hair.UpdateSimulationParameters(wl.settings, np.float32(1.0 / synth.FRAME_RATE / wl.substeps))


def oracle_pair(asset):
    return oracle.OracleHair(asset, np.float32), oracle.OracleHair(asset, np.float64)
