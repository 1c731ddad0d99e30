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
COND_FACTOR = 4.0


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


def conditioned_tolerance(asset, st0, advance, seed=0, cond_factor=COND_FACTOR):
    """Per-vertex tolerance against the FP64 oracle for one substep.

    `advance(hair64)` runs the substep on an oracle.OracleHair(float64).  The
    FP64 oracle is run from st0 and from st0 perturbed by one float32 rounding;
    the relative spread `sens` is the error any FP32 evaluation must be allowed
    at that vertex (its inputs are themselves rounded).  Tolerance per vertex =
    max(1e-5, 4 * sens): well-conditioned vertices are held to 1e-5, vertices
    on nearly collapsed segments (normalising a vector of length << |p|
    amplifies rounding by |p|/length) are held to their conditioning."""
    import oracle
    ref = oracle.OracleHair(asset, np.float64)
    ref.set_state(st0)
    advance(ref)
    pert = oracle.OracleHair(asset, np.float64)
    pert.set_state(perturbed_state(st0, seed))
    advance(pert)
    sens = {k: rel_err(getattr(pert, k), getattr(ref, k)) for k in ("pos", "prev")}
    tol = {k: np.maximum(REL_TOL, cond_factor * sens[k]) for k in sens}
    return ref, tol


def check_fp64(gpu_st, ref, tol, what):
    """Asserts the conditioned bound; returns (worst error, number of vertices
    whose tolerance was widened by their conditioning)."""
    worst, widened = 0.0, 0
    for k in ("pos", "prev"):
        e = rel_err(gpu_st[k], getattr(ref, k))
        bad = e > tol[k]
        assert not bad.any(), (f"{what}: {k} rel err {e[bad].max():.3e} at vertex {int(np.argmax(bad))} "
                               f"exceeds its tolerance {tol[k][bad].max():.3e}")
        worst = max(worst, float(e.max()))
        widened += int((tol[k] > REL_TOL).sum())
    return worst, widened


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
