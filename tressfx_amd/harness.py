"""Frame driver reproducing the reference's per-frame order
(src/Simulation.cpp:50-131, TressFXSample.cpp:795-810):

    skin/animate collision mesh -> SDF Update -> per substep:
        UpdateBones + UpdateSimulationParameters -> Simulate -> CollideWithHair

with a substep defined as one Simulate at dt/substeps followed by the
collision, the SDF rebuilt once per frame and bones evaluated at the substep's
end time.  Device memory for the collision meshes comes from torch (plumbing);
all compute is the HIP library.
"""
import numpy as np

from . import SdfMeshDesc, TressFXHairObject, TressFXSDFCollision, TressFXSDFCollisionSystem, TressFXSimulation
from . import synth


class ResidentMeshFrames:
    """Collision-mesh vertex buffers for a window of frames, uploaded to HBM
    before the timed region (StandardVertex float32[M, 8] per frame)."""

    def __init__(self, torch, mesh, bone_fn, frames, frame_rate=synth.FRAME_RATE):
        self.mesh = mesh
        host = np.stack([mesh.vertices(f / frame_rate, bone_fn(f / frame_rate) if bone_fn else None) for f in frames])
        self.frames = list(frames)
        self.verts_d = torch.from_numpy(host).cuda()
        self.idx_d = torch.from_numpy(mesh.indices.reshape(-1).astype(np.int32)).cuda()
        p = host[..., :3]
        self.bmin = p.min(axis=1)
        self.bmax = p.max(axis=1)
        self.nv = host.shape[1]

    def desc(self, k):
        d = SdfMeshDesc()
        d.vertices = self.verts_d[k].data_ptr()
        d.indices = self.idx_d.data_ptr()
        d.num_vertices = self.nv
        d.num_triangles = self.mesh.indices.shape[0]
        for a in range(3):
            d.bbox_min[a], d.bbox_max[a] = float(self.bmin[k][a]), float(self.bmax[k][a])
            d.initial_bbox_min[a], d.initial_bbox_max[a] = float(self.bmin[0][a]), float(self.bmax[0][a])
        return d


class FrameDriver:
    """One workload (hair object + collider) stepped frame by frame with either
    the reference dispatch structure (`fused=False`) or the fused substep."""

    def __init__(self, torch, wl, frames, fused=True, sim=None, system=None):
        self.wl = wl
        self.fused = fused
        self.sim = sim or TressFXSimulation()
        self.system = system or TressFXSDFCollisionSystem()
        self.hair = TressFXHairObject(wl.asset)
        if wl.capsules is not None:
            self.hair.SetCollisionCapsules(*wl.capsules)
        self.meshes = None
        self.sdf = None
        self.first = frames[0]
        n = wl.substeps
        dt = 1.0 / synth.FRAME_RATE
        self.bones = {}
        self.wind = {}
        for f in frames:
            for s in range(n):
                t = (f + (s + 1) / n) * dt
                self.bones[(f, s)] = np.ascontiguousarray(wl.bone_fn(t), np.float32).reshape(1, 16)
                self.wind[(f, s)] = wl.wind_fn(t) if wl.wind_fn else None
        self.dt_sub = np.float32(dt / n)
        if wl.mesh is not None:
            self.meshes = ResidentMeshFrames(torch, wl.mesh, wl.bone_fn, frames)
            origin, cell, dims = wl.sdf_grid
            self.sdf = TressFXSDFCollision(self.meshes.desc(0), collision_margin=wl.collision_margin,
                                           explicit=(origin, cell, dims))

    def frame(self, f, stream=None):
        wl = self.wl
        if self.sdf is not None:
            self.sdf.SetMesh(self.meshes.desc(f - self.first))
            self.sdf.Update(self.system, stream)
        for s in range(wl.substeps):
            self.hair.UpdateBoneMatrices(self.bones[(f, s)])
            w = self.wind[(f, s)]
            if w is not None:
                wl.settings.wind_magnitude = w
            self.hair.UpdateSimulationParameters(wl.settings, self.dt_sub)
            if self.fused:
                self.sim.SimulateFused([self.hair], [self.sdf] if self.sdf is not None else None, stream)
            else:
                self.sim.Simulate([self.hair], stream)
                if self.sdf is not None:
                    self.sdf.CollideWithHair(self.system, self.hair, stream)

    @property
    def vertices(self):
        a = self.wl.asset
        return a.num_total_vertices


class BatchDriver:
    """Many workloads stepped together through the batched form (HairBatch):
    per frame every collider's mesh is set, every member's substep inputs are
    staged, and the frame is run -- directly or as a replayed hipGraph.
    Per-frame host inputs (mesh descriptors, bone matrices, settings) are
    built for the whole window of frames before any frame runs."""

    def __init__(self, torch, workloads, frames, launch_shapes=None, meshes=None):
        """meshes: ResidentMeshFrames per workload to share (same frames), or
        None to upload them here."""
        from . import HairBatch, SimulationSettings
        self.wls = list(workloads)
        self.frames = list(frames)
        self.first = self.frames[0]
        self.hairs, self.sdfs, self.meshes = [], [], []
        self.batch = HairBatch()
        for i, wl in enumerate(self.wls):
            hair = TressFXHairObject(wl.asset)
            if wl.capsules is not None:
                hair.SetCollisionCapsules(*wl.capsules)
            sdf = m = None
            if wl.mesh is not None:
                m = meshes[i] if meshes is not None else ResidentMeshFrames(torch, wl.mesh, wl.bone_fn, self.frames)
                origin, cell, dims = wl.sdf_grid
                sdf = TressFXSDFCollision(m.desc(0), collision_margin=wl.collision_margin,
                                          explicit=(origin, cell, dims))
            self.hairs.append(hair)
            self.sdfs.append(sdf)
            self.meshes.append(m)
            self.batch.add(hair, sdf)
        for V, variant in (launch_shapes or {}).items():
            self.batch.set_launch_shape(V, variant)
        subs = {wl.substeps for wl in self.wls}
        assert len(subs) == 1, "members of one batch share the substep count"
        self.substeps = subs.pop()
        dt = 1.0 / synth.FRAME_RATE
        self.dt_sub = np.float32(dt / self.substeps)
        n = len(self.wls)
        self.descs = [[m.desc(k) if m is not None else None for m in self.meshes] for k in range(len(self.frames))]
        self.inputs = {}
        for f in self.frames:
            for s in range(self.substeps):
                t = (f + (s + 1) / self.substeps) * dt
                bones = np.stack([np.ascontiguousarray(wl.bone_fn(t), np.float32).reshape(16) for wl in self.wls])
                st = (SimulationSettings * n)()
                for m, wl in enumerate(self.wls):
                    st[m] = wl.settings
                    if wl.wind_fn is not None:
                        st[m].wind_magnitude = wl.wind_fn(t)
                self.inputs[(f, s)] = (bones, st)

    @property
    def vertices(self):
        return sum(wl.asset.num_total_vertices for wl in self.wls)

    def frame(self, f, stream=None, use_graph=False):
        k = f - self.first
        for sdf, d in zip(self.sdfs, self.descs[k]):
            if sdf is not None:
                sdf.SetMesh(d)
        self.batch.begin_frame()
        for s in range(self.substeps):
            bones, st = self.inputs[(f, s)]
            self.batch.stage_substep(bones, st, self.dt_sub)
        self.batch.run_frame(stream, use_graph)


class PlainMultiDriver:
    """The same workloads stepped with the reference's per-object dispatch:
    per frame each collider's SDF Update, then per substep each object's
    inputs, Simulate over the objects and CollideWithHair per object."""

    def __init__(self, torch, workloads, frames, meshes=None):
        self.wls = list(workloads)
        self.frames = list(frames)
        self.first = self.frames[0]
        self.sim = TressFXSimulation()
        self.system = TressFXSDFCollisionSystem()
        self.hairs, self.sdfs, self.meshes = [], [], []
        for i, wl in enumerate(self.wls):
            hair = TressFXHairObject(wl.asset)
            if wl.capsules is not None:
                hair.SetCollisionCapsules(*wl.capsules)
            sdf = m = None
            if wl.mesh is not None:
                m = meshes[i] if meshes is not None else ResidentMeshFrames(torch, wl.mesh, wl.bone_fn, self.frames)
                origin, cell, dims = wl.sdf_grid
                sdf = TressFXSDFCollision(m.desc(0), collision_margin=wl.collision_margin,
                                          explicit=(origin, cell, dims))
            self.hairs.append(hair)
            self.sdfs.append(sdf)
            self.meshes.append(m)
        self.substeps = self.wls[0].substeps
        dt = 1.0 / synth.FRAME_RATE
        self.dt_sub = np.float32(dt / self.substeps)
        self.descs = [[m.desc(k) if m is not None else None for m in self.meshes] for k in range(len(self.frames))]
        self.inputs = {}
        for f in self.frames:
            for s in range(self.substeps):
                t = (f + (s + 1) / self.substeps) * dt
                self.inputs[(f, s)] = [(np.ascontiguousarray(wl.bone_fn(t), np.float32).reshape(1, 16),
                                        wl.wind_fn(t) if wl.wind_fn else None) for wl in self.wls]

    @property
    def vertices(self):
        return sum(wl.asset.num_total_vertices for wl in self.wls)

    def frame(self, f, stream=None):
        k = f - self.first
        for sdf, d in zip(self.sdfs, self.descs[k]):
            if sdf is not None:
                sdf.SetMesh(d)
                sdf.Update(self.system, stream)
        for s in range(self.substeps):
            for hair, wl, (bones, wind) in zip(self.hairs, self.wls, self.inputs[(f, s)]):
                hair.UpdateBoneMatrices(bones)
                if wind is not None:
                    wl.settings.wind_magnitude = wind
                hair.UpdateSimulationParameters(wl.settings, self.dt_sub)
            self.sim.Simulate(self.hairs, stream)
            for hair, sdf in zip(self.hairs, self.sdfs):
                if sdf is not None:
                    sdf.CollideWithHair(self.system, hair, stream)
