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

    def frame(self, f, stream=None, use_graph=False, autotune_repeats=0):
        """One frame; with autotune_repeats > 0 the batch autotuner runs on it
        first (records in self.autotune_records; the results are unaffected)."""
        k = f - self.first
        for sdf, d in zip(self.sdfs, self.descs[k]):
            if sdf is not None:
                sdf.SetMesh(d)
        self.batch.begin_frame()
        for s in range(self.substeps):
            bones, st = self.inputs[(f, s)]
            self.batch.stage_substep(bones, st, self.dt_sub)
        if autotune_repeats > 0:
            self.autotune_records = self.batch.autotune(stream, repeats=autotune_repeats)
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


def served_mesh(wl):
    """The collision mesh of a workload as the served form takes it: a body
    with its own skinning (FurBody) as it is, else the rest mesh (static
    noise) skinned every frame by the head bone (bone 0)."""
    if hasattr(wl.mesh, "skinned_mesh"):
        return wl.mesh.skinned_mesh()
    verts = wl.mesh.vertices(0.0)
    skin = np.zeros((verts.shape[0], 8), np.float32)
    skin[:, 4] = 1.0
    return dict(positions=verts[:, :3], normals=verts[:, 4:7], skinning=skin, indices=wl.mesh.indices, follow_bone=0)


def frame_inputs(wl, horizon):
    """Per-frame inputs of a workload for frames 0..horizon-1: hair bones at
    the end of each of its substeps [H, substeps, (nb,) 16], the collision
    mesh's bones at the frame's start [H, (nb,) 16] and wind [H, substeps]."""
    sub = wl.substeps
    f = np.arange(horizon, dtype=np.float64)
    t_sub = (f[:, None] + (np.arange(sub)[None, :] + 1.0) / sub) / synth.FRAME_RATE
    hb = synth.trajectory_bones(wl.bone_fn, t_sub)
    mb = synth.trajectory_bones(wl.bone_fn, f / synth.FRAME_RATE)
    wind = (np.vectorize(wl.wind_fn)(t_sub) if wl.wind_fn is not None
            else np.full(t_sub.shape, wl.settings.wind_magnitude)).astype(np.float32)
    return np.ascontiguousarray(hb), np.ascontiguousarray(mb), np.ascontiguousarray(wind)


class SkinnedFrameDriver:
    """One workload whose collider (a FurBody) is skinned on the GPU every
    frame -- TressFXBoneSkinning::Update, then the SDF Update -- stepped with
    the reference dispatch (`mode="plain"`: per substep Simulate +
    CollideWithHair), the fused substep (`"fused"`) or the batched form with
    the frame as a replayed hipGraph (`"graph"`).  Inputs for every frame of
    `frames` are evaluated beforehand."""

    def __init__(self, torch, wl, frames, mode="graph"):
        from . import HairBatch, TressFXBoneSkinning
        assert mode in ("plain", "fused", "graph")
        self.wl, self.mode = wl, mode
        self.frames = list(frames)
        self.first = self.frames[0]
        m = wl.mesh.skinned_mesh()
        self.skin = TressFXBoneSkinning(m["positions"], m["skinning"], m["indices"], normals=m["normals"],
                                        follow_bone=m["follow_bone"])
        origin, cell, dims = wl.sdf_grid
        self.sdf = TressFXSDFCollision(self.skin.mesh(), collision_margin=wl.collision_margin,
                                       explicit=(origin, cell, dims))
        self.hair = TressFXHairObject(wl.asset)
        self.sim = TressFXSimulation()
        self.system = TressFXSDFCollisionSystem()
        self.substeps = wl.substeps
        self.dt_sub = np.float32(1.0 / synth.FRAME_RATE / self.substeps)
        hb, mb, wind = frame_inputs(wl, self.frames[-1] + 1)
        self.hb = hb[self.first:].reshape(len(self.frames), self.substeps, -1)
        self.mb = mb[self.first:].reshape(len(self.frames), -1, 16)
        self.wind = wind[self.first:]
        self.batch = None
        if mode == "graph":
            self.batch = HairBatch()
            self.batch.add(self.hair, self.sdf)

    @property
    def vertices(self):
        return self.wl.asset.num_total_vertices

    def frame(self, f, stream=None):
        k = f - self.first
        wl = self.wl
        self.skin.Update(self.mb[k], stream)
        if self.batch is not None:
            self.sdf.SetMesh(self.skin.mesh())
            self.batch.begin_frame()
            for s in range(self.substeps):
                self.batch.stage_substep(self.hb[k, s][None], wl.settings, self.dt_sub,
                                         self.wind[k, s:s + 1] if wl.wind_fn is not None else None)
            self.batch.run_frame(stream, use_graph=True)
            return
        self.sdf.SetMesh(self.skin.mesh())
        self.sdf.Update(self.system, stream)
        for s in range(self.substeps):
            self.hair.UpdateBoneMatrices(self.hb[k, s].reshape(-1, 16))
            if wl.wind_fn is not None:
                wl.settings.wind_magnitude = float(self.wind[k, s])
            self.hair.UpdateSimulationParameters(wl.settings, self.dt_sub)
            if self.mode == "fused":
                self.sim.SimulateFused([self.hair], [self.sdf], stream)
            else:
                self.sim.Simulate([self.hair], stream)
                self.sdf.CollideWithHair(self.system, self.hair, stream)


class ServedRun:
    """A client of the served form: instances arriving as a Poisson process
    (each drawn from `workloads`, so the sizes follow their distribution) with
    exponential lifetimes, and step requests of k frames (k uniform over `ks`)
    arriving as a Poisson process on random resident instances.  A refused
    admission waits in a FIFO and is retried as memory frees up.  Per-frame
    inputs (bones for every substep, the collision mesh's bones, wind) are
    evaluated for `horizon` frames per workload before the run.

    `tenants`: workloads resident for the whole run outside the arrival
    process, each driven by its own client at `tenant_hz` requests of one
    frame per second; their request latencies are reported separately."""

    def __init__(self, server, workloads, initial, duration_s, inst_rate=20.0, lifetime_s=30.0, req_rate=5000.0,
                 ks=(1, 2, 4), seed=4, horizon=512, tenants=(), tenant_hz=60.0):
        self.server = server
        self.wls = list(workloads)
        self.tenants = list(tenants)
        self.tenant_hz = float(tenant_hz)
        self.initial = int(initial)
        self.duration = float(duration_s)
        self.ks = tuple(ks)
        self.substeps = server.substeps
        self.horizon = int(horizon)
        rng = np.random.Generator(np.random.PCG64(seed))
        self.rng = rng
        # arrivals and requests on [0, duration)
        n_arr = rng.poisson(inst_rate * self.duration)
        self.arrivals = np.sort(rng.uniform(0.0, self.duration, n_arr))
        self.arrival_wl = rng.integers(0, len(self.wls), n_arr)
        n_req = rng.poisson(req_rate * self.duration)
        self.requests = np.sort(rng.uniform(0.0, self.duration, n_req))
        self.req_k = rng.choice(self.ks, n_req)
        self.lifetime_s = float(lifetime_s)
        # per-workload inputs for `horizon` frames
        self.inputs = [frame_inputs(wl, self.horizon) for wl in self.wls]
        self.meshes = [served_mesh(wl) for wl in self.wls]
        self.tenant_inputs = [frame_inputs(wl, self.horizon) for wl in self.tenants]
        self.tenant_meshes = [served_mesh(wl) for wl in self.tenants]

    def _admit_wl(self, wl, mesh):
        try:
            return self.server.admit(wl.asset, wl.settings, mesh=mesh, sdf_grid=wl.sdf_grid,
                                     collision_margin=wl.collision_margin, capsules=wl.capsules,
                                     substeps=wl.substeps if wl.substeps != self.substeps else 0)
        except Exception as e:  # noqa: BLE001 - admission refusals come back as TFXError code 4
            if getattr(e, "code", None) == 4:
                return None
            raise

    def _admit(self, w):
        return self._admit_wl(self.wls[w], self.meshes[w])

    def burst(self, frames):
        """The served form's capacity: every workload admitted once, `frames`
        frames of each submitted at once (requests outrun the GPU, so every
        pass coalesces one frame of every instance), then drained.  Returns
        the server statistics over the drain and its wall time."""
        import time
        srv = self.server
        ids = []
        for w in range(len(self.wls)):
            iid = self._admit(w)
            if iid is None:
                raise RuntimeError("the burst's instances do not fit the admission budget")
            ids.append((iid, w))
        srv.drain()
        srv.stats(reset=True)
        t0 = time.perf_counter()
        for iid, w in ids:
            hb, mb, wind = self.inputs[w]
            srv.submit(iid, hb[:frames], mb[:frames], wind[:frames], track=False)
        t_sub = time.perf_counter() - t0
        srv.drain()
        t1 = time.perf_counter() - t0
        for iid, _ in ids:
            srv.retire(iid)
        srv.drain()  # their memory released before the caller's next measurement
        return {"window_s": t1, "submit_s": t_sub, "instances": len(ids), "frames_per_instance": int(frames),
                **srv.stats()}

    def run(self, log=None):
        import time
        srv = self.server
        resident = []          # instance ids
        info = {}              # id -> [workload, next frame, departure time]
        refused = 0
        pending = []           # (workload) waiting for admission
        tenant_ids = []
        for wl, mesh in zip(self.tenants, self.tenant_meshes):
            iid = self._admit_wl(wl, mesh)
            if iid is None:
                raise RuntimeError("a tenant does not fit the admission budget")
            tenant_ids.append(iid)
        tenant_next = [0] * len(tenant_ids)     # next frame
        tenant_tickets = [[] for _ in tenant_ids]
        tenant_period = 1.0 / self.tenant_hz if self.tenant_hz > 0 else np.inf
        next_tenant = 0.0
        for k in range(self.initial):
            w = k % len(self.wls)
            iid = self._admit(w)
            if iid is None:
                refused += 1
                pending.append(w)
                continue
            resident.append(iid)
            info[iid] = [w, 0, self.rng.exponential(self.lifetime_s)]
        srv.stats(reset=True)
        t0 = time.perf_counter()
        ia = ir = 0
        submitted = frames = retired = 0
        next_retry = 0.0
        last_log = 0.0
        while ia < len(self.arrivals) or ir < len(self.requests):
            now = time.perf_counter() - t0
            ta = self.arrivals[ia] if ia < len(self.arrivals) else np.inf
            tr = self.requests[ir] if ir < len(self.requests) else np.inf
            due = min(ta, tr)
            if tenant_ids and now >= next_tenant:
                for j, iid in enumerate(tenant_ids):
                    hb, mb, wind = self.tenant_inputs[j]
                    f = tenant_next[j] % self.horizon
                    tenant_tickets[j].append(srv.submit(iid, hb[f:f + 1], mb[f:f + 1], wind[f:f + 1]))
                    tenant_next[j] += 1
                next_tenant += tenant_period
            if due > now:
                time.sleep(max(0.0, min(due - now, next_tenant - now if tenant_ids else 1.0, 0.002)))
                continue
            # departures
            if resident and now >= next_retry:
                for iid in [i for i in resident if info[i][2] <= now]:
                    srv.retire(iid)
                    resident.remove(iid)
                    del info[iid]
                    retired += 1
                while pending:
                    iid = self._admit(pending[0])
                    if iid is None:
                        break
                    w = pending.pop(0)
                    resident.append(iid)
                    info[iid] = [w, 0, now + self.rng.exponential(self.lifetime_s)]
                next_retry = now + 0.05
            if ta <= tr:
                w = int(self.arrival_wl[ia])
                ia += 1
                iid = self._admit(w)
                if iid is None:
                    refused += 1
                    pending.append(w)
                else:
                    resident.append(iid)
                    info[iid] = [w, 0, now + self.rng.exponential(self.lifetime_s)]
            else:
                k = int(self.req_k[ir])
                ir += 1
                if not resident:
                    continue
                iid = resident[int(self.rng.integers(0, len(resident)))]
                w, f, _ = info[iid]
                hb, mb, wind = self.inputs[w]
                sl = np.arange(f, f + k) % self.horizon
                srv.submit(iid, hb[sl], mb[sl], wind[sl], track=False)
                info[iid][1] = f + k
                submitted += 1
                frames += k
            if log and now - last_log > 1.0:
                last_log = now
                log(f"served t={now:.1f}s resident={len(resident)} submitted={submitted}")
        t_sub = time.perf_counter() - t0
        srv.drain()
        t1 = time.perf_counter() - t0
        st = srv.stats()
        out = {
            "window_s": t1, "submit_window_s": t_sub, "requests_submitted": submitted, "frames_submitted": frames,
            "offered_request_rate": submitted / t_sub if t_sub > 0 else 0.0, "instances_retired": retired,
            "admissions_refused": refused, "resident_at_end": len(resident), **st,
        }
        # instances still resident: workload index and frames run (inputs frames 0..n-1 of that
        # workload while n stays within the horizon)
        self.resident = {int(iid): (int(info[iid][0]), int(info[iid][1])) for iid in resident}
        if tenant_ids:
            tl = []
            for j, iid in enumerate(tenant_ids):
                lat = np.array([srv.wait(t) for t in tenant_tickets[j]])
                wl = self.tenants[j]
                tl.append({"frames": int(tenant_next[j]), "vertices": int(wl.asset.num_total_vertices),
                           "substeps": int(wl.substeps),
                           "vertex_substeps_per_s": tenant_next[j] * wl.asset.num_total_vertices * wl.substeps / t1,
                           "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                                          "max": float(lat.max())}})
            out["tenants"] = tl
        return out
