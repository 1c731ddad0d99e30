// tfx_server.hip — the served form: resident instances advanced by a worker
// thread with continuous batching.
//
// Each pass takes the next queued frame of every instance that has one and runs
// them together on one stream: one upload of the pass's staging block, the
// collision meshes skinned in one launch, their SDFs built in one launch per
// build phase, then per substep one persistent fused launch per strand length
// over all members' strand blocks.  Requests that arrive while passes run join
// the next pass; an instance's frames run in submission order because passes
// are stream-ordered.  Up to `max_inflight` passes are queued on the GPU so the
// host prepares pass n+1 while pass n runs.  Per instance, a pass is the
// reference's per-frame order (src/Simulation.cpp:50-131): skin the collision
// mesh, SDF Update, then per substep UpdateBoneMatrices +
// UpdateSimulationParameters, Simulate, CollideWithHair.
#include "tfx_internal.h"

#include <chrono>
#include <condition_variable>
#include <deque>
#include <thread>
#include <atomic>
#include <unordered_map>

namespace {

using Clock = std::chrono::steady_clock;
inline double now_s() { return std::chrono::duration<double>(Clock::now().time_since_epoch()).count(); }

struct Request {
    unsigned long long ticket = 0;
    int frames = 0;
    int issued = 0;  // frames staged into passes
    int done = 0;    // frames completed on the GPU
    int nb = 0, nmb = 0;
    std::vector<float> hair_bones;  // [frames][instance substeps][nb * 16]
    std::vector<float> mesh_bones;  // [frames][nmb * 16] (empty: keep)
    std::vector<float> wind;        // [frames][instance substeps] (empty: settings)
    double t_submit = 0.0;
    bool keep_result = true;  // false: counted in the statistics only
};

struct Instance {
    int id = 0;
    tfx_hair* hair = nullptr;
    tfx_sdf* sdf = nullptr;
    tfx_mesh_skinning* mesh = nullptr;
    tfx_simulation_settings settings{};
    int substeps = 1;     // substeps per frame
    float dt = 0.0f;      // substep length
    long long bytes = 0;
    std::deque<Request*> queue;  // unfinished requests, oldest first
    int inflight = 0;            // frames staged but not completed
    bool retiring = false;
};

struct Result {
    bool finished = false;
    double latency_ms = 0.0;
    int status = TFX_OK;
};

struct Pass {
    hipEvent_t ev = nullptr;
    int slot = 0;
    std::vector<std::pair<Request*, Instance*>> frames;
    long long vertex_substeps = 0;
};

struct Slot {
    void* host = nullptr;
    void* dev = nullptr;
    size_t bytes = 0;
    hipEvent_t ev = nullptr;  // recorded after the last pass that used the slot
};

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

long long hair_bytes(const tfx_asset_desc* a) {
    const long long S = (long long)a->num_guide_strands * (a->num_follow_strands_per_guide + 1);
    const long long N = S * a->num_vertices_per_strand;
    return N * (5 * 16 + 4) + S * (16 + 32 + 48) + 16ll * TFX_MAX_NUM_BONES * 4 + (long long)sizeof(DevParams);
}

long long mesh_bytes(const tfx_skinned_mesh_desc* m) {
    return (long long)m->num_vertices * (32 + 32 + 32) + 12ll * m->num_triangles + 16ll * TFX_MAX_NUM_BONES * 4;
}

long long sdf_bytes(long long cells) {
    const long long bricks = cells / 64 + 1;
    return cells * 4 + bricks + bricks / 8 + 64;
}

}  // namespace

struct tfx_server {
    tfx_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv_work;  // worker: new work or stop
    std::condition_variable cv_done;  // clients: a request finished / idle
    std::unordered_map<int, Instance*> instances;
    std::unordered_map<unsigned long long, Result> results;
    int next_id = 1;
    unsigned long long next_ticket = 1;
    long long resident_bytes = 0;
    int unfinished = 0;  // requests not yet finished
    bool stop = false;
    int error = TFX_OK;
    std::string error_msg;
    std::thread worker;
    hipStream_t stream = nullptr;
    std::vector<Slot> slots;
    int next_slot = 0;
    std::deque<Pass> inflight;
    std::vector<hipEvent_t> event_pool;
    tfx_simulation shape;  // launch-shape policy of the fused launches
    std::vector<Instance*> graveyard;  // retired instances to free (worker only)
    bool freeing = false;              // the worker is freeing a batch taken from the graveyard
    // per pass and strand length, as the batched form; off by default: the
    // served passes are small, and deferring gained nothing there (the extra
    // launch per substep cost about as much as it saved)
    std::atomic<int> defer_sdf{TFX_SDF_DEFER_NEVER};
    // statistics since the last reset
    std::vector<double> latencies;
    long long completed_requests = 0, completed_frames = 0, vertex_substeps = 0, passes = 0;
    double members_sum = 0.0;
    double t_reset = 0.0;
};

namespace {

void free_instance(Instance* in) {
    if (!in) return;
    tfx_hair_destroy(in->hair);
    tfx_sdf_destroy(in->sdf);
    tfx_mesh_skinning_destroy(in->mesh);
    delete in;
}

hipEvent_t take_event(tfx_server* s) {
    if (!s->event_pool.empty()) {
        hipEvent_t e = s->event_pool.back();
        s->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
    return e;
}

// Marks the oldest pass complete (its event has completed).  Lock held.
void retire_pass(tfx_server* s) {
    Pass p = std::move(s->inflight.front());
    s->inflight.pop_front();
    const double t = now_s();
    s->passes += 1;
    s->members_sum += (double)p.frames.size();
    s->vertex_substeps += p.vertex_substeps;
    s->completed_frames += (long long)p.frames.size();
    for (auto& fr : p.frames) {
        Request* r = fr.first;
        Instance* in = fr.second;
        in->inflight -= 1;
        r->done += 1;
        if (r->done == r->frames) {
            const double ms = (t - r->t_submit) * 1e3;
            if (r->keep_result) {
                Result& res = s->results[r->ticket];
                res.finished = true;
                res.latency_ms = ms;
            }
            s->latencies.push_back(ms);
            s->completed_requests += 1;
            s->unfinished -= 1;
            // requests of one instance finish in submission order
            if (!in->queue.empty() && in->queue.front() == r) in->queue.pop_front();
            delete r;
        }
    }
    s->event_pool.push_back(p.ev);
    // retired instances with nothing left go to the graveyard
    for (auto it = s->instances.begin(); it != s->instances.end();) {
        Instance* in = it->second;
        if (in->retiring && in->queue.empty() && in->inflight == 0) {
            s->resident_bytes -= in->bytes;
            s->graveyard.push_back(in);
            it = s->instances.erase(it);
        } else {
            ++it;
        }
    }
    s->cv_done.notify_all();
}

void fail_all(tfx_server* s, int code, const std::string& msg) {
    s->error = code;
    s->error_msg = msg;
    for (auto& kv : s->instances) {
        for (Request* r : kv.second->queue) {
            if (r->keep_result) {
                Result& res = s->results[r->ticket];
                res.finished = true;
                res.status = code;
            }
            s->unfinished -= 1;
            delete r;
        }
        kv.second->queue.clear();
        kv.second->inflight = 0;
    }
    s->inflight.clear();
    s->cv_done.notify_all();
}

int ensure_slot(Slot& sl, size_t bytes) {
    if (bytes <= sl.bytes) return TFX_OK;
    if (sl.ev) HIP_TRY(hipEventSynchronize(sl.ev));
    if (sl.host) (void)hipHostFree(sl.host);
    if (sl.dev) (void)hipFree(sl.dev);
    sl.host = sl.dev = nullptr;
    sl.bytes = 0;
    const size_t cap = std::max(bytes + bytes / 2, size_t(1) << 20);
    HIP_TRY(hipHostMalloc(&sl.host, cap, hipHostMallocDefault));
    HIP_TRY(hipMalloc(&sl.dev, cap));
    sl.bytes = cap;
    return TFX_OK;
}

// Stages and launches one pass over `frames` (one frame of each listed
// request).  Runs without the lock: only the worker touches instance objects
// and a request's inputs are immutable after submission.
int launch_pass(tfx_server* s, Pass& pass) {
    const int n = (int)pass.frames.size();
    // Substep `sub` advances the members with more than `sub` substeps.
    int subs = 1;
    for (auto& fr : pass.frames) subs = std::max(subs, fr.second->substeps);
    // ---- layout
    std::vector<int> mesh_members, sdf_members;
    size_t mesh_bone_floats = 0, hair_bone_floats = 0;
    std::vector<size_t> mesh_bone_off(n, 0), hair_bone_off(n, 0);
    for (int m = 0; m < n; ++m) {
        Instance* in = pass.frames[m].second;
        if (in->mesh) {
            mesh_members.push_back(m);
            mesh_bone_off[m] = mesh_bone_floats;
            mesh_bone_floats += 16 * (size_t)in->mesh->bones_used;
        }
        if (in->sdf) sdf_members.push_back(m);
        hair_bone_off[m] = hair_bone_floats;
        hair_bone_floats += 16 * (size_t)in->hair->bones_used;
    }
    const size_t off_mjobs = 0;
    const size_t off_mbones = align256(off_mjobs + mesh_members.size() * sizeof(tfx::MeshSkinJob));
    const size_t off_sjobs = align256(off_mbones + mesh_bone_floats * sizeof(float));
    const size_t off_sub = align256(off_sjobs + sdf_members.size() * sizeof(tfx::SdfJob));
    const size_t sub_params = align256((size_t)n * sizeof(DevParams));
    const size_t sub_bytes = sub_params + align256(hair_bone_floats * sizeof(float));
    const size_t off_tables = off_sub + (size_t)subs * sub_bytes;
    const size_t off_items = align256(off_tables + (size_t)subs * n * sizeof(DevHair));
    // work items per strand length, members with more substeps first, so
    // that substep `sub` runs a prefix of each strand length's items
    std::vector<tfx::WorkItem> items;
    int vbegin[65] = {0};
    std::vector<int> vcount((size_t)65 * subs, 0);  // [V][sub]
    Shape shapes[65];
    std::vector<tfx_hair*> hairs(n);
    for (int m = 0; m < n; ++m) hairs[m] = pass.frames[m].second->hair;
    const int cus = cu_count(s->cfg.device);
    for (int V = 4; V <= 64; V *= 2) {
        vbegin[V] = (int)items.size();
        std::vector<int> mem;
        for (int m = 0; m < n; ++m)
            if (hairs[m]->V == V) mem.push_back(m);
        if (mem.empty()) continue;
        std::stable_sort(mem.begin(), mem.end(), [&](int a, int b) {
            return pass.frames[a].second->substeps > pass.frames[b].second->substeps;
        });
        shapes[V] = choose_shape(&s->shape, V, hairs.data(), n, cus);
        const int B = kl_fused_block_vertices(V, shapes[V].variant) / V;
        for (int m : mem) {
            for (int g0 = 0; g0 < hairs[m]->G; g0 += B) items.push_back(tfx::WorkItem{m, g0});
            const int upto = (int)items.size() - vbegin[V];
            for (int sub = 0; sub < pass.frames[m].second->substeps; ++sub) vcount[(size_t)V * subs + sub] = upto;
        }
    }
    // SDF collision deferred to a collide pass after each substep's fused
    // launches for the strand lengths whose blocks fit in one round of
    // resident blocks (as the batched form); collide items (member, first
    // vertex) ordered by member substeps, so substep `sub` runs a prefix
    int defer[65] = {0};
    const int dmode = s->defer_sdf.load();
    for (int V = 4; V <= 64; V *= 2) {
        const int cnt = vcount[(size_t)V * subs];
        defer[V] = cnt > 0 && (dmode == TFX_SDF_DEFER_ALWAYS || (dmode == TFX_SDF_DEFER_AUTO && cnt <= shapes[V].resident));
    }
    const int cbegin = (int)items.size();
    std::vector<int> ccount(subs, 0);
    {
        std::vector<int> cm;
        for (int m = 0; m < n; ++m)
            if (defer[hairs[m]->V] && pass.frames[m].second->sdf) cm.push_back(m);
        std::stable_sort(cm.begin(), cm.end(), [&](int a, int b) {
            return pass.frames[a].second->substeps > pass.frames[b].second->substeps;
        });
        const int chunk = kl_collide_chunk();
        for (int m : cm) {
            for (int v0 = 0; v0 < hairs[m]->N; v0 += chunk) items.push_back(tfx::WorkItem{m, v0});
            for (int sub = 0; sub < pass.frames[m].second->substeps; ++sub) ccount[sub] = (int)items.size() - cbegin;
        }
    }
    const size_t total = off_items + items.size() * sizeof(tfx::WorkItem);
    // ---- slot
    Slot& sl = s->slots[s->next_slot];
    pass.slot = s->next_slot;
    s->next_slot = (s->next_slot + 1) % (int)s->slots.size();
    if (sl.ev) HIP_TRY(hipEventSynchronize(sl.ev));
    int rc = ensure_slot(sl, total);
    if (rc) return rc;
    char* H = (char*)sl.host;
    char* D = (char*)sl.dev;
    // ---- collision meshes: bones of the frame, follow-bone bbox, skinning jobs
    int max_mesh_verts = 0, max_cells = 0, max_tris = 0;
    tfx::MeshSkinJob* mjobs = reinterpret_cast<tfx::MeshSkinJob*>(H + off_mjobs);
    for (size_t j = 0; j < mesh_members.size(); ++j) {
        const int m = mesh_members[j];
        Request* r = pass.frames[m].first;
        tfx_mesh_skinning* ms = pass.frames[m].second->mesh;
        if (!r->mesh_bones.empty()) {
            const int nb = std::min(r->nmb, TFX_MAX_NUM_BONES);
            std::memcpy(ms->bones_h.data(), &r->mesh_bones[(size_t)r->issued * r->nmb * 16],
                        (size_t)nb * 16 * sizeof(float));
        }
        mesh_follow_bbox(ms);
        float* mb = reinterpret_cast<float*>(H + off_mbones) + mesh_bone_off[m];
        std::memcpy(mb, ms->bones_h.data(), (size_t)16 * ms->bones_used * sizeof(float));
        tfx::MeshSkinJob& jb = mjobs[j];
        jb.out = ms->out;
        jb.rest = ms->rest;
        jb.skin = ms->skin;
        jb.bones = reinterpret_cast<const float*>(D + off_mbones) + mesh_bone_off[m];
        jb.n = ms->n;
        jb.pad = 0;
        max_mesh_verts = std::max(max_mesh_verts, ms->n);
    }
    // ---- SDFs: mesh of the frame, grid origin, build jobs
    tfx::SdfJob* sjobs = reinterpret_cast<tfx::SdfJob*>(H + off_sjobs);
    for (size_t j = 0; j < sdf_members.size(); ++j) {
        const int m = sdf_members[j];
        Instance* in = pass.frames[m].second;
        tfx_sdf* sd = in->sdf;
        if (in->mesh) {
            tfx_sdf_mesh_desc d;
            tfx_mesh_skinning_get_mesh(in->mesh, &d);
            sd->mesh = d;
        }
        if (!sd->explicit_grid)
            for (int a = 0; a < 3; ++a) sd->origin[a] = sd->mesh.bbox_min[a] - sd->pad_boundary;
        tfx::SdfJob& jb = sjobs[j];
        jb.cells = sd->cells;
        jb.verts = sd->mesh.vertices;
        jb.idx = sd->mesh.indices;
        jb.ntri = (sd->mesh.vertices && sd->mesh.indices) ? sd->mesh.num_triangles : 0;
        jb.ncell = sd->nx * sd->ny * sd->nz;
        jb.g = sdf_grid(sd);
        max_cells = std::max(max_cells, jb.ncell);
        max_tris = std::max(max_tris, jb.ntri);
    }
    const bool sparse = tfx::sdf_multi_sparse(max_cells, (int)sdf_members.size());
    for (size_t j = 0; j < sdf_members.size(); ++j) sdf_band_stage(pass.frames[sdf_members[j]].second->sdf, sjobs[j], sparse);
    // ---- substeps: parameters and bones of every member (UpdateBoneMatrices,
    // UpdateSimulationParameters, then IncreaseSimulationFrame)
    for (int sub = 0; sub < subs; ++sub) {
        char* sb = H + off_sub + (size_t)sub * sub_bytes;
        DevParams* params = reinterpret_cast<DevParams*>(sb);
        float* bone_block = reinterpret_cast<float*>(sb + sub_params);
        for (int m = 0; m < n; ++m) {
            Request* r = pass.frames[m].first;
            Instance* in = pass.frames[m].second;
            if (sub >= in->substeps) continue;  // no work item reads this member's parameters
            tfx_hair* h = in->hair;
            const size_t at = (size_t)r->issued * in->substeps + sub;
            if (r->nb > 0) {
                rc = tfx_hair_update_bone_matrices(h, &r->hair_bones[at * (size_t)r->nb * 16],
                                                   std::min(r->nb, TFX_MAX_NUM_BONES));
                if (rc) return rc;
            }
            tfx_simulation_settings st = in->settings;
            if (!r->wind.empty()) st.wind_magnitude = r->wind[at];
            rc = tfx_hair_update_simulation_parameters(h, &st, in->dt);
            if (rc) return rc;
            if (in->sdf && in->sdf->collide_variant != TFX_SDF_COLLIDE_FORWARD)
                return fail(TFX_ERR_UNSUPPORTED, "tfx_server: instance %d has a mixed-variant SDF; the served form "
                                                 "collides with forward differences only", in->id);
            pack_dev_params(h, in->sdf, params[m]);
            std::memcpy(bone_block + hair_bone_off[m], h->bones_h.data(), (size_t)16 * h->bones_used * sizeof(float));
            h->bones_dirty = TFX_MAX_NUM_BONES;
            tfx_hair_increase_simulation_frame(h);
        }
        DevHair* tables = reinterpret_cast<DevHair*>(H + off_tables) + (size_t)sub * n;
        const char* dsb = D + off_sub + (size_t)sub * sub_bytes;
        for (int m = 0; m < n; ++m) {
            Instance* in = pass.frames[m].second;
            DevHair dh = dev_hair(in->hair, in->sdf);
            dh.params = reinterpret_cast<const DevParams*>(dsb) + m;
            dh.bones = reinterpret_cast<const float*>(dsb + sub_params) + hair_bone_off[m];
            tables[m] = dh;
        }
    }
    if (!items.empty()) std::memcpy(H + off_items, items.data(), items.size() * sizeof(tfx::WorkItem));
    for (int m = 0; m < n; ++m) {
        Request* r = pass.frames[m].first;
        tfx_hair* h = pass.frames[m].second->hair;
        pass.vertex_substeps += (long long)h->N * pass.frames[m].second->substeps;
        r->issued += 1;
    }
    // ---- launches
    hipStream_t st = s->stream;
    HIP_TRY(hipMemcpyAsync(sl.dev, sl.host, total, hipMemcpyHostToDevice, st));
    if (!mesh_members.empty())
        HIP_TRY(tfx_launch_mesh_skin_multi(reinterpret_cast<const tfx::MeshSkinJob*>(D + off_mjobs),
                                           (int)mesh_members.size(), max_mesh_verts, st));
    if (!sdf_members.empty()) {
        const tfx::SdfJob* dj = reinterpret_cast<const tfx::SdfJob*>(D + off_sjobs);
        for (int phase = 0; phase < 4; ++phase)
            HIP_TRY(kl_launch_sdf_multi(phase, dj, (int)sdf_members.size(), max_cells, max_tris, st));
    }
    const DevHair* dtables = reinterpret_cast<const DevHair*>(D + off_tables);
    const tfx::WorkItem* ditems = reinterpret_cast<const tfx::WorkItem*>(D + off_items);
    for (int sub = 0; sub < subs; ++sub) {
        for (int V = 4; V <= 64; V *= 2) {
            const int cnt = vcount[(size_t)V * subs + sub];
            if (!cnt) continue;
            const int grid = std::min(cnt, shapes[V].resident);
            HIP_TRY(kl_launch_fused(V, shapes[V].variant, dtables + (size_t)sub * n, ditems + vbegin[V], cnt, grid,
                                     defer[V], s->shape.queue_d, st, nullptr, nullptr));
        }
        if (ccount[sub])
            HIP_TRY(kl_launch_collide_items(dtables + (size_t)sub * n, ditems + cbegin, ccount[sub], st, nullptr,
                                             nullptr));
    }
    HIP_TRY(hipEventRecord(pass.ev, st));
    if (!sl.ev) HIP_TRY(hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming));
    HIP_TRY(hipEventRecord(sl.ev, st));
    return TFX_OK;
}

void worker_main(tfx_server* s) {
    (void)hipSetDevice(s->cfg.device);
    g_device = s->cfg.device;
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        // free retired instances (outside the lock: frees may synchronise)
        if (!s->graveyard.empty()) {
            std::vector<Instance*> g;
            g.swap(s->graveyard);
            s->freeing = true;
            lk.unlock();
            for (Instance* in : g) free_instance(in);
            lk.lock();
            s->freeing = false;
            s->cv_done.notify_all();
            // instances retired while the lock was released: their
            // notifications found no waiter, so look again before sleeping
            continue;
        }
        // completed passes
        while (!s->inflight.empty() && hipEventQuery(s->inflight.front().ev) == hipSuccess) retire_pass(s);
        if (s->error) {
            if (s->stop) break;
            s->cv_work.wait(lk);
            continue;
        }
        // the next frame of every instance that has one
        Pass pass;
        if ((int)s->inflight.size() < s->cfg.max_inflight) {
            for (auto& kv : s->instances) {
                Instance* in = kv.second;
                for (Request* r : in->queue)
                    if (r->issued < r->frames) {
                        pass.frames.push_back({r, in});
                        break;
                    }
            }
        }
        if (pass.frames.empty()) {
            if (!s->inflight.empty()) {
                // wait for the oldest pass (without the lock)
                hipEvent_t ev = s->inflight.front().ev;
                lk.unlock();
                hipError_t e = hipEventSynchronize(ev);
                lk.lock();
                if (e != hipSuccess) fail_all(s, TFX_ERR_HIP, std::string("pass failed: ") + hipGetErrorString(e));
                continue;
            }
            if (!s->graveyard.empty()) continue;  // retire_pass above may have added some
            if (s->stop) break;
            s->cv_done.notify_all();  // idle
            s->cv_work.wait(lk);
            continue;
        }
        pass.ev = take_event(s);
        if (!pass.ev) {
            fail_all(s, TFX_ERR_HIP, "hipEventCreate failed");
            continue;
        }
        // instances of this pass keep their objects until the pass completes
        for (auto& fr : pass.frames) fr.second->inflight += 1;
        lk.unlock();
        const int rc = launch_pass(s, pass);
        lk.lock();
        if (rc) {
            // the pass may have advanced SDF band histories it never built
            for (auto& fr : pass.frames)
                if (fr.second->sdf) fr.second->sdf->band_valid = 0;
            fail_all(s, rc, tfx_last_error());
            continue;
        }
        s->inflight.push_back(std::move(pass));
    }
}

}  // namespace

extern "C" int tfx_server_create(const tfx_server_config* cfg, tfx_server** out) {
    CHECK_ARG(cfg && out, "tfx_server_create: null argument");
    CHECK_ARG(cfg->substeps_per_frame >= 1 && cfg->substeps_per_frame <= 16, "tfx_server_create: substeps_per_frame");
    CHECK_ARG(cfg->max_instances >= 1 && cfg->hbm_budget_bytes > 0, "tfx_server_create: limits");
    CHECK_ARG(cfg->substep_dt > 0.0f, "tfx_server_create: substep_dt");
    *out = nullptr;
    tfx_server* s = new tfx_server();
    s->cfg = *cfg;
    s->cfg.max_inflight = std::min(std::max(cfg->max_inflight, 1), 4);
    HIP_TRY(hipSetDevice(cfg->device));
    s->shape.device = cfg->device;
    for (int v = 0; v <= 64; ++v) s->shape.variant[v] = s->shape.per_cu[v] = -1;
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc((void**)&s->shape.queue_d, tfx::kFusedQueueWords * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(s->shape.queue_d, 0, tfx::kFusedQueueWords * sizeof(uint32_t));
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        if (s->stream) (void)hipStreamDestroy(s->stream);
        (void)hipFree(s->shape.queue_d);
        delete s;
        return fail(TFX_ERR_HIP, "tfx_server_create: %s", hipGetErrorString(e));
    }
    s->slots.resize(s->cfg.max_inflight + 1);
    s->t_reset = now_s();
    s->worker = std::thread(worker_main, s);
    *out = s;
    return TFX_OK;
}

extern "C" int tfx_server_destroy(tfx_server* s) {
    if (!s) return TFX_OK;
    tfx_server_drain(s);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    (void)hipStreamSynchronize(s->stream);
    for (auto& kv : s->instances) free_instance(kv.second);
    for (Instance* in : s->graveyard) free_instance(in);
    for (Slot& sl : s->slots) {
        if (sl.ev) (void)hipEventDestroy(sl.ev);
        if (sl.host) (void)hipHostFree(sl.host);
        if (sl.dev) (void)hipFree(sl.dev);
    }
    for (hipEvent_t e : s->event_pool) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(s->stream);
    (void)hipFree(s->shape.queue_d);
    delete s;
    return TFX_OK;
}

extern "C" int tfx_server_admit(tfx_server* s, const tfx_instance_desc* d, int* id) {
    CHECK_ARG(s && d && d->asset && id, "tfx_server_admit: null argument");
    CHECK_ARG(d->substeps >= 0 && d->substeps <= 16, "tfx_server_admit: substeps must be in [0, 16]");
    long long bytes = hair_bytes(d->asset);
    long long cells = 0;
    if (d->mesh) {
        bytes += mesh_bytes(d->mesh);
        if (d->sdf_dims[0] > 0) {
            CHECK_ARG(d->sdf_dims[1] > 0 && d->sdf_dims[2] > 0 && d->sdf_cell > 0.0f, "tfx_server_admit: SDF grid");
            cells = (long long)d->sdf_dims[0] * d->sdf_dims[1] * d->sdf_dims[2];
        } else {
            CHECK_ARG(d->sdf_num_cells_in_x > 0, "tfx_server_admit: SDF grid");
            const long long nx = d->sdf_num_cells_in_x;
            cells = (long long)(1.4 * (double)((nx * 26) / 10) * ((nx * 26) / 10) * ((nx * 26) / 10));  // upper bound
        }
        bytes += sdf_bytes(cells);
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->error) return fail(s->error, "tfx_server: %s", s->error_msg.c_str());
        if ((int)s->instances.size() >= s->cfg.max_instances)
            return fail(TFX_ERR_ADMISSION, "tfx_server_admit: %d instances resident (limit)", (int)s->instances.size());
        if (s->resident_bytes + bytes > s->cfg.hbm_budget_bytes)
            return fail(TFX_ERR_ADMISSION, "tfx_server_admit: %lld + %lld bytes exceed the budget of %lld",
                        s->resident_bytes, bytes, s->cfg.hbm_budget_bytes);
        s->resident_bytes += bytes;  // reserved while the objects are created
    }
    HIP_TRY(hipSetDevice(s->cfg.device));
    g_device = s->cfg.device;
    Instance* in = new Instance();
    in->settings = d->settings;
    in->bytes = bytes;
    // a frame lasts substeps_per_frame x substep_dt whatever the instance's substeps
    in->substeps = d->substeps > 0 ? d->substeps : s->cfg.substeps_per_frame;
    in->dt = in->substeps == s->cfg.substeps_per_frame
                 ? s->cfg.substep_dt
                 : (float)((double)s->cfg.substep_dt * s->cfg.substeps_per_frame / in->substeps);
    int rc = tfx_hair_create(d->asset, &in->hair);
    if (!rc && d->num_capsules > 0)
        rc = tfx_hair_set_collision_capsules(in->hair, d->capsules0, d->capsules1, d->num_capsules);
    if (!rc && d->mesh) {
        rc = tfx_mesh_skinning_create(d->mesh, &in->mesh);
        tfx_sdf_mesh_desc md;
        if (!rc) rc = tfx_mesh_skinning_get_mesh(in->mesh, &md);
        if (!rc) {
            if (d->sdf_dims[0] > 0)
                rc = tfx_sdf_create_explicit(&md, d->sdf_origin, d->sdf_cell, d->sdf_dims[0], d->sdf_dims[1],
                                             d->sdf_dims[2], d->collision_margin, &in->sdf);
            else
                rc = tfx_sdf_create(&md, d->sdf_num_cells_in_x, d->collision_margin, &in->sdf);
        }
    }
    std::lock_guard<std::mutex> lk(s->mu);
    if (rc) {
        s->resident_bytes -= bytes;
        std::string msg = tfx_last_error();
        free_instance(in);
        return fail(rc, "%s", msg.c_str());
    }
    in->id = s->next_id++;
    s->instances[in->id] = in;
    *id = in->id;
    return TFX_OK;
}

extern "C" int tfx_server_set_sdf_deferral(tfx_server* s, int mode) {
    CHECK_ARG(s, "tfx_server_set_sdf_deferral: null server");
    CHECK_ARG(mode >= TFX_SDF_DEFER_AUTO && mode <= TFX_SDF_DEFER_ALWAYS,
              "tfx_server_set_sdf_deferral: mode must be TFX_SDF_DEFER_AUTO, _NEVER or _ALWAYS");
    s->defer_sdf.store(mode);  // passes launched from now on
    return TFX_OK;
}

extern "C" int tfx_server_retire(tfx_server* s, int id) {
    CHECK_ARG(s, "tfx_server_retire: null server");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->instances.find(id);
        if (it == s->instances.end() || it->second->retiring)
            return fail(TFX_ERR_NOT_FOUND, "tfx_server_retire: no instance %d", id);
        Instance* in = it->second;
        in->retiring = true;
        if (in->queue.empty() && in->inflight == 0) {
            s->resident_bytes -= in->bytes;
            s->graveyard.push_back(in);
            s->instances.erase(it);
        }
    }
    s->cv_work.notify_all();
    return TFX_OK;
}

extern "C" int tfx_server_submit(tfx_server* s, int id, int frames, const float* hair_bones, int num_bones,
                                 const float* mesh_bones, int num_mesh_bones, const float* wind,
                                 unsigned long long* ticket) {
    CHECK_ARG(s && frames > 0, "tfx_server_submit: bad argument");
    CHECK_ARG(num_bones >= 0 && num_mesh_bones >= 0 && (hair_bones || num_bones == 0) &&
                  (mesh_bones || num_mesh_bones == 0),
              "tfx_server_submit: bad bone arguments");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->error) return fail(s->error, "tfx_server: %s", s->error_msg.c_str());
        auto it = s->instances.find(id);
        if (it == s->instances.end() || it->second->retiring)
            return fail(TFX_ERR_NOT_FOUND, "tfx_server_submit: no instance %d", id);
        // the inputs are sized by the instance's substeps (fixed at admission)
        const int subs = it->second->substeps;
        Request* r = new Request();
        r->frames = frames;
        r->nb = num_bones;
        r->nmb = mesh_bones ? num_mesh_bones : 0;
        if (num_bones) r->hair_bones.assign(hair_bones, hair_bones + (size_t)frames * subs * num_bones * 16);
        if (r->nmb) r->mesh_bones.assign(mesh_bones, mesh_bones + (size_t)frames * r->nmb * 16);
        if (wind) r->wind.assign(wind, wind + (size_t)frames * subs);
        r->ticket = s->next_ticket++;
        r->t_submit = now_s();
        r->keep_result = ticket != nullptr;
        it->second->queue.push_back(r);
        if (ticket) {
            s->results[r->ticket] = Result{};
            *ticket = r->ticket;
        }
        s->unfinished += 1;
    }
    s->cv_work.notify_all();
    return TFX_OK;
}

extern "C" int tfx_server_poll(tfx_server* s, unsigned long long ticket, int* done, double* latency_ms) {
    CHECK_ARG(s && done, "tfx_server_poll: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->results.find(ticket);
    if (it == s->results.end()) return fail(TFX_ERR_NOT_FOUND, "tfx_server_poll: unknown ticket");
    *done = it->second.finished ? 1 : 0;
    if (latency_ms) *latency_ms = it->second.latency_ms;
    if (it->second.finished) {
        const int st = it->second.status;
        s->results.erase(it);
        if (st) return fail(st, "tfx_server: request failed: %s", s->error_msg.c_str());
    }
    return TFX_OK;
}

extern "C" int tfx_server_wait(tfx_server* s, unsigned long long ticket, double* latency_ms) {
    CHECK_ARG(s, "tfx_server_wait: null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->results.find(ticket);
    if (it == s->results.end()) return fail(TFX_ERR_NOT_FOUND, "tfx_server_wait: unknown ticket");
    s->cv_done.wait(lk, [&] { return s->results[ticket].finished; });
    it = s->results.find(ticket);
    if (latency_ms) *latency_ms = it->second.latency_ms;
    const int st = it->second.status;
    s->results.erase(it);
    if (st) return fail(st, "tfx_server: request failed: %s", s->error_msg.c_str());
    return TFX_OK;
}

extern "C" int tfx_server_drain(tfx_server* s) {
    CHECK_ARG(s, "tfx_server_drain: null server");
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_work.notify_all();
    // retired instances' memory is released too, so that frees (which
    // synchronise the device) do not land in the work submitted next
    s->cv_done.wait(lk, [&] {
        return (s->unfinished == 0 && s->inflight.empty() && s->graveyard.empty() && !s->freeing) ||
               s->error != TFX_OK;
    });
    if (s->error) return fail(s->error, "tfx_server: %s", s->error_msg.c_str());
    return TFX_OK;
}

extern "C" int tfx_server_stats_read(tfx_server* s, tfx_server_stats* o, int reset) {
    CHECK_ARG(s && o, "tfx_server_stats_read: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    std::memset(o, 0, sizeof(*o));
    o->completed_requests = s->completed_requests;
    o->completed_frames = s->completed_frames;
    o->vertex_substeps = s->vertex_substeps;
    o->passes = s->passes;
    o->mean_instances_per_pass = s->passes ? s->members_sum / (double)s->passes : 0.0;
    if (!s->latencies.empty()) {
        std::vector<double> l = s->latencies;
        std::sort(l.begin(), l.end());
        auto pct = [&](double q) {
            const double pos = q * (double)(l.size() - 1);
            const size_t lo = (size_t)pos;
            const size_t hi = std::min(lo + 1, l.size() - 1);
            return l[lo] + (l[hi] - l[lo]) * (pos - (double)lo);
        };
        o->latency_p50_ms = pct(0.50);
        o->latency_p99_ms = pct(0.99);
        double sum = 0.0;
        for (double x : l) sum += x;
        o->latency_mean_ms = sum / (double)l.size();
        o->latency_max_ms = l.back();
    }
    o->elapsed_s = now_s() - s->t_reset;
    o->resident_bytes = s->resident_bytes;
    o->resident_instances = (int)s->instances.size();
    o->queued_requests = s->unfinished;
    if (reset) {
        s->latencies.clear();
        s->completed_requests = s->completed_frames = s->vertex_substeps = s->passes = 0;
        s->members_sum = 0.0;
        s->t_reset = now_s();
    }
    return TFX_OK;
}

extern "C" int tfx_server_get_objects(tfx_server* s, int id, tfx_hair** hair, tfx_sdf** sdf) {
    CHECK_ARG(s, "tfx_server_get_objects: null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->instances.find(id);
    if (it == s->instances.end()) return fail(TFX_ERR_NOT_FOUND, "tfx_server_get_objects: no instance %d", id);
    Instance* in = it->second;
    CHECK_ARG(in->queue.empty() && in->inflight == 0, "tfx_server_get_objects: instance has queued frames");
    if (hair) *hair = in->hair;
    if (sdf) *sdf = in->sdf;
    return TFX_OK;
}
