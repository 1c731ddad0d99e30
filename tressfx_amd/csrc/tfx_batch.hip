// tfx_batch.hip — the batched form: many hair objects of uneven size advance
// together.  Per substep, one persistent fused launch per strand-length class
// covers every member through an (object, strand block) work list; per frame,
// every member's SDF is rebuilt by one launch per build phase.  The frame —
// one parameter upload, the SDF build and all substeps — can be captured once
// into a hipGraph and replayed while membership is unchanged.
//
// The per-frame order of the reference (src/Simulation.cpp:50-131) is kept
// for every member: SDF Update, then per substep Simulate followed by
// CollideWithHair (each member collides with its own SDF).
#include "tfx_internal.h"

namespace {

constexpr int kBatchMaxSubsteps = 16;

struct Member {
    tfx_hair* hair;
    tfx_sdf* sdf;
};

}  // namespace

struct tfx_batch {
    int device = 0;
    std::vector<Member> members;
    bool dirty = true;                 // membership changed: tables, work lists and graphs are stale
    int cap = 0;                       // substep slots laid out
    tfx_simulation shape;              // launch-shape policy (variants per strand length)

    // Slot layout (host pinned slot and its device copy are identical):
    //   SdfJob[nsdf] | substep 0 | substep 1 | ... | substep cap-1
    // with substep s = DevParams[n] | bones float[bone_floats], so that one
    // copy of the first off_sub + staged * sub_bytes bytes moves a frame.
    size_t off_sub = 0, sub_params = 0, sub_bytes = 0, slot_bytes = 0;
    std::vector<int> bone_off;         // per member, floats into one substep's bone block
    int bone_floats = 0;
    std::vector<int> sdf_members;      // per distinct SDF, the first member that collides with it
    int max_cells = 0, max_tris = 0;

    void* dev_slot = nullptr;          // device copy of a slot
    DevHair* dev_tables = nullptr;     // [cap][n]
    tfx::WorkItem* dev_items = nullptr;
    uint32_t* dev_queue = nullptr;     // the fused launches' work queues (kFusedQueueWords)
    int vbegin[65] = {0}, vcount[65] = {0}, vvariant[65] = {0}, vresident[65] = {0};
    int vdefer[65] = {0};              // class V's fused launch leaves the SDF collision to the collide pass
    int cbegin = 0, ccount = 0;        // collide pass work items (member, first vertex) in dev_items

    void* host[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    size_t host_bytes = 0;
    int slot = -1;                     // slot being staged, -1 = none
    int next_slot = 0;
    int staged = 0;                    // substeps staged into `slot`

    hipGraphExec_t exec[2] = {nullptr, nullptr};
    int exec_substeps[2] = {0, 0};     // substeps captured in exec[k]
    int exec_tris[2] = {0, 0};         // SDF construct grid (largest mesh) captured in exec[k]
    int exec_arith[2] = {0, 0};        // kernel build (TFX_ARITH_*) captured in exec[k]

    // SDF build beside the first substep's deferring fused launches (which
    // read no SDF): a second stream forked from and joined to the frame's
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};

static void release_device(tfx_batch* b) {
    (void)hipFree(b->dev_slot);
    (void)hipFree(b->dev_tables);
    (void)hipFree(b->dev_items);
    (void)hipFree(b->dev_queue);
    b->dev_slot = nullptr;
    b->dev_tables = nullptr;
    b->dev_items = nullptr;
    b->dev_queue = nullptr;
    for (int k = 0; k < 2; ++k) {
        if (b->exec[k]) (void)hipGraphExecDestroy(b->exec[k]);
        b->exec[k] = nullptr;
        b->exec_substeps[k] = 0;
        b->exec_tris[k] = 0;
    }
}

static void release_host(tfx_batch* b) {
    for (int k = 0; k < 2; ++k) {
        if (b->ev[k]) { (void)hipEventSynchronize(b->ev[k]); (void)hipEventDestroy(b->ev[k]); b->ev[k] = nullptr; }
        if (b->host[k]) { (void)hipHostFree(b->host[k]); b->host[k] = nullptr; }
    }
    b->host_bytes = 0;
}

// Work items (member, first strand) of the strand blocks of every member of
// strand length V, for blocks of B strands: large objects first, so their
// blocks start early and the tail is made of the small ones.
static void class_items(const tfx_batch* b, int V, int B, std::vector<tfx::WorkItem>& items) {
    const int n = (int)b->members.size();
    std::vector<int> order;
    for (int m = 0; m < n; ++m)
        if (b->members[m].hair->V == V) order.push_back(m);
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return b->members[x].hair->G > b->members[y].hair->G; });
    for (int m : order)
        for (int g0 = 0; g0 < b->members[m].hair->G; g0 += B) items.push_back(tfx::WorkItem{m, g0});
}

// Work lists per strand-length class (identical for every substep), for the
// launch shapes the batch's policy gives.
static int build_work_lists(tfx_batch* b) {
    const int n = (int)b->members.size();
    std::vector<tfx_hair*> hairs(n);
    for (int m = 0; m < n; ++m) hairs[m] = b->members[m].hair;
    std::vector<tfx::WorkItem> items;
    const int cus = cu_count(b->device);
    for (int V = 4; V <= 64; V *= 2) {
        b->vbegin[V] = (int)items.size();
        b->vcount[V] = 0;
        long long guides = 0;
        for (int m = 0; m < n; ++m)
            if (hairs[m]->V == V) guides += hairs[m]->G;
        if (!guides) continue;
        const Shape sh = choose_shape(&b->shape, V, hairs.data(), n, cus);
        b->vvariant[V] = sh.variant;
        b->vresident[V] = sh.resident;
        class_items(b, V, kl_fused_block_vertices(V, sh.variant) / V, items);
        b->vcount[V] = (int)items.size() - b->vbegin[V];
    }
    // SDF collision deferred to a collide pass after each substep's fused
    // launches, per class: by default where the class's strand blocks fit in
    // one round of resident blocks, so that the fused launch's duration is one
    // block's latency and the lookups, taken inside a strand block one group
    // after another, would add to it in full.
    b->cbegin = (int)items.size();
    const int chunk = kl_collide_chunk();
    for (int V = 4; V <= 64; V *= 2) {
        const int mode = b->shape.defer_sdf;
        b->vdefer[V] = b->vcount[V] > 0 && (mode == 1 || (mode < 0 && b->vcount[V] <= b->vresident[V]));
        if (!b->vdefer[V]) continue;
        for (int m = 0; m < n; ++m) {
            const tfx_hair* h = hairs[m];
            if (h->V != V || !b->members[m].sdf) continue;
            const long long nv = (long long)h->G * (h->F + 1) * V;
            for (long long v0 = 0; v0 < nv; v0 += chunk) items.push_back(tfx::WorkItem{m, (int)v0});
        }
    }
    b->ccount = (int)items.size() - b->cbegin;
    if (b->dev_items) {
        HIP_TRY(hipDeviceSynchronize());
        HIP_TRY(hipFree(b->dev_items));
        b->dev_items = nullptr;
    }
    if (!items.empty()) {
        HIP_TRY(hipMalloc((void**)&b->dev_items, items.size() * sizeof(tfx::WorkItem)));
        HIP_TRY(hipMemcpy(b->dev_items, items.data(), items.size() * sizeof(tfx::WorkItem), hipMemcpyHostToDevice));
    }
    return TFX_OK;
}

// Lays out slots, tables and work lists for the current membership.
static int prepare(tfx_batch* b) {
    if (!b->dirty) return TFX_OK;
    HIP_TRY(hipDeviceSynchronize());
    release_device(b);
    const int n = (int)b->members.size();
    b->cap = kBatchMaxSubsteps;
    b->bone_off.assign(n, 0);
    b->bone_floats = 0;
    b->sdf_members.clear();
    b->max_cells = b->max_tris = 0;
    for (int m = 0; m < n; ++m) {
        b->bone_off[m] = b->bone_floats;
        b->bone_floats += 16 * b->members[m].hair->bones_used;
        // One build per distinct SDF: members may share one (several hair
        // objects colliding with one head), and it is built once per frame.
        tfx_sdf* s = b->members[m].sdf;
        bool first = s != nullptr;
        for (int j : b->sdf_members) first = first && b->members[j].sdf != s;
        if (first) {
            b->sdf_members.push_back(m);
            b->max_cells = std::max(b->max_cells, s->nx * s->ny * s->nz);
        }
    }
    auto align = [](size_t x) { return (x + 255) & ~size_t(255); };
    b->off_sub = align(b->sdf_members.size() * sizeof(tfx::SdfJob));
    b->sub_params = align((size_t)n * sizeof(DevParams));
    b->sub_bytes = b->sub_params + align((size_t)b->bone_floats * sizeof(float));
    b->slot_bytes = b->off_sub + (size_t)b->cap * b->sub_bytes;
    if (b->slot_bytes > b->host_bytes) {
        release_host(b);
        for (int k = 0; k < 2; ++k) {
            HIP_TRY(hipHostMalloc(&b->host[k], b->slot_bytes, hipHostMallocDefault));
            HIP_TRY(hipEventCreateWithFlags(&b->ev[k], hipEventDisableTiming));
        }
        b->host_bytes = b->slot_bytes;
    }
    if (n > 0) HIP_TRY(hipMalloc(&b->dev_slot, b->slot_bytes));
    // DevHair tables: member m at substep s reads its parameters and bones from slot s.
    std::vector<DevHair> tables((size_t)b->cap * n);
    for (int s = 0; s < b->cap; ++s)
        for (int m = 0; m < n; ++m) {
            DevHair d = dev_hair(b->members[m].hair, b->members[m].sdf);
            char* sub = (char*)b->dev_slot + b->off_sub + (size_t)s * b->sub_bytes;
            d.params = reinterpret_cast<const DevParams*>(sub) + m;
            d.bones = reinterpret_cast<const float*>(sub + b->sub_params) + b->bone_off[m];
            tables[(size_t)s * n + m] = d;
        }
    if (n > 0) {
        HIP_TRY(hipMalloc((void**)&b->dev_tables, tables.size() * sizeof(DevHair)));
        HIP_TRY(hipMemcpy(b->dev_tables, tables.data(), tables.size() * sizeof(DevHair), hipMemcpyHostToDevice));
        HIP_TRY(hipMalloc((void**)&b->dev_queue, tfx::kFusedQueueWords * sizeof(uint32_t)));
        HIP_TRY(hipMemset(b->dev_queue, 0, tfx::kFusedQueueWords * sizeof(uint32_t)));
        HIP_TRY(hipDeviceSynchronize());
    }
    int rc = build_work_lists(b);
    if (rc) return rc;
    b->dirty = false;
    return TFX_OK;
}

extern "C" int tfx_batch_create(tfx_batch** out) {
    CHECK_ARG(out, "tfx_batch_create: null argument");
    tfx_batch* b = new tfx_batch();
    b->device = g_device;
    b->shape.device = g_device;
    for (int v = 0; v <= 64; ++v) b->shape.variant[v] = b->shape.per_cu[v] = -1;
    // created here, not inside a frame that may be under graph capture
    if (hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->join, hipEventDisableTiming) != hipSuccess) {
        tfx_batch_destroy(b);
        return fail(TFX_ERR_HIP, "tfx_batch_create: stream or event creation failed");
    }
    *out = b;
    return TFX_OK;
}

extern "C" int tfx_batch_destroy(tfx_batch* b) {
    if (!b) return TFX_OK;
    (void)hipDeviceSynchronize();
    release_device(b);
    release_host(b);
    if (b->side) (void)hipStreamDestroy(b->side);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->join) (void)hipEventDestroy(b->join);
    delete b;
    return TFX_OK;
}

extern "C" int tfx_batch_add(tfx_batch* b, tfx_hair* hair, tfx_sdf* sdf) {
    CHECK_ARG(b && hair, "tfx_batch_add: null argument");
    CHECK_ARG(b->slot < 0, "tfx_batch_add: a frame is being staged");
    for (const Member& m : b->members)
        if (m.hair == hair) return fail(TFX_ERR_INVALID_ARGUMENT, "tfx_batch_add: hair already a member");
    b->members.push_back(Member{hair, sdf});
    b->dirty = true;
    return TFX_OK;
}

extern "C" int tfx_batch_remove(tfx_batch* b, tfx_hair* hair) {
    CHECK_ARG(b && hair, "tfx_batch_remove: null argument");
    CHECK_ARG(b->slot < 0, "tfx_batch_remove: a frame is being staged");
    for (size_t i = 0; i < b->members.size(); ++i)
        if (b->members[i].hair == hair) {
            b->members.erase(b->members.begin() + i);
            b->dirty = true;
            return TFX_OK;
        }
    return fail(TFX_ERR_NOT_FOUND, "tfx_batch_remove: not a member");
}

extern "C" int tfx_batch_count(const tfx_batch* b, int* n) {
    CHECK_ARG(b && n, "tfx_batch_count: null argument");
    *n = (int)b->members.size();
    return TFX_OK;
}

extern "C" int tfx_batch_set_sdf_deferral(tfx_batch* b, int mode) {
    CHECK_ARG(b, "tfx_batch_set_sdf_deferral: null batch");
    CHECK_ARG(mode >= TFX_SDF_DEFER_AUTO && mode <= TFX_SDF_DEFER_ALWAYS,
              "tfx_batch_set_sdf_deferral: mode must be TFX_SDF_DEFER_AUTO, _NEVER or _ALWAYS");
    b->shape.defer_sdf = mode;
    b->dirty = true;
    return TFX_OK;
}

extern "C" int tfx_batch_set_launch_shape(tfx_batch* b, int V, int variant, int blocks_per_cu) {
    CHECK_ARG(b, "tfx_batch_set_launch_shape: null batch");
    int rc = tfx_simulation_set_launch_shape(&b->shape, V, variant, blocks_per_cu);
    if (rc) return rc;
    b->dirty = true;
    return TFX_OK;
}

// Starts a frame: takes a free staging slot and moves every member's SDF grid
// to its mesh's current bounding box (TressFXSDFCollision::UpdateSDFGrid,
// TressFXSDFCollision.cpp:88-93).
extern "C" int tfx_batch_begin_frame(tfx_batch* b) {
    CHECK_ARG(b, "tfx_batch_begin_frame: null batch");
    CHECK_ARG(b->slot < 0, "tfx_batch_begin_frame: previous frame not run");
    for (size_t m = 0; m < b->members.size(); ++m)
        if (b->members[m].sdf && b->members[m].sdf->collide_variant != TFX_SDF_COLLIDE_FORWARD)
            return fail(TFX_ERR_UNSUPPORTED, "tfx_batch_begin_frame: the batched form collides with forward "
                                             "differences only (member %d has a mixed-variant SDF)", (int)m);
    int rc = prepare(b);
    if (rc) return rc;
    const int k = b->next_slot;
    b->next_slot ^= 1;
    if (b->ev[k]) HIP_TRY(hipEventSynchronize(b->ev[k]));
    b->slot = k;
    b->staged = 0;
    tfx::SdfJob* jobs = reinterpret_cast<tfx::SdfJob*>(b->host[k]);
    b->max_tris = 0;
    for (size_t j = 0; j < b->sdf_members.size(); ++j) {
        tfx_sdf* s = b->members[b->sdf_members[j]].sdf;
        if (!s->explicit_grid)
            for (int a = 0; a < 3; ++a) s->origin[a] = s->mesh.bbox_min[a] - s->pad_boundary;
        tfx::SdfJob& jb = jobs[j];
        jb.cells = s->cells;
        jb.verts = s->mesh.vertices;
        jb.idx = s->mesh.indices;
        jb.ntri = (s->mesh.vertices && s->mesh.indices) ? s->mesh.num_triangles : 0;
        jb.ncell = s->nx * s->ny * s->nz;
        jb.g = sdf_grid(s);
        b->max_tris = std::max(b->max_tris, jb.ntri);
    }
    return TFX_OK;
}

// Stages one substep for every member: bone matrices and
// UpdateSimulationParameters (TressFXHairObject.cpp:366-437) per member, the
// packed block copied into the slot, then IncreaseSimulationFrame.
//   bones: member m's matrices at bones + m * bones_stride (floats), up to its bones_used
//   settings: one per member (n_settings == member count) or one for all
//   wind_magnitude: per member override of settings.wind_magnitude, or NULL
extern "C" int tfx_batch_stage_substep(tfx_batch* b, const float* bones, int bones_stride,
                                       const tfx_simulation_settings* settings, int n_settings, float time_step,
                                       const float* wind_magnitude) {
    CHECK_ARG(b && settings, "tfx_batch_stage_substep: null argument");
    CHECK_ARG(b->slot >= 0, "tfx_batch_stage_substep: call tfx_batch_begin_frame first");
    CHECK_ARG(b->staged < b->cap, "tfx_batch_stage_substep: too many substeps in one frame");
    const int n = (int)b->members.size();
    CHECK_ARG(n_settings == 1 || n_settings == n, "tfx_batch_stage_substep: settings count");
    CHECK_ARG(!bones || bones_stride >= 16, "tfx_batch_stage_substep: bones stride below one matrix");
    char* sub = (char*)b->host[b->slot] + b->off_sub + (size_t)b->staged * b->sub_bytes;
    DevParams* params = reinterpret_cast<DevParams*>(sub);
    float* bone_block = reinterpret_cast<float*>(sub + b->sub_params);
    for (int m = 0; m < n; ++m) {
        tfx_hair* h = b->members[m].hair;
        if (bones) {
            const int nb = std::min(bones_stride / 16, TFX_MAX_NUM_BONES);
            int rc = tfx_hair_update_bone_matrices(h, bones + (size_t)m * bones_stride, nb);
            if (rc) return rc;
        }
        tfx_simulation_settings st = settings[n_settings == 1 ? 0 : m];
        if (wind_magnitude) st.wind_magnitude = wind_magnitude[m];
        int rc = tfx_hair_update_simulation_parameters(h, &st, time_step);
        if (rc) return rc;
        pack_dev_params(h, b->members[m].sdf, params[m]);
        std::memcpy(bone_block + b->bone_off[m], h->bones_h.data(), (size_t)16 * h->bones_used * sizeof(float));
        h->bones_dirty = TFX_MAX_NUM_BONES;  // the object's own table is stale for the plain path
        tfx_hair_increase_simulation_frame(h);
    }
    b->staged++;
    return TFX_OK;
}

static int launch_frame(tfx_batch* b, hipStream_t s, int substeps, int timers) {
    const int n = (int)b->members.size();
    HIP_TRY(hipMemcpyAsync(b->dev_slot, b->host[b->slot], b->off_sub + (size_t)substeps * b->sub_bytes,
                           hipMemcpyHostToDevice, s));
    const int nsdf = (int)b->sdf_members.size();
    bool any_defer = false;
    for (int V = 4; V <= 64; V *= 2) any_defer = any_defer || (b->vcount[V] && b->vdefer[V]);
    const bool overlap = nsdf && any_defer;
    hipStream_t sb = s;  // stream of the SDF build
    if (overlap) {
        HIP_TRY(hipEventRecord(b->fork, s));
        HIP_TRY(hipStreamWaitEvent(b->side, b->fork, 0));
        sb = b->side;
    }
    if (nsdf) {
        const tfx::SdfJob* jobs = reinterpret_cast<const tfx::SdfJob*>(b->dev_slot);
        static const char* kNames[4] = {"InitializeSignedDistanceField", "ConstructSignedDistanceField",
                                        "FinalizeSignedDistanceField", "FinalizeSignedDistanceField"};
        for (int phase = 0; phase < 4; ++phase) {
            PassTimer tm(timers ? kNames[phase] : nullptr, sb);
            HIP_TRY(kl_launch_sdf_multi(phase, jobs, nsdf, b->max_cells, b->max_tris, sb));
        }
    }
    if (overlap) HIP_TRY(hipEventRecord(b->join, b->side));
    for (int sub = 0; sub < substeps; ++sub) {
        const DevHair* tables = b->dev_tables + (size_t)sub * n;
        // substep 0: the deferring classes first, beside the SDF build; the
        // build joins before any launch that reads the SDF
        if (overlap && sub == 0) {
            for (int V = 4; V <= 64; V *= 2) {
                if (!b->vcount[V] || !b->vdefer[V]) continue;
                const int grid = std::min(b->vcount[V], b->vresident[V]);
                PassTimer tm(timers ? fused_label(V, kl_fused_block_vertices(V, b->vvariant[V])) : nullptr, s, true);
                HIP_TRY(tm.launched(kl_launch_fused(V, b->vvariant[V], tables, b->dev_items + b->vbegin[V],
                                                    b->vcount[V], grid, 1, b->dev_queue, s, tm.start(),
                                                    tm.stop())));
            }
            HIP_TRY(hipStreamWaitEvent(s, b->join, 0));
        }
        for (int V = 4; V <= 64; V *= 2) {
            if (!b->vcount[V]) continue;
            if (overlap && sub == 0 && b->vdefer[V]) continue;
            const int grid = std::min(b->vcount[V], b->vresident[V]);
            PassTimer tm(timers ? fused_label(V, kl_fused_block_vertices(V, b->vvariant[V])) : nullptr, s, true);
            HIP_TRY(tm.launched(kl_launch_fused(V, b->vvariant[V], tables, b->dev_items + b->vbegin[V],
                                                b->vcount[V], grid, b->vdefer[V], b->dev_queue, s, tm.start(),
                                                tm.stop())));
        }
        if (b->ccount) {
            PassTimer tm(timers ? "CollideHairVerticesWithSdf" : nullptr, s, true);
            HIP_TRY(tm.launched(
                kl_launch_collide_items(tables, b->dev_items + b->cbegin, b->ccount, s, tm.start(), tm.stop())));
        }
    }
    // A frame without substeps still joins the forked build (substep 0 joins it otherwise).
    if (overlap && substeps == 0) HIP_TRY(hipStreamWaitEvent(s, b->join, 0));
    return TFX_OK;
}

// Autotuner of the launch shapes: per strand-length class, every distinct
// strand-block size and every persistent-blocks-per-CU count up to the
// kernel's occupancy runs the class's fused launch of the staged frame's first
// substep (followed by the class's collide pass where that shape defers the
// SDF collision) `repeats` times (HIP events on the stream, fastest run kept), each
// from the members' state as it was before the call (a snapshot, restored
// before every run and at the end), and the fastest shape is pinned for the
// class.
extern "C" int tfx_batch_autotune(tfx_batch* b, void* stream, int repeats, tfx_autotune_record* out, int max_out,
                                  int* num_out) {
    CHECK_ARG(b && num_out && (out || max_out == 0), "tfx_batch_autotune: null argument");
    CHECK_ARG(b->slot >= 0 && b->staged > 0, "tfx_batch_autotune: stage a frame first (begin_frame, stage_substep)");
    CHECK_ARG(repeats >= 1 && repeats <= 16, "tfx_batch_autotune: repeats must be 1..16");
    *num_out = 0;
    const int n = (int)b->members.size();
    if (n == 0) return TFX_OK;
    hipStream_t s = (hipStream_t)stream;
    // the staged parameters of substep 0 on the device (run_frame uploads them again)
    HIP_TRY(hipMemcpyAsync(b->dev_slot, b->host[b->slot], b->off_sub + b->sub_bytes, hipMemcpyHostToDevice, s));
    // snapshot of the state the fused kernel writes
    struct Buf { float4* p; size_t bytes; size_t at; int V; };
    std::vector<Buf> bufs;
    size_t total = 0;
    for (int m = 0; m < n; ++m) {
        tfx_hair* h = b->members[m].hair;
        for (float4* p : {h->pos, h->prev, h->prevprev, h->tangent}) {
            bufs.push_back(Buf{p, (size_t)h->N * sizeof(float4), total, h->V});
            total += (size_t)h->N * sizeof(float4);
        }
    }
    char* snap = nullptr;
    tfx::WorkItem* items_d = nullptr;
    tfx::WorkItem* citems_d = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = TFX_OK;
    auto cleanup = [&]() {
        if (snap) {
            for (const Buf& x : bufs) (void)hipMemcpyAsync(x.p, snap + x.at, x.bytes, hipMemcpyDeviceToDevice, s);
            (void)hipStreamSynchronize(s);
        }
        (void)hipFree(snap);
        (void)hipFree(items_d);
        (void)hipFree(citems_d);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    };
    auto hip_fail = [&](hipError_t e, const char* what) {
        cleanup();
        return fail(e == hipErrorOutOfMemory ? TFX_ERR_OUT_OF_MEMORY : TFX_ERR_HIP, "tfx_batch_autotune: %s: %s",
                    what, hipGetErrorString(e));
    };
    hipError_t e = hipMalloc((void**)&snap, std::max<size_t>(total, 1));
    if (e != hipSuccess) { snap = nullptr; return hip_fail(e, "snapshot"); }
    for (const Buf& x : bufs)
        if ((e = hipMemcpyAsync(snap + x.at, x.p, x.bytes, hipMemcpyDeviceToDevice, s)) != hipSuccess)
            return hip_fail(e, "snapshot");
    if ((e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess)
        return hip_fail(e, "events");
    std::vector<tfx_hair*> hairs(n);
    for (int m = 0; m < n; ++m) hairs[m] = b->members[m].hair;
    const int cus = cu_count(b->device);
    size_t items_cap = 0, citems_cap = 0;
    const int chunk = kl_collide_chunk();
    for (int V = 4; V <= 64; V *= 2) {
        bool any = false;
        for (int m = 0; m < n; ++m) any = any || hairs[m]->V == V;
        if (!any) continue;
        const Shape def = choose_shape(&b->shape, V, hairs.data(), n, cus);
        const int def_per_cu = std::max(1, def.resident / std::max(cus, 1));
        tfx_autotune_record rec{V, def.variant, def_per_cu, 0, 0.0f, 0.0f, 0};
        // the class's collide pass, for the shapes that defer the SDF collision
        std::vector<tfx::WorkItem> citems;
        for (int m = 0; m < n; ++m)
            if (hairs[m]->V == V && b->members[m].sdf)
                for (long long v0 = 0; v0 < (long long)hairs[m]->N; v0 += chunk)
                    citems.push_back(tfx::WorkItem{m, (int)v0});
        if (citems.size() > citems_cap) {
            (void)hipFree(citems_d);
            citems_d = nullptr;
            if ((e = hipMalloc((void**)&citems_d, citems.size() * sizeof(tfx::WorkItem))) != hipSuccess)
                return hip_fail(e, "work list");
            citems_cap = citems.size();
        }
        if (!citems.empty() && (e = hipMemcpy(citems_d, citems.data(), citems.size() * sizeof(tfx::WorkItem),
                                              hipMemcpyHostToDevice)) != hipSuccess)
            return hip_fail(e, "work list");
        float best = -1.0f;
        for (int variant = 0; variant <= 2; ++variant) {
            const int vb = kl_fused_block_vertices(V, variant);
            if (variant > 0 && vb == kl_fused_block_vertices(V, variant - 1)) continue;
            std::vector<tfx::WorkItem> items;
            class_items(b, V, vb / V, items);
            if (items.size() > items_cap) {
                (void)hipFree(items_d);
                items_d = nullptr;
                if ((e = hipMalloc((void**)&items_d, items.size() * sizeof(tfx::WorkItem))) != hipSuccess)
                    return hip_fail(e, "work list");
                items_cap = items.size();
            }
            if ((e = hipMemcpy(items_d, items.data(), items.size() * sizeof(tfx::WorkItem), hipMemcpyHostToDevice)) !=
                hipSuccess)
                return hip_fail(e, "work list");
            const int occ = std::max(1, kl_fused_occupancy(V, variant));
            for (int per_cu = 1; per_cu <= occ; ++per_cu) {
                const int grid = std::min((int)items.size(), per_cu * cus);
                const int mode = b->shape.defer_sdf;
                const int defer = !citems.empty() && (mode == TFX_SDF_DEFER_ALWAYS ||
                                                      (mode == TFX_SDF_DEFER_AUTO && (int)items.size() <= per_cu * cus));
                float t = -1.0f;
                for (int r = 0; r < repeats; ++r) {
                    for (const Buf& x : bufs)
                        if (x.V == V && (e = hipMemcpyAsync(x.p, snap + x.at, x.bytes, hipMemcpyDeviceToDevice, s)) !=
                                            hipSuccess)
                            return hip_fail(e, "restore");
                    if ((e = hipEventRecord(e0, s)) != hipSuccess) return hip_fail(e, "event");
                    if ((e = kl_launch_fused(V, variant, b->dev_tables, items_d, (int)items.size(), grid, defer,
                                             b->dev_queue, s, nullptr, nullptr)) != hipSuccess ||
                        (defer && (e = kl_launch_collide_items(b->dev_tables, citems_d, (int)citems.size(), s,
                                                               nullptr, nullptr)) != hipSuccess))
                        return hip_fail(e, "launch");
                    if ((e = hipEventRecord(e1, s)) != hipSuccess || (e = hipEventSynchronize(e1)) != hipSuccess)
                        return hip_fail(e, "event");
                    float ms = 0.0f;
                    if ((e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) return hip_fail(e, "event");
                    if (t < 0.0f || ms < t) t = ms;
                }
                rec.candidates += 1;
                if (variant == def.variant && per_cu == def_per_cu) rec.default_us = 1e3f * t;
                if (best < 0.0f || t < best) {
                    best = t;
                    rec.variant = variant;
                    rec.blocks_per_cu = per_cu;
                    rec.sdf_deferred = defer;
                }
            }
        }
        rec.chosen_us = 1e3f * best;
        b->shape.variant[V] = rec.variant;
        b->shape.per_cu[V] = rec.blocks_per_cu;
        if (*num_out < max_out) out[*num_out] = rec;
        *num_out += 1;
    }
    cleanup();
    // the work lists and captured frames follow the pinned shapes
    for (int k = 0; k < 2; ++k) {
        if (b->exec[k]) (void)hipGraphExecDestroy(b->exec[k]);
        b->exec[k] = nullptr;
    }
    rc = build_work_lists(b);
    return rc;
}

// Runs the staged frame: parameter upload, batched SDF build, and the staged
// substeps.  use_graph: the launches are captured once per (slot, substep
// count, membership, kernel arithmetic) into a hipGraph and replayed.
extern "C" int tfx_batch_run_frame(tfx_batch* b, void* stream, int use_graph) {
    CHECK_ARG(b, "tfx_batch_run_frame: null batch");
    CHECK_ARG(b->slot >= 0, "tfx_batch_run_frame: no frame staged");
    hipStream_t s = (hipStream_t)stream;
    const int k = b->slot;
    const int substeps = b->staged;
    int rc = TFX_OK;
    if (b->members.empty()) {
        b->slot = -1;
        return TFX_OK;
    }
    // Band pointers of each SDF build (the sparse initialise/finalise), in
    // launch order; a frame that fails to launch leaves the history unknown.
    tfx::SdfJob* jobs = reinterpret_cast<tfx::SdfJob*>(b->host[k]);
    const bool sparse = tfx::sdf_multi_sparse(b->max_cells, (int)b->sdf_members.size());
    for (size_t j = 0; j < b->sdf_members.size(); ++j)
        sdf_band_stage(b->members[b->sdf_members[j]].sdf, jobs[j], sparse);
    auto failed = [&](int code) {
        for (int j : b->sdf_members) b->members[j].sdf->band_valid = 0;
        b->slot = -1;
        return code;
    };
    if (use_graph) {
        CHECK_ARG(s != nullptr, "tfx_batch_run_frame: graph replay needs a non-default stream");
        const int arith = g_arith;
        if (!b->exec[k] || b->exec_substeps[k] != substeps || b->exec_tris[k] != b->max_tris ||
            b->exec_arith[k] != arith) {
            if (b->exec[k]) { (void)hipGraphExecDestroy(b->exec[k]); b->exec[k] = nullptr; }
            hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
            if (e != hipSuccess) return failed(fail(TFX_ERR_HIP, "hipStreamBeginCapture: %s", hipGetErrorString(e)));
            rc = launch_frame(b, s, substeps, 0);
            hipGraph_t graph = nullptr;
            e = hipStreamEndCapture(s, &graph);
            if (rc) { if (graph) (void)hipGraphDestroy(graph); return failed(rc); }
            if (e != hipSuccess) return failed(fail(TFX_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e)));
            e = hipGraphInstantiate(&b->exec[k], graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) return failed(fail(TFX_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e)));
            b->exec_substeps[k] = substeps;
            b->exec_tris[k] = b->max_tris;
            b->exec_arith[k] = arith;
        }
        PassTimer tm("BatchFrame", s);
        hipError_t e = hipGraphLaunch(b->exec[k], s);
        if (e != hipSuccess) return failed(fail(TFX_ERR_HIP, "hipGraphLaunch: %s", hipGetErrorString(e)));
    } else {
        PassTimer tm("BatchFrame", s);
        rc = launch_frame(b, s, substeps, 1);
    }
    if (rc) return failed(rc);
    hipError_t e = hipEventRecord(b->ev[k], s);
    b->slot = -1;
    if (e != hipSuccess) return fail(TFX_ERR_HIP, "hipEventRecord: %s", hipGetErrorString(e));
    return TFX_OK;
}
