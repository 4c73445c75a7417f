// kscr_api_batch.hip -- batched grooms of the C ABI (kscr_batch_*): one
// launch per step over the packs of many handles, replayed as a hipGraph.
#include <algorithm>
#include <cstring>

#include "kscr_api_internal.h"

using namespace kscr;
using namespace kscr::api;

// ------------------------------------------------------------------ batch
struct kscr_batch {
    std::vector<kscr_handle*> grooms;
    int device = 0;
    int block = 0;
    int num_packs = 0;
    DevBuf<StepArgs> d_args;
    DevBuf<int4> d_packs;  // groom by groom, strands in memory order
    DevBuf<int> d_next;    // work counter of the persistent mapping
    std::vector<StepArgs> args;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    bool use_graph = true;
    hipGraphExec_t graph = nullptr;
    // recorded after each batch_step; every groom's own stream waits on it
    hipEvent_t done = nullptr;
    int graph_steps = 0;
    float graph_dt = 0.f;
    std::string err;
    int longest = 0;
    // KSCR_BATCH_GRID: one workgroup per pack; KSCR_BATCH_PERSISTENT:
    // persistent_grid workgroups walk the packs
    int mapping = KSCR_BATCH_GRID;
    int persistent_grid = 0;
    int fail(int code, const std::string& m) {
        err = m;
        return code;
    }
    int hip_fail(hipError_t e, const char* what) {
        err = std::string(what) + ": " + hipGetErrorString(e);
        return KSCR_ERR_HIP;
    }
    void drop_graph() {
        if (graph) (void)hipGraphExecDestroy(graph);
        graph = nullptr;
        graph_steps = 0;
    }
};

namespace {
// Packs every groom's strands into workgroups of `block` vertices; a pack is
// (begin, end, groom, 0), in memory order (consecutive workgroups touch
// consecutive memory).
int batch_set_block(kscr_batch* b, int block) {
    std::vector<int4> packs;
    for (size_t g = 0; g < b->grooms.size(); g++) {
        std::vector<int2> p2;
        const int rc = build_packs(b->grooms[g], block, p2);
        if (rc) return rc;
        for (const int2& pk : p2) packs.push_back(make_int4(pk.x, pk.y, (int)g, 0));
    }
    if ((b->persistent_grid = batch_persistent_grid(block)) <= 0)
        return b->fail(KSCR_ERR_HIP, "occupancy of the persistent batched kernel");
    hipError_t e;
    if (!b->d_next.p && (e = b->d_next.alloc(1)) != hipSuccess) return b->hip_fail(e, "hipMalloc batch work counter");
    if (b->d_packs.n < packs.size() && (e = b->d_packs.alloc(packs.size())) != hipSuccess)
        return b->hip_fail(e, "hipMalloc batch packs");
    if ((e = hipMemcpyAsync(b->d_packs.p, packs.data(), packs.size() * sizeof(int4), hipMemcpyHostToDevice, b->stream)) !=
            hipSuccess ||
        (e = hipStreamSynchronize(b->stream)) != hipSuccess)
        return b->hip_fail(e, "upload batch packs");
    b->block = block;
    b->num_packs = (int)packs.size();
    b->drop_graph();
    return KSCR_OK;
}
}  // namespace

extern "C" {

int kscr_batch_create(kscr_handle* const* handles, int num_handles, int block_size, kscr_batch** out) {
    if (!handles || num_handles <= 0 || !out) return KSCR_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    auto* b = new kscr_batch;
    int longest = 0;
    for (int g = 0; g < num_handles; g++) {
        kscr_handle* h = handles[g];
        if (!h || !h->uploaded) {
            delete b;
            return KSCR_ERR_NOT_CONFIGURED;
        }
        if (global_path(h) || h->p.schedule != KSCR_SCHEDULE_JACOBI) {
            delete b;
            return KSCR_ERR_UNSUPPORTED;  // the batched kernel is the contact-free Jacobi step
        }
        if (g > 0 && h->device != handles[0]->device) {
            delete b;
            return KSCR_ERR_INVALID_ARGUMENT;
        }
        longest = std::max(longest, longest_strand(h));
        b->grooms.push_back(h);
    }
    b->device = handles[0]->device;
    int block = block_size;
    if (block <= 0) {
        for (int c : kBlockSizes)
            if (c >= longest && c >= 512) { block = c; break; }
        if (block <= 0)
            for (int c : kBlockSizes)
                if (c >= longest) { block = c; break; }
    }
    bool ok = false;
    for (int c : kBlockSizes) ok |= (c == block);
    if (!ok || block < longest) {
        delete b;
        return KSCR_ERR_STRAND_TOO_LONG;
    }
    b->longest = longest;
    hipError_t e;
    if ((e = hipSetDevice(b->device)) != hipSuccess || (e = b->d_args.alloc(num_handles)) != hipSuccess) {
        delete b;
        return KSCR_ERR_HIP;
    }
    if ((e = hipStreamCreate(&b->stream)) != hipSuccess) {
        delete b;
        return KSCR_ERR_HIP;
    }
    b->own_stream = true;
    if ((e = hipEventCreateWithFlags(&b->done, hipEventDisableTiming)) != hipSuccess) {
        (void)hipStreamDestroy(b->stream);
        delete b;
        return KSCR_ERR_HIP;
    }
    const int rc = batch_set_block(b, block);
    if (rc) {
        (void)hipEventDestroy(b->done);
        (void)hipStreamDestroy(b->stream);
        delete b;
        return rc;
    }
    b->args.resize(num_handles);
    *out = b;
    return KSCR_OK;
}

void kscr_batch_destroy(kscr_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    b->drop_graph();
    if (b->done) (void)hipEventDestroy(b->done);
    if (b->own_stream && b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int kscr_batch_set_graph(kscr_batch* b, int use_graph) {
    if (!b) return KSCR_ERR_INVALID_ARGUMENT;
    b->use_graph = use_graph != 0;
    return KSCR_OK;
}

int kscr_batch_step(kscr_batch* b, float dt, int nsteps, int with_sdf) {
    if (!b || nsteps < 0) return KSCR_ERR_INVALID_ARGUMENT;
    if (nsteps == 0 || !(dt > 0)) return KSCR_OK;
    hipError_t e;
    if ((e = hipSetDevice(b->device)) != hipSuccess) return b->hip_fail(e, "hipSetDevice");
    for (size_t g = 0; g < b->grooms.size(); g++) {
        kscr_handle* h = b->grooms[g];
        if (!h->uploaded) return b->fail(KSCR_ERR_NOT_CONFIGURED, "groom " + std::to_string(g) + " is not uploaded");
        // kscr_set_params may have switched a groom to a path the batched
        // kernel does not run (it is the contact-free Jacobi step)
        if (global_path(h) || h->p.schedule != KSCR_SCHEDULE_JACOBI)
            return b->fail(KSCR_ERR_UNSUPPORTED, "groom " + std::to_string(g) +
                                                     ": contacts, step size limits, glued vertices and the coloured "
                                                     "schedule are not batched");
        if (with_sdf && h->strain_on && h->sdf_on && h->sdf_set)
            return b->fail(KSCR_ERR_UNSUPPORTED, "groom " + std::to_string(g) +
                                                     ": strain limiting after the SDF pass is not batched");
        int rc = validate_meta(h);
        if (rc) return b->fail(rc, "groom " + std::to_string(g) + ": " + h->err);
        // the handles' own streams may still hold work on their state
        if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return b->hip_fail(e, "hipStreamSynchronize");
        b->args[g] = step_args(h, dt, with_sdf != 0);
        // later steps of this call: the clock kernel advances the device copy;
        // the host mirror follows the same float arithmetic
        for (int k = 1; k < nsteps; k++) {
            h->last_h = h->h;
            h->time += dt;
            h->step_counter += b->args[g].num_substeps;
        }
    }
    if ((e = hipMemcpyAsync(b->d_args.p, b->args.data(), b->args.size() * sizeof(StepArgs), hipMemcpyHostToDevice,
                            b->stream)) != hipSuccess)
        return b->hip_fail(e, "hipMemcpyAsync batch arguments");
    const int ng = (int)b->grooms.size();
    auto launch = [&]() -> hipError_t {
        if (b->mapping != KSCR_BATCH_PERSISTENT)
            return launch_step_batch(b->d_args.p, b->d_packs.p, b->num_packs, b->block, b->stream);
        // the work counter starts at 0 (a kernel, not a memset: the launches
        // may be captured into a graph)
        hipError_t r = fill_i32(b->d_next.p, 0, 1, b->stream);
        if (r != hipSuccess) return r;
        return launch_step_batch_persistent(b->d_args.p, b->d_packs.p, b->num_packs, b->block, b->persistent_grid,
                                            b->d_next.p, b->stream);
    };
    // later work queued on a groom's own stream (kscr_get_pos_vel,
    // kscr_advance, ...) is ordered after this batch's kernels
    auto order_handles = [&]() -> int {
        hipError_t r;
        if ((r = hipEventRecord(b->done, b->stream)) != hipSuccess) return b->hip_fail(r, "hipEventRecord");
        for (kscr_handle* h : b->grooms)
            if ((r = hipStreamWaitEvent(h->stream, b->done, 0)) != hipSuccess)
                return b->hip_fail(r, "hipStreamWaitEvent");
        return KSCR_OK;
    };
    if (!b->use_graph) {
        for (int k = 0; k < nsteps; k++) {
            if ((e = launch()) != hipSuccess) return b->hip_fail(e, "batch step launch");
            if ((e = launch_batch_clock(b->d_args.p, ng, dt, b->stream)) != hipSuccess)
                return b->hip_fail(e, "batch clock launch");
        }
        return order_handles();
    }
    if (!b->graph || b->graph_steps != nsteps || b->graph_dt != dt) {
        b->drop_graph();
        hipGraph_t graph;
        if ((e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal)) != hipSuccess)
            return b->hip_fail(e, "hipStreamBeginCapture");
        hipError_t le = hipSuccess;
        for (int k = 0; k < nsteps && le == hipSuccess; k++) {
            le = launch();
            if (le == hipSuccess) le = launch_batch_clock(b->d_args.p, ng, dt, b->stream);
        }
        if ((e = hipStreamEndCapture(b->stream, &graph)) != hipSuccess) return b->hip_fail(e, "hipStreamEndCapture");
        if (le != hipSuccess) {
            (void)hipGraphDestroy(graph);
            return b->hip_fail(le, "batch launch under capture");
        }
        e = hipGraphInstantiate(&b->graph, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            b->graph = nullptr;
            return b->hip_fail(e, "hipGraphInstantiate");
        }
        b->graph_steps = nsteps;
        b->graph_dt = dt;
    }
    if ((e = hipGraphLaunch(b->graph, b->stream)) != hipSuccess) return b->hip_fail(e, "hipGraphLaunch");
    return order_handles();
}

int kscr_batch_autotune(kscr_batch* b, float dt, int steps, int* best_block) {
    if (!b || steps <= 0 || !(dt > 0)) return KSCR_ERR_INVALID_ARGUMENT;
    hipError_t e;
    if ((e = hipSetDevice(b->device)) != hipSuccess) return b->hip_fail(e, "hipSetDevice");
    // save every groom's state and clock
    struct Saved {
        DevBuf<float> x, v, u;
        DevBuf<float4> q;
        float h, last_h, time;
        long counter;
    };
    std::vector<Saved> saved(b->grooms.size());
    const hipMemcpyKind d2d = hipMemcpyDeviceToDevice;
    auto copy_state = [&](kscr_handle* h, Saved& sv, bool save) -> hipError_t {
        const size_t n3 = (size_t)h->n * 3 * 4, nq = (size_t)h->n * 16;
        hipError_t r;
        if (save) {
            if ((r = sv.x.alloc((size_t)h->n * 3)) != hipSuccess || (r = sv.v.alloc((size_t)h->n * 3)) != hipSuccess ||
                (r = sv.u.alloc((size_t)h->n * 3)) != hipSuccess || (r = sv.q.alloc(h->n)) != hipSuccess)
                return r;
            sv.h = h->h, sv.last_h = h->last_h, sv.time = h->time, sv.counter = h->step_counter;
            if ((r = hipMemcpyAsync(sv.x.p, h->d_x.p, n3, d2d, b->stream)) != hipSuccess) return r;
            if ((r = hipMemcpyAsync(sv.v.p, h->d_v.p, n3, d2d, b->stream)) != hipSuccess) return r;
            if ((r = hipMemcpyAsync(sv.u.p, h->d_u.p, n3, d2d, b->stream)) != hipSuccess) return r;
            return hipMemcpyAsync(sv.q.p, h->d_q.p, nq, d2d, b->stream);
        }
        h->h = sv.h, h->last_h = sv.last_h, h->time = sv.time, h->step_counter = sv.counter;
        if ((r = hipMemcpyAsync(h->d_x.p, sv.x.p, n3, d2d, b->stream)) != hipSuccess) return r;
        if ((r = hipMemcpyAsync(h->d_v.p, sv.v.p, n3, d2d, b->stream)) != hipSuccess) return r;
        if ((r = hipMemcpyAsync(h->d_u.p, sv.u.p, n3, d2d, b->stream)) != hipSuccess) return r;
        return hipMemcpyAsync(h->d_q.p, sv.q.p, nq, d2d, b->stream);
    };
    for (size_t g = 0; g < b->grooms.size(); g++) {
        if ((e = hipStreamSynchronize(b->grooms[g]->stream)) != hipSuccess) return b->hip_fail(e, "hipStreamSynchronize");
        if ((e = copy_state(b->grooms[g], saved[g], true)) != hipSuccess) return b->hip_fail(e, "autotune save");
    }
    const bool graph_was = b->use_graph;
    b->use_graph = false;
    hipEvent_t ev0, ev1;
    if ((e = hipEventCreate(&ev0)) != hipSuccess || (e = hipEventCreate(&ev1)) != hipSuccess)
        return b->hip_fail(e, "hipEventCreate");
    const int start_block = b->block, start_mapping = b->mapping;
    int best = -1, best_mapping = start_mapping, rc = KSCR_OK;
    float best_ms = 0;
    for (int c : kBlockSizes) {
        if (c < b->longest) continue;
        if ((rc = batch_set_block(b, c)) != KSCR_OK) break;
        for (int m : {KSCR_BATCH_GRID, KSCR_BATCH_PERSISTENT}) {
            b->mapping = m;
            if ((rc = kscr_batch_step(b, dt, 1, 0)) != KSCR_OK) break;  // warm-up
            (void)hipEventRecord(ev0, b->stream);
            if ((rc = kscr_batch_step(b, dt, steps, 0)) != KSCR_OK) break;
            (void)hipEventRecord(ev1, b->stream);
            if ((e = hipEventSynchronize(ev1)) != hipSuccess) {
                rc = b->hip_fail(e, "hipEventSynchronize");
                break;
            }
            float ms = 0;
            (void)hipEventElapsedTime(&ms, ev0, ev1);
            if (best < 0 || ms < best_ms) best = c, best_mapping = m, best_ms = ms;
        }
        if (rc != KSCR_OK) break;
    }
    b->mapping = (rc == KSCR_OK && best > 0) ? best_mapping : start_mapping;
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    b->use_graph = graph_was;
    for (size_t g = 0; g < b->grooms.size(); g++)
        if ((e = copy_state(b->grooms[g], saved[g], false)) != hipSuccess && rc == KSCR_OK)
            rc = b->hip_fail(e, "autotune restore");
    const int rc2 = batch_set_block(b, (rc == KSCR_OK && best > 0) ? best : start_block);
    if ((e = hipStreamSynchronize(b->stream)) != hipSuccess && rc == KSCR_OK) rc = b->hip_fail(e, "hipStreamSynchronize");
    if (rc) return rc;
    if (rc2) return rc2;
    if (best_block) *best_block = b->block;
    return KSCR_OK;
}

int kscr_batch_synchronize(kscr_batch* b) {
    if (!b) return KSCR_ERR_INVALID_ARGUMENT;
    hipError_t e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return b->hip_fail(e, "hipStreamSynchronize");
    return KSCR_OK;
}

int kscr_batch_block_size(const kscr_batch* b) { return b ? b->block : -1; }

int kscr_batch_set_mapping(kscr_batch* b, int mapping) {
    if (!b || (mapping != KSCR_BATCH_GRID && mapping != KSCR_BATCH_PERSISTENT)) return KSCR_ERR_INVALID_ARGUMENT;
    if (mapping != b->mapping) b->drop_graph();
    b->mapping = mapping;
    return KSCR_OK;
}

int kscr_batch_mapping(const kscr_batch* b) { return b ? b->mapping : -1; }

int kscr_batch_last_error(const kscr_batch* b, char* buf, size_t n) {
    if (!b || !buf || n == 0) return KSCR_ERR_INVALID_ARGUMENT;
    std::snprintf(buf, n, "%s", b->err.c_str());
    return KSCR_OK;
}

}  // extern "C"

