// kscr_api.hip -- host side of the C ABI declared in include/kscr.h.
//
// The handle mirrors YarnBall::Sim (YB/YBSimulator.h:18-99): a host copy of
// the per-vertex state that setup code fills (configure, setYoung, setK*),
// device buffers owned by the handle, and the step driver.  Setup follows the
// reference's host code expression for expression, including its float and
// double mixing:
//   Sim::Sim                 YB/YBSimulator.cu:81-127
//   createFromCurves         YB/io/reader.cpp:16-50
//   Sim::configure           YB/YBSimulator.cu:161-274
//   setKStretch/KBend/Young  YB/YBSimulator.cu:276-320
//   uploadMeta validation    YB/YBSimulator.cu:322-335
//   Sim::advance             YB/sim/step.cpp:38-62
//
// The differentiation entry points are in kscr_api_adjoint.hip, the batch in
// kscr_api_batch.hip and the LBVH in kscr_api_lbvh.hip; the handle and the
// helpers they share in kscr_api_internal.h.
#include <cstdio>
#include <cstring>

#include "kscr_api_internal.h"

using namespace kscr;
using namespace kscr::api;

static_assert(sizeof(kscr_params) == 27 * 4, "kscr_params layout is part of the ABI");

namespace {
const double kPi = 3.141592653589793238462643;

thread_local std::string g_create_error;
}  // namespace

namespace kscr {
namespace api {

// memcpy of large host arrays split over host threads
void host_copy(void* dst, const void* src, size_t bytes) {
    const size_t chunk = (size_t)8 << 20;
    const unsigned workers = std::min<unsigned>(16, std::max(1u, std::thread::hardware_concurrency()));
    if (bytes < 2 * chunk || workers == 1) {
        std::memcpy(dst, src, bytes);
        return;
    }
    const size_t per = (bytes + workers - 1) / workers;
    std::vector<std::thread> pool;
    for (unsigned w = 0; w < workers; w++) {
        const size_t a = w * per, b = std::min(bytes, a + per);
        if (a >= b) break;
        pool.emplace_back([=] { std::memcpy((char*)dst + a, (const char*)src + a, b - a); });
    }
    for (auto& t : pool) t.join();
}

// uploadMeta (YBSimulator.cu:322-335) on a parameter set p for handle h
// (h->min_seg is the shortest rest segment once configured).
int validate_params(kscr_handle* h, const kscr_params& p) {
    if (p.detection_scaler <= 1) return h->fail(KSCR_ERR_DETECTION_SCALER, "Detection scaler must be greater than 1.");
    if (2 * h->min_seg < 2 * p.radius + p.barrier_thickness)
        return h->fail(KSCR_ERR_SEGMENT_TOO_SHORT,
                       "Use thinner yarn or use longer segments. (2 * min_seg_length must be > 2 * radius + "
                       "barrier_thickness)");
    if (p.num_itr < 0) return h->fail(KSCR_ERR_INVALID_ARGUMENT, "num_itr must be >= 0");
    if (!(p.max_h > 0)) return h->fail(KSCR_ERR_INVALID_ARGUMENT, "max_h must be > 0");
    if (p.schedule != KSCR_SCHEDULE_JACOBI && p.schedule != KSCR_SCHEDULE_COLORED)
        return h->fail(KSCR_ERR_INVALID_ARGUMENT, "unknown schedule");
    if (p.schedule == KSCR_SCHEDULE_COLORED && (p.enable_hair_hair_coll || p.use_step_size_limit || h->has_glue))
        return h->fail(KSCR_ERR_UNSUPPORTED, "the coloured schedule covers the contact-free step (no contacts, no step "
                                             "size limit, no glued vertices)");
    return KSCR_OK;
}

int validate_meta(kscr_handle* h) { return validate_params(h, h->p); }

bool colored(const kscr_handle* h) { return h->p.schedule == KSCR_SCHEDULE_COLORED; }

// Jacobi strands longer than the largest workgroup are split over packs that
// overlap by a one-vertex halo at each cut (iterate_kernel), as the
// reference runs chains of any length in overlapping blocks
// (cosserat.cu:30-37).
bool split_strands(const kscr_handle* h) { return !colored(h) && h->longest > kMaxBlock; }

// The step runs as per-substep and per-iteration kernels (kscr_contact.hip)
// with hair-hair contacts, step size limits, glued vertices (a glued vertex
// reads its partner's iterate, which may lie in another workgroup) or split
// strands (a cut's neighbours lie in another workgroup).
bool global_path(const kscr_handle* h) {
    return h->p.enable_hair_hair_coll || h->p.use_step_size_limit || h->has_glue || split_strands(h);
}

// Vertices a workgroup of `block` threads holds: one per lane, or two for the
// coloured step (whose pair alignment can leave one lane slot unused).
int pack_capacity(const kscr_handle* h, int block) { return colored(h) ? 2 * block - 1 : block; }

// Largest instantiated block of the step kernel in use.
int max_block(const kscr_handle* h) { return colored(h) ? 512 : 1024; }

// Workgroups the greedy packing of whole strands gives at `capacity`
// vertices per pack (the count build_packs produces).
int count_packs(const kscr_handle* h, int capacity) {
    const int ns = (int)h->strand_off.size() - 1;
    int packs = 0, s = 0;
    while (s < ns) {
        const int beg = h->strand_off[s];
        const int s0 = s;
        while (s < ns && h->strand_off[s + 1] - beg <= capacity) s++;
        if (s == s0) return 0;  // a strand longer than the pack
        packs++;
    }
    return packs;
}

int choose_block(const kscr_handle* h) {
    const int longest = longest_strand(h);
    if (h->p.block_size > 0) return h->p.block_size <= max_block(h) ? h->p.block_size : -1;
    // split strands: the contact path's pack size; longer strands are cut
    if (split_strands(h)) return 256;
    // Smallest instantiated block that holds the longest strand, but holding
    // at least 512 vertices so that several ~100-vertex strands share a
    // workgroup (fewer idle lanes).  The contact path prefers 256: smaller
    // packs leave more packs free of contacts (those run all iterations in
    // one launch) and shorten the per-iteration launches (100k-strand
    // lognormal groom with contacts: 37.2 ms/step at 256, 41.1 at 512,
    // 42.6 at 1024).
    // A small groom gets smaller packs, down to the smallest that holds its
    // longest strand, until it has kMinPacks workgroups: one 512-vertex
    // workgroup per CU leaves the step latency-bound (2000 strands x 64:
    // forward 0.061 ms/step with 512-vertex packs, 0.054 with 64; the
    // adjoint's backward 0.169 and 0.150 ms/step).
    const int min_cap = global_path(h) ? 255 : 511;
    int chosen = -1;
    for (int b : kBlockSizes)
        if (b <= max_block(h) && pack_capacity(h, b) >= longest && pack_capacity(h, b) >= min_cap) {
            chosen = b;
            break;
        }
    if (chosen > 0 && !colored(h)) {
        constexpr int kMinPacks = 1024;
        while (chosen > kBlockSizes[0] && pack_capacity(h, chosen / 2) >= longest &&
               count_packs(h, pack_capacity(h, chosen)) < kMinPacks)
            chosen /= 2;
        return chosen;
    }
    if (chosen > 0) return chosen;
    for (int b : kBlockSizes)
        if (b <= max_block(h) && pack_capacity(h, b) >= longest) return b;
    return -1;
}

// Greedy packing of whole strands into workgroups of at most `block`
// vertices.  With split strands, a strand longer than `block` gets packs of
// its own: each owns a run of vertices and adds the vertex before it (when
// the strand continues before) and the vertex after it (when it continues
// after) as halos, so the pack [x, y) starts with hasPrev or ends with hasNext
// exactly where it is cut.
int build_packs(kscr_handle* h, int block, std::vector<int2>& packs) {
    packs.clear();
    const int ns = (int)h->strand_off.size() - 1;
    const bool split = split_strands(h);
    int s = 0;
    while (s < ns) {
        const int beg = h->strand_off[s];
        int end = beg;
        while (s < ns && h->strand_off[s + 1] - beg <= block) {
            end = h->strand_off[s + 1];
            s++;
        }
        if (end == beg && split && block >= 3) {
            const int e = h->strand_off[s + 1];
            for (int own = beg; own < e;) {
                const int before = own > beg ? 1 : 0;
                const int own_end = (e - own) + before <= block ? e : own + (block - before - 1);
                packs.push_back(make_int2(own - before, own_end + (own_end < e ? 1 : 0)));
                own = own_end;
            }
            s++;
            continue;
        }
        if (end == beg)
            return h->fail(KSCR_ERR_STRAND_TOO_LONG, "strand " + std::to_string(s) + " has " +
                                                         std::to_string(h->strand_off[s + 1] - h->strand_off[s]) +
                                                         " vertices, more than the workgroup size " + std::to_string(block));
        packs.push_back(make_int2(beg, end));
    }
    if (split) return KSCR_OK;
    // The fused kernels read neighbours at t-1 / t+1 only through hasPrev /
    // hasNext; a pack must therefore start without hasPrev and end without
    // hasNext.
    for (const int2& pk : packs) {
        if (h->flags[pk.x] & FLAG_PREV) return h->fail(KSCR_ERR_INVALID_ARGUMENT, "pack starts inside a strand");
        if (h->flags[pk.y - 1] & FLAG_NEXT) return h->fail(KSCR_ERR_INVALID_ARGUMENT, "pack ends inside a strand");
    }
    return KSCR_OK;
}

int upload_all(kscr_handle* h) {
    const int n = h->n;
    hipError_t e;
    if ((e = hipSetDevice(h->device)) != hipSuccess) return h->hip_fail(e, "hipSetDevice");
    std::vector<float4> mat(n), qv(n), qr(n);
    for (int i = 0; i < n; i++) {
        float fb;
        std::memcpy(&fb, &h->flags[i], 4);
        mat[i] = make_float4(h->inv_mass[i], h->l_rest[i], h->k_stretch[i], fb);
        qv[i] = make_float4(h->q[i].x, h->q[i].y, h->q[i].z, h->q[i].w);
        qr[i] = make_float4(h->q_rest[i].x, h->q_rest[i].y, h->q_rest[i].z, h->q_rest[i].w);
    }
    int block = choose_block(h);
    if (block < 0)
        return h->fail(KSCR_ERR_STRAND_TOO_LONG, "no workgroup size holds the longest strand (at most " +
                                                     std::to_string(pack_capacity(h, max_block(h))) +
                                                     " vertices for this schedule)");
    std::vector<int2> packs;
    int rc = build_packs(h, pack_capacity(h, block), packs);
    if (rc) return rc;
    h->block = block;
    h->num_packs = (int)packs.size();
    if ((e = h->d_packs.alloc(packs.size())) != hipSuccess) return h->hip_fail(e, "hipMalloc packs");
    const size_t n3 = (size_t)n * 3 * sizeof(float);
    if ((e = hipMemcpyAsync(h->d_x.p, h->pos.data(), n3, hipMemcpyHostToDevice, h->stream)) != hipSuccess ||
        (e = hipMemcpyAsync(h->d_v.p, h->vel.data(), n3, hipMemcpyHostToDevice, h->stream)) != hipSuccess ||
        (e = hipMemcpyAsync(h->d_q.p, qv.data(), n * sizeof(float4), hipMemcpyHostToDevice, h->stream)) != hipSuccess ||
        (e = hipMemcpyAsync(h->d_qr.p, qr.data(), n * sizeof(float4), hipMemcpyHostToDevice, h->stream)) != hipSuccess ||
        (e = hipMemcpyAsync(h->d_mat.p, mat.data(), n * sizeof(float4), hipMemcpyHostToDevice, h->stream)) != hipSuccess ||
        (e = hipMemcpyAsync(h->d_packs.p, packs.data(), packs.size() * sizeof(int2), hipMemcpyHostToDevice, h->stream)) !=
            hipSuccess)
        return h->hip_fail(e, "hipMemcpyAsync upload");
    if (h->has_glue &&
        ((e = h->d_conn.alloc(n)) != hipSuccess ||
         (e = hipMemcpyAsync(h->d_conn.p, h->conn.data(), n * sizeof(int), hipMemcpyHostToDevice, h->stream)) !=
             hipSuccess))
        return h->hip_fail(e, "upload connections");
    if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return h->hip_fail(e, "hipStreamSynchronize");
    h->uploaded = true;
    return KSCR_OK;
}

// Per-groom kernel arguments of one step (substep split of Sim::advance,
// YB/sim/step.cpp:41-43); advances the handle's clock.
StepArgs step_args(kscr_handle* h, float dt, bool with_sdf) {
    const int steps = std::max(1, (int)std::ceil(dt / h->p.max_h));
    h->last_h = h->h;
    h->h = dt / steps;
    StepArgs a{};
    a.x = h->d_x.p;
    a.v = h->d_v.p;
    a.u = h->d_u.p;
    a.q = h->d_q.p;
    a.q_rest = h->d_qr.p;
    a.mat = h->d_mat.p;
    a.num_substeps = steps;
    a.num_itr = h->p.num_itr;
    a.h = h->h;
    a.last_h = h->last_h;
    a.time = h->time;
    for (int k = 0; k < 3; k++) a.gravity[k] = h->p.gravity[k];
    a.damping = h->p.damping;
    a.velocity_decay = h->p.velocity_decay;
    a.acceleration_ratio = h->p.acceleration_ratio;
    a.external_force_multiplier = h->p.external_force_multiplier;
    a.wind_peak = h->p.wind_peak;
    a.wind_y_freq = h->p.wind_y_freq;
    a.wind_z_freq = h->p.wind_z_freq;
    a.wind_time_freq = h->p.wind_time_freq;
    a.wind_sharpness = h->p.wind_sharpness;
    a.sdf_on = (with_sdf && h->sdf_on && h->sdf_set) ? 1 : 0;
    a.sdf = h->sdf;
    a.conn = h->has_glue ? h->d_conn.p : nullptr;
    h->step_counter += steps;
    h->time += dt;
    return a;
}

// Step with hair-hair contacts: the reference's per-substep sequence
// (YB/sim/step.cpp:49-57) as separate kernels, see kscr_contact.hip.
int launch_contact_step(kscr_handle* h, float dt, bool with_sdf) {
    hipError_t e;
    if ((e = contact_reserve(h->contact, h->n)) != hipSuccess) return h->hip_fail(e, "contact buffers");
    const long counter0 = h->step_counter;
    const StepArgs a = step_args(h, dt, with_sdf);
    ContactParams cp{};
    cp.radius = h->p.radius;
    cp.barrier = h->p.barrier_thickness;
    cp.k_collision = h->p.k_collision;
    cp.k_friction = h->p.k_friction;
    cp.friction_coeff = h->p.friction_coeff;
    cp.detection_radius = h->p.radius + 0.5f * h->p.barrier_thickness;  // uploadMeta (YBSimulator.cu:323-324)
    cp.scaled_detection_radius = cp.detection_radius * h->p.detection_scaler;
    cp.detection_scaler = h->p.detection_scaler;
    cp.use_step_size_limit = h->p.use_step_size_limit;
    cp.use_velocity_radius = h->p.use_velocity_radius;
    ContactDev& c = h->contact;
    int cur = -1;  // where the previous substep's last iterate is (none yet)
    for (int s = 0; s < a.num_substeps; s++) {
        const bool detect =
            h->p.enable_hair_hair_coll && h->p.detection_period > 0 && (counter0 + s) % h->p.detection_period == 0;
        // the previous substep's endItr, this one's initItr and, when due, the segment boxes
        if ((e = contact_begin(a, cp, c, cur, detect, h->stream)) != hipSuccess) return h->hip_fail(e, "substep start");
        if (detect) {
            const bool rebuild = h->last_bvh_rebuild >= h->p.bvh_rebuild_period;
            if ((e = contact_detect(a, cp, c, rebuild, h->stream)) != hipSuccess)
                return h->hip_fail(e, "collision detection");
            if (rebuild)
                h->last_bvh_rebuild = 0;
            else
                h->last_bvh_rebuild += h->h * h->p.detection_period;
        }
        if ((e = contact_step_limit(a, cp, c, h->stream)) != hipSuccess) return h->hip_fail(e, "step limit");
        // packs without contacts run all iterations in one launch; the
        // per-iteration launches below cover the packs with contacts
        if ((e = contact_classify(a, c, h->d_packs.p, h->num_packs, h->block, h->stream)) != hipSuccess)
            return h->hip_fail(e, "contact classification");
        float4* q_fin = (a.num_itr % 2 == 0) ? h->d_q.p : c.q2;  // where the iteration ping-pong ends
        if ((e = contact_free_iterate(a, c, h->d_q.p, q_fin, h->d_packs.p, h->num_packs, h->block, h->stream)) !=
            hipSuccess)
            return h->hip_fail(e, "contact-free iterations");
        cur = 0;
        for (int k = 0; k < a.num_itr; k++) {
            if ((e = contact_iterate(a, cp, c, cur, h->d_q.p, c.q2, h->d_packs.p, h->num_packs, h->block,
                                     h->stream)) != hipSuccess)
                return h->hip_fail(e, "contact iteration");
            std::swap(h->d_q.p, c.q2);  // the new orientations become the current ones
            cur ^= 1;
        }
    }
    // the last substep's endItr (the others' ran at the start of the next substep)
    if (cur >= 0 && (e = contact_end(a, c, cur, h->stream)) != hipSuccess) return h->hip_fail(e, "contact end");
    StepArgs a2 = a;
    a2.q = h->d_q.p;
    if (a.sdf_on && (e = sdf_pass(a2, h->n, h->stream)) != hipSuccess) return h->hip_fail(e, "SDF pass");
    return strain_pass(h, a);
}

// Strain limiting after the step's SDF pass (SDFBodyCollisionSolver.cpp:
// 182-187: StrainLimiting runs when UseStrainLimiting is set).
int strain_pass(kscr_handle* h, const StepArgs& a) {
    if (!a.sdf_on || !h->strain_on) return KSCR_OK;
    const hipError_t e = launch_strain_limit(h->d_x.p, h->d_strand_off.p, h->d_strain_rest.p,
                                             (int)h->strand_off.size() - 1, h->strain_error, h->stream);
    if (e != hipSuccess) return h->hip_fail(e, "strain limiting");
    return KSCR_OK;
}

// Reads and clears the device error word of the contact path.
int check_device_error(kscr_handle* h) {
    if (!h->contact.error) return KSCR_OK;
    int word = 0;
    hipError_t e;
    if ((e = hipMemcpyAsync(&word, h->contact.error, sizeof(int), hipMemcpyDeviceToHost, h->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(h->stream)) != hipSuccess)
        return h->hip_fail(e, "read device error word");
    if (!word) return KSCR_OK;
    if ((e = fill_i32(h->contact.error, 0, 1, h->stream)) != hipSuccess) return h->hip_fail(e, "reset error word");
    if (word & DEV_ERR_MAX_COLLISIONS)
        return h->fail(KSCR_ERR_MAX_COLLISIONS, "MAX_COLLISIONS_PER_SEGMENT exceeded. Current simulation state may be corrupted!");
    if (word & DEV_ERR_PAIR_OVERFLOW) return h->fail(KSCR_ERR_MAX_COLLISIONS, "broad-phase pair buffer overflow");
    return h->fail(KSCR_ERR_MAX_COLLISIONS, "LBVH query deeper than its stack");
}

int launch_step(kscr_handle* h, float dt, bool with_sdf) {
    if (!h->uploaded) return h->fail(KSCR_ERR_NOT_CONFIGURED, "upload() has not been called");
    if (!(dt > 0)) return KSCR_OK;  // Sim::advance returns for h <= 0
    int rc = validate_meta(h);
    if (rc) return rc;
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return h->hip_fail(e, "hipSetDevice");
    if (global_path(h)) return launch_contact_step(h, dt, with_sdf);
    const StepArgs a = step_args(h, dt, with_sdf);
    e = colored(h) ? launch_step_colored(a, h->d_packs.p, h->num_packs, h->block, h->stream)
                   : launch_step_fused(a, h->d_packs.p, h->num_packs, h->block, h->stream);
    if (e != hipSuccess) return h->hip_fail(e, "step kernel launch");
    return strain_pass(h, a);
}

int longest_strand(const kscr_handle* h) { return h->longest; }

// Re-uploads (invMass, lRest, kStretch, flags) and qRest from the host mirror,
// leaving the dynamic state on the device untouched.
int upload_material(kscr_handle* h) {
    const int n = h->n;
    std::vector<float4> mat(n), qr(n);
    for (int i = 0; i < n; i++) {
        float fb;
        std::memcpy(&fb, &h->flags[i], 4);
        mat[i] = make_float4(h->inv_mass[i], h->l_rest[i], h->k_stretch[i], fb);
        qr[i] = make_float4(h->q_rest[i].x, h->q_rest[i].y, h->q_rest[i].z, h->q_rest[i].w);
    }
    hipError_t e;
    if ((e = hipMemcpyAsync(h->d_mat.p, mat.data(), n * sizeof(float4), hipMemcpyHostToDevice, h->stream)) != hipSuccess ||
        (e = hipMemcpyAsync(h->d_qr.p, qr.data(), n * sizeof(float4), hipMemcpyHostToDevice, h->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(h->stream)) != hipSuccess)
        return h->hip_fail(e, "upload material");
    return KSCR_OK;
}

// Copies the dynamic state (x, v, u, q) into checkpoint slot k (save) or back.

}  // namespace api
}  // namespace kscr

extern "C" {

int kscr_params_default(kscr_params* p) {
    if (!p) return KSCR_ERR_INVALID_ARGUMENT;
    std::memset(p, 0, sizeof(*p));
    p->gravity[0] = 0.f;
    p->gravity[1] = -9.8f;
    p->gravity[2] = 0.f;
    p->num_itr = 8;
    p->max_h = 4e-4f;
    p->velocity_decay = 0.2f;
    p->damping = 1e-6f;
    p->radius = 1e-4f;
    p->barrier_thickness = 8e-4f;
    p->acceleration_ratio = 1.f;
    p->k_collision = 1e-5f;
    p->detection_scaler = 1.2f;
    p->friction_coeff = 0.1f;
    p->k_friction = 5.f;
    p->external_force_multiplier = 1.f;
    p->detection_period = 1;
    p->use_step_size_limit = 0;
    p->use_velocity_radius = 1;
    p->bvh_rebuild_period = 1 / 10.f;
    p->enable_hair_hair_coll = 0;
    p->wind_peak = 0.f;
    p->wind_y_freq = 0.f;
    p->wind_z_freq = 0.f;
    p->wind_time_freq = 0.f;
    p->wind_sharpness = 1.f;
    p->schedule = KSCR_SCHEDULE_JACOBI;
    p->block_size = 0;
    return KSCR_OK;
}

const char* kscr_create_error(void) { return g_create_error.c_str(); }

int kscr_create(const kscr_params* p, int num_verts, const float* positions, int num_strands, const int* strand_offsets,
                int device, kscr_handle** out) {
    g_create_error.clear();
    if (!p || !positions || !strand_offsets || !out || num_strands <= 0) {
        g_create_error = "null argument or no strands";
        return KSCR_ERR_INVALID_ARGUMENT;
    }
    *out = nullptr;
    if (num_verts < 3) {
        g_create_error = "Too little vertices";
        return KSCR_ERR_TOO_FEW_VERTICES;
    }
    if (strand_offsets[0] != 0 || strand_offsets[num_strands] != num_verts) {
        g_create_error = "strand_offsets must start at 0 and end at num_verts";
        return KSCR_ERR_INVALID_ARGUMENT;
    }
    for (int s = 0; s < num_strands; s++)
        if (strand_offsets[s + 1] <= strand_offsets[s]) {
            g_create_error = "strand_offsets must be strictly increasing";
            return KSCR_ERR_INVALID_ARGUMENT;
        }
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) {
        g_create_error = std::string("no HIP device: ") + hipGetErrorString(e);
        return KSCR_ERR_HIP;
    }
    if (device < 0) device = count - 1;
    if (device >= count) {
        g_create_error = "device ordinal out of range";
        return KSCR_ERR_INVALID_ARGUMENT;
    }
    kscr_handle* h = new kscr_handle;
    h->p = *p;
    h->n = num_verts;
    h->device = device;
    h->strand_off.assign(strand_offsets, strand_offsets + num_strands + 1);
    for (int s = 0; s < num_strands; s++) h->longest = std::max(h->longest, strand_offsets[s + 1] - strand_offsets[s]);
    const size_t n3 = (size_t)num_verts * 3;
    if (h->pos.resize(n3) != hipSuccess || h->vel.resize(n3) != hipSuccess) {
        g_create_error = "hipHostMalloc of the host mirror failed";
        delete h;
        return KSCR_ERR_OUT_OF_MEMORY;
    }
    std::memcpy(h->pos.data(), positions, n3 * sizeof(float));
    std::memset(h->vel.data(), 0, n3 * sizeof(float));
    // Sim::Sim initial values (YBSimulator.cu:118-126)
    h->inv_mass.assign(num_verts, 1.f);
    h->l_rest.assign(num_verts, 1.f);
    h->k_stretch.assign(num_verts, 100.f);
    h->flags.assign(num_verts, FLAG_NEXT);
    h->conn.assign(num_verts, -1);
    h->q.assign(num_verts, f4{0, 0, 0, 1});
    h->q_rest.assign(num_verts, f4{0, 0, 0, 1});
    h->flags[num_verts - 1] = 0;
    // createFromCurves: cut off the end of each curve
    for (int s = 0; s < num_strands; s++) h->flags[strand_offsets[s + 1] - 1] = 0;
    h->h = p->max_h;
    h->last_h = p->max_h;
    *out = h;
    return KSCR_OK;
}

void kscr_destroy(kscr_handle* h) {
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    delete h;
}

int kscr_configure(kscr_handle* h, float density) {
    if (!h) return KSCR_ERR_INVALID_ARGUMENT;
    const int n = h->n;
    h->density = density;
    h->max_seg = 0;
    h->min_seg = FLT_MAX;
    f4 lastQ{0, 0, 0, 1};
    for (int i = 0; i < n; i++) {
        if (i < n - 1) {
            const bool hasNext = (h->flags[i] & FLAG_NEXT) != 0;
            h->flags[i + 1] = hasNext ? (h->flags[i + 1] | FLAG_PREV) : (h->flags[i + 1] & ~FLAG_PREV);
            const bool hasNextNext = hasNext && (h->flags[i + 1] & FLAG_NEXT);
            h->flags[i] = hasNextNext ? (h->flags[i] | FLAG_NEXT_ORI) : (h->flags[i] & ~FLAG_NEXT_ORI);
            if (!hasNext) h->flags[i] |= FLAG_FIX_ORI;
        }
        if (!(h->flags[i] & FLAG_PREV) && !(h->flags[i] & FLAG_NEXT))
            return h->fail(KSCR_ERR_DANGLING_SEGMENT, "Dangling segment. Yarns must be atleast 2 segments long");
        h->l_rest[i] = 1.f / n;
        float mass = 0;
        if (h->flags[i] & FLAG_PREV) mass += h->l_rest[i - 1];
        if (h->flags[i] & FLAG_NEXT) {
            const f3 a{h->pos[3 * i], h->pos[3 * i + 1], h->pos[3 * i + 2]};
            const f3 b{h->pos[3 * i + 3], h->pos[3 * i + 4], h->pos[3 * i + 5]};
            const f3 seg = b - a;
            h->l_rest[i] = length(seg);
            if (h->l_rest[i] == 0 || !std::isfinite(h->l_rest[i]))
                return h->fail(KSCR_ERR_ZERO_LENGTH_SEGMENT, "0 length segment at vertex " + std::to_string(i));
            lastQ = from_to(f3{1, 0, 0}, normalize(seg));
            mass += h->l_rest[i];
            h->max_seg = std::max(h->max_seg, h->l_rest[i]);
            h->min_seg = std::min(h->min_seg, h->l_rest[i]);
        }
        h->q[i] = lastQ;
        // mass = (l_{i-1} + l_i) * r^2 * pi * density / 2, with the reference's
        // float * double promotion
        mass = (float)((double)mass * ((double)(h->p.radius * h->p.radius) * kPi * (double)density * (double)0.5f));
        if (mass != 0)
            h->inv_mass[i] *= 1 / mass;
        else
            h->inv_mass[i] = 0;
    }
    for (int i = 0; i < n - 1; i++) h->q_rest[i] = normalize(rmul(conj(h->q[i]), h->q[i + 1]));

    hipError_t e;
    if ((e = hipSetDevice(h->device)) != hipSuccess) return h->hip_fail(e, "hipSetDevice");
    if (!h->stream) {
        // the null stream unless the caller installs one with kscr_set_stream
    }
    if ((e = h->d_x.alloc((size_t)n * 3)) != hipSuccess || (e = h->d_v.alloc((size_t)n * 3)) != hipSuccess ||
        (e = h->d_u.alloc((size_t)n * 3)) != hipSuccess || (e = h->d_q.alloc(n)) != hipSuccess ||
        (e = h->d_qr.alloc(n)) != hipSuccess || (e = h->d_mat.alloc(n)) != hipSuccess)
        return h->hip_fail(e, "hipMalloc state");
    // lastVels start at zero (YBSimulator.cu:253-254)
    if ((e = hipMemsetAsync(h->d_u.p, 0, (size_t)n * 3 * sizeof(float), h->stream)) != hipSuccess)
        return h->hip_fail(e, "hipMemsetAsync");
    if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return h->hip_fail(e, "hipStreamSynchronize");
    h->configured = true;
    h->uploaded = false;
    return KSCR_OK;
}

int kscr_set_young(kscr_handle* h, float young, float deviation) {
    if (!h) return KSCR_ERR_INVALID_ARGUMENT;
    if (!h->configured) return h->fail(KSCR_ERR_NOT_CONFIGURED, "No rest length. Must call configure()");
    const float mmrad = h->p.radius * 1000;
    for (int i = 0; i < h->n; i++) {
        const float len = h->l_rest[i];
        const float ss = (float)((double)young * kPi * (double)mmrad * (double)mmrad);
        const float bt = (float)((double)(ss * deviation * mmrad * mmrad) * 0.25 / (double)(len * len));
        h->k_stretch[i] = ss;
        h->q_rest[i] = bt * h->q_rest[i];
    }
    h->uploaded = false;
    return KSCR_OK;
}

int kscr_set_kstretch(kscr_handle* h, float k) {
    if (!h) return KSCR_ERR_INVALID_ARGUMENT;
    if (!h->configured) return h->fail(KSCR_ERR_NOT_CONFIGURED, "No rest length. Must call configure()");
    for (int i = 0; i < h->n; i++) h->k_stretch[i] = k * h->l_rest[i];
    h->uploaded = false;
    return KSCR_OK;
}

int kscr_set_kbend(kscr_handle* h, float k) {
    if (!h) return KSCR_ERR_INVALID_ARGUMENT;
    if (!h->configured) return h->fail(KSCR_ERR_NOT_CONFIGURED, "No rest length. Must call configure()");
    k *= 4;
    for (int i = 0; i < h->n; i++) h->q_rest[i] = (k / h->l_rest[i]) * normalize(h->q_rest[i]);
    h->uploaded = false;
    return KSCR_OK;
}

int kscr_sag_free(kscr_handle* h, const int* roots, int num_roots, int iterations) {
    if (!h || iterations < 0 || (roots && num_roots <= 0)) return KSCR_ERR_INVALID_ARGUMENT;
    if (!h->configured) return h->fail(KSCR_ERR_NOT_CONFIGURED, "No rest length. Must call configure()");
    std::vector<int> starts;
    if (roots) {
        starts.assign(roots, roots + num_roots);
        for (int s = 0; s < num_roots; s++)
            if (starts[s] < 0 || starts[s] >= h->n || (s > 0 && starts[s] <= starts[s - 1]))
                return h->fail(KSCR_ERR_INVALID_ARGUMENT, "roots must be ascending vertex ids");
    } else {
        starts.assign(h->strand_off.begin(), h->strand_off.end() - 1);
    }
    if (iterations == 0) return KSCR_OK;
    static_assert(sizeof(f4) == 4 * sizeof(float), "rotors are 4 packed floats");
    sag_free_groom(h->pos.data(), h->inv_mass.data(), h->l_rest.data(), h->k_stretch.data(),
                   reinterpret_cast<float*>(h->q.data()), reinterpret_cast<float*>(h->q_rest.data()), h->p.gravity,
                   starts.data(), (int)starts.size(), h->n, iterations);
    h->uploaded = false;
    return KSCR_OK;
}

int kscr_pin_vertices(kscr_handle* h, const int* ids, int n) {
    if (!h || (n > 0 && !ids)) return KSCR_ERR_INVALID_ARGUMENT;
    for (int k = 0; k < n; k++) {
        if (ids[k] < 0 || ids[k] >= h->n) return h->fail(KSCR_ERR_INVALID_ARGUMENT, "vertex id out of range");
        h->inv_mass[ids[k]] = 0;
    }
    h->uploaded = false;
    return KSCR_OK;
}

int kscr_fix_orientation(kscr_handle* h, const int* ids, int n) {
    if (!h || (n > 0 && !ids)) return KSCR_ERR_INVALID_ARGUMENT;
    for (int k = 0; k < n; k++) {
        if (ids[k] < 0 || ids[k] >= h->n) return h->fail(KSCR_ERR_INVALID_ARGUMENT, "vertex id out of range");
        h->flags[ids[k]] |= FLAG_FIX_ORI;
    }
    h->uploaded = false;
    return KSCR_OK;
}

int kscr_set_connections(kscr_handle* h, const int* conn) {
    if (!h) return KSCR_ERR_INVALID_ARGUMENT;
    if (conn) {
        for (int i = 0; i < h->n; i++)
            if (conn[i] < -1 || conn[i] >= h->n || conn[i] == i)
                return h->fail(KSCR_ERR_INVALID_ARGUMENT,
                               "connection of vertex " + std::to_string(i) + " is neither -1 nor another vertex");
        for (int i = 0; i < h->n; i++)
            if (conn[i] >= 0 && conn[conn[i]] != i)
                return h->fail(KSCR_ERR_INVALID_ARGUMENT, "connections must be symmetric (vertex " +
                                                              std::to_string(i) + " is glued to " +
                                                              std::to_string(conn[i]) + ", which is not glued back)");
    }
    bool any = false;
    for (int i = 0; i < h->n; i++) any |= conn && conn[i] >= 0;
    if (any && h->p.schedule == KSCR_SCHEDULE_COLORED)
        return h->fail(KSCR_ERR_UNSUPPORTED, "the coloured schedule covers strands without glued vertices");
    for (int i = 0; i < h->n; i++) h->conn[i] = conn ? conn[i] : -1;
    h->has_glue = any;
    h->uploaded = false;
    return KSCR_OK;
}

int kscr_set_params(kscr_handle* h, const kscr_params* p) {
    if (!h || !p) return KSCR_ERR_INVALID_ARGUMENT;
    // a rejected parameter set leaves the handle's parameters as they were
    if (h->configured) {
        int rc = validate_params(h, *p);
        if (rc) return rc;
    }
    const int old_block = h->p.block_size, old_schedule = h->p.schedule;
    h->p = *p;
    // the packing depends on both: upload() rebuilds it (the automatic block
    // size's preference for the contact path applies from the next upload;
    // both paths run with any block)
    if (p->block_size != old_block || p->schedule != old_schedule) h->uploaded = false;
    return KSCR_OK;
}

int kscr_get_params(const kscr_handle* h, kscr_params* p) {
    if (!h || !p) return KSCR_ERR_INVALID_ARGUMENT;
    *p = h->p;
    return KSCR_OK;
}

int kscr_upload(kscr_handle* h) {
    if (!h) return KSCR_ERR_INVALID_ARGUMENT;
    if (!h->configured) return h->fail(KSCR_ERR_NOT_CONFIGURED, "Must call configure() before upload()");
    int rc = validate_meta(h);
    if (rc) return rc;
    return upload_all(h);
}

int kscr_set_pos_vel(kscr_handle* h, const float* positions, const float* velocities) {
    if (!h) return KSCR_ERR_INVALID_ARGUMENT;
    if (!h->uploaded) return h->fail(KSCR_ERR_NOT_CONFIGURED, "upload() has not been called");
    const size_t n3 = (size_t)h->n * 3;
    hipError_t e;
    if ((e = hipSetDevice(h->device)) != hipSuccess) return h->hip_fail(e, "hipSetDevice");
    if (positions) {
        host_copy(h->pos.data(), positions, n3 * sizeof(float));
        if ((e = hipMemcpyAsync(h->d_x.p, h->pos.data(), n3 * sizeof(float), hipMemcpyHostToDevice, h->stream)) !=
            hipSuccess)
            return h->hip_fail(e, "hipMemcpyAsync");
    }
    if (velocities) {
        host_copy(h->vel.data(), velocities, n3 * sizeof(float));
        if ((e = hipMemcpyAsync(h->d_v.p, h->vel.data(), n3 * sizeof(float), hipMemcpyHostToDevice, h->stream)) !=
            hipSuccess)
            return h->hip_fail(e, "hipMemcpyAsync");
    }
    if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return h->hip_fail(e, "hipStreamSynchronize");
    return KSCR_OK;
}

int kscr_get_pos_vel(kscr_handle* h, float* positions, float* velocities) {
    if (!h) return KSCR_ERR_INVALID_ARGUMENT;
    if (!h->uploaded) return h->fail(KSCR_ERR_NOT_CONFIGURED, "upload() has not been called");
    const size_t n3 = (size_t)h->n * 3;
    hipError_t e;
    if ((e = hipSetDevice(h->device)) != hipSuccess) return h->hip_fail(e, "hipSetDevice");
    if ((e = hipMemcpyAsync(h->pos.data(), h->d_x.p, n3 * sizeof(float), hipMemcpyDeviceToHost, h->stream)) !=
            hipSuccess ||
        (e = hipMemcpyAsync(h->vel.data(), h->d_v.p, n3 * sizeof(float), hipMemcpyDeviceToHost, h->stream)) !=
            hipSuccess)
        return h->hip_fail(e, "hipMemcpyAsync");
    if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return h->hip_fail(e, "hipStreamSynchronize");
    if (positions) host_copy(positions, h->pos.data(), n3 * sizeof(float));
    if (velocities) host_copy(velocities, h->vel.data(), n3 * sizeof(float));
    return check_device_error(h);
}

int kscr_download_pos_vel(kscr_handle* h, const float** positions, const float** velocities) {
    if (!h) return KSCR_ERR_INVALID_ARGUMENT;
    if (!h->uploaded) return h->fail(KSCR_ERR_NOT_CONFIGURED, "upload() has not been called");
    const size_t n3 = (size_t)h->n * 3;
    hipError_t e;
    if ((e = hipSetDevice(h->device)) != hipSuccess) return h->hip_fail(e, "hipSetDevice");
    if ((e = hipMemcpyAsync(h->pos.data(), h->d_x.p, n3 * sizeof(float), hipMemcpyDeviceToHost, h->stream)) !=
            hipSuccess ||
        (e = hipMemcpyAsync(h->vel.data(), h->d_v.p, n3 * sizeof(float), hipMemcpyDeviceToHost, h->stream)) !=
            hipSuccess)
        return h->hip_fail(e, "hipMemcpyAsync");
    if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return h->hip_fail(e, "hipStreamSynchronize");
    if (positions) *positions = h->pos.data();
    if (velocities) *velocities = h->vel.data();
    return check_device_error(h);
}

int kscr_get_last_vel(kscr_handle* h, float* last_vel) {
    if (!h || !last_vel) return KSCR_ERR_INVALID_ARGUMENT;
    if (!h->uploaded) return h->fail(KSCR_ERR_NOT_CONFIGURED, "upload() has not been called");
    hipError_t e;
    if ((e = hipSetDevice(h->device)) != hipSuccess) return h->hip_fail(e, "hipSetDevice");
    if ((e = hipMemcpyAsync(last_vel, h->d_u.p, (size_t)h->n * 3 * sizeof(float), hipMemcpyDeviceToHost, h->stream)) !=
        hipSuccess)
        return h->hip_fail(e, "hipMemcpyAsync");
    if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return h->hip_fail(e, "hipStreamSynchronize");
    return KSCR_OK;
}

int kscr_get_q(kscr_handle* h, float* q) {
    if (!h || !q) return KSCR_ERR_INVALID_ARGUMENT;
    if (!h->uploaded) return h->fail(KSCR_ERR_NOT_CONFIGURED, "upload() has not been called");
    hipError_t e;
    if ((e = hipSetDevice(h->device)) != hipSuccess) return h->hip_fail(e, "hipSetDevice");
    if ((e = hipMemcpyAsync(q, h->d_q.p, (size_t)h->n * sizeof(float4), hipMemcpyDeviceToHost, h->stream)) != hipSuccess)
        return h->hip_fail(e, "hipMemcpyAsync");
    if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return h->hip_fail(e, "hipStreamSynchronize");
    for (int i = 0; i < h->n; i++) h->q[i] = f4{q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3]};
    return KSCR_OK;
}

int kscr_set_q(kscr_handle* h, const float* q) {
    if (!h || !q) return KSCR_ERR_INVALID_ARGUMENT;
    if (!h->uploaded) return h->fail(KSCR_ERR_NOT_CONFIGURED, "upload() has not been called");
    for (int i = 0; i < h->n; i++) h->q[i] = f4{q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3]};
    hipError_t e;
    if ((e = hipSetDevice(h->device)) != hipSuccess) return h->hip_fail(e, "hipSetDevice");
    if ((e = hipMemcpyAsync(h->d_q.p, h->q.data(), (size_t)h->n * sizeof(float4), hipMemcpyHostToDevice, h->stream)) !=
        hipSuccess)
        return h->hip_fail(e, "hipMemcpyAsync");
    if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return h->hip_fail(e, "hipStreamSynchronize");
    return KSCR_OK;
}

int kscr_get_rest(const kscr_handle* h, float* inv_mass, float* l_rest, float* k_stretch, uint32_t* flags,
                  float* q_rest) {
    if (!h) return KSCR_ERR_INVALID_ARGUMENT;
    for (int i = 0; i < h->n; i++) {
        if (inv_mass) inv_mass[i] = h->inv_mass[i];
        if (l_rest) l_rest[i] = h->l_rest[i];
        if (k_stretch) k_stretch[i] = h->k_stretch[i];
        if (flags) flags[i] = h->flags[i];
        if (q_rest) {
            q_rest[4 * i] = h->q_rest[i].x;
            q_rest[4 * i + 1] = h->q_rest[i].y;
            q_rest[4 * i + 2] = h->q_rest[i].z;
            q_rest[4 * i + 3] = h->q_rest[i].w;
        }
    }
    return KSCR_OK;
}

int kscr_get_contacts(kscr_handle* h, int* num, int* cols) {
    if (!h || !num || !cols) return KSCR_ERR_INVALID_ARGUMENT;
    const int n = h->n;
    if (!h->contact.num_cols) {
        for (int i = 0; i < n; i++) {
            num[i] = 0;
            for (int k = 0; k < kMaxCols; k++) cols[kMaxCols * i + k] = -1;
        }
        return KSCR_OK;
    }
    hipError_t e;
    if ((e = hipSetDevice(h->device)) != hipSuccess) return h->hip_fail(e, "hipSetDevice");
    if ((e = hipMemcpyAsync(num, h->contact.num_cols, n * sizeof(int), hipMemcpyDeviceToHost, h->stream)) != hipSuccess ||
        (e = hipMemcpyAsync(cols, h->contact.cols, (size_t)kMaxCols * n * sizeof(int), hipMemcpyDeviceToHost,
                            h->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(h->stream)) != hipSuccess)
        return h->hip_fail(e, "hipMemcpyAsync contacts");
    for (int i = 0; i < n; i++)
        for (int k = num[i]; k < kMaxCols; k++) cols[kMaxCols * i + k] = -1;
    return KSCR_OK;
}

// SdfArgs of a grid over [grid_min, grid_max] (SDFBodyCollisionSolver::InitParams, :148-159)
static SdfArgs sdf_args(const int dim[3], const float grid_min[3], const float grid_max[3], float threshold,
                        const float rotation[4], const float translation[3]) {
    SdfArgs s{};
    for (int a = 0; a < 3; a++) {
        s.dim[a] = dim[a];
        s.gmin[a] = grid_min[a];
        const float dist = grid_max[a] - grid_min[a];
        s.ds[a] = dist / (1.f * dim[a]);
        s.inv_ds[a] = 1.f / s.ds[a];
    }
    s.thresh = threshold;
    for (int a = 0; a < 4; a++) s.rot[a] = rotation ? rotation[a] : (a == 3 ? 1.f : 0.f);
    for (int a = 0; a < 3; a++) s.trans[a] = translation ? translation[a] : 0.f;
    return s;
}

int kscr_set_sdf(kscr_handle* h, const int dim[3], const float grid_min[3], const float grid_max[3], const float* phi,
                 const float* grad, float threshold, const float rotation[4], const float translation[3]) {
    if (!h || !dim || !grid_min || !grid_max || !phi || !grad) return KSCR_ERR_INVALID_ARGUMENT;
    size_t cells = 1;
    for (int a = 0; a < 3; a++) {
        if (dim[a] <= 0) return h->fail(KSCR_ERR_INVALID_ARGUMENT, "SDF dimensions must be positive");
        cells *= (size_t)dim[a];
    }
    SdfArgs s = sdf_args(dim, grid_min, grid_max, threshold, rotation, translation);
    std::vector<float4> g(cells);
    for (size_t c = 0; c < cells; c++) g[c] = make_float4(grad[3 * c], grad[3 * c + 1], grad[3 * c + 2], phi[c]);
    hipError_t e;
    if ((e = hipSetDevice(h->device)) != hipSuccess) return h->hip_fail(e, "hipSetDevice");
    if ((e = h->d_grid.alloc(cells)) != hipSuccess) return h->hip_fail(e, "hipMalloc SDF grid");
    if ((e = hipMemcpyAsync(h->d_grid.p, g.data(), cells * sizeof(float4), hipMemcpyHostToDevice, h->stream)) !=
        hipSuccess)
        return h->hip_fail(e, "hipMemcpyAsync SDF grid");
    if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return h->hip_fail(e, "hipStreamSynchronize");
    s.grid = h->d_grid.p;
    h->sdf = s;
    h->sdf_set = true;
    h->sdf_on = true;
    return KSCR_OK;
}

int kscr_mesh_sdf(int device, const float* vertices, int num_vertices, const int* triangles, int num_triangles,
                  const int dim[3], float* phi, float* grad, float grid_min[3], float grid_max[3], char* err,
                  size_t err_len) {
    std::string msg;
    auto finish = [&](int rc) {
        if (rc != KSCR_OK && err && err_len) {
            std::strncpy(err, msg.c_str(), err_len - 1);
            err[err_len - 1] = 0;
        }
        return rc;
    };
    if (!dim || dim[0] < 1 || dim[1] < 1 || dim[2] < 1) {
        msg = "mesh SDF: grid dimensions must be positive";
        return finish(KSCR_ERR_INVALID_ARGUMENT);
    }
    hipError_t e;
    if (device >= 0 && (e = hipSetDevice(device)) != hipSuccess) {
        msg = std::string("hipSetDevice: ") + hipGetErrorString(e);
        return finish(KSCR_ERR_HIP);
    }
    const size_t cells = (size_t)dim[0] * dim[1] * dim[2];
    DevBuf<float4> out;
    if ((size_t)dim[0] * dim[1] * dim[2] > (size_t)INT32_MAX) {
        msg = "mesh SDF: more than 2^31 - 1 grid nodes";
        return finish(KSCR_ERR_INVALID_ARGUMENT);
    }
    if ((e = out.alloc(cells)) != hipSuccess) {
        msg = std::string("hipMalloc: ") + hipGetErrorString(e);
        return finish(KSCR_ERR_OUT_OF_MEMORY);
    }
    float lo[3], hi[3];
    int rc = kscr::mesh_sdf_build(vertices, num_vertices, triangles, num_triangles, dim, out.p, lo, hi, nullptr, msg);
    if (rc != KSCR_OK) return finish(rc);
    std::vector<float4> g(cells);
    if ((e = hipMemcpy(g.data(), out.p, cells * sizeof(float4), hipMemcpyDeviceToHost)) != hipSuccess) {
        msg = std::string("hipMemcpy: ") + hipGetErrorString(e);
        return finish(KSCR_ERR_HIP);
    }
    for (size_t c = 0; c < cells; c++) {
        if (phi) phi[c] = g[c].w;
        if (grad) grad[3 * c] = g[c].x, grad[3 * c + 1] = g[c].y, grad[3 * c + 2] = g[c].z;
    }
    for (int a = 0; a < 3; a++) {
        if (grid_min) grid_min[a] = lo[a];
        if (grid_max) grid_max[a] = hi[a];
    }
    return finish(KSCR_OK);
}

int kscr_set_sdf_from_mesh(kscr_handle* h, const float* vertices, int num_vertices, const int* triangles,
                           int num_triangles, const int dim[3], float threshold, const float rotation[4],
                           const float translation[3]) {
    if (!h || !dim || dim[0] < 1 || dim[1] < 1 || dim[2] < 1) return KSCR_ERR_INVALID_ARGUMENT;
    const size_t cells = (size_t)dim[0] * dim[1] * dim[2];
    if (cells > (size_t)INT32_MAX) return h->fail(KSCR_ERR_INVALID_ARGUMENT, "more than 2^31 - 1 SDF grid nodes");
    hipError_t e;
    if ((e = hipSetDevice(h->device)) != hipSuccess) return h->hip_fail(e, "hipSetDevice");
    if ((e = h->d_grid.alloc(cells)) != hipSuccess) return h->hip_fail(e, "hipMalloc SDF grid");
    float lo[3], hi[3];
    std::string msg;
    const int rc = kscr::mesh_sdf_build(vertices, num_vertices, triangles, num_triangles, dim, h->d_grid.p, lo, hi,
                                        h->stream, msg);
    if (rc != KSCR_OK) {
        h->sdf_set = false;
        h->sdf_on = false;
        return h->fail(rc, msg);
    }
    SdfArgs s = sdf_args(dim, lo, hi, threshold, rotation, translation);
    s.grid = h->d_grid.p;
    h->sdf = s;
    h->sdf_set = true;
    h->sdf_on = true;
    return KSCR_OK;
}

int kscr_enable_sdf(kscr_handle* h, int enable) {
    if (!h) return KSCR_ERR_INVALID_ARGUMENT;
    h->sdf_on = enable != 0;
    return KSCR_OK;
}

int kscr_set_strain_limit(kscr_handle* h, int enable, float strain_error, const float* rest_lengths) {
    if (!h) return KSCR_ERR_INVALID_ARGUMENT;
    if (!enable) {
        h->strain_on = false;
        return KSCR_OK;
    }
    if (!(strain_error >= 0) || !std::isfinite(strain_error))
        return h->fail(KSCR_ERR_INVALID_ARGUMENT, "strain_error must be finite and >= 0");
    const int n = h->n;
    const int ns = (int)h->strand_off.size() - 1;
    // rest length of segment (i, i + 1); the last vertex of a strand has none
    std::vector<float> rest(n, 0.0f);
    for (int s = 0; s < ns; s++)
        for (int i = h->strand_off[s]; i + 1 < h->strand_off[s + 1]; i++) {
            if (rest_lengths) {
                rest[i] = rest_lengths[i];
            } else {
                // InitAdditionalHairData (SDFBodyCollisionSolver.cpp:108-120):
                // glm length of the segment on the host, no FMA
                const float dx = h->pos[3 * (i + 1)] - h->pos[3 * i];
                const float dy = h->pos[3 * (i + 1) + 1] - h->pos[3 * i + 1];
                const float dz = h->pos[3 * (i + 1) + 2] - h->pos[3 * i + 2];
                rest[i] = std::sqrt((dx * dx + dy * dy) + dz * dz);
            }
            if (!(rest[i] >= 0) || !std::isfinite(rest[i]))
                return h->fail(KSCR_ERR_INVALID_ARGUMENT, "rest lengths must be finite and >= 0");
        }
    hipError_t e;
    if ((e = hipSetDevice(h->device)) != hipSuccess) return h->hip_fail(e, "hipSetDevice");
    if ((e = h->d_strand_off.alloc(ns + 1)) != hipSuccess || (e = h->d_strain_rest.alloc(n)) != hipSuccess)
        return h->hip_fail(e, "hipMalloc strain limiting");
    if ((e = hipMemcpyAsync(h->d_strand_off.p, h->strand_off.data(), (ns + 1) * sizeof(int), hipMemcpyHostToDevice,
                            h->stream)) != hipSuccess ||
        (e = hipMemcpyAsync(h->d_strain_rest.p, rest.data(), n * sizeof(float), hipMemcpyHostToDevice, h->stream)) !=
            hipSuccess ||
        (e = hipStreamSynchronize(h->stream)) != hipSuccess)
        return h->hip_fail(e, "upload strain limiting");
    h->strain_error = strain_error;
    h->strain_on = true;
    return KSCR_OK;
}

int kscr_advance(kscr_handle* h, float dt) {
    if (!h) return KSCR_ERR_INVALID_ARGUMENT;
    return launch_step(h, dt, false);
}

int kscr_solve_next_step(kscr_handle* h, float dt) {
    if (!h) return KSCR_ERR_INVALID_ARGUMENT;
    return launch_step(h, dt, true);
}

int kscr_set_stream(kscr_handle* h, void* stream) {
    if (!h) return KSCR_ERR_INVALID_ARGUMENT;
    h->stream = (hipStream_t)stream;
    return KSCR_OK;
}

int kscr_synchronize(kscr_handle* h) {
    if (!h) return KSCR_ERR_INVALID_ARGUMENT;
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return h->hip_fail(e, "hipStreamSynchronize");
    return check_device_error(h);
}

int kscr_num_verts(const kscr_handle* h) { return h ? h->n : -1; }
int kscr_num_strands(const kscr_handle* h) { return h ? (int)h->strand_off.size() - 1 : -1; }
int kscr_block_size(const kscr_handle* h) { return h ? h->block : -1; }

int kscr_last_error(const kscr_handle* h, char* buf, size_t n) {
    if (!h || !buf || n == 0) return KSCR_ERR_INVALID_ARGUMENT;
    std::snprintf(buf, n, "%s", h->err.c_str());
    return KSCR_OK;
}

// ------------------------------------------------------------------ autotuner
int kscr_autotune(kscr_handle* h, float dt, int steps, int* best_block) {
    if (!h || steps <= 0 || !(dt > 0)) return KSCR_ERR_INVALID_ARGUMENT;
    if (!h->uploaded) return h->fail(KSCR_ERR_NOT_CONFIGURED, "upload() has not been called");
    hipError_t e;
    if ((e = hipSetDevice(h->device)) != hipSuccess) return h->hip_fail(e, "hipSetDevice");
    const size_t n3 = (size_t)h->n * 3;
    DevBuf<float> sx, sv, su;
    DevBuf<float4> sq;
    if ((e = sx.alloc(n3)) != hipSuccess || (e = sv.alloc(n3)) != hipSuccess || (e = su.alloc(n3)) != hipSuccess ||
        (e = sq.alloc(h->n)) != hipSuccess)
        return h->hip_fail(e, "hipMalloc autotune scratch");
    auto save = [&](bool restore) -> hipError_t {
        hipError_t r;
        const hipMemcpyKind k = hipMemcpyDeviceToDevice;
        if (!restore) {
            if ((r = hipMemcpyAsync(sx.p, h->d_x.p, n3 * 4, k, h->stream)) != hipSuccess) return r;
            if ((r = hipMemcpyAsync(sv.p, h->d_v.p, n3 * 4, k, h->stream)) != hipSuccess) return r;
            if ((r = hipMemcpyAsync(su.p, h->d_u.p, n3 * 4, k, h->stream)) != hipSuccess) return r;
            return hipMemcpyAsync(sq.p, h->d_q.p, (size_t)h->n * 16, k, h->stream);
        }
        if ((r = hipMemcpyAsync(h->d_x.p, sx.p, n3 * 4, k, h->stream)) != hipSuccess) return r;
        if ((r = hipMemcpyAsync(h->d_v.p, sv.p, n3 * 4, k, h->stream)) != hipSuccess) return r;
        if ((r = hipMemcpyAsync(h->d_u.p, su.p, n3 * 4, k, h->stream)) != hipSuccess) return r;
        return hipMemcpyAsync(h->d_q.p, sq.p, (size_t)h->n * 16, k, h->stream);
    };
    if ((e = save(false)) != hipSuccess) return h->hip_fail(e, "autotune save");
    // With contacts, the lists of the last detection are used until the next
    // one (detection_period > 1): they are restored with the state, and so is
    // the rebuild clock.  The tree left by the timing runs may differ from
    // the one before them; a refit of it yields the same overlapping pairs,
    // and the lists are sorted, so the steps that follow are unchanged.
    DevBuf<int> s_num, s_cols;
    const bool with_lists = h->contact.num_cols != nullptr && h->contact.n == h->n;
    if (with_lists) {
        if ((e = s_num.alloc(h->n)) != hipSuccess || (e = s_cols.alloc((size_t)h->n * kMaxCols)) != hipSuccess)
            return h->hip_fail(e, "hipMalloc autotune contact lists");
        if ((e = hipMemcpyAsync(s_num.p, h->contact.num_cols, (size_t)h->n * 4, hipMemcpyDeviceToDevice,
                                h->stream)) != hipSuccess ||
            (e = hipMemcpyAsync(s_cols.p, h->contact.cols, (size_t)h->n * kMaxCols * 4, hipMemcpyDeviceToDevice,
                                h->stream)) != hipSuccess)
            return h->hip_fail(e, "autotune save contact lists");
    }
    const float h0 = h->h, lh0 = h->last_h, t0 = h->time, rebuild0 = h->last_bvh_rebuild;
    const long c0 = h->step_counter;
    const int longest = longest_strand(h);
    const int saved_block = h->p.block_size;
    hipEvent_t ev0, ev1;
    if ((e = hipEventCreate(&ev0)) != hipSuccess || (e = hipEventCreate(&ev1)) != hipSuccess)
        return h->hip_fail(e, "hipEventCreate");
    int best = -1;
    float best_ms = 0;
    int rc = KSCR_OK;
    for (int b : kBlockSizes) {
        if (b > max_block(h) || (pack_capacity(h, b) < longest && !split_strands(h))) continue;
        h->p.block_size = b;
        if ((rc = upload_all(h)) != KSCR_OK) break;
        if ((e = save(true)) != hipSuccess) { rc = h->hip_fail(e, "autotune restore"); break; }
        h->h = h0, h->last_h = lh0, h->time = t0, h->step_counter = c0, h->last_bvh_rebuild = rebuild0;
        if ((rc = launch_step(h, dt, false)) != KSCR_OK) break;  // warm-up
        (void)hipEventRecord(ev0, h->stream);
        for (int k = 0; k < steps && rc == KSCR_OK; k++) rc = launch_step(h, dt, false);
        (void)hipEventRecord(ev1, h->stream);
        if ((e = hipEventSynchronize(ev1)) != hipSuccess) { rc = h->hip_fail(e, "hipEventSynchronize"); break; }
        float ms = 0;
        (void)hipEventElapsedTime(&ms, ev0, ev1);
        if (best < 0 || ms < best_ms) best = b, best_ms = ms;
    }
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    h->p.block_size = (rc == KSCR_OK && best > 0) ? best : saved_block;
    int rc2 = upload_all(h);
    if ((e = save(true)) != hipSuccess) return h->hip_fail(e, "autotune restore");
    if (with_lists && h->contact.num_cols != nullptr && h->contact.n == h->n) {
        if ((e = hipMemcpyAsync(h->contact.num_cols, s_num.p, (size_t)h->n * 4, hipMemcpyDeviceToDevice,
                                h->stream)) != hipSuccess ||
            (e = hipMemcpyAsync(h->contact.cols, s_cols.p, (size_t)h->n * kMaxCols * 4, hipMemcpyDeviceToDevice,
                                h->stream)) != hipSuccess)
            return h->hip_fail(e, "autotune restore contact lists");
    }
    if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return h->hip_fail(e, "hipStreamSynchronize");
    h->h = h0, h->last_h = lh0, h->time = t0, h->step_counter = c0, h->last_bvh_rebuild = rebuild0;
    if (rc) return rc;
    if (rc2) return rc2;
    if (best_block) *best_block = h->block;
    return KSCR_OK;
}

}  // extern "C"

