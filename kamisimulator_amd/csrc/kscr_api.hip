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
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/kscr.h"
#include "kscr_contact.h"
#include "kscr_device.h"
#include "kscr_lbvh.h"
#include "kscr_math.h"

using namespace kscr;

static_assert(sizeof(kscr_params) == 27 * 4, "kscr_params layout is part of the ABI");

namespace {
const double kPi = 3.141592653589793238462643;

thread_local std::string g_create_error;

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    hipError_t alloc(size_t count) {
        release();
        if (count == 0) return hipSuccess;
        hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
        if (e == hipSuccess) n = count;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { release(); }
};
// Page-locked host array (hipHostMalloc): the host mirror of the exchanged
// state, so that device-host copies run at full link speed.
template <class T>
struct PinnedBuf {
    T* p = nullptr;
    size_t n = 0;
    hipError_t resize(size_t count) {
        if (count == n) return hipSuccess;
        release();
        if (count == 0) return hipSuccess;
        hipError_t e = hipHostMalloc((void**)&p, count * sizeof(T), hipHostMallocDefault);
        if (e == hipSuccess) n = count;
        else p = nullptr;
        return e;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
    }
    T* data() { return p; }
    const T* data() const { return p; }
    T& operator[](size_t i) { return p[i]; }
    const T& operator[](size_t i) const { return p[i]; }
    ~PinnedBuf() { release(); }
};

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

// Device state of a differentiable rollout (kscr_adjoint_forward/backward).
struct AdjointRun {
    float dt = 0;
    int nsteps = 0, every = 0, substeps = 0, num_itr = 0;
    bool ready = false;
    // checkpoints: state at steps 0, every, 2 every, ... and at the end
    DevBuf<float> ck_x, ck_v, ck_u;
    DevBuf<float4> ck_q;
    struct Clock {
        float h, last_h, time;
        long counter;
    };
    std::vector<Clock> clocks;
    Clock end_clock{};
    // tape of one segment of `every` steps
    DevBuf<float> t_x, t_y, t_dx;
    DevBuf<unsigned char> t_sigma;
    DevBuf<float4> t_q;
    DevBuf<float> t_ilw;
    // adjoint state
    DevBuf<float> xb, vb, ub, ksb;
    DevBuf<float4> qb, qrb;
    // the forward pass records the last segment's tape itself
    bool last_recorded = false;
    std::vector<float> last_h_of, last_lh_of;  // h, last_h of the last segment's steps
    float record_ms = 0, sweep_ms = 0, backward_ms = 0;
    // the backward pass re-records segment j - 1 on this stream while the
    // handle's stream sweeps segment j (tapes in two slots by segment parity)
    hipStream_t rec_stream = nullptr;
    AdjointRun() = default;
    AdjointRun(const AdjointRun&) = delete;
    AdjointRun& operator=(const AdjointRun&) = delete;
    ~AdjointRun() {
        if (rec_stream) (void)hipStreamDestroy(rec_stream);
    }
};
}  // namespace

struct kscr_handle {
    kscr_params p{};
    int n = 0;
    int device = 0;
    std::vector<int> strand_off;
    int longest = 0;  // vertices of the longest strand
    // host mirror (YarnBall::Sim::verts/vels/qs/qRests); positions and
    // velocities page-locked for the per-step exchange
    PinnedBuf<float> pos, vel;
    std::vector<float> inv_mass, l_rest, k_stretch;
    std::vector<uint32_t> flags;
    std::vector<int> conn;  // Vertex::connectionIndex (YarnBall.h:56), -1 if none
    bool has_glue = false;
    std::vector<f4> q, q_rest;
    float min_seg = FLT_MAX, max_seg = 0, density = 0;
    bool configured = false;
    // step state
    float h = 4e-4f, last_h = 4e-4f, time = 0.f;
    long step_counter = 0;
    // device
    DevBuf<float> d_x, d_v, d_u;
    DevBuf<float4> d_q, d_qr, d_mat;
    DevBuf<int> d_conn;
    DevBuf<int2> d_packs;
    DevBuf<float4> d_grid;
    int num_packs = 0;
    int block = 0;
    bool uploaded = false;
    hipStream_t stream = nullptr;
    // SDF
    bool sdf_set = false, sdf_on = false;
    SdfArgs sdf{};
    // hair-hair contacts
    ContactDev contact;
    float last_bvh_rebuild = INFINITY;
    AdjointRun adj;
    std::string err;
    ~kscr_handle() { contact_release(contact); }

    int fail(int code, const std::string& msg) {
        err = msg;
        return code;
    }
    int hip_fail(hipError_t e, const char* what) {
        err = std::string(what) + ": " + hipGetErrorString(e);
        return KSCR_ERR_HIP;
    }
};

namespace {

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

// Largest workgroup of the step kernels: a strand longer than this is split
// over several packs.
constexpr int kMaxBlock = 1024;

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

int longest_strand(const kscr_handle* h);

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
    return KSCR_OK;
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
hipError_t checkpoint_copy(kscr_handle* h, size_t k, bool save, hipStream_t stream) {
    AdjointRun& A = h->adj;
    const size_t n = h->n, n3 = 3 * n;
    const hipMemcpyKind kind = hipMemcpyDeviceToDevice;
    float* x = A.ck_x.p + k * n3;
    float* v = A.ck_v.p + k * n3;
    float* u = A.ck_u.p + k * n3;
    float4* q = A.ck_q.p + k * n;
    hipError_t e;
    if (save) {
        if ((e = hipMemcpyAsync(x, h->d_x.p, n3 * 4, kind, stream)) != hipSuccess) return e;
        if ((e = hipMemcpyAsync(v, h->d_v.p, n3 * 4, kind, stream)) != hipSuccess) return e;
        if ((e = hipMemcpyAsync(u, h->d_u.p, n3 * 4, kind, stream)) != hipSuccess) return e;
        return hipMemcpyAsync(q, h->d_q.p, n * 16, kind, stream);
    }
    if ((e = hipMemcpyAsync(h->d_x.p, x, n3 * 4, kind, stream)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(h->d_v.p, v, n3 * 4, kind, stream)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(h->d_u.p, u, n3 * 4, kind, stream)) != hipSuccess) return e;
    return hipMemcpyAsync(h->d_q.p, q, n * 16, kind, stream);
}

AdjointRun::Clock clock_of(const kscr_handle* h) { return AdjointRun::Clock{h->h, h->last_h, h->time, h->step_counter}; }
void set_clock(kscr_handle* h, const AdjointRun::Clock& c) {
    h->h = c.h, h->last_h = c.last_h, h->time = c.time, h->step_counter = c.counter;
}

// Tapes of two checkpoint segments (slots 0 and 1), each `every` steps of
// `substeps` substeps.
int reserve_tape(kscr_handle* h) {
    AdjointRun& A = h->adj;
    const size_t n = h->n, S = (size_t)A.substeps, I1 = (size_t)A.num_itr + 1, every = 2 * (size_t)A.every;
    const size_t K = (size_t)A.num_itr;
    const size_t per_sub = S * n * every, per_dx = S * I1 * n * every, per_q = S * I1 * n * every;
    const size_t per_ilw = S * K * n * every;
    if (A.t_dx.n >= per_dx * 3 && A.t_q.n >= per_q && A.t_x.n >= per_sub * 3 && A.t_ilw.n >= per_ilw)
        return KSCR_OK;
    hipError_t e;
    if ((e = A.t_x.alloc(per_sub * 3)) != hipSuccess || (e = A.t_y.alloc(per_sub * 3)) != hipSuccess ||
        (e = A.t_sigma.alloc(per_sub)) != hipSuccess || (e = A.t_dx.alloc(per_dx * 3)) != hipSuccess ||
        (e = A.t_q.alloc(per_q)) != hipSuccess || (e = A.t_ilw.alloc(std::max<size_t>(per_ilw, 1))) != hipSuccess)
        return h->hip_fail(e, "hipMalloc tape");
    return KSCR_OK;
}

// Tape of step j of the segment in tape slot `slot`.
TapeArgs tape_of(kscr_handle* h, size_t slot, size_t j) {
    AdjointRun& A = h->adj;
    const size_t n = h->n, S = (size_t)A.substeps, K = (size_t)A.num_itr;
    j += slot * (size_t)A.every;
    TapeArgs t{};
    t.x = A.t_x.p + j * S * n * 3;
    t.y = A.t_y.p + j * S * n * 3;
    t.sigma_free = A.t_sigma.p + j * S * n;
    t.dx = A.t_dx.p + j * S * (K + 1) * n * 3;
    t.q = A.t_q.p + j * S * (K + 1) * n;
    t.ilw = A.t_ilw.p + j * S * K * n;
    t.n = (int)n;
    return t;
}

int adjoint_supported(kscr_handle* h) {
    if (h->p.enable_hair_hair_coll || h->p.use_step_size_limit)
        return h->fail(KSCR_ERR_UNSUPPORTED, "the adjoint covers the contact-free step (no contacts, no step size limit)");
    if (h->has_glue) return h->fail(KSCR_ERR_UNSUPPORTED, "the adjoint covers strands without glued vertices");
    if (split_strands(h))
        return h->fail(KSCR_ERR_UNSUPPORTED, "the adjoint covers strands of at most " + std::to_string(kMaxBlock) +
                                                 " vertices");
    if (h->p.wind_peak != 0) return h->fail(KSCR_ERR_UNSUPPORTED, "the adjoint covers the step without wind");
    if (colored(h)) return h->fail(KSCR_ERR_UNSUPPORTED, "the adjoint covers the Jacobi schedule");
    return KSCR_OK;
}

}  // namespace

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

int kscr_set_material(kscr_handle* h, const float* k_stretch, const float* q_rest) {
    if (!h) return KSCR_ERR_INVALID_ARGUMENT;
    if (!h->configured) return h->fail(KSCR_ERR_NOT_CONFIGURED, "No rest length. Must call configure()");
    for (int i = 0; i < h->n; i++) {
        if (k_stretch) h->k_stretch[i] = k_stretch[i];
        if (q_rest) h->q_rest[i] = f4{q_rest[4 * i], q_rest[4 * i + 1], q_rest[4 * i + 2], q_rest[4 * i + 3]};
    }
    if (!h->uploaded) return KSCR_OK;
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return h->hip_fail(e, "hipSetDevice");
    return upload_material(h);
}

// ------------------------------------------------------------------ adjoint
int kscr_adjoint_forward(kscr_handle* h, float dt, int nsteps, int checkpoint_every) {
    if (!h || !(dt > 0) || nsteps <= 0 || checkpoint_every <= 0) return KSCR_ERR_INVALID_ARGUMENT;
    if (!h->uploaded) return h->fail(KSCR_ERR_NOT_CONFIGURED, "upload() has not been called");
    int rc = validate_meta(h);
    if (rc) return rc;
    if ((rc = adjoint_supported(h))) return rc;
    hipError_t e;
    if ((e = hipSetDevice(h->device)) != hipSuccess) return h->hip_fail(e, "hipSetDevice");
    AdjointRun& A = h->adj;
    A.ready = false;
    const size_t n = h->n;
    const size_t slots = (size_t)(nsteps + checkpoint_every - 1) / checkpoint_every + 1;  // + end state
    if (A.ck_q.n < slots * n) {
        if ((e = A.ck_x.alloc(slots * n * 3)) != hipSuccess || (e = A.ck_v.alloc(slots * n * 3)) != hipSuccess ||
            (e = A.ck_u.alloc(slots * n * 3)) != hipSuccess || (e = A.ck_q.alloc(slots * n)) != hipSuccess)
            return h->hip_fail(e, "hipMalloc checkpoints");
    }
    A.dt = dt;
    A.nsteps = nsteps;
    A.every = checkpoint_every;
    A.substeps = std::max(1, (int)std::ceil(dt / h->p.max_h));
    A.num_itr = h->p.num_itr;
    A.last_recorded = false;
    if ((rc = reserve_tape(h)) != KSCR_OK) return rc;
    // steps of the last segment write their tape as they run
    const int last_s0 = ((nsteps - 1) / checkpoint_every) * checkpoint_every;
    A.last_h_of.assign(checkpoint_every, 0.f);
    A.last_lh_of.assign(checkpoint_every, 0.f);
    A.clocks.clear();
    for (int s = 0; s < nsteps; s++) {
        if (s % checkpoint_every == 0) {
            if ((e = checkpoint_copy(h, (size_t)(s / checkpoint_every), true, h->stream)) != hipSuccess)
                return h->hip_fail(e, "save checkpoint");
            A.clocks.push_back(clock_of(h));
        }
        if (s < last_s0) {
            if ((rc = launch_step(h, dt, false)) != KSCR_OK) return rc;
            continue;
        }
        const StepArgs sa = step_args(h, dt, false);
        A.last_h_of[s - last_s0] = sa.h;
        A.last_lh_of[s - last_s0] = sa.last_h;
        if ((e = launch_step_record(sa, tape_of(h, 0, (size_t)(s - last_s0)), h->d_packs.p, h->num_packs, h->block,
                                    h->stream)) != hipSuccess)
            return h->hip_fail(e, "record kernel launch");
    }
    A.last_recorded = true;
    A.ready = true;
    return KSCR_OK;
}

int kscr_adjoint_backward(kscr_handle* h, const float* x_bar, const float* v_bar, float* k_stretch_bar,
                          float* q_rest_bar) {
    if (!h || !k_stretch_bar || !q_rest_bar) return KSCR_ERR_INVALID_ARGUMENT;
    AdjointRun& A = h->adj;
    if (!A.ready) return h->fail(KSCR_ERR_NOT_CONFIGURED, "kscr_adjoint_forward has not been run");
    if (h->p.num_itr != A.num_itr || std::max(1, (int)std::ceil(A.dt / h->p.max_h)) != A.substeps)
        return h->fail(KSCR_ERR_INVALID_ARGUMENT, "num_itr or max_h changed since the forward pass");
    int rc = adjoint_supported(h);
    if (rc) return rc;
    hipError_t e;
    if ((e = hipSetDevice(h->device)) != hipSuccess) return h->hip_fail(e, "hipSetDevice");
    const size_t n = h->n, n3 = 3 * n;
    if ((rc = reserve_tape(h)) != KSCR_OK) return rc;
    if (A.qb.n != n) {
        if ((e = A.xb.alloc(n3)) != hipSuccess || (e = A.vb.alloc(n3)) != hipSuccess || (e = A.ub.alloc(n3)) != hipSuccess ||
            (e = A.ksb.alloc(n)) != hipSuccess || (e = A.qb.alloc(n)) != hipSuccess || (e = A.qrb.alloc(n)) != hipSuccess)
            return h->hip_fail(e, "hipMalloc adjoint state");
    }
    const hipMemcpyKind h2d = hipMemcpyHostToDevice;
    if ((e = x_bar ? hipMemcpyAsync(A.xb.p, x_bar, n3 * 4, h2d, h->stream) : hipMemsetAsync(A.xb.p, 0, n3 * 4, h->stream)) !=
            hipSuccess ||
        (e = v_bar ? hipMemcpyAsync(A.vb.p, v_bar, n3 * 4, h2d, h->stream) : hipMemsetAsync(A.vb.p, 0, n3 * 4, h->stream)) !=
            hipSuccess ||
        (e = hipMemsetAsync(A.ub.p, 0, n3 * 4, h->stream)) != hipSuccess ||
        (e = hipMemsetAsync(A.ksb.p, 0, n * 4, h->stream)) != hipSuccess ||
        (e = hipMemsetAsync(A.qb.p, 0, n * 16, h->stream)) != hipSuccess ||
        (e = hipMemsetAsync(A.qrb.p, 0, n * 16, h->stream)) != hipSuccess)
        return h->hip_fail(e, "seed adjoint state");
    const int nseg = (int)A.clocks.size();
    const size_t end_slot = (size_t)nseg;
    if ((e = checkpoint_copy(h, end_slot, true, h->stream)) != hipSuccess) return h->hip_fail(e, "save end state");
    const AdjointRun::Clock end_clock = clock_of(h);
    if (!A.rec_stream && (e = hipStreamCreateWithFlags(&A.rec_stream, hipStreamNonBlocking)) != hipSuccess)
        return h->hip_fail(e, "hipStreamCreate");
    hipStream_t ms = h->stream, rs = A.rec_stream;

    AdjointArgs aa{};
    aa.mat = h->d_mat.p;
    aa.q_rest = h->d_qr.p;
    aa.num_substeps = A.substeps;
    aa.num_itr = A.num_itr;
    aa.damping = h->p.damping;
    aa.velocity_decay = h->p.velocity_decay;
    aa.acceleration_ratio = h->p.acceleration_ratio;
    for (int k = 0; k < 3; k++) aa.ext[k] = h->p.gravity[k] * h->p.external_force_multiplier;
    // |ext|^2 as the device dot product forms it
    aa.ext2 = std::fma(aa.ext[2], aa.ext[2], std::fma(aa.ext[1], aa.ext[1], aa.ext[0] * aa.ext[0]));
    aa.xb = A.xb.p;
    aa.vb = A.vb.p;
    aa.ub = A.ub.p;
    aa.qb = A.qb.p;
    aa.ksb = A.ksb.p;
    aa.qrb = A.qrb.p;
    // events: per segment record begin/end (record stream) and sweep
    // begin/end (handle stream); then start, record-stream done, end
    std::vector<hipEvent_t> ev(4 * (size_t)nseg + 3, nullptr);
    for (auto& x : ev)
        if ((e = hipEventCreate(&x)) != hipSuccess) {
            rc = h->hip_fail(e, "hipEventCreate");
            break;
        }
    auto rec_begin = [&](int seg) { return ev[4 * seg]; };
    auto rec_end = [&](int seg) { return ev[4 * seg + 1]; };
    auto sweep_begin = [&](int seg) { return ev[4 * seg + 2]; };
    auto sweep_end = [&](int seg) { return ev[4 * seg + 3]; };
    hipEvent_t ev_start = ev[4 * nseg], ev_rec_done = ev[4 * nseg + 1], ev_end = ev[4 * nseg + 2];
    // Segment j's tape lives in slot (nseg - 1 - j) % 2: the last segment in
    // slot 0, where the forward pass recorded it.  Re-recording segment j - 1
    // (record stream) overlaps the sweep of segment j (handle stream); it
    // waits for the sweep of segment j + 1, the last user of its slot.
    auto slot_of = [&](int seg) { return (size_t)((nseg - 1 - seg) % 2); };
    std::vector<std::vector<float>> hs(nseg), lhs(nseg);
    std::vector<bool> rerecorded(nseg, false);
    if (rc == KSCR_OK) {
        (void)hipEventRecord(ev_start, ms);
        if ((e = hipStreamWaitEvent(rs, ev_start, 0)) != hipSuccess) rc = h->hip_fail(e, "hipStreamWaitEvent");
    }
    for (int seg = nseg - 1; seg >= 0 && rc == KSCR_OK; seg--) {
        const int s0 = seg * A.every, s1 = std::min(A.nsteps, s0 + A.every);
        if (seg == nseg - 1 && A.last_recorded) {
            hs[seg] = A.last_h_of;  // recorded by the forward pass
            lhs[seg] = A.last_lh_of;
        } else {
            rerecorded[seg] = true;
            hs[seg].assign(A.every, 0.f);
            lhs[seg].assign(A.every, 0.f);
            if (seg + 2 <= nseg - 1 && (e = hipStreamWaitEvent(rs, sweep_end(seg + 2), 0)) != hipSuccess) {
                rc = h->hip_fail(e, "hipStreamWaitEvent");
                break;
            }
            (void)hipEventRecord(rec_begin(seg), rs);
            if ((e = checkpoint_copy(h, (size_t)seg, false, rs)) != hipSuccess) {
                rc = h->hip_fail(e, "restore checkpoint");
                break;
            }
            set_clock(h, A.clocks[seg]);
            for (int s = s0; s < s1; s++) {
                const StepArgs sa = step_args(h, A.dt, false);
                hs[seg][s - s0] = sa.h;
                lhs[seg][s - s0] = sa.last_h;
                if ((e = launch_step_record(sa, tape_of(h, slot_of(seg), (size_t)(s - s0)), h->d_packs.p,
                                            h->num_packs, h->block, rs)) != hipSuccess) {
                    rc = h->hip_fail(e, "record kernel launch");
                    break;
                }
            }
            (void)hipEventRecord(rec_end(seg), rs);
            if (rc == KSCR_OK && (e = hipStreamWaitEvent(ms, rec_end(seg), 0)) != hipSuccess)
                rc = h->hip_fail(e, "hipStreamWaitEvent");
        }
        (void)hipEventRecord(sweep_begin(seg), ms);
        for (int s = s1 - 1; s >= s0 && rc == KSCR_OK; s--) {
            aa.tape = tape_of(h, slot_of(seg), (size_t)(s - s0));
            aa.h = hs[seg][s - s0];
            aa.last_h = lhs[seg][s - s0];
            if ((e = launch_step_adjoint(aa, h->d_packs.p, h->num_packs, h->block, ms)) != hipSuccess)
                rc = h->hip_fail(e, "adjoint kernel launch");
        }
        (void)hipEventRecord(sweep_end(seg), ms);
    }
    // the record stream's work (it writes the state arrays) ends before the
    // end state is restored
    (void)hipEventRecord(ev_rec_done, rs);
    (void)hipStreamWaitEvent(ms, ev_rec_done, 0);
    if (nseg > 1) A.last_recorded = false;  // earlier segments reused the tape
    if ((e = checkpoint_copy(h, end_slot, false, ms)) != hipSuccess && rc == KSCR_OK) rc = h->hip_fail(e, "restore end state");
    set_clock(h, end_clock);
    (void)hipEventRecord(ev_end, ms);
    if (rc == KSCR_OK &&
        ((e = hipMemcpyAsync(k_stretch_bar, A.ksb.p, n * 4, hipMemcpyDeviceToHost, ms)) != hipSuccess ||
         (e = hipMemcpyAsync(q_rest_bar, A.qrb.p, n * 16, hipMemcpyDeviceToHost, ms)) != hipSuccess))
        rc = h->hip_fail(e, "download gradient");
    if ((e = hipStreamSynchronize(ms)) != hipSuccess && rc == KSCR_OK) rc = h->hip_fail(e, "hipStreamSynchronize");
    if ((e = hipStreamSynchronize(rs)) != hipSuccess && rc == KSCR_OK) rc = h->hip_fail(e, "hipStreamSynchronize");
    A.record_ms = A.sweep_ms = A.backward_ms = 0;
    if (rc == KSCR_OK) {
        for (int seg = 0; seg < nseg; seg++) {
            float r = 0, w = 0;
            if (rerecorded[seg]) (void)hipEventElapsedTime(&r, rec_begin(seg), rec_end(seg));
            (void)hipEventElapsedTime(&w, sweep_begin(seg), sweep_end(seg));
            A.record_ms += r;
            A.sweep_ms += w;
        }
        (void)hipEventElapsedTime(&A.backward_ms, ev_start, ev_end);
    }
    for (auto& x : ev)
        if (x) (void)hipEventDestroy(x);
    return rc;
}

int kscr_adjoint_last_time(const kscr_handle* h, float* record_ms, float* sweep_ms, float* backward_ms) {
    if (!h) return KSCR_ERR_INVALID_ARGUMENT;
    if (record_ms) *record_ms = h->adj.record_ms;
    if (sweep_ms) *sweep_ms = h->adj.sweep_ms;
    if (backward_ms) *backward_ms = h->adj.backward_ms;
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

// ------------------------------------------------------------------ batch
struct kscr_batch {
    std::vector<kscr_handle*> grooms;
    int device = 0;
    int block = 0;
    int num_packs = 0;
    DevBuf<StepArgs> d_args;
    DevBuf<int4> d_packs;
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
// (begin, end, groom, 0).
int batch_set_block(kscr_batch* b, int block) {
    std::vector<int4> packs;
    for (size_t g = 0; g < b->grooms.size(); g++) {
        std::vector<int2> p2;
        const int rc = build_packs(b->grooms[g], block, p2);
        if (rc) return rc;
        for (const int2& pk : p2) packs.push_back(make_int4(pk.x, pk.y, (int)g, 0));
    }
    hipError_t e;
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
            if ((e = launch_step_batch(b->d_args.p, b->d_packs.p, b->num_packs, b->block, b->stream)) != hipSuccess)
                return b->hip_fail(e, "batch step launch");
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
            le = launch_step_batch(b->d_args.p, b->d_packs.p, b->num_packs, b->block, b->stream);
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
    const int start_block = b->block;
    int best = -1, rc = KSCR_OK;
    float best_ms = 0;
    for (int c : kBlockSizes) {
        if (c < b->longest) continue;
        if ((rc = batch_set_block(b, c)) != KSCR_OK) break;
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
        if (best < 0 || ms < best_ms) best = c, best_ms = ms;
    }
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

int kscr_batch_last_error(const kscr_batch* b, char* buf, size_t n) {
    if (!b || !buf || n == 0) return KSCR_ERR_INVALID_ARGUMENT;
    std::snprintf(buf, n, "%s", b->err.c_str());
    return KSCR_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ LBVH
struct kscr_lbvh {
    int device = 0;
    LbvhDev d;  // its counter block holds the query's error word at d.count[2]
    int2* pairs = nullptr;
    size_t pair_cap = 0;
    ~kscr_lbvh() {
        lbvh_release(d);
        if (pairs) (void)hipFree(pairs);
    }
};

extern "C" {

int kscr_lbvh_create(int device, kscr_lbvh** out) {
    if (!out) return KSCR_ERR_INVALID_ARGUMENT;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return KSCR_ERR_HIP;
    if (device < 0) device = count - 1;
    if (device >= count) return KSCR_ERR_INVALID_ARGUMENT;
    auto* b = new kscr_lbvh;
    b->device = device;
    if (hipSetDevice(device) != hipSuccess) {
        delete b;
        return KSCR_ERR_HIP;
    }
    *out = b;
    return KSCR_OK;
}

void kscr_lbvh_destroy(kscr_lbvh* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    (void)hipDeviceSynchronize();
    delete b;
}

static int lbvh_upload(kscr_lbvh* b, const float* aabbs, int n) {
    if (hipSetDevice(b->device) != hipSuccess) return KSCR_ERR_HIP;
    if (lbvh_reserve(b->d, n) != hipSuccess) return KSCR_ERR_OUT_OF_MEMORY;
    if (hipMemcpy(b->d.aabb, aabbs, (size_t)n * 6 * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
        return KSCR_ERR_HIP;
    return KSCR_OK;
}

int kscr_lbvh_compute(kscr_lbvh* b, const float* aabbs, int n) {
    if (!b || !aabbs || n <= 0) return KSCR_ERR_INVALID_ARGUMENT;
    int rc = lbvh_upload(b, aabbs, n);
    if (rc) return rc;
    if (lbvh_build(b->d, n, nullptr) != hipSuccess || hipDeviceSynchronize() != hipSuccess) return KSCR_ERR_HIP;
    return KSCR_OK;
}

int kscr_lbvh_refit(kscr_lbvh* b, const float* aabbs) {
    if (!b || !aabbs || !b->d.built) return KSCR_ERR_INVALID_ARGUMENT;
    int rc = lbvh_upload(b, aabbs, b->d.n);
    if (rc) return rc;
    if (lbvh_refit(b->d, nullptr) != hipSuccess || hipDeviceSynchronize() != hipSuccess) return KSCR_ERR_HIP;
    return KSCR_OK;
}

int kscr_lbvh_query(kscr_lbvh* b, int* pairs, size_t cap, size_t* count) {
    if (!b || !count || !b->d.built || (cap > 0 && !pairs)) return KSCR_ERR_INVALID_ARGUMENT;
    if (hipSetDevice(b->device) != hipSuccess) return KSCR_ERR_HIP;
    if (cap > b->pair_cap) {
        if (b->pairs) (void)hipFree(b->pairs);
        b->pairs = nullptr;
        if (hipMalloc((void**)&b->pairs, cap * sizeof(int2)) != hipSuccess) return KSCR_ERR_OUT_OF_MEMORY;
        b->pair_cap = cap;
    }
    if (lbvh_query(b->d, b->pairs, (int)cap, -1, b->d.count + 2, nullptr) != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess)
        return KSCR_ERR_HIP;
    int cnt[3] = {0, 0, 0};  // pairs found, redo count, error word
    if (hipMemcpy(cnt, b->d.count, sizeof cnt, hipMemcpyDeviceToHost) != hipSuccess) return KSCR_ERR_HIP;
    const int found = cnt[0], err = cnt[2];
    const size_t m = std::min((size_t)found, cap);
    if (m && hipMemcpy(pairs, b->pairs, m * sizeof(int2), hipMemcpyDeviceToHost) != hipSuccess) return KSCR_ERR_HIP;
    *count = (size_t)found;
    return err ? KSCR_ERR_MAX_COLLISIONS : KSCR_OK;
}

int kscr_lbvh_compute_device(kscr_lbvh* b, const float* d_aabbs, int n, void* stream) {
    if (!b || !d_aabbs || n <= 0) return KSCR_ERR_INVALID_ARGUMENT;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    if (hipSetDevice(b->device) != hipSuccess) return KSCR_ERR_HIP;
    if (lbvh_reserve(b->d, n) != hipSuccess) return KSCR_ERR_OUT_OF_MEMORY;
    if (hipMemcpyAsync(b->d.aabb, d_aabbs, (size_t)n * 6 * sizeof(float), hipMemcpyDeviceToDevice, s) != hipSuccess ||
        lbvh_build(b->d, n, s) != hipSuccess)
        return KSCR_ERR_HIP;
    return KSCR_OK;
}

int kscr_lbvh_refit_device(kscr_lbvh* b, const float* d_aabbs, void* stream) {
    if (!b || !d_aabbs || !b->d.built) return KSCR_ERR_INVALID_ARGUMENT;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    if (hipSetDevice(b->device) != hipSuccess) return KSCR_ERR_HIP;
    if (hipMemcpyAsync(b->d.aabb, d_aabbs, (size_t)b->d.n * 6 * sizeof(float), hipMemcpyDeviceToDevice, s) !=
            hipSuccess ||
        lbvh_refit(b->d, s) != hipSuccess)
        return KSCR_ERR_HIP;
    return KSCR_OK;
}

int kscr_lbvh_query_device(kscr_lbvh* b, int* d_pairs, size_t cap, size_t* count, void* stream) {
    if (!b || !count || !b->d.built || (cap > 0 && !d_pairs) || cap > (size_t)INT32_MAX)
        return KSCR_ERR_INVALID_ARGUMENT;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    if (hipSetDevice(b->device) != hipSuccess) return KSCR_ERR_HIP;
    if (lbvh_query(b->d, reinterpret_cast<int2*>(d_pairs), (int)cap, -1, b->d.count + 2, s) != hipSuccess)
        return KSCR_ERR_HIP;
    int cnt[3] = {0, 0, 0};  // pairs found, redo count, error word: one copy back
    if (hipMemcpyAsync(cnt, b->d.count, sizeof cnt, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return KSCR_ERR_HIP;
    *count = (size_t)cnt[0];
    return cnt[2] ? KSCR_ERR_MAX_COLLISIONS : KSCR_OK;
}

int kscr_lbvh_get(kscr_lbvh* b, uint32_t* codes, uint32_t* ids, uint32_t* links, float* bounds, uint32_t* leaf_parent,
                  float* root, int* max_stack) {
    if (!b || !b->d.built) return KSCR_ERR_INVALID_ARGUMENT;
    const int n = b->d.n;
    if (hipSetDevice(b->device) != hipSuccess) return KSCR_ERR_HIP;
    if (codes && hipMemcpy(codes, b->d.codes, n * 4, hipMemcpyDeviceToHost) != hipSuccess) return KSCR_ERR_HIP;
    if (ids && hipMemcpy(ids, b->d.ids, n * 4, hipMemcpyDeviceToHost) != hipSuccess) return KSCR_ERR_HIP;
    if (leaf_parent && hipMemcpy(leaf_parent, b->d.leaf_parent, n * 4, hipMemcpyDeviceToHost) != hipSuccess)
        return KSCR_ERR_HIP;
    if (root && hipMemcpy(root, b->d.root, 6 * 4, hipMemcpyDeviceToHost) != hipSuccess) return KSCR_ERR_HIP;
    if (max_stack && hipMemcpy(max_stack, b->d.max_depth, 4, hipMemcpyDeviceToHost) != hipSuccess) return KSCR_ERR_HIP;
    if ((links || bounds) && n > 1) {
        std::vector<LbvhNode> nodes(n - 1);
        if (hipMemcpy(nodes.data(), b->d.nodes, (size_t)(n - 1) * sizeof(LbvhNode), hipMemcpyDeviceToHost) != hipSuccess)
            return KSCR_ERR_HIP;
        for (int i = 0; i < n - 1; i++) {
            if (links) {
                links[4 * i] = nodes[i].parent;
                links[4 * i + 1] = nodes[i].left;
                links[4 * i + 2] = nodes[i].right;
                links[4 * i + 3] = nodes[i].fence;
            }
            if (bounds)
                for (int k = 0; k < 12; k++) bounds[12 * i + k] = nodes[i].b[k / 6][k % 6];
        }
    }
    return KSCR_OK;
}

}  // extern "C"
