// kscr_api_internal.h -- the handle behind the C ABI of include/kscr.h and
// the helpers shared by the ABI's source files: kscr_api.hip (handle setup,
// state exchange, the step, the autotuner), kscr_api_adjoint.hip
// (differentiation), kscr_api_batch.hip (batched grooms).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <string>
#include <thread>
#include <vector>

#include "../../include/kscr.h"
#include "kscr_contact.h"
#include "kscr_device.h"
#include "kscr_lbvh.h"
#include "kscr_math.h"

namespace kscr {
namespace api {

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
void host_copy(void* dst, const void* src, size_t bytes);

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
    std::vector<float> last_h_of, last_lh_of, last_t_of;  // h, last_h, time of the last segment's steps
    bool with_sdf = false;  // steps are kscr_solve_next_step's (the SDF pass after each)
    bool wind = false;      // the tape holds the adaptive guess's acceleration
    DevBuf<float> t_acc;
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

}  // namespace api
}  // namespace kscr

struct kscr_handle {
    kscr_params p{};
    int n = 0;
    int device = 0;
    std::vector<int> strand_off;
    int longest = 0;  // vertices of the longest strand
    // host mirror (YarnBall::Sim::verts/vels/qs/qRests); positions and
    // velocities page-locked for the per-step exchange
    kscr::api::PinnedBuf<float> pos, vel;
    std::vector<float> inv_mass, l_rest, k_stretch;
    std::vector<uint32_t> flags;
    std::vector<int> conn;  // Vertex::connectionIndex (YarnBall.h:56), -1 if none
    bool has_glue = false;
    std::vector<kscr::f4> q, q_rest;
    float min_seg = FLT_MAX, max_seg = 0, density = 0;
    bool configured = false;
    // step state
    float h = 4e-4f, last_h = 4e-4f, time = 0.f;
    long step_counter = 0;
    // device
    kscr::api::DevBuf<float> d_x, d_v, d_u;
    kscr::api::DevBuf<float4> d_q, d_qr, d_mat;
    kscr::api::DevBuf<int> d_conn;
    kscr::api::DevBuf<int2> d_packs;
    kscr::api::DevBuf<float4> d_grid;
    int num_packs = 0;
    int block = 0;
    bool uploaded = false;
    hipStream_t stream = nullptr;
    // SDF
    bool sdf_set = false, sdf_on = false;
    kscr::SdfArgs sdf{};
    // strain limiting after the SDF pass (kscr_set_strain_limit): strand
    // offsets and per-segment rest lengths on the device
    bool strain_on = false;
    float strain_error = 0.1f;
    kscr::api::DevBuf<int> d_strand_off;
    kscr::api::DevBuf<float> d_strain_rest;
    // hair-hair contacts
    kscr::ContactDev contact;
    float last_bvh_rebuild = INFINITY;
    kscr::api::AdjointRun adj;
    std::string err;
    ~kscr_handle() { kscr::contact_release(contact); }

    int fail(int code, const std::string& msg) {
        err = msg;
        return code;
    }
    int hip_fail(hipError_t e, const char* what) {
        err = std::string(what) + ": " + hipGetErrorString(e);
        return KSCR_ERR_HIP;
    }
};

namespace kscr {
namespace api {

// Setup and step helpers (kscr_api.hip).
int validate_params(kscr_handle* h, const kscr_params& p);
int validate_meta(kscr_handle* h);
bool colored(const kscr_handle* h);
constexpr int kMaxBlock = 1024;  // largest workgroup of the step kernels
bool split_strands(const kscr_handle* h);
bool global_path(const kscr_handle* h);
int pack_capacity(const kscr_handle* h, int block);
int max_block(const kscr_handle* h);
int longest_strand(const kscr_handle* h);
int choose_block(const kscr_handle* h);
int build_packs(kscr_handle* h, int block, std::vector<int2>& packs);
int upload_all(kscr_handle* h);
StepArgs step_args(kscr_handle* h, float dt, bool with_sdf);
int launch_contact_step(kscr_handle* h, float dt, bool with_sdf);
int check_device_error(kscr_handle* h);
int launch_step(kscr_handle* h, float dt, bool with_sdf);
int strain_pass(kscr_handle* h, const StepArgs& a);
int upload_material(kscr_handle* h);

}  // namespace api
}  // namespace kscr
