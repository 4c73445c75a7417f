// kscr_contact.h -- the step with hair-hair contacts (per-substep kernels).
//
// With contacts a vertex's solve reads the dx of partner segments of other
// strands every iteration, so the step cannot stay inside one workgroup: it
// runs as a sequence of kernels per substep (the reference's own structure,
// YB/sim/step.cpp:9-62), with the Jacobi snapshot kept by ping-pong buffers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kscr_device.h"
#include "kscr_lbvh.h"

namespace kscr {

constexpr int kMaxCols = 16;  // MAX_COLLISIONS_PER_SEGMENT (YarnBall.h:19)

enum : int { DEV_ERR_NONE = 0, DEV_ERR_MAX_COLLISIONS = 1, DEV_ERR_PAIR_OVERFLOW = 2, DEV_ERR_QUERY_STACK = 4 };

struct ContactDev {
    int n = 0;
    float* dx[2] = {nullptr, nullptr};  // float[3n] ping-pong iterates
    float* y = nullptr;                 // float[3n] inertial target (as a delta)
    float4* q2 = nullptr;               // ping-pong partner of the handle's q
    float* max_step = nullptr;          // float[n]
    int* num_cols = nullptr;            // int[n]
    int* cols = nullptr;                // int[16n], 16 slots per segment, sorted
    int2* pairs = nullptr;
    int pair_cap = 0;
    int* error = nullptr;  // device error word (DEV_ERR_* bits)
    uint8_t* contact_pack = nullptr;  // uint8[n] (>= packs): pack p holds a contact this substep
    LbvhDev bvh;
};

hipError_t contact_reserve(ContactDev& c, int n);
void contact_release(ContactDev& c);

struct ContactParams {
    float radius, barrier, k_collision, k_friction, friction_coeff;
    float detection_radius, scaled_detection_radius, detection_scaler;
    int use_step_size_limit, use_velocity_radius;
};

// The start of a substep: endItr of the previous one from c.dx[prev] (prev
// < 0: none), initItr into c.dx[0] and c.y (and the lastVel update), and with
// `boxes` the padded segment boxes into c.bvh.aabb (buildAABBs).
hipError_t contact_begin(const StepArgs& a, const ContactParams& cp, ContactDev& c, int prev, bool boxes,
                         hipStream_t s);
// detectCollisions after contact_begin wrote the boxes: LBVH build or refit,
// query, narrow phase, sorted lists.
hipError_t contact_detect(const StepArgs& a, const ContactParams& cp, ContactDev& c, bool rebuild, hipStream_t s);
// recomputeStepLimitKernel
hipError_t contact_step_limit(const StepArgs& a, const ContactParams& cp, ContactDev& c, hipStream_t s);
// One Jacobi iteration, dx[cur] -> dx[cur^1] and q_in -> q_out, over the packs
// of whole strands (at most `block` vertices each, one lane per vertex).
hipError_t contact_iterate(const StepArgs& a, const ContactParams& cp, ContactDev& c, int cur, const float4* q_in,
                           float4* q_out, const int2* packs, int num_packs, int block, hipStream_t s);
// endItr from dx[cur]
// contact_pack from the current contact lists (after detection).
hipError_t contact_classify(const StepArgs& a, ContactDev& c, const int2* packs, int num_packs, int block,
                            hipStream_t s);
// All iterations of the substep for the packs without contacts: reads dx[0]
// and q_in, writes dx[num_itr & 1] and q_fin (the buffers the last of the
// contact_iterate calls writes).
hipError_t contact_free_iterate(const StepArgs& a, ContactDev& c, const float4* q_in, float4* q_fin, const int2* packs,
                                int num_packs, int block, hipStream_t s);
hipError_t contact_end(const StepArgs& a, ContactDev& c, int cur, hipStream_t s);
// SDF pass over all non-root vertices
hipError_t sdf_pass(const StepArgs& a, int n, hipStream_t s);

}  // namespace kscr
