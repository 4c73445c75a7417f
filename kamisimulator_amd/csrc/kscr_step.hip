// kscr_step.hip -- fused Stable Cosserat Rods step for gfx950.
//
// One workgroup owns a pack of whole strands (at most BLOCK vertices, one
// vertex per thread).  A launch performs a whole SolveNextStep(dt):
//   for each substep:  initItr; numItr x (position solve; orientation solve);
//                      endItr
//   then the SDF body-collision pass (when enabled).
// Vertex state lives in registers for the whole launch; the only values that
// cross threads (a neighbour's dx, q and the force/Hessian a segment adds to
// its upper vertex) go through LDS, double buffered so that every read of
// iteration k sees iteration k-1 values ("Jacobi snapshot").  HBM is touched
// once per launch: 84 B read and 52 B written per vertex.
//
// Reference semantics followed (file:line in the reference):
//   initItr                YB/sim/iteration.cu:8-59
//   position solve         YB/sim/cosserat.cu:112-287 (no hair-hair contacts)
//   orientation solve      YB/sim/cosserat.cu:290-337
//   endItr                 YB/sim/iteration.cu:62-74
//   substep split          YB/sim/step.cpp:38-62 (computed on the host)
//   SDF pass               SDF/cudaSDFCollision.cu:88-155, :205-268
#include <algorithm>

#include "kscr_device.h"
#include "kscr_math.h"
#include "kscr_solve.h"

namespace kscr {
namespace {

template <int BLOCK>
struct StepLds {
    float x[3][BLOCK];
    float dx[2][3][BLOCK];
    float4 q[2][BLOCK];
    float4 f[BLOCK];
};

// Tape stores are non-temporal: the tape of a checkpoint segment is read back
// only after the whole segment has been recorded, far beyond the caches.
__device__ inline void tape3(float* base, int slot, int n, int i, f3 v) {
    float* p = base + (size_t)slot * n * 3 + 3 * (size_t)i;
    __builtin_nontemporal_store(v.x, p);
    __builtin_nontemporal_store(v.y, p + 1);
    __builtin_nontemporal_store(v.z, p + 2);
}
typedef float tape_f4 __attribute__((ext_vector_type(4)));
__device__ inline void tape4(float4* p, f4 v) {
    __builtin_nontemporal_store(tape_f4{v.x, v.y, v.z, v.w}, reinterpret_cast<tape_f4*>(p));
}

// Advances the strands of one pack [vbeg, vend) through a whole step; with
// RECORD it also writes the tape of the step.
template <int BLOCK, bool RECORD = false>
__device__ __forceinline__ void run_pack(const StepArgs& a, const int vbeg, const int vend, StepLds<BLOCK>& L,
                                         const TapeArgs* tp = nullptr) {
    float(&s_x)[3][BLOCK] = L.x;
    float(&s_dx)[2][3][BLOCK] = L.dx;
    float4(&s_q)[2][BLOCK] = L.q;
    float4(&s_f)[BLOCK] = L.f;
    const int t = threadIdx.x;
    const int i = vbeg + t;
    const bool act = i < vend;
    const int tm = t > 0 ? t - 1 : 0;
    const int tp1 = t < BLOCK - 1 ? t + 1 : t;

    f3 x{0, 0, 0}, v{0, 0, 0}, u{0, 0, 0};
    f4 q{0, 0, 0, 1}, qr{0, 0, 0, 1}, qrp{0, 0, 0, 1};
    float im = 0.f, lr = 1.f, ks = 0.f, lrp = 1.f, ksp = 0.f;
    uint32_t fl = 0;
    if (act) {
        x = ld3(a.x, i);
        v = ld3(a.v, i);
        u = ld3(a.u, i);
        q = to4(a.q[i]);
        qr = to4(a.q_rest[i]);
        const float4 m = a.mat[i];
        im = m.x;
        lr = m.y;
        ks = m.z;
        fl = __float_as_uint(m.w);
        if (fl & FLAG_PREV) {
            qrp = to4(a.q_rest[i - 1]);
            const float4 mp = a.mat[i - 1];
            lrp = mp.y;
            ksp = mp.z;
        }
    }
    const bool hasNext = (fl & FLAG_NEXT) != 0;
    const bool hasPrev = (fl & FLAG_PREV) != 0;
    const bool free = act && im != 0.f;
    const bool rotates = act && hasNext && !(fl & FLAG_FIX_ORI);

    const float h = a.h;
    const float damp = a.damping / h;
    const float invl = 1.0f / lr;
    const float kk = ks * invl;
    const float dd = kk * invl;
    const float a_in = 1.0f / (im * h * h);
    const float f_in = 1.0f / (h * h * im);
    // Diagonal of the position-solve matrix: a_in, plus the own segment's
    // stretch term, plus the previous segment's (the value the previous
    // vertex computes as dd (1 + damp)); constant over the step, so its
    // inverse is formed once.
    float hd = a_in;
    if (hasNext) hd = hd + dd * (1 + damp);
    if (hasPrev) {
        const float invlp = 1.0f / lrp;
        const float kkp = ksp * invlp;
        const float ddp = kkp * invlp;
        hd = hd + ddp * (1 + damp);
    }
    const InertiaInv hinv = inertia_inverse(hd, a_in);

    for (int s = 0; s < a.num_substeps; s++) {
        // ---- initItr: inertial target y and VBD adaptive initial guess
        f3 y, dx;
        if (RECORD && act) tape3(tp->x, s, tp->n, i, x);
        bool sigma_free;
        f3 acc;
        init_vertex(a, free, x, v, u, y, dx, sigma_free, acc);
        if (RECORD && act) {
            tape3(tp->y, s, tp->n, i, y);
            if (tp->acc) tape3(tp->acc, s, tp->n, i, acc);
            tp->sigma_free[(size_t)s * tp->n + i] = sigma_free ? 1 : 0;
            tape3(tp->dx, s * (a.num_itr + 1), tp->n, i, dx);
            tape4(tp->q + (size_t)s * (a.num_itr + 1) * tp->n + i, q);
        }
        __syncthreads();  // previous substep's LDS reads are complete
        s_x[0][t] = x.x;
        s_x[1][t] = x.y;
        s_x[2][t] = x.z;
        s_dx[0][0][t] = dx.x;
        s_dx[0][1][t] = dx.y;
        s_dx[0][2][t] = dx.z;
        s_q[0][t] = make_float4(q.x, q.y, q.z, q.w);
        __syncthreads();
        const f3 xj = hasNext ? f3{s_x[0][t + 1], s_x[1][t + 1], s_x[2][t + 1]} : x;

        for (int k = 0; k < a.num_itr; k++) {
            const int cur = k & 1;
            // ---- position solve, part 1: own segment (i, i+1)
            // fo stays -0 without a segment: x + (-0) == x for every float x
            // (also x = -0), so it is added unconditionally below
            f3 fo{-0.0f, -0.0f, -0.0f};
            if (hasNext) {
                const f3 dxj{s_dx[cur][0][t + 1], s_dx[cur][1][t + 1], s_dx[cur][2][t + 1]};
                const f3 c = ((xj - x) + (dxj - dx)) * invl - rotate_e1(q);
                fo = kk * c - (damp * dd) * dx;
                const f3 f2 = (-kk) * c - (damp * dd) * dxj;
                s_f[t] = make_float4(f2.x, f2.y, f2.z, 0.0f);
            }
            __syncthreads();
            // ---- position solve, part 2: local 3x3 Newton step
            f3 dxn = dx;
            if (free) {
                f3 f = f_in * (y - dx);
                f = f + fo;
                if (hasPrev) {
                    const float4 p = s_f[t - 1];
                    f = f + f3{p.x, p.y, p.z};
                }
                dxn = dx + a.acceleration_ratio * inertia_solve(hinv, f);
            }
            s_dx[cur ^ 1][0][t] = dxn.x;
            s_dx[cur ^ 1][1][t] = dxn.y;
            s_dx[cur ^ 1][2][t] = dxn.z;
            __syncthreads();
            // ---- orientation solve (closed form, SCR Eq. 18)
            f4 qn = q;
            float inv_norm = 0.0f;
            if (rotates) {
                const f3 dxj{s_dx[cur ^ 1][0][t + 1], s_dx[cur ^ 1][1][t + 1], s_dx[cur ^ 1][2][t + 1]};
                // read unconditionally (clamped into the pack); orientation_update
                // uses qm only with FLAG_PREV and qp only with FLAG_NEXT_ORI
                const f4 qm = to4(s_q[cur][tm]);
                const f4 qp = to4(s_q[cur][tp1]);
                qn = orientation_update(x, xj, dxn, dxj, lr, invl, ks, q, qm, qp, qrp, qr, fl, inv_norm);
            }
            s_q[cur ^ 1][t] = make_float4(qn.x, qn.y, qn.z, qn.w);
            dx = dxn;
            q = qn;
            if (RECORD && act) {
                tape3(tp->dx, s * (a.num_itr + 1) + k + 1, tp->n, i, dx);
                tape4(tp->q + (size_t)(s * (a.num_itr + 1) + k + 1) * tp->n + i, q);
                __builtin_nontemporal_store(inv_norm, tp->ilw + (size_t)(s * a.num_itr + k) * tp->n + i);
            }
            // The next iteration's first LDS write (s_f) is ordered after this
            // iteration's s_f reads by the barrier above; its s_dx/s_q writes
            // target the buffers this iteration only read before that barrier.
        }
        // ---- endItr
        end_vertex(a, free, dx, x, v);
    }

    if (a.sdf_on && act && hasPrev) sdf_project(a.sdf, x, v);

    if (act) {
        st3(a.x, i, x);
        st3(a.v, i, v);
        st3(a.u, i, u);
        a.q[i] = make_float4(q.x, q.y, q.z, q.w);
    }
}

// At 512 threads the LDS of three workgroups fits a CU (6 waves per SIMD);
// reg_waves keeps the register allocation within the 80 VGPRs that allow
// those 6 waves (at 1024 threads the LDS admits one workgroup, 4 waves).
constexpr int reg_waves(int block) { return block >= 1024 ? 4 : 6; }

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(reg_waves(BLOCK)))) void step_fused_kernel(
    StepArgs a, const int2* __restrict__ packs) {
    __shared__ StepLds<BLOCK> lds;
    const int2 pk = packs[blockIdx.x];
    run_pack<BLOCK>(a, pk.x, pk.y, lds);
}

// The record kernel runs beside the reverse sweep (two streams) and is held
// at the plain step's register budget so that two waves of each fit a SIMD
// (2 x 80 + 2 x 168 VGPRs); it then spills a few values to scratch.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(reg_waves(BLOCK)))) void step_record_kernel(
    StepArgs a, TapeArgs tape, const int2* __restrict__ packs) {
    __shared__ StepLds<BLOCK> lds;
    const int2 pk = packs[blockIdx.x];
    run_pack<BLOCK, true>(a, pk.x, pk.y, lds, &tape);
}

// Batched form: every workgroup reads its groom's arguments (pointers,
// substep lengths, gravity, SDF grid) from a device table.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(reg_waves(BLOCK)))) void step_batch_kernel(
    const StepArgs* __restrict__ grooms, const int4* __restrict__ packs) {
    __shared__ StepLds<BLOCK> lds;
    const int4 pk = packs[blockIdx.x];
    const StepArgs a = grooms[pk.z];
    run_pack<BLOCK>(a, pk.x, pk.y, lds);
}

// Persistent form of the batched step: a grid of as many workgroups as the
// GPU holds at once, each taking the next pack from a device counter (reset
// to 0 before the launch) until the counter passes the last pack.  Packs are
// taken in memory order, and a workgroup that starts late only finds the
// queue shorter.  Packs are independent within a step, so the order in which
// they are taken changes nothing in the results.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(reg_waves(BLOCK)))) void step_batch_persistent_kernel(
    const StepArgs* __restrict__ grooms, const int4* __restrict__ packs, int num_packs, int* __restrict__ next) {
    __shared__ StepLds<BLOCK> lds;
    __shared__ int s_pack;
    for (;;) {
        if (threadIdx.x == 0) s_pack = atomicAdd(next, 1);
        __syncthreads();
        const int p = s_pack;
        __syncthreads();  // every lane has read s_pack before lane 0 takes the next one
        if (p >= num_packs) break;
        const int4 pk = packs[p];
        const StepArgs a = grooms[pk.z];
        run_pack<BLOCK>(a, pk.x, pk.y, lds);
    }
}

// After a batched step: last_h <- h and time <- time + dt for every groom
// (Sim::advance, YB/sim/step.cpp:41-43 and :59).
__global__ void batch_clock_kernel(StepArgs* grooms, int n, float dt) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    grooms[g].last_h = grooms[g].h;
    grooms[g].time = grooms[g].time + dt;
}

}  // namespace

hipError_t launch_step_fused(const StepArgs& a, const int2* packs, int num_packs, int block, hipStream_t stream) {
    if (num_packs <= 0) return hipSuccess;
    const dim3 grid(num_packs);
    switch (block) {
#define KSCR_CASE(B)                                                                      \
    case B:                                                                               \
        hipLaunchKernelGGL(step_fused_kernel<B>, grid, dim3(B), 0, stream, a, packs);     \
        break;
        KSCR_CASE(64)
        KSCR_CASE(128)
        KSCR_CASE(256)
        KSCR_CASE(512)
        KSCR_CASE(1024)
#undef KSCR_CASE
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_step_record(const StepArgs& a, const TapeArgs& tape, const int2* packs, int num_packs, int block,
                              hipStream_t stream) {
    if (num_packs <= 0) return hipSuccess;
    const dim3 grid(num_packs);
    switch (block) {
#define KSCR_CASE(B)                                                                              \
    case B:                                                                                       \
        hipLaunchKernelGGL(step_record_kernel<B>, grid, dim3(B), 0, stream, a, tape, packs);      \
        break;
        KSCR_CASE(64)
        KSCR_CASE(128)
        KSCR_CASE(256)
        KSCR_CASE(512)
        KSCR_CASE(1024)
#undef KSCR_CASE
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_step_batch(const StepArgs* grooms, const int4* packs, int num_packs, int block, hipStream_t stream) {
    if (num_packs <= 0) return hipSuccess;
    const dim3 grid(num_packs);
    switch (block) {
#define KSCR_CASE(B)                                                                       \
    case B:                                                                                \
        hipLaunchKernelGGL(step_batch_kernel<B>, grid, dim3(B), 0, stream, grooms, packs); \
        break;
        KSCR_CASE(64)
        KSCR_CASE(128)
        KSCR_CASE(256)
        KSCR_CASE(512)
        KSCR_CASE(1024)
#undef KSCR_CASE
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_step_batch_persistent(const StepArgs* grooms, const int4* packs, int num_packs, int block, int grid,
                                        int* next, hipStream_t stream) {
    if (num_packs <= 0) return hipSuccess;
    const dim3 g(std::max(1, std::min(grid, num_packs)));
    switch (block) {
#define KSCR_CASE(B)                                                                                                 \
    case B:                                                                                                          \
        hipLaunchKernelGGL(step_batch_persistent_kernel<B>, g, dim3(B), 0, stream, grooms, packs, num_packs, next);  \
        break;
        KSCR_CASE(64)
        KSCR_CASE(128)
        KSCR_CASE(256)
        KSCR_CASE(512)
        KSCR_CASE(1024)
#undef KSCR_CASE
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

int batch_persistent_grid(int block) {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    hipError_t e = hipErrorInvalidValue;
    switch (block) {
#define KSCR_CASE(B)                                                                                        \
    case B:                                                                                                 \
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, step_batch_persistent_kernel<B>, B, 0);   \
        break;
        KSCR_CASE(64)
        KSCR_CASE(128)
        KSCR_CASE(256)
        KSCR_CASE(512)
        KSCR_CASE(1024)
#undef KSCR_CASE
        default:
            break;
    }
    if (e != hipSuccess || per_cu <= 0) return 0;
    return cus * per_cu;
}

hipError_t launch_batch_clock(StepArgs* grooms, int n, float dt, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(batch_clock_kernel, dim3((n + 63) / 64), dim3(64), 0, stream, grooms, n, dt);
    return hipGetLastError();
}

}  // namespace kscr
