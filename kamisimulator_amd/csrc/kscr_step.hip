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
#include "kscr_device.h"
#include "kscr_math.h"

namespace kscr {
namespace {

__device__ inline f3 ld3(const float* p, int i) { return f3{p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
__device__ inline void st3(float* p, int i, f3 a) {
    p[3 * i] = a.x;
    p[3 * i + 1] = a.y;
    p[3 * i + 2] = a.z;
}
__device__ inline f4 to4(float4 a) { return f4{a.x, a.y, a.z, a.w}; }

// Trilinear SDF lookup and projection of one particle; mirrors
// SDFHeadCollision with its float/double index arithmetic.
__device__ inline void sdf_project(const SdfArgs& s, f3& x, f3& v) {
    const f4 invq{-s.rot[0], -s.rot[1], -s.rot[2], s.rot[3]};
    const f3 lp = rotate(invq, x - f3{s.trans[0], s.trans[1], s.trans[2]}) - f3{s.gmin[0], s.gmin[1], s.gmin[2]};
    const float l[3] = {lp.x, lp.y, lp.z};
    int idx[3];
    float ud[3];
#pragma unroll
    for (int a = 0; a < 3; a++) idx[a] = (int)((double)(l[a] * s.inv_ds[a]) + 0.5);
    if (idx[0] < 0 || idx[0] >= s.dim[0] || idx[1] < 0 || idx[1] >= s.dim[1] || idx[2] < 0 || idx[2] >= s.dim[2]) return;
#pragma unroll
    for (int a = 0; a < 3; a++)
        ud[a] = (float)((double)s.inv_ds[a] * ((double)l[a] - ((double)idx[a] - 0.5) * (double)s.ds[a]));
    // corners c000 c001 c010 c011 c100 c101 c110 c111 (bits x y z)
    f4 c[8];
#pragma unroll
    for (int o = 0; o < 8; o++) {
        const int cx = idx[0] + ((o >> 2) & 1), cy = idx[1] + ((o >> 1) & 1), cz = idx[2] + (o & 1);
        const bool in = cx < s.dim[0] && cy < s.dim[1] && cz < s.dim[2];  // cx, cy, cz >= 0 here
        const int kx = min(cx, s.dim[0] - 1), ky = min(cy, s.dim[1] - 1), kz = min(cz, s.dim[2] - 1);
        const float4 g = s.grid[(size_t)kx + (size_t)ky * s.dim[0] + (size_t)kz * s.dim[0] * s.dim[1]];
        // gradient is zero outside the grid, phi is clamped to the border
        c[o] = in ? f4{g.x, g.y, g.z, g.w} : f4{0.0f, 0.0f, 0.0f, g.w};
    }
    f4 r;
    float* rp = &r.x;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const float c000 = (&c[0].x)[k], c001 = (&c[1].x)[k], c010 = (&c[2].x)[k], c011 = (&c[3].x)[k];
        const float c100 = (&c[4].x)[k], c101 = (&c[5].x)[k], c110 = (&c[6].x)[k], c111 = (&c[7].x)[k];
        const float c00 = c000 * (1.f - ud[0]) + c100 * ud[0];
        const float c01 = c001 * (1.f - ud[0]) + c101 * ud[0];
        const float c10 = c010 * (1.f - ud[0]) + c110 * ud[0];
        const float c11 = c011 * (1.f - ud[0]) + c111 * ud[0];
        const float c0 = c00 * (1.f - ud[1]) + c10 * ud[1];
        const float c1 = c01 * (1.f - ud[1]) + c11 * ud[1];
        rp[k] = c0 * (1.f - ud[2]) + c1 * ud[2];
    }
    f3 n = vec(r);
    n = n * (1.0f / sqrtf(dot(n, n)));
    const float phi = r.w - s.thresh;
    if (phi < 0) {
        x = x - phi * n;
        const float vl = length(v);
        if ((double)vl >= 1e-6) {
            const f3 vn = v / vl;
            v = v - dot(vn, n) * n;
        }
    }
}

template <int BLOCK>
struct StepLds {
    float x[3][BLOCK];
    float dx[2][3][BLOCK];
    float4 q[2][BLOCK];
    float4 f[BLOCK];
};

// Advances the strands of one pack [vbeg, vend) through a whole step.
template <int BLOCK>
__device__ __forceinline__ void run_pack(const StepArgs& a, const int vbeg, const int vend, StepLds<BLOCK>& L) {
    float(&s_x)[3][BLOCK] = L.x;
    float(&s_dx)[2][3][BLOCK] = L.dx;
    float4(&s_q)[2][BLOCK] = L.q;
    float4(&s_f)[BLOCK] = L.f;
    const int t = threadIdx.x;
    const int i = vbeg + t;
    const bool act = i < vend;

    f3 x{0, 0, 0}, v{0, 0, 0}, u{0, 0, 0};
    f4 q{0, 0, 0, 1}, qr{0, 0, 0, 1}, qrp{0, 0, 0, 1};
    float im = 0.f, lr = 1.f, ks = 0.f;
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
        if (fl & FLAG_PREV) qrp = to4(a.q_rest[i - 1]);
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
    const f3 g{a.gravity[0], a.gravity[1], a.gravity[2]};

    for (int s = 0; s < a.num_substeps; s++) {
        // ---- initItr: inertial target y and VBD adaptive initial guess
        f3 y = v;
        f3 dx;
        {
            float windx = 0.f;
            if (a.wind_peak != 0.f) {
                const float wy = x.y > 0 ? (float)cos(2 * M_PI * (double)(a.wind_y_freq * x.y + a.wind_time_freq * a.time))
                                         : 0.f;
                const float wz = (float)cos(2 * M_PI * (double)(a.wind_z_freq * fabsf(x.z) + a.wind_time_freq * a.time));
                const float yz = wy * wz;
                const float sg = (float)((0.f < yz) - (yz < 0.f));
                windx = a.wind_peak * fmaxf(0.f, powf(yz, a.wind_sharpness) * sg);
            }
            const f3 ext = (g + f3{windx, 0.f, 0.f}) * a.external_force_multiplier;
            dx = h * v;
            const f3 lastVel = u;
            u = v;
            if (free) {
                y = dx + (h * h) * ext;
                const float ext2 = dot(ext, ext);
                if (ext2 > 0) {
                    const f3 acc = (v - lastVel) / a.last_h;
                    const float sc = fminf(fmaxf(dot(acc, ext) / ext2, 0.f), 1.f);
                    dx = dx + (h * h * sc) * ext;
                }
            }
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
            f3 fo{0, 0, 0};
            if (hasNext) {
                const f3 dxj{s_dx[cur][0][t + 1], s_dx[cur][1][t + 1], s_dx[cur][2][t + 1]};
                const f3 c = ((xj - x) + (dxj - dx)) * invl - rotate_e1(q);
                fo = kk * c - (damp * dd) * dx;
                const f3 f2 = (-kk) * c - (damp * dd) * dxj;
                s_f[t] = make_float4(f2.x, f2.y, f2.z, dd * (1 + damp));
            }
            __syncthreads();
            // ---- position solve, part 2: local 3x3 Newton step
            f3 dxn = dx;
            if (free) {
                float hd = a_in;
                f3 f = f_in * (y - dx);
                if (hasNext) {
                    f = f + fo;
                    hd = hd + dd * (1 + damp);
                }
                if (hasPrev) {
                    const float4 p = s_f[t - 1];
                    f = f + f3{p.x, p.y, p.z};
                    hd = hd + p.w;
                }
                dxn = dx + a.acceleration_ratio * sym_solve(hd, hd, hd, a_in, a_in, a_in, f);
            }
            s_dx[cur ^ 1][0][t] = dxn.x;
            s_dx[cur ^ 1][1][t] = dxn.y;
            s_dx[cur ^ 1][2][t] = dxn.z;
            __syncthreads();
            // ---- orientation solve (closed form, SCR Eq. 18)
            f4 qn = q;
            if (rotates) {
                const f3 dxj{s_dx[cur ^ 1][0][t + 1], s_dx[cur ^ 1][1][t + 1], s_dx[cur ^ 1][2][t + 1]};
                f3 vv = ((xj - x) + (dxj - dxn)) / lr;
                vv = vv * (-2.0f * ks);
                f4 b{0, 0, 0, 0};
                if (hasPrev) {
                    const f4 qj = to4(s_q[cur][t - 1]);
                    const float phi = dot(rmul(conj(qj), q), qrp) > 0 ? 1.f : -1.f;
                    b = b + phi * rmul(qj, qrp);
                }
                if (fl & FLAG_NEXT_ORI) {
                    const f4 qj = to4(s_q[cur][t + 1]);
                    const float phi = dot(rmul(conj(q), qj), qr) > 0 ? 1.f : -1.f;
                    b = b + phi * rmul(qj, conj(qr));
                }
                const float lambda = length(vv) + length(b);
                qn = normalize(rmul(rmul(mk4(vv, 0.f), b), f4{1.f, 0.f, 0.f, 0.f}) + lambda * b);
            }
            s_q[cur ^ 1][t] = make_float4(qn.x, qn.y, qn.z, qn.w);
            dx = dxn;
            q = qn;
            // The next iteration's first LDS write (s_f) is ordered after this
            // iteration's s_f reads by the barrier above; its s_dx/s_q writes
            // target the buffers this iteration only read before that barrier.
        }
        // ---- endItr
        if (free) v = dx * (1.0f / h) * (1 - a.velocity_decay * h);
        x = x + dx;
    }

    if (a.sdf_on && act && hasPrev) sdf_project(a.sdf, x, v);

    if (act) {
        st3(a.x, i, x);
        st3(a.v, i, v);
        st3(a.u, i, u);
        a.q[i] = make_float4(q.x, q.y, q.z, q.w);
    }
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void step_fused_kernel(StepArgs a, const int2* __restrict__ packs) {
    __shared__ StepLds<BLOCK> lds;
    const int2 pk = packs[blockIdx.x];
    run_pack<BLOCK>(a, pk.x, pk.y, lds);
}

// Batched form: every workgroup reads its groom's arguments (pointers,
// substep lengths, gravity, SDF grid) from a device table.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void step_batch_kernel(const StepArgs* __restrict__ grooms,
                                                           const int4* __restrict__ packs) {
    __shared__ StepLds<BLOCK> lds;
    const int4 pk = packs[blockIdx.x];
    const StepArgs a = grooms[pk.z];
    run_pack<BLOCK>(a, pk.x, pk.y, lds);
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

hipError_t launch_batch_clock(StepArgs* grooms, int n, float dt, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(batch_clock_kernel, dim3((n + 63) / 64), dim3(64), 0, stream, grooms, n, dt);
    return hipGetLastError();
}

}  // namespace kscr
