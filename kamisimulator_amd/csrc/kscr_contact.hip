// kscr_contact.hip -- Stable Cosserat Rods step with hair-hair contacts.
//
// Per substep (YB/sim/step.cpp:49-57 and the captured graph of :17-30):
//   substep_begin_kernel       endItr of the previous substep (iteration.cu:62-74),
//                              initItr (iteration.cu:8-59) and, when
//                              detection is due, buildAABBs (collision.cu:7-34)
//   detection (when due)       LBVH, query,                (YBSimulator.cu:424-457,
//                              buildCollisionList          collision.cu:36-92)
//   step_limit_kernel          recomputeStepLimitKernel    (collision.cu:94-131)
//   numItr x iterate_kernel    cosseratItr with IPC barrier and friction,
//                              then quaternionLambdaItr    (cosserat.cu:25-337)
// and end_kernel (endItr) after the last substep.
// The reference's races are replaced by ping-pong buffers (Jacobi snapshot)
// and its atomically ordered contact lists by ascending lists, so results are
// deterministic and equal to the CPU restatement's.
#include "kscr_contact.h"
#include "kscr_math.h"
#include "kscr_solve.h"
#include "kscr_sync.h"

namespace kscr {
namespace {

constexpr int kThreads = 256;

inline dim3 grid_for(int n) { return dim3((unsigned)((n + kThreads - 1) / kThreads)); }

// buildAABBs<USE_VEL_RADIUS> (collision.cu:7-34): padded box of the segment
// from p0 to p1 with initial guesses d0, d1 (no segment: the empty box).
__device__ inline void store_segment_box(const ContactParams& cp, bool has_segment, f3 p0, f3 p1, f3 d0, f3 d1,
                                         float* __restrict__ out) {
    float b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    if (has_segment) {
        const float P0[3] = {p0.x, p0.y, p0.z}, P1[3] = {p1.x, p1.y, p1.z};
        for (int k = 0; k < 3; k++) {
            b[k] = gmin(gmin(b[k], P0[k]), P1[k]);
            b[3 + k] = gmax(gmax(b[3 + k], P0[k]), P1[k]);
        }
        float pad = cp.scaled_detection_radius;
        if (cp.use_velocity_radius) pad += sqrtf(gmax(dot(d0, d0), dot(d1, d1)));
        for (int k = 0; k < 3; k++) {
            b[k] -= pad;
            b[3 + k] += pad;
        }
    }
    for (int k = 0; k < 6; k++) out[k] = b[k];
}

// The start of a substep, one lane per vertex: endItr of the previous
// substep (END; iteration.cu:62-74, from its last iterate dx_prev), initItr
// (iteration.cu:8-59) into dx0 and y, and with BOX the padded segment boxes
// of buildAABBs.  A segment box needs the next vertex's new position and
// initial guess: they come through LDS from the next lane; the last lane of
// a workgroup leaves a segment that continues into the next workgroup to
// box_fixup_kernel, which runs after this kernel (reading another
// workgroup's vertex here could see it before or after its update).
// dx_prev may alias dx0: each lane reads its own vertex before writing it.
template <bool END, bool BOX>
__global__ __launch_bounds__(kThreads) void substep_begin_kernel(StepArgs a, ContactParams cp, int n,
                                                                 const float* dx_prev, float* dx0,
                                                                 float* __restrict__ yout, float* __restrict__ box) {
    __shared__ float s_x[3][BOX ? kThreads : 1];
    __shared__ float s_d[3][BOX ? kThreads : 1];
    const int t = threadIdx.x;
    const int i = blockIdx.x * kThreads + t;
    const bool act = i < n;
    f3 x{0, 0, 0}, dx{0, 0, 0};
    uint32_t fl = 0;
    if (act) {
        x = ld3(a.x, i);
        f3 v = ld3(a.v, i);
        const float4 m = a.mat[i];
        fl = __float_as_uint(m.w);
        const bool free = m.x != 0.f;
        if (END) {
            end_vertex(a, free, ld3(dx_prev, i), x, v);
            st3(a.x, i, x);
            st3(a.v, i, v);
        }
        f3 u = ld3(a.u, i), y;
        init_vertex(a, free, x, v, u, y, dx);
        st3(a.u, i, u);
        st3(dx0, i, dx);
        st3(yout, i, y);
    }
    if (BOX) {
        s_x[0][t] = x.x;
        s_x[1][t] = x.y;
        s_x[2][t] = x.z;
        s_d[0][t] = dx.x;
        s_d[1][t] = dx.y;
        s_d[2][t] = dx.z;
        lds_barrier();  // the state stores above are not read here
        const bool seg = (fl & FLAG_NEXT) != 0;
        if (act && (!seg || t + 1 < kThreads)) {
            const int tn = seg ? t + 1 : t;
            const f3 xn{s_x[0][tn], s_x[1][tn], s_x[2][tn]}, dn{s_d[0][tn], s_d[1][tn], s_d[2][tn]};
            store_segment_box(cp, seg, x, xn, dx, dn, box + 6 * (size_t)i);
        }
    }
}

// The segment boxes substep_begin_kernel leaves: segments from the last
// vertex of a workgroup to the first of the next.
__global__ void box_fixup_kernel(StepArgs a, ContactParams cp, int n, const float* __restrict__ dx0,
                                 float* __restrict__ box) {
    const int i = (blockIdx.x * blockDim.x + threadIdx.x) * kThreads + (kThreads - 1);
    if (i >= n) return;
    if (!(__float_as_uint(a.mat[i].w) & FLAG_NEXT)) return;
    store_segment_box(cp, true, ld3(a.x, i), ld3(a.x, i + 1), ld3(dx0, i), ld3(dx0, i + 1), box + 6 * (size_t)i);
}

// buildCollisionList<USE_VEL_RADIUS> (collision.cu:36-92) over a device-sized
// pair list; grid-stride so no host read of the pair count is needed.
__global__ void narrow_kernel(StepArgs a, ContactParams cp, const int2* __restrict__ pairs, const int* __restrict__ count,
                              int cap, const float* __restrict__ dx, int* __restrict__ num_cols, int* __restrict__ cols,
                              int* __restrict__ error) {
    const int total = *count;
    if (total > cap && blockIdx.x == 0 && threadIdx.x == 0) atomicOr(error, DEV_ERR_PAIR_OVERFLOW);
    const int m = total < cap ? total : cap;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < m; k += gridDim.x * blockDim.x) {
        const int2 ids = pairs[k];
        const int ia = ids.x, ib = ids.y;
        if (a.conn) {  // glued segments (collision.cu:44-49)
            const int c0 = a.conn[ia], c1 = a.conn[ia + 1];
            if (c0 == ib || c1 == ib || c0 == ib + 1 || c1 == ib + 1) continue;
        }
        if (abs(ib - ia) <= 2) continue;  // neighbouring segments
        const f3 a0 = ld3(a.x, ia);
        const f3 a1 = ld3(a.x, ia + 1) - a0;
        const f3 b0 = ld3(a.x, ib) - a0;
        const f3 b1 = ld3(a.x, ib + 1) - a0;
        float u, v;
        segment_closest(f3{0, 0, 0}, a1, b0, b1, u, v);
        const f3 nrm = u * a1 - mix3(b0, b1, v);
        const float d2 = dot(nrm, nrm);
        float r = 2 * cp.scaled_detection_radius;
        if (cp.use_velocity_radius) {
            const f3 da0 = ld3(dx, ia), da1 = ld3(dx, ia + 1), db0 = ld3(dx, ib), db1 = ld3(dx, ib + 1);
            r += sqrtf(gmax(dot(da0, da0), dot(da1, da1)));
            r += sqrtf(gmax(dot(db0, db0), dot(db1, db1)));
        }
        if (d2 < r * r) {
            int slot = atomicAdd(&num_cols[ia], 1);
            if (slot >= kMaxCols)
                atomicOr(error, DEV_ERR_MAX_COLLISIONS);
            else
                cols[kMaxCols * (size_t)ia + slot] = ib;
            slot = atomicAdd(&num_cols[ib], 1);
            if (slot >= kMaxCols)
                atomicOr(error, DEV_ERR_MAX_COLLISIONS);
            else
                cols[kMaxCols * (size_t)ib + slot] = ia;
        }
    }
}

// ascending contact list per segment (deterministic summation order)
__global__ void sort_lists_kernel(int n, int* __restrict__ num_cols, int* __restrict__ cols) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int m = num_cols[i];
    if (m > kMaxCols) {
        m = kMaxCols;
        num_cols[i] = m;
    }
    int v[kMaxCols];
    for (int k = 0; k < kMaxCols; k++) v[k] = k < m ? cols[kMaxCols * (size_t)i + k] : 0;
    for (int k = 1; k < m; k++) {
        const int key = v[k];
        int j = k - 1;
        while (j >= 0 && v[j] > key) {
            v[j + 1] = v[j];
            j--;
        }
        v[j + 1] = key;
    }
    for (int k = 0; k < m; k++) cols[kMaxCols * (size_t)i + k] = v[k];
}

// recomputeStepLimitKernel<LIMIT, USE_VEL_RADIUS> (collision.cu:94-131)
__global__ void step_limit_kernel(StepArgs a, ContactParams cp, int n, const float* __restrict__ dx,
                                  const int* __restrict__ num_cols, const int* __restrict__ cols,
                                  float* __restrict__ max_step) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float m = INFINITY;
    const uint32_t fl = __float_as_uint(a.mat[i].w);
    if (cp.use_step_size_limit && (fl & FLAG_NEXT)) {
        const f3 a0 = ld3(a.x, i);
        const f3 a1 = ld3(a.x, i + 1) - a0;
        m = cp.detection_radius * (cp.detection_scaler - 1);
        if (cp.use_velocity_radius) m += length(ld3(dx, i));
        const float margin = (1 - 0.1f) * 0.5f;
        const int nc = num_cols[i];
        for (int k = 0; k < nc; k++) {
            const int c = cols[kMaxCols * (size_t)i + k];
            const f3 b0 = ld3(a.x, c) - a0;
            const f3 b1 = ld3(a.x, c + 1) - a0;
            float u, v;
            segment_closest(f3{0, 0, 0}, a1, b0, b1, u, v);
            const f3 nrm = u * a1 - mix3(b0, b1, v);
            m = gmin(m, margin * length(nrm));
        }
    }
    max_step[i] = m;
}

// One contact of segment i with segment c (cosserat.cu:171-227).
__device__ inline void contact_terms(const StepArgs& a, const ContactParams& cp, const float* __restrict__ dx, int i,
                                     int c, float damp, f3& f, s3& H, f3& f2, s3& H2) {
    const float invb = 1.0f / cp.barrier;
    const float radius = 2.0f * cp.radius;
    const float kCol = cp.k_collision * invb;
    const f3 xi = ld3(a.x, i), xj = ld3(a.x, i + 1), dxi = ld3(dx, i), dxj = ld3(dx, i + 1);
    const f3 b0 = ld3(a.x, c), b1 = ld3(a.x, c + 1), db0 = ld3(dx, c), db1 = ld3(dx, c + 1);
    float ux, uy;
    segment_closest(f3{0, 0, 0}, (xj - xi) + (dxj - dxi), (b0 - xi) + (db0 - dxi), (b1 - xi) + (db1 - dxi), ux, uy);
    const f3 dpos = mix3(xi, xj, ux) - mix3(b0, b1, uy);
    const f3 ddpos = mix3(dxi, dxj, ux) - mix3(db0, db1, uy);
    f3 nrm = dpos + ddpos;
    float d = length(nrm);
    if (!(d > 0.0f)) return;  // closest points coincide: no contact normal (the reference divides by 0)
    nrm = nrm * (1.0f / d);
    const float wj = ux;
    const float wi = 1.0f - ux;
    d = d - radius;
    d = d * invb;
    if (d > 1) return;
    d = gmax(d, 1e-3f);
    const float invd = 1.0f / d;
    const float logd = det_logf(d);
    float dH = (-3.0f + (2.0f + invd) * invd - 2.0f * logd) * kCol * invb;
    const float ff = -(1.0f - d) * (d - 1.0f + 2.0f * d * logd) * invd * kCol;
    f = f + (ff * wi - damp * dH * wi * wi * dot(nrm, dxi)) * nrm;
    f2 = f2 + (ff * wj - damp * dH * wj * wj * dot(nrm, dxj)) * nrm;
    dH = dH * (1.0f + damp);
    s3 op = souter(nrm);
    sadd(H, sscale(op, dH * wi * wi));
    sadd(H2, sscale(op, dH * wj * wj));
    const f3 u = ddpos - dot(nrm, ddpos) * nrm;
    const float ul = length(u);
    if (ul > 0) {
        const float cand = cp.friction_coeff * ff / ul;
        const float f1 = cand < cp.k_friction ? cand : cp.k_friction;
        op.d[0] -= 1;
        op.d[1] -= 1;
        op.d[2] -= 1;
        f = f - (f1 * wi) * u;
        ssub(H, sscale(op, (wi * wi) * f1));
        f2 = f2 - (f1 * wj) * u;
        ssub(H2, sscale(op, (wj * wj) * f1));
    }
}

struct SegSums {
    f3 f0, f1, f2a, f2b;
    s3 H0, H1, H2a, H2b;
    float limit;
};

// The stretch terms of segment (i, i+1) (cosserat.cu:141-162): added to
// vertex i's sums (f0, H0) and to vertex i+1's share (f2a, H2a).
__device__ inline void stretch_sums(const float4 mi, f3 xi, f3 xj, f3 dxi, f3 dxj, f4 qi, float damp, SegSums& s) {
    const float invl = 1.0f / mi.y;
    const f3 c = ((xj - xi) + (dxj - dxi)) * invl - rotate_e1(qi);
    const float k = mi.z * invl;
    float d = k * invl;
    s.f0 = s.f0 + (k * c - (damp * d) * dxi);
    s.f2a = s.f2a + ((-k) * c - (damp * d) * dxj);
    d = d * (1 + damp);
    s.H0.d[0] += d;
    s.H0.d[1] += d;
    s.H0.d[2] += d;
    s.H2a.d[0] += d;
    s.H2a.d[1] += d;
    s.H2a.d[2] += d;
}

// Everything segment i adds, by sector (sector 0: stretch + even-position
// contacts on top of vertex i's inertia terms; sector 1: odd positions).
__device__ inline void segment_sums(const StepArgs& a, const ContactParams& cp, const float* __restrict__ dx,
                                    const float4* __restrict__ q, const int* __restrict__ num_cols,
                                    const int* __restrict__ cols, const float* __restrict__ max_step, int i, float damp,
                                    SegSums& s) {
    s.f1 = f3{0, 0, 0};
    s.f2a = f3{0, 0, 0};
    s.f2b = f3{0, 0, 0};
    s.H1 = s3{{0, 0, 0, 0, 0, 0}};
    s.H2a = s3{{0, 0, 0, 0, 0, 0}};
    s.H2b = s3{{0, 0, 0, 0, 0, 0}};
    s.limit = INFINITY;
    const float4 mi = a.mat[i];
    if (!(__float_as_uint(mi.w) & FLAG_NEXT)) return;
    s.limit = max_step[i];
    stretch_sums(mi, ld3(a.x, i), ld3(a.x, i + 1), ld3(dx, i), ld3(dx, i + 1), to4(q[i]), damp, s);
    const int nc = num_cols[i];
    const int* L = cols + kMaxCols * (size_t)i;
    for (int k2 = 0; k2 < nc; k2 += 2) contact_terms(a, cp, dx, i, L[k2], damp, s.f0, s.H0, s.f2a, s.H2a);
    for (int k2 = 1; k2 < nc; k2 += 2) contact_terms(a, cp, dx, i, L[k2], damp, s.f1, s.H1, s.f2b, s.H2b);
}

// One VBD iteration for a pack of whole strands, one lane per vertex:
// cosseratItr with contacts (the position solve), then quaternionLambdaItr
// (the orientation solve) on the new positions of the same iteration.  The
// stretch and contact sums of segment (i, i+1) are formed once, by lane i;
// vertex i+1's share reaches lane i+1 through LDS and is added there in the
// order the sequential sum adds it (own segment, then the previous one).
// The previous iterate (dx, q) is read from global memory: contact partners
// lie in other workgroups (Jacobi snapshot).  Only the packs that hold a
// contact in this substep take part (contact_pack); a partner's pack holds
// the contact too, so no lane reads the iterate of a pack left out.
// A pack cut out of a long strand starts with a halo vertex (hasPrev at lane
// 0) and/or ends with one (hasNext at the last vertex).  The first halo's
// lane forms its segment's sums, whose share its right neighbour needs; the
// last halo's lane runs its position solve, whose new dx the last owned
// vertex's orientation solve needs.  Both repeat, operation for operation,
// what the owning pack computes for those vertices; only owned vertices are
// written.
// amdgpu_waves_per_eu(5): 96 VGPRs (3 values spilled to scratch, reloaded
// in the loop over a segment's contacts) instead of 103, 5 waves per SIMD
// instead of 4; with 256-thread
// workgroups that is one more workgroup per CU (100k-strand contact groom:
// 37.2 -> 35.9 ms/step; a cap at 6 waves spills 18 and measures 36.9).
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(5))) void iterate_kernel(StepArgs a, ContactParams cp, const int2* __restrict__ packs,
                                                        const float* __restrict__ dx, const float4* __restrict__ q_in,
                                                        const float* __restrict__ y, const int* __restrict__ num_cols,
                                                        const int* __restrict__ cols,
                                                        const float* __restrict__ max_step,
                                                        const uint8_t* __restrict__ contact_pack,
                                                        float* __restrict__ dx_out, float4* __restrict__ q_out) {
    __shared__ float s_f2[3][BLOCK];
    __shared__ float s_H2[6][BLOCK];
    __shared__ float s_lim[BLOCK];
    __shared__ float s_dxn[3][BLOCK];
    if (!contact_pack[blockIdx.x]) return;  // a pack without contacts is free_iterate_kernel's
    const int2 pk = packs[blockIdx.x];
    const int t = threadIdx.x;
    const int i = pk.x + t;
    const bool act = i < pk.y;
    const float4 mi = act ? a.mat[i] : make_float4(0.f, 1.f, 0.f, 0.f);
    const uint32_t fl = act ? __float_as_uint(mi.w) : 0u;
    const bool halo_before = t == 0 && (fl & FLAG_PREV);
    const bool own = act && !halo_before && !(i == pk.y - 1 && (fl & FLAG_NEXT));
    const bool free = act && mi.x != 0.f;
    const f3 dxi = act ? ld3(dx, i) : f3{0, 0, 0};
    const float h = a.h;
    const float damp = a.damping / h;
    SegSums s;
    s.H0 = s3{{0, 0, 0, 0, 0, 0}};
    s.f0 = f3{0, 0, 0};
    if (free) {  // the vertex's inertia terms open its sum
        const float a_in = 1.0f / (mi.x * h * h);
        s.H0 = s3{{a_in, a_in, a_in, a_in, a_in, a_in}};
        s.f0 = (1.0f / (h * h * mi.x)) * (ld3(y, i) - dxi);
        const int g = a.conn ? a.conn[i] : -1;
        if (g >= 0) {  // the glue to the connected vertex (cosserat.cu:123-130)
            const float stiffness = 40.0f;
            s.f0 = s.f0 - stiffness * (((ld3(a.x, i) - ld3(a.x, g)) + (dxi - ld3(dx, g))) + damp * dxi);
            const float hg = stiffness * (1 + damp);
            s.H0.d[0] += hg;
            s.H0.d[1] += hg;
            s.H0.d[2] += hg;
        }
    }
    if (act) {
        segment_sums(a, cp, dx, q_in, num_cols, cols, max_step, i, damp, s);
    } else {
        s.f1 = s.f2a = s.f2b = f3{0, 0, 0};
        s.H1 = s.H2a = s.H2b = s3{{0, 0, 0, 0, 0, 0}};
        s.limit = INFINITY;
    }
    {
        const f3 f2 = s.f2a + s.f2b;
        s3 H2 = s.H2a;
        sadd(H2, s.H2b);
        s_f2[0][t] = f2.x;
        s_f2[1][t] = f2.y;
        s_f2[2][t] = f2.z;
        for (int c = 0; c < 6; c++) s_H2[c][t] = H2.d[c];
        s_lim[t] = s.limit;
    }
    // the kernel reads none of its global writes: its barriers order LDS only
    lds_barrier();
    f3 out = dxi;  // a pinned vertex keeps its iterate
    if (free) {
        f3 f = s.f0 + s.f1;
        s3 H = s.H0;
        sadd(H, s.H1);
        float limit = s.limit;
        if ((fl & FLAG_PREV) && !halo_before) {
            const f3 f2{s_f2[0][t - 1], s_f2[1][t - 1], s_f2[2][t - 1]};
            s3 H2;
            for (int c = 0; c < 6; c++) H2.d[c] = s_H2[c][t - 1];
            limit = gmin(limit, s_lim[t - 1]);
            f = f + f2;
            sadd(H, H2);
        }
        out = dxi + a.acceleration_ratio * sym_solve(H.d[0], H.d[1], H.d[2], H.d[3], H.d[4], H.d[5], f);
        // A contact Hessian can make H numerically singular in FP32 (the
        // adjugate determinant cancels to 0); the reference's step is then
        // inf/NaN, which no step limit removes.  Such a step is skipped.
        if (!isfinite(out.x) || !isfinite(out.y) || !isfinite(out.z)) {
            out = dxi;
        } else {
            const float l = length(out);
            if (l > limit && l > 0) out = out * (limit / l);
        }
    }
    if (own) st3(dx_out, i, out);
    s_dxn[0][t] = out.x;
    s_dxn[1][t] = out.y;
    s_dxn[2][t] = out.z;
    lds_barrier();
    if (!own) return;
    const f4 q = to4(q_in[i]);
    f4 qn = q;
    if ((fl & FLAG_NEXT) && !(fl & FLAG_FIX_ORI)) {
        const f4 qm = (fl & FLAG_PREV) ? to4(q_in[i - 1]) : f4{0, 0, 0, 1};
        const f4 qp = (fl & FLAG_NEXT_ORI) ? to4(q_in[i + 1]) : f4{0, 0, 0, 1};
        const f4 qrp = (fl & FLAG_PREV) ? to4(a.q_rest[i - 1]) : f4{0, 0, 0, 1};
        const f3 dxj{s_dxn[0][t + 1], s_dxn[1][t + 1], s_dxn[2][t + 1]};
        qn = orientation_update(ld3(a.x, i), ld3(a.x, i + 1), out, dxj, mi.y, 1.0f / mi.y, mi.z, q, qm, qp, qrp,
                                to4(a.q_rest[i]), fl);
    }
    q_out[i] = make_float4(qn.x, qn.y, qn.z, qn.w);
}

// contact_pack[p] = 1 when any segment of pack p has a contact in the current
// lists (the lists are symmetric: both packs of a contact are marked).
// A pack holding a glued vertex is marked too: the glue partner's iterate is
// read from global memory (connections are symmetric, so the partner's pack
// is marked as well).  So is every pack cut out of a long strand (it starts
// with hasPrev or ends with hasNext): its neighbours across a cut lie in
// another pack, which is cut too.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void classify_kernel(const int2* __restrict__ packs,
                                                         const float4* __restrict__ mat,
                                                         const int* __restrict__ num_cols,
                                                         const int* __restrict__ conn,
                                                         uint8_t* __restrict__ contact_pack) {
    const int2 pk = packs[blockIdx.x];
    const int i = pk.x + (int)threadIdx.x;
    bool mark = false;
    if (i < pk.y) {
        const uint32_t fl = __float_as_uint(mat[i].w);
        const bool cut = (i == pk.x && (fl & FLAG_PREV)) || (i == pk.y - 1 && (fl & FLAG_NEXT));
        mark = cut || num_cols[i] > 0 || (conn && conn[i] >= 0);
    }
    const int any = __syncthreads_or(mark);
    if (threadIdx.x == 0) contact_pack[blockIdx.x] = any ? 1 : 0;
}

// All numItr iterations of the packs without contacts in one launch: the
// arithmetic of iterate_kernel for segments with no contacts (stretch terms,
// the same sums and 3x3 solve, the step limit, the orientation solve), with
// the previous iterate of a vertex's neighbours taken from LDS instead of
// global memory, since no other pack enters.  dx0 / q_in hold the substep's
// initial iterate, dx_fin / q_fin receive the last one (they may be the same
// buffers: each lane reads its own vertex before any lane writes).
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void free_iterate_kernel(StepArgs a, const int2* __restrict__ packs,
                                                             const uint8_t* __restrict__ contact_pack,
                                                             const float* dx0, const float4* q_in,
                                                             const float* __restrict__ y,
                                                             const float* __restrict__ max_step, float* dx_fin,
                                                             float4* q_fin) {
    // Without contacts the share a segment passes to its upper vertex has
    // H2 = (h, h, h, +0, +0, +0): one float crosses LDS, and the previous
    // vertex's step limit is loaded once (72 B of LDS per lane: 8 waves per
    // SIMD at 256 threads).
    __shared__ float s_f2[3][BLOCK];
    __shared__ float s_h2[BLOCK];
    __shared__ float s_dx[2][3][BLOCK];
    __shared__ float4 s_q[2][BLOCK];
    if (contact_pack[blockIdx.x]) return;  // iterate_kernel's, one launch per iteration
    const int2 pk = packs[blockIdx.x];
    const int t = threadIdx.x;
    const int i = pk.x + t;
    const bool act = i < pk.y;
    const int tm = t > 0 ? t - 1 : 0;
    const int tp1 = t < BLOCK - 1 ? t + 1 : t;
    const float4 mi = act ? a.mat[i] : make_float4(0.f, 1.f, 0.f, 0.f);
    const uint32_t fl = act ? __float_as_uint(mi.w) : 0u;
    const bool free = act && mi.x != 0.f;
    const bool next = (fl & FLAG_NEXT) != 0;
    const f3 xi = act ? ld3(a.x, i) : f3{0, 0, 0};
    const f3 xj = next ? ld3(a.x, i + 1) : xi;
    const f3 yi = free ? ld3(y, i) : f3{0, 0, 0};
    const f4 qr = act ? to4(a.q_rest[i]) : f4{0, 0, 0, 1};
    const f4 qrp = (fl & FLAG_PREV) ? to4(a.q_rest[i - 1]) : f4{0, 0, 0, 1};
    const float lim = next ? max_step[i] : INFINITY;
    const float lim_prev = (fl & FLAG_PREV) ? max_step[i - 1] : INFINITY;  // vertex i-1 has a segment
    const float h = a.h;
    const float damp = a.damping / h;
    const float a_in = free ? 1.0f / (mi.x * h * h) : 0.0f;
    const float f_in = free ? 1.0f / (h * h * mi.x) : 0.0f;
    f3 dx = act ? ld3(dx0, i) : f3{0, 0, 0};
    f4 q = act ? to4(q_in[i]) : f4{0, 0, 0, 1};
    s_dx[0][0][t] = dx.x;
    s_dx[0][1][t] = dx.y;
    s_dx[0][2][t] = dx.z;
    s_q[0][t] = make_float4(q.x, q.y, q.z, q.w);
    __syncthreads();
    for (int k = 0; k < a.num_itr; k++) {
        const int cur = k & 1;
        SegSums s;
        s.H0 = s3{{0, 0, 0, 0, 0, 0}};
        s.f0 = f3{0, 0, 0};
        if (free) {
            s.H0 = s3{{a_in, a_in, a_in, a_in, a_in, a_in}};
            s.f0 = f_in * (yi - dx);
        }
        s.f1 = s.f2a = s.f2b = f3{0, 0, 0};
        s.H1 = s.H2a = s.H2b = s3{{0, 0, 0, 0, 0, 0}};
        s.limit = lim;
        if (next) {
            const f3 dxj{s_dx[cur][0][t + 1], s_dx[cur][1][t + 1], s_dx[cur][2][t + 1]};
            stretch_sums(mi, xi, xj, dx, dxj, q, damp, s);
        }
        {
            const f3 f2 = s.f2a + s.f2b;
            s3 H2 = s.H2a;
            sadd(H2, s.H2b);
            s_f2[0][t] = f2.x;
            s_f2[1][t] = f2.y;
            s_f2[2][t] = f2.z;
            s_h2[t] = H2.d[0];  // == d[1] == d[2]; d[3..5] == +0
        }
        __syncthreads();
        f3 out = dx;
        if (free) {
            f3 f = s.f0 + s.f1;
            s3 H = s.H0;
            sadd(H, s.H1);
            float limit = s.limit;
            if (fl & FLAG_PREV) {
                const f3 f2{s_f2[0][t - 1], s_f2[1][t - 1], s_f2[2][t - 1]};
                const float h2 = s_h2[t - 1];
                const s3 H2{{h2, h2, h2, 0.0f, 0.0f, 0.0f}};
                limit = gmin(limit, lim_prev);
                f = f + f2;
                sadd(H, H2);
            }
            out = dx + a.acceleration_ratio * sym_solve(H.d[0], H.d[1], H.d[2], H.d[3], H.d[4], H.d[5], f);
            if (!isfinite(out.x) || !isfinite(out.y) || !isfinite(out.z)) {
                out = dx;
            } else {
                const float l = length(out);
                if (l > limit && l > 0) out = out * (limit / l);
            }
        }
        s_dx[cur ^ 1][0][t] = out.x;
        s_dx[cur ^ 1][1][t] = out.y;
        s_dx[cur ^ 1][2][t] = out.z;
        __syncthreads();
        f4 qn = q;
        if (next && !(fl & FLAG_FIX_ORI)) {
            const f4 qm = to4(s_q[cur][tm]);   // used only with FLAG_PREV
            const f4 qp = to4(s_q[cur][tp1]);  // used only with FLAG_NEXT_ORI
            const f3 dxj{s_dx[cur ^ 1][0][t + 1], s_dx[cur ^ 1][1][t + 1], s_dx[cur ^ 1][2][t + 1]};
            qn = orientation_update(xi, xj, out, dxj, mi.y, 1.0f / mi.y, mi.z, q, qm, qp, qrp, qr, fl);
        }
        s_q[cur ^ 1][t] = make_float4(qn.x, qn.y, qn.z, qn.w);
        dx = out;
        q = qn;
    }
    if (act) {
        st3(dx_fin, i, dx);
        q_fin[i] = make_float4(q.x, q.y, q.z, q.w);
    }
}

__global__ void end_kernel(StepArgs a, int n, const float* __restrict__ dx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    f3 x = ld3(a.x, i), v = ld3(a.v, i);
    end_vertex(a, a.mat[i].x != 0.f, ld3(dx, i), x, v);
    st3(a.x, i, x);
    st3(a.v, i, v);
}

__global__ void sdf_kernel(StepArgs a, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (!(__float_as_uint(a.mat[i].w) & FLAG_PREV)) return;  // strand root
    f3 x = ld3(a.x, i), v = ld3(a.v, i);
    sdf_project(a.sdf, x, v);
    st3(a.x, i, x);
    st3(a.v, i, v);
}

template <class T>
hipError_t dalloc(T*& p, size_t count) {
    return hipMalloc((void**)&p, count * sizeof(T) + 64);
}

}  // namespace

void contact_release(ContactDev& c) {
    void* ps[] = {c.dx[0], c.dx[1], c.y, c.q2, c.max_step, c.num_cols, c.cols, c.pairs, c.error, c.contact_pack};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    lbvh_release(c.bvh);
    c = ContactDev{};
}

hipError_t contact_reserve(ContactDev& c, int n) {
    if (c.n == n && c.dx[0]) return hipSuccess;
    contact_release(c);
    hipError_t e;
    const size_t n3 = (size_t)n * 3;
    // Broad-phase pair buffer of numVerts * MAX_COLLISIONS_PER_SEGMENT pairs,
    // the size YarnBall gives it (YBSimulator.cu:449).  The reference truncates
    // silently beyond it; here an overflow is flagged (KSCR_ERR_MAX_COLLISIONS).
    if (n > INT32_MAX / kMaxCols) return hipErrorInvalidValue;
    c.pair_cap = kMaxCols * n;
    if ((e = dalloc(c.dx[0], n3)) != hipSuccess || (e = dalloc(c.dx[1], n3)) != hipSuccess ||
        (e = dalloc(c.y, n3)) != hipSuccess || (e = dalloc(c.q2, n)) != hipSuccess ||
        (e = dalloc(c.max_step, n)) != hipSuccess || (e = dalloc(c.num_cols, n)) != hipSuccess ||
        (e = dalloc(c.cols, (size_t)kMaxCols * n)) != hipSuccess || (e = dalloc(c.pairs, (size_t)c.pair_cap)) != hipSuccess ||
        (e = dalloc(c.error, 1)) != hipSuccess || (e = dalloc(c.contact_pack, n)) != hipSuccess ||
        (e = lbvh_reserve(c.bvh, n)) != hipSuccess) {
        contact_release(c);
        return e;
    }
    c.n = n;
    if ((e = fill_i32(c.num_cols, 0, n, nullptr)) != hipSuccess) return e;
    if ((e = fill_i32(c.error, 0, 1, nullptr)) != hipSuccess) return e;
    return hipDeviceSynchronize();
}

hipError_t contact_begin(const StepArgs& a, const ContactParams& cp, ContactDev& c, int prev, bool boxes, hipStream_t s) {
    const dim3 grid = grid_for(c.n), block(kThreads);
    float* const dx_prev = prev >= 0 ? c.dx[prev] : nullptr;
    if (prev < 0 && !boxes)
        hipLaunchKernelGGL((substep_begin_kernel<false, false>), grid, block, 0, s, a, cp, c.n, dx_prev, c.dx[0], c.y,
                           c.bvh.aabb);
    else if (prev < 0)
        hipLaunchKernelGGL((substep_begin_kernel<false, true>), grid, block, 0, s, a, cp, c.n, dx_prev, c.dx[0], c.y,
                           c.bvh.aabb);
    else if (!boxes)
        hipLaunchKernelGGL((substep_begin_kernel<true, false>), grid, block, 0, s, a, cp, c.n, dx_prev, c.dx[0], c.y,
                           c.bvh.aabb);
    else
        hipLaunchKernelGGL((substep_begin_kernel<true, true>), grid, block, 0, s, a, cp, c.n, dx_prev, c.dx[0], c.y,
                           c.bvh.aabb);
    if (boxes) {
        const int seams = (c.n + kThreads - 1) / kThreads;
        hipLaunchKernelGGL(box_fixup_kernel, dim3((unsigned)((seams + 255) / 256)), dim3(256), 0, s, a, cp, c.n,
                           c.dx[0], c.bvh.aabb);
    }
    return hipGetLastError();
}

hipError_t contact_detect(const StepArgs& a, const ContactParams& cp, ContactDev& c, bool rebuild, hipStream_t s) {
    hipError_t e = rebuild ? lbvh_build(c.bvh, c.n, s) : lbvh_refit(c.bvh, s);
    if (e != hipSuccess) return e;
    if ((e = lbvh_query(c.bvh, c.pairs, c.pair_cap, 2, c.error, s)) != hipSuccess) return e;
    if ((e = fill_i32(c.num_cols, 0, c.n, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(narrow_kernel, dim3(1024), dim3(kThreads), 0, s, a, cp, c.pairs, c.bvh.count, c.pair_cap, c.dx[0],
                       c.num_cols, c.cols, c.error);
    hipLaunchKernelGGL(sort_lists_kernel, grid_for(c.n), dim3(kThreads), 0, s, c.n, c.num_cols, c.cols);
    return hipGetLastError();
}

hipError_t contact_step_limit(const StepArgs& a, const ContactParams& cp, ContactDev& c, hipStream_t s) {
    hipLaunchKernelGGL(step_limit_kernel, grid_for(c.n), dim3(kThreads), 0, s, a, cp, c.n, c.dx[0], c.num_cols, c.cols,
                       c.max_step);
    return hipGetLastError();
}

hipError_t contact_iterate(const StepArgs& a, const ContactParams& cp, ContactDev& c, int cur, const float4* q_in,
                           float4* q_out, const int2* packs, int num_packs, int block, hipStream_t s) {
    if (num_packs <= 0) return hipSuccess;
    const dim3 grid(num_packs);
    switch (block) {
#define KSCR_CASE(B)                                                                                              \
    case B:                                                                                                       \
        hipLaunchKernelGGL(iterate_kernel<B>, grid, dim3(B), 0, s, a, cp, packs, c.dx[cur], q_in, c.y, c.num_cols, \
                           c.cols, c.max_step, c.contact_pack, c.dx[cur ^ 1], q_out);                         \
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

hipError_t contact_classify(const StepArgs& a, ContactDev& c, const int2* packs, int num_packs, int block,
                            hipStream_t s) {
    if (num_packs <= 0) return hipSuccess;
    const dim3 grid(num_packs);
    switch (block) {
#define KSCR_CASE(B)                                                                                       \
    case B:                                                                                                \
        hipLaunchKernelGGL(classify_kernel<B>, grid, dim3(B), 0, s, packs, a.mat, c.num_cols, a.conn, c.contact_pack); \
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

hipError_t contact_free_iterate(const StepArgs& a, ContactDev& c, const float4* q_in, float4* q_fin, const int2* packs,
                                int num_packs, int block, hipStream_t s) {
    if (num_packs <= 0 || a.num_itr <= 0) return hipSuccess;
    const dim3 grid(num_packs);
    float* dx_fin = c.dx[a.num_itr & 1];
    switch (block) {
#define KSCR_CASE(B)                                                                                             \
    case B:                                                                                                      \
        hipLaunchKernelGGL(free_iterate_kernel<B>, grid, dim3(B), 0, s, a, packs, c.contact_pack, c.dx[0], q_in, \
                           c.y, c.max_step, dx_fin, q_fin);                                                      \
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

hipError_t contact_end(const StepArgs& a, ContactDev& c, int cur, hipStream_t s) {
    hipLaunchKernelGGL(end_kernel, grid_for(c.n), dim3(kThreads), 0, s, a, c.n, c.dx[cur]);
    return hipGetLastError();
}

hipError_t sdf_pass(const StepArgs& a, int n, hipStream_t s) {
    hipLaunchKernelGGL(sdf_kernel, grid_for(n), dim3(kThreads), 0, s, a, n);
    return hipGetLastError();
}

}  // namespace kscr
