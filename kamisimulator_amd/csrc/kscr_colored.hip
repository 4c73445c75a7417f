// kscr_colored.hip -- the Stable Cosserat Rods step with an even/odd vertex
// colouring (Gauss-Seidel over the two colours) instead of the reference's
// Jacobi snapshot.
//
// Colour = parity of the global vertex index: along a strand neighbouring
// vertices have different colours, so all vertices of one colour can be
// solved at once.  One iteration is
//   position solve of the even vertices (every dx at iteration k),
//   position solve of the odd vertices (even dx already updated),
//   orientation solve of the even segments (neighbouring rotors at k),
//   orientation solve of the odd segments (even rotors already updated),
// each solve being the reference's per-vertex update (YB/sim/cosserat.cu:
// 112-287 without contacts, :290-337) on the values it reads.  initItr,
// endItr and the SDF pass are those of the Jacobi step.
//
// Work mapping: one lane owns the vertex pair (2P, 2P+1), P global, so both
// colours keep every lane busy and one solve of each colour is one pass of
// the lane.  A workgroup owns a pack of whole strands of at most
// 2 BLOCK - 1 vertices.  Values crossing lanes go through LDS: the next
// lane's even x, dx and q, the previous lane's odd q, and the contribution of
// an odd vertex's segment to the next even vertex.  Three barriers per
// iteration (one and a half per vertex update).
#include "kscr_device.h"
#include "kscr_math.h"
#include "kscr_solve.h"

namespace kscr {
namespace {

template <int BLOCK>
struct ColoredLds {
    float x[3][2 * BLOCK];
    float dx[3][2 * BLOCK];  // even vertices only
    float4 q[2 * BLOCK];
    float4 f[2 * BLOCK];  // odd vertex's segment: contribution to the next (even) vertex
};

// One vertex of a lane's pair and its per-launch constants.
struct Vtx {
    f3 x, v, u, y, dx;
    f4 q, qr, qrp;
    float im, lr, ks, invl, kk, dd, f_in;
    uint32_t fl;
    bool act, free, has_next, has_prev, rotates;
    InertiaInv hinv;
};

__device__ inline void load_vertex(const StepArgs& a, int i, bool act, float damp, Vtx& s) {
    s.act = act;
    s.x = s.v = s.u = f3{0, 0, 0};
    s.q = s.qr = s.qrp = f4{0, 0, 0, 1};
    s.im = 0.f, s.lr = 1.f, s.ks = 0.f;
    s.fl = 0;
    float lrp = 1.f, ksp = 0.f;
    if (act) {
        s.x = ld3(a.x, i);
        s.v = ld3(a.v, i);
        s.u = ld3(a.u, i);
        s.q = to4(a.q[i]);
        s.qr = to4(a.q_rest[i]);
        const float4 m = a.mat[i];
        s.im = m.x;
        s.lr = m.y;
        s.ks = m.z;
        s.fl = __float_as_uint(m.w);
        if (s.fl & FLAG_PREV) {
            s.qrp = to4(a.q_rest[i - 1]);
            const float4 mp = a.mat[i - 1];
            lrp = mp.y;
            ksp = mp.z;
        }
    }
    s.has_next = (s.fl & FLAG_NEXT) != 0;
    s.has_prev = (s.fl & FLAG_PREV) != 0;
    s.free = act && s.im != 0.f;
    s.rotates = act && s.has_next && !(s.fl & FLAG_FIX_ORI);
    const float h = a.h;
    s.invl = 1.0f / s.lr;
    s.kk = s.ks * s.invl;
    s.dd = s.kk * s.invl;
    const float a_in = 1.0f / (s.im * h * h);
    s.f_in = 1.0f / (h * h * s.im);
    // diagonal of the position-solve matrix, as the Jacobi step forms it
    float hd = a_in;
    if (s.has_next) hd = hd + s.dd * (1 + damp);
    if (s.has_prev) {
        const float invlp = 1.0f / lrp;
        const float kkp = ksp * invlp;
        const float ddp = kkp * invlp;
        hd = hd + ddp * (1 + damp);
    }
    s.hinv = inertia_inverse(hd, a_in);
}

__device__ inline void store_vertex(const StepArgs& a, int i, const Vtx& s) {
    if (!s.act) return;
    st3(a.x, i, s.x);
    st3(a.v, i, s.v);
    st3(a.u, i, s.u);
    a.q[i] = make_float4(s.q.x, s.q.y, s.q.z, s.q.w);
}

template <int BLOCK>
__device__ __forceinline__ void run_pack_colored(const StepArgs& a, const int vbeg, const int vend, ColoredLds<BLOCK>& L) {
    const int t = threadIdx.x;
    const int base = vbeg & ~1;  // even vertex of the first pair
    const int i0 = base + 2 * t, i1 = i0 + 1;
    const int l0 = 2 * t, l1 = 2 * t + 1;  // LDS slots
    const float damp = a.damping / a.h;
    Vtx E, O;
    load_vertex(a, i0, i0 >= vbeg && i0 < vend, damp, E);
    load_vertex(a, i1, i1 >= vbeg && i1 < vend, damp, O);
    // E.has_next means O is the next vertex of the same strand (packs hold
    // whole strands), and O.has_prev that E is its previous one.

    for (int s = 0; s < a.num_substeps; s++) {
        init_vertex(a, E.free, E.x, E.v, E.u, E.y, E.dx);
        init_vertex(a, O.free, O.x, O.v, O.u, O.y, O.dx);
        __syncthreads();  // previous substep's LDS reads are complete
        L.x[0][l0] = E.x.x, L.x[1][l0] = E.x.y, L.x[2][l0] = E.x.z;
        L.x[0][l1] = O.x.x, L.x[1][l1] = O.x.y, L.x[2][l1] = O.x.z;
        L.dx[0][l0] = E.dx.x, L.dx[1][l0] = E.dx.y, L.dx[2][l0] = E.dx.z;
        L.q[l0] = make_float4(E.q.x, E.q.y, E.q.z, E.q.w);
        L.q[l1] = make_float4(O.q.x, O.q.y, O.q.z, O.q.w);
        __syncthreads();
        const f3 xjE = E.has_next ? O.x : E.x;
        const f3 xjO = O.has_next ? f3{L.x[0][l1 + 1], L.x[1][l1 + 1], L.x[2][l1 + 1]} : O.x;

        for (int k = 0; k < a.num_itr; k++) {
            // rotors are fixed during the position solves
            const f3 rE = rotate_e1(E.q), rO = rotate_e1(O.q);
            // odd vertex's segment (o, o+1) at iteration k: its part of the
            // next even vertex's force
            if (O.has_next) {
                const f3 dxj{L.dx[0][l1 + 1], L.dx[1][l1 + 1], L.dx[2][l1 + 1]};
                const f3 c = ((xjO - O.x) + (dxj - O.dx)) * O.invl - rO;
                const f3 f2 = (-O.kk) * c - (damp * O.dd) * dxj;
                L.f[l1] = make_float4(f2.x, f2.y, f2.z, 0.0f);
            }
            __syncthreads();
            // ---- even vertices
            f3 dxE = E.dx;
            if (E.free) {
                f3 f = E.f_in * (E.y - E.dx);
                if (E.has_next) {
                    const f3 c = ((xjE - E.x) + (O.dx - E.dx)) * E.invl - rE;
                    f = f + (E.kk * c - (damp * E.dd) * E.dx);
                }
                if (E.has_prev) {
                    const float4 p = L.f[l0 - 1];
                    f = f + f3{p.x, p.y, p.z};
                }
                dxE = E.dx + a.acceleration_ratio * inertia_solve(E.hinv, f);
            }
            L.dx[0][l0] = dxE.x, L.dx[1][l0] = dxE.y, L.dx[2][l0] = dxE.z;
            __syncthreads();
            // ---- odd vertices, with the new even dx
            const f3 dxjO = O.has_next ? f3{L.dx[0][l1 + 1], L.dx[1][l1 + 1], L.dx[2][l1 + 1]} : O.dx;
            f3 dxO = O.dx;
            if (O.free) {
                f3 f = O.f_in * (O.y - O.dx);
                if (O.has_next) {
                    const f3 c = ((xjO - O.x) + (dxjO - O.dx)) * O.invl - rO;
                    f = f + (O.kk * c - (damp * O.dd) * O.dx);
                }
                if (O.has_prev) {  // segment (e, o), formed as the even vertex's upper contribution
                    const f3 c = ((xjE - E.x) + (O.dx - dxE)) * E.invl - rE;
                    f = f + ((-E.kk) * c - (damp * E.dd) * O.dx);
                }
                dxO = O.dx + a.acceleration_ratio * inertia_solve(O.hinv, f);
            }
            // ---- even segments' orientation (neighbouring rotors at k)
            f4 qE = E.q;
            if (E.rotates) {
                const f4 qm = E.has_prev ? to4(L.q[l0 - 1]) : f4{0, 0, 0, 1};
                const f4 qp = (E.fl & FLAG_NEXT_ORI) ? O.q : f4{0, 0, 0, 1};
                qE = orientation_update(E.x, xjE, dxE, dxO, E.lr, E.invl, E.ks, E.q, qm, qp, E.qrp, E.qr, E.fl);
            }
            L.q[l0] = make_float4(qE.x, qE.y, qE.z, qE.w);
            __syncthreads();
            // ---- odd segments' orientation, with the new even rotors
            f4 qO = O.q;
            if (O.rotates) {
                const f4 qm = O.has_prev ? qE : f4{0, 0, 0, 1};
                const f4 qp = (O.fl & FLAG_NEXT_ORI) ? to4(L.q[l1 + 1]) : f4{0, 0, 0, 1};
                qO = orientation_update(O.x, xjO, dxO, dxjO, O.lr, O.invl, O.ks, O.q, qm, qp, O.qrp, O.qr, O.fl);
            }
            L.q[l1] = make_float4(qO.x, qO.y, qO.z, qO.w);
            E.dx = dxE, O.dx = dxO;
            E.q = qE, O.q = qO;
            // Next iteration: L.f[l1] is rewritten after the barriers that
            // order its reads here; L.dx[l0] after the next first barrier,
            // which orders the previous lane's read of the old value; L.q
            // entries are read only behind the barriers that follow their
            // writes.
        }
        end_vertex(a, E.free, E.dx, E.x, E.v);
        end_vertex(a, O.free, O.dx, O.x, O.v);
    }

    if (a.sdf_on && E.act && E.has_prev) sdf_project(a.sdf, E.x, E.v);
    if (a.sdf_on && O.act && O.has_prev) sdf_project(a.sdf, O.x, O.v);
    store_vertex(a, i0, E);
    store_vertex(a, i1, O);
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void step_colored_kernel(StepArgs a, const int2* __restrict__ packs) {
    __shared__ ColoredLds<BLOCK> lds;
    const int2 pk = packs[blockIdx.x];
    run_pack_colored<BLOCK>(a, pk.x, pk.y, lds);
}

}  // namespace

hipError_t launch_step_colored(const StepArgs& a, const int2* packs, int num_packs, int block, hipStream_t stream) {
    if (num_packs <= 0) return hipSuccess;
    const dim3 grid(num_packs);
    switch (block) {
#define KSCR_CASE(B)                                                                      \
    case B:                                                                               \
        hipLaunchKernelGGL(step_colored_kernel<B>, grid, dim3(B), 0, stream, a, packs);   \
        break;
        KSCR_CASE(64)
        KSCR_CASE(128)
        KSCR_CASE(256)
        KSCR_CASE(512)
#undef KSCR_CASE
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace kscr
