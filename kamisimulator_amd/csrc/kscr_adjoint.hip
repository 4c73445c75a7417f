// kscr_adjoint.hip -- reverse mode (adjoint) of the contact-free Stable
// Cosserat Rods step, one fused kernel per step.
//
// The forward step is the Jacobi-snapshot VBD step of kscr_step.hip
// (YB/sim/iteration.cu:8-74, YB/sim/cosserat.cu:25-337).  Its tape -- x, v, u
// and y of every substep and every iterate dx^k, q^k -- is written by
// step_record_kernel.  This kernel walks the substeps and iterations of one
// step backwards, strands packed in workgroups exactly as in the forward, the
// adjoint state of each vertex in registers and the neighbour contributions
// exchanged through LDS.  The per-vertex accumulation order is the order in
// which the sequential restatement adds the same terms, so FP32 results are
// reproducible bit for bit on the CPU.
//
// Differentiated quantities: the state (x, v, u, q) and the per-vertex
// parameters kStretch (stretch stiffness) and qRest (rest Darboux rotor, whose
// norm is the bend stiffness).  The quaternion sign switch of
// quaternionLambdaItr (cosserat.cu:320, :328) and the clamp of the adaptive
// initial guess (iteration.cu:50) are piecewise constant.
#include "kscr_device.h"
#include "kscr_math.h"
#include "kscr_solve.h"

namespace kscr {
namespace {

KS_FN f3 neg(f3 a) { return f3{-a.x, -a.y, -a.z}; }
KS_FN f4 zero4() { return f4{0.f, 0.f, 0.f, 0.f}; }
KS_FN f3 zero3() { return f3{0.f, 0.f, 0.f}; }

// Transposed Jacobian of rotate(q, e1) = (w^2 + x^2 - y^2 - z^2, 2(xy + wz),
// 2(xz - wy)) with respect to q = (x, y, z, w), applied to rb.
__device__ inline f4 rot_e1_vjp(f4 q, f3 rb) {
    const float wb = 2.0f * fmaf(q.w, rb.x, fmaf(q.z, rb.y, -(q.y * rb.z)));
    const float xb = 2.0f * dot(rb, vec(q));
    const float yb = 2.0f * fmaf(q.x, rb.y, -fmaf(q.y, rb.x, q.w * rb.z));
    const float zb = 2.0f * fmaf(q.x, rb.z, fmaf(q.w, rb.y, -(q.z * rb.x)));
    return f4{xb, yb, zb, wb};
}

// a (x) conj(e1) = -(a (x) e1)
KS_FN f4 rmul_conj_e1(f4 a) { return f4{-a.w, -a.z, a.y, a.x}; }

// Accumulations of the sweep's adjoint-only terms, each component one FMA
// chain (the oracle's adjoint forms them the same way):
// acc + s a
KS_FN f3 fma3(float s, f3 a, f3 acc) { return f3{fmaf(s, a.x, acc.x), fmaf(s, a.y, acc.y), fmaf(s, a.z, acc.z)}; }
KS_FN f4 fma4(float s, f4 a, f4 acc) {
    return f4{fmaf(s, a.x, acc.x), fmaf(s, a.y, acc.y), fmaf(s, a.z, acc.z), fmaf(s, a.w, acc.w)};
}
// acc + a . b
KS_FN float dot_acc(f3 a, f3 b, float acc) { return fmaf(a.z, b.z, fmaf(a.y, b.y, fmaf(a.x, b.x, acc))); }
// c ? a : b per component
KS_FN f3 sel3(bool c, f3 a, f3 b) { return f3{c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z}; }
KS_FN f4 sel4(bool c, f4 a, f4 b) { return f4{c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z, c ? a.w : b.w}; }
// acc + (v, 0) (x) b = acc + (b.w v + v x vec(b), -v . vec(b))
KS_FN f4 rmul_pure_acc(f3 v, f4 b, f4 acc) {
    return f4{fmaf(b.w, v.x, fmaf(v.y, b.z, fmaf(-b.y, v.z, acc.x))),
              fmaf(b.w, v.y, fmaf(v.z, b.x, fmaf(-b.z, v.x, acc.y))),
              fmaf(b.w, v.z, fmaf(v.x, b.y, fmaf(-b.x, v.y, acc.z))),
              fmaf(-v.z, b.z, fmaf(-v.y, b.y, fmaf(-v.x, b.x, acc.w)))};
}

__device__ inline f4 ld4(const float4* p, size_t i) { return to4(p[i]); }
__device__ inline void st4(float4* p, size_t i, f4 a) { p[i] = make_float4(a.x, a.y, a.z, a.w); }

// Values of one tape iterate a thread reads: its own and the next vertex's
// dx, the rotors of its segment and of the two neighbouring ones, and the
// normalization factor of the orientation solve that starts from it.
struct Iterate {
    f3 dx, dx_n;
    f4 q, qm, qp;
    float ilw;
};

template <int BLOCK>
struct AdjLds {
    float eb[3][BLOCK];   // d(segment vector e) of the orientation solve
    // contributions to the next vertex's rotor adjoint and to the previous
    // vertex's rotor and qRest adjoints, before their sign switch phi: the
    // reader adds phi times them (phi = +-1, so phi times a product is the
    // product of phi times a factor, exactly)
    float4 qpb[BLOCK];
    float4 qmb[BLOCK];
    float4 qrp[BLOCK];
    // 1: orientation solved, 2: previous term, 4: next term; 8 / 16: the
    // previous / next term's phi is -1
    int mask[BLOCK];
    float4 fbD[BLOCK];    // adjoint of the position solve's force, and of D
    float c[3][BLOCK];    // stretch constraint c of the segment at iterate k
};

// EXT: the step has wind or the SDF pass (their adjoints are compiled in);
// without them the kernel holds fewer registers.
template <int BLOCK, bool EXT>
__device__ __forceinline__ void adjoint_pack(const AdjointArgs& a, const int vbeg, const int vend, AdjLds<BLOCK>& L) {
    const int t = threadIdx.x;
    const int i = vbeg + t;
    const bool act = i < vend;
    const int n = a.tape.n;
    const int K = a.num_itr;
    const float4 mi = act ? a.mat[i] : make_float4(0.f, 1.f, 0.f, 0.f);
    const uint32_t fl = act ? __float_as_uint(mi.w) : 0u;
    const bool free = mi.x != 0.0f;
    const bool has_prev = (fl & FLAG_PREV) != 0, has_next = (fl & FLAG_NEXT) != 0;
    const bool has_next_ori = (fl & FLAG_NEXT_ORI) != 0;
    const bool pass = (fl & FLAG_FIX_ORI) || !has_next;
    const float4 mp = has_prev ? a.mat[i - 1] : make_float4(0.f, 1.f, 0.f, 0.f);
    const float im = mi.x, lr = mi.y, ks = mi.z;
    const float lrp = mp.y, ksp = mp.z;
    const f4 qr = act ? ld4(a.q_rest, i) : zero4();
    const f4 qrm = has_prev ? ld4(a.q_rest, i - 1) : zero4();

    // adjoint state of the step's outputs
    f3 xb = act ? ld3(a.xb, i) : zero3();
    f3 vb = act ? ld3(a.vb, i) : zero3();
    f3 ub = act ? ld3(a.ub, i) : zero3();
    f4 qb = act ? ld4(a.qb, i) : zero4();
    float ksb = act ? a.ksb[i] : 0.0f;
    f4 qrb = act ? ld4(a.qrb, i) : zero4();

    const float hh = a.h;
    const float damp = a.damping / hh;
    const float invH = 1.0f / hh;
    const float decay = 1.0f - a.velocity_decay * hh;
    // per-vertex constants of the step, with the forward's operations
    const float il = 1.0f / lr, kk = ks * il, dd = kk * il;          // own segment
    const float ilp = 1.0f / lrp, kkp = ksp * ilp, ddp = kkp * ilp;  // previous segment
    const float sc = (-2.0f * ks) / lr;                              // d(v)/d(e) of the orientation solve
    const float m2lr = -2.0f / lr;
    // coefficients of the position solve's adjoint terms in kStretch and of
    // the damping terms
    const float il2 = il * il;
    const float k_damp = -(damp * il2), k_diag = (1.0f + damp) * il2;
    const float nd = -(damp * dd), ndp = -(damp * ddp);
    const bool unit_ratio = a.acceleration_ratio == 1.0f;  // ratio * dxb == dxb exactly
    const float a_in = 1.0f / (im * hh * hh);
    const float c1 = 1.0f / (hh * hh * im);
    float hd = a_in;
    if (has_next) hd = hd + dd * (1.0f + damp);
    if (has_prev) hd = hd + ddp * (1.0f + damp);
    const InertiaInv hinv = inertia_inverse(hd, a_in);
    // Tape reads use indices clamped into [0, n): a thread past the pack or a
    // vertex without a neighbour loads some valid entry and never uses it
    // (every use is under the same act / has_* condition), so the loads need
    // no branches.
    const int ic = min(i, n - 1);
    const int inx = min(ic + 1, n - 1), ipv = max(ic - 1, 0);

    // ---- the SDF pass after the last substep (non-root vertices), at the
    // state endItr left: x + dx^K and dx^K / h * decay (free) or v (pinned,
    // which initItr's y holds)
    if (EXT && a.sdf_on && act && has_prev) {
        const size_t s1 = (size_t)(a.num_substeps - 1);
        const f3 dxK = ld3(a.tape.dx + (s1 * (K + 1) * n + (size_t)K * n) * 3, ic);
        f3 xp = ld3(a.tape.x + s1 * n * 3, ic);
        const f3 vp = free ? dxK * (1.0f / hh) * (1 - a.velocity_decay * hh) : ld3(a.tape.y + s1 * n * 3, ic);
        xp = xp + dxK;
        sdf_project_vjp(a.sdf, xp, vp, xb, vb);
    }

    for (int sub = a.num_substeps - 1; sub >= 0; sub--) {
        const size_t so = (size_t)sub * n;
        const f3 x = ld3(a.tape.x + so * 3, ic);
        const f3 xn = ld3(a.tape.x + so * 3, inx);
        const f3 y = ld3(a.tape.y + so * 3, ic);

        // ---- endItr: x' = x + dx^K, v' = dx^K / h * decay (free vertices)
        f3 dxb = xb + (free ? vb * (invH * decay) : zero3());
        if (free) vb = zero3();
        f3 yb = zero3();

        // Iterates of the tape: k + 1 is carried over from the previous pass
        // of the loop, k - 1 is loaded one pass ahead (the loads stay in
        // flight across the barriers, which wait on LDS traffic only).
        const size_t o_sub = (size_t)sub * (K + 1) * n;  // dx^k and q^k: K + 1 per substep
        const size_t w_sub = (size_t)sub * K * n;        // 1 / |w| of iteration k: K per substep
        // Reverse iteration k reads iterate k + 1 (dx, dx_n) and iterate k
        // (dx, dx_n and the rotors) and prefetches iterate k - 1 (iterate 0
        // again on the last pass).
        const auto load_iterate = [&](int k, Iterate& it) __attribute__((always_inline)) {
            const float* DXk = a.tape.dx + (o_sub + (size_t)k * n) * 3;
            const float4* Qk = a.tape.q + o_sub + (size_t)k * n;
            it.dx = ld3(DXk, ic), it.dx_n = ld3(DXk, inx);
            it.q = ld4(Qk, ic), it.qm = ld4(Qk, ipv), it.qp = ld4(Qk, inx);
            it.ilw = a.tape.ilw[w_sub + (size_t)k * n + ic];
        };
        const auto iteration = [&](int k, const Iterate& nx, const Iterate& cu, Iterate& pf)
                                   __attribute__((always_inline)) {
            load_iterate(max(k - 1, 0), pf);
            const f3 dxn = nx.dx, dxn_n = nx.dx_n;
            const f3 dx = cu.dx, dx_n = cu.dx_n;
            const f4 Q = cu.q, Qm = cu.qm, Qp = cu.qp;

            // stretch constraint of segment i at iterate k (the forward's c);
            // the next vertex reads it for its previous segment (a vertex
            // without a segment forms a value that nobody reads)
            const f3 c_next = ((xn - x) + (dx_n - dx)) * il - rotate_e1(Q);
            L.c[0][t] = c_next.x;
            L.c[1][t] = c_next.y;
            L.c[2][t] = c_next.z;

            // The iteration has no branches: every lane forms every term from
            // valid inputs (tape indices and neighbour lanes clamped into
            // range) and keeps or drops it with a select, which leaves the
            // scheduler one block to interleave; the kept values are those of
            // the term-by-term conditions below.
            const int tm = t > 0 ? t - 1 : 0, tp = t < BLOCK - 1 ? t + 1 : t;

            // ---- phase A: orientation solve of segment i (quaternionLambdaItr)
            const bool solve = act && !pass;
            const f3 e = (xn - x) + (dxn_n - dxn);
            const f3 vv = e * sc;
            // b = phi_m Q_{i-1} qRest_{i-1} + phi_p Q_{i+1} conj(qRest_i)
            const f4 bm = rmul(Qm, qrm);
            const f4 bp = rmul(Qp, conj(qr));
            const float phim = dot(Q, bm) > 0 ? 1.0f : -1.0f;
            const float phip = dot(Q, bp) > 0 ? 1.0f : -1.0f;
            const f4 bms = phim * bm, bps = phip * bp;
            const f4 b = has_prev ? sel4(has_next_ori, bms + bps, bms) : sel4(has_next_ori, bps, zero4());
            const float lv = sqrt_cr(clamp_sq(dot(vv, vv))), lb = sqrt_cr(clamp_sq(dot(b, b)));
            const float lambda = lv + lb;
            // the solve's result q^{k+1} = w / |w| and 1 / |w| come from the
            // tape (the forward's own values), not from a recomputation
            const f4 qn = nx.q;
            const f4 g = qb;
            const f4 wb = cu.ilw * fma4(-dot(qn, g), qn, g);
            const float lamb = dot(wb, b);
            // w = (vv, 0) (x) b (x) e1 + lambda b: mb = wb (x) conj(e1)
            const f4 mb = rmul_conj_e1(wb);
            f3 vvb = vec(rmul(mb, conj(b)));
            f4 bb = rmul_pure_acc(neg(vv), mb, lambda * wb);
            // lambda = |v| + |b|
            vvb = fma3(lamb * rcp_cr(lv), vv, vvb);
            bb = fma4(lamb * rcp_cr(lb), b, bb);
            // the neighbours' terms, before their sign switch: they read them
            // only under the mask bits and add phi times them
            const f4 qmb = rmul(bb, conj(qrm));
            const f4 qrp = rmul(conj(Qm), bb);
            const f4 qpb = rmul(bb, qr);
            // qrb + conj(phi_p conj(Q_{i+1}) (x) bb)
            const f4 cqb = rmul(conj(Qp), bb);
            qrb = sel4(solve && has_next_ori,
                       f4{fmaf(-phip, cqb.x, qrb.x), fmaf(-phip, cqb.y, qrb.y), fmaf(-phip, cqb.z, qrb.z),
                          fmaf(phip, cqb.w, qrb.w)},
                       qrb);
            const f3 eb = vvb * sc;
            ksb = solve ? fmaf(dot(vvb, e), m2lr, ksb) : ksb;
            const int mask = solve ? (1 | (has_prev ? (phim < 0 ? (2 | 8) : 2) : 0) |
                                      (has_next_ori ? (phip < 0 ? (4 | 16) : 4) : 0))
                                   : 0;
            L.qmb[t] = make_float4(qmb.x, qmb.y, qmb.z, qmb.w);
            L.qrp[t] = make_float4(qrp.x, qrp.y, qrp.z, qrp.w);
            L.qpb[t] = make_float4(qpb.x, qpb.y, qpb.z, qpb.w);
            L.eb[0][t] = eb.x;
            L.eb[1][t] = eb.y;
            L.eb[2][t] = eb.z;
            L.mask[t] = mask;
            __syncthreads();
            // gather, in the order the sequential sweep adds the terms
            const int mm = L.mask[tm], mn = L.mask[tp];
            const f4 qpbm = to4(L.qpb[tm]);
            const f3 ebm{L.eb[0][tm], L.eb[1][tm], L.eb[2][tm]};
            const f4 qmbn = to4(L.qmb[tp]), qrpn = to4(L.qrp[tp]);
            f4 qb2 = zero4();
            qb2 = sel4(has_prev && (mm & 4), fma4((mm & 16) ? -1.0f : 1.0f, qpbm, qb2), qb2);
            const bool from_prev = has_prev && (mm & 1);
            dxb = sel3(from_prev, dxb + ebm, dxb);
            xb = sel3(from_prev, xb + ebm, xb);
            qb2 = sel4(act && pass, qb2 + qb, qb2);
            dxb = sel3(solve, dxb - eb, dxb);
            xb = sel3(solve, xb - eb, xb);
            const bool from_next = has_next && (mn & 2);
            const float ph = (mn & 8) ? -1.0f : 1.0f;
            qb2 = sel4(from_next, fma4(ph, qmbn, qb2), qb2);
            qrb = sel4(from_next, fma4(ph, qrpn, qrb), qrb);

            // ---- phase B1: position solve of vertex i (cosseratItr)
            const bool moves = act && free;
            const f3 cprev{L.c[0][tm], L.c[1][tm], L.c[2][tm]};
            f3 f = c1 * (y - dx);
            f = sel3(has_next, f + (kk * c_next - (damp * dd) * dx), f);
            f = sel3(has_prev, f + ((-kkp) * cprev - (damp * ddp) * dx), f);
            const f3 delta = inertia_solve(hinv, f);
            const f3 fbm = inertia_solve(hinv, unit_ratio ? dxb : a.acceleration_ratio * dxb);
            const f3 fb = sel3(moves, fbm, zero3());
            const float Db = moves ? -dot(fbm, delta) : 0.0f;
            yb = sel3(moves, fma3(c1, fbm, yb), yb);
            f3 dxb2 = sel3(moves, fma3(-c1, fbm, dxb), dxb);
            L.fbD[t] = make_float4(fb.x, fb.y, fb.z, Db);
            __syncthreads();

            // ---- phase B2: stretch constraint of segment i - 1, recomputed
            // from the published adjoints with the operations vertex i - 1
            // uses (no third exchange), and of segment i
            const float4 pb = L.fbD[tm], nb = L.fbD[tp];
            const f3 fbp{pb.x, pb.y, pb.z};
            const f3 fbn{nb.x, nb.y, nb.z};
            const float Dbn = nb.w;
            const f3 s_prev = (kkp * (fbp - fb)) * ilp;
            xb = sel3(has_prev, xb + s_prev, xb);
            dxb2 = sel3(has_prev, fma3(ndp, fb, dxb2 + s_prev), dxb2);
            const f3 fd = fb - fbn;
            const f3 cb = kk * fd;
            const f3 s_own = cb * il;
            // kStretch through c's 1 / l, the damping term and the diagonal
            float ksn = fmaf(dot(c_next, fd), il, ksb);
            ksn = fmaf(k_damp, dot_acc(dx_n, fbn, dot(dx, fb)), ksn);
            ksn = fmaf(k_diag, Db + Dbn, ksn);
            ksb = has_next ? ksn : ksb;
            xb = sel3(has_next, xb - s_own, xb);
            qb2 = sel4(has_next, qb2 + rot_e1_vjp(Q, neg(cb)), qb2);
            dxb2 = sel3(has_next, fma3(nd, fb, dxb2 - s_own), dxb2);
            dxb = dxb2;
            qb = qb2;
        };
        Iterate A, B, C;
        {
            const float* DXK = a.tape.dx + (o_sub + (size_t)K * n) * 3;
            A.dx = ld3(DXK, ic), A.dx_n = ld3(DXK, inx);
            A.q = ld4(a.tape.q + o_sub + (size_t)K * n, ic);
        }
        if (K > 0) load_iterate(K - 1, B);
        // The loop is not unrolled: the roles rotate by register copies.
        // Unrolled by three (no copies) the kernel needs ~210 VGPRs instead
        // of ~160, and two of its waves then leave no room on a SIMD for the
        // re-record kernel that runs beside it on the other stream.
#pragma nounroll
        for (int k = K - 1; k >= 0; k--) {
            iteration(k, A, B, C);
            A = B;
            B = C;
        }

        // ---- initItr: y = x + h v + h^2 ext (free), dx^0 = h v (+ h^2 s ext), u' = v
        f3 vbn = vb + ub;
        vbn = vbn + hh * dxb;
        f3 ubn = zero3();
        if (EXT && act && free && a.wind_peak != 0) {
            // ext depends on x through the wind: ext = (g + (wind(x), 0, 0)) efm
            vbn = vbn + hh * yb;
            const WindForce w = wind_force<true>(a.wind_peak, a.wind_y_freq, a.wind_z_freq, a.wind_time_freq,
                                                 a.wind_sharpness, a.time, x);
            const f3 g{a.gravity[0], a.gravity[1], a.gravity[2]};
            const f3 ext = (g + f3{w.fx, 0.f, 0.f}) * a.external_force_multiplier;
            const float ext2 = dot(ext, ext);
            f3 extb = (hh * hh) * yb;
            if (ext2 > 0) {
                // s = clamp(raw, 0, 1), raw = acc . ext / |ext|^2 (acc from the tape)
                const f3 acc = ld3(a.tape.acc + so * 3, ic);
                const float raw = dot(acc, ext) / ext2;
                extb = extb + (hh * hh * clamp01(raw)) * dxb;
                if (a.tape.sigma_free[so + i]) {
                    const float sb = hh * hh * dot(dxb, ext);
                    const f3 ab = (sb / ext2) * ext;
                    vbn = vbn + ab / a.last_h;
                    ubn = ubn - ab / a.last_h;
                    extb = extb + sb * (acc / ext2 - ((2.0f * raw) / ext2) * ext);
                }
            }
            const float wb = extb.x * a.external_force_multiplier;
            xb.y = xb.y + wb * w.d_dy;
            xb.z = xb.z + wb * w.d_dz;
        } else if (act && free) {
            vbn = vbn + hh * yb;
            if (a.ext2 > 0) {
                const f3 ext{a.ext[0], a.ext[1], a.ext[2]};
                // the forward's sigma = clamp(a . ext / |ext|^2, 0, 1), inside
                // its clamp range
                if (a.tape.sigma_free[so + i]) {
                    const float sb = hh * hh * dot(dxb, ext);
                    const f3 ab = (sb / a.ext2) * ext;
                    vbn = vbn + ab / a.last_h;
                    ubn = ubn - ab / a.last_h;
                }
            }
        }
        vb = vbn;
        ub = ubn;
    }
    if (act) {
        st3(a.xb, i, xb);
        st3(a.vb, i, vb);
        st3(a.ub, i, ub);
        st4(a.qb, i, qb);
        a.ksb[i] = ksb;
        st4(a.qrb, i, qrb);
    }
}

template <int BLOCK, bool EXT>
__global__ __launch_bounds__(BLOCK) void step_adjoint_kernel(AdjointArgs a, const int2* __restrict__ packs) {
    __shared__ AdjLds<BLOCK> lds;
    const int2 pk = packs[blockIdx.x];
    adjoint_pack<BLOCK, EXT>(a, pk.x, pk.y, lds);
}

}  // namespace

hipError_t launch_step_adjoint(const AdjointArgs& a, const int2* packs, int num_packs, int block, hipStream_t stream) {
    if (num_packs <= 0) return hipSuccess;
    const dim3 grid(num_packs);
    const bool ext = a.wind_peak != 0 || a.sdf_on;
    switch (block) {
#define KSCR_CASE(B)                                                                                  \
    case B:                                                                                           \
        if (ext)                                                                                      \
            hipLaunchKernelGGL((step_adjoint_kernel<B, true>), grid, dim3(B), 0, stream, a, packs);   \
        else                                                                                          \
            hipLaunchKernelGGL((step_adjoint_kernel<B, false>), grid, dim3(B), 0, stream, a, packs);  \
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

}  // namespace kscr
