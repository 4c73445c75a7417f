// adjoint_oracle.hpp -- discrete adjoint (reverse mode) of the contact-free
// Stable Cosserat Rods step of scr_oracle.hpp.
//
// TEST INFRASTRUCTURE ONLY (see scr_oracle.hpp).  The reference has no
// adjoint; this differentiates the oracle's own Jacobi-snapshot step, in exact
// arithmetic, treating the quaternion sign switch phi (cosserat.cu:320, :328)
// and the clamp of the adaptive initial guess (iteration.cu:50) as piecewise
// constant.  T = double is checked against central finite differences of the
// FP64 forward (tests/test_oracle_adjoint.py); T = float is the restatement
// the HIP adjoint is compared with.
//
// Tape of one substep: x, v, u at its start, y, and every iterate dx^k, q^k.
#pragma once

#include <vector>

#include "scr_oracle.hpp"

namespace orc {

template <class T>
struct SubstepTape {
    std::vector<V3<T>> x, v, u, y;
    std::vector<std::vector<V3<T>>> dx;  // numItr + 1 iterates
    std::vector<std::vector<R4<T>>> q;   // numItr + 1 iterates
};

template <class T>
struct StepTape {
    std::vector<SubstepTape<T>> sub;
    T h, lastH, time;
};

template <class T>
struct Adjoint {
    // adjoints of the state (x, v, u = lastVel, q)
    std::vector<V3<T>> xb, vb, ub;
    std::vector<R4<T>> qb;
    // parameter adjoints: per-vertex stretch stiffness and rest rotor
    std::vector<T> ksb;
    std::vector<R4<T>> qrb;
    void resize(int n) {
        xb.assign(n, V3<T>{0, 0, 0});
        vb.assign(n, V3<T>{0, 0, 0});
        ub.assign(n, V3<T>{0, 0, 0});
        qb.assign(n, R4<T>{0, 0, 0, 0});
        ksb.assign(n, T(0));
        qrb.assign(n, R4<T>{0, 0, 0, 0});
    }
};

template <class T>
inline T clamp_sq(T x) {
    return std::min(std::max(x, T(0x1p-62)), T(0x1p62));
}

// rotate(q, e1) = (w^2 + x^2 - y^2 - z^2, 2(xy + wz), 2(xz - wy)) for
// q = (x, y, z, w); its transposed Jacobian with respect to q applied to rb.
template <class T>
inline R4<T> rot_e1_vjp(R4<T> q, V3<T> rb) {
    const T wb = T(2) * std::fma(q.w, rb.x, std::fma(q.z, rb.y, -(q.y * rb.z)));
    const T xb = T(2) * dot(rb, q.q());
    const T yb = T(2) * std::fma(q.x, rb.y, -std::fma(q.y, rb.x, q.w * rb.z));
    const T zb = T(2) * std::fma(q.x, rb.z, std::fma(q.w, rb.y, -(q.z * rb.x)));
    return R4<T>{xb, yb, zb, wb};
}

// a (x) conj(e1) = -(a (x) e1)
template <class T>
inline R4<T> rmul_conj_e1(R4<T> a) {
    return R4<T>{-a.w, -a.z, a.y, a.x};
}

// Accumulations of the sweep's adjoint-only terms, one FMA chain per
// component (the HIP sweep forms them the same way): acc + s a, acc + a . b,
// acc + (v, 0) (x) b.
template <class T>
inline V3<T> fma3(T s, V3<T> a, V3<T> acc) {
    return V3<T>{std::fma(s, a.x, acc.x), std::fma(s, a.y, acc.y), std::fma(s, a.z, acc.z)};
}
template <class T>
inline R4<T> fma4(T s, R4<T> a, R4<T> acc) {
    return R4<T>{std::fma(s, a.x, acc.x), std::fma(s, a.y, acc.y), std::fma(s, a.z, acc.z), std::fma(s, a.w, acc.w)};
}
template <class T>
inline T dot_acc(V3<T> a, V3<T> b, T acc) {
    return std::fma(a.z, b.z, std::fma(a.y, b.y, std::fma(a.x, b.x, acc)));
}
template <class T>
inline R4<T> rmul_pure_acc(V3<T> v, R4<T> b, R4<T> acc) {
    return R4<T>{std::fma(b.w, v.x, std::fma(v.y, b.z, std::fma(-b.y, v.z, acc.x))),
                 std::fma(b.w, v.y, std::fma(v.z, b.x, std::fma(-b.z, v.x, acc.y))),
                 std::fma(b.w, v.z, std::fma(v.x, b.y, std::fma(-b.x, v.y, acc.z))),
                 std::fma(-v.z, b.z, std::fma(-v.y, b.y, std::fma(-v.x, b.x, acc.w)))};
}

// The wind's force along x at position x (init_itr's expressions) and its
// derivatives with respect to x.y and x.z on the branch x lies on; sin(t) is
// cos(t - pi / 2) by the same cos as the forward's.
template <class T>
struct WindForce {
    T fx, d_dy, d_dz;
};
template <class T>
inline WindForce<T> wind_force(const Params& p, T time, V3<T> x) {
    WindForce<T> w{0, 0, 0};
    const double thy = 2 * M_PI * (double)((T)p.wind_y_freq * x.y + (T)p.wind_time_freq * time);
    const double thz = 2 * M_PI * (double)((T)p.wind_z_freq * std::fabs(x.z) + (T)p.wind_time_freq * time);
    const T wy = x.y > 0 ? wind_cos_t<T>(thy) : T(0);
    const T wz = wind_cos_t<T>(thz);
    const T yz = wy * wz;
    const T sg = (T)((T(0) < yz) - (yz < T(0)));
    const T pw = wind_pow_t(yz, (T)p.wind_sharpness) * sg;
    const T mx = std::isnan(pw) ? T(0) : std::max(T(0), pw);
    w.fx = (T)p.wind_peak * mx;
    if (pw > T(0)) {
        const T s = (T)p.wind_sharpness;
        const T dmx = sg * (s * wind_pow_t(yz, s - T(1)));
        const T dwy = x.y > 0 ? -wind_cos_t<T>(thy - 0.5 * M_PI) * ((T)(2 * M_PI) * (T)p.wind_y_freq) : T(0);
        const T sz = (T)((T(0) < x.z) - (x.z < T(0)));
        const T dwz = -wind_cos_t<T>(thz - 0.5 * M_PI) * (((T)(2 * M_PI) * (T)p.wind_z_freq) * sz);
        w.d_dy = (((T)p.wind_peak * dmx) * wz) * dwy;
        w.d_dz = (((T)p.wind_peak * dmx) * wy) * dwz;
    }
    return w;
}

template <class T>
struct AdjointSim : Sim<T> {
    using Sim<T>::n;
    using Sim<T>::p;
    using Sim<T>::x;
    using Sim<T>::v;
    using Sim<T>::u;
    using Sim<T>::q;
    using Sim<T>::qRest;
    using Sim<T>::invMass;
    using Sim<T>::lRest;
    using Sim<T>::kStretch;
    using Sim<T>::flags;
    using Sim<T>::y;
    using Sim<T>::dx;
    using Sim<T>::h;
    using Sim<T>::lastH;
    using Sim<T>::hF;
    using Sim<T>::lastHF;
    using Sim<T>::timeF;
    using Sim<T>::time;
    using Sim<T>::step_counter;
    using Sim<T>::sdf;

    // Forward step (advance) recording the tape.
    void record_step(float dtF, StepTape<T>& tape) {
        const int steps = std::max(1, (int)std::ceil(dtF / p.max_h));
        lastHF = hF;
        hF = dtF / steps;
        lastH = (T)lastHF;
        h = (T)hF;
        tape.h = h;
        tape.lastH = lastH;
        tape.time = time;
        tape.sub.assign(steps, SubstepTape<T>{});
        for (int s = 0; s < steps; s++, step_counter++) {
            SubstepTape<T>& t = tape.sub[s];
            t.x = x;
            t.v = v;
            t.u = u;
            this->init_itr();
            this->recompute_step_limit();
            t.y = y;
            t.dx.push_back(dx);
            t.q.push_back(q);
            for (int k = 0; k < p.num_itr; k++) {
                this->iterate();
                t.dx.push_back(dx);
                t.q.push_back(q);
            }
            this->end_itr();
        }
        timeF += dtF;
        time = (T)timeF;
    }

    // 1 / |w| of the forward's orientation solve of vertex i in iteration k
    // (orientation_solve's operations on the tape's values)
    T forward_inv_norm(const SubstepTape<T>& t, int k, int i) const {
        const uint32_t fl = flags[i];
        const std::vector<V3<T>>& DXn = t.dx[k + 1];
        const std::vector<R4<T>>& Q = t.q[k];
        V3<T> vv = ((t.x[i + 1] - t.x[i]) + (DXn[i + 1] - DXn[i])) / lRest[i];
        vv = vv * (T(-2) * kStretch[i]);
        R4<T> b{0, 0, 0, 0};
        if (fl & F_PREV) {
            const R4<T> bm = rmul(Q[i - 1], qRest[i - 1]);
            T phi = dot4(Q[i], bm) > 0 ? T(1) : T(-1);
            b = add4(b, scale4(phi, bm));
        }
        if (fl & F_NEXT_ORI) {
            const R4<T> bp = rmul(Q[i + 1], conj(qRest[i]));
            T phi = dot4(Q[i], bp) > 0 ? T(1) : T(-1);
            b = add4(b, scale4(phi, bp));
        }
        T lambda = length(vv) + length4(b);
        const R4<T> w = add4(rmul_e1(rmul_pure(vv, b)), scale4(lambda, b));
        return T(1) / std::sqrt(dot4(w, w));
    }

    // Reverse of apply_sdf at the particle (x, v) it was applied to: xb, vb
    // hold the adjoints of the projected position and velocity on entry and
    // of x, v on exit (the cell choice and the inside test count as
    // constant; within a cell the weights are affine in x).
    void sdf_vjp(V3<T> x, V3<T> v, V3<T>& xb, V3<T>& vb) const {
        const R4<T> invq{-(T)sdf.rot[0], -(T)sdf.rot[1], -(T)sdf.rot[2], (T)sdf.rot[3]};
        const R4<T> rq{(T)sdf.rot[0], (T)sdf.rot[1], (T)sdf.rot[2], (T)sdf.rot[3]};
        const V3<T> tr{(T)sdf.trans[0], (T)sdf.trans[1], (T)sdf.trans[2]};
        const V3<T> mn{(T)sdf.gmin[0], (T)sdf.gmin[1], (T)sdf.gmin[2]};
        const V3<T> lp = rotate(invq, x - tr) - mn;
        const T l[3] = {lp.x, lp.y, lp.z};
        int idx[3];
        T ud[3];
        for (int a = 0; a < 3; a++) idx[a] = (int)((double)(l[a] * (T)sdf.inv_ds[a]) + 0.5);
        if (!this->sdf_inside(idx[0], idx[1], idx[2])) return;
        for (int a = 0; a < 3; a++)
            ud[a] = (T)((double)sdf.inv_ds[a] * ((double)l[a] - ((double)idx[a] - 0.5) * (double)sdf.ds[a]));
        T c[8][4];
        for (int o = 0; o < 8; o++)
            this->sdf_corner(idx[0] + ((o >> 2) & 1), idx[1] + ((o >> 1) & 1), idx[2] + (o & 1), c[o]);
        T res[4];
        for (int k = 0; k < 4; k++) {
            T c00 = c[0][k] * (1.f - ud[0]) + c[4][k] * ud[0];
            T c01 = c[1][k] * (1.f - ud[0]) + c[5][k] * ud[0];
            T c10 = c[2][k] * (1.f - ud[0]) + c[6][k] * ud[0];
            T c11 = c[3][k] * (1.f - ud[0]) + c[7][k] * ud[0];
            T c0 = c00 * (1.f - ud[1]) + c10 * ud[1];
            T c1 = c01 * (1.f - ud[1]) + c11 * ud[1];
            res[k] = c0 * (1.f - ud[2]) + c1 * ud[2];
        }
        const V3<T> m{res[0], res[1], res[2]};
        const T inv_m = T(1) / std::sqrt(dot(m, m));
        const V3<T> nrm = m * inv_m;
        const T phi = res[3] - (T)sdf.thresh;
        if (!(phi < 0)) return;
        V3<T> nb = (-phi) * xb;
        const T phib = -dot(nrm, xb);
        const T vl = length(v);
        if ((double)vl >= 1e-6) {
            const V3<T> vn = v / vl;
            const T beta = dot(vn, nrm);
            const T betab = -dot(nrm, vb);
            nb = nb - beta * vb;
            nb = nb + betab * vn;
            vb = vb + (betab / vl) * (nrm - beta * vn);
        }
        const V3<T> mb = (nb - dot(nrm, nb) * nrm) * inv_m;
        const T rb[4] = {mb.x, mb.y, mb.z, phib};
        T udb[3] = {0, 0, 0};
        for (int k = 0; k < 4; k++) {
            T c00 = c[0][k] * (1.f - ud[0]) + c[4][k] * ud[0];
            T c01 = c[1][k] * (1.f - ud[0]) + c[5][k] * ud[0];
            T c10 = c[2][k] * (1.f - ud[0]) + c[6][k] * ud[0];
            T c11 = c[3][k] * (1.f - ud[0]) + c[7][k] * ud[0];
            T c0 = c00 * (1.f - ud[1]) + c10 * ud[1];
            T c1 = c01 * (1.f - ud[1]) + c11 * ud[1];
            const T c0b = rb[k] * (1.f - ud[2]), c1b = rb[k] * ud[2];
            udb[2] += rb[k] * (c1 - c0);
            const T c00b = c0b * (1.f - ud[1]), c10b = c0b * ud[1], c01b = c1b * (1.f - ud[1]), c11b = c1b * ud[1];
            udb[1] += c0b * (c10 - c00) + c1b * (c11 - c01);
            udb[0] += ((c00b * (c[4][k] - c[0][k]) + c01b * (c[5][k] - c[1][k])) + c10b * (c[6][k] - c[2][k])) +
                      c11b * (c[7][k] - c[3][k]);
        }
        const V3<T> lb{udb[0] * (T)sdf.inv_ds[0], udb[1] * (T)sdf.inv_ds[1], udb[2] * (T)sdf.inv_ds[2]};
        xb = xb + rotate(rq, lb);
    }

    // Reverse of one substep: adj holds the adjoints of the substep's outputs
    // (x', v', u', q') on entry and of its inputs (x, v, u, q) on exit.
    void backward_substep(const SubstepTape<T>& t, Adjoint<T>& adj) {
        const int K = p.num_itr;
        const T hh = h;
        const T damp = T(p.damping) / hh;
        const T invH = T(1) / hh;
        const T decay = T(1) - T(p.velocity_decay) * hh;
        std::vector<V3<T>> dxb(n), dxb2(n), fb(n), yb(n, V3<T>{0, 0, 0});
        std::vector<T> Db(n);
        std::vector<R4<T>> qb2(n);
        // endItr: x' = x + dx^K; v' = dx^K invH decay (free) or v (pinned)
        for (int i = 0; i < n; i++) {
            const bool free = invMass[i] != 0;
            dxb[i] = adj.xb[i] + (free ? adj.vb[i] * (invH * decay) : V3<T>{0, 0, 0});
            if (free) adj.vb[i] = V3<T>{0, 0, 0};  // pinned: v' = v keeps its adjoint
        }
        for (int k = K - 1; k >= 0; k--) {
            const std::vector<V3<T>>& DX = t.dx[k];
            const std::vector<V3<T>>& DXn = t.dx[k + 1];
            const std::vector<R4<T>>& Q = t.q[k];
            // ---- orientation solve (uses DXn and Q)
            for (int i = 0; i < n; i++) qb2[i] = R4<T>{0, 0, 0, 0};
            for (int i = 0; i < n; i++) {
                const uint32_t fl = flags[i];
                if ((fl & F_FIX_ORI) || !(fl & F_NEXT)) {
                    qb2[i] = add4(qb2[i], adj.qb[i]);
                    continue;
                }
                const V3<T> e = (t.x[i + 1] - t.x[i]) + (DXn[i + 1] - DXn[i]);
                const T sc = T(-2) * kStretch[i] / lRest[i];
                const V3<T> vv = e * sc;
                // b = phi_m Q[i-1] qRest[i-1] + phi_p Q[i+1] conj(qRest[i])
                R4<T> b{0, 0, 0, 0};
                T phim = 1, phip = 1;
                if (fl & F_PREV) {
                    const R4<T> bm = rmul(Q[i - 1], qRest[i - 1]);
                    phim = dot4(Q[i], bm) > 0 ? T(1) : T(-1);
                    b = scale4(phim, bm);
                }
                if (fl & F_NEXT_ORI) {
                    const R4<T> bp = rmul(Q[i + 1], conj(qRest[i]));
                    phip = dot4(Q[i], bp) > 0 ? T(1) : T(-1);
                    b = (fl & F_PREV) ? add4(b, scale4(phip, bp)) : scale4(phip, bp);
                }
                // lengths from the squared lengths clamped into [2^-62, 2^62]
                // (the HIP sweep's definition; no effect on a rod's lengths)
                const T lv = std::sqrt(clamp_sq(dot(vv, vv))), lb = std::sqrt(clamp_sq(dot4(b, b)));
                const T lambda = lv + lb;
                // normalize: the forward's result qn = w / |w| (the tape's
                // q^{k+1}) and its factor 1 / |w|
                const R4<T> qn = t.q[k + 1][i];
                const T ilw = forward_inv_norm(t, k, i);
                const R4<T> g = adj.qb[i];
                const R4<T> wb = scale4(ilw, fma4(-dot4(qn, g), qn, g));
                const T lamb = dot4(wb, b);
                // w = t + lambda b, t = m (x) e1: mbar = wbar (x) conj(e1)
                const R4<T> mb = rmul_conj_e1(wb);
                // m = (vv, 0) (x) b: vvbar = vec(mbar (x) conj(b)), bbar = lambda wbar + (-vv, 0) (x) mbar
                V3<T> vvb = rmul(mb, conj(b)).q();
                R4<T> bb = rmul_pure_acc(-vv, mb, scale4(lambda, wb));
                // lambda = |v| + |b|
                vvb = fma3(lamb * (T(1) / lv), vv, vvb);
                bb = fma4(lamb * (T(1) / lb), b, bb);
                // b terms, bm = Q[i-1] (x) qRest[i-1], bp = Q[i+1] (x) conj(qRest[i]):
                // phi times the product with bb (phi = +-1 commutes with the
                // rounding, so this is the product with phi bb)
                if (fl & F_PREV) {
                    qb2[i - 1] = fma4(phim, rmul(bb, conj(qRest[i - 1])), qb2[i - 1]);
                    adj.qrb[i - 1] = fma4(phim, rmul(conj(Q[i - 1]), bb), adj.qrb[i - 1]);
                }
                if (fl & F_NEXT_ORI) {
                    qb2[i + 1] = fma4(phip, rmul(bb, qRest[i]), qb2[i + 1]);
                    const R4<T> cqb = rmul(conj(Q[i + 1]), bb);
                    const R4<T> r = adj.qrb[i];
                    adj.qrb[i] = R4<T>{std::fma(-phip, cqb.x, r.x), std::fma(-phip, cqb.y, r.y),
                                       std::fma(-phip, cqb.z, r.z), std::fma(phip, cqb.w, r.w)};
                }
                // v = e * sc
                const V3<T> eb = vvb * sc;
                adj.ksb[i] = std::fma(dot(vvb, e), T(-2) / lRest[i], adj.ksb[i]);
                dxb[i + 1] = dxb[i + 1] + eb;
                dxb[i] = dxb[i] - eb;
                adj.xb[i + 1] = adj.xb[i + 1] + eb;
                adj.xb[i] = adj.xb[i] - eb;
            }
            // ---- position solve (uses DX and Q): dx^{k+1} = dx^k + accel H^-1 f
            for (int i = 0; i < n; i++) {
                dxb2[i] = dxb[i];
                fb[i] = V3<T>{0, 0, 0};
                Db[i] = 0;
            }
            for (int i = 0; i < n; i++) {
                if (invMass[i] == 0) continue;
                const T a_in = T(1) / (invMass[i] * hh * hh);
                // diagonal accumulated as the forward's H = H0 + H1 + H2
                T hd = a_in;
                V3<T> f = (T(1) / (hh * hh * invMass[i])) * (t.y[i] - DX[i]);
                if (flags[i] & F_NEXT) {
                    const T il = T(1) / lRest[i];
                    const V3<T> c = ((t.x[i + 1] - t.x[i]) + (DX[i + 1] - DX[i])) * il - rotate_e1(Q[i]);
                    const T kk = kStretch[i] * il, dd = kk * il;
                    f = f + (kk * c - (damp * dd) * DX[i]);
                    hd = hd + dd * (1 + damp);
                }
                if (flags[i] & F_PREV) {
                    const int j = i - 1;
                    const T il = T(1) / lRest[j];
                    const V3<T> c = ((t.x[i] - t.x[j]) + (DX[i] - DX[j])) * il - rotate_e1(Q[j]);
                    const T kk = kStretch[j] * il, dd = kk * il;
                    f = f + ((-kk) * c - (damp * dd) * DX[i]);
                    hd = hd + dd * (1 + damp);
                }
                const Sym3<T> H{{hd, hd, hd, a_in, a_in, a_in}};
                const V3<T> delta = sym_solve(H, f);
                fb[i] = sym_solve(H, T(p.acceleration_ratio) * dxb[i]);
                Db[i] = -dot(fb[i], delta);  // D enters the diagonal of H: Hbar = -fbar delta^T
                const T c1 = T(1) / (hh * hh * invMass[i]);
                yb[i] = fma3(c1, fb[i], yb[i]);
                dxb2[i] = fma3(-c1, fb[i], dxb2[i]);
            }
            for (int j = 0; j < n; j++) {
                if (!(flags[j] & F_NEXT)) continue;
                const T il = T(1) / lRest[j];
                const V3<T> c = ((t.x[j + 1] - t.x[j]) + (DX[j + 1] - DX[j])) * il - rotate_e1(Q[j]);
                const T kk = kStretch[j] * il, dd = kk * il;
                const V3<T> cb = kk * (fb[j] - fb[j + 1]);
                const V3<T> s = cb * il;
                adj.xb[j + 1] = adj.xb[j + 1] + s;
                adj.xb[j] = adj.xb[j] - s;
                dxb2[j + 1] = dxb2[j + 1] + s;
                dxb2[j] = dxb2[j] - s;
                const R4<T> qrot = rot_e1_vjp(Q[j], -cb);
                qb2[j] = add4(qb2[j], qrot);
                const T nd = -(damp * dd);
                dxb2[j] = fma3(nd, fb[j], dxb2[j]);
                dxb2[j + 1] = fma3(nd, fb[j + 1], dxb2[j + 1]);
                // kStretch through c's 1 / l, the damping term and the diagonal:
                // dot(c, fb_j - fb_j+1) il - damp il^2 (DX_j fb_j + DX_j+1 fb_j+1) + (1 + damp) il^2 (D_j + D_j+1)
                const T il2 = il * il;
                adj.ksb[j] = std::fma(dot(c, fb[j] - fb[j + 1]), il, adj.ksb[j]);
                adj.ksb[j] = std::fma(-(damp * il2), dot_acc(DX[j + 1], fb[j + 1], dot(DX[j], fb[j])), adj.ksb[j]);
                adj.ksb[j] = std::fma((T(1) + damp) * il2, Db[j] + Db[j + 1], adj.ksb[j]);
            }
            dxb.swap(dxb2);
            adj.qb.swap(qb2);
        }
        // ---- initItr: dx^0 = h v (+ h^2 sigma ext), y = h v + h^2 ext, u' = v
        const V3<T> g{(T)p.gravity[0], (T)p.gravity[1], (T)p.gravity[2]};
        const V3<T> ext = g * (T)p.external_force_multiplier;  // without wind
        const T ext2 = length2(ext);
        std::vector<V3<T>> ubn(n, V3<T>{0, 0, 0});
        for (int i = 0; i < n; i++) {
            V3<T> vb = adj.vb[i] + adj.ub[i];  // u' = v
            vb = vb + hh * dxb[i];
            if (invMass[i] != 0 && p.wind_peak != 0) {
                // ext depends on x through the wind: ext = (g + (wind(x), 0, 0)) efm
                vb = vb + hh * yb[i];
                const WindForce<T> w = wind_force<T>(p, time, t.x[i]);
                const V3<T> ew = (g + V3<T>{w.fx, 0, 0}) * (T)p.external_force_multiplier;
                const T ew2 = length2(ew);
                V3<T> extb = (hh * hh) * yb[i];
                if (ew2 > 0) {
                    const V3<T> acc = (t.v[i] - t.u[i]) / lastH;
                    const T raw = dot(acc, ew) / ew2;
                    extb = extb + (hh * hh * std::min(std::max(raw, T(0)), T(1))) * dxb[i];
                    if (raw > 0 && raw < 1) {
                        const T sb = hh * hh * dot(dxb[i], ew);
                        const V3<T> ab = (sb / ew2) * ew;
                        vb = vb + ab / lastH;
                        ubn[i] = ubn[i] - ab / lastH;
                        extb = extb + sb * (acc / ew2 - ((T(2) * raw) / ew2) * ew);
                    }
                }
                const T wb = extb.x * (T)p.external_force_multiplier;
                adj.xb[i].y = adj.xb[i].y + wb * w.d_dy;
                adj.xb[i].z = adj.xb[i].z + wb * w.d_dz;
            } else if (invMass[i] != 0) {
                vb = vb + hh * yb[i];
                if (ext2 > 0) {
                    const V3<T> a = (t.v[i] - t.u[i]) / lastH;
                    const T sraw = dot(a, ext) / ext2;
                    if (sraw > 0 && sraw < 1) {
                        const T sb = hh * hh * dot(dxb[i], ext);
                        const V3<T> ab = (sb / ext2) * ext;
                        vb = vb + ab / lastH;
                        ubn[i] = ubn[i] - ab / lastH;
                    }
                }
            }
            adj.vb[i] = vb;
        }
        adj.ub.swap(ubn);
        // x: the substep's x enters through x' = x + dx (xb passes through) and
        // through the stretch/orientation terms (accumulated above).
    }

    struct Checkpoint {
        std::vector<V3<T>> x, v, u;
        std::vector<R4<T>> q;
        T h, lastH, time;
        float hF, lastHF, timeF;
        long counter;
    };
    Checkpoint save() const {
        return Checkpoint{x, v, u, q, h, lastH, time, hF, lastHF, timeF, step_counter};
    }
    void restore(const Checkpoint& c) {
        x = c.x, v = c.v, u = c.u, q = c.q;
        h = c.h, lastH = c.lastH, time = c.time;
        hF = c.hF, lastHF = c.lastHF, timeF = c.timeF;
        step_counter = c.counter;
    }

    // Loss L = sum_s |x_tip(s) - target_s|^2 after nsteps steps of dtF, and its
    // gradient with respect to every vertex's kStretch and qRest.  Forward with
    // checkpoints every K steps; backward segment by segment, recomputing the
    // segment's steps from its checkpoint to record their tapes.  With
    // with_sdf every step is followed by the SDF pass (solve_next_step).  x0b,
    // v0b (optional) receive dL/dx and dL/dv of the initial state.
    T gradient(float dtF, int nsteps, int K, const std::vector<int>& tips, const std::vector<T>& targets,
               std::vector<T>& ksb, std::vector<R4<T>>& qrb, bool with_sdf = false,
               std::vector<V3<T>>* x0b = nullptr, std::vector<V3<T>>* v0b = nullptr) {
        std::vector<Checkpoint> ck;
        for (int s = 0; s < nsteps; s++) {
            if (s % K == 0) ck.push_back(save());
            this->advance(dtF);
            if (with_sdf) this->apply_sdf();
        }
        T loss = 0;
        Adjoint<T> adj;
        adj.resize(n);
        for (size_t k = 0; k < tips.size(); k++) {
            const V3<T> d = x[tips[k]] - V3<T>{targets[3 * k], targets[3 * k + 1], targets[3 * k + 2]};
            loss += dot(d, d);
            adj.xb[tips[k]] = T(2) * d;
        }
        const Checkpoint end = save();
        for (int seg = (int)ck.size() - 1; seg >= 0; seg--) {
            restore(ck[seg]);
            const int s0 = seg * K, s1 = std::min(nsteps, s0 + K);
            std::vector<StepTape<T>> tapes(s1 - s0);
            for (int s = s0; s < s1; s++) {
                record_step(dtF, tapes[s - s0]);
                if (with_sdf) this->apply_sdf();
            }
            for (int s = s1 - 1; s >= s0; s--) backward_step(tapes[s - s0], adj, with_sdf);
        }
        restore(end);
        ksb = adj.ksb;
        qrb = adj.qrb;
        if (x0b) *x0b = adj.xb;
        if (v0b) *v0b = adj.vb;
        return loss;
    }

    void backward_step(const StepTape<T>& tape, Adjoint<T>& adj, bool with_sdf = false) {
        const T h_save = h, lh_save = lastH, t_save = time;
        h = tape.h;
        lastH = tape.lastH;
        time = tape.time;
        if (with_sdf && sdf.enabled) {
            // the SDF pass after the last substep, at the state end_itr left
            const SubstepTape<T>& t = tape.sub.back();
            const int K = p.num_itr;
            const T invH = T(1) / h;
            for (int i = 0; i < n; i++) {
                if (!(flags[i] & F_PREV)) continue;  // strand roots are skipped
                const V3<T> vp = invMass[i] != 0 ? t.dx[K][i] * invH * (T(1) - T(p.velocity_decay) * h) : t.v[i];
                const V3<T> xp = t.x[i] + t.dx[K][i];
                sdf_vjp(xp, vp, adj.xb[i], adj.vb[i]);
            }
        }
        for (int s = (int)tape.sub.size() - 1; s >= 0; s--) backward_substep(tape.sub[s], adj);
        h = h_save;
        lastH = lh_save;
        time = t_save;
    }
};

}  // namespace orc
