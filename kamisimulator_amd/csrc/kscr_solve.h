// kscr_solve.h -- per-vertex pieces of the step shared by the fused kernel
// and the contact-path kernels, so both evaluate the identical FP32 sequence.
#pragma once

#include "kscr_device.h"
#include "kscr_math.h"

namespace kscr {

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


// The wind's force along x at position x (YB/sim/iteration.cu:16-29): cos of
// the double argument and pow in float (the reference's cos(2 * M_PI *
// float) and pow(float, float)), by the deterministic sequences of
// kscr_math.h.  With DERIV, d_dy and d_dz receive its derivatives with
// respect to x.y and x.z on the branch x lies on (the kinks at y = 0, z = 0
// and where the clamp at 0 switches count as constant); sin(t) is formed as
// cos(t - pi / 2) by the same sequence.
struct WindForce {
    float fx, d_dy, d_dz;
};
template <bool DERIV>
__device__ inline WindForce wind_force(float peak, float y_freq, float z_freq, float time_freq, float sharpness,
                                       float time, f3 x) {
    WindForce w{0.f, 0.f, 0.f};
    const double thy = 2 * M_PI * (double)(y_freq * x.y + time_freq * time);
    const double thz = 2 * M_PI * (double)(z_freq * fabsf(x.z) + time_freq * time);
    const float wy = x.y > 0 ? (float)det_cos(thy) : 0.f;
    const float wz = (float)det_cos(thz);
    const float yz = wy * wz;
    const float sg = (float)((0.f < yz) - (yz < 0.f));
    const float pw = det_powf(yz, sharpness) * sg;
    w.fx = peak * fmaxf(0.f, pw);
    if (DERIV && pw > 0.f) {  // d max(0, pw) / d yz = sg * s * yz^(s-1) where pw > 0 (false for NaN)
        const float dmx = sg * (sharpness * det_powf(yz, sharpness - 1.0f));
        const float dwy = x.y > 0 ? -(float)det_cos(thy - 0.5 * M_PI) * ((float)(2 * M_PI) * y_freq) : 0.f;
        const float sz = (float)((0.f < x.z) - (x.z < 0.f));
        const float dwz = -(float)det_cos(thz - 0.5 * M_PI) * (((float)(2 * M_PI) * z_freq) * sz);
        w.d_dy = ((peak * dmx) * wz) * dwy;
        w.d_dz = ((peak * dmx) * wy) * dwz;
    }
    return w;
}

// Reverse of sdf_project at the particle (x, v) it was applied to: xb, vb hold
// the adjoints of the projected position and velocity on entry and of x, v
// on exit.  Cell choice and the outside-grid corners are piecewise constant;
// within a cell the weights are affine in x, and the projection is
// x' = x - phi n, v' = v - ((v . n) / |v|) n with n = g / |g| for the
// interpolated gradient g.
__device__ inline void sdf_project_vjp(const SdfArgs& s, f3 x, f3 v, f3& xb, f3& vb) {
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
    f4 c[8];
#pragma unroll
    for (int o = 0; o < 8; o++) {
        const int cx = idx[0] + ((o >> 2) & 1), cy = idx[1] + ((o >> 1) & 1), cz = idx[2] + (o & 1);
        const bool in = cx < s.dim[0] && cy < s.dim[1] && cz < s.dim[2];
        const int kx = min(cx, s.dim[0] - 1), ky = min(cy, s.dim[1] - 1), kz = min(cz, s.dim[2] - 1);
        const float4 g = s.grid[(size_t)kx + (size_t)ky * s.dim[0] + (size_t)kz * s.dim[0] * s.dim[1]];
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
    const f3 m = vec(r);
    const float inv_m = 1.0f / sqrtf(dot(m, m));
    const f3 n = m * inv_m;
    const float phi = r.w - s.thresh;
    if (!(phi < 0)) return;
    // v' = v - beta n, beta = (v / |v|) . n (when |v| >= 1e-6)
    f3 nb = (-phi) * xb;  // from x' = x - phi n
    const float phib = -dot(n, xb);
    const float vl = length(v);
    if ((double)vl >= 1e-6) {
        const f3 vn = v / vl;
        const float beta = dot(vn, n);
        const float betab = -dot(n, vb);
        nb = nb - beta * vb;
        nb = nb + betab * vn;
        vb = vb + (betab / vl) * (n - beta * vn);
    }
    // n = m / |m|
    const f3 mb = (nb - dot(n, nb) * n) * inv_m;
    const float rb[4] = {mb.x, mb.y, mb.z, phib};
    float udb[3] = {0.f, 0.f, 0.f};
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
        const float c0b = rb[k] * (1.f - ud[2]), c1b = rb[k] * ud[2];
        udb[2] += rb[k] * (c1 - c0);
        const float c00b = c0b * (1.f - ud[1]), c10b = c0b * ud[1], c01b = c1b * (1.f - ud[1]), c11b = c1b * ud[1];
        udb[1] += c0b * (c10 - c00) + c1b * (c11 - c01);
        udb[0] += ((c00b * (c100 - c000) + c01b * (c101 - c001)) + c10b * (c110 - c010)) + c11b * (c111 - c011);
    }
    const f3 lb{udb[0] * s.inv_ds[0], udb[1] * s.inv_ds[1], udb[2] * s.inv_ds[2]};
    xb = xb + rotate(f4{s.rot[0], s.rot[1], s.rot[2], s.rot[3]}, lb);
}


// initItr for one vertex (YB/sim/iteration.cu:8-59): inertial target y (as a
// displacement), VBD adaptive initial guess dx, lastVel update.  sigma_free
// reports whether the guess's factor was strictly inside (0, 1); acc
// receives the acceleration the guess projects (0 where there is no guess).
__device__ inline void init_vertex(const StepArgs& a, bool free, const f3& x, const f3& v, f3& u, f3& y, f3& dx,
                                   bool& sigma_free, f3& acc) {
    sigma_free = false;
    acc = f3{0.f, 0.f, 0.f};
    const float h = a.h;
    const f3 g{a.gravity[0], a.gravity[1], a.gravity[2]};
    float windx = 0.f;
    if (a.wind_peak != 0.f)
        windx = wind_force<false>(a.wind_peak, a.wind_y_freq, a.wind_z_freq, a.wind_time_freq, a.wind_sharpness,
                                  a.time, x).fx;
    const f3 ext = (g + f3{windx, 0.f, 0.f}) * a.external_force_multiplier;
    y = v;
    dx = h * v;
    const f3 lastVel = u;
    u = v;
    if (free) {
        y = dx + (h * h) * ext;
        const float ext2 = dot(ext, ext);
        if (ext2 > 0) {
            acc = (v - lastVel) / a.last_h;
            const float raw = dot(acc, ext) / ext2;
            sigma_free = raw > 0 && raw < 1;
            const float sc = clamp01(raw);
            dx = dx + (h * h * sc) * ext;
        }
    }
}

__device__ inline void init_vertex(const StepArgs& a, bool free, const f3& x, const f3& v, f3& u, f3& y, f3& dx,
                                   bool& sigma_free) {
    f3 acc;
    init_vertex(a, free, x, v, u, y, dx, sigma_free, acc);
}

__device__ inline void init_vertex(const StepArgs& a, bool free, const f3& x, const f3& v, f3& u, f3& y, f3& dx) {
    bool sigma_free;
    init_vertex(a, free, x, v, u, y, dx, sigma_free);
}

// Orientation solve of one segment (YB/sim/cosserat.cu:303-335, SCR Eq. 18):
// x, xj positions of the segment ends, dxn, dxj their new dx, qm / qp the
// previous-iteration orientations of the neighbouring segments, qrp / qr the
// rest rotors of segments i-1 and i.
// The sign switches phi (cosserat.cu:320, :328) test <conj(qm) q, qrp> and
// <conj(q) qp, qr>.  With <a, b> = Re(a conj(b)) and Re(ABC) = Re(BCA) these
// equal <q, qm qrp> and <q, qp conj(qr)>, the dot products of q with the two
// rotors b is built from, which are formed here anyway.
// ilr = 1 / lr (rounded to nearest); inv_norm receives the factor 1 / |w|
// that normalizes w to the result.
__device__ inline f4 orientation_update(f3 x, f3 xj, f3 dxn, f3 dxj, float lr, float ilr, float ks, f4 q, f4 qm,
                                        f4 qp, f4 qrp, f4 qr, uint32_t fl, float& inv_norm) {
    f3 vv = div_rn((xj - x) + (dxj - dxn), lr, ilr);  // the three divisions by lr
    vv = vv * (-2.0f * ks);
    f4 b{0, 0, 0, 0};
    if (fl & FLAG_PREV) {
        const f4 bm = rmul(qm, qrp);
        const float phi = dot(q, bm) > 0 ? 1.f : -1.f;
        b = b + phi * bm;
    }
    if (fl & FLAG_NEXT_ORI) {
        const f4 bp = rmul(qp, conj(qr));
        const float phi = dot(q, bp) > 0 ? 1.f : -1.f;
        b = b + phi * bp;
    }
    const float lambda = length(vv) + length(b);
    const f4 w = rmul_e1(rmul_pure(vv, b)) + lambda * b;
    inv_norm = rsqrt_rn(dot(w, w));
    return f4{w.x * inv_norm, w.y * inv_norm, w.z * inv_norm, w.w * inv_norm};
}
__device__ inline f4 orientation_update(f3 x, f3 xj, f3 dxn, f3 dxj, float lr, float ilr, float ks, f4 q, f4 qm,
                                        f4 qp, f4 qrp, f4 qr, uint32_t fl) {
    float inv_norm;
    return orientation_update(x, xj, dxn, dxj, lr, ilr, ks, q, qm, qp, qrp, qr, fl, inv_norm);
}

// endItr (YB/sim/iteration.cu:62-74)
__device__ inline void end_vertex(const StepArgs& a, bool free, f3 dx, f3& x, f3& v) {
    if (free) v = dx * (1.0f / a.h) * (1 - a.velocity_decay * a.h);
    x = x + dx;
}

}  // namespace kscr
