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


// initItr for one vertex (YB/sim/iteration.cu:8-59): inertial target y (as a
// displacement), VBD adaptive initial guess dx, lastVel update.  sigma_free
// reports whether the guess's factor was strictly inside (0, 1).
__device__ inline void init_vertex(const StepArgs& a, bool free, const f3& x, const f3& v, f3& u, f3& y, f3& dx,
                                   bool& sigma_free) {
    sigma_free = false;
    const float h = a.h;
    const f3 g{a.gravity[0], a.gravity[1], a.gravity[2]};
    float windx = 0.f;
    if (a.wind_peak != 0.f) {
        // cos of the double argument and pow in float (the reference's
        // cos(2 * M_PI * float) and pow(float, float)), by the deterministic
        // sequences of kscr_math.h
        const float wy =
            x.y > 0 ? (float)det_cos(2 * M_PI * (double)(a.wind_y_freq * x.y + a.wind_time_freq * a.time)) : 0.f;
        const float wz = (float)det_cos(2 * M_PI * (double)(a.wind_z_freq * fabsf(x.z) + a.wind_time_freq * a.time));
        const float yz = wy * wz;
        const float sg = (float)((0.f < yz) - (yz < 0.f));
        windx = a.wind_peak * fmaxf(0.f, det_powf(yz, a.wind_sharpness) * sg);
    }
    const f3 ext = (g + f3{windx, 0.f, 0.f}) * a.external_force_multiplier;
    y = v;
    dx = h * v;
    const f3 lastVel = u;
    u = v;
    if (free) {
        y = dx + (h * h) * ext;
        const float ext2 = dot(ext, ext);
        if (ext2 > 0) {
            const f3 acc = (v - lastVel) / a.last_h;
            const float raw = dot(acc, ext) / ext2;
            sigma_free = raw > 0 && raw < 1;
            const float sc = clamp01(raw);
            dx = dx + (h * h * sc) * ext;
        }
    }
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
