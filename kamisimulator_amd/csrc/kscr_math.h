// kscr_math.h -- small vector, rotor and symmetric-3x3 algebra for the
// Stable Cosserat Rods step, usable from host C++ and from gfx950 device code.
//
// Operation order follows the formulas the reference evaluates (glm 0.9.9 and
// KittenEngine's Rotor/SymMat), so that FP32 results agree with the
// reference's to rounding:
//   quaternion product   KE/includes/modules/Rotor.h:210-212
//   rotation of a vector KE/includes/modules/Rotor.h:114-123
//   fromTo               KE/includes/modules/Rotor.h:79-88
//   3x3 inverse          extern/glm/glm/detail/func_matrix.inl:272-294
//   mat3 * vec3          extern/glm/glm/detail/type_mat3x3.inl:418-424
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define KS_FN __host__ __device__ inline
#else
#define KS_FN inline
#endif

namespace kscr {

struct f3 {
    float x, y, z;
};
struct f4 {
    float x, y, z, w;
};

// glm's min/max/clamp (comparison based: a NaN first argument propagates)
KS_FN float gmin(float x, float y) { return (y < x) ? y : x; }
KS_FN float gmax(float x, float y) { return (x < y) ? y : x; }
KS_FN float clamp01(float x) { return gmin(gmax(x, 0.0f), 1.0f); }

KS_FN f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
KS_FN f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
KS_FN f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
KS_FN f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
KS_FN f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
KS_FN f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
// a / b rounded to nearest, from y = 1 / b rounded to nearest: the product
// a y and two FMA residual corrections (Markstein's theorem: once q is within
// one ulp of a / b, fma(fma(-q, b, a), y, q) is the correctly rounded
// quotient; the first correction brings a y within half an ulp and a bit).
// This holds while no step underflows or overflows, which |a| in
// [2^-64, 2^64] and a normal b of moderate exponent (a rest length) ensure;
// any other a -- zeros included, whose sign the corrections would lose --
// takes the division itself.  Equal to a / b bit for bit, in ~6 instead of
// ~11 VALU operations.
KS_FN float div_rn_unguarded(float a, float b, float y) {
    float q = a * y;
    float r = fmaf(-q, b, a);
    q = fmaf(r, y, q);
    r = fmaf(-q, b, a);
    return fmaf(r, y, q);
}
// one range test for the three components (a NaN component passes it and
// stays NaN)
KS_FN f3 div_rn(f3 a, float b, float y) {
    const float lo = fminf(fabsf(a.x), fminf(fabsf(a.y), fabsf(a.z)));
    const float hi = fmaxf(fabsf(a.x), fmaxf(fabsf(a.y), fabsf(a.z)));
    if (!(lo >= 0x1p-64f && hi <= 0x1p64f)) return a / b;
    return f3{div_rn_unguarded(a.x, b, y), div_rn_unguarded(a.y, b, y), div_rn_unguarded(a.z, b, y)};
}
// Step-time algebra uses fused multiply-adds (one rounding for a*b + c), the
// way the reference's CUDA compiler contracts its FP32 arithmetic.  The
// grouping is fixed here and restated in the CPU oracle, so both sides round
// identically.
KS_FN float dot(f3 a, f3 b) { return fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x)); }
KS_FN f3 cross(f3 a, f3 b) {
    return f3{fmaf(a.y, b.z, -(b.y * a.z)), fmaf(a.z, b.x, -(b.z * a.x)), fmaf(a.x, b.y, -(b.x * a.y))};
}
// sqrtf(x), and 1.0f / sqrtf(x), with fewer operations on the device.
// sqrt_rn is hipcc's correctly rounded sqrtf without its final special-case
// select: inputs below 2^-64 scaled by 2^32 (exact) and the root by 2^-16,
// the hardware root s moved to s -+ 1 ulp by the signs of the FMA residuals
// x - s' s; zeros, infinities and NaN come out of the fix-up as they are.
// rcp_rn is the hardware reciprocal with one FMA correction, correctly
// rounded for every argument in [2^-126, 2^126).  Over all 2^32 inputs,
// sqrt_rn equals sqrtf and rsqrt_rn equals 1.0f / sqrtf except at +-0 and
// +inf, where it gives NaN for infinity or zero -- and a vector scaled by it
// is NaN either way (0 * inf, inf * 0).
KS_FN float sqrt_rn(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    const bool small = x < 0x1p-64f;
    const float xs = small ? x * 0x1p32f : x;
    const float s = __builtin_amdgcn_sqrtf(xs);
    const float sm = __uint_as_float(__float_as_uint(s) - 1u);
    const float rm = fmaf(-sm, s, xs);
    const float sp = __uint_as_float(__float_as_uint(s) + 1u);
    const float rp = fmaf(-sp, s, xs);
    const float t0 = rm <= 0.0f ? sm : s;
    const float t = rp > 0.0f ? sp : t0;
    return small ? t * 0x1p-16f : t;
#else
    return sqrtf(x);
#endif
}
KS_FN float rsqrt_rn(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    const float s = sqrt_rn(x);
    const float y = __builtin_amdgcn_rcpf(s);
    return fmaf(fmaf(-y, s, 1.0f), y, y);
#else
    return 1.0f / sqrtf(x);
#endif
}
#if defined(__HIPCC__)
// |a| and 1 / a of the reverse sweep, for arguments clamped into the range
// where the hardware square root and reciprocal with their FMA corrections
// are correctly rounded (compared with sqrtf and 1.0f / x over every
// significand at exponents -62 to 63: equal): ~4 and ~4 VALU operations
// instead of ~16 and ~11, branch free.  The clamp, [2^-62, 2^62] for the
// squared length, is part of the adjoint's definition (the oracle applies it
// too); it changes nothing for a rod's lengths and only keeps a zero length
// from dividing by zero.
__device__ inline float clamp_sq(float x) { return __builtin_amdgcn_fmed3f(x, 0x1p-62f, 0x1p62f); }
__device__ inline float sqrt_cr(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sm = __uint_as_float(__float_as_uint(s) - 1u);
    const float rm = fmaf(-sm, s, x);
    const float sp = __uint_as_float(__float_as_uint(s) + 1u);
    const float rp = fmaf(-sp, s, x);
    const float t = rm <= 0.0f ? sm : s;
    return rp > 0.0f ? sp : t;
}
__device__ inline float rcp_cr(float b) {
    const float y = __builtin_amdgcn_rcpf(b);
    return fmaf(fmaf(-y, b, 1.0f), y, y);
}

#endif
KS_FN float length(f3 a) { return sqrt_rn(dot(a, a)); }
KS_FN f3 normalize(f3 a) { return a * rsqrt_rn(dot(a, a)); }

KS_FN float dot(f4 a, f4 b) { return fmaf(a.w, b.w, fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x))); }
KS_FN float length(f4 a) { return sqrt_rn(dot(a, a)); }
KS_FN f4 normalize(f4 a) {
    float s = rsqrt_rn(dot(a, a));
    return f4{a.x * s, a.y * s, a.z * s, a.w * s};
}
KS_FN f4 operator+(f4 a, f4 b) { return f4{a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
KS_FN f4 operator*(float s, f4 a) { return f4{s * a.x, s * a.y, s * a.z, s * a.w}; }
KS_FN f3 vec(f4 a) { return f3{a.x, a.y, a.z}; }
KS_FN f4 mk4(f3 v, float w) { return f4{v.x, v.y, v.z, w}; }

// conjugate == inverse of a unit rotor
KS_FN f4 conj(f4 a) { return f4{-a.x, -a.y, -a.z, a.w}; }

// Hamilton product a*b
KS_FN f4 rmul(f4 a, f4 b) {
    const f3 aq = vec(a), bq = vec(b);
    const f3 cr = cross(aq, bq);
    return f4{fmaf(a.w, bq.x, fmaf(b.w, aq.x, cr.x)), fmaf(a.w, bq.y, fmaf(b.w, aq.y, cr.y)),
              fmaf(a.w, bq.z, fmaf(b.w, aq.z, cr.z)), fmaf(a.w, b.w, -dot(aq, bq))};
}

// r * v * r^-1
KS_FN f3 rotate(f4 r, f3 v) {
    const f3 q = vec(r);
    const f3 c1 = cross(q, v);
    const f3 a{fmaf(r.w, v.x, c1.x), fmaf(r.w, v.y, c1.y), fmaf(r.w, v.z, c1.z)};
    const float c = dot(v, q);
    const f3 c2 = cross(q, a);
    return f3{fmaf(r.w, a.x, fmaf(c, q.x, c2.x)), fmaf(r.w, a.y, fmaf(c, q.y, c2.y)), fmaf(r.w, a.z, fmaf(c, q.z, c2.z))};
}

// rotate(r, (1,0,0)): the general formula with the products by 0 and 1 of
// e1 done exactly (r.w e1 + q x e1 = (w, q.z, -q.y), e1.q = q.x); the values
// are the general formula's, up to the sign of exact zeros.
KS_FN f3 rotate_e1(f4 r) {
    const f3 q = vec(r);
    const f3 a{r.w, q.z, -q.y};
    const float c = q.x;
    const f3 c2 = cross(q, a);
    return f3{fmaf(r.w, a.x, fmaf(c, q.x, c2.x)), fmaf(r.w, a.y, fmaf(c, q.y, c2.y)), fmaf(r.w, a.z, fmaf(c, q.z, c2.z))};
}

// m (x) (e1, 0) = (m.w, m.z, -m.y, -m.x), and (v, 0) (x) b: rmul with the
// zero and unit factors done exactly (same values up to the sign of zeros).
KS_FN f4 rmul_e1(f4 m) { return f4{m.w, m.z, -m.y, -m.x}; }
KS_FN f4 rmul_pure(f3 v, f4 b) {
    const f3 bq = vec(b);
    const f3 cr = cross(v, bq);
    return f4{fmaf(b.w, v.x, cr.x), fmaf(b.w, v.y, cr.y), fmaf(b.w, v.z, cr.z), -dot(v, bq)};
}

// Second column of the orthonormal basis built around n (used for the
// antiparallel case of from_to).
KS_FN f3 ortho_basis_x1(f3 n) {
    if (n.z >= n.y) {
        const float a = 1.0f / (1.0f + n.z);
        const float b = -n.x * n.y * a;
        return f3{b, 1.0f - n.y * n.y * a, -n.y};
    }
    const float a = 1.0f / (1.0f + n.y);
    return f3{1.0f - n.x * n.x * a, -n.x, -n.x * n.z * a};
}

// Shortest-arc rotor taking unit vector a to unit vector b.
KS_FN f4 from_to(f3 a, f3 b) {
    f3 h = (a + b) * 0.5f;
    float l = dot(h, h);
    if (l > 0)
        h = h * (1.0f / sqrtf(l));
    else
        h = ortho_basis_x1(a);
    return mk4(cross(a, h), dot(a, h));
}

// Symmetric 3x3 stored as diagonal (d0, d1, d2) and off-diagonal
// (m01, m02, m12).  Solves H x = f through the adjugate/determinant formula.
KS_FN f3 sym_solve(float m00, float m11, float m22, float m01, float m02, float m12, f3 f) {
    const float m10 = m01, m20 = m02, m21 = m12;
    // 2x2 minors (glm inverse's cofactors), each a*b - c*d with one fma
    const float c00 = fmaf(m11, m22, -(m21 * m12));
    const float c01 = fmaf(m01, m22, -(m21 * m02));
    const float c02 = fmaf(m01, m12, -(m11 * m02));
    const float c10 = fmaf(m10, m22, -(m20 * m12));
    const float c20 = fmaf(m10, m21, -(m20 * m11));
    const float c11 = fmaf(m00, m22, -(m20 * m02));
    const float c21 = fmaf(m00, m21, -(m20 * m01));
    const float c12 = fmaf(m00, m12, -(m10 * m02));
    const float c22 = fmaf(m00, m11, -(m10 * m01));
    const float ood = 1.0f / fmaf(m20, c02, fmaf(-m10, c01, m00 * c00));
    const float i00 = c00 * ood, i10 = -c10 * ood, i20 = c20 * ood;
    const float i01 = -c01 * ood, i11 = c11 * ood, i21 = -c21 * ood;
    const float i02 = c02 * ood, i12 = -c12 * ood, i22 = c22 * ood;
    return f3{fmaf(i20, f.z, fmaf(i10, f.y, i00 * f.x)), fmaf(i21, f.z, fmaf(i11, f.y, i01 * f.x)),
              fmaf(i22, f.z, fmaf(i12, f.y, i02 * f.x))};
}

// The contact-free position-solve matrix has every entry a (the reference's
// hess3(1/(m h^2)) fills all six entries) and diagonal hd (a plus the stretch
// terms).  sym_solve's cofactors of it take three distinct values (c00 = c11
// = c22, c01 = c10 = c12 = c21, c02 = c20, fma being symmetric in its
// factors), so its inverse is fixed by three numbers; computed once, they give
// exactly sym_solve(hd, hd, hd, a, a, a, f) for every f.
struct InertiaInv {
    float p, q, r;  // diagonal, entries (0,1) (1,2), entries (0,2)
};
KS_FN InertiaInv inertia_inverse(float hd, float a) {
    const float c00 = fmaf(hd, hd, -(a * a));
    const float c01 = fmaf(a, hd, -(a * a));
    const float c02 = fmaf(a, a, -(hd * a));
    const float ood = 1.0f / fmaf(a, c02, fmaf(-a, c01, hd * c00));
    return InertiaInv{c00 * ood, -c01 * ood, c02 * ood};
}
KS_FN f3 inertia_solve(InertiaInv m, f3 f) {
    return f3{fmaf(m.r, f.z, fmaf(m.q, f.y, m.p * f.x)), fmaf(m.q, f.z, fmaf(m.p, f.y, m.q * f.x)),
              fmaf(m.p, f.z, fmaf(m.q, f.y, m.r * f.x))};
}

// Natural logarithm by a fixed sequence of IEEE float operations: d = m 2^e
// with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s), s = (m-1)/(m+1), by its
// odd series to s^13, plus e ln2 with ln2 split into a high and a low part.
// Host and device evaluate the identical sequence (used by the contact
// barrier, whose argument lies in [1e-3, 1]).
KS_FN float det_logf(float d) {
    int e = 0;
    float m = frexpf(d, &e);
    if (m < 0.70710678f) {
        m = m * 2.0f;
        e = e - 1;
    }
    const float s = (m - 1.0f) / (m + 1.0f);
    const float s2 = s * s;
    float p = 1.0f / 13.0f;
    p = p * s2 + 1.0f / 11.0f;
    p = p * s2 + 1.0f / 9.0f;
    p = p * s2 + 1.0f / 7.0f;
    p = p * s2 + 1.0f / 5.0f;
    p = p * s2 + 1.0f / 3.0f;
    p = p * s2 + 1.0f;
    const float lm = 2.0f * s * p;
    const float fe = (float)e;
    return fe * 6.93145752e-1f + (lm + fe * 1.42860677e-6f);
}

// cos of a double by a fixed sequence of IEEE double operations (host and
// device give identical bits): x = k pi/2 + r with pi/2 split into parts of
// 30, 30 and 53 bits (Cody-Waite reduction, exact for |k| < 2^22), |r| <= pi/4,
// then the Taylor series of cos r (to r^18) or sin r (to r^19) by Horner.
KS_FN double det_cos(double x) {
    const double k = rint(x * 0.6366197723675814);
    const double r = ((x - k * 1.570796325802803) - k * 9.920935791635221e-10) - k * 5.170182981794105e-19;
    const double r2 = r * r;
    double c = 1.5619206968586225e-16;  // 1/18!
    c = c * r2 - 4.779477332387385e-14;
    c = c * r2 + 1.1470745597729725e-11;
    c = c * r2 - 2.08767569878681e-09;
    c = c * r2 + 2.755731922398589e-07;
    c = c * r2 - 2.48015873015873e-05;
    c = c * r2 + 0.001388888888888889;
    c = c * r2 - 0.041666666666666664;
    c = c * r2 + 0.5;
    c = 1.0 - c * r2;
    double sn = 8.22063524662433e-18;  // 1/19!
    sn = sn * r2 - 2.8114572543455206e-15;
    sn = sn * r2 + 7.647163731819816e-13;
    sn = sn * r2 - 1.6059043836821613e-10;
    sn = sn * r2 + 2.505210838544172e-08;
    sn = sn * r2 - 2.7557319223985893e-06;
    sn = sn * r2 + 0.0001984126984126984;
    sn = sn * r2 - 0.008333333333333333;
    sn = sn * r2 + 0.16666666666666666;
    sn = r - (sn * r2) * r;
    switch ((long long)k & 3) {
        case 0: return c;
        case 1: return -sn;
        case 2: return -c;
        default: return sn;
    }
}

// Natural log of a finite double > 0: x = m 2^e, m in [sqrt(1/2), sqrt(2)),
// log m = 2 atanh(s), s = (m - 1)/(m + 1), odd series to s^25, plus e ln 2
// with ln 2 split into a 30-bit and a 53-bit part.
KS_FN double det_log(double x) {
    int e = 0;
    double m = frexp(x, &e);
    if (m < 0.7071067811865476) {
        m = m * 2.0;
        e = e - 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double s2 = s * s;
    double p = 0.04;  // 1/25
    p = p * s2 + 0.043478260869565216;
    p = p * s2 + 0.047619047619047616;
    p = p * s2 + 0.05263157894736842;
    p = p * s2 + 0.058823529411764705;
    p = p * s2 + 0.06666666666666667;
    p = p * s2 + 0.07692307692307693;
    p = p * s2 + 0.09090909090909091;
    p = p * s2 + 0.1111111111111111;
    p = p * s2 + 0.14285714285714285;
    p = p * s2 + 0.2;
    p = p * s2 + 0.3333333333333333;
    p = p * s2 + 1.0;
    const double de = (double)e;
    return de * 0.6931471796706319 + (de * 8.893134238886677e-10 + 2.0 * s * p);
}

// exp of a double: y = k ln 2 + r, |r| <= ln2/2, Taylor series of e^r to
// r^20, scaled by 2^k.
KS_FN double det_exp(double y) {
    const double k = rint(y * 1.4426950408889634);
    const double r = (y - k * 0.6931471796706319) - k * 8.893134238886677e-10;
    double p = 4.110317623312165e-19;  // 1/20!
    p = p * r + 8.22063524662433e-18;
    p = p * r + 1.5619206968586225e-16;
    p = p * r + 2.8114572543455206e-15;
    p = p * r + 4.779477332387385e-14;
    p = p * r + 7.647163731819816e-13;
    p = p * r + 1.1470745597729725e-11;
    p = p * r + 1.6059043836821613e-10;
    p = p * r + 2.08767569878681e-09;
    p = p * r + 2.505210838544172e-08;
    p = p * r + 2.755731922398589e-07;
    p = p * r + 2.7557319223985893e-06;
    p = p * r + 2.48015873015873e-05;
    p = p * r + 0.0001984126984126984;
    p = p * r + 0.001388888888888889;
    p = p * r + 0.008333333333333333;
    p = p * r + 0.041666666666666664;
    p = p * r + 0.16666666666666666;
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    return ldexp(p, (int)k);
}

// pow(b, e) for floats through det_log / det_exp in double, rounded once to
// float.  A negative base needs an integral exponent (else NaN), as powf.
KS_FN float det_powf(float b, float e) {
    if (e == 0.0f) return 1.0f;
    if (b != b || e != e) return NAN;
    bool neg = false;
    if (b < 0.0f) {
        const float ei = rintf(e);
        if (ei != e) return NAN;
        neg = fmodf(ei, 2.0f) != 0.0f;
        b = -b;
    }
    if (b == 0.0f) return e > 0.0f ? (neg ? -0.0f : 0.0f) : (neg ? -INFINITY : INFINITY);
    const float f = (float)det_exp((double)e * det_log((double)b));
    return neg ? -f : f;
}

// Clamped closest parameters (u on segment a, v on segment b) of two segments
// a0-a1 and b0-b1, with the 0.5 fallback for parallel segments.
KS_FN void segment_closest(f3 a0, f3 a1, f3 b0, f3 b1, float& u, float& v) {
    const f3 A0 = a1 - a0, A1 = b1 - b0;
    const float a = dot(A0, A0);
    const float b = dot(A0, A1);
    const float c = dot(A1, A1);
    const f3 diff = b0 - a0;
    float uy = b * dot(A0, diff) - a * dot(A1, diff);
    uy = uy / (a * c - b * b);
    if (!(uy - uy == 0.0f)) uy = 0.5f;  // not finite
    uy = clamp01(uy);
    const f3 pb = b0 + uy * (b1 - b0);
    const float ux = clamp01(dot(pb - a0, A0) / a);
    const f3 pa = a0 + ux * (a1 - a0);
    uy = clamp01(dot(pa - b0, A1) / c);
    u = ux;
    v = uy;
}

KS_FN f3 mix3(f3 a, f3 b, float t) { return a + t * (b - a); }

// Packed symmetric 3x3 (d0, d1, d2, m01, m02, m12)
struct s3 {
    float d[6];
};
KS_FN void sadd(s3& a, const s3& b) {
    for (int k = 0; k < 6; k++) a.d[k] += b.d[k];
}
KS_FN void ssub(s3& a, const s3& b) {
    for (int k = 0; k < 6; k++) a.d[k] -= b.d[k];
}
KS_FN s3 sscale(const s3& a, float s) {
    s3 r;
    for (int k = 0; k < 6; k++) r.d[k] = a.d[k] * s;
    return r;
}
KS_FN s3 souter(f3 v) { return s3{{v.x * v.x, v.y * v.y, v.z * v.z, v.x * v.y, v.x * v.z, v.y * v.z}}; }

// Vertex flag bits (same values as the reference's VertexFlags).
enum : uint32_t { FLAG_PREV = 1u, FLAG_NEXT = 2u, FLAG_NEXT_ORI = 4u, FLAG_FIX_ORI = 8u };

}  // namespace kscr
