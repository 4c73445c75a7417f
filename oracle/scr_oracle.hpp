// scr_oracle.hpp -- CPU restatement of KamiSimulator's Stable Cosserat Rods step.
//
// TEST INFRASTRUCTURE ONLY.  This is the oracle that the HIP product path is
// checked against.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and only as the checker (or as the CPU
// baseline that is timed beside the GPU); the product never calls it.
//
// Semantics: "Jacobi snapshot".  The reference kernels are racy (blocks
// overlap by a halo vertex, q is updated in place), so this restatement fixes
// a deterministic order: every read of iteration k sees iteration k-1 values,
// except that the orientation half-step sees the new dx of the same iteration
// (it is a separate launch after the position kernel, YBSimulator.cu:415-418).
//
// T = float is the FP32 restatement (also the CPU baseline), T = double is the
// FP64 oracle.  The float/double mixing of the host setup code is kept
// expression for expression (cited per function); the device formulas are
// evaluated in T.  Build with -ffp-contract=off so every operation rounds once.
//
// Pinning of each part (tests/test_oracle_*.py):
//   rotor algebra, segment closest points, Morton codes, 3x3 inverse:
//       vectors from the reference's own header-only math (oracle/_ref probe)
//       committed under tests/golden/, plus closed forms;
//   initItr/endItr: closed form of a lone free vertex (symplectic Euler with
//       decay) and the adaptive-initial-guess clamp cases;
//   position solve: rest-state invariance, rigid-motion equivariance, the
//       two-vertex scalar closed form;
//   orientation solve: the rest fixed point, unit norm, q -> -q symmetry;
//   setup: mass/stiffness closed forms and Rotor::fromTo vectors;
//   SDF: affine-field exactness of the half-cell convention, sphere projection;
//   float vs double: the FP64 error is reported per config.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

namespace orc {
inline uint32_t morton3(int x, int y, int z);
}
#include "lbvh_oracle.hpp"

namespace orc {

// ---------------------------------------------------------------- vectors
template <class T>
struct V3 {
    T x, y, z;
};
template <class T>
inline V3<T> v3(T x, T y, T z) {
    return V3<T>{x, y, z};
}
template <class T>
inline V3<T> operator+(V3<T> a, V3<T> b) {
    return {a.x + b.x, a.y + b.y, a.z + b.z};
}
template <class T>
inline V3<T> operator-(V3<T> a, V3<T> b) {
    return {a.x - b.x, a.y - b.y, a.z - b.z};
}
template <class T>
inline V3<T> operator-(V3<T> a) {
    return {-a.x, -a.y, -a.z};
}
template <class T>
inline V3<T> operator*(T s, V3<T> a) {
    return {s * a.x, s * a.y, s * a.z};
}
template <class T>
inline V3<T> operator*(V3<T> a, T s) {
    return {a.x * s, a.y * s, a.z * s};
}
// glm divides componentwise by the scalar (no reciprocal)
template <class T>
inline V3<T> operator/(V3<T> a, T s) {
    return {a.x / s, a.y / s, a.z / s};
}
// The step-time algebra (dot, cross, rotor product and rotation, the 3x3
// solve) is evaluated with fused multiply-adds, one rounding per a*b + c, as
// the reference's CUDA compiler contracts FP32 arithmetic; the grouping is the
// device code's (kamisimulator_amd/csrc/kscr_math.h), so FP32 results agree
// bit for bit.  glm dot (func_geometric.inl) as fma(z, z', fma(y, y', x x')).
template <class T>
inline T dot(V3<T> a, V3<T> b) {
    return std::fma(a.z, b.z, std::fma(a.y, b.y, a.x * b.x));
}
template <class T>
inline V3<T> cross(V3<T> a, V3<T> b) {
    return {std::fma(a.y, b.z, -(b.y * a.z)), std::fma(a.z, b.x, -(b.z * a.x)), std::fma(a.x, b.y, -(b.x * a.y))};
}
template <class T>
inline T length2(V3<T> a) {
    return dot(a, a);
}
template <class T>
inline T length(V3<T> a) {
    return std::sqrt(dot(a, a));
}
// glm normalize = v * inversesqrt(dot(v,v)), inversesqrt = 1/sqrt
template <class T>
inline V3<T> normalize(V3<T> a) {
    return a * (T(1) / std::sqrt(dot(a, a)));
}
// glm mix(x, y, a) = x + a*(y - x)   (func_common.inl:149)
template <class T>
inline V3<T> mix(V3<T> a, V3<T> b, T t) {
    return a + t * (b - a);
}

// Rotor (quaternion) laid out (x, y, z, w), KE/includes/modules/Rotor.h:39-48
template <class T>
struct R4 {
    T x, y, z, w;
    V3<T> q() const { return {x, y, z}; }
};
template <class T>
inline R4<T> r4(V3<T> v, T w) {
    return R4<T>{v.x, v.y, v.z, w};
}
template <class T>
inline T dot4(R4<T> a, R4<T> b) {
    return std::fma(a.w, b.w, std::fma(a.z, b.z, std::fma(a.y, b.y, a.x * b.x)));
}
template <class T>
inline T length4(R4<T> a) {
    return std::sqrt(dot4(a, a));
}
template <class T>
inline R4<T> normalize4(R4<T> a) {
    T s = T(1) / std::sqrt(dot4(a, a));
    return {a.x * s, a.y * s, a.z * s, a.w * s};
}
template <class T>
inline R4<T> add4(R4<T> a, R4<T> b) {
    return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w};
}
template <class T>
inline R4<T> scale4(T s, R4<T> a) {
    return {s * a.x, s * a.y, s * a.z, s * a.w};
}
// Rotor::inverse = conjugate (Rotor.h:109)
template <class T>
inline R4<T> conj(R4<T> a) {
    return {-a.x, -a.y, -a.z, a.w};
}
// Hamilton product (Rotor.h:210-212)
template <class T>
inline R4<T> rmul(R4<T> a, R4<T> b) {
    const V3<T> aq = a.q(), bq = b.q();
    const V3<T> cr = cross(aq, bq);
    return R4<T>{std::fma(a.w, bq.x, std::fma(b.w, aq.x, cr.x)), std::fma(a.w, bq.y, std::fma(b.w, aq.y, cr.y)),
                 std::fma(a.w, bq.z, std::fma(b.w, aq.z, cr.z)), std::fma(a.w, b.w, -dot(aq, bq))};
}
// Rotor::rotate (Rotor.h:114-123)
template <class T>
inline V3<T> rotate(R4<T> r, V3<T> v) {
    const V3<T> q = r.q();
    const V3<T> c1 = cross(q, v);
    const V3<T> a{std::fma(r.w, v.x, c1.x), std::fma(r.w, v.y, c1.y), std::fma(r.w, v.z, c1.z)};
    const T c = dot(v, q);
    const V3<T> c2 = cross(q, a);
    return {std::fma(r.w, a.x, std::fma(c, q.x, c2.x)), std::fma(r.w, a.y, std::fma(c, q.y, c2.y)),
            std::fma(r.w, a.z, std::fma(c, q.z, c2.z))};
}
// rotate(r, e1) with e1's zero and unit factors done exactly: r.w e1 + q x e1
// = (w, q.z, -q.y) and e1.q = q.x, then the general formula (same values up to
// the sign of exact zeros).
template <class T>
inline V3<T> rotate_e1(R4<T> r) {
    const V3<T> q = r.q();
    const V3<T> a{r.w, q.z, -q.y};
    const T c = q.x;
    const V3<T> c2 = cross(q, a);
    return {std::fma(r.w, a.x, std::fma(c, q.x, c2.x)), std::fma(r.w, a.y, std::fma(c, q.y, c2.y)),
            std::fma(r.w, a.z, std::fma(c, q.z, c2.z))};
}
// m (x) (e1, 0) = (m.w, m.z, -m.y, -m.x); (v, 0) (x) b without the products
// by the zero scalar part.
template <class T>
inline R4<T> rmul_e1(R4<T> m) {
    return R4<T>{m.w, m.z, -m.y, -m.x};
}
template <class T>
inline R4<T> rmul_pure(V3<T> v, R4<T> b) {
    const V3<T> bq = b.q();
    const V3<T> cr = cross(v, bq);
    return R4<T>{std::fma(b.w, v.x, cr.x), std::fma(b.w, v.y, cr.y), std::fma(b.w, v.z, cr.z), -dot(v, bq)};
}
// orthoBasisX(n)[1] (KE/includes/modules/Common.h:220-236), used by fromTo for
// antiparallel vectors.
template <class T>
inline V3<T> ortho_basis_x1(V3<T> n) {
    if (n.z >= n.y) {
        const T a = T(1) / (T(1) + n.z);
        const T b = -n.x * n.y * a;
        return {b, T(1) - n.y * n.y * a, -n.y};
    } else {
        const T a = T(1) / (T(1) + n.y);
        return {T(1) - n.x * n.x * a, -n.x, -n.x * n.z * a};
    }
}
// Rotor::fromTo (Rotor.h:79-88)
template <class T>
inline R4<T> from_to(V3<T> from, V3<T> to) {
    V3<T> h = (from + to) * T(0.5f);
    T l = length2(h);
    if (l > 0)
        h = h * (T(1) / std::sqrt(l));
    else
        h = ortho_basis_x1(from);
    return r4(cross(from, h), dot(from, h));
}

// Packed symmetric 3x3 (KE/includes/modules/SymMat.h:6-87):
// d[0..2] = diagonal, d[3] = m01, d[4] = m02, d[5] = m12.
template <class T>
struct Sym3 {
    T d[6];
};
// glm inverse(mat3) (glm/detail/func_matrix.inl:272-294) applied to the
// symmetric matrix, followed by glm mat3*vec3 (type_mat3x3.inl:418-424).
template <class T>
inline V3<T> sym_solve(const Sym3<T>& H, V3<T> f) {
    // m[c][r]; symmetric so m[c][r] = m[r][c]
    const T m00 = H.d[0], m11 = H.d[1], m22 = H.d[2];
    const T m01 = H.d[3], m02 = H.d[4], m12 = H.d[5];
    const T m10 = m01, m20 = m02, m21 = m12;
    const T c00 = std::fma(m11, m22, -(m21 * m12));
    const T c01 = std::fma(m01, m22, -(m21 * m02));
    const T c02 = std::fma(m01, m12, -(m11 * m02));
    const T c10 = std::fma(m10, m22, -(m20 * m12));
    const T c20 = std::fma(m10, m21, -(m20 * m11));
    const T c11 = std::fma(m00, m22, -(m20 * m02));
    const T c21 = std::fma(m00, m21, -(m20 * m01));
    const T c12 = std::fma(m00, m12, -(m10 * m02));
    const T c22 = std::fma(m00, m11, -(m10 * m01));
    const T ood = T(1) / std::fma(m20, c02, std::fma(-m10, c01, m00 * c00));
    const T i00 = c00 * ood, i10 = -c10 * ood, i20 = c20 * ood;
    const T i01 = -c01 * ood, i11 = c11 * ood, i21 = -c21 * ood;
    const T i02 = c02 * ood, i12 = -c12 * ood, i22 = c22 * ood;
    return {std::fma(i20, f.z, std::fma(i10, f.y, i00 * f.x)), std::fma(i21, f.z, std::fma(i11, f.y, i01 * f.x)),
            std::fma(i22, f.z, std::fma(i12, f.y, i02 * f.x))};
}

// Kit::segmentClosestPoints (KE/includes/modules/Common.h:294-312)
template <class T>
inline void segment_closest_points(V3<T> a0, V3<T> a1, V3<T> b0, V3<T> b1, T& u, T& v) {
    V3<T> A0 = a1 - a0, A1 = b1 - b0;
    T a = length2(A0);
    T b = dot(A0, A1);
    T c = length2(A1);
    V3<T> diff = b0 - a0;
    T uy = b * dot(A0, diff) - a * dot(A1, diff);
    uy /= a * c - b * b;
    if (!std::isfinite(uy)) uy = T(0.5f);
    uy = std::min(std::max(uy, T(0)), T(1));
    T ux = std::min(std::max(dot(mix(b0, b1, uy) - a0, A0) / a, T(0)), T(1));
    uy = std::min(std::max(dot(mix(a0, a1, ux) - b0, A1) / c, T(0)), T(1));
    u = ux;
    v = uy;
}

// Kit::mortonExpandBits / getMorton (KE/includes/modules/Common.h:566-579)
inline uint32_t morton_expand_bits(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}
inline uint32_t morton3(int x, int y, int z) {
    return morton_expand_bits((uint32_t)x) * 4 + morton_expand_bits((uint32_t)y) * 2 + morton_expand_bits((uint32_t)z);
}

// Natural logarithm evaluated by a fixed sequence of IEEE float operations
// (frexp-style split d = m 2^e with m in [sqrt(1/2), sqrt(2)), then
// log m = 2 atanh(s), s = (m-1)/(m+1), by its odd series to s^13, plus e ln 2
// with ln 2 split into a high and a low part).  The device code uses the same
// sequence, so the barrier energy rounds identically on both sides.
inline float det_logf(float d) {
    int e = 0;
    float m = std::frexp(d, &e);  // d = m 2^e, m in [0.5, 1)
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
inline double det_logf(double d) { return std::log(d); }

// Wind terms of initItr (YB/sim/iteration.cu:22-25): cos of a double argument
// and pow of floats.  The FP32 restatement evaluates them by fixed sequences
// of IEEE double operations, the same the device uses:
//   cos: x = k pi/2 + r, pi/2 in parts of 30, 30 and 53 bits, |r| <= pi/4,
//        Taylor series of cos r to r^18 / sin r to r^19;
//   log: x = m 2^e, m in [sqrt(1/2), sqrt(2)), 2 atanh((m-1)/(m+1)) to s^25;
//   exp: y = k ln2 + r, Taylor series to r^20, times 2^k;
//   powf(b, e) = (float) exp(e log b), negative bases need integral e.
// Pinned against the C library (tests/test_oracle_primitives.py).  The FP64
// oracle uses std::cos / std::pow.
inline double wind_cos(double x) {
    const double k = std::rint(x * 0.6366197723675814);
    const double r = ((x - k * 1.570796325802803) - k * 9.920935791635221e-10) - k * 5.170182981794105e-19;
    const double r2 = r * r;
    static const double cc[9] = {1.5619206968586225e-16, -4.779477332387385e-14, 1.1470745597729725e-11,
                                 -2.08767569878681e-09,  2.755731922398589e-07,  -2.48015873015873e-05,
                                 0.001388888888888889,   -0.041666666666666664,  0.5};
    static const double sc[9] = {8.22063524662433e-18,  -2.8114572543455206e-15, 7.647163731819816e-13,
                                 -1.6059043836821613e-10, 2.505210838544172e-08,  -2.7557319223985893e-06,
                                 0.0001984126984126984,  -0.008333333333333333,   0.16666666666666666};
    double c = cc[0];
    for (int j = 1; j < 9; j++) c = c * r2 + cc[j];
    c = 1.0 - c * r2;
    double sn = sc[0];
    for (int j = 1; j < 9; j++) sn = sn * r2 + sc[j];
    sn = r - (sn * r2) * r;
    const long long q = (long long)k & 3;
    return q == 0 ? c : (q == 1 ? -sn : (q == 2 ? -c : sn));
}
inline double wind_log(double x) {
    int e = 0;
    double m = std::frexp(x, &e);
    if (m < 0.7071067811865476) {
        m = m * 2.0;
        e = e - 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double s2 = s * s;
    double p = 0.04;
    for (int d = 23; d >= 3; d -= 2) p = p * s2 + 1.0 / (double)d;
    p = p * s2 + 1.0;
    const double de = (double)e;
    return de * 0.6931471796706319 + (de * 8.893134238886677e-10 + 2.0 * s * p);
}
inline double wind_exp(double y) {
    const double k = std::rint(y * 1.4426950408889634);
    const double r = (y - k * 0.6931471796706319) - k * 8.893134238886677e-10;
    static const double inv_fact[21] = {4.110317623312165e-19, 8.22063524662433e-18,   1.5619206968586225e-16,
                                        2.8114572543455206e-15, 4.779477332387385e-14, 7.647163731819816e-13,
                                        1.1470745597729725e-11, 1.6059043836821613e-10, 2.08767569878681e-09,
                                        2.505210838544172e-08,  2.755731922398589e-07,  2.7557319223985893e-06,
                                        2.48015873015873e-05,   0.0001984126984126984,  0.001388888888888889,
                                        0.008333333333333333,   0.041666666666666664,   0.16666666666666666,
                                        0.5,                    1.0,                    1.0};
    double p = inv_fact[0];
    for (int j = 1; j < 21; j++) p = p * r + inv_fact[j];
    return std::ldexp(p, (int)k);
}
inline float wind_powf(float b, float e) {
    if (e == 0.0f) return 1.0f;
    if (std::isnan(b) || std::isnan(e)) return NAN;
    bool neg = false;
    if (b < 0.0f) {
        const float ei = std::rint(e);
        if (ei != e) return NAN;
        neg = std::fmod(ei, 2.0f) != 0.0f;
        b = -b;
    }
    if (b == 0.0f) return e > 0.0f ? (neg ? -0.0f : 0.0f) : (neg ? -INFINITY : INFINITY);
    const float f = (float)wind_exp((double)e * wind_log((double)b));
    return neg ? -f : f;
}
template <class T>
inline T wind_cos_t(double x) {
    return (T)wind_cos(x);
}
template <>
inline double wind_cos_t<double>(double x) {
    return std::cos(x);
}
inline float wind_pow_t(float b, float e) { return wind_powf(b, e); }
inline double wind_pow_t(double b, double e) { return std::pow(b, e); }

// Symmetric 3x3 accumulation helpers (hess3 operators, SymMat.h:31-64)
template <class T>
inline void sym_add(Sym3<T>& a, const Sym3<T>& b) {
    for (int k = 0; k < 6; k++) a.d[k] += b.d[k];
}
template <class T>
inline Sym3<T> sym_outer(V3<T> v) {
    return Sym3<T>{{v.x * v.x, v.y * v.y, v.z * v.z, v.x * v.y, v.x * v.z, v.y * v.z}};
}
template <class T>
inline Sym3<T> sym_scale(const Sym3<T>& a, T s) {
    Sym3<T> r;
    for (int k = 0; k < 6; k++) r.d[k] = a.d[k] * s;
    return r;
}
template <class T>
inline void sym_sub(Sym3<T>& a, const Sym3<T>& b) {
    for (int k = 0; k < 6; k++) a.d[k] -= b.d[k];
}

// ---------------------------------------------------------------- params
// Same field order as kscr_params (include/kscr.h); MetaData in the reference
// (YB/YarnBall.h:61-124).
struct Params {
    float gravity[3];
    int32_t num_itr;
    float max_h;
    float damping;
    float velocity_decay;
    float radius;
    float barrier_thickness;
    float acceleration_ratio;
    float k_collision;
    float k_friction;
    float friction_coeff;
    float detection_scaler;
    float bvh_rebuild_period;
    int32_t detection_period;
    int32_t use_step_size_limit;
    int32_t use_velocity_radius;
    int32_t enable_hair_hair_coll;
    float external_force_multiplier;
    float wind_peak;
    float wind_y_freq;
    float wind_z_freq;
    float wind_time_freq;
    float wind_sharpness;
    int32_t schedule;
    int32_t block_size;
};

enum : uint32_t { F_PREV = 1, F_NEXT = 2, F_NEXT_ORI = 4, F_FIX_ORI = 8 };

static const double kPi = 3.141592653589793238462643;  // YBSimulator.cu:55

struct Sdf {
    bool enabled = false;
    int dim[3] = {0, 0, 0};
    float gmin[3], gmax[3];
    float ds[3], inv_ds[3];
    float thresh = 0;
    float rot[4] = {0, 0, 0, 1};
    float trans[3] = {0, 0, 0};
    std::vector<float> phi;   // dim^3
    std::vector<float> grad;  // 3*dim^3
};

// One contact of segment (xi, xj) with segment (b0, b1), both displaced by
// their current dx (cosseratItr, YB/sim/cosserat.cu:171-227): IPC log barrier
// on the normalised gap d = (|closest-point distance| - 2 radius) / barrier,
// force -b'(d) kCol and Gauss-Newton Hessian b''(d) kCol / barrier along the
// contact normal, b(d) = -(d - 1)^2 log d, split between the segment's two
// vertices by the closest-point parameter; Rayleigh damping; and friction on
// the tangential relative displacement, capped by the Coulomb cone.
// Adds vertex i's terms to (f, H) and vertex i+1's to (f2, H2).
template <class T>
inline void contact_pair(const Params& p, V3<T> xi, V3<T> xj, V3<T> dxi, V3<T> dxj, V3<T> b0, V3<T> b1, V3<T> db0,
                     V3<T> db1, T damp, V3<T>& f, Sym3<T>& H, V3<T>& f2, Sym3<T>& H2) {
    const T invb = T(1) / T(p.barrier_thickness);
    const T radius = T(2) * T(p.radius);
    const T kCol = T(p.k_collision) * invb;
    T ux, uy;

    segment_closest_points(V3<T>{0, 0, 0}, (xj - xi) + (dxj - dxi), (b0 - xi) + (db0 - dxi), (b1 - xi) + (db1 - dxi),
                           ux, uy);
    const V3<T> dpos = mix(xi, xj, ux) - mix(b0, b1, uy);
    const V3<T> ddpos = mix(dxi, dxj, ux) - mix(db0, db1, uy);
    V3<T> nrm = dpos + ddpos;
    T d = length(nrm);
    if (!(d > 0)) return;  // closest points coincide: no contact normal (the reference divides by 0)
    nrm = nrm * (T(1) / d);
    const T wj = ux;
    const T wi = T(1) - ux;
    d = d - radius;
    d = d * invb;
    if (d > 1) return;
    d = std::max(d, T(1e-3f));
    const T invd = T(1) / d;
    const T logd = det_logf(d);
    T dH = (T(-3) + (T(2) + invd) * invd - T(2) * logd) * kCol * invb;
    const T ff = -(T(1) - d) * (d - T(1) + T(2) * d * logd) * invd * kCol;
    f = f + (ff * wi - damp * dH * wi * wi * dot(nrm, dxi)) * nrm;
    f2 = f2 + (ff * wj - damp * dH * wj * wj * dot(nrm, dxj)) * nrm;
    dH = dH * (T(1) + damp);
    Sym3<T> op = sym_outer(nrm);
    sym_add(H, sym_scale(op, dH * wi * wi));
    sym_add(H2, sym_scale(op, dH * wj * wj));
    const V3<T> u = ddpos - dot(nrm, ddpos) * nrm;
    const T ul = length(u);
    if (ul > 0) {
        const T cand = T(p.friction_coeff) * ff / ul;
        const T f1 = cand < T(p.k_friction) ? cand : T(p.k_friction);  // glm::min(fricK, cand)
        op.d[0] -= 1;
        op.d[1] -= 1;
        op.d[2] -= 1;
        f = f - (f1 * wi) * u;
        sym_sub(H, sym_scale(op, (wi * wi) * f1));
        f2 = f2 - (f1 * wj) * u;
        sym_sub(H2, sym_scale(op, (wj * wj) * f1));
    }
    }


// ---------------------------------------------------------------- simulator
template <class T>
struct Sim {
    int n = 0;
    Params p;
    // Host-side state as in YarnBall::Sim (YBSimulator.h:18-47).  Setup quantities
    // are stored in float exactly as the reference computes them; the dynamic
    // state is in T.
    std::vector<float> pos0;  // positions used by configure (float)
    std::vector<int> strandOff;  // first vertex of each strand, and n
    std::vector<float> invMassF, lRestF, kStretchF;
    std::vector<uint32_t> flags;
    // Vertex::connectionIndex (YarnBall.h:56): the vertex this one is glued to,
    // -1 if none (closed curves, reader.cpp:40-44; glueEndpoints,
    // jsonBuilder.cpp:163-188)
    std::vector<int> conn;
    std::vector<R4<float>> qF, qRestF;
    bool configured = false;
    float density = 0, minSegLen = 0, maxSegLen = 0;

    // dynamic state in T
    std::vector<V3<T>> x, v, u;  // pos, vels, lastVels
    std::vector<R4<T>> q;
    std::vector<R4<T>> qRest;
    std::vector<T> invMass, lRest, kStretch;
    // per-substep scratch
    std::vector<V3<T>> y, dx, dxNew;
    std::vector<R4<T>> qNew;

    // hair-hair contacts (YB/sim/collision.cu, YBSimulator.cu:424-470)
    static const int kMaxCols = 16;  // MAX_COLLISIONS_PER_SEGMENT (YarnBall.h:19)
    std::vector<std::vector<int>> cols;  // per segment, sorted ascending
    std::vector<T> maxStep;
    Lbvh bvh;
    float lastBVHRebuild = std::numeric_limits<float>::infinity();
    int error = 0;     // 8 = more than kMaxCols contacts on a segment
    long num_pairs = 0;  // broad-phase pairs of the last detection

    T h = T(4e-4f), lastH = T(4e-4f), time = 0;
    float hF = 4e-4f, lastHF = 4e-4f, timeF = 0;
    long step_counter = 0;
    Sdf sdf;
    int threads = 1;

    // Sim::Sim + createFromCurves (YBSimulator.cu:81-127; YB/io/reader.cpp:16-50)
    int create(const Params& prm, int nv, const float* P, int ns, const int* off) {
        if (nv < 3) return 2;
        n = nv;
        p = prm;
        pos0.assign(P, P + 3 * (size_t)nv);
        strandOff.assign(off, off + ns + 1);
        invMassF.assign(nv, 1.f);
        lRestF.assign(nv, 1.f);
        kStretchF.assign(nv, 100.f);
        flags.assign(nv, F_NEXT);
        conn.assign(nv, -1);
        qF.assign(nv, R4<float>{0, 0, 0, 1});
        qRestF.assign(nv, R4<float>{0, 0, 0, 1});
        flags[nv - 1] = 0;
        for (int s = 0; s < ns; s++) flags[off[s + 1] - 1] = 0;  // cut off the end of each curve
        x.resize(nv);
        v.assign(nv, V3<T>{0, 0, 0});
        u.assign(nv, V3<T>{0, 0, 0});
        for (int i = 0; i < nv; i++) x[i] = V3<T>{(T)P[3 * i], (T)P[3 * i + 1], (T)P[3 * i + 2]};
        h = T(prm.max_h);
        hF = prm.max_h;
        return 0;
    }

    // Sim::configure (YBSimulator.cu:161-241), float arithmetic with the double
    // sub-expression of the lumped mass (:229).
    int configure(float dens) {
        density = dens;
        maxSegLen = 0;
        minSegLen = std::numeric_limits<float>::max();
        R4<float> lastQ{0, 0, 0, 1};
        for (int i = 0; i < n; i++) {
            if (i < n - 1) {
                bool hasNext = (flags[i] & F_NEXT) != 0;
                flags[i + 1] = hasNext ? (flags[i + 1] | F_PREV) : (flags[i + 1] & ~F_PREV);
                bool hasNextNext = hasNext && (flags[i + 1] & F_NEXT);
                flags[i] = hasNextNext ? (flags[i] | F_NEXT_ORI) : (flags[i] & ~F_NEXT_ORI);
                if (!hasNext) flags[i] |= F_FIX_ORI;
            }
            if (!(flags[i] & F_PREV) && !(flags[i] & F_NEXT)) return 3;
            lRestF[i] = 1.f / n;
            float mass = 0;
            if (flags[i] & F_PREV) mass += lRestF[i - 1];
            if (flags[i] & F_NEXT) {
                V3<float> a{pos0[3 * i], pos0[3 * i + 1], pos0[3 * i + 2]};
                V3<float> b{pos0[3 * i + 3], pos0[3 * i + 4], pos0[3 * i + 5]};
                V3<float> seg = b - a;
                lRestF[i] = length(seg);
                if (lRestF[i] == 0 || !std::isfinite(lRestF[i])) return 4;
                lastQ = from_to(V3<float>{1, 0, 0}, normalize(seg));
                mass += lRestF[i];
                maxSegLen = std::max(maxSegLen, lRestF[i]);
                minSegLen = std::min(minSegLen, lRestF[i]);
            }
            qF[i] = lastQ;
            mass = (float)((double)mass * ((double)(p.radius * p.radius) * kPi * (double)dens * (double)0.5f));
            if (mass != 0)
                invMassF[i] *= 1 / mass;
            else
                invMassF[i] = 0;
        }
        for (int i = 0; i < n - 1; i++) qRestF[i] = normalize4(rmul(conj(qF[i]), qF[i + 1]));
        configured = true;
        return 0;
    }

    // Sim::setYoung (YBSimulator.cu:300-320)
    int set_young(float young, float deviation) {
        if (!configured) return 7;
        float mmrad = p.radius * 1000;
        for (int i = 0; i < n; i++) {
            float len = lRestF[i];
            float ss = (float)((double)young * kPi * (double)mmrad * (double)mmrad);
            float bt = (float)((double)(ss * deviation * mmrad * mmrad) * 0.25 / (double)(len * len));
            kStretchF[i] = ss;
            qRestF[i] = R4<float>{qRestF[i].x * bt, qRestF[i].y * bt, qRestF[i].z * bt, qRestF[i].w * bt};
        }
        return 0;
    }
    // Sim::setKStretch (YBSimulator.cu:276-284)
    int set_kstretch(float k) {
        if (!configured) return 7;
        for (int i = 0; i < n; i++) kStretchF[i] = k * lRestF[i];
        return 0;
    }
    // Sim::setKBend (YBSimulator.cu:286-298)
    int set_kbend(float k) {
        if (!configured) return 7;
        k *= 4;
        for (int i = 0; i < n; i++) {
            float s = k / lRestF[i];
            R4<float> nq = normalize4(qRestF[i]);
            qRestF[i] = R4<float>{s * nq.x, s * nq.y, s * nq.z, s * nq.w};
        }
        return 0;
    }
    // Per-vertex stretch stiffness and rest rotors (per-strand materials).
    // Per-vertex materials (kscr_set_material): the host mirror, and the live
    // state too once uploaded.
    void set_material(const float* ks, const float* qr) {
        const bool live = (int)kStretch.size() == n;
        for (int i = 0; i < n; i++) {
            if (ks) kStretchF[i] = ks[i];
            if (qr) qRestF[i] = R4<float>{qr[4 * i], qr[4 * i + 1], qr[4 * i + 2], qr[4 * i + 3]};
            if (live && ks) kStretch[i] = (T)kStretchF[i];
            if (live && qr) qRest[i] = R4<T>{(T)qRestF[i].x, (T)qRestF[i].y, (T)qRestF[i].z, (T)qRestF[i].w};
        }
    }
    // Same, in the state's precision (after upload; for the FP64 gradient checks).
    void set_material_t(const double* ks, const double* qr) {
        for (int i = 0; i < n; i++) {
            if (ks) kStretch[i] = (T)ks[i];
            if (qr) qRest[i] = R4<T>{(T)qr[4 * i], (T)qr[4 * i + 1], (T)qr[4 * i + 2], (T)qr[4 * i + 3]};
        }
    }
    void pin(const int* ids, int m) {
        for (int k = 0; k < m; k++) invMassF[ids[k]] = 0;
    }
    void fix_orientation(const int* ids, int m) {
        for (int k = 0; k < m; k++) flags[ids[k]] |= F_FIX_ORI;
    }

    // ------------------------------------------------------------ SagFree
    // Sim::SagFree / SingleSagFree (YBSimulator.cu:480-676): host code of the
    // reference, run between setYoung/root pinning and upload
    // (StableCosseratRodsSolver.cpp:203-209).  Float and double mixed as
    // written there (glm host arithmetic: no FMA; glm's dot/length/normalize
    // forms), C library sqrt/acos/sinf/cosf, fitAngle's long-double fmod.
    // Pass 1 (tip to root) shortens each rest length so that the stretch force
    // carries the weight below the segment; pass 2 (root to tip) rotates the
    // orientations and rest rotors so that bending carries the torque of the
    // weight.
    struct SfV {
        float x, y, z;
    };
    static SfV sf_add(SfV a, SfV b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
    static SfV sf_sub(SfV a, SfV b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
    static SfV sf_mul(float s, SfV a) { return {s * a.x, s * a.y, s * a.z}; }
    static float sf_dot(SfV a, SfV b) { return a.x * b.x + a.y * b.y + a.z * b.z; }  // glm: tmp.x + tmp.y + tmp.z
    static SfV sf_cross(SfV a, SfV b) { return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
    static float sf_length(SfV a) { return std::sqrt(sf_dot(a, a)); }
    static SfV sf_normalize(SfV a) { return sf_mul(1.0f / std::sqrt(sf_dot(a, a)), a); }
    static float sf_dot4(R4<float> a, R4<float> b) { return (a.x * b.x + a.y * b.y) + (a.z * b.z + a.w * b.w); }
    static SfV sf_q(R4<float> r) { return {r.x, r.y, r.z}; }
    // Rotor product (Rotor.h:210-212) and rotate (Rotor.h:114-123)
    static R4<float> sf_rmul(R4<float> a, R4<float> b) {
        const SfV v = sf_add(sf_add(sf_mul(a.w, sf_q(b)), sf_mul(b.w, sf_q(a))), sf_cross(sf_q(a), sf_q(b)));
        return {v.x, v.y, v.z, a.w * b.w - sf_dot(sf_q(a), sf_q(b))};
    }
    static SfV sf_rotate(R4<float> r, SfV v) {
        const SfV a = sf_add(sf_mul(r.w, v), sf_cross(sf_q(r), v));
        const float c = sf_dot(v, sf_q(r));
        return sf_add(sf_add(sf_mul(r.w, a), sf_cross(sf_q(r), a)), sf_mul(c, sf_q(r)));
    }
    static SfV sf_d3(R4<float> r) { return sf_normalize(sf_rotate(r, SfV{1, 0, 0})); }  // getD3 (Rotor.h:206)
    // fitAngle (YBSimulator.cu:59-78): wrap into [-pi, pi) through a long-double fmod
    static double sf_fmod(double x, double y) {
        const long double lx = x, ly = y;
        if (ly == 0) return 0;
        return (double)(lx - (long long)(lx / ly) * ly);
    }
    static double sf_fit_angle(double a) {
        const double pi = 3.141592653589793238462643, two_pi = 6.2831853071795864769252867665590057683943;
        return sf_fmod(sf_fmod(a + pi, two_pi) + two_pi, two_pi) - pi;
    }
    static double sf_sign(double x) { return (double)((0 < x) - (x < 0)); }

    void sag_free(const int* roots, int num_roots, int iterations) {
        for (int it = 0; it < iterations; it++) single_sag_free(roots, num_roots);
    }
    void single_sag_free(const int* roots, int num_roots) {
        const SfV g{p.gravity[0], p.gravity[1], p.gravity[2]};
        auto pos = [&](int i) { return SfV{pos0[3 * i], pos0[3 * i + 1], pos0[3 * i + 2]}; };
        for (int s = 0; s < num_roots; s++) {
            const float eps = 1e-6f;
            const int rootv = roots[s];
            const int nextv = (s == num_roots - 1) ? n : roots[s + 1];
            std::vector<SfV> f_srs((size_t)(nextv - rootv), SfV{0, 0, 0});

            // 1. rest lengths against the stretch force
            float massFromTip = 1 / invMassF[nextv - 1];
            for (int v = 0; rootv + v < nextv - 1; ++v) {
                const int vId = nextv - 2 - v;
                const SfV axis = sf_d3(qF[vId]);
                const float kSt = kStretchF[vId];
                const float oldRest = lRestF[vId];
                if (invMassF[vId] != 0)
                    massFromTip = (float)((double)massFromTip + 0.5 / (double)invMassF[vId]);
                else
                    massFromTip = (float)((double)massFromTip +
                                          0.5 * (double)(lRestF[vId] / lRestF[vId + 1]) / (double)invMassF[vId + 1]);
                const double f_s_g = (double)(sf_dot(axis, g) * massFromTip);
                const double f_s = -f_s_g;
                float rest;
                if (f_s == 0) {
                    rest = oldRest;
                } else if (1 - 4 * f_s * (double)oldRest / (double)kSt < 0) {
                    rest = 1.5f * oldRest;
                } else {
                    const double fsk = f_s / (double)kSt;
                    const double fskinv = (double)kSt / f_s;
                    const double disc = std::sqrt(1 - (double)(4 * oldRest) * fsk);
                    const double ans1 = (1 + disc) * 0.5 * fskinv;
                    const double ans2 = (1 - disc) * 0.5 * fskinv;
                    rest = (float)(std::fabs((double)oldRest - ans1) >= std::fabs((double)oldRest - ans2) ? ans2 : ans1);
                }
                rest = std::min(std::max(rest, eps), 1.5f * oldRest);
                lRestF[vId] = rest;
                if (rootv + v < nextv - 2) f_srs[vId - rootv - 1] = sf_mul((float)-f_s, axis);
                massFromTip = (float)((double)massFromTip + 0.5 / (double)invMassF[vId]);
            }

            // 2. orientations and rest rotors against the bending torque
            const float nearest = 1e-4f;
            float massToTip = 0;
            for (int v = 1; rootv + v < nextv; ++v) {
                if (invMassF[rootv + v] != 0)
                    massToTip += 1 / invMassF[rootv + v];
                else
                    massToTip += (lRestF[rootv + v] / lRestF[rootv + v - 1]) / invMassF[rootv + v - 1];
            }
            R4<float> accum{0, 0, 0, 1};
            for (int v = 1; rootv + v < nextv - 1; ++v) {
                const int vId = rootv + v;
                const R4<float> cQk = sf_rmul(accum, qF[vId]);
                if (invMassF[vId] != 0)
                    massToTip = (float)((double)massToTip - 0.5 / (double)invMassF[vId]);
                else
                    massToTip = (float)((double)massToTip -
                                        0.5 * (double)(lRestF[vId] / lRestF[vId + 1]) / (double)invMassF[vId + 1]);
                const SfV axisk = sf_d3(cQk);
                const SfV f_b = sf_add(f_srs[v], sf_mul(massToTip, g));
                if (sf_dot(f_b, axisk) < 1 - nearest) {
                    const float kBend = std::sqrt(sf_dot4(qRestF[vId], qRestF[vId]));
                    const SfV planeNormal = sf_normalize(sf_cross(axisk, f_b));
                    const SfV torqueDir = sf_cross(planeNormal, axisk);
                    const float torque = sf_dot(torqueDir, f_b) * sf_length(sf_sub(pos(vId + 1), pos(vId)));
                    const float Tk = torque / kBend;
                    const float lfb = sf_length(f_b);
                    const SfV fbn = {f_b.x / lfb, f_b.y / lfb, f_b.z / lfb};
                    const double d = ((double)axisk.x * -(double)fbn.x + (double)axisk.y * -(double)fbn.y) +
                                     (double)axisk.z * -(double)fbn.z;
                    const double maxrad = std::acos(d);
                    const double maxTheta1 = sf_fit_angle(maxrad + (double)nearest);
                    const double maxTheta2 = sf_fit_angle(maxrad - (double)nearest);
                    const double maxTheta = std::fabs(maxTheta1) >= std::fabs(maxTheta2) ? maxTheta2 : maxTheta1;
                    const float left = maxTheta >= 0 ? 0.f : (float)maxTheta;
                    const float right = maxTheta >= 0 ? (float)maxTheta : 0.f;
                    double resTheta = 0;
                    if (std::fabs(Tk) > 2) {
                        resTheta = maxTheta;
                    } else {
                        const double half_pi = 1.5707963267948966192313216916397;
                        double theta1 = 2 * std::acos(1 - (0.5 * (double)Tk * (double)Tk));
                        double theta2 = 2 * std::acos(-1 + (0.5 * (double)Tk * (double)Tk));
                        theta1 = sf_fit_angle(theta1);
                        theta2 = sf_fit_angle(theta2);
                        double cT1 = -sf_sign(maxTheta) * 100, cT2 = -sf_sign(maxTheta) * 100;
                        if (std::fabs(theta1) <= half_pi) cT1 = theta1;
                        if (std::fabs(theta2) > half_pi) cT2 = theta2;
                        if (sf_sign(cT1) == sf_sign(maxTheta) && sf_sign(cT2) == sf_sign(maxTheta)) {
                            resTheta = std::min(cT1, cT2);
                        } else {
                            if (sf_sign(cT1) == sf_sign(maxTheta)) resTheta = cT1;
                            if (sf_sign(cT2) == sf_sign(maxTheta)) resTheta = cT2;
                        }
                    }
                    resTheta = std::min(std::max(resTheta, (double)left), (double)right);
                    // Rotor::angleAxis (Rotor.h:51-54) in float
                    float rad = (float)resTheta;
                    rad = (float)(rad * 0.5);
                    const float sn = std::sin(rad), cs = std::cos(rad);
                    const R4<float> qTheta{sn * planeNormal.x, sn * planeNormal.y, sn * planeNormal.z, cs};
                    accum = sf_rmul(accum, qTheta);
                    qF[vId] = sf_rmul(accum, qF[vId]);
                    const R4<float> rel = sf_rmul(R4<float>{-qF[vId - 1].x, -qF[vId - 1].y, -qF[vId - 1].z, qF[vId - 1].w}, qF[vId]);
                    const float inv = 1.0f / std::sqrt(sf_dot4(rel, rel));
                    const float len = std::sqrt(sf_dot4(qRestF[vId - 1], qRestF[vId - 1]));
                    qRestF[vId - 1] = R4<float>{rel.x * inv * len, rel.y * inv * len, rel.z * inv * len, rel.w * inv * len};
                }
                massToTip = (float)((double)massToTip - 0.5 / (double)invMassF[vId]);
            }
        }
    }

    // uploadMeta validation (YBSimulator.cu:322-335)
    int validate() const {
        if (p.detection_scaler <= 1) return 5;
        if (2 * minSegLen < 2 * p.radius + p.barrier_thickness) return 6;
        return 0;
    }

    // Sim::upload: host mirror -> state in T
    void upload() {
        q.resize(n);
        qRest.resize(n);
        invMass.resize(n);
        lRest.resize(n);
        kStretch.resize(n);
        for (int i = 0; i < n; i++) {
            q[i] = R4<T>{(T)qF[i].x, (T)qF[i].y, (T)qF[i].z, (T)qF[i].w};
            qRest[i] = R4<T>{(T)qRestF[i].x, (T)qRestF[i].y, (T)qRestF[i].z, (T)qRestF[i].w};
            invMass[i] = (T)invMassF[i];
            lRest[i] = (T)lRestF[i];
            kStretch[i] = (T)kStretchF[i];
        }
        y.assign(n, V3<T>{0, 0, 0});
        cols.assign(n, std::vector<int>());
        maxStep.assign(n, T(1));
        dx.assign(n, V3<T>{0, 0, 0});
        dxNew.assign(n, V3<T>{0, 0, 0});
        qNew.assign(n, R4<T>{0, 0, 0, 1});
    }

    // ------------------------------------------------------------ kernels
    // initItr (YB/sim/iteration.cu:8-59)
    // Branch of the adaptive initial guess's clamp per vertex and substep
    // (appended by init_itr while log_clamp is set): 0 = sigma clamped to 0,
    // 1 = interior, 2 = clamped to 1, 3 = no guess (pinned or no acceleration).
    bool log_clamp = false;
    std::vector<uint8_t> clamp_log;

    void init_itr() {
        const T hh = h;
        const V3<T> g{(T)p.gravity[0], (T)p.gravity[1], (T)p.gravity[2]};
        const size_t log0 = clamp_log.size();
        if (log_clamp) clamp_log.resize(log0 + n, 3);
#pragma omp parallel for schedule(static) num_threads(threads)
        for (int i = 0; i < n; i++) {
            const V3<T> vel = v[i];
            T windx = 0;
            if (p.wind_peak != 0) {
                // cos in double of a float argument (M_PI is double), pow in T
                T yp = x[i].y, zp = x[i].z;
                T wy = yp > 0 ? wind_cos_t<T>(2 * M_PI * (double)((T)p.wind_y_freq * yp + (T)p.wind_time_freq * time)) : T(0);
                T wz = wind_cos_t<T>(2 * M_PI * (double)((T)p.wind_z_freq * std::fabs(zp) + (T)p.wind_time_freq * time));
                T yz = wy * wz;
                T sg = (T)((T(0) < yz) - (yz < T(0)));
                T pw = wind_pow_t(yz, (T)p.wind_sharpness) * sg;
                T mx = std::isnan(pw) ? T(0) : std::max(T(0), pw);  // fmax(0, NaN) = 0
                windx = (T)p.wind_peak * mx;
            }
            V3<T> ext = (g + V3<T>{windx, 0, 0}) * (T)p.external_force_multiplier;
            V3<T> d = hh * vel;
            V3<T> lastVel = u[i];
            u[i] = vel;
            if (invMass[i] != 0) {
                y[i] = d + (hh * hh) * ext;
                T ext2 = length2(ext);
                if (ext2 > 0) {
                    V3<T> a = (vel - lastVel) / lastH;
                    const T raw = dot(a, ext) / ext2;
                    if (log_clamp) clamp_log[log0 + i] = raw <= T(0) ? 0 : (raw >= T(1) ? 2 : 1);
                    T s = std::min(std::max(raw, T(0)), T(1));
                    d = d + (hh * hh * s) * ext;
                }
            } else {
                y[i] = vel;  // the reference leaves vels untouched for pinned vertices
            }
            dx[i] = d;
        }
    }

    // Position solve (YB/sim/cosserat.cu:112-287), Jacobi: every read sees the
    // iteration-k dx and q.  The reference splits a vertex's work into two
    // "sectors": sector 0 starts from the inertia terms and adds the stretch
    // term of segment i and the contacts at even list positions; sector 1 sums
    // the contacts at odd positions from zero; then sector 1 is added to
    // sector 0, and finally the contribution of segment i-1 (computed the same
    // way by vertex i-1) is added.  The summation order below is that order.
    struct SegSums {
        V3<T> f0, f1, f2a, f2b;  // own sector 0 / sector 1; upper vertex sector 0 / 1
        Sym3<T> H0, H1, H2a, H2b;
        T limit;
    };

    // One contact of segment i with segment c (cosserat.cu:171-227).
    void contact_terms(int i, int c, T damp, V3<T>& f, Sym3<T>& H, V3<T>& f2, Sym3<T>& H2) const {
        contact_pair<T>(p, x[i], x[i + 1], dx[i], dx[i + 1], x[c], x[c + 1], dx[c], dx[c + 1], damp, f, H, f2, H2);
    }
    // Substep preamble only (initItr, a forced detection, step limits), for
    // the contact tests: x is unchanged, dx holds the initial guess.
    void detect_only(float dtF) {
        const int steps = std::max(1, (int)std::ceil(dtF / p.max_h));
        lastHF = hF;
        hF = dtF / steps;
        lastH = (T)lastHF;
        h = (T)hF;
        init_itr();
        lastBVHRebuild = std::numeric_limits<T>::infinity();
        detect_collisions();
        recompute_step_limit();
    }

    // Everything segment i (between vertices i and i+1) adds, sector by sector.
    // f0/H0 enter holding vertex i's inertia terms.
    void segment_sums(int i, T damp, SegSums& s) const {
        s.f1 = V3<T>{0, 0, 0};
        s.f2a = V3<T>{0, 0, 0};
        s.f2b = V3<T>{0, 0, 0};
        s.H1 = Sym3<T>{{0, 0, 0, 0, 0, 0}};
        s.H2a = Sym3<T>{{0, 0, 0, 0, 0, 0}};
        s.H2b = Sym3<T>{{0, 0, 0, 0, 0, 0}};
        s.limit = std::numeric_limits<T>::infinity();
        if (!(flags[i] & F_NEXT)) return;
        s.limit = maxStep[i];
        const T invl = T(1) / lRest[i];
        const V3<T> c = ((x[i + 1] - x[i]) + (dx[i + 1] - dx[i])) * invl - rotate_e1(q[i]);
        const T k = kStretch[i] * invl;
        T d = k * invl;
        s.f0 = s.f0 + (k * c - (damp * d) * dx[i]);
        s.f2a = s.f2a + ((-k) * c - (damp * d) * dx[i + 1]);
        d = d * (1 + damp);
        s.H0.d[0] += d;
        s.H0.d[1] += d;
        s.H0.d[2] += d;
        s.H2a.d[0] += d;
        s.H2a.d[1] += d;
        s.H2a.d[2] += d;
        const std::vector<int>& L = cols[i];
        for (size_t k2 = 0; k2 < L.size(); k2 += 2) contact_terms(i, L[k2], damp, s.f0, s.H0, s.f2a, s.H2a);
        for (size_t k2 = 1; k2 < L.size(); k2 += 2) contact_terms(i, L[k2], damp, s.f1, s.H1, s.f2b, s.H2b);
    }

    V3<T> position_solve(int i) const {
        if (invMass[i] == 0) return dx[i];
        const T hh = h;
        const T damp = T(p.damping) / hh;
        const V3<T> dxi = dx[i];
        const T a_in = T(1) / (invMass[i] * hh * hh);
        SegSums s;
        s.H0 = Sym3<T>{{a_in, a_in, a_in, a_in, a_in, a_in}};
        s.f0 = (T(1) / (hh * hh * invMass[i])) * (y[i] - dxi);
        if (conn[i] >= 0) {
            // the glue to the connected vertex (cosserat.cu:123-130)
            const int j = conn[i];
            const T stiffness = T(40.0f);
            s.f0 = s.f0 - stiffness * (((x[i] - x[j]) + (dxi - dx[j])) + damp * dxi);
            const T g = stiffness * (1 + damp);
            s.H0.d[0] += g;
            s.H0.d[1] += g;
            s.H0.d[2] += g;
        }
        segment_sums(i, damp, s);
        V3<T> f = s.f0 + s.f1;
        Sym3<T> H = s.H0;
        sym_add(H, s.H1);
        T limit = s.limit;
        if (flags[i] & F_PREV) {
            SegSums t;
            t.f0 = V3<T>{0, 0, 0};
            t.H0 = Sym3<T>{{0, 0, 0, 0, 0, 0}};
            segment_sums(i - 1, damp, t);
            const V3<T> f2 = t.f2a + t.f2b;
            Sym3<T> H2 = t.H2a;
            sym_add(H2, t.H2b);
            limit = std::min(limit, t.limit);
            f = f + f2;
            sym_add(H, H2);
        }
        V3<T> out = dxi + T(p.acceleration_ratio) * sym_solve(H, f);
        // A contact Hessian can make H numerically singular in FP32 (its
        // adjugate determinant cancels to 0): the reference's step is then
        // inf/NaN, which no step limit removes.  Such a step is skipped.
        if (!std::isfinite(out.x) || !std::isfinite(out.y) || !std::isfinite(out.z)) return dxi;
        const T l = length(out);
        if (l > limit && l > 0) out = out * (limit / l);
        return out;
    }

    // ------------------------------------------------------------ contacts
    // buildAABBs<USE_VEL_RADIUS> (YB/sim/collision.cu:7-34), in float.
    std::vector<Aabb> segment_aabbs() const {
        const float detR = p.radius + 0.5f * p.barrier_thickness;
        const float scaled = detR * p.detection_scaler;
        std::vector<Aabb> out(n, aabb_empty());
        for (int i = 0; i < n; i++) {
            if (!(flags[i] & F_NEXT)) continue;
            const float p0[3] = {(float)x[i].x, (float)x[i].y, (float)x[i].z};
            const float p1[3] = {(float)x[i + 1].x, (float)x[i + 1].y, (float)x[i + 1].z};
            Aabb a = aabb_empty();
            aabb_absorb(a, p0);
            aabb_absorb(a, p1);
            float pad = scaled;
            if (p.use_velocity_radius) {
                const V3<float> d0{(float)dx[i].x, (float)dx[i].y, (float)dx[i].z};
                const V3<float> d1{(float)dx[i + 1].x, (float)dx[i + 1].y, (float)dx[i + 1].z};
                pad += std::sqrt(std::max(length2(d0), length2(d1)));
            }
            for (int k = 0; k < 3; k++) {
                a.mn[k] -= pad;
                a.mx[k] += pad;
            }
            out[i] = a;
        }
        return out;
    }

    // Sim::detectCollisions (YBSimulator.cu:424-457) with the reference's two
    // bugs fixed (bvh.reset() nulling the tree; the pair buffer sized in ints)
    // and deterministic, ascending per-segment contact lists.
    void detect_collisions() {
        const std::vector<Aabb> boxes = segment_aabbs();
        if (lastBVHRebuild >= p.bvh_rebuild_period) {
            bvh.build(boxes);
            lastBVHRebuild = 0;
        } else {
            bvh.refit(boxes);
            lastBVHRebuild += hF * p.detection_period;
        }
        const auto pairs = bvh.query(boxes);
        num_pairs = (long)pairs.size();
        const float detR = p.radius + 0.5f * p.barrier_thickness;
        const T scaled = T(detR * p.detection_scaler);
        for (auto& L : cols) L.clear();
        for (const auto& pr : pairs) {
            const int a = pr.first, b = pr.second;
            // glued segments (collision.cu:44-49)
            const int c0 = conn[a], c1 = conn[a + 1];
            if (c0 == b || c1 == b || c0 == b + 1 || c1 == b + 1) continue;
            if (std::abs(b - a) <= 2) continue;  // neighbouring segments
            const V3<T> a0 = x[a];
            const V3<T> a1 = x[a + 1] - a0;
            const V3<T> b0 = x[b] - a0;
            const V3<T> b1 = x[b + 1] - a0;
            T u, v;
            segment_closest_points(V3<T>{0, 0, 0}, a1, b0, b1, u, v);
            const V3<T> nrm = u * a1 - mix(b0, b1, v);
            const T d2 = length2(nrm);
            T r = T(2) * scaled;
            if (p.use_velocity_radius) {
                r += std::sqrt(std::max(length2(dx[a]), length2(dx[a + 1])));
                r += std::sqrt(std::max(length2(dx[b]), length2(dx[b + 1])));
            }
            if (d2 < r * r) {
                if ((int)cols[a].size() >= kMaxCols || (int)cols[b].size() >= kMaxCols) error = 8;
                if ((int)cols[a].size() < kMaxCols) cols[a].push_back(b);
                if ((int)cols[b].size() < kMaxCols) cols[b].push_back(a);
            }
        }
        for (auto& L : cols) std::sort(L.begin(), L.end());
    }

    // recomputeStepLimitKernel<LIMIT, USE_VEL_RADIUS> (YB/sim/collision.cu:94-131)
    void recompute_step_limit() {
        const T inf = std::numeric_limits<T>::infinity();
        const T detR = T(p.radius + 0.5f * p.barrier_thickness);
        const T margin = T((1 - 0.1f) * 0.5f);
        for (int i = 0; i < n; i++) {
            T m = inf;
            if (p.use_step_size_limit && (flags[i] & F_NEXT)) {
                const V3<T> a0 = x[i];
                const V3<T> a1 = x[i + 1] - a0;
                m = detR * (T(p.detection_scaler) - 1);
                if (p.use_velocity_radius) m += length(dx[i]);
                for (int c : cols[i]) {
                    const V3<T> b0 = x[c] - a0;
                    const V3<T> b1 = x[c + 1] - a0;
                    T u, v;
                    segment_closest_points(V3<T>{0, 0, 0}, a1, b0, b1, u, v);
                    const V3<T> nrm = u * a1 - mix(b0, b1, v);
                    m = std::min(m, margin * length(nrm));
                }
            }
            maxStep[i] = m;
        }
    }

    // Orientation solve of vertex i (YB/sim/cosserat.cu:290-337), Jacobi: reads
    // the new dx and the previous-iteration q of the neighbours.
    R4<T> orientation_solve(int i) const {
        const uint32_t fl = flags[i];
        if ((fl & F_FIX_ORI) || !(fl & F_NEXT)) return q[i];
        V3<T> vv = ((x[i + 1] - x[i]) + (dxNew[i + 1] - dxNew[i])) / lRest[i];
        vv = vv * (T(-2) * kStretch[i]);
        R4<T> b{0, 0, 0, 0};
        const R4<T> qi = q[i];
        // phi = sign <conj(q_{i-1}) q_i, r_{i-1}> (cosserat.cu:320) is
        // sign <q_i, q_{i-1} r_{i-1}>, and sign <conj(q_i) q_{i+1}, r_i>
        // (:328) is sign <q_i, q_{i+1} conj(r_i)>: <a, b> = Re(a conj(b)) and
        // Re(ABC) = Re(BCA).  Both rotors are the terms of b.
        if (fl & F_PREV) {
            const R4<T> bm = rmul(q[i - 1], qRest[i - 1]);
            T phi = dot4(qi, bm) > 0 ? T(1) : T(-1);
            b = add4(b, scale4(phi, bm));
        }
        if (fl & F_NEXT_ORI) {
            const R4<T> bp = rmul(q[i + 1], conj(qRest[i]));
            T phi = dot4(qi, bp) > 0 ? T(1) : T(-1);
            b = add4(b, scale4(phi, bp));
        }
        T lambda = length(vv) + length4(b);
        R4<T> t = rmul_e1(rmul_pure(vv, b));
        return normalize4(add4(t, scale4(lambda, b)));
    }

    // endItr (YB/sim/iteration.cu:62-74)
    void end_itr() {
        const T invH = T(1) / h;
#pragma omp parallel for schedule(static) num_threads(threads)
        for (int i = 0; i < n; i++) {
            if (invMass[i] != 0) v[i] = dx[i] * invH * (T(1) - T(p.velocity_decay) * h);
            x[i] = x[i] + dx[i];
        }
    }

    // One VBD iteration: cosseratItr then quaternionLambdaItr
    // (YBSimulator.cu:415-418) with snapshot reads.
    void iterate() {
        if (p.schedule == 1) {
            iterate_colored();
            return;
        }
#pragma omp parallel for schedule(static) num_threads(threads)
        for (int i = 0; i < n; i++) dxNew[i] = position_solve(i);
#pragma omp parallel for schedule(static) num_threads(threads)
        for (int i = 0; i < n; i++) qNew[i] = orientation_solve(i);
        std::swap(dx, dxNew);
        std::swap(q, qNew);
    }

    // Even/odd coloured iteration (no reference counterpart): vertices and
    // segments coloured by the parity of their global index; each colour is
    // solved with the values the previous colour left (Gauss-Seidel over the
    // two colours), using the same per-vertex updates as the Jacobi step.
    void iterate_colored() {
        for (int c = 0; c < 2; c++) {
#pragma omp parallel for schedule(static) num_threads(threads)
            for (int i = c; i < n; i += 2) dxNew[i] = position_solve(i);
#pragma omp parallel for schedule(static) num_threads(threads)
            for (int i = c; i < n; i += 2) dx[i] = dxNew[i];
        }
        // dxNew holds the new dx of every vertex: the orientation solve reads it
        for (int c = 0; c < 2; c++) {
#pragma omp parallel for schedule(static) num_threads(threads)
            for (int i = c; i < n; i += 2) qNew[i] = orientation_solve(i);
#pragma omp parallel for schedule(static) num_threads(threads)
            for (int i = c; i < n; i += 2) q[i] = qNew[i];
        }
    }

    // Sim::advance (YB/sim/step.cpp:38-62)
    void advance(float dtF) {
        if (dtF <= 0) return;
        int steps = std::max(1, (int)std::ceil(dtF / p.max_h));
        lastHF = hF;
        hF = dtF / steps;
        // The FP64 oracle uses the same (float) substep lengths, so that the
        // float-vs-double difference measures arithmetic rounding only.
        lastH = (T)lastHF;
        h = (T)hF;
        for (int s = 0; s < steps; s++, step_counter++) {
            init_itr();
            if (p.enable_hair_hair_coll && p.detection_period > 0 && step_counter % p.detection_period == 0)
                detect_collisions();
            recompute_step_limit();
            for (int k = 0; k < p.num_itr; k++) iterate();
            end_itr();
        }
        timeF += dtF;
        time = (T)timeF;
    }

    // ------------------------------------------------------------ SDF pass
    // SDFHeadCollision (SDF/cudaSDFCollision.cu:88-155, :205-268).  The index
    // and the unit-cube weights are evaluated with the same float/double
    // mixing as the reference kernel.
    bool sdf_inside(int ix, int iy, int iz) const {
        return ix >= 0 && ix < sdf.dim[0] && iy >= 0 && iy < sdf.dim[1] && iz >= 0 && iz < sdf.dim[2];
    }
    size_t sdf_idx(int ix, int iy, int iz) const {
        return (size_t)ix + (size_t)iy * sdf.dim[0] + (size_t)iz * sdf.dim[0] * sdf.dim[1];
    }
    void sdf_corner(int ix, int iy, int iz, T c[4]) const {
        if (sdf_inside(ix, iy, iz)) {
            size_t k = sdf_idx(ix, iy, iz);
            c[0] = sdf.grad[3 * k];
            c[1] = sdf.grad[3 * k + 1];
            c[2] = sdf.grad[3 * k + 2];
        } else {
            c[0] = c[1] = c[2] = 0;
        }
        int cx = std::max(0, std::min(ix, sdf.dim[0] - 1));
        int cy = std::max(0, std::min(iy, sdf.dim[1] - 1));
        int cz = std::max(0, std::min(iz, sdf.dim[2] - 1));
        c[3] = sdf.phi[sdf_idx(cx, cy, cz)];
    }
    void apply_sdf() {
        if (!sdf.enabled) return;
        const R4<T> invq{-(T)sdf.rot[0], -(T)sdf.rot[1], -(T)sdf.rot[2], (T)sdf.rot[3]};
        const V3<T> tr{(T)sdf.trans[0], (T)sdf.trans[1], (T)sdf.trans[2]};
        const V3<T> mn{(T)sdf.gmin[0], (T)sdf.gmin[1], (T)sdf.gmin[2]};
#pragma omp parallel for schedule(static) num_threads(threads)
        for (int i = 0; i < n; i++) {
            if (!(flags[i] & F_PREV)) continue;  // LocalIdx == 0: strand root
            V3<T> lp = rotate(invq, x[i] - tr) - mn;
            T l[3] = {lp.x, lp.y, lp.z};
            int idx[3];
            T ud[3];
            for (int a = 0; a < 3; a++) idx[a] = (int)((double)(l[a] * (T)sdf.inv_ds[a]) + 0.5);
            if (!sdf_inside(idx[0], idx[1], idx[2])) continue;
            for (int a = 0; a < 3; a++)
                ud[a] = (T)((double)sdf.inv_ds[a] * ((double)l[a] - ((double)idx[a] - 0.5) * (double)sdf.ds[a]));
            T c[8][4];
            // corner order c000 c001 c010 c011 c100 c101 c110 c111 (x y z)
            for (int o = 0; o < 8; o++) sdf_corner(idx[0] + ((o >> 2) & 1), idx[1] + ((o >> 1) & 1), idx[2] + (o & 1), c[o]);
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
            V3<T> nrm{res[0], res[1], res[2]};
            // helper_math normalize: v * rsqrtf(dot(v, v))
            nrm = nrm * (T(1) / std::sqrt(dot(nrm, nrm)));
            T phi = res[3] - (T)sdf.thresh;
            if (phi < 0) {
                x[i] = x[i] - phi * nrm;
                T vl = length(v[i]);
                if ((double)vl >= 1e-6) {  // double literal in the reference
                    V3<T> vn = v[i] / vl;
                    v[i] = v[i] - dot(vn, nrm) * nrm;
                }
            }
        }
    }

    // ----------------------------------------------------- strain limiting
    // StrainLimiting (SDF/cudaSDFCollision.cu:176-199), run by
    // SDFBodyCollisionSolver::SolveNextStep right after the SDF pass when
    // UseStrainLimiting is set (SDFBodyCollisionSolver.cpp:182-187): each
    // strand walked from its root; a segment whose length l departs from its
    // rest length l0 by more than StrainError * l0 is given length l0 by moving
    // its upper vertex along the segment (helper_math length and normalize);
    // velocities are untouched.
    bool strainOn = false;
    float strainError = 0.1f;
    std::vector<float> strainRest;  // rest length of segment (i, i + 1)
    // Rest lengths of InitAdditionalHairData (SDFBodyCollisionSolver.cpp:
    // 108-120): glm length of the current segments, float, no FMA.
    void strain_rest_from_positions() {
        strainRest.assign(n, 0.f);
        for (size_t s = 0; s + 1 < strandOff.size(); s++)
            for (int i = strandOff[s]; i + 1 < strandOff[s + 1]; i++) {
                const float dx = (float)x[i + 1].x - (float)x[i].x;
                const float dy = (float)x[i + 1].y - (float)x[i].y;
                const float dz = (float)x[i + 1].z - (float)x[i].z;
                strainRest[i] = std::sqrt((dx * dx + dy * dy) + dz * dz);
            }
    }
    void apply_strain_limit() {
        if (!strainOn || !sdf.enabled) return;
        const T err = (T)strainError;
        for (size_t s = 0; s + 1 < strandOff.size(); s++) {
            V3<T> cur = x[strandOff[s]];
            for (int i = strandOff[s]; i + 1 < strandOff[s + 1]; i++) {
                V3<T> nxt = x[i + 1];
                const T l0 = (T)strainRest[i];
                const T l = length(nxt - cur);
                if (std::fabs(l - l0) > err * l0) {
                    const V3<T> dir = normalize(cur - nxt);
                    nxt = nxt + (l - l0) * dir;
                    x[i + 1] = nxt;
                }
                cur = nxt;
            }
        }
    }

    // StableCosseratRodsSolver::SolveNextStep (StableCosseratRodsSolver.cpp:216-234)
    void solve_next_step(float dt) {
        advance(dt);
        apply_sdf();
        apply_strain_limit();
    }
};

}  // namespace orc
