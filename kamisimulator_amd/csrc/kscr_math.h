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

KS_FN f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
KS_FN f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
KS_FN f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
KS_FN f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
KS_FN f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
KS_FN f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
KS_FN float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
KS_FN f3 cross(f3 a, f3 b) { return f3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
KS_FN float length(f3 a) { return sqrtf(dot(a, a)); }
KS_FN f3 normalize(f3 a) { return a * (1.0f / sqrtf(dot(a, a))); }

KS_FN float dot(f4 a, f4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
KS_FN float length(f4 a) { return sqrtf(dot(a, a)); }
KS_FN f4 normalize(f4 a) {
    float s = 1.0f / sqrtf(dot(a, a));
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
    f3 aq = vec(a), bq = vec(b);
    f3 v = a.w * bq + b.w * aq + cross(aq, bq);
    return mk4(v, a.w * b.w - dot(aq, bq));
}

// r * v * r^-1
KS_FN f3 rotate(f4 r, f3 v) {
    f3 q = vec(r);
    f3 a = r.w * v + cross(q, v);
    float c = dot(v, q);
    return r.w * a + cross(q, a) + c * q;
}

// rotate(r, (1,0,0)) with the same operation order as the general case
KS_FN f3 rotate_e1(f4 r) { return rotate(r, f3{1.0f, 0.0f, 0.0f}); }

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
    const float ood =
        1.0f / (+m00 * (m11 * m22 - m21 * m12) - m10 * (m01 * m22 - m21 * m02) + m20 * (m01 * m12 - m11 * m02));
    const float i00 = +(m11 * m22 - m21 * m12) * ood;
    const float i10 = -(m10 * m22 - m20 * m12) * ood;
    const float i20 = +(m10 * m21 - m20 * m11) * ood;
    const float i01 = -(m01 * m22 - m21 * m02) * ood;
    const float i11 = +(m00 * m22 - m20 * m02) * ood;
    const float i21 = -(m00 * m21 - m20 * m01) * ood;
    const float i02 = +(m01 * m12 - m11 * m02) * ood;
    const float i12 = -(m00 * m12 - m10 * m02) * ood;
    const float i22 = +(m00 * m11 - m10 * m01) * ood;
    return f3{i00 * f.x + i10 * f.y + i20 * f.z, i01 * f.x + i11 * f.y + i21 * f.z,
              i02 * f.x + i12 * f.y + i22 * f.z};
}

// Vertex flag bits (same values as the reference's VertexFlags).
enum : uint32_t { FLAG_PREV = 1u, FLAG_NEXT = 2u, FLAG_NEXT_ORI = 4u, FLAG_FIX_ORI = 8u };

}  // namespace kscr
