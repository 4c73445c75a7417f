// kscr_meshsdf.hip -- body SDF grid built from a triangle mesh on gfx950.
//
// The reference builds the grid on the host, once per body mesh
// (PNMesh::GenerateSDF + StoreSDFQueries, src/PNMesh.cpp:100-171, through
// Kami::SDF::Init, src/utilities/Geometry.cpp:5-59):
//   1. the mesh's signed distance (Discregrid TriangleMeshDistance: Eberly's
//      point-triangle distance, sign from angle-weighted pseudo-normals) at
//      every node of a cubic Lagrange grid (Discregrid
//      CubicLagrangeDiscreteGrid, `dim` cells over the vertex box padded by
//      0.1; corner nodes plus two nodes on every cell edge);
//   2. that grid's interpolant at the body grid nodes gmin + i ds of the
//      vertex box, ds = extent / dim: the value, and inside the body the
//      normalized interpolant gradient, outside the direction from the grid
//      centre.
// Here both stages run on the device in double, with the reference's
// arithmetic sequences, so the float grid matches the reference bit for bit
// (apart from the choice among triangles at exactly equal distance, which
// leaves the distance unchanged).
//
//   node_distance_kernel  one lane per cubic grid node of the cells the body
//                         grid samples (the nodes of the other cells, in the
//                         0.1 padding, are never read): nearest triangle by a
//                         depth-first search of a linear BVH over the
//                         triangle boxes (nearer child first, children and
//                         stacked subtrees pruned by their box distance), then
//                         the pseudo-normal sign.  Ties of the squared
//                         distance go to the smaller triangle index, so the
//                         result does not depend on the traversal order.  A
//                         lane whose search passes 1024 leaf tests (near the
//                         medial axis nearly every triangle is a candidate;
//                         at a sphere's centre all are) lists its node for
//                         the next kernel instead of holding its wave.
//   node_brute_kernel     one workgroup per listed node: every triangle,
//                         split over the lanes, then a reduction to the
//                         smallest (squared distance, index) -- the same
//                         minimum.
//   sample_kernel         one lane per body grid node: the cell's 32 node
//                         values against the shape functions, written as the
//                         (grad, phi) float4 grid the SDF pass reads.
//
// The pruning keeps a subtree whose box distance is within `slack` of the
// best squared distance so far.  Box distances are computed to a few ulps;
// the triangle's quadratic-form distance carries an absolute error of a few
// ulps of |p - v0|^2 (cancellation), which is far below slack =
// 1e-12 * (domain diagonal)^2.  So no subtree that holds the minimum is
// dropped and the search returns the brute-force minimum.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <thread>
#include <vector>

#include "kscr.h"
#include "kscr_device.h"
#include "kscr_lbvh.h"

namespace kscr {
namespace {

struct Dbl3 {
    double x, y, z;
};
__host__ __device__ inline Dbl3 operator+(Dbl3 a, Dbl3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline Dbl3 operator-(Dbl3 a, Dbl3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline Dbl3 operator*(double s, Dbl3 a) { return {s * a.x, s * a.y, s * a.z}; }
__host__ __device__ inline double dot3(Dbl3 a, Dbl3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Dbl3 cross3(Dbl3 a, Dbl3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
// v / |v|, componentwise division
inline Dbl3 unit3(Dbl3 a) {
    const double l = std::sqrt(dot3(a, a));
    return {a.x / l, a.y / l, a.z / l};
}

// nearest feature of a triangle; the pseudo-normal slot of each
enum Feature { kV0 = 0, kV1, kV2, kE01, kE12, kE02, kFace };
constexpr int kNormalSlots = 7;  // face, e01, e12, e02, v0, v1, v2
__device__ inline int normal_slot(int f) {
    return f == kFace ? 0 : f == kE01 ? 1 : f == kE12 ? 2 : f == kE02 ? 3 : 4 + f;
}

// Squared distance from p to triangle (v0, v1, v2), the nearest point and
// its feature: the minimum of |v0 + s e0 + t e1 - p|^2 over the triangle by
// the region of the unconstrained minimum (TriangleMeshDistance.h:557-749),
// the squared distance from the quadratic's value there, clamped at 0.
__device__ inline double point_triangle_d2(Dbl3 p, Dbl3 v0, Dbl3 v1, Dbl3 v2, Dbl3& nearest, int& feature) {
    const Dbl3 d = v0 - p, e0 = v1 - v0, e1 = v2 - v0;
    const double a00 = dot3(e0, e0), a01 = dot3(e0, e1), a11 = dot3(e1, e1);
    const double b0 = dot3(d, e0), b1 = dot3(d, e1), c = dot3(d, d);
    const double det = fabs(a00 * a11 - a01 * a01);
    double s = a01 * b1 - a11 * b0, t = a01 * b0 - a00 * b1, d2;
    auto q_st = [&]() { return s * (a00 * s + a01 * t + 2 * b0) + t * (a01 * s + a11 * t + 2 * b1) + c; };
    auto at_v0 = [&]() { s = 0, t = 0, feature = kV0, d2 = c; };
    auto at_v1 = [&]() { s = 1, t = 0, feature = kV1, d2 = a00 + 2 * b0 + c; };
    auto at_v2 = [&]() { s = 0, t = 1, feature = kV2, d2 = a11 + 2 * b1 + c; };
    auto edge01 = [&]() {
        if (b0 >= 0) at_v0();
        else if (-b0 >= a00) at_v1();
        else t = 0, s = -b0 / a00, feature = kE01, d2 = b0 * s + c;
    };
    auto edge02 = [&]() {
        if (b1 >= 0) at_v0();
        else if (-b1 >= a11) at_v2();
        else s = 0, t = -b1 / a11, feature = kE02, d2 = b1 * t + c;
    };
    auto edge12 = [&](double numer, bool s_param) {
        const double denom = a00 - 2 * a01 + a11;
        if (numer >= denom) {
            if (s_param) at_v1();
            else at_v2();
            return;
        }
        if (s_param) s = numer / denom, t = 1 - s;
        else t = numer / denom, s = 1 - t;
        feature = kE12;
        d2 = q_st();
    };
    if (s + t <= det) {
        if (s < 0) {
            if (t < 0 && b0 < 0) edge01();
            else edge02();
        } else if (t < 0) {
            edge01();
        } else {
            const double inv = 1.0 / det;
            s *= inv, t *= inv;
            feature = kFace;
            d2 = q_st();
        }
    } else if (s < 0) {
        const double tmp0 = a01 + b0, tmp1 = a11 + b1;
        if (tmp1 > tmp0) edge12(tmp1 - tmp0, true);
        else if (tmp1 <= 0) at_v2();
        else edge02();
    } else if (t < 0) {
        const double tmp0 = a01 + b1, tmp1 = a00 + b0;
        if (tmp1 > tmp0) edge12(tmp1 - tmp0, false);
        else if (tmp1 <= 0) at_v1();
        else edge01();
    } else {
        const double numer = a11 + b1 - a01 - b0;
        if (numer <= 0) at_v2();
        else edge12(numer, true);
    }
    if (d2 < 0) d2 = 0;
    nearest = v0 + s * e0 + t * e1;
    return d2;
}

// The cubic grid: `n` cells per axis from lo, cell size and its reciprocal;
// node counts of each kind.
struct CubicGrid {
    unsigned n[3];
    double lo[3], cell[3], inv_cell[3];
    unsigned long long nv, ne[3];
    __host__ __device__ unsigned long long num_nodes() const { return nv + 2 * (ne[0] + ne[1] + ne[2]); }
};

// Node l: corner nodes first (x fastest), then two nodes per x edge (x
// fastest, then y, z), per y edge (y, z, x) and per z edge (z, x, y), at 1/3
// and 2/3 of the edge (cubic_lagrange_discrete_grid.cpp:496-548).
__device__ inline Dbl3 node_position(const CubicGrid& g, unsigned long long l) {
    const unsigned long long n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
    unsigned long long ijk[3];
    int axis = -1;
    unsigned long long e = 0;
    if (l < g.nv) {
        ijk[2] = l / ((n1 + 1) * (n0 + 1));
        const unsigned long long r = l % ((n1 + 1) * (n0 + 1));
        ijk[1] = r / (n0 + 1);
        ijk[0] = r % (n0 + 1);
    } else if (l < g.nv + 2 * g.ne[0]) {
        axis = 0, e = l - g.nv;
        const unsigned long long w = e / 2, r = w % ((n1 + 1) * n0);
        ijk[2] = w / ((n1 + 1) * n0);
        ijk[1] = r / n0;
        ijk[0] = r % n0;
    } else if (l < g.nv + 2 * (g.ne[0] + g.ne[1])) {
        axis = 1, e = l - g.nv - 2 * g.ne[0];
        const unsigned long long w = e / 2, r = w % ((n2 + 1) * n1);
        ijk[0] = w / ((n2 + 1) * n1);
        ijk[2] = r / n1;
        ijk[1] = r % n1;
    } else {
        axis = 2, e = l - g.nv - 2 * (g.ne[0] + g.ne[1]);
        const unsigned long long w = e / 2, r = w % ((n0 + 1) * n2);
        ijk[1] = w / ((n0 + 1) * n2);
        ijk[0] = r / n2;
        ijk[2] = r % n2;
    }
    double x[3];
    for (int a = 0; a < 3; a++) x[a] = g.lo[a] + g.cell[a] * (double)ijk[a];
    if (axis >= 0) x[axis] += (1.0 + (double)(e % 2)) / 3.0 * g.cell[axis];
    return Dbl3{x[0], x[1], x[2]};
}

// The cells the body grid samples: a box of cells [c0, c0 + m) (every body
// node's cell lies in it; the cubic grid's padding cells are never read).
// Its nodes are enumerated compactly -- corners, x-edge, y-edge, z-edge
// nodes, each x fastest -- and mapped to their ids in the full layout.
struct SubGrid {
    unsigned c0[3], m[3];
    unsigned long long nc, nx, ny, nz;  // node counts of each kind in the box
    __host__ __device__ unsigned long long count() const { return nc + nx + ny + nz; }
};

__device__ inline unsigned long long sub_node_id(const CubicGrid& g, const SubGrid& s, unsigned long long q) {
    const unsigned long long n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
    unsigned long long e[3];  // extent of the index box of this kind, per axis
    int kind;
    if (q < s.nc) {
        kind = 0;
        for (int a = 0; a < 3; a++) e[a] = s.m[a] + 1ull;
    } else {
        q -= s.nc;
        kind = 1;
        if (q >= s.nx) q -= s.nx, kind = 2;
        if (kind == 2 && q >= s.ny) q -= s.ny, kind = 3;
        for (int a = 0; a < 3; a++) e[a] = s.m[a] + (a == kind - 1 ? 0ull : 1ull);
    }
    const unsigned long long h = kind ? q % 2 : 0;
    if (kind) q /= 2;
    const unsigned long long i = s.c0[0] + q % e[0], j = s.c0[1] + (q / e[0]) % e[1], k = s.c0[2] + q / (e[0] * e[1]);
    switch (kind) {
        case 0: return (n0 + 1) * (n1 + 1) * k + (n0 + 1) * j + i;
        case 1: return g.nv + 2 * (n0 * (n1 + 1) * k + n0 * j + i) + h;
        case 2: return g.nv + 2 * g.ne[0] + 2 * (n1 * (n2 + 1) * i + n1 * k + j) + h;
        default: return g.nv + 2 * (g.ne[0] + g.ne[1]) + 2 * (n2 * (n0 + 1) * j + n2 * i + k) + h;
    }
}

struct MeshDev {
    const float4* tri;  // three vertices per triangle (w unused)
    const double* pn;   // kNormalSlots pseudo-normals (xyz) per triangle
    int nt;
};

// squared distance from p to a float box (min xyz, max xyz)
__device__ inline double box_d2(const float* b, Dbl3 p) {
    const double c[3] = {p.x, p.y, p.z};
    double s = 0;
    for (int a = 0; a < 3; a++) {
        const double lo = b[a], hi = b[3 + a];
        const double d = c[a] < lo ? lo - c[a] : (c[a] > hi ? c[a] - hi : 0.0);
        s += d * d;
    }
    return s;
}

constexpr int kDistBlock = 64;
constexpr uint32_t kLeafBit = 0x80000000u, kNoNode = 0xFFFFFFFFu;

template <int STACK>
__global__ __launch_bounds__(kDistBlock) void node_distance_kernel(MeshDev m, const LbvhNode* __restrict__ nodes,
                                                                   const uint32_t* __restrict__ ids, CubicGrid g,
                                                                   SubGrid sub, double slack, int budget,
                                                                   double* __restrict__ coef,
                                                                   unsigned long long* __restrict__ overflow,
                                                                   unsigned* __restrict__ overflow_count,
                                                                   int* __restrict__ error) {
    __shared__ uint32_t st_node[STACK][kDistBlock];
    __shared__ float st_d2[STACK][kDistBlock];
    const unsigned long long q = (unsigned long long)blockIdx.x * kDistBlock + threadIdx.x;
    if (q >= sub.count()) return;
    const unsigned long long l = sub_node_id(g, sub, q);
    const int lane = threadIdx.x;
    const Dbl3 p = node_position(g, l);
    double best = INFINITY;
    int best_k = 0x7fffffff, best_f = kFace;
    Dbl3 best_q = p;
    int tests = 0;
    auto test = [&](int k) {
        const float4 a = m.tri[3 * k], b = m.tri[3 * k + 1], c = m.tri[3 * k + 2];
        Dbl3 q;
        int f;
        const double d2 = point_triangle_d2(p, Dbl3{a.x, a.y, a.z}, Dbl3{b.x, b.y, b.z}, Dbl3{c.x, c.y, c.z}, q, f);
        if (d2 < best || (d2 == best && k < best_k)) best = d2, best_k = k, best_f = f, best_q = q;
        tests++;
    };
    if (m.nt == 1) {
        test(0);
    } else {
        uint32_t node = 0;  // the root
        int sp = 0;
        for (;;) {
            const float4* np = reinterpret_cast<const float4*>(nodes + node);
            const float4 w0 = np[0], w1 = np[1], w2 = np[2], w3 = np[3];
            const uint32_t lc = __float_as_uint(w0.y), rc = __float_as_uint(w0.z);
            const float b0[6] = {w1.x, w1.y, w1.z, w1.w, w2.x, w2.y};
            const float b1[6] = {w2.z, w2.w, w3.x, w3.y, w3.z, w3.w};
            const double d0 = box_d2(b0, p), d1 = box_d2(b1, p);
            bool go0 = d0 <= best + slack, go1 = d1 <= best + slack;
            // leaf children are tested in place, the nearer first
            if (d1 < d0) {
                if (go1 && (rc & kLeafBit)) test((int)ids[rc & ~kLeafBit]), go1 = false;
                if (go0 && (lc & kLeafBit)) test((int)ids[lc & ~kLeafBit]), go0 = false;
            } else {
                if (go0 && (lc & kLeafBit)) test((int)ids[lc & ~kLeafBit]), go0 = false;
                if (go1 && (rc & kLeafBit)) test((int)ids[rc & ~kLeafBit]), go1 = false;
            }
            go0 = go0 && d0 <= best + slack;
            go1 = go1 && d1 <= best + slack;
            uint32_t next = kNoNode;
            if (go0 && go1) {
                const bool right_first = d1 < d0;
                if (sp < STACK) {
                    // the box distance rounded down, so the pop-time test never drops more
                    st_node[sp][lane] = right_first ? lc : rc;
                    st_d2[sp][lane] = __double2float_rd(right_first ? d0 : d1);
                    sp++;
                } else {
                    atomicOr(error, 1);  // deeper than the stack: reported, the result is not used
                }
                next = right_first ? rc : lc;
            } else if (go0) {
                next = lc;
            } else if (go1) {
                next = rc;
            }
            while (next == kNoNode && sp > 0) {
                --sp;
                if ((double)st_d2[sp][lane] <= best + slack) next = st_node[sp][lane];
            }
            if (next == kNoNode) break;
            if (tests > budget) {
                // near the medial axis almost every triangle is a candidate:
                // left to node_brute_kernel, which splits it over a workgroup
                overflow[atomicAdd(overflow_count, 1u)] = l;
                return;
            }
            node = next;
        }
    }
    if (best_k == 0x7fffffff) {  // no triangle reached (cannot happen for finite input): reported
        atomicOr(error, 2);
        coef[l] = NAN;
        return;
    }
    const double* n = m.pn + 3 * ((size_t)kNormalSlots * best_k + normal_slot(best_f));
    const Dbl3 u = p - best_q;
    const double sign = (u.x * n[0] + u.y * n[1] + u.z * n[2]) >= 0.0 ? 1.0 : -1.0;
    coef[l] = sign * sqrt(best);
}

// The nodes whose search exceeded its budget: one workgroup per node tests
// every triangle (a strided share per lane), then reduces to the smallest
// (squared distance, triangle index), the same minimum the search returns.
constexpr int kBruteBlock = 256;
__global__ __launch_bounds__(kBruteBlock) void node_brute_kernel(MeshDev m, CubicGrid g,
                                                                 const unsigned long long* __restrict__ list,
                                                                 const unsigned* __restrict__ count,
                                                                 double* __restrict__ coef) {
    __shared__ double s_d2[kBruteBlock];
    __shared__ int s_k[kBruteBlock];
    const unsigned n = *count;
    const int tid = threadIdx.x;
    auto tri_d2 = [&](int k, Dbl3 p, Dbl3& q, int& f) {
        const float4 a = m.tri[3 * k], b = m.tri[3 * k + 1], c = m.tri[3 * k + 2];
        return point_triangle_d2(p, Dbl3{a.x, a.y, a.z}, Dbl3{b.x, b.y, b.z}, Dbl3{c.x, c.y, c.z}, q, f);
    };
    for (unsigned i = blockIdx.x; i < n; i += gridDim.x) {  // n is the same for every lane
        const unsigned long long l = list[i];
        const Dbl3 p = node_position(g, l);
        double best = INFINITY;
        int best_k = 0x7fffffff;
        for (int k = tid; k < m.nt; k += kBruteBlock) {
            Dbl3 q;
            int f;
            const double d2 = tri_d2(k, p, q, f);
            if (d2 < best) best = d2, best_k = k;  // k increases: the first of equals is the smallest
        }
        s_d2[tid] = best;
        s_k[tid] = best_k;
        __syncthreads();
        for (int w = kBruteBlock / 2; w > 0; w >>= 1) {
            if (tid < w) {
                const double d = s_d2[tid + w];
                const int k = s_k[tid + w];
                if (d < s_d2[tid] || (d == s_d2[tid] && k < s_k[tid])) s_d2[tid] = d, s_k[tid] = k;
            }
            __syncthreads();
        }
        if (tid == 0) {
            const int k = s_k[0];
            Dbl3 q;
            int f;
            tri_d2(k, p, q, f);  // the winner's nearest point and feature, recomputed
            const double* nrm = m.pn + 3 * ((size_t)kNormalSlots * k + normal_slot(f));
            const Dbl3 u = p - q;
            const double sign = (u.x * nrm[0] + u.y * nrm[1] + u.z * nrm[2]) >= 0.0 ? 1.0 : -1.0;
            coef[l] = sign * sqrt(s_d2[0]);
        }
        __syncthreads();
    }
}

// The body grid: float nodes gmin + i ds, the vertex box's centre.
struct BodyGrid {
    int dim[3];
    float gmin[3], ds[3], centre[3];
};

// One body grid node: the cubic cell holding it (clamped to the last cell),
// its local coordinates in [-1, 1]^3, the 32 node values against the
// serendipity shape functions (each value the product sequence of
// shape_function_, cubic_lagrange_discrete_grid.cpp:286-477, summed in node
// order as :823-894 sums them), then PNMesh::StoreSDFQueries' convention.
__global__ __launch_bounds__(256) void sample_kernel(const double* __restrict__ coef, CubicGrid g, BodyGrid b,
                                                     float4* __restrict__ out) {
    const long long total = (long long)b.dim[0] * b.dim[1] * b.dim[2];
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int ix = (int)(i % b.dim[0]), iy = (int)((i / b.dim[0]) % b.dim[1]),
              iz = (int)(i / ((long long)b.dim[0] * b.dim[1]));
    const float pf[3] = {b.gmin[0] + b.ds[0] * (float)ix, b.gmin[1] + b.ds[1] * (float)iy,
                         b.gmin[2] + b.ds[2] * (float)iz};
    const double x[3] = {pf[0], pf[1], pf[2]};
    unsigned mi[3];
    double xi[3], c0[3];
    for (int a = 0; a < 3; a++) {
        const double t = (x[a] - g.lo[a]) * g.inv_cell[a];
        mi[a] = t > 0.0 ? (unsigned)t : 0u;
        if (mi[a] >= g.n[a]) mi[a] = g.n[a] - 1;
        const double smin = g.lo[a] + (double)mi[a] * g.cell[a];
        const double smax = smin + g.cell[a];
        const double denom = smax - smin;
        c0[a] = 2.0 / denom;
        const double c1 = (smax + smin) / denom;
        xi[a] = c0[a] * x[a] - c1;
    }
    const unsigned long long nx = g.n[0], ny = g.n[1], nz = g.n[2];
    const unsigned long long ci = mi[0], cj = mi[1], ck = mi[2];
    double phi = 0.0, gr[3] = {0.0, 0.0, 0.0};
    auto acc = [&](unsigned long long id, double N, double dx, double dy, double dz) {
        const double c = coef[id];
        phi += c * N;
        gr[0] += c * dx;
        gr[1] += c * dy;
        gr[2] += c * dz;
    };
    const double X = xi[0], Y = xi[1], Z = xi[2];
    const double x2 = X * X, y2 = Y * Y, z2 = Z * Z;
    const double mlin[3] = {1.0 - X, 1.0 - Y, 1.0 - Z}, plin[3] = {1.0 + X, 1.0 + Y, 1.0 + Z};
    const double fac = 1.0 / 64.0 * (9.0 * (x2 + y2 + z2) - 19.0);
    const double gx = 9.0 * (3.0 * x2 + y2 + z2) - 19.0, gy = 9.0 * (x2 + 3.0 * y2 + z2) - 19.0,
                 gz = 9.0 * (x2 + y2 + 3.0 * z2) - 19.0;
    const double t18[3] = {18.0 * X, 18.0 * Y, 18.0 * Z};
    // corner nodes, x bit 0, y bit 1, z bit 2
    const unsigned long long plane = (nx + 1) * (ny + 1);
#pragma unroll
    for (int c = 0; c < 8; c++) {
        const bool sx = c & 1, sy = c & 2, sz = c & 4;
        const double lx = sx ? plin[0] : mlin[0], ly = sy ? plin[1] : mlin[1], lz = sz ? plin[2] : mlin[2];
        const unsigned long long id = plane * (ck + (sz ? 1 : 0)) + (nx + 1) * (cj + (sy ? 1 : 0)) + ci + (sx ? 1 : 0);
        acc(id, fac * (lx * ly) * lz, (sx ? t18[0] + gx : t18[0] - gx) * (ly * lz) / 64.0,
            (lx * lz) * (sy ? t18[1] + gy : t18[1] - gy) / 64.0, (lx * ly) * (sz ? t18[2] + gz : t18[2] - gz) / 64.0);
    }
    const double sq[3] = {x2, y2, z2}, co[3] = {X, Y, Z};
    unsigned long long off = g.nv;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const int bb = a == 0 ? 1 : 0, cc = a == 2 ? 1 : 2;  // transverse axes, increasing
        const double one_m = 1.0 - sq[a];
        const double f = 9.0 / 64.0 * one_m;
        const double lm3 = 1.0 - 3.0 * co[a], lp3 = 1.0 + 3.0 * co[a];
        const double three = 3.0 - 9.0 * sq[a], two = 2.0 * co[a];
        const double dm = -three - two, dp = three - two;
        const double om_m = one_m * lm3, om_p = one_m * lp3;
#pragma unroll
        for (int e = 0; e < 4; e++) {
            bool pb, pc;  // transverse sides
            unsigned long long base;
            if (a == 0) {  // x edge at (j + pb, k + pc); edge order (y, z) with z fastest
                pb = e & 2, pc = e & 1;
                base = off + 2 * (nx * (ny + 1) * (ck + pc) + nx * (cj + pb) + ci);
            } else if (a == 1) {  // y edge at (i + pb, k + pc); (x, z) with x fastest
                pb = e & 1, pc = e & 2;
                base = off + 2 * (ny * (nz + 1) * (ci + pb) + ny * (ck + pc) + cj);
            } else {  // z edge at (i + pb, j + pc); (x, y) with y fastest
                pb = e & 2, pc = e & 1;
                base = off + 2 * (nz * (nx + 1) * (cj + pc) + nz * (ci + pb) + ck);
            }
            const double lb = pb ? plin[bb] : mlin[bb], lc = pc ? plin[cc] : mlin[cc];
            const double tr = lb * lc;
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const double om = h ? om_p : om_m;
                double d[3];
                d[a] = (h ? dp : dm) * tr * (9.0 / 64.0);
                d[bb] = ((pb ? om : -om) * lc) * (9.0 / 64.0);
                d[cc] = ((pc ? om : -om) * lb) * (9.0 / 64.0);
                acc(base + h, (f * (h ? lp3 : lm3)) * tr, d[0], d[1], d[2]);
            }
        }
        off += 2 * g.ne[a];
    }
    const float dist = (float)phi;
    float4 o;
    o.w = dist;
    if (dist < 0.0f) {
        const double gg[3] = {gr[0] * c0[0], gr[1] * c0[1], gr[2] * c0[2]};
        const double l = sqrt(gg[0] * gg[0] + gg[1] * gg[1] + gg[2] * gg[2]);
        o.x = (float)(gg[0] / l), o.y = (float)(gg[1] / l), o.z = (float)(gg[2] / l);
    } else {
        const float t[3] = {b.centre[0] - pf[0], b.centre[1] - pf[1], b.centre[2] - pf[2]};
        const float l = sqrtf(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
        o.x = -t[0] / l, o.y = -t[1] / l, o.z = -t[2] / l;
    }
    out[i] = o;
}

__global__ void fill_nan_kernel(double* p, unsigned long long n) {
    const unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = NAN;
}

// f(begin, end) over [0, n) split across up to 16 host threads
template <class F>
void parallel_ranges(int n, F f) {
    const int workers = (int)std::max(1u, std::min(std::thread::hardware_concurrency(), 16u));
    const int chunks = std::min(workers, std::max(1, n / 4096));
    if (chunks <= 1) {
        f(0, n);
        return;
    }
    std::vector<std::thread> pool;
    for (int c = 0; c < chunks; c++)
        pool.emplace_back(f, (int)((long long)n * c / chunks), (int)((long long)n * (c + 1) / chunks));
    for (auto& t : pool) t.join();
}

template <class T>
struct Scoped {
    T* p = nullptr;
    hipError_t alloc(size_t n) { return hipMalloc((void**)&p, n * sizeof(T)); }
    ~Scoped() {
        if (p) (void)hipFree(p);
    }
};
struct ScopedLbvh {
    LbvhDev d;
    ~ScopedLbvh() { lbvh_release(d); }
};

}  // namespace

int mesh_sdf_build(const float* verts, int nv, const int* tris, int nt, const int dim[3], float4* out, float gmin[3],
                   float gmax[3], hipStream_t stream, std::string& err) {
    if (!verts || !tris || !dim || !out || nv < 3 || nt < 1) {
        err = "mesh SDF: needs at least 3 vertices and 1 triangle";
        return KSCR_ERR_INVALID_ARGUMENT;
    }
    long long cells = 1;
    for (int a = 0; a < 3; a++) {
        if (dim[a] < 1 || dim[a] > 4096) {
            err = "mesh SDF: grid dimensions must lie in [1, 4096]";
            return KSCR_ERR_INVALID_ARGUMENT;
        }
        cells *= dim[a];
    }
    if (cells > (1ll << 31) - 1) {
        err = "mesh SDF: more than 2^31 - 1 grid nodes";
        return KSCR_ERR_INVALID_ARGUMENT;
    }
    for (long long i = 0; i < 3ll * nv; i++)
        if (!std::isfinite(verts[i])) {
            err = "mesh SDF: non-finite vertex coordinate";
            return KSCR_ERR_INVALID_ARGUMENT;
        }
    for (long long i = 0; i < 3ll * nt; i++)
        if (tris[i] < 0 || tris[i] >= nv) {
            err = "mesh SDF: triangle index out of range";
            return KSCR_ERR_INVALID_ARGUMENT;
        }

    // vertex box (float, as aiMesh::mAABB) and the padded cubic grid domain
    double lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
        gmin[a] = gmax[a] = verts[a];
        for (int i = 1; i < nv; i++) gmin[a] = std::min(gmin[a], verts[3 * i + a]), gmax[a] = std::max(gmax[a], verts[3 * i + a]);
        lo[a] = (double)gmin[a] - 0.1;
        hi[a] = (double)gmax[a] + 0.1;
    }
    CubicGrid g;
    double diag2 = 0;
    for (int a = 0; a < 3; a++) {
        g.n[a] = (unsigned)dim[a];
        g.lo[a] = lo[a];
        g.cell[a] = (hi[a] - lo[a]) / (double)dim[a];
        g.inv_cell[a] = 1.0 / g.cell[a];
        diag2 += (hi[a] - lo[a]) * (hi[a] - lo[a]);
    }
    const unsigned long long n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
    g.nv = (n0 + 1) * (n1 + 1) * (n2 + 1);
    g.ne[0] = n0 * (n1 + 1) * (n2 + 1);
    g.ne[1] = (n0 + 1) * n1 * (n2 + 1);
    g.ne[2] = (n0 + 1) * (n1 + 1) * n2;
    const unsigned long long n_nodes = g.num_nodes();

    // triangles, boxes and angle-weighted pseudo-normals
    // (TriangleMeshDistance.h:342-441), in double on the host.  The per-
    // triangle work runs on a thread pool; the sums over triangles (vertex
    // and edge normals) run in triangle order, as the reference adds them.
    std::vector<float4> tri(3 * (size_t)nt);
    std::vector<float> box(6 * (size_t)nt);
    std::vector<double> pn(3 * (size_t)kNormalSlots * nt);
    std::vector<Dbl3> n_vert(nv, Dbl3{0, 0, 0}), n_face(nt);
    std::vector<double> w(3 * (size_t)nt);
    std::vector<char> degenerate(nt, 0);
    auto vtx = [&](int i) { return Dbl3{verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]}; };
    parallel_ranges(nt, [&](int k0, int k1) {
        for (int k = k0; k < k1; k++) {
            const int* t = tris + 3 * k;
            const Dbl3 a = vtx(t[0]), b = vtx(t[1]), c = vtx(t[2]);
            const Dbl3 cr = cross3(b - a, c - a);
            if (!(dot3(cr, cr) > 0)) {
                degenerate[k] = 1;
                continue;
            }
            n_face[k] = unit3(cr);
            w[3 * k] = std::acos(std::fabs(dot3(unit3(b - a), unit3(c - a))));
            w[3 * k + 1] = std::acos(std::fabs(dot3(unit3(a - b), unit3(c - b))));
            w[3 * k + 2] = std::acos(std::fabs(dot3(unit3(b - c), unit3(a - c))));
            for (int j = 0; j < 3; j++) {
                const float* q = verts + 3 * t[j];
                tri[3 * k + j] = make_float4(q[0], q[1], q[2], 0.f);
                for (int a2 = 0; a2 < 3; a2++) {
                    float& mn = box[6 * k + a2];
                    float& mx = box[6 * k + 3 + a2];
                    if (j == 0 || q[a2] < mn) mn = q[a2];
                    if (j == 0 || q[a2] > mx) mx = q[a2];
                }
            }
        }
    });
    for (int k = 0; k < nt; k++)
        if (degenerate[k]) {
            err = "mesh SDF: triangle " + std::to_string(k) + " has zero area";
            return KSCR_ERR_INVALID_ARGUMENT;
        }
    for (int k = 0; k < nt; k++) {
        const int* t = tris + 3 * k;
        for (int j = 0; j < 3; j++) n_vert[t[j]] = n_vert[t[j]] + w[3 * k + j] * n_face[k];
    }
    // edge normals: the face normals of the triangles sharing an edge summed
    // in triangle order (edges 01, 12, 02 within a triangle).  Edges are
    // grouped by their smaller vertex with a stable counting sort, then by
    // the larger one within the group.
    const size_t ne = 3 * (size_t)nt;
    std::vector<uint32_t> start(nv + 1, 0), order(ne);
    auto edge_ends = [&](size_t e, int& lo, int& hi) {
        const int* t = tris + 3 * (e / 3);
        const int s = (int)(e % 3);
        const int i = t[s == 2 ? 0 : s], j = t[s == 0 ? 1 : 2];
        lo = std::min(i, j), hi = std::max(i, j);
    };
    for (size_t e = 0; e < ne; e++) {
        int lo, hi;
        edge_ends(e, lo, hi);
        start[lo + 1]++;
    }
    for (int i = 0; i < nv; i++) start[i + 1] += start[i];
    {
        std::vector<uint32_t> fill(start.begin(), start.end() - 1);
        for (size_t e = 0; e < ne; e++) {
            int lo, hi;
            edge_ends(e, lo, hi);
            order[fill[lo]++] = (uint32_t)e;
        }
    }
    std::vector<Dbl3> edge_n(ne);  // per triangle edge slot: the normalized sum
    parallel_ranges(nv, [&](int v0, int v1) {
        std::vector<char> done;
        for (int i = v0; i < v1; i++) {
            const uint32_t b0 = start[i], b1 = start[i + 1];
            done.assign(b1 - b0, 0);
            for (uint32_t x = b0; x < b1; x++) {
                if (done[x - b0]) continue;
                int lo, hi, lo2, hi2;
                edge_ends(order[x], lo, hi);
                Dbl3 sum = n_face[order[x] / 3];
                for (uint32_t y = x + 1; y < b1; y++) {
                    edge_ends(order[y], lo2, hi2);
                    if (hi2 == hi) sum = sum + n_face[order[y] / 3];
                }
                const Dbl3 n = unit3(sum);
                for (uint32_t y = x; y < b1; y++) {
                    edge_ends(order[y], lo2, hi2);
                    if (hi2 == hi) edge_n[order[y]] = n, done[y - b0] = 1;
                }
            }
        }
    });
    parallel_ranges(nv, [&](int v0, int v1) {
        for (int i = v0; i < v1; i++) n_vert[i] = unit3(n_vert[i]);
    });
    parallel_ranges(nt, [&](int k0, int k1) {
        for (int k = k0; k < k1; k++) {
            const int* t = tris + 3 * k;
            const Dbl3 slot[kNormalSlots] = {n_face[k],         edge_n[3 * (size_t)k], edge_n[3 * (size_t)k + 1],
                                             edge_n[3 * (size_t)k + 2], n_vert[t[0]], n_vert[t[1]],
                                             n_vert[t[2]]};
            for (int s2 = 0; s2 < kNormalSlots; s2++) {
                double* d = pn.data() + 3 * ((size_t)kNormalSlots * k + s2);
                d[0] = slot[s2].x, d[1] = slot[s2].y, d[2] = slot[s2].z;
            }
        }
    });

    hipError_t e;
    auto fail = [&](hipError_t code, const char* what) {
        err = std::string("mesh SDF: ") + what + ": " + hipGetErrorString(code);
        return code == hipErrorOutOfMemory ? KSCR_ERR_OUT_OF_MEMORY : KSCR_ERR_HIP;
    };
    Scoped<float4> d_tri;
    Scoped<double> d_pn, d_coef;
    Scoped<int> d_err;
    Scoped<unsigned long long> d_over;
    ScopedLbvh bvh;
    if ((e = d_tri.alloc(tri.size())) != hipSuccess || (e = d_pn.alloc(pn.size())) != hipSuccess ||
        (e = d_coef.alloc(n_nodes)) != hipSuccess || (e = d_err.alloc(2)) != hipSuccess ||
        (e = lbvh_reserve(bvh.d, nt)) != hipSuccess)
        return fail(e, "hipMalloc");
    if ((e = hipMemcpyAsync(d_tri.p, tri.data(), tri.size() * sizeof(float4), hipMemcpyHostToDevice, stream)) !=
            hipSuccess ||
        (e = hipMemcpyAsync(d_pn.p, pn.data(), pn.size() * sizeof(double), hipMemcpyHostToDevice, stream)) !=
            hipSuccess ||
        (e = hipMemcpyAsync(bvh.d.aabb, box.data(), box.size() * sizeof(float), hipMemcpyHostToDevice, stream)) !=
            hipSuccess ||
        (e = fill_i32(d_err.p, 0, 2, stream)) != hipSuccess)
        return fail(e, "upload");
    // nodes outside the sampled cells stay NaN: a read of one would show in the grid
    hipLaunchKernelGGL(fill_nan_kernel, dim3((unsigned)((n_nodes + 255) / 256)), dim3(256), 0, stream, d_coef.p,
                       n_nodes);
    if ((e = lbvh_build(bvh.d, nt, stream)) != hipSuccess || (e = lbvh_read_depth(bvh.d, stream)) != hipSuccess)
        return fail(e, "triangle BVH build");
    if (bvh.d.max_depth_host > 64) {
        err = "mesh SDF: triangle BVH deeper than 64 levels";
        return KSCR_ERR_UNSUPPORTED;
    }
    MeshDev m{d_tri.p, d_pn.p, nt};
    const double slack = 1e-12 * diag2;
    BodyGrid b;
    for (int a = 0; a < 3; a++) {
        b.dim[a] = dim[a];
        b.gmin[a] = gmin[a];
        b.ds[a] = (gmax[a] - gmin[a]) / (float)dim[a];
        b.centre[a] = (gmax[a] + gmin[a]) / 2;
    }
    // the cells the body nodes fall in: along each axis the first and the last
    // node's cell (the cell index is monotone in the node index), with the
    // sample kernel's arithmetic
    SubGrid sub;
    for (int a = 0; a < 3; a++) {
        auto cell_of = [&](int i) {
            const float x = b.gmin[a] + b.ds[a] * (float)i;
            const double t = ((double)x - g.lo[a]) * g.inv_cell[a];
            unsigned c = t > 0.0 ? (unsigned)t : 0u;
            return c >= g.n[a] ? g.n[a] - 1 : c;
        };
        const unsigned lo_c = cell_of(0), hi_c = cell_of(dim[a] - 1);
        sub.c0[a] = lo_c;
        sub.m[a] = hi_c - lo_c + 1;
    }
    {
        const unsigned long long m0 = sub.m[0], m1 = sub.m[1], m2 = sub.m[2];
        sub.nc = (m0 + 1) * (m1 + 1) * (m2 + 1);
        sub.nx = 2 * m0 * (m1 + 1) * (m2 + 1);
        sub.ny = 2 * (m0 + 1) * m1 * (m2 + 1);
        sub.nz = 2 * (m0 + 1) * (m1 + 1) * m2;
    }
    const dim3 grid((unsigned)((sub.count() + kDistBlock - 1) / kDistBlock));
    if ((e = d_over.alloc(sub.count())) != hipSuccess) return fail(e, "hipMalloc");
    unsigned* over_count = reinterpret_cast<unsigned*>(d_err.p + 1);
    // leaf tests a lane may spend before its node goes to node_brute_kernel
    const int budget = 1024;
    if (bvh.d.max_depth_host <= 32)
        hipLaunchKernelGGL(node_distance_kernel<32>, grid, dim3(kDistBlock), 0, stream, m, bvh.d.nodes, bvh.d.ids, g,
                           sub, slack, budget, d_coef.p, d_over.p, over_count, d_err.p);
    else
        hipLaunchKernelGGL(node_distance_kernel<64>, grid, dim3(kDistBlock), 0, stream, m, bvh.d.nodes, bvh.d.ids, g,
                           sub, slack, budget, d_coef.p, d_over.p, over_count, d_err.p);
    if ((e = hipGetLastError()) != hipSuccess) return fail(e, "node_distance_kernel");
    hipLaunchKernelGGL(node_brute_kernel, dim3(2048), dim3(kBruteBlock), 0, stream, m, g, d_over.p, over_count,
                       d_coef.p);
    if ((e = hipGetLastError()) != hipSuccess) return fail(e, "node_brute_kernel");
    hipLaunchKernelGGL(sample_kernel, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, stream, d_coef.p, g, b, out);
    if ((e = hipGetLastError()) != hipSuccess) return fail(e, "sample_kernel");
    int flag = 0;
    if ((e = hipMemcpyAsync(&flag, d_err.p, sizeof(int), hipMemcpyDeviceToHost, stream)) != hipSuccess ||
        (e = hipStreamSynchronize(stream)) != hipSuccess)
        return fail(e, "hipStreamSynchronize");
    if (flag) {
        err = (flag & 1) ? "mesh SDF: triangle BVH search stack overflow" : "mesh SDF: node without a nearest triangle";
        return KSCR_ERR_HIP;
    }
    return KSCR_OK;
}

}  // namespace kscr
