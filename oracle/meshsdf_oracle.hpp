// meshsdf_oracle.hpp -- body SDF grid built from a triangle mesh, CPU
// restatement for the GPU builder (kscr_meshsdf.hip).
//
// TEST INFRASTRUCTURE ONLY (see scr_oracle.hpp).
//
// The reference builds the body SDF in two stages:
//   Kami::SDF::Init (src/utilities/Geometry.cpp:5-35): Discregrid's
//     TriangleMeshDistance (extern/Discregrid/geometry/TriangleMeshDistance.h)
//     gives the signed distance -- the nearest point by Eberly's
//     point-triangle distance (:557-749), the sign from the angle-weighted
//     pseudo-normal of the nearest vertex, edge or face (:280-317, :342-441);
//     a CubicLagrangeDiscreteGrid (extern/Discregrid/
//     cubic_lagrange_discrete_grid.cpp) of `dim` cells over the vertex box
//     padded by 0.1 stores that distance at its 32-node serendipity cells
//     (corner nodes and two nodes per cell edge, :496-548, :645-757);
//   PNMesh::StoreSDFQueries (src/PNMesh.cpp:109-171) samples the interpolant
//     (:823-894) at the float nodes meshAABB.min + i ds, ds = extent / dim:
//     inside (value < 0) the normalized interpolant gradient, outside the
//     direction from the grid centre.
// This restatement follows those arithmetic sequences in double, so that the
// GPU builder can be compared bit for bit.  The nearest triangle is found by
// brute force over all triangles; ties of the squared distance go to the
// smaller triangle index (the reference's sphere tree takes the first one its
// traversal meets: the distance is the same, the sign can differ only at
// points equidistant from features with different pseudo-normals).
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <unordered_map>
#include <vector>

namespace orc {

struct D3 {
    double x, y, z;
};
inline D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 operator*(double s, D3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double ddot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline D3 dcross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
// v / |v| (componentwise division, TriangleMeshDistance.h:83-90)
inline D3 dunit(D3 a) {
    const double l = std::sqrt(ddot(a, a));
    return {a.x / l, a.y / l, a.z / l};
}

// nearest feature of a triangle: vertex 0..2, edge 01 / 12 / 02, face
enum MeshFeature { kV0 = 0, kV1, kV2, kE01, kE12, kE02, kFace };

// Squared distance from p to triangle (v0, v1, v2): the minimum of
// |v0 + s e0 + t e1 - p|^2 over s, t >= 0, s + t <= 1 by the region of the
// (s, t) plane the unconstrained minimum falls in, with the squared distance
// taken from the quadratic's value at the minimum (as
// TriangleMeshDistance.h:557-749 computes it, clamped at 0).
inline double point_triangle_d2(D3 p, D3 v0, D3 v1, D3 v2, D3& nearest, int& feature) {
    const D3 d = v0 - p, e0 = v1 - v0, e1 = v2 - v0;
    const double a00 = ddot(e0, e0), a01 = ddot(e0, e1), a11 = ddot(e1, e1);
    const double b0 = ddot(d, e0), b1 = ddot(d, e1), c = ddot(d, d);
    const double det = std::fabs(a00 * a11 - a01 * a01);
    double s = a01 * b1 - a11 * b0, t = a01 * b0 - a00 * b1, d2;
    // the quadratic at (s, t) and at the three corners
    auto q_st = [&]() { return s * (a00 * s + a01 * t + 2 * b0) + t * (a01 * s + a11 * t + 2 * b1) + c; };
    auto at_v0 = [&]() { s = 0, t = 0, feature = kV0, d2 = c; };
    auto at_v1 = [&]() { s = 1, t = 0, feature = kV1, d2 = a00 + 2 * b0 + c; };
    auto at_v2 = [&]() { s = 0, t = 1, feature = kV2, d2 = a11 + 2 * b1 + c; };
    // minimum on the edge t = 0 / s = 0
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
    // interior minimum of the edge s + t = 1 at parameter numer / denom from
    // v2 (toward v1: s) or from v1 (toward v2: t)
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

struct MeshSdfOracle {
    std::vector<D3> v;
    std::vector<std::array<int, 3>> tri;
    std::vector<D3> n_face, n_vert;
    std::vector<std::array<D3, 3>> n_edge;  // per triangle: edges 01, 12, 02

    // angle-weighted pseudo-normals (TriangleMeshDistance.h:342-441)
    void setup(const float* verts, int nv, const int* tris, int nt) {
        v.resize(nv);
        for (int i = 0; i < nv; i++) v[i] = D3{verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]};
        tri.resize(nt);
        for (int i = 0; i < nt; i++) tri[i] = {tris[3 * i], tris[3 * i + 1], tris[3 * i + 2]};
        n_face.resize(nt);
        n_vert.assign(nv, D3{0, 0, 0});
        n_edge.resize(nt);
        std::unordered_map<uint64_t, D3> edge_sum;
        auto key = [&](int i, int j) { return (uint64_t)std::min(i, j) * (uint64_t)nv + (uint64_t)std::max(i, j); };
        auto add_edge = [&](int i, int j, D3 n) {
            auto it = edge_sum.find(key(i, j));
            if (it == edge_sum.end()) edge_sum[key(i, j)] = n;
            else it->second = it->second + n;
        };
        for (int k = 0; k < nt; k++) {
            const D3 a = v[tri[k][0]], b = v[tri[k][1]], c = v[tri[k][2]];
            const D3 n = dunit(dcross(b - a, c - a));
            n_face[k] = n;
            const double w0 = std::acos(std::fabs(ddot(dunit(b - a), dunit(c - a))));
            const double w1 = std::acos(std::fabs(ddot(dunit(a - b), dunit(c - b))));
            const double w2 = std::acos(std::fabs(ddot(dunit(b - c), dunit(a - c))));
            n_vert[tri[k][0]] = n_vert[tri[k][0]] + w0 * n;
            n_vert[tri[k][1]] = n_vert[tri[k][1]] + w1 * n;
            n_vert[tri[k][2]] = n_vert[tri[k][2]] + w2 * n;
            add_edge(tri[k][0], tri[k][1], n);
            add_edge(tri[k][1], tri[k][2], n);
            add_edge(tri[k][0], tri[k][2], n);
        }
        for (auto& n : n_vert) n = dunit(n);
        for (int k = 0; k < nt; k++) {
            n_edge[k][0] = dunit(edge_sum[key(tri[k][0], tri[k][1])]);
            n_edge[k][1] = dunit(edge_sum[key(tri[k][1], tri[k][2])]);
            n_edge[k][2] = dunit(edge_sum[key(tri[k][0], tri[k][2])]);
        }
    }

    D3 pseudo_normal(int k, int feature) const {
        switch (feature) {
            case kV0: return n_vert[tri[k][0]];
            case kV1: return n_vert[tri[k][1]];
            case kV2: return n_vert[tri[k][2]];
            case kE01: return n_edge[k][0];
            case kE12: return n_edge[k][1];
            case kE02: return n_edge[k][2];
            default: return n_face[k];
        }
    }

    // signed distance at p (TriangleMeshDistance::signed_distance), the
    // nearest point and the triangle it lies on
    double signed_distance(D3 p, D3& nearest, int* tri_id = nullptr) const {
        double best = INFINITY;
        int best_k = -1, best_f = kFace;
        nearest = p;
        for (int k = 0; k < (int)tri.size(); k++) {
            D3 q;
            int f;
            const double d2 = point_triangle_d2(p, v[tri[k][0]], v[tri[k][1]], v[tri[k][2]], q, f);
            if (d2 < best) best = d2, best_k = k, best_f = f, nearest = q;
        }
        if (tri_id) *tri_id = best_k;
        const D3 u = p - nearest;
        const double sign = ddot(u, pseudo_normal(best_k, best_f)) >= 0.0 ? 1.0 : -1.0;
        return sign * std::sqrt(best);
    }
};

// Discregrid's CubicLagrangeDiscreteGrid with one field, restated.
struct CubicGridOracle {
    unsigned n[3] = {0, 0, 0};
    double lo[3], cell[3], inv_cell[3];
    std::vector<double> coef;

    size_t num_corner() const { return (size_t)(n[0] + 1) * (n[1] + 1) * (n[2] + 1); }
    size_t num_edge(int a) const {
        size_t m = 1;
        for (int b = 0; b < 3; b++) m *= (b == a) ? n[b] : n[b] + 1;
        return m;
    }
    size_t num_nodes() const { return num_corner() + 2 * (num_edge(0) + num_edge(1) + num_edge(2)); }

    // domain: the vertex box padded by 0.1 (Geometry.cpp:24-29); cell size
    // extent / resolution and its reciprocal (discrete_grid.hpp:17-22)
    void setup(const std::vector<D3>& verts, const unsigned res[3]) {
        double hi[3];
        for (int a = 0; a < 3; a++) lo[a] = INFINITY, hi[a] = -INFINITY;
        for (const D3& p : verts) {
            const double c[3] = {p.x, p.y, p.z};
            for (int a = 0; a < 3; a++) lo[a] = std::min(lo[a], c[a]), hi[a] = std::max(hi[a], c[a]);
        }
        for (int a = 0; a < 3; a++) {
            hi[a] += 0.1;
            lo[a] -= 0.1;
            n[a] = res[a];
            cell[a] = (hi[a] - lo[a]) / (double)res[a];
            inv_cell[a] = 1.0 / cell[a];
        }
    }

    // indexToNodePosition (cubic_lagrange_discrete_grid.cpp:496-548): corner
    // nodes x fastest, then two nodes per x edge (edge index x fastest), per
    // y edge (y fastest, then z) and per z edge (z fastest, then x), at 1/3
    // and 2/3 of the edge
    D3 node_position(size_t l) const {
        const size_t nv = num_corner(), ex = num_edge(0), ey = num_edge(1);
        size_t ijk[3];
        int axis = -1;
        size_t e = 0;
        if (l < nv) {
            ijk[2] = l / ((size_t)(n[1] + 1) * (n[0] + 1));
            const size_t r = l % ((size_t)(n[1] + 1) * (n[0] + 1));
            ijk[1] = r / (n[0] + 1);
            ijk[0] = r % (n[0] + 1);
        } else if (l < nv + 2 * ex) {
            axis = 0, e = l - nv;
            const size_t w = e / 2;
            ijk[2] = w / ((size_t)(n[1] + 1) * n[0]);
            const size_t r = w % ((size_t)(n[1] + 1) * n[0]);
            ijk[1] = r / n[0];
            ijk[0] = r % n[0];
        } else if (l < nv + 2 * (ex + ey)) {
            axis = 1, e = l - nv - 2 * ex;
            const size_t w = e / 2;
            ijk[0] = w / ((size_t)(n[2] + 1) * n[1]);
            const size_t r = w % ((size_t)(n[2] + 1) * n[1]);
            ijk[2] = r / n[1];
            ijk[1] = r % n[1];
        } else {
            axis = 2, e = l - nv - 2 * (ex + ey);
            const size_t w = e / 2;
            ijk[1] = w / ((size_t)(n[0] + 1) * n[2]);
            const size_t r = w % ((size_t)(n[0] + 1) * n[2]);
            ijk[0] = r / n[2];
            ijk[2] = r % n[2];
        }
        double x[3];
        for (int a = 0; a < 3; a++) x[a] = lo[a] + cell[a] * (double)ijk[a];
        if (axis >= 0) x[axis] += (1.0 + (double)(e % 2)) / 3.0 * cell[axis];
        return D3{x[0], x[1], x[2]};
    }

    // the 32 node ids of cell (i, j, k) in the shape-function order
    // (cubic_lagrange_discrete_grid.cpp:698-740)
    void cell_nodes(unsigned i, unsigned j, unsigned k, size_t id[32]) const {
        const size_t nx = n[0], ny = n[1], nz = n[2];
        const size_t pl = (nx + 1) * (ny + 1);
        for (int c = 0; c < 8; c++) id[c] = pl * (k + (c >> 2)) + (nx + 1) * (j + ((c >> 1) & 1)) + i + (c & 1);
        size_t off = num_corner();
        for (int c = 0; c < 4; c++) {  // x edges at (j, k), (j, k+1), (j+1, k), (j+1, k+1)
            const size_t jj = j + (c >> 1), kk = k + (c & 1);
            id[8 + 2 * c] = off + 2 * (nx * (ny + 1) * kk + nx * jj + i);
            id[9 + 2 * c] = id[8 + 2 * c] + 1;
        }
        off += 2 * num_edge(0);
        for (int c = 0; c < 4; c++) {  // y edges at (i, k), (i+1, k), (i, k+1), (i+1, k+1)
            const size_t ii = i + (c & 1), kk = k + (c >> 1);
            id[16 + 2 * c] = off + 2 * (ny * (nz + 1) * ii + ny * kk + j);
            id[17 + 2 * c] = id[16 + 2 * c] + 1;
        }
        off += 2 * num_edge(1);
        for (int c = 0; c < 4; c++) {  // z edges at (i, j), (i, j+1), (i+1, j), (i+1, j+1)
            const size_t ii = i + (c >> 1), jj = j + (c & 1);
            id[24 + 2 * c] = off + 2 * (nz * (nx + 1) * jj + nz * ii + k);
            id[25 + 2 * c] = id[24 + 2 * c] + 1;
        }
    }

    // Serendipity cubic shape functions of the 32-node cell on [-1, 1]^3 and
    // their derivatives, node order as cell_nodes.  Each value is the product
    // sequence of shape_function_ (cubic_lagrange_discrete_grid.cpp:286-477),
    // so that the roundings agree: corner (sx, sy, sz) is
    // fac * ((1 + sx x)(1 + sy y)) * (1 + sz z), fac = (9 |xi|^2 - 19) / 64;
    // the edge node at 1 + 3 s x along x is (9/64 (1 - x^2) (1 + 3 s x)) *
    // ((1 + sy y)(1 + sz z)), and likewise along y (transverse x, z) and z
    // (transverse x, y).
    static void shape(const double xi[3], double N[32], double dN[32][3]) {
        const double x = xi[0], y = xi[1], z = xi[2];
        const double x2 = x * x, y2 = y * y, z2 = z * z;
        const double m[3] = {1.0 - x, 1.0 - y, 1.0 - z}, p[3] = {1.0 + x, 1.0 + y, 1.0 + z};
        auto lin = [&](int a, int s) { return s > 0 ? p[a] : m[a]; };
        const double fac = 1.0 / 64.0 * (9.0 * (x2 + y2 + z2) - 19.0);
        const double gx = 9.0 * (3.0 * x2 + y2 + z2) - 19.0, gy = 9.0 * (x2 + 3.0 * y2 + z2) - 19.0,
                     gz = 9.0 * (x2 + y2 + 3.0 * z2) - 19.0;
        const double t18[3] = {18.0 * x, 18.0 * y, 18.0 * z};
        for (int c = 0; c < 8; c++) {
            const int sx = (c & 1) ? 1 : -1, sy = (c & 2) ? 1 : -1, sz = (c & 4) ? 1 : -1;
            N[c] = fac * (lin(0, sx) * lin(1, sy)) * lin(2, sz);
            dN[c][0] = (sx > 0 ? t18[0] + gx : t18[0] - gx) * (lin(1, sy) * lin(2, sz)) / 64.0;
            dN[c][1] = (lin(0, sx) * lin(2, sz)) * (sy > 0 ? t18[1] + gy : t18[1] - gy) / 64.0;
            dN[c][2] = (lin(0, sx) * lin(1, sy)) * (sz > 0 ? t18[2] + gz : t18[2] - gz) / 64.0;
        }
        const double sq[3] = {x2, y2, z2}, co[3] = {x, y, z};
        // edges along axis a; transverse axes (b, c) in increasing order
        for (int a = 0; a < 3; a++) {
            const int b = a == 0 ? 1 : 0, c = a == 2 ? 1 : 2;
            const double one_m = 1.0 - sq[a];
            const double f = 9.0 / 64.0 * one_m;
            const double lm3 = 1.0 - 3.0 * co[a], lp3 = 1.0 + 3.0 * co[a];
            const double three = 3.0 - 9.0 * sq[a], two = 2.0 * co[a];
            const double dm = -three - two, dp = three - two;  // d/da of (1 - a^2)(1 -+ 3a)
            const double om_m = one_m * lm3, om_p = one_m * lp3;
            for (int e = 0; e < 4; e++) {
                // edge order: x edges (y, z) = (-,-), (-,+), (+,-), (+,+) with z
                // fastest; y edges (x, z) with x fastest; z edges (x, y) with y fastest
                int sb, sc;
                if (a == 0) sb = (e & 2) ? 1 : -1, sc = (e & 1) ? 1 : -1;
                else if (a == 1) sb = (e & 1) ? 1 : -1, sc = (e & 2) ? 1 : -1;
                else sb = (e & 2) ? 1 : -1, sc = (e & 1) ? 1 : -1;
                const double tr = lin(b, sb) * lin(c, sc);
                for (int h = 0; h < 2; h++) {
                    const int j = 8 + 8 * a + 2 * e + h;
                    const double om = h ? om_p : om_m;
                    N[j] = (f * (h ? lp3 : lm3)) * tr;
                    dN[j][a] = (h ? dp : dm) * tr * (9.0 / 64.0);
                    dN[j][b] = ((sb > 0 ? om : -om) * lin(c, sc)) * (9.0 / 64.0);
                    dN[j][c] = ((sc > 0 ? om : -om) * lin(b, sb)) * (9.0 / 64.0);
                }
            }
        }
    }

    // interpolate with gradient (cubic_lagrange_discrete_grid.cpp:823-894):
    // the cell holding x (clamped to the last cell), x mapped to [-1, 1]^3,
    // the coefficients summed in node order
    double interpolate(const double x[3], double grad[3]) const {
        unsigned mi[3];
        for (int a = 0; a < 3; a++) {
            mi[a] = (unsigned)((x[a] - lo[a]) * inv_cell[a]);
            if (mi[a] >= n[a]) mi[a] = n[a] - 1;
        }
        double xi[3], c0[3];
        for (int a = 0; a < 3; a++) {
            const double smin = lo[a] + (double)mi[a] * cell[a];
            const double smax = smin + cell[a];
            const double denom = smax - smin;
            c0[a] = 2.0 / denom;
            const double c1 = (smax + smin) / denom;
            xi[a] = c0[a] * x[a] - c1;
        }
        size_t id[32];
        cell_nodes(mi[0], mi[1], mi[2], id);
        double N[32], dN[32][3];
        shape(xi, N, dN);
        double phi = 0.0, g[3] = {0.0, 0.0, 0.0};
        for (int j = 0; j < 32; j++) {
            const double c = coef[id[j]];
            phi += c * N[j];
            for (int a = 0; a < 3; a++) g[a] += c * dN[j][a];
        }
        for (int a = 0; a < 3; a++) grad[a] = g[a] * c0[a];
        return phi;
    }
};

// The whole reference pipeline: grid coefficients from the mesh's signed
// distance, then PNMesh::StoreSDFQueries on the nodes gmin + i ds of the
// vertex box (float node positions); phi[n], grad[3n], x fastest.
struct BodySdfOracle {
    MeshSdfOracle mesh;
    CubicGridOracle grid;
    float gmin[3], gmax[3];

    void build(const float* verts, int nv, const int* tris, int nt, const int dim[3]) {
        mesh.setup(verts, nv, tris, nt);
        const unsigned res[3] = {(unsigned)dim[0], (unsigned)dim[1], (unsigned)dim[2]};
        grid.setup(mesh.v, res);
        for (int a = 0; a < 3; a++) gmin[a] = INFINITY, gmax[a] = -INFINITY;
        for (int i = 0; i < nv; i++)
            for (int a = 0; a < 3; a++)
                gmin[a] = std::min(gmin[a], verts[3 * i + a]), gmax[a] = std::max(gmax[a], verts[3 * i + a]);
        grid.coef.resize(grid.num_nodes());
#pragma omp parallel for schedule(dynamic, 64)
        for (long long l = 0; l < (long long)grid.coef.size(); l++) {
            D3 c{0, 0, 0};
            grid.coef[l] = mesh.signed_distance(grid.node_position((size_t)l), c);
        }
    }

    void sample(const int dim[3], float* phi, float* grad) const {
        const float ds[3] = {(gmax[0] - gmin[0]) / (float)dim[0], (gmax[1] - gmin[1]) / (float)dim[1],
                             (gmax[2] - gmin[2]) / (float)dim[2]};
        float centre[3];
        for (int a = 0; a < 3; a++) centre[a] = (gmax[a] + gmin[a]) / 2;
        const long long n = (long long)dim[0] * dim[1] * dim[2];
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < n; i++) {
            const int ix = (int)(i % dim[0]), iy = (int)((i / dim[0]) % dim[1]),
                      iz = (int)(i / ((long long)dim[0] * dim[1]));
            const float p[3] = {gmin[0] + ds[0] * (float)ix, gmin[1] + ds[1] * (float)iy, gmin[2] + ds[2] * (float)iz};
            const double x[3] = {p[0], p[1], p[2]};
            double g[3];
            const double d = grid.interpolate(x, g);
            const float dist = (float)d;
            phi[i] = dist;
            if (dist < 0.0f) {
                // Eigen's normalize(): divide by the square root of the squared norm
                const double l = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
                for (int a = 0; a < 3; a++) grad[3 * i + a] = (float)(g[a] / l);
            } else {
                const float t[3] = {centre[0] - p[0], centre[1] - p[1], centre[2] - p[2]};
                const float l = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
                for (int a = 0; a < 3; a++) grad[3 * i + a] = -t[a] / l;
            }
        }
    }
};

}  // namespace orc
