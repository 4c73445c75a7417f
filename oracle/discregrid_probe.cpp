// discregrid_probe.cpp -- runs the reference's body SDF construction on a
// mesh and writes what it computes.
//
// TEST INFRASTRUCTURE ONLY.  Built by `make -C oracle ref` from the
// reference's own sources where they lie (src/utilities/Geometry.cpp, the
// vendored Discregrid and Eigen), output into oracle/_ref/.  PNMesh.cpp needs
// assimp, which the image lacks, so the node loop of PNMesh::StoreSDFQueries
// (src/PNMesh.cpp:115-168) is spelled out below around the reference's own
// Kami::SDF::CollisionTest / Distance.  tests/golden/make_mesh_sdf_vectors.py
// runs it and stores the results as tests/golden/mesh_sdf_vectors.npz.
//
//   discregrid_probe <in> <out>
//   in:  i32 nv, nt, dim[3], nq; f32 vertices[3 nv]; i32 triangles[3 nt];
//        f64 query points[3 nq]
//   out: u64 n_nodes; f64 node coefficients[n_nodes];
//        f32 phi[dim], grad[3 dim] (x fastest);
//        f64 signed distance[nq], nearest point[3 nq]; i32 triangle[nq]
#include <cstdint>
#include <cstdio>
#include <limits>
#include <vector>

#include "utilities/Geometry.h"

template <class T>
static bool rd(FILE* f, T* p, size_t n) {
    return fread(p, sizeof(T), n, f) == n;
}
template <class T>
static void wr(FILE* f, const T* p, size_t n) {
    fwrite(p, sizeof(T), n, f);
}

int main(int argc, char** argv) {
    if (argc != 3) {
        fprintf(stderr, "usage: discregrid_probe <in> <out>\n");
        return 2;
    }
    FILE* in = fopen(argv[1], "rb");
    if (!in) return 2;
    int32_t hdr[6];
    if (!rd(in, hdr, 6)) return 2;
    const int nv = hdr[0], nt = hdr[1], nq = hdr[5];
    const glm::uvec3 dim(hdr[2], hdr[3], hdr[4]);
    std::vector<float> v(3 * (size_t)nv);
    std::vector<int32_t> t(3 * (size_t)nt);
    std::vector<double> q(3 * (size_t)nq);
    if (!rd(in, v.data(), v.size()) || !rd(in, t.data(), t.size()) || !rd(in, q.data(), q.size())) return 2;
    fclose(in);

    // PNMesh's arrays and mesh box (src/PNMesh.cpp:30-63)
    auto positions = std::make_unique<glm::vec3[]>(nv);
    auto tris = std::make_unique<glm::uvec3[]>(nt);
    glm::vec3 lo(std::numeric_limits<float>::max()), hi(-std::numeric_limits<float>::max());
    for (int i = 0; i < nv; i++) {
        positions[i] = glm::vec3(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
        lo = glm::min(lo, positions[i]);
        hi = glm::max(hi, positions[i]);
    }
    for (int i = 0; i < nt; i++) tris[i] = glm::uvec3(t[3 * i], t[3 * i + 1], t[3 * i + 2]);

    Kami::SDF sdf;
    sdf.Init(positions, tris, nv, nt, dim);
    const std::vector<double>& coef = sdf.GetGridNodes()[0];

    // PNMesh::StoreSDFQueries
    const int total = (int)(dim.x * dim.y * dim.z);
    std::vector<float> phi(total), grad(3 * (size_t)total);
    const glm::vec3 ds = glm::vec3((hi.x - lo.x) / dim.x, (hi.y - lo.y) / dim.y, (hi.z - lo.z) / dim.z);
#pragma omp parallel for
    for (int i = 0; i < total; i++) {
        auto pos = lo;
        const int ix = i % dim.x, iy = (i / dim.x) % dim.y, iz = i / (dim.x * dim.y);
        pos += ds.x * (ix) * glm::vec3(1.f, 0.f, 0.f);
        pos += ds.y * (iy) * glm::vec3(0.f, 1.f, 0.f);
        pos += ds.z * (iz) * glm::vec3(0.f, 0.f, 1.f);
        glm::vec3 cp, n;
        float dist;
        if (sdf.CollisionTest(pos, cp, n, dist)) {
            phi[i] = dist;
            grad[3 * i] = n.x, grad[3 * i + 1] = n.y, grad[3 * i + 2] = n.z;
        } else {
            phi[i] = (float)sdf.Distance(pos);
            glm::vec3 centre = (hi + lo) / 2.f;
            glm::vec3 to = centre - pos;
            const float l = glm::length(to);
            grad[3 * i] = -to.x / l, grad[3 * i + 1] = -to.y / l, grad[3 * i + 2] = -to.z / l;
        }
    }

    // TriangleMeshDistance at the query points, on the mesh as Geometry.cpp builds it
    std::vector<Eigen::Vector3d> pos(nv);
    std::vector<std::array<unsigned int, 3>> tri(nt);
    for (int i = 0; i < nv; i++) pos[i] = Eigen::Vector3d(positions[i].x, positions[i].y, positions[i].z);
    for (int i = 0; i < nt; i++) tri[i] = {tris[i].x, tris[i].y, tris[i].z};
    Discregrid::TriangleMesh mesh(pos, tri);
    Discregrid::TriangleMeshDistance md(mesh);
    std::vector<double> sd(nq), near(3 * (size_t)nq);
    std::vector<int32_t> tid(nq);
    for (int i = 0; i < nq; i++) {
        const auto r = md.signed_distance(Eigen::Vector3d(q[3 * i], q[3 * i + 1], q[3 * i + 2]));
        sd[i] = r.distance;
        for (int a = 0; a < 3; a++) near[3 * i + a] = r.nearest_point[a];
        tid[i] = r.triangle_id;
    }

    FILE* out = fopen(argv[2], "wb");
    if (!out) return 2;
    const uint64_t nn = coef.size();
    wr(out, &nn, 1);
    wr(out, coef.data(), coef.size());
    wr(out, phi.data(), phi.size());
    wr(out, grad.data(), grad.size());
    wr(out, sd.data(), sd.size());
    wr(out, near.data(), near.size());
    wr(out, tid.data(), tid.size());
    fclose(out);
    return 0;
}
