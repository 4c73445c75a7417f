// oracle_capi.cpp -- extern "C" access to the CPU restatement (scr_oracle.hpp).
//
// TEST INFRASTRUCTURE ONLY (see the header of scr_oracle.hpp).  The entry
// points mirror include/kscr.h so that parity tests drive the oracle and the
// HIP library with the same call sequence.
#include <cstdint>
#include <cstring>
#include <memory>

#include "adjoint_oracle.hpp"
#include "meshsdf_oracle.hpp"
#include "scr_oracle.hpp"

namespace {
struct Base {
    virtual ~Base() = default;
    virtual int create(const orc::Params& p, int nv, const float* P, int ns, const int* off) = 0;
    virtual int configure(float d) = 0;
    virtual int set_young(float y, float dev) = 0;
    virtual int set_kstretch(float k) = 0;
    virtual int set_kbend(float k) = 0;
    virtual void pin(const int* ids, int m) = 0;
    virtual void fix(const int* ids, int m) = 0;
    virtual int sag_free(const int* roots, int m, int iterations) = 0;
    virtual int set_params(const orc::Params& p) = 0;
    virtual int upload() = 0;
    virtual void set_pos_vel(const float* p, const float* v) = 0;
    virtual void set_pos_vel_t(const double* p, const double* v) = 0;
    virtual void set_last_vel(const float* u) = 0;
    virtual void set_connections(const int* conn) = 0;
    virtual void get_pos_vel(float* p, float* v) const = 0;
    virtual void get_pos_vel64(double* p, double* v) const = 0;
    virtual void get_q(float* q) const = 0;
    virtual void get_q64(double* q) const = 0;
    virtual void set_q(const float* q) = 0;
    virtual void get_rest(float* im, float* lr, float* ks, uint32_t* fl, float* qr) const = 0;
    virtual orc::Sdf& sdf() = 0;
    virtual void set_strain_limit(int on, float err, const float* rest) = 0;
    virtual void advance(float dt) = 0;
    virtual void solve_next_step(float dt) = 0;
    virtual void set_threads(int t) = 0;
    virtual int nverts() const = 0;
    virtual long num_pairs() const = 0;
    virtual int error() const = 0;
    virtual void contacts(int* num, int* cols) const = 0;
    virtual void set_material(const float* ks, const float* qr) = 0;
    virtual void detect_only(float dt) = 0;
    virtual void max_step(float* out) const = 0;
    virtual void get_dx(float* out) const = 0;
    virtual void set_material_t(const double* ks, const double* qr) = 0;
    virtual double gradient(float dt, int nsteps, int K, const int* tips, int ntips, const double* targets, double* ksb,
                            double* qrb, int with_sdf, double* x0b, double* v0b) = 0;
    virtual long clamp_log(int enable, uint8_t* out, long cap) = 0;
};

template <class T>
struct Impl : Base {
    orc::AdjointSim<T> s;
    int create(const orc::Params& p, int nv, const float* P, int ns, const int* off) override {
        return s.create(p, nv, P, ns, off);
    }
    int configure(float d) override { return s.configure(d); }
    int set_young(float y, float dev) override { return s.set_young(y, dev); }
    int set_kstretch(float k) override { return s.set_kstretch(k); }
    int set_kbend(float k) override { return s.set_kbend(k); }
    void pin(const int* ids, int m) override { s.pin(ids, m); }
    void fix(const int* ids, int m) override { s.fix_orientation(ids, m); }
    int sag_free(const int* roots, int m, int iterations) override {
        if (!s.configured) return 7;
        s.sag_free(roots, m, iterations);
        return 0;
    }
    int set_params(const orc::Params& p) override {
        s.p = p;
        return s.configured ? s.validate() : 0;
    }
    int upload() override {
        if (!s.configured) return 7;
        int e = s.validate();
        if (e) return e;
        s.upload();
        return 0;
    }
    void set_pos_vel(const float* p, const float* v) override {
        for (int i = 0; i < s.n; i++) {
            if (p) s.x[i] = orc::V3<T>{(T)p[3 * i], (T)p[3 * i + 1], (T)p[3 * i + 2]};
            if (v) s.v[i] = orc::V3<T>{(T)v[3 * i], (T)v[3 * i + 1], (T)v[3 * i + 2]};
        }
    }
    void set_pos_vel_t(const double* p, const double* v) override {
        for (int i = 0; i < s.n; i++) {
            if (p) s.x[i] = orc::V3<T>{(T)p[3 * i], (T)p[3 * i + 1], (T)p[3 * i + 2]};
            if (v) s.v[i] = orc::V3<T>{(T)v[3 * i], (T)v[3 * i + 1], (T)v[3 * i + 2]};
        }
    }
    void set_connections(const int* conn) override {
        for (int i = 0; i < s.n; i++) s.conn[i] = conn ? conn[i] : -1;
    }
    void set_last_vel(const float* u) override {
        for (int i = 0; i < s.n; i++) s.u[i] = orc::V3<T>{(T)u[3 * i], (T)u[3 * i + 1], (T)u[3 * i + 2]};
    }
    void get_pos_vel(float* p, float* v) const override {
        for (int i = 0; i < s.n; i++) {
            if (p) {
                p[3 * i] = (float)s.x[i].x;
                p[3 * i + 1] = (float)s.x[i].y;
                p[3 * i + 2] = (float)s.x[i].z;
            }
            if (v) {
                v[3 * i] = (float)s.v[i].x;
                v[3 * i + 1] = (float)s.v[i].y;
                v[3 * i + 2] = (float)s.v[i].z;
            }
        }
    }
    void get_pos_vel64(double* p, double* v) const override {
        for (int i = 0; i < s.n; i++) {
            if (p) {
                p[3 * i] = (double)s.x[i].x;
                p[3 * i + 1] = (double)s.x[i].y;
                p[3 * i + 2] = (double)s.x[i].z;
            }
            if (v) {
                v[3 * i] = (double)s.v[i].x;
                v[3 * i + 1] = (double)s.v[i].y;
                v[3 * i + 2] = (double)s.v[i].z;
            }
        }
    }
    void get_q(float* q) const override {
        for (int i = 0; i < s.n; i++) {
            q[4 * i] = (float)s.q[i].x;
            q[4 * i + 1] = (float)s.q[i].y;
            q[4 * i + 2] = (float)s.q[i].z;
            q[4 * i + 3] = (float)s.q[i].w;
        }
    }
    void get_q64(double* q) const override {
        for (int i = 0; i < s.n; i++) {
            q[4 * i] = (double)s.q[i].x;
            q[4 * i + 1] = (double)s.q[i].y;
            q[4 * i + 2] = (double)s.q[i].z;
            q[4 * i + 3] = (double)s.q[i].w;
        }
    }
    void set_q(const float* q) override {
        for (int i = 0; i < s.n; i++) s.q[i] = orc::R4<T>{(T)q[4 * i], (T)q[4 * i + 1], (T)q[4 * i + 2], (T)q[4 * i + 3]};
    }
    void get_rest(float* im, float* lr, float* ks, uint32_t* fl, float* qr) const override {
        for (int i = 0; i < s.n; i++) {
            if (im) im[i] = s.invMassF[i];
            if (lr) lr[i] = s.lRestF[i];
            if (ks) ks[i] = s.kStretchF[i];
            if (fl) fl[i] = s.flags[i];
            if (qr) {
                qr[4 * i] = s.qRestF[i].x;
                qr[4 * i + 1] = s.qRestF[i].y;
                qr[4 * i + 2] = s.qRestF[i].z;
                qr[4 * i + 3] = s.qRestF[i].w;
            }
        }
    }
    orc::Sdf& sdf() override { return s.sdf; }
    void set_strain_limit(int on, float err, const float* rest) override {
        s.strainOn = on != 0;
        s.strainError = err;
        if (!on) return;
        if (rest)
            s.strainRest.assign(rest, rest + s.n);
        else
            s.strain_rest_from_positions();
    }
    void advance(float dt) override { s.advance(dt); }
    void solve_next_step(float dt) override { s.solve_next_step(dt); }
    void set_threads(int t) override { s.threads = t < 1 ? 1 : t; }
    int nverts() const override { return s.n; }
    long num_pairs() const override { return s.num_pairs; }
    int error() const override { return s.error; }
    void set_material(const float* ks, const float* qr) override { s.set_material(ks, qr); }
    void detect_only(float dt) override { s.detect_only(dt); }
    void max_step(float* out) const override {
        for (int i = 0; i < s.n; i++) out[i] = (float)s.maxStep[i];
    }
    void get_dx(float* out) const override {
        for (int i = 0; i < s.n; i++) out[3 * i] = (float)s.dx[i].x, out[3 * i + 1] = (float)s.dx[i].y, out[3 * i + 2] = (float)s.dx[i].z;
    }
    void set_material_t(const double* ks, const double* qr) override { s.set_material_t(ks, qr); }
    double gradient(float dt, int nsteps, int K, const int* tips, int ntips, const double* targets, double* ksb,
                    double* qrb, int with_sdf, double* x0b, double* v0b) override {
        std::vector<int> tp(tips, tips + ntips);
        std::vector<T> tg(3 * ntips);
        for (int k = 0; k < 3 * ntips; k++) tg[k] = (T)targets[k];
        std::vector<T> kb;
        std::vector<orc::R4<T>> qb;
        std::vector<orc::V3<T>> xb, vb;
        const T loss = s.gradient(dt, nsteps, K, tp, tg, kb, qb, with_sdf != 0, &xb, &vb);
        for (int i = 0; i < s.n; i++) {
            if (x0b) x0b[3 * i] = (double)xb[i].x, x0b[3 * i + 1] = (double)xb[i].y, x0b[3 * i + 2] = (double)xb[i].z;
            if (v0b) v0b[3 * i] = (double)vb[i].x, v0b[3 * i + 1] = (double)vb[i].y, v0b[3 * i + 2] = (double)vb[i].z;
        }
        for (int i = 0; i < s.n; i++) {
            ksb[i] = (double)kb[i];
            qrb[4 * i] = (double)qb[i].x, qrb[4 * i + 1] = (double)qb[i].y, qrb[4 * i + 2] = (double)qb[i].z,
                     qrb[4 * i + 3] = (double)qb[i].w;
        }
        return (double)loss;
    }
    long clamp_log(int enable, uint8_t* out, long cap) override {
        const long m = (long)s.clamp_log.size();
        if (out)
            for (long k = 0; k < std::min(m, cap); k++) out[k] = s.clamp_log[k];
        if (enable >= 0) {
            s.log_clamp = enable != 0;
            if (enable) s.clamp_log.clear();
        }
        return m;
    }
    void contacts(int* num, int* cols) const override {
        for (int i = 0; i < s.n; i++) {
            num[i] = (int)s.cols[i].size();
            for (int k = 0; k < 16; k++) cols[16 * i + k] = k < num[i] ? s.cols[i][k] : -1;
        }
    }
};
}  // namespace

static_assert(sizeof(orc::Params) == 27 * 4, "orc::Params must mirror kscr_params");

extern "C" {

struct oracle_handle {
    std::unique_ptr<Base> b;
};

// precision: 32 (float restatement) or 64 (FP64 oracle)
int oracle_create(int precision, const void* params, int nv, const float* pos, int ns, const int* off,
                  oracle_handle** out) {
    if (!params || !pos || !off || !out || nv <= 0 || ns <= 0) return 1;
    auto* h = new oracle_handle;
    if (precision == 64)
        h->b.reset(new Impl<double>);
    else
        h->b.reset(new Impl<float>);
    orc::Params p;
    std::memcpy(&p, params, sizeof(p));
    int e = h->b->create(p, nv, pos, ns, off);
    if (e) {
        delete h;
        return e;
    }
    *out = h;
    return 0;
}
void oracle_destroy(oracle_handle* h) { delete h; }
int oracle_configure(oracle_handle* h, float density) { return h->b->configure(density); }
int oracle_set_young(oracle_handle* h, float y, float dev) { return h->b->set_young(y, dev); }
int oracle_set_kstretch(oracle_handle* h, float k) { return h->b->set_kstretch(k); }
int oracle_set_kbend(oracle_handle* h, float k) { return h->b->set_kbend(k); }
int oracle_pin_vertices(oracle_handle* h, const int* ids, int n) {
    h->b->pin(ids, n);
    return 0;
}
int oracle_fix_orientation(oracle_handle* h, const int* ids, int n) {
    h->b->fix(ids, n);
    return 0;
}
int oracle_sag_free(oracle_handle* h, const int* roots, int n, int iterations) {
    return h->b->sag_free(roots, n, iterations);
}
// body SDF grid of a triangle mesh (meshsdf_oracle.hpp): phi[dim], grad[3 dim]
// on the vertex box (written to gmin/gmax); coef (may be NULL) receives the
// cubic grid's node values if it holds max_coef of them.  Returns the node count.
long oracle_mesh_sdf(const float* verts, int nv, const int* tris, int nt, const int* dim, float* phi, float* grad,
                     float* gmin, float* gmax, double* coef, long max_coef) {
    orc::BodySdfOracle b;
    b.build(verts, nv, tris, nt, dim);
    b.sample(dim, phi, grad);
    for (int a = 0; a < 3; a++) gmin[a] = b.gmin[a], gmax[a] = b.gmax[a];
    const long n = (long)b.grid.coef.size();
    if (coef && max_coef >= n) std::memcpy(coef, b.grid.coef.data(), sizeof(double) * n);
    return n;
}
// signed distance, nearest point and nearest triangle of single points (double)
int oracle_mesh_distance(const float* verts, int nv, const int* tris, int nt, const double* p, int np, double* dist,
                         double* nearest, int* tri) {
    orc::MeshSdfOracle m;
    m.setup(verts, nv, tris, nt);
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < np; i++) {
        orc::D3 c{0, 0, 0};
        dist[i] = m.signed_distance(orc::D3{p[3 * i], p[3 * i + 1], p[3 * i + 2]}, c, tri + i);
        nearest[3 * i] = c.x, nearest[3 * i + 1] = c.y, nearest[3 * i + 2] = c.z;
    }
    return 0;
}
// the cubic cell's shape functions at xi in [-1, 1]^3: N[32], dN[32][3]
void oracle_cubic_shape(const double* xi, double* N, double* dN) {
    orc::CubicGridOracle::shape(xi, N, reinterpret_cast<double(*)[3]>(dN));
}
// node positions of the cubic grid over the padded vertex box: pos[3 n]
long oracle_cubic_nodes(const float* verts, int nv, const int* dim, double* pos, long max_nodes) {
    std::vector<orc::D3> v(nv);
    for (int i = 0; i < nv; i++) v[i] = orc::D3{verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]};
    orc::CubicGridOracle g;
    const unsigned res[3] = {(unsigned)dim[0], (unsigned)dim[1], (unsigned)dim[2]};
    g.setup(v, res);
    const long n = (long)g.num_nodes();
    if (pos && max_nodes >= n)
        for (long l = 0; l < n; l++) {
            const orc::D3 x = g.node_position((size_t)l);
            pos[3 * l] = x.x, pos[3 * l + 1] = x.y, pos[3 * l + 2] = x.z;
        }
    return n;
}
// the cubic grid over the padded vertex box with the given node values,
// interpolated at points: value[np], grad[3 np]
int oracle_cubic_interpolate(const float* verts, int nv, const int* dim, const double* coef, const double* p, int np,
                             double* value, double* grad) {
    std::vector<orc::D3> v(nv);
    for (int i = 0; i < nv; i++) v[i] = orc::D3{verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]};
    orc::CubicGridOracle g;
    const unsigned res[3] = {(unsigned)dim[0], (unsigned)dim[1], (unsigned)dim[2]};
    g.setup(v, res);
    g.coef.assign(coef, coef + g.num_nodes());
    for (int i = 0; i < np; i++) value[i] = g.interpolate(p + 3 * i, grad + 3 * i);
    return 0;
}
int oracle_set_params(oracle_handle* h, const void* params) {
    orc::Params p;
    std::memcpy(&p, params, sizeof(p));
    return h->b->set_params(p);
}
int oracle_upload(oracle_handle* h) { return h->b->upload(); }
int oracle_set_pos_vel(oracle_handle* h, const float* p, const float* v) {
    h->b->set_pos_vel(p, v);
    return 0;
}
// Vertex::connectionIndex of every vertex (NULL: none), as kscr_set_connections.
int oracle_set_connections(oracle_handle* h, const int* conn) {
    h->b->set_connections(conn);
    return 0;
}
int oracle_set_last_vel(oracle_handle* h, const float* u) {
    h->b->set_last_vel(u);
    return 0;
}
int oracle_get_pos_vel(oracle_handle* h, float* p, float* v) {
    h->b->get_pos_vel(p, v);
    return 0;
}
int oracle_get_pos_vel64(oracle_handle* h, double* p, double* v) {
    h->b->get_pos_vel64(p, v);
    return 0;
}
int oracle_get_q(oracle_handle* h, float* q) {
    h->b->get_q(q);
    return 0;
}
int oracle_get_q64(oracle_handle* h, double* q) {
    h->b->get_q64(q);
    return 0;
}
int oracle_set_q(oracle_handle* h, const float* q) {
    h->b->set_q(q);
    return 0;
}
int oracle_get_rest(oracle_handle* h, float* im, float* lr, float* ks, uint32_t* fl, float* qr) {
    h->b->get_rest(im, lr, ks, fl, qr);
    return 0;
}
// Grid spacing as in SDFBodyCollisionSolver::InitParams (:142-159):
// ds = (max - min) / (1.f * dim), invDs = 1.f / ds, all in float.
int oracle_set_sdf(oracle_handle* h, const int* dim, const float* gmin, const float* gmax, const float* phi,
                   const float* grad, float thresh, const float* rot, const float* trans) {
    orc::Sdf& s = h->b->sdf();
    size_t cells = 1;
    for (int a = 0; a < 3; a++) {
        if (dim[a] <= 0) return 1;
        s.dim[a] = dim[a];
        s.gmin[a] = gmin[a];
        s.gmax[a] = gmax[a];
        float dist = gmax[a] - gmin[a];
        s.ds[a] = dist / (1.f * dim[a]);
        s.inv_ds[a] = 1.f / s.ds[a];
        cells *= (size_t)dim[a];
    }
    s.thresh = thresh;
    for (int a = 0; a < 4; a++) s.rot[a] = rot ? rot[a] : (a == 3 ? 1.f : 0.f);
    for (int a = 0; a < 3; a++) s.trans[a] = trans ? trans[a] : 0.f;
    s.phi.assign(phi, phi + cells);
    s.grad.assign(grad, grad + 3 * cells);
    s.enabled = true;
    return 0;
}
int oracle_enable_sdf(oracle_handle* h, int on) {
    h->b->sdf().enabled = on != 0;
    return 0;
}
// strain limiting after the SDF pass; rest float[n] or null for the
// current segment lengths
int oracle_set_strain_limit(oracle_handle* h, int on, float err, const float* rest) {
    h->b->set_strain_limit(on, err, rest);
    return 0;
}
int oracle_advance(oracle_handle* h, float dt) {
    h->b->advance(dt);
    return 0;
}
int oracle_solve_next_step(oracle_handle* h, float dt) {
    h->b->solve_next_step(dt);
    return 0;
}
int oracle_set_threads(oracle_handle* h, int t) {
    h->b->set_threads(t);
    return 0;
}
int oracle_num_verts(oracle_handle* h) { return h->b->nverts(); }
long oracle_num_pairs(oracle_handle* h) { return h->b->num_pairs(); }
int oracle_error(oracle_handle* h) { return h->b->error(); }
void oracle_contacts(oracle_handle* h, int* num, int* cols) { h->b->contacts(num, cols); }
// initItr + forced detection + step limits of one substep of dt (no iterations)
void oracle_detect_only(oracle_handle* h, float dt) { h->b->detect_only(dt); }
void oracle_max_step(oracle_handle* h, float* out) { h->b->max_step(out); }
void oracle_dx(oracle_handle* h, float* out) { h->b->get_dx(out); }
int oracle_set_material(oracle_handle* h, const float* ks, const float* qr) {
    h->b->set_material(ks, qr);
    return 0;
}
// Positions and velocities in double (the FP64 oracle's state without a
// float rounding; either pointer may be null).
int oracle_set_pos_vel64(oracle_handle* h, const double* p, const double* v) {
    h->b->set_pos_vel_t(p, v);
    return 0;
}
int oracle_set_material64(oracle_handle* h, const double* ks, const double* qr) {
    h->b->set_material_t(ks, qr);
    return 0;
}
// Loss and dLoss/d(kStretch, qRest) of nsteps steps (checkpoints every K);
// with_sdf: each step followed by the SDF pass; x0b, v0b (may be null):
// dLoss/d(initial x, v).
double oracle_gradient(oracle_handle* h, float dt, int nsteps, int K, const int* tips, int ntips, const double* targets,
                       double* ksb, double* qrb, int with_sdf, double* x0b, double* v0b) {
    return h->b->gradient(dt, nsteps, K, tips, ntips, targets, ksb, qrb, with_sdf, x0b, v0b);
}

// Branches of the adaptive initial guess's clamp (scr_oracle.hpp init_itr),
// one byte per vertex and substep, logged while enabled: enable = 1 clears
// the log and starts logging; enable = 0 stops; enable < 0 changes nothing.
// Copies up to cap bytes of
// the log as it was before the call into out (may be NULL) and returns its
// length.
long oracle_clamp_log(oracle_handle* h, int enable, uint8_t* out, long cap) {
    return h->b->clamp_log(enable, out, cap);
}

// ---- LBVH restatement (lbvh_oracle.hpp) for the bit-exact tree tests
struct oracle_lbvh {
    orc::Lbvh t;
    std::vector<orc::Aabb> boxes;
};
static std::vector<orc::Aabb> to_boxes(const float* a, int n) {
    std::vector<orc::Aabb> b(n);
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) b[i].mn[k] = a[6 * i + k], b[i].mx[k] = a[6 * i + 3 + k];
    return b;
}
oracle_lbvh* oracle_lbvh_build(const float* aabbs, int n) {
    auto* o = new oracle_lbvh;
    o->boxes = to_boxes(aabbs, n);
    o->t.build(o->boxes);
    return o;
}
void oracle_lbvh_refit(oracle_lbvh* o, const float* aabbs) {
    o->boxes = to_boxes(aabbs, o->t.n);
    o->t.refit(o->boxes);
}
void oracle_lbvh_destroy(oracle_lbvh* o) { delete o; }
long oracle_lbvh_query(oracle_lbvh* o, int* pairs, long cap) {
    auto pr = o->t.query(o->boxes);
    for (long k = 0; k < (long)pr.size() && k < cap; k++) pairs[2 * k] = pr[k].first, pairs[2 * k + 1] = pr[k].second;
    return (long)pr.size();
}
void oracle_lbvh_get(oracle_lbvh* o, uint32_t* codes, uint32_t* ids, uint32_t* links, float* bounds, uint32_t* lp,
                     float* root, int* max_stack) {
    const int n = o->t.n;
    for (int i = 0; i < n; i++) {
        if (codes) codes[i] = o->t.codes[i];
        if (ids) ids[i] = o->t.ids[i];
        if (lp) lp[i] = o->t.leaf_parent[i];
    }
    for (int i = 0; i < n - 1; i++) {
        const orc::LbvhNode& nd = o->t.nodes[i];
        if (links) links[4 * i] = nd.parent, links[4 * i + 1] = nd.left, links[4 * i + 2] = nd.right, links[4 * i + 3] = nd.fence;
        if (bounds)
            for (int c = 0; c < 2; c++)
                for (int k = 0; k < 3; k++) bounds[12 * i + 6 * c + k] = nd.b[c].mn[k], bounds[12 * i + 6 * c + 3 + k] = nd.b[c].mx[k];
    }
    if (root)
        for (int k = 0; k < 3; k++) root[k] = o->t.root.mn[k], root[3 + k] = o->t.root.mx[k];
    if (max_stack) *max_stack = o->t.max_stack;
}
// Contact state of the step oracle after the last detection.
long oracle_num_pairs(oracle_handle* h);
int oracle_error(oracle_handle* h);

// Primitive access for unit tests.
void oracle_rotor_mul(const float* a, const float* b, float* out) {
    orc::R4<float> r = orc::rmul(orc::R4<float>{a[0], a[1], a[2], a[3]}, orc::R4<float>{b[0], b[1], b[2], b[3]});
    out[0] = r.x, out[1] = r.y, out[2] = r.z, out[3] = r.w;
}
void oracle_rotor_rotate(const float* q, const float* v, float* out) {
    orc::V3<float> r = orc::rotate(orc::R4<float>{q[0], q[1], q[2], q[3]}, orc::V3<float>{v[0], v[1], v[2]});
    out[0] = r.x, out[1] = r.y, out[2] = r.z;
}
void oracle_rotor_from_to(const float* a, const float* b, float* out) {
    orc::R4<float> r = orc::from_to(orc::V3<float>{a[0], a[1], a[2]}, orc::V3<float>{b[0], b[1], b[2]});
    out[0] = r.x, out[1] = r.y, out[2] = r.z, out[3] = r.w;
}
void oracle_segment_closest_points(const float* a0, const float* a1, const float* b0, const float* b1, float* uv) {
    orc::segment_closest_points(orc::V3<float>{a0[0], a0[1], a0[2]}, orc::V3<float>{a1[0], a1[1], a1[2]},
                                orc::V3<float>{b0[0], b0[1], b0[2]}, orc::V3<float>{b1[0], b1[1], b1[2]}, uv[0], uv[1]);
}
uint32_t oracle_morton3(int x, int y, int z) { return orc::morton3(x, y, z); }
// Solves H x = f with H packed (d0,d1,d2,m01,m02,m12), glm inverse semantics.
void oracle_sym_solve(const float* H, const float* f, float* out) {
    orc::Sym3<float> S{{H[0], H[1], H[2], H[3], H[4], H[5]}};
    orc::V3<float> r = orc::sym_solve(S, orc::V3<float>{f[0], f[1], f[2]});
    out[0] = r.x, out[1] = r.y, out[2] = r.z;
}
// One contact pair (orc::contact_pair) in double: in = xi, xj, dxi, dxj, b0,
// b1, db0, db1 (24 values); out = f (3), H (6), f2 (3), H2 (6), starting from 0.
void oracle_contact_pair(const void* params, const double* in, double damp, double* out) {
    const orc::Params& p = *(const orc::Params*)params;
    orc::V3<double> v[8];
    for (int k = 0; k < 8; k++) v[k] = orc::V3<double>{in[3 * k], in[3 * k + 1], in[3 * k + 2]};
    orc::V3<double> f{0, 0, 0}, f2{0, 0, 0};
    orc::Sym3<double> H{{0, 0, 0, 0, 0, 0}}, H2{{0, 0, 0, 0, 0, 0}};
    orc::contact_pair<double>(p, v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], damp, f, H, f2, H2);
    out[0] = f.x, out[1] = f.y, out[2] = f.z;
    for (int k = 0; k < 6; k++) out[3 + k] = H.d[k];
    out[9] = f2.x, out[10] = f2.y, out[11] = f2.z;
    for (int k = 0; k < 6; k++) out[12 + k] = H2.d[k];
}
// det_logf of the FP32 restatement (vectorised).
void oracle_det_logf(const float* d, float* out, long n) {
    for (long i = 0; i < n; i++) out[i] = orc::det_logf(d[i]);
}
// Deterministic wind functions of the FP32 restatement (vectorised over n).
void oracle_wind_cos(const double* x, double* out, long n) {
    for (long i = 0; i < n; i++) out[i] = orc::wind_cos(x[i]);
}
void oracle_wind_powf(const float* b, const float* e, float* out, long n) {
    for (long i = 0; i < n; i++) out[i] = orc::wind_powf(b[i], e[i]);
}
}  // extern "C"
