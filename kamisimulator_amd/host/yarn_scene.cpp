// yarn_scene.cpp -- see yarn_scene.h.
#include "yarn_scene.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <fstream>
#include <limits>
#include <stdexcept>
#include <unordered_map>

#include "curve_io.h"
#include "json_lite.h"

namespace kami_amd {

namespace {

bool file_exists(const std::string& p) { return !p.empty() && std::ifstream(p).good(); }

bool ends_with(const std::string& s, const char* tail) {
    const std::string t(tail);
    return s.size() > t.size() && s.compare(s.size() - t.size(), t.size(), t) == 0;
}

// glm::length2 / glm::length of a float vec3: (x*x + y*y) + z*z
inline float length2(const float* a, const float* b) {
    const float x = a[0] - b[0], y = a[1] - b[1], z = a[2] - b[2];
    return (x * x + y * y) + z * z;
}

// Uniform grid over vertex positions for the endpoint search: any vertex
// within `radius` of a point lies in the 27 cells around the point's cell
// (cell edge >= radius).  Cells are hashed; a hash collision only adds
// candidates, which the distance test sorts out.
struct VertexGrid {
    double lo[3], cell;
    std::unordered_map<uint64_t, std::vector<int>> cells;

    VertexGrid(const std::vector<float>& P, float radius) {
        const size_t n = P.size() / 3;
        double hi[3];
        for (int a = 0; a < 3; a++) lo[a] = hi[a] = n ? P[a] : 0.0;
        for (size_t i = 0; i < n; i++)
            for (int a = 0; a < 3; a++) {
                lo[a] = std::min(lo[a], (double)P[3 * i + a]);
                hi[a] = std::max(hi[a], (double)P[3 * i + a]);
            }
        const double extent = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-30});
        // at most ~1e6 cells per axis keeps the integer cell coordinates small
        cell = std::max((double)radius * 1.001, extent * 1e-6);
        for (size_t i = 0; i < n; i++) insert((int)i, &P[3 * i]);
    }
    int64_t coord(double x, int a) const { return (int64_t)std::floor((x - lo[a]) / cell); }
    static uint64_t key(int64_t x, int64_t y, int64_t z) {
        return (uint64_t)x * 0x9E3779B97F4A7C15ull ^ (uint64_t)y * 0xC2B2AE3D27D4EB4Full ^
               (uint64_t)z * 0x165667B19E3779F9ull;
    }
    uint64_t key_of(const float* p) const { return key(coord(p[0], 0), coord(p[1], 1), coord(p[2], 2)); }
    void insert(int i, const float* p) { cells[key_of(p)].push_back(i); }
    void erase(int i, const float* p) {
        auto& v = cells[key_of(p)];
        v.erase(std::find(v.begin(), v.end(), i));
    }
};

}  // namespace

void glue_endpoints(std::vector<float>& P, const std::vector<int>& off, std::vector<int>& conn, float radius) {
    const int n = (int)(P.size() / 3);
    if ((int)conn.size() != n) throw std::runtime_error("glue_endpoints: one connection per vertex expected");
    if (!(radius >= 0)) return;
    // VertexFlags::hasNext: every vertex but the last of its strand
    std::vector<char> has_next(n, 1);
    for (size_t s = 1; s < off.size(); s++)
        if (off[s] >= 1 && off[s] <= n) has_next[off[s] - 1] = 0;
    std::vector<int> ends;
    for (int i = 0; i < n; i++) {
        const bool has_prev = i && has_next[i - 1];
        if (has_prev != (bool)has_next[i]) ends.push_back(i);
    }
    VertexGrid grid(P, radius);
    for (int i : ends) {
        const float pos[3] = {P[3 * i], P[3 * i + 1], P[3 * i + 2]};
        // the closest vertex with |i - j| > 2, the lowest index on ties; only
        // one within the radius can be glued, and those are all in the grid
        // cells around the endpoint
        float min_d = std::numeric_limits<float>::infinity();
        int closest = -1;
        int64_t c[3];
        for (int a = 0; a < 3; a++) c[a] = grid.coord(pos[a], a);
        for (int64_t dx = -1; dx <= 1; dx++)
            for (int64_t dy = -1; dy <= 1; dy++)
                for (int64_t dz = -1; dz <= 1; dz++) {
                    auto it = grid.cells.find(VertexGrid::key(c[0] + dx, c[1] + dy, c[2] + dz));
                    if (it == grid.cells.end()) continue;
                    for (int j : it->second) {
                        if (std::abs(i - j) <= 2) continue;
                        const float d = std::sqrt(length2(pos, &P[3 * (size_t)j]));
                        if (d < min_d || (d == min_d && j < closest)) {
                            min_d = d;
                            closest = j;
                        }
                    }
                }
        if (closest < 0 || !(min_d <= radius)) continue;
        if (conn[closest] >= 0 || conn[i] >= 0) continue;
        conn[closest] = i;
        conn[i] = closest;
        float mid[3];
        for (int a = 0; a < 3; a++) mid[a] = 0.5f * (P[3 * i + a] + P[3 * (size_t)closest + a]);
        for (int v : {i, closest}) {
            grid.erase(v, &P[3 * (size_t)v]);
            for (int a = 0; a < 3; a++) P[3 * (size_t)v + a] = mid[a];
            grid.insert(v, &P[3 * (size_t)v]);
        }
    }
}

YarnScene load_yarn_scene(const std::string& json_path) {
    const JsonFile f = JsonFile::load(json_path);
    const JsonValue& root = f.root();
    YarnScene sc;
    sc.curve_file = root["curveFile"].as_string("curveFile");
    std::string data_path = sc.curve_file;
    if (!file_exists(data_path) && !data_path.empty() && data_path[0] != '/') {
        const size_t slash = json_path.rfind('/');
        if (slash != std::string::npos) {
            const std::string beside = json_path.substr(0, slash + 1) + data_path;
            if (file_exists(beside)) data_path = beside;
        }
    }
    // mat4 transform(1) scaled by 0.01; "transform" gives rows 0-2 of a 3x4
    Transform tf;
    tf.m[0] = tf.m[5] = tf.m[10] = 0.01f;
    const JsonValue& t = root["transform"];
    if (!t.is_null())
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 4; j++) tf.m[j * 4 + i] = t[i][j].as_float("transform element");
    std::copy(tf.m, tf.m + 16, sc.transform);

    sc.resample = !root["resampleLength"].is_null();
    sc.resample_length = root["resampleLength"].as_float("resampleLength");
    sc.break_up_closed_curves = root["breakUpClosedCurves"].as_bool("breakUpClosedCurves");
    Curves cv;
    if (ends_with(data_path, ".poly"))
        cv = read_poly(data_path, sc.resample_length, tf, sc.break_up_closed_curves, sc.resample);
    else if (ends_with(data_path, ".obj"))
        cv = read_obj(data_path, sc.resample_length, tf, sc.break_up_closed_curves, sc.resample);
    else
        cv = read_bcc(data_path, sc.resample_length, tf, sc.break_up_closed_curves, sc.resample);

    // createFromCurves (reader.cpp:16-50)
    sc.strand_offsets.assign(1, 0);
    for (size_t k = 0; k < cv.points.size(); k++) {
        for (const Vec3& p : cv.points[k]) {
            if (!std::isfinite(p.x) || !std::isfinite(p.y) || !std::isfinite(p.z))
                throw std::runtime_error("curve " + std::to_string(k) + " has a non-finite point");
            sc.positions.insert(sc.positions.end(), {p.x, p.y, p.z});
        }
        sc.strand_offsets.push_back((int)(sc.positions.size() / 3));
        sc.closed.push_back(cv.closed[k] ? 1 : 0);
    }
    const int n = (int)(sc.positions.size() / 3);
    if (n < 3) throw std::runtime_error("Too little vertices");
    sc.connections.assign(n, -1);
    for (size_t k = 0; k < cv.points.size(); k++)
        if (cv.closed[k]) {
            const int a = sc.strand_offsets[k], b = sc.strand_offsets[k + 1] - 1;
            sc.connections[a] = b;
            sc.connections[b] = a;
        }

    kscr_params_default(&sc.params);
    const JsonValue& cr = root["curveRadius"];
    if (!cr.is_null()) {
        const float r = cr.as_float("curveRadius");
        constexpr float ratio = 0.05f;
        sc.params.radius = ratio * r;
        sc.params.barrier_thickness = 2 * (1 - ratio) * r;
    }
    const JsonValue& sim = root["simulation"];
    double density = 1., k_stretch = 5e5, k_bend = 1e-1;
    if (!sim.is_null()) {
        kscr_params& p = sc.params;
        auto num = [&](const char* k, float& dst) {
            if (!sim[k].is_null()) dst = sim[k].as_float(k);
        };
        auto dnum = [&](const char* k, double& dst) {
            if (!sim[k].is_null()) dst = sim[k].as_float(k);
        };
        dnum("density", density);
        num("frictionCoeff", p.friction_coeff);
        dnum("kStretch", k_stretch);
        dnum("kBend", k_bend);
        num("kCollision", p.k_collision);
        num("kFriction", p.k_friction);
        num("damping", p.damping);
        const JsonValue& g = sim["gravity"];
        if (!g.is_null())
            for (int a = 0; a < 3; a++) p.gravity[a] = g[a].as_float("gravity element");
        num("drag", p.velocity_decay);
        num("maxTimeStep", p.max_h);
        if (!sim["numIterations"].is_null()) p.num_itr = sim["numIterations"].as_int("numIterations");
        if (!sim["detectionPeriod"].is_null()) p.detection_period = sim["detectionPeriod"].as_int("detectionPeriod");
        num("detectionScaler", p.detection_scaler);
        if (!sim["stepLimit"].is_null()) p.use_step_size_limit = sim["stepLimit"].as_bool("stepLimit") ? 1 : 0;
        if (!sim["velStepLimit"].is_null())
            p.use_velocity_radius = sim["velStepLimit"].as_bool("velStepLimit") ? 1 : 0;
        num("bvhRebuildPeriod", p.bvh_rebuild_period);
    }
    sc.density = (float)density;
    sc.k_stretch = (float)k_stretch;
    sc.k_bend = (float)k_bend;

    sc.glued_positions = sc.positions;
    if (!root["glueEndpoints"].is_null()) {
        sc.glue_radius = root["glueEndpoints"].as_float("glueEndpoints");
        if (sc.glue_radius >= 0) glue_endpoints(sc.glued_positions, sc.strand_offsets, sc.connections, sc.glue_radius);
    }
    const std::vector<float>& P = sc.glued_positions;
    std::vector<char> pin(n, 0);

    const JsonValue& borders = root["fixBorders"];
    if (borders.is_array() && borders.size() == 6) {
        float border[6];
        for (int i = 0; i < 6; i++) border[i] = borders[i].as_float("fixBorders element");
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int i = 0; i < n; i++)
            for (int a = 0; a < 3; a++) {
                lo[a] = std::min(lo[a], P[3 * i + a]);
                hi[a] = std::max(hi[a], P[3 * i + a]);
            }
        const float pad = 0.01f * sc.params.radius;
        for (int a = 0; a < 3; a++) {
            lo[a] -= pad;
            hi[a] += pad;
        }
        hi[0] -= border[0];
        lo[0] += border[1];
        hi[1] -= border[2];
        lo[1] += border[3];
        hi[2] -= border[4];
        lo[2] += border[5];
        for (int i = 0; i < n; i++) {
            bool inside = true;
            for (int a = 0; a < 3; a++) inside = inside && lo[a] <= P[3 * i + a] && hi[a] >= P[3 * i + a];
            if (!inside) pin[i] = 1;
        }
    }

    const JsonValue& fix = root["fixVertex"];
    if (fix.is_array())
        for (size_t k = 0; k < fix.size(); k++) {
            const JsonValue& e = fix[k];
            if (!e.is_array()) {
                const int id = e.as_int("fixVertex index");
                if (id < 0 || id >= n)
                    throw std::runtime_error("fixVertex index " + std::to_string(id) + " is not a vertex (" +
                                             std::to_string(n) + " vertices)");
                pin[id] = 1;
                continue;
            }
            if (e.size() < 3) continue;
            const float c[3] = {e[0].as_float("fixVertex x"), e[1].as_float("fixVertex y"),
                                e[2].as_float("fixVertex z")};
            if (e.size() >= 4) {
                float r2 = e[3].as_float("fixVertex radius");
                r2 *= r2;
                for (int j = 0; j < n; j++)
                    if (length2(&P[3 * (size_t)j], c) < r2) pin[j] = 1;
            } else {
                float min_d = INFINITY;
                int closest = -1;
                for (int j = 0; j < n; j++) {
                    const float d = length2(&P[3 * (size_t)j], c);
                    if (d < min_d) {
                        min_d = d;
                        closest = j;
                    }
                }
                if (closest >= 0) pin[closest] = 1;
            }
        }
    for (int i = 0; i < n; i++)
        if (pin[i]) sc.pinned.push_back(i);
    return sc;
}

kscr_handle* build_yarn_sim(const YarnScene& sc, int device, int schedule, int block_size) {
    kscr_params p = sc.params;
    p.schedule = schedule;
    p.block_size = block_size;
    const int n = (int)(sc.positions.size() / 3);
    const int ns = (int)sc.strand_offsets.size() - 1;
    kscr_handle* h = nullptr;
    if (kscr_create(&p, n, sc.positions.data(), ns, sc.strand_offsets.data(), device, &h) != KSCR_OK)
        throw std::runtime_error(std::string("kscr_create failed: ") + kscr_create_error());
    auto check = [&](int rc, const char* what) {
        if (rc == KSCR_OK) return;
        char buf[512] = {0};
        kscr_last_error(h, buf, sizeof buf);
        kscr_destroy(h);
        throw std::runtime_error(std::string(what) + " failed (status " + std::to_string(rc) + "): " + buf);
    };
    check(kscr_configure(h, sc.density), "configure");
    check(kscr_set_kstretch(h, sc.k_stretch), "setKStretch");
    check(kscr_set_kbend(h, sc.k_bend), "setKBend");
    if (std::any_of(sc.connections.begin(), sc.connections.end(), [](int c) { return c >= 0; }))
        check(kscr_set_connections(h, sc.connections.data()), "connections");
    check(kscr_pin_vertices(h, sc.pinned.data(), (int)sc.pinned.size()), "pin vertices");
    check(kscr_upload(h), "upload");
    // glueEndpoints moved vertices after configure took the rest state
    if (sc.glued_positions != sc.positions) check(kscr_set_pos_vel(h, sc.glued_positions.data(), nullptr), "glued positions");
    return h;
}

}  // namespace kami_amd
