// curve_io.cpp -- groom and curve file formats (see curve_io.h).
#include "curve_io.h"

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <unordered_map>

namespace kami_amd {

namespace {

struct File {
    FILE* f = nullptr;
    File(const std::string& path, const char* mode) : f(std::fopen(path.c_str(), mode)) {
        if (!f) throw std::runtime_error("cannot open " + path);
    }
    ~File() {
        if (f) std::fclose(f);
    }
    void read(void* p, size_t bytes, const std::string& what) {
        if (bytes && std::fread(p, 1, bytes, f) != bytes) throw std::runtime_error("short file reading " + what);
    }
    void write(const void* p, size_t bytes) {
        if (bytes && std::fwrite(p, 1, bytes, f) != bytes) throw std::runtime_error("write failed");
    }
};

// CheckShrinkenIndexOrNot (USCHairsalonImporter.cpp:10-12); a zero stride,
// which the reference divides by, keeps every index
bool keep_index(uint32_t query, uint32_t shrunk, uint32_t from_file) {
    if (query == 0) return true;
    const uint32_t stride = shrunk ? from_file / shrunk : 0;
    return stride == 0 || query % stride == 0;
}

// the BCC header (YB/YarnBall.h:22-31)
#pragma pack(push, 1)
struct BccHeader {
    char sign[3];
    unsigned char byte_count;
    char curve_type[2];
    char dimensions;
    char up_dimension;
    uint64_t curve_count;
    uint64_t total_control_point_count;
    char file_info[40];
};
#pragma pack(pop)
static_assert(sizeof(BccHeader) == 64, "BCC header is 64 bytes");

std::vector<Vec3> maybe_resample(const std::vector<Vec3>& pts, bool resample, float tar) {
    if (!resample || pts.size() < 4) return pts;
    return resample_cmr(pts, 1.0, (double)pts.size() - 2.0, (double)tar);
}

}  // namespace

Vec3 Transform::apply(Vec3 p) const {
    // glm mat4 * vec4: (m0 x + m1 y) + (m2 z + m3 w) per row, columns m0..m3
    float r[3];
    for (int i = 0; i < 3; i++)
        r[i] = (m[0 + i] * p.x + m[4 + i] * p.y) + (m[8 + i] * p.z + m[12 + i] * 1.0f);
    return Vec3{r[0], r[1], r[2]};
}

Hair import_usc_hairsalon(const std::string& path, uint32_t max_strands, uint32_t max_length) {
    File in(path, "rb");
    uint32_t num = 0;
    in.read(&num, 4, "strand count");
    const bool full = max_strands == 10000 && max_length == 100;
    const uint32_t hx = full ? num : std::min(max_strands, num);
    const uint32_t hy = full ? 100u : std::min(100u, max_length);
    Hair h(hx, hy);
    if (hx == 0) return h;
    const uint32_t iden = num / hx;
    std::vector<Vec3> strand;
    for (uint32_t i = 0; i < num; i++) {
        uint32_t length = 0;
        in.read(&length, 4, "strand length");
        strand.resize(length);
        in.read(strand.data(), sizeof(Vec3) * length, "strand vertices");
        if (full) {
            if (length > hy)
                throw std::runtime_error("strand " + std::to_string(i) + " has " + std::to_string(length) +
                                         " vertices, more than the importer's 100");
            h.strand_vert_count[i] = length;
            h.vertex_count += length;
            for (uint32_t j = 0; j < length; j++) h.set(h.initial, i, j, strand[j]);
            continue;
        }
        if (!keep_index(i, hx, num)) continue;
        const uint32_t u = i / iden;
        if (u >= hx) continue;  // past the shrunk array (the reference writes beyond it)
        h.strand_vert_count[u] = std::min(hy, length);
        h.vertex_count += h.strand_vert_count[u];
        for (uint32_t j = 0; j < length; j++) {
            if (!keep_index(j, max_length, length)) continue;
            const uint32_t jden = (length > max_length) ? (length / max_length) : 1;
            const uint32_t v = j / jden;
            if (v < hy) h.set(h.initial, u, v, strand[j]);
        }
    }
    h.current = h.initial;
    return h;
}

void export_usc_hairsalon(const std::string& name, const Hair& hair) {
    File out(name + ".data", "wb");
    out.write(&hair.num_strands, 4);
    for (uint32_t s = 0; s < hair.num_strands; s++) {
        const uint32_t n = hair.strand_vert_count[s];
        out.write(&n, 4);
        for (uint32_t v = 0; v < n; v++) {
            const Vec3 p = hair.get(hair.current, s, v);
            out.write(&p, sizeof p);
        }
    }
}

Curves read_bcc(const std::string& path, float tar, const Transform& tf, bool break_closed, bool allow_resample) {
    File in(path, "rb");
    BccHeader hd;
    in.read(&hd, sizeof hd, "BCC header");
    if (hd.sign[0] != 'B' || hd.sign[1] != 'C' || hd.sign[2] != 'C' || hd.byte_count != 0x44)
        throw std::runtime_error("Unsupported BCC file");
    const bool polyline = hd.curve_type[0] == 'P' && hd.curve_type[1] == 'L';
    if (!polyline && (hd.curve_type[0] != 'C' || hd.curve_type[1] != '0'))
        throw std::runtime_error("Not polyline or uniform CMR spline");
    if (hd.dimensions != 3) throw std::runtime_error("Only curves in 3D are supported");
    Curves c;
    for (uint64_t i = 0; i < hd.curve_count; i++) {
        int32_t num = 0;
        in.read(&num, 4, "curve length");
        const bool closed = num < 0;
        const uint32_t count = (uint32_t)std::abs(num);
        std::vector<Vec3> pts(count);
        in.read(pts.data(), sizeof(Vec3) * count, "curve points");
        for (auto& p : pts) p = tf.apply(p);
        if (count < 3) continue;
        if (allow_resample && !polyline) pts = resample_cmr(pts, 1.0, (double)pts.size() - 2.0, (double)tar);
        c.points.push_back(pts);
        c.closed.push_back(closed && !break_closed);
    }
    return c;
}

Curves read_obj(const std::string& path, float tar, const Transform& tf, bool break_closed, bool allow_resample) {
    std::ifstream file(path);
    if (!file) throw std::runtime_error("cannot open " + path);
    std::vector<Vec3> verts;
    std::vector<std::vector<int>> lines;
    std::string line;
    while (std::getline(file, line)) {
        std::istringstream iss(line);
        std::string prefix;
        iss >> prefix;
        if (prefix == "v") {
            Vec3 p{0, 0, 0};
            iss >> p.x >> p.y >> p.z;
            verts.push_back(tf.apply(p));
        } else if (prefix == "l") {
            std::vector<int> idx;
            int k;
            while (iss >> k) idx.push_back(k);
            if (!idx.empty()) lines.push_back(idx);
        }
    }
    Curves c;
    for (const auto& l : lines) {
        std::vector<Vec3> curve;
        curve.reserve(l.size());
        bool closed = l.front() == l.back();
        for (int k : l) {
            if (k < 1 || k > (int)verts.size()) throw std::runtime_error("OBJ line index out of range");
            curve.push_back(verts[(size_t)k - 1]);
        }
        if (break_closed && closed) {
            curve.pop_back();
            closed = false;
        }
        curve = maybe_resample(curve, allow_resample, tar);
        if (curve.size() < 4) continue;
        c.points.push_back(curve);
        c.closed.push_back(closed);
    }
    return c;
}

Curves read_poly(const std::string& path, float tar, const Transform& tf, bool break_closed, bool allow_resample) {
    std::ifstream file(path);
    if (!file) throw std::runtime_error("cannot open " + path);
    std::string line;
    bool points_section = true;
    std::vector<Vec3> pts;
    std::vector<std::pair<int, int>> segs;
    while (std::getline(file, line)) {
        std::istringstream iss(line);
        if (line == "POINTS") {
            points_section = true;
            continue;
        } else if (line == "POLYS") {
            points_section = false;
            continue;
        } else if (line == "END") {
            break;
        }
        char colon = 0;
        if (points_section) {
            uint32_t id;
            float x, y, z;
            if (!(iss >> id >> colon >> x >> y >> z) || colon != ':')
                throw std::runtime_error("Error parsing .poly. Point parse error.");
            pts.push_back(tf.apply(Vec3{x, y, z}));
            if (id != pts.size()) throw std::runtime_error("Error parsing .poly. Point id mismatch.");
        } else {
            int id, i, j;
            if (!(iss >> id >> colon >> i >> j) || colon != ':')
                throw std::runtime_error("Error parsing .poly. Segment parse error.");
            if (i < 1 || j < 1 || i > (int)pts.size() || j > (int)pts.size())
                throw std::runtime_error("Error parsing .poly. Segment index out of range.");
            segs.push_back({i - 1, j - 1});
        }
    }
    std::unordered_multimap<int, int> links;
    for (auto s : segs) {
        links.insert({s.first, s.second});
        links.insert({s.second, s.first});
    }
    Curves c;
    std::vector<bool> visited(pts.size(), false);
    // open curves, from their ends (every open curve is resampled, as in the
    // reference)
    for (uint32_t i = 0; i < pts.size(); i++) {
        if (visited[i]) continue;
        const size_t con = links.count((int)i);
        if (con > 2) throw std::runtime_error("Graphs not allowed.");
        if (con > 1) continue;
        std::vector<Vec3> curve;
        int cur = (int)i;
        while (!visited[cur]) {
            visited[cur] = true;
            curve.push_back(pts[cur]);
            auto range = links.equal_range(cur);
            if (range.first == range.second) break;
            for (auto it = range.first; it != range.second; ++it)
                if (!visited[it->second]) {
                    cur = it->second;
                    break;
                }
        }
        if (curve.size() < 4) continue;
        curve = resample_cmr(curve, 1.0, (double)curve.size() - 2.0, (double)tar);
        c.points.push_back(curve);
        c.closed.push_back(false);
    }
    // closed curves
    for (uint32_t i = 0; i < pts.size(); i++) {
        if (visited[i]) continue;
        if (links.count((int)i) != 2) throw std::runtime_error("Graphs not allowed.");
        std::vector<Vec3> curve;
        int cur = (int)i;
        for (;;) {
            visited[cur] = true;
            curve.push_back(pts[cur]);
            bool found = false, at_start = false;
            auto range = links.equal_range(cur);
            if (range.first == range.second) break;
            for (auto it = range.first; it != range.second; ++it) {
                if (!visited[it->second]) {
                    cur = it->second;
                    found = true;
                    break;
                }
                if ((uint32_t)it->second == i) at_start = true;
            }
            if (!found) {
                if (!at_start) throw std::runtime_error("Closed curve not closed.");
                break;
            }
        }
        curve = maybe_resample(curve, allow_resample, tar);
        if (curve.size() < 4) continue;
        c.points.push_back(curve);
        c.closed.push_back(!break_closed);
    }
    return c;
}

Curves read_curves(const std::string& path, float tar, const Transform& tf, bool break_closed, bool allow_resample) {
    const size_t dot = path.rfind('.');
    const std::string ext = dot == std::string::npos ? "" : path.substr(dot + 1);
    if (ext == "bcc") return read_bcc(path, tar, tf, break_closed, allow_resample);
    if (ext == "obj") return read_obj(path, tar, tf, break_closed, allow_resample);
    if (ext == "poly") return read_poly(path, tar, tf, break_closed, allow_resample);
    throw std::runtime_error("unknown curve file extension \"" + ext + "\"");
}

void write_bcc(const std::string& path, const std::vector<float>& pos, const std::vector<int>& off, bool as_polyline) {
    BccHeader hd;
    std::memset(&hd, 0, sizeof hd);
    hd.sign[0] = 'B', hd.sign[1] = 'C', hd.sign[2] = 'C';
    hd.byte_count = 0x44;
    hd.curve_type[0] = as_polyline ? 'P' : 'C';
    hd.curve_type[1] = as_polyline ? 'L' : '0';
    hd.dimensions = 3;
    hd.up_dimension = 1;
    // every strand with a segment is one curve (runs of hasNext vertices)
    for (size_t s = 0; s + 1 < off.size(); s++) {
        const int n = off[s + 1] - off[s];
        if (n < 2) continue;
        hd.curve_count++;
        hd.total_control_point_count += (uint64_t)n;
    }
    File out(path, "wb");
    out.write(&hd, sizeof hd);
    for (size_t s = 0; s + 1 < off.size(); s++) {
        const int32_t n = off[s + 1] - off[s];
        if (n < 2) continue;
        out.write(&n, 4);
        out.write(&pos[3 * (size_t)off[s]], 12 * (size_t)n);
    }
}

void write_obj(const std::string& path, const std::vector<float>& pos, const std::vector<int>& off) {
    File out(path, "w");
    std::fprintf(out.f, "# YarnBall Sim\n\n");
    std::fprintf(out.f, "o YarnBall\n\n");
    std::fprintf(out.f, "# Vertices (in meters)\n");
    const size_t n = pos.size() / 3;
    for (size_t i = 0; i < n; i++)
        std::fprintf(out.f, "v %.16f %.16f %.16f\n", (double)pos[3 * i], (double)pos[3 * i + 1], (double)pos[3 * i + 2]);
    std::fprintf(out.f, "\n# Curves\n");
    for (size_t s = 0; s + 1 < off.size(); s++) {
        if (off[s + 1] - off[s] < 2) continue;
        std::fprintf(out.f, "l %d", off[s] + 1);
        for (int i = off[s]; i + 1 < off[s + 1]; i++) std::fprintf(out.f, " %d", i + 2);
        std::fprintf(out.f, "\n");
    }
}

Hair hair_from_curves(const Curves& c) {
    uint32_t mv = 0;
    for (const auto& p : c.points) mv = std::max(mv, (uint32_t)p.size());
    Hair h((uint32_t)c.points.size(), mv);
    for (uint32_t s = 0; s < h.num_strands; s++) {
        h.strand_vert_count[s] = (uint32_t)c.points[s].size();
        h.vertex_count += c.points[s].size();
        for (uint32_t v = 0; v < h.strand_vert_count[s]; v++) h.set(h.initial, s, v, c.points[s][v]);
    }
    h.current = h.initial;
    return h;
}

void read_obj_mesh(const std::string& path, std::vector<float>& vertices, std::vector<int>& triangles) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open mesh file " + path);
    vertices.clear();
    triangles.clear();
    std::string line;
    std::vector<int> face;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string tag;
        if (!(ls >> tag)) continue;
        if (tag == "v") {
            float x, y, z;
            if (!(ls >> x >> y >> z)) throw std::runtime_error("mesh file " + path + ": bad vertex line");
            vertices.insert(vertices.end(), {x, y, z});
        } else if (tag == "f") {
            face.clear();
            std::string item;
            const long nv = (long)vertices.size() / 3;
            while (ls >> item) {
                const long k = std::strtol(item.c_str(), nullptr, 10);  // the position index before any '/'
                const long idx = k > 0 ? k - 1 : nv + k;
                if (k == 0 || idx < 0 || idx >= nv) throw std::runtime_error("mesh file " + path + ": bad face index");
                face.push_back((int)idx);
            }
            if (face.size() < 3) throw std::runtime_error("mesh file " + path + ": face with fewer than 3 corners");
            for (size_t c = 1; c + 1 < face.size(); c++) triangles.insert(triangles.end(), {face[0], face[c], face[c + 1]});
        }
    }
    if (triangles.empty()) throw std::runtime_error("mesh file " + path + ": no faces");
}

}  // namespace kami_amd
