// solver_cli.cpp -- command-line driver of the host solver (tests and
// integration checks).
//
//   kscr_solver_cli parse <settings.json>
//       prints the parsed JsonParams as JSON (no GPU)
//   kscr_solver_cli prep <settings.json> <hair.in> <hair.out>
//       curve preparation of Construct only (no GPU); prints rootVertices and
//       strandRefs as JSON
//   kscr_solver_cli run <settings.json> <hair.in|-> <sdf.bin|mesh.obj|-> <steps> <dt> <hair.out>
//       Construct, then `steps` SolveNextStep(dt); writes the Hair.  A .obj
//       body is a triangle mesh whose SDF grid (64 x 128 x 64) is built on
//       the device.  When the settings name a YarnBall scene (YBJsonPath),
//       the Hair may be "-" and the output is the simulation state instead:
//       i32 n, f32 positions[3n], f32 velocities[3n]
//   kscr_solver_cli ybscene <scene.json> <state.out>
//       the host part of a YarnBall scene (no GPU): prints the parameters,
//       curve offsets, closed flags, connections and pinned vertices as JSON;
//       writes i32 n, f32 positions[3n] (as read), f32 positions[3n] (after
//       glueEndpoints)
//   kscr_solver_cli mesh2sdf <mesh.obj> <dim_x> <dim_y> <dim_z> <sdf.out>
//       the body SDF grid of a triangle mesh, built on the device
//   kscr_solver_cli usc2hair <in.data> <hair.out> [max_strands max_length]
//   kscr_solver_cli hair2usc <hair.in> <name>            (writes name.data)
//   kscr_solver_cli curves2hair <in.bcc|.obj|.poly> <hair.out> <target_seg_len> <resample 0|1>
//                               <break_closed 0|1> [m0,...,m15]
//       prints the closed flags as JSON; m: column-major transform
//   kscr_solver_cli hair2bcc <hair.in> <out.bcc> <polyline 0|1>
//   kscr_solver_cli hair2obj <hair.in> <out.obj>
//   kscr_solver_cli cmrcoords <points.f32> <start> <end> <tar>
//       one two-direction resampling pass over the control points (raw
//       float32 xyz): prints the spline parameters as JSON
//   (no GPU for the file conversions)
//
// Hair files: see Hair::read.  SDF files: i32 dim[3], f32 min[3], f32 max[3],
// f32 phi[n], f32 grad[3n].
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <vector>

#include "curve_io.h"
#include "scr_solver.h"

using namespace kami_amd;

namespace {

BodySdf read_sdf(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("cannot open SDF file " + path);
    BodySdf s;
    in.read((char*)s.dim, sizeof s.dim);
    in.read((char*)s.grid_min, sizeof s.grid_min);
    in.read((char*)s.grid_max, sizeof s.grid_max);
    const size_t n = (size_t)s.dim[0] * s.dim[1] * s.dim[2];
    s.phi.resize(n);
    s.grad.resize(3 * n);
    in.read((char*)s.phi.data(), 4 * (std::streamsize)n);
    in.read((char*)s.grad.data(), 12 * (std::streamsize)n);
    if (!in) throw std::runtime_error("short SDF file " + path);
    return s;
}

void write_sdf(const std::string& path, const BodySdf& s) {
    std::ofstream out(path, std::ios::binary);
    if (!out) throw std::runtime_error("cannot write SDF file " + path);
    out.write((const char*)s.dim, sizeof s.dim);
    out.write((const char*)s.grid_min, sizeof s.grid_min);
    out.write((const char*)s.grid_max, sizeof s.grid_max);
    out.write((const char*)s.phi.data(), 4 * (std::streamsize)s.phi.size());
    out.write((const char*)s.grad.data(), 4 * (std::streamsize)s.grad.size());
    if (!out) throw std::runtime_error("cannot write SDF file " + path);
}

bool ends_with(const std::string& s, const std::string& tail) {
    return s.size() >= tail.size() && s.compare(s.size() - tail.size(), tail.size(), tail) == 0;
}

void print_params(const StableCosseratRodsSolver::JsonParams& p) {
    std::printf(
        "{\"YBJsonPath\": \"%s\", \"EnableResample\": %d, \"EnableSDF\": %d, \"EnableHairHairColl\": %d, "
        "\"UseYoungsModule\": %d, \"NumItr\": %d, \"SagFreeItr\": %d, \"SubdivCoeff\": %d, \"StretchK\": %.9g, "
        "\"BendK\": %.9g, \"YoungsModule\": %.9g, \"StiffnessDeviationCoeff\": %.9g, \"CurveRadius\": %.9g, "
        "\"Density\": %.9g, \"ResampleLen\": %.9g, \"VelocityDecay\": %.9g, \"ExternalForceMultiplier\": %.9g, "
        "\"WindPeak\": %.9g, \"WindYFreq\": %.9g, \"WindZFreq\": %.9g, \"WindTimeFreq\": %.9g, \"WindSharpness\": %.9g, "
        "\"HasGravity\": %d, \"Gravity\": [%.9g, %.9g, %.9g], \"SDFThreshold\": %.9g, \"StrainError\": %.9g, "
        "\"UseStrainLimiting\": %d, \"MaxSubstep\": %.9g, \"Schedule\": %d, \"BlockSize\": %d, "
        "\"DeviceResidentHair\": %d}\n",
        p.YBJsonPath.c_str(), p.EnableResample, p.EnableSDF, p.EnableHairHairColl, p.UseYoungsModule, p.NumItr,
        p.SagFreeItr, p.SubdivCoeff, p.StretchK, p.BendK, p.YoungsModule, p.StiffnessDeviationCoeff, p.CurveRadius,
        p.Density, p.ResampleLen, p.VelocityDecay, p.ExternalForceMultiplier, p.WindPeak, p.WindYFreq, p.WindZFreq,
        p.WindTimeFreq, p.WindSharpness, p.HasGravity, p.Gravity[0], p.Gravity[1], p.Gravity[2], p.SDFThreshold,
        p.StrainError, p.UseStrainLimiting, p.MaxSubstep, p.Schedule, p.BlockSize, p.DeviceResidentHair);
}

void print_refs(const StableCosseratRodsSolver& s) {
    std::printf("{\"rootVertices\": [");
    for (size_t i = 0; i < s.RootVertices().size(); i++) std::printf("%s%d", i ? ", " : "", s.RootVertices()[i]);
    std::printf("], \"strandRefs\": [");
    for (size_t i = 0; i < s.StrandRefs().size(); i++) std::printf("%s%d", i ? ", " : "", s.StrandRefs()[i]);
    std::printf("], \"numVerts\": %d}\n", s.NumSimVertices());
}

// the used vertices of a Hair as a flat state: positions and strand offsets
void flatten(const Hair& h, std::vector<float>& pos, std::vector<int>& off) {
    off.assign(1, 0);
    pos.clear();
    for (uint32_t s = 0; s < h.num_strands; s++) {
        for (uint32_t v = 0; v < h.strand_vert_count[s]; v++) {
            const Vec3 p = h.get(h.current, s, v);
            pos.insert(pos.end(), {p.x, p.y, p.z});
        }
        off.push_back(off.back() + (int)h.strand_vert_count[s]);
    }
}

Transform parse_transform(const char* text) {
    Transform t;
    std::string str(text);
    std::stringstream ss(str);
    std::string item;
    int k = 0;
    while (std::getline(ss, item, ',')) {
        if (k >= 16) throw std::runtime_error("transform: 16 numbers expected");
        t.m[k++] = std::stof(item);
    }
    if (k != 16) throw std::runtime_error("transform: 16 numbers expected");
    return t;
}

void write_state(const std::string& path, int n, const float* a, const float* b) {
    std::ofstream out(path, std::ios::binary);
    if (!out) throw std::runtime_error("cannot write " + path);
    out.write((const char*)&n, 4);
    out.write((const char*)a, 12 * (std::streamsize)n);
    out.write((const char*)b, 12 * (std::streamsize)n);
    if (!out) throw std::runtime_error("cannot write " + path);
}

template <class T>
void print_list(const char* key, const std::vector<T>& v, const char* fmt, bool last = false) {
    std::printf("\"%s\": [", key);
    for (size_t i = 0; i < v.size(); i++) {
        if (i) std::printf(", ");
        std::printf(fmt, v[i]);
    }
    std::printf("]%s", last ? "" : ", ");
}

void print_yarn_scene(const YarnScene& sc) {
    const kscr_params& p = sc.params;
    std::printf("{\"curveFile\": \"%s\", \"resample\": %d, \"resampleLength\": %.9g, \"breakUpClosedCurves\": %d, ",
                sc.curve_file.c_str(), (int)sc.resample, sc.resample_length, (int)sc.break_up_closed_curves);
    std::printf("\"transform\": [");
    for (int k = 0; k < 16; k++) std::printf("%s%.9g", k ? ", " : "", sc.transform[k]);
    std::printf("], \"density\": %.9g, \"kStretch\": %.9g, \"kBend\": %.9g, \"glueRadius\": %.9g, ", sc.density,
                sc.k_stretch, sc.k_bend, sc.glue_radius);
    std::printf(
        "\"params\": {\"gravity\": [%.9g, %.9g, %.9g], \"num_itr\": %d, \"max_h\": %.9g, \"damping\": %.9g, "
        "\"velocity_decay\": %.9g, \"radius\": %.9g, \"barrier_thickness\": %.9g, \"acceleration_ratio\": %.9g, "
        "\"k_collision\": %.9g, \"k_friction\": %.9g, \"friction_coeff\": %.9g, \"detection_scaler\": %.9g, "
        "\"bvh_rebuild_period\": %.9g, \"detection_period\": %d, \"use_step_size_limit\": %d, "
        "\"use_velocity_radius\": %d, \"enable_hair_hair_coll\": %d, \"external_force_multiplier\": %.9g, "
        "\"wind_peak\": %.9g, \"wind_y_freq\": %.9g, \"wind_z_freq\": %.9g, \"wind_time_freq\": %.9g, "
        "\"wind_sharpness\": %.9g}, ",
        p.gravity[0], p.gravity[1], p.gravity[2], p.num_itr, p.max_h, p.damping, p.velocity_decay, p.radius,
        p.barrier_thickness, p.acceleration_ratio, p.k_collision, p.k_friction, p.friction_coeff, p.detection_scaler,
        p.bvh_rebuild_period, p.detection_period, p.use_step_size_limit, p.use_velocity_radius,
        p.enable_hair_hair_coll, p.external_force_multiplier, p.wind_peak, p.wind_y_freq, p.wind_z_freq,
        p.wind_time_freq, p.wind_sharpness);
    print_list("strandOffsets", sc.strand_offsets, "%d");
    print_list("closed", sc.closed, "%d");
    print_list("connections", sc.connections, "%d");
    print_list("pinned", sc.pinned, "%d", true);
    std::printf("}\n");
}

int usage() {
    std::fprintf(stderr,
                 "usage: kscr_solver_cli parse <json> | prep <json> <hair.in> <hair.out> | "
                 "run <json> <hair.in|-> <sdf.bin|mesh.obj|-> <steps> <dt> <out> | ybscene <scene.json> <state.out> | "
                 "mesh2sdf <mesh.obj> <dx> <dy> <dz> <sdf.out> | (file conversions: see the source)\n");
    return 2;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) return usage();
    const std::string cmd = argv[1];
    try {
        if (cmd == "parse") {
            print_params(StableCosseratRodsSolver::LoadSettingJsonContent(argv[2]));
            return 0;
        }
        if (cmd == "prep" && argc == 5) {
            auto hair = std::make_shared<Hair>(Hair::read(argv[3]));
            StableCosseratRodsSolver s(hair);
            s.LoadSettings(argv[2]);
            s.PrepareCurves();
            hair->write(argv[4]);
            print_refs(s);
            return 0;
        }
        if (cmd == "usc2hair" && (argc == 4 || argc == 6)) {
            const uint32_t ms = argc == 6 ? (uint32_t)std::strtoul(argv[4], nullptr, 10) : 10000u;
            const uint32_t ml = argc == 6 ? (uint32_t)std::strtoul(argv[5], nullptr, 10) : 100u;
            import_usc_hairsalon(argv[2], ms, ml).write(argv[3]);
            return 0;
        }
        if (cmd == "hair2usc" && argc == 4) {
            export_usc_hairsalon(argv[3], Hair::read(argv[2]));
            return 0;
        }
        if (cmd == "curves2hair" && (argc == 7 || argc == 8)) {
            const Transform tf = argc == 8 ? parse_transform(argv[7]) : Transform{};
            const Curves c = read_curves(argv[2], std::stof(argv[4]), tf, std::atoi(argv[6]) != 0, std::atoi(argv[5]) != 0);
            hair_from_curves(c).write(argv[3]);
            std::printf("{\"closed\": [");
            for (size_t i = 0; i < c.closed.size(); i++) std::printf("%s%d", i ? ", " : "", (int)c.closed[i]);
            std::printf("]}\n");
            return 0;
        }
        if ((cmd == "hair2bcc" && argc == 5) || (cmd == "hair2obj" && argc == 4)) {
            std::vector<float> pos;
            std::vector<int> off;
            flatten(Hair::read(argv[2]), pos, off);
            if (cmd == "hair2bcc")
                write_bcc(argv[3], pos, off, std::atoi(argv[4]) != 0);
            else
                write_obj(argv[3], pos, off);
            return 0;
        }
        if (cmd == "mesh2sdf" && argc == 7) {
            std::vector<float> v;
            std::vector<int> t;
            read_obj_mesh(argv[2], v, t);
            BodySdf g;
            for (int a = 0; a < 3; a++) g.dim[a] = std::atoi(argv[3 + a]);
            const size_t n = (size_t)std::max(g.dim[0], 0) * std::max(g.dim[1], 0) * std::max(g.dim[2], 0);
            g.phi.resize(n);
            g.grad.resize(3 * n);
            char err[512] = {0};
            const int rc = kscr_mesh_sdf(-1, v.data(), (int)v.size() / 3, t.data(), (int)t.size() / 3, g.dim,
                                         g.phi.data(), g.grad.data(), g.grid_min, g.grid_max, err, sizeof err);
            if (rc != KSCR_OK) throw std::runtime_error(std::string("kscr_mesh_sdf: ") + err);
            write_sdf(argv[6], g);
            return 0;
        }
        if (cmd == "cmrcoords" && argc == 6) {
            std::ifstream in(argv[2], std::ios::binary);
            if (!in) throw std::runtime_error(std::string("cannot open ") + argv[2]);
            std::vector<Vec3> pts;
            Vec3 p;
            while (in.read((char*)&p, sizeof p)) pts.push_back(p);
            const std::vector<double> ts =
                resample_cmr_coords(pts, std::atof(argv[3]), std::atof(argv[4]), std::atof(argv[5]));
            std::printf("[");
            for (size_t i = 0; i < ts.size(); i++) std::printf("%s%.17g", i ? ", " : "", ts[i]);
            std::printf("]\n");
            return 0;
        }
        if (cmd == "ybscene" && argc == 4) {
            const YarnScene sc = load_yarn_scene(argv[2]);
            write_state(argv[3], (int)sc.positions.size() / 3, sc.positions.data(), sc.glued_positions.data());
            print_yarn_scene(sc);
            return 0;
        }
        if (cmd == "run" && argc == 8) {
            auto hair = std::string(argv[3]) == "-" ? nullptr : std::make_shared<Hair>(Hair::read(argv[3]));
            StableCosseratRodsSolver s(hair);
            const std::string body = argv[4];
            if (ends_with(body, ".obj") || ends_with(body, ".OBJ")) {
                BodyMesh m;
                read_obj_mesh(body, m.vertices, m.triangles);
                s.SubstituteBodyMesh(m);
            } else if (body != "-") {
                s.SubstituteBodySdf(read_sdf(body));
            }
            s.Construct(argv[2]);
            const int steps = std::atoi(argv[5]);
            const float dt = (float)std::atof(argv[6]);
            for (int k = 0; k < steps; k++) s.SolveNextStep(dt);
            if (s.FromYarnScene()) {
                const int n = s.NumSimVertices();
                std::vector<float> pos(3 * (size_t)n), vel(3 * (size_t)n);
                if (kscr_get_pos_vel(s.handle(), pos.data(), vel.data()) != KSCR_OK)
                    throw std::runtime_error("kscr_get_pos_vel failed");
                write_state(argv[7], n, pos.data(), vel.data());
                print_refs(s);
                std::fprintf(stderr, "%s", s.GetInfoString().c_str());
                return 0;
            }
            if (!hair) throw std::runtime_error("a Hair file is needed unless the settings name a YarnBall scene");
            s.DownloadHair();  // the driver's draw-frame download (KamiSimulator.cpp:389)
            hair->write(argv[7]);
            print_refs(s);
            std::fprintf(stderr, "%s", s.GetInfoString().c_str());
            return 0;
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return usage();
}
