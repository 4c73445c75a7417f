// scr_solver.cpp -- see scr_solver.h.
#include "scr_solver.h"

#include <algorithm>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <thread>

#include "json_lite.h"

namespace kami_amd {

// Runs body(i) for i in [0, n) over the host's threads (strands are
// independent in the Hair exchange).
template <class F>
static void ParallelStrands(size_t n, const F& body) {
    const unsigned workers = std::min<unsigned>(16, std::max(1u, std::thread::hardware_concurrency()));
    if (n < 256 || workers == 1) {
        for (size_t i = 0; i < n; i++) body(i);
        return;
    }
    std::vector<std::thread> pool;
    for (unsigned w = 0; w < workers; w++)
        pool.emplace_back([&, w] {
            for (size_t i = w; i < n; i += workers) body(i);
        });
    for (auto& t : pool) t.join();
}

StableCosseratRodsSolver::StableCosseratRodsSolver(std::shared_ptr<Hair> hair, int device)
    : hair_(std::move(hair)), device_(device) {}

StableCosseratRodsSolver::~StableCosseratRodsSolver() {
    if (sim_) kscr_destroy(sim_);
}

void StableCosseratRodsSolver::check(int rc, const char* what) const {
    if (rc == KSCR_OK) return;
    char buf[512] = {0};
    if (sim_) kscr_last_error(sim_, buf, sizeof buf);
    throw std::runtime_error(std::string(what) + " failed (status " + std::to_string(rc) + "): " + buf);
}

StableCosseratRodsSolver::JsonParams StableCosseratRodsSolver::LoadSettingJsonContent(const std::string& path,
                                                                                      std::string* dump) {
    const JsonFile j = JsonFile::load(path);
    if (dump) *dump = j.dump();
    JsonParams p{};
    if (j.CheckKey("YBJsonPath")) p.YBJsonPath = j.GetString("YBJsonPath");
    if (j.CheckKey("EnableResample")) p.EnableResample = j.GetBool("EnableResample");
    if (j.CheckKey("EnableSDF")) p.EnableSDF = j.GetBool("EnableSDF");
    if (j.CheckKey("EnableHairHairColl")) p.EnableHairHairColl = j.GetBool("EnableHairHairColl");
    if (j.CheckKey("UseYoungsModule")) p.UseYoungsModule = j.GetBool("UseYoungsModule");
    if (j.CheckKey("NumItr")) p.NumItr = j.GetInteger("NumItr");
    if (j.CheckKey("SagFreeItr")) p.SagFreeItr = j.GetInteger("SagFreeItr");
    if (j.CheckKey("SubdivCoeff")) p.SubdivCoeff = j.GetInteger("SubdivCoeff");
    if (j.CheckKey("BendK")) p.BendK = j.GetFloat("BendK");
    if (j.CheckKey("StretchK")) p.StretchK = j.GetFloat("StretchK");
    if (j.CheckKey("CurveRadius")) p.CurveRadius = j.GetFloat("CurveRadius");
    if (j.CheckKey("Density")) p.Density = j.GetFloat("Density");
    if (j.CheckKey("YoungsModule")) p.YoungsModule = j.GetFloat("YoungsModule");
    if (j.CheckKey("StiffnessDeviationCoeff")) p.StiffnessDeviationCoeff = j.GetFloat("StiffnessDeviationCoeff");
    if (j.CheckKey("ResampleLen")) p.ResampleLen = j.GetFloat("ResampleLen");
    if (j.CheckKey("VelocityDecay")) p.VelocityDecay = j.GetFloat("VelocityDecay");
    if (j.CheckKey("ExternalForceMultiplier")) p.ExternalForceMultiplier = j.GetFloat("ExternalForceMultiplier");
    if (j.CheckKey("WindPeak")) p.WindPeak = j.GetFloat("WindPeak");
    if (j.CheckKey("WindYFreq")) p.WindYFreq = j.GetFloat("WindYFreq");
    if (j.CheckKey("WindZFreq")) p.WindZFreq = j.GetFloat("WindZFreq");
    if (j.CheckKey("WindTimeFreq")) p.WindTimeFreq = j.GetFloat("WindTimeFreq");
    if (j.CheckKey("WindSharpness")) p.WindSharpness = j.GetFloat("WindSharpness");
    if (j.CheckKey("Gravity")) {
        const std::vector<double> g = j.GetNumbers("Gravity");
        if (g.size() != 3) throw std::runtime_error("Gravity must be an array of 3 numbers");
        p.HasGravity = true;
        for (int k = 0; k < 3; k++) p.Gravity[k] = (float)g[k];
    }
    if (j.CheckKey("SDFThreshold")) p.SDFThreshold = j.GetFloat("SDFThreshold");
    if (j.CheckKey("StrainError")) p.StrainError = j.GetFloat("StrainError");
    if (j.CheckKey("UseStrainLimiting")) p.UseStrainLimiting = j.GetBool("UseStrainLimiting");
    if (j.CheckKey("MaxSubstep")) p.MaxSubstep = j.GetFloat("MaxSubstep");
    if (j.CheckKey("Schedule")) {
        const std::string s = j.GetString("Schedule");
        if (s == "jacobi")
            p.Schedule = KSCR_SCHEDULE_JACOBI;
        else if (s == "colored")
            p.Schedule = KSCR_SCHEDULE_COLORED;
        else
            throw std::runtime_error("Schedule must be \"jacobi\" or \"colored\"");
    }
    if (j.CheckKey("BlockSize")) p.BlockSize = j.GetInteger("BlockSize");
    if (j.CheckKey("DeviceResidentHair")) p.DeviceResidentHair = j.GetBool("DeviceResidentHair");
    return p;
}

void StableCosseratRodsSolver::SubstituteBodySdf(const BodySdf& sdf) {
    const size_t nodes = (size_t)sdf.dim[0] * sdf.dim[1] * sdf.dim[2];
    if (sdf.phi.size() != nodes || sdf.grad.size() != 3 * nodes) throw std::runtime_error("SDF grid size mismatch");
    sdf_.reset(new BodySdf(sdf));
    mesh_.reset();
    if (sim_) apply_body();
}

void StableCosseratRodsSolver::SubstituteBodyMesh(const BodyMesh& mesh) {
    if (mesh.vertices.size() % 3 || mesh.triangles.size() % 3 || mesh.triangles.empty())
        throw std::runtime_error("body mesh: xyz per vertex and three indices per triangle expected");
    mesh_.reset(new BodyMesh(mesh));
    sdf_.reset();
    if (sim_) apply_body();
}

// hands the body to the simulation: a grid as it is, a mesh built on the device
void StableCosseratRodsSolver::apply_body() {
    if (sdf_)
        check(kscr_set_sdf(sim_, sdf_->dim, sdf_->grid_min, sdf_->grid_max, sdf_->phi.data(), sdf_->grad.data(),
                           jp_.SDFThreshold, sdf_->rotation, sdf_->translation),
              "kscr_set_sdf");
    else if (mesh_)
        check(kscr_set_sdf_from_mesh(sim_, mesh_->vertices.data(), (int)mesh_->vertices.size() / 3,
                                     mesh_->triangles.data(), (int)mesh_->triangles.size() / 3, mesh_->dim,
                                     jp_.SDFThreshold, mesh_->rotation, mesh_->translation),
              "kscr_set_sdf_from_mesh");
}

// StableCosseratRodsSolver::BuildYarnBallHair (StableCosseratRodsSolver.cpp:112-173),
// curve part.
void StableCosseratRodsSolver::PrepareCurves() {
    if (!hair_) throw std::runtime_error("no hair substituted");
    Hair& h = *hair_;
    if (jp_.SubdivCoeff > 1) divide_hair(h, (uint32_t)jp_.SubdivCoeff);
    root_vertices_.clear();
    strand_refs_.clear();
    flat_pos_.clear();
    flat_vel_.clear();
    strand_offsets_.assign(1, 0);
    int num_verts = 0;
    for (uint32_t s = 0; s < h.num_strands; s++) {
        std::vector<Vec3> curve;
        for (uint32_t v = 0; v < h.strand_vert_count[s]; v++) curve.push_back(h.get(h.current, s, v));
        if (curve.size() < 4) continue;
        if (jp_.EnableResample) {
            curve = resample_cmr(curve, 1, (double)curve.size() - 2, jp_.ResampleLen);
            const size_t new_len = curve.size();
            if (h.max_verts < new_len) {
                // cap at hairSize.y keeping both ends
                std::vector<Vec3> capped(h.max_verts);
                const float stride = (float)(new_len - 2) / (float)(h.max_verts - 2);
                capped[0] = curve[0];
                for (uint32_t v = 1; v + 1 < h.max_verts; v++) capped[v] = curve[(uint32_t)(stride * v)];
                capped[h.max_verts - 1] = curve.back();
                curve = capped;
            }
            h.vertex_count += (uint64_t)curve.size() - h.strand_vert_count[s];
            h.strand_vert_count[s] = (uint32_t)curve.size();
            for (uint32_t v = 0; v < curve.size(); v++) h.set(h.current, s, v, curve[v]);
        }
        root_vertices_.push_back(num_verts);
        strand_refs_.push_back((int)s);
        num_verts += (int)h.strand_vert_count[s];
        for (const Vec3& p : curve) {
            flat_pos_.push_back(p.x);
            flat_pos_.push_back(p.y);
            flat_pos_.push_back(p.z);
        }
        strand_offsets_.push_back(num_verts);
    }
    flat_vel_.assign(flat_pos_.size(), 0.f);
}

void StableCosseratRodsSolver::build_sim() {
    if (flat_pos_.empty()) throw std::runtime_error("no strand has 4 or more vertices");
    kscr_params p;
    kscr_params_default(&p);
    const float r = jp_.CurveRadius;
    const float ratio = 0.05f;
    p.radius = ratio * r;
    p.barrier_thickness = 2 * (1 - ratio) * r;
    p.num_itr = jp_.NumItr;
    p.velocity_decay = jp_.VelocityDecay;
    p.use_step_size_limit = jp_.EnableHairHairColl;
    p.use_velocity_radius = 1;
    p.enable_hair_hair_coll = jp_.EnableHairHairColl;
    p.external_force_multiplier = jp_.ExternalForceMultiplier;
    p.wind_peak = jp_.WindPeak;
    p.wind_y_freq = jp_.WindYFreq;
    p.wind_z_freq = jp_.WindZFreq;
    p.wind_time_freq = jp_.WindTimeFreq;
    p.wind_sharpness = jp_.WindSharpness;
    if (jp_.HasGravity)
        for (int k = 0; k < 3; k++) p.gravity[k] = jp_.Gravity[k];
    p.max_h = jp_.MaxSubstep;
    p.schedule = jp_.Schedule;
    p.block_size = jp_.BlockSize;
    const int nv = NumSimVertices();
    const int ns = (int)strand_offsets_.size() - 1;
    int rc = kscr_create(&p, nv, flat_pos_.data(), ns, strand_offsets_.data(), device_, &sim_);
    if (rc) throw std::runtime_error(std::string("kscr_create failed: ") + kscr_create_error());
    check(kscr_configure(sim_, jp_.Density), "configure");
    if (jp_.UseYoungsModule) {
        check(kscr_set_young(sim_, jp_.YoungsModule, jp_.StiffnessDeviationCoeff), "setYoung");
    } else {
        check(kscr_set_kbend(sim_, jp_.BendK), "setKBend");
        check(kscr_set_kstretch(sim_, jp_.StretchK), "setKStretch");
    }
    check(kscr_pin_vertices(sim_, root_vertices_.data(), (int)root_vertices_.size()), "pin roots");
    if (jp_.SagFreeItr > 0)
        check(kscr_sag_free(sim_, root_vertices_.data(), (int)root_vertices_.size(), jp_.SagFreeItr), "SagFree");
    check(kscr_upload(sim_), "upload");
    apply_body();
    // the SDF solver's strain limiting (SDFBodyCollisionSolver.cpp:182-187),
    // with rest lengths from the positions at construction, as its
    // InitAdditionalHairData takes them from the Hair (:93-157)
    if (jp_.EnableSDF && jp_.UseStrainLimiting)
        check(kscr_set_strain_limit(sim_, 1, jp_.StrainError, nullptr), "set_strain_limit");
}

// StableCosseratRodsSolver::Construct (StableCosseratRodsSolver.cpp:33-61)
void StableCosseratRodsSolver::LoadSettings(const std::string& settingJsonPath) {
    if (!settingJsonPath.empty() && std::ifstream(settingJsonPath).good())
        jp_ = LoadSettingJsonContent(settingJsonPath, &json_dump_);
    use_yarn_scene_ = !jp_.YBJsonPath.empty() && std::ifstream(jp_.YBJsonPath).good();
    if (jp_.SagFreeItr < 0) throw std::runtime_error("SagFreeItr must be >= 0");
}

void StableCosseratRodsSolver::Construct(const std::string& settingJsonPath) {
    LoadSettings(settingJsonPath);
    if (sim_) {
        kscr_destroy(sim_);
        sim_ = nullptr;
    }
    yarn_.reset();
    if (!use_yarn_scene_) {
        PrepareCurves();
        build_sim();
        return;
    }
    // DivideHair runs before the scene choice (StableCosseratRodsSolver.cpp:40-46)
    if (hair_ && jp_.SubdivCoeff > 1) divide_hair(*hair_, (uint32_t)jp_.SubdivCoeff);
    yarn_.reset(new YarnScene(load_yarn_scene(jp_.YBJsonPath)));
    root_vertices_.clear();
    strand_refs_.clear();
    strand_offsets_ = yarn_->strand_offsets;
    flat_pos_ = yarn_->glued_positions;
    flat_vel_.assign(flat_pos_.size(), 0.f);
    sim_ = build_yarn_sim(*yarn_, device_, jp_.Schedule, jp_.BlockSize);
}

// StableCosseratRodsSolver::SolveNextStep (StableCosseratRodsSolver.cpp:178-198):
// advance, then the SDF pass (here on the device), then the Hair is updated.
void StableCosseratRodsSolver::SolveNextStep(float dt) {
    if (!sim_) throw std::runtime_error("Construct() has not been called");
    if (yarn_) {
        // no strand is exchanged with the Hair, so the body pass cannot reach the yarn
        check(kscr_advance(sim_, dt), "advance");
        return;
    }
    if (jp_.EnableSDF) {
        if (!sdf_ && !mesh_)
            throw std::runtime_error(
                "EnableSDF is set but no body was substituted (SubstituteBodySdf / SubstituteBodyMesh)");
        check(kscr_solve_next_step(sim_, dt), "solve_next_step");
    } else {
        check(kscr_advance(sim_, dt), "advance");
    }
    // The reference downloads every step because its SDF pass runs on the
    // host; here the state stays on the device and a device-resident Hair is
    // brought up to date by the driver's DownloadHair at draw frames.
    if (!jp_.DeviceResidentHair) DownloadHair();
}

// StableCosseratRodsSolver::UploadHair (StableCosseratRodsSolver.cpp:202-221)
void StableCosseratRodsSolver::UploadHair() {
    if (!sim_) throw std::runtime_error("Construct() has not been called");
    // a yarn scene exchanges no strand: the reference re-uploads the state it
    // downloaded last, which leaves the simulation as it is
    if (yarn_) return;
    const Hair& h = *hair_;
    const int nv = NumSimVertices();
    ParallelStrands(strand_refs_.size(), [&](size_t i) {
        const uint32_t u = (uint32_t)strand_refs_[i];
        const int root = root_vertices_[i];
        const int next = (i + 1 == strand_refs_.size()) ? nv : root_vertices_[i + 1];
        for (int v = 0; root + v < next; v++) {
            const Vec3 p = h.get(h.current, u, (uint32_t)v), w = h.get(h.velocity, u, (uint32_t)v);
            float* dp = &flat_pos_[3 * (size_t)(root + v)];
            float* dv = &flat_vel_[3 * (size_t)(root + v)];
            dp[0] = p.x, dp[1] = p.y, dp[2] = p.z;
            dv[0] = w.x, dv[1] = w.y, dv[2] = w.z;
        }
    });
    check(kscr_set_pos_vel(sim_, flat_pos_.data(), flat_vel_.data()), "upload positions and velocities");
}

// StableCosseratRodsSolver::DownloadHair (StableCosseratRodsSolver.cpp:223-243)
void StableCosseratRodsSolver::DownloadHair() {
    if (!sim_) throw std::runtime_error("Construct() has not been called");
    // straight from the library's page-locked mirror into the padded Hair
    // layout, strands split over host threads (the reference's OpenMP loop)
    const float *P = nullptr, *V = nullptr;
    check(kscr_download_pos_vel(sim_, &P, &V), "download positions and velocities");
    if (yarn_) return;  // no strand to exchange
    Hair& h = *hair_;
    const int nv = NumSimVertices();
    ParallelStrands(strand_refs_.size(), [&](size_t i) {
        const uint32_t u = (uint32_t)strand_refs_[i];
        const int root = root_vertices_[i];
        const int next = (i + 1 == strand_refs_.size()) ? nv : root_vertices_[i + 1];
        for (int v = 0; root + v < next; v++) {
            const float* sp = P + 3 * (size_t)(root + v);
            const float* sv = V + 3 * (size_t)(root + v);
            h.set(h.current, u, (uint32_t)v, Vec3{sp[0], sp[1], sp[2]});
            h.set(h.velocity, u, (uint32_t)v, Vec3{sv[0], sv[1], sv[2]});
        }
    });
}

// StableCosseratRodsSolver::GetInfoString (StableCosseratRodsSolver.cpp:245-269)
std::string StableCosseratRodsSolver::GetInfoString() const {
    std::stringstream s;
    s << "Solver: Stable Cosserat Rods (MI355X HIP)" << std::endl;
    s << "SubdivCoeff: " << jp_.SubdivCoeff << std::endl;
    if (jp_.UseYoungsModule) {
        s << "Use Youngs: Yes" << std::endl;
        s << "Youngs: " << jp_.YoungsModule << std::endl;
        s << "Deviation: " << jp_.StiffnessDeviationCoeff << std::endl;
    } else {
        s << "Use Youngs: No" << std::endl;
        s << "StretchK: " << jp_.StretchK << std::endl;
        s << "BendK: " << jp_.BendK << std::endl;
    }
    s << "Sag-free itr: " << jp_.SagFreeItr << std::endl;
    if (yarn_) s << "YarnBall scene: " << jp_.YBJsonPath << " (" << yarn_->curve_file << ")" << std::endl;
    if (sim_) s << "Simulated vertices: " << NumSimVertices() << ", workgroup " << kscr_block_size(sim_) << std::endl;
    s << "Json:" << std::endl << json_dump_;
    return s.str();
}

}  // namespace kami_amd
