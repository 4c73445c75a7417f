// scr_solver.h -- host-side Stable Cosserat Rods solver over the HIP library.
//
// Mirrors the reference's plugin (src/solvers/StableCosseratRods/
// StableCosseratRodsSolver.h:26-135, the Solver interface of
// src/solvers/Solver.h:13-85): the same virtual methods (Construct,
// SolveNextStep, UploadHair, DownloadHair, GetInfoString), the same JSON
// setting keys with the same defaults, and the same Hair exchange (strands
// with fewer than 4 vertices are skipped and remembered through rootVertices
// and strandRefs).  Simulation calls go through the C ABI of include/kscr.h.
//
// Differences, all additive: the body SDF is handed over as a sampled grid
// (SubstituteBodySdf) or as a triangle mesh whose grid the library builds on
// the device (SubstituteBodyMesh) instead of a Body object; SolveNextStep
// runs the SDF pass on the device instead of the host round trip; the
// settings accept Gravity, SDFThreshold, StrainError, UseStrainLimiting (the
// SDF solver's keys), MaxSubstep, Schedule, BlockSize and DeviceResidentHair (SolveNextStep then leaves the Hair to DownloadHair,
// which the driver calls once per draw frame).
//
// YBJsonPath: when the setting names an existing YarnBall scene file, the
// simulation is built from that scene (yarn_scene.h) instead of the Hair, as
// the reference does (StableCosseratRodsSolver.cpp:44-45): no strand is
// exchanged with the Hair (rootVertices and strandRefs stay empty), the
// scene's own parameters apply, and the settings keys that configure the
// Hair-built simulation are not used (of the additive keys, Schedule and
// BlockSize still apply).  The reference's SolveNextStep still runs its host
// SDF pass over the (unsimulated) Hair, which leaves the yarn untouched; here
// SolveNextStep advances the yarn only.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "hair.h"
#include "kscr.h"
#include "yarn_scene.h"

namespace kami_amd {

// A body SDF sampled on a regular grid (kscr_set_sdf layout).
struct BodySdf {
    int dim[3] = {0, 0, 0};
    float grid_min[3] = {0, 0, 0}, grid_max[3] = {0, 0, 0};
    std::vector<float> phi;   // dim.x * dim.y * dim.z, x fastest
    std::vector<float> grad;  // 3 per node
    float rotation[4] = {0, 0, 0, 1};
    float translation[3] = {0, 0, 0};
};

// A closed triangle body mesh whose SDF grid the library builds on the device
// (PNMesh::GenerateSDF + StoreSDFQueries, src/PNMesh.cpp:100-171); dim is
// the grid resolution (PNMESH_DEFAULT_SDF_SIZE, src/PNMesh.cpp:18).
struct BodyMesh {
    std::vector<float> vertices;  // xyz per vertex
    std::vector<int> triangles;   // three vertex indices per triangle
    int dim[3] = {64, 128, 64};
    float rotation[4] = {0, 0, 0, 1};
    float translation[3] = {0, 0, 0};
};

class StableCosseratRodsSolver {
   public:
    // StableCosseratRodsSolver::JsonParams (StableCosseratRodsSolver.h:40-109):
    // the element names are the JSON keys.
    struct JsonParams {
        std::string YBJsonPath;
        bool EnableResample = true;
        bool EnableSDF = true;
        bool EnableHairHairColl = false;
        bool UseYoungsModule = true;
        int NumItr = 8;
        int SagFreeItr = 0;
        int SubdivCoeff = 3;
        float StretchK = 3e3f;
        float BendK = 6e-4f;
        float YoungsModule = 5;
        float StiffnessDeviationCoeff = 1;
        float CurveRadius = 0.00004f;
        float Density = 1300;
        float ResampleLen = 1e-3f;
        float VelocityDecay = 0.05f;
        float ExternalForceMultiplier = 1;
        float WindPeak = 40.f;
        float WindYFreq = 1.f;
        float WindZFreq = 0.7f;
        float WindTimeFreq = 0.4f;
        float WindSharpness = 3.f;
        // additive keys
        bool HasGravity = false;
        float Gravity[3] = {0.f, -9.8f, 0.f};
        float SDFThreshold = 0.0f;
        // the SDF solver's strain limiting (SDFBodyCollisionSolver.h:26-27)
        float StrainError = 0.1f;
        bool UseStrainLimiting = false;
        float MaxSubstep = 4e-4f;  // Sim::maxH
        int Schedule = KSCR_SCHEDULE_JACOBI;
        int BlockSize = 0;
        bool DeviceResidentHair = false;
    };

    explicit StableCosseratRodsSolver(std::shared_ptr<Hair> hair, int device = -1);
    ~StableCosseratRodsSolver();
    StableCosseratRodsSolver(const StableCosseratRodsSolver&) = delete;
    StableCosseratRodsSolver& operator=(const StableCosseratRodsSolver&) = delete;

    // Solver::SubstituteHair / SubstituteBody
    void SubstituteHair(std::shared_ptr<Hair> hair) { hair_ = std::move(hair); }
    void SubstituteBodySdf(const BodySdf& sdf);
    // The body as a mesh: its SDF grid is built on the device (replaces a
    // substituted grid, and the other way round).
    void SubstituteBodyMesh(const BodyMesh& mesh);

    // Solver interface
    void Construct(const std::string& settingJsonPath = "");
    void SolveNextStep(float dt);  // dt in seconds
    void UploadHair();
    void DownloadHair();
    std::string GetInfoString() const;

    // Reads the JSON keys (StableCosseratRodsSolver.cpp:76-110 plus the
    // additive ones); *dump receives the file text.
    static JsonParams LoadSettingJsonContent(const std::string& jsonPath, std::string* dump = nullptr);
    // The settings part of Construct: reads the JSON file when it exists,
    // checks SagFreeItr and notes whether YBJsonPath names an existing scene.
    void LoadSettings(const std::string& settingJsonPath);
    // Curve preparation of Construct without the simulation: DivideHair when
    // SubdivCoeff > 1, then per strand (>= 4 vertices) the Catmull-Rom
    // resample capped at hairSize.y; fills the flat curve arrays.
    void PrepareCurves();

    const JsonParams& params() const { return jp_; }
    const std::vector<int>& RootVertices() const { return root_vertices_; }
    const std::vector<int>& StrandRefs() const { return strand_refs_; }
    int NumSimVertices() const { return (int)flat_pos_.size() / 3; }
    kscr_handle* handle() const { return sim_; }
    // true when the simulation was built from the YBJsonPath scene
    bool FromYarnScene() const { return yarn_ != nullptr; }
    const YarnScene* Scene() const { return yarn_.get(); }

   private:
    void check(int rc, const char* what) const;
    void build_sim();

    std::shared_ptr<Hair> hair_;
    int device_;
    JsonParams jp_{};
    std::string json_dump_;
    kscr_handle* sim_ = nullptr;
    std::unique_ptr<BodySdf> sdf_;
    std::unique_ptr<BodyMesh> mesh_;
    std::unique_ptr<YarnScene> yarn_;
    bool use_yarn_scene_ = false;
    void apply_body();
    std::vector<int> root_vertices_, strand_refs_;
    std::vector<float> flat_pos_, flat_vel_;
    std::vector<int> strand_offsets_;
};

}  // namespace kami_amd
