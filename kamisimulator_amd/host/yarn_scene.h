// yarn_scene.h -- YarnBall JSON scenes (the YBJsonPath setting of the
// reference plugin, StableCosseratRodsSolver.cpp:44-45): a curve file,
// its transform, the rod radius, the material and simulation parameters,
// glued endpoints and pinned vertices, built into a simulation over the C ABI.
//
// Follows YarnBall::buildFromJSON and Sim::glueEndpoints
// (YarnBall/io/jsonBuilder.cpp:7-188).  The scene is read in two parts: a
// host-only description (load_yarn_scene: no device needed) and the
// simulation built from it (build_yarn_sim).
#pragma once

#include <string>
#include <vector>

#include "kscr.h"

namespace kami_amd {

struct YarnScene {
    std::string curve_file;  // as named by "curveFile"
    float transform[16];     // column-major; default diag(0.01, 0.01, 0.01, 1)
    bool resample = false;   // "resampleLength" present
    float resample_length = 0.f;
    bool break_up_closed_curves = false;
    kscr_params params;  // Sim::Sim defaults with "curveRadius" and "simulation" applied
    float density = 1.f, k_stretch = 5e5f, k_bend = 1e-1f;
    float glue_radius = -1.f;  // "glueEndpoints"; < 0: no gluing
    // The curves as read (after transform and resampling): positions of the
    // vertices of every curve one after another, curve offsets, closed flags.
    std::vector<float> positions;
    std::vector<int> strand_offsets;
    std::vector<int> closed;
    // Vertex::connectionIndex after createFromCurves and glueEndpoints.
    std::vector<int> connections;
    // Positions after glueEndpoints (glued pairs moved to their midpoint);
    // the rest state is taken from `positions` (configure runs first).
    std::vector<float> glued_positions;
    // Vertices with invMass = 0 (fixBorders and fixVertex), ascending.
    std::vector<int> pinned;
};

// Reads the scene file and the curve file it names, and applies the host
// steps of buildFromJSON: the parameters, glueEndpoints, fixBorders and
// fixVertex.  A relative curveFile that does not exist from the working
// directory is also looked up next to the scene file.  Throws
// std::runtime_error on unreadable files, malformed values and vertex
// indices out of range.
YarnScene load_yarn_scene(const std::string& json_path);

// Sim::glueEndpoints(radius) over positions (xyz per vertex), strand offsets
// and connections, in place.  Endpoints are the first and last vertex of
// every strand, visited in vertex order; each takes the closest vertex j with
// |i - j| > 2 (the lowest index among equal distances) and, when it lies
// within `radius` and neither vertex is connected yet, the two are glued and
// both move to their midpoint.  The reference runs the endpoints over OpenMP
// threads; this is its result with one thread.
void glue_endpoints(std::vector<float>& positions, const std::vector<int>& strand_offsets,
                    std::vector<int>& connections, float radius);

// Sim::Sim(numVerts) + configure(density) + setKStretch + setKBend + the
// connections + the pins + upload, then the glued positions.  Returns the
// handle (owned by the caller); throws std::runtime_error with the
// library's message on failure.
kscr_handle* build_yarn_sim(const YarnScene& scene, int device, int schedule = KSCR_SCHEDULE_JACOBI,
                            int block_size = 0);

}  // namespace kami_amd
