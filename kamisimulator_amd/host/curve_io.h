// curve_io.h -- groom and curve file formats of KamiSimulator and YarnBall,
// for loading real grooms into the solver and dumping trajectories.
//
//   USC-HairSalon .data: u32 strand count, then per strand u32 n and n xyz
//   floats (src/importers/USCHairsalonImporter.cpp:60-85,
//   src/exporters/USCHairsalonExporter.cpp:6-18).
//   BCC (binary curve collection): 64-byte header {"BCC", 0x44, "PL" |
//   "C0", dimensions 3, up dimension, u64 curve count, u64 control point
//   count, 40 info bytes}, then per curve i32 n (negative: closed) and n xyz
//   floats (YB/io/reader.cpp:52-97, YB/io/fileExport.cpp:6-57).
//   OBJ polylines ("v x y z", "l i j k ..."; YB/io/reader.cpp:99-150,
//   fileExport.cpp:59-92) and the .poly point/segment lists
//   (YB/io/reader.cpp:152-285).
#pragma once

#include <string>
#include <vector>

#include "hair.h"

namespace kami_amd {

// ImportUSCHairSalonToHair with its defaults (maxStrandCount 10000, maxLength
// 100) reads every strand into hairSize.y = 100; any other pair takes every
// (count / max)-th strand and vertex (ShrinkenImportUSCHairSalon).  A strand
// longer than hairSize.y, which the reference writes past its arrays, is an
// error here.
Hair import_usc_hairsalon(const std::string& path, uint32_t max_strands = 10000, uint32_t max_length = 100);
// ExportUSCHairCore: the current vertices of every strand to `name` + ".data".
void export_usc_hairsalon(const std::string& name, const Hair& hair);

// Curves as YarnBall's readers hand them to createFromCurves.
struct Curves {
    std::vector<std::vector<Vec3>> points;
    std::vector<bool> closed;
};

// A column-major 4x4 matrix applied to every point as glm's mat4 * vec4(p, 1).
struct Transform {
    float m[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    Vec3 apply(Vec3 p) const;
};

// readFromBCC / readFromOBJ / readFromPoly: uniform Catmull-Rom curves
// ("C0", OBJ and .poly when allow_resample) resampled at target_seg_len
// between their second and second-to-last points; short curves dropped as in
// the reference (BCC fewer than 3 points, OBJ and .poly fewer than 4).
Curves read_bcc(const std::string& path, float target_seg_len, const Transform& tf, bool break_closed,
                bool allow_resample);
Curves read_obj(const std::string& path, float target_seg_len, const Transform& tf, bool break_closed,
                bool allow_resample);
Curves read_poly(const std::string& path, float target_seg_len, const Transform& tf, bool break_closed,
                 bool allow_resample);
Curves read_curves(const std::string& path, float target_seg_len, const Transform& tf, bool break_closed,
                   bool allow_resample);  // by extension

// Sim::exportToBCC / exportToOBJ of a flat simulation state: positions
// (float[3n]) and strand offsets (every strand one open curve).
void write_bcc(const std::string& path, const std::vector<float>& positions, const std::vector<int>& strand_offsets,
               bool as_polyline);
void write_obj(const std::string& path, const std::vector<float>& positions, const std::vector<int>& strand_offsets);

// Triangle mesh of an OBJ file ("v x y z" and "f a b c ..." with 1-based or
// negative indices, "a/b/c" forms allowed): polygons split into fans from
// their first corner, as assimp's aiProcess_Triangulate does for the convex
// faces of body meshes (PNMesh loads with it, src/PNMesh.cpp:14).
void read_obj_mesh(const std::string& path, std::vector<float>& vertices, std::vector<int>& triangles);

// Curves into the padded Hair layout (hairSize.y = the longest curve).
Hair hair_from_curves(const Curves& c);

}  // namespace kami_amd
