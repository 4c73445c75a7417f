// kscr_device.h -- device-side data layout of one simulation handle and the
// launchers of the step kernels.
//
// HBM layout (structure of arrays, one entry per vertex, strands contiguous):
//   x, v, u     float[3N]  position, velocity, velocity at the previous
//                          substep start ("lastVels")          read/write
//   q           float4[N]  segment orientation (x, y, z, w)    read/write
//   q_rest      float4[N]  rest Darboux rotor; its norm is the bend stiffness
//   mat         float4[N]  (invMass, lRest, kStretch, flags as bits)
//   packs       int2[P]    vertex range [begin, end) of each workgroup; a
//                          range always holds whole strands
//   sdf grid    float4[dim.x*dim.y*dim.z]  (dphi/dx, dphi/dy, dphi/dz, phi)
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

namespace kscr {

struct SdfArgs {
    const float4* grid;  // (grad, phi) per node, x fastest
    int dim[3];
    float gmin[3];
    float ds[3];
    float inv_ds[3];
    float thresh;
    float rot[4];
    float trans[3];
};

struct StepArgs {
    float* x;
    float* v;
    float* u;
    float4* q;
    const float4* q_rest;
    const float4* mat;
    int num_substeps;
    int num_itr;
    float h;
    float last_h;
    float time;
    float gravity[3];
    float damping;
    float velocity_decay;
    float acceleration_ratio;
    float external_force_multiplier;
    float wind_peak, wind_y_freq, wind_z_freq, wind_time_freq, wind_sharpness;
    int sdf_on;
    SdfArgs sdf;
    // Vertex::connectionIndex per vertex (the glued vertex or -1), or null
    // when no vertex is glued; read by the per-iteration kernels only
    const int* conn;
};

// Tape of the differentiable step: for every substep of one recorded step,
// x at the substep start, y, whether the adaptive initial guess's factor
// lies strictly inside (0, 1) (its clamp is then inactive), every iterate
// dx^k (k = 0..numItr) and the rotors q^k the iterations start from
// (k = 0..numItr-1).  Arrays are indexed [(sub * (num_itr + 1) + k) * n + i]
// (dx), [(sub * num_itr + k) * n + i] (q) and [sub * n + i] (per-substep
// values), float3 packed as 3 floats.
struct TapeArgs {
    float* x;
    float* y;
    unsigned char* sigma_free;
    float* dx;
    float4* q;   // q^0 .. q^K per substep
    float* ilw;  // per iteration: the orientation solve's 1 / |w| (w normalized to q^{k+1})
    float* acc;  // per substep, with wind (else null): the adaptive guess's acceleration (v - u) / lastH
    int n;
};

// Arguments of the reverse sweep over one recorded step.  xb..qrb hold the
// adjoints of the step's outputs on entry and of its inputs on exit (ksb and
// qrb accumulate over all steps).
struct AdjointArgs {
    const float4* mat;
    const float4* q_rest;
    TapeArgs tape;
    int num_substeps;
    int num_itr;
    float h;
    float last_h;
    float damping;
    float velocity_decay;
    float acceleration_ratio;
    float ext[3];  // gravity * external_force_multiplier
    float ext2;    // |ext|^2
    // wind (wind_peak != 0: the tape holds acc) at the step's time
    float gravity[3];
    float external_force_multiplier;
    float wind_peak, wind_y_freq, wind_z_freq, wind_time_freq, wind_sharpness, time;
    // the SDF pass at the end of the step (sdf_on)
    int sdf_on;
    SdfArgs sdf;
    float* xb;
    float* vb;
    float* ub;
    float4* qb;
    float* ksb;
    float4* qrb;
};

// One launch advances every strand of the packs by num_substeps substeps (and
// applies the SDF pass afterwards when sdf_on).  block is the workgroup size
// in vertices; packs[p] = [begin, end) of workgroup p.
hipError_t launch_step_fused(const StepArgs& a, const int2* packs, int num_packs, int block, hipStream_t stream);
// Even/odd coloured step (kscr_colored.hip): block threads, two vertices per
// lane, packs of at most 2 block - 1 vertices; block <= 512.
hipError_t launch_step_colored(const StepArgs& a, const int2* packs, int num_packs, int block, hipStream_t stream);
// SagFree initialisation on the host mirror of a groom (kscr_sagfree.cpp);
// q and q_rest are (x, y, z, w) rotors, 4 floats each.
void sag_free_groom(const float* pos, const float* inv_mass, float* l_rest, const float* k_stretch, float* q,
                    float* q_rest, const float gravity[3], const int* roots, int num_roots, int num_verts, int iterations);
// Body SDF grid of a triangle mesh (kscr_meshsdf.hip): the (grad, phi) float4
// grid of dim[0]*dim[1]*dim[2] nodes into the device array out, on stream;
// gmin / gmax receive the vertex box the grid spans.  Returns a kscr_status,
// with a message in err on failure; synchronises the stream.
int mesh_sdf_build(const float* verts, int nv, const int* tris, int nt, const int dim[3], float4* out, float gmin[3],
                   float gmax[3], hipStream_t stream, std::string& err);
// Same step, writing the tape of the step (contact-free step, no SDF pass).
hipError_t launch_step_record(const StepArgs& a, const TapeArgs& tape, const int2* packs, int num_packs, int block,
                              hipStream_t stream);
// Reverse sweep over one recorded step.
hipError_t launch_step_adjoint(const AdjointArgs& a, const int2* packs, int num_packs, int block, hipStream_t stream);
// Batched form: packs[p] = (begin, end, groom, 0) index the device table of
// per-groom arguments.
hipError_t launch_step_batch(const StepArgs* grooms, const int4* packs, int num_packs, int block, hipStream_t stream);
// Strain limiting after the SDF pass (kscr_strain.hip): strands
// [strand_off[s], strand_off[s + 1]), rest[i] the rest length of segment
// (i, i + 1).
hipError_t launch_strain_limit(float* x, const int* strand_off, const float* rest, int num_strands,
                               float strain_error, hipStream_t stream);
// Persistent form: `grid` workgroups take the packs from the device counter
// *next, which must be 0 at the launch; batch_persistent_grid(block) is the
// number of workgroups the current device holds at once (0 if unknown).
hipError_t launch_step_batch_persistent(const StepArgs* grooms, const int4* packs, int num_packs, int block, int grid,
                                        int* next, hipStream_t stream);
int batch_persistent_grid(int block);
// last_h <- h, time <- time + dt for n grooms of a device table.
hipError_t launch_batch_clock(StepArgs* grooms, int n, float dt, hipStream_t stream);

// Block sizes the fused kernel is instantiated for.
constexpr int kBlockSizes[] = {64, 128, 256, 512, 1024};

}  // namespace kscr
