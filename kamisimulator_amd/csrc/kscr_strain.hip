// kscr_strain.hip -- strain limiting after the body SDF pass.
//
// StrainLimiting (SDF/cudaSDFCollision.cu:176-199), launched by
// SDFBodyCollisionSolver::SolveNextStep right after the SDF projection when
// UseStrainLimiting is set (SDFBodyCollisionSolver.cpp:182-187): every strand
// is walked from its root; a segment whose length l departs from its rest
// length l0 by more than StrainError * l0 is given length l0 again by moving
// its upper vertex along the segment, and the walk continues from the moved
// vertex.  Velocities are left as they are.
//
// The walk is sequential along a strand, so one thread owns one strand.  It
// runs once per step, after the step kernel; each thread streams its strand's
// positions and rest lengths once.
#include "kscr_device.h"
#include "kscr_math.h"
#include "kscr_solve.h"

namespace kscr {
namespace {

constexpr int kStrainThreads = 256;

__global__ __launch_bounds__(kStrainThreads) void strain_limit_kernel(float* __restrict__ x,
                                                                       const int* __restrict__ strand_off,
                                                                       const float* __restrict__ rest, int num_strands,
                                                                       float strain_error) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= num_strands) return;
    const int end = strand_off[s + 1];
    int i = strand_off[s];
    f3 cur = ld3(x, i);
    for (; i + 1 < end; i++) {
        f3 nxt = ld3(x, i + 1);
        const float l0 = rest[i];
        const float l = length(nxt - cur);
        if (fabsf(l - l0) > strain_error * l0) {
            const f3 dir = normalize(cur - nxt);
            nxt = nxt + (l - l0) * dir;
            st3(x, i + 1, nxt);
        }
        cur = nxt;
    }
}

}  // namespace

hipError_t launch_strain_limit(float* x, const int* strand_off, const float* rest, int num_strands,
                               float strain_error, hipStream_t stream) {
    if (num_strands <= 0) return hipSuccess;
    const dim3 grid((num_strands + kStrainThreads - 1) / kStrainThreads);
    hipLaunchKernelGGL(strain_limit_kernel, grid, dim3(kStrainThreads), 0, stream, x, strand_off, rest, num_strands,
                       strain_error);
    return hipGetLastError();
}

}  // namespace kscr
