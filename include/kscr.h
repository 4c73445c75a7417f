/*
 * kscr.h - C ABI of the MI355X-native Stable Cosserat Rods step.
 *
 * This is the inner seam below KamiSimulator's Solver plugin interface
 * (src/solvers/Solver.h:13-85).  It mirrors the lifecycle of the reference's
 * YarnBall::Sim (src/solvers/StableCosseratRods/YarnBall/YBSimulator.h:18-99)
 * and of the SDF body-collision pass
 * (src/solvers/SDFCollisionSolver/SDFBodyCollisionSolver.cpp:122-263), with
 * plain pointers and sizes only.  Every entry point returns a kscr_status;
 * no C++ exception crosses this boundary and nothing calls exit().
 *
 * Layout conventions (host arrays, caller owned, copied in and out):
 *   positions / velocities : float[num_verts*3], xyz interleaved
 *   orientations q, q_rest : float[num_verts*4], (x, y, z, w), w is the scalar
 *   strand_offsets         : int[num_strands+1], vertex ranges of each strand
 * Device buffers are owned by the handle.  Calls on one handle must be
 * serialised by the caller; distinct handles are independent.
 */
#ifndef KSCR_H_
#define KSCR_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kscr_handle kscr_handle;

/* Status codes.  KSCR_ERR_MAX_COLLISIONS mirrors the device error word of
 * src/solvers/StableCosseratRods/YarnBall/YarnBall.h:16; the validation codes
 * mirror the throws of YBSimulator.cu:82, :185-186, :197, :326-332. */
typedef enum kscr_status {
    KSCR_OK = 0,
    KSCR_ERR_INVALID_ARGUMENT = 1,
    KSCR_ERR_TOO_FEW_VERTICES = 2,     /* Sim::Sim, numVerts < 3            */
    KSCR_ERR_DANGLING_SEGMENT = 3,     /* Sim::configure, 1-vertex strand   */
    KSCR_ERR_ZERO_LENGTH_SEGMENT = 4,  /* Sim::configure, |x_{i+1}-x_i| = 0 */
    KSCR_ERR_DETECTION_SCALER = 5,     /* uploadMeta, detectionScaler <= 1  */
    KSCR_ERR_SEGMENT_TOO_SHORT = 6,    /* uploadMeta, 2 minSegLen < 2r + b  */
    KSCR_ERR_NOT_CONFIGURED = 7,       /* setYoung/setK* before configure   */
    KSCR_ERR_MAX_COLLISIONS = 8,       /* > 16 contacts on one segment      */
    KSCR_ERR_STRAND_TOO_LONG = 9,      /* strand longer than the coloured
                                        * schedule's workgroup (1023) or
                                        * than an explicit block_size on a
                                        * path that cannot cut strands     */
    KSCR_ERR_HIP = 10,                 /* HIP runtime failure               */
    KSCR_ERR_OUT_OF_MEMORY = 11,
    KSCR_ERR_UNSUPPORTED = 12
} kscr_status;

/* Iteration schedule (no reference counterpart for COLORED, an addition of
 * this build).  JACOBI: every solve of iteration k reads iteration-k values,
 * the reference's semantics.  COLORED: vertices and segments coloured by the
 * parity of their global index; the even colour is solved first and the odd
 * colour reads its new values (position solves, then orientation solves).
 * Per iteration it costs what JACOBI costs and converges about as far as two
 * JACOBI iterations.  It covers the contact-free step: with hair-hair contacts
 * or step size limits, with the adjoint, and in a batch it is
 * KSCR_ERR_UNSUPPORTED; strands of up to 1023 vertices. */
enum {
    KSCR_SCHEDULE_JACOBI = 0,
    KSCR_SCHEDULE_COLORED = 1
};

/* Vertex flags, same bit values as YarnBall::VertexFlags (YarnBall.h:33-39). */
enum {
    KSCR_FLAG_HAS_PREV = 1,
    KSCR_FLAG_HAS_NEXT = 2,
    KSCR_FLAG_HAS_NEXT_ORIENTATION = 4,
    KSCR_FLAG_FIX_ORIENTATION = 8
};

/* Simulation parameters: the MetaData fields the host sets
 * (YarnBall.h:61-124) plus Sim::maxH (YBSimulator.h:26) and the two
 * additions of this build (schedule, block_size).  kscr_params_default()
 * fills the values of Sim::Sim (YBSimulator.cu:81-112). */
typedef struct kscr_params {
    float gravity[3];
    int32_t num_itr;
    float max_h;
    float damping;
    float velocity_decay;
    float radius;
    float barrier_thickness;
    float acceleration_ratio;
    float k_collision;
    float k_friction;
    float friction_coeff;
    float detection_scaler;
    float bvh_rebuild_period;
    int32_t detection_period;
    int32_t use_step_size_limit;
    int32_t use_velocity_radius;
    int32_t enable_hair_hair_coll;
    float external_force_multiplier;
    float wind_peak;
    float wind_y_freq;
    float wind_z_freq;
    float wind_time_freq;
    float wind_sharpness;
    int32_t schedule;    /* KSCR_SCHEDULE_*                         */
    int32_t block_size;  /* vertices per workgroup; 0 = automatic:
                          * the smallest size holding the longest
                          * strand and at least 512 vertices (256 when
                          * contacts or step limits are on), halved
                          * while the groom has fewer than 1024
                          * workgroups, taken at upload;
                          * kscr_autotune measures instead */
} kscr_params;

/* Sim::Sim defaults (YBSimulator.cu:81-112). */
int kscr_params_default(kscr_params* p);

/* Sim::Sim(numVerts) + YarnBall::createFromCurves (YB/io/reader.cpp:16-50):
 * every vertex gets hasNext except the last one of each strand.  Strands
 * must be stored contiguously; strand_offsets[0] = 0 and
 * strand_offsets[num_strands] = num_verts.  device < 0 selects the last
 * visible device, as SelectCuda does (StableCosseratRodsSolver.cpp:64-74). */
int kscr_create(const kscr_params* p, int num_verts, const float* positions, int num_strands,
                const int* strand_offsets, int device, kscr_handle** out);
void kscr_destroy(kscr_handle* h);

/* Sim::configure (YBSimulator.cu:161-274): flags, rest lengths, lumped
 * masses, initial and rest orientations; allocates device memory. */
int kscr_configure(kscr_handle* h, float density);
/* Sim::setYoung (YBSimulator.cu:300-320), young in GPa. */
int kscr_set_young(kscr_handle* h, float young_gpa, float deviation);
/* Sim::setKStretch / Sim::setKBend (YBSimulator.cu:276-298). */
int kscr_set_kstretch(kscr_handle* h, float k);
int kscr_set_kbend(kscr_handle* h, float k);
/* Pins vertices: invMass = 0 (StableCosseratRodsSolver.cpp:203-205). */
int kscr_pin_vertices(kscr_handle* h, const int* ids, int n);
/* Sets VertexFlags::fixOrientation on vertices (a true clamp of a root
 * segment; the flag is honoured by cosserat.cu:303). */
int kscr_fix_orientation(kscr_handle* h, const int* ids, int n);
/* Vertex::connectionIndex (YarnBall.h:56) of every vertex: conn[i] is the
 * vertex that vertex i is glued to, or -1 (conn = NULL: none).  The
 * reference sets it for the two ends of a closed curve (createFromCurves,
 * YB/io/reader.cpp:40-44) and for curve endpoints within a search radius of
 * another vertex (glueEndpoints, YB/io/jsonBuilder.cpp:163-188); both
 * set both directions of a pair, and so must conn (else
 * KSCR_ERR_INVALID_ARGUMENT).  The position solve then adds a spring of stiffness
 * 40 towards the glued vertex (cosserat.cu:123-130), and the narrow phase
 * skips the segment pairs a glue joins (collision.cu:44-49).  A groom with
 * glued vertices steps through the per-iteration kernels of the contact path
 * (the partner may lie in another workgroup); the coloured schedule, the
 * batch and the adjoint reject it (KSCR_ERR_UNSUPPORTED).  Marks the handle
 * not uploaded. */
int kscr_set_connections(kscr_handle* h, const int* conn);
/* SagFree initialisation, as Sim::SagFree(rootVertices, iteration)
 * (YBSimulator.cu:480-676), called by the SCR solver when SagFreeItr > 0
 * (StableCosseratRodsSolver.cpp:207): per strand, rest lengths shortened so
 * that stretch carries the weight below each segment, then frames and rest
 * rotors turned so that bending carries the weight's torque.  Host code, like
 * the reference's; call it after kscr_configure, kscr_set_young and
 * kscr_pin_vertices and before kscr_upload (it marks the handle not
 * uploaded).  roots: the first vertex of every strand, ascending, each strand
 * ending where the next begins (NULL = the strands of kscr_create);
 * gravity is the one in the handle's parameters. */
int kscr_sag_free(kscr_handle* h, const int* roots, int num_roots, int iterations);
/* Replaces the simulation parameters; runs the uploadMeta validation
 * (YBSimulator.cu:322-335). */
int kscr_set_params(kscr_handle* h, const kscr_params* p);
int kscr_get_params(const kscr_handle* h, kscr_params* p);
/* Sim::upload (YBSimulator.cu:345-353): host mirror -> device. */
int kscr_upload(kscr_handle* h);

/* Sim::uploadPosAndVel / downloadPosAndVel (YBSimulator.cu:375-386). */
int kscr_set_pos_vel(kscr_handle* h, const float* positions, const float* velocities);
int kscr_get_pos_vel(kscr_handle* h, float* positions, float* velocities);
/* The same download without the copy out: positions and velocities (either
 * may be NULL) receive pointers to the handle's page-locked host mirror
 * (float[3n] each), valid until the next call on the handle.  No reference
 * counterpart: the exchange of Sim::downloadPosAndVel at link speed, for a
 * caller that scatters the state into its own layout (the host solver's
 * DownloadHair). */
int kscr_download_pos_vel(kscr_handle* h, const float** positions, const float** velocities);
/* Velocities of the previous substep (float[num_verts*3]), the state the
 * adaptive initial guess of initItr differentiates: Sim's d_lastVels
 * (YarnBall.h:68, iteration.cu:47-50).  Synchronous. */
int kscr_get_last_vel(kscr_handle* h, float* last_vel);
/* Orientation state q (float[num_verts*4]) in and out. */
int kscr_get_q(kscr_handle* h, float* q);
int kscr_set_q(kscr_handle* h, const float* q);
/* Rest state after configure/setYoung (host mirror; synchronous). */
int kscr_get_rest(const kscr_handle* h, float* inv_mass, float* l_rest, float* k_stretch, uint32_t* flags,
                  float* q_rest);

/* Body SDF on a regular grid (SDFBodyCollisionSolver.cpp:122-167 and
 * src/PNMesh.cpp:109-171): node (i,j,k) sits at min + (i,j,k)*ds with
 * ds = (max - min)/dim; phi[dim.x*dim.y*dim.z] and grad[3*...] are laid out
 * x fastest.  rotation (x,y,z,w) and translation place the body grid in the
 * world.  threshold is SDFThreshold (SDFBodyCollisionSolver.h:25). */
int kscr_set_sdf(kscr_handle* h, const int dim[3], const float grid_min[3], const float grid_max[3],
                 const float* phi, const float* grad, float threshold, const float rotation[4],
                 const float translation[3]);
int kscr_enable_sdf(kscr_handle* h, int enable);
/* Strain limiting after the body SDF pass: SDFBodyCollisionSolver's
 * EnableStrainLimiting (SDFBodyCollisionSolver.cpp:189), StrainError
 * (SDFBodyCollisionSolver.h:26, default 0.1) and the StrainLimiting kernel
 * (SDF/cudaSDFCollision.cu:176-199) it launches after SDFHeadCollision
 * (SDFBodyCollisionSolver.cpp:182-187).  After every step that applies the
 * SDF pass (kscr_solve_next_step), each strand is walked from its root and a
 * segment whose length l differs from its rest length l0 by more than
 * strain_error * l0 is given length l0 by moving its upper vertex along the
 * segment; velocities are unchanged.  rest_lengths float[num_verts] holds the
 * rest length of segment (i, i + 1) at i (the last vertex of a strand has
 * none); NULL takes the segment lengths of the handle's host positions now,
 * as InitAdditionalHairData does from the Hair (SDFBodyCollisionSolver.cpp:
 * 93-157).  enable = 0 switches it off.  The batched step and the adjoint
 * return KSCR_ERR_UNSUPPORTED with it on. */
int kscr_set_strain_limit(kscr_handle* h, int enable, float strain_error, const float* rest_lengths);
/* Body SDF grid from a closed triangle mesh, built on the device:
 * PNMesh::GenerateSDF + PNMesh::StoreSDFQueries (src/PNMesh.cpp:100-171)
 * through Kami::SDF::Init / CollisionTest / Distance
 * (src/utilities/Geometry.cpp:5-59).  The mesh's signed distance (Discregrid
 * TriangleMeshDistance: nearest point by Eberly's point-triangle distance,
 * sign from angle-weighted pseudo-normals) fills a cubic Lagrange grid of dim
 * cells over the vertex box padded by 0.1 (Discregrid
 * CubicLagrangeDiscreteGrid); its interpolant is sampled at the nodes
 * min + (i,j,k)*ds of the vertex box, ds = (max - min)/dim: phi, and grad =
 * the normalized interpolant gradient where phi < 0, else the direction from
 * the grid centre.  vertices float[3*num_vertices]; triangles
 * int[3*num_triangles], counter-clockwise seen from outside; a triangle of
 * zero area is KSCR_ERR_INVALID_ARGUMENT.  phi float[dim.x*dim.y*dim.z] and
 * grad float[3*...] (x fastest; either may be NULL) and grid_min/grid_max
 * (the vertex box; may be NULL) are written; they are kscr_set_sdf's inputs.
 * device < 0 uses the current device.  err (may be NULL) receives a message
 * on failure. */
int kscr_mesh_sdf(int device, const float* vertices, int num_vertices, const int* triangles, int num_triangles,
                  const int dim[3], float* phi, float* grad, float grid_min[3], float grid_max[3], char* err,
                  size_t err_len);
/* The same grid built straight into the handle's body SDF, followed by what
 * kscr_set_sdf does with it (no host round trip). */
int kscr_set_sdf_from_mesh(kscr_handle* h, const float* vertices, int num_vertices, const int* triangles,
                           int num_triangles, const int dim[3], float threshold, const float rotation[4],
                           const float translation[3]);

/* Sim::advance (YB/sim/step.cpp:38-62): splits dt [s] into
 * S = max(1, ceil(dt/max_h)) substeps of numItr VBD iterations. */
int kscr_advance(kscr_handle* h, float dt);
/* StableCosseratRodsSolver::SolveNextStep (StableCosseratRodsSolver.cpp:216-234)
 * without the host round trips: advance, then the SDF pass on device. */
int kscr_solve_next_step(kscr_handle* h, float dt);

/* Execution control.  stream is a hipStream_t (NULL = the null stream). */
int kscr_set_stream(kscr_handle* h, void* stream);
int kscr_synchronize(kscr_handle* h);
int kscr_num_verts(const kscr_handle* h);
int kscr_num_strands(const kscr_handle* h);
/* Vertices per workgroup actually used by the step kernel. */
int kscr_block_size(const kscr_handle* h);
/* Autotuner: times `steps` steps of dt for every instantiated workgroup size
 * that holds the longest strand (state is saved and restored), keeps the
 * fastest and returns it in best_block (may be NULL). */
int kscr_autotune(kscr_handle* h, float dt, int steps, int* best_block);

/* Per-vertex material: stretch stiffness (float[n]) and rest Darboux rotor
 * (float[4n], x y z w; its norm is the bend stiffness).  Writes the values
 * Sim::setYoung / setKStretch / setKBend compute into Vertex::kStretch and
 * Sim::qRests (YarnBall.h:55, YBSimulator.h:22), for materials that vary per
 * strand or per vertex.  Either pointer may be NULL (kept).  Takes effect on
 * the device at once when the handle is uploaded. */
int kscr_set_material(kscr_handle* h, const float* k_stretch, const float* q_rest);

/* ---- Differentiation (no counterpart in the reference, which has no
 * adjoint).  The reverse mode of the contact-free step, with
 * checkpoint-and-recompute.
 *
 * kscr_adjoint_forward advances nsteps steps of dt (kscr_advance semantics)
 * and keeps a device copy of the state every checkpoint_every steps.
 * kscr_adjoint_backward then takes the adjoints of the final positions and
 * velocities (float[3n] each, NULL = zero) back through the rollout: for each
 * segment of checkpoint_every steps, last first, it restores the segment's
 * checkpoint, re-runs the segment recording its tape (the last segment's tape
 * is kept from kscr_adjoint_forward), and sweeps its steps backwards.  It returns d(loss)/d(kStretch) (float[n]) and d(loss)/d(qRest)
 * (float[4n]) and leaves the state at the end of the rollout.  Wind is
 * differentiated through its dependence on the positions; returns
 * KSCR_ERR_UNSUPPORTED for contacts, step size limits, glued vertices or
 * strands longer than 1024 vertices.  Steps follow kscr_advance (no SDF
 * pass). */
int kscr_adjoint_forward(kscr_handle* h, float dt, int nsteps, int checkpoint_every);
/* The same with steps that follow kscr_solve_next_step: the SDF pass after
 * each step, differentiated as the projection it is within a grid cell (the
 * cell choice and the inside/outside test count as constant). */
int kscr_adjoint_forward_sdf(kscr_handle* h, float dt, int nsteps, int checkpoint_every);
int kscr_adjoint_backward(kscr_handle* h, const float* x_bar, const float* v_bar, float* k_stretch_bar,
                          float* q_rest_bar);
/* Device times (ms) of the last kscr_adjoint_backward, from HIP events:
 * re-recording the segments (summed over the segments, on the library's
 * recording stream), the reverse sweeps (summed, on the handle's stream) and
 * the whole backward pass from its start to the restored end state (the
 * re-recording of segment j - 1 overlaps the sweep of segment j, so this is
 * less than the sum).  Any pointer may be NULL. */
int kscr_adjoint_last_time(const kscr_handle* h, float* record_ms, float* sweep_ms, float* backward_ms);

/* ---- Batched form (no counterpart in the reference: its closest analogue is
 * CantileverTestSolver, which steps independent solvers one after another,
 * src/solvers/CantileverTestSolver.cpp:106-111).  A batch advances many
 * independent handles (grooms, each with its own gravity, SDF and parameters)
 * in one kernel launch per step, replayed from a captured hipGraph.  The
 * handles keep ownership of their state; kscr_get_pos_vel etc. stay valid. */
typedef struct kscr_batch kscr_batch;
/* block_size 0 = the same choice as kscr_create's (>= the longest strand). */
int kscr_batch_create(kscr_handle* const* handles, int num_handles, int block_size, kscr_batch** out);
void kscr_batch_destroy(kscr_batch* b);
/* nsteps steps of dt for every groom; with_sdf applies each groom's SDF pass
 * after each step (kscr_solve_next_step semantics) where one is set. */
int kscr_batch_step(kscr_batch* b, float dt, int nsteps, int with_sdf);
/* use_graph = 0 launches the kernels directly instead of replaying a graph. */
int kscr_batch_set_graph(kscr_batch* b, int use_graph);
int kscr_batch_synchronize(kscr_batch* b);
/* Mapping of the packs (whole strands of one groom, at most the workgroup
 * size in vertices) onto workgroups: KSCR_BATCH_GRID launches one workgroup
 * per pack; KSCR_BATCH_PERSISTENT launches as many workgroups as the device
 * holds at once, each taking the next pack from a device work counter until
 * none is left.  The results are the same. */
enum { KSCR_BATCH_GRID = 0, KSCR_BATCH_PERSISTENT = 1 };
int kscr_batch_set_mapping(kscr_batch* b, int mapping);
int kscr_batch_mapping(const kscr_batch* b);
/* Autotuner of the batched launch: times `steps` steps of dt for every
 * instantiated workgroup size that holds the longest strand of any groom,
 * under both mappings, keeps the fastest pair (every groom's state and clock
 * are saved and restored) and returns the workgroup size in best_block (may
 * be NULL; kscr_batch_mapping gives the mapping). */
int kscr_batch_autotune(kscr_batch* b, float dt, int steps, int* best_block);
int kscr_batch_block_size(const kscr_batch* b);
int kscr_batch_last_error(const kscr_batch* b, char* buf, size_t n);

/* Contact lists of the last collision detection (hair-hair contacts only):
 * num[num_verts] contacts per segment and cols[16*num_verts] partner segment
 * ids, ascending, -1 in unused slots (YB/YarnBall.h:74-76, :19). */
int kscr_get_contacts(kscr_handle* h, int* num, int* cols);

/* ---- Linear BVH (Kitten::LBVH, KE/KittenGpuLBVH/lbvh.cuh:14-83).  Boxes are
 * float[6*n] (min xyz, max xyz); an empty box (+inf, -inf) is ignored. */
typedef struct kscr_lbvh kscr_lbvh;
int kscr_lbvh_create(int device, kscr_lbvh** out);
void kscr_lbvh_destroy(kscr_lbvh* b);
/* LBVH::compute (lbvh.cu:357-394): root box, Morton codes, stable sort,
 * Karras build, refit. */
int kscr_lbvh_compute(kscr_lbvh* b, const float* aabbs, int n);
/* LBVH::refit (lbvh.cu:348-356) with new boxes for the same objects. */
int kscr_lbvh_refit(kscr_lbvh* b, const float* aabbs);
/* LBVH::query (lbvh.cu:411-426): unique overlapping pairs as int[2*k]
 * (found object, query object); *count receives the number found, which may
 * exceed cap (only cap pairs are written). */
int kscr_lbvh_query(kscr_lbvh* b, int* pairs, size_t cap, size_t* count);
/* The same three operations on boxes and pairs that live in device memory,
 * as the reference's own interface takes them (LBVH::compute(Bound*, size),
 * LBVH::refit(), LBVH::query(ivec2*, size): lbvh.cu:348-426): d_aabbs is
 * float[6*n] and d_pairs int[2*cap] on the handle's device, and the work runs
 * on `stream` (a hipStream_t; NULL = the null stream).  d_aabbs is read in
 * place, as the reference keeps the pointer compute is given: it must hold
 * the boxes until the work queued by the next compute_device, refit_device or
 * query_device has run.  compute_device and refit_device queue their work and
 * return; query_device waits for it and returns the number of pairs found
 * (which may exceed cap), as LBVH::query does (lbvh.cu:425). */
int kscr_lbvh_compute_device(kscr_lbvh* b, const float* d_aabbs, int n, void* stream);
int kscr_lbvh_refit_device(kscr_lbvh* b, const float* d_aabbs, void* stream);
int kscr_lbvh_query_device(kscr_lbvh* b, int* d_pairs, size_t cap, size_t* count, void* stream);
/* No counterpart in the reference: the LBVH self query without device
 * allocation, for callers that cannot allocate (code under stream capture,
 * serving loops, harnesses that forbid hipMalloc).  kscr_lbvh_compute_device
 * then kscr_lbvh_query_device over d_aabbs (float[6*n]) into d_pairs
 * (int[2*cap]), on the current device and `stream`, with the tree in a
 * fixed device array of the library (one per device, 24 MiB): at most
 * 131072 boxes, more is KSCR_ERR_INVALID_ARGUMENT.  *count receives the
 * number of pairs found (which may exceed cap).  Calls that overlap on one
 * device share that array and race.  err (may be NULL) receives a message on
 * failure. */
int kscr_lbvh_self_query_static(const float* d_aabbs, int n, int* d_pairs, size_t cap, size_t* count, void* stream,
                                char* err, size_t err_len);
/* No counterpart in the reference (a tuning and test hook): how many of the
 * nodes whose leaf ranges cross a 512-leaf tile boundary the fit of a tree
 * over at most 131072 boxes keeps in LDS, 0 .. 4096 (the default).  Trees
 * with more such nodes (about 2k at 100k boxes) fit them from global memory.
 * Applies from the next compute or refit; the tree is the same either way. */
int kscr_lbvh_set_fit_slots(kscr_lbvh* b, int slots);
/* Tree contents: sorted Morton codes and object ids (n), internal nodes as
 * (parent, left, right, fence) (4 per node) and child boxes (12 per node),
 * leaf parents (n), root box (6), DFS stack bound. */
int kscr_lbvh_get(kscr_lbvh* b, uint32_t* codes, uint32_t* ids, uint32_t* links, float* bounds,
                  uint32_t* leaf_parent, float* root, int* max_stack);

/* Copies the last error message (NUL terminated) into buf. */
int kscr_last_error(const kscr_handle* h, char* buf, size_t n);

/* Self-check of the fast correctly rounded arithmetic the kernels use in
 * place of IEEE division and square root, run on `device` (< 0: current)
 * against the compiler's IEEE sequences; no reference counterpart.  counts
 * receives the number of differing results: [0] sqrt over all 2^32 float
 * bit patterns (expected 0); [1] 1/sqrt over all 2^32 (expected 3: +0, -0,
 * +inf, where a vector scaled by either is NaN); [2] the reverse sweep's
 * clamped square root and reciprocal over all non-negative patterns
 * (expected 0); [3] the division by a rest length over all 2^32 numerators
 * for each of `divisors` (expected 0).  Takes about a second per divisor. */
int kscr_selftest_rounding(int device, const float* divisors, int num_divisors, unsigned long long counts[4]);
/* Library-wide message of the last kscr_create failure. */
const char* kscr_create_error(void);

#ifdef __cplusplus
}
#endif

#endif /* KSCR_H_ */
