/*
 * pdc.h — C ABI of the MI355X per-timestep rigid-body contact pipeline.
 *
 * The pipeline is: grid-SDF contact generation, per-contact primal-dual
 * friction terms, assembly of the 6-DOF block-sparse Newton system, a
 * block-Jacobi preconditioned CG on 6x6 block-CSR, and the merit line search.
 * Every entry point is plain C: integer status codes, plain pointers and
 * sizes, no exceptions cross the boundary.  All host arrays are caller-owned;
 * the library owns device buffers per context.  A context is not thread-safe:
 * use one context per host thread.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   - Solver plugin  ForceSolver::fillForces   src/Solver/ForceSolverBase.h:78-82
 *                    PrimalDualForceSolver     src/Solver/PrimalDual/PrimalDualForceSolver.cpp:79-130
 *   - configuration  fillFromYAML              src/Solver/PrimalDual/PrimalDualForceSolver.cpp:650-675,
 *                                              src/Solver/PrimalDual/NonSmoothForces/NonSmoothContactForce.cpp:320-334
 *   - contacts       DynamicSystem::contactDetection  src/Dynamics/DynamicSystem.cpp:106-122
 *                    ContactGeneratorBase::operator() src/Contact/ContactGenerator/ContactGeneratorBase.h:85-95
 *   - timestep       DynamicSystem::step       src/Dynamics/DynamicSystem.cpp:130-218
 *   - timings        printTimings              src/Solver/PrimalDual/PrimalDualForceSolver.cpp:636-648
 */
#ifndef PDC_H
#define PDC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------- */
enum {
	PDC_OK = 0,
	PDC_ERR_INVALID_ARGUMENT = 1, /* null pointer, negative size, bad enum */
	PDC_ERR_CONFIG = 2,           /* configuration rejected (the reference throws std::runtime_error) */
	PDC_ERR_DEVICE = 3,           /* a HIP call failed */
	PDC_ERR_NO_DEVICE = 4,        /* no MI355X (gfx950) device visible */
	PDC_ERR_CAPACITY = 5,         /* request exceeds a fixed capacity or the HBM budget */
	PDC_ERR_STATE = 6             /* call order violated (e.g. step before bodies) */
};

/* ---- enums --------------------------------------------------------------- */
/* Shapes: the primitives (src/CollisionObject/Cube.cpp, Sphere.cpp,
 * Ellipsoid.cpp) and triangle meshes (Mesh.cpp, Convex.cpp; see
 * pdc_register_mesh). */
enum { PDC_SHAPE_BOX = 0, PDC_SHAPE_SPHERE = 1, PDC_SHAPE_ELLIPSOID = 2, PDC_SHAPE_MESH = 3 };

/* non_smooth_force.name: "non_smooth_contact" | "only_normal"
 * (PrimalDualForceSolver.cpp:664-673) */
enum { PDC_NON_SMOOTH_CONTACT = 0, PDC_ONLY_NORMAL = 1 };

/* Storage of the inner PCG's operator: FP32 6x6 blocks (PDC_PCG_FP32) or FP64
 * blocks.  The PCG vectors, dot products and scalars are FP64 on both, and the
 * Newton outer loop, residuals, barrier terms, centering parameter and
 * line-search merit are FP64 always. */
enum { PDC_PCG_FP64 = 0, PDC_PCG_FP32 = 1 };

/* ---- parameters ---------------------------------------------------------- */
/* Defaults (pdc_default_params) follow the reference's code defaults:
 *   simulation.timestep 1e-2, gravity (0,-9.81,0)      DynamicSystem.cpp:386, DynamicSystem.h:241
 *   solver.tolerance 1e-6, solver.iterations 100        PrimalDualForceSolver.cpp:651-652
 *   static_coeff 0.5, restitution 0.0 (0.5 only_normal) NonSmoothContactForce.cpp:329-333, OnlyNormalForce.cpp:227
 *   spring_k 1e6, spring_damp 0.01 when spring based   NonSmoothContactForce.cpp:324-327
 */
typedef struct pdc_params {
	double timestep;
	double gravity[3];
	int32_t force_model;        /* PDC_NON_SMOOTH_CONTACT | PDC_ONLY_NORMAL */
	int32_t spring_based_force; /* 0 | 1 */
	double friction_coeff;      /* global mu (static_coeff) */
	double restitution;
	double spring_k;
	double spring_damp;
	double tolerance;           /* Newton exit tolerance on exitErr */
	int32_t iterations;         /* Newton iteration cap */
	int32_t pcg_precision;      /* PDC_PCG_FP64 | PDC_PCG_FP32 */
	double pcg_tol;             /* relative residual |r|/|b| of the scaled system; at or above
	                               1e-9 the single-reduction PCG recurrences are used, below it
	                               the standard PCG iteration (both one launch per iteration) */
	int32_t pcg_max_iter;       /* the reference's maxCGIt = 1000 (PrimalDualForceSolver.cpp:381) */
	int32_t grid_n;             /* SDF grid nodes per axis (64, 96, ...) */
	double grid_pad_frac;       /* grid box = shape half-extent * (1 + grid_pad_frac) */
	double sample_spacing;      /* target spacing of surface samples */
} pdc_params;

/* Per-call solver statistics (the reference logs these through Corrade::Debug
 * and accumulates them into its .info timings). */
typedef struct pdc_solve_stats {
	int32_t n_contacts;
	int32_t n_local_bodies;     /* bodies in contact (the local solve set) */
	int32_t newton_iterations;
	int32_t pcg_iterations;     /* summed over Newton iterations */
	int32_t pcg_maxiter_hits;   /* solves that reached pcg_max_iter (Hu *= 1.3 rule at 2000 or
	                               more bodies in contact, where the reference solves iteratively) */
	int32_t gd_fallbacks;       /* line-search failures that fell back to gradient descent */
	int32_t line_search_trials; /* trial points evaluated */
	int32_t ties;               /* contact samples within the tie band of the threshold */
	int32_t n_blocks;           /* 6x6 blocks of the Newton matrix */
	int32_t nonfinite;          /* Newton iterations whose error or inner solve was not finite:
	                               the exit test never passes on a NaN error (as with the
	                               reference's std::max), and a solve that breaks down gives
	                               du = 0 (as the reference's failed factorisation,
	                               PrimalDualForceSolver.cpp:390-394) */
	double exit_error;
	double char_mass;
	double char_speed;
} pdc_solve_stats;

/* Timings in seconds accumulated since the last pdc_reset_timings, measured
 * with HIP events on the context's stream. */
typedef struct pdc_timings {
	double contact_detection;   /* (a) */
	double force_process;       /* whole solve, like m_forceProcessTimer */
	double build_system;        /* (b)+(c), like m_systemBuildTimer */
	double solve_system;        /* (d),      like m_systemSolveTimer */
	double line_search;         /* (e),      like m_lineSearchTimer */
	int64_t steps;
	int64_t newton_iterations;
	int64_t contacts;
} pdc_timings;

/* ---- library ------------------------------------------------------------- */
typedef struct pdc_context pdc_context;

/* Fills *p with the reference defaults listed above. */
void pdc_default_params(pdc_params* p);

/* Number of usable gfx950 devices (0 when none); never fails. */
int pdc_device_count(void);

/* Creates a context on HIP device `device`.  Returns PDC_ERR_NO_DEVICE when
 * no gfx950 device is visible — there is no CPU fallback. */
int pdc_create(int device, const pdc_params* params, pdc_context** out);
void pdc_destroy(pdc_context* ctx);

/* Replaces the configuration.  Rejects what the reference rejects
 * (PrimalDualForceSolver.cpp:655-661) plus out-of-range numbers. */
int pdc_configure(pdc_context* ctx, const pdc_params* params);

/* Message of the last failing call on this context ("" if none). */
const char* pdc_last_error(const pdc_context* ctx);

/* ---- bodies (RigidObjectBase inputs: RigidObjectBase.cpp:45-119) ---------- */
/* Uploads the body table: shape type and dimensions (box half extents, or
 * sphere radius in dims[3i], or ellipsoid semi-axes, or a mesh id), mass,
 * principal moments of inertia in the body frame, static flags, and whether
 * the body's surface samples are tested against other bodies (NULL: all
 * are).  Builds the FP32 SDF grid of every body and the surface samples of
 * every body that emits them on the device. */
int pdc_upload_bodies(pdc_context* ctx, int32_t n_bodies,
                      const int32_t* shape_type, const double* dims,
                      const double* mass, const double* inertia_diag,
                      const int32_t* is_static, const int32_t* emit_samples);

/* Triangle meshes (Mesh::initializeFCLObject src/CollisionObject/Mesh.cpp:18-110,
 * Mesh::signedDistance :125-167, MeshMeshContactGenerator::operator()
 * src/Contact/ContactGenerator/MeshMeshContactGenerator.cpp:7-119).
 * Registers a triangle mesh with the context: n_vertices vertices (3 doubles
 * each) in the body frame (centre of mass at the origin, axes along the
 * principal axes of inertia) and n_triangles triangles (3 vertex indices
 * each, counter-clockwise seen from outside).  *mesh_id receives its id.  A
 * body of shape PDC_SHAPE_MESH names its mesh by dims[3i] = id (dims[3i+1]
 * and dims[3i+2] are not read).  Its SDF grid holds the mesh's signed
 * distance, negative behind the angle-weighted pseudonormal of the closest
 * feature (libigl's signed_distance_pseudonormal in the reference), over the
 * box of its vertices padded by grid_pad_frac; its surface samples are the
 * mesh's vertices (the points the reference tests against the other mesh).
 * Meshes stay registered for the context's life.  PDC_ERR_INVALID_ARGUMENT
 * for an index out of range, a non-finite coordinate, or a mesh whose
 * triangles all have zero area. */
int pdc_register_mesh(pdc_context* ctx, int32_t n_vertices, const double* vertices,
                      int32_t n_triangles, const int32_t* triangles, int32_t* mesh_id);

/* Body state: pos (3N), quat (4N, x y z w as in the reference's NPZ output),
 * linear and angular velocity (3N each).  Kept device-resident. */
int pdc_set_state(pdc_context* ctx, const double* pos, const double* quat_xyzw,
                  const double* vel, const double* angvel);
int pdc_get_state(pdc_context* ctx, double* pos, double* quat_xyzw,
                  double* vel, double* angvel);
/* The same for bodies [first, first + n) only (arrays of n entries). */
int pdc_set_state_range(pdc_context* ctx, int32_t first, int32_t n, const double* pos,
                        const double* quat_xyzw, const double* vel, const double* angvel);
int pdc_get_state_range(pdc_context* ctx, int32_t first, int32_t n, double* pos,
                        double* quat_xyzw, double* vel, double* angvel);

/* Resident scenes that come and go (no counterpart in the reference, which
 * loads one scene per process, FileLoader.cpp:14-259).
 * pdc_append_bodies adds n bodies at the end of the uploaded table as a new
 * last scene (grids and samples built for them only, every other body and
 * scene kept, the other scenes' friction kept); *first_body receives the
 * index of the first.  pdc_replace_bodies overwrites bodies [first, first +
 * n) in place (the bodies keep their scene): their surface samples must fit
 * the sample range the replaced bodies held, and when they do not fill it
 * the last body must emit no samples (it keeps the rest, which are never
 * tested).  Both take the new bodies' state (a NULL array gives the identity
 * pose or zero velocities).  A static body far from everything that emits no
 * samples is inert: it changes no result of its scene. */
int pdc_append_bodies(pdc_context* ctx, int32_t n, const int32_t* shape_type,
                      const double* dims, const double* mass, const double* inertia_diag,
                      const int32_t* is_static, const int32_t* emit_samples,
                      const double* pos, const double* quat_xyzw, const double* vel,
                      const double* angvel, int32_t* first_body);
int pdc_replace_bodies(pdc_context* ctx, int32_t first, int32_t n, const int32_t* shape_type,
                       const double* dims, const double* mass, const double* inertia_diag,
                       const int32_t* is_static, const int32_t* emit_samples,
                       const double* pos, const double* quat_xyzw, const double* vel,
                       const double* angvel);
/* Grows the capacity of the body table to `bodies` bodies and `samples`
 * surface samples ahead of appends (so that appends do not copy the table). */
int pdc_reserve_bodies(pdc_context* ctx, int64_t bodies, int64_t samples);
/* Device memory held by the context's buffers, in bytes (the body table,
 * grids, samples, contacts and solver state as they stand). */
int pdc_device_bytes(const pdc_context* ctx, int64_t* bytes);
/* Diagnostics, no counterpart in the reference: the body groups of the last
 * solve's preconditioner (built by solves with pcg_tol below 1e-9, whose
 * standard PCG form is preconditioned by the inverse blocks of groups of up
 * to 8 strongly coupled bodies).  group_key (one entry per uploaded body)
 * receives, for every body in the last solve's contacts, the global id of
 * its group's key body (its smallest id), and -1 for the others.
 * PDC_ERR_STATE when no such solve has run since the last upload (or the
 * last call was pdc_fill_forces_noalloc, which keeps nothing). */
int pdc_get_groups(pdc_context* ctx, int32_t n_bodies, int32_t* group_key);

/* ---- (a) contact generation ----------------------------------------------- */
/* Runs broadphase, grid-SDF sample tests, compaction, the reference's
 * per-pair position merge (DynamicSystem.cpp:341-381) and frame construction
 * (ContactGeneratorBase.h:21-49) on the device-resident state.  The contact
 * list stays on the device; *n_contacts receives its length and *n_ties the
 * number of samples whose signed distance lies within the tie band. */
int pdc_detect_contacts(pdc_context* ctx, int32_t* n_contacts, int32_t* n_ties);

/* Diagnostics of the last contact detection (no counterpart in the
 * reference): candidate pairs, raw sample hits, tie-band samples, the
 * samples of the emitting sides summed over pairs, those of them in the
 * sample cells kept for the partner (the ones tested), the grid nodes under
 * each partner's box summed over pairs and sides, pairs with more than 2048
 * hits (merged in global memory), and the tested samples that fall inside
 * the partner's grid box (each reads 8 grid nodes). */
typedef struct pdc_detection_stats {
	int64_t pairs, raw_hits, ties;
	int64_t samples_all, samples_read, brick_nodes, large_pairs;
	int64_t samples_in_grid;
} pdc_detection_stats;
int pdc_get_detection_stats(const pdc_context* ctx, pdc_detection_stats* out);

/* Diagnostics, no counterpart in the reference: the FP32 SDF grid of body
 * `body` (grid_n^3 nodes, x fastest, over the body-frame box
 * [-h, h] per axis, h = the shape's half extent times 1 + grid_pad_frac).
 * PDC_ERR_CAPACITY when capacity < grid_n^3. */
int pdc_get_grid_nodes(pdc_context* ctx, int32_t body, int64_t capacity, float* nodes);

/* Copies the current contact list to the host in the reference's
 * ContactInformation layout (ContactInfo.h:138-144): obj_ids (2C, ordered
 * objId[0] < objId[1]), pos (3C), depth (C), transformation (36C: per contact
 * transformation[0] then transformation[1], each 6x3 column-major), and
 * local_force (3C) from the last solve.  Any pointer may be NULL. */
int pdc_get_contacts(pdc_context* ctx, int32_t capacity, int32_t* obj_ids,
                     double* pos, double* depth, double* transformation,
                     double* local_force);

/* ---- (b)-(e) solver: ForceSolver::fillForces drop-in ---------------------- */
/* Computes contact impulses for the given contacts and state, exactly the
 * contract of PrimalDualForceSolver::fillForces (PrimalDualForceSolver.cpp:79-130):
 * out_acc (6N) receives every body's (acceleration, angularAcc) velocity
 * increment — g*dt and 0 for bodies without contacts — and out_local_force
 * (3C) receives each contact's force in its local frame times charSpeed *
 * charMass.  Contact ids use the global numbering of the uploaded bodies. */
int pdc_fill_forces(pdc_context* ctx, int32_t n_contacts,
                    const int32_t* obj_ids, const double* transformation,
                    const double* depth, const double* quat_xyzw,
                    const double* vel, const double* angvel, double* out_acc,
                    double* out_local_force, pdc_solve_stats* stats);

/* The same computation as pdc_fill_forces, allocating no device memory.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * (code that must not call hipMalloc or hipFree, such as tight serving loops
 * or harnesses that forbid it).  Every device buffer the call needs beyond the
 * uploaded body table comes from a fixed scratch array of 64 MiB per device
 * inside the library, held until the process ends.  Limits: one scene in the
 * context, at most 32768 contacts and 4096 uploaded bodies; larger requests,
 * or a call whose buffers do not fit the scratch, return PDC_ERR_CAPACITY.
 * Calls on one device share that scratch: calls that overlap in time on one
 * GPU (from several threads or contexts) race.  The call's contact list is
 * not kept in the context (pdc_get_contacts does not return it), and kernel
 * timing is suspended for the call. */
int pdc_fill_forces_noalloc(pdc_context* ctx, int32_t n_contacts,
                            const int32_t* obj_ids, const double* transformation,
                            const double* depth, const double* quat_xyzw,
                            const double* vel, const double* angvel, double* out_acc,
                            double* out_local_force, pdc_solve_stats* stats);

/* ---- whole timestep on the device-resident state ---------------------------
 * DynamicSystem::step without adaptive substeps (DynamicSystem.cpp:130-161,
 * 207-210): contact detection, fillForces, then semi-implicit integration.
 * `stats` (may be NULL) receives the statistics of the last step (summed over
 * scenes when the bodies are partitioned into several). */
int pdc_step(pdc_context* ctx, int32_t n_steps, pdc_solve_stats* stats);

/* ---- independent scenes: batched and served stepping ----------------------
 * No counterpart in the reference, which simulates one scene per process
 * (DynamicSystem.cpp:130-218).  The uploaded body table is partitioned into
 * n_scenes independent scenes of consecutive bodies: scene k holds bodies
 * [body_begin[k], body_begin[k+1]), body_begin[0] = 0, body_begin[n] = N.
 * Bodies of different scenes never touch.  Every scene runs the same
 * per-timestep pipeline as a single scene, with its own Newton iteration,
 * PCG and line search; all scenes advance in the same kernel launches, and a
 * scene's result does not depend on which other scenes are stepped with it.
 * pdc_upload_bodies resets the partition to one scene. */
int pdc_set_scenes(pdc_context* ctx, int32_t n_scenes, const int32_t* body_begin);
int pdc_scene_count(const pdc_context* ctx);
/* Per-scene friction coefficient (n_scenes entries); pdc_configure with a new
 * friction_coeff resets every scene to it. */
int pdc_set_scene_friction(pdc_context* ctx, const double* mu);
/* n_steps timesteps of the selected scenes (all scenes when n_selected is 0);
 * the other scenes keep their state.  stats (may be NULL) receives each
 * selected scene's statistics of the last step (n_scenes entries when
 * n_selected is 0).  pdc_step is pdc_step_scenes over all scenes, its stats
 * summed over scenes (Newton iterations and exit error: the maximum). */
int pdc_step_scenes(pdc_context* ctx, int32_t n_steps, int32_t n_selected,
                    const int32_t* scene_ids, pdc_solve_stats* stats);

/* Work shapes, per scene size class (no counterpart in the reference, whose
 * CPU loops have none).  A scene's class is its count of bodies that are
 * dynamic or emit surface samples: < 64, < 256, < 1024, >= 1024.  contact_tile: contacts per workgroup of the per-contact
 * kernels (64, 128 or 256); body_waves: bodies (and body pairs) per workgroup
 * of the one-wavefront-per-body kernels (1, 2 or 4); pcg_chunk: PCG
 * launches the host queues at a time (8, 16 or 32, for every class; 0:
 * automatic, 8 for a single scene and 16 for several).  A scene's results depend on its own
 * class's shapes only, so they stay the same alone and in any mix that uses
 * the same table.  Defaults: 256, 4, 0. */
typedef struct pdc_shapes {
	int32_t contact_tile[4];
	int32_t body_waves[4];
	int32_t pcg_chunk;
} pdc_shapes;
int pdc_get_shapes(const pdc_context* ctx, pdc_shapes* out);
int pdc_set_shapes(pdc_context* ctx, const pdc_shapes* shapes);
/* Autotuner: for every size class present, times each candidate shape on
 * the class's scenes alone (steps timesteps from the current state, which
 * is restored afterwards) and keeps the fastest; then the PCG chunk length
 * over all scenes.  The tuned table is kept only if all scenes together run
 * faster with it than with the table tuning started from.  Returns the
 * chosen table in *chosen (may be NULL). */
int pdc_autotune(pdc_context* ctx, int32_t steps, pdc_shapes* chosen);

int pdc_get_timings(const pdc_context* ctx, pdc_timings* out);
void pdc_reset_timings(pdc_context* ctx);

/* Kernel-level timing used by the benchmark: when enabled, every launch of
 * the kernel class `which` is bracketed by HIP events on the context's stream.
 * pdc_kernel_timing returns the summed milliseconds, the launch count and the
 * summed algorithmic bytes (compulsory traffic: each input read once, each
 * output written once) of those launches. */
/* Classes: SPMV = the PCG iteration kernel (k_pcg_iter or k_pcgs_iter:
 * ELL block SpMV, preconditioner apply, vector updates, partial dot
 * products), timed per launch with the iterations launched directly (no
 * graph replay) while enabled, active iterations only; PCG = the same kernel
 * timed by one event pair around each solve's launches, every launch
 * counted; CONTACT_TERMS = constraint + contact-residual kernels of a
 * Newton iteration; ASSEMBLY = the assembly kernels and the group inverses;
 * LINE_SEARCH = the evaluation kernel of every line-search pass (its trials'
 * contact and body parts, the gradient-descent pass included); CONTACTS =
 * the cell cull and the sample-test kernel.  Algorithmic bytes are counted
 * by the library for SPMV, PCG, CONTACT_TERMS and CONTACTS. */
enum { PDC_KT_NONE = 0, PDC_KT_SPMV = 1, PDC_KT_CONTACT_TERMS = 2,
       PDC_KT_ASSEMBLY = 3, PDC_KT_LINE_SEARCH = 4, PDC_KT_CONTACTS = 5,
       PDC_KT_PCG = 6 };
int pdc_enable_kernel_timing(pdc_context* ctx, int32_t which);
int pdc_kernel_timing(pdc_context* ctx, double* total_ms, int64_t* launches,
                      double* algorithmic_bytes);

#ifdef __cplusplus
}
#endif
#endif /* PDC_H */
