/*
 * orc.h — ORACLE, TEST INFRASTRUCTURE ONLY.
 *
 * C entry points of the CPU FP64 restatement of the reference's per-timestep
 * contact pipeline (pdc_oracle.cpp).  Only the test suite, the smoke check and
 * the benchmark's CPU-baseline leg load this library, as the checker and as
 * the timed CPU path.  The product library never links or calls it.
 */
#ifndef ORC_H
#define ORC_H

#include <stdint.h>
#include "../include/pdc.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_scene orc_scene;

/* Linear solver of the Newton step.  ORC_SOLVER_REFERENCE follows the
 * reference's rule: direct factorisation when fewer than 2000 bodies are in
 * contact, PCG otherwise (PrimalDualForceSolver.cpp:385). */
enum { ORC_SOLVER_REFERENCE = 0, ORC_SOLVER_DIRECT = 1, ORC_SOLVER_PCG = 2 };

orc_scene* orc_create(const pdc_params* p, int32_t n, const int32_t* shape_type,
                      const double* dims, const double* mass,
                      const double* inertia_diag, const int32_t* is_static,
                      const int32_t* emit_samples);
/* The same with triangle meshes (bodies of shape PDC_SHAPE_MESH name mesh
 * dims[3i] = index): n_meshes meshes, mesh m with mesh_nv[m] vertices (body
 * frame) and mesh_nt[m] triangles, concatenated in mesh_v / mesh_tri. */
orc_scene* orc_create_meshes(const pdc_params* p, int32_t n, const int32_t* shape_type,
                             const double* dims, const double* mass, const double* inertia_diag,
                             const int32_t* is_static, const int32_t* emit_samples, int32_t n_meshes,
                             const int32_t* mesh_nv, const double* mesh_v, const int32_t* mesh_nt,
                             const int32_t* mesh_tri);
void orc_destroy(orc_scene* s);
void orc_set_params(orc_scene* s, const pdc_params* p);
/* threads <= 0: the reference's rules (1 thread below 1000 unknowns, else all
 * threads).  threads > 0: fixed count. */
void orc_set_threads(orc_scene* s, int32_t threads);
void orc_set_linear_solver(orc_scene* s, int32_t which);
void orc_set_state(orc_scene* s, const double* pos, const double* quat,
                   const double* vel, const double* angvel);
void orc_get_state(const orc_scene* s, double* pos, double* quat, double* vel,
                   double* angvel);

/* (a) contacts on the current state; returns the contact count. */
int32_t orc_detect_contacts(orc_scene* s, int32_t* n_ties);
int32_t orc_get_contacts(const orc_scene* s, int32_t cap, int32_t* ids,
                         double* pos, double* depth, double* transformation,
                         double* local_force);
/* Raw (pre-merge) sample hits of the last detection: count only. */
int32_t orc_raw_contact_count(const orc_scene* s);

/* fillForces on caller-given contacts and state (bodies from orc_create). */
int32_t orc_fill_forces(orc_scene* s, int32_t C, const int32_t* ids,
                        const double* transformation, const double* depth,
                        const double* quat, const double* vel,
                        const double* angvel, double* out_acc,
                        double* out_local_force, pdc_solve_stats* stats);

/* n_steps of detect + fillForces + integrate.  `stats` gets the last step. */
int32_t orc_step(orc_scene* s, int32_t n_steps, pdc_solve_stats* stats);

/* Per-iteration trace of the last fillForces: for each Newton iteration
 * 8 doubles {err, exitErr, eps, Hu(before update), stepSize, pcgIters, gap,
 * sacErr}.  Returns the number of iterations written. */
int32_t orc_trace(const orc_scene* s, int32_t cap, double* out);

/* ---- building blocks, exposed for unit tests ---- */
void orc_lorentz(double mu, const double f[3], double* c, double g[2],
                 double H[4], double* radius);
void orc_lorentz_project(double mu, const double in[3], double out[3]);
void orc_normal_to_transform(const double pos[3], const double n[3],
                             const double x0[3], const double x1[3],
                             double T[36]);
/* Grid-SDF value and gradient of body b at a point in its local frame.
 * Returns 0 when the point is outside the grid box. */
int32_t orc_sdf_query(const orc_scene* s, int32_t b, const double q[3],
                      double* sd, double grad[3]);
/* FP32 node value of body b's grid at node (i,j,k). */
float orc_sdf_node(const orc_scene* s, int32_t b, int32_t i, int32_t j,
                   int32_t k);
int32_t orc_samples(const orc_scene* s, int32_t b, int32_t cap, double* out);
/* Analytic signed distance of the ellipsoid with semi-axes dims at n points
 * (xyz rows). */
void orc_ellipsoid_sdf(const double dims[3], int32_t n, const double* pts, double* out);
/* Signed distance of n points to the triangle mesh (nv vertices, nt
 * triangles) by the pseudonormal rule. */
void orc_mesh_sdf(int32_t nv, const double* v, int32_t nt, const int32_t* tri, int32_t n, const double* pts,
                  double* out);

/* Assembled Newton system of the first iteration of fillForces on the given
 * contacts/state (before Jacobi scaling): dense 6N x 6N matrix (row-major)
 * and rhs.  Returns 6N, or -1 if cap is too small.  Local numbering is the
 * reference's: order of first appearance in the contact list. */
int32_t orc_first_system(orc_scene* s, int32_t C, const int32_t* ids,
                         const double* transformation, const double* depth,
                         const double* quat, const double* vel,
                         const double* angvel, int32_t cap, double* A,
                         double* rhs, int32_t* local_to_global);

/* Dense SPD solve by the oracle's block-sparse Cholesky (for tests):
 * A is n x n row-major, n a multiple of 6.  Returns 0 on success. */
int32_t orc_block_cholesky_solve(int32_t n, const double* A, const double* b,
                                 double* x);
/* PCG on a dense SPD matrix (for tests): from x as given when `warm` is
 * nonzero, else from zero; preconditioned by block Jacobi when rep is NULL,
 * else by the inverse blocks of the groups of block rows given by rep (per
 * block row, the row holding its group's key).  Returns the iteration count. */
int32_t orc_pcg_solve(int32_t n, const double* A, const double* b, double tol,
                      int32_t maxit, double* x, int32_t warm, const int32_t* rep);
/* The body groups of the standard PCG form's preconditioner (per body the
 * index of the body holding its group's key): n bodies with keys, static
 * flags and contact counts w; np pairs (pa, pb) with pn contacts each. */
int32_t orc_group_bodies(int32_t n, const int64_t* key, const int32_t* stat, const int64_t* w,
                         int32_t np, const int32_t* pa, const int32_t* pb, const int32_t* pn,
                         int32_t* rep);

#ifdef __cplusplus
}
#endif
#endif
