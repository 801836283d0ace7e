/*
 * Tycho2 CPU oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's one-rank FP64 sweep path
 * (SweepType TraverseGraph, GaussElim NoPivot/Original) used as the checker
 * for the HIP sweep engine.  Only tests, the smoke check and the benchmark's
 * CPU-baseline leg may load it; the product never calls it.
 *
 * Layouts follow the reference:
 *   psi / source  [cell][angle][vertex][group]   (src/PsiData.hh:162)
 *   phi           [cell][vertex][group]          (src/PsiData.hh:369)
 *   sigma         [cell][group]  (the reference is per cell only,
 *                                 src/Problem.cc:149-172; groups are
 *                                 decoupled, so a per-group value is the same
 *                                 as one reference run per group)
 * Mesh indices are int64 with -1 for the reference's UINT64_MAX sentinels.
 *
 * Parity status of each function is stated in oracle/README.md and pinned by
 * tests/test_oracle_*.py.
 */
#ifndef TYCHO2_ORACLE_H
#define TYCHO2_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* cell solvers, the reference's GaussElim deck values (src/Global.hh:99-105);
 * both Cramer values select one adjugate solve (see or_cell_solve) */
enum {
    OR_GAUSS_ORIGINAL = 0,
    OR_GAUSS_NOPIVOT = 1,
    OR_GAUSS_CRAMER_GLU = 2,
    OR_GAUSS_CRAMER_INTEL = 3
};

typedef struct {
    int64_t nc;
    const int64_t *adj;     /* [nc][4]   neighbour cell or -1            */
    const int64_t *c2f;     /* [nc][4][4] cell vertex -> face vertex, -1 on k */
    const int64_t *nbrv;    /* [nc][4][3] face vertex -> neighbour vertex */
    const double *volume;   /* [nc]                                        */
    const double *area;     /* [nc][4]                                     */
    const double *normal;   /* [nc][4][3] outward unit normals             */
} or_mesh;

/* Build connectivity and geometry from one-partition serial-mesh arrays. */
int or_mesh_build(int64_t nc, int64_t nf, int64_t nn,
                  const int64_t *cell_faces, const int64_t *cell_nodes,
                  const int64_t *face_cells, const int64_t *face_nodes,
                  const double *coords,
                  int64_t *adj, int64_t *f2c, int64_t *c2f, int64_t *nbrv,
                  double *volume, double *area, double *normal);

/* Product quadrature of order sn: 2 sn^2 angles. Returns 0 or -1. */
int or_quadrature(int sn, double *xi, double *eta, double *mu, double *wt);

/* One cell solve (for unit tests): abar[k] = A_k (Omega . n_k),
 * upwind[k][j] = psi entering through face k at cell vertex j. */
void or_cell_solve(int method, double volume, const double abar[4],
                   double sigma_t, const double q[4],
                   const double upwind[4][4], double out[4]);

/* One full transport sweep psi = L^-1 source over all angles and groups.
 * Cyclic dependencies are broken by lagging faces (see the .c file); the
 * number of lagged (angle, cell, face) triples is returned in *n_lagged. */
int or_sweep(const or_mesh *m, int64_t na, const double *omega, int64_t ng,
             const double *sigma_t, const double *source, double *psi,
             int method, int nthreads, int64_t *n_lagged);

/* Sweep a subset of angles [a0, a1) only (CPU-baseline sampling). */
/* Dependency rule of every following sweep: tol < 0 (default) the graph
 * traverser's, tol >= 0 the precomputed schedule's (faces with
 * -tol <= Omega.n <= 0 lagged from the start).  Process-wide. */
void or_set_dependency_tolerance(double tol);

int or_sweep_angles(const or_mesh *m, int64_t na, const double *omega,
                    int64_t ng, const double *sigma_t, const double *source,
                    double *psi, int method, int nthreads, int64_t a0,
                    int64_t a1, int64_t *n_lagged);

/* Defect r = q - L psi of every (cell, angle, group) for a given psi and
 * total source, as the nodal density whose weak form is r (the full
 * operator: no face is lagged).  source, psi, density: [c][a][v][g]. */
void or_defect(const or_mesh *m, int64_t na, const double *omega, int64_t ng,
               const double *sigma_t, const double *source, const double *psi,
               double *density);

void or_psi_to_phi(int64_t nc, int64_t na, int64_t ng, const double *wt,
                   const double *psi, double *phi);

void or_total_source(int64_t nc, int64_t na, int64_t ng,
                     const double *source, const double *sigma_s,
                     const double *phi, double *total);

/* Source iteration (reference SourceIteration::fixedPoint).  psi is in/out
 * (start from zero for the reference's behaviour); phi gets the final scalar
 * flux; err_hist (length iter_max, may be NULL) the L1-relative changes. */
int or_source_iteration(const or_mesh *m, int64_t na, const double *omega,
                        const double *wt, int64_t ng, const double *sigma_t,
                        const double *sigma_s, const double *source,
                        int64_t iter_max, double err_max, int method,
                        int nthreads, double *psi, double *phi,
                        double *err_hist, int64_t *n_iter);

/* Restarted GMRES(restart) on (I - D L^-1 M S) phi = D L^-1 q with x0 = 0,
 * then one final sweep for psi (reference SourceIteration::krylov). */
int or_krylov(const or_mesh *m, int64_t na, const double *omega,
              const double *wt, int64_t ng, const double *sigma_t,
              const double *sigma_s, const double *source, int64_t iter_max,
              double rtol, int64_t restart, int method, int nthreads,
              double *psi, double *phi, double *res_hist, int64_t *n_iter);

/* Dense GMRES on a small explicit matrix (row-major), for textbook checks. */
int or_gmres_dense(int64_t n, const double *a, const double *b, double *x,
                   int64_t restart, int64_t iter_max, double rtol,
                   int64_t *n_iter);

/* The reference's manufactured hat source on [0,100]^3. */
void or_hat_source(int64_t nc, const int64_t *cell_nodes, const double *coords,
                   int64_t na, const double *xi, const double *eta,
                   const double *mu, int64_t ng, const double *sigma_t_cell,
                   const double *sigma_s_cell, double *source);

#ifdef __cplusplus
}
#endif
#endif
