/*
 * tycho2_sweep.h -- C ABI of the MI355X sweep engine (libtycho2_sweep.so).
 *
 * The engine replaces, for one call of SweeperAbstract::sweep(psi, source)
 * and for the outer iterations that drive it, the reference's sweep stack:
 *   SweeperAbstract / SweeperTraverse  src/SweeperAbstract.hh:47-70,
 *                                      src/SweeperTraverse.cc:55-89
 *   GraphTraverser + Priorities        src/GraphTraverser.cc:416-759,
 *                                      src/Priorities.cc:396-435
 *   SweepData + Transport cell solve   src/SweepData.hh:139-174,
 *                                      src/Transport.cc:687-776
 *   PsiData / PhiData storage          src/PsiData.hh:59-371
 *   moments + outer iterations         src/Util.cc:85-159,
 *                                      src/SourceIteration.cc:76-285
 *
 * Conventions
 *   - Plain pointers and sizes only.  Host arrays use the reference's layouts:
 *       psi, source  [cell][angle][vertex][group]  (src/PsiData.hh:162)
 *       phi          [cell][vertex][group]         (src/PsiData.hh:369)
 *       sigma        [cell][group]
 *     Mesh indices are int64 with -1 where the reference uses UINT64_MAX.
 *   - Every entry point returns T2_OK (0) or a negative status; the message
 *     of the last failure on the calling thread is t2_last_error().  No C++
 *     exception crosses the boundary (the reference aborts instead,
 *     src/Assert.cc:51-65).
 *   - One host thread at a time per handle (the reference is
 *     MPI_THREAD_SINGLE, src/Main.cc:197-201).
 */
#ifndef TYCHO2_SWEEP_H
#define TYCHO2_SWEEP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define T2_ABI_VERSION 3

/* status codes */
#define T2_OK 0
#define T2_ERR_ARG -1      /* bad argument or inconsistent mesh            */
#define T2_ERR_HIP -2      /* HIP runtime failure                          */
#define T2_ERR_NOMEM -3    /* device or host allocation failed             */
#define T2_ERR_STATE -4    /* call not valid in the handle's current state */
#define T2_ERR_NODEV -5    /* no usable GPU                                */
#define T2_ERR_NOCONV -6   /* iteration hit its limit without converging   */
#define T2_ERR_TIMEOUT -7  /* a ready-queue wait ran past its poll limit   */

/* sweep arithmetic (also the angular-flux storage type) */
#define T2_FP64 0
#define T2_FP32 1

/* cell solvers, deck key GaussElim (src/Main.cc:167-176, values as
 * src/Global.hh:99-105): elimination with partial pivoting, without
 * pivoting (the fastest; every engine), Cramer's rule (both Cramer options
 * select one adjugate solve).  Solvers other than NoPivot run on the level
 * engine with one group per lane. */
#define T2_GAUSS_ORIGINAL 0
#define T2_GAUSS_NOPIVOT 1
#define T2_GAUSS_CRAMER_GLU 2
#define T2_GAUSS_CRAMER_INTEL 3

typedef struct t2_sweeper t2_sweeper;

typedef struct {
    int64_t n_cells, n_angles, n_groups;
    int64_t n_steps;          /* sweep steps (max levels over angles)      */
    int64_t min_levels;       /* fewest levels of any angle                */
    int64_t n_lagged;         /* (angle, cell, face) triples lagged        */
    int64_t psi_bytes;        /* device bytes of the angular flux          */
    int64_t device_bytes;     /* all device memory held by the handle      */
    int32_t precision;        /* T2_FP64 / T2_FP32                         */
    int32_t gauss;            /* T2_GAUSS_*                                */
    int64_t sweeps;           /* full sweeps issued by the handle so far   */
    int32_t engine;           /* T2_ENGINE_*                               */
    int32_t angle_batches;    /* schedule knob (t2_set_schedule)           */
    int32_t items_per_block;  /* launch knobs (t2_set_launch)              */
    int32_t groups_per_lane;
    int32_t chunk_rounds;     /* queue knob (t2_set_engine)                */
    int32_t queue_first_step; /* first step run by the queue (-1: none)    */
    int32_t intra_angle_p;    /* priority schedule rules (t2_set_priority; */
    int32_t inter_angle_p;    /*   -1 and 0: level sets)                   */
    int64_t max_cells_per_step;
    int32_t concurrent_batches; /* angle batches on parallel streams     */
} t2_info;

/* Message of the last failure on this thread ("" if none). */
const char *t2_last_error(void);
int t2_abi_version(void);

/* ------------------------------------------------------------ host helpers
 * Pure host functions (no device work); usable without a GPU. */

/* Cell volumes, face areas and outward unit normals as the reference's
 * TychoMesh constructor computes them (src/TychoMesh.cc:52-208), with the
 * face vertex order f2c[c][k][j] (src/TychoMesh.hh:99-100). */
int t2_mesh_geometry(int64_t n_cells, int64_t n_nodes, const double *coords,
                     const int64_t *cell_nodes, const int64_t *f2c,
                     double *volume, double *area, double *normal);

/* Level of every (angle, cell) in the per-angle sweep order, levels[a][c]
 * (replaces GraphTraverser's ready queue, src/GraphTraverser.cc:467-644).
 * Dependency cycles, on which the reference's traversal never ends, are
 * broken by lagging the weakest unmet face (smallest A |Omega.n|); the number
 * of lagged (angle, cell, face) triples is returned in *n_lagged. */
int t2_sweep_levels(int64_t n_cells, const int64_t *adj_cell,
                    const double *area, const double *normal,
                    int64_t n_angles, const double *omega, int32_t *levels,
                    int64_t *n_lagged);
/* The same with the dependency rule of t2_set_dependency_tolerance:
 * dep_tol >= 0 makes only faces with Omega.n < -dep_tol dependencies and
 * lags the other interior faces with Omega.n <= 0 from the start (counted
 * in *n_lagged); dep_tol < 0 is t2_sweep_levels. */
int t2_sweep_levels_tol(int64_t n_cells, const int64_t *adj_cell,
                        const double *area, const double *normal,
                        int64_t n_angles, const double *omega,
                        double dep_tol, int32_t *levels, int64_t *n_lagged);

/* Traversal priorities of the reference's graph traverser on one rank,
 * priorities[a][c] (Priorities::calcPriorities, src/Priorities.cc:396-435;
 * deck keys intraAngleP, interAngleP, src/Main.cc:111-112):
 *   intra_angle_p  0 random (a seeded 31-bit draw per item), 1 b-level
 *                  (longest downwind path), 2 BFDS, 3 DFDS, 4 DFHDS;
 *   inter_angle_p  0 interleaved, 1 globally prioritised angles (higher
 *                  index first), 2 locally prioritised angles.
 * Unsigned 64-bit arithmetic as in the reference; with no partition sides
 * on one rank, DFDS and DFHDS coincide.  Directions with lagged faces use
 * the dependency graph left after lagging. */
int t2_sweep_priorities(int64_t n_cells, const int64_t *adj_cell,
                        const double *area, const double *normal,
                        int64_t n_angles, const double *omega,
                        int intra_angle_p, int inter_angle_p,
                        uint64_t *priorities);

/* The priority list schedule t2_set_priority runs: each step holds up to
 * max_cells_per_step (0: no limit) of the (angle, cell) items ready when it
 * starts, highest priority first (ties: lower angle, then lower cell) --
 * GraphTraverser::traverse's ready heap popped maxCellsPerStep times per
 * step (src/GraphTraverser.cc:518-536).  items[n_cells * n_angles] receives
 * angle * n_cells + cell in schedule order, step_off[n_steps + 1] (room for
 * n_cells * n_angles + 1) the step boundaries; order as t2_set_order. */
int t2_priority_schedule(int64_t n_cells, const int64_t *adj_cell,
                         const double *area, const double *normal,
                         int64_t n_angles, const double *omega,
                         int intra_angle_p, int inter_angle_p,
                         int64_t max_cells_per_step, int order,
                         int64_t *items, int64_t *step_off, int64_t *n_steps);

/* ------------------------------------------------------------ setup
 * Replaces the reference's per-mesh set-up (TychoMesh, Quadrature,
 * GraphTraverser constructor, Priorities::calcPriorities; src/Main.cc:
 * 245-281).  Mesh: node coordinates, cell nodes, cell adjacency (-1 on the
 * boundary) and face-to-cell vertex map as TychoMesh exposes them
 * (getNodeCoord, getCellNode, getAdjCell, getFaceToCellVrtx).  Quadrature:
 * omega[a][3] and weights[a].  Cross sections sigma[cell][group]. */
int t2_sweeper_create(t2_sweeper **out, int device,
                      int64_t n_cells, int64_t n_nodes, const double *coords,
                      const int64_t *cell_nodes, const int64_t *adj_cell,
                      const int64_t *f2c, int64_t n_angles,
                      const double *omega, const double *weights,
                      int64_t n_groups, const double *sigma_t,
                      const double *sigma_s, int precision, int gauss);
/* Environment: T2_NO_GRAPH=1 makes the level engine launch its per-step
 * kernels directly instead of replaying a captured hipGraph (a debugging
 * aid; results are identical). */
int t2_sweeper_destroy(t2_sweeper *s);
int t2_sweeper_info(const t2_sweeper *s, t2_info *info);

/* Launch shape: cells per block and groups per lane (autotuner knobs);
 * 0 keeps the default. */
int t2_set_launch(t2_sweeper *s, int items_per_block, int groups_per_lane);

/* Items per lane of the level engine's step launches (1 or 2; NoPivot cell
 * solver and one group per lane only): with 2 a wavefront solves two sets
 * of items, all of both sets' loads issued before either set is solved, so
 * that twice the loads are in flight at the same occupancy.  Results are
 * bitwise those of one item per lane (autotuner knob; the ready queue
 * ignores it).  No reference counterpart: it maps the reference's cell
 * loop (src/GraphTraverser.cc:467-759) onto wavefront lanes. */
int t2_set_items_per_lane(t2_sweeper *s, int items);

/* Sweep schedule: the angles are swept in `angle_batches` consecutive
 * batches (octant by octant, then by index), each batch level by level.  One
 * batch runs every angle concurrently; more batches shrink each step's
 * working set so that upwind reads stay in cache (autotuner knob).  Results
 * do not depend on it. */
int t2_set_schedule(t2_sweeper *s, int angle_batches);

/* Level engine: sweep the angle batches of t2_set_schedule concurrently,
 * batch b > 0 on a stream of its own forked from and joined back to the
 * handle's stream (parallel branches of the captured hipGraph), so that one
 * batch's launch gaps are filled with another's work; 0 (default) sweeps
 * them one after the other.  Batches are independent within a sweep, so
 * results do not depend on it.  No reference counterpart (the reference's
 * angle groups are OpenMP threads, src/GraphTraverser.cc:151-176). */
int t2_set_concurrent_batches(t2_sweeper *s, int concurrent);

/* Dependency rule of the sweep schedule.  tol < 0 (default): the graph
 * traverser's, every interior face with Omega.n <= 0 waits for its upwind
 * cell (src/TychoMesh.hh:107-110, src/GraphTraverser.cc:433-457).
 * tol >= 0: the precomputed schedule's of SweepType OriginalTycho1/2
 * (TOL = 1e-10, src/SweepSchedule.cc:57,108-125): only faces with
 * Omega.n < -tol are dependencies; an interior face with
 * -tol <= Omega.n < 0 reads its upwind cell's psi from before the sweep, as
 * a lagged face does (the reference's Sweeper reads whatever its psi array
 * holds when the cell is solved, src/Sweeper.cc:281-300, which is that value
 * or this sweep's depending on the order; here it is order independent).
 * Faces with Omega.n = 0 couple nothing either way (src/Transport.cc:113,
 * 184).  Cycles that remain are broken as in t2_sweep_levels. */
int t2_set_dependency_tolerance(t2_sweeper *s, double tol);

/* Order of the items within a step (any order is valid; results do not
 * depend on it): T2_ORDER_CELL by angle then cell; T2_ORDER_FACES (default)
 * by angle, then the cell's pattern of incoming interior faces, then cell,
 * so that the items sharing a wavefront read the same faces. */
#define T2_ORDER_CELL 0
#define T2_ORDER_FACES 1
int t2_set_order(t2_sweeper *s, int order);

/* Priority schedule in place of the level sets (deck keys intraAngleP,
 * interAngleP, maxCellsPerStep; src/Main.cc:110-112): the list schedule of
 * t2_priority_schedule.  Every step is a set of independent items, so each
 * engine runs it; for the ready queue it is the ticket order, the
 * reference's heap policy.  intra_angle_p = -1 returns to level sets
 * (default).  angle_batches does not apply to a priority schedule.  Results
 * do not depend on it. */
int t2_set_priority(t2_sweeper *s, int intra_angle_p, int inter_angle_p,
                    int64_t max_cells_per_step);

/* Sweep engine (GraphTraverser::traverse, src/GraphTraverser.cc:467-759):
 *   T2_ENGINE_LEVELS  one kernel launch per schedule step, the whole sweep
 *                     captured once in a hipGraph and replayed (default);
 *   T2_ENGINE_QUEUE   one persistent launch: blocks take chunks of items in
 *                     schedule order from a device ticket counter, each item
 *                     waits on its atomic pending-upwind counter, solves, and
 *                     decrements the counters of the cells that read it;
 *   T2_ENGINE_HYBRID  per-step launches while a step fills the GPU, then one
 *                     ready-queue launch for the thin tail of the schedule
 *                     (the steps that hold fewer items than the queue kernel
 *                     keeps resident; t2_info.queue_first_step).
 * Results are bitwise the same.  chunk_rounds: items per lane slot in one
 * queue chunk (a chunk is chunk_rounds x items_per_block items; 0 = 4).
 * spin_limit bounds every queue wait (polls; 0 = default); a wait that runs
 * out abandons the sweep, and the next synchronising call on the handle
 * returns T2_ERR_TIMEOUT (the counters are re-armed, so later sweeps are
 * valid). */
#define T2_ENGINE_LEVELS 0
#define T2_ENGINE_QUEUE 1
#define T2_ENGINE_HYBRID 2
int t2_set_engine(t2_sweeper *s, int engine, int chunk_rounds,
                  int64_t spin_limit);
/* Hybrid engine: first step run by the ready queue (steps before it are
 * launched one by one); -1 (default) picks the thin tail automatically. */
int t2_set_queue_split(t2_sweeper *s, int64_t first_step);

/* ------------------------------------------------------------ drop-in sweep
 * SweeperAbstract::sweep(psi, source, zeroPsiBound) on host arrays in the
 * reference layout: psi = L^-1 source for all angles and groups.  psi is
 * written (its old contents are read only through lagged faces).
 * zero_psi_bound has no effect on one device (no partition faces). */
int t2_sweep_host(t2_sweeper *s, double *psi, const double *source,
                  int zero_psi_bound);

/* ------------------------------------------------------------ device loop
 * The fixed source of SweeperAbstract::solve(): isotropic q[cell][vertex]
 * [group], or the reference's angle-dependent source[cell][angle][vertex]
 * [group] (host arrays; kept on the device).  Setting one clears the other;
 * NULL clears both. */
int t2_set_source_isotropic(t2_sweeper *s, const double *q);
int t2_set_source_angular(t2_sweeper *s, const double *source);

/* Source iteration on the device (src/SourceIteration.cc:127-207): starts
 * from psi = 0, stops when the L1-relative phi change <= err_max or after
 * iter_max sweeps.  err_hist (host, length iter_max) may be NULL.
 * Returns T2_ERR_NOCONV when iter_max stopped it above err_max (the
 * reference stops there silently); phi, psi and *n_iter then hold the last
 * iterate. */
int t2_solve_source_iteration(t2_sweeper *s, int64_t iter_max, double err_max,
                              int64_t *n_iter, double *err_hist);

/* t2_solve_source_iteration without any device allocation.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * once their handle exists (serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, arithmetic and results, bit for bit; the
 * L1-change reduction keeps its partial sums in a fixed device array of the
 * library (2 x 1024 + 2 doubles, 16 KiB, one copy per GPU, held until the
 * process ends) instead of the handle's buffers, and the call makes no
 * allocation, graph instantiation or stream creation: a sweep whose hipGraph
 * is not captured yet is launched step by step.  The array's size does not
 * depend on the problem (the reduction runs 1024 blocks over any length),
 * so no problem size exceeds it.  Returns T2_ERR_STATE when a resource
 * would have to be created first: the ready-queue tables of
 * T2_ENGINE_QUEUE / T2_ENGINE_HYBRID and the streams of concurrent angle
 * batches are made by the handle's first ordinary sweep or solve.  Every
 * call on one GPU uses that GPU's one array: calls that overlap in time on
 * the same GPU (from several host threads) race and must not be made. */
int t2_solve_source_iteration_noalloc(t2_sweeper *s, int64_t iter_max,
                                      double err_max, int64_t *n_iter,
                                      double *err_hist);

/* GMRES(restart) on (I - D L^-1 M S) phi = D L^-1 q with x0 = 0, then one
 * sweep for psi (src/SourceIteration.cc:216-285).  Arnoldi, Gram-Schmidt and
 * the iterate are FP64 on the device whatever the sweep precision.
 * res_hist (host, length iter_max) may be NULL.  Returns T2_ERR_NOCONV when
 * iter_max operator applies left the relative residual above rtol (phi and
 * psi hold the last iterate). */
int t2_solve_krylov(t2_sweeper *s, int64_t iter_max, double rtol,
                    int64_t restart, int64_t *n_iter, double *res_hist);

/* The two operator applies of the reference's Krylov path, for callers that
 * keep their own Krylov solver (PETSc's KSP in src/KrylovSolver.cc:79-169)
 * and hand it only phi-sized vectors, host [c][v][g] (src/PsiData.hh:369):
 *   t2_krylov_apply: b = (I - D L^-1 M S) x, lhsOperator of
 *     src/SourceIteration.cc:76-114 (one sweep with source M S x alone;
 *     lagged faces read the handle's psi from before the sweep);
 *   t2_krylov_rhs:   b = D L^-1 q, the right-hand side of
 *     src/SourceIteration.cc:237-244, with the handle's source.
 * Both leave the swept psi on the device and the handle's phi unchanged.
 * After the solve, t2_set_phi(x) and t2_sweep_step give
 * psi = L^-1 (M S x + q) (src/SourceIteration.cc:254-260). */
int t2_krylov_apply(t2_sweeper *s, const double *x, double *b);
int t2_krylov_rhs(t2_sweeper *s, double *b);
/* t2_krylov_apply on device vectors (memory of the handle's device, n_cells
 * x 4 x n_groups doubles each, not overlapping), for a Krylov solver that
 * keeps its vectors on the GPU: the work is enqueued on the handle's stream
 * (t2_get_stream) after everything already there, and the call returns
 * without waiting.  The caller orders its own writes of x before the call
 * and its reads of b after it (e.g. t2_synchronize). */
int t2_krylov_apply_device(t2_sweeper *s, const double *x, double *b);

/* Mixed-precision solve (handles created with T2_FP32): FP32 sweeps with an
 * FP32 angular flux inside FP64 defect correction.
 *   1. GMRES(restart) in FP64 on phi with FP32 sweeps, to rtol_sweep;
 *      psi32 = L^-1 (q + M S phi) by one more sweep, phi1 = D psi32 (FP64).
 *   2. r = q + M S phi1 - L psi32, the defect of the coupled equations,
 *      evaluated in FP64 from FP64 cell coefficients (a parallel operator
 *      apply, no sweep), stored in FP32 as a nodal source density.
 *   3. GMRES on the correction (I - D L^-1 M S) D delta = D L^-1 r with FP32
 *      sweeps, to rtol_correction; phi = phi1 + D delta in FP64.
 *   4. psi = psi32 + delta: psi32 is swept again (step 3's sweeps reused its
 *      buffer), and delta = L^-1 (r + M S D delta) by one FP32 sweep written
 *      over the residual.  t2_get_psi / t2_get_psi_angle then return
 *      psi32 + delta added in FP64: the angular flux to the accuracy of phi,
 *      not of FP32, until the next sweep or t2_defect replaces either part.
 * Memory: psi32 plus an FP32 residual (psi-sized each).  phi1 (host,
 * [cell][vertex][group], may be NULL) receives the scalar flux after step 1
 * so callers can report the error of each step.  Every step runs; the call
 * returns T2_ERR_NOCONV if either GMRES stage stopped at iter_max above its
 * tolerance (info holds both stages' residuals). */
typedef struct {
    int64_t iters_sweep;       /* GMRES iterations of step 1              */
    int64_t iters_correction;  /* GMRES iterations of step 3              */
    int64_t sweeps;            /* FP32 sweeps in total                    */
    double residual_sweep;     /* relative GMRES residual after step 1    */
    double residual_correction;
    double correction_rel;     /* ||D delta||_2 / ||phi||_2               */
} t2_refine_info;

int t2_solve_refined(t2_sweeper *s, int64_t iter_max, double rtol_sweep,
                     double rtol_correction, int64_t restart,
                     t2_refine_info *info, double *phi1);

/* FP64 defect of the transport equations for an angular flux: the operator
 * apply of step 2 above, on its own.  r = q + M S phi - L psi for every
 * (cell, angle, group), with q the handle's source, phi its scalar flux
 * (t2_set_phi or the last solve) and L the cell systems of
 * src/Transport.cc:54-271 (src/Transport.cc:687-733 per group) applied, not
 * inverted, with every incoming interior face reading the upwind cell's psi
 * (no face lagged).  The reference forms no defect; its sweep
 * (src/SweeperAbstract.hh:50-52) inverts this operator.
 * psi: host [c][a][v][g] (src/PsiData.hh:162), stored as the handle's
 * angular flux first (rounded to FP32 on FP32 handles), or NULL for the
 * handle's own.  density: host [c][a][v][g], the nodal source density d with
 * weak form V/20 (I + J) d = r on each cell, so that a sweep with source d
 * gives L^-1 r; rounded to FP32 on FP32 handles. */
int t2_defect(t2_sweeper *s, const double *psi, double *density);

/* One sweep with the current scalar flux: psi = L^-1 (q + S phi) and
 * phi <- D psi (one source-iteration step; the benchmark's unit of work). */
int t2_sweep_step(t2_sweeper *s);

/* One sweep psi = L^-1 (current total source) without the source update or
 * the moment (times the sweep kernels alone). */
int t2_sweep_only(t2_sweeper *s);

/* ------------------------------------------------------------ batches
 * Several independent problems (handles on one device with one precision and
 * isotropic sources; meshes, quadratures and group counts may differ) swept
 * together.  Members keep their own psi, sources and phi; each batch sweep
 * gives every member exactly what its own sweep would, bit for bit.
 *   T2_ENGINE_LEVELS  members with the same group count share one merged
 *                     step schedule (one launch per merged step and group
 *                     count, replayed from a hipGraph); the merged records
 *                     are a copy of the members' records;
 *   T2_ENGINE_QUEUE   one persistent launch for all members: chunks of the
 *                     members' items interleaved step by step behind one
 *                     ticket counter, each waiting on its own problem's
 *                     pending-upwind counters (no copy of the records).
 * Members must outlive the batch; a member whose schedule or source kind
 * changes afterwards makes batch calls fail with T2_ERR_STATE.  Wait limits
 * behave as for one handle (t2_batch_synchronize reports T2_ERR_TIMEOUT).
 * No reference counterpart (the reference runs one problem per process). */
typedef struct t2_batch t2_batch;
int t2_batch_create(t2_batch **out, t2_sweeper *const *members, int count);
int t2_batch_destroy(t2_batch *b);
int t2_batch_set_engine(t2_batch *b, int engine, int chunk_rounds,
                        int64_t spin_limit);
/* Level form: `chains` >= 1 sweeps independent merged schedules on
 * parallel streams (parallel branches of the captured graph): one per group
 * count, or, for chains > 1, up to `chains` per group count with that group
 * count's members dealt out by work; 0 (default) interleaves every launch
 * on one stream.  Results do not depend on it. */
int t2_batch_set_concurrent(t2_batch *b, int chains);
/* Level form: threads per block of the merged step launches (a multiple of
 * 64 in [64, 256]; default 256), one group per lane.  Results do not depend
 * on it (autotuner knob). */
int t2_batch_set_launch(t2_batch *b, int block);
/* t2_sweep_step / t2_sweep_only of every member in one batched sweep */
int t2_batch_sweep_step(t2_batch *b);
int t2_batch_sweep_only(t2_batch *b);
/* longest member schedule (steps) and total (angle, cell) items */
int t2_batch_info(const t2_batch *b, int64_t *n_steps, int64_t *n_items);
/* t2_solve_krylov / t2_solve_source_iteration of every member, advanced in
 * lock step: each round, every unfinished member's next operator apply (or
 * SI step) shares one batched sweep; finished members drop out of the sweeps.
 * Each member's phi and psi equal its own solve's bit for bit.  Per member
 * (arrays of the member count, may be NULL): iterations, and the wall time
 * from the call to the member's converged phi (seconds).  T2_ERR_NOCONV when
 * any member stopped at iter_max above its tolerance (outputs are filled). */
int t2_batch_solve_krylov(t2_batch *b, int64_t iter_max, double rtol,
                          int64_t restart, int64_t *n_iter, double *seconds);
int t2_batch_solve_source_iteration(t2_batch *b, int64_t iter_max,
                                    double err_max, int64_t *n_iter,
                                    double *seconds);
/* t2_solve_refined of every FP32 member, each of its stages in lock step
 * (the correction's right-hand side sweep carries each member's own FP32
 * residual source, and the final correction sweep writes each member's
 * residual buffer in place).  info and seconds (arrays of the member count,
 * may be NULL): each member's t2_refine_info, and the wall time from the
 * call to the end of its correction GMRES.  T2_ERR_NOCONV as for one
 * handle, when any member's stage stopped above its tolerance. */
int t2_batch_solve_refined(t2_batch *b, int64_t iter_max, double rtol_sweep,
                           double rtol_correction, int64_t restart,
                           t2_refine_info *info, double *seconds);
int t2_batch_synchronize(t2_batch *b);
int t2_batch_get_stream(const t2_batch *b, void **stream);

/* Testing hook for the out-of-memory paths: the nth device allocation of a
 * sweeper from now on (nth >= 1, counted over all handles) fails with
 * T2_ERR_NOMEM as hipMalloc running out of memory would; 0 disables it.
 * Everything allocated before the failure is released by the failing call
 * (creation) or stays owned by its handle.  No reference counterpart (the
 * reference aborts on any failure, src/Assert.cc:51-65). */
int t2_set_alloc_fail(int64_t nth);

/* Copy results to the host in the reference layouts (after a refined
 * solve, psi is psi32 + its FP32 correction added in FP64). */
int t2_get_phi(const t2_sweeper *s, double *phi);
int t2_get_psi(const t2_sweeper *s, double *psi);
/* One angle of psi, psi_a[cell][vertex][group] (e.g. to inspect full-size
 * problems whose whole psi does not fit the host). */
int t2_get_psi_angle(const t2_sweeper *s, int64_t angle, double *psi_a);
/* Replace the device scalar flux (host phi[cell][vertex][group]). */
int t2_set_phi(t2_sweeper *s, const double *phi);

/* Synchronize the handle's stream; returns the HIP stream it launches on
 * in *stream (a hipStream_t) for event timing by the caller. */
int t2_synchronize(t2_sweeper *s);
int t2_get_stream(const t2_sweeper *s, void **stream);

#ifdef __cplusplus
}
#endif
#endif
