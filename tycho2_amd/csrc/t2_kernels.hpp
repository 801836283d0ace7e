// Device kernels of the sweep engine and their launch wrappers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace t2 {

// Per-cell geometry record on the device (set-up only): 4 outward unit
// normals, 4 face areas, the volume, padded to 160 bytes.
constexpr int kGeoStride = 20;
constexpr int kGeoNormal = 0;
constexpr int kGeoArea = 12;
constexpr int kGeoVolume = 16;

// Largest block of the sweep kernels (their __launch_bounds__).
constexpr int kMaxBlock = 256;

// One work item = one (angle, cell) of the sweep, stored in step order and
// read as a stream by the sweep kernels.
//   a12[k] = A_k (Omega . n_k) / 12: < 0 incoming, > 0 outgoing, 0 neither
//   v20    = V / 20
//   sig    = sigma_t of the cell (used when sigma_t does not vary by group)
template <typename T>
struct alignas(16) ItemGeo {
    T a12[4];
    T v20;
    T sig;
};

// up[k]: upwind cell through face k (>= 0), -1 when face k is outgoing,
// parallel or on the vacuum boundary, or -2 - i for lagged face i (its
// values come from the pre-sweep snapshot lag_buf[i]).  meta holds, for each
// face k, the neighbour vertex facing cell vertex j (j != k, ascending) in
// bits 6k + 2*slot.  problem indexes the problem table of a batched launch
// (0 for a single problem).
struct alignas(16) ItemTopo {
    int32_t cell;
    int32_t angle;
    int32_t up[4];
    uint32_t meta;
    int32_t problem;
};

// Per-problem pointers of a batched launch (several independent problems
// with the same group count and precision swept by one launch per step).
template <typename T>
struct ProblemDesc {
    int64_t nc;
    int sigma_mode;
    const T *sigma;
    const T *src_iso;
    const T *src_ang_t;            // [na][nc][4][G] in T, or null
    T *psi;
    const T *lag_buf;
};

enum SigmaMode { kSigmaCellGroup = 0, kSigmaGroup = 1, kSigmaCell = 2 };

// Cell solvers: elimination without pivoting (0, what a zero-initialised
// SweepArgs selects), with partial pivoting, and Cramer's rule (adjugate).
enum CellSolver { kSolveNoPivot = 0, kSolveOriginal = 1, kSolveCramer = 2 };

template <typename T>
struct SweepArgs {
    const ItemGeo<T> *geo;     // step slice of the item records
    const ItemTopo *topo;
    int64_t n_items;
    int64_t nc;
    int sigma_mode;
    const T *sigma;            // [nc][G] (cell-group) or [G] (group)
    const T *src_iso;          // [nc][4][G] isotropic total source
    const double *src_ang;     // [na][nc][4][G] angle-dependent part or null
    const T *src_ang_t;        // [na][nc][4][G] same, in T, or null
    T *psi;                    // [na][nc][4][G]
    const T *lag_buf;          // [n_lag][4][G] upwind values held pre-sweep
    // batched launch: per-problem pointers replace nc, sigma, src_iso, psi
    // and lag_buf (src_ang and src_ang_t must be null); items of problems
    // whose active[] flag is 0 are skipped (active may be null: all)
    const ProblemDesc<T> *problems;
    const int32_t *active;
    // CellSolver; the level engine's kernels take any, the ready-queue
    // kernels kSolveNoPivot
    int solver;
    // level engine, kSolveNoPivot, one group per lane, one problem (no
    // problems / active tables): 2 solves two items per lane, their loads
    // issued before either solve (0 or 1: one item per lane)
    int items_per_lane;
};

// Defect of the coupled transport equations for a stored angular flux,
// r = q + S phi - L psi, evaluated in FP64 from FP64 item records (lagged
// faces read the upwind cell's current psi: the operator is the full one),
// stored (rounded to the flux's type T) as the nodal source density whose
// weak form is r, so that a sweep with it as source gives L^-1 r.
template <typename T>
struct ResidualArgs {
    const ItemGeo<double> *geo;
    const ItemTopo *topo;
    int64_t n_items;
    int64_t nc;
    int sigma_mode;
    const double *sigma;       // FP64, layout by sigma_mode
    const double *src_iso;     // [nc][4][G] FP64 total isotropic source
    const double *src_ang;     // [na][nc][4][G] FP64 or null
    const int32_t *lag_up;     // upwind cell of lagged face i
    const T *psi;
    T *r;                      // [na][nc][4][G]
};

template <typename T>
hipError_t launch_residual(const ResidualArgs<T> &args, int groups,
                           hipStream_t stream);

// Builds the item records from the step-ordered work list; sets *err to a
// nonzero value if an item has more than three incoming interior faces.
template <typename T>
hipError_t launch_build_items(const uint32_t *work, int64_t n, int cell_bits,
                              int64_t nc, const double *omega,
                              const double *geo, const int4 *adj,
                              const uint32_t *nbr, const double *sigma_cell,
                              const int64_t *lag_key, int64_t n_lag,
                              ItemGeo<T> *out_geo, ItemTopo *out_topo,
                              int *err, hipStream_t stream);

// Copies one problem's step-ordered records into a batch's merged schedule
// (src_off / dst_off: [n_steps + 1] step offsets on the device).
template <typename T>
hipError_t launch_merge_items(const ItemGeo<T> *src_geo,
                              const ItemTopo *src_topo,
                              const int64_t *src_off, const int64_t *dst_off,
                              int64_t n_steps, int64_t n_items,
                              int32_t problem, ItemGeo<T> *dst_geo,
                              ItemTopo *dst_topo, hipStream_t stream);

// Sweep of one step: every item's cell solve for all groups.
template <typename T>
hipError_t launch_sweep_step(const SweepArgs<T> &args, int groups,
                             int groups_per_lane, int block,
                             hipStream_t stream);

// Ready-queue form of a sweep: one persistent launch over the step-ordered
// items from first_item on, split into chunks that never span two steps.
// The queued items' tables are indexed by item - first_item.
struct QueueState {
    const int64_t *chunk_off;      // [n_chunks + 1] item ranges
    int64_t n_chunks;
    int32_t *pending;              // upwind queued items not yet solved
    const int4 *down;              // items reading the item's psi (-1 none)
    const int32_t *arm;            // counters' start values; null: all
                                   // upwind interior faces (first_item 0)
    int64_t first_item;
    unsigned long long *ticket;    // next chunk to hand out (0 at the start)
    int32_t *abort;                // set when a wait exceeds spin_limit
    int64_t spin_limit;            // polls per wait before giving up
    const int32_t *active;         // batched launch: problems to sweep
};

template <typename T>
hipError_t launch_queue_sweep(const SweepArgs<T> &args, const QueueState &q,
                              int groups, int groups_per_lane, int block,
                              int64_t grid, hipStream_t stream);
// Blocks of the queue kernel resident on the current device at once.
template <typename T>
hipError_t queue_grid(int groups, int groups_per_lane, int block,
                      int64_t *grid);

// Batched ready queue: per-problem tables and the interleaved chunk list.
template <typename T>
struct BatchProblem {
    SweepArgs<T> a;                // records (whole sweep), arrays, nc
    int32_t *pending;              // the problem's tables, as in QueueState
    const int4 *down;
    const int32_t *arm;
    int64_t first_item;
    int32_t groups;                // 1, 2, 4, ..., 64
};
struct BatchChunk {
    int64_t lo;                    // first item, in the problem's step order
    int32_t n;                     // items
    int32_t problem;
};
template <typename T>
hipError_t launch_queue_batch(const BatchProblem<T> *problems,
                              const BatchChunk *chunks, const QueueState &q,
                              int64_t grid, hipStream_t stream);
template <typename T>
hipError_t queue_batch_grid(int64_t *grid);

// Links the queued items >= first_item: down[] and pending[] (and arm[],
// if not null) from the item records; pos (na * nc) and dcount (queued
// items) are int32 scratch; *err is set if an item had more than 4 readers.
hipError_t launch_queue_links(const ItemTopo *topo, int64_t n_items,
                              int64_t nc, int64_t first_item, int32_t *pos,
                              int32_t *dcount, int4 *down, int32_t *pending,
                              int32_t *arm, int *err, hipStream_t stream);
// Resets the queued items' counters to their start values (after an
// abandoned sweep).
hipError_t launch_queue_rearm(const ItemTopo *topo, int64_t n_items,
                              int64_t first_item, const int32_t *arm,
                              int32_t *pending, hipStream_t stream);

// lag_buf[i][j][g] = psi[a_i][u_i][vertex of u_i facing c_i's vertex j][g]
template <typename T>
hipError_t launch_lag_gather(const int64_t *lag_key, const int32_t *lag_up,
                             int64_t n_lag, int64_t nc, int groups,
                             const uint32_t *nbr, const T *psi, T *lag_buf,
                             hipStream_t stream);

// phi[c][v][g] = sum_a w_a psi[a][c][v][g]  (FP64 accumulation)
template <typename T>
hipError_t launch_psi_to_phi(const T *psi, const double *wt, int64_t na,
                             int64_t n_phi, double *phi, hipStream_t stream);

// out[c][v][g] = (q ? q[c][v][g] : 0) + sigma_s[c][g] / (4 pi) * phi[c][v][g]
template <typename T>
hipError_t launch_total_source(const double *q, const double *sigma_s,
                               const double *phi, int64_t nc, int groups,
                               T *out, hipStream_t stream);

// Layout changes between the reference's [c][a][v][g] and the engine's
// [a][c][v][g] (vg = 4 * groups contiguous values per (cell, angle)).
template <typename T>
hipError_t launch_ca_to_ac(const double *in, T *out, int64_t nc,
                           int64_t na, int64_t vg, hipStream_t stream);
// out = in (+ corr, when not null, added in FP64)
template <typename T>
hipError_t launch_ac_to_ca(const T *in, double *out, int64_t nc, int64_t na,
                           int64_t vg, hipStream_t stream,
                           const T *corr = nullptr);

// Deterministic reductions: partial sums per block into `partial`
// (kRedBlocks entries per result), then one block adds them in order.
constexpr int kRedBlocks = 1024;
constexpr int kRedThreads = 256;
// out[0] = sum |x - y|, out[1] = sum |x|
hipError_t launch_l1_change(const double *x, const double *y, int64_t n,
                            double *partial, double *out, hipStream_t stream);
// out[j] = sum_i V[j][i] w[i], j < m  (V holds m vectors of length n)
hipError_t launch_multi_dot(const double *V, int64_t ldv, int m,
                            const double *w, int64_t n, double *partial,
                            double *out, hipStream_t stream);
// w -= sum_j h[j] V[j]  (h on the device)
hipError_t launch_multi_axpy(const double *V, int64_t ldv, int m,
                             const double *h, double *w, int64_t n,
                             hipStream_t stream);
// y = alpha * x + beta * y  (x and y distinct)
hipError_t launch_axpby(double alpha, const double *x, double beta, double *y,
                        int64_t n, hipStream_t stream);
// x = alpha * x
hipError_t launch_scale(double *x, double alpha, int64_t n,
                        hipStream_t stream);
hipError_t launch_fill(double *x, double v, int64_t n, hipStream_t stream);

}  // namespace t2
