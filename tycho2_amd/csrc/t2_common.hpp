// Shared host-side declarations of the sweep engine.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "tycho2_sweep.h"

namespace t2 {

// Records the failure message for t2_last_error() and returns `code`.
int fail(int code, const std::string &msg);
void clear_error();

// Per-cell geometry and connectivity in the engine's host layout.
struct HostMesh {
    int64_t nc = 0;
    std::vector<double> volume;    // [nc]
    std::vector<double> area;      // [nc][4]
    std::vector<double> normal;    // [nc][4][3]
    std::vector<int64_t> adj;      // [nc][4], -1 on the boundary
    // neighbour vertex through face k at cell vertex j (j != k), two bits at
    // position 2*(4k+j); the j == k slots are zero.
    std::vector<uint32_t> nbr_pack;
};

// Geometry of the reference's TychoMesh constructor.
int mesh_geometry(int64_t nc, int64_t nn, const double *coords,
                  const int64_t *cell_nodes, const int64_t *f2c,
                  double *volume, double *area, double *normal);

int build_host_mesh(int64_t nc, int64_t nn, const double *coords,
                    const int64_t *cell_nodes, const int64_t *adj,
                    const int64_t *f2c, HostMesh &m);

// Omega . n with every product and sum rounded separately (this file is
// built with -ffp-contract=off; the device evaluates the same expression
// with explicitly rounded operations, so both sides agree on which faces are
// incoming).
inline double omega_dot_n(const double *om, const double *n)
{
    return om[0] * n[0] + om[1] * n[1] + om[2] * n[2];
}

struct LaggedFace {
    int32_t angle;
    int32_t cell;
    int32_t face;
    int32_t upwind;
};

// Sweep schedule: every (angle, cell) pair in level order.
struct Schedule {
    int64_t n_steps = 0;
    int64_t min_levels = 0;
    int cell_bits = 0;                 // work item = cell | angle << cell_bits
    std::vector<int64_t> step_off;     // [n_steps + 1]
    std::vector<uint32_t> work;        // [nc * na]
    std::vector<LaggedFace> lagged;    // sorted by (angle, cell, face)
    std::vector<int64_t> batch_steps;  // angle batch b: steps [b], [b + 1]
};

// Dependency rule of the level sets.  dep_tol < 0: the graph traverser's,
// every interior face with Omega.n <= 0 is a dependency
// (src/TychoMesh.hh:107-110).  dep_tol >= 0: the precomputed schedule's,
// only faces with Omega.n < -dep_tol are (src/SweepSchedule.cc:57,108-125);
// the other interior faces with Omega.n <= 0 are lagged from the start.
constexpr double kTraverseRule = -1.0;

// Levels of one angle; returns the lagged faces appended to `lagged`.
void angle_levels(const HostMesh &m, const double *om, int32_t angle,
                  int32_t *level, std::vector<LaggedFace> &lagged,
                  std::vector<int32_t> &scratch_indeg,
                  std::vector<int32_t> &scratch_frontier,
                  double dep_tol = kTraverseRule);

int build_levels(const HostMesh &m, int64_t na, const double *omega,
                 int32_t *levels, std::vector<LaggedFace> &lagged,
                 double dep_tol = kTraverseRule);

// Angles swept together: octant by octant (sign pattern of Omega), then by
// index, cut into `batches` nearly equal consecutive groups.
std::vector<int32_t> angle_batches(int64_t na, const double *omega,
                                   int batches);

// Items in step order: by angle batch and level, then within a step by
// angle and, for order T2_ORDER_FACES, the pattern of incoming interior
// faces, then by cell.
int build_schedule(const HostMesh &m, int64_t na, const double *omega,
                   int batches, int order, Schedule &s,
                   double dep_tol = kTraverseRule);

// Traversal priorities of the reference's graph traverser on one rank
// (src/Priorities.cc:396-435): intra-angle rule 0 random, 1 b-level,
// 2 BFDS, 3 DFDS, 4 DFHDS; inter-angle rule 0 interleaved, 1 globally
// prioritised angles, 2 locally prioritised angles.  prio[a][c], uint64
// with the reference's wrap-around arithmetic.  `levels` and `lagged` are
// what build_levels returns for the same mesh and directions.
int sweep_priorities(const HostMesh &m, int64_t na, const double *omega,
                     const int32_t *levels,
                     const std::vector<LaggedFace> &lagged, int intra,
                     int inter, std::vector<uint64_t> &prio);

// Items in priority list-schedule order: each step takes up to `width`
// (0: all) of the items ready when it starts, highest priority first (ties:
// lower angle, then lower cell), so every step is a set of mutually
// independent items and every item comes after its upwind items.  Within a
// step, items stay in that order (T2_ORDER_CELL) or are grouped stably by
// their pattern of incoming interior faces (T2_ORDER_FACES).
int build_priority_schedule(const HostMesh &m, int64_t na,
                            const double *omega, int intra, int inter,
                            int64_t width, int order, Schedule &s,
                            double dep_tol = kTraverseRule);

}  // namespace t2
