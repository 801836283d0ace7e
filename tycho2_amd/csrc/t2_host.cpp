// Host-side set-up of the sweep engine: cell geometry and per-angle sweep
// schedules.  Pure C++ (no device calls); built with -ffp-contract=off so the
// floating-point results match the reference's TychoMesh bit for bit.
#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <random>
#include <thread>

#include "t2_common.hpp"

namespace t2 {

static thread_local std::string g_error;

int fail(int code, const std::string &msg)
{
    g_error = msg;
    return code;
}

void clear_error() { g_error.clear(); }

const char *last_error_cstr() { return g_error.c_str(); }

// ---------------------------------------------------------------- geometry
namespace {

struct V3 {
    double x, y, z;
};

inline V3 load(const double *p) { return {p[0], p[1], p[2]}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 cross(V3 a, V3 b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z,
            a.x * b.y - a.y * b.x};
}
inline bool same(V3 a, V3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

}  // namespace

// Volume |det|/6, triangle areas |cross|/2 and unit normals pointing away
// from the opposite vertex, in the reference's operation order
// (src/TychoMesh.cc:68-187).
int mesh_geometry(int64_t nc, int64_t nn, const double *coords,
                  const int64_t *cell_nodes, const int64_t *f2c,
                  double *volume, double *area, double *normal)
{
    for (int64_t c = 0; c < nc; c++) {
        V3 p[4];
        for (int v = 0; v < 4; v++) {
            int64_t node = cell_nodes[4 * c + v];
            if (node < 0 || node >= nn)
                return fail(T2_ERR_ARG, "cell node index out of range");
            p[v] = load(coords + 3 * node);
        }
        V3 e1 = p[1] - p[0], e2 = p[2] - p[0], e3 = p[3] - p[0];
        V3 x = cross(e1, e2);
        volume[c] = 1.0 / 6.0 * std::fabs(x.x * e3.x + x.y * e3.y + x.z * e3.z);
        for (int k = 0; k < 4; k++) {
            V3 f[3];
            // face k holds the three cell vertices other than k, each once,
            // so that the face and cell vertex maps are inverse
            // (src/TychoMesh.cc:227-241)
            unsigned seen = 0;
            for (int j = 0; j < 3; j++) {
                int64_t cv = f2c[(4 * c + k) * 3 + j];
                if (cv < 0 || cv > 3 || cv == k || (seen >> cv & 1u))
                    return fail(T2_ERR_ARG, "bad face-to-cell vertex map");
                seen |= 1u << cv;
                f[j] = p[cv];
            }
            V3 n = cross(f[1] - f[0], f[2] - f[0]);
            double len2 = n.x * n.x + n.y * n.y + n.z * n.z;
            area[4 * c + k] = 0.5 * std::sqrt(len2);
            double len = std::sqrt(len2);
            V3 u = {n.x / len, n.y / len, n.z / len};
            int off = 0;
            for (; off < 4; off++)
                if (!same(p[off], f[0]) && !same(p[off], f[1]) &&
                    !same(p[off], f[2]))
                    break;
            if (off == 4)
                return fail(T2_ERR_ARG, "degenerate cell");
            V3 d = p[off] - f[0];
            if (u.x * d.x + u.y * d.y + u.z * d.z > 0)
                u = {-u.x, -u.y, -u.z};
            double *out = normal + (4 * c + k) * 3;
            out[0] = u.x;
            out[1] = u.y;
            out[2] = u.z;
        }
    }
    return T2_OK;
}

int build_host_mesh(int64_t nc, int64_t nn, const double *coords,
                    const int64_t *cell_nodes, const int64_t *adj,
                    const int64_t *f2c, HostMesh &m)
{
    m.nc = nc;
    m.volume.resize(nc);
    m.area.resize(4 * nc);
    m.normal.resize(12 * nc);
    m.adj.assign(adj, adj + 4 * nc);
    m.nbr_pack.assign(nc, 0u);
    int rc = mesh_geometry(nc, nn, coords, cell_nodes, f2c, m.volume.data(),
                           m.area.data(), m.normal.data());
    if (rc)
        return rc;
    for (int64_t c = 0; c < nc; c++) {
        uint32_t pack = 0;
        for (int k = 0; k < 4; k++) {
            int64_t u = adj[4 * c + k];
            if (u < -1 || u >= nc || u == c)
                return fail(T2_ERR_ARG, "adjacent cell index out of range");
            if (u < 0)
                continue;
            bool back = false;  // u must list c as a neighbour too
            for (int kk = 0; kk < 4; kk++)
                back |= adj[4 * u + kk] == c;
            if (!back)
                return fail(T2_ERR_ARG, "adjacency is not symmetric");
            for (int j = 0; j < 4; j++) {
                if (j == k)
                    continue;
                int64_t node = cell_nodes[4 * c + j];
                int found = -1;
                for (int v = 0; v < 4; v++)
                    if (cell_nodes[4 * u + v] == node)
                        found = v;
                if (found < 0)
                    return fail(T2_ERR_ARG, "neighbour does not share face");
                pack |= uint32_t(found) << (2 * (4 * k + j));
            }
        }
        m.nbr_pack[c] = pack;
    }
    return T2_OK;
}

// ---------------------------------------------------------------- levels
// Level of each cell for one direction: sources (no interior face with
// Omega.n <= 0, the reference's incoming test, src/TychoMesh.hh:107-110 and
// src/GraphTraverser.cc:433-457) are level 0; a cell is one level above its
// last upwind neighbour.
//
// When the frontier empties before every cell has a level, the remaining
// cells sit on or behind a dependency cycle (the reference waits forever,
// src/GraphTraverser.cc:518).  The weakest unmet face -- smallest
// A |Omega.n| over the incoming interior faces of unfinished cells not yet
// released, lowest (cell, face) on ties -- is lagged: during a sweep it reads
// the angular flux its upwind cell held before the sweep.  The set of
// finished cells at a stall is the same for every traversal order, so the
// lagged set is the one the oracle's FIFO traversal finds.
//
// With dep_tol >= 0 (the precomputed schedule's rule) only faces with
// Omega.n < -dep_tol are dependencies; an interior face with
// -dep_tol <= Omega.n <= 0 is lagged from the start (its upwind value is the
// pre-sweep one whatever the order; the reference's Sweeper reads whatever
// its psi array holds at that moment, src/Sweeper.cc:281-300).
void angle_levels(const HostMesh &m, const double *om, int32_t angle,
                  int32_t *level, std::vector<LaggedFace> &lagged,
                  std::vector<int32_t> &indeg, std::vector<int32_t> &frontier,
                  double dep_tol)
{
    const int64_t nc = m.nc;
    // met[c][k]: 0 unmet dependency, 1 released, 2 lagged, 3 no dependency
    std::vector<uint8_t> met(4 * nc);
    indeg.assign(nc, 0);
    frontier.clear();
    for (int64_t c = 0; c < nc; c++) {
        int d = 0;
        for (int k = 0; k < 4; k++) {
            const int64_t u = m.adj[4 * c + k];
            const double dn = omega_dot_n(om, &m.normal[(4 * c + k) * 3]);
            const bool in = u >= 0 && !(dn > 0);
            const bool dep = in && (dep_tol < 0 || dn < -dep_tol);
            met[4 * c + k] = dep ? 0 : in ? 2 : 3;
            if (in && !dep)
                lagged.push_back({angle, int32_t(c), k, int32_t(u)});
            d += dep;
        }
        indeg[c] = d;
        level[c] = -1;
    }
    for (int64_t c = 0; c < nc; c++)
        if (indeg[c] == 0) {
            frontier.push_back(int32_t(c));
            level[c] = 0;
        }
    int64_t done = 0;
    int32_t lev = 0;
    std::vector<int32_t> next;
    while (done < nc) {
        while (frontier.empty()) {
            int64_t bc = -1;
            int bk = -1;
            double bw = 0.0;
            for (int64_t c = 0; c < nc; c++) {
                if (level[c] >= 0)
                    continue;
                for (int k = 0; k < 4; k++) {
                    if (met[4 * c + k] != 0)
                        continue;
                    const double w =
                        m.area[4 * c + k] *
                        std::fabs(omega_dot_n(om, &m.normal[(4 * c + k) * 3]));
                    if (bc < 0 || w < bw) {
                        bc = c;
                        bk = k;
                        bw = w;
                    }
                }
            }
            met[4 * bc + bk] = 2;
            lagged.push_back({angle, int32_t(bc), bk,
                              int32_t(m.adj[4 * bc + bk])});
            if (--indeg[bc] == 0) {
                level[bc] = lev;
                frontier.push_back(int32_t(bc));
            }
        }
        next.clear();
        for (int32_t c : frontier) {
            done++;
            for (int k = 0; k < 4; k++) {
                const int64_t u = m.adj[4 * int64_t(c) + k];
                if (u < 0 || level[u] >= 0 ||
                    !(omega_dot_n(om, &m.normal[(4 * int64_t(c) + k) * 3]) > 0))
                    continue;
                int back = 0;  // u's face towards c
                while (m.adj[4 * u + back] != c)
                    back++;
                if (met[4 * u + back] != 0)
                    continue;
                met[4 * u + back] = 1;
                if (--indeg[u] == 0) {
                    level[u] = lev + 1;
                    next.push_back(int32_t(u));
                }
            }
        }
        frontier.swap(next);
        lev++;
    }
}

static int n_workers(int64_t jobs)
{
    unsigned hw = std::thread::hardware_concurrency();
    int64_t n = hw ? hw : 4;
    n = std::min<int64_t>(n, 32);
    return int(std::max<int64_t>(1, std::min(n, jobs)));
}

int build_levels(const HostMesh &m, int64_t na, const double *omega,
                 int32_t *levels, std::vector<LaggedFace> &lagged,
                 double dep_tol)
{
    const int nw = n_workers(na);
    std::vector<std::vector<LaggedFace>> lag(nw);
    std::atomic<int64_t> next{0};
    auto work = [&](int w) {
        std::vector<int32_t> indeg, frontier;
        for (int64_t a; (a = next.fetch_add(1)) < na;)
            angle_levels(m, omega + 3 * a, int32_t(a), levels + a * m.nc,
                         lag[w], indeg, frontier, dep_tol);
    };
    std::vector<std::thread> pool;
    for (int w = 1; w < nw; w++)
        pool.emplace_back(work, w);
    work(0);
    for (auto &t : pool)
        t.join();
    lagged.clear();
    for (auto &l : lag)
        lagged.insert(lagged.end(), l.begin(), l.end());
    std::sort(lagged.begin(), lagged.end(),
              [](const LaggedFace &x, const LaggedFace &y) {
                  if (x.angle != y.angle) return x.angle < y.angle;
                  if (x.cell != y.cell) return x.cell < y.cell;
                  return x.face < y.face;
              });
    return T2_OK;
}

std::vector<int32_t> angle_batches(int64_t na, const double *omega,
                                   int batches)
{
    std::vector<int32_t> order(na);
    for (int64_t a = 0; a < na; a++)
        order[a] = int32_t(a);
    auto octant = [&](int32_t a) {
        return (omega[3 * a] > 0) + 2 * (omega[3 * a + 1] > 0) +
               4 * (omega[3 * a + 2] > 0);
    };
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
        return octant(x) < octant(y);
    });
    const int64_t nb = std::max<int64_t>(1, std::min<int64_t>(batches, na));
    std::vector<int32_t> batch(na);
    for (int64_t i = 0; i < na; i++)
        batch[order[i]] = int32_t(i * nb / na);
    return batch;
}

// Work list in step order.  Batch b occupies steps [off_b, off_b + L_b) with
// L_b the most levels of any of its angles; step off_b + l holds every
// (angle, cell) of the batch at level l, angle-major and cell-ascending.
int build_schedule(const HostMesh &m, int64_t na, const double *omega,
                   int batches, int order, Schedule &s, double dep_tol)
{
    const int64_t nc = m.nc;
    int cb = 1;
    while ((int64_t(1) << cb) < nc)
        cb++;
    int ab = 1;
    while ((int64_t(1) << ab) < na)
        ab++;
    if (cb + ab > 32)
        return fail(T2_ERR_ARG, "cells x angles exceed the 32-bit work item");
    s.cell_bits = cb;
    std::vector<int32_t> levels;
    try {
        levels.resize(size_t(na) * nc);
    } catch (...) {
        return fail(T2_ERR_NOMEM, "host memory for level sets");
    }
    s.lagged.clear();
    int rc = build_levels(m, na, omega, levels.data(), s.lagged, dep_tol);
    if (rc)
        return rc;
    const std::vector<int32_t> batch = angle_batches(na, omega, batches);
    const int nb = 1 + *std::max_element(batch.begin(), batch.end());
    std::vector<int32_t> nlev(na), blev(nb, 0);
    int32_t min_lev = INT32_MAX;
    for (int64_t a = 0; a < na; a++) {
        int32_t mx = 0;
        const int32_t *l = levels.data() + a * nc;
        for (int64_t c = 0; c < nc; c++)
            mx = std::max(mx, l[c]);
        nlev[a] = mx + 1;
        min_lev = std::min(min_lev, mx + 1);
        blev[batch[a]] = std::max(blev[batch[a]], mx + 1);
    }
    std::vector<int64_t> boff(nb + 1, 0);
    for (int b = 0; b < nb; b++)
        boff[b + 1] = boff[b] + blev[b];
    s.n_steps = boff[nb];
    s.batch_steps = boff;
    s.min_levels = min_lev;
    s.step_off.assign(s.n_steps + 1, 0);
    for (int64_t a = 0; a < na; a++) {
        const int32_t *l = levels.data() + a * nc;
        for (int64_t c = 0; c < nc; c++)
            s.step_off[boff[batch[a]] + l[c] + 1]++;
    }
    for (int64_t st = 0; st < s.n_steps; st++)
        s.step_off[st + 1] += s.step_off[st];
    std::vector<int64_t> pos(s.step_off.begin(), s.step_off.end() - 1);
    s.work.resize(size_t(na) * nc);
    std::vector<int32_t> cells(nc);
    for (int64_t a = 0; a < na; a++) {
        const int32_t *l = levels.data() + a * nc;
        for (int64_t c = 0; c < nc; c++)
            cells[c] = int32_t(c);
        if (order == T2_ORDER_FACES) {
            // stable counting sort of the cells by incoming-face pattern, so
            // that the items sharing a wavefront take the same upwind faces
            const double *om = omega + 3 * a;
            std::vector<uint8_t> mask(nc);
            std::array<int64_t, 17> start{};
            for (int64_t c = 0; c < nc; c++) {
                unsigned mk = 0;
                for (int k = 0; k < 4; k++)
                    if (m.adj[4 * c + k] >= 0 &&
                        omega_dot_n(om, &m.normal[(4 * c + k) * 3]) < 0)
                        mk |= 1u << k;
                mask[c] = uint8_t(mk);
                start[mk + 1]++;
            }
            for (int b = 0; b < 16; b++)
                start[b + 1] += start[b];
            for (int64_t c = 0; c < nc; c++)
                cells[start[mask[c]]++] = int32_t(c);
        }
        for (int64_t i = 0; i < nc; i++) {
            const int64_t c = cells[i];
            s.work[pos[boff[batch[a]] + l[c]]++] =
                uint32_t(c) | (uint32_t(a) << cb);
        }
    }
    return T2_OK;
}

// ------------------------------------------------------------ priorities
namespace {

// Runs body(a) for every angle on the host's cores.
template <typename F>
void for_each_angle(int64_t na, F body)
{
    const int nw = n_workers(na);
    std::atomic<int64_t> next{0};
    auto work = [&]() {
        for (int64_t a; (a = next.fetch_add(1)) < na;)
            body(a);
    };
    std::vector<std::thread> pool;
    for (int w = 1; w < nw; w++)
        pool.emplace_back(work);
    work();
    for (auto &t : pool)
        t.join();
}

// The sweep dependency graph of every direction as the level sets resolve
// it: (a, c) -> (a, u) when face k of c is outgoing (Omega.n > 0, the
// reference's isOutgoing, src/TychoMesh.hh:107-108), u = adj[c][k], and
// u's dependency through its face towards c is not lagged.  Stored as one
// byte per (angle, cell): bits 0-3 the downwind faces, bits 4-6 the number
// of (unlagged) upwind dependencies.
struct SweepDag {
    const HostMesh &m;
    std::vector<uint8_t> bits;  // [na][nc]

    SweepDag(const HostMesh &m_, int64_t na, const double *omega,
             const std::vector<LaggedFace> &lagged)
        : m(m_), bits(size_t(na) * m_.nc, 0)
    {
        const int64_t nc = m.nc;
        std::vector<uint8_t> lag(size_t(na) * nc, 0);
        for (const LaggedFace &f : lagged)
            lag[size_t(f.angle) * nc + f.cell] |= uint8_t(1u << f.face);
        for_each_angle(na, [&](int64_t a) {
            const double *om = omega + 3 * a;
            const uint8_t *lg = lag.data() + a * nc;
            uint8_t *b = bits.data() + a * nc;
            for (int64_t c = 0; c < nc; c++) {
                unsigned down = 0, deps = 0;
                for (int k = 0; k < 4; k++) {
                    const int64_t u = m.adj[4 * c + k];
                    if (u < 0)
                        continue;
                    if (!(omega_dot_n(om, &m.normal[(4 * c + k) * 3]) > 0)) {
                        deps += !(lg[c] >> k & 1u);
                        continue;
                    }
                    int back = 0;
                    while (back < 4 && m.adj[4 * u + back] != c)
                        back++;
                    if (back < 4 && !(lg[u] >> back & 1u) &&
                        !(omega_dot_n(om, &m.normal[(4 * u + back) * 3]) > 0))
                        down |= 1u << k;
                }
                b[c] = uint8_t(down | deps << 4);
            }
        });
    }

    int down(int64_t a, int64_t c, int64_t out[4]) const
    {
        const unsigned d = bits[size_t(a) * m.nc + c] & 15u;
        int n = 0;
        for (int k = 0; k < 4; k++)
            if (d >> k & 1u)
                out[n++] = m.adj[4 * c + k];
        return n;
    }

    int in_degree(int64_t a, int64_t c) const
    {
        return bits[size_t(a) * m.nc + c] >> 4;
    }
};

// Cells of one angle in decreasing level order (a reverse topological order
// of its dependency graph).
void reverse_level_order(const int32_t *level, int64_t nc,
                         std::vector<int32_t> &cells)
{
    int32_t top = 0;
    for (int64_t c = 0; c < nc; c++)
        top = std::max(top, level[c]);
    std::vector<int64_t> start(size_t(top) + 2, 0);
    for (int64_t c = 0; c < nc; c++)
        start[size_t(top - level[c]) + 1]++;
    for (int32_t l = 0; l <= top; l++)
        start[l + 1] += start[l];
    cells.resize(nc);
    for (int64_t c = 0; c < nc; c++)
        cells[start[size_t(top - level[c])]++] = int32_t(c);
}

}  // namespace

int sweep_priorities(const HostMesh &m, int64_t na, const double *omega,
                     const int32_t *levels,
                     const std::vector<LaggedFace> &lagged, int intra,
                     int inter, std::vector<uint64_t> &prio)
{
    if (intra < 0 || intra > 4 || inter < 0 || inter > 2)
        return fail(T2_ERR_ARG, "intraAngleP must be in [0, 4] and "
                                "interAngleP in [0, 2]");
    const int64_t nc = m.nc;
    const SweepDag dag(m, na, omega, lagged);
    std::vector<uint64_t> blev;
    try {
        prio.assign(size_t(na) * nc, 0);
        blev.assign(size_t(na) * nc, 0);
    } catch (...) {
        return fail(T2_ERR_NOMEM, "host memory for sweep priorities");
    }
    // b-level: longest path to a cell with no downwind dependant, by the
    // backward traversal of src/Priorities.cc:120-143 (one rank: no
    // partition sides)
    std::vector<uint64_t> top(na, 0);
    for_each_angle(na, [&](int64_t a) {
        std::vector<int32_t> cells;
        reverse_level_order(levels + a * nc, nc, cells);
        uint64_t *b = blev.data() + a * nc;
        for (int32_t c : cells) {
            int64_t d[4];
            const int n = dag.down(a, c, d);
            uint64_t v = 0;
            for (int i = 0; i < n; i++)
                v = std::max(v, b[d[i]] + 1);
            b[c] = v;
            top[a] = std::max(top[a], v);
        }
    });
    const uint64_t max_blevel = *std::max_element(top.begin(), top.end());
    switch (intra) {
    case 0: {  // random (src/Priorities.cc:283-291), a seeded 31-bit draw
        std::mt19937 gen(0);
        for (int64_t a = 0; a < na; a++)
            for (int64_t c = 0; c < nc; c++)
                prio[size_t(a) * nc + c] = uint64_t(gen() >> 1);
        break;
    }
    case 1:  // b-levels (src/Priorities.cc:297-305)
        prio = blev;
        break;
    default: {
        // BFDS / DFDS / DFHDS (src/Priorities.cc:170-262, 311-322): on one
        // rank no face is a partition side, so only the parent shift (0 for
        // BFDS, -1 otherwise, added in unsigned 64-bit arithmetic) acts
        const uint64_t shift = intra == 2 ? 0 : ~uint64_t(0);
        for_each_angle(na, [&](int64_t a) {
            std::vector<int32_t> cells;
            reverse_level_order(levels + a * nc, nc, cells);
            uint64_t *p = prio.data() + a * nc;
            for (int32_t c : cells) {
                int64_t d[4];
                const int n = dag.down(a, c, d);
                uint64_t v = 0;
                for (int i = 0; i < n; i++)
                    v = std::max(v, p[d[i]] + shift);
                p[c] = v;
            }
        });
        break;
    }
    }
    // inter-angle rule (src/Priorities.cc:328-375); nlevels is the largest
    // b-level (src/Priorities.cc:400-401)
    const uint64_t span = max_blevel * uint64_t(nc);
    if (inter == 1) {
        for (int64_t a = 0; a < na; a++)
            for (int64_t c = 0; c < nc; c++)
                prio[size_t(a) * nc + c] += uint64_t(a) * span;
    } else if (inter == 2) {
        std::vector<std::pair<uint64_t, uint64_t>> high(na);
        for (int64_t a = 0; a < na; a++) {
            uint64_t h = 0;
            for (int64_t c = 0; c < nc; c++)
                h = std::max(h, prio[size_t(a) * nc + c]);
            high[a] = {h, uint64_t(a)};
        }
        std::sort(high.begin(), high.end());
        uint64_t rank = 0;
        for (auto it = high.rbegin(); it != high.rend(); ++it) {
            ++rank;
            const int64_t a = int64_t(it->second);
            for (int64_t c = 0; c < nc; c++)
                prio[size_t(a) * nc + c] += (uint64_t(na) - rank) * span;
        }
    }
    return T2_OK;
}

int build_priority_schedule(const HostMesh &m, int64_t na,
                            const double *omega, int intra, int inter,
                            int64_t width, int order, Schedule &s,
                            double dep_tol)
{
    const int64_t nc = m.nc;
    int cb = 1;
    while ((int64_t(1) << cb) < nc)
        cb++;
    int ab = 1;
    while ((int64_t(1) << ab) < na)
        ab++;
    if (cb + ab > 32)
        return fail(T2_ERR_ARG, "cells x angles exceed the 32-bit work item");
    if (width < 0)
        return fail(T2_ERR_ARG, "max cells per step must be >= 0");
    s.cell_bits = cb;
    std::vector<int32_t> levels;
    try {
        levels.resize(size_t(na) * nc);
    } catch (...) {
        return fail(T2_ERR_NOMEM, "host memory for level sets");
    }
    s.lagged.clear();
    int rc = build_levels(m, na, omega, levels.data(), s.lagged, dep_tol);
    if (rc)
        return rc;
    int32_t min_lev = INT32_MAX;
    for (int64_t a = 0; a < na; a++)
        min_lev = std::min(min_lev,
                           1 + *std::max_element(levels.begin() + a * nc,
                                                 levels.begin() + (a + 1) * nc));
    s.min_levels = min_lev;
    std::vector<uint64_t> prio;
    rc = sweep_priorities(m, na, omega, levels.data(), s.lagged, intra, inter,
                          prio);
    if (rc)
        return rc;
    levels = std::vector<int32_t>();
    const SweepDag dag(m, na, omega, s.lagged);
    struct Ready {
        uint64_t p;
        int32_t a, c;
    };
    // max-heap: highest priority, then lowest angle, then lowest cell
    auto below = [](const Ready &x, const Ready &y) {
        if (x.p != y.p)
            return x.p < y.p;
        if (x.a != y.a)
            return x.a > y.a;
        return x.c > y.c;
    };
    std::vector<uint8_t> indeg;
    std::vector<Ready> heap;
    try {
        indeg.resize(size_t(na) * nc);
        s.work.clear();
        s.work.reserve(size_t(na) * nc);
    } catch (...) {
        return fail(T2_ERR_NOMEM, "host memory for the priority schedule");
    }
    for (int64_t a = 0; a < na; a++)
        for (int64_t c = 0; c < nc; c++) {
            const int d = dag.in_degree(a, c);
            indeg[size_t(a) * nc + c] = uint8_t(d);
            if (d == 0)
                heap.push_back({prio[size_t(a) * nc + c], int32_t(a),
                                int32_t(c)});
        }
    std::make_heap(heap.begin(), heap.end(), below);
    s.step_off.assign(1, 0);
    std::vector<Ready> step;
    std::vector<uint32_t> sorted;
    while (!heap.empty()) {
        step.clear();
        while (!heap.empty() && (width == 0 || int64_t(step.size()) < width)) {
            std::pop_heap(heap.begin(), heap.end(), below);
            step.push_back(heap.back());
            heap.pop_back();
        }
        const size_t first = s.work.size();
        for (const Ready &r : step)
            s.work.push_back(uint32_t(r.c) | (uint32_t(r.a) << cb));
        if (order == T2_ORDER_FACES) {
            // stable grouping by incoming-face pattern within the step
            std::array<size_t, 17> start{};
            std::vector<uint8_t> mask(step.size());
            for (size_t i = 0; i < step.size(); i++) {
                unsigned mk = 0;
                for (int k = 0; k < 4; k++)
                    if (m.adj[4 * int64_t(step[i].c) + k] >= 0 &&
                        omega_dot_n(omega + 3 * int64_t(step[i].a),
                                    &m.normal[(4 * int64_t(step[i].c) + k) *
                                              3]) < 0)
                        mk |= 1u << k;
                mask[i] = uint8_t(mk);
                start[mk + 1]++;
            }
            for (int b = 0; b < 16; b++)
                start[b + 1] += start[b];
            sorted.resize(step.size());
            for (size_t i = 0; i < step.size(); i++)
                sorted[start[mask[i]]++] = s.work[first + i];
            std::copy(sorted.begin(), sorted.end(), s.work.begin() + first);
        }
        s.step_off.push_back(int64_t(s.work.size()));
        for (const Ready &r : step) {
            int64_t d[4];
            const int n = dag.down(r.a, r.c, d);
            for (int i = 0; i < n; i++) {
                uint8_t &k = indeg[size_t(r.a) * nc + d[i]];
                if (--k == 0) {
                    heap.push_back({prio[size_t(r.a) * nc + d[i]], r.a,
                                    int32_t(d[i])});
                    std::push_heap(heap.begin(), heap.end(), below);
                }
            }
        }
    }
    if (int64_t(s.work.size()) != na * nc)
        return fail(T2_ERR_ARG, "priority schedule: dependency graph left "
                                "items unscheduled");
    s.n_steps = int64_t(s.step_off.size()) - 1;
    s.batch_steps = {0, s.n_steps};
    return T2_OK;
}

}  // namespace t2

// ---------------------------------------------------------------- C ABI
extern "C" {

const char *t2_last_error(void) { return t2::last_error_cstr(); }

int t2_abi_version(void) { return T2_ABI_VERSION; }

int t2_mesh_geometry(int64_t n_cells, int64_t n_nodes, const double *coords,
                     const int64_t *cell_nodes, const int64_t *f2c,
                     double *volume, double *area, double *normal)
{
    t2::clear_error();
    if (n_cells <= 0 || n_nodes <= 0 || !coords || !cell_nodes || !f2c ||
        !volume || !area || !normal)
        return t2::fail(T2_ERR_ARG, "t2_mesh_geometry: bad argument");
    return t2::mesh_geometry(n_cells, n_nodes, coords, cell_nodes, f2c,
                             volume, area, normal);
}

int t2_sweep_levels_tol(int64_t n_cells, const int64_t *adj_cell,
                        const double *area, const double *normal,
                        int64_t n_angles, const double *omega,
                        double dep_tol, int32_t *levels, int64_t *n_lagged)
{
    t2::clear_error();
    if (n_cells <= 0 || n_angles <= 0 || !adj_cell || !area || !normal ||
        !omega || !levels || std::isnan(dep_tol))
        return t2::fail(T2_ERR_ARG, "t2_sweep_levels: bad argument");
    t2::HostMesh m;
    m.nc = n_cells;
    m.adj.assign(adj_cell, adj_cell + 4 * n_cells);
    m.area.assign(area, area + 4 * n_cells);
    m.normal.assign(normal, normal + 12 * n_cells);
    std::vector<t2::LaggedFace> lagged;
    int rc = t2::build_levels(m, n_angles, omega, levels, lagged,
                              dep_tol < 0 ? t2::kTraverseRule : dep_tol);
    if (n_lagged)
        *n_lagged = int64_t(lagged.size());
    return rc;
}

int t2_sweep_levels(int64_t n_cells, const int64_t *adj_cell,
                    const double *area, const double *normal,
                    int64_t n_angles, const double *omega, int32_t *levels,
                    int64_t *n_lagged)
{
    return t2_sweep_levels_tol(n_cells, adj_cell, area, normal, n_angles,
                               omega, t2::kTraverseRule, levels, n_lagged);
}


int t2_sweep_priorities(int64_t n_cells, const int64_t *adj_cell,
                        const double *area, const double *normal,
                        int64_t n_angles, const double *omega,
                        int intra_angle_p, int inter_angle_p,
                        uint64_t *priorities)
{
    t2::clear_error();
    if (n_cells <= 0 || n_angles <= 0 || !adj_cell || !area || !normal ||
        !omega || !priorities)
        return t2::fail(T2_ERR_ARG, "t2_sweep_priorities: bad argument");
    t2::HostMesh m;
    m.nc = n_cells;
    m.adj.assign(adj_cell, adj_cell + 4 * n_cells);
    m.area.assign(area, area + 4 * n_cells);
    m.normal.assign(normal, normal + 12 * n_cells);
    std::vector<int32_t> levels(size_t(n_angles) * n_cells);
    std::vector<t2::LaggedFace> lagged;
    int rc = t2::build_levels(m, n_angles, omega, levels.data(), lagged);
    if (rc)
        return rc;
    std::vector<uint64_t> prio;
    rc = t2::sweep_priorities(m, n_angles, omega, levels.data(), lagged,
                              intra_angle_p, inter_angle_p, prio);
    if (rc)
        return rc;
    std::copy(prio.begin(), prio.end(), priorities);
    return T2_OK;
}

int t2_priority_schedule(int64_t n_cells, const int64_t *adj_cell,
                         const double *area, const double *normal,
                         int64_t n_angles, const double *omega,
                         int intra_angle_p, int inter_angle_p,
                         int64_t max_cells_per_step, int order,
                         int64_t *items, int64_t *step_off, int64_t *n_steps)
{
    t2::clear_error();
    if (n_cells <= 0 || n_angles <= 0 || !adj_cell || !area || !normal ||
        !omega || !items || !step_off || !n_steps)
        return t2::fail(T2_ERR_ARG, "t2_priority_schedule: bad argument");
    if (order != T2_ORDER_CELL && order != T2_ORDER_FACES)
        return t2::fail(T2_ERR_ARG, "order must be T2_ORDER_CELL or "
                                    "T2_ORDER_FACES");
    t2::HostMesh m;
    m.nc = n_cells;
    m.adj.assign(adj_cell, adj_cell + 4 * n_cells);
    m.area.assign(area, area + 4 * n_cells);
    m.normal.assign(normal, normal + 12 * n_cells);
    t2::Schedule s;
    int rc = t2::build_priority_schedule(m, n_angles, omega, intra_angle_p,
                                         inter_angle_p, max_cells_per_step,
                                         order, s);
    if (rc)
        return rc;
    const uint32_t cmask = (uint32_t(1) << s.cell_bits) - 1;
    for (size_t i = 0; i < s.work.size(); i++)
        items[i] = int64_t(s.work[i] >> s.cell_bits) * n_cells +
                   int64_t(s.work[i] & cmask);
    std::copy(s.step_off.begin(), s.step_off.end(), step_off);
    *n_steps = s.n_steps;
    return T2_OK;
}

}  // extern "C"
