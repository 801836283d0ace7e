// Host-side set-up of the sweep engine: cell geometry and per-angle sweep
// schedules.  Pure C++ (no device calls); built with -ffp-contract=off so the
// floating-point results match the reference's TychoMesh bit for bit.
#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
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
            for (int j = 0; j < 3; j++) {
                int64_t cv = f2c[(4 * c + k) * 3 + j];
                if (cv < 0 || cv > 3 || cv == k)
                    return fail(T2_ERR_ARG, "bad face-to-cell vertex map");
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
void angle_levels(const HostMesh &m, const double *om, int32_t angle,
                  int32_t *level, std::vector<LaggedFace> &lagged,
                  std::vector<int32_t> &indeg, std::vector<int32_t> &frontier)
{
    const int64_t nc = m.nc;
    // met[c][k]: 0 unmet dependency, 1 released, 2 lagged, 3 no dependency
    std::vector<uint8_t> met(4 * nc);
    indeg.assign(nc, 0);
    frontier.clear();
    for (int64_t c = 0; c < nc; c++) {
        int d = 0;
        for (int k = 0; k < 4; k++) {
            const bool dep = m.adj[4 * c + k] >= 0 &&
                             !(omega_dot_n(om, &m.normal[(4 * c + k) * 3]) > 0);
            met[4 * c + k] = dep ? 0 : 3;
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
                 int32_t *levels, std::vector<LaggedFace> &lagged)
{
    const int nw = n_workers(na);
    std::vector<std::vector<LaggedFace>> lag(nw);
    std::atomic<int64_t> next{0};
    auto work = [&](int w) {
        std::vector<int32_t> indeg, frontier;
        for (int64_t a; (a = next.fetch_add(1)) < na;)
            angle_levels(m, omega + 3 * a, int32_t(a), levels + a * m.nc,
                         lag[w], indeg, frontier);
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
                   int batches, int order, Schedule &s)
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
    int rc = build_levels(m, na, omega, levels.data(), s.lagged);
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

int t2_sweep_levels(int64_t n_cells, const int64_t *adj_cell,
                    const double *area, const double *normal,
                    int64_t n_angles, const double *omega, int32_t *levels,
                    int64_t *n_lagged)
{
    t2::clear_error();
    if (n_cells <= 0 || n_angles <= 0 || !adj_cell || !area || !normal ||
        !omega || !levels)
        return t2::fail(T2_ERR_ARG, "t2_sweep_levels: bad argument");
    t2::HostMesh m;
    m.nc = n_cells;
    m.adj.assign(adj_cell, adj_cell + 4 * n_cells);
    m.area.assign(area, area + 4 * n_cells);
    m.normal.assign(normal, normal + 12 * n_cells);
    std::vector<t2::LaggedFace> lagged;
    int rc = t2::build_levels(m, n_angles, omega, levels, lagged);
    if (n_lagged)
        *n_lagged = int64_t(lagged.size());
    return rc;
}

}  // extern "C"
