// sweep_hip.x -- C++ host driver of the MI355X sweep engine.
//
// The same structure as the reference's executable (src/Main.cc): read a
// serial mesh and an input deck, build the quadrature and cross sections,
// construct a sweeper through the SweeperAbstract interface
// (src/SweeperAbstract.hh:47-70) and call solve().  SweeperHIP::sweep goes
// through the engine's C ABI (include/tycho2_sweep.h); the host-side source
// iteration below calls it exactly as the reference's
// SourceIteration::fixedPoint does (src/SourceIteration.cc:127-207), and
// "DeviceLoop true" in the deck runs the outer iteration on the device;
// "Precision FP32" and "SweepEngine Queue" select the engine's FP32 sweeps
// and its persistent ready-queue launch; "SweepSchedule Priorities" orders
// the sweep by the reference's traversal priorities (deck keys intraAngleP,
// interAngleP, maxCellsPerStep) instead of level sets.
//
// Usage: sweep_hip.x <mesh.smesh | mesh.pmesh> <input.deck>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "tycho2_sweep.h"

namespace {

// ------------------------------------------------------------------ data
// Angular flux in the reference's layout [cell][angle][vertex][group]
// (src/PsiData.hh:162); scalar flux [cell][vertex][group] (:369).
struct PsiData {
    int64_t nc, na, ng;
    std::vector<double> v;
    PsiData(int64_t c, int64_t a, int64_t g) : nc(c), na(a), ng(g),
        v(size_t(c) * a * 4 * g, 0.0) {}
    double &operator()(int64_t g, int64_t vx, int64_t a, int64_t c)
    {
        return v[((c * na + a) * 4 + vx) * ng + g];
    }
};

struct PhiData {
    int64_t nc, ng;
    std::vector<double> v;
    PhiData(int64_t c, int64_t g) : nc(c), ng(g), v(size_t(c) * 4 * g, 0.0) {}
};

void check(int rc)
{
    if (rc != T2_OK)
        throw std::runtime_error(std::string("engine: ") + t2_last_error());
}

// A solve that stopped at iterMax keeps its last iterate, as the reference's
// loops do (src/SourceIteration.cc:144-176); the engine reports it, and the
// run goes on with a warning.
void check_solve(int rc)
{
    if (rc == T2_ERR_NOCONV) {
        std::fprintf(stderr, "sweep_hip.x: warning: %s\n", t2_last_error());
        return;
    }
    check(rc);
}

// ------------------------------------------------------------------ deck
// key value pairs, '#' comments (the reference's KeyValueReader format)
struct Deck {
    std::map<std::string, std::string> kv;
    explicit Deck(const std::string &path)
    {
        std::ifstream in(path);
        if (!in)
            throw std::runtime_error("cannot open deck " + path);
        std::string line;
        while (std::getline(in, line)) {
            const size_t hash = line.find('#');
            if (hash != std::string::npos)
                line.resize(hash);
            std::istringstream ls(line);
            std::string k, v;
            if (ls >> k >> v)
                kv[k] = v;
        }
    }
    std::string str(const std::string &k, const std::string &def = "") const
    {
        auto it = kv.find(k);
        if (it == kv.end()) {
            if (def.empty())
                throw std::runtime_error("deck key not found: " + k);
            return def;
        }
        return it->second;
    }
    double num(const std::string &k) const { return std::stod(str(k)); }
    double num(const std::string &k, double def) const
    {
        return kv.count(k) ? std::stod(kv.at(k)) : def;
    }
    bool flag(const std::string &k, const std::string &def = "false") const
    {
        return str(k, def) == "true";
    }
};

// ------------------------------------------------------------------ mesh
// .smesh v2 (src/SerialMesh.hh:43-68) and the per-cell connectivity the
// engine takes (src/TychoMeshIO.cc:153-296): local face k is opposite cell
// vertex k, face vertex j is the face record's node j.
struct Mesh {
    int64_t nc = 0, nf = 0, nn = 0;
    std::vector<int64_t> cell_faces, cell_nodes, face_cells, face_nodes;
    std::vector<int64_t> material;   // per cell
    std::vector<double> coords;
    std::vector<int64_t> adj, f2c;
};

// Adjacency and face-to-cell vertex maps from the cell/face records.
void build_connectivity(Mesh &m)
{
    m.adj.assign(4 * m.nc, -1);
    m.f2c.assign(12 * m.nc, -1);
    for (int64_t c = 0; c < m.nc; c++) {
        const int64_t *cn = &m.cell_nodes[4 * c];
        for (int s = 0; s < 4; s++) {
            const int64_t f = m.cell_faces[4 * c + s];
            const int64_t *fn = &m.face_nodes[3 * f];
            int lf = -1;
            for (int v = 0; v < 4; v++)
                if (cn[v] != fn[0] && cn[v] != fn[1] && cn[v] != fn[2])
                    lf = v;
            if (lf < 0)
                throw std::runtime_error("mesh: a face is not a cell face");
            const int64_t c0 = m.face_cells[2 * f], c1 = m.face_cells[2 * f + 1];
            m.adj[4 * c + lf] = c1 < 0 ? -1 : (c0 == c ? c1 : c0);
            for (int j = 0; j < 3; j++)
                for (int v = 0; v < 4; v++)
                    if (fn[j] == cn[v])
                        m.f2c[(4 * c + lf) * 3 + j] = v;
        }
    }
}

Mesh read_smesh(std::ifstream &in, const std::string &path)
{
    uint64_t hdr[4];
    in.read(reinterpret_cast<char *>(hdr), sizeof(hdr));
    if (!in || hdr[0] != 2)
        throw std::runtime_error(path + ": not a Tycho 2 serial mesh v2");
    Mesh m;
    m.nc = int64_t(hdr[1]);
    m.nf = int64_t(hdr[2]);
    m.nn = int64_t(hdr[3]);
    std::vector<int64_t> cells(size_t(m.nc) * 9), faces(size_t(m.nf) * 5);
    m.coords.resize(size_t(m.nn) * 3);
    in.read(reinterpret_cast<char *>(cells.data()), cells.size() * 8);
    in.read(reinterpret_cast<char *>(faces.data()), faces.size() * 8);
    in.read(reinterpret_cast<char *>(m.coords.data()), m.coords.size() * 8);
    if (!in)
        throw std::runtime_error(path + ": truncated");
    m.cell_faces.resize(4 * m.nc);
    m.cell_nodes.resize(4 * m.nc);
    m.material.resize(m.nc);
    for (int64_t c = 0; c < m.nc; c++) {
        for (int i = 0; i < 4; i++) {
            m.cell_faces[4 * c + i] = cells[9 * c + i];
            m.cell_nodes[4 * c + i] = cells[9 * c + 4 + i];
        }
        m.material[c] = cells[9 * c + 8];
    }
    m.face_cells.resize(2 * m.nf);
    m.face_nodes.resize(3 * m.nf);
    for (int64_t f = 0; f < m.nf; f++) {
        m.face_cells[2 * f] = faces[5 * f];
        m.face_cells[2 * f + 1] = faces[5 * f + 1];
        for (int j = 0; j < 3; j++)
            m.face_nodes[3 * f + j] = faces[5 * f + 2 + j];
    }
    return m;
}

// .pmesh v2 (src/ParallelMesh.hh:43-81): every partition read and merged
// into one mesh numbered by the global IDs the file stores (one GPU holds
// the whole mesh; the reference gives each MPI rank one partition).
Mesh read_pmesh(std::ifstream &in, const std::string &path)
{
    const uint64_t inval = ~uint64_t(0);
    uint64_t hdr[2];
    in.read(reinterpret_cast<char *>(hdr), sizeof(hdr));
    if (!in || hdr[0] != 2)
        throw std::runtime_error(path + ": not a Tycho 2 parallel mesh v2");
    const uint64_t np = hdr[1];
    std::vector<uint64_t> off(np + 1);
    in.read(reinterpret_cast<char *>(off.data()), off.size() * 8);
    struct Part {
        std::vector<uint64_t> cell, face, node;
    };
    std::vector<Part> parts(np);
    uint64_t nc = 0, nf = 0, nn = 0;
    for (uint64_t p = 0; p < np; p++) {
        in.seekg(std::streamoff(off[p]));
        uint64_t n[3];
        in.read(reinterpret_cast<char *>(n), sizeof(n));
        Part &pt = parts[p];
        pt.cell.resize(n[0] * 10);
        pt.face.resize(n[1] * 9);
        pt.node.resize(n[2] * 4);
        in.read(reinterpret_cast<char *>(pt.cell.data()), pt.cell.size() * 8);
        in.read(reinterpret_cast<char *>(pt.face.data()), pt.face.size() * 8);
        in.read(reinterpret_cast<char *>(pt.node.data()), pt.node.size() * 8);
        if (!in)
            throw std::runtime_error(path + ": truncated");
        nc += n[0];
        for (uint64_t f = 0; f < n[1]; f++)
            nf = std::max(nf, pt.face[9 * f + 5] + 1);
        for (uint64_t v = 0; v < n[2]; v++)
            nn = std::max(nn, pt.node[4 * v + 3] + 1);
    }
    Mesh m;
    m.nc = int64_t(nc);
    m.nf = int64_t(nf);
    m.nn = int64_t(nn);
    m.cell_faces.assign(4 * nc, -1);
    m.cell_nodes.assign(4 * nc, -1);
    m.material.assign(nc, 0);
    m.face_cells.assign(2 * nf, -1);
    m.face_nodes.assign(3 * nf, -1);
    m.coords.assign(3 * nn, 0.0);
    auto gid = [&](uint64_t part, uint64_t local_cell) -> int64_t {
        if (part >= np || local_cell * 10 + 8 >= parts[part].cell.size())
            throw std::runtime_error(path + ": bad cell reference");
        return int64_t(parts[part].cell[local_cell * 10 + 8]);
    };
    for (uint64_t p = 0; p < np; p++) {
        const Part &pt = parts[p];
        const uint64_t pc = pt.cell.size() / 10, pf = pt.face.size() / 9;
        const uint64_t pn = pt.node.size() / 4;
        for (uint64_t c = 0; c < pc; c++) {
            const uint64_t *r = &pt.cell[10 * c];
            const uint64_t g = r[8];
            if (g >= nc)
                throw std::runtime_error(path + ": bad cell global ID");
            m.material[g] = int64_t(r[9]);
            for (int i = 0; i < 4; i++) {
                if (r[i] >= pf || r[4 + i] >= pn)
                    throw std::runtime_error(path + ": bad cell record");
                m.cell_faces[4 * g + i] = int64_t(pt.face[9 * r[i] + 5]);
                m.cell_nodes[4 * g + i] = int64_t(pt.node[4 * r[4 + i] + 3]);
            }
        }
        for (uint64_t f = 0; f < pf; f++) {
            const uint64_t *r = &pt.face[9 * f];
            const uint64_t g = r[5];
            for (int s = 0; s < 2; s++)
                m.face_cells[2 * g + s] =
                    r[s] == inval ? -1 : gid(r[6 + s], r[s]);
            for (int j = 0; j < 3; j++) {
                if (r[2 + j] >= pn)
                    throw std::runtime_error(path + ": bad face record");
                m.face_nodes[3 * g + j] = int64_t(pt.node[4 * r[2 + j] + 3]);
            }
        }
        for (uint64_t v = 0; v < pn; v++)
            std::memcpy(&m.coords[3 * pt.node[4 * v + 3]], &pt.node[4 * v],
                        3 * sizeof(double));
    }
    return m;
}

// A .smesh or a .pmesh file, told apart by its format name.
Mesh read_mesh(const std::string &path)
{
    std::ifstream in(path, std::ios::binary);
    if (!in)
        throw std::runtime_error("cannot open mesh " + path);
    char name[33] = {};
    in.read(name, 32);
    if (!in)
        throw std::runtime_error(path + ": not a Tycho 2 mesh");
    Mesh m;
    if (std::strncmp(name, "Tycho 2 Serial Mesh", 32) == 0)
        m = read_smesh(in, path);
    else if (std::strncmp(name, "Tycho 2 Parallel Mesh", 32) == 0)
        m = read_pmesh(in, path);
    else
        throw std::runtime_error(path + ": not a Tycho 2 mesh");
    for (int64_t f : m.cell_faces)
        if (f < 0 || f >= m.nf)
            throw std::runtime_error(path + ": bad face index");
    build_connectivity(m);
    return m;
}

// ------------------------------------------------------------------ quadrature
// Chebyshev-Legendre product set (src/Quadrature.cc:150-244) with
// Gauss-Legendre nodes and weights from Newton's method in long double,
// rounded to 16 decimals as the reference's tables are.
void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w)
{
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    for (int i = 0; i < n; i++) {
        long double r = std::cos(M_PI * (i + 0.75) / (n + 0.5));
        long double dp = 0;
        for (int it = 0; it < 100; it++) {
            long double p0 = 1, p1 = r;
            for (int k = 2; k <= n; k++) {
                long double p2 = ((2 * k - 1) * r * p1 - (k - 1) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (r * p1 - p0) / (r * r - 1);
            long double dx = p1 / dp;
            r -= dx;
            if (std::fabs(dx) < 1e-19L)
                break;
        }
        long double p0 = 1, p1 = r;
        for (int k = 2; k <= n; k++) {
            long double p2 = ((2 * k - 1) * r * p1 - (k - 1) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        dp = n * (r * p1 - p0) / (r * r - 1);
        long double wt = 2 / ((1 - r * r) * dp * dp);
        char buf[64];
        std::snprintf(buf, sizeof(buf), "%.16Lf", r);
        x[n - 1 - i] = std::strtod(buf, nullptr);
        std::snprintf(buf, sizeof(buf), "%.16Lf", wt);
        w[n - 1 - i] = std::strtod(buf, nullptr);
    }
}

void quadrature(int sn, std::vector<double> &omega, std::vector<double> &wt)
{
    if (sn < 2 || sn > 30 || sn % 2)
        throw std::runtime_error("Quadrature order not supported.");
    std::vector<double> gn, gw;
    gauss_legendre(sn, gn, gw);
    const int na = 2 * sn * sn;
    omega.assign(3 * na, 0.0);
    wt.assign(na, 0.0);
    for (int i = 0; i < sn; i++)
        for (int k = 0; k < 2 * sn; k++) {
            const double phi = (2.0 * M_PI) / (2 * sn) * (k + 0.5);
            const int idx = i * 2 * sn + k;
            omega[3 * idx] = std::sqrt(1.0 - gn[i] * gn[i]) * std::cos(phi);
            omega[3 * idx + 1] = std::sqrt(1.0 - gn[i] * gn[i]) * std::sin(phi);
            omega[3 * idx + 2] = gn[i];
            wt[idx] = gw[i] * (2.0 * M_PI) / (2 * sn);
        }
}

// ------------------------------------------------------------------ sweeper
class SweeperAbstract {
public:
    virtual ~SweeperAbstract() = default;
    virtual void sweep(PsiData &psi, const PsiData &source,
                       bool zeroPsiBound = false) = 0;
    virtual void solve() = 0;
};

struct Problem {
    Mesh mesh;
    int sn = 0;
    int64_t na = 0, ng = 0;
    std::vector<double> omega, wt, sigma_t, sigma_s;
};

// The reference's manufactured hat source on [0,100]^3 (src/Problem.cc:47-82)
void hat_source(const Problem &p, PsiData &src)
{
    for (int64_t c = 0; c < p.mesh.nc; c++)
        for (int64_t a = 0; a < p.na; a++)
            for (int v = 0; v < 4; v++) {
                const double *x = &p.mesh.coords[3 * p.mesh.cell_nodes[4 * c + v]];
                const double dx = x[0] - 50.0, dy = x[1] - 50.0, dz = x[2] - 50.0;
                const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
                double val = 0.0;
                if (r <= 30.0)
                    val = -dx / (30.0 * r) * p.omega[3 * a] -
                          dy / (30.0 * r) * p.omega[3 * a + 1] -
                          dz / (30.0 * r) * p.omega[3 * a + 2] +
                          (p.sigma_t[c * p.ng] - p.sigma_s[c * p.ng]) *
                              (1.0 - r / 30.0);
                for (int64_t g = 0; g < p.ng; g++)
                    src(g, v, a, c) = val;
            }
}

void psi_to_phi(const Problem &p, const PsiData &psi, PhiData &phi)
{
    std::fill(phi.v.begin(), phi.v.end(), 0.0);
    for (int64_t c = 0; c < p.mesh.nc; c++)
        for (int64_t a = 0; a < p.na; a++)
            for (int64_t i = 0; i < 4 * p.ng; i++)
                phi.v[c * 4 * p.ng + i] +=
                    psi.v[(c * p.na + a) * 4 * p.ng + i] * p.wt[a];
}

class SweeperHIP : public SweeperAbstract {
public:
    SweeperHIP(const Problem &p, const Deck &deck) : c_p(p), c_deck(deck)
    {
        const int prec = deck.str("Precision", "FP64") == "FP32" ? T2_FP32
                                                                 : T2_FP64;
        if (prec == T2_FP32 && !deck.flag("DeviceLoop"))
            throw std::runtime_error(
                "Precision FP32 runs the refined solve on the device; set "
                "DeviceLoop true");
        // "SweepEngine Queue" selects the persistent ready-queue launch
        // (default: per-step launches replayed from a hipGraph)
        const std::string engine = deck.str("SweepEngine", "Levels");
        if (engine != "Levels" && engine != "Queue")
            throw std::runtime_error("SweepEngine must be Levels or Queue");
        // "SweepSchedule Priorities": the reference's traversal priorities
        // (src/Priorities.cc:396-435) with maxCellsPerStep ready pairs per
        // step (src/GraphTraverser.cc:518-536) instead of level sets
        const std::string sched = deck.str("SweepSchedule", "Levels");
        if (sched != "Levels" && sched != "Priorities")
            throw std::runtime_error(
                "SweepSchedule must be Levels or Priorities");
        // GaussElim as the reference reads it (src/Main.cc:167-176)
        const std::string ge = deck.str("GaussElim", "NoPivot");
        int gauss = T2_GAUSS_NOPIVOT;
        if (ge == "Original")
            gauss = T2_GAUSS_ORIGINAL;
        else if (ge == "CramerGlu")
            gauss = T2_GAUSS_CRAMER_GLU;
        else if (ge == "CramerIntel")
            gauss = T2_GAUSS_CRAMER_INTEL;
        else if (ge != "NoPivot")
            throw std::runtime_error("GaussElim not recognized: " + ge);
        check(t2_sweeper_create(&c_engine, 0, p.mesh.nc, p.mesh.nn,
                                p.mesh.coords.data(), p.mesh.cell_nodes.data(),
                                p.mesh.adj.data(), p.mesh.f2c.data(), p.na,
                                p.omega.data(), p.wt.data(), p.ng,
                                p.sigma_t.data(), p.sigma_s.data(), prec,
                                gauss));
        check(t2_set_engine(c_engine,
                            engine == "Queue" ? T2_ENGINE_QUEUE
                                              : T2_ENGINE_LEVELS,
                            0, 0));
        const std::string type = deck.str("SweepType");
        if (sched == "Priorities" || type != "HIP")
            check(t2_set_priority(c_engine, int(deck.num("intraAngleP")),
                                  int(deck.num("interAngleP")),
                                  int64_t(deck.num("maxCellsPerStep"))));
        if (type == "OriginalTycho1" || type == "OriginalTycho2")
            check(t2_set_dependency_tolerance(c_engine, 1e-10));
    }
    ~SweeperHIP() override { t2_sweeper_destroy(c_engine); }

    void sweep(PsiData &psi, const PsiData &source, bool zeroPsiBound) override
    {
        check(t2_sweep_host(c_engine, psi.v.data(), source.v.data(),
                            zeroPsiBound ? 1 : 0));
    }

    void solve() override;

    PsiData psi{0, 0, 0};
    int64_t iterations = 0;

private:
    void fixed_point(PsiData &psi, const PsiData &source);

    const Problem &c_p;
    const Deck &c_deck;
    t2_sweeper *c_engine = nullptr;
};

// Source iteration through sweep() (src/SourceIteration.cc:127-207):
// total source = q + sigma_s/(4 pi) phi, sweep, phi = D psi, stop on the
// L1-relative phi change.
void SweeperHIP::fixed_point(PsiData &psi, const PsiData &source)
{
    const Problem &p = c_p;
    PsiData total(p.mesh.nc, p.na, p.ng);
    PhiData phi_new(p.mesh.nc, p.ng), phi_old(p.mesh.nc, p.ng);
    psi_to_phi(p, psi, phi_new);
    const int64_t iter_max = int64_t(c_deck.num("iterMax"));
    const double err_max = c_deck.num("errMax");
    double err = 1.0;
    iterations = 0;
    while (iterations < iter_max && err > err_max) {
        phi_old.v = phi_new.v;
        for (int64_t c = 0; c < p.mesh.nc; c++)
            for (int64_t a = 0; a < p.na; a++)
                for (int v = 0; v < 4; v++)
                    for (int64_t g = 0; g < p.ng; g++)
                        total(g, v, a, c) =
                            source.v[((c * p.na + a) * 4 + v) * p.ng + g] +
                            p.sigma_s[c * p.ng + g] / (4.0 * M_PI) *
                                phi_old.v[(c * 4 + v) * p.ng + g];
        sweep(psi, total, false);
        psi_to_phi(p, psi, phi_new);
        double num = 0.0, den = 0.0;
        for (size_t i = 0; i < phi_new.v.size(); i++) {
            num += std::fabs(phi_new.v[i] - phi_old.v[i]);
            den += std::fabs(phi_new.v[i]);
        }
        err = num / den;
        std::printf("   iteration: %lld   error: %e\n", (long long)iterations,
                    err);
        iterations++;
    }
}

void SweeperHIP::solve()
{
    const Problem &p = c_p;
    PsiData source(p.mesh.nc, p.na, p.ng);
    hat_source(p, source);
    psi = PsiData(p.mesh.nc, p.na, p.ng);
    if (!c_deck.flag("DeviceLoop")) {
        if (!c_deck.flag("SourceIteration", "true"))
            throw std::runtime_error("the host loop runs source iteration; "
                                     "set DeviceLoop true for Krylov");
        fixed_point(psi, source);
        return;
    }
    check(t2_set_source_angular(c_engine, source.v.data()));
    const int64_t iter_max = int64_t(c_deck.num("iterMax"));
    if (c_deck.str("Precision", "FP64") == "FP32") {
        // FP32 sweeps inside FP64 defect correction: GMRES stages to the
        // deck's RefineTolSweep / RefineTolCorrection (defaults 1e-6 / 1e-5,
        // phi and psi to ~1e-12 of the FP64 solution); psi comes back as
        // psi32 plus its correction, in FP64
        t2_refine_info info{};
        check_solve(t2_solve_refined(
            c_engine, iter_max, c_deck.num("RefineTolSweep", 1e-6),
            c_deck.num("RefineTolCorrection", 1e-5), 30, &info, nullptr));
        iterations = info.iters_sweep + info.iters_correction;
        std::printf("   refined solve: %lld + %lld GMRES iterations, "
                    "%lld FP32 sweeps, correction %e\n",
                    (long long)info.iters_sweep,
                    (long long)info.iters_correction, (long long)info.sweeps,
                    info.correction_rel);
    } else if (c_deck.flag("SourceIteration", "true")) {
        check_solve(t2_solve_source_iteration(c_engine, iter_max,
                                              c_deck.num("errMax"),
                                              &iterations, nullptr));
    } else {
        check_solve(t2_solve_krylov(c_engine, iter_max, c_deck.num("errMax"),
                                    30, &iterations, nullptr));
    }
    check(t2_get_psi(c_engine, psi.v.data()));
}

// End-of-run checks the reference's executable prints (src/Main.cc:340-345).
// L2 relative error of the cell-average psi against the hat solution at the
// cell centroid of [0,100]^3 (src/Problem.cc:91-134).
double hat_l2_error(const Problem &p, PsiData &psi)
{
    double diff = 0.0, norm = 0.0;
    for (int64_t c = 0; c < p.mesh.nc; c++)
        for (int64_t a = 0; a < p.na; a++)
            for (int64_t g = 0; g < p.ng; g++) {
                double x = 0.0, y = 0.0, z = 0.0, val = 0.0;
                for (int v = 0; v < 4; v++) {
                    const double *n = &p.mesh.coords[3 * p.mesh.cell_nodes[4 * c + v]];
                    x += n[0] - 50.0;
                    y += n[1] - 50.0;
                    z += n[2] - 50.0;
                    val += psi(g, v, a, c);
                }
                x /= 4.0;
                y /= 4.0;
                z /= 4.0;
                val /= 4.0;
                const double r = std::sqrt(x * x + y * y + z * z);
                const double act = r <= 30.0 ? 1.0 - r / 30.0 : 0.0;
                norm += val * val;
                diff += (val - act) * (val - act);
            }
    return std::sqrt(diff / norm);
}

// Largest |psi_g - psi_0| (g > 0) relative to the largest |psi_0|
// (src/Util.cc:55-79).
double diff_between_groups(PsiData &psi)
{
    double max_diff = 0.0, max_entry = 0.0;
    for (int64_t c = 0; c < psi.nc; c++)
        for (int64_t a = 0; a < psi.na; a++)
            for (int v = 0; v < 4; v++) {
                const double p0 = psi(0, v, a, c);
                max_entry = std::max(max_entry, std::fabs(p0));
                for (int64_t g = 1; g < psi.ng; g++)
                    max_diff = std::max(max_diff, std::fabs(p0 - psi(g, v, a, c)));
            }
    return max_diff / max_entry;
}

// "Tycho 2 Psi Output" v1 (src/PsiData.cc:46-106)
void write_psi(const std::string &path, const PsiData &psi)
{
    std::ofstream out(path, std::ios::binary);
    char name[32] = {0};
    std::strncpy(name, "Tycho 2 Psi Output", sizeof(name));
    uint64_t hdr[4] = {1, uint64_t(psi.nc), uint64_t(psi.na), uint64_t(psi.ng)};
    out.write(name, 32);
    out.write(reinterpret_cast<const char *>(hdr), sizeof(hdr));
    out.write(reinterpret_cast<const char *>(psi.v.data()), psi.v.size() * 8);
    if (!out)
        throw std::runtime_error("cannot write " + path);
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc < 3) {
        std::printf("Usage: ./sweep_hip.x <serial or parallel mesh> <input deck>\n");
        return 0;
    }
    try {
        Deck deck(argv[2]);
        // SweepType (src/Main.cc:143-164): HIP is the engine's level sets;
        // the reference's single-rank sweepers map onto the engine with
        // their schedules: TraverseGraph follows the deck's priorities,
        // OriginalTycho1/2 follow them with the precomputed schedule's
        // dependency tolerance (src/SweepSchedule.cc:57).  The multi-rank
        // domain-decomposition sweepers (PBJ, Schur) are not built here.
        const std::string type = deck.str("SweepType");
        if (type != "HIP" && type != "TraverseGraph" &&
            type != "OriginalTycho1" && type != "OriginalTycho2")
            throw std::runtime_error(
                "Sweep type not recognized: " + type +
                " (this executable runs SweepType HIP, TraverseGraph, "
                "OriginalTycho1 and OriginalTycho2)");
        Problem p;
        p.mesh = read_mesh(argv[1]);
        p.sn = int(deck.num("snOrder"));
        p.ng = int64_t(deck.num("nGroups"));
        quadrature(p.sn, p.omega, p.wt);
        p.na = int64_t(p.wt.size());
        // per-material cross sections (src/Problem.cc:149-172)
        const double st[2] = {deck.num("sigmaT1"), deck.num("sigmaT2")};
        const double ss[2] = {deck.num("sigmaS1"), deck.num("sigmaS2")};
        const std::vector<int64_t> &material = p.mesh.material;
        p.sigma_t.resize(p.mesh.nc * p.ng);
        p.sigma_s.resize(p.mesh.nc * p.ng);
        for (int64_t c = 0; c < p.mesh.nc; c++) {
            if (material[c] != 1 && material[c] != 2)
                throw std::runtime_error("Material index out of bounds.");
            for (int64_t g = 0; g < p.ng; g++) {
                p.sigma_t[c * p.ng + g] = st[material[c] - 1];
                p.sigma_s[c * p.ng + g] = ss[material[c] - 1];
            }
        }
        SweeperHIP sweeper(p, deck);
        SweeperAbstract &abstract = sweeper;
        const auto t0 = std::chrono::steady_clock::now();
        abstract.solve();
        const std::chrono::duration<double> dt =
            std::chrono::steady_clock::now() - t0;
        std::printf("iterations: %lld\n", (long long)sweeper.iterations);
        std::printf("Total time: %.2f\n", dt.count());
        std::printf("L2 Relative Error: %e\n", hat_l2_error(p, sweeper.psi));
        std::printf("Diff between groups: %e\n",
                    diff_between_groups(sweeper.psi));
        if (deck.flag("OutputFile"))
            write_psi(deck.str("OutputFilename"), sweeper.psi);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "sweep_hip.x: %s\n", e.what());
        return 1;
    }
    return 0;
}
