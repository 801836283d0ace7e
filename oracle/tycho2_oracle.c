/*
 * Tycho2 CPU oracle -- TEST INFRASTRUCTURE ONLY (see tycho2_oracle.h).
 *
 * Restates, in plain C, the reference's one-rank FP64 path:
 *   mesh connectivity/geometry  src/TychoMeshIO.cc:153-296, src/TychoMesh.cc:52-260
 *   quadrature                  src/Quadrature.cc:150-244
 *   cell solve                  src/Transport.cc:54-410, 687-733
 *   upwind gather               src/Transport.cc:740-776
 *   traversal                   src/GraphTraverser.cc:416-459, 467-644
 *   moments / SI / Krylov       src/Util.cc:85-159, src/SourceIteration.cc:76-285
 *   hat source                  src/Problem.cc:47-82
 * The arithmetic is written so that each floating-point operation happens in
 * the reference's order (no FMA contraction: build with -ffp-contract=off),
 * which makes the results bitwise reproducible against it.
 */
#include "tycho2_oracle.h"
#include "gl_tables.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------ mesh */
static void cross3(const double a[3], const double b[3], double c[3])
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

static void sub3(const double *p, const double *q, double r[3])
{
    r[0] = p[0] - q[0];
    r[1] = p[1] - q[1];
    r[2] = p[2] - q[2];
}

/* TychoMesh.cc:96-117 */
static double tet_volume(const double *p[4])
{
    double a[3], b[3], c[3], x[3];
    sub3(p[1], p[0], a);
    sub3(p[2], p[0], b);
    sub3(p[3], p[0], c);
    cross3(a, b, x);
    return 1.0 / 6.0 * fabs(x[0] * c[0] + x[1] * c[1] + x[2] * c[2]);
}

/* TychoMesh.cc:68-86 */
static double tri_area(const double *p[3])
{
    double a[3], b[3], x[3];
    sub3(p[1], p[0], a);
    sub3(p[2], p[0], b);
    cross3(a, b, x);
    return 0.5 * sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
}

/* TychoMesh.cc:125-187: unit normal of the face, flipped to point away from
 * the cell vertex that is not on the face (found by coordinate equality). */
static void outward_normal(const double *fp[3], const double *cp[4],
                           double n[3])
{
    double a[3], b[3], x[3];
    sub3(fp[1], fp[0], a);
    sub3(fp[2], fp[0], b);
    cross3(a, b, x);
    double len = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    n[0] = x[0] / len;
    n[1] = x[1] / len;
    n[2] = x[2] / len;
    int off = 0;
    for (off = 0; off < 4; off++) {
        int on = 0;
        for (int j = 0; j < 3; j++)
            if (cp[off][0] == fp[j][0] && cp[off][1] == fp[j][1] &&
                cp[off][2] == fp[j][2])
                on = 1;
        if (!on)
            break;
    }
    double d = n[0] * (cp[off][0] - fp[0][0]) + n[1] * (cp[off][1] - fp[0][1]) +
               n[2] * (cp[off][2] - fp[0][2]);
    if (d > 0) {
        n[0] = -n[0];
        n[1] = -n[1];
        n[2] = -n[2];
    }
}

int or_mesh_build(int64_t nc, int64_t nf, int64_t nn,
                  const int64_t *cell_faces, const int64_t *cell_nodes,
                  const int64_t *face_cells, const int64_t *face_nodes,
                  const double *coords,
                  int64_t *adj, int64_t *f2c, int64_t *c2f, int64_t *nbrv,
                  double *volume, double *area, double *normal)
{
    for (int64_t c = 0; c < nc; c++) {
        const int64_t *cn = cell_nodes + 4 * c;
        for (int s = 0; s < 4; s++) {
            int64_t f = cell_faces[4 * c + s];
            if (f < 0 || f >= nf)
                return -1;
            const int64_t *fn = face_nodes + 3 * f;
            /* local face = index of the cell vertex not on the face
             * (TychoMeshIO.cc:129-145) */
            int lf = -1;
            for (int v = 0; v < 4; v++)
                if (cn[v] != fn[0] && cn[v] != fn[1] && cn[v] != fn[2]) {
                    lf = v;
                    break;
                }
            if (lf < 0)
                return -2;
            int64_t c0 = face_cells[2 * f], c1 = face_cells[2 * f + 1];
            adj[4 * c + lf] = (c1 < 0) ? -1 : (c0 == c ? c1 : c0);
            /* face vertex j <-> face node j (TychoMeshIO.cc:275-296) */
            for (int j = 0; j < 3; j++) {
                int cv = -1;
                for (int v = 0; v < 4; v++)
                    if (fn[j] == cn[v])
                        cv = v;
                if (cv < 0)
                    return -3;
                f2c[(4 * c + lf) * 3 + j] = cv;
            }
        }
        for (int k = 0; k < 4; k++) {
            for (int v = 0; v < 4; v++)
                c2f[(4 * c + k) * 4 + v] = -1;
            for (int j = 0; j < 3; j++)
                c2f[(4 * c + k) * 4 + f2c[(4 * c + k) * 3 + j]] = j;
        }
    }
    for (int64_t c = 0; c < nc; c++) {
        const int64_t *cn = cell_nodes + 4 * c;
        const double *cp[4];
        for (int v = 0; v < 4; v++) {
            if (cn[v] < 0 || cn[v] >= nn)
                return -4;
            cp[v] = coords + 3 * cn[v];
        }
        volume[c] = tet_volume(cp);
        for (int k = 0; k < 4; k++) {
            const double *fp[3];
            for (int j = 0; j < 3; j++)
                fp[j] = coords + 3 * cn[f2c[(4 * c + k) * 3 + j]];
            area[4 * c + k] = tri_area(fp);
            outward_normal(fp, cp, normal + (4 * c + k) * 3);
            /* neighbour vertex sharing the node (TychoMesh.cc:245-259) */
            int64_t nb = adj[4 * c + k];
            for (int j = 0; j < 3; j++) {
                int64_t *dst = nbrv + (4 * c + k) * 3 + j;
                *dst = -1;
                if (nb < 0)
                    continue;
                int64_t node = cn[f2c[(4 * c + k) * 3 + j]];
                for (int v = 0; v < 4; v++)
                    if (cell_nodes[4 * nb + v] == node)
                        *dst = v;
                if (*dst < 0)
                    return -5;
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------ quadrature */
/* Quadrature.cc:226-243: index i*2N+k, polar cosine = GL node i, azimuth
 * (2 pi)/(2N) (k + 1/2), weight gw_i (2 pi)/(2N). */
int or_quadrature(int sn, double *xi, double *eta, double *mu, double *wt)
{
    const double *gn = or_gl_nodes(sn), *gw = or_gl_weights(sn);
    if (!gn)
        return -1;
    for (int i = 0; i < sn; i++)
        for (int k = 0; k < 2 * sn; k++) {
            double phi = (2.0 * M_PI) / (2 * sn) * (k + 0.5);
            int idx = i * 2 * sn + k;
            xi[idx] = sqrt(1.0 - gn[i] * gn[i]) * cos(phi);
            eta[idx] = sqrt(1.0 - gn[i] * gn[i]) * sin(phi);
            mu[idx] = gn[i];
            wt[idx] = gw[i] * (2.0 * M_PI) / (2 * sn);
        }
    return 0;
}

/* ------------------------------------------------------------ cell solve */
/* Row-operation elimination without pivoting (Transport.cc:338-410):
 * rows 0..2 are scaled by the reciprocal of their pivot, the last pivot
 * divides. */
static void solve_nopivot(double m[4][4], double b[4])
{
    for (int r = 0; r < 3; r++) {
        double inv = 1.0 / m[r][r];
        m[r][r] = 1.0;
        for (int c = r + 1; c < 4; c++)
            m[r][c] = m[r][c] * inv;
        b[r] = b[r] * inv;
        for (int row = r + 1; row < 4; row++) {
            double t = m[row][r];
            m[row][r] = 0.0;
            for (int c = r + 1; c < 4; c++)
                m[row][c] = m[row][c] - m[r][c] * t;
            b[row] = b[row] - b[r] * t;
        }
    }
    b[3] = b[3] / m[3][3];
    for (int r = 2; r >= 0; r--)
        for (int c = 3; c > r; c--)
            b[r] = b[r] - m[r][c] * b[c];
}

/* Left-looking LU with partial pivoting (Transport.cc:284-334). */
static void solve_pivot(double m[4][4], double b[4])
{
    const int n = 4;
    for (int col = 0; col < n - 1; col++) {
        int best = col;
        double big = fabs(m[col][col]);
        for (int row = col + 1; row < n; row++) {
            double t = fabs(m[row][col]);
            if (t > big) {
                best = row;
                big = t;
            }
        }
        if (best != col) {
            for (int c = 0; c < n; c++) {
                double t = m[best][c];
                m[best][c] = m[col][c];
                m[col][c] = t;
            }
            double t = b[best];
            b[best] = b[col];
            b[col] = t;
        }
        m[col][col] = 1. / m[col][col];
        for (int row = col + 1; row < n; row++)
            m[row][col] *= m[col][col];
        for (int k = 0; k <= col; k++)
            for (int row = k + 1; row < n; row++)
                m[row][col + 1] -= m[row][k] * m[k][col + 1];
    }
    m[n - 1][n - 1] = 1. / m[n - 1][n - 1];
    for (int col = 0; col < n - 1; col++)
        for (int row = col + 1; row < n; row++)
            b[row] -= m[row][col] * b[col];
    for (int col = n - 1; col >= 0; col--) {
        b[col] *= m[col][col];
        for (int row = col - 1; row >= 0; row--)
            b[row] -= m[row][col] * b[col];
    }
}

/* Cramer's rule, x = adj(A) b / det(A), for the reference's CramerGlu and
 * CramerIntel options (src/Transport.cc:414-673), which form the adjugate
 * by two different cofactor expansions.  Here the adjugate comes from the
 * Laplace expansion over complementary 2x2 minors of rows {0,1} and {2,3}
 * (D. Eberly, "The Laplace Expansion Theorem: Computing the Determinants and
 * Inverses of Matrices", Geometric Tools, 2008); all three agree up to
 * rounding. */
static void solve_cramer(double a[4][4], double b[4])
{
    const double s0 = a[0][0] * a[1][1] - a[1][0] * a[0][1];
    const double s1 = a[0][0] * a[1][2] - a[1][0] * a[0][2];
    const double s2 = a[0][0] * a[1][3] - a[1][0] * a[0][3];
    const double s3 = a[0][1] * a[1][2] - a[1][1] * a[0][2];
    const double s4 = a[0][1] * a[1][3] - a[1][1] * a[0][3];
    const double s5 = a[0][2] * a[1][3] - a[1][2] * a[0][3];
    const double c5 = a[2][2] * a[3][3] - a[3][2] * a[2][3];
    const double c4 = a[2][1] * a[3][3] - a[3][1] * a[2][3];
    const double c3 = a[2][1] * a[3][2] - a[3][1] * a[2][2];
    const double c2 = a[2][0] * a[3][3] - a[3][0] * a[2][3];
    const double c1 = a[2][0] * a[3][2] - a[3][0] * a[2][2];
    const double c0 = a[2][0] * a[3][1] - a[3][0] * a[2][1];
    const double det = s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 +
                       s5 * c0;
    double adj[4][4];
    adj[0][0] = a[1][1] * c5 - a[1][2] * c4 + a[1][3] * c3;
    adj[0][1] = -a[0][1] * c5 + a[0][2] * c4 - a[0][3] * c3;
    adj[0][2] = a[3][1] * s5 - a[3][2] * s4 + a[3][3] * s3;
    adj[0][3] = -a[2][1] * s5 + a[2][2] * s4 - a[2][3] * s3;
    adj[1][0] = -a[1][0] * c5 + a[1][2] * c2 - a[1][3] * c1;
    adj[1][1] = a[0][0] * c5 - a[0][2] * c2 + a[0][3] * c1;
    adj[1][2] = -a[3][0] * s5 + a[3][2] * s2 - a[3][3] * s1;
    adj[1][3] = a[2][0] * s5 - a[2][2] * s2 + a[2][3] * s1;
    adj[2][0] = a[1][0] * c4 - a[1][1] * c2 + a[1][3] * c0;
    adj[2][1] = -a[0][0] * c4 + a[0][1] * c2 - a[0][3] * c0;
    adj[2][2] = a[3][0] * s4 - a[3][1] * s2 + a[3][3] * s0;
    adj[2][3] = -a[2][0] * s4 + a[2][1] * s2 - a[2][3] * s0;
    adj[3][0] = -a[1][0] * c3 + a[1][1] * c1 - a[1][2] * c0;
    adj[3][1] = a[0][0] * c3 - a[0][1] * c1 + a[0][2] * c0;
    adj[3][2] = -a[3][0] * s3 + a[3][1] * s1 - a[3][2] * s0;
    adj[3][3] = a[2][0] * s3 - a[2][1] * s1 + a[2][2] * s0;
    const double inv = 1.0 / det;
    double x[4];
    for (int i = 0; i < 4; i++)
        x[i] = (adj[i][0] * b[0] + adj[i][1] * b[1] + adj[i][2] * b[2] +
                adj[i][3] * b[3]) * inv;
    for (int i = 0; i < 4; i++)
        b[i] = x[i];
}

/* Linear-DG system of one tet (Transport.cc:54-271, 687-733).  With
 * a_k = abar_k / 12 and s = sigma_t V / 20:
 *   M[i][j] = a_i + (1 + d_ij) s + sum_{k: abar_k > 0, k != i,j} (1 + d_ij) a_k
 *   b[i]    = V/20 (q_i + sum_j q_j)
 *             - sum_{k: abar_k < 0, k != i} sum_{j != k} (1 + d_ij) a_k psi_k,j
 * The additions happen in the reference's order (faces ascending, then
 * columns ascending). */
static void cell_system(double volume, const double abar[4], double sigma_t,
                        const double q[4], const double up[4][4],
                        double m[4][4], double b[4])
{
    double s = sigma_t * volume / 20.0;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            m[i][j] = abar[i] / 12.0 + (i == j ? 2.0 * s : s);
    for (int k = 0; k < 4; k++) {
        if (!(abar[k] > 0))
            continue;
        double a = abar[k] / 12.0;
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++)
                if (i != k && j != k)
                    m[i][j] += (i == j ? 2.0 * a : a);
    }
    double v20 = volume / 20.0;
    for (int i = 0; i < 4; i++) {
        double sum = 0.0;
        for (int j = 0; j < 4; j++)
            sum += (i == j ? 2.0 * q[j] : q[j]);
        b[i] = v20 * sum;
    }
    for (int k = 0; k < 4; k++) {
        if (!(abar[k] < 0))
            continue;
        double a = abar[k] / 12.0;
        for (int i = 0; i < 4; i++) {
            if (i == k)
                continue;
            for (int j = 0; j < 4; j++)
                if (j != k)
                    b[i] -= (i == j ? 2.0 * a : a) * up[k][j];
        }
    }
}

void or_cell_solve(int method, double volume, const double abar[4],
                   double sigma_t, const double q[4],
                   const double upwind[4][4], double out[4])
{
    double m[4][4], b[4];
    cell_system(volume, abar, sigma_t, q, upwind, m, b);
    if (method == OR_GAUSS_ORIGINAL)
        solve_pivot(m, b);
    else if (method == OR_GAUSS_CRAMER_GLU || method == OR_GAUSS_CRAMER_INTEL)
        solve_cramer(m, b);
    else
        solve_nopivot(m, b);
    for (int i = 0; i < 4; i++)
        out[i] = b[i];
}

/* ----------------------------------------------------------------- sweep */
static inline double omega_dot_n(const double *om, const double *n)
{
    return om[0] * n[0] + om[1] * n[1] + om[2] * n[2];
}

/* Lagged faces of one angle: (cell, face) and the upwind values the face
 * reads, psi[upwind][.][a][.] as held before the sweep. */
typedef struct {
    int64_t cell;
    int face;
    double *vals; /* [4][ng], entry j = k unused */
} or_lag;

/* Update of one (cell, angle): SweepData::update (SweepData.hh:139-174) with
 * the upwind gather of Transport::populateLocalPsiBound (Transport.cc:740-776)
 * folded in.  Face k's value at cell vertex j is the neighbour's psi at
 * nbrv[c][k][c2f[c][k][j]]; exterior faces give zero (vacuum); lagged faces
 * read their snapshot. */
static void update_cell(const or_mesh *m, int64_t c, int64_t a, int64_t na,
                        const double *om, int64_t ng, const double *sigma_t,
                        const double *source, double *psi, int method,
                        const or_lag *lag, int64_t n_lag)
{
    double abar[4], odn[4];
    for (int k = 0; k < 4; k++) {
        odn[k] = omega_dot_n(om, m->normal + (4 * c + k) * 3);
        abar[k] = m->area[4 * c + k] * odn[k];
    }
    const double vol = m->volume[c];
    for (int64_t g = 0; g < ng; g++) {
        double q[4], up[4][4], out[4];
        for (int v = 0; v < 4; v++)
            q[v] = source[((c * na + a) * 4 + v) * ng + g];
        for (int k = 0; k < 4; k++) {
            for (int j = 0; j < 4; j++)
                up[k][j] = 0.0;
            int64_t nb = m->adj[4 * c + k];
            if (odn[k] > 0 || nb < 0)
                continue;
            const or_lag *l = NULL;
            for (int64_t i = 0; i < n_lag; i++)
                if (lag[i].cell == c && lag[i].face == k)
                    l = lag + i;
            for (int j = 0; j < 4; j++) {
                if (j == k)
                    continue;
                if (l) {
                    up[k][j] = l->vals[j * ng + g];
                    continue;
                }
                int64_t fv = m->c2f[(4 * c + k) * 4 + j];
                int64_t nv = m->nbrv[(4 * c + k) * 3 + fv];
                up[k][j] = psi[((nb * na + a) * 4 + nv) * ng + g];
            }
        }
        or_cell_solve(method, vol, abar, sigma_t[c * ng + g], q, up, out);
        for (int v = 0; v < 4; v++)
            psi[((c * na + a) * 4 + v) * ng + g] = out[v];
    }
}

/*
 * Sweep of one angle: Kahn's traversal of the cell graph whose edges are the
 * interior faces with Omega.n <= 0 as seen from the downwind cell
 * (GraphTraverser.cc:416-459 counts, 611-627 releases).  Any order of ready
 * cells gives the same psi.
 *
 * Cycles (the reference loops forever on them, GraphTraverser.cc:518): when
 * no cell is ready, the weakest unmet face -- smallest A |Omega.n| over every
 * interior incoming face of an unfinished cell whose upwind cell is
 * unfinished, lowest (cell, face) on ties -- is lagged, and the traversal
 * goes on.  A lagged face reads the upwind psi held before this sweep, as the
 * reference's PBJ sweeps read last iterate's partition faces
 * (SweeperPBJ.cc:168-219); at a converged iteration that is the current psi.
 * The set of finished cells at a stall does not depend on the traversal
 * order, so neither does the lagged set.
 *
 * Pass 1 fixes the order and the lagged faces, the snapshot of the lagged
 * faces' upwind values is taken, and pass 2 computes in that order.
 */
/* Dependency rule: < 0 the graph traverser's (every interior face with
 * Omega.n <= 0 is a dependency, TychoMesh.hh:107-110); >= 0 the precomputed
 * schedule's of SweepType OriginalTycho1/2 (SweepSchedule.cc:57,108-125:
 * only Omega.n < -tol is a dependency), the other interior faces with
 * Omega.n <= 0 lagged from the start. */
static double g_dep_tol = -1.0;

void or_set_dependency_tolerance(double tol) { g_dep_tol = tol; }

static void add_lag(or_lag **lag, int64_t *n_lag, int64_t *cap, int64_t c,
                    int k)
{
    if (*n_lag == *cap) {
        *cap = *cap ? 2 * *cap : 16;
        *lag = realloc(*lag, sizeof(or_lag) * *cap);
        if (!*lag)
            abort();
    }
    (*lag)[*n_lag].cell = c;
    (*lag)[*n_lag].face = k;
    (*lag)[*n_lag].vals = NULL;
    (*n_lag)++;
}

static int64_t sweep_angle(const or_mesh *m, int64_t a, int64_t na,
                           const double *om, int64_t ng,
                           const double *sigma_t, const double *source,
                           double *psi, int method, int64_t *indeg,
                           int64_t *queue, unsigned char *queued,
                           unsigned char *met)
{
    const int64_t nc = m->nc;
    int64_t head = 0, tail = 0, n_lag = 0, cap = 0;
    or_lag *lag = NULL;
    const double tol = g_dep_tol;
    /* met[c][k]: 0 unmet incoming interior face, 1 released by the upwind
     * cell, 2 lagged, 3 not a dependency */
    for (int64_t c = 0; c < nc; c++) {
        int64_t d = 0;
        for (int k = 0; k < 4; k++) {
            double dn = omega_dot_n(om, m->normal + (4 * c + k) * 3);
            int in = m->adj[4 * c + k] >= 0 && !(dn > 0);
            int dep = in && (tol < 0 || dn < -tol);
            met[4 * c + k] = dep ? 0 : in ? 2 : 3;
            if (in && !dep)
                add_lag(&lag, &n_lag, &cap, c, k);
            d += dep;
        }
        indeg[c] = d;
        queued[c] = 0;
    }
    for (int64_t c = 0; c < nc; c++)
        if (indeg[c] == 0) {
            queue[tail++] = c;
            queued[c] = 1;
        }
    while (head < nc) {
        if (head == tail) {
            int64_t bc = -1;
            int bk = -1;
            double bw = 0.0;
            for (int64_t c = 0; c < nc; c++) {
                if (queued[c])
                    continue;
                for (int k = 0; k < 4; k++) {
                    if (met[4 * c + k] != 0)
                        continue;
                    double d = omega_dot_n(om, m->normal + (4 * c + k) * 3);
                    double wgt = m->area[4 * c + k] * fabs(d);
                    if (bc < 0 || wgt < bw) {
                        bc = c;
                        bk = k;
                        bw = wgt;
                    }
                }
            }
            if (bc < 0)
                abort(); /* unreachable: a stalled cell has an unmet face */
            add_lag(&lag, &n_lag, &cap, bc, bk);
            met[4 * bc + bk] = 2;
            if (--indeg[bc] == 0) {
                queue[tail++] = bc;
                queued[bc] = 1;
            }
            continue;
        }
        int64_t c = queue[head++];
        for (int k = 0; k < 4; k++) {
            int64_t nb = m->adj[4 * c + k];
            if (nb < 0 || queued[nb] ||
                !(omega_dot_n(om, m->normal + (4 * c + k) * 3) > 0))
                continue;
            int back = 0; /* nb's face towards c */
            while (m->adj[4 * nb + back] != c)
                back++;
            if (met[4 * nb + back] != 0)
                continue;
            met[4 * nb + back] = 1;
            if (--indeg[nb] == 0) {
                queue[tail++] = nb;
                queued[nb] = 1;
            }
        }
    }
    /* snapshot of the lagged faces' upwind values */
    for (int64_t i = 0; i < n_lag; i++) {
        const int64_t c = lag[i].cell;
        const int k = lag[i].face;
        const int64_t nb = m->adj[4 * c + k];
        lag[i].vals = calloc(4 * ng, sizeof(double));
        for (int j = 0; j < 4; j++) {
            if (j == k)
                continue;
            int64_t fv = m->c2f[(4 * c + k) * 4 + j];
            int64_t nv = m->nbrv[(4 * c + k) * 3 + fv];
            for (int64_t g = 0; g < ng; g++)
                lag[i].vals[j * ng + g] =
                    psi[((nb * na + a) * 4 + nv) * ng + g];
        }
    }
    for (int64_t i = 0; i < nc; i++)
        update_cell(m, queue[i], a, na, om, ng, sigma_t, source, psi, method,
                    lag, n_lag);
    for (int64_t i = 0; i < n_lag; i++)
        free(lag[i].vals);
    free(lag);
    return n_lag;
}

int or_sweep_angles(const or_mesh *m, int64_t na, const double *omega,
                    int64_t ng, const double *sigma_t, const double *source,
                    double *psi, int method, int nthreads, int64_t a0,
                    int64_t a1, int64_t *n_lagged)
{
    const int64_t nc = m->nc;
    int64_t lag_total = 0;
    int err = 0;
#ifdef _OPENMP
    if (nthreads <= 0)
        nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads) reduction(+ : lag_total)
#endif
    {
        int64_t *indeg = malloc(sizeof(int64_t) * nc);
        int64_t *queue = malloc(sizeof(int64_t) * nc);
        unsigned char *queued = malloc(nc);
        unsigned char *lagged = malloc(4 * nc);
        if (!indeg || !queue || !queued || !lagged) {
#ifdef _OPENMP
#pragma omp atomic write
#endif
            err = -1;
        } else {
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
            for (int64_t a = a0; a < a1; a++)
                lag_total += sweep_angle(m, a, na, omega + 3 * a, ng, sigma_t,
                                         source, psi, method, indeg, queue,
                                         queued, lagged);
        }
        free(indeg);
        free(queue);
        free(queued);
        free(lagged);
    }
    (void)nthreads;
    if (n_lagged)
        *n_lagged = lag_total;
    return err;
}

int or_sweep(const or_mesh *m, int64_t na, const double *omega, int64_t ng,
             const double *sigma_t, const double *source, double *psi,
             int method, int nthreads, int64_t *n_lagged)
{
    return or_sweep_angles(m, na, omega, ng, sigma_t, source, psi, method,
                           nthreads, 0, na, n_lagged);
}

/* ---------------------------------------------------------------- defect */
/* Defect r = b - M psi_c of every (cell, angle, group), with M and b the
 * cell system above (Transport.cc:54-271, applied instead of inverted) and
 * every interior incoming face reading the upwind cell's psi (no lagging:
 * the full operator).  Stored as the nodal density d = (20/V)(r - sum(r)/5),
 * whose weak form V/20 (I + J) d is r, so that a sweep with d as its source
 * gives the correction psi* - psi.  source, psi, density: [c][a][v][g]. */
void or_defect(const or_mesh *m, int64_t na, const double *omega, int64_t ng,
               const double *sigma_t, const double *source, const double *psi,
               double *density)
{
    for (int64_t c = 0; c < m->nc; c++)
        for (int64_t a = 0; a < na; a++) {
            const double *om = omega + 3 * a;
            double abar[4], odn[4];
            for (int k = 0; k < 4; k++) {
                odn[k] = omega_dot_n(om, m->normal + (4 * c + k) * 3);
                abar[k] = m->area[4 * c + k] * odn[k];
            }
            const double vol = m->volume[c];
            for (int64_t g = 0; g < ng; g++) {
                double q[4], up[4][4], mat[4][4], b[4], r[4];
                for (int v = 0; v < 4; v++)
                    q[v] = source[((c * na + a) * 4 + v) * ng + g];
                for (int k = 0; k < 4; k++) {
                    for (int j = 0; j < 4; j++)
                        up[k][j] = 0.0;
                    int64_t nb = m->adj[4 * c + k];
                    if (odn[k] > 0 || nb < 0)
                        continue;
                    for (int j = 0; j < 4; j++) {
                        if (j == k)
                            continue;
                        int64_t fv = m->c2f[(4 * c + k) * 4 + j];
                        int64_t nv = m->nbrv[(4 * c + k) * 3 + fv];
                        up[k][j] = psi[((nb * na + a) * 4 + nv) * ng + g];
                    }
                }
                cell_system(vol, abar, sigma_t[c * ng + g], q, up, mat, b);
                double sum = 0.0;
                for (int i = 0; i < 4; i++) {
                    r[i] = b[i];
                    for (int j = 0; j < 4; j++)
                        r[i] -= mat[i][j] * psi[((c * na + a) * 4 + j) * ng + g];
                    sum += r[i];
                }
                for (int v = 0; v < 4; v++)
                    density[((c * na + a) * 4 + v) * ng + g] =
                        (r[v] - 0.2 * sum) / (vol / 20.0);
            }
        }
}

/* --------------------------------------------------------------- moments */
/* Util.cc:85-97: angle-ascending accumulation for each (cell, vertex, group) */
void or_psi_to_phi(int64_t nc, int64_t na, int64_t ng, const double *wt,
                   const double *psi, double *phi)
{
    for (int64_t c = 0; c < nc; c++) {
        double *p = phi + c * 4 * ng;
        for (int64_t i = 0; i < 4 * ng; i++)
            p[i] = 0.0;
        for (int64_t a = 0; a < na; a++) {
            const double *s = psi + (c * na + a) * 4 * ng;
            for (int64_t i = 0; i < 4 * ng; i++)
                p[i] += s[i] * wt[a];
        }
    }
}

/* Util.cc:118-130 with sigma_s per (cell, group) */
void or_total_source(int64_t nc, int64_t na, int64_t ng,
                     const double *source, const double *sigma_s,
                     const double *phi, double *total)
{
    for (int64_t c = 0; c < nc; c++)
        for (int64_t a = 0; a < na; a++)
            for (int64_t v = 0; v < 4; v++)
                for (int64_t g = 0; g < ng; g++) {
                    int64_t i = ((c * na + a) * 4 + v) * ng + g;
                    total[i] = (source ? source[i] : 0.0) +
                               sigma_s[c * ng + g] / (4.0 * M_PI) *
                                   phi[(c * 4 + v) * ng + g];
                }
}

static void phi_to_psi_scaled(int64_t nc, int64_t na, int64_t ng,
                              const double *sigma_s, const double *x,
                              double *total)
{
    /* Util::operatorS then Util::phiToPsi (Util.cc:103-112, 151-159) */
    for (int64_t c = 0; c < nc; c++)
        for (int64_t a = 0; a < na; a++)
            for (int64_t v = 0; v < 4; v++)
                for (int64_t g = 0; g < ng; g++)
                    total[((c * na + a) * 4 + v) * ng + g] =
                        sigma_s[c * ng + g] / (4.0 * M_PI) *
                        x[(c * 4 + v) * ng + g];
}

/* SourceIteration.cc:127-207 */
int or_source_iteration(const or_mesh *m, int64_t na, const double *omega,
                        const double *wt, int64_t ng, const double *sigma_t,
                        const double *sigma_s, const double *source,
                        int64_t iter_max, double err_max, int method,
                        int nthreads, double *psi, double *phi,
                        double *err_hist, int64_t *n_iter)
{
    const int64_t nc = m->nc, nphi = nc * 4 * ng;
    double *total = malloc(sizeof(double) * nc * na * 4 * ng);
    double *old = malloc(sizeof(double) * nphi);
    if (!total || !old) {
        free(total);
        free(old);
        return -1;
    }
    or_psi_to_phi(nc, na, ng, wt, psi, phi);
    int64_t iter = 0;
    double err = 1.0;
    int rc = 0;
    while (iter < iter_max && err > err_max) {
        memcpy(old, phi, sizeof(double) * nphi);
        or_total_source(nc, na, ng, source, sigma_s, old, total);
        rc = or_sweep(m, na, omega, ng, sigma_t, total, psi, method, nthreads,
                      NULL);
        if (rc)
            break;
        or_psi_to_phi(nc, na, ng, wt, psi, phi);
        double num = 0.0, den = 0.0;
        for (int64_t i = 0; i < nphi; i++) {
            num += fabs(phi[i] - old[i]);
            den += fabs(phi[i]);
        }
        err = num / den;
        if (err_hist)
            err_hist[iter] = err;
        iter++;
    }
    if (n_iter)
        *n_iter = iter;
    free(total);
    free(old);
    return rc;
}

/* ----------------------------------------------------------------- GMRES */
/* Restarted GMRES with modified Gram-Schmidt and Givens rotations, x0 = 0,
 * stopping when the residual estimate falls to rtol * ||b|| or after
 * iter_max iterations in total.  The reference hands this to PETSc's
 * KSPGMRES (KrylovSolver.cc:79-169), which is not part of the reference
 * sources; parity is pinned on the converged answer, not on the iterates. */
typedef void (*or_matvec)(void *ctx, const double *x, double *y);

static double dot(int64_t n, const double *x, const double *y)
{
    double s = 0.0;
    for (int64_t i = 0; i < n; i++)
        s += x[i] * y[i];
    return s;
}

static int gmres(int64_t n, or_matvec op, void *ctx, const double *b,
                 double *x, int64_t restart, int64_t iter_max, double rtol,
                 double *res_hist, int64_t *n_iter)
{
    const int64_t mr = restart;
    double *V = malloc(sizeof(double) * n * (mr + 1));
    double *H = calloc((size_t)(mr + 1) * mr, sizeof(double));
    double *cs = malloc(sizeof(double) * mr), *sn = malloc(sizeof(double) * mr);
    double *gv = malloc(sizeof(double) * (mr + 1)), *y = malloc(sizeof(double) * mr);
    double *w = malloc(sizeof(double) * n);
    if (!V || !H || !cs || !sn || !gv || !y || !w) {
        free(V); free(H); free(cs); free(sn); free(gv); free(y); free(w);
        return -1;
    }
    for (int64_t i = 0; i < n; i++)
        x[i] = 0.0;
    const double bnorm = sqrt(dot(n, b, b));
    const double target = rtol * bnorm;
    int64_t it = 0;
    double rnorm = bnorm;
    while (it < iter_max && rnorm > target) {
        /* r = b - A x */
        op(ctx, x, w);
        for (int64_t i = 0; i < n; i++)
            V[i] = b[i] - w[i];
        double beta = sqrt(dot(n, V, V));
        rnorm = beta;
        if (beta <= target || beta == 0.0)
            break;
        for (int64_t i = 0; i < n; i++)
            V[i] /= beta;
        for (int64_t i = 0; i <= mr; i++)
            gv[i] = 0.0;
        gv[0] = beta;
        int64_t j = 0;
        for (; j < mr && it < iter_max; j++) {
            double *vj1 = V + (j + 1) * n;
            op(ctx, V + j * n, vj1);
            for (int64_t i = 0; i <= j; i++) {
                double h = dot(n, vj1, V + i * n);
                H[i * mr + j] = h;
                for (int64_t k = 0; k < n; k++)
                    vj1[k] -= h * V[i * n + k];
            }
            double hn = sqrt(dot(n, vj1, vj1));
            H[(j + 1) * mr + j] = hn;
            if (hn != 0.0)
                for (int64_t k = 0; k < n; k++)
                    vj1[k] /= hn;
            for (int64_t i = 0; i < j; i++) {
                double t = cs[i] * H[i * mr + j] + sn[i] * H[(i + 1) * mr + j];
                H[(i + 1) * mr + j] =
                    -sn[i] * H[i * mr + j] + cs[i] * H[(i + 1) * mr + j];
                H[i * mr + j] = t;
            }
            double hjj = H[j * mr + j], hj1 = H[(j + 1) * mr + j];
            double r = sqrt(hjj * hjj + hj1 * hj1);
            cs[j] = hjj / r;
            sn[j] = hj1 / r;
            H[j * mr + j] = r;
            H[(j + 1) * mr + j] = 0.0;
            gv[j + 1] = -sn[j] * gv[j];
            gv[j] = cs[j] * gv[j];
            rnorm = fabs(gv[j + 1]);
            if (res_hist)
                res_hist[it] = rnorm / (bnorm > 0 ? bnorm : 1.0);
            it++;
            if (rnorm <= target) {
                j++;
                break;
            }
        }
        /* x += V y with H y = g */
        for (int64_t i = j - 1; i >= 0; i--) {
            double s = gv[i];
            for (int64_t k = i + 1; k < j; k++)
                s -= H[i * mr + k] * y[k];
            y[i] = s / H[i * mr + i];
        }
        for (int64_t i = 0; i < j; i++)
            for (int64_t k = 0; k < n; k++)
                x[k] += y[i] * V[i * n + k];
    }
    if (n_iter)
        *n_iter = it;
    free(V); free(H); free(cs); free(sn); free(gv); free(y); free(w);
    return 0;
}

struct dense_ctx {
    int64_t n;
    const double *a;
};

static void dense_op(void *ctx, const double *x, double *y)
{
    const struct dense_ctx *d = ctx;
    for (int64_t i = 0; i < d->n; i++)
        y[i] = dot(d->n, d->a + i * d->n, x);
}

int or_gmres_dense(int64_t n, const double *a, const double *b, double *x,
                   int64_t restart, int64_t iter_max, double rtol,
                   int64_t *n_iter)
{
    struct dense_ctx d = {n, a};
    return gmres(n, dense_op, &d, b, x, restart, iter_max, rtol, NULL, n_iter);
}

struct transport_ctx {
    const or_mesh *m;
    int64_t na, ng;
    const double *omega, *wt, *sigma_t, *sigma_s;
    double *psi, *total;
    int method, nthreads;
};

/* b = (I - D L^-1 M S) x  (SourceIteration.cc:76-114) */
static void transport_op(void *ctx, const double *x, double *y)
{
    struct transport_ctx *t = ctx;
    const int64_t nc = t->m->nc, n = nc * 4 * t->ng;
    phi_to_psi_scaled(nc, t->na, t->ng, t->sigma_s, x, t->total);
    or_sweep(t->m, t->na, t->omega, t->ng, t->sigma_t, t->total, t->psi,
             t->method, t->nthreads, NULL);
    or_psi_to_phi(nc, t->na, t->ng, t->wt, t->psi, y);
    for (int64_t i = 0; i < n; i++)
        y[i] = x[i] - y[i];
}

/* SourceIteration.cc:216-285 */
int or_krylov(const or_mesh *m, int64_t na, const double *omega,
              const double *wt, int64_t ng, const double *sigma_t,
              const double *sigma_s, const double *source, int64_t iter_max,
              double rtol, int64_t restart, int method, int nthreads,
              double *psi, double *phi, double *res_hist, int64_t *n_iter)
{
    const int64_t nc = m->nc, n = nc * 4 * ng;
    double *total = malloc(sizeof(double) * nc * na * 4 * ng);
    double *b = malloc(sizeof(double) * n);
    if (!total || !b) {
        free(total);
        free(b);
        return -1;
    }
    int rc = or_sweep(m, na, omega, ng, sigma_t, source, psi, method,
                      nthreads, NULL);
    or_psi_to_phi(nc, na, ng, wt, psi, b);
    struct transport_ctx t = {m, na, ng, omega, wt, sigma_t, sigma_s,
                              psi, total, method, nthreads};
    if (!rc)
        rc = gmres(n, transport_op, &t, b, phi, restart, iter_max, rtol,
                   res_hist, n_iter);
    if (!rc) {
        or_total_source(nc, na, ng, source, sigma_s, phi, total);
        rc = or_sweep(m, na, omega, ng, sigma_t, total, psi, method,
                      nthreads, NULL);
    }
    free(total);
    free(b);
    return rc;
}

/* ------------------------------------------------------------ hat source */
/* Problem.cc:47-82 on the cube [0,100]^3, sigma per cell. */
void or_hat_source(int64_t nc, const int64_t *cell_nodes, const double *coords,
                   int64_t na, const double *xi, const double *eta,
                   const double *mu, int64_t ng, const double *sigma_t_cell,
                   const double *sigma_s_cell, double *source)
{
    const double half = 100.0 / 2.0;
    for (int64_t c = 0; c < nc; c++)
        for (int64_t a = 0; a < na; a++)
            for (int64_t v = 0; v < 4; v++) {
                const double *p = coords + 3 * cell_nodes[4 * c + v];
                double x = p[0] - half, y = p[1] - half, z = p[2] - half;
                double r = sqrt(x * x + y * y + z * z);
                for (int64_t g = 0; g < ng; g++) {
                    double val = 0.0;
                    if (r <= 30.0)
                        val = -x / (30.0 * r) * xi[a] - y / (30.0 * r) * eta[a] -
                              z / (30.0 * r) * mu[a] +
                              (sigma_t_cell[c] - sigma_s_cell[c]) *
                                  (1.0 - r / 30.0);
                    source[((c * na + a) * 4 + v) * ng + g] = val;
                }
            }
}
