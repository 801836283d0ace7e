// Device kernels of the sweep engine (gfx950).
//
// The cell solve restates the reference's linear-DG update of one
// (cell, angle) for all groups (src/SweepData.hh:139-174,
// src/Transport.cc:54-410, 687-776) in registers.  One work item is one
// (angle, cell); its geometry is precomputed once per mesh into a compact
// record read as a stream in step order, and the lanes of a wavefront cover
// consecutive groups of an item, so every flux load and store touches
// group-contiguous memory.
#include "t2_kernels.hpp"

#include <algorithm>

namespace t2 {
namespace {

constexpr double kInv12 = 1.0 / 12.0;
// s_waitcnt operand (gfx9 encoding) waiting for vmcnt = 0 only: expcnt and
// lgkmcnt fields at their maxima.
constexpr int kWaitVmcnt0 = 0x0F70;
// Accelerator complex dies (XCDs) of an MI355X, each with its own L2.
constexpr int kXcds = 8;
constexpr double kInv20 = 1.0 / 20.0;

// Omega . n with each product and sum rounded separately: the host level
// builder evaluates the same expression, so both agree on face directions.
__device__ __forceinline__ double omega_dot_n(double ox, double oy, double oz,
                                              const double *n)
{
#pragma clang fp contract(off)
    return ox * n[0] + oy * n[1] + oz * n[2];
}

// Reciprocal: hardware estimate refined by two Newton steps in FP64 (within
// an ulp or so of 1/x), the hardware estimate alone in FP32.
template <typename T>
__device__ __forceinline__ T recip(T x);
template <>
__device__ __forceinline__ double recip<double>(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}
template <>
__device__ __forceinline__ float recip<float>(float x)
{
    return __builtin_amdgcn_rcpf(x);
}

// Index of `key` in the sorted array, or -1.
__device__ __forceinline__ int64_t find_key(const int64_t *keys, int64_t n,
                                            int64_t key)
{
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (keys[mid] < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    return (lo < n && keys[lo] == key) ? lo : -1;
}

// 4x4 solve by row elimination without pivoting (src/Transport.cc:338-410):
// rows 0..2 scaled by their pivot's reciprocal, then back substitution.
template <typename T>
__device__ __forceinline__ void solve_nopivot(T m[4][4], T b[4])
{
#pragma unroll
    for (int r = 0; r < 3; r++) {
        const T inv = recip<T>(m[r][r]);
#pragma unroll
        for (int c = r + 1; c < 4; c++)
            m[r][c] *= inv;
        b[r] *= inv;
#pragma unroll
        for (int row = r + 1; row < 4; row++) {
            const T t = m[row][r];
#pragma unroll
            for (int c = r + 1; c < 4; c++)
                m[row][c] -= m[r][c] * t;
            b[row] -= b[r] * t;
        }
    }
    b[3] *= recip<T>(m[3][3]);
#pragma unroll
    for (int r = 2; r >= 0; r--)
#pragma unroll
        for (int c = 3; c > r; c--)
            b[r] -= m[r][c] * b[c];
}

// 4x4 solve by elimination with partial pivoting (src/Transport.cc:284-334):
// the first row holding the largest |pivot| is swapped in, then a
// column-by-column LU and the two triangular solves.  Rows are swapped by
// selects over compile-time indices (no dynamically indexed registers).
template <typename T>
__device__ __forceinline__ void solve_pivot(T m[4][4], T b[4])
{
#pragma unroll
    for (int col = 0; col < 3; col++) {
        int best = col;
        T big = m[col][col] < T(0) ? -m[col][col] : m[col][col];
#pragma unroll
        for (int row = col + 1; row < 4; row++) {
            const T t = m[row][col] < T(0) ? -m[row][col] : m[row][col];
            if (t > big) {
                best = row;
                big = t;
            }
        }
#pragma unroll
        for (int row = col + 1; row < 4; row++) {
            const bool sw = best == row;
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const T x = m[row][c], y = m[col][c];
                m[row][c] = sw ? y : x;
                m[col][c] = sw ? x : y;
            }
            const T x = b[row], y = b[col];
            b[row] = sw ? y : x;
            b[col] = sw ? x : y;
        }
        m[col][col] = recip<T>(m[col][col]);
#pragma unroll
        for (int row = col + 1; row < 4; row++)
            m[row][col] *= m[col][col];
#pragma unroll
        for (int k = 0; k <= col; k++)
#pragma unroll
            for (int row = k + 1; row < 4; row++)
                m[row][col + 1] -= m[row][k] * m[k][col + 1];
    }
    m[3][3] = recip<T>(m[3][3]);
#pragma unroll
    for (int col = 0; col < 3; col++)
#pragma unroll
        for (int row = col + 1; row < 4; row++)
            b[row] -= m[row][col] * b[col];
#pragma unroll
    for (int col = 3; col >= 0; col--) {
        b[col] *= m[col][col];
#pragma unroll
        for (int row = col - 1; row >= 0; row--)
            b[row] -= m[row][col] * b[col];
    }
}

// Cramer's rule x = adj(A) b / det(A) (the reference's CramerGlu and
// CramerIntel options, src/Transport.cc:414-673), the adjugate by Laplace
// expansion over complementary 2x2 minors of rows {0,1} and {2,3}
// (D. Eberly, "The Laplace Expansion Theorem", Geometric Tools, 2008).
template <typename T>
__device__ __forceinline__ void solve_cramer(const T a[4][4], T b[4])
{
    const T s0 = a[0][0] * a[1][1] - a[1][0] * a[0][1];
    const T s1 = a[0][0] * a[1][2] - a[1][0] * a[0][2];
    const T s2 = a[0][0] * a[1][3] - a[1][0] * a[0][3];
    const T s3 = a[0][1] * a[1][2] - a[1][1] * a[0][2];
    const T s4 = a[0][1] * a[1][3] - a[1][1] * a[0][3];
    const T s5 = a[0][2] * a[1][3] - a[1][2] * a[0][3];
    const T c5 = a[2][2] * a[3][3] - a[3][2] * a[2][3];
    const T c4 = a[2][1] * a[3][3] - a[3][1] * a[2][3];
    const T c3 = a[2][1] * a[3][2] - a[3][1] * a[2][2];
    const T c2 = a[2][0] * a[3][3] - a[3][0] * a[2][3];
    const T c1 = a[2][0] * a[3][2] - a[3][0] * a[2][2];
    const T c0 = a[2][0] * a[3][1] - a[3][0] * a[2][1];
    const T inv = recip<T>(s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 +
                           s5 * c0);
    const T b0 = b[0], b1 = b[1], b2 = b[2], b3 = b[3];
    b[0] = ((a[1][1] * c5 - a[1][2] * c4 + a[1][3] * c3) * b0 +
            (-a[0][1] * c5 + a[0][2] * c4 - a[0][3] * c3) * b1 +
            (a[3][1] * s5 - a[3][2] * s4 + a[3][3] * s3) * b2 +
            (-a[2][1] * s5 + a[2][2] * s4 - a[2][3] * s3) * b3) * inv;
    b[1] = ((-a[1][0] * c5 + a[1][2] * c2 - a[1][3] * c1) * b0 +
            (a[0][0] * c5 - a[0][2] * c2 + a[0][3] * c1) * b1 +
            (-a[3][0] * s5 + a[3][2] * s2 - a[3][3] * s1) * b2 +
            (a[2][0] * s5 - a[2][2] * s2 + a[2][3] * s1) * b3) * inv;
    b[2] = ((a[1][0] * c4 - a[1][1] * c2 + a[1][3] * c0) * b0 +
            (-a[0][0] * c4 + a[0][1] * c2 - a[0][3] * c0) * b1 +
            (a[3][0] * s4 - a[3][1] * s2 + a[3][3] * s0) * b2 +
            (-a[2][0] * s4 + a[2][1] * s2 - a[2][3] * s0) * b3) * inv;
    b[3] = ((-a[1][0] * c3 + a[1][1] * c1 - a[1][2] * c0) * b0 +
            (a[0][0] * c3 - a[0][1] * c1 + a[0][2] * c0) * b1 +
            (-a[3][0] * s3 + a[3][1] * s1 - a[3][2] * s0) * b2 +
            (a[2][0] * s3 - a[2][1] * s1 + a[2][2] * s0) * b3) * inv;
}

// ---------------------------------------------------------------- set-up
template <typename T>
__global__ void build_items_kernel(const uint32_t *__restrict__ work,
                                   int64_t n, int cell_bits, int64_t nc,
                                   const double *__restrict__ omega,
                                   const double *__restrict__ geo,
                                   const int4 *__restrict__ adj,
                                   const uint32_t *__restrict__ nbr,
                                   const double *__restrict__ sigma_cell,
                                   const int64_t *__restrict__ lag_key,
                                   int64_t n_lag, ItemGeo<T> *out_geo,
                                   ItemTopo *out_topo, int *err)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t w = work[i];
    const int64_t c = int64_t(w & ((1u << cell_bits) - 1u));
    const int64_t a = int64_t(w >> cell_bits);
    const double ox = omega[3 * a], oy = omega[3 * a + 1],
                 oz = omega[3 * a + 2];
    const double *g = geo + c * kGeoStride;
    const int4 nb4 = adj[c];
    const int32_t nb[4] = {nb4.x, nb4.y, nb4.z, nb4.w};
    const uint32_t np = nbr[c];
    ItemGeo<T> rg;
    ItemTopo rt;
    rt.cell = int32_t(c);
    rt.angle = int32_t(a);
    rt.meta = 0u;
    rt.problem = 0;
    int n_in = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        // Abar_k = A_k (Omega . n_k)  (src/Transport.cc:697-704)
        const double abar = g[kGeoArea + k] * omega_dot_n(ox, oy, oz, g + 3 * k);
        rg.a12[k] = T(abar * kInv12);
        rt.up[k] = -1;
        if (abar < 0.0 && nb[k] >= 0) {
            n_in++;
            const int64_t lag =
                n_lag > 0 ? find_key(lag_key, n_lag, (a * nc + c) * 4 + k) : -1;
            rt.up[k] = lag >= 0 ? int32_t(-2 - lag) : nb[k];
            int slot = 0;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                if (j == k)
                    continue;
                const uint32_t vtx = (np >> (2 * (4 * k + j))) & 3u;
                rt.meta |= vtx << (6 * k + 2 * slot);
                slot++;
            }
        }
    }
    rg.v20 = T(g[kGeoVolume] * kInv20);
    rg.sig = sigma_cell ? T(sigma_cell[c]) : T(0);
    if (n_in > 3)
        *err = 1;
    out_geo[i] = rg;
    out_topo[i] = rt;
}

// ---------------------------------------------------------------- sweep
// Cell solve of work item `item` (topology record t) for the groups of lane
// gl: gl, gl + LPI, ... (LPI = G / GPL lanes per item).  COH: upwind psi is
// read and psi written at the device coherence point (sc1 loads, write-
// through stores), as the ready-queue kernel needs; otherwise plain cached
// accesses (kernel boundaries order the per-step launches).
template <typename T, int G, int GPL, bool COH = false,
          int SOLVER = kSolveNoPivot>
__device__ __forceinline__ void solve_item(const SweepArgs<T> &p,
                                           int64_t item, int gl,
                                           const ItemTopo &t)
{
    constexpr int LPI = G / GPL;          // lanes per work item
    constexpr uint32_t CG = 4 * G;        // values per (cell, angle)
    // record fields as values (a local copy of the whole record would be
    // placed in scratch memory by the sigma selection below)
    const ItemGeo<T> *gp = p.geo + item;
    const T ge_a12[4] = {gp->a12[0], gp->a12[1], gp->a12[2], gp->a12[3]};
    const T ge_v20 = gp->v20;
    const T ge_sig = gp->sig;
    int64_t nc = p.nc;
    int sigma_mode = p.sigma_mode;
    const T *sigma = p.sigma, *src_iso = p.src_iso, *lag_buf = p.lag_buf;
    T *psi = p.psi;
    const T *src_ang_t = p.src_ang_t;
    if (p.problems) {
        const ProblemDesc<T> &d = p.problems[t.problem];
        nc = d.nc;
        sigma_mode = d.sigma_mode;
        sigma = d.sigma;
        src_iso = d.src_iso;
        src_ang_t = d.src_ang_t;
        lag_buf = d.lag_buf;
        psi = d.psi;
    }
    const size_t plane = size_t(nc) * CG;
    const T *psi_a = psi + size_t(t.angle) * plane;
    const uint32_t own = uint32_t(t.cell) * CG;

    // outgoing-face sums of the streaming matrix
    T o[4];
    T P = T(0);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        o[k] = ge_a12[k] > T(0) ? ge_a12[k] : T(0);
        P += o[k];
    }

#pragma unroll
    for (int jg = 0; jg < GPL; jg++) {
        const int g = gl + jg * LPI;
        T sig = ge_sig;
        if (sigma_mode != kSigmaCell)
            sig = sigma[sigma_mode == kSigmaGroup ? uint32_t(g)
                                                  : uint32_t(t.cell) * G + g];
        const T s = sig * ge_v20;

        // source: V/20 (q_i + sum_j q_j)  (Transport.cc:58-72)
        T q[4];
#pragma unroll
        for (int v = 0; v < 4; v++)
            q[v] = src_iso[own + v * G + g];
        if (p.src_ang) {
            const double *qa = p.src_ang + size_t(t.angle) * plane + own + g;
#pragma unroll
            for (int v = 0; v < 4; v++)
                q[v] += T(qa[v * G]);
        }
        if (src_ang_t) {
            const T *qa = src_ang_t + size_t(t.angle) * plane + own + g;
#pragma unroll
            for (int v = 0; v < 4; v++)
                q[v] += qa[v * G];
        }
        const T sumq = (q[0] + q[1]) + (q[2] + q[3]);
        T b[4];
#pragma unroll
        for (int v = 0; v < 4; v++)
            b[v] = ge_v20 * (sumq + q[v]);

        // upwind faces: b_i -= a_k (sum_j psi_j + psi_i), i, j != k
        // (Transport.cc:175-271, 740-776); vacuum faces contribute zero
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int32_t u = t.up[k];
            if (u == -1)
                continue;
            T uv[4];
            if (u >= 0) {
                const T *pn = psi_a + uint32_t(u) * CG + g;
                int slot = 0;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    if (j == k) {
                        uv[j] = T(0);
                        continue;
                    }
                    const uint32_t vtx = (t.meta >> (6 * k + 2 * slot)) & 3u;
                    uv[j] = COH ? __hip_atomic_load(pn + vtx * G, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT)
                                : pn[vtx * G];
                    slot++;
                }
            } else {
                const T *pl = lag_buf + size_t(-2 - u) * CG + g;
#pragma unroll
                for (int j = 0; j < 4; j++)
                    uv[j] = (j == k) ? T(0) : pl[j * G];
            }
            const T sum = (uv[0] + uv[1]) + (uv[2] + uv[3]);
#pragma unroll
            for (int i = 0; i < 4; i++)
                if (i != k)
                    b[i] -= ge_a12[k] * (sum + uv[i]);
        }

        // M[i][j] = a_i + (1 + d_ij)(s + sum_{k out, k != i, j} a_k)
        // (Transport.cc:79-168)
        T m[4][4];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++)
                m[i][j] = (i == j) ? ge_a12[i] + T(2) * (s + (P - o[i]))
                                   : ge_a12[i] + s + ((P - o[i]) - o[j]);
        if constexpr (SOLVER == kSolveOriginal)
            solve_pivot<T>(m, b);
        else if constexpr (SOLVER == kSolveCramer)
            solve_cramer<T>(m, b);
        else
            solve_nopivot<T>(m, b);

        T *dst = psi + size_t(t.angle) * plane + own + g;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            if constexpr (COH)
                __hip_atomic_store(dst + v * G, b[v], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            else  // streamed: keeps psi from evicting cached inputs
                __builtin_nontemporal_store(b[v], dst + v * G);
        }
    }
}

template <typename T, int G, int GPL, int SOLVER = kSolveNoPivot>
__global__ void __launch_bounds__(kMaxBlock)
sweep_step_kernel(const SweepArgs<T> p)
{
    constexpr int LPI = G / GPL;
    // Blocks go to the XCDs round robin; XCD x takes the x-th contiguous
    // slice of the step (the grid is a multiple of kXcds), so the same
    // angles land on the same XCD step after step.
    const int64_t per_xcd = gridDim.x / kXcds;
    const int64_t lb =
        int64_t(blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
    const int64_t tid = lb * blockDim.x + threadIdx.x;
    const int64_t item = tid / LPI;
    const int gl = int(tid % LPI);
    if (item >= p.n_items)
        return;
    const ItemTopo t = p.topo[item];
    if (p.active && !p.active[t.problem])
        return;
    solve_item<T, G, GPL, false, SOLVER>(p, item, gl, t);
}

// ------------------------------------------------------- two items per lane
// Inputs of one item for one group, loaded ahead of its solve.
template <typename T>
struct ItemLoad {
    T a12[4], v20, sig;
    T q[4];
    T uw[4][4];        // upwind values per face (0 where face k is not read)
    int32_t up[4];
};

template <typename T, int G>
__device__ __forceinline__ void load_item(const SweepArgs<T> &p, int64_t item,
                                          int g, ItemLoad<T> &in)
{
    constexpr uint32_t CG = 4 * G;
    const ItemTopo t = p.topo[item];
    const ItemGeo<T> *gp = p.geo + item;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        in.a12[k] = gp->a12[k];
        in.up[k] = t.up[k];
    }
    in.v20 = gp->v20;
    in.sig = gp->sig;
    if (p.sigma_mode != kSigmaCell)
        in.sig = p.sigma[p.sigma_mode == kSigmaGroup
                             ? uint32_t(g)
                             : uint32_t(t.cell) * G + g];
    const size_t plane = size_t(p.nc) * CG;
    const T *psi_a = p.psi + size_t(t.angle) * plane;
    const uint32_t own = uint32_t(t.cell) * CG;
#pragma unroll
    for (int v = 0; v < 4; v++)
        in.q[v] = p.src_iso[own + v * G + g];
    if (p.src_ang) {
        const double *qa = p.src_ang + size_t(t.angle) * plane + own + g;
#pragma unroll
        for (int v = 0; v < 4; v++)
            in.q[v] += T(qa[v * G]);
    }
    if (p.src_ang_t) {
        const T *qa = p.src_ang_t + size_t(t.angle) * plane + own + g;
#pragma unroll
        for (int v = 0; v < 4; v++)
            in.q[v] += qa[v * G];
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
#pragma unroll
        for (int j = 0; j < 4; j++)
            in.uw[k][j] = T(0);
        const int32_t u = t.up[k];
        if (u == -1)
            continue;
        if (u >= 0) {
            const T *pn = psi_a + uint32_t(u) * CG + g;
            int slot = 0;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                if (j == k)
                    continue;
                const uint32_t vtx = (t.meta >> (6 * k + 2 * slot)) & 3u;
                in.uw[k][j] = pn[vtx * G];
                slot++;
            }
        } else {
            const T *pl = p.lag_buf + size_t(-2 - u) * CG + g;
#pragma unroll
            for (int j = 0; j < 4; j++)
                if (j != k)
                    in.uw[k][j] = pl[j * G];
        }
    }
}

// The cell solve of solve_item on preloaded inputs (same arithmetic, same
// order), and the store of psi.
template <typename T, int G>
__device__ __forceinline__ void solve_store_item(const SweepArgs<T> &p,
                                                 int64_t item, int g,
                                                 const ItemLoad<T> &in)
{
    constexpr uint32_t CG = 4 * G;
    T o[4];
    T P = T(0);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        o[k] = in.a12[k] > T(0) ? in.a12[k] : T(0);
        P += o[k];
    }
    const T s = in.sig * in.v20;
    const T sumq = (in.q[0] + in.q[1]) + (in.q[2] + in.q[3]);
    T b[4];
#pragma unroll
    for (int v = 0; v < 4; v++)
        b[v] = in.v20 * (sumq + in.q[v]);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (in.up[k] == -1)
            continue;
        const T *uv = in.uw[k];
        const T sum = (uv[0] + uv[1]) + (uv[2] + uv[3]);
#pragma unroll
        for (int i = 0; i < 4; i++)
            if (i != k)
                b[i] -= in.a12[k] * (sum + uv[i]);
    }
    T m[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
            m[i][j] = (i == j) ? in.a12[i] + T(2) * (s + (P - o[i]))
                               : in.a12[i] + s + ((P - o[i]) - o[j]);
    solve_nopivot<T>(m, b);
    const ItemTopo t = p.topo[item];
    T *dst = p.psi + size_t(t.angle) * size_t(p.nc) * CG +
             uint32_t(t.cell) * CG + g;
#pragma unroll
    for (int v = 0; v < 4; v++)
        __builtin_nontemporal_store(b[v], dst + v * G);
}

// Two items per lane: a wavefront covers items [w 2I, w 2I + 2I) with
// I = 64 / G items per set, lane (i, g) solving group g of items w 2I + i and
// w 2I + I + i.  Both items' loads are issued before either is solved and
// stored (the compiler cannot move loads of psi past stores to psi), so a
// wavefront has twice the loads in flight of one item per lane, each access
// with one item per lane's pattern.  Single-problem launches, NoPivot.
template <typename T, int G>
__global__ void __launch_bounds__(kMaxBlock)
sweep_step2_kernel(const SweepArgs<T> p)
{
    constexpr int I = 64 / G;
    const int64_t per_xcd = gridDim.x / kXcds;
    const int64_t lb =
        int64_t(blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
    const int64_t tid = lb * blockDim.x + threadIdx.x;
    const int64_t wave = tid / 64;
    const int lane = int(tid % 64);
    const int g = lane % G;
    const int64_t a = wave * 2 * I + lane / G;
    const int64_t b = a + I;
    if (a >= p.n_items)
        return;
    ItemLoad<T> in_a, in_b;
    load_item<T, G>(p, a, g, in_a);
    const bool has_b = b < p.n_items;
    if (has_b)
        load_item<T, G>(p, b, g, in_b);
    solve_store_item<T, G>(p, a, g, in_a);
    if (has_b)
        solve_store_item<T, G>(p, b, g, in_b);
}

// ---------------------------------------------------------------- ready queue
// The whole sweep in one persistent launch.  Blocks take chunks of work items
// from a global ticket counter in step order (a chunk never spans two steps)
// and split each chunk between their wavefronts.  A wavefront first waits
// until the pending-upwind counter of each of its items reaches zero, then
// solves them all, then decrements the counters of the items that read their
// psi.  Every dependency of a chunk lies in an earlier chunk, already held by
// a running block whose own dependencies are earlier still, so the lowest
// outstanding ticket can always proceed: nothing assumes how many blocks are
// resident.  Each wait is bounded by spin_limit polls; a wait that runs out
// sets the abort flag, after which every block stops taking work and exits.
//
// Cross-XCD visibility: psi is written through to the coherence point and
// upwind psi read from it (sc1 stores and loads), the counters are device-
// scope atomics, and the stores are complete (vmcnt drained) before any
// decrement is issued.  No L2-wide write-back or invalidate is needed, which
// agent-scope release/acquire fences would issue per wavefront and chunk.
// One wavefront's share of a chunk [lo, hi) of one problem's items: wait,
// solve, release.
template <typename T, int G, int GPL>
__device__ __forceinline__ void queue_chunk(const SweepArgs<T> &p,
                                            int32_t *pending,
                                            const int4 *down,
                                            const int32_t *arm,
                                            int64_t first_item, int64_t lo,
                                            int64_t hi, const QueueState &q)
{
    constexpr int LPI = G / GPL;          // lanes per item
    constexpr int IPR = 64 / LPI;         // items per wavefront round
    const int wave = int(threadIdx.x / 64), lane = int(threadIdx.x % 64);
    const int gl = lane % LPI;
    const int n_waves = int(blockDim.x / 64);
    // this wavefront's share of the chunk: an even split, so that a chunk
    // smaller than a block round still spreads over every wavefront
    const int64_t span = (hi - lo + n_waves - 1) / n_waves;
    const int64_t w_lo = lo + wave * span;
    const int64_t w_hi = w_lo + span < hi ? w_lo + span : hi;
    const int64_t first = w_lo + lane / LPI;
    bool ok = true;
    for (int64_t it = first; ok && it < w_hi; it += IPR) {
        const int32_t *mine = pending + (it - first_item);
        int64_t spins = 0;
        int nap = 1;
        while (__hip_atomic_load(mine, __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT) > 0) {
            if (++spins > q.spin_limit) {
                __hip_atomic_store(q.abort, 1, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
                ok = false;
                break;
            }
            if ((spins & 15) == 0 &&
                __hip_atomic_load(q.abort, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT) != 0) {
                ok = false;
                break;
            }
            // back off (up to ~1 us between polls): thousands of waiting
            // wavefronts polling at full rate would load the memory system
            // the solving ones need
            for (int k = 0; k < nap; k++)
                __builtin_amdgcn_s_sleep(2);
            nap = nap < 16 ? 2 * nap : 16;
        }
    }
    // no upwind load may move above the waits (compiler ordering; the loads
    // themselves read the coherence point)
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    for (int64_t it = first; ok && it < w_hi; it += IPR) {
        const ItemTopo t = p.topo[it];
        solve_item<T, G, GPL, true>(p, it, gl, t);
    }
    // every psi store of this wavefront is complete before a decrement
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
    if (ok && gl == 0) {
        for (int64_t it = first; it < w_hi; it += IPR) {
            int n_up;
            if (arm) {
                n_up = arm[it - first_item];
            } else {
                const ItemTopo t = p.topo[it];
                n_up = (t.up[0] >= 0) + (t.up[1] >= 0) + (t.up[2] >= 0) +
                       (t.up[3] >= 0);
            }
            // re-armed for the next sweep (nothing else touches it now)
            __hip_atomic_store(pending + (it - first_item), n_up,
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const int4 d = down[it - first_item];
            const int32_t dn[4] = {d.x, d.y, d.z, d.w};
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (dn[k] >= 0)
                    __hip_atomic_fetch_add(pending + (dn[k] - first_item), -1,
                                           __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// Next chunk for the whole block, or a value >= n_chunks once the work is
// handed out or the sweep was abandoned.
__device__ __forceinline__ int64_t next_chunk(const QueueState &q,
                                              int64_t *s_chunk)
{
    if (threadIdx.x == 0) {
        int64_t c = int64_t(__hip_atomic_fetch_add(
            q.ticket, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        if (__hip_atomic_load(q.abort, __ATOMIC_RELAXED,
                              __HIP_MEMORY_SCOPE_AGENT) != 0)
            c = q.n_chunks;
        *s_chunk = c;
    }
    __syncthreads();
    const int64_t c = *s_chunk;
    __syncthreads();
    return c;
}

template <typename T, int G, int GPL>
__global__ void __launch_bounds__(kMaxBlock)
queue_sweep_kernel(const SweepArgs<T> p, const QueueState q)
{
    __shared__ int64_t s_chunk;
    for (;;) {
        const int64_t c = next_chunk(q, &s_chunk);
        if (c >= q.n_chunks)
            break;
        queue_chunk<T, G, GPL>(p, q.pending, q.down, q.arm, q.first_item,
                               q.chunk_off[c], q.chunk_off[c + 1], q);
    }
}

// Several independent problems in one persistent launch: the chunk table
// interleaves the problems step by step, each chunk names its problem, and
// the problem's group count selects the lane mapping.
template <typename T>
__global__ void __launch_bounds__(kMaxBlock)
queue_batch_kernel(const BatchProblem<T> *problems, const BatchChunk *chunks,
                   const QueueState q)
{
    __shared__ int64_t s_chunk;
    for (;;) {
        const int64_t c = next_chunk(q, &s_chunk);
        if (c >= q.n_chunks)
            break;
        const BatchChunk ch = chunks[c];
        if (q.active && !q.active[ch.problem])
            continue;  // whole chunk skipped: its counters stay armed
        const BatchProblem<T> &bp = problems[ch.problem];
        const int64_t lo = ch.lo, hi = ch.lo + ch.n;
        switch (bp.groups) {
        case 1: queue_chunk<T, 1, 1>(bp.a, bp.pending, bp.down, bp.arm, bp.first_item, lo, hi, q); break;
        case 2: queue_chunk<T, 2, 1>(bp.a, bp.pending, bp.down, bp.arm, bp.first_item, lo, hi, q); break;
        case 4: queue_chunk<T, 4, 1>(bp.a, bp.pending, bp.down, bp.arm, bp.first_item, lo, hi, q); break;
        case 8: queue_chunk<T, 8, 1>(bp.a, bp.pending, bp.down, bp.arm, bp.first_item, lo, hi, q); break;
        case 16: queue_chunk<T, 16, 1>(bp.a, bp.pending, bp.down, bp.arm, bp.first_item, lo, hi, q); break;
        case 32: queue_chunk<T, 32, 1>(bp.a, bp.pending, bp.down, bp.arm, bp.first_item, lo, hi, q); break;
        case 64: queue_chunk<T, 64, 1>(bp.a, bp.pending, bp.down, bp.arm, bp.first_item, lo, hi, q); break;
        default: break;
        }
    }
}

// pos[a][c] = item of (a, c) for every item; the queued items' readers
// tables cleared.
__global__ void queue_index_kernel(const ItemTopo *__restrict__ topo,
                                   int64_t n, int64_t nc, int64_t first_item,
                                   int32_t *__restrict__ pos,
                                   int32_t *__restrict__ dcount,
                                   int4 *__restrict__ down)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const ItemTopo t = topo[i];
    pos[size_t(t.angle) * size_t(nc) + t.cell] = int32_t(i);
    if (i >= first_item) {
        dcount[i - first_item] = 0;
        down[i - first_item] = make_int4(-1, -1, -1, -1);
    }
}

// Counters of the queued items back to their start values.
__global__ void queue_rearm_kernel(const ItemTopo *__restrict__ topo,
                                   int64_t n, int64_t first_item,
                                   const int32_t *__restrict__ arm,
                                   int32_t *__restrict__ pending)
{
    const int64_t i = first_item + int64_t(blockIdx.x) * blockDim.x +
                      threadIdx.x;
    if (i >= n)
        return;
    int v;
    if (arm) {
        v = arm[i - first_item];
    } else {
        const ItemTopo t = topo[i];
        v = (t.up[0] >= 0) + (t.up[1] >= 0) + (t.up[2] >= 0) + (t.up[3] >= 0);
    }
    pending[i - first_item] = v;
}

// Links the queued items: down[j - first_item] collects the items that read
// item j's psi, and each queued item's counter (and arm[] copy, if given)
// is the number of its upwind items that are queued too (the others finished
// in earlier launches).
__global__ void queue_down_kernel(const ItemTopo *__restrict__ topo,
                                  int64_t n, int64_t nc, int64_t first_item,
                                  const int32_t *__restrict__ pos,
                                  int32_t *__restrict__ dcount,
                                  int4 *__restrict__ down,
                                  int32_t *__restrict__ pending,
                                  int32_t *__restrict__ arm, int *err)
{
    const int64_t i = first_item + int64_t(blockIdx.x) * blockDim.x +
                      threadIdx.x;
    if (i >= n)
        return;
    const ItemTopo t = topo[i];
    int count = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (t.up[k] < 0)
            continue;
        const int32_t j = pos[size_t(t.angle) * size_t(nc) + t.up[k]];
        if (j < first_item)
            continue;
        count++;
        const int slot = atomicAdd(dcount + (j - first_item), 1);
        if (slot < 4)
            reinterpret_cast<int32_t *>(down + (j - first_item))[slot] =
                int32_t(i);
        else
            *err = 1;
    }
    pending[i - first_item] = count;
    if (arm)
        arm[i - first_item] = count;
}

template <typename T, int G>
hipError_t launch_g(const SweepArgs<T> &args, int gpl, int block,
                    hipStream_t stream)
{
    if (block < 64 || block > kMaxBlock || block % 64)
        return hipErrorInvalidConfiguration;
    if (args.n_items <= 0)
        return hipSuccess;
    const int64_t lanes_per_item = G / gpl;
    const int64_t threads = args.n_items * lanes_per_item;
    const int64_t blocks = (threads + block - 1) / block;
    const unsigned grid = unsigned((blocks + kXcds - 1) / kXcds * kXcds);
    if (args.items_per_lane == 2) {
        if (gpl != 1 || args.solver != kSolveNoPivot || args.problems ||
            args.active)
            return hipErrorInvalidValue;
        // a wavefront covers 2 x 64 / G items
        const int64_t waves = (args.n_items * G + 127) / 128;
        const int64_t blocks2 = (waves * 64 + block - 1) / block;
        const unsigned grid2 = unsigned((blocks2 + kXcds - 1) / kXcds * kXcds);
        hipLaunchKernelGGL((sweep_step2_kernel<T, G>), dim3(grid2),
                           dim3(block), 0, stream, args);
        return hipGetLastError();
    }
    if (args.solver != kSolveNoPivot) {
        // the other solvers run one group per lane
        if (gpl != 1)
            return hipErrorInvalidValue;
        if (args.solver == kSolveOriginal)
            hipLaunchKernelGGL((sweep_step_kernel<T, G, 1, kSolveOriginal>),
                               dim3(grid), dim3(block), 0, stream, args);
        else if (args.solver == kSolveCramer)
            hipLaunchKernelGGL((sweep_step_kernel<T, G, 1, kSolveCramer>),
                               dim3(grid), dim3(block), 0, stream, args);
        else
            return hipErrorInvalidValue;
        return hipGetLastError();
    }
    switch (gpl) {
    case 1:
        hipLaunchKernelGGL((sweep_step_kernel<T, G, 1>), dim3(grid),
                           dim3(block), 0, stream, args);
        break;
    case 2:
        if constexpr (G >= 2) {
            hipLaunchKernelGGL((sweep_step_kernel<T, G, 2>), dim3(grid),
                               dim3(block), 0, stream, args);
            break;
        }
        return hipErrorInvalidValue;
    case 4:
        if constexpr (G >= 4) {
            hipLaunchKernelGGL((sweep_step_kernel<T, G, 4>), dim3(grid),
                               dim3(block), 0, stream, args);
            break;
        }
        return hipErrorInvalidValue;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------- batching
// Copies one problem's item records into the merged schedule of a batch:
// item i of step s goes to dst_off[s] + (i - src_off[s]).
template <typename T>
__global__ void merge_items_kernel(const ItemGeo<T> *__restrict__ src_geo,
                                   const ItemTopo *__restrict__ src_topo,
                                   const int64_t *__restrict__ src_off,
                                   const int64_t *__restrict__ dst_off,
                                   int64_t n_steps, int32_t problem,
                                   ItemGeo<T> *dst_geo, ItemTopo *dst_topo)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= src_off[n_steps])
        return;
    int64_t lo = 0, hi = n_steps;   // step s with src_off[s] <= i < src_off[s+1]
    while (hi - lo > 1) {
        const int64_t mid = (lo + hi) >> 1;
        if (src_off[mid] <= i)
            lo = mid;
        else
            hi = mid;
    }
    const int64_t d = dst_off[lo] + (i - src_off[lo]);
    ItemTopo t = src_topo[i];
    t.problem = problem;
    dst_topo[d] = t;
    dst_geo[d] = src_geo[i];
}

// ---------------------------------------------------------------- residual
// r_i = V/20 (s_i + sum_j s_j) - sum_{k in} a_k (sum_j u_kj + u_ki)
//       - sum_j M_ij psi_j   for every (angle, cell, group), in FP64.
template <typename T, int G>
__global__ void __launch_bounds__(kMaxBlock)
residual_kernel(const ResidualArgs<T> p)
{
    constexpr uint32_t CG = 4 * G;
    const int64_t tid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t item = tid / G;
    const int g = int(tid % G);
    if (item >= p.n_items)
        return;
    const ItemTopo t = p.topo[item];
    const ItemGeo<double> *gp = p.geo + item;
    const double ge_a12[4] = {gp->a12[0], gp->a12[1], gp->a12[2], gp->a12[3]};
    const double ge_v20 = gp->v20;
    const double ge_sig = gp->sig;
    const size_t plane = size_t(p.nc) * CG;
    const T *psi_a = p.psi + size_t(t.angle) * plane;
    const uint32_t own = uint32_t(t.cell) * CG;
    double o[4], P = 0.0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        o[k] = ge_a12[k] > 0.0 ? ge_a12[k] : 0.0;
        P += o[k];
    }
    double sig = ge_sig;
    if (p.sigma_mode != kSigmaCell)
        sig = p.sigma[p.sigma_mode == kSigmaGroup ? uint32_t(g)
                                                  : uint32_t(t.cell) * G + g];
    const double s = sig * ge_v20;
    double q[4], x[4];
#pragma unroll
    for (int v = 0; v < 4; v++) {
        q[v] = p.src_iso[own + v * G + g];
        if (p.src_ang)
            q[v] += p.src_ang[size_t(t.angle) * plane + own + v * G + g];
        x[v] = double(psi_a[own + v * G + g]);   // FP32 psi: exact widening
    }
    const double sumq = (q[0] + q[1]) + (q[2] + q[3]);
    double r[4];
#pragma unroll
    for (int v = 0; v < 4; v++)
        r[v] = ge_v20 * (sumq + q[v]);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        int32_t u = t.up[k];
        if (u == -1)
            continue;
        if (u < -1)
            u = p.lag_up[-2 - u];
        const T *pn = psi_a + uint32_t(u) * CG + g;
        double uv[4];
        int slot = 0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            if (j == k) {
                uv[j] = 0.0;
                continue;
            }
            const uint32_t vtx = (t.meta >> (6 * k + 2 * slot)) & 3u;
            uv[j] = double(pn[vtx * G]);
            slot++;
        }
        const double sum = (uv[0] + uv[1]) + (uv[2] + uv[3]);
#pragma unroll
        for (int i = 0; i < 4; i++)
            if (i != k)
                r[i] -= ge_a12[k] * (sum + uv[i]);
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const double m = (i == j) ? ge_a12[i] + 2.0 * (s + (P - o[i]))
                                      : ge_a12[i] + s + ((P - o[i]) - o[j]);
            r[i] -= m * x[j];
        }
    // r is a weak-form (integrated) defect; the sweep takes nodal source
    // densities q with weak form V/20 (I + J) q, so store the density
    // (20 / V)(I - J / 5) r that reproduces r exactly.
    const double rs = ((r[0] + r[1]) + (r[2] + r[3])) * 0.2;
    T *dst = p.r + size_t(t.angle) * plane + own + g;
#pragma unroll
    for (int v = 0; v < 4; v++)
        dst[v * G] = T((r[v] - rs) / ge_v20);
}

// ---------------------------------------------------------------- lagging
template <typename T>
__global__ void lag_gather_kernel(const int64_t *lag_key,
                                  const int32_t *lag_up, int64_t n_lag,
                                  int64_t nc, int G, const uint32_t *nbr,
                                  const T *psi, T *lag_buf)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n_lag * 4 * G)
        return;
    const int g = int(i % G);
    const int j = int((i / G) % 4);
    const int64_t l = i / (4 * G);
    const int64_t key = lag_key[l];
    const int k = int(key & 3);
    const int64_t ac = key >> 2;
    const int64_t c = ac % nc, a = ac / nc;
    T v = T(0);
    if (j != k) {
        const int vtx = int((nbr[c] >> (2 * (4 * k + j))) & 3u);
        v = psi[((size_t(a) * nc + lag_up[l]) * 4 + vtx) * G + g];
    }
    lag_buf[i] = v;
}

// ---------------------------------------------------------------- moments
template <typename T>
struct alignas(16) Vec16 {
    T v[16 / sizeof(T)];
};

template <typename T>
__global__ void psi_to_phi_kernel(const T *__restrict__ psi,
                                  const double *__restrict__ wt, int64_t na,
                                  int64_t n_phi, double *__restrict__ phi)
{
#pragma clang fp contract(off)
    constexpr int V = 16 / sizeof(T);
    const int64_t nv = n_phi / V;
    const Vec16<T> *src = reinterpret_cast<const Vec16<T> *>(psi);
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < nv;
         i += int64_t(gridDim.x) * blockDim.x) {
        double acc[V];
#pragma unroll
        for (int e = 0; e < V; e++)
            acc[e] = 0.0;
        // angles added in ascending order; unrolled so that several angle
        // planes' loads are in flight at once
#pragma unroll 8
        for (int64_t a = 0; a < na; a++) {
            const Vec16<T> x = src[size_t(a) * nv + i];
            const double w = wt[a];
#pragma unroll
            for (int e = 0; e < V; e++)
                acc[e] += double(x.v[e]) * w;
        }
#pragma unroll
        for (int e = 0; e < V; e++)
            phi[i * V + e] = acc[e];
    }
}

template <typename T>
__global__ void total_source_kernel(const double *__restrict__ q,
                                    const double *__restrict__ sigma_s,
                                    const double *__restrict__ phi,
                                    int64_t nc, int G, T *__restrict__ out)
{
#pragma clang fp contract(off)
    const int64_t n = nc * 4 * G;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x) {
        const int64_t c = i / (4 * G);
        const int g = int(i % G);
        const double s = sigma_s[c * G + g] / (4.0 * 3.14159265358979323846);
        const double v = (q ? q[i] : 0.0) + s * (phi ? phi[i] : 0.0);
        out[i] = T(v);
    }
}

template <typename T>
__global__ void ca_to_ac_kernel(const double *__restrict__ in,
                                T *__restrict__ out, int64_t nc,
                                int64_t na, int64_t vg)
{
    const int64_t n = nc * na * vg;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x) {
        const int64_t e = i % vg;
        const int64_t ca = i / vg;   // output index (a, c)
        const int64_t c = ca % nc, a = ca / nc;
        out[i] = T(in[(c * na + a) * vg + e]);
    }
}

template <typename T>
__global__ void ac_to_ca_kernel(const T *__restrict__ in,
                                const T *__restrict__ corr,
                                double *__restrict__ out, int64_t nc,
                                int64_t na, int64_t vg)
{
    const int64_t n = nc * na * vg;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x) {
        const int64_t e = i % vg;
        const int64_t ca = i / vg;   // output index (c, a)
        const int64_t a = ca % na, c = ca / na;
        const int64_t j = (a * nc + c) * vg + e;
        double v = double(in[j]);
        if (corr)   // the correction is added in FP64
            v += double(corr[j]);
        out[i] = v;
    }
}

// ---------------------------------------------------------------- reductions
__device__ __forceinline__ double block_sum(double v, double *lds)
{
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        lds[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < int(blockDim.x >> 6); i++)
            s += lds[i];
    __syncthreads();
    return s;  // valid in thread 0
}

__global__ void l1_change_kernel(const double *__restrict__ x,
                                 const double *__restrict__ y, int64_t n,
                                 double *partial)
{
    __shared__ double lds[16];
    double s0 = 0.0, s1 = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x) {
        s0 += fabs(x[i] - y[i]);
        s1 += fabs(x[i]);
    }
    s0 = block_sum(s0, lds);
    s1 = block_sum(s1, lds);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = s0;
        partial[gridDim.x + blockIdx.x] = s1;
    }
}

// out[r] = sum of partial[r * count .. (r+1) * count) in index order
__global__ void finish_sums_kernel(const double *partial, int count,
                                   double *out)
{
    __shared__ double lds[16];
    const double *p = partial + size_t(blockIdx.x) * count;
    double s = 0.0;
    for (int i = threadIdx.x; i < count; i += blockDim.x)
        s += p[i];
    s = block_sum(s, lds);
    if (threadIdx.x == 0)
        out[blockIdx.x] = s;
}

constexpr int kDotChunk = 8;

__global__ void multi_dot_kernel(const double *__restrict__ V, int64_t ldv,
                                 int m, const double *__restrict__ w,
                                 int64_t n, double *partial)
{
    __shared__ double lds[16];
    const int j0 = blockIdx.y * kDotChunk;
    double acc[kDotChunk];
#pragma unroll
    for (int j = 0; j < kDotChunk; j++)
        acc[j] = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x) {
        const double wi = w[i];
#pragma unroll
        for (int j = 0; j < kDotChunk; j++)
            if (j0 + j < m)
                acc[j] += V[size_t(j0 + j) * ldv + i] * wi;
    }
#pragma unroll
    for (int j = 0; j < kDotChunk; j++) {
        const double s = block_sum(acc[j], lds);
        if (threadIdx.x == 0 && j0 + j < m)
            partial[size_t(j0 + j) * gridDim.x + blockIdx.x] = s;
    }
}

__global__ void multi_axpy_kernel(const double *__restrict__ V, int64_t ldv,
                                  int m, const double *__restrict__ h,
                                  double *__restrict__ w, int64_t n)
{
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x) {
        double s = w[i];
        for (int j = 0; j < m; j++)
            s -= h[j] * V[size_t(j) * ldv + i];
        w[i] = s;
    }
}

__global__ void axpby_kernel(double alpha, const double *__restrict__ x,
                             double beta, double *__restrict__ y, int64_t n)
{
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x)
        y[i] = alpha * x[i] + beta * y[i];
}

__global__ void scale_kernel(double *x, double alpha, int64_t n)
{
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x)
        x[i] *= alpha;
}

__global__ void fill_kernel(double *x, double v, int64_t n)
{
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x)
        x[i] = v;
}

inline unsigned stream_grid(int64_t n, int block = 256)
{
    int64_t g = (n + block - 1) / block;
    if (g > 65536)
        g = 65536;
    if (g < 1)
        g = 1;
    return unsigned(g);
}

}  // namespace

template <typename T>
hipError_t launch_build_items(const uint32_t *work, int64_t n, int cell_bits,
                              int64_t nc, const double *omega,
                              const double *geo, const int4 *adj,
                              const uint32_t *nbr, const double *sigma_cell,
                              const int64_t *lag_key, int64_t n_lag,
                              ItemGeo<T> *out_geo, ItemTopo *out_topo,
                              int *err, hipStream_t stream)
{
    if (n <= 0)
        return hipSuccess;
    const unsigned grid = unsigned((n + 255) / 256);
    hipLaunchKernelGGL((build_items_kernel<T>), dim3(grid), dim3(256), 0,
                       stream, work, n, cell_bits, nc, omega, geo, adj, nbr,
                       sigma_cell, lag_key, n_lag, out_geo, out_topo, err);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_sweep_step(const SweepArgs<T> &args, int groups,
                             int groups_per_lane, int block,
                             hipStream_t stream)
{
    switch (groups) {
    case 1: return launch_g<T, 1>(args, groups_per_lane, block, stream);
    case 2: return launch_g<T, 2>(args, groups_per_lane, block, stream);
    case 4: return launch_g<T, 4>(args, groups_per_lane, block, stream);
    case 8: return launch_g<T, 8>(args, groups_per_lane, block, stream);
    case 16: return launch_g<T, 16>(args, groups_per_lane, block, stream);
    case 32: return launch_g<T, 32>(args, groups_per_lane, block, stream);
    case 64: return launch_g<T, 64>(args, groups_per_lane, block, stream);
    default: return hipErrorInvalidValue;
    }
}

template <typename T, int G, int GPL>
hipError_t queue_dispatch(const SweepArgs<T> *args, const QueueState *q,
                          int block, int64_t grid, hipStream_t stream,
                          int64_t *resident)
{
    if (resident) {
        int dev = 0, cus = 0, per_cu = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess)
            e = hipDeviceGetAttribute(
                &cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e == hipSuccess)
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per_cu,
                reinterpret_cast<const void *>(&queue_sweep_kernel<T, G, GPL>),
                block, 0);
        *resident = int64_t(std::max(1, per_cu)) * std::max(1, cus);
        return e;
    }
    grid = std::min<int64_t>(grid, std::max<int64_t>(1, q->n_chunks));
    hipLaunchKernelGGL((queue_sweep_kernel<T, G, GPL>),
                       dim3(unsigned(std::max<int64_t>(1, grid))), dim3(block),
                       0, stream, *args, *q);
    return hipGetLastError();
}

template <typename T, int G>
hipError_t queue_g(const SweepArgs<T> *args, const QueueState *q, int gpl,
                   int block, int64_t grid, hipStream_t stream,
                   int64_t *resident)
{
    if (block < 64 || block > kMaxBlock || block % 64)
        return hipErrorInvalidConfiguration;
    switch (gpl) {
    case 1:
        return queue_dispatch<T, G, 1>(args, q, block, grid, stream, resident);
    case 2:
        if constexpr (G >= 2)
            return queue_dispatch<T, G, 2>(args, q, block, grid, stream,
                                           resident);
        return hipErrorInvalidValue;
    case 4:
        if constexpr (G >= 4)
            return queue_dispatch<T, G, 4>(args, q, block, grid, stream,
                                           resident);
        return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
    }
}

template <typename T>
hipError_t queue_any(const SweepArgs<T> *args, const QueueState *q,
                     int groups, int gpl, int block, int64_t grid,
                     hipStream_t stream, int64_t *resident)
{
    switch (groups) {
    case 1: return queue_g<T, 1>(args, q, gpl, block, grid, stream, resident);
    case 2: return queue_g<T, 2>(args, q, gpl, block, grid, stream, resident);
    case 4: return queue_g<T, 4>(args, q, gpl, block, grid, stream, resident);
    case 8: return queue_g<T, 8>(args, q, gpl, block, grid, stream, resident);
    case 16: return queue_g<T, 16>(args, q, gpl, block, grid, stream, resident);
    case 32: return queue_g<T, 32>(args, q, gpl, block, grid, stream, resident);
    case 64: return queue_g<T, 64>(args, q, gpl, block, grid, stream, resident);
    default: return hipErrorInvalidValue;
    }
}

template <typename T>
hipError_t launch_queue_sweep(const SweepArgs<T> &args, const QueueState &q,
                              int groups, int groups_per_lane, int block,
                              int64_t grid, hipStream_t stream)
{
    if (q.n_chunks <= 0)
        return hipSuccess;
    return queue_any<T>(&args, &q, groups, groups_per_lane, block, grid,
                        stream, nullptr);
}

template <typename T>
hipError_t queue_grid(int groups, int groups_per_lane, int block,
                      int64_t *grid)
{
    return queue_any<T>(nullptr, nullptr, groups, groups_per_lane, block, 0,
                        nullptr, grid);
}

template <typename T>
hipError_t launch_queue_batch(const BatchProblem<T> *problems,
                              const BatchChunk *chunks, const QueueState &q,
                              int64_t grid, hipStream_t stream)
{
    if (q.n_chunks <= 0)
        return hipSuccess;
    grid = std::min<int64_t>(std::max<int64_t>(1, grid), q.n_chunks);
    hipLaunchKernelGGL((queue_batch_kernel<T>), dim3(unsigned(grid)),
                       dim3(kMaxBlock), 0, stream, problems, chunks, q);
    return hipGetLastError();
}

template <typename T>
hipError_t queue_batch_grid(int64_t *grid)
{
    int dev = 0, cus = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess)
        e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                                  dev);
    if (e == hipSuccess)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void *>(&queue_batch_kernel<T>),
            kMaxBlock, 0);
    *grid = int64_t(std::max(1, per_cu)) * std::max(1, cus);
    return e;
}

hipError_t launch_queue_links(const ItemTopo *topo, int64_t n_items,
                              int64_t nc, int64_t first_item, int32_t *pos,
                              int32_t *dcount, int4 *down, int32_t *pending,
                              int32_t *arm, int *err, hipStream_t stream)
{
    if (n_items <= first_item)
        return hipSuccess;
    hipLaunchKernelGGL(queue_index_kernel, dim3(unsigned((n_items + 255) / 256)),
                       dim3(256), 0, stream, topo, n_items, nc, first_item,
                       pos, dcount, down);
    hipLaunchKernelGGL(queue_down_kernel,
                       dim3(unsigned((n_items - first_item + 255) / 256)),
                       dim3(256), 0, stream, topo, n_items, nc, first_item,
                       pos, dcount, down, pending, arm, err);
    return hipGetLastError();
}

hipError_t launch_queue_rearm(const ItemTopo *topo, int64_t n_items,
                              int64_t first_item, const int32_t *arm,
                              int32_t *pending, hipStream_t stream)
{
    if (n_items <= first_item)
        return hipSuccess;
    hipLaunchKernelGGL(queue_rearm_kernel,
                       dim3(unsigned((n_items - first_item + 255) / 256)),
                       dim3(256), 0, stream, topo, n_items, first_item, arm,
                       pending);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_merge_items(const ItemGeo<T> *src_geo,
                              const ItemTopo *src_topo,
                              const int64_t *src_off, const int64_t *dst_off,
                              int64_t n_steps, int64_t n_items,
                              int32_t problem, ItemGeo<T> *dst_geo,
                              ItemTopo *dst_topo, hipStream_t stream)
{
    if (n_items <= 0)
        return hipSuccess;
    hipLaunchKernelGGL((merge_items_kernel<T>),
                       dim3(unsigned((n_items + 255) / 256)), dim3(256), 0,
                       stream, src_geo, src_topo, src_off, dst_off, n_steps,
                       problem, dst_geo, dst_topo);
    return hipGetLastError();
}

template <typename T, int G>
hipError_t launch_residual_g(const ResidualArgs<T> &args, hipStream_t stream)
{
    if (args.n_items <= 0)
        return hipSuccess;
    const int64_t threads = args.n_items * G;
    const unsigned grid = unsigned((threads + kMaxBlock - 1) / kMaxBlock);
    hipLaunchKernelGGL((residual_kernel<T, G>), dim3(grid), dim3(kMaxBlock),
                       0, stream, args);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_residual(const ResidualArgs<T> &args, int groups,
                           hipStream_t stream)
{
    switch (groups) {
    case 1: return launch_residual_g<T, 1>(args, stream);
    case 2: return launch_residual_g<T, 2>(args, stream);
    case 4: return launch_residual_g<T, 4>(args, stream);
    case 8: return launch_residual_g<T, 8>(args, stream);
    case 16: return launch_residual_g<T, 16>(args, stream);
    case 32: return launch_residual_g<T, 32>(args, stream);
    case 64: return launch_residual_g<T, 64>(args, stream);
    default: return hipErrorInvalidValue;
    }
}

template <typename T>
hipError_t launch_lag_gather(const int64_t *lag_key, const int32_t *lag_up,
                             int64_t n_lag, int64_t nc, int groups,
                             const uint32_t *nbr, const T *psi, T *lag_buf,
                             hipStream_t stream)
{
    if (n_lag <= 0)
        return hipSuccess;
    const int64_t n = n_lag * 4 * groups;
    hipLaunchKernelGGL((lag_gather_kernel<T>), dim3(stream_grid(n)),
                       dim3(256), 0, stream, lag_key, lag_up, n_lag, nc,
                       groups, nbr, psi, lag_buf);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_psi_to_phi(const T *psi, const double *wt, int64_t na,
                             int64_t n_phi, double *phi, hipStream_t stream)
{
    const int64_t nv = n_phi / (16 / sizeof(T));
    hipLaunchKernelGGL((psi_to_phi_kernel<T>), dim3(stream_grid(nv)),
                       dim3(256), 0, stream, psi, wt, na, n_phi, phi);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_total_source(const double *q, const double *sigma_s,
                               const double *phi, int64_t nc, int groups,
                               T *out, hipStream_t stream)
{
    const int64_t n = nc * 4 * groups;
    hipLaunchKernelGGL((total_source_kernel<T>), dim3(stream_grid(n)),
                       dim3(256), 0, stream, q, sigma_s, phi, nc, groups,
                       out);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_ca_to_ac(const double *in, T *out, int64_t nc,
                           int64_t na, int64_t vg, hipStream_t stream)
{
    hipLaunchKernelGGL((ca_to_ac_kernel<T>), dim3(stream_grid(nc * na * vg)),
                       dim3(256), 0, stream, in, out, nc, na, vg);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_ac_to_ca(const T *in, double *out, int64_t nc, int64_t na,
                           int64_t vg, hipStream_t stream, const T *corr)
{
    hipLaunchKernelGGL((ac_to_ca_kernel<T>), dim3(stream_grid(nc * na * vg)),
                       dim3(256), 0, stream, in, corr, out, nc, na, vg);
    return hipGetLastError();
}

hipError_t launch_l1_change(const double *x, const double *y, int64_t n,
                            double *partial, double *out, hipStream_t stream)
{
    hipLaunchKernelGGL(l1_change_kernel, dim3(kRedBlocks), dim3(kRedThreads),
                       0, stream, x, y, n, partial);
    hipLaunchKernelGGL(finish_sums_kernel, dim3(2), dim3(kRedThreads), 0,
                       stream, partial, kRedBlocks, out);
    return hipGetLastError();
}

hipError_t launch_multi_dot(const double *V, int64_t ldv, int m,
                            const double *w, int64_t n, double *partial,
                            double *out, hipStream_t stream)
{
    if (m <= 0)
        return hipSuccess;
    const unsigned chunks = unsigned((m + kDotChunk - 1) / kDotChunk);
    hipLaunchKernelGGL(multi_dot_kernel, dim3(kRedBlocks, chunks),
                       dim3(kRedThreads), 0, stream, V, ldv, m, w, n,
                       partial);
    hipLaunchKernelGGL(finish_sums_kernel, dim3(m), dim3(kRedThreads), 0,
                       stream, partial, kRedBlocks, out);
    return hipGetLastError();
}

hipError_t launch_multi_axpy(const double *V, int64_t ldv, int m,
                             const double *h, double *w, int64_t n,
                             hipStream_t stream)
{
    hipLaunchKernelGGL(multi_axpy_kernel, dim3(stream_grid(n)), dim3(256), 0,
                       stream, V, ldv, m, h, w, n);
    return hipGetLastError();
}

hipError_t launch_axpby(double alpha, const double *x, double beta, double *y,
                        int64_t n, hipStream_t stream)
{
    hipLaunchKernelGGL(axpby_kernel, dim3(stream_grid(n)), dim3(256), 0,
                       stream, alpha, x, beta, y, n);
    return hipGetLastError();
}

hipError_t launch_scale(double *x, double alpha, int64_t n,
                        hipStream_t stream)
{
    hipLaunchKernelGGL(scale_kernel, dim3(stream_grid(n)), dim3(256), 0,
                       stream, x, alpha, n);
    return hipGetLastError();
}

hipError_t launch_fill(double *x, double v, int64_t n, hipStream_t stream)
{
    hipLaunchKernelGGL(fill_kernel, dim3(stream_grid(n)), dim3(256), 0,
                       stream, x, v, n);
    return hipGetLastError();
}

#define T2_INSTANTIATE(T)                                                     \
    template hipError_t launch_build_items<T>(                                \
        const uint32_t *, int64_t, int, int64_t, const double *,              \
        const double *, const int4 *, const uint32_t *, const double *,       \
        const int64_t *, int64_t, ItemGeo<T> *, ItemTopo *, int *,            \
        hipStream_t);                                                         \
    template hipError_t launch_sweep_step<T>(const SweepArgs<T> &, int, int,  \
                                             int, hipStream_t);               \
    template hipError_t launch_residual<T>(const ResidualArgs<T> &, int,      \
                                           hipStream_t);                      \
    template hipError_t launch_queue_sweep<T>(const SweepArgs<T> &,           \
                                              const QueueState &, int, int,   \
                                              int, int64_t, hipStream_t);     \
    template hipError_t queue_grid<T>(int, int, int, int64_t *);              \
    template hipError_t launch_queue_batch<T>(const BatchProblem<T> *,        \
                                              const BatchChunk *,             \
                                              const QueueState &, int64_t,    \
                                              hipStream_t);                   \
    template hipError_t queue_batch_grid<T>(int64_t *);                       \
    template hipError_t launch_merge_items<T>(                                \
        const ItemGeo<T> *, const ItemTopo *, const int64_t *,                \
        const int64_t *, int64_t, int64_t, int32_t, ItemGeo<T> *,             \
        ItemTopo *, hipStream_t);                                             \
    template hipError_t launch_lag_gather<T>(const int64_t *, const int32_t *,\
                                             int64_t, int64_t, int,           \
                                             const uint32_t *, const T *,     \
                                             T *, hipStream_t);               \
    template hipError_t launch_psi_to_phi<T>(const T *, const double *,       \
                                             int64_t, int64_t, double *,      \
                                             hipStream_t);                    \
    template hipError_t launch_total_source<T>(const double *,                \
                                               const double *,                \
                                               const double *, int64_t, int,  \
                                               T *, hipStream_t);             \
    template hipError_t launch_ac_to_ca<T>(const T *, double *, int64_t,      \
                                           int64_t, int64_t, hipStream_t,     \
                                           const T *);                        \
    template hipError_t launch_ca_to_ac<T>(const double *, T *, int64_t,      \
                                           int64_t, int64_t, hipStream_t);

T2_INSTANTIATE(double)
T2_INSTANTIATE(float)

}  // namespace t2
