// C ABI of the sweep engine: device state, sweeps and outer iterations.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "t2_common.hpp"
#include "t2_kernels.hpp"

// Device scratch of t2_solve_source_iteration_noalloc: the L1-change
// reduction's per-block partial sums (2 x kRedBlocks) and its two totals.
// Every GPU the library runs on holds one copy until the process ends.  Host
// code uses only the per-device address that hipGetSymbolAddress returns
// (noalloc_scratch below), never the array's name.
constexpr int kNoAllocScratch = 2 * t2::kRedBlocks + 2;
static __device__ __align__(256) double scratch[kNoAllocScratch];

namespace {

using t2::fail;

#define T2_HIP(expr)                                                        \
    do {                                                                    \
        hipError_t e_ = (expr);                                             \
        if (e_ != hipSuccess)                                               \
            return fail(e_ == hipErrorOutOfMemory ? T2_ERR_NOMEM : T2_ERR_HIP, \
                        std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define T2_TRY(expr)        \
    do {                    \
        int rc_ = (expr);   \
        if (rc_ != T2_OK)   \
            return rc_;     \
    } while (0)

// A captured sweep, keyed by the angle-dependent sources it reads and the
// angular-flux buffer it writes.
struct GraphEntry {
    const double *ang;     // FP64 angle-dependent source or null
    const void *ang_t;     // psi-precision angle-dependent source or null
    const void *out;       // angular flux written (the handle's psi or the
                           // refined solve's correction buffer)
    hipGraphExec_t exec;
};

}  // namespace

struct t2_sweeper {
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t nc = 0, na = 0, ng = 0;
    int precision = T2_FP64, gauss = T2_GAUSS_NOPIVOT;
    int block = 256, gpl = 1, batches = 1, order = T2_ORDER_FACES;
    int ipl = 1;                     // items per lane (level engine)
    int intra = -1, inter = 0;       // priority schedule (intra < 0: levels)
    int64_t step_width = 0;
    double dep_tol = t2::kTraverseRule;   // dependency rule of the schedule
    bool use_graph = true;
    bool no_alloc = false;           // inside an allocation-free entry point
    bool concurrent = false;         // angle batches on parallel streams
    std::vector<hipStream_t> side;   // their extra streams
    std::vector<hipEvent_t> join_ev;
    hipEvent_t fork_ev = nullptr;
    void *h_pin[2] = {nullptr, nullptr};   // pinned host staging (host link)
    hipEvent_t pin_ev[2] = {nullptr, nullptr};
    t2::HostMesh hm;                 // kept for schedule rebuilds
    std::vector<double> omega_h, sigma_t_h;
    t2::Schedule sched;
    size_t bytes = 0;
    std::vector<void *> allocs;

    // static device data
    double *d_omega = nullptr, *d_wt = nullptr;
    double *d_sigs = nullptr;        // FP64 [nc][G]
    int sigma_mode = t2::kSigmaCellGroup;
    void *d_sigt = nullptr;          // T, layout by sigma_mode
    double *d_sigt64 = nullptr;      // FP64 copy (FP32 handles: residual)
    uint32_t *d_nbr = nullptr;
    // schedule-dependent
    void *d_item_geo = nullptr;      // ItemGeo<T> in step order
    t2::ItemGeo<double> *d_item_geo64 = nullptr;  // FP32 handles: residual
    t2::ItemTopo *d_item_topo = nullptr;
    int64_t *d_lag_key = nullptr;
    int32_t *d_lag_up = nullptr;
    void *d_lag_buf = nullptr;
    // fluxes and sources
    void *d_psi = nullptr;           // T [na][nc][4][G]
    void *d_src = nullptr;           // T [nc][4][G]
    double *d_src64 = nullptr;       // FP64 [nc][4][G] (residual)
    float *d_res = nullptr;          // FP32 [na][nc][4][G] (residual; after
                                     // a refined solve, the correction of psi)
    bool psi_corrected = false;      // psi = d_psi + d_res (refined solve)
    double *d_qiso = nullptr;        // FP64 [nc][4][G]
    double *d_qang = nullptr;        // FP64 [na][nc][4][G]
    double *d_stage = nullptr;       // FP64 psi-sized staging (host path)
    double *d_hostang = nullptr;     // FP64 [na][nc][4][G] (host path)
    double *d_phi = nullptr, *d_phi_old = nullptr;
    double *d_kvec = nullptr;        // 2 phi-sized vectors (Krylov apply)
    double *d_red = nullptr, *d_scal = nullptr;
    double *d_kry = nullptr;         // Krylov basis + work vectors
    int64_t kry_cols = 0;

    std::vector<GraphEntry> graphs;

    // ready-queue engine, built on first use for the schedule and launch
    // shape (n_chunks < 0: not built)
    int engine = T2_ENGINE_LEVELS;
    int chunk_rounds = 4;
    int64_t spin_limit = 0;
    int64_t n_sweeps = 0;            // full sweeps issued
    int64_t n_chunks = -1, queue_grid = 0;
    int64_t queue_step = -1;         // first queued step (-1: not built)
    int64_t split = -1;              // hybrid's first queued step (-1: auto)
    int64_t queue_item = 0;          // first queued item
    int64_t n_queued = 0;            // queued items
    int64_t *d_chunk_off = nullptr;
    int32_t *d_pending = nullptr;    // [n_queued] counters
    int4 *d_down = nullptr;          // [n_queued] readers
    int32_t *d_arm = nullptr;        // [n_queued] start values (hybrid)
    unsigned long long *d_ticket = nullptr;
    int32_t *d_abort = nullptr;

    int64_t n_phi() const { return nc * 4 * ng; }
    int64_t n_psi() const { return na * n_phi(); }
    size_t tsize() const { return precision == T2_FP64 ? 8 : 4; }
};

namespace {

// Allocation-failure hook (t2_set_alloc_fail): the device allocation that
// brings the count down to zero fails as an out-of-memory would.
std::atomic<int64_t> g_alloc_fail{0};

int dalloc(t2_sweeper *s, void **p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0)
        bytes = 8;
    if (g_alloc_fail.load(std::memory_order_relaxed) > 0 &&
        --g_alloc_fail == 0)
        return fail(T2_ERR_NOMEM, "hipMalloc(" + std::to_string(bytes) +
                                      " B): failed by t2_set_alloc_fail");
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(T2_ERR_NOMEM, "hipMalloc(" + std::to_string(bytes) +
                                      " B): " + hipGetErrorString(e));
    }
    s->allocs.push_back(*p);
    s->bytes += bytes;
    return T2_OK;
}

template <typename P>
int dalloc_t(t2_sweeper *s, P **p, size_t bytes)
{
    void *v = nullptr;
    int rc = dalloc(s, &v, bytes);
    *p = static_cast<P *>(v);
    return rc;
}

// Address of this device's copy of `scratch`, looked up once per device
// (the device must be current).  A fixed table: the lookup allocates nothing.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_dev[kMaxDevices] = {};

int noalloc_scratch(int device, double **out)
{
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices)
        return fail(T2_ERR_ARG, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        void *p = nullptr;
        T2_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p)
            return fail(T2_ERR_HIP, "hipGetSymbolAddress(scratch): null");
        g_scratch_dev[device] = static_cast<double *>(p);
    }
    *out = g_scratch_dev[device];
    return T2_OK;
}

template <typename P>
void dfree(t2_sweeper *s, P *&p, size_t bytes)
{
    if (!p)
        return;
    auto it = std::find(s->allocs.begin(), s->allocs.end(),
                        static_cast<void *>(p));
    if (it != s->allocs.end()) {
        s->allocs.erase(it);
        s->bytes -= bytes;
    }
    (void)hipFree(static_cast<void *>(p));
    p = nullptr;
}

// The deck's GaussElim choice as the kernels' solver (both Cramer options
// are one adjugate solve).
int cell_solver(int gauss)
{
    switch (gauss) {
    case T2_GAUSS_ORIGINAL: return t2::kSolveOriginal;
    case T2_GAUSS_CRAMER_GLU:
    case T2_GAUSS_CRAMER_INTEL: return t2::kSolveCramer;
    default: return t2::kSolveNoPivot;
    }
}

int set_device(const t2_sweeper *s)
{
    T2_HIP(hipSetDevice(s->device));
    return T2_OK;
}

int check(const t2_sweeper *s)
{
    if (!s)
        return fail(T2_ERR_ARG, "null sweeper handle");
    return set_device(s);
}

template <typename P>
int upload(P *dst, const P *src, size_t count)
{
    T2_HIP(hipMemcpy(dst, src, sizeof(P) * count, hipMemcpyHostToDevice));
    return T2_OK;
}

void drop_graphs(t2_sweeper *s)
{
    for (auto &g : s->graphs)
        (void)hipGraphExecDestroy(g.exec);
    s->graphs.clear();
}

// Default poll limit of a ready-queue wait (polls of ~1 us: seconds, far
// beyond any wait of a correct sweep, which is bounded by the sweep itself).
constexpr int64_t kDefaultSpinLimit = int64_t(1) << 22;

void free_queue(t2_sweeper *s)
{
    const size_t nq = size_t(s->n_queued);
    drop_graphs(s);
    if (s->n_chunks >= 0)
        dfree(s, s->d_chunk_off, sizeof(int64_t) * size_t(s->n_chunks + 1));
    dfree(s, s->d_pending, sizeof(int32_t) * nq);
    dfree(s, s->d_down, sizeof(int4) * nq);
    dfree(s, s->d_arm, sizeof(int32_t) * nq);
    s->n_chunks = -1;
    s->queue_step = -1;
    s->queue_item = 0;
    s->n_queued = 0;
}

// ------------------------------------------------------------------ sweeps
// Streams and events for angle batches swept concurrently (created once).
int ensure_side_streams(t2_sweeper *s, int64_t n)
{
    if (!s->fork_ev)
        T2_HIP(hipEventCreateWithFlags(&s->fork_ev, hipEventDisableTiming));
    while (int64_t(s->side.size()) < n) {
        hipStream_t st = nullptr;
        hipEvent_t ev = nullptr;
        T2_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        s->side.push_back(st);
        T2_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        s->join_ev.push_back(ev);
    }
    return T2_OK;
}

// psi_out = L^-1 (src + ang + ang_t).  Lagged faces read a snapshot of the
// handle's psi taken before the sweep; a sweep into another buffer (the
// refined solve's correction, whose previous value is zero) reads zeros
// there.
template <typename T>
int enqueue_sweep(t2_sweeper *s, const double *ang, const void *ang_t,
                  void *psi_out)
{
    const int64_t n_lag = int64_t(s->sched.lagged.size());
    if (n_lag > 0 && psi_out == s->d_psi)
        T2_HIP(t2::launch_lag_gather<T>(
            s->d_lag_key, s->d_lag_up, n_lag, s->nc, int(s->ng), s->d_nbr,
            static_cast<const T *>(s->d_psi), static_cast<T *>(s->d_lag_buf),
            s->stream));
    else if (n_lag > 0)   // a whole number of doubles: 4 G values per face
        T2_HIP(t2::launch_fill(static_cast<double *>(s->d_lag_buf), 0.0,
                               int64_t(sizeof(T)) * n_lag * 4 * s->ng / 8,
                               s->stream));
    t2::SweepArgs<T> a{};
    a.nc = s->nc;
    a.sigma_mode = s->sigma_mode;
    a.sigma = static_cast<const T *>(s->d_sigt);
    a.src_iso = static_cast<const T *>(s->d_src);
    a.src_ang = ang;
    a.src_ang_t = static_cast<const T *>(ang_t);
    a.psi = static_cast<T *>(psi_out);
    a.lag_buf = static_cast<const T *>(s->d_lag_buf);
    a.solver = cell_solver(s->gauss);
    a.items_per_lane = s->ipl;
    const auto *geo = static_cast<const t2::ItemGeo<T> *>(s->d_item_geo);
    auto launch_steps = [&](int64_t lo, int64_t hi, hipStream_t stream) {
        for (int64_t st = lo; st < hi; st++) {
            const int64_t off = s->sched.step_off[st];
            a.geo = geo + off;
            a.topo = s->d_item_topo + off;
            a.n_items = s->sched.step_off[st + 1] - off;
            T2_HIP(t2::launch_sweep_step<T>(a, int(s->ng), s->gpl, s->block,
                                            stream));
        }
        return T2_OK;
    };
    const int64_t nb = int64_t(s->sched.batch_steps.size()) - 1;
    if (s->engine == T2_ENGINE_LEVELS && s->concurrent && nb > 1) {
        // angle batches are independent: batch b > 0 runs on its own
        // stream, forked from and joined back to the handle's stream (in a
        // captured graph: parallel branches), so one batch's launch gaps
        // are filled by another's work
        T2_TRY(ensure_side_streams(s, nb - 1));
        T2_HIP(hipEventRecord(s->fork_ev, s->stream));
        for (int64_t b = 1; b < nb; b++) {
            hipStream_t sb = s->side[b - 1];
            T2_HIP(hipStreamWaitEvent(sb, s->fork_ev, 0));
            T2_TRY(launch_steps(s->sched.batch_steps[b],
                                s->sched.batch_steps[b + 1], sb));
            T2_HIP(hipEventRecord(s->join_ev[b - 1], sb));
        }
        T2_TRY(launch_steps(s->sched.batch_steps[0], s->sched.batch_steps[1],
                            s->stream));
        for (int64_t b = 1; b < nb; b++)
            T2_HIP(hipStreamWaitEvent(s->stream, s->join_ev[b - 1], 0));
    } else {
        // per-step launches: the whole schedule (level engine) or the steps
        // before the queued tail (hybrid)
        T2_TRY(launch_steps(0,
                            s->engine == T2_ENGINE_LEVELS ? s->sched.n_steps
                                                          : s->queue_step,
                            s->stream));
    }
    if (s->engine != T2_ENGINE_LEVELS && s->n_chunks > 0) {
        a.geo = geo;
        a.topo = s->d_item_topo;
        a.n_items = s->nc * s->na;
        t2::QueueState q{};
        q.chunk_off = s->d_chunk_off;
        q.n_chunks = s->n_chunks;
        q.pending = s->d_pending;
        q.down = s->d_down;
        q.arm = s->d_arm;
        q.first_item = s->queue_item;
        q.ticket = s->d_ticket;
        q.abort = s->d_abort;
        q.spin_limit = s->spin_limit > 0 ? s->spin_limit : kDefaultSpinLimit;
        // the ticket is cleared by a kernel (graph-safe), 8 bytes of zeros
        T2_HIP(t2::launch_fill(reinterpret_cast<double *>(s->d_ticket), 0.0,
                               1, s->stream));
        T2_HIP(t2::launch_queue_sweep<T>(a, q, int(s->ng), s->gpl, s->block,
                                         s->queue_grid, s->stream));
    }
    return T2_OK;
}

int enqueue_sweep_any(t2_sweeper *s, const double *ang, const void *ang_t,
                      void *psi_out)
{
    return s->precision == T2_FP64
               ? enqueue_sweep<double>(s, ang, ang_t, psi_out)
               : enqueue_sweep<float>(s, ang, ang_t, psi_out);
}

int queue_resident(t2_sweeper *s, int64_t *grid)
{
    if (s->precision == T2_FP64)
        T2_HIP(t2::queue_grid<double>(int(s->ng), s->gpl, s->block, grid));
    else
        T2_HIP(t2::queue_grid<float>(int(s->ng), s->gpl, s->block, grid));
    return T2_OK;
}

// First queued step of the current engine: 0 for the ready queue; for the
// hybrid, the first step of the thin tail -- the steps after it all hold
// fewer items than the queue kernel keeps resident in one round, so their
// per-step launches would leave the GPU mostly idle.
int queue_first_step(t2_sweeper *s, int64_t *first)
{
    *first = 0;
    if (s->engine != T2_ENGINE_HYBRID)
        return T2_OK;
    if (s->split >= 0) {
        *first = std::min(s->split, s->sched.n_steps);
        return T2_OK;
    }
    int64_t grid = 0;
    T2_TRY(queue_resident(s, &grid));
    const int64_t resident_items = grid * (s->block / (s->ng / s->gpl));
    int64_t st = s->sched.n_steps;
    while (st > 0 && s->sched.step_off[st] - s->sched.step_off[st - 1] <
                         resident_items)
        st--;
    *first = st;
    return T2_OK;
}

// Ready-queue tables for the schedule from step first_step on and the
// current launch shape: chunks of chunk_rounds rounds of one block's items
// that never span two steps, the readers of every queued item and the
// pending-upwind counters.
int build_queue(t2_sweeper *s, int64_t first_step)
{
    free_queue(s);
    const int64_t nc = s->nc, na = s->na;
    const int64_t n_items = nc * na;
    if (n_items >= (int64_t(1) << 31))
        return fail(T2_ERR_ARG, "ready queue: more than 2^31 items");
    const int64_t first = s->sched.step_off[first_step];
    const int64_t nq = n_items - first;
    const int64_t per_chunk =
        int64_t(s->block / (s->ng / s->gpl)) * s->chunk_rounds;
    // a level-set step too small to give every resident block a full chunk
    // is cut into more, smaller chunks (down to one wavefront round of
    // items), so that thin levels still spread over the GPU; the steps of a
    // priority schedule keep full chunks (they are cut by max_cells_per_step
    // and the next step's items are mostly ready already)
    int64_t resident = 0;
    T2_TRY(queue_resident(s, &resident));
    const int64_t per_wave =
        s->intra < 0 ? 64 / (s->ng / s->gpl) : per_chunk;
    std::vector<int64_t> off(1, first);
    for (int64_t st = first_step; st < s->sched.n_steps; st++) {
        const int64_t lo = s->sched.step_off[st], hi = s->sched.step_off[st + 1];
        const int64_t fair = (hi - lo + resident - 1) / std::max<int64_t>(
                                                           1, resident);
        const int64_t c = std::max(per_wave, std::min(per_chunk, fair));
        for (int64_t i = lo; i < hi; i += c)
            off.push_back(std::min(i + c, hi));
    }
    T2_TRY(dalloc_t(s, &s->d_chunk_off, sizeof(int64_t) * off.size()));
    // from here on free_queue releases what was allocated
    s->n_chunks = int64_t(off.size()) - 1;
    s->queue_step = first_step;
    s->queue_item = first;
    s->n_queued = nq;
    T2_TRY(upload(s->d_chunk_off, off.data(), off.size()));
    T2_TRY(dalloc_t(s, &s->d_pending, sizeof(int32_t) * size_t(nq)));
    T2_TRY(dalloc_t(s, &s->d_down, sizeof(int4) * size_t(nq)));
    if (first > 0)
        T2_TRY(dalloc_t(s, &s->d_arm, sizeof(int32_t) * size_t(nq)));
    if (!s->d_ticket) {
        T2_TRY(dalloc_t(s, &s->d_ticket, sizeof(unsigned long long)));
        T2_TRY(dalloc_t(s, &s->d_abort, sizeof(int32_t)));
        T2_HIP(hipMemset(s->d_abort, 0, sizeof(int32_t)));
    }
    int32_t *pos = nullptr, *dcount = nullptr;
    int *d_err = nullptr;
    auto cleanup = [&]() {
        dfree(s, pos, sizeof(int32_t) * size_t(n_items));
        dfree(s, dcount, sizeof(int32_t) * size_t(nq));
        dfree(s, d_err, sizeof(int));
    };
    int rc = T2_OK, err = 0;
    if ((rc = dalloc_t(s, &pos, sizeof(int32_t) * size_t(n_items))) ||
        (rc = dalloc_t(s, &dcount, sizeof(int32_t) * size_t(nq))) ||
        (rc = dalloc_t(s, &d_err, sizeof(int)))) {
        cleanup();
        return rc;
    }
    hipError_t e = hipMemsetAsync(d_err, 0, sizeof(int), s->stream);
    if (e == hipSuccess)
        e = t2::launch_queue_links(s->d_item_topo, n_items, nc, first, pos,
                                   dcount, s->d_down, s->d_pending, s->d_arm,
                                   d_err, s->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(&err, d_err, sizeof(int), hipMemcpyDeviceToHost,
                           s->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(s->stream);
    cleanup();
    if (e != hipSuccess)
        return fail(T2_ERR_HIP, std::string("ready-queue set-up: ") +
                                    hipGetErrorString(e));
    if (err)
        return fail(T2_ERR_ARG, "ready-queue set-up: a cell feeds more than "
                                "four cells of one angle");
    return queue_resident(s, &s->queue_grid);
}

// Tables of the current engine, rebuilt when stale.
int ensure_queue(t2_sweeper *s)
{
    if (s->engine == T2_ENGINE_LEVELS)
        return T2_OK;
    int64_t first = 0;
    T2_TRY(queue_first_step(s, &first));
    if (s->n_chunks >= 0 && s->queue_step == first)
        return T2_OK;
    T2_HIP(hipStreamSynchronize(s->stream));
    return build_queue(s, first);
}

// After an abandoned ready-queue sweep: reports it once and re-arms the
// counters so that later sweeps are valid.
int check_abort(t2_sweeper *s)
{
    if (!s->d_abort)
        return T2_OK;
    int32_t flag = 0;
    T2_HIP(hipStreamSynchronize(s->stream));
    T2_HIP(hipMemcpy(&flag, s->d_abort, sizeof(flag), hipMemcpyDeviceToHost));
    if (!flag)
        return T2_OK;
    if (s->d_pending)
        T2_HIP(t2::launch_queue_rearm(s->d_item_topo, s->nc * s->na,
                                      s->queue_item, s->d_arm, s->d_pending,
                                      s->stream));
    T2_HIP(hipMemsetAsync(s->d_abort, 0, sizeof(int32_t), s->stream));
    T2_HIP(hipStreamSynchronize(s->stream));
    return fail(T2_ERR_TIMEOUT,
                "a ready-queue wait ran past its poll limit; the sweep was "
                "abandoned and its results are invalid");
}

int after_queue(t2_sweeper *s, int rc)
{
    if (!s)
        return rc;
    const int a = check_abort(s);
    return a != T2_OK ? a : rc;
}

// psi = L^-1 (src + ang + ang_t): every step of the schedule, captured once
// per source combination into a hipGraph and replayed.
// Most step launches one level-engine sweep may issue (one hipGraph node
// each).  Priority schedules with a small max_cells_per_step can have
// hundreds of thousands of steps; sweeps of up to this many launches are the
// ones known to run.
constexpr int64_t kMaxStepLaunches = 24576;

int check_step_launches(int64_t launches)
{
    if (launches > kMaxStepLaunches)
        return fail(T2_ERR_ARG,
                    "the level engine issues at most " +
                        std::to_string(kMaxStepLaunches) +
                        " step launches per sweep; this schedule needs " +
                        std::to_string(launches) +
                        " (use the ready-queue engine or a larger "
                        "max_cells_per_step)");
    return T2_OK;
}

int run_sweep(t2_sweeper *s, const double *ang, const void *ang_t = nullptr,
              void *psi_out = nullptr)
{
    if (!psi_out)
        psi_out = s->d_psi;
    if (psi_out == s->d_psi)   // psi is replaced: no correction applies
        s->psi_corrected = false;
    const int64_t n_side = int64_t(s->sched.batch_steps.size()) - 2;
    if (s->no_alloc) {
        // nothing may be created here: the ready-queue tables and the
        // streams of concurrent batches must exist already
        int64_t first = 0;
        if (s->engine != T2_ENGINE_LEVELS) {
            T2_TRY(queue_first_step(s, &first));
            if (s->n_chunks < 0 || s->queue_step != first)
                return fail(T2_ERR_STATE,
                            "allocation-free call: the ready-queue tables "
                            "of this engine and schedule are not built "
                            "(an ordinary sweep or solve builds them)");
        }
        if (s->concurrent && n_side > 0 &&
            (!s->fork_ev || int64_t(s->side.size()) < n_side))
            return fail(T2_ERR_STATE,
                        "allocation-free call: the streams of concurrent "
                        "angle batches are not created (an ordinary sweep "
                        "or solve creates them)");
    }
    T2_TRY(ensure_queue(s));
    T2_TRY(check_step_launches(s->engine == T2_ENGINE_LEVELS
                                   ? s->sched.n_steps
                                   : s->queue_step));
    if (s->concurrent && !s->no_alloc)   // created outside any capture
        T2_TRY(ensure_side_streams(s, n_side));
    s->n_sweeps++;
    if (!s->use_graph)
        return enqueue_sweep_any(s, ang, ang_t, psi_out);
    hipGraphExec_t exec = nullptr;
    for (auto &g : s->graphs)
        if (g.ang == ang && g.ang_t == ang_t && g.out == psi_out)
            exec = g.exec;
    if (!exec && s->no_alloc)   // no capture: the same launches, direct
        return enqueue_sweep_any(s, ang, ang_t, psi_out);
    if (!exec) {
        hipGraph_t g = nullptr;
        T2_HIP(hipStreamBeginCapture(s->stream,
                                     hipStreamCaptureModeThreadLocal));
        int rc = enqueue_sweep_any(s, ang, ang_t, psi_out);
        hipError_t e = hipStreamEndCapture(s->stream, &g);
        if (rc != T2_OK) {
            if (g)
                (void)hipGraphDestroy(g);
            return rc;
        }
        T2_HIP(e);
        e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        T2_HIP(e);
        s->graphs.push_back({ang, ang_t, psi_out, exec});
    }
    T2_HIP(hipGraphLaunch(exec, s->stream));
    return T2_OK;
}

// src (T) = q + sigma_s / (4 pi) phi   (either may be null)
int total_source(t2_sweeper *s, const double *q, const double *phi)
{
    if (s->precision == T2_FP64)
        T2_HIP(t2::launch_total_source<double>(
            q, s->d_sigs, phi, s->nc, int(s->ng),
            static_cast<double *>(s->d_src), s->stream));
    else
        T2_HIP(t2::launch_total_source<float>(
            q, s->d_sigs, phi, s->nc, int(s->ng),
            static_cast<float *>(s->d_src), s->stream));
    return T2_OK;
}

int psi_to_phi(t2_sweeper *s, double *phi)
{
    if (s->precision == T2_FP64)
        T2_HIP(t2::launch_psi_to_phi<double>(
            static_cast<const double *>(s->d_psi), s->d_wt, s->na,
            s->n_phi(), phi, s->stream));
    else
        T2_HIP(t2::launch_psi_to_phi<float>(
            static_cast<const float *>(s->d_psi), s->d_wt, s->na, s->n_phi(),
            phi, s->stream));
    return T2_OK;
}

int zero_psi(t2_sweeper *s)
{
    // psi is a whole number of doubles in both precisions (4 values/item)
    const int64_t n = int64_t(s->n_psi() * s->tsize() / 8);
    T2_HIP(t2::launch_fill(static_cast<double *>(s->d_psi), 0.0, n,
                           s->stream));
    return T2_OK;
}

// psi = 0 at the start of a solve whose first step is a full sweep: before
// that sweep has written all of psi only the lagged faces read it (their
// pre-sweep snapshot), so without lagged faces the fill is skipped.
int zero_psi_before_sweep(t2_sweeper *s)
{
    return s->sched.lagged.empty() ? T2_OK : zero_psi(s);
}

// ---------------------------------------------------------- host link
// psi-sized transfers between the caller's ordinary (pageable) arrays and
// the device go through two pinned chunks: host threads copy one chunk
// while the DMA engine moves the other, instead of the runtime's serial
// staging of pageable memory.
constexpr size_t kPinChunk = size_t(64) << 20;

int ensure_pinned(t2_sweeper *s)
{
    for (int b = 0; b < 2; b++) {
        if (!s->h_pin[b])
            T2_HIP(hipHostMalloc(&s->h_pin[b], kPinChunk));
        if (!s->pin_ev[b])
            T2_HIP(hipEventCreateWithFlags(&s->pin_ev[b],
                                           hipEventDisableTiming));
    }
    return T2_OK;
}

void parallel_copy(void *dst, const void *src, size_t bytes)
{
    const size_t min_part = size_t(4) << 20;
    unsigned hw = std::thread::hardware_concurrency();
    size_t parts = std::min<size_t>(hw ? std::min(hw, 8u) : 4u,
                                    std::max<size_t>(1, bytes / min_part));
    if (parts <= 1) {
        std::memcpy(dst, src, bytes);
        return;
    }
    const size_t step = (bytes + parts - 1) / parts;
    std::vector<std::thread> pool;
    for (size_t i = 1; i < parts; i++) {
        const size_t lo = i * step;
        if (lo >= bytes)
            break;
        const size_t len = std::min(step, bytes - lo);
        pool.emplace_back([=]() {
            std::memcpy(static_cast<char *>(dst) + lo,
                        static_cast<const char *>(src) + lo, len);
        });
    }
    std::memcpy(dst, src, std::min(step, bytes));
    for (auto &t : pool)
        t.join();
}

// Host -> device, ordered on the handle's stream (later work sees it).
int upload_host(t2_sweeper *s, void *dev, const void *host, size_t bytes)
{
    T2_TRY(ensure_pinned(s));
    for (size_t off = 0, k = 0; off < bytes; off += kPinChunk, k++) {
        const size_t len = std::min(kPinChunk, bytes - off);
        const int b = int(k & 1);
        T2_HIP(hipEventSynchronize(s->pin_ev[b]));   // its last DMA is done
        parallel_copy(s->h_pin[b], static_cast<const char *>(host) + off, len);
        T2_HIP(hipMemcpyAsync(static_cast<char *>(dev) + off, s->h_pin[b],
                              len, hipMemcpyHostToDevice, s->stream));
        T2_HIP(hipEventRecord(s->pin_ev[b], s->stream));
    }
    return T2_OK;
}

// Device -> host after all earlier work on the handle's stream; returns
// when the host array is complete.
int download_host(t2_sweeper *s, void *host, const void *dev, size_t bytes)
{
    T2_TRY(ensure_pinned(s));
    size_t prev_off = 0, prev_len = 0;
    int prev_b = -1;
    for (size_t off = 0, k = 0; off < bytes; off += kPinChunk, k++) {
        const size_t len = std::min(kPinChunk, bytes - off);
        const int b = int(k & 1);
        T2_HIP(hipMemcpyAsync(s->h_pin[b], static_cast<const char *>(dev) + off,
                              len, hipMemcpyDeviceToHost, s->stream));
        T2_HIP(hipEventRecord(s->pin_ev[b], s->stream));
        if (prev_b >= 0) {
            T2_HIP(hipEventSynchronize(s->pin_ev[prev_b]));
            parallel_copy(static_cast<char *>(host) + prev_off,
                          s->h_pin[prev_b], prev_len);
        }
        prev_off = off;
        prev_len = len;
        prev_b = b;
    }
    if (prev_b >= 0) {
        T2_HIP(hipEventSynchronize(s->pin_ev[prev_b]));
        parallel_copy(static_cast<char *>(host) + prev_off, s->h_pin[prev_b],
                      prev_len);
    }
    return T2_OK;
}

int fetch(t2_sweeper *s, double *host, const double *dev, int64_t n)
{
    T2_HIP(hipMemcpyAsync(host, dev, sizeof(double) * n,
                          hipMemcpyDeviceToHost, s->stream));
    T2_HIP(hipStreamSynchronize(s->stream));
    return T2_OK;
}

int dot(t2_sweeper *s, const double *x, const double *y, int64_t n,
        double *out)
{
    T2_HIP(t2::launch_multi_dot(x, n, 1, y, n, s->d_red, s->d_scal,
                                s->stream));
    return fetch(s, out, s->d_scal, 1);
}

// ------------------------------------------------------------------ set-up
// sigma_t layout on the device: per cell (inside the item records) when
// every group agrees, per group when every cell agrees, else [cell][group].
int pick_sigma_mode(int64_t nc, int64_t ng, const double *sig)
{
    bool by_group = true, by_cell = true;
    for (int64_t c = 0; c < nc && (by_group || by_cell); c++)
        for (int64_t g = 0; g < ng; g++) {
            by_group &= sig[c * ng + g] == sig[g];
            by_cell &= sig[c * ng + g] == sig[c * ng];
        }
    if (by_cell)
        return t2::kSigmaCell;
    return by_group ? t2::kSigmaGroup : t2::kSigmaCellGroup;
}

template <typename T>
int build_items(t2_sweeper *s, const double *d_geo, const int4 *d_adj,
                const uint32_t *d_work, const double *d_sig_cell,
                t2::ItemGeo<T> *out_geo)
{
    const int64_t n = s->nc * s->na;
    int *d_err = nullptr;
    T2_TRY(dalloc_t(s, &d_err, sizeof(int)));
    T2_HIP(hipMemset(d_err, 0, sizeof(int)));
    T2_HIP(t2::launch_build_items<T>(
        d_work, n, s->sched.cell_bits, s->nc, s->d_omega, d_geo, d_adj,
        s->d_nbr, d_sig_cell, s->d_lag_key, int64_t(s->sched.lagged.size()),
        out_geo, s->d_item_topo, d_err, s->stream));
    int err = 0;
    T2_HIP(hipMemcpyAsync(&err, d_err, sizeof(int), hipMemcpyDeviceToHost,
                          s->stream));
    T2_HIP(hipStreamSynchronize(s->stream));
    dfree(s, d_err, sizeof(int));
    if (err)
        return fail(T2_ERR_ARG, "a cell has four incoming faces");
    return T2_OK;
}

size_t geo_record_size(const t2_sweeper *s)
{
    return s->precision == T2_FP64 ? sizeof(t2::ItemGeo<double>)
                                   : sizeof(t2::ItemGeo<float>);
}

void free_schedule(t2_sweeper *s)
{
    const size_t n_items = size_t(s->nc) * s->na;
    const int64_t n_lag = int64_t(s->sched.lagged.size());
    drop_graphs(s);
    free_queue(s);
    dfree(s, s->d_item_geo, geo_record_size(s) * n_items);
    dfree(s, s->d_item_geo64, sizeof(t2::ItemGeo<double>) * n_items);
    dfree(s, s->d_item_topo, sizeof(t2::ItemTopo) * n_items);
    dfree(s, s->d_lag_key, sizeof(int64_t) * n_lag);
    dfree(s, s->d_lag_up, sizeof(int32_t) * n_lag);
    dfree(s, s->d_lag_buf, s->tsize() * size_t(n_lag) * 4 * s->ng);
}

// Level sets, lagged faces and item records for the current batch count.
int build_schedule_device(t2_sweeper *s)
{
    const int64_t nc = s->nc, na = s->na, ng = s->ng;
    free_schedule(s);
    if (s->intra >= 0)
        T2_TRY(t2::build_priority_schedule(s->hm, na, s->omega_h.data(),
                                           s->intra, s->inter, s->step_width,
                                           s->order, s->sched, s->dep_tol));
    else
        T2_TRY(t2::build_schedule(s->hm, na, s->omega_h.data(), s->batches,
                                  s->order, s->sched, s->dep_tol));
    const size_t n_items = size_t(nc) * na;
    std::vector<double> geo(size_t(nc) * t2::kGeoStride, 0.0);
    std::vector<int4> adj(nc);
    std::vector<double> sig_cell(nc);
    for (int64_t c = 0; c < nc; c++) {
        double *g = &geo[size_t(c) * t2::kGeoStride];
        for (int i = 0; i < 12; i++)
            g[t2::kGeoNormal + i] = s->hm.normal[12 * c + i];
        for (int k = 0; k < 4; k++)
            g[t2::kGeoArea + k] = s->hm.area[4 * c + k];
        g[t2::kGeoVolume] = s->hm.volume[c];
        adj[c] = make_int4(int(s->hm.adj[4 * c]), int(s->hm.adj[4 * c + 1]),
                           int(s->hm.adj[4 * c + 2]),
                           int(s->hm.adj[4 * c + 3]));
        sig_cell[c] = s->sigma_t_h[c * ng];
    }
    const int64_t n_lag = int64_t(s->sched.lagged.size());
    if (n_lag > 0) {
        std::vector<int64_t> key(n_lag);
        std::vector<int32_t> up(n_lag);
        for (int64_t i = 0; i < n_lag; i++) {
            const auto &l = s->sched.lagged[i];
            key[i] = (int64_t(l.angle) * nc + l.cell) * 4 + l.face;
            up[i] = l.upwind;
        }
        T2_TRY(dalloc_t(s, &s->d_lag_key, sizeof(int64_t) * n_lag));
        T2_TRY(dalloc_t(s, &s->d_lag_up, sizeof(int32_t) * n_lag));
        T2_TRY(dalloc(s, &s->d_lag_buf, s->tsize() * size_t(n_lag) * 4 * ng));
        T2_TRY(upload(s->d_lag_key, key.data(), n_lag));
        T2_TRY(upload(s->d_lag_up, up.data(), n_lag));
    }
    double *d_geo = nullptr, *d_sig_cell = nullptr;
    int4 *d_adj = nullptr;
    uint32_t *d_work = nullptr;
    int rc = T2_OK;
    auto cleanup = [&]() {
        dfree(s, d_geo, sizeof(double) * geo.size());
        dfree(s, d_adj, sizeof(int4) * nc);
        dfree(s, d_work, sizeof(uint32_t) * n_items);
        dfree(s, d_sig_cell, sizeof(double) * nc);
    };
    if ((rc = dalloc_t(s, &d_geo, sizeof(double) * geo.size())) ||
        (rc = dalloc_t(s, &d_adj, sizeof(int4) * nc)) ||
        (rc = dalloc_t(s, &d_work, sizeof(uint32_t) * n_items)) ||
        (rc = dalloc_t(s, &d_sig_cell, sizeof(double) * nc)) ||
        (rc = upload(d_geo, geo.data(), geo.size())) ||
        (rc = upload(d_adj, adj.data(), nc)) ||
        (rc = upload(d_work, s->sched.work.data(), n_items)) ||
        (rc = upload(d_sig_cell, sig_cell.data(), nc)) ||
        (rc = dalloc(s, &s->d_item_geo, geo_record_size(s) * n_items)) ||
        (rc = dalloc_t(s, &s->d_item_topo, sizeof(t2::ItemTopo) * n_items))) {
        cleanup();
        return rc;
    }
    std::vector<uint32_t>().swap(s->sched.work);
    if (s->precision == T2_FP64) {
        rc = build_items<double>(
            s, d_geo, d_adj, d_work, d_sig_cell,
            static_cast<t2::ItemGeo<double> *>(s->d_item_geo));
    } else {
        rc = build_items<float>(
            s, d_geo, d_adj, d_work, d_sig_cell,
            static_cast<t2::ItemGeo<float> *>(s->d_item_geo));
        // FP64 records of the same items, for the FP64 defect
        if (rc == T2_OK)
            rc = dalloc_t(s, &s->d_item_geo64,
                          sizeof(t2::ItemGeo<double>) * n_items);
        if (rc == T2_OK)
            rc = build_items<double>(s, d_geo, d_adj, d_work, d_sig_cell,
                                     s->d_item_geo64);
    }
    cleanup();
    return rc;
}

int setup_device(t2_sweeper *s, const double *weights, const double *sigma_s)
{
    const int64_t nc = s->nc, na = s->na, ng = s->ng;
    const size_t ts = s->tsize();
    const int64_t n_phi = nc * 4 * ng;
    if (n_phi >= (int64_t(1) << 31))
        return fail(T2_ERR_ARG, "cells x 4 x groups must stay below 2^31");
    T2_TRY(dalloc_t(s, &s->d_omega, sizeof(double) * 3 * na));
    T2_TRY(dalloc_t(s, &s->d_wt, sizeof(double) * na));
    T2_TRY(dalloc_t(s, &s->d_sigs, sizeof(double) * nc * ng));
    T2_TRY(upload(s->d_omega, s->omega_h.data(), 3 * na));
    T2_TRY(upload(s->d_wt, weights, na));
    T2_TRY(upload(s->d_sigs, sigma_s, nc * ng));

    const double *sigma_t = s->sigma_t_h.data();
    s->sigma_mode = pick_sigma_mode(nc, ng, sigma_t);
    const int64_t n_sig = s->sigma_mode == t2::kSigmaGroup       ? ng
                          : s->sigma_mode == t2::kSigmaCellGroup ? nc * ng
                                                                 : 1;
    std::vector<char> sig(size_t(n_sig) * ts);
    std::vector<double> sig64(n_sig);
    for (int64_t i = 0; i < n_sig; i++) {
        const double v = s->sigma_mode == t2::kSigmaCell ? 0.0 : sigma_t[i];
        sig64[i] = v;
        if (s->precision == T2_FP64)
            reinterpret_cast<double *>(sig.data())[i] = v;
        else
            reinterpret_cast<float *>(sig.data())[i] = float(v);
    }
    T2_TRY(dalloc(s, &s->d_sigt, sig.size()));
    T2_HIP(hipMemcpy(s->d_sigt, sig.data(), sig.size(),
                     hipMemcpyHostToDevice));
    if (s->precision == T2_FP32) {
        T2_TRY(dalloc_t(s, &s->d_sigt64, sizeof(double) * n_sig));
        T2_TRY(upload(s->d_sigt64, sig64.data(), n_sig));
    }
    T2_TRY(dalloc_t(s, &s->d_nbr, sizeof(uint32_t) * nc));
    T2_TRY(upload(s->d_nbr, s->hm.nbr_pack.data(), nc));
    T2_TRY(build_schedule_device(s));

    T2_TRY(dalloc(s, &s->d_psi, ts * size_t(na) * n_phi));
    T2_TRY(dalloc(s, &s->d_src, ts * size_t(n_phi)));
    T2_TRY(dalloc_t(s, &s->d_phi, sizeof(double) * n_phi));
    T2_TRY(dalloc_t(s, &s->d_phi_old, sizeof(double) * n_phi));
    T2_TRY(dalloc_t(s, &s->d_red, sizeof(double) * t2::kRedBlocks * 64));
    T2_TRY(dalloc_t(s, &s->d_scal, sizeof(double) * 64));
    T2_TRY(zero_psi(s));
    T2_TRY(total_source(s, nullptr, nullptr));
    T2_HIP(t2::launch_fill(s->d_phi, 0.0, n_phi, s->stream));
    T2_HIP(hipStreamSynchronize(s->stream));
    return T2_OK;
}

// ------------------------------------------------------------------ GMRES
// Krylov storage: basis V[0..m] and `extra` work vectors after it.
int krylov_space(t2_sweeper *s, int64_t m, int extra, double **V,
                 double **work)
{
    const int64_t n = s->n_phi();
    const int64_t cols = m + 1 + extra;
    if (s->kry_cols < cols) {
        dfree(s, s->d_kry, sizeof(double) * n * s->kry_cols);
        s->kry_cols = 0;
        T2_TRY(dalloc_t(s, &s->d_kry, sizeof(double) * n * cols));
        s->kry_cols = cols;
    }
    *V = s->d_kry;
    *work = s->d_kry + n * (m + 1);
    return T2_OK;
}

// y = (I - D L^-1 M S) v: the operator of the scalar-flux equations
// (src/SourceIteration.cc:76-114), one sweep with no fixed source.
int transport_apply(t2_sweeper *s, const double *v, double *y)
{
    T2_TRY(total_source(s, nullptr, v));
    T2_TRY(run_sweep(s, nullptr));
    T2_TRY(psi_to_phi(s, y));
    T2_HIP(t2::launch_axpby(1.0, v, -1.0, y, s->n_phi(), s->stream));
    return T2_OK;
}

// Restarted GMRES(m) from x = 0: classical Gram-Schmidt applied twice and
// Givens rotations, all in FP64 on the device; stops at rtol * ||b|| or after
// iter_max matrix applications.  The reference delegates this to PETSc's
// KSPGMRES (src/KrylovSolver.cc:79-169).
//
// Written as a state machine so that one driver can advance many solves
// with one batched sweep per step: next() runs until the iteration needs
// y = A v (returns with *need = true and req_in / req_out set; the caller
// applies the operator and calls next() again) or until it is finished.
struct Gmres {
    enum State { kCycle, kResidual, kStep, kArnoldi, kDone };
    t2_sweeper *s = nullptr;
    const double *b = nullptr;
    double *x = nullptr, *V = nullptr, *hist = nullptr;
    int64_t n = 0, m = 0, iter_max = 0, it = 0, j = 0;
    double bnorm = 0, target = 0, rnorm = 0;
    bool x_zero = true;
    State state = kCycle;
    std::vector<double> H, cs, sn, g, y, h, h2;
    const double *req_in = nullptr;
    double *req_out = nullptr;

    int init(t2_sweeper *s_, const double *b_, double *x_, int64_t m_,
             double rtol, int64_t iter_max_, double *V_, double *hist_)
    {
        s = s_;
        b = b_;
        x = x_;
        m = m_;
        iter_max = iter_max_;
        V = V_;
        hist = hist_;
        n = s->n_phi();
        it = 0;
        j = 0;
        x_zero = true;
        double bb = 0.0;
        T2_TRY(dot(s, b, b, n, &bb));
        bnorm = std::sqrt(bb);
        target = rtol * bnorm;
        rnorm = bnorm;
        T2_HIP(t2::launch_fill(x, 0.0, n, s->stream));
        H.assign(size_t(m + 1) * m, 0.0);
        cs.assign(m, 0.0);
        sn.assign(m, 0.0);
        g.assign(m + 1, 0.0);
        y.assign(m, 0.0);
        h.assign(m + 1, 0.0);
        h2.assign(m + 1, 0.0);
        state = kCycle;
        return T2_OK;
    }

    int end_cycle()
    {
        for (int64_t i = j - 1; i >= 0; i--) {
            double acc = g[i];
            for (int64_t k = i + 1; k < j; k++)
                acc -= H[i * m + k] * y[k];
            y[i] = acc / H[i * m + i];
        }
        for (int64_t i = 0; i < j; i++)
            y[i] = -y[i];
        T2_HIP(hipMemcpyAsync(s->d_scal, y.data(), sizeof(double) * j,
                              hipMemcpyHostToDevice, s->stream));
        T2_HIP(t2::launch_multi_axpy(V, n, int(j), s->d_scal, x, n,
                                     s->stream));
        T2_HIP(hipStreamSynchronize(s->stream));
        x_zero = false;
        state = kCycle;
        return T2_OK;
    }

    int next(bool *need)
    {
        *need = false;
        for (;;) {
            switch (state) {
            case kCycle:
                if (!(it < iter_max && rnorm > target)) {
                    state = kDone;
                    break;
                }
                state = kResidual;
                if (x_zero) {
                    // r = b for x = 0: no operator apply (a whole sweep)
                    T2_HIP(hipMemcpyAsync(V, b, sizeof(double) * n,
                                          hipMemcpyDeviceToDevice,
                                          s->stream));
                    break;
                }
                req_in = x;
                req_out = V;
                *need = true;
                return T2_OK;
            case kResidual: {
                if (!x_zero)  // V = b - A x
                    T2_HIP(t2::launch_axpby(1.0, b, -1.0, V, n, s->stream));
                double beta2 = 0.0;
                T2_TRY(dot(s, V, V, n, &beta2));
                const double beta = std::sqrt(beta2);
                rnorm = beta;
                if (beta <= target || beta == 0.0) {
                    state = kDone;
                    break;
                }
                T2_HIP(t2::launch_scale(V, 1.0 / beta, n, s->stream));
                std::fill(g.begin(), g.end(), 0.0);
                g[0] = beta;
                j = 0;
                state = kStep;
                break;
            }
            case kStep:
                if (!(j < m && it < iter_max)) {
                    T2_TRY(end_cycle());
                    break;
                }
                req_in = V + n * j;
                req_out = V + n * (j + 1);
                state = kArnoldi;
                *need = true;
                return T2_OK;
            case kArnoldi: {
                double *w = V + n * (j + 1);
                for (int pass = 0; pass < 2; pass++) {
                    T2_HIP(t2::launch_multi_dot(V, n, int(j + 1), w, n,
                                                s->d_red, s->d_scal,
                                                s->stream));
                    T2_HIP(t2::launch_multi_axpy(V, n, int(j + 1), s->d_scal,
                                                 w, n, s->stream));
                    T2_TRY(fetch(s, pass ? h2.data() : h.data(), s->d_scal,
                                 j + 1));
                }
                for (int64_t i = 0; i <= j; i++)
                    H[i * m + j] = h[i] + h2[i];
                double ww = 0.0;
                T2_TRY(dot(s, w, w, n, &ww));
                const double hn = std::sqrt(ww);
                H[(j + 1) * m + j] = hn;
                if (hn != 0.0)
                    T2_HIP(t2::launch_scale(w, 1.0 / hn, n, s->stream));
                for (int64_t i = 0; i < j; i++) {
                    const double t =
                        cs[i] * H[i * m + j] + sn[i] * H[(i + 1) * m + j];
                    H[(i + 1) * m + j] =
                        -sn[i] * H[i * m + j] + cs[i] * H[(i + 1) * m + j];
                    H[i * m + j] = t;
                }
                const double hjj = H[j * m + j], hj1 = H[(j + 1) * m + j];
                const double r = std::sqrt(hjj * hjj + hj1 * hj1);
                cs[j] = hjj / r;
                sn[j] = hj1 / r;
                H[j * m + j] = r;
                H[(j + 1) * m + j] = 0.0;
                g[j + 1] = -sn[j] * g[j];
                g[j] = cs[j] * g[j];
                rnorm = std::fabs(g[j + 1]);
                if (hist)
                    hist[it] = rnorm / (bnorm > 0 ? bnorm : 1.0);
                it++;
                j++;
                if (rnorm <= target)
                    T2_TRY(end_cycle());
                else
                    state = kStep;
                break;
            }
            case kDone:
                return T2_OK;
            }
        }
    }

    double rel_res() const { return rnorm / (bnorm > 0 ? bnorm : 1.0); }
    // false when the iteration limit stopped the solve above its target
    bool converged() const { return rnorm <= target; }
};

int gmres(t2_sweeper *s, const double *b, double *x, int64_t m, double rtol,
          int64_t iter_max, double *V, int64_t *n_iter, double *hist,
          double *rel_res, bool *converged)
{
    Gmres gm;
    T2_TRY(gm.init(s, b, x, m, rtol, iter_max, V, hist));
    for (;;) {
        bool need = false;
        T2_TRY(gm.next(&need));
        if (!need)
            break;
        T2_TRY(transport_apply(s, gm.req_in, gm.req_out));
    }
    if (n_iter)
        *n_iter = gm.it;
    if (rel_res)
        *rel_res = gm.rel_res();
    if (converged)
        *converged = gm.converged();
    return T2_OK;
}

// Status of a solve that ran to its end: T2_ERR_NOCONV (with the stage's
// figures in the message) when its iteration limit stopped it above its
// tolerance.  The solution arrays hold the last iterate either way.
int noconv(const std::string &what, int64_t iters, double reached,
           double wanted)
{
    char buf[160];
    std::snprintf(buf, sizeof(buf),
                  ": %lld iterations reached %.3e, above the tolerance %.3e",
                  (long long)iters, reached, wanted);
    return fail(T2_ERR_NOCONV, what + buf);
}

}  // namespace

// ====================================================================== C ABI
extern "C" {

int t2_sweeper_create(t2_sweeper **out, int device, int64_t n_cells,
                      int64_t n_nodes, const double *coords,
                      const int64_t *cell_nodes, const int64_t *adj_cell,
                      const int64_t *f2c, int64_t n_angles,
                      const double *omega, const double *weights,
                      int64_t n_groups, const double *sigma_t,
                      const double *sigma_s, int precision, int gauss)
{
    t2::clear_error();
    if (!out)
        return fail(T2_ERR_ARG, "t2_sweeper_create: out is null");
    *out = nullptr;
    if (n_cells <= 0 || n_nodes <= 0 || n_angles <= 0 || !coords ||
        !cell_nodes || !adj_cell || !f2c || !omega || !weights || !sigma_t ||
        !sigma_s)
        return fail(T2_ERR_ARG, "t2_sweeper_create: bad argument");
    if (n_groups != 1 && n_groups != 2 && n_groups != 4 && n_groups != 8 &&
        n_groups != 16 && n_groups != 32 && n_groups != 64)
        return fail(T2_ERR_ARG, "n_groups must be a power of two <= 64");
    if (precision != T2_FP64 && precision != T2_FP32)
        return fail(T2_ERR_ARG, "precision must be T2_FP64 or T2_FP32");
    if (gauss != T2_GAUSS_NOPIVOT && gauss != T2_GAUSS_ORIGINAL &&
        gauss != T2_GAUSS_CRAMER_GLU && gauss != T2_GAUSS_CRAMER_INTEL)
        return fail(T2_ERR_ARG, "gauss must be one of T2_GAUSS_*");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        return fail(T2_ERR_NODEV, "no HIP device");
    }
    if (device < 0 || device >= ndev)
        return fail(T2_ERR_ARG, "device index out of range");

    auto *s = new (std::nothrow) t2_sweeper();
    if (!s)
        return fail(T2_ERR_NOMEM, "host allocation");
    int rc = t2::build_host_mesh(n_cells, n_nodes, coords, cell_nodes,
                                 adj_cell, f2c, s->hm);
    if (rc) {
        delete s;
        return rc;
    }
    s->device = device;
    s->nc = n_cells;
    s->na = n_angles;
    s->ng = n_groups;
    s->precision = precision;
    s->gauss = gauss;
    s->omega_h.assign(omega, omega + 3 * n_angles);
    s->sigma_t_h.assign(sigma_t, sigma_t + n_cells * n_groups);
    if (const char *e = std::getenv("T2_NO_GRAPH"))
        s->use_graph = !(e[0] == '1');
    if (hipSetDevice(device) != hipSuccess ||
        hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) !=
            hipSuccess) {
        delete s;
        return fail(T2_ERR_HIP, "stream creation failed");
    }
    double *scr = nullptr;   // this device's scratch address, once
    rc = noalloc_scratch(device, &scr);
    if (rc == T2_OK)
        rc = setup_device(s, weights, sigma_s);
    if (rc != T2_OK) {
        std::string msg = t2_last_error();
        t2_sweeper_destroy(s);
        return fail(rc, msg);
    }
    *out = s;
    return T2_OK;
}

int t2_set_alloc_fail(int64_t nth)
{
    t2::clear_error();
    if (nth < 0)
        return fail(T2_ERR_ARG, "t2_set_alloc_fail: nth < 0");
    g_alloc_fail.store(nth);
    return T2_OK;
}

int t2_sweeper_destroy(t2_sweeper *s)
{
    if (!s)
        return T2_OK;
    (void)hipSetDevice(s->device);
    if (s->stream)
        (void)hipStreamSynchronize(s->stream);
    drop_graphs(s);
    for (void *p : s->allocs)
        (void)hipFree(p);
    for (int b = 0; b < 2; b++) {
        if (s->h_pin[b])
            (void)hipHostFree(s->h_pin[b]);
        if (s->pin_ev[b])
            (void)hipEventDestroy(s->pin_ev[b]);
    }
    for (size_t i = 0; i < s->side.size(); i++) {
        (void)hipStreamSynchronize(s->side[i]);
        (void)hipStreamDestroy(s->side[i]);
        (void)hipEventDestroy(s->join_ev[i]);
    }
    if (s->fork_ev)
        (void)hipEventDestroy(s->fork_ev);
    if (s->stream)
        (void)hipStreamDestroy(s->stream);
    delete s;
    return T2_OK;
}

int t2_sweeper_info(const t2_sweeper *s, t2_info *info)
{
    t2::clear_error();
    if (!s || !info)
        return fail(T2_ERR_ARG, "t2_sweeper_info: bad argument");
    info->n_cells = s->nc;
    info->n_angles = s->na;
    info->n_groups = s->ng;
    info->n_steps = s->sched.n_steps;
    info->min_levels = s->sched.min_levels;
    info->n_lagged = int64_t(s->sched.lagged.size());
    info->psi_bytes = int64_t(s->n_psi() * s->tsize());
    info->device_bytes = int64_t(s->bytes);
    info->precision = s->precision;
    info->gauss = s->gauss;
    info->sweeps = s->n_sweeps;
    info->engine = s->engine;
    info->angle_batches = s->batches;
    info->intra_angle_p = s->intra;
    info->inter_angle_p = s->inter;
    info->max_cells_per_step = s->step_width;
    info->concurrent_batches = s->concurrent ? 1 : 0;
    info->items_per_block = int32_t(s->block / (s->ng / s->gpl));
    info->groups_per_lane = s->gpl;
    info->chunk_rounds = s->chunk_rounds;
    info->queue_first_step = int32_t(s->queue_step);
    return T2_OK;
}

int t2_set_launch(t2_sweeper *s, int items_per_block, int groups_per_lane)
{
    t2::clear_error();
    if (!s)
        return fail(T2_ERR_ARG, "null sweeper handle");
    int gpl = groups_per_lane ? groups_per_lane : s->gpl;
    if (gpl != 1 && gpl != 2 && gpl != 4)
        return fail(T2_ERR_ARG, "groups_per_lane must be 1, 2 or 4");
    if (gpl != 1 && s->gauss != T2_GAUSS_NOPIVOT)
        return fail(T2_ERR_ARG, "the pivoting and Cramer cell solvers run "
                                "one group per lane");
    if (gpl > s->ng)
        return fail(T2_ERR_ARG, "groups_per_lane exceeds the group count");
    if (gpl != 1 && s->ipl != 1)
        return fail(T2_ERR_ARG, "several groups per lane need one item per "
                                "lane (t2_set_items_per_lane 1)");
    int block = s->block;
    if (items_per_block) {
        const int64_t lanes = int64_t(items_per_block) * (s->ng / gpl);
        if (lanes < 64 || lanes > t2::kMaxBlock || lanes % 64)
            return fail(T2_ERR_ARG,
                        "items_per_block x lanes per item must be a "
                        "multiple of 64 in [64, 256]");
        block = int(lanes);
    }
    if (block != s->block || gpl != s->gpl) {
        T2_TRY(set_device(s));
        T2_HIP(hipStreamSynchronize(s->stream));
        free_queue(s);
    }
    s->block = block;
    s->gpl = gpl;
    return T2_OK;
}

int t2_set_items_per_lane(t2_sweeper *s, int items)
{
    t2::clear_error();
    T2_TRY(check(s));
    if (items != 1 && items != 2)
        return fail(T2_ERR_ARG, "items per lane must be 1 or 2");
    if (items != 1 && (s->gauss != T2_GAUSS_NOPIVOT || s->gpl != 1))
        return fail(T2_ERR_ARG, "two items per lane need the NoPivot cell "
                                "solver and one group per lane");
    if (items == s->ipl)
        return T2_OK;
    T2_HIP(hipStreamSynchronize(s->stream));
    drop_graphs(s);
    s->ipl = items;
    return T2_OK;
}

int t2_set_order(t2_sweeper *s, int order)
{
    t2::clear_error();
    T2_TRY(check(s));
    if (order != T2_ORDER_CELL && order != T2_ORDER_FACES)
        return fail(T2_ERR_ARG, "order must be T2_ORDER_CELL or "
                                "T2_ORDER_FACES");
    if (order == s->order)
        return T2_OK;
    T2_HIP(hipStreamSynchronize(s->stream));
    s->order = order;
    return build_schedule_device(s);
}

int t2_set_dependency_tolerance(t2_sweeper *s, double tol)
{
    t2::clear_error();
    T2_TRY(check(s));
    if (std::isnan(tol))
        return fail(T2_ERR_ARG, "dependency tolerance is NaN");
    const double rule = tol < 0 ? t2::kTraverseRule : tol;
    if (rule == s->dep_tol)
        return T2_OK;
    T2_HIP(hipStreamSynchronize(s->stream));
    s->dep_tol = rule;
    return build_schedule_device(s);
}

int t2_set_priority(t2_sweeper *s, int intra_angle_p, int inter_angle_p,
                    int64_t max_cells_per_step)
{
    t2::clear_error();
    T2_TRY(check(s));
    if (intra_angle_p < -1 || intra_angle_p > 4 || inter_angle_p < 0 ||
        inter_angle_p > 2 || max_cells_per_step < 0)
        return fail(T2_ERR_ARG, "intra_angle_p must be in [-1, 4], "
                                "inter_angle_p in [0, 2] and "
                                "max_cells_per_step >= 0");
    if (intra_angle_p < 0) {
        inter_angle_p = 0;
        max_cells_per_step = 0;
    }
    if (intra_angle_p == s->intra && inter_angle_p == s->inter &&
        max_cells_per_step == s->step_width)
        return T2_OK;
    T2_HIP(hipStreamSynchronize(s->stream));
    s->intra = intra_angle_p;
    s->inter = inter_angle_p;
    s->step_width = max_cells_per_step;
    return build_schedule_device(s);
}

int t2_set_concurrent_batches(t2_sweeper *s, int concurrent)
{
    t2::clear_error();
    T2_TRY(check(s));
    if ((concurrent != 0) == s->concurrent)
        return T2_OK;
    T2_HIP(hipStreamSynchronize(s->stream));
    drop_graphs(s);
    s->concurrent = concurrent != 0;
    return T2_OK;
}

int t2_set_schedule(t2_sweeper *s, int angle_batches)
{
    t2::clear_error();
    T2_TRY(check(s));
    if (angle_batches < 1 || angle_batches > s->na)
        return fail(T2_ERR_ARG, "angle_batches must be in [1, n_angles]");
    if (angle_batches == s->batches)
        return T2_OK;
    T2_HIP(hipStreamSynchronize(s->stream));
    s->batches = angle_batches;
    return build_schedule_device(s);
}

static int sweep_host_body(t2_sweeper *s, double *psi, const double *source,
                  int zero_psi_bound)
{
    (void)zero_psi_bound;  // no partition faces on one device
    t2::clear_error();
    T2_TRY(check(s));
    if (!psi || !source)
        return fail(T2_ERR_ARG, "t2_sweep_host: null array");
    const int64_t n = s->n_psi();
    const int64_t vg = 4 * s->ng;
    if (!s->d_stage)
        T2_TRY(dalloc_t(s, &s->d_stage, sizeof(double) * n));
    if (!s->d_hostang)
        T2_TRY(dalloc_t(s, &s->d_hostang, sizeof(double) * n));
    T2_TRY(upload_host(s, s->d_stage, source, sizeof(double) * n));
    T2_HIP(t2::launch_ca_to_ac(s->d_stage, s->d_hostang, s->nc, s->na, vg,
                               s->stream));
    if (!s->sched.lagged.empty()) {
        // lagged faces read the caller's previous psi
        if (s->precision != T2_FP64)
            return fail(T2_ERR_STATE,
                        "host sweep with lagged faces needs FP64 storage");
        T2_TRY(upload_host(s, s->d_stage, psi, sizeof(double) * n));
        T2_HIP(t2::launch_ca_to_ac(s->d_stage,
                                   static_cast<double *>(s->d_psi), s->nc,
                                   s->na, vg, s->stream));
    }
    T2_TRY(total_source(s, nullptr, nullptr));
    T2_TRY(run_sweep(s, s->d_hostang));
    if (s->precision == T2_FP64)
        T2_HIP(t2::launch_ac_to_ca<double>(
            static_cast<const double *>(s->d_psi), s->d_stage, s->nc, s->na,
            vg, s->stream));
    else
        T2_HIP(t2::launch_ac_to_ca<float>(static_cast<const float *>(s->d_psi),
                                          s->d_stage, s->nc, s->na, vg,
                                          s->stream));
    return download_host(s, psi, s->d_stage, sizeof(double) * n);
}

int t2_set_source_isotropic(t2_sweeper *s, const double *q)
{
    t2::clear_error();
    T2_TRY(check(s));
    T2_HIP(hipStreamSynchronize(s->stream));
    if (s->d_qang) {
        drop_graphs(s);
        dfree(s, s->d_qang, sizeof(double) * s->n_psi());
    }
    if (!q) {
        dfree(s, s->d_qiso, sizeof(double) * s->n_phi());
        return T2_OK;
    }
    if (!s->d_qiso)
        T2_TRY(dalloc_t(s, &s->d_qiso, sizeof(double) * s->n_phi()));
    T2_TRY(upload(s->d_qiso, q, s->n_phi()));
    return T2_OK;
}

int t2_set_source_angular(t2_sweeper *s, const double *source)
{
    t2::clear_error();
    T2_TRY(check(s));
    T2_TRY(t2_set_source_isotropic(s, nullptr));
    if (!source)
        return T2_OK;
    const int64_t n = s->n_psi();
    T2_TRY(dalloc_t(s, &s->d_qang, sizeof(double) * n));
    if (!s->d_stage)
        T2_TRY(dalloc_t(s, &s->d_stage, sizeof(double) * n));
    T2_TRY(upload_host(s, s->d_stage, source, sizeof(double) * n));
    T2_HIP(t2::launch_ca_to_ac(s->d_stage, s->d_qang, s->nc, s->na,
                               4 * s->ng, s->stream));
    T2_HIP(hipStreamSynchronize(s->stream));
    return T2_OK;
}

int t2_sweep_step(t2_sweeper *s)
{
    t2::clear_error();
    T2_TRY(check(s));
    T2_TRY(total_source(s, s->d_qiso, s->d_phi));
    T2_TRY(run_sweep(s, s->d_qang));
    std::swap(s->d_phi, s->d_phi_old);
    T2_TRY(psi_to_phi(s, s->d_phi));
    return T2_OK;
}

int t2_sweep_only(t2_sweeper *s)
{
    t2::clear_error();
    T2_TRY(check(s));
    return run_sweep(s, s->d_qang);
}

// partial / sums: the L1-change reduction's 2 x kRedBlocks partial sums and
// its two totals (the handle's, or the allocation-free entry point's scratch)
static int solve_si_body(t2_sweeper *s, int64_t iter_max, double err_max,
                         int64_t *n_iter, double *err_hist, double *partial,
                         double *sums_dev)
{
    t2::clear_error();
    T2_TRY(check(s));
    if (iter_max < 0)
        return fail(T2_ERR_ARG, "iter_max < 0");
    T2_TRY(zero_psi(s));
    T2_HIP(t2::launch_fill(s->d_phi, 0.0, s->n_phi(), s->stream));
    int64_t it = 0;
    double err = 1.0;
    while (it < iter_max && err > err_max) {
        T2_TRY(t2_sweep_step(s));
        T2_HIP(t2::launch_l1_change(s->d_phi, s->d_phi_old, s->n_phi(),
                                    partial, sums_dev, s->stream));
        double sums[2];
        T2_TRY(fetch(s, sums, sums_dev, 2));
        err = sums[0] / sums[1];
        if (err_hist)
            err_hist[it] = err;
        it++;
    }
    if (n_iter)
        *n_iter = it;
    if (err > err_max)
        return noconv("source iteration", it, err, err_max);
    return T2_OK;
}

static int solve_krylov_body(t2_sweeper *s, int64_t iter_max, double rtol,
                    int64_t restart, int64_t *n_iter, double *res_hist)
{
    t2::clear_error();
    T2_TRY(check(s));
    if (restart < 1 || restart > 60 || iter_max < 0)
        return fail(T2_ERR_ARG, "restart must be in [1, 60]");
    double *V = nullptr, *work = nullptr;
    T2_TRY(krylov_space(s, restart, 1, &V, &work));
    double *b = work;
    T2_TRY(zero_psi_before_sweep(s));
    // b = D L^-1 q
    T2_TRY(total_source(s, s->d_qiso, nullptr));
    T2_TRY(run_sweep(s, s->d_qang));
    T2_TRY(psi_to_phi(s, b));
    bool conv = true;
    int64_t its = 0;
    double rel = 0.0;
    T2_TRY(gmres(s, b, s->d_phi, restart, rtol, iter_max, V, &its, res_hist,
                 &rel, &conv));
    if (n_iter)
        *n_iter = its;
    // psi = L^-1 (q + M S phi)
    T2_TRY(total_source(s, s->d_qiso, s->d_phi));
    T2_TRY(run_sweep(s, s->d_qang));
    T2_HIP(hipStreamSynchronize(s->stream));
    if (!conv)
        return noconv("GMRES", its, rel, rtol);
    return T2_OK;
}

static int solve_refined_body(t2_sweeper *s, int64_t iter_max, double rtol_sweep,
                     double rtol_correction, int64_t restart,
                     t2_refine_info *info, double *phi1)
{
    t2::clear_error();
    T2_TRY(check(s));
    if (s->precision != T2_FP32)
        return fail(T2_ERR_STATE, "the refined solve needs an FP32 handle");
    if (restart < 1 || restart > 60 || iter_max < 0)
        return fail(T2_ERR_ARG, "restart must be in [1, 60]");
    const int64_t n = s->n_phi();
    double *V = nullptr, *work = nullptr;
    T2_TRY(krylov_space(s, restart, 2, &V, &work));
    double *b = work, *y = work + n;
    if (!s->d_res)
        T2_TRY(dalloc_t(s, &s->d_res, sizeof(float) * s->n_psi()));
    if (!s->d_src64)
        T2_TRY(dalloc_t(s, &s->d_src64, sizeof(double) * n));
    t2_refine_info st{};
    const int64_t sweeps0 = s->n_sweeps;
    T2_TRY(zero_psi_before_sweep(s));

    // 1. FP32-sweep GMRES on phi, then psi32 and phi1 = D psi32
    T2_TRY(total_source(s, s->d_qiso, nullptr));
    T2_TRY(run_sweep(s, s->d_qang));
    T2_TRY(psi_to_phi(s, b));
    bool conv1 = true, conv2 = true;
    T2_TRY(gmres(s, b, s->d_phi, restart, rtol_sweep, iter_max, V,
                 &st.iters_sweep, nullptr, &st.residual_sweep, &conv1));
    T2_TRY(total_source(s, s->d_qiso, s->d_phi));
    T2_TRY(run_sweep(s, s->d_qang));
    double *phi_1 = s->d_phi_old;
    T2_TRY(psi_to_phi(s, phi_1));
    if (phi1)
        T2_TRY(fetch(s, phi1, phi_1, n));

    // 2. FP64 defect r = q + M S phi1 - L psi32
    T2_HIP(t2::launch_total_source<double>(s->d_qiso, s->d_sigs, phi_1,
                                           s->nc, int(s->ng), s->d_src64,
                                           s->stream));
    t2::ResidualArgs<float> ra{};
    ra.geo = s->d_item_geo64;
    ra.topo = s->d_item_topo;
    ra.n_items = s->nc * s->na;
    ra.nc = s->nc;
    ra.sigma_mode = s->sigma_mode;
    ra.sigma = s->d_sigt64;
    ra.src_iso = s->d_src64;
    ra.src_ang = s->d_qang;
    ra.lag_up = s->d_lag_up;
    ra.psi = static_cast<const float *>(s->d_psi);
    ra.r = s->d_res;
    T2_HIP(t2::launch_residual<float>(ra, int(s->ng), s->stream));

    // 3. correction: GMRES on (I - D L^-1 M S) D delta = D L^-1 r
    T2_TRY(total_source(s, nullptr, nullptr));
    T2_TRY(run_sweep(s, nullptr, s->d_res));
    T2_TRY(psi_to_phi(s, b));
    T2_TRY(gmres(s, b, y, restart, rtol_correction, iter_max, V,
                 &st.iters_correction, nullptr, &st.residual_correction,
                 &conv2));
    // 4. psi to FP64 accuracy, psi = psi32 + dpsi.  The correction's
    // sweeps used the psi buffer, so psi32 = L^-1 (q + M S phi) is swept
    // again from step 1's phi (bitwise the psi32 of the defect); then
    // dpsi = L^-1 (r + M S D delta), y being D delta to the correction's
    // tolerance, is swept in FP32 over the residual in place (each item reads
    // its own r before it stores its dpsi, and upwind reads see dpsi).
    T2_TRY(total_source(s, s->d_qiso, s->d_phi));
    T2_TRY(run_sweep(s, s->d_qang));
    T2_TRY(total_source(s, nullptr, y));
    T2_TRY(run_sweep(s, nullptr, s->d_res, s->d_res));
    s->psi_corrected = true;
    // phi = phi1 + y
    double corr = 0.0, base = 0.0;
    T2_TRY(dot(s, y, y, n, &corr));
    T2_HIP(t2::launch_axpby(1.0, phi_1, 1.0, y, n, s->stream));
    T2_HIP(hipMemcpyAsync(s->d_phi, y, sizeof(double) * n,
                          hipMemcpyDeviceToDevice, s->stream));
    T2_TRY(dot(s, s->d_phi, s->d_phi, n, &base));
    st.correction_rel = base > 0 ? std::sqrt(corr / base) : 0.0;
    T2_HIP(hipStreamSynchronize(s->stream));
    st.sweeps = s->n_sweeps - sweeps0;
    if (info)
        *info = st;
    if (!conv1)
        return noconv("refined solve, FP32-sweep GMRES stage", st.iters_sweep,
                      st.residual_sweep, rtol_sweep);
    if (!conv2)
        return noconv("refined solve, correction GMRES stage",
                      st.iters_correction, st.residual_correction,
                      rtol_correction);
    return T2_OK;
}

int t2_get_phi(const t2_sweeper *s, double *phi)
{
    t2::clear_error();
    T2_TRY(check(s));
    if (!phi)
        return fail(T2_ERR_ARG, "null phi");
    T2_HIP(hipStreamSynchronize(s->stream));
    T2_HIP(hipMemcpy(phi, s->d_phi, sizeof(double) * s->n_phi(),
                     hipMemcpyDeviceToHost));
    return check_abort(const_cast<t2_sweeper *>(s));
}

int t2_set_phi(t2_sweeper *s, const double *phi)
{
    t2::clear_error();
    T2_TRY(check(s));
    if (!phi)
        return fail(T2_ERR_ARG, "null phi");
    T2_HIP(hipStreamSynchronize(s->stream));
    T2_HIP(hipMemcpy(s->d_phi, phi, sizeof(double) * s->n_phi(),
                     hipMemcpyHostToDevice));
    return T2_OK;
}

int t2_get_psi(const t2_sweeper *s_const, double *psi)
{
    t2::clear_error();
    t2_sweeper *s = const_cast<t2_sweeper *>(s_const);
    T2_TRY(check(s));
    if (!psi)
        return fail(T2_ERR_ARG, "null psi");
    const int64_t n = s->n_psi();
    if (!s->d_stage)
        T2_TRY(dalloc_t(s, &s->d_stage, sizeof(double) * n));
    if (s->precision == T2_FP64)
        T2_HIP(t2::launch_ac_to_ca<double>(
            static_cast<const double *>(s->d_psi), s->d_stage, s->nc, s->na,
            4 * s->ng, s->stream));
    else   // after a refined solve: psi32 + its FP32 correction, in FP64
        T2_HIP(t2::launch_ac_to_ca<float>(
            static_cast<const float *>(s->d_psi), s->d_stage, s->nc, s->na,
            4 * s->ng, s->stream, s->psi_corrected ? s->d_res : nullptr));
    return download_host(s, psi, s->d_stage, sizeof(double) * n);
}

int t2_defect(t2_sweeper *s, const double *psi, double *density)
{
    t2::clear_error();
    T2_TRY(check(s));
    if (!density)
        return fail(T2_ERR_ARG, "t2_defect: null density");
    const int64_t n = s->n_psi();
    const int64_t vg = 4 * s->ng;
    const bool fp64 = s->precision == T2_FP64;
    if (!s->d_stage)
        T2_TRY(dalloc_t(s, &s->d_stage, sizeof(double) * n));
    if (!s->d_src64)
        T2_TRY(dalloc_t(s, &s->d_src64, sizeof(double) * s->n_phi()));
    if (fp64 && !s->d_hostang)
        T2_TRY(dalloc_t(s, &s->d_hostang, sizeof(double) * n));
    if (!fp64 && !s->d_res)
        T2_TRY(dalloc_t(s, &s->d_res, sizeof(float) * n));
    if (!fp64)   // the residual buffer is overwritten
        s->psi_corrected = false;
    if (psi) {
        T2_TRY(upload_host(s, s->d_stage, psi, sizeof(double) * n));
        if (fp64)
            T2_HIP(t2::launch_ca_to_ac(s->d_stage,
                                       static_cast<double *>(s->d_psi), s->nc,
                                       s->na, vg, s->stream));
        else
            T2_HIP(t2::launch_ca_to_ac(s->d_stage,
                                       static_cast<float *>(s->d_psi), s->nc,
                                       s->na, vg, s->stream));
    }
    // q + M S phi in FP64, then the operator apply of the refined solve
    T2_HIP(t2::launch_total_source<double>(s->d_qiso, s->d_sigs, s->d_phi,
                                           s->nc, int(s->ng), s->d_src64,
                                           s->stream));
    if (fp64) {
        t2::ResidualArgs<double> ra{};
        ra.geo = static_cast<const t2::ItemGeo<double> *>(s->d_item_geo);
        ra.topo = s->d_item_topo;
        ra.n_items = s->nc * s->na;
        ra.nc = s->nc;
        ra.sigma_mode = s->sigma_mode;
        ra.sigma = static_cast<const double *>(s->d_sigt);
        ra.src_iso = s->d_src64;
        ra.src_ang = s->d_qang;
        ra.lag_up = s->d_lag_up;
        ra.psi = static_cast<const double *>(s->d_psi);
        ra.r = s->d_hostang;
        T2_HIP(t2::launch_residual<double>(ra, int(s->ng), s->stream));
        T2_HIP(t2::launch_ac_to_ca<double>(s->d_hostang, s->d_stage, s->nc,
                                           s->na, vg, s->stream));
    } else {
        t2::ResidualArgs<float> ra{};
        ra.geo = s->d_item_geo64;
        ra.topo = s->d_item_topo;
        ra.n_items = s->nc * s->na;
        ra.nc = s->nc;
        ra.sigma_mode = s->sigma_mode;
        ra.sigma = s->d_sigt64;
        ra.src_iso = s->d_src64;
        ra.src_ang = s->d_qang;
        ra.lag_up = s->d_lag_up;
        ra.psi = static_cast<const float *>(s->d_psi);
        ra.r = s->d_res;
        T2_HIP(t2::launch_residual<float>(ra, int(s->ng), s->stream));
        T2_HIP(t2::launch_ac_to_ca<float>(s->d_res, s->d_stage, s->nc, s->na,
                                          vg, s->stream));
    }
    return download_host(s, density, s->d_stage, sizeof(double) * n);
}

static int krylov_vectors(t2_sweeper *s)
{
    if (!s->d_kvec)
        T2_TRY(dalloc_t(s, &s->d_kvec, 2 * sizeof(double) * s->n_phi()));
    return T2_OK;
}

static int krylov_apply_body(t2_sweeper *s, const double *x, double *b)
{
    t2::clear_error();
    T2_TRY(check(s));
    if (!x || !b)
        return fail(T2_ERR_ARG, "t2_krylov_apply: null vector");
    const int64_t n = s->n_phi();
    T2_TRY(krylov_vectors(s));
    double *dx = s->d_kvec, *dy = s->d_kvec + n;
    T2_HIP(hipStreamSynchronize(s->stream));
    T2_HIP(hipMemcpy(dx, x, sizeof(double) * n, hipMemcpyHostToDevice));
    T2_TRY(transport_apply(s, dx, dy));
    return fetch(s, b, dy, n);
}

static int krylov_rhs_body(t2_sweeper *s, double *b)
{
    t2::clear_error();
    T2_TRY(check(s));
    if (!b)
        return fail(T2_ERR_ARG, "t2_krylov_rhs: null vector");
    const int64_t n = s->n_phi();
    T2_TRY(krylov_vectors(s));
    T2_TRY(total_source(s, s->d_qiso, nullptr));
    T2_TRY(run_sweep(s, s->d_qang));
    T2_TRY(psi_to_phi(s, s->d_kvec));
    return fetch(s, b, s->d_kvec, n);
}

int t2_krylov_apply(t2_sweeper *s, const double *x, double *b)
{
    return after_queue(s, krylov_apply_body(s, x, b));
}

int t2_krylov_apply_device(t2_sweeper *s, const double *x, double *b)
{
    t2::clear_error();
    T2_TRY(check(s));
    if (!x || !b)
        return fail(T2_ERR_ARG, "t2_krylov_apply_device: null vector");
    if (x == b)
        return fail(T2_ERR_ARG, "t2_krylov_apply_device: x and b alias");
    const int64_t n = s->n_phi();
    if ((x < b && x + n > b) || (b < x && b + n > x))
        return fail(T2_ERR_ARG, "t2_krylov_apply_device: x and b overlap");
    return transport_apply(s, x, b);
}

int t2_krylov_rhs(t2_sweeper *s, double *b)
{
    return after_queue(s, krylov_rhs_body(s, b));
}

int t2_get_psi_angle(const t2_sweeper *s, int64_t angle, double *psi_a)
{
    t2::clear_error();
    T2_TRY(check(s));
    if (!psi_a || angle < 0 || angle >= s->na)
        return fail(T2_ERR_ARG, "t2_get_psi_angle: bad argument");
    const int64_t n = s->n_phi();
    T2_HIP(hipStreamSynchronize(s->stream));
    if (s->precision == T2_FP64) {
        T2_HIP(hipMemcpy(psi_a, static_cast<const double *>(s->d_psi) + angle * n,
                         sizeof(double) * n, hipMemcpyDeviceToHost));
    } else {
        std::vector<float> tmp(n), corr;
        T2_HIP(hipMemcpy(tmp.data(),
                         static_cast<const float *>(s->d_psi) + angle * n,
                         sizeof(float) * n, hipMemcpyDeviceToHost));
        if (s->psi_corrected) {   // psi32 + its correction, in FP64
            corr.resize(n);
            T2_HIP(hipMemcpy(corr.data(), s->d_res + angle * n,
                             sizeof(float) * n, hipMemcpyDeviceToHost));
        }
        for (int64_t i = 0; i < n; i++)
            psi_a[i] = double(tmp[i]) + (corr.empty() ? 0.0 : double(corr[i]));
    }
    return T2_OK;
}

int t2_synchronize(t2_sweeper *s)
{
    t2::clear_error();
    T2_TRY(check(s));
    T2_HIP(hipStreamSynchronize(s->stream));
    return check_abort(s);
}

int t2_set_engine(t2_sweeper *s, int engine, int chunk_rounds,
                  int64_t spin_limit)
{
    t2::clear_error();
    T2_TRY(check(s));
    if (engine != T2_ENGINE_LEVELS && engine != T2_ENGINE_QUEUE &&
        engine != T2_ENGINE_HYBRID)
        return fail(T2_ERR_ARG, "engine must be T2_ENGINE_LEVELS, "
                                "T2_ENGINE_QUEUE or T2_ENGINE_HYBRID");
    if (spin_limit < 0)
        return fail(T2_ERR_ARG, "spin_limit < 0");
    if (chunk_rounds < 0 || chunk_rounds > 256)
        return fail(T2_ERR_ARG, "chunk_rounds must be in [0, 256]");
    if (engine != T2_ENGINE_LEVELS && s->gauss != T2_GAUSS_NOPIVOT)
        return fail(T2_ERR_ARG, "the ready-queue engines use the NoPivot "
                                "cell solver");
    const int rounds = chunk_rounds ? chunk_rounds : 4;
    T2_HIP(hipStreamSynchronize(s->stream));
    if (engine != s->engine || spin_limit != s->spin_limit)
        drop_graphs(s);
    if (rounds != s->chunk_rounds)
        free_queue(s);
    s->engine = engine;
    s->chunk_rounds = rounds;
    s->spin_limit = spin_limit;
    return T2_OK;
}

int t2_set_queue_split(t2_sweeper *s, int64_t first_step)
{
    t2::clear_error();
    T2_TRY(check(s));
    if (first_step < -1)
        return fail(T2_ERR_ARG, "first_step must be >= -1");
    if (first_step != s->split) {
        T2_HIP(hipStreamSynchronize(s->stream));
        drop_graphs(s);
    }
    s->split = first_step;
    return T2_OK;
}

int t2_sweep_host(t2_sweeper *s, double *psi, const double *source,
                  int zero_psi_bound)
{
    return after_queue(s, sweep_host_body(s, psi, source, zero_psi_bound));
}

int t2_solve_source_iteration(t2_sweeper *s, int64_t iter_max, double err_max,
                              int64_t *n_iter, double *err_hist)
{
    if (!s)
        return solve_si_body(s, iter_max, err_max, n_iter, err_hist, nullptr,
                             nullptr);
    return after_queue(s, solve_si_body(s, iter_max, err_max, n_iter,
                                        err_hist, s->d_red, s->d_scal));
}

int t2_solve_source_iteration_noalloc(t2_sweeper *s, int64_t iter_max,
                                      double err_max, int64_t *n_iter,
                                      double *err_hist)
{
    t2::clear_error();
    T2_TRY(check(s));
    double *scr = nullptr;
    T2_TRY(noalloc_scratch(s->device, &scr));
    // the partial sums and totals are written in full by every reduction
    // before they are read: nothing in the scratch needs clearing
    s->no_alloc = true;
    const int rc = solve_si_body(s, iter_max, err_max, n_iter, err_hist, scr,
                                 scr + 2 * t2::kRedBlocks);
    s->no_alloc = false;
    return after_queue(s, rc);
}

int t2_solve_krylov(t2_sweeper *s, int64_t iter_max, double rtol,
                    int64_t restart, int64_t *n_iter, double *res_hist)
{
    return after_queue(s, solve_krylov_body(s, iter_max, rtol, restart,
                                            n_iter, res_hist));
}

int t2_solve_refined(t2_sweeper *s, int64_t iter_max, double rtol_sweep,
                     double rtol_correction, int64_t restart,
                     t2_refine_info *info, double *phi1)
{
    return after_queue(s, solve_refined_body(s, iter_max, rtol_sweep,
                                             rtol_correction, restart, info,
                                             phi1));
}

int t2_get_stream(const t2_sweeper *s, void **stream)
{
    t2::clear_error();
    if (!s || !stream)
        return fail(T2_ERR_ARG, "t2_get_stream: bad argument");
    *stream = static_cast<void *>(s->stream);
    return T2_OK;
}

}  // extern "C"

// ====================================================================== batch
namespace {

// Members with one group count, swept by one merged level schedule.
struct LevelGroup {
    int ng = 0;
    std::vector<int> idx;                // member indices
    int32_t *d_active = nullptr;         // active flags in group order
    std::vector<int64_t> step_off;       // merged schedule
    int64_t n_items = 0;
    void *d_geo = nullptr;
    t2::ItemTopo *d_topo = nullptr;
    void *d_desc = nullptr;
};

}  // namespace

struct t2_batch {
    std::vector<t2_sweeper *> members;
    int device = 0;
    int precision = T2_FP64;
    hipStream_t stream = nullptr;
    hipEvent_t order = nullptr;          // orders member and batch streams
    int engine = T2_ENGINE_LEVELS;
    int concurrent = 0;                  // level form: chains per group count
                                         // (0: all launches on one stream)
    int block = t2::kMaxBlock;           // level form: threads per block
    std::vector<hipStream_t> side;       // their extra streams
    std::vector<hipEvent_t> join_ev;
    hipEvent_t fork_ev = nullptr;
    int chunk_rounds = 4;
    int64_t spin_limit = 0;
    int64_t n_steps = 0, n_items = 0;
    // level form (built on first use)
    bool levels_built = false;
    std::vector<LevelGroup> groups;
    // queue form (built on first use)
    int64_t n_chunks = -1, grid = 0;
    void *d_problems = nullptr;
    t2::BatchChunk *d_chunks = nullptr;
    unsigned long long *d_ticket = nullptr;
    int32_t *d_abort = nullptr;
    std::vector<const void *> ang_t;     // members' T-typed angular sources
    int32_t *d_active = nullptr;         // members swept by the next sweep
    std::vector<int32_t> active;         // host copy (kept alive for copies)
    std::vector<int32_t> uploaded;       // flags the device holds
    hipGraphExec_t graph = nullptr;
    // correction sweeps of the refined solve: each member's sweep writes its
    // residual buffer in place (its angle-dependent source) and lagged faces
    // read zero; the tables point there while corr is set, and the sweep is
    // captured in a graph of its own
    bool corr = false;
    hipGraphExec_t graph_corr = nullptr;
    std::vector<std::vector<const void *>> snap;  // members' buffers
};

namespace {

std::vector<const void *> member_buffers(const t2_sweeper *m)
{
    return {m->d_item_geo, m->d_item_topo, m->d_psi, m->d_src, m->d_lag_buf,
            m->d_lag_key, m->d_sigt, m->d_qang, m->d_pending, m->d_down,
            reinterpret_cast<const void *>(intptr_t(m->sched.n_steps))};
}

void batch_snapshot(t2_batch *b)
{
    b->snap.clear();
    for (auto *m : b->members)
        b->snap.push_back(member_buffers(m));
}

// A member whose schedule, buffers or source kind changed since the batch
// last built its tables would be swept with stale records.
int batch_check(const t2_batch *b)
{
    if (!b)
        return fail(T2_ERR_ARG, "null batch");
    for (size_t i = 0; i < b->members.size(); i++)
        if (member_buffers(b->members[i]) != b->snap[i])
            return fail(T2_ERR_STATE,
                        "batch member " + std::to_string(i) +
                            " changed its schedule or source since the "
                            "batch was built");
    return T2_OK;
}

int batch_fail_cleanup(t2_batch *b, int rc)
{
    std::string msg = t2_last_error();
    t2_batch_destroy(b);
    return fail(rc, msg);
}

void free_levels(t2_batch *b)
{
    for (auto &g : b->groups) {
        if (g.d_active)
            (void)hipFree(g.d_active);
        if (g.d_geo)
            (void)hipFree(g.d_geo);
        if (g.d_topo)
            (void)hipFree(g.d_topo);
        if (g.d_desc)
            (void)hipFree(g.d_desc);
    }
    b->groups.clear();
    b->levels_built = false;
}

void free_batch_queue(t2_batch *b)
{
    if (b->d_problems)
        (void)hipFree(b->d_problems);
    if (b->d_chunks)
        (void)hipFree(b->d_chunks);
    b->d_problems = nullptr;
    b->d_chunks = nullptr;
    b->n_chunks = -1;
}

void drop_batch_graph(t2_batch *b)
{
    if (b->graph)
        (void)hipGraphExecDestroy(b->graph);
    if (b->graph_corr)
        (void)hipGraphExecDestroy(b->graph_corr);
    b->graph = nullptr;
    b->graph_corr = nullptr;
}

// Angular flux a member's batched sweep writes.
void *member_psi_out(const t2_batch *b, const t2_sweeper *m)
{
    return b->corr ? static_cast<void *>(m->d_res) : m->d_psi;
}

template <typename T>
t2::ProblemDesc<T> problem_desc(const t2_sweeper *m, const void *ang_t,
                                void *psi_out)
{
    t2::ProblemDesc<T> d{};
    d.nc = m->nc;
    d.sigma_mode = m->sigma_mode;
    d.sigma = static_cast<const T *>(m->d_sigt);
    d.src_iso = static_cast<const T *>(m->d_src);
    d.src_ang_t = static_cast<const T *>(ang_t);
    d.psi = static_cast<T *>(psi_out);
    d.lag_buf = static_cast<const T *>(m->d_lag_buf);
    return d;
}

template <typename T>
t2::BatchProblem<T> batch_problem(const t2_sweeper *m, const void *ang_t,
                                  void *psi_out)
{
    t2::BatchProblem<T> bp{};
    bp.a.geo = static_cast<const t2::ItemGeo<T> *>(m->d_item_geo);
    bp.a.topo = m->d_item_topo;
    bp.a.n_items = m->nc * m->na;
    bp.a.nc = m->nc;
    bp.a.sigma_mode = m->sigma_mode;
    bp.a.sigma = static_cast<const T *>(m->d_sigt);
    bp.a.src_iso = static_cast<const T *>(m->d_src);
    bp.a.src_ang_t = static_cast<const T *>(ang_t);
    bp.a.psi = static_cast<T *>(psi_out);
    bp.a.lag_buf = static_cast<const T *>(m->d_lag_buf);
    bp.pending = m->d_pending;
    bp.down = m->d_down;
    bp.arm = m->d_arm;
    bp.first_item = m->queue_item;
    bp.groups = int32_t(m->ng);
    return bp;
}

// Rewrites the device problem tables that exist (after a change of the
// members' angle-dependent sources); the captured sweeps read them.
template <typename T>
int write_tables(t2_batch *b)
{
    for (auto &g : b->groups) {
        if (!g.d_desc)
            continue;
        std::vector<t2::ProblemDesc<T>> desc(g.idx.size());
        for (size_t j = 0; j < g.idx.size(); j++)
            desc[j] = problem_desc<T>(
                b->members[g.idx[j]], b->ang_t[g.idx[j]],
                member_psi_out(b, b->members[g.idx[j]]));
        T2_HIP(hipMemcpy(g.d_desc, desc.data(),
                         sizeof(t2::ProblemDesc<T>) * desc.size(),
                         hipMemcpyHostToDevice));
    }
    if (b->d_problems) {
        std::vector<t2::BatchProblem<T>> probs(b->members.size());
        for (size_t i = 0; i < b->members.size(); i++)
            probs[i] = batch_problem<T>(b->members[i], b->ang_t[i],
                                        member_psi_out(b, b->members[i]));
        T2_HIP(hipMemcpy(b->d_problems, probs.data(),
                         sizeof(t2::BatchProblem<T>) * probs.size(),
                         hipMemcpyHostToDevice));
    }
    return T2_OK;
}

// Sets every member's angle-dependent source, and whether the following
// sweeps are correction sweeps (into the residual buffers).
int batch_set_ang_t(t2_batch *b, const std::vector<const void *> &ang_t,
                    bool corr = false)
{
    if (ang_t == b->ang_t && corr == b->corr)
        return T2_OK;
    T2_HIP(hipStreamSynchronize(b->stream));
    b->ang_t = ang_t;
    b->corr = corr;
    return b->precision == T2_FP64 ? write_tables<double>(b)
                                   : write_tables<float>(b);
}

// Level form: members grouped by group count; each group's schedules merged
// step by step into one record stream with a problem table.
template <typename T>
int build_level_group(t2_batch *b, LevelGroup &g)
{
    int64_t n_steps = 0;
    for (int i : g.idx)
        n_steps = std::max(n_steps, b->members[i]->sched.n_steps);
    std::vector<int64_t> count(n_steps, 0);
    for (int i : g.idx) {
        const t2_sweeper *m = b->members[i];
        for (int64_t st = 0; st < m->sched.n_steps; st++)
            count[st] += m->sched.step_off[st + 1] - m->sched.step_off[st];
    }
    g.step_off.assign(n_steps + 1, 0);
    for (int64_t st = 0; st < n_steps; st++)
        g.step_off[st + 1] = g.step_off[st] + count[st];
    g.n_items = g.step_off[n_steps];
    const size_t n = size_t(std::max<int64_t>(1, g.n_items));
    T2_HIP(hipMalloc(&g.d_geo, sizeof(t2::ItemGeo<T>) * n));
    T2_HIP(hipMalloc(reinterpret_cast<void **>(&g.d_topo),
                     sizeof(t2::ItemTopo) * n));
    std::vector<int64_t> filled(n_steps, 0);
    std::vector<t2::ProblemDesc<T>> desc(g.idx.size());
    int64_t *d_off = nullptr;
    T2_HIP(hipMalloc(reinterpret_cast<void **>(&d_off),
                     sizeof(int64_t) * 2 * size_t(n_steps + 1)));
    for (size_t j = 0; j < g.idx.size(); j++) {
        t2_sweeper *m = b->members[g.idx[j]];
        const int64_t ms = m->sched.n_steps;
        std::vector<int64_t> dst(ms + 1, 0);
        for (int64_t st = 0; st < ms; st++) {
            dst[st] = g.step_off[st] + filled[st];
            filled[st] += m->sched.step_off[st + 1] - m->sched.step_off[st];
        }
        hipError_t e = hipMemcpy(d_off, m->sched.step_off.data(),
                                 sizeof(int64_t) * (ms + 1),
                                 hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(d_off + ms + 1, dst.data(),
                          sizeof(int64_t) * (ms + 1), hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = t2::launch_merge_items<T>(
                static_cast<const t2::ItemGeo<T> *>(m->d_item_geo),
                m->d_item_topo, d_off, d_off + ms + 1, ms,
                m->sched.step_off[ms], int32_t(j),
                static_cast<t2::ItemGeo<T> *>(g.d_geo), g.d_topo, b->stream);
        if (e == hipSuccess)
            e = hipStreamSynchronize(b->stream);
        if (e != hipSuccess) {
            (void)hipFree(d_off);
            return fail(T2_ERR_HIP, std::string("batch merge: ") +
                                        hipGetErrorString(e));
        }
        desc[j] = problem_desc<T>(m, b->ang_t[g.idx[j]],
                                  member_psi_out(b, m));
    }
    (void)hipFree(d_off);
    T2_HIP(hipMalloc(reinterpret_cast<void **>(&g.d_active),
                     sizeof(int32_t) * g.idx.size()));
    T2_HIP(hipMalloc(&g.d_desc, sizeof(t2::ProblemDesc<T>) * desc.size()));
    T2_HIP(hipMemcpy(g.d_desc, desc.data(),
                     sizeof(t2::ProblemDesc<T>) * desc.size(),
                     hipMemcpyHostToDevice));
    return T2_OK;
}

int build_levels(t2_batch *b)
{
    free_levels(b);
    std::vector<int> ngs;
    for (auto *m : b->members)
        if (std::find(ngs.begin(), ngs.end(), int(m->ng)) == ngs.end())
            ngs.push_back(int(m->ng));
    std::sort(ngs.begin(), ngs.end());
    // one merged schedule per group count, or, for parallel chains, up to
    // `concurrent` per group count with the members dealt out by work
    // (largest first, each to the lightest chain)
    const int chains = std::max(1, b->concurrent);
    for (int ng : ngs) {
        std::vector<int> idx;
        for (size_t i = 0; i < b->members.size(); i++)
            if (b->members[i]->ng == ng)
                idx.push_back(int(i));
        auto work = [&](int i) {
            return b->members[i]->nc * b->members[i]->na * b->members[i]->ng;
        };
        std::stable_sort(idx.begin(), idx.end(),
                         [&](int x, int y) { return work(x) > work(y); });
        const size_t nch = std::min<size_t>(size_t(chains), idx.size());
        std::vector<LevelGroup> part(nch);
        std::vector<int64_t> load(nch, 0);
        for (int i : idx) {
            const size_t c = size_t(
                std::min_element(load.begin(), load.end()) - load.begin());
            part[c].idx.push_back(i);
            load[c] += work(i);
        }
        for (auto &g : part) {
            g.ng = ng;
            std::sort(g.idx.begin(), g.idx.end());
            b->groups.push_back(g);
        }
    }
    for (auto &g : b->groups)
        T2_TRY(b->precision == T2_FP64 ? build_level_group<double>(b, g)
                                       : build_level_group<float>(b, g));
    b->levels_built = true;
    return T2_OK;
}

// Queue form: every member's pending counters and readers table, one
// problem table, and chunks interleaved step by step over the members.
template <typename T>
int build_batch_queue(t2_batch *b)
{
    free_batch_queue(b);
    const int nm = int(b->members.size());
    std::vector<t2::BatchProblem<T>> probs(nm);
    for (int i = 0; i < nm; i++) {
        t2_sweeper *m = b->members[i];
        if (m->n_chunks < 0 || m->queue_step != 0) {
            T2_HIP(hipStreamSynchronize(m->stream));
            T2_TRY(build_queue(m, 0));
        }
        probs[i] = batch_problem<T>(m, b->ang_t[i], member_psi_out(b, m));
    }
    // chunks of full size while a merged step gives every resident block
    // one; thinner steps are cut finer (down to one wavefront round of a
    // member's items), measured in lanes (items x groups) over all members
    T2_HIP(t2::queue_batch_grid<T>(&b->grid));
    const int64_t resident = std::max<int64_t>(1, b->grid);
    std::vector<t2::BatchChunk> chunks;
    for (int64_t st = 0; st < b->n_steps; st++) {
        int64_t lanes = 0;
        for (const t2_sweeper *m : b->members)
            if (st < m->sched.n_steps)
                lanes += (m->sched.step_off[st + 1] - m->sched.step_off[st]) *
                         m->ng;
        const int64_t fair = (lanes + resident - 1) / resident;
        for (int i = 0; i < nm; i++) {
            const t2_sweeper *m = b->members[i];
            if (st >= m->sched.n_steps)
                continue;
            const int64_t per = int64_t(t2::kMaxBlock / m->ng) *
                                b->chunk_rounds;
            const int64_t per_wave = m->intra < 0 ? 64 / m->ng : per;
            const int64_t c_n = std::max(
                per_wave, std::min(per, (fair + m->ng - 1) / m->ng));
            const int64_t lo = m->sched.step_off[st];
            const int64_t hi = m->sched.step_off[st + 1];
            for (int64_t c = lo; c < hi; c += c_n)
                chunks.push_back({c, int32_t(std::min(c_n, hi - c)), i});
        }
    }
    T2_HIP(hipMalloc(&b->d_problems, sizeof(t2::BatchProblem<T>) * nm));
    T2_HIP(hipMemcpy(b->d_problems, probs.data(),
                     sizeof(t2::BatchProblem<T>) * nm, hipMemcpyHostToDevice));
    T2_HIP(hipMalloc(reinterpret_cast<void **>(&b->d_chunks),
                     sizeof(t2::BatchChunk) * std::max<size_t>(1, chunks.size())));
    T2_HIP(hipMemcpy(b->d_chunks, chunks.data(),
                     sizeof(t2::BatchChunk) * chunks.size(),
                     hipMemcpyHostToDevice));
    if (!b->d_ticket) {
        T2_HIP(hipMalloc(reinterpret_cast<void **>(&b->d_ticket),
                         sizeof(unsigned long long)));
        T2_HIP(hipMalloc(reinterpret_cast<void **>(&b->d_abort),
                         sizeof(int32_t)));
        T2_HIP(hipMemset(b->d_abort, 0, sizeof(int32_t)));
    }
    b->n_chunks = int64_t(chunks.size());
    return T2_OK;
}

template <typename T>
int batch_enqueue(t2_batch *b)
{
    for (auto *m : b->members) {
        const int64_t n_lag = int64_t(m->sched.lagged.size());
        if (n_lag > 0 && !b->corr)
            T2_HIP(t2::launch_lag_gather<T>(
                m->d_lag_key, m->d_lag_up, n_lag, m->nc, int(m->ng),
                m->d_nbr, static_cast<const T *>(m->d_psi),
                static_cast<T *>(m->d_lag_buf), b->stream));
        else if (n_lag > 0)   // a correction's previous value is zero
            T2_HIP(t2::launch_fill(static_cast<double *>(m->d_lag_buf), 0.0,
                                   int64_t(sizeof(T)) * n_lag * 4 * m->ng / 8,
                                   b->stream));
    }
    if (b->engine == T2_ENGINE_QUEUE) {
        t2::QueueState q{};
        q.n_chunks = b->n_chunks;
        q.ticket = b->d_ticket;
        q.abort = b->d_abort;
        q.spin_limit = b->spin_limit > 0 ? b->spin_limit : kDefaultSpinLimit;
        q.active = b->d_active;
        T2_HIP(t2::launch_fill(reinterpret_cast<double *>(b->d_ticket), 0.0,
                               1, b->stream));
        T2_HIP(t2::launch_queue_batch<T>(
            static_cast<const t2::BatchProblem<T> *>(b->d_problems),
            b->d_chunks, q, b->grid, b->stream));
        return T2_OK;
    }
    auto launch = [&](LevelGroup &g, int64_t st, hipStream_t stream) {
        t2::SweepArgs<T> a{};
        a.problems = static_cast<const t2::ProblemDesc<T> *>(g.d_desc);
        a.active = g.d_active;
        a.geo = static_cast<const t2::ItemGeo<T> *>(g.d_geo) + g.step_off[st];
        a.topo = g.d_topo + g.step_off[st];
        a.n_items = g.step_off[st + 1] - g.step_off[st];
        T2_HIP(t2::launch_sweep_step<T>(a, g.ng, 1, b->block, stream));
        return T2_OK;
    };
    const size_t ng = b->groups.size();
    if (b->concurrent > 0 && ng > 1) {
        // merged schedules share nothing: group g > 0 runs its merged steps
        // on a stream of its own (a parallel branch of the captured graph),
        // forked from and joined to the batch's
        T2_HIP(hipEventRecord(b->fork_ev, b->stream));
        for (size_t i = 1; i < ng; i++) {
            LevelGroup &g = b->groups[i];
            T2_HIP(hipStreamWaitEvent(b->side[i - 1], b->fork_ev, 0));
            for (int64_t st = 0; st + 1 < int64_t(g.step_off.size()); st++)
                T2_TRY(launch(g, st, b->side[i - 1]));
            T2_HIP(hipEventRecord(b->join_ev[i - 1], b->side[i - 1]));
        }
        LevelGroup &g0 = b->groups[0];
        for (int64_t st = 0; st + 1 < int64_t(g0.step_off.size()); st++)
            T2_TRY(launch(g0, st, b->stream));
        for (size_t i = 1; i < ng; i++)
            T2_HIP(hipStreamWaitEvent(b->stream, b->join_ev[i - 1], 0));
        return T2_OK;
    }
    int64_t longest = 0;
    for (auto &g : b->groups)
        longest = std::max<int64_t>(longest, int64_t(g.step_off.size()) - 1);
    for (int64_t st = 0; st < longest; st++)
        for (auto &g : b->groups) {
            if (st + 1 >= int64_t(g.step_off.size()))
                continue;
            T2_TRY(launch(g, st, b->stream));
        }
    return T2_OK;
}

// Work already queued on the members' streams precedes the batch's work...
int batch_after_members(t2_batch *b)
{
    for (auto *m : b->members) {
        T2_HIP(hipEventRecord(b->order, m->stream));
        T2_HIP(hipStreamWaitEvent(b->stream, b->order, 0));
    }
    return T2_OK;
}

// ...and later work on the members' streams follows it.
int members_after_batch(t2_batch *b)
{
    T2_HIP(hipEventRecord(b->order, b->stream));
    for (auto *m : b->members)
        T2_HIP(hipStreamWaitEvent(m->stream, b->order, 0));
    return T2_OK;
}

int batch_sweep(t2_batch *b)
{
    if (b->engine == T2_ENGINE_QUEUE && b->n_chunks < 0) {
        drop_batch_graph(b);
        T2_HIP(hipStreamSynchronize(b->stream));
        T2_TRY(b->precision == T2_FP64 ? build_batch_queue<double>(b)
                                       : build_batch_queue<float>(b));
        batch_snapshot(b);
    }
    if (b->engine == T2_ENGINE_LEVELS && !b->levels_built) {
        // one launch per merged step and group count
        std::vector<std::pair<int64_t, int64_t>> longest;  // (groups, steps)
        for (auto *m : b->members) {
            auto it = std::find_if(
                longest.begin(), longest.end(),
                [&](const std::pair<int64_t, int64_t> &x) {
                    return x.first == m->ng;
                });
            if (it == longest.end())
                longest.push_back({m->ng, m->sched.n_steps});
            else
                it->second = std::max(it->second, m->sched.n_steps);
        }
        int64_t launches = 0;
        for (auto &x : longest)
            launches += x.second;
        T2_TRY(check_step_launches(launches));
        drop_batch_graph(b);
        T2_HIP(hipStreamSynchronize(b->stream));
        T2_TRY(build_levels(b));
        batch_snapshot(b);
        b->uploaded.clear();
    }
    // active flags (all members unless a solve masks some), uploaded when
    // they change; the kernels of the captured sweep read them
    for (size_t i = 0; i < b->members.size(); i++)
        if (b->active[i])
            b->members[i]->n_sweeps++;
    if (b->active != b->uploaded) {
        T2_HIP(hipMemcpyAsync(b->d_active, b->active.data(),
                              sizeof(int32_t) * b->active.size(),
                              hipMemcpyHostToDevice, b->stream));
        for (auto &g : b->groups) {
            std::vector<int32_t> ga(g.idx.size());
            for (size_t j = 0; j < g.idx.size(); j++)
                ga[j] = b->active[g.idx[j]];
            T2_HIP(hipMemcpyAsync(g.d_active, ga.data(),
                                  sizeof(int32_t) * ga.size(),
                                  hipMemcpyHostToDevice, b->stream));
            T2_HIP(hipStreamSynchronize(b->stream));
        }
        T2_HIP(hipStreamSynchronize(b->stream));
        b->uploaded = b->active;
    }
    if (b->concurrent > 0 && b->engine == T2_ENGINE_LEVELS) {
        // streams and events of the parallel branches, outside any capture
        if (!b->fork_ev)
            T2_HIP(hipEventCreateWithFlags(&b->fork_ev,
                                           hipEventDisableTiming));
        while (b->side.size() + 1 < b->groups.size()) {
            hipStream_t st = nullptr;
            hipEvent_t ev = nullptr;
            T2_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
            b->side.push_back(st);
            T2_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
            b->join_ev.push_back(ev);
        }
    }
    hipGraphExec_t &exec = b->corr ? b->graph_corr : b->graph;
    if (!exec) {
        hipGraph_t g = nullptr;
        T2_HIP(hipStreamBeginCapture(b->stream,
                                     hipStreamCaptureModeThreadLocal));
        int rc = b->precision == T2_FP64 ? batch_enqueue<double>(b)
                                         : batch_enqueue<float>(b);
        hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (rc != T2_OK) {
            if (g)
                (void)hipGraphDestroy(g);
            return rc;
        }
        T2_HIP(e);
        e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        T2_HIP(e);
    }
    T2_HIP(hipGraphLaunch(exec, b->stream));
    if (!b->corr)   // the members' psi are replaced
        for (size_t i = 0; i < b->members.size(); i++)
            if (b->active[i])
                b->members[i]->psi_corrected = false;
    return T2_OK;
}

// After an abandoned batched ready-queue sweep: reports it once and re-arms
// every member's counters.
int batch_check_abort(t2_batch *b)
{
    if (!b->d_abort)
        return T2_OK;
    int32_t flag = 0;
    T2_HIP(hipStreamSynchronize(b->stream));
    T2_HIP(hipMemcpy(&flag, b->d_abort, sizeof(flag), hipMemcpyDeviceToHost));
    if (!flag)
        return T2_OK;
    for (auto *m : b->members)
        if (m->d_pending)
            T2_HIP(t2::launch_queue_rearm(m->d_item_topo, m->nc * m->na,
                                          m->queue_item, m->d_arm,
                                          m->d_pending, b->stream));
    T2_HIP(hipMemsetAsync(b->d_abort, 0, sizeof(int32_t), b->stream));
    T2_HIP(hipStreamSynchronize(b->stream));
    return fail(T2_ERR_TIMEOUT,
                "a ready-queue wait ran past its poll limit; the batched "
                "sweep was abandoned and its results are invalid");
}

}  // namespace

extern "C" {

int t2_batch_create(t2_batch **out, t2_sweeper *const *members, int count)
{
    t2::clear_error();
    if (!out || !members || count < 1)
        return fail(T2_ERR_ARG, "t2_batch_create: bad argument");
    *out = nullptr;
    for (int i = 0; i < count; i++) {
        const t2_sweeper *m = members[i];
        if (!m)
            return fail(T2_ERR_ARG, "t2_batch_create: null member");
        if (m->device != members[0]->device ||
            m->precision != members[0]->precision)
            return fail(T2_ERR_ARG, "batch members need the same device and "
                                    "precision");
        if (m->d_qang)
            return fail(T2_ERR_ARG, "batch members need isotropic sources");
        if (m->gauss != T2_GAUSS_NOPIVOT)
            return fail(T2_ERR_ARG, "batch members use the NoPivot cell "
                                    "solver");
        for (int j = 0; j < i; j++)
            if (members[j] == m)
                return fail(T2_ERR_ARG, "a handle appears twice in a batch");
    }
    auto *b = new (std::nothrow) t2_batch();
    if (!b)
        return fail(T2_ERR_NOMEM, "host allocation");
    b->members.assign(members, members + count);
    b->device = members[0]->device;
    b->precision = members[0]->precision;
    for (auto *m : b->members) {
        b->n_steps = std::max(b->n_steps, m->sched.n_steps);
        b->n_items += m->nc * m->na;
    }
    if (hipSetDevice(b->device) != hipSuccess ||
        hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) !=
            hipSuccess) {
        delete b;
        return fail(T2_ERR_HIP, "stream creation failed");
    }
    if (hipEventCreateWithFlags(&b->order, hipEventDisableTiming) !=
        hipSuccess)
        return batch_fail_cleanup(b, fail(T2_ERR_HIP, "event creation"));
    b->active.assign(count, 1);
    b->ang_t.assign(count, nullptr);
    if (hipMalloc(reinterpret_cast<void **>(&b->d_active),
                  sizeof(int32_t) * count) != hipSuccess)
        return batch_fail_cleanup(b, fail(T2_ERR_NOMEM, "batch flags"));
    for (auto *m : b->members)
        if (hipStreamSynchronize(m->stream) != hipSuccess)
            return batch_fail_cleanup(b, fail(T2_ERR_HIP, "member sync"));
    batch_snapshot(b);
    *out = b;
    return T2_OK;
}

int t2_batch_destroy(t2_batch *b)
{
    if (!b)
        return T2_OK;
    (void)hipSetDevice(b->device);
    if (b->stream)
        (void)hipStreamSynchronize(b->stream);
    drop_batch_graph(b);
    free_levels(b);
    free_batch_queue(b);
    if (b->d_ticket)
        (void)hipFree(b->d_ticket);
    if (b->d_abort)
        (void)hipFree(b->d_abort);
    if (b->d_active)
        (void)hipFree(b->d_active);
    if (b->order)
        (void)hipEventDestroy(b->order);
    for (size_t i = 0; i < b->side.size(); i++) {
        (void)hipStreamSynchronize(b->side[i]);
        (void)hipStreamDestroy(b->side[i]);
        (void)hipEventDestroy(b->join_ev[i]);
    }
    if (b->fork_ev)
        (void)hipEventDestroy(b->fork_ev);
    if (b->stream)
        (void)hipStreamDestroy(b->stream);
    delete b;
    return T2_OK;
}

int t2_batch_set_launch(t2_batch *b, int block)
{
    t2::clear_error();
    if (!b)
        return fail(T2_ERR_ARG, "null batch");
    if (block < 64 || block > t2::kMaxBlock || block % 64)
        return fail(T2_ERR_ARG, "block must be a multiple of 64 in [64, " +
                                    std::to_string(t2::kMaxBlock) + "]");
    if (block == b->block)
        return T2_OK;
    T2_HIP(hipSetDevice(b->device));
    T2_HIP(hipStreamSynchronize(b->stream));
    drop_batch_graph(b);
    b->block = block;
    return T2_OK;
}

int t2_batch_set_concurrent(t2_batch *b, int chains)
{
    t2::clear_error();
    if (!b)
        return fail(T2_ERR_ARG, "null batch");
    if (chains < 0 || chains > 64)
        return fail(T2_ERR_ARG, "chains must be in [0, 64]");
    if (chains == b->concurrent)
        return T2_OK;
    T2_HIP(hipSetDevice(b->device));
    T2_HIP(hipStreamSynchronize(b->stream));
    drop_batch_graph(b);
    // the member split of the merged schedules follows the chain count
    if (std::max(1, chains) != std::max(1, b->concurrent)) {
        free_levels(b);
        b->levels_built = false;
    }
    b->concurrent = chains;
    return T2_OK;
}

int t2_batch_set_engine(t2_batch *b, int engine, int chunk_rounds,
                        int64_t spin_limit)
{
    t2::clear_error();
    if (!b)
        return fail(T2_ERR_ARG, "null batch");
    if (engine != T2_ENGINE_LEVELS && engine != T2_ENGINE_QUEUE)
        return fail(T2_ERR_ARG, "engine must be T2_ENGINE_LEVELS or "
                                "T2_ENGINE_QUEUE");
    if (spin_limit < 0)
        return fail(T2_ERR_ARG, "spin_limit < 0");
    if (chunk_rounds < 0 || chunk_rounds > 256)
        return fail(T2_ERR_ARG, "chunk_rounds must be in [0, 256]");
    const int rounds = chunk_rounds ? chunk_rounds : 4;
    T2_HIP(hipSetDevice(b->device));
    T2_HIP(hipStreamSynchronize(b->stream));
    if (engine != b->engine || spin_limit != b->spin_limit ||
        rounds != b->chunk_rounds)
        drop_batch_graph(b);
    if (rounds != b->chunk_rounds)
        free_batch_queue(b);
    b->engine = engine;
    b->chunk_rounds = rounds;
    b->spin_limit = spin_limit;
    return T2_OK;
}

int t2_batch_sweep_only(t2_batch *b)
{
    t2::clear_error();
    T2_TRY(batch_check(b));
    T2_HIP(hipSetDevice(b->device));
    T2_TRY(batch_after_members(b));
    T2_TRY(batch_sweep(b));
    return members_after_batch(b);
}

int t2_batch_sweep_step(t2_batch *b)
{
    t2::clear_error();
    T2_TRY(batch_check(b));
    T2_HIP(hipSetDevice(b->device));
    T2_TRY(batch_after_members(b));
    for (auto *m : b->members) {
        if (b->precision == T2_FP64)
            T2_HIP(t2::launch_total_source<double>(
                m->d_qiso, m->d_sigs, m->d_phi, m->nc, int(m->ng),
                static_cast<double *>(m->d_src), b->stream));
        else
            T2_HIP(t2::launch_total_source<float>(
                m->d_qiso, m->d_sigs, m->d_phi, m->nc, int(m->ng),
                static_cast<float *>(m->d_src), b->stream));
    }
    T2_TRY(batch_sweep(b));
    for (auto *m : b->members) {
        std::swap(m->d_phi, m->d_phi_old);
        if (b->precision == T2_FP64)
            T2_HIP(t2::launch_psi_to_phi<double>(
                static_cast<const double *>(m->d_psi), m->d_wt, m->na,
                m->n_phi(), m->d_phi, b->stream));
        else
            T2_HIP(t2::launch_psi_to_phi<float>(
                static_cast<const float *>(m->d_psi), m->d_wt, m->na,
                m->n_phi(), m->d_phi, b->stream));
    }
    return members_after_batch(b);
}

// One batched sweep of the members flagged in act (their total sources set
// on their own streams beforehand).
static int batch_sweep_some(t2_batch *b, const std::vector<int32_t> &act)
{
    b->active = act;
    T2_TRY(batch_after_members(b));
    int rc = batch_sweep(b);
    std::fill(b->active.begin(), b->active.end(), 1);
    T2_TRY(rc);
    return members_after_batch(b);
}

static double seconds_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() -
                                         t0)
        .count();
}

// Advances the members' GMRES solves in lock step: every unfinished member
// runs to its next operator apply y = (I - D L^-1 M S) v, one batched sweep
// serves all the applies of a round, and a member whose GMRES needs no more
// applies drops out of the sweeps.  done_at[i] (may be null) receives the
// time from t0 at which member i finished.
static int batch_gmres(t2_batch *b, std::vector<Gmres> &gm,
                       std::chrono::steady_clock::time_point t0,
                       double *done_at)
{
    const size_t nm = b->members.size();
    std::vector<int32_t> act(nm, 0);
    std::vector<bool> done(nm, false);
    for (;;) {
        bool any = false;
        for (size_t i = 0; i < nm; i++) {
            act[i] = 0;
            if (done[i])
                continue;
            bool need = false;
            T2_TRY(gm[i].next(&need));
            if (need) {
                act[i] = 1;
                any = true;
                T2_TRY(total_source(b->members[i], nullptr, gm[i].req_in));
            } else {
                done[i] = true;
                if (done_at)
                    done_at[i] = seconds_since(t0);
            }
        }
        if (!any)
            return T2_OK;
        T2_TRY(batch_sweep_some(b, act));
        for (size_t i = 0; i < nm; i++) {
            if (!act[i])
                continue;
            t2_sweeper *m = b->members[i];
            T2_TRY(psi_to_phi(m, gm[i].req_out));
            T2_HIP(t2::launch_axpby(1.0, gm[i].req_in, -1.0, gm[i].req_out,
                                    m->n_phi(), m->stream));
        }
    }
}

int t2_batch_solve_krylov(t2_batch *b, int64_t iter_max, double rtol,
                          int64_t restart, int64_t *n_iter, double *seconds)
{
    t2::clear_error();
    T2_TRY(batch_check(b));
    if (restart < 1 || restart > 60 || iter_max < 0)
        return fail(T2_ERR_ARG, "restart must be in [1, 60]");
    T2_HIP(hipSetDevice(b->device));
    const auto t0 = std::chrono::steady_clock::now();
    const size_t nm = b->members.size();
    std::vector<double *> V(nm), rhs(nm);
    for (size_t i = 0; i < nm; i++) {
        t2_sweeper *m = b->members[i];
        double *work = nullptr;
        T2_TRY(krylov_space(m, restart, 1, &V[i], &work));
        rhs[i] = work;
        T2_TRY(zero_psi_before_sweep(m));
        T2_TRY(total_source(m, m->d_qiso, nullptr));  // b = D L^-1 q
    }
    std::vector<int32_t> act(nm, 1);
    T2_TRY(batch_sweep_some(b, act));
    std::vector<Gmres> gm(nm);
    for (size_t i = 0; i < nm; i++) {
        t2_sweeper *m = b->members[i];
        T2_TRY(psi_to_phi(m, rhs[i]));
        T2_TRY(gm[i].init(m, rhs[i], m->d_phi, restart, rtol, iter_max, V[i],
                          nullptr));
    }
    std::vector<double> done_at(nm, 0.0);
    T2_TRY(batch_gmres(b, gm, t0, done_at.data()));
    for (size_t i = 0; i < nm; i++) {
        if (n_iter)
            n_iter[i] = gm[i].it;
        if (seconds)
            seconds[i] = done_at[i];
    }
    // psi = L^-1 (q + M S phi) for every member
    for (auto *m : b->members)
        T2_TRY(total_source(m, m->d_qiso, m->d_phi));
    T2_TRY(batch_sweep_some(b, act));
    for (auto *m : b->members)
        T2_HIP(hipStreamSynchronize(m->stream));
    T2_TRY(batch_check_abort(b));
    for (size_t i = 0; i < nm; i++)
        if (!gm[i].converged())
            return noconv("batched GMRES, member " + std::to_string(i),
                          gm[i].it, gm[i].rel_res(), rtol);
    return T2_OK;
}

// The steps of the one-handle refined solve for every member, each stage in
// lock step; per member the device work is the same sequence of operations,
// so results are bitwise those of t2_solve_refined on the member alone.
static int batch_refined_body(t2_batch *b, int64_t iter_max,
                              double rtol_sweep, double rtol_correction,
                              int64_t restart, t2_refine_info *info,
                              double *seconds)
{
    const auto t0 = std::chrono::steady_clock::now();
    const size_t nm = b->members.size();
    std::vector<double *> V(nm), rhs(nm), y(nm);
    std::vector<int64_t> sweeps0(nm);
    for (size_t i = 0; i < nm; i++) {
        t2_sweeper *m = b->members[i];
        const int64_t n = m->n_phi();
        double *work = nullptr;
        T2_TRY(krylov_space(m, restart, 2, &V[i], &work));
        rhs[i] = work;
        y[i] = work + n;
        if (!m->d_res)
            T2_TRY(dalloc_t(m, &m->d_res, sizeof(float) * m->n_psi()));
        if (!m->d_src64)
            T2_TRY(dalloc_t(m, &m->d_src64, sizeof(double) * n));
        sweeps0[i] = m->n_sweeps;
        T2_TRY(zero_psi_before_sweep(m));
        T2_TRY(total_source(m, m->d_qiso, nullptr));
    }
    std::vector<int32_t> all(nm, 1);
    std::vector<t2_refine_info> st(nm);
    std::vector<Gmres> gm(nm);

    // 1. FP32-sweep GMRES on phi, then psi32 and phi1 = D psi32
    T2_TRY(batch_sweep_some(b, all));
    for (size_t i = 0; i < nm; i++) {
        t2_sweeper *m = b->members[i];
        T2_TRY(psi_to_phi(m, rhs[i]));
        T2_TRY(gm[i].init(m, rhs[i], m->d_phi, restart, rtol_sweep, iter_max,
                          V[i], nullptr));
    }
    T2_TRY(batch_gmres(b, gm, t0, nullptr));
    std::vector<bool> conv1(nm);
    for (size_t i = 0; i < nm; i++) {
        t2_sweeper *m = b->members[i];
        st[i].iters_sweep = gm[i].it;
        st[i].residual_sweep = gm[i].rel_res();
        conv1[i] = gm[i].converged();
        T2_TRY(total_source(m, m->d_qiso, m->d_phi));
    }
    T2_TRY(batch_sweep_some(b, all));

    // 2. FP64 defect r = q + M S phi1 - L psi32 of each member
    std::vector<const void *> res(nm);
    for (size_t i = 0; i < nm; i++) {
        t2_sweeper *m = b->members[i];
        double *phi_1 = m->d_phi_old;
        T2_TRY(psi_to_phi(m, phi_1));
        T2_HIP(t2::launch_total_source<double>(m->d_qiso, m->d_sigs, phi_1,
                                               m->nc, int(m->ng), m->d_src64,
                                               m->stream));
        t2::ResidualArgs<float> ra{};
        ra.geo = m->d_item_geo64;
        ra.topo = m->d_item_topo;
        ra.n_items = m->nc * m->na;
        ra.nc = m->nc;
        ra.sigma_mode = m->sigma_mode;
        ra.sigma = m->d_sigt64;
        ra.src_iso = m->d_src64;
        ra.src_ang = nullptr;
        ra.lag_up = m->d_lag_up;
        ra.psi = static_cast<const float *>(m->d_psi);
        ra.r = m->d_res;
        T2_HIP(t2::launch_residual<float>(ra, int(m->ng), m->stream));
        res[i] = m->d_res;
        T2_TRY(total_source(m, nullptr, nullptr));
    }

    // 3. correction: GMRES on (I - D L^-1 M S) D delta = D L^-1 r
    T2_TRY(batch_set_ang_t(b, res));
    T2_TRY(batch_sweep_some(b, all));
    T2_TRY(batch_set_ang_t(b, std::vector<const void *>(nm, nullptr)));
    for (size_t i = 0; i < nm; i++) {
        t2_sweeper *m = b->members[i];
        T2_TRY(psi_to_phi(m, rhs[i]));
        T2_TRY(gm[i].init(m, rhs[i], y[i], restart, rtol_correction,
                          iter_max, V[i], nullptr));
    }
    std::vector<double> done_at(nm, 0.0);
    T2_TRY(batch_gmres(b, gm, t0, done_at.data()));
    for (size_t i = 0; i < nm; i++) {
        t2_sweeper *m = b->members[i];
        st[i].iters_correction = gm[i].it;
        st[i].residual_correction = gm[i].rel_res();
        T2_TRY(total_source(m, m->d_qiso, m->d_phi));
    }
    // 4. psi32 swept again from step 1's phi, then every member's
    // correction dpsi = L^-1 (r + M S D delta) over its residual in place
    T2_TRY(batch_sweep_some(b, all));
    for (size_t i = 0; i < nm; i++) {
        T2_TRY(total_source(b->members[i], nullptr, y[i]));
        res[i] = b->members[i]->d_res;
    }
    T2_TRY(batch_set_ang_t(b, res, true));
    T2_TRY(batch_sweep_some(b, all));
    T2_TRY(batch_set_ang_t(b, std::vector<const void *>(nm, nullptr)));
    bool converged = true;
    for (size_t i = 0; i < nm; i++) {
        t2_sweeper *m = b->members[i];
        const int64_t n = m->n_phi();
        m->psi_corrected = true;
        // phi = phi1 + y
        double corr = 0.0, base = 0.0;
        T2_TRY(dot(m, y[i], y[i], n, &corr));
        T2_HIP(t2::launch_axpby(1.0, m->d_phi_old, 1.0, y[i], n, m->stream));
        T2_HIP(hipMemcpyAsync(m->d_phi, y[i], sizeof(double) * n,
                              hipMemcpyDeviceToDevice, m->stream));
        T2_TRY(dot(m, m->d_phi, m->d_phi, n, &base));
        st[i].correction_rel = base > 0 ? std::sqrt(corr / base) : 0.0;
        T2_HIP(hipStreamSynchronize(m->stream));
        st[i].sweeps = m->n_sweeps - sweeps0[i];
        converged = converged && conv1[i] && gm[i].converged();
        if (info)
            info[i] = st[i];
        if (seconds)
            seconds[i] = done_at[i];
    }
    if (!converged)
        return fail(T2_ERR_NOCONV,
                    "batched refined solve: a member's GMRES stage reached "
                    "its iteration limit above its tolerance (see its "
                    "residual_sweep / residual_correction)");
    return T2_OK;
}

int t2_batch_solve_refined(t2_batch *b, int64_t iter_max, double rtol_sweep,
                           double rtol_correction, int64_t restart,
                           t2_refine_info *info, double *seconds)
{
    t2::clear_error();
    T2_TRY(batch_check(b));
    if (b->precision != T2_FP32)
        return fail(T2_ERR_STATE, "the refined solve needs FP32 members");
    if (restart < 1 || restart > 60 || iter_max < 0)
        return fail(T2_ERR_ARG, "restart must be in [1, 60]");
    T2_HIP(hipSetDevice(b->device));
    int rc = batch_refined_body(b, iter_max, rtol_sweep, rtol_correction,
                                restart, info, seconds);
    if (rc != T2_OK) {
        // later sweeps must not see a residual source left in the tables
        const std::string msg = t2_last_error();
        (void)hipStreamSynchronize(b->stream);
        b->ang_t.assign(b->members.size(), nullptr);
        b->corr = false;
        (void)(b->precision == T2_FP64 ? write_tables<double>(b)
                                       : write_tables<float>(b));
        if (rc == T2_ERR_NOCONV) {   // results are whole; a timeout wins
            const int a = batch_check_abort(b);
            return a != T2_OK ? a : fail(rc, msg);
        }
        return fail(rc, msg);
    }
    return batch_check_abort(b);
}

int t2_batch_solve_source_iteration(t2_batch *b, int64_t iter_max,
                                    double err_max, int64_t *n_iter,
                                    double *seconds)
{
    t2::clear_error();
    T2_TRY(batch_check(b));
    if (iter_max < 0)
        return fail(T2_ERR_ARG, "iter_max < 0");
    T2_HIP(hipSetDevice(b->device));
    const auto t0 = std::chrono::steady_clock::now();
    const size_t nm = b->members.size();
    if (n_iter)
        std::fill(n_iter, n_iter + nm, 0);
    if (seconds)
        std::fill(seconds, seconds + nm, 0.0);
    std::vector<int64_t> it(nm, 0);
    std::vector<double> err(nm, 1.0);
    for (auto *m : b->members) {
        T2_TRY(zero_psi(m));
        T2_HIP(t2::launch_fill(m->d_phi, 0.0, m->n_phi(), m->stream));
    }
    std::vector<int32_t> act(nm, 0);
    for (;;) {
        bool any = false;
        for (size_t i = 0; i < nm; i++) {
            t2_sweeper *m = b->members[i];
            act[i] = it[i] < iter_max && err[i] > err_max;
            if (act[i]) {
                any = true;
                T2_TRY(total_source(m, m->d_qiso, m->d_phi));
            }
        }
        if (!any)
            break;
        T2_TRY(batch_sweep_some(b, act));
        for (size_t i = 0; i < nm; i++) {
            if (!act[i])
                continue;
            t2_sweeper *m = b->members[i];
            std::swap(m->d_phi, m->d_phi_old);
            T2_TRY(psi_to_phi(m, m->d_phi));
            T2_HIP(t2::launch_l1_change(m->d_phi, m->d_phi_old, m->n_phi(),
                                        m->d_red, m->d_scal, m->stream));
            double sums[2];
            T2_TRY(fetch(m, sums, m->d_scal, 2));
            err[i] = sums[0] / sums[1];
            it[i]++;
            if (!(it[i] < iter_max && err[i] > err_max)) {
                if (n_iter)
                    n_iter[i] = it[i];
                if (seconds)
                    seconds[i] = seconds_since(t0);
            }
        }
    }
    T2_TRY(batch_check_abort(b));
    for (size_t i = 0; i < nm; i++)
        if (err[i] > err_max)
            return noconv("batched source iteration, member " +
                              std::to_string(i),
                          it[i], err[i], err_max);
    return T2_OK;
}

int t2_batch_info(const t2_batch *b, int64_t *n_steps, int64_t *n_items)
{
    t2::clear_error();
    if (!b)
        return fail(T2_ERR_ARG, "null batch");
    if (n_steps)
        *n_steps = b->n_steps;
    if (n_items)
        *n_items = b->n_items;
    return T2_OK;
}

int t2_batch_synchronize(t2_batch *b)
{
    t2::clear_error();
    if (!b)
        return fail(T2_ERR_ARG, "null batch");
    T2_HIP(hipSetDevice(b->device));
    T2_HIP(hipStreamSynchronize(b->stream));
    return batch_check_abort(b);
}

int t2_batch_get_stream(const t2_batch *b, void **stream)
{
    t2::clear_error();
    if (!b || !stream)
        return fail(T2_ERR_ARG, "t2_batch_get_stream: bad argument");
    *stream = static_cast<void *>(b->stream);
    return T2_OK;
}

}  // extern "C"
