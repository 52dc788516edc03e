// C ABI entry points (include/ktplace_amd.h): context, netlist upload,
// positions, SparseSystem access, solve wrapper and kernel timers.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "kt_internal.h"
#include "lal_dev.h"

namespace kt {

// A roctx range named after the reference's phase timers (ScopedTimer in
// simpl_placer.cc:411-520: simpl-init, simpl-build, simpl-solve,
// simpl-legalize), so a rocprofv3 --marker-trace timeline shows the phases.
struct PhaseRange {
    explicit PhaseRange(const char *name) { roctxRangePushA(name); }
    ~PhaseRange() { roctxRangePop(); }
    PhaseRange(const PhaseRange &) = delete;
    PhaseRange &operator=(const PhaseRange &) = delete;
};

thread_local std::string lastError;

void setError(const std::string &msg) { lastError = msg; }

int hipFail(hipError_t e, const char *what, const char *file, int line) {
    lastError = std::string(what) + ": " + hipGetErrorString(e) + " (" + file + ":" + std::to_string(line) + ")";
    return e == hipErrorOutOfMemory ? KT_ENOMEM : KT_EHIP;
}

int buildB2B(kt_ctx *ctx, double alpha, double floorX, double floorY);
int solveFp64(kt_ctx *ctx, const kt_solve_opts &o, kt_solve_result &res);
int solveMixed(kt_ctx *ctx, const kt_solve_opts &o, kt_solve_result &res);
int solveExact(kt_ctx *ctx, const kt_solve_opts &o, kt_solve_result &res);
int solveHistory(kt_ctx *ctx, const kt_solve_result &res, std::vector<double> &hist, uint64_t &cap);
int multiply(kt_ctx *ctx, int axis, const double *dIn, double *dOut);
int lalSetDesign(kt_ctx *ctx, const std::vector<lal::RowSpec> &rows, const lal::Box &die);
int lalBuildGrid(kt_ctx *ctx, uint64_t perAxis, double densityLimit);
int lalOccupancy(kt_ctx *ctx, std::vector<double> *occHost, std::vector<uint32_t> *countHost = nullptr);
int lalOverflow(kt_ctx *ctx, double *out);
int lalHpwl(kt_ctx *ctx, double *out);
int loopReset(kt_ctx *ctx, uint64_t referenceIter, double tight, double relaxed, uint64_t stall);
int anchorsFromPositions(kt_ctx *ctx);
int loopUpdate(kt_ctx *ctx, uint64_t iteration, kt_loop_state *out);
int lalLegalize(kt_ctx *ctx);

int ensureSystem(kt_ctx *ctx, int axis, uint64_t rows, uint64_t nnzCap) {
    Sys &s = ctx->sys[axis];
    KT_CHECK(rows < 0xffffffffull && nnzCap < 0xffffffffull, KT_EINVAL, "system too large for 32-bit indices");
    KT_TRY(s.rowPtr.reserve(rows + 1));
    // four spare entries: the SpMV's wide loads read whole groups of four
    KT_TRY(s.col.reserve(nnzCap + kQuadPad));
    KT_TRY(s.val.reserve(nnzCap + kQuadPad));
    KT_TRY(s.diag.reserve(std::max<uint64_t>(rows, 1)));
    KT_TRY(s.rhs.reserve(std::max<uint64_t>(rows, 1)));
    s.rows = rows;
    s.nnzCap = nnzCap;
    return KT_OK;
}

int systemNnz(kt_ctx *ctx, int axis, uint64_t *nnz) {
    Sys &s = ctx->sys[axis];
    if (s.nnzHost < 0) {
        uint32_t v = 0;
        if (s.rowPtr.p) {
            KT_HIP(hipMemcpyAsync(&v, s.rowPtr.p + s.rows, sizeof(v), hipMemcpyDeviceToHost, ctx->stream));
            KT_HIP(hipStreamSynchronize(ctx->stream));
        }
        s.nnzHost = v;
    }
    *nnz = static_cast<uint64_t>(s.nnzHost);
    return KT_OK;
}

int timedBegin(kt_ctx *ctx, int kernel, hipEvent_t *start) {
    *start = nullptr;
    if (!ctx->profile) return KT_OK;
    KT_HIP(hipEventCreate(start));
    KT_HIP(hipEventRecord(*start, ctx->stream));
    return KT_OK;
}

int timedEnd(kt_ctx *ctx, int kernel, hipEvent_t start) {
    if (!ctx->profile || start == nullptr) return KT_OK;
    hipEvent_t stop;
    KT_HIP(hipEventCreate(&stop));
    KT_HIP(hipEventRecord(stop, ctx->stream));
    ctx->timers[kernel].pending.push_back(start);
    ctx->timers[kernel].pending.push_back(stop);
    return KT_OK;
}

// Events for a launch that records them itself (hipExtLaunchKernelGGL): the
// pair is queued for collectTimers like timedBegin / timedEnd's.
int timedKernelEvents(kt_ctx *ctx, int kernel, hipEvent_t *start, hipEvent_t *stop) {
    *start = *stop = nullptr;
    if (!ctx->profile) return KT_OK;
    KT_HIP(hipEventCreate(start));
    KT_HIP(hipEventCreate(stop));
    ctx->timers[kernel].pending.push_back(*start);
    ctx->timers[kernel].pending.push_back(*stop);
    return KT_OK;
}

int collectTimers(kt_ctx *ctx) {
    KT_HIP(hipStreamSynchronize(ctx->stream));
    for (auto &t : ctx->timers) {
        for (size_t i = 0; i + 1 < t.pending.size(); i += 2) {
            float ms = 0.0f;
            KT_HIP(hipEventElapsedTime(&ms, t.pending[i], t.pending[i + 1]));
            t.ms += ms;
            ++t.launches;
            (void)hipEventDestroy(t.pending[i]);
            (void)hipEventDestroy(t.pending[i + 1]);
        }
        t.pending.clear();
    }
    return KT_OK;
}

namespace {

template <typename T>
int upload(kt_ctx *ctx, DevBuf<T> &dst, const T *src, size_t count) {
    KT_TRY(dst.reserve(std::max<size_t>(count, 1)));
    if (count) KT_HIP(hipMemcpyAsync(dst.p, src, count * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    return KT_OK;
}

int narrow(const uint64_t *src, size_t count, std::vector<uint32_t> &out, const char *what) {
    out.resize(count);
    for (size_t i = 0; i < count; ++i) {
        if (src[i] > 0xffffffffull) {
            setError(std::string(what) + " exceeds 32 bits");
            return KT_EINVAL;
        }
        out[i] = static_cast<uint32_t>(src[i]);
    }
    return KT_OK;
}

int uploadNetlist(kt_ctx *ctx, const kt_netlist &nl) {
    const uint32_t V = nl.num_vertices, n = nl.num_movable, N = nl.num_nets;
    const uint64_t P = nl.num_pins, Pm = nl.num_cell_pins;
    KT_CHECK(P < 0xffffffffull && Pm < 0xffffffffull, KT_EINVAL, "upload: too many pins for 32-bit indices");
    KT_CHECK(nl.net_start && nl.var_of_vertex && nl.pin_start, KT_EINVAL, "upload: null netlist array");
    KT_CHECK(nl.net_start[N] == P, KT_EINVAL, "upload: net_start[N] != num_pins");
    KT_CHECK(nl.pin_start[n] == Pm, KT_EINVAL, "upload: pin_start[n] != num_cell_pins");
    // validate indices on the host so no kernel can read out of bounds
    for (uint32_t j = 0; j < N; ++j)
        KT_CHECK(nl.net_start[j] <= nl.net_start[j + 1], KT_EINVAL, "upload: net_start not monotone");
    for (uint32_t i = 0; i < n; ++i)
        KT_CHECK(nl.pin_start[i] <= nl.pin_start[i + 1], KT_EINVAL, "upload: pin_start not monotone");
    for (uint64_t k = 0; k < P; ++k) KT_CHECK(nl.pin_cell[k] < V, KT_EINVAL, "upload: pin cell out of range");
    for (uint32_t v = 0; v < V; ++v)
        KT_CHECK(nl.var_of_vertex[v] == kNoVar || nl.var_of_vertex[v] < n, KT_EINVAL, "upload: var out of range");
    for (uint32_t i = 0; i < n; ++i) {
        KT_CHECK(nl.mov_vertex[i] < V && nl.var_of_vertex[nl.mov_vertex[i]] == i, KT_EINVAL,
                 "upload: mov_vertex / var_of_vertex mismatch");
    }
    std::vector<uint32_t> netMov(N, 0);
    // per pin: its vertex's variable and the movable pins before it in its net
    std::vector<uint32_t> varMov(2 * P);
    uint64_t bound = 0;
    for (uint32_t j = 0; j < N; ++j) {
        const uint64_t p = nl.net_start[j + 1] - nl.net_start[j];
        for (uint64_t k = nl.net_start[j]; k < nl.net_start[j + 1]; ++k) {
            const uint32_t var = nl.var_of_vertex[nl.pin_cell[k]];
            varMov[2 * k] = var;
            varMov[2 * k + 1] = netMov[j];
            netMov[j] += var != kNoVar ? 1u : 0u;
        }
        if (p >= 2) bound += 4 * p - 6;  // 2 bound pins x (p-1) + (p-2) inner pins x 2
    }
    // every pinsOfCell entry of row i must be a pin of cell i: the B2B count and
    // fill passes rely on it to stay inside the row's CSR slots
    for (uint32_t i = 0; i < n; ++i) {
        for (uint64_t e = nl.pin_start[i]; e < nl.pin_start[i + 1]; ++e) {
            const uint32_t j = nl.cell_pin_net[e];
            KT_CHECK(j < N, KT_EINVAL, "upload: pinsOfCell net out of range");
            KT_CHECK(nl.cell_pin_idx[e] < nl.net_start[j + 1] - nl.net_start[j], KT_EINVAL,
                     "upload: pinsOfCell pin out of range");
            KT_CHECK(nl.pin_cell[nl.net_start[j] + nl.cell_pin_idx[e]] == nl.mov_vertex[i], KT_EINVAL,
                     "upload: pinsOfCell entry is not a pin of its row's cell");
        }
    }
    KT_CHECK(bound < 0xffffffffull, KT_EINVAL, "upload: B2B entries exceed 32-bit indices");
    // Cells with no net path to a fixed vertex: union-find over vertices, each
    // net joining its pins; a component holding a fixed vertex is anchored.
    std::vector<uint32_t> parent(V);
    for (uint32_t v = 0; v < V; ++v) parent[v] = v;
    const auto root = [&parent](uint32_t v) {
        while (parent[v] != v) {
            parent[v] = parent[parent[v]];  // path halving
            v = parent[v];
        }
        return v;
    };
    for (uint32_t j = 0; j < N; ++j) {
        const uint64_t s0 = nl.net_start[j], s1 = nl.net_start[j + 1];
        if (s1 - s0 < 2) continue;
        const uint32_t first = root(nl.pin_cell[s0]);
        for (uint64_t k = s0 + 1; k < s1; ++k) {
            const uint32_t r = root(nl.pin_cell[k]), f = root(first);
            if (r != f) parent[std::max(r, f)] = std::min(r, f);
        }
    }
    std::vector<uint8_t> anchoredRoot(V, 0);
    for (uint32_t v = 0; v < V; ++v)
        if (nl.var_of_vertex[v] == kNoVar) anchoredRoot[root(v)] = 1;
    std::vector<uint8_t> floating(n);
    uint32_t floatingCells = 0;
    for (uint32_t i = 0; i < n; ++i) {
        floating[i] = anchoredRoot[root(nl.mov_vertex[i])] ? 0 : 1;
        floatingCells += floating[i];
    }
    std::vector<uint32_t> netStart32, pinStart32;
    KT_TRY(narrow(nl.net_start, N + 1ull, netStart32, "net_start"));
    KT_TRY(narrow(nl.pin_start, n + 1ull, pinStart32, "pin_start"));
    KT_TRY(upload(ctx, ctx->netStart, netStart32.data(), netStart32.size()));
    KT_TRY(upload(ctx, ctx->pinStart, pinStart32.data(), pinStart32.size()));
    KT_TRY(upload(ctx, ctx->pinCell, nl.pin_cell, P));
    KT_TRY(upload(ctx, ctx->offX, nl.pin_off_x, P));
    KT_TRY(upload(ctx, ctx->offY, nl.pin_off_y, P));
    KT_TRY(upload(ctx, ctx->weight, nl.net_weight, N));
    KT_TRY(upload(ctx, ctx->varOfVertex, nl.var_of_vertex, V));
    KT_TRY(upload(ctx, ctx->movVertex, nl.mov_vertex, n));
    KT_TRY(upload(ctx, ctx->fixedX, nl.fixed_x, V));
    KT_TRY(upload(ctx, ctx->fixedY, nl.fixed_y, V));
    KT_TRY(upload(ctx, ctx->refNet, nl.cell_pin_net, Pm));
    KT_TRY(upload(ctx, ctx->refPin, nl.cell_pin_idx, Pm));
    KT_TRY(upload(ctx, ctx->netMov, netMov.data(), N));
    // the nets whose extremes a wave finds (the B2B build's netExtremesWide)
    std::vector<uint32_t> wideNets;
    for (uint32_t j = 0; j < N; ++j)
        if (nl.net_start[j + 1] - nl.net_start[j] > kWideNetPins) wideNets.push_back(j);
    KT_TRY(upload(ctx, ctx->wideNets, wideNets.data(), wideNets.size()));
    ctx->numWideNets = static_cast<uint32_t>(wideNets.size());
    KT_TRY(upload(ctx, ctx->pinVar, varMov.data(), 2 * P));
    KT_TRY(upload(ctx, ctx->floating, floating.data(), n));
    if (nl.width) KT_TRY(upload(ctx, ctx->width, nl.width, n));
    if (nl.height) KT_TRY(upload(ctx, ctx->height, nl.height, n));
    if (nl.area) KT_TRY(upload(ctx, ctx->area, nl.area, n));
    ctx->obstacles.clear();
    for (uint32_t o = 0; o < nl.num_obstacles; ++o) {
        const double *b = nl.obstacles + 4ull * o;
        ctx->obstacles.push_back({b[0], b[1], b[2], b[3]});
    }
    KT_HIP(hipStreamSynchronize(ctx->stream));
    ctx->V = V;
    ctx->n = n;
    ctx->N = N;
    ctx->P = P;
    ctx->Pm = Pm;
    ctx->nnzBound = bound;
    ctx->floatingCells = floatingCells;
    ctx->haveNetlist = true;
    ctx->haveAnchors = false;
    ctx->posCount = 0;
    return KT_OK;
}

int setPositions(kt_ctx *ctx, const double *x, const double *y, uint64_t count) {
    KT_CHECK(x && y, KT_EINVAL, "positions: null pointer");
    KT_CHECK(count < 0xffffffffull, KT_EINVAL, "positions: too many");
    KT_TRY(ctx->posX.reserve(std::max<uint64_t>(count, 1)));
    KT_TRY(ctx->posY.reserve(std::max<uint64_t>(count, 1)));
    if (count) {
        KT_HIP(hipMemcpyAsync(ctx->posX.p, x, count * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        KT_HIP(hipMemcpyAsync(ctx->posY.p, y, count * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    }
    ctx->posCount = count;
    return KT_OK;
}

int getPositions(kt_ctx *ctx, double *x, double *y, uint64_t count) {
    KT_CHECK(x && y, KT_EINVAL, "positions: null pointer");
    KT_CHECK(count == ctx->posCount, KT_EINVAL, "positions: count differs from the stored positions");
    if (count) {
        KT_HIP(hipMemcpyAsync(x, ctx->posX.p, count * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        KT_HIP(hipMemcpyAsync(y, ctx->posY.p, count * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    }
    KT_HIP(hipStreamSynchronize(ctx->stream));
    return KT_OK;
}

int setAnchors(kt_ctx *ctx, const double *ax, const double *ay) {
    KT_CHECK(ctx->haveNetlist, KT_ESTATE, "anchors: no netlist uploaded");
    KT_CHECK(ax && ay, KT_EINVAL, "anchors: null pointer");
    KT_TRY(upload(ctx, ctx->anchorX, ax, ctx->n));
    KT_TRY(upload(ctx, ctx->anchorY, ay, ctx->n));
    ctx->haveAnchors = true;
    return KT_OK;
}

int solve(kt_ctx *ctx, const kt_solve_opts *opts, kt_solve_result *out) {
    KT_CHECK(opts && out, KT_EINVAL, "solve: null argument");
    KT_CHECK(opts->precision == KT_PREC_FP64 || opts->precision == KT_PREC_MIXED ||
                 opts->precision == KT_PREC_FP64_EXACT,
             KT_EINVAL, "solve: unknown precision");
    KT_CHECK(opts->warm_start == KT_WARM_PREVIOUS || opts->warm_start == KT_WARM_LEARNED, KT_EINVAL,
             "solve: unknown warm start");
    ctx->profile = opts->profile != 0;
    kt_solve_result res{};
    const bool learned = opts->warm_start == KT_WARM_LEARNED;
    if (learned) KT_TRY(warmStart(ctx));
    int rc;
    if (opts->precision == KT_PREC_FP64) {
        rc = solveFp64(ctx, *opts, res);
    } else if (opts->precision == KT_PREC_MIXED) {
        rc = solveMixed(ctx, *opts, res);
    } else {
        for (int a = 0; a < 2; ++a) {
            uint64_t nnz;
            KT_TRY(systemNnz(ctx, a, &nnz));
        }
        rc = solveExact(ctx, *opts, res);
    }
    if (rc != KT_OK) return rc;
    if (ctx->profile) KT_TRY(collectTimers(ctx));
    if (learned) {
        KT_HIP(hipStreamSynchronize(ctx->stream));
        res.warm_gamma[0] = ctx->pinned[62];
        res.warm_gamma[1] = ctx->pinned[63];
    }
    *out = res;
    return KT_OK;
}

}  // namespace
}  // namespace kt

namespace kt {
namespace {
__global__ void __launch_bounds__(256) kt_copy_probe(const uint4 *__restrict__ in, uint4 *__restrict__ out,
                                                      uint64_t n) {
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += static_cast<uint64_t>(gridDim.x) * 256)
        out[i] = in[i];
}
}  // namespace
}  // namespace kt

using namespace kt;

#define KT_GUARD(ctx)                                                      \
    do {                                                                   \
        if ((ctx) == nullptr) {                                            \
            setError("null context");                                      \
            return KT_EINVAL;                                              \
        }                                                                  \
        hipError_t kt_se_ = hipSetDevice((ctx)->device);                   \
        if (kt_se_ != hipSuccess) return hipFail(kt_se_, "hipSetDevice", __FILE__, __LINE__); \
    } while (0)

#define KT_CATCH(...)                                         \
    try {                                                     \
        __VA_ARGS__                                           \
    } catch (const std::bad_alloc &) {                        \
        setError("host out of memory");                       \
        return KT_ENOMEM;                                     \
    } catch (const std::exception &e) {                       \
        setError(e.what());                                   \
        return KT_EINVAL;                                     \
    }

extern "C" {

const char *kt_last_error(void) { return lastError.c_str(); }

int kt_ctx_create(kt_ctx **out, int device) {
    KT_CHECK(out != nullptr, KT_EINVAL, "kt_ctx_create: null out");
    *out = nullptr;
    int count = 0;
    KT_HIP(hipGetDeviceCount(&count));
    KT_CHECK(device >= 0 && device < count, KT_EINVAL, "kt_ctx_create: no such device");
    KT_HIP(hipSetDevice(device));
    KT_CATCH({
        auto *ctx = new kt_ctx();
        ctx->device = device;
        // KTPLACE_NO_GRAPHS=1 launches every CG kernel directly instead of
        // replaying captured hipGraphs (for tools that trace graph launches badly)
        const char *noGraphs = std::getenv("KTPLACE_NO_GRAPHS");
        if (noGraphs && noGraphs[0] == '1') ctx->tune.useGraphs = false;
        hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void **>(&ctx->ctlHost), 2 * sizeof(AxisCtl));
        if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void **>(&ctx->pinned), 64 * sizeof(double));
        if (e != hipSuccess) {
            if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
            delete ctx;
            return hipFail(e, "kt_ctx_create", __FILE__, __LINE__);
        }
        // this device's copy of kt_hpwl_noalloc's fixed scratch
        const int rc = noAllocInit(device);
        if (rc != KT_OK) {
            (void)hipStreamDestroy(ctx->stream);
            (void)hipHostFree(ctx->ctlHost);
            (void)hipHostFree(ctx->pinned);
            delete ctx;
            return rc;
        }
        *out = ctx;
        return KT_OK;
    })
}

int kt_ctx_destroy(kt_ctx *ctx) {
    if (ctx == nullptr) return KT_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &t : ctx->timers)
        for (hipEvent_t ev : t.pending) (void)hipEventDestroy(ev);
    ctx->netStart.release(); ctx->pinCell.release(); ctx->varOfVertex.release(); ctx->movVertex.release();
    ctx->pinStart.release(); ctx->refNet.release(); ctx->refPin.release(); ctx->netMov.release();
    ctx->wideNets.release();
    ctx->floating.release(); ctx->hpwlScratch.release(); ctx->loopDev.release();
    ctx->b2bEdgeOff.release(); ctx->b2bEntOff.release(); ctx->b2bLong.release(); ctx->b2bRowEdge.release();
    for (int a = 0; a < 2; ++a) ctx->b2bEdges[a].release();
    ctx->offX.release(); ctx->offY.release(); ctx->weight.release(); ctx->fixedX.release();
    ctx->fixedY.release(); ctx->width.release(); ctx->height.release(); ctx->area.release();
    ctx->posX.release(); ctx->posY.release(); ctx->saveX.release(); ctx->saveY.release();
    ctx->anchorX.release(); ctx->anchorY.release(); ctx->vxy.release();
    ctx->ext.release(); ctx->scanTmp.release(); ctx->pinVar.release();
    for (int a = 0; a < 2; ++a) ctx->pinAt[a].release();
    for (auto &s : ctx->sys) {
        s.rowPtr.release(); s.col.release(); s.val.release(); s.diag.release(); s.rhs.release();
        s.val32.release(); s.diag32.release(); s.isd.release();
    }
    for (int a = 0; a < 2; ++a) {
        ctx->r[a].release(); ctx->p[a].release(); ctx->ap[a].release(); ctx->xs[a].release();
        ctx->r32[a].release(); ctx->p32[a].release(); ctx->ap32[a].release(); ctx->x32[a].release();
    }
    ctx->partials.release(); ctx->ctl.release(); ctx->history.release();
    for (auto &byPrecision : ctx->graphs)
        for (auto &g : byPrecision) {
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
            if (g.graph) (void)hipGraphDestroy(g.graph);
        }
    if (ctx->ctlHost) (void)hipHostFree(ctx->ctlHost);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    if (ctx->loopHost) (void)hipHostFree(ctx->loopHost);
    if (ctx->lalPinned) (void)hipHostFree(ctx->lalPinned);
    if (ctx->lalLand) (void)hipHostFree(ctx->lalLand);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return KT_OK;
}

int kt_sync(kt_ctx *ctx) {
    KT_GUARD(ctx);
    KT_HIP(hipStreamSynchronize(ctx->stream));
    return KT_OK;
}

int kt_ctx_device_bytes(kt_ctx *ctx, uint64_t *bytes) {
    KT_GUARD(ctx);
    KT_CHECK(bytes, KT_EINVAL, "null out");
    uint64_t b = 0;
    b += ctx->netStart.bytes() + ctx->pinCell.bytes() + ctx->varOfVertex.bytes() + ctx->movVertex.bytes();
    b += ctx->pinStart.bytes() + ctx->refNet.bytes() + ctx->refPin.bytes() + ctx->netMov.bytes() + ctx->wideNets.bytes();
    b += ctx->floating.bytes() + ctx->hpwlScratch.bytes() + ctx->loopDev.bytes();
    b += ctx->b2bEdgeOff.bytes() + ctx->b2bEntOff.bytes() + ctx->b2bLong.bytes() + ctx->b2bRowEdge.bytes() + ctx->b2bEdges[0].bytes() +
         ctx->b2bEdges[1].bytes();
    b += ctx->offX.bytes() + ctx->offY.bytes() + ctx->weight.bytes() + ctx->fixedX.bytes() + ctx->fixedY.bytes();
    b += ctx->width.bytes() + ctx->height.bytes() + ctx->area.bytes();
    b += ctx->posX.bytes() + ctx->posY.bytes() + ctx->saveX.bytes() + ctx->saveY.bytes();
    b += ctx->anchorX.bytes() + ctx->anchorY.bytes() + ctx->vxy.bytes() + ctx->ext.bytes();
    b += ctx->pinAt[0].bytes() + ctx->pinAt[1].bytes();
    b += ctx->pinVar.bytes();
    b += ctx->scanTmp.bytes();
    for (auto &s : ctx->sys)
        b += s.rowPtr.bytes() + s.col.bytes() + s.val.bytes() + s.diag.bytes() + s.rhs.bytes() + s.val32.bytes() +
             s.diag32.bytes() + s.isd.bytes();
    for (int a = 0; a < 2; ++a)
        b += ctx->r[a].bytes() + ctx->p[a].bytes() + ctx->ap[a].bytes() + ctx->xs[a].bytes() + ctx->r32[a].bytes() +
             ctx->p32[a].bytes() + ctx->ap32[a].bytes() + ctx->x32[a].bytes();
    b += ctx->partials.bytes() + ctx->ctl.bytes() + ctx->history.bytes();
    *bytes = b;
    return KT_OK;
}

int kt_upload_netlist(kt_ctx *ctx, const kt_netlist *nl) {
    KT_GUARD(ctx);
    KT_CHECK(nl, KT_EINVAL, "kt_upload_netlist: null netlist");
    KT_CATCH(return uploadNetlist(ctx, *nl);)
}

int kt_set_positions(kt_ctx *ctx, const double *x, const double *y, uint64_t count) {
    KT_GUARD(ctx);
    return setPositions(ctx, x, y, count);
}

int kt_get_positions(kt_ctx *ctx, double *x, double *y, uint64_t count) {
    KT_GUARD(ctx);
    return getPositions(ctx, x, y, count);
}

int kt_set_anchors(kt_ctx *ctx, const double *ax, const double *ay) {
    KT_GUARD(ctx);
    return setAnchors(ctx, ax, ay);
}

int kt_save_positions(kt_ctx *ctx) {
    KT_GUARD(ctx);
    const uint64_t c = ctx->posCount;
    KT_TRY(ctx->saveX.reserve(std::max<uint64_t>(c, 1)));
    KT_TRY(ctx->saveY.reserve(std::max<uint64_t>(c, 1)));
    if (c) {
        KT_HIP(hipMemcpyAsync(ctx->saveX.p, ctx->posX.p, c * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        KT_HIP(hipMemcpyAsync(ctx->saveY.p, ctx->posY.p, c * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    }
    return KT_OK;
}

int kt_restore_positions(kt_ctx *ctx) {
    KT_GUARD(ctx);
    const uint64_t c = ctx->posCount;
    KT_CHECK(ctx->saveX.cap >= c, KT_ESTATE, "kt_restore_positions: nothing saved");
    if (c) {
        KT_HIP(hipMemcpyAsync(ctx->posX.p, ctx->saveX.p, c * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        KT_HIP(hipMemcpyAsync(ctx->posY.p, ctx->saveY.p, c * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    }
    return KT_OK;
}

int kt_build_b2b(kt_ctx *ctx, const double *x, const double *y, double alpha, const double *anchor_x,
                 const double *anchor_y, double floor_x, double floor_y) {
    KT_GUARD(ctx);
    const PhaseRange range("simpl-build");
    KT_CHECK(ctx->haveNetlist, KT_ESTATE, "kt_build_b2b: no netlist uploaded");
    KT_TRY(setPositions(ctx, x, y, ctx->n));
    if (alpha > 0.0) {
        KT_CHECK(anchor_x && anchor_y, KT_EINVAL, "kt_build_b2b: alpha > 0 needs anchors");
        KT_TRY(setAnchors(ctx, anchor_x, anchor_y));
    }
    KT_TRY(buildB2B(ctx, alpha, floor_x, floor_y));
    KT_HIP(hipStreamSynchronize(ctx->stream));
    return KT_OK;
}

int kt_build_star(kt_ctx *ctx, uint64_t *size) {
    KT_GUARD(ctx);
    const PhaseRange range("simpl-init");
    KT_CATCH(return buildStar(ctx, size);)
}

int kt_build_b2b_resident(kt_ctx *ctx, double alpha, double floor_x, double floor_y) {
    KT_GUARD(ctx);
    const PhaseRange range("simpl-build");
    return buildB2B(ctx, alpha, floor_x, floor_y);
}

int kt_system_size(kt_ctx *ctx, int axis, uint64_t *rows, uint64_t *nnz) {
    KT_GUARD(ctx);
    KT_CHECK(axis == 0 || axis == 1, KT_EINVAL, "axis must be 0 or 1");
    KT_CHECK(rows && nnz, KT_EINVAL, "null out");
    *rows = ctx->sys[axis].rows;
    return systemNnz(ctx, axis, nnz);
}

int kt_get_system(kt_ctx *ctx, int axis, uint64_t *row_ptr, uint32_t *col, double *val, double *diag,
                  double *rhs) {
    KT_GUARD(ctx);
    KT_CHECK(axis == 0 || axis == 1, KT_EINVAL, "axis must be 0 or 1");
    Sys &s = ctx->sys[axis];
    uint64_t nnz;
    KT_TRY(systemNnz(ctx, axis, &nnz));
    KT_CATCH({
        std::vector<uint32_t> rp(s.rows + 1);
        if (s.rows + 1 > 0 && s.rowPtr.p)
            KT_HIP(hipMemcpyAsync(rp.data(), s.rowPtr.p, rp.size() * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                  ctx->stream));
        if (nnz) {
            if (col) KT_HIP(hipMemcpyAsync(col, s.col.p, nnz * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
            if (val) KT_HIP(hipMemcpyAsync(val, s.val.p, nnz * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        }
        if (s.rows) {
            if (diag) KT_HIP(hipMemcpyAsync(diag, s.diag.p, s.rows * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
            if (rhs) KT_HIP(hipMemcpyAsync(rhs, s.rhs.p, s.rows * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        }
        KT_HIP(hipStreamSynchronize(ctx->stream));
        if (row_ptr)
            for (size_t i = 0; i < rp.size(); ++i) row_ptr[i] = rp[i];
        return KT_OK;
    })
}

int kt_set_system(kt_ctx *ctx, int axis, uint64_t rows, const uint64_t *row_ptr, const uint32_t *col,
                  const double *val, const double *diag, const double *rhs) {
    KT_GUARD(ctx);
    KT_CHECK(axis == 0 || axis == 1, KT_EINVAL, "axis must be 0 or 1");
    KT_CHECK(row_ptr && diag && rhs, KT_EINVAL, "kt_set_system: null array");
    const uint64_t nnz = row_ptr[rows];
    KT_CHECK(nnz == 0 || (col && val), KT_EINVAL, "kt_set_system: null col/val");
    KT_CHECK(row_ptr[0] == 0, KT_EINVAL, "kt_set_system: row_ptr[0] != 0");
    for (uint64_t i = 0; i < rows; ++i)
        KT_CHECK(row_ptr[i] <= row_ptr[i + 1], KT_EINVAL, "kt_set_system: row_ptr not monotone");
    for (uint64_t e = 0; e < nnz; ++e) KT_CHECK(col[e] < rows, KT_EINVAL, "kt_set_system: column out of range");
    KT_TRY(ensureSystem(ctx, axis, rows, nnz));
    Sys &s = ctx->sys[axis];
    KT_CATCH({
        std::vector<uint32_t> rp;
        KT_TRY(narrow(row_ptr, rows + 1, rp, "row_ptr"));
        KT_HIP(hipMemcpyAsync(s.rowPtr.p, rp.data(), rp.size() * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
        if (nnz) {
            KT_HIP(hipMemcpyAsync(s.col.p, col, nnz * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
            KT_HIP(hipMemcpyAsync(s.val.p, val, nnz * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        }
        if (rows) {
            KT_HIP(hipMemcpyAsync(s.diag.p, diag, rows * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
            KT_HIP(hipMemcpyAsync(s.rhs.p, rhs, rows * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        }
        KT_HIP(hipStreamSynchronize(ctx->stream));
        s.nnzHost = static_cast<int64_t>(nnz);
        s.have32 = false;
        s.have24 = false;
        s.havePk = false;
        return KT_OK;
    })
}

int kt_multiply(kt_ctx *ctx, int axis, const double *in, double *out) {
    KT_GUARD(ctx);
    KT_CHECK(axis == 0 || axis == 1, KT_EINVAL, "axis must be 0 or 1");
    KT_CHECK(in && out, KT_EINVAL, "kt_multiply: null pointer");
    const uint64_t rows = ctx->sys[axis].rows;
    uint64_t nnz;
    KT_TRY(systemNnz(ctx, axis, &nnz));
    double *dIn = nullptr, *dOut = nullptr;
    KT_HIP(hipMalloc(reinterpret_cast<void **>(&dIn), std::max<uint64_t>(rows, 1) * sizeof(double)));
    hipError_t e = hipMalloc(reinterpret_cast<void **>(&dOut), std::max<uint64_t>(rows, 1) * sizeof(double));
    if (e != hipSuccess) {
        (void)hipFree(dIn);
        return hipFail(e, "hipMalloc", __FILE__, __LINE__);
    }
    int rc = KT_OK;
    e = hipMemcpyAsync(dIn, in, rows * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) rc = multiply(ctx, axis, dIn, dOut);
    if (e == hipSuccess && rc == KT_OK)
        e = hipMemcpyAsync(out, dOut, rows * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(dIn);
    (void)hipFree(dOut);
    if (e != hipSuccess) return hipFail(e, "kt_multiply", __FILE__, __LINE__);
    return rc;
}

int kt_solve_both_resident(kt_ctx *ctx, const kt_solve_opts *opts, kt_solve_result *out) {
    KT_GUARD(ctx);
    const PhaseRange range("simpl-solve");
    KT_CATCH(return solve(ctx, opts, out);)
}

int kt_solve_both(kt_ctx *ctx, double *x, double *y, uint64_t count, const kt_solve_opts *opts,
                  kt_observer observe, void *user, kt_solve_result *out) {
    KT_GUARD(ctx);
    const PhaseRange range("simpl-solve");
    KT_CATCH({
        KT_TRY(setPositions(ctx, x, y, count));
        kt_solve_result res{};
        KT_TRY(solve(ctx, opts, &res));
        KT_TRY(getPositions(ctx, x, y, count));
        if (observe && res.iterations > 0) {
            std::vector<double> hist;
            uint64_t cap = 0;
            if (opts->precision == KT_PREC_FP64_EXACT) {
                // host history is [iteration][axis]; lay it out [axis][iteration]
                cap = res.iterations;
                hist.assign(2 * cap, 0.0);
                for (uint64_t it = 0; it < cap; ++it) {
                    hist[it] = ctx->exactHistory[2 * it];
                    hist[cap + it] = ctx->exactHistory[2 * it + 1];
                }
            } else {
                KT_TRY(solveHistory(ctx, res, hist, cap));
            }
            const AxisCtl *c = ctx->ctlHost;
            for (uint64_t it = 0; it < res.iterations; ++it) {
                double r[2];
                for (int a = 0; a < 2; ++a) {
                    const uint64_t st = res.steps[a];
                    r[a] = (st == 0 || it >= st || it >= cap) ? res.resid[a] : hist[a * cap + it];
                    if (st > 0 && it >= st && st - 1 < cap) r[a] = hist[a * cap + st - 1];
                }
                bool last = false;
                if (it + 1 == res.iterations) {
                    const bool liveX = c[0].s[0].active && res.resid[0] > opts->tol;
                    const bool liveY = c[1].s[0].active && res.resid[1] > opts->tol;
                    last = !liveX && !liveY;
                }
                observe(user, it + 1, r[0], r[1], last ? 1 : 0);
            }
        }
        *out = res;
        return KT_OK;
    })
}

int kt_set_rows(kt_ctx *ctx, uint32_t num_rows, const double *row_y, const double *row_height, const double *row_x0,
                const double *row_x1, const double *row_pitch, const double die[4]) {
    KT_GUARD(ctx);
    KT_CHECK(die && (num_rows == 0 || (row_y && row_height && row_x0 && row_x1 && row_pitch)), KT_EINVAL,
             "kt_set_rows: null array");
    KT_CATCH({
        std::vector<lal::RowSpec> rows(num_rows);
        for (uint32_t r = 0; r < num_rows; ++r) rows[r] = {row_y[r], row_height[r], row_x0[r], row_x1[r], row_pitch[r]};
        return lalSetDesign(ctx, rows, lal::Box{die[0], die[1], die[2], die[3]});
    })
}

int kt_build_grid(kt_ctx *ctx, uint64_t bins_per_axis, double density_limit, uint64_t *bins_out) {
    KT_GUARD(ctx);
    uint64_t per = bins_per_axis;
    if (per == 0)
        per = static_cast<uint64_t>(std::clamp(std::sqrt(static_cast<double>(std::max<uint32_t>(ctx->n, 1)) / 51.0),
                                               16.0, 256.0));
    KT_CATCH(KT_TRY(lalBuildGrid(ctx, per, density_limit));)
    if (bins_out) *bins_out = per;
    return KT_OK;
}

int kt_grid_avail(kt_ctx *ctx, double *avail, double *total_avail) {
    KT_GUARD(ctx);
    KT_CHECK(ctx->grid.size() > 0, KT_ESTATE, "kt_grid_avail: no grid");
    if (avail) std::copy(ctx->grid.avail.begin(), ctx->grid.avail.end(), avail);
    if (total_avail) *total_avail = ctx->grid.totalAvail;
    return KT_OK;
}

int kt_occupancy(kt_ctx *ctx, double *occ) {
    KT_GUARD(ctx);
    KT_CHECK(occ, KT_EINVAL, "kt_occupancy: null out");
    KT_CHECK(ctx->posCount == ctx->n, KT_ESTATE, "kt_occupancy: positions must have one entry per movable cell");
    KT_CATCH({
        std::vector<double> o;
        KT_TRY(lalOccupancy(ctx, &o, nullptr));
        std::copy(o.begin(), o.end(), occ);
        return KT_OK;
    })
}

int kt_overflow(kt_ctx *ctx, double *overflow) {
    KT_GUARD(ctx);
    KT_CHECK(overflow, KT_EINVAL, "kt_overflow: null out");
    KT_CHECK(ctx->posCount == ctx->n, KT_ESTATE, "kt_overflow: positions must have one entry per movable cell");
    KT_CATCH(return lalOverflow(ctx, overflow);)
}

int kt_hpwl(kt_ctx *ctx, double *hpwl) {
    KT_GUARD(ctx);
    KT_CHECK(hpwl, KT_EINVAL, "kt_hpwl: null out");
    KT_CATCH(return lalHpwl(ctx, hpwl);)
}

int kt_set_lal_options(kt_ctx *ctx, uint64_t max_level, uint64_t min_cells_to_split, double stripe_area_fraction) {
    KT_GUARD(ctx);
    KT_CHECK(stripe_area_fraction >= 0.0, KT_EINVAL, "kt_set_lal_options: negative stripe fraction");
    if (max_level) ctx->lalOpts.maxLevel = max_level;
    if (min_cells_to_split) ctx->lalOpts.minCellsToSplit = min_cells_to_split;
    if (stripe_area_fraction > 0.0) ctx->lalOpts.stripeAreaFraction = stripe_area_fraction;
    return KT_OK;
}

int kt_lal_resident(kt_ctx *ctx) {
    KT_GUARD(ctx);
    const PhaseRange range("simpl-legalize");
    KT_CATCH(return lalLegalize(ctx);)
}

int kt_lal(kt_ctx *ctx, double *x, double *y) {
    KT_GUARD(ctx);
    const PhaseRange range("simpl-legalize");
    KT_CHECK(x && y, KT_EINVAL, "kt_lal: null positions");
    KT_CATCH({
        KT_TRY(setPositions(ctx, x, y, ctx->n));
        KT_TRY(lalLegalize(ctx));
        return getPositions(ctx, x, y, ctx->n);
    })
}

int kt_lal_batches(kt_ctx *ctx, uint64_t *batches) {
    KT_GUARD(ctx);
    KT_CHECK(batches, KT_EINVAL, "kt_lal_batches: null output");
    *batches = ctx->lalBatches;
    return KT_OK;
}

int kt_loop_reset(kt_ctx *ctx, uint64_t reference_iter, double tight_frac, double relaxed_frac,
                  uint64_t stall_iters) {
    KT_GUARD(ctx);
    KT_CATCH(return loopReset(ctx, reference_iter, tight_frac, relaxed_frac, stall_iters);)
}

int kt_anchors_from_positions(kt_ctx *ctx) {
    KT_GUARD(ctx);
    KT_CATCH(return anchorsFromPositions(ctx);)
}

int kt_get_anchors(kt_ctx *ctx, double *x, double *y) {
    KT_GUARD(ctx);
    KT_CHECK(x && y, KT_EINVAL, "kt_get_anchors: null output");
    KT_CHECK(ctx->haveAnchors, KT_ESTATE, "kt_get_anchors: no anchors");
    if (ctx->n) {
        KT_HIP(hipMemcpyAsync(x, ctx->anchorX.p, ctx->n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        KT_HIP(hipMemcpyAsync(y, ctx->anchorY.p, ctx->n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    }
    KT_HIP(hipStreamSynchronize(ctx->stream));
    return KT_OK;
}

int kt_loop_update(kt_ctx *ctx, uint64_t iteration, kt_loop_state *out) {
    KT_GUARD(ctx);
    KT_CHECK(out, KT_EINVAL, "kt_loop_update: null output");
    KT_CATCH(return loopUpdate(ctx, iteration, out);)
}

int kt_floating_cells(kt_ctx *ctx, uint32_t *count, uint8_t *flags) {
    KT_GUARD(ctx);
    KT_CHECK(count, KT_EINVAL, "kt_floating_cells: null count");
    KT_CHECK(ctx->haveNetlist, KT_ESTATE, "kt_floating_cells: upload a netlist first");
    *count = ctx->floatingCells;
    if (flags && ctx->n) {
        KT_HIP(hipMemcpyAsync(flags, ctx->floating.p, ctx->n, hipMemcpyDeviceToHost, ctx->stream));
        KT_HIP(hipStreamSynchronize(ctx->stream));
    }
    return KT_OK;
}

int kt_lal_reruns(kt_ctx *ctx, uint64_t *reruns) {
    KT_GUARD(ctx);
    KT_CHECK(reruns, KT_EINVAL, "kt_lal_reruns: null output");
    *reruns = ctx->lalReruns;
    return KT_OK;
}

int kt_set_lal_fused(kt_ctx *ctx, uint32_t max_cells) {
    KT_GUARD(ctx);
    KT_CHECK(max_cells <= kLalFusedCells, KT_EINVAL, "kt_set_lal_fused: at most 1024 cells");
    ctx->lalFusedCells = max_cells;
    return KT_OK;
}

int kt_lal_fused_batches(kt_ctx *ctx, uint64_t *batches) {
    KT_GUARD(ctx);
    KT_CHECK(batches, KT_EINVAL, "kt_lal_fused_batches: null output");
    *batches = ctx->lalFusedBatches;
    return KT_OK;
}

int kt_set_warm_start(kt_ctx *ctx, const float *params, uint64_t count) {
    KT_GUARD(ctx);
    KT_CATCH(return warmSetParams(ctx, params, count);)
}

int kt_warm_features(kt_ctx *ctx, float *features, double *scale) {
    KT_GUARD(ctx);
    KT_CATCH(return warmFeaturesHost(ctx, features, scale);)
}

int kt_warm_predict(kt_ctx *ctx, double *dx, double *dy) {
    KT_GUARD(ctx);
    KT_CHECK(dx && dy, KT_EINVAL, "kt_warm_predict: null array");
    KT_CATCH(return warmPredictHost(ctx, dx, dy);)
}

int kt_probe_copy(kt_ctx *ctx, uint64_t bytes, int reps, double *ms_per_copy) {
    KT_GUARD(ctx);
    KT_CHECK(bytes >= 16 && bytes % 16 == 0 && reps > 0 && ms_per_copy, KT_EINVAL,
             "kt_probe_copy: bytes must be a positive multiple of 16");
    void *a = nullptr, *b = nullptr;
    KT_HIP(hipMalloc(&a, bytes));
    hipError_t e = hipMalloc(&b, bytes);
    if (e != hipSuccess) {
        (void)hipFree(a);
        return hipFail(e, "hipMalloc", __FILE__, __LINE__);
    }
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (e == hipSuccess) e = hipMemsetAsync(a, 1, bytes, ctx->stream);
    if (e == hipSuccess) e = hipEventCreate(&t0);
    if (e == hipSuccess) e = hipEventCreate(&t1);
    const uint64_t n = bytes / 16;
    if (e == hipSuccess) e = hipEventRecord(t0, ctx->stream);
    for (int r = 0; r < reps && e == hipSuccess; ++r) {
        kt_copy_probe<<<dim3(4096), dim3(256), 0, ctx->stream>>>(static_cast<const uint4 *>(a), static_cast<uint4 *>(b), n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(t1, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(t1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (e != hipSuccess) return hipFail(e, "kt_probe_copy", __FILE__, __LINE__);
    *ms_per_copy = ms / reps;
    return KT_OK;
}

int kt_set_tuning(kt_ctx *ctx, int lanes, uint32_t tile, int use_graphs, double refresh_factor) {
    return kt_set_tuning2(ctx, lanes, tile, use_graphs, refresh_factor, -1, 0);
}

int kt_set_tuning2(kt_ctx *ctx, int lanes, uint32_t tile, int use_graphs, double refresh_factor, int xcd_split,
                   int rows_in_flight) {
    KT_GUARD(ctx);
    KT_CHECK(lanes == 0 || lanes == 1 || lanes == 2 || lanes == 4 || lanes == 8 || lanes == 16, KT_EINVAL,
             "kt_set_tuning: lanes must be 0, 1, 2, 4, 8 or 16");
    KT_CHECK(tile % 1024 == 0, KT_EINVAL, "kt_set_tuning: tile must be a multiple of 1024");
    KT_CHECK(rows_in_flight == 0 || rows_in_flight == 2 || rows_in_flight == 4 || rows_in_flight == 8, KT_EINVAL,
             "kt_set_tuning: rows in flight must be 0, 2, 4 or 8");
    ctx->tune.rows = rows_in_flight;
    KT_CHECK(refresh_factor >= 0.0 && refresh_factor < 1.0, KT_EINVAL, "kt_set_tuning: refresh factor in [0, 1)");
    ctx->tune.lanes = lanes;
    ctx->tune.tile = tile;
    const char *noGraphs = std::getenv("KTPLACE_NO_GRAPHS");
    ctx->tune.useGraphs = use_graphs != 0 && !(noGraphs && noGraphs[0] == '1');
    ctx->tune.refreshFactor = refresh_factor;
    KT_CHECK(xcd_split <= 2, KT_EINVAL, "kt_set_tuning: xcd_split must be -1, 0, 1 or 2");
    ctx->tune.xcdSplit = xcd_split < 0 ? -1 : xcd_split;
    return KT_OK;
}

int kt_set_spmv_halo(kt_ctx *ctx, uint32_t halo) {
    KT_GUARD(ctx);
    KT_CHECK(halo % 256 == 0, KT_EINVAL, "kt_set_spmv_halo: halo must be a multiple of 256");
    KT_CHECK(halo <= 65536, KT_EINVAL, "kt_set_spmv_halo: halo above 65536 entries");
    ctx->tune.halo = halo;
    return KT_OK;
}

int kt_set_spmv_columns(kt_ctx *ctx, int bytes) {
    KT_GUARD(ctx);
    KT_CHECK(bytes == 0 || bytes == 3 || bytes == 4, KT_EINVAL, "kt_set_spmv_columns: bytes must be 0, 3 or 4");
    ctx->tune.colBytes = bytes;
    return KT_OK;
}

int kt_set_spmv_wide_loads(kt_ctx *ctx, int on) {
    KT_GUARD(ctx);
    KT_CHECK(on == 0 || on == 1, KT_EINVAL, "kt_set_spmv_wide_loads: on must be 0 or 1");
    ctx->tune.wideLoads = on;
    return KT_OK;
}

int kt_set_mixed_scaling(kt_ctx *ctx, int on) {
    KT_GUARD(ctx);
    KT_CHECK(on == 0 || on == 1, KT_EINVAL, "kt_set_mixed_scaling: on must be 0 or 1");
    ctx->tune.mixedScaling = on;
    return KT_OK;
}

int kt_set_mixed_fused_update(kt_ctx *ctx, int on) {
    KT_GUARD(ctx);
    KT_CHECK(on == 0 || on == 1, KT_EINVAL, "kt_set_mixed_fused_update: on must be 0 or 1");
    ctx->tune.fusedUpdate = on;
    return KT_OK;
}

int kt_set_spmv_layout(kt_ctx *ctx, int layout, uint32_t block_entries, int unroll, int row_lanes,
                       int block_threads) {
    KT_GUARD(ctx);
    KT_CHECK(layout == KT_SPMV_ROW_GROUPS || layout == KT_SPMV_ENTRY_BLOCKS, KT_EINVAL,
             "kt_set_spmv_layout: unknown layout");
    KT_CHECK(block_threads == 0 || block_threads == 256 || block_threads == 512 || block_threads == 1024,
             KT_EINVAL, "kt_set_spmv_layout: block threads must be 0, 256, 512 or 1024");
    const int nt = block_threads == 0 ? 1024 : block_threads;
    // the entry-block sizes built for each block size
    const bool entriesOk = block_entries == 0 ||
                           (nt == 1024 && (block_entries == 4096 || block_entries == 8192 || block_entries == 16384)) ||
                           (nt == 512 && block_entries == 4096) || (nt == 256 && block_entries == 2048);
    KT_CHECK(entriesOk, KT_EINVAL,
             "kt_set_spmv_layout: block entries must be 0 or, for 1024 threads, 4096, 8192 or 16384; "
             "for 512 threads 4096; for 256 threads 2048");
    KT_CHECK(unroll == 0 || unroll == 4 || unroll == 8, KT_EINVAL, "kt_set_spmv_layout: unroll must be 0, 4 or 8");
    KT_CHECK(unroll != 8 || nt != 512, KT_EINVAL, "kt_set_spmv_layout: unroll 8 needs 256 or 1024 threads per block");
    KT_CHECK(row_lanes == 0 || row_lanes == 1 || row_lanes == 4, KT_EINVAL,
             "kt_set_spmv_layout: row lanes must be 0, 1 or 4");
    ctx->tune.flatThreads = block_threads;
    ctx->tune.flatRowLanes = row_lanes;
    ctx->tune.layout = layout;
    ctx->tune.flatEntries = block_entries;
    ctx->tune.flatUnroll = unroll;
    return KT_OK;
}

int kt_kernel_times(kt_ctx *ctx, int kernel, uint64_t *launches, double *total_ms) {
    KT_GUARD(ctx);
    KT_CHECK(kernel >= 0 && kernel < KT_KERNEL_COUNT, KT_EINVAL, "unknown kernel id");
    KT_TRY(collectTimers(ctx));
    if (launches) *launches = ctx->timers[kernel].launches;
    if (total_ms) *total_ms = ctx->timers[kernel].ms;
    return KT_OK;
}

int kt_reset_kernel_times(kt_ctx *ctx) {
    KT_GUARD(ctx);
    KT_TRY(collectTimers(ctx));
    for (auto &t : ctx->timers) {
        t.launches = 0;
        t.ms = 0.0;
    }
    return KT_OK;
}

}  // extern "C"

namespace kt {

// Host copies of a context's system and positions (for kt_batch_add_ctx).
int getSystemHost(kt_ctx *ctx, int axis, std::vector<uint64_t> &rowPtr, std::vector<uint32_t> &col,
                  std::vector<double> &val, std::vector<double> &diag, std::vector<double> &rhs) {
    uint64_t rows = 0, nnz = 0;
    KT_TRY(kt_system_size(ctx, axis, &rows, &nnz));
    rowPtr.resize(rows + 1);
    col.resize(nnz);
    val.resize(nnz);
    diag.resize(rows);
    rhs.resize(rows);
    return kt_get_system(ctx, axis, rowPtr.data(), col.data(), val.data(), diag.data(), rhs.data());
}

int getPositionsHost(kt_ctx *ctx, std::vector<double> &x, std::vector<double> &y) {
    KT_CHECK(ctx, KT_EINVAL, "null context");
    x.resize(ctx->posCount);
    y.resize(ctx->posCount);
    return kt_get_positions(ctx, x.data(), y.data(), ctx->posCount);
}

}  // namespace kt
