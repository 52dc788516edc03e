// solveBothAxes in the reference's exact arithmetic order (KT_PREC_FP64_EXACT).
//
// Every floating-point operation of simpl_lss.cc:203-290 is reproduced in the
// same order: the SpMV sums diag * in first and then the row's entries in CSR
// order (SparseSystem::multiply, simpl_lss.h:49-57); every dot product sums
// 4096-element chunks sequentially and then the chunk partials in order
// (getDeterministicSum, simpl_parallel.h:35-51); the vector updates are the
// same multiply-then-add expressions (the library is built without FMA
// contraction).  The result is bit-identical to the reference's CPU solve, so
// a SimPL run on the GPU follows the reference's trajectory exactly.  It is
// a parity mode: a chunk sum runs on one thread, so it is far slower than the
// default solver.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "kt_internal.h"

namespace kt {
namespace {

constexpr int kB = 256;
constexpr uint32_t kChunk = 4096;  // simpl_parallel.h:12 grain

struct ExactAxis {
    const uint32_t *rowPtr;
    const uint32_t *col;
    const double *val;
    const double *diag;
    const double *rhs;
    double *x, *r, *z, *p, *ap;
    double *part;   // [chunks]
    uint32_t rows;
};

struct ExactArgs {
    ExactAxis ax[2];
    double *scal;   // per axis: [0] bNorm [1] rho [2] resid [3] dot out [4] active [5] stepped [6] tmp
};

// out = A in, SparseSystem::multiply order
__global__ void __launch_bounds__(kB) exMultiply(ExactArgs a, int which, const int *live) {
    const int axis = blockIdx.y;
    if (live && !live[axis]) return;
    const ExactAxis &s = a.ax[axis];
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= s.rows) return;
    const double *in = which == 0 ? s.x : s.p;
    double sum = s.diag[i] * in[i];
    for (uint32_t e = s.rowPtr[i]; e < s.rowPtr[i + 1]; ++e) sum += s.val[e] * in[s.col[e]];
    s.ap[i] = sum;
}

// chunk partials of a.b: one thread per 4096-element chunk, in order
__global__ void __launch_bounds__(kB) exChunkDot(ExactArgs a, int which, const int *live) {
    const int axis = blockIdx.y;
    if (live && !live[axis]) return;
    const ExactAxis &s = a.ax[axis];
    const uint32_t c = blockIdx.x * kB + threadIdx.x;
    const uint32_t chunks = (s.rows + kChunk - 1) / kChunk;
    if (c >= chunks) return;
    const double *u, *v;
    switch (which) {
        case 0: u = s.rhs; v = s.rhs; break;  // b.b
        case 1: u = s.r; v = s.z; break;      // r.z
        case 2: u = s.r; v = s.r; break;      // r.r
        case 3: u = s.p; v = s.ap; break;     // p.Ap
        default: u = nullptr; v = nullptr; break;  // (b - Ap).(b - Ap) after Ap = A x
    }
    const uint32_t end = min(s.rows, (c + 1) * kChunk);
    double sum = 0.0;
    if (u) {
        for (uint32_t i = c * kChunk; i < end; ++i) sum += u[i] * v[i];
    } else {
        for (uint32_t i = c * kChunk; i < end; ++i) {
            const double t = s.rhs[i] - s.ap[i];
            sum += t * t;
        }
    }
    s.part[c] = sum;
}

__global__ void exChunkTotal(ExactArgs a, const int *live, int slot) {
    const int axis = threadIdx.x;
    if (axis >= 2 || (live && !live[axis])) return;
    const ExactAxis &s = a.ax[axis];
    const uint32_t chunks = (s.rows + kChunk - 1) / kChunk;
    double t = 0.0;
    for (uint32_t c = 0; c < chunks; ++c) t += s.part[c];
    a.scal[axis * 8 + slot] = t;
}

// r = b - Ap; z = r / diag; p = z (simpl_lss.cc:221-230)
__global__ void __launch_bounds__(kB) exStartVectors(ExactArgs a) {
    const int axis = blockIdx.y;
    const ExactAxis &s = a.ax[axis];
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= s.rows) return;
    s.r[i] = s.rhs[i] - s.ap[i];
    s.z[i] = s.r[i] / s.diag[i];
    s.p[i] = s.z[i];
}

// x += a p; r -= a Ap; z = r / diag (simpl_lss.cc:247-251)
__global__ void __launch_bounds__(kB) exUpdate(ExactArgs a, const int *live, const double *alpha) {
    const int axis = blockIdx.y;
    if (!live[axis]) return;
    const ExactAxis &s = a.ax[axis];
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= s.rows) return;
    const double al = alpha[axis];
    s.x[i] += al * s.p[i];
    s.r[i] -= al * s.ap[i];
    s.z[i] = s.r[i] / s.diag[i];
}

// p = z + beta p (simpl_lss.cc:259-261)
__global__ void __launch_bounds__(kB) exPUpdate(ExactArgs a, const int *live, const double *beta) {
    const int axis = blockIdx.y;
    if (!live[axis]) return;
    const ExactAxis &s = a.ax[axis];
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= s.rows) return;
    s.p[i] = s.z[i] + beta[axis] * s.p[i];
}

}  // namespace

// Host-driven loop: the scalar steps run on the host on values read back from
// the device (the same IEEE double operations as the reference's host code).
int solveExact(kt_ctx *ctx, const kt_solve_opts &o, kt_solve_result &res) {
    const uint64_t rows = ctx->sys[0].rows;
    KT_CHECK(ctx->sys[1].rows == rows, KT_ESTATE, "solve: x and y systems differ in size");
    KT_CHECK(ctx->posCount == rows, KT_ESTATE, "solve: positions do not match the systems");
    ExactArgs a{};
    const uint32_t chunks = static_cast<uint32_t>((rows + kChunk - 1) / kChunk);
    KT_TRY(ctx->exact.reserve(2ull * (5 * std::max<uint64_t>(rows, 1) + chunks + 1) + 64));
    double *base = ctx->exact.p;
    for (int ax = 0; ax < 2; ++ax) {
        Sys &s = ctx->sys[ax];
        ExactAxis &e = a.ax[ax];
        e.rowPtr = s.rowPtr.p;
        e.col = s.col.p;
        e.val = s.val.p;
        e.diag = s.diag.p;
        e.rhs = s.rhs.p;
        e.x = ax == 0 ? ctx->posX.p : ctx->posY.p;
        e.r = base;
        e.z = base + rows;
        e.p = base + 2 * rows;
        e.ap = base + 3 * rows;
        e.part = base + 4 * rows;
        e.rows = static_cast<uint32_t>(rows);
        base += 4 * rows + chunks + 1;
    }
    a.scal = base;  // 16 doubles
    int *liveDev = reinterpret_cast<int *>(base + 16);
    double *alphaDev = base + 20;
    double *betaDev = base + 22;
    res = kt_solve_result{};
    if (rows == 0) return KT_OK;
    const dim3 gRows((static_cast<uint32_t>(rows) + kB - 1) / kB, 2);
    const dim3 gChunks((chunks + kB - 1) / kB, 2);
    double *host = ctx->pinned;        // 16 scalars read back
    int *stepHost = reinterpret_cast<int *>(ctx->pinned + 16);  // 3 x 2 flags
    double *coefHost = ctx->pinned + 24;                        // alpha[2], beta[2]
    const auto dot = [&](int which, const int *live, int slot) -> int {
        exChunkDot<<<gChunks, kB, 0, ctx->stream>>>(a, which, live);
        KT_HIP(hipGetLastError());
        exChunkTotal<<<1, 64, 0, ctx->stream>>>(a, live, slot);
        KT_HIP(hipGetLastError());
        return KT_OK;
    };
    const auto readScal = [&]() -> int {
        KT_HIP(hipMemcpyAsync(host, a.scal, 16 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        KT_HIP(hipStreamSynchronize(ctx->stream));
        return KT_OK;
    };
    // start (simpl_lss.cc:216-236)
    exMultiply<<<gRows, kB, 0, ctx->stream>>>(a, 0, nullptr);
    KT_HIP(hipGetLastError());
    exStartVectors<<<gRows, kB, 0, ctx->stream>>>(a);
    KT_HIP(hipGetLastError());
    KT_TRY(dot(0, nullptr, 0));
    KT_TRY(dot(1, nullptr, 1));
    KT_TRY(dot(2, nullptr, 2));
    KT_TRY(readScal());
    struct Axis {
        double bNorm = 0.0, rho = 0.0, resid = 0.0;
        bool active = false;
        uint64_t steps = 0;
    } ax[2];
    for (int i = 0; i < 2; ++i) {
        ax[i].bNorm = std::sqrt(host[i * 8 + 0]);
        if (!(ax[i].bNorm > 0.0)) continue;
        ax[i].rho = host[i * 8 + 1];
        ax[i].resid = std::sqrt(host[i * 8 + 2]) / ax[i].bNorm;
        ax[i].active = true;
    }
    const auto live = [&](const Axis &x) { return x.active && x.resid > o.tol; };
    std::vector<double> hist;
    uint64_t it = 0;
    for (; it < o.max_iter && (live(ax[0]) || live(ax[1])); ++it) {
        int *step0 = stepHost, *step1 = stepHost + 2, *step2 = stepHost + 4;
        step0[0] = live(ax[0]) ? 1 : 0;
        step0[1] = live(ax[1]) ? 1 : 0;
        KT_HIP(hipMemcpyAsync(liveDev, step0, 2 * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
        exMultiply<<<gRows, kB, 0, ctx->stream>>>(a, 1, liveDev);
        KT_HIP(hipGetLastError());
        KT_TRY(dot(3, liveDev, 3));
        KT_TRY(readScal());
        double *alpha = coefHost;
        step1[0] = step0[0];
        step1[1] = step0[1];
        for (int i = 0; i < 2; ++i) {
            alpha[i] = 0.0;
            if (!step1[i]) continue;
            ++ax[i].steps;
            const double pap = host[i * 8 + 3];
            if (!(pap > 0.0)) {
                ax[i].active = false;
                step1[i] = 0;
                continue;
            }
            alpha[i] = ax[i].rho / pap;
        }
        KT_HIP(hipMemcpyAsync(liveDev, step1, 2 * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
        KT_HIP(hipMemcpyAsync(alphaDev, alpha, 2 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        exUpdate<<<gRows, kB, 0, ctx->stream>>>(a, liveDev, alphaDev);
        KT_HIP(hipGetLastError());
        KT_TRY(dot(2, liveDev, 2));
        KT_TRY(dot(1, liveDev, 1));
        KT_TRY(readScal());
        double *beta = coefHost + 2;
        step2[0] = step1[0];
        step2[1] = step1[1];
        for (int i = 0; i < 2; ++i) {
            beta[i] = 0.0;
            if (!step2[i]) continue;
            ax[i].resid = std::sqrt(host[i * 8 + 2]) / ax[i].bNorm;
            const double rhoNext = host[i * 8 + 1];
            if (!(rhoNext > 0.0)) {
                ax[i].active = false;
                step2[i] = 0;
                continue;
            }
            beta[i] = rhoNext / ax[i].rho;
            ax[i].rho = rhoNext;
        }
        KT_HIP(hipMemcpyAsync(liveDev + 2, step2, 2 * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
        KT_HIP(hipMemcpyAsync(betaDev, beta, 2 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        exPUpdate<<<gRows, kB, 0, ctx->stream>>>(a, liveDev + 2, betaDev);
        KT_HIP(hipGetLastError());
        hist.push_back(ax[0].resid);
        hist.push_back(ax[1].resid);
    }
    // true residual of the final iterate
    exMultiply<<<gRows, kB, 0, ctx->stream>>>(a, 0, nullptr);
    KT_HIP(hipGetLastError());
    KT_TRY(dot(4, nullptr, 4));
    KT_TRY(readScal());
    for (int i = 0; i < 2; ++i) {
        res.resid[i] = ax[i].resid;
        res.true_resid[i] = ax[i].bNorm > 0.0 ? std::sqrt(host[i * 8 + 4]) / ax[i].bNorm : 0.0;
        res.steps[i] = ax[i].steps;
        ctx->ctlHost[i].s[0].active = ax[i].active ? 1 : 0;
    }
    res.iterations = it;
    ctx->exactHistory = std::move(hist);
    return KT_OK;
}

}  // namespace kt
