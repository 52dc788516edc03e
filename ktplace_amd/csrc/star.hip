// Star net model for the initial placement rounds on gfx950.
//
// Follows LinearSystems::buildStar / buildStarAxis (simpl_lss.cc:116-201) and
// produces the same CSR bit for bit.  Variables: the n movable cells, then one
// star variable per net with at least three pins and at least one movable pin,
// in net order.  A 2-pin net is a spring of weight w between its pins; a net
// with a star joins every pin to the star with w = weight * p / (p - 1).
// Springs to fixed pins go to diag / rhs only.  Entries of a row appear in
// the order the reference appends them: a cell's row in ascending net order
// (its pinsOfCell order), a star's row in the net's pin order; diag and rhs
// accumulate in the same order with the same expressions.  The systems do not
// depend on positions; x and y differ only in their right-hand sides.
//
// Kernels: starFlags (nets that get a star), scan -> star index, starNets
// (star -> net), starCounts (row lengths), scan -> rowPtr, starFill.
#include <algorithm>

#include "kt_internal.h"

namespace kt {

int scanInclusive(kt_ctx *ctx, uint32_t *data, uint64_t m);
int scanRowPtr(kt_ctx *ctx, uint32_t *rowPtr, uint64_t rows);

namespace {

constexpr int kB = 256;

inline dim3 gridFor(uint64_t n, uint32_t y = 1) {
    return dim3(static_cast<uint32_t>(std::max<uint64_t>((n + kB - 1) / kB, 1)), y);
}

__global__ void __launch_bounds__(kB) starFlags(const uint32_t *__restrict__ netStart,
                                                const uint32_t *__restrict__ netMov, uint32_t N,
                                                uint32_t *__restrict__ flag) {
    const uint32_t j = blockIdx.x * kB + threadIdx.x;
    if (j >= N) return;
    flag[j] = (netStart[j + 1] - netStart[j] >= 3 && netMov[j] > 0) ? 1u : 0u;
}

// flag holds the inclusive scan: net j has star n + flag[j] - 1 when it has one
__device__ __forceinline__ uint32_t starOf(const uint32_t *__restrict__ netStart, const uint32_t *__restrict__ netMov,
                                           const uint32_t *__restrict__ scan, uint32_t j, uint32_t n) {
    const bool has = netStart[j + 1] - netStart[j] >= 3 && netMov[j] > 0;
    return has ? n + scan[j] - 1 : kNoVar;
}

__global__ void __launch_bounds__(kB) starNets(const uint32_t *__restrict__ netStart,
                                               const uint32_t *__restrict__ netMov,
                                               const uint32_t *__restrict__ scan, uint32_t N, uint32_t n,
                                               uint32_t *__restrict__ starNet) {
    const uint32_t j = blockIdx.x * kB + threadIdx.x;
    if (j >= N) return;
    const uint32_t s = starOf(netStart, netMov, scan, j, n);
    if (s != kNoVar) starNet[s - n] = j;
}

struct StarArgs {
    const uint32_t *pinStart, *refNet, *refPin;
    const uint32_t *netStart, *pinCell, *varOfVertex, *netMov, *scan, *starNet;
    const double *weight;
    const double *off[2];
    const double *fixed[2];
    uint32_t *rowPtr[2];
    uint32_t *col[2];
    double *val[2];
    double *diag[2];
    double *rhs[2];
    uint32_t n, size;
};

__global__ void __launch_bounds__(kB) starCounts(StarArgs a) {
    const uint32_t r = blockIdx.x * kB + threadIdx.x;
    if (r >= a.size) return;
    uint32_t count = 0;
    if (r < a.n) {
        for (uint32_t e = a.pinStart[r]; e < a.pinStart[r + 1]; ++e) {
            const uint32_t net = a.refNet[e];
            const uint32_t b = a.netStart[net], p = a.netStart[net + 1] - b;
            if (p == 2) {
                const uint32_t other = a.pinCell[b + (a.refPin[e] == 0 ? 1 : 0)];
                count += a.varOfVertex[other] != kNoVar ? 1u : 0u;
            } else if (p >= 3) {
                count += 1;  // the net has a star: this cell is movable
            }
        }
    } else {
        count = a.netMov[a.starNet[r - a.n]];
    }
    for (int axis = 0; axis < 2; ++axis) {
        a.rowPtr[axis][r + 1] = count;
        if (r == 0) a.rowPtr[axis][0] = 0;
    }
}

// std::max(a, b) is (a < b) ? b : a
__device__ __forceinline__ double refMax(double a, double b) { return a < b ? b : a; }

__global__ void __launch_bounds__(kB) starFill(StarArgs a) {
    const uint32_t r = blockIdx.x * kB + threadIdx.x;
    const int axis = blockIdx.y;
    if (r >= a.size) return;
    const double *__restrict__ off = a.off[axis];
    const double *__restrict__ fixed = a.fixed[axis];
    uint32_t *__restrict__ col = a.col[axis];
    double *__restrict__ val = a.val[axis];
    uint32_t slot = a.rowPtr[axis][r];
    double diag = 0.0, rhs = 0.0;
    if (r < a.n) {
        // pull(self = this cell, other) for each of its nets, in net order
        for (uint32_t e = a.pinStart[r]; e < a.pinStart[r + 1]; ++e) {
            const uint32_t net = a.refNet[e];
            const uint32_t b = a.netStart[net], p = a.netStart[net + 1] - b;
            const uint32_t q = a.refPin[e];
            const double selfOffset = off[b + q];
            if (p == 2) {
                const uint32_t oq = q == 0 ? 1 : 0;
                const uint32_t cell = a.pinCell[b + oq];
                const uint32_t var = a.varOfVertex[cell];
                const double w = a.weight[net];
                diag += w;
                if (var != kNoVar) {
                    col[slot] = var;
                    val[slot] = -w;
                    ++slot;
                    rhs += w * (off[b + oq] - selfOffset);
                } else {
                    rhs += w * (fixed[cell] + off[b + oq] - selfOffset);
                }
            } else if (p >= 3) {
                const double w = a.weight[net] * static_cast<double>(p) / static_cast<double>(p - 1);
                diag += w;
                col[slot] = starOf(a.netStart, a.netMov, a.scan, net, a.n);
                val[slot] = -w;
                ++slot;
                rhs += w * (0.0 - selfOffset);
            }
        }
    } else {
        // pull(self = star, other = each pin of the net), in pin order
        const uint32_t net = a.starNet[r - a.n];
        const uint32_t b = a.netStart[net], p = a.netStart[net + 1] - b;
        const double w = a.weight[net] * static_cast<double>(p) / static_cast<double>(p - 1);
        for (uint32_t q = 0; q < p; ++q) {
            const uint32_t cell = a.pinCell[b + q];
            const uint32_t var = a.varOfVertex[cell];
            diag += w;
            if (var != kNoVar) {
                col[slot] = var;
                val[slot] = -w;
                ++slot;
                rhs += w * (off[b + q] - 0.0);
            } else {
                rhs += w * (fixed[cell] + off[b + q] - 0.0);
            }
        }
    }
    a.diag[axis][r] = refMax(diag, 1e-12);
    a.rhs[axis][r] = rhs;
}

}  // namespace

int ensureSystem(kt_ctx *ctx, int axis, uint64_t rows, uint64_t nnzCap);

// Builds both star systems; *size = n + stars.
int buildStar(kt_ctx *ctx, uint64_t *sizeOut) {
    KT_CHECK(ctx->haveNetlist, KT_ESTATE, "buildStar: no netlist uploaded");
    const uint32_t n = ctx->n, N = ctx->N;
    KT_TRY(ctx->starScan.reserve(std::max<uint32_t>(N, 1)));
    uint32_t stars = 0;
    if (N) {
        starFlags<<<gridFor(N), kB, 0, ctx->stream>>>(ctx->netStart.p, ctx->netMov.p, N, ctx->starScan.p);
        KT_HIP(hipGetLastError());
        KT_TRY(scanInclusive(ctx, ctx->starScan.p, N));
        KT_HIP(hipMemcpyAsync(&stars, ctx->starScan.p + N - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
        KT_HIP(hipStreamSynchronize(ctx->stream));
    }
    const uint64_t size = static_cast<uint64_t>(n) + stars;
    KT_CHECK(size < 0xffffffffull, KT_EINVAL, "buildStar: too many variables");
    KT_TRY(ctx->starNet.reserve(std::max<uint32_t>(stars, 1)));
    if (N) {
        starNets<<<gridFor(N), kB, 0, ctx->stream>>>(ctx->netStart.p, ctx->netMov.p, ctx->starScan.p, N, n,
                                                    ctx->starNet.p);
        KT_HIP(hipGetLastError());
    }
    // at most one entry per pin in a cell row and one per pin in a star row
    const uint64_t cap = std::max<uint64_t>(2 * ctx->P, 1);
    for (int axis = 0; axis < 2; ++axis) KT_TRY(ensureSystem(ctx, axis, size, cap));
    StarArgs a{};
    a.pinStart = ctx->pinStart.p;
    a.refNet = ctx->refNet.p;
    a.refPin = ctx->refPin.p;
    a.netStart = ctx->netStart.p;
    a.pinCell = ctx->pinCell.p;
    a.varOfVertex = ctx->varOfVertex.p;
    a.netMov = ctx->netMov.p;
    a.scan = ctx->starScan.p;
    a.starNet = ctx->starNet.p;
    a.weight = ctx->weight.p;
    a.off[0] = ctx->offX.p;
    a.off[1] = ctx->offY.p;
    a.fixed[0] = ctx->fixedX.p;
    a.fixed[1] = ctx->fixedY.p;
    for (int axis = 0; axis < 2; ++axis) {
        Sys &s = ctx->sys[axis];
        a.rowPtr[axis] = s.rowPtr.p;
        a.col[axis] = s.col.p;
        a.val[axis] = s.val.p;
        a.diag[axis] = s.diag.p;
        a.rhs[axis] = s.rhs.p;
        s.nnzHost = -1;
        s.have32 = false;
        s.have24 = false;
        s.havePk = false;
    }
    a.n = n;
    a.size = static_cast<uint32_t>(size);
    ctx->lastAlpha = 0.0;
    if (size == 0) {
        for (int axis = 0; axis < 2; ++axis)
            KT_HIP(hipMemsetAsync(ctx->sys[axis].rowPtr.p, 0, sizeof(uint32_t), ctx->stream));
    } else {
        starCounts<<<gridFor(size), kB, 0, ctx->stream>>>(a);
        KT_HIP(hipGetLastError());
        for (int axis = 0; axis < 2; ++axis) KT_TRY(scanRowPtr(ctx, ctx->sys[axis].rowPtr.p, size));
        starFill<<<gridFor(size, 2), kB, 0, ctx->stream>>>(a);
        KT_HIP(hipGetLastError());
    }
    KT_HIP(hipStreamSynchronize(ctx->stream));
    if (sizeOut) *sizeOut = size;
    return KT_OK;
}

}  // namespace kt
