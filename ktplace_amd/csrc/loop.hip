// Device-resident bookkeeping of SimPL's outer loop: the HPWL of the lower
// and upper bounds (getHpwl, simpl_placer.cc:256-275), the overflow of the
// upper bound (getOccupancy / getOverflow, :282-297) and Convergence::update
// (:450-461), computed on the context's stream where the placements already
// are.  The upper bound lives in the anchor buffers (it is the next B2B
// build's anchor set, simpl_placer.cc:516) and the lower bound in the
// positions, so an outer iteration moves no placement across PCIe: the host
// reads back one small state record per iteration.
#include <algorithm>
#include <cfloat>

#include "kt_internal.h"

namespace kt {

int lalHpwlInto(kt_ctx *ctx, const double *x, const double *y, double *dst);
int lalExcessInto(kt_ctx *ctx, const double *x, const double *y, double *dst);

namespace {

// Layout of ctx->loopDev (doubles; counters are stored as exact doubles).
enum LoopSlot : int {
    kUpperHpwl = 0,  // HPWL of the anchors (upper bound)
    kLowerHpwl,      // HPWL of the positions (lower bound)
    kExcessUpper,    // overflow excess of the upper bound
    kExcessLower,    // overflow excess of the lower bound (measure-only calls)
    kGap,
    kOverflowUpper,
    kOverflowLower,
    kGapReference,
    kBestUpper,
    kStalledFor,
    kIteration,
    kConverged,
    kReferenceIter,  // options
    kTightFrac,
    kRelaxedFrac,
    kStallIters,
    kLoopSlots
};

__global__ void loopReset(double *s, double referenceIter, double tight, double relaxed, double stall) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int k = 0; k < kLoopSlots; ++k) s[k] = 0.0;
    s[kBestUpper] = DBL_MAX;
    s[kReferenceIter] = referenceIter;
    s[kTightFrac] = tight;
    s[kRelaxedFrac] = relaxed;
    s[kStallIters] = stall;
}

// Convergence::update for `iteration` (>= 1) on the HPWLs just computed, and
// the overflows (excess / total free area, getOverflow).  iteration 0 only
// fills in the measurements.
__global__ void loopUpdate(double *s, double iteration, double totalAvail, int withLower) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double upper = s[kUpperHpwl], lower = s[kLowerHpwl];
    const double gap = upper - lower;
    s[kGap] = gap;
    s[kOverflowUpper] = totalAvail > 0.0 ? s[kExcessUpper] / totalAvail : 0.0;
    s[kOverflowLower] = withLower ? (totalAvail > 0.0 ? s[kExcessLower] / totalAvail : 0.0) : 0.0;
    s[kIteration] = iteration;
    if (iteration == 0.0) {
        s[kConverged] = 0.0;
        return;
    }
    if (iteration == s[kReferenceIter]) s[kGapReference] = gap;
    s[kStalledFor] = upper < s[kBestUpper] ? 0.0 : s[kStalledFor] + 1.0;
    s[kBestUpper] = upper < s[kBestUpper] ? upper : s[kBestUpper];
    const double ref = s[kGapReference];
    const bool stalled = s[kStalledFor] >= s[kStallIters];
    const bool converged = iteration > s[kReferenceIter] && ref > 0.0 &&
                           (gap < s[kTightFrac] * ref || (gap <= s[kRelaxedFrac] * ref && stalled));
    s[kConverged] = converged ? 1.0 : 0.0;
}

__global__ void copyPair(const double *__restrict__ ax, const double *__restrict__ ay, double *__restrict__ bx,
                         double *__restrict__ by, uint32_t n) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    bx[i] = ax[i];
    by[i] = ay[i];
}

int loopBuffers(kt_ctx *ctx) {
    KT_TRY(ctx->loopDev.reserve(kLoopSlots));
    if (ctx->loopHost == nullptr) KT_HIP(hipHostMalloc(reinterpret_cast<void **>(&ctx->loopHost), kLoopSlots * sizeof(double)));
    return KT_OK;
}

}  // namespace

int loopReset(kt_ctx *ctx, uint64_t referenceIter, double tight, double relaxed, uint64_t stall) {
    KT_TRY(loopBuffers(ctx));
    loopReset<<<1, 64, 0, ctx->stream>>>(ctx->loopDev.p, static_cast<double>(referenceIter), tight, relaxed,
                                         static_cast<double>(stall));
    KT_HIP(hipGetLastError());
    return KT_OK;
}

int anchorsFromPositions(kt_ctx *ctx) {
    KT_CHECK(ctx->haveNetlist && ctx->posCount == ctx->n, KT_ESTATE,
             "anchors from positions: positions must have one entry per movable cell");
    const uint32_t n = ctx->n;
    KT_TRY(ctx->anchorX.reserve(std::max<uint32_t>(n, 1)));
    KT_TRY(ctx->anchorY.reserve(std::max<uint32_t>(n, 1)));
    if (n) {
        copyPair<<<(n + 255) / 256, 256, 0, ctx->stream>>>(ctx->posX.p, ctx->posY.p, ctx->anchorX.p, ctx->anchorY.p,
                                                           n);
        KT_HIP(hipGetLastError());
    }
    ctx->haveAnchors = true;
    return KT_OK;
}

int loopUpdate(kt_ctx *ctx, uint64_t iteration, kt_loop_state *out) {
    KT_CHECK(ctx->haveNetlist && ctx->posCount == ctx->n, KT_ESTATE,
             "loop update: positions must have one entry per movable cell");
    KT_CHECK(ctx->haveAnchors, KT_ESTATE, "loop update: no upper bound (anchors) yet");
    KT_CHECK(ctx->grid.size() > 0, KT_ESTATE, "loop update: build the grid first");
    KT_CHECK(ctx->loopDev.cap >= kLoopSlots && ctx->loopHost, KT_ESTATE, "loop update: reset the loop state first");
    double *s = ctx->loopDev.p;
    KT_TRY(lalHpwlInto(ctx, ctx->anchorX.p, ctx->anchorY.p, s + kUpperHpwl));
    KT_TRY(lalHpwlInto(ctx, ctx->posX.p, ctx->posY.p, s + kLowerHpwl));
    KT_TRY(lalExcessInto(ctx, ctx->anchorX.p, ctx->anchorY.p, s + kExcessUpper));
    const bool withLower = iteration == 0;
    if (withLower) KT_TRY(lalExcessInto(ctx, ctx->posX.p, ctx->posY.p, s + kExcessLower));
    loopUpdate<<<1, 64, 0, ctx->stream>>>(s, static_cast<double>(iteration), ctx->grid.totalAvail, withLower ? 1 : 0);
    KT_HIP(hipGetLastError());
    KT_HIP(hipMemcpyAsync(ctx->loopHost, s, kLoopSlots * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    KT_HIP(hipStreamSynchronize(ctx->stream));
    const double *h = ctx->loopHost;
    out->lower_hpwl = h[kLowerHpwl];
    out->upper_hpwl = h[kUpperHpwl];
    out->gap = h[kGap];
    out->overflow_upper = h[kOverflowUpper];
    out->overflow_lower = h[kOverflowLower];
    out->gap_reference = h[kGapReference];
    out->best_upper = h[kBestUpper];
    out->iteration = static_cast<uint64_t>(h[kIteration]);
    out->stalled_for = static_cast<uint64_t>(h[kStalledFor]);
    out->converged = h[kConverged] != 0.0 ? 1 : 0;
    out->reserved = 0;
    return KT_OK;
}

}  // namespace kt
