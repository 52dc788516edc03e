// Bound2Bound Laplacian assembly for x and y on gfx950.
//
// Follows LinearSystems::buildB2B / buildB2BAxis (simpl_lss.cc:11-114) and
// produces the same CSR bit for bit: the reference's row order, entry order
// within a row (pinsOfCell order; bound pin -> every other pin in index order,
// inner pin -> lo then hi), duplicates kept, fixed neighbours dropped from the
// CSR but kept in diag/rhs, and the same IEEE-754 double expressions in the
// same order (the library is compiled with -ffp-contract=off, so no multiply
// is fused into an add).
//
// Kernels (both axes per launch, blockIdx.y = axis where it applies):
//   vertexPositions  px/py = fixed positions with movable cells overwritten (:14-19)
//   pinStage         per pin: position (vertex + offset), rhs term, variable
//                    (then one record per net: pin range, weight base, lo/hi)
//   netExtremes      first argmin / first argmax of pos + offset per net  (:35-45)
//   rowCounts        movable neighbours per row                              (:75-82)
//   scan             rowPtr = inclusive scan of the counts                  (:83)
//   rowFill          col, val = -w, diag, rhs, anchor term                  (:89-113)
#include "kt_internal.h"

namespace kt {
namespace {

constexpr int kBlock = 256;
constexpr int kScanItems = 8;
constexpr uint64_t kScanTile = kBlock * kScanItems;

__global__ void __launch_bounds__(kBlock)
vertexPositions(const double *__restrict__ x, const double *__restrict__ y,
                const uint32_t *__restrict__ varOfVertex, const double *__restrict__ fixedX,
                const double *__restrict__ fixedY, double *__restrict__ vx, double *__restrict__ vy,
                uint32_t V) {
    const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
    if (v >= V) return;
    const uint32_t var = varOfVertex[v];
    vx[v] = var != kNoVar ? x[var] : fixedX[v];
    vy[v] = var != kNoVar ? y[var] : fixedY[v];
}

// One thread per pin, both axes.  The pin's position (pinAt: vertex position
// + offset, simpl_lss.cc:32-34), its variable (kNoVar for a fixed vertex) and
// the term an edge to it adds to rhs before the owner's offset is taken off:
// position - vertex position for a movable vertex, the position itself for a
// fixed one (simpl_lss.cc:94-104).  These are the same double expressions the
// per-entry code would evaluate, once per pin instead of once per CSR entry
// that names the pin, so the entry loops below read them in net order
// instead of gathering vertex data at random.
// Stored per axis as one (position, rhs term) pair per pin: the row kernels
// read both at once, at random for a net's lo/hi pins and a row's own pin.
__global__ void __launch_bounds__(kBlock)
pinStage(const uint32_t *__restrict__ pinCell, const uint32_t *__restrict__ varOfVertex,
         const double *__restrict__ offX, const double *__restrict__ offY, const double *__restrict__ vx,
         const double *__restrict__ vy, double2 *__restrict__ pinX, double2 *__restrict__ pinY,
         uint32_t *__restrict__ pinVar, uint64_t P) {
    const uint64_t k = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (k >= P) return;
    const uint32_t c = pinCell[k];
    const uint32_t var = varOfVertex[c];
    const double cx = vx[c], cy = vy[c];
    const double px = cx + offX[k], py = cy + offY[k];
    pinX[k] = make_double2(px, var != kNoVar ? px - cx : px);
    pinY[k] = make_double2(py, var != kNoVar ? py - cy : py);
    pinVar[k] = var;
}

// Everything a row's reference to a net reads about the net, in one 64-byte
// record (one memory sector): its pin range, movable pins and edge weight
// base (weight / (pins - 1), simpl_lss.cc:55), and this build's first argmin /
// argmax per axis with their pins' variables.  The row kernels visit nets in
// random order, so one sector per visit instead of one per array.
struct alignas(64) NetRec {
    uint32_t start, size, mov, pad;
    uint32_t lox, hix, loy, hiy;
    uint32_t varLoX, varHiX, varLoY, varHiY;
    double base;
    double unused;
};
static_assert(sizeof(NetRec) == 64, "one sector per net");

// One thread per net; the running best value is the reference's pinAt(lo),
// recomputed there from the same operands, so the comparisons are identical.
__global__ void __launch_bounds__(kBlock)
netExtremes(const uint32_t *__restrict__ netStart, const uint32_t *__restrict__ netMov,
            const double *__restrict__ weight, const double2 *__restrict__ pinX, const double2 *__restrict__ pinY,
            const uint32_t *__restrict__ pinVar, NetRec *__restrict__ rec, uint32_t N) {
    const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= N) return;
    const uint32_t b = netStart[j];
    const uint32_t p = netStart[j + 1] - b;
    uint32_t lox = 0, hix = 0, loy = 0, hiy = 0;
    if (p > 0) {
        double bestLoX = pinX[b].x, bestHiX = bestLoX;
        double bestLoY = pinY[b].x, bestHiY = bestLoY;
        for (uint32_t q = 1; q < p; ++q) {
            const double px = pinX[b + q].x;
            const double py = pinY[b + q].x;
            if (px < bestLoX) { lox = q; bestLoX = px; }
            if (px > bestHiX) { hix = q; bestHiX = px; }
            if (py < bestLoY) { loy = q; bestLoY = py; }
            if (py > bestHiY) { hiy = q; bestHiY = py; }
        }
    }
    uint4 *dst = reinterpret_cast<uint4 *>(rec + j);
    dst[0] = make_uint4(b, p, netMov[j], 0u);
    dst[1] = make_uint4(lox, hix, loy, hiy);
    dst[2] = p > 0 ? make_uint4(pinVar[b + lox], pinVar[b + hix], pinVar[b + loy], pinVar[b + hiy])
                   : make_uint4(kNoVar, kNoVar, kNoVar, kNoVar);
    // used only where lo != hi, i.e. p >= 2
    rec[j].base = weight[j] / static_cast<double>(p - 1);
}

struct RowArgs {
    const uint32_t *pinStart;
    const uint32_t *refNet;
    const uint32_t *refPin;
    const uint32_t *pinVar;
    const NetRec *rec;
    const double2 *pin[2];  // per pin: position, rhs term (pinStage)
    const double *pos[2];
    const double *anchor[2];
    uint32_t *rowPtr[2];
    uint32_t *col[2];
    double *val[2];
    double *diag[2];
    double *rhs[2];
    double floor[2];
    double alpha;
    uint32_t n;
};

__device__ __forceinline__ uint32_t extLo(const uint4 &e, int axis) { return axis ? e.z : e.x; }
__device__ __forceinline__ uint32_t extHi(const uint4 &e, int axis) { return axis ? e.w : e.y; }
__device__ __forceinline__ uint4 recHead(const NetRec *rec, uint32_t net) {
    return reinterpret_cast<const uint4 *>(rec + net)[0];  // start, size, mov
}
__device__ __forceinline__ uint4 recExt(const NetRec *rec, uint32_t net) {
    return reinterpret_cast<const uint4 *>(rec + net)[1];  // lo_x, hi_x, lo_y, hi_y
}
__device__ __forceinline__ uint4 recVars(const NetRec *rec, uint32_t net) {
    return reinterpret_cast<const uint4 *>(rec + net)[2];  // variables of lo_x, hi_x, lo_y, hi_y
}

// std::max(a, b) is (a < b) ? b : a; written out so NaN behaves the same.
__device__ __forceinline__ double refMax(double a, double b) { return a < b ? b : a; }

__global__ void __launch_bounds__(kBlock) rowCounts(RowArgs a) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const int axis = blockIdx.y;
    if (i >= a.n) return;
    uint32_t count = 0;
    for (uint32_t e = a.pinStart[i]; e < a.pinStart[i + 1]; ++e) {
        const uint32_t net = a.refNet[e];
        const uint32_t q = a.refPin[e];
        const uint4 head = recHead(a.rec, net), x = recExt(a.rec, net);
        const uint32_t lo = extLo(x, axis), hi = extHi(x, axis);
        if (lo == hi) continue;
        if (q == lo || q == hi) {
            count += head.z - 1;  // every other movable pin; q itself is movable
        } else {
            const uint4 v = recVars(a.rec, net);
            count += (extLo(v, axis) != kNoVar ? 1u : 0u) + (extHi(v, axis) != kNoVar ? 1u : 0u);
        }
    }
    a.rowPtr[axis][i + 1] = count;
    if (i == 0) a.rowPtr[axis][0] = 0;
}

__global__ void __launch_bounds__(kBlock) rowFill(RowArgs a) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const int axis = blockIdx.y;
    if (i >= a.n) return;
    const double2 *__restrict__ pin = a.pin[axis];
    const uint32_t *__restrict__ pinVar = a.pinVar;
    const double lengthFloor = a.floor[axis];
    uint32_t *__restrict__ col = a.col[axis];
    double *__restrict__ val = a.val[axis];
    uint32_t slot = a.rowPtr[axis][i];
    double diag = 0.0, rhs = 0.0;
    for (uint32_t e = a.pinStart[i]; e < a.pinStart[i + 1]; ++e) {
        const uint32_t net = a.refNet[e];
        const uint32_t q = a.refPin[e];
        const uint4 head = recHead(a.rec, net), x = recExt(a.rec, net);
        const uint32_t lo = extLo(x, axis), hi = extHi(x, axis);
        if (lo == hi) continue;
        const uint32_t b = head.x;
        const uint32_t p = head.y;
        const double base = a.rec[net].base;  // weight / (p - 1)
        // q's vertex is this movable row: .y = own - its vertex position
        const double2 ownPin = pin[b + q];
        const double own = ownPin.x;
        const double ownOffset = ownPin.y;
        auto edge = [&](uint32_t other, uint32_t var) {
            const double2 there = pin[b + other];
            const double w = base / refMax(fabs(own - there.x), lengthFloor);
            diag += w;
            if (var != kNoVar) {
                col[slot] = var;
                val[slot] = -w;
                ++slot;
            }
            // movable: (there - cellPos) - ownOffset; fixed: there - ownOffset
            rhs += w * (there.y - ownOffset);
        };
        if (q == lo || q == hi) {
            for (uint32_t other = 0; other < p; ++other) {
                if (other != q) edge(other, pinVar[b + other]);
            }
        } else {
            const uint4 v = recVars(a.rec, net);
            edge(lo, extLo(v, axis));
            edge(hi, extHi(v, axis));
        }
    }
    if (a.alpha > 0.0) {
        const double anchor = a.anchor[axis][i];
        const double w = a.alpha / refMax(fabs(a.pos[axis][i] - anchor), lengthFloor);
        diag += w;
        rhs += w * anchor;
    }
    a.diag[axis][i] = refMax(diag, 1e-12);
    a.rhs[axis][i] = rhs;
}

// Inclusive in-place scan of one tile of kScanTile elements; tile totals go to sums.
template <typename T>
__global__ void __launch_bounds__(kBlock)
scanTile(T *__restrict__ data, uint64_t m, T *__restrict__ sums) {
    __shared__ T waveSums[kBlock / 64];
    const uint64_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    T v[kScanItems];
    T s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint64_t idx = base + k;
        s += idx < m ? data[idx] : T(0);
        v[k] = s;
    }
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    T x = s;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const T y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) waveSums[wave] = x;
    __syncthreads();
    T waveOffset = 0;
    for (int w = 0; w < wave; ++w) waveOffset += waveSums[w];
    const T excl = waveOffset + x - s;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint64_t idx = base + k;
        if (idx < m) data[idx] = v[k] + excl;
    }
    if (threadIdx.x == kBlock - 1) sums[blockIdx.x] = waveOffset + x;
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
scanAdd(T *__restrict__ data, uint64_t m, const T *__restrict__ sums) {
    if (blockIdx.x == 0) return;
    const T add = sums[blockIdx.x - 1];
    const uint64_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint64_t idx = base + k;
        if (idx < m) data[idx] += add;
    }
}

template <typename T>
int scanLevel(kt_ctx *ctx, T *data, uint64_t m, T *tmp, uint64_t tmpCap) {
    if (m == 0) return KT_OK;
    const uint64_t blocks = (m + kScanTile - 1) / kScanTile;
    KT_CHECK(blocks <= tmpCap, KT_EINVAL, "scan: scratch too small");
    T *sums = tmp;
    scanTile<T><<<dim3(static_cast<uint32_t>(blocks)), dim3(kBlock), 0, ctx->stream>>>(data, m, sums);
    KT_HIP(hipGetLastError());
    if (blocks > 1) {
        KT_TRY(scanLevel<T>(ctx, sums, blocks, tmp + blocks, tmpCap - blocks));
        scanAdd<T><<<dim3(static_cast<uint32_t>(blocks)), dim3(kBlock), 0, ctx->stream>>>(data, m, sums);
        KT_HIP(hipGetLastError());
    }
    return KT_OK;
}

}  // namespace

uint64_t scanScratch(uint64_t m) {
    uint64_t total = 0;
    while (m > 1) {
        m = (m + kScanTile - 1) / kScanTile;
        total += m;
    }
    return total + 1;
}

// Inclusive in-place scans of m values.
int scanInclusive(kt_ctx *ctx, uint32_t *data, uint64_t m) {
    const uint64_t cap = scanScratch(m);
    KT_TRY(ctx->scanTmp.reserve(cap));
    return scanLevel<uint32_t>(ctx, data, m, ctx->scanTmp.p, cap);
}

int scanInclusive64(kt_ctx *ctx, uint64_t *data, uint64_t m) {
    const uint64_t cap = scanScratch(m);
    KT_TRY(ctx->scanTmp64.reserve(cap));
    return scanLevel<uint64_t>(ctx, data, m, ctx->scanTmp64.p, cap);
}

int scanRowPtr(kt_ctx *ctx, uint32_t *rowPtr, uint64_t rows) {
    return scanInclusive(ctx, rowPtr + 1, rows);
}

// px/py: fixed positions with the movable cells' positions written over
int vertexPositionsOf(kt_ctx *ctx, const double *x, const double *y) {
    KT_TRY(ctx->vx.reserve(std::max<uint32_t>(ctx->V, 1)));
    KT_TRY(ctx->vy.reserve(std::max<uint32_t>(ctx->V, 1)));
    if (ctx->V == 0) return KT_OK;
    vertexPositions<<<dim3((ctx->V + kBlock - 1) / kBlock), dim3(kBlock), 0, ctx->stream>>>(
        x, y, ctx->varOfVertex.p, ctx->fixedX.p, ctx->fixedY.p, ctx->vx.p, ctx->vy.p, ctx->V);
    KT_HIP(hipGetLastError());
    return KT_OK;
}

int vertexPositionsLaunch(kt_ctx *ctx) { return vertexPositionsOf(ctx, ctx->posX.p, ctx->posY.p); }

int buildB2B(kt_ctx *ctx, double alpha, double floorX, double floorY) {
    KT_CHECK(ctx->haveNetlist, KT_ESTATE, "buildB2B: no netlist uploaded");
    KT_CHECK(ctx->posCount == ctx->n, KT_ESTATE, "buildB2B: positions must have one entry per movable cell");
    KT_CHECK(!(alpha > 0.0) || ctx->haveAnchors, KT_ESTATE, "buildB2B: alpha > 0 needs anchors");
    const uint32_t n = ctx->n;
    for (int axis = 0; axis < 2; ++axis) KT_TRY(ensureSystem(ctx, axis, n, ctx->nnzBound));
    KT_TRY(ctx->ext.reserve(std::max<uint64_t>(16ull * ctx->N, 16)));  // one NetRec per net
    KT_TRY(vertexPositionsLaunch(ctx));
    const uint64_t pins = std::max<uint64_t>(ctx->P, 1);
    for (int axis = 0; axis < 2; ++axis) KT_TRY(ctx->pinAt[axis].reserve(2 * pins));
    KT_TRY(ctx->pinVar.reserve(pins));
    double2 *pinX = reinterpret_cast<double2 *>(ctx->pinAt[0].p);
    double2 *pinY = reinterpret_cast<double2 *>(ctx->pinAt[1].p);
    if (ctx->P > 0) {
        pinStage<<<dim3(static_cast<uint32_t>((ctx->P + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream>>>(
            ctx->pinCell.p, ctx->varOfVertex.p, ctx->offX.p, ctx->offY.p, ctx->vx.p, ctx->vy.p, pinX, pinY,
            ctx->pinVar.p, ctx->P);
        KT_HIP(hipGetLastError());
    }
    NetRec *rec = reinterpret_cast<NetRec *>(ctx->ext.p);
    if (ctx->N > 0) {
        netExtremes<<<dim3((ctx->N + kBlock - 1) / kBlock), dim3(kBlock), 0, ctx->stream>>>(
            ctx->netStart.p, ctx->netMov.p, ctx->weight.p, pinX, pinY, ctx->pinVar.p, rec, ctx->N);
        KT_HIP(hipGetLastError());
    }
    RowArgs a{};
    a.pinStart = ctx->pinStart.p;
    a.refNet = ctx->refNet.p;
    a.refPin = ctx->refPin.p;
    a.pinVar = ctx->pinVar.p;
    a.rec = rec;
    a.pin[0] = pinX;
    a.pin[1] = pinY;
    a.pos[0] = ctx->posX.p;
    a.pos[1] = ctx->posY.p;
    a.anchor[0] = ctx->anchorX.p;
    a.anchor[1] = ctx->anchorY.p;
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
    }
    a.floor[0] = floorX;
    a.floor[1] = floorY;
    a.alpha = alpha;
    ctx->lastAlpha = alpha;
    a.n = n;
    if (n == 0) {
        for (int axis = 0; axis < 2; ++axis)
            KT_HIP(hipMemsetAsync(ctx->sys[axis].rowPtr.p, 0, sizeof(uint32_t), ctx->stream));
        return KT_OK;
    }
    const dim3 grid((n + kBlock - 1) / kBlock, 2);
    rowCounts<<<grid, dim3(kBlock), 0, ctx->stream>>>(a);
    KT_HIP(hipGetLastError());
    for (int axis = 0; axis < 2; ++axis) KT_TRY(scanRowPtr(ctx, ctx->sys[axis].rowPtr.p, n));
    hipEvent_t t0 = nullptr;
    KT_TRY(timedBegin(ctx, KT_KERNEL_B2B_FILL, &t0));
    rowFill<<<grid, dim3(kBlock), 0, ctx->stream>>>(a);
    KT_HIP(hipGetLastError());
    KT_TRY(timedEnd(ctx, KT_KERNEL_B2B_FILL, t0));
    return KT_OK;
}

}  // namespace kt
