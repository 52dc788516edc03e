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
//   refCounts        per pinsOfCell reference: edges and CSR entries     (:47-82)
//   scans            each reference's first edge and CSR slot; rowPtr     (:83)
//   edgeFill         w and rhs term per edge, col / val = -w              (:89-104)
//                    and diag / rhs of the rows inside its edge block     (:105-113)
//   rowSumsRest      diag / rhs of the rows across edge blocks or with no edges
#include "kt_internal.h"

namespace kt {
namespace {

constexpr int kBlock = 256;
constexpr int kScanItems = 8;
constexpr uint64_t kScanTile = kBlock * kScanItems;

__global__ void __launch_bounds__(kBlock)
vertexPositions(const double *__restrict__ x, const double *__restrict__ y,
                const uint32_t *__restrict__ varOfVertex, const double *__restrict__ fixedX,
                const double *__restrict__ fixedY, double2 *__restrict__ vxy, uint32_t V) {
    const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
    if (v >= V) return;
    const uint32_t var = varOfVertex[v];
    vxy[v] = var != kNoVar ? make_double2(x[var], y[var]) : make_double2(fixedX[v], fixedY[v]);
}

// One thread per pin, both axes.  The pin's position (vertex position +
// offset, simpl_lss.cc:32-34) and the term an edge to it adds to rhs before
// the owner's offset is taken off: position - vertex position for a movable
// vertex, the position itself for a fixed one (simpl_lss.cc:94-104).  These
// are the same double expressions the per-edge code would evaluate, once per
// pin instead of once per edge that names the pin.  Both axes share one
// 32-byte record, and the pin's variable and its movable rank in the net one
// 8-byte pair, so an edge reads one line of each for both axes.
struct alignas(32) PinRec {
    double posX, termX, posY, termY;
};
__device__ __forceinline__ double pinPos(const PinRec &r, int axis) { return axis ? r.posY : r.posX; }
__device__ __forceinline__ double pinTerm(const PinRec &r, int axis) { return axis ? r.termY : r.termX; }

__global__ void __launch_bounds__(kBlock)
pinStage(const uint32_t *__restrict__ pinCell, const uint2 *__restrict__ varMov, const double *__restrict__ offX,
         const double *__restrict__ offY, const double2 *__restrict__ vxy, PinRec *__restrict__ pin, uint64_t P) {
    const uint64_t k = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (k >= P) return;
    const uint32_t c = pinCell[k];
    const bool movable = varMov[k].x != kNoVar;
    const double2 cv = vxy[c];
    const double cx = cv.x, cy = cv.y;
    const double px = cx + offX[k], py = cy + offY[k];
    pin[k] = PinRec{px, movable ? px - cx : px, py, movable ? py - cy : py};
}

// Everything a row's reference to a net reads about the net, in one 128-byte
// record (one cache line): its pin range, movable pins and edge weight base
// (weight / (pins - 1), simpl_lss.cc:55), this build's first argmin / argmax
// per axis with their pins' variables, and those two pins' position and rhs
// term per axis.  The references visit nets in random order, so one line per
// visit instead of one per array; and an edge to lo or hi (every edge of an
// inner pin, the one edge of a two-pin net) finds the far pin in the record
// instead of a pin record of its own.
struct alignas(128) NetRec {
    uint32_t start, size, mov, pad;
    uint32_t lox, hix, loy, hiy;
    uint32_t varLoX, varHiX, varLoY, varHiY;
    double base;
    double unused;
    double ext[2][4];  // per axis: position and rhs term of lo, then of hi
};
static_assert(sizeof(NetRec) == 128, "one line per net");

__device__ __forceinline__ void writeNetRec(NetRec *r, uint32_t b, uint32_t p, uint32_t mov, double weight,
                                            const uint2 *__restrict__ varMov, uint32_t lox, uint32_t hix,
                                            uint32_t loy, uint32_t hiy, const PinRec &pLoX, const PinRec &pHiX,
                                            const PinRec &pLoY, const PinRec &pHiY) {
    uint4 *dst = reinterpret_cast<uint4 *>(r);
    dst[0] = make_uint4(b, p, mov, 0u);
    dst[1] = make_uint4(lox, hix, loy, hiy);
    dst[2] = p > 0 ? make_uint4(varMov[b + lox].x, varMov[b + hix].x, varMov[b + loy].x, varMov[b + hiy].x)
                   : make_uint4(kNoVar, kNoVar, kNoVar, kNoVar);
    // used only where lo != hi, i.e. p >= 2
    reinterpret_cast<double2 *>(dst)[3] = make_double2(weight / static_cast<double>(p - 1), 0.0);
    double2 *ext = reinterpret_cast<double2 *>(r->ext[0]);
    ext[0] = make_double2(pLoX.posX, pLoX.termX);
    ext[1] = make_double2(pHiX.posX, pHiX.termX);
    ext[2] = make_double2(pLoY.posY, pLoY.termY);
    ext[3] = make_double2(pHiY.posY, pHiY.termY);
}

// One thread per net; the running best value is the reference's pinAt(lo),
// recomputed there from the same operands, so the comparisons are identical.
__global__ void __launch_bounds__(kBlock)
netExtremes(const uint32_t *__restrict__ netStart, const uint32_t *__restrict__ netMov,
            const double *__restrict__ weight, const PinRec *__restrict__ pin, const uint2 *__restrict__ varMov,
            NetRec *__restrict__ rec, uint32_t N) {
    const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= N) return;
    const uint32_t b = netStart[j];
    const uint32_t p = netStart[j + 1] - b;
    if (p > kWideNetPins) return;  // netExtremesWide
    uint32_t lox = 0, hix = 0, loy = 0, hiy = 0;
    PinRec pLoX{}, pHiX{}, pLoY{}, pHiY{};
    if (p > 0) {
        const PinRec first = pin[b];
        pLoX = pHiX = pLoY = pHiY = first;
        // the pins' loads go out four at a time (a 200-pin net is one
        // thread's loop), the comparisons run in pin order
        constexpr uint32_t kBatch = 4;
        for (uint32_t q0 = 1; q0 < p; q0 += kBatch) {
            PinRec c[kBatch];
#pragma unroll
            for (uint32_t u = 0; u < kBatch; ++u) c[u] = pin[b + (q0 + u < p ? q0 + u : 0u)];
#pragma unroll
            for (uint32_t u = 0; u < kBatch; ++u) {
                const uint32_t q = q0 + u;
                if (q >= p) break;
                if (c[u].posX < pLoX.posX) { lox = q; pLoX = c[u]; }
                if (c[u].posX > pHiX.posX) { hix = q; pHiX = c[u]; }
                if (c[u].posY < pLoY.posY) { loy = q; pLoY = c[u]; }
                if (c[u].posY > pHiY.posY) { hiy = q; pHiY = c[u]; }
            }
        }
    }
    writeNetRec(rec + j, b, p, netMov[j], weight[j], varMov, lox, hix, loy, hiy, pLoX, pHiX, pLoY, pHiY);
}

// The nets of more than kWideNetPins pins, one wave each: every lane keeps the
// first argmin / argmax of its pins (lane, lane + 64, ...), then the wave
// reduces them; the smaller value wins, equal values (neither is less) go to
// the smaller index, which is the sequential scan's first argmin / argmax.
struct Ext {
    double v;
    uint32_t i;
};
__device__ __forceinline__ Ext extLess(Ext a, Ext b) { return b.v < a.v || (!(a.v < b.v) && b.i < a.i) ? b : a; }
__device__ __forceinline__ Ext extMore(Ext a, Ext b) { return b.v > a.v || (!(a.v > b.v) && b.i < a.i) ? b : a; }
__device__ __forceinline__ Ext waveExt(Ext e, bool less) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const Ext o{__shfl_xor(e.v, off, 64), static_cast<uint32_t>(__shfl_xor(static_cast<int>(e.i), off, 64))};
        e = less ? extLess(e, o) : extMore(e, o);
    }
    return e;
}

__global__ void __launch_bounds__(256)
netExtremesWide(const uint32_t *__restrict__ wide, uint32_t count, const uint32_t *__restrict__ netStart,
                const uint32_t *__restrict__ netMov, const double *__restrict__ weight,
                const PinRec *__restrict__ pin, const uint2 *__restrict__ varMov, NetRec *__restrict__ rec) {
    const uint32_t lane = threadIdx.x & 63u, w = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (w >= count) return;  // the whole wave
    const uint32_t j = wide[w];
    const uint32_t b = netStart[j], p = netStart[j + 1] - b;  // p > kWideNetPins
    constexpr double kInf = __builtin_huge_val();
    Ext lx{kInf, ~0u}, hx{-kInf, ~0u}, ly{kInf, ~0u}, hy{-kInf, ~0u};
    for (uint32_t q = lane; q < p; q += 64) {
        const PinRec c = pin[b + q];
        // strict comparisons in pin order within the lane: its first extreme
        if (lx.i == ~0u || c.posX < lx.v) lx = Ext{c.posX, q};
        if (hx.i == ~0u || c.posX > hx.v) hx = Ext{c.posX, q};
        if (ly.i == ~0u || c.posY < ly.v) ly = Ext{c.posY, q};
        if (hy.i == ~0u || c.posY > hy.v) hy = Ext{c.posY, q};
    }
    lx = waveExt(lx, true);
    hx = waveExt(hx, false);
    ly = waveExt(ly, true);
    hy = waveExt(hy, false);
    if (lane != 0) return;
    writeNetRec(rec + j, b, p, netMov[j], weight[j], varMov, lx.i, hx.i, ly.i, hy.i, pin[b + lx.i], pin[b + hx.i],
                pin[b + ly.i], pin[b + hy.i]);
}

// The assembly is organised around the pinsOfCell references (e: row i's
// reference to pin q of net j, in the reference's row walk order) instead of
// one thread walking each row:
//   refCounts  per reference and axis, the edges the row walk emits for it
//              (bound pin: every other pin; inner pin: lo and hi; a net with
//              lo == hi: none, simpl_lss.cc:47-73) and how many of them reach
//              movable vertices (CSR entries); packed x | y << 32
//   scans      exclusive scans of both counts over the references: each
//              reference's first edge slot and first CSR slot, so rowPtr[i]
//              is the CSR scan at the row's first reference
//   edgeFill   each edge's weight w and rhs term w (there - own offset), in
//              row-walk order, and col / val = -w of the movable ones straight
//              into their CSR slots, one lane per edge
//   row sums   diag and rhs of each row as sequential sums over its edges in
//              the row walk's order, then the anchor term: in edgeFill, from
//              LDS, for the rows whose run lies in one edge block; in
//              rowSumsRest, from the edge list, for the few others
// Every double is the reference's expression and every sum keeps its order,
// so rowPtr, col, val, diag and rhs are the reference's bit for bit.
struct RefArgs {
    const uint32_t *pinStart;
    const uint32_t *refNet;
    const uint32_t *refPin;
    const uint2 *varMov;        // per pin: variable, movable pins before it in its net
    const NetRec *rec;
    const PinRec *pin;          // per pin: position, rhs term per axis (pinStage)
    const double *pos[2];
    const double *anchor[2];
    unsigned long long *edgeOff;  // [Pm + 1] edges: counts at e + 1, then the scan (x | y << 32)
    unsigned long long *entOff;   // [Pm + 1] CSR entries, likewise
    double2 *edges[2];          // per axis: (w, rhs term) per edge in row-walk order
    uint32_t *rowEdge[2];       // per axis: first edge of each row's run (rows [0, n])
    uint32_t *rowPtr[2];
    uint32_t *col[2];
    double *val[2];
    double *diag[2];
    double *rhs[2];
    double floor[2];
    double alpha;
    uint32_t n;
    uint32_t Pm;
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
__device__ __forceinline__ uint32_t half(unsigned long long v, int axis) {
    return static_cast<uint32_t>(axis ? v >> 32 : v);
}

// std::max(a, b) is (a < b) ? b : a; written out so NaN behaves the same.
__device__ __forceinline__ double refMax(double a, double b) { return a < b ? b : a; }

__global__ void __launch_bounds__(kBlock) refCounts(RefArgs a) {
    const uint32_t e = blockIdx.x * kBlock + threadIdx.x;
    if (e == 0) {
        a.edgeOff[0] = 0ull;
        a.entOff[0] = 0ull;
    }
    if (e >= a.Pm) return;
    const uint32_t net = a.refNet[e];
    const uint32_t q = a.refPin[e];
    const uint4 head = recHead(a.rec, net), x = recExt(a.rec, net), v = recVars(a.rec, net);
    uint32_t edges[2], ents[2];
#pragma unroll
    for (int axis = 0; axis < 2; ++axis) {
        const uint32_t lo = extLo(x, axis), hi = extHi(x, axis);
        if (lo == hi) {
            edges[axis] = ents[axis] = 0;
        } else if (q == lo || q == hi) {
            edges[axis] = head.y - 1;  // every other pin
            ents[axis] = head.z - 1;   // every other movable pin; q itself is movable
        } else {
            edges[axis] = 2;
            ents[axis] = (extLo(v, axis) != kNoVar ? 1u : 0u) + (extHi(v, axis) != kNoVar ? 1u : 0u);
        }
    }
    a.edgeOff[e + 1] = static_cast<unsigned long long>(edges[0]) | (static_cast<unsigned long long>(edges[1]) << 32);
    a.entOff[e + 1] = static_cast<unsigned long long>(ents[0]) | (static_cast<unsigned long long>(ents[1]) << 32);
}

// rowPtr[i] = the CSR scan at row i's first reference, and the row's first
// edge per axis (rows [0, n]).
__global__ void __launch_bounds__(kBlock) refRowPtr(RefArgs a) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i > a.n) return;
    const uint32_t e = a.pinStart[i];
    const unsigned long long s = a.entOff[e], t = a.edgeOff[e];
    a.rowPtr[0][i] = half(s, 0);
    a.rowPtr[1][i] = half(s, 1);
    a.rowEdge[0][i] = half(t, 0);
    a.rowEdge[1][i] = half(t, 1);
}

// Edge-parallel fill: the edges of each axis (in row-walk order) are cut into
// blocks of kFillEdges; lane k of a block takes edge start + k, finds its
// reference (the last one whose first edge is at or before it) and writes
// that edge.  Consecutive lanes write consecutive edge-list slots and, mostly,
// consecutive CSR slots, and lanes on one reference share its net record and
// read the net's pins side by side; no lane walks a long run alone.  The block
// keeps its edges in LDS and sums the rows whose run it holds whole; only the
// edges of rows across block boundaries go to the edge list, for rowSumsRest.
constexpr uint32_t kFillEdges = 512, kFillThreads = 256;
constexpr uint32_t kFillRefs = 2048;  // references a block keeps in LDS for the lookup

// The first edge of row i's run on one axis (rows [0, n]).
__device__ __forceinline__ uint32_t rowEdge(const RefArgs &a, uint32_t i, int axis) { return a.rowEdge[axis][i]; }

// first[axis][b] = the last reference whose first edge is at or before edge
// b * kFillEdges of that axis; first[2 blocks + axis blocks + b] = the first
// row whose run starts at or after that edge.
__global__ void __launch_bounds__(kBlock) edgeBlockStart(RefArgs a, uint32_t blocks, uint32_t *__restrict__ first) {
    const uint32_t b = blockIdx.x * kBlock + threadIdx.x;
    const int axis = blockIdx.y;
    if (b >= blocks) return;
    const uint32_t start = b * kFillEdges;
    uint32_t lo = 0, hi = a.Pm + 1;  // first e in [0, Pm] with edgeOff[e] > start
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (half(a.edgeOff[mid], axis) <= start) lo = mid + 1;
        else hi = mid;
    }
    first[axis * blocks + b] = lo - 1;  // edgeOff[0] = 0 <= start, so lo >= 1
    lo = 0;
    hi = a.n;  // first i in [0, n] with rowEdge(i) >= start (rowEdge(n) = all edges)
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (rowEdge(a, mid, axis) < start) lo = mid + 1;
        else hi = mid;
    }
    first[2 * blocks + axis * blocks + b] = lo;
}

// diag and rhs of row i from its edges' (w, rhs term) in the row walk's order
// (simpl_lss.cc:105-113): sequential sums, then the anchor term and the floor.
template <typename Edge>
__device__ __forceinline__ void rowTotals(const RefArgs &a, int axis, uint32_t i, Edge edge, uint32_t e0,
                                          uint32_t e1) {
    // the loads of eight edges are issued together (long runs: a bound pin
    // of a 200-pin net), the sums still run one edge after the other
    constexpr uint32_t kBatch = 8;
    double diag = 0.0, rhs = 0.0;
    for (uint32_t k0 = e0; k0 < e1; k0 += kBatch) {
        double2 d[kBatch];
#pragma unroll
        for (uint32_t u = 0; u < kBatch; ++u) d[u] = edge(k0 + u < e1 ? k0 + u : e1 - 1);
#pragma unroll
        for (uint32_t u = 0; u < kBatch; ++u) {
            if (k0 + u < e1) {
                diag += d[u].x;
                rhs += d[u].y;
            }
        }
    }
    if (a.alpha > 0.0) {
        const double anchor = a.anchor[axis][i];
        const double w = a.alpha / refMax(fabs(a.pos[axis][i] - anchor), a.floor[axis]);
        diag += w;
        rhs += w * anchor;
    }
    a.diag[axis][i] = refMax(diag, 1e-12);
    a.rhs[axis][i] = rhs;
}

__global__ void __launch_bounds__(kFillThreads) edgeFill(RefArgs a, uint32_t blocks, const uint32_t *__restrict__ first) {
    __shared__ uint32_t refOff[kFillRefs];
    __shared__ uint32_t refCount;
    __shared__ double2 stage[kFillEdges];  // the block's edges, for the row sums below
    __shared__ uint32_t headEnd, tailStart;  // edges of rows across the block's start / end
    const int axis = blockIdx.y;
    const uint32_t total = half(a.edgeOff[a.Pm], axis);
    const uint32_t start = blockIdx.x * kFillEdges;
    if (start >= total) return;  // block-uniform
    const uint32_t end = min(start + kFillEdges, total);
    const uint32_t e0 = first[axis * blocks + blockIdx.x];
    // the first rows whose run starts in the block: their run bounds are loaded
    // now, so they arrive while the edges are filled
    const uint32_t r0 = first[2 * blocks + axis * blocks + blockIdx.x];
    const uint32_t rFirst = r0 + threadIdx.x < a.n ? r0 + threadIdx.x : a.n - 1;  // n > 0 here
    const uint32_t pa = rowEdge(a, rFirst, axis), pb = rowEdge(a, rFirst + 1, axis);
    // the block's references e0, e0 + 1, ... up to the one holding edge end - 1
    if (threadIdx.x == 0) {
        refCount = kFillRefs + 1;
        tailStart = end;
    }
    __syncthreads();
    for (uint32_t base = 0; base < kFillRefs; base += kFillThreads) {
        const uint32_t r = base + threadIdx.x;
        const uint32_t e = e0 + r;
        if (r < kFillRefs && e <= a.Pm) {
            const uint32_t o = half(a.edgeOff[e], axis);
            refOff[r] = o;
            if (o >= end) atomicMin(&refCount, r);  // references from here on start past the block
        }
        __syncthreads();
        if (refCount <= base + kFillThreads) break;  // block-uniform
    }
    const uint32_t R = refCount;  // the block's references are e0 .. e0 + R - 1 when R <= kFillRefs
    for (uint32_t k = start + threadIdx.x; k < end; k += kFillThreads) {
        uint32_t e, t;
        if (R <= kFillRefs) {  // largest r < R with refOff[r] <= k
            uint32_t lo = 0, hi = R;
            while (hi - lo > 1) {
                const uint32_t mid = lo + (hi - lo) / 2;
                if (refOff[mid] <= k) lo = mid;
                else hi = mid;
            }
            e = e0 + lo;
            t = k - refOff[lo];
        } else {  // more zero-edge references than the table holds: search memory
            uint32_t lo = e0, hi = a.Pm + 1;
            while (lo < hi) {
                const uint32_t mid = lo + (hi - lo) / 2;
                if (half(a.edgeOff[mid], axis) <= k) lo = mid + 1;
                else hi = mid;
            }
            e = lo - 1;
            t = k - half(a.edgeOff[e], axis);
        }
        const uint32_t net = a.refNet[e];
        const uint32_t q = a.refPin[e];
        const uint4 head = recHead(a.rec, net), x = recExt(a.rec, net), v = recVars(a.rec, net);
        const uint32_t lo = extLo(x, axis), hi = extHi(x, axis);
        const uint32_t b = head.x;
        const bool bound = q == lo || q == hi;
        const double base = a.rec[net].base;  // weight / (p - 1)
        const double2 *ext = reinterpret_cast<const double2 *>(a.rec[net].ext[axis]);
        const double2 eLo = ext[0], eHi = ext[1];  // (position, rhs term) of lo and of hi
        double own, ownOffset;
        if (bound) {
            own = q == lo ? eLo.x : eHi.x;
            ownOffset = q == lo ? eLo.y : eHi.y;
        } else {
            const PinRec ownPin = a.pin[b + q];
            own = pinPos(ownPin, axis);
            ownOffset = pinTerm(ownPin, axis);
        }
        const uint32_t other = bound ? (t < q ? t : t + 1) : (t == 0 ? lo : hi);
        double there, thereTerm;
        uint32_t var, mi;
        if (!bound || head.y == 2) {
            // the far pin is lo or hi (a two-pin net's bound pins are its lo and
            // hi, lo != hi here): all from the record.  Its CSR slot within the
            // reference's run: inner pin, lo first (if movable); a two-pin net's
            // bound pin has the one entry
            const bool isLo = other == lo;
            there = isLo ? eLo.x : eHi.x;
            thereTerm = isLo ? eLo.y : eHi.y;
            var = isLo ? extLo(v, axis) : extHi(v, axis);
            mi = bound || t == 0 ? 0u : (extLo(v, axis) != kNoVar ? 1u : 0u);
        } else {
            // bound pin of a bigger net: the lanes of the reference read the
            // net's pins side by side
            const PinRec th = a.pin[b + other];
            there = pinPos(th, axis);
            thereTerm = pinTerm(th, axis);
            const uint2 vm = a.varMov[b + other];
            var = vm.x;
            mi = vm.y - (other > q ? 1u : 0u);
        }
        const double w = base / refMax(fabs(own - there), a.floor[axis]);
        // movable: (there - cellPos) - ownOffset; fixed: there - ownOffset
        stage[k - start] = make_double2(w, w * (thereTerm - ownOffset));
        if (var != kNoVar) {
            const uint32_t slot = half(a.entOff[e], axis) + mi;
            a.col[axis][slot] = var;
            a.val[axis][slot] = -w;
        }
    }
    if (threadIdx.x == 0) headEnd = r0 < a.n ? min(pa, end) : end;
    __syncthreads();
    // the rows whose whole run lies in this block, one thread each, from LDS;
    // rows with no edges or a run across a block boundary are rowSumsRest's
    for (uint32_t i = r0 + threadIdx.x; i < a.n; i += kFillThreads) {
        const bool pre = i == r0 + threadIdx.x;
        const uint32_t ea = pre ? pa : rowEdge(a, i, axis);
        if (ea >= end) break;  // runs start in order: the later rows are past the block too
        const uint32_t eb = pre ? pb : rowEdge(a, i + 1, axis);
        if (ea == eb) continue;
        if (eb > end) {  // the one row that runs past the block's end
            tailStart = ea;
            continue;
        }
        rowTotals(a, axis, i, [&](uint32_t k) { return stage[k - start]; }, ea, eb);
    }
    __syncthreads();
    // only the rows across the block's boundaries are summed from the edge
    // list: the edges before the first row that starts here, and those of
    // the row that runs past the end
    double2 *__restrict__ edges = a.edges[axis];
    const uint32_t h = headEnd, ts = tailStart;
    for (uint32_t k = start + threadIdx.x; k < h; k += kFillThreads) edges[k] = stage[k - start];
    for (uint32_t k = ts + threadIdx.x; k < end; k += kFillThreads) edges[k] = stage[k - start];
}

// diag and rhs of the rows edgeFill left: rows with no edges, and rows whose
// run crosses a boundary of the fill's edge blocks (at most one per boundary),
// summed from the edge list in memory.  Every row is computed once, by
// edgeFill or here.
__global__ void __launch_bounds__(kBlock) rowSumsRest(RefArgs a) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const int axis = blockIdx.y;
    if (i >= a.n) return;
    const uint32_t ea = rowEdge(a, i, axis), eb = rowEdge(a, i + 1, axis);
    if (ea < eb && ea / kFillEdges == (eb - 1) / kFillEdges) return;  // summed by edgeFill
    const double2 *__restrict__ edges = a.edges[axis];
    rowTotals(a, axis, i, [&](uint32_t k) { return edges[k]; }, ea, eb);
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

// Inclusive in-place scans of m values: small ones by the tile scan above,
// large ones by rocPRIM's single-pass device scan (scan.hip).  Integer sums,
// so either gives the same values.
int deviceScan32(kt_ctx *ctx, uint32_t *data, uint64_t m);
int deviceScan64(kt_ctx *ctx, uint64_t *data, uint64_t m);
constexpr uint64_t kDeviceScanMin = 1ull << 16;

int scanInclusive(kt_ctx *ctx, uint32_t *data, uint64_t m) {
    if (m >= kDeviceScanMin) return deviceScan32(ctx, data, m);
    const uint64_t cap = scanScratch(m);
    KT_TRY(ctx->scanTmp.reserve(cap));
    return scanLevel<uint32_t>(ctx, data, m, ctx->scanTmp.p, cap);
}

int scanInclusive64(kt_ctx *ctx, uint64_t *data, uint64_t m) {
    if (m >= kDeviceScanMin) return deviceScan64(ctx, data, m);
    const uint64_t cap = scanScratch(m);
    KT_TRY(ctx->scanTmp64.reserve(cap));
    return scanLevel<uint64_t>(ctx, data, m, ctx->scanTmp64.p, cap);
}

int scanRowPtr(kt_ctx *ctx, uint32_t *rowPtr, uint64_t rows) {
    return scanInclusive(ctx, rowPtr + 1, rows);
}

// px/py: fixed positions with the movable cells' positions written over
int vertexPositionsOf(kt_ctx *ctx, const double *x, const double *y) {
    KT_TRY(ctx->vxy.reserve(std::max<uint32_t>(ctx->V, 1)));
    if (ctx->V == 0) return KT_OK;
    vertexPositions<<<dim3((ctx->V + kBlock - 1) / kBlock), dim3(kBlock), 0, ctx->stream>>>(
        x, y, ctx->varOfVertex.p, ctx->fixedX.p, ctx->fixedY.p, ctx->vxy.p, ctx->V);
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
    KT_TRY(ctx->ext.reserve(std::max<uint64_t>(32ull * ctx->N, 32)));  // one NetRec (128 bytes) per net
    KT_TRY(vertexPositionsLaunch(ctx));
    const uint64_t pins = std::max<uint64_t>(ctx->P, 1);
    KT_TRY(ctx->pinAt[0].reserve(4 * pins));  // one PinRec (4 doubles) per pin
    PinRec *pin = reinterpret_cast<PinRec *>(ctx->pinAt[0].p);
    const uint2 *varMov = reinterpret_cast<const uint2 *>(ctx->pinVar.p);  // uploaded with the netlist
    if (ctx->P > 0) {
        pinStage<<<dim3(static_cast<uint32_t>((ctx->P + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream>>>(
            ctx->pinCell.p, varMov, ctx->offX.p, ctx->offY.p, ctx->vxy.p, pin, ctx->P);
        KT_HIP(hipGetLastError());
    }
    NetRec *rec = reinterpret_cast<NetRec *>(ctx->ext.p);
    if (ctx->N > 0) {
        netExtremes<<<dim3((ctx->N + kBlock - 1) / kBlock), dim3(kBlock), 0, ctx->stream>>>(
            ctx->netStart.p, ctx->netMov.p, ctx->weight.p, pin, varMov, rec, ctx->N);
        KT_HIP(hipGetLastError());
        if (ctx->numWideNets) {
            netExtremesWide<<<dim3((ctx->numWideNets + 3) / 4), dim3(256), 0, ctx->stream>>>(
                ctx->wideNets.p, ctx->numWideNets, ctx->netStart.p, ctx->netMov.p, ctx->weight.p, pin, varMov, rec);
            KT_HIP(hipGetLastError());
        }
    }
    RefArgs a{};
    a.pinStart = ctx->pinStart.p;
    a.refNet = ctx->refNet.p;
    a.refPin = ctx->refPin.p;
    a.varMov = varMov;
    a.rec = rec;
    a.pin = pin;
    a.pos[0] = ctx->posX.p;
    a.pos[1] = ctx->posY.p;
    a.anchor[0] = ctx->anchorX.p;
    a.anchor[1] = ctx->anchorY.p;
    const uint64_t Pm = ctx->Pm;
    KT_TRY(ctx->b2bEdgeOff.reserve(Pm + 1));
    KT_TRY(ctx->b2bEntOff.reserve(Pm + 1));
    for (int axis = 0; axis < 2; ++axis) KT_TRY(ctx->b2bEdges[axis].reserve(2 * std::max<uint64_t>(ctx->nnzBound, 1)));
    KT_TRY(ctx->b2bRowEdge.reserve(2ull * (n + 1)));
    a.rowEdge[0] = ctx->b2bRowEdge.p;
    a.rowEdge[1] = ctx->b2bRowEdge.p + (n + 1);
    a.edgeOff = reinterpret_cast<unsigned long long *>(ctx->b2bEdgeOff.p);
    a.entOff = reinterpret_cast<unsigned long long *>(ctx->b2bEntOff.p);
    for (int axis = 0; axis < 2; ++axis) {
        Sys &s = ctx->sys[axis];
        a.edges[axis] = reinterpret_cast<double2 *>(ctx->b2bEdges[axis].p);
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
    a.floor[0] = floorX;
    a.floor[1] = floorY;
    a.alpha = alpha;
    ctx->lastAlpha = alpha;
    a.n = n;
    a.Pm = static_cast<uint32_t>(Pm);
    if (n == 0) {
        for (int axis = 0; axis < 2; ++axis)
            KT_HIP(hipMemsetAsync(ctx->sys[axis].rowPtr.p, 0, sizeof(uint32_t), ctx->stream));
        return KT_OK;
    }
    hipEvent_t t0 = nullptr;
    KT_TRY(timedBegin(ctx, KT_KERNEL_B2B_FILL, &t0));
    const uint32_t refBlocks = static_cast<uint32_t>(std::max<uint64_t>((Pm + kBlock - 1) / kBlock, 1));
    refCounts<<<dim3(refBlocks), dim3(kBlock), 0, ctx->stream>>>(a);
    KT_HIP(hipGetLastError());
    KT_TRY(scanInclusive64(ctx, reinterpret_cast<uint64_t *>(a.edgeOff), Pm + 1));
    KT_TRY(scanInclusive64(ctx, reinterpret_cast<uint64_t *>(a.entOff), Pm + 1));
    refRowPtr<<<dim3((n + 1 + kBlock - 1) / kBlock), dim3(kBlock), 0, ctx->stream>>>(a);
    KT_HIP(hipGetLastError());
    if (Pm > 0) {
        // edges per axis are at most nnzBound (every edge of every net)
        const uint32_t blocks = static_cast<uint32_t>(std::max<uint64_t>((ctx->nnzBound + kFillEdges - 1) / kFillEdges, 1));
        KT_TRY(ctx->b2bLong.reserve(4ull * blocks));
        edgeBlockStart<<<dim3((blocks + kBlock - 1) / kBlock, 2), dim3(kBlock), 0, ctx->stream>>>(a, blocks,
                                                                                                   ctx->b2bLong.p);
        KT_HIP(hipGetLastError());
        edgeFill<<<dim3(blocks, 2), dim3(kFillThreads), 0, ctx->stream>>>(a, blocks, ctx->b2bLong.p);
        KT_HIP(hipGetLastError());
    }
    rowSumsRest<<<dim3((n + kBlock - 1) / kBlock, 2), dim3(kBlock), 0, ctx->stream>>>(a);
    KT_HIP(hipGetLastError());
    KT_TRY(timedEnd(ctx, KT_KERNEL_B2B_FILL, t0));
    return KT_OK;
}

}  // namespace kt
