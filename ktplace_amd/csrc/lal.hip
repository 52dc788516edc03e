// Look-ahead legalization, HPWL, occupancy and overflow on gfx950.
//
// Follows LookAheadLegalizer::legalize (simpl_lal.cc:43-372) and the
// bookkeeping of simplPlacer::Impl (simpl_placer.cc:256-297).
//
// Split of the work
//   GPU  every per-cell operation: bin of each cell and the occupancy grid
//        (fixed-point atomics, exact and order-free whenever cell areas are
//        multiples of 2^-16), region membership, and for every level of the
//        recursive bisection of every region of a batch at once: the sort of
//        the cells of each block by (centre, id), the split at half the area,
//        the second sort by distance from the cut, the greedy stripe filling,
//        the per-stripe extent and the affine map; clamp to the die; HPWL
//        with the reference's chunked summation order (bit-exact); overflow.
//   Host the per-block geometry (free-area profile, cutline, stripes:
//        lal_geom.cc), the flood fill of overfilled bins and the growth of
//        clusters, and the batch schedule.  As in the reference, occupancy
//        is updated after each batch from the cells that changed bin only
//        (the GPU returns them), and rectangle sums use exact fixed-point row
//        prefixes patched on the rows that changed, so a batch costs work in
//        proportion to its cells, not to the grid.
//
// Level-synchronous recursion: all blocks of one level of all regions of a
// batch are processed together.  Cells are kept as lists of ids ordered by
// (block, key, id) with two stable radix sorts (key, then block) started from
// id order, or, for up to 2048 cells, one workgroup's bitonic sort on the
// same triple; either reproduces the reference's total order
// (simpl_parallel.h:26-33).  Area prefix sums use 64-bit fixed point, so the split index and the
// stripe boundaries are exact for areas that are multiples of 2^-16.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "kt_internal.h"
#include "lal_dev.h"
#include "lal_geom.h"

namespace kt {

int radixSortPairs(kt_ctx *ctx, uint64_t *keys, uint32_t *vals, uint64_t *keysTmp, uint32_t *valsTmp, uint64_t n,
                   int bits, DevBuf<uint32_t> &counts, DevBuf<uint32_t> &raw);
int scanInclusive(kt_ctx *ctx, uint32_t *data, uint64_t m);
int scanInclusive64(kt_ctx *ctx, uint64_t *data, uint64_t m);
bool smallSortFits(uint64_t n);
int smallSortBySegment(kt_ctx *ctx, uint32_t *ids, const uint64_t *key, const uint32_t *segById, uint64_t n);

namespace {

constexpr int kB = 256;

inline dim3 gridFor(uint64_t n) { return dim3(static_cast<uint32_t>(std::max<uint64_t>((n + kB - 1) / kB, 1))); }

__device__ __forceinline__ unsigned long long waveSumU64(unsigned long long v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// bin of each cell centre (simpl_placer.cc:277-289) and fixed-point occupancy.
// The lanes of a wave first combine their areas per bin (one pass per
// distinct bin among the wave's cells), one lane adds each sum to the
// block's table in LDS, and the block adds its table to memory at the end: a
// clumped placement puts most cells into a few bins, where one atomic per
// cell, or even per wave, would serialise on those bins.  Integer sums, so
// the result does not depend on the grouping.
__global__ void __launch_bounds__(kB)
cellBins(const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ w,
         const double *__restrict__ h, const double *__restrict__ area, GridDev g, uint32_t n,
         uint32_t *__restrict__ binOf, unsigned long long *__restrict__ occFix, uint32_t *__restrict__ cellCount) {
    // the block's bins, area sums and cell counts, open addressing; a bin
    // that finds no free slot within kProbes is added to memory directly
    constexpr uint32_t kSlots = 256, kEmpty = 0xffffffffu;
    constexpr int kProbes = 8;
    __shared__ uint32_t keys[kSlots];
    __shared__ unsigned long long sums[kSlots];
    __shared__ uint32_t cells[kSlots];
    for (uint32_t t = threadIdx.x; t < kSlots; t += kB) {
        keys[t] = kEmpty;
        sums[t] = 0ull;
        cells[t] = 0u;
    }
    __syncthreads();
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kB;
    for (uint64_t base = static_cast<uint64_t>(blockIdx.x) * kB; base < n; base += stride) {  // block-uniform
        const bool valid = base + threadIdx.x < n;
        const uint32_t i = valid ? static_cast<uint32_t>(base + threadIdx.x) : 0u;
        uint32_t k = 0;
        unsigned long long v = 0;
        if (valid) {
            k = binOfPoint(g, x[i] + 0.5 * w[i], y[i] + 0.5 * h[i]);
            binOf[i] = k;
            v = static_cast<unsigned long long>(llrint(area[i] * kFixScale));
        }
        bool todo = valid;
        unsigned long long pending = __ballot(todo);
        while (pending) {
            const int leader = __ffsll(static_cast<long long>(pending)) - 1;
            const uint32_t lk = __shfl(k, leader);
            const bool mine = todo && k == lk;
            const unsigned long long s = waveSumU64(mine ? v : 0ull);
            const uint32_t c = static_cast<uint32_t>(__popcll(__ballot(mine)));
            if (static_cast<int>(threadIdx.x & 63) == leader) {
                uint32_t slot = (lk * 2654435761u) & (kSlots - 1);
                bool kept = false;
                for (int probe = 0; probe < kProbes; ++probe) {
                    const uint32_t prev = atomicCAS(&keys[slot], kEmpty, lk);
                    if (prev == kEmpty || prev == lk) {
                        atomicAdd(&sums[slot], s);
                        atomicAdd(&cells[slot], c);
                        kept = true;
                        break;
                    }
                    slot = (slot + 1) & (kSlots - 1);
                }
                if (!kept) {
                    atomicAdd(occFix + lk, s);
                    atomicAdd(cellCount + lk, c);
                }
            }
            todo = todo && !mine;
            pending = __ballot(todo);
        }
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < kSlots; t += kB)
        if (keys[t] != kEmpty) {
            atomicAdd(occFix + keys[t], sums[t]);
            atomicAdd(cellCount + keys[t], cells[t]);
        }
}

__global__ void __launch_bounds__(kB)
fixToDouble(const unsigned long long *__restrict__ fix, double *__restrict__ out, uint32_t m) {
    const uint32_t k = blockIdx.x * kB + threadIdx.x;
    if (k < m) out[k] = static_cast<double>(fix[k]) / kFixScale;
}

// overflow: sum of max(occ - g avail, 0), fixed tree per block, then in order
__global__ void __launch_bounds__(kB)
overflowPartials(const double *__restrict__ occ, const double *__restrict__ avail, double g, uint32_t m,
                 double *__restrict__ part) {
    __shared__ double s[kB];
    const uint32_t k = blockIdx.x * kB + threadIdx.x;
    const double ex = k < m ? occ[k] - g * avail[k] : 0.0;
    s[threadIdx.x] = ex < 0.0 ? 0.0 : ex;
    __syncthreads();
    for (int off = kB / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ void sumInOrder(const double *__restrict__ part, uint32_t m, double *__restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double t = 0.0;
    for (uint32_t k = 0; k < m; ++k) t += part[k];
    *out = t;
}

// ------------------------------------------------------------------ HPWL
// simpl_placer.cc:256-275: BBox of pos + off over the net's pins, 0 for nets
// with fewer than two pins.
__global__ void __launch_bounds__(kB)
netHpwl(const uint32_t *__restrict__ netStart, const uint32_t *__restrict__ pinCell, const double *__restrict__ offX,
        const double *__restrict__ offY, const double2 *__restrict__ vxy, uint32_t N, double *__restrict__ out) {
    const uint32_t j = blockIdx.x * kB + threadIdx.x;
    if (j >= N) return;
    const uint32_t b = netStart[j], e = netStart[j + 1];
    if (e - b < 2) {
        out[j] = 0.0;
        return;
    }
    double lx = INFINITY, ly = INFINITY, ux = -INFINITY, uy = -INFINITY;
    for (uint32_t q = b; q < e; ++q) {
        const uint32_t c = pinCell[q];
        const double2 v = vxy[c];
        const double px = v.x + offX[q], py = v.y + offY[q];
        // BBox::extend: std::min / std::max (kt_geometry.cc:141-146)
        lx = px < lx ? px : lx;
        ly = py < ly ? py : ly;
        ux = ux < px ? px : ux;
        uy = uy < py ? py : uy;
    }
    out[j] = (lx > ux || ly > uy) ? 0.0 : (ux - lx) + (uy - ly);
}

// getDeterministicSum (simpl_parallel.h:35-51): one thread per 4096-chunk.
__global__ void __launch_bounds__(kB)
chunkSums(const double *__restrict__ v, uint32_t n, double *__restrict__ part) {
    const uint32_t c = blockIdx.x * kB + threadIdx.x;
    const uint32_t chunks = (n + 4095) / 4096;
    if (c >= chunks) return;
    const uint32_t end = min(n, (c + 1) * 4096u);
    double s = 0.0;
    for (uint32_t i = c * 4096u; i < end; ++i) s += v[i];
    part[c] = s;
}

// ------------------------------------------------------------------ LAL per-cell kernels
__global__ void __launch_bounds__(kB)
regionFlags(const uint32_t *__restrict__ binOf, const uint32_t *__restrict__ binRegion, uint32_t n,
            uint32_t *__restrict__ seg, uint32_t *__restrict__ flag) {
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= n) return;
    const uint32_t r = binRegion[binOf[i]];
    seg[i] = r;
    flag[i] = r != kNoVar ? 1u : 0u;
}

__global__ void __launch_bounds__(kB)
activeFlags(const uint32_t *__restrict__ seg, uint32_t n, uint32_t *__restrict__ flag) {
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i < n) flag[i] = seg[i] != kNoVar ? 1u : 0u;
}

// ids of flagged cells in id order (flag scanned inclusively into pos)
__global__ void __launch_bounds__(kB)
compactIds(const uint32_t *__restrict__ seg, const uint32_t *__restrict__ incl, uint32_t n,
           uint32_t *__restrict__ ids) {
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= n || seg[i] == kNoVar) return;
    ids[incl[i] - 1] = i;
}

// The same over the cells of one batch (list, in id order).
__global__ void __launch_bounds__(kB)
listFlags(const uint32_t *__restrict__ list, const uint32_t *__restrict__ seg, uint32_t m,
          uint32_t *__restrict__ flag) {
    const uint32_t j = blockIdx.x * kB + threadIdx.x;
    if (j < m) flag[j] = seg[list[j]] != kNoVar ? 1u : 0u;
}

__global__ void __launch_bounds__(kB)
listCompact(const uint32_t *__restrict__ list, const uint32_t *__restrict__ seg, const uint32_t *__restrict__ incl,
            uint32_t m, uint32_t *__restrict__ ids) {
    const uint32_t j = blockIdx.x * kB + threadIdx.x;
    if (j >= m) return;
    const uint32_t i = list[j];
    if (seg[i] != kNoVar) ids[incl[j] - 1] = i;
}

__global__ void __launch_bounds__(kB)
listFill(const uint32_t *__restrict__ list, uint32_t m, uint32_t *__restrict__ a, uint32_t v) {
    const uint32_t j = blockIdx.x * kB + threadIdx.x;
    if (j < m) a[list[j]] = v;
}

__global__ void __launch_bounds__(kB)
listRemap(const uint32_t *__restrict__ list, uint32_t m, const uint32_t *__restrict__ half,
          const uint32_t *__restrict__ remap, uint32_t *__restrict__ seg) {
    const uint32_t j = blockIdx.x * kB + threadIdx.x;
    if (j >= m) return;
    const uint32_t i = list[j];
    const uint32_t h = half[i];
    seg[i] = h == kNoVar ? kNoVar : remap[h];
}

// bin -> region for the bins of each region of a batch (one block per
// region), or back to kNoVar
__global__ void __launch_bounds__(kB)
paintRegions(const uint4 *__restrict__ rects, uint32_t nx, uint32_t *__restrict__ binRegion, int clear,
             uint32_t *__restrict__ counters) {
    if (counters && blockIdx.x == 0 && threadIdx.x < 4) counters[threadIdx.x] = threadIdx.x == 0 ? kNoVar : 0u;
    const uint4 r = rects[blockIdx.x];  // x0, x1, y0, y1
    const uint32_t w = r.y - r.x + 1, cells = w * (r.w - r.z + 1);
    const uint32_t v = clear ? kNoVar : blockIdx.x;
    for (uint32_t k = threadIdx.x; k < cells; k += kB) binRegion[(r.z + k / w) * nx + r.x + k % w] = v;
}

// The same for a batch of at most kRectsByValue regions, passed in the
// launch arguments (no copy to the device first).
constexpr uint32_t kRectsByValue = 8;
struct RectSet {
    uint4 r[kRectsByValue];
};
__global__ void __launch_bounds__(kB)
paintRegionsArg(RectSet rects, uint32_t nx, uint32_t *__restrict__ binRegion, int clear,
                uint32_t *__restrict__ counters) {
    if (counters && blockIdx.x == 0 && threadIdx.x < 4) counters[threadIdx.x] = threadIdx.x == 0 ? kNoVar : 0u;
    const uint4 r = rects.r[blockIdx.x];
    const uint32_t w = r.y - r.x + 1, cells = w * (r.w - r.z + 1);
    const uint32_t v = clear ? kNoVar : blockIdx.x;
    for (uint32_t k = threadIdx.x; k < cells; k += kB) binRegion[(r.z + k / w) * nx + r.x + k % w] = v;
}

// Cells of the batch whose bin changed: new bin stored, (old, new, area)
// appended for the host's occupancy update.
using Moved = LalMoved;

// Per-cell state of a batch for the in-order rule (see lalRunBatch).
struct BatchState {
    uint32_t *regionOf;           // region whose run last took the cell
    double *saveX, *saveY;        // position before that run
    uint32_t *saveBin;            // bin before that run
    uint32_t *binStart;           // bin at the start of the batch
    double *x, *y;
    const double *w, *h;
    uint32_t *binOf;
    const uint32_t *binRegion;    // bin -> region of the batch (kNoVar elsewhere)
    uint32_t *seg;
    GridDev g;
};

__global__ void __launch_bounds__(kB) saveBatch(const uint32_t *__restrict__ list, uint32_t m, BatchState s) {
    const uint32_t j = blockIdx.x * kB + threadIdx.x;
    if (j >= m) return;
    const uint32_t i = list[j];
    s.regionOf[i] = s.seg[i];
    s.saveX[i] = s.x[i];
    s.saveY[i] = s.y[i];
    s.saveBin[i] = s.binOf[i];
    s.binStart[i] = s.binOf[i];
}

// The batch's cells in any order (the small path sorts them on a total
// order itself), with their restore points as saveBatch records them: seg
// as regionFlags, and each wave appends its members to list with one
// atomic on count.
__global__ void __launch_bounds__(kB)
regionMembers(uint32_t n, BatchState s, uint32_t *__restrict__ list, uint32_t *__restrict__ count) {
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    const bool valid = i < n;
    const uint32_t bin = valid ? s.binOf[i] : 0u;
    const uint32_t r = valid ? s.binRegion[bin] : kNoVar;
    if (valid) s.seg[i] = r;
    const bool member = r != kNoVar;
    const unsigned long long mask = __ballot(member);
    if (mask == 0ull) return;
    const int lane = static_cast<int>(threadIdx.x & 63);
    const int leader = __ffsll(static_cast<long long>(mask)) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(count, static_cast<uint32_t>(__popcll(mask)));
    base = __shfl(base, leader);
    const uint32_t below = static_cast<uint32_t>(__popcll(mask & ((1ull << lane) - 1ull)));
    if (member) {
        list[base + below] = i;
        s.regionOf[i] = r;
        s.saveX[i] = s.x[i];
        s.saveY[i] = s.y[i];
        s.saveBin[i] = bin;
        s.binStart[i] = bin;
    }
}

__device__ __forceinline__ uint32_t binNow(const BatchState &s, uint32_t i) {
    return binOfPoint(s.g, s.x[i] + 0.5 * s.w[i], s.y[i] + 0.5 * s.h[i]);
}

// The end of a pass over the batch's cells: each cell's current bin, computed
// once, is stored in binOf; cnt[0] = the first region, among those after its
// own, that a cell of a region >= first moved into (stays kNoVar when there is
// none); a cell whose bin differs from its bin at the batch start appends
// (old, new, area) to out, counted in cnt[2], for the host's occupancy update.
__global__ void __launch_bounds__(kB)
passEnd(const uint32_t *__restrict__ list, uint32_t m, uint32_t first, BatchState s,
        const double *__restrict__ area, uint32_t *__restrict__ cnt, Moved *__restrict__ out) {
    const uint32_t j = blockIdx.x * kB + threadIdx.x;
    if (j >= m) return;
    const uint32_t i = list[j];
    const uint32_t k = binNow(s, i);
    const uint32_t r = s.regionOf[i];
    if (r != kNoVar && r >= first) {
        const uint32_t rj = s.binRegion[k];
        if (rj != kNoVar && rj > r) atomicMin(cnt, rj);
    }
    s.binOf[i] = k;
    const uint32_t old = s.binStart[i];
    if (k == old) return;
    const uint32_t slot = atomicAdd(cnt + 2, 1u);
    out[slot] = Moved{old, k, i, 0u, static_cast<unsigned long long>(llrint(area[i] * kFixScale))};
}

// regions >= bad back to their restore points; earlier ones keep their result
__global__ void __launch_bounds__(kB)
repairBatch(const uint32_t *__restrict__ list, uint32_t m, uint32_t bad, BatchState s) {
    const uint32_t j = blockIdx.x * kB + threadIdx.x;
    if (j >= m) return;
    const uint32_t i = list[j];
    const uint32_t r = s.regionOf[i];
    if (r != kNoVar && r >= bad) {
        s.x[i] = s.saveX[i];
        s.y[i] = s.saveY[i];
        s.binOf[i] = s.saveBin[i];
    } else {
        s.binOf[i] = binNow(s, i);
    }
}

// cells now in the bins of regions >= bad take part in the run again; cells
// that arrive from an earlier region get that region as their restore point
__global__ void __launch_bounds__(kB)
resegBatch(const uint32_t *__restrict__ list, uint32_t m, uint32_t bad, BatchState s, uint32_t *__restrict__ count) {
    const uint32_t j = blockIdx.x * kB + threadIdx.x;
    if (j >= m) return;
    const uint32_t i = list[j];
    const uint32_t b = s.binOf[i];
    const uint32_t rj = s.binRegion[b];
    if (rj != kNoVar && rj >= bad) {
        s.seg[i] = rj;
        atomicAdd(count, 1u);
        const uint32_t r = s.regionOf[i];
        if (r == kNoVar || r < bad) {
            s.regionOf[i] = rj;
            s.saveX[i] = s.x[i];
            s.saveY[i] = s.y[i];
            s.saveBin[i] = b;
        }
    } else {
        s.seg[i] = kNoVar;
    }
}

struct CentreArgs {
    const double *pos;   // positions along the level's axis
    const double *size;  // widths (x) or heights (y)
};

__device__ __forceinline__ double centreOf(const CentreArgs &c, uint32_t i) { return c.pos[i] + 0.5 * c.size[i]; }

__global__ void __launch_bounds__(kB)
centreKeys(const uint32_t *__restrict__ ids, uint32_t A, CentreArgs c, uint64_t *__restrict__ keys) {
    const uint32_t k = blockIdx.x * kB + threadIdx.x;
    if (k < A) keys[k] = orderKey(centreOf(c, ids[k]));
}

__global__ void __launch_bounds__(kB)
segmentKeys(const uint32_t *__restrict__ ids, const uint32_t *__restrict__ seg, uint32_t A,
            uint64_t *__restrict__ keys) {
    const uint32_t k = blockIdx.x * kB + threadIdx.x;
    if (k < A) keys[k] = seg[ids[k]];
}

// segment ranges of a (segment, ...) sorted id list; begin/end pre-filled with 0
__device__ __forceinline__ void segmentBoundsAt(uint32_t k, const uint32_t *__restrict__ ids,
                                                const uint32_t *__restrict__ seg, uint32_t A,
                                                uint32_t *__restrict__ begin, uint32_t *__restrict__ end) {
    const uint32_t s = seg[ids[k]];
    if (k == 0 || seg[ids[k - 1]] != s) begin[s] = k;
    if (k + 1 == A || seg[ids[k + 1]] != s) end[s] = k + 1;
}

__global__ void __launch_bounds__(kB)
segmentBounds(const uint32_t *__restrict__ ids, const uint32_t *__restrict__ seg, uint32_t A,
              uint32_t *__restrict__ begin, uint32_t *__restrict__ end) {
    const uint32_t k = blockIdx.x * kB + threadIdx.x;
    if (k < A) segmentBoundsAt(k, ids, seg, A, begin, end);
}

__global__ void __launch_bounds__(kB)
areaFix(const uint32_t *__restrict__ ids, const double *__restrict__ area, uint32_t A, uint64_t *__restrict__ out) {
    const uint32_t k = blockIdx.x * kB + threadIdx.x;
    if (k < A) out[k] = areaFixOf(area, ids[k]);
}

using SegStat = LalSegStat;

// Split of each block at half its cell area (simpl_lal.cc:225-236), from the
// inclusive fixed-point prefix P of the sorted areas.
__device__ __forceinline__ void splitBlockAt(uint32_t s, const uint32_t *__restrict__ begin,
                                             const uint32_t *__restrict__ end, const uint64_t *__restrict__ P,
                                             SegStat *__restrict__ out) {
    const uint32_t b = begin[s], e = end[s];
    SegStat st{};
    st.count = e - b;
    if (st.count > 0) {
        const uint64_t before = b > 0 ? P[b - 1] : 0ull;
        st.cellArea = P[e - 1] - before;
        const double half = 0.5 * (static_cast<double>(st.cellArea) / kFixScale);
        // smallest m >= 1 with sum of the first m areas >= half, at most count - 1
        uint32_t lo = 1, hi = st.count > 1 ? st.count - 1 : 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) / 2;
            const double low = static_cast<double>(P[b + mid - 1] - before) / kFixScale;
            if (low < half) lo = mid + 1; else hi = mid;
        }
        st.m = st.count > 1 ? lo : 0;
        st.lowArea = st.m > 0 ? P[b + st.m - 1] - before : 0ull;
    }
    out[s] = st;
}

__global__ void __launch_bounds__(kB)
splitBlocks(const uint32_t *__restrict__ begin, const uint32_t *__restrict__ end, const uint64_t *__restrict__ P,
            uint32_t B, SegStat *__restrict__ out) {
    const uint32_t s = blockIdx.x * kB + threadIdx.x;
    if (s < B) splitBlockAt(s, begin, end, P, out);
}

struct HalfArgs {
    const uint32_t *begin;     // per block, sorted by (block, centre, id)
    const SegStat *stat;
    const uint8_t *split;      // block splits this level
    const double *cut;         // per block cutline
    CentreArgs c;
};

// Half of each cell (low: first m cells of its block), key = -fromCut(centre)
// (simpl_lal.cc:284-287); cells of blocks that do not split leave the level.
__device__ __forceinline__ void halfKeyAt(uint32_t k, const uint32_t *__restrict__ ids,
                                          const uint32_t *__restrict__ seg, const HalfArgs &h,
                                          uint32_t *__restrict__ half, uint64_t *__restrict__ key2) {
    const uint32_t i = ids[k];
    const uint32_t s = seg[i];
    if (!h.split[s]) {
        half[i] = kNoVar;
        return;
    }
    const bool low = (k - h.begin[s]) < h.stat[s].m;
    const double c = centreOf(h.c, i);
    const double cut = h.cut[s];
    const double fromCut = low ? cut - c : c - cut;
    half[i] = 2 * s + (low ? 0u : 1u);
    key2[i] = orderKey(-fromCut);
}

__global__ void __launch_bounds__(kB)
halfKeys(const uint32_t *__restrict__ ids, const uint32_t *__restrict__ seg, uint32_t A, HalfArgs h,
         uint32_t *__restrict__ half, uint64_t *__restrict__ key2) {
    const uint32_t k = blockIdx.x * kB + threadIdx.x;
    if (k < A) halfKeyAt(k, ids, seg, h, half, key2);
}

__global__ void __launch_bounds__(kB)
gatherKeys(const uint32_t *__restrict__ ids, const uint64_t *__restrict__ byCell, uint32_t A,
           uint64_t *__restrict__ keys) {
    const uint32_t k = blockIdx.x * kB + threadIdx.x;
    if (k < A) keys[k] = byCell[ids[k]];
}

struct StripeArgs {
    const uint32_t *stripeStart;  // [halves + 1] stripes of each half, fill order
    const double *span;           // [stripes][2] start, end
    const double *cap;            // [stripes]
    uint32_t *stripeEnd;          // [stripes] exclusive end index into the sorted list
    unsigned long long *lo;       // [stripes] min centre (order key)
    unsigned long long *hi;       // [stripes] max centre (order key)
};

// Greedy stripe filling (simpl_lal.cc:288-295): cells in order of distance
// from the cut, a stripe takes cells while its used area is below capacity.
// With the inclusive fixed-point prefix P over the sorted list, stripe s ends
// at the first cell index k with (P[k-1] - P[start_s - 1]) >= cap_s; the last
// stripe takes the rest.
__device__ __forceinline__ void stripeRangeAt(uint32_t hh, const uint32_t *__restrict__ begin,
                                              const uint32_t *__restrict__ end, const uint64_t *__restrict__ P,
                                              const StripeArgs &sa) {
    const uint32_t s0 = sa.stripeStart[hh], s1 = sa.stripeStart[hh + 1];
    if (s0 == s1) return;
    const uint32_t b = begin[hh], e = end[hh];
    uint32_t start = b;
    for (uint32_t s = s0; s < s1; ++s) {
        uint32_t stop = e;
        if (s + 1 < s1) {
            const uint64_t base = start > 0 ? P[start - 1] : 0ull;
            const double cap = sa.cap[s];
            // first k in [start, e) with used(k) = area of cells [start, k) >= cap
            uint32_t lo = start, hiK = e;
            while (lo < hiK) {
                const uint32_t mid = (lo + hiK) / 2;
                const uint64_t before = mid > 0 ? P[mid - 1] : 0ull;
                const double used = static_cast<double>(before - base) / kFixScale;
                if (used >= cap) hiK = mid; else lo = mid + 1;
            }
            stop = lo;
        }
        sa.stripeEnd[s] = stop;
        start = stop;
    }
}

__global__ void __launch_bounds__(kB)
stripeRanges(const uint32_t *__restrict__ begin, const uint32_t *__restrict__ end, const uint64_t *__restrict__ P,
             uint32_t halves, StripeArgs sa) {
    const uint32_t hh = blockIdx.x * kB + threadIdx.x;
    if (hh < halves) stripeRangeAt(hh, begin, end, P, sa);
}

__device__ __forceinline__ void stripeOfCellAt(uint32_t k, const uint32_t *__restrict__ ids,
                                               const uint32_t *__restrict__ half, const StripeArgs &sa,
                                               const CentreArgs &c, uint32_t *__restrict__ stripeOf) {
    const uint32_t i = ids[k];
    const uint32_t hh = half[i];
    uint32_t s = sa.stripeStart[hh];
    const uint32_t s1 = sa.stripeStart[hh + 1];
    while (s + 1 < s1 && sa.stripeEnd[s] <= k) ++s;
    stripeOf[k] = s;
    const unsigned long long key = orderKey(centreOf(c, i));
    atomicMin(sa.lo + s, key);
    atomicMax(sa.hi + s, key);
}

__global__ void __launch_bounds__(kB)
stripeOfCell(const uint32_t *__restrict__ ids, const uint32_t *__restrict__ half, uint32_t A, StripeArgs sa,
             CentreArgs c, uint32_t *__restrict__ stripeOf) {
    const uint32_t k = blockIdx.x * kB + threadIdx.x;
    if (k < A) stripeOfCellAt(k, ids, half, sa, c, stripeOf);
}

// Affine map inside each stripe (simpl_lal.cc:297-314), stripe extent given.
__device__ __forceinline__ void scaleCellWith(uint32_t k, const uint32_t *__restrict__ ids, uint32_t s,
                                              unsigned long long loKey, unsigned long long hiKey, const StripeArgs &sa,
                                              const CentreArgs &c, double *__restrict__ posOut) {
    const uint32_t i = ids[k];
    const double lo = fromOrderKey(loKey), hi = fromOrderKey(hiKey);
    const double ci = centreOf(c, i);
    const double t = hi > lo ? (ci - lo) / (hi - lo) : 0.5;
    const double size = c.size[i];
    const double start = sa.span[2 * s], stop = sa.span[2 * s + 1];
    const double cc = start + t * (stop - start);
    const double upper = start < stop - size ? stop - size : start;
    const double v = cc - 0.5 * size;
    posOut[i] = v < start ? start : (upper < v ? upper : v);
}

__global__ void __launch_bounds__(kB)
scaleCells(const uint32_t *__restrict__ ids, uint32_t A, const uint32_t *__restrict__ stripeOf, StripeArgs sa,
           CentreArgs c, double *__restrict__ posOut) {
    const uint32_t k = blockIdx.x * kB + threadIdx.x;
    if (k >= A) return;
    const uint32_t s = stripeOf[k];
    scaleCellWith(k, ids, s, sa.lo[s], sa.hi[s], sa, c, posOut);
}

// ------------------------------------------------------------ small batches
// A batch of at most kSmallCells cells runs each half of a level in one
// workgroup: the same per-element routines as the multi-kernel path, with
// workgroup barriers between the steps instead of kernel boundaries, and the
// orders taken by one bitonic sort of (block, key, id) in LDS.  Inactive cells
// sort last, so the list needs no compaction.
constexpr uint32_t kSmallCells = 2048;
constexpr int kSmallThreads = 1024;

// Threads for a small batch sorted at `width` (a power of two): width / 2,
// at least one wave.  Every loop of the small kernels strides by blockDim.x,
// and the block scan covers 2 * blockDim.x >= width elements, so a batch of a
// few cells runs in one wave, whose barriers cost next to nothing, instead of
// sixteen.
inline uint32_t smallThreads(uint32_t width) {
    return std::min<uint32_t>(kSmallThreads, std::max<uint32_t>(64u, width / 2));
}

__device__ void bitonicTriples(uint32_t *ss, uint64_t *sk, uint32_t *si, uint32_t width) {
    for (uint32_t k = 2; k <= width; k <<= 1) {
        for (uint32_t d = k >> 1; d > 0; d >>= 1) {
            for (uint32_t t = threadIdx.x; t < width; t += blockDim.x) {
                const uint32_t u = t ^ d;
                if (u > t) {
                    const bool greater = ss[t] != ss[u] ? ss[t] > ss[u] : sk[t] != sk[u] ? sk[t] > sk[u] : si[t] > si[u];
                    if (greater == ((t & k) == 0)) {
                        const uint32_t s0 = ss[t], i0 = si[t];
                        const uint64_t k0 = sk[t];
                        ss[t] = ss[u];
                        sk[t] = sk[u];
                        si[t] = si[u];
                        ss[u] = s0;
                        sk[u] = k0;
                        si[u] = i0;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// Inclusive prefix sum in place of v[0, n), n <= 2 * blockDim.x (exact:
// integer sums do not depend on the order).
__device__ void blockScanU64(uint64_t *v, uint32_t n, uint64_t *waveTotals) {
    const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const uint64_t a0 = 2 * t < n ? v[2 * t] : 0ull, a1 = 2 * t + 1 < n ? v[2 * t + 1] : 0ull;
    const uint64_t local = a0 + a1;
    uint64_t x = local;
    for (int off = 1; off < 64; off <<= 1) {
        const uint64_t y = __shfl_up(x, off, 64);
        if (lane >= static_cast<uint32_t>(off)) x += y;
    }
    if (lane == 63) waveTotals[wave] = x;
    __syncthreads();
    uint64_t offset = 0;
    for (uint32_t w = 0; w < wave; ++w) offset += waveTotals[w];
    const uint64_t excl = offset + x - local;
    if (2 * t < n) v[2 * t] = excl + a0;
    if (2 * t + 1 < n) v[2 * t + 1] = excl + a0 + a1;
    __syncthreads();
}

struct SmallArgs {
    const uint32_t *list;   // the batch's cells
    uint32_t M, width;      // cells, sort width (power of two >= M)
    uint32_t *seg, *half;   // per cell
    uint64_t *key2;         // per cell
    CentreArgs c;
    const double *area;
    uint32_t *ids;          // sorted cell ids of the level
    uint32_t *begin, *end;  // per block / half
    SegStat *stats;
    uint32_t *stripeOf;
    uint32_t B, A;          // blocks and active cells of the level
};

// centre order, block ranges, area prefix, split (first half of a level)
__global__ void __launch_bounds__(kSmallThreads) smallLevelA(SmallArgs a) {
    __shared__ uint32_t ss[kSmallCells], si[kSmallCells];
    __shared__ uint64_t sk[kSmallCells];
    __shared__ uint64_t sP[kSmallCells];  // the area prefix, searched by the split
    __shared__ uint64_t waveTotals[kSmallThreads / 64];
    for (uint32_t j = threadIdx.x; j < a.width; j += blockDim.x) {
        if (j < a.M) {
            const uint32_t id = a.list[j], sg = a.seg[id];
            ss[j] = sg;
            si[j] = id;
            sk[j] = sg != kNoVar ? orderKey(centreOf(a.c, id)) : ~0ull;
        } else {
            ss[j] = kNoVar;
            si[j] = kNoVar;
            sk[j] = ~0ull;
        }
    }
    __syncthreads();
    bitonicTriples(ss, sk, si, a.width);
    for (uint32_t k = threadIdx.x; k < a.A; k += blockDim.x) a.ids[k] = si[k];
    for (uint32_t q = threadIdx.x; q < 2 * a.B + 2; q += blockDim.x) {
        a.begin[q] = 0;
        a.end[q] = 0;
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < a.A; k += blockDim.x) {  // ss[k], si[k]: block and id of cell k
        const uint32_t sg = ss[k];
        if (k == 0 || ss[k - 1] != sg) a.begin[sg] = k;
        if (k + 1 == a.A || ss[k + 1] != sg) a.end[sg] = k + 1;
        sP[k] = areaFixOf(a.area, si[k]);
    }
    __syncthreads();
    blockScanU64(sP, a.A, waveTotals);
    for (uint32_t s = threadIdx.x; s < a.B; s += blockDim.x) splitBlockAt(s, a.begin, a.end, sP, a.stats);
}

// halves, distance order, stripes, affine map, next level's blocks
__global__ void __launch_bounds__(kSmallThreads)
smallLevelB(SmallArgs a, HalfArgs ha, StripeArgs sa, const uint32_t *__restrict__ remap, uint32_t H, uint32_t S,
            uint32_t A2, double *__restrict__ posOut) {
    __shared__ uint32_t ss[kSmallCells], si[kSmallCells];
    __shared__ uint64_t sk[kSmallCells];
    __shared__ uint64_t sP[kSmallCells];  // the area prefix, searched by the stripe filling
    __shared__ uint64_t waveTotals[kSmallThreads / 64];
    for (uint32_t j = threadIdx.x; j < a.M; j += blockDim.x) a.half[a.list[j]] = kNoVar;
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < a.A; k += blockDim.x) halfKeyAt(k, a.ids, a.seg, ha, a.half, a.key2);
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < a.width; j += blockDim.x) {
        if (j < a.M) {
            const uint32_t id = a.list[j], h = a.half[id];
            ss[j] = h;
            si[j] = id;
            sk[j] = h != kNoVar ? a.key2[id] : ~0ull;
        } else {
            ss[j] = kNoVar;
            si[j] = kNoVar;
            sk[j] = ~0ull;
        }
    }
    __syncthreads();
    bitonicTriples(ss, sk, si, a.width);
    for (uint32_t k = threadIdx.x; k < A2; k += blockDim.x) a.ids[k] = si[k];
    for (uint32_t q = threadIdx.x; q < 2 * a.B + 2; q += blockDim.x) {
        a.begin[q] = 0;
        a.end[q] = 0;
    }
    for (uint32_t q = threadIdx.x; q < S + 1; q += blockDim.x) {
        sa.lo[q] = ~0ull;
        sa.hi[q] = 0ull;
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < A2; k += blockDim.x) {  // ss[k], si[k]: half and id of cell k
        const uint32_t hh = ss[k];
        if (k == 0 || ss[k - 1] != hh) a.begin[hh] = k;
        if (k + 1 == A2 || ss[k + 1] != hh) a.end[hh] = k + 1;
        sP[k] = areaFixOf(a.area, si[k]);
    }
    __syncthreads();
    blockScanU64(sP, A2, waveTotals);
    for (uint32_t hh = threadIdx.x; hh < H; hh += blockDim.x) stripeRangeAt(hh, a.begin, a.end, sP, sa);
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < A2; k += blockDim.x) {
        const uint32_t hh = ss[k];
        uint32_t st = sa.stripeStart[hh];
        const uint32_t s1 = sa.stripeStart[hh + 1];
        while (st + 1 < s1 && sa.stripeEnd[st] <= k) ++st;
        a.stripeOf[k] = st;
        const unsigned long long key = orderKey(centreOf(a.c, si[k]));
        atomicMin(sa.lo + st, key);
        atomicMax(sa.hi + st, key);
    }
    __threadfence();
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < A2; k += blockDim.x) {
        // the stripe extents were formed by atomics: read them back at device scope
        const uint32_t st = a.stripeOf[k];
        const unsigned long long lo = __hip_atomic_load(sa.lo + st, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long hi = __hip_atomic_load(sa.hi + st, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        scaleCellWith(k, si, st, lo, hi, sa, a.c, posOut);
    }
    for (uint32_t j = threadIdx.x; j < a.M; j += blockDim.x) {
        const uint32_t id = a.list[j], h = a.half[id];
        a.seg[id] = h == kNoVar ? kNoVar : remap[h];
    }
}

// simpl_lal.cc:365-372
__global__ void __launch_bounds__(kB)
clampToDie(double *__restrict__ x, double *__restrict__ y, const double *__restrict__ w,
           const double *__restrict__ h, double x0, double y0, double x1, double y1, uint32_t n) {
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= n) return;
    const double ux = x0 < x1 - w[i] ? x1 - w[i] : x0;
    const double uy = y0 < y1 - h[i] ? y1 - h[i] : y0;
    x[i] = x[i] < x0 ? x0 : (ux < x[i] ? ux : x[i]);
    y[i] = y[i] < y0 ? y0 : (uy < y[i] ? uy : y[i]);
}

__global__ void __launch_bounds__(kB)
remapSegments(const uint32_t *__restrict__ half, const uint32_t *__restrict__ remap, uint32_t n,
              uint32_t *__restrict__ seg) {
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= n) return;
    const uint32_t h = half[i];
    seg[i] = h == kNoVar ? kNoVar : remap[h];
}

__global__ void __launch_bounds__(kB) fillU64(unsigned long long *p, unsigned long long v, uint64_t n) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(kB) + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void __launch_bounds__(kB) fillU32(uint32_t *p, uint32_t v, uint64_t n) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(kB) + threadIdx.x;
    if (i < n) p[i] = v;
}

template <typename F>
void parallelFor(size_t n, F f) {
    const size_t threads = std::min<size_t>(std::max(1u, std::thread::hardware_concurrency()), 16);
    if (n < 64 || threads <= 1) {
        for (size_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::vector<std::thread> pool;
    std::atomic<size_t> next{0};
    for (size_t t = 0; t < threads; ++t)
        pool.emplace_back([&] {
            for (size_t i; (i = next.fetch_add(16)) < n;)
                for (size_t k = i; k < std::min(n, i + 16); ++k) f(k);
        });
    for (auto &th : pool) th.join();
}

int bitsFor(uint64_t v) {
    int b = 0;
    while (b < 64 && (v >> b) != 0) ++b;
    return std::max(b, 1);
}

}  // namespace

// ------------------------------------------------------------------ host entry points
int lalSetDesign(kt_ctx *ctx, const std::vector<lal::RowSpec> &rows, const lal::Box &die) {
    ctx->lalRows = rows;
    ctx->lalDie = die;
    ctx->haveDesign = true;
    return KT_OK;
}

int lalBuildGrid(kt_ctx *ctx, uint64_t perAxis, double densityLimit) {
    KT_CHECK(ctx->haveNetlist && ctx->haveDesign, KT_ESTATE, "build_grid: netlist and rows needed");
    KT_CHECK(perAxis > 0 && perAxis <= 4096, KT_EINVAL, "build_grid: bins per axis out of range");
    lal::buildGrid(ctx->grid, ctx->freeSpace, ctx->lalRows, ctx->obstacles, ctx->lalDie, perAxis);
    ctx->lalTables.buildAvail(ctx->grid);
    ctx->densityLimit = std::clamp(densityLimit, 0.05, 1.0);
    KT_TRY(ctx->gridAvail.reserve(ctx->grid.size()));
    KT_HIP(hipMemcpyAsync(ctx->gridAvail.p, ctx->grid.avail.data(), ctx->grid.size() * sizeof(double),
                          hipMemcpyHostToDevice, ctx->stream));
    KT_HIP(hipStreamSynchronize(ctx->stream));
    return KT_OK;
}

namespace {

GridDev gridDev(const lal::Grid &g) {
    return GridDev{g.ox, g.oy, g.dx, g.dy, static_cast<uint32_t>(g.nx), static_cast<uint32_t>(g.ny)};
}

}  // namespace

// binOf and occupancy of positions (x, y) on the device (enqueued, no
// synchronisation).
int lalOccupancyOf(kt_ctx *ctx, const double *x, const double *y) {
    KT_CHECK(ctx->grid.size() > 0 && ctx->gridAvail.cap >= ctx->grid.size(), KT_ESTATE, "occupancy: no grid");
    const uint32_t n = ctx->n;
    const uint64_t m = ctx->grid.size();
    KT_TRY(ctx->binOf.reserve(std::max<uint32_t>(n, 1)));
    KT_TRY(ctx->occFix.reserve(m));
    KT_TRY(ctx->occ.reserve(m));
    KT_TRY(ctx->binCells.reserve(m));
    KT_HIP(hipMemsetAsync(ctx->occFix.p, 0, m * sizeof(uint64_t), ctx->stream));
    KT_HIP(hipMemsetAsync(ctx->binCells.p, 0, m * sizeof(uint32_t), ctx->stream));
    if (n) {
        // a few cells per thread: each block's bins reach memory once
        const dim3 grid(std::min<uint32_t>(gridFor(n).x, 2048u));
        cellBins<<<grid, kB, 0, ctx->stream>>>(x, y, ctx->width.p, ctx->height.p, ctx->area.p, gridDev(ctx->grid), n,
                                                ctx->binOf.p, reinterpret_cast<unsigned long long *>(ctx->occFix.p),
                                                ctx->binCells.p);
        KT_HIP(hipGetLastError());
    }
    fixToDouble<<<gridFor(m), kB, 0, ctx->stream>>>(reinterpret_cast<unsigned long long *>(ctx->occFix.p),
                                                     ctx->occ.p, static_cast<uint32_t>(m));
    KT_HIP(hipGetLastError());
    return KT_OK;
}

// Overflow excess sum (occ - g avail)+ of positions (x, y) into *dst on the
// device (enqueued; getOverflow divides it by the total free area).
int lalExcessInto(kt_ctx *ctx, const double *x, const double *y, double *dst) {
    KT_TRY(lalOccupancyOf(ctx, x, y));
    const uint64_t m = ctx->grid.size();
    const uint32_t blocks = static_cast<uint32_t>((m + kB - 1) / kB);
    KT_TRY(ctx->lalScratch.reserve(blocks + 1));
    overflowPartials<<<blocks, kB, 0, ctx->stream>>>(ctx->occ.p, ctx->gridAvail.p, ctx->densityLimit,
                                                     static_cast<uint32_t>(m), ctx->lalScratch.p);
    KT_HIP(hipGetLastError());
    sumInOrder<<<1, 64, 0, ctx->stream>>>(ctx->lalScratch.p, blocks, dst);
    KT_HIP(hipGetLastError());
    return KT_OK;
}

// binOf and occupancy of the context's positions (device), occupancy copied
// to host `occ` when non-null.
int lalOccupancy(kt_ctx *ctx, std::vector<double> *occHost, std::vector<uint32_t> *countHost) {
    KT_TRY(lalOccupancyOf(ctx, ctx->posX.p, ctx->posY.p));
    const uint64_t m = ctx->grid.size();
    if (occHost) {
        occHost->resize(m);
        KT_HIP(hipMemcpyAsync(occHost->data(), ctx->occ.p, m * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    }
    if (countHost) {
        countHost->resize(m);
        KT_HIP(hipMemcpyAsync(countHost->data(), ctx->binCells.p, m * sizeof(uint32_t), hipMemcpyDeviceToHost,
                              ctx->stream));
    }
    if (occHost || countHost) KT_HIP(hipStreamSynchronize(ctx->stream));
    return KT_OK;
}

int lalOverflow(kt_ctx *ctx, double *out) {
    const uint64_t m = ctx->grid.size();
    const uint32_t blocks = static_cast<uint32_t>((m + kB - 1) / kB);
    KT_TRY(ctx->lalScratch.reserve(blocks + 1));
    KT_TRY(lalExcessInto(ctx, ctx->posX.p, ctx->posY.p, ctx->lalScratch.p + blocks));
    double excess = 0.0;
    KT_HIP(hipMemcpyAsync(&excess, ctx->lalScratch.p + blocks, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    KT_HIP(hipStreamSynchronize(ctx->stream));
    *out = ctx->grid.totalAvail > 0.0 ? excess / ctx->grid.totalAvail : 0.0;
    return KT_OK;
}

int vertexPositionsLaunch(kt_ctx *ctx);
int vertexPositionsOf(kt_ctx *ctx, const double *x, const double *y);

// getHpwl of positions (x, y) into *dst on the device (enqueued).
int lalHpwlInto(kt_ctx *ctx, const double *x, const double *y, double *dst) {
    KT_CHECK(ctx->haveNetlist, KT_ESTATE, "hpwl: no netlist");
    KT_TRY(vertexPositionsOf(ctx, x, y));
    const uint32_t N = ctx->N;
    const uint32_t chunks = (N + 4095) / 4096;
    KT_TRY(ctx->hpwlScratch.reserve(static_cast<uint64_t>(N) + chunks + 1));
    double *perNet = ctx->hpwlScratch.p;
    double *part = perNet + N;
    if (N) {
        netHpwl<<<gridFor(N), kB, 0, ctx->stream>>>(ctx->netStart.p, ctx->pinCell.p, ctx->offX.p, ctx->offY.p,
                                                    ctx->vxy.p, N, perNet);
        KT_HIP(hipGetLastError());
        chunkSums<<<gridFor(chunks), kB, 0, ctx->stream>>>(perNet, N, part);
        KT_HIP(hipGetLastError());
    }
    sumInOrder<<<1, 64, 0, ctx->stream>>>(part, chunks, dst);
    KT_HIP(hipGetLastError());
    return KT_OK;
}

int lalHpwl(kt_ctx *ctx, double *out) {
    KT_CHECK(ctx->posCount == ctx->n, KT_ESTATE, "hpwl: positions must have one entry per movable cell");
    KT_TRY(ctx->lalScratch.reserve(1));
    KT_TRY(lalHpwlInto(ctx, ctx->posX.p, ctx->posY.p, ctx->lalScratch.p));
    double h = 0.0;
    KT_HIP(hipMemcpyAsync(&h, ctx->lalScratch.p, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    KT_HIP(hipStreamSynchronize(ctx->stream));
    *out = h;
    return KT_OK;
}

// One batch of non-touching regions (simpl_lal.cc:122-151, 178-316).
// Pinned host staging and its device twin for the per-level tables.
int lalStaging(kt_ctx *ctx, size_t bytes) {
    if (bytes > ctx->lalPinnedCap) {
        KT_HIP(hipStreamSynchronize(ctx->stream));
        if (ctx->lalPinned) (void)hipHostFree(ctx->lalPinned);
        ctx->lalPinned = nullptr;
        ctx->lalPinnedCap = 0;
        const size_t cap = std::max<size_t>(bytes, 1 << 16);
        KT_HIP(hipHostMalloc(reinterpret_cast<void **>(&ctx->lalPinned), cap));
        ctx->lalPinnedCap = cap;
    }
    KT_TRY(ctx->lalUp.reserve((bytes + 7) / 8));
    return KT_OK;
}

// Pinned landing area for a batch's small transfers: a batch makes several
// round trips to the host (member count, per-level block stats, escape
// checks, moved cells), and a transfer to or from pageable memory is staged
// by the runtime at extra cost each time.  Layout: kLandScalars 32-bit slots,
// then one variable part used in turn by the region rectangles, the block
// stats of a level and the moved cells (each read or sent completely before
// the next one is staged).  It grows only after a synchronisation, when
// nothing in flight uses it.
constexpr size_t kLandScalars = 16;
enum LandSlot { kLandActive = 2, kLandNone = 4 };

int lalLandReserve(kt_ctx *ctx, size_t varBytes) {
    const size_t bytes = kLandScalars * sizeof(uint32_t) + varBytes;
    if (bytes <= ctx->lalLandCap) return KT_OK;
    KT_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->lalLand) (void)hipHostFree(ctx->lalLand);
    ctx->lalLand = nullptr;
    ctx->lalLandCap = 0;
    const size_t cap = std::max<size_t>(2 * bytes, 1 << 16);
    KT_HIP(hipHostMalloc(reinterpret_cast<void **>(&ctx->lalLand), cap));
    ctx->lalLandCap = cap;
    return KT_OK;
}
inline uint32_t *landSlots(kt_ctx *ctx) { return reinterpret_cast<uint32_t *>(ctx->lalLand); }
inline uint8_t *landVar(kt_ctx *ctx) { return ctx->lalLand + kLandScalars * sizeof(uint32_t); }
uint8_t *lalLandVar(kt_ctx *ctx) { return landVar(ctx); }

int lalRunBatchFused(kt_ctx *ctx, const std::vector<lal::BinRect> &regions,
                     const std::vector<std::pair<uint32_t, uint32_t>> &cells, std::vector<LalMoved> &moved, bool *ran);

// Legalizes one batch of non-touching regions; the cells whose bin changed
// are returned in `moved` (their device binOf is updated).
int lalRunBatch(kt_ctx *ctx, const std::vector<lal::BinRect> &regions, const std::vector<uint32_t> &binCells,
                std::vector<Moved> &moved) {
    const uint32_t n = ctx->n;
    const lal::Grid &g = ctx->grid;
    moved.clear();
    if (regions.empty() || n == 0) return KT_OK;
    struct Block {
        lal::Box box;
        double regionRoom;
    };
    std::vector<Block> blocks(regions.size());
    std::vector<uint4> rects(regions.size());
    for (size_t r = 0; r < regions.size(); ++r) {
        const lal::BinRect &b = regions[r];
        rects[r] = make_uint4(static_cast<uint32_t>(b.x0), static_cast<uint32_t>(b.x1), static_cast<uint32_t>(b.y0),
                              static_cast<uint32_t>(b.y1));
        blocks[r] = {g.box(b), g.sum(g.avail, b)};
    }
    const uint32_t R = static_cast<uint32_t>(regions.size());
    // per-cell scratch
    for (auto *b : {&ctx->lalSeg, &ctx->lalFlag, &ctx->lalIds, &ctx->lalVals, &ctx->lalValsTmp, &ctx->lalHalf,
                    &ctx->lalStripeOf, &ctx->lalList})
        KT_TRY(b->reserve(n));
    for (auto *b : {&ctx->lalKeys, &ctx->lalKeysTmp, &ctx->lalKey2, &ctx->lalPrefix}) KT_TRY(b->reserve(n));
    // the batch's cell count from the host's cells per bin (the regions do
    // not overlap): no round trip; the device's own count is checked against
    // it with the batch's last transfer
    uint64_t cellsInRegions = 0;
    bool splits = false;
    for (const lal::BinRect &b : regions) {
        uint64_t inRegion = 0;
        for (size_t iy = b.y0; iy <= b.y1; ++iy)
            for (size_t ix = b.x0; ix <= b.x1; ++ix) inRegion += binCells[g.at(ix, iy)];
        cellsInRegions += inRegion;
        splits = splits || inRegion > ctx->lalOpts.minCellsToSplit;
    }
    KT_CHECK(cellsInRegions <= n, KT_ESTATE, "lal: cells per bin out of step with the device");
    const uint32_t M = static_cast<uint32_t>(cellsInRegions);
    // A region of at most minCellsToSplit cells is a block that does not
    // split (simpl_lal.cc:216-218): its cells keep their positions and bins.
    // A batch of such regions (or with a level cap of 1) moves no cell, so it
    // has nothing to do on the device.
    if (!splits || ctx->lalOpts.maxLevel <= 1) return KT_OK;
    // batches of up to kSmallCells cells: one workgroup per half level
    static const bool smallPath = [] {
        const char *v = std::getenv("KTPLACE_LAL_SMALL");
        return !(v && v[0] == '0');
    }();
    const bool small = smallPath && M <= kSmallCells;
    // the batch's cells: for the small path in any order, appended in one
    // pass; otherwise in id order (the stable radix sorts start from it).
    // lalMoved holds the batch's four counters (two words) and then up to M
    // moved records (three words each), so one transfer returns both.  The
    // counters: [0] first region a cell escaped into (kNoVar: none), [1]
    // cells of a rerun, [2] moved records, [3] the device's count of the
    // batch's cells; the paint kernel sets them to {kNoVar, 0, 0, 0}.
    KT_TRY(ctx->lalMoved.reserve(2ull + 3ull * M));
    uint32_t *cnt = reinterpret_cast<uint32_t *>(ctx->lalMoved.p);
    Moved *dMoved = reinterpret_cast<Moved *>(ctx->lalMoved.p + 2);
    KT_TRY(lalLandReserve(ctx, 0));  // the landing area exists before its first use
    // bin -> region for the batch's bins (every other bin holds kNoVar); a
    // few regions travel in the launch arguments, more through a copy.  The
    // paint's first block also sets the counters.
    const bool rectsByValue = R <= kRectsByValue;
    RectSet rectSet{};
    const uint4 *dRects = nullptr;
    if (rectsByValue) {
        for (uint32_t r = 0; r < R; ++r) rectSet.r[r] = rects[r];
        paintRegionsArg<<<R, kB, 0, ctx->stream>>>(rectSet, static_cast<uint32_t>(g.nx), ctx->binRegion.p, 0, cnt);
    } else {
        KT_TRY(ctx->lalRects.reserve(4ull * R));
        KT_TRY(lalLandReserve(ctx, R * sizeof(uint4)));
        std::memcpy(landVar(ctx), rects.data(), R * sizeof(uint4));
        KT_HIP(hipMemcpyAsync(ctx->lalRects.p, landVar(ctx), R * sizeof(uint4), hipMemcpyHostToDevice, ctx->stream));
        dRects = reinterpret_cast<const uint4 *>(ctx->lalRects.p);
        paintRegions<<<R, kB, 0, ctx->stream>>>(dRects, static_cast<uint32_t>(g.nx), ctx->binRegion.p, 0, cnt);
    }
    KT_HIP(hipGetLastError());
    // restore points of the batch's cells and their bins at the start
    for (auto *b : {&ctx->lalRegionOf, &ctx->lalSaveBin, &ctx->lalBinStart}) KT_TRY(b->reserve(n));
    for (auto *b : {&ctx->lalSaveX, &ctx->lalSaveY}) KT_TRY(b->reserve(n));
    BatchState bs{ctx->lalRegionOf.p, ctx->lalSaveX.p, ctx->lalSaveY.p, ctx->lalSaveBin.p, ctx->lalBinStart.p,
                  ctx->posX.p, ctx->posY.p, ctx->width.p, ctx->height.p, ctx->binOf.p, ctx->binRegion.p,
                  ctx->lalSeg.p, gridDev(g)};
    if (small) {
        regionMembers<<<gridFor(n), kB, 0, ctx->stream>>>(n, bs, ctx->lalList.p, cnt + 3);
        KT_HIP(hipGetLastError());
    } else {
        regionFlags<<<gridFor(n), kB, 0, ctx->stream>>>(ctx->binOf.p, ctx->binRegion.p, n, ctx->lalSeg.p,
                                                        ctx->lalFlag.p);
        KT_HIP(hipGetLastError());
        KT_TRY(scanInclusive(ctx, ctx->lalFlag.p, n));
        compactIds<<<gridFor(n), kB, 0, ctx->stream>>>(ctx->lalSeg.p, ctx->lalFlag.p, n, ctx->lalList.p);
        KT_HIP(hipGetLastError());
        // kept aside: the levels reuse lalFlag
        KT_HIP(hipMemcpyAsync(cnt + 3, ctx->lalFlag.p + n - 1, sizeof(uint32_t), hipMemcpyDeviceToDevice,
                              ctx->stream));
    }
    const auto paint = [&](int clear) -> int {
        if (rectsByValue)
            paintRegionsArg<<<R, kB, 0, ctx->stream>>>(rectSet, static_cast<uint32_t>(g.nx), ctx->binRegion.p, clear,
                                                       nullptr);
        else
            paintRegions<<<R, kB, 0, ctx->stream>>>(dRects, static_cast<uint32_t>(g.nx), ctx->binRegion.p, clear,
                                                    nullptr);
        KT_HIP(hipGetLastError());
        return KT_OK;
    };
    const auto unpaint = [&]() -> int { return paint(1); };
    if (M == 0) return unpaint();
    const uint32_t *list = ctx->lalList.p;
    if (!small) {  // the small path's member pass recorded them
        saveBatch<<<gridFor(M), kB, 0, ctx->stream>>>(list, M, bs);
        KT_HIP(hipGetLastError());
    }
    uint32_t width = 2;
    while (width < M) width <<= 1;

    // The recursion over the levels for the cells whose lalSeg names a region.
    const std::vector<Block> regionBlocks = blocks;
    const auto runLevels = [&](uint32_t A) -> int {
        std::vector<Block> blocks = regionBlocks;
        bool vertical = true;
        // cells of the level: those with a region at level 1, then those of
        // the blocks that split (known on the host from the block statistics)
        for (size_t level = 1;; ++level) {
            if (A == 0) break;
            const uint32_t B = static_cast<uint32_t>(blocks.size());
            const CentreArgs ca{vertical ? ctx->posX.p : ctx->posY.p, vertical ? ctx->width.p : ctx->height.p};
            KT_TRY(ctx->lalBegin.reserve(2ull * B + 2));
            KT_TRY(ctx->lalEnd.reserve(2ull * B + 2));
            KT_TRY(ctx->lalStats.reserve(B));
            SmallArgs sm{};
            if (small) {
                sm.list = list;
                sm.M = M;
                sm.width = width;
                sm.seg = ctx->lalSeg.p;
                sm.half = ctx->lalHalf.p;
                sm.key2 = ctx->lalKey2.p;
                sm.c = ca;
                sm.area = ctx->area.p;
                sm.ids = ctx->lalIds.p;
                sm.begin = ctx->lalBegin.p;
                sm.end = ctx->lalEnd.p;
                sm.stats = ctx->lalStats.p;
                sm.stripeOf = ctx->lalStripeOf.p;
                sm.B = B;
                sm.A = A;
                smallLevelA<<<1, smallThreads(width), 0, ctx->stream>>>(sm);
                KT_HIP(hipGetLastError());
            } else {
                // active cells in id order
                listFlags<<<gridFor(M), kB, 0, ctx->stream>>>(list, ctx->lalSeg.p, M, ctx->lalFlag.p);
                KT_HIP(hipGetLastError());
                KT_TRY(scanInclusive(ctx, ctx->lalFlag.p, M));
                listCompact<<<gridFor(M), kB, 0, ctx->stream>>>(list, ctx->lalSeg.p, ctx->lalFlag.p, M, ctx->lalIds.p);
                KT_HIP(hipGetLastError());
                // sort by (block, centre, id): stable sorts by centre, then by block
                centreKeys<<<gridFor(A), kB, 0, ctx->stream>>>(ctx->lalIds.p, A, ca, ctx->lalKeys.p);
                KT_HIP(hipGetLastError());
                if (smallSortFits(A)) {
                    KT_TRY(smallSortBySegment(ctx, ctx->lalIds.p, ctx->lalKeys.p, ctx->lalSeg.p, A));
                } else {
                    KT_TRY(radixSortPairs(ctx, ctx->lalKeys.p, ctx->lalIds.p, ctx->lalKeysTmp.p, ctx->lalValsTmp.p, A,
                                          64, ctx->sortCounts, ctx->sortRaw));
                    segmentKeys<<<gridFor(A), kB, 0, ctx->stream>>>(ctx->lalIds.p, ctx->lalSeg.p, A, ctx->lalKeys.p);
                    KT_HIP(hipGetLastError());
                    KT_TRY(radixSortPairs(ctx, ctx->lalKeys.p, ctx->lalIds.p, ctx->lalKeysTmp.p, ctx->lalValsTmp.p, A,
                                          bitsFor(B), ctx->sortCounts, ctx->sortRaw));
                }
                // block ranges, area prefix, split
                fillU32<<<gridFor(2ull * B + 2), kB, 0, ctx->stream>>>(ctx->lalBegin.p, 0u, 2ull * B + 2);
                fillU32<<<gridFor(2ull * B + 2), kB, 0, ctx->stream>>>(ctx->lalEnd.p, 0u, 2ull * B + 2);
                segmentBounds<<<gridFor(A), kB, 0, ctx->stream>>>(ctx->lalIds.p, ctx->lalSeg.p, A, ctx->lalBegin.p,
                                                                  ctx->lalEnd.p);
                areaFix<<<gridFor(A), kB, 0, ctx->stream>>>(ctx->lalIds.p, ctx->area.p, A, ctx->lalPrefix.p);
                KT_TRY(scanInclusive64(ctx, ctx->lalPrefix.p, A));
                splitBlocks<<<gridFor(B), kB, 0, ctx->stream>>>(ctx->lalBegin.p, ctx->lalEnd.p, ctx->lalPrefix.p, B,
                                                                ctx->lalStats.p);
                KT_HIP(hipGetLastError());
            }
            std::vector<SegStat> stats(B);
            KT_TRY(lalLandReserve(ctx, B * sizeof(SegStat)));
            KT_HIP(hipMemcpyAsync(landVar(ctx), ctx->lalStats.p, B * sizeof(SegStat), hipMemcpyDeviceToHost, ctx->stream));
            KT_HIP(hipStreamSynchronize(ctx->stream));
            std::memcpy(stats.data(), landVar(ctx), B * sizeof(SegStat));

            // host geometry for every splitting block (simpl_lal.cc:213-263)
            std::vector<uint8_t> split(B, 0);
            std::vector<double> cut(B, 0.0);
            std::vector<std::vector<lal::Stripe>> halfStripes(2ull * B);
            std::vector<Block> next(2ull * B);
            const lal::FreeSpace &fs = ctx->freeSpace;
            const double g2 = ctx->densityLimit;
            const double frac = ctx->lalOpts.stripeAreaFraction;
            parallelFor(B, [&](size_t s) {
                const SegStat &st = stats[s];
                if (level >= ctx->lalOpts.maxLevel || st.count <= ctx->lalOpts.minCellsToSplit) return;
                split[s] = 1;
                const Block &bl = blocks[s];
                const double cellArea = static_cast<double>(st.cellArea) / kFixScale;
                const double lowArea = static_cast<double>(st.lowArea) / kFixScale;
                const double lo = vertical ? bl.box.x0 : bl.box.y0;
                const double hi = vertical ? bl.box.x1 : bl.box.y1;
                const lal::Profile room = fs.profile(bl.box, vertical);
                const double share = cellArea > 0.0 ? lowArea / cellArea : 0.5;
                const double cutline = room.total() > 0.0 ? room.coordinateAt(share * room.total()) : lo + share * (hi - lo);
                cut[s] = cutline;
                lal::Box lowBox = bl.box, highBox = bl.box;
                (vertical ? lowBox.x1 : lowBox.y1) = cutline;
                (vertical ? highBox.x0 : highBox.y0) = cutline;
                const double highArea = static_cast<double>(st.cellArea - st.lowArea) / kFixScale;
                if (st.m > 0)
                    halfStripes[2 * s] = lal::stripes(fs, ctx->obstacles, lowBox, vertical, lowArea, bl.regionRoom, g2,
                                                      frac, cutline, true);
                if (st.count > st.m)
                    halfStripes[2 * s + 1] = lal::stripes(fs, ctx->obstacles, highBox, vertical, highArea, bl.regionRoom,
                                                          g2, frac, cutline, false);
                // a zero-length half box yields no stripe; give its cells one
                // stripe at the box edge so they stay inside it
                if (st.m > 0 && halfStripes[2 * s].empty())
                    halfStripes[2 * s].push_back({vertical ? lowBox.x0 : lowBox.y0, vertical ? lowBox.x1 : lowBox.y1, 0.0});
                if (st.count > st.m && halfStripes[2 * s + 1].empty())
                    halfStripes[2 * s + 1].push_back(
                        {vertical ? highBox.x0 : highBox.y0, vertical ? highBox.x1 : highBox.y1, 0.0});
                next[2 * s] = {lowBox, bl.regionRoom};
                next[2 * s + 1] = {highBox, bl.regionRoom};
            });
            uint32_t splits = 0, A2 = 0;
            for (uint32_t s = 0; s < B; ++s)
                if (split[s]) {
                    ++splits;
                    A2 += stats[s].count;
                }
            if (splits == 0) break;
            // stripe tables
            const uint32_t H = 2 * B;
            std::vector<uint32_t> stripeStart(H + 1, 0);
            for (uint32_t h = 0; h < H; ++h) stripeStart[h + 1] = stripeStart[h] + halfStripes[h].size();
            const uint32_t S = stripeStart[H];
            std::vector<double> span(2ull * S + 2), cap(S + 1);
            for (uint32_t h = 0; h < H; ++h)
                for (size_t q = 0; q < halfStripes[h].size(); ++q) {
                    const uint32_t s = stripeStart[h] + static_cast<uint32_t>(q);
                    span[2 * s] = halfStripes[h][q].start;
                    span[2 * s + 1] = halfStripes[h][q].end;
                    cap[s] = halfStripes[h][q].capacity;
                }
            // next level: the halves that hold cells, renumbered densely
            std::vector<uint32_t> remap(H, kNoVar);
            std::vector<Block> kept;
            for (uint32_t s = 0; s < B; ++s) {
                if (!split[s]) continue;
                if (stats[s].m > 0) {
                    remap[2 * s] = static_cast<uint32_t>(kept.size());
                    kept.push_back(next[2 * s]);
                }
                if (stats[s].count > stats[s].m) {
                    remap[2 * s + 1] = static_cast<uint32_t>(kept.size());
                    kept.push_back(next[2 * s + 1]);
                }
            }
            // one upload of the level's host tables through pinned memory
            const auto up8 = [](size_t bytes) { return (bytes + 7) & ~size_t{7}; };
            const size_t oCut = 0, oSpan = oCut + up8(B * sizeof(double)), oCap = oSpan + up8(span.size() * sizeof(double)),
                         oStart = oCap + up8(cap.size() * sizeof(double)),
                         oRemap = oStart + up8((H + 1) * sizeof(uint32_t)), oSplit = oRemap + up8(H * sizeof(uint32_t)),
                         total = oSplit + up8(B);
            KT_TRY(lalStaging(ctx, total));
            uint8_t *host = ctx->lalPinned;
            std::memcpy(host + oCut, cut.data(), B * sizeof(double));
            std::memcpy(host + oSpan, span.data(), span.size() * sizeof(double));
            std::memcpy(host + oCap, cap.data(), cap.size() * sizeof(double));
            std::memcpy(host + oStart, stripeStart.data(), (H + 1) * sizeof(uint32_t));
            std::memcpy(host + oRemap, remap.data(), H * sizeof(uint32_t));
            std::memcpy(host + oSplit, split.data(), B);
            uint8_t *dev = reinterpret_cast<uint8_t *>(ctx->lalUp.p);
            KT_HIP(hipMemcpyAsync(dev, host, total, hipMemcpyHostToDevice, ctx->stream));
            const double *dCut = reinterpret_cast<const double *>(dev + oCut);
            const double *dSpan = reinterpret_cast<const double *>(dev + oSpan);
            const double *dCap = reinterpret_cast<const double *>(dev + oCap);
            const uint32_t *dStart = reinterpret_cast<const uint32_t *>(dev + oStart);
            const uint32_t *dRemap = reinterpret_cast<const uint32_t *>(dev + oRemap);
            const uint8_t *dSplit = dev + oSplit;
            KT_TRY(ctx->lalStripeEnd.reserve(S + 1));
            KT_TRY(ctx->lalStripeLo.reserve(S + 1));
            KT_TRY(ctx->lalStripeHi.reserve(S + 1));
            // halves and the second order: by (half, -fromCut(centre), id)
            HalfArgs ha{ctx->lalBegin.p, ctx->lalStats.p, dSplit, dCut, ca};
            StripeArgs sa{dStart, dSpan, dCap, ctx->lalStripeEnd.p,
                          reinterpret_cast<unsigned long long *>(ctx->lalStripeLo.p),
                          reinterpret_cast<unsigned long long *>(ctx->lalStripeHi.p)};
            if (small) {
                smallLevelB<<<1, smallThreads(width), 0, ctx->stream>>>(sm, ha, sa, dRemap, H, S, A2,
                                                                         const_cast<double *>(ca.pos));
                KT_HIP(hipGetLastError());
            } else {
                listFill<<<gridFor(M), kB, 0, ctx->stream>>>(list, M, ctx->lalHalf.p, kNoVar);
                halfKeys<<<gridFor(A), kB, 0, ctx->stream>>>(ctx->lalIds.p, ctx->lalSeg.p, A, ha, ctx->lalHalf.p,
                                                             ctx->lalKey2.p);
                KT_HIP(hipGetLastError());
                // cells of non-splitting blocks leave; the rest are ordered by half
                listFlags<<<gridFor(M), kB, 0, ctx->stream>>>(list, ctx->lalHalf.p, M, ctx->lalFlag.p);
                KT_HIP(hipGetLastError());
                KT_TRY(scanInclusive(ctx, ctx->lalFlag.p, M));
                listCompact<<<gridFor(M), kB, 0, ctx->stream>>>(list, ctx->lalHalf.p, ctx->lalFlag.p, M, ctx->lalIds.p);
                gatherKeys<<<gridFor(A2), kB, 0, ctx->stream>>>(ctx->lalIds.p, ctx->lalKey2.p, A2, ctx->lalKeys.p);
                KT_HIP(hipGetLastError());
                if (smallSortFits(A2)) {
                    KT_TRY(smallSortBySegment(ctx, ctx->lalIds.p, ctx->lalKeys.p, ctx->lalHalf.p, A2));
                } else {
                    KT_TRY(radixSortPairs(ctx, ctx->lalKeys.p, ctx->lalIds.p, ctx->lalKeysTmp.p, ctx->lalValsTmp.p, A2, 64,
                                          ctx->sortCounts, ctx->sortRaw));
                    segmentKeys<<<gridFor(A2), kB, 0, ctx->stream>>>(ctx->lalIds.p, ctx->lalHalf.p, A2, ctx->lalKeys.p);
                    KT_HIP(hipGetLastError());
                    KT_TRY(radixSortPairs(ctx, ctx->lalKeys.p, ctx->lalIds.p, ctx->lalKeysTmp.p, ctx->lalValsTmp.p, A2,
                                          bitsFor(H), ctx->sortCounts, ctx->sortRaw));
                }
                fillU32<<<gridFor(2ull * B + 2), kB, 0, ctx->stream>>>(ctx->lalBegin.p, 0u, 2ull * B + 2);
                fillU32<<<gridFor(2ull * B + 2), kB, 0, ctx->stream>>>(ctx->lalEnd.p, 0u, 2ull * B + 2);
                segmentBounds<<<gridFor(A2), kB, 0, ctx->stream>>>(ctx->lalIds.p, ctx->lalHalf.p, A2, ctx->lalBegin.p,
                                                                   ctx->lalEnd.p);
                areaFix<<<gridFor(A2), kB, 0, ctx->stream>>>(ctx->lalIds.p, ctx->area.p, A2, ctx->lalPrefix.p);
                KT_TRY(scanInclusive64(ctx, ctx->lalPrefix.p, A2));
                fillU64<<<gridFor(S + 1), kB, 0, ctx->stream>>>(sa.lo, ~0ull, S + 1);
                fillU64<<<gridFor(S + 1), kB, 0, ctx->stream>>>(sa.hi, 0ull, S + 1);
                stripeRanges<<<gridFor(H), kB, 0, ctx->stream>>>(ctx->lalBegin.p, ctx->lalEnd.p, ctx->lalPrefix.p, H, sa);
                KT_HIP(hipGetLastError());
                stripeOfCell<<<gridFor(A2), kB, 0, ctx->stream>>>(ctx->lalIds.p, ctx->lalHalf.p, A2, sa, ca,
                                                                  ctx->lalStripeOf.p);
                KT_HIP(hipGetLastError());
                scaleCells<<<gridFor(A2), kB, 0, ctx->stream>>>(ctx->lalIds.p, A2, ctx->lalStripeOf.p, sa, ca,
                                                                const_cast<double *>(ca.pos));
                KT_HIP(hipGetLastError());
                listRemap<<<gridFor(M), kB, 0, ctx->stream>>>(list, M, ctx->lalHalf.p, dRemap, ctx->lalSeg.p);
                KT_HIP(hipGetLastError());
            }
            blocks.swap(kept);
            vertical = !vertical;
            A = A2;
            // The next level would split no block when every half holds at
            // most minCellsToSplit cells (their counts are the split's m and
            // count - m) or the level cap is reached: it would only sort and
            // measure, so it and its round trip are skipped.
            bool more = level + 1 < ctx->lalOpts.maxLevel;
            if (more) {
                more = false;
                for (uint32_t s = 0; s < B && !more; ++s)
                    more = split[s] && (stats[s].m > ctx->lalOpts.minCellsToSplit ||
                                        stats[s].count - stats[s].m > ctx->lalOpts.minCellsToSplit);
            }
            if (!more) break;
        }
        return KT_OK;
    };
    // The regions of a batch run together on the cells they held at the start.
    // The reference runs them one after the other: a cell that a region moves
    // into the bins of a later region of the batch is legalized there too.  So
    // when that happens, the regions from the first such one on are restored
    // and run again with the cells their bins hold now (the earlier regions'
    // results stand), until no cell lands in a later region.
    // Each pass: the levels, then the escape check and (speculatively) the
    // moved cells in one kernel, brought back in one round trip.  A cell that
    // moved into a later region (bad != kNoVar) discards the speculative
    // moved list: the repair rewrites every batch cell's bin and the regions
    // from `bad` run again.
    constexpr uint32_t kMovedInline = 4096;  // up to this many cells, the moved list comes with the counters
    const bool inlineMoved = M <= kMovedInline;
    const size_t headBytes = 4 * sizeof(uint32_t);
    uint32_t active = M, moves = 0;
    for (uint32_t first = 0;;) {
        KT_TRY(runLevels(active));
        const size_t back = headBytes + (inlineMoved ? static_cast<size_t>(M) * sizeof(Moved) : 0);
        KT_TRY(lalLandReserve(ctx, back));
        passEnd<<<gridFor(M), kB, 0, ctx->stream>>>(list, M, first, bs, ctx->area.p, cnt, dMoved);
        KT_HIP(hipGetLastError());
        // the regions' bins are cleared before the transfer, on the bet that
        // this pass is the last (a rerun paints them again)
        KT_TRY(unpaint());
        KT_HIP(hipMemcpyAsync(landVar(ctx), cnt, back, hipMemcpyDeviceToHost, ctx->stream));
        KT_HIP(hipStreamSynchronize(ctx->stream));
        const uint32_t *got = reinterpret_cast<const uint32_t *>(landVar(ctx));
        KT_CHECK(got[3] == M, KT_ESTATE, "lal: the batch's cell count differs from the cells per bin");
        const uint32_t bad = got[0];
        moves = got[2];
        if (bad == kNoVar) break;
        KT_TRY(paint(0));
        repairBatch<<<gridFor(M), kB, 0, ctx->stream>>>(list, M, bad, bs);
        KT_HIP(hipGetLastError());
        resegBatch<<<gridFor(M), kB, 0, ctx->stream>>>(list, M, bad, bs, cnt + 1);
        KT_HIP(hipGetLastError());
        uint32_t *slot = landSlots(ctx);
        KT_HIP(hipMemcpyAsync(slot + kLandActive, cnt + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
        KT_HIP(hipStreamSynchronize(ctx->stream));
        active = slot[kLandActive];
        first = bad;
        ++ctx->lalReruns;
        // counters back to {kNoVar, 0, 0} for the rerun ([3] stays)
        slot[kLandNone] = kNoVar;
        slot[kLandNone + 1] = slot[kLandNone + 2] = 0u;
        KT_HIP(hipMemcpyAsync(cnt, slot + kLandNone, 3 * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
    }
    // the batch's cells whose bin changed (binOf holds every batch cell's bin)
    moved.resize(moves);
    if (!inlineMoved && moves) {
        KT_TRY(lalLandReserve(ctx, headBytes + moves * sizeof(Moved)));
        KT_HIP(hipMemcpyAsync(landVar(ctx) + headBytes, dMoved, moves * sizeof(Moved), hipMemcpyDeviceToHost,
                              ctx->stream));
        KT_HIP(hipStreamSynchronize(ctx->stream));
    }
    if (moves) std::memcpy(moved.data(), landVar(ctx) + headBytes, moves * sizeof(Moved));
    return KT_OK;
}

// LookAheadLegalizer::legalize (simpl_lal.cc:43-124) on the context's positions.
int lalLegalize(kt_ctx *ctx) {
    KT_CHECK(ctx->haveNetlist && ctx->grid.size() > 0, KT_ESTATE, "lal: build the grid first");
    KT_CHECK(ctx->posCount == ctx->n, KT_ESTATE, "lal: positions must have one entry per movable cell");
    ctx->lalReruns = 0;
    const lal::Grid &g = ctx->grid;
    const uint32_t n = ctx->n;
    std::vector<double> occ;
    std::vector<uint32_t> cells;  // cells per bin, kept in step with the device's binOf
    KT_TRY(lalOccupancy(ctx, &occ, &cells));
    // the members of each bin, kept on the host like the occupancy (from the
    // cells that change bin), as doubly linked lists: a batch of the fused
    // kernel takes its cells from them
    const uint32_t fusedCells = std::min(ctx->lalFusedCells, kLalFusedCells);
    const bool members = fusedCells > 0;  // only the fused path reads them
    std::vector<uint32_t> binOfHost, head, next, prev;
    if (members) {
        binOfHost.resize(n);
        head.assign(g.size(), kNoVar);
        next.assign(n, kNoVar);
        prev.assign(n, kNoVar);
    }
    if (members && n) {
        KT_HIP(hipMemcpyAsync(binOfHost.data(), ctx->binOf.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
        KT_HIP(hipStreamSynchronize(ctx->stream));
    }
    const auto link = [&](uint32_t i, uint32_t k) {
        next[i] = head[k];
        prev[i] = kNoVar;
        if (head[k] != kNoVar) prev[head[k]] = i;
        head[k] = i;
    };
    const auto unlink = [&](uint32_t i, uint32_t k) {
        if (prev[i] != kNoVar) next[prev[i]] = next[i]; else head[k] = next[i];
        if (next[i] != kNoVar) prev[next[i]] = prev[i];
    };
    for (uint32_t i = members ? n : 0; i-- > 0;) {
        KT_CHECK(binOfHost[i] < g.size(), KT_ESTATE, "lal: a cell's bin is outside the grid");
        link(i, binOfHost[i]);
    }
    // occupancy is kept on the host from here on and updated with the cells
    // that change bin, as the reference does (simpl_lal.cc:86-93)
    std::vector<uint64_t> occFix(g.size());
    for (size_t k = 0; k < g.size(); ++k) occFix[k] = static_cast<uint64_t>(std::llround(occ[k] * kFixScale));
    lal::AreaTables &tables = ctx->lalTables;  // free-area table built with the grid
    tables.setOcc(occFix);
    KT_TRY(ctx->binRegion.reserve(g.size()));
    fillU32<<<gridFor(g.size()), kB, 0, ctx->stream>>>(ctx->binRegion.p, kNoVar, g.size());
    KT_HIP(hipGetLastError());
    std::vector<Moved> moved;
    std::vector<std::pair<uint32_t, uint32_t>> batchCells;
    std::vector<uint8_t> rowDirty(g.ny, 0);
    const auto reachOf = [&](const lal::BinRect &r) {
        return lal::BinRect{r.x0 > 0 ? r.x0 - 1 : 0, std::min(r.x1 + 1, g.nx - 1), r.y0 > 0 ? r.y0 - 1 : 0,
                            std::min(r.y1 + 1, g.ny - 1)};
    };
    const auto touches = [](const lal::BinRect &a, const lal::BinRect &b) {
        return a.x0 <= b.x1 && b.x0 <= a.x1 && a.y0 <= b.y1 && b.y0 <= a.y1;
    };
    std::vector<lal::BinRect> batch, reach;
    uint64_t batches = 0, fused = 0;
    const auto runBatch = [&]() -> int {
        if (batch.empty()) return KT_OK;
        bool ran = false;
        if (fusedCells > 0 && batch.size() <= kLalFusedRegions && ctx->lalOpts.maxLevel <= 64) {
            // the batch's cells from the bins' members, when few enough
            uint64_t count = 0;
            bool splits = false;
            for (const lal::BinRect &b : batch) {
                uint64_t inRegion = 0;
                for (size_t iy = b.y0; iy <= b.y1; ++iy)
                    for (size_t ix = b.x0; ix <= b.x1; ++ix) inRegion += cells[g.at(ix, iy)];
                count += inRegion;
                splits = splits || inRegion > ctx->lalOpts.minCellsToSplit;
            }
            if (count <= fusedCells && splits && ctx->lalOpts.maxLevel > 1) {
                batchCells.clear();
                for (uint32_t r = 0; r < batch.size(); ++r) {
                    const lal::BinRect &b = batch[r];
                    for (size_t iy = b.y0; iy <= b.y1; ++iy)
                        for (size_t ix = b.x0; ix <= b.x1; ++ix)
                            for (uint32_t i = head[g.at(ix, iy)]; i != kNoVar; i = next[i]) batchCells.push_back({i, r});
                }
                KT_CHECK(batchCells.size() == count, KT_ESTATE, "lal: bin members out of step with the cell counts");
                std::sort(batchCells.begin(), batchCells.end());
                KT_TRY(lalRunBatchFused(ctx, batch, batchCells, moved, &ran));
                fused += ran ? 1 : 0;
            }
        }
        if (!ran) KT_TRY(lalRunBatch(ctx, batch, cells, moved));
        ++batches;
        batch.clear();
        reach.clear();
        for (const Moved &mv : moved) {
            KT_CHECK(mv.cell < n && mv.oldBin < g.size() && mv.newBin < g.size(), KT_ESTATE,
                     "lal: a moved cell's record is out of range");
            --cells[mv.oldBin];
            ++cells[mv.newBin];
            occFix[mv.oldBin] -= mv.area;
            occFix[mv.newBin] += mv.area;
            occ[mv.oldBin] = static_cast<double>(occFix[mv.oldBin]) / kFixScale;
            occ[mv.newBin] = static_cast<double>(occFix[mv.newBin]) / kFixScale;
            rowDirty[mv.oldBin / g.nx] = 1;
            rowDirty[mv.newBin / g.nx] = 1;
            if (members) {
                unlink(mv.cell, mv.oldBin);
                link(mv.cell, mv.newBin);
            }
        }
        for (size_t iy = 0; iy < g.ny && !moved.empty(); ++iy)
            if (rowDirty[iy]) {
                tables.updateRow(occFix, iy);
                rowDirty[iy] = 0;
            }
        return KT_OK;
    };
    for (const lal::BinRect &cluster : lal::overfilledClusters(g, occ, ctx->densityLimit)) {
        lal::BinRect bins = lal::grow(g, tables, occ, cluster, ctx->densityLimit);
        bool clash = false;
        for (const lal::BinRect &r : reach) clash = clash || touches(r, reachOf(bins));
        if (clash) {
            KT_TRY(runBatch());
            bins = lal::grow(g, tables, occ, cluster, ctx->densityLimit);
        }
        batch.push_back(bins);
        reach.push_back(reachOf(bins));
    }
    KT_TRY(runBatch());
    if (n) {
        clampToDie<<<gridFor(n), kB, 0, ctx->stream>>>(ctx->posX.p, ctx->posY.p, ctx->width.p, ctx->height.p,
                                                       ctx->lalDie.x0, ctx->lalDie.y0, ctx->lalDie.x1, ctx->lalDie.y1, n);
        KT_HIP(hipGetLastError());
    }
    ctx->lalBatches = batches;
    ctx->lalFusedBatches = fused;
    return KT_OK;
}

}  // namespace kt
