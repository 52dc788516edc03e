// One batch of look-ahead legalization in one launch (simpl_lal.cc:102-124,
// 126-363): every level of the recursive bisection of every region of the
// batch, with the block geometry (free-area profile, cutline, stripes) on
// the device, the in-order rule for cells that land in a later region of the
// batch, and the list of cells whose bin changed.  One workgroup of four
// waves runs it; the host sends one packed record (the regions, the rows,
// spans and obstacles that meet them, the batch's cells) and reads back one
// small result, so a batch costs one launch and one round trip whatever the
// number of levels.
//
// Per level, as the multi-kernel path in lal.hip does:
//   1. the active cells sorted by (block, centre, id) with a bitonic sort in
//      LDS (the batch's cells are indexed in id order, so the local index
//      breaks ties as the id does, simpl_parallel.h:26-33);
//   2. block ranges, the fixed-point area prefix, the split at half the area;
//   3. per splitting block, one wave: the profile of the block, the cutline at
//      the same share of free area, the stripes of both halves in fill order
//      (lal_seq.h: the reference's sequential expressions, run lane-uniform;
//      the lanes collect and sort the steps and obstacle cuts);
//   4. halves ordered by distance from the cut, greedy stripe filling on the
//      prefix, per-stripe extents, the affine map into each stripe;
//   5. the halves that hold cells become the next level's blocks.
//
// A batch larger than the kernel's tables never reaches it: the host checks
// the record against the capacities below.  The kernel checks them again and,
// if one would overflow, puts every cell of the batch back where it started
// and reports it, and the host runs the batch on the multi-kernel path.
#include <algorithm>
#include <cstring>
#include <vector>

#include "kt_internal.h"
#include "lal_dev.h"
#include "lal_geom.h"
#include "lal_seq.h"

namespace kt {

namespace {

using namespace lalseq;

constexpr uint32_t kFCells = kLalFusedCells;
constexpr int kFThreads = 256;
constexpr int kFWaves = kFThreads / 64;
constexpr uint32_t kFRegions = kLalFusedRegions;
constexpr uint32_t kFRows = 384;      // rows meeting the regions (each region's own copy)
constexpr uint32_t kFSpans = 1536;    // free spans of those rows
constexpr uint32_t kFObs = 128;       // obstacles meeting the regions
constexpr uint32_t kFSteps = 1024;    // steps of one vertical profile (a power of two)
constexpr uint32_t kFEdges = kFSteps + 8;  // >= steps + 2 (vertical), 2 rows + 2 (horizontal)
constexpr uint32_t kLanes = 64;
constexpr uint32_t kFCuts = 128;      // lo, hi and obstacle edges of one half box (a power of two)
constexpr uint32_t kFStripes = 64;    // stripes of one half box
constexpr uint32_t kFStack = 32;      // depth of the stripe subdivision
constexpr uint32_t kNone = 0xffffffffu;

// status bits of a run
constexpr uint32_t kOverflow = 1u;

struct FusedLds {
    // regions of the batch
    uint32_t rect[kFRegions][4];  // bins x0, x1, y0, y1
    double box[kFRegions][4];     // x0, y0, x1, y1
    double room[kFRegions];       // free area of the region's bins
    uint32_t rrow[kFRegions][2];  // first row, rows
    uint32_t robs[kFRegions][2];  // first obstacle, obstacles
    // rows and spans meeting them (region by region, rows ascending)
    double rowY0[kFRows], rowY1[kFRows];
    uint32_t rowSpan[kFRows][2];  // first span, spans
    double spanA[kFSpans], spanB[kFSpans];
    uint32_t spanRow[kFSpans];
    double obs[kFObs][4];
    // cells by local index j (ascending cell id)
    uint32_t id[kFCells], seg[kFCells], half[kFCells];
    uint64_t key2[kFCells], P[kFCells];
    // sort
    uint32_t ss[kFCells], si[kFCells];
    uint64_t sk[kFCells];
    // one wave's staging for the register path
    struct Wave {
        double at[kLanes], dh[kLanes], cuts[kLanes];
    } wave[kFWaves];
    // the LDS path, for boxes beyond the register path: one wave at a time
    struct Big {
        double at[kFSteps], dh[kFSteps];
        double edges[kFEdges], prefix[kFEdges];
        double cuts[kFCuts], fin[kFStripes + 1], stack[kFStack], fromCut[kFStripes];
    } big;
    uint32_t bigBlocks[kFCells];  // blocks of this level left for the LDS path
    uint64_t waveTotals[kFWaves];
    uint32_t anySplit, more, status, bad, count, moves, nextBlocks, nBig;
};

struct FusedScratch {
    double *startX, *startY, *saveX, *saveY;     // [kFCells]
    uint32_t *startBin, *saveBin, *regionOf;     // [kFCells]
    double *blockBox[2];                         // [kFCells][4]
    double *blockRoom[2];                        // [kFCells]
    uint32_t *blockRegion[2];                    // [kFCells]
    LalSegStat *stats;                           // [kFCells]
    uint32_t *split;                             // [kFCells]
    double *cut;                                 // [kFCells]
    uint32_t *begin, *end;                       // [2 kFCells]
    uint32_t *nStripes;                          // [2 kFCells]
    double *stripe;                              // [2 kFCells][kFStripes][3] start, end, capacity
    uint32_t *stripeEnd;                         // [2 kFCells][kFStripes]
    unsigned long long *stripeLo, *stripeHi;     // [2 kFCells][kFStripes]
    double *halfBox;                             // [2 kFCells][4]
    uint32_t *stripeOf;                          // [kFCells]
    uint32_t *remap;                             // [2 kFCells]
};

struct FusedRecord {
    const uint8_t *base;
    uint32_t oRect, oBox, oRoom, oRRow, oRObs, oRowY0, oRowY1, oRowSpan, oSpanA, oSpanB, oSpanRow, oObs, oCell, oCReg;
    uint32_t M, R, rows, spans, obs;
};

struct FusedArgs {
    FusedRecord rec;
    GridDev g;
    double *x, *y;
    const double *w, *h, *area;
    uint32_t *binOf;
    uint32_t maxLevel, minCells;
    double frac, density;
    FusedScratch s;
    unsigned long long *out;  // [0] status, [1] moved, [2] reruns, [3] cells; then LalMoved records
};

template <typename T>
__device__ __forceinline__ const T *recArray(const FusedRecord &r, uint32_t off) {
    return reinterpret_cast<const T *>(r.base + off);
}

__device__ __forceinline__ void waveSync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t laneBelow(unsigned long long mask, uint32_t lane) {
    return static_cast<uint32_t>(__popcll(mask & ((1ull << lane) - 1ull)));
}

// Inclusive prefix sum in place of v[0, n) by the whole workgroup.
__device__ void scanU64(uint64_t *v, uint32_t n, FusedLds &L) {
    const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
    uint64_t carry = 0;
    for (uint32_t base = 0; base < n; base += 2 * kFThreads) {
        const uint32_t i0 = base + 2 * t, i1 = i0 + 1;
        const uint64_t a0 = i0 < n ? v[i0] : 0ull, a1 = i1 < n ? v[i1] : 0ull;
        const uint64_t local = a0 + a1;
        uint64_t x = local;
        for (int off = 1; off < 64; off <<= 1) {
            const uint64_t y = __shfl_up(x, off, 64);
            if (lane >= static_cast<uint32_t>(off)) x += y;
        }
        if (lane == 63) L.waveTotals[wave] = x;
        __syncthreads();
        uint64_t offset = carry, total = 0;
        for (int w = 0; w < kFWaves; ++w) {
            if (w < static_cast<int>(wave)) offset += L.waveTotals[w];
            total += L.waveTotals[w];
        }
        const uint64_t excl = offset + x - local;
        if (i0 < n) v[i0] = excl + a0;
        if (i1 < n) v[i1] = excl + a0 + a1;
        carry += total;
        __syncthreads();
    }
}

// Bitonic sort of (ss, sk, si) triples over width (a power of two), ascending.
__device__ void sortTriples(FusedLds &L, uint32_t width) {
    for (uint32_t k = 2; k <= width; k <<= 1) {
        for (uint32_t d = k >> 1; d > 0; d >>= 1) {
            for (uint32_t t = threadIdx.x; t < width; t += kFThreads) {
                const uint32_t u = t ^ d;
                if (u > t) {
                    const bool greater = L.ss[t] != L.ss[u]   ? L.ss[t] > L.ss[u]
                                         : L.sk[t] != L.sk[u] ? L.sk[t] > L.sk[u]
                                                              : L.si[t] > L.si[u];
                    if (greater == ((t & k) == 0)) {
                        const uint32_t s0 = L.ss[t], i0 = L.si[t];
                        const uint64_t k0 = L.sk[t];
                        L.ss[t] = L.ss[u];
                        L.sk[t] = L.sk[u];
                        L.si[t] = L.si[u];
                        L.ss[u] = s0;
                        L.sk[u] = k0;
                        L.si[u] = i0;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// Bitonic sort of width (a power of two) entries of one wave's arrays;
// pairs (a, b) lexicographically when b is given.
__device__ void waveSortPairs(double *a, double *b, uint32_t width, uint32_t lane) {
    for (uint32_t k = 2; k <= width; k <<= 1) {
        for (uint32_t d = k >> 1; d > 0; d >>= 1) {
            for (uint32_t t = lane; t < width; t += 64) {
                const uint32_t u = t ^ d;
                if (u > t) {
                    const bool greater = a[t] > a[u] || (b != nullptr && a[t] == a[u] && b[t] > b[u]);
                    if (greater == ((t & k) == 0)) {
                        const double a0 = a[t];
                        a[t] = a[u];
                        a[u] = a0;
                        if (b != nullptr) {
                            const double b0 = b[t];
                            b[t] = b[u];
                            b[u] = b0;
                        }
                    }
                }
            }
            waveSync();
        }
    }
}

__device__ __forceinline__ uint32_t pow2AtLeast(uint32_t v) {
    uint32_t w = 2;
    while (w < v) w <<= 1;
    return w;
}

// FreeSpace::getProfile (simpl_grid.h:179-231) of box b by one wave, over the
// rows of region `reg`.  The rows are the region's, ascending; a sub-box of
// the region meets no other row.  Lane-uniform result in the wave's arrays.
__device__ __forceinline__ FlatProfile waveProfile(FusedLds &L, FusedLds::Big &W, const double b[4], bool vertical, uint32_t reg,
                                   uint32_t lane) {
    const uint32_t r0 = L.rrow[reg][0], r1 = r0 + L.rrow[reg][1];
    uint32_t lo = r0, hi = r1;  // upper_bound of b.y0 on the rows' top edges
    while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2;
        if (b[1] < L.rowY1[mid]) hi = mid; else lo = mid + 1;
    }
    const uint32_t first = lo;
    hi = r1;  // first row from there that starts at or above the box's top
    while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2;
        if (L.rowY0[mid] < b[3]) lo = mid + 1; else hi = mid;
    }
    const uint32_t last = lo;
    const double axisLo = vertical ? b[0] : b[1];
    const double axisHi = vertical ? b[2] : b[3];
    FlatProfile p{W.edges, W.prefix, 0, kFEdges, true};
    profileStart(p, axisLo);
    if (vertical) {
        uint32_t s0 = 0, s1 = 0;
        if (first < last) {
            s0 = L.rowSpan[first][0];
            s1 = L.rowSpan[last - 1][0] + L.rowSpan[last - 1][1];
        }
        uint32_t n = 0;
        for (uint32_t base = s0; base < s1; base += 64) {
            const uint32_t t = base + lane;
            bool emit = false;
            double a = 0.0, z = 0.0, hgt = 0.0;
            if (t < s1) {
                const uint32_t rr = L.spanRow[t];
                hgt = smin(b[3], L.rowY1[rr]) - smax(b[1], L.rowY0[rr]);
                a = smax(L.spanA[t], axisLo);
                z = smin(L.spanB[t], axisHi);
                emit = hgt > 0.0 && z > a;
            }
            const unsigned long long mask = __ballot(emit);
            if (emit) {
                const uint32_t k = n + 2 * laneBelow(mask, lane);
                if (k + 1 < kFSteps) {
                    W.at[k] = a;
                    W.dh[k] = hgt;
                    W.at[k + 1] = z;
                    W.dh[k + 1] = -hgt;
                }
            }
            n += 2 * static_cast<uint32_t>(__popcll(mask));
        }
        if (n > kFSteps) {
            p.ok = false;
            return p;
        }
        const uint32_t width = pow2AtLeast(n);
        for (uint32_t t = n + lane; t < width; t += 64) {
            W.at[t] = __builtin_inf();
            W.dh[t] = __builtin_inf();
        }
        waveSync();
        waveSortPairs(W.at, W.dh, width, lane);
        profileSweepSteps(p, W.at, W.dh, n);
    } else {
        for (uint32_t base = first; base < last; base += 64) {
            const uint32_t r = base + lane;
            int valid = 0;
            double a = 0.0, z = 0.0, len = 0.0;
            if (r < last) {
                a = smax(axisLo, L.rowY0[r]);
                z = smin(axisHi, L.rowY1[r]);
                valid = z > a ? 1 : 0;
                if (valid) {
                    const uint32_t sp0 = L.rowSpan[r][0], sp1 = sp0 + L.rowSpan[r][1];
                    for (uint32_t sp = sp0; sp < sp1; ++sp) len += smax(0.0, smin(b[2], L.spanB[sp]) - smax(b[0], L.spanA[sp]));
                }
            }
            const uint32_t chunk = last - base < 64 ? last - base : 64;
            for (uint32_t q = 0; q < chunk; ++q) {
                if (__shfl(valid, static_cast<int>(q)) == 0) continue;
                const double aq = __shfl(a, static_cast<int>(q));
                const double zq = __shfl(z, static_cast<int>(q));
                const double lq = __shfl(len, static_cast<int>(q));
                profileExtend(p, aq, 0.0);
                profileExtend(p, zq, lq * (zq - aq));
            }
        }
    }
    profileFinish(p, axisHi);
    waveSync();
    return p;
}

// getStripes (simpl_lal.cc:238-263) of one half box and their fill order
// (:275-282), written to the half's slots.  Returns the number of stripes,
// or -1 when a table would overflow.
__device__ __forceinline__ int waveStripes(FusedLds &L, FusedLds::Big &W, const FusedArgs &a, const double box[4], bool vertical,
                           double cellArea, double regionRoom, double cutline, bool belowCut, uint32_t reg,
                           uint32_t slot, uint32_t lane) {
    const FlatProfile room = waveProfile(L, W, box, vertical, reg, lane);
    if (!room.ok) return -1;
    const double lo = vertical ? box[0] : box[1];
    const double hi = vertical ? box[2] : box[3];
    W.cuts[0] = lo;
    W.cuts[1] = hi;
    uint32_t nc = 2;
    const uint32_t o0 = L.robs[reg][0], o1 = o0 + L.robs[reg][1];
    for (uint32_t base = o0; base < o1; base += 64) {
        const uint32_t t = base + lane;
        bool e0ok = false, e1ok = false;
        double e0 = 0.0, e1 = 0.0;
        if (t < o1) {
            const double *o = L.obs[t];
            if (!(o[2] <= box[0] || o[0] >= box[2] || o[3] <= box[1] || o[1] >= box[3])) {
                e0 = vertical ? o[0] : o[1];
                e1 = vertical ? o[2] : o[3];
                e0ok = e0 > lo && e0 < hi;
                e1ok = e1 > lo && e1 < hi;
            }
        }
        const unsigned long long m0 = __ballot(e0ok), m1 = __ballot(e1ok);
        const uint32_t k0 = nc + laneBelow(m0, lane);
        const uint32_t k1 = nc + static_cast<uint32_t>(__popcll(m0)) + laneBelow(m1, lane);
        if (e0ok && k0 < kFCuts) W.cuts[k0] = e0;
        if (e1ok && k1 < kFCuts) W.cuts[k1] = e1;
        nc += static_cast<uint32_t>(__popcll(m0) + __popcll(m1));
    }
    if (nc > kFCuts) return -1;
    const uint32_t width = pow2AtLeast(nc);
    for (uint32_t t = nc + lane; t < width; t += 64) W.cuts[t] = __builtin_inf();
    waveSync();
    waveSortPairs(W.cuts, nullptr, width, lane);
    uint32_t nu = 1;  // std::unique
    for (uint32_t t = 1; t < nc; ++t) {
        const double c = W.cuts[t];
        if (!(W.cuts[nu - 1] == c)) W.cuts[nu++] = c;
    }
    waveSync();
    const double total = profileTotal(room);
    const double limit = a.frac * (regionRoom > 0.0 ? regionRoom : total);
    const double fill = total > 0.0 ? smin(a.density, cellArea / total) : a.density;
    const uint32_t nf = stripeCuts(room, W.cuts, nu, limit, W.fin, kFStripes + 1, W.stack, kFStack);
    if (nf == 0) return -1;
    waveSync();
    const uint32_t ns = nf - 1;
    for (uint32_t s = lane; s < ns; s += 64) W.fromCut[s] = stripeFromCut(W.fin[s], W.fin[s + 1], cutline, belowCut);
    waveSync();
    for (uint32_t s = lane; s < ns; s += 64) {
        const uint32_t q = slot * kFStripes + stripeRank(W.fromCut, ns, s);
        a.s.stripe[3 * q] = W.fin[s];
        a.s.stripe[3 * q + 1] = W.fin[s + 1];
        a.s.stripe[3 * q + 2] = fill * profileBetween(room, W.fin[s], W.fin[s + 1]);
    }
    return static_cast<int>(ns);
}

// ------------------------------------------------------------ register path
// A profile of at most 64 edges held across the wave: lane k holds edge k and
// the free area before it, and the last edge and prefix are kept uniform.
// Queries walk the same binary searches as lal_seq.h with readlane instead of
// LDS loads, the steps, cuts and stripe lists live in lanes too, and the
// arithmetic is lal_seq.h's expression for expression.  A box that needs more
// than 64 steps, edges, cuts or stripes goes to the LDS path above.
__device__ __forceinline__ double laneValue(double v, uint32_t l) {
    const uint64_t b = static_cast<uint64_t>(__double_as_longlong(v));
    const uint32_t lo = static_cast<uint32_t>(
        __builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(b)), static_cast<int>(l)));
    const uint32_t hi = static_cast<uint32_t>(
        __builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(b >> 32)), static_cast<int>(l)));
    return __longlong_as_double(static_cast<long long>((static_cast<uint64_t>(hi) << 32) | lo));
}

__device__ __forceinline__ uint32_t laneWord(uint32_t v, uint32_t l) {
    return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), static_cast<int>(l)));
}

struct RegProfile {
    double e, p;         // this lane's edge and the free area before it (lanes < n)
    double back, backP;  // the last edge and its prefix (uniform): the total
    uint32_t n;
    bool ok;
};

__device__ __forceinline__ void rpStart(RegProfile &r, double lo) {
    r.e = lo;
    r.p = 0.0;
    r.back = lo;
    r.backP = 0.0;
    r.n = 1;
    r.ok = true;
}

// profileExtend
__device__ __forceinline__ void rpExtend(RegProfile &r, double to, double area, uint32_t lane) {
    if (to > r.back) {
        if (r.n >= kLanes) {
            r.ok = false;
            return;
        }
        r.backP = r.backP + area;
        if (lane == r.n) {
            r.e = to;
            r.p = r.backP;
        }
        r.back = to;
        ++r.n;
    }
}

// profileFinish
__device__ __forceinline__ void rpFinish(RegProfile &r, double hi, uint32_t lane) {
    rpExtend(r, hi, 0.0, lane);
    if (r.ok && r.n == 1) {
        if (lane == 1) {
            r.e = hi;
            r.p = 0.0;
        }
        r.back = hi;
        r.backP = 0.0;
        r.n = 2;
    }
}

// profileBefore
__device__ __forceinline__ double rpBefore(const RegProfile &r, double c) {
    if (c <= laneValue(r.e, 0)) return 0.0;
    if (c >= r.back) return r.backP;
    uint32_t lo = 0, hi = r.n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2;
        if (c < laneValue(r.e, mid)) hi = mid; else lo = mid + 1;
    }
    const uint32_t k = lo - 1;
    const double ek = laneValue(r.e, k), ek1 = laneValue(r.e, k + 1);
    const double pk = laneValue(r.p, k), pk1 = laneValue(r.p, k + 1);
    const double f = (c - ek) / (ek1 - ek);
    return pk + f * (pk1 - pk);
}

// profileCoordinateAt
__device__ __forceinline__ double rpCoordinateAt(const RegProfile &r, double target) {
    if (target <= 0.0) return laneValue(r.e, 0);
    if (target >= r.backP) return r.back;
    uint32_t lo = 0, hi = r.n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2;
        if (laneValue(r.p, mid) < target) lo = mid + 1; else hi = mid;
    }
    const uint32_t k = lo;
    const double pk = laneValue(r.p, k), pk0 = laneValue(r.p, k - 1);
    const double ek = laneValue(r.e, k), ek0 = laneValue(r.e, k - 1);
    const double step = pk - pk0;
    const double f = step > 0.0 ? (target - pk0) / step : 1.0;
    return ek0 + f * (ek - ek0);
}

// Bitonic sort across the first `width` lanes (a power of two <= 64),
// ascending: of each compared pair, the lane below keeps the smaller element
// where the run ascends and the larger where it descends.
__device__ __forceinline__ void regSortPairs(double &a, double &b, uint32_t width, uint32_t lane) {
    for (uint32_t k = 2; k <= width; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            const double pa = __shfl_xor(a, static_cast<int>(j)), pb = __shfl_xor(b, static_cast<int>(j));
            const bool pLess = pa < a || (pa == a && pb < b);
            const bool meLess = a < pa || (a == pa && b < pb);
            const bool lower = (lane & j) == 0, up = (lane & k) == 0;
            if (lower == up ? pLess : meLess) {
                a = pa;
                b = pb;
            }
        }
}

__device__ __forceinline__ void regSortTriples(uint32_t &s, uint64_t &k, uint32_t &i, uint32_t width, uint32_t lane) {
    for (uint32_t kk = 2; kk <= width; kk <<= 1)
        for (uint32_t j = kk >> 1; j > 0; j >>= 1) {
            const uint32_t ps = __shfl_xor(s, static_cast<int>(j)), pi = __shfl_xor(i, static_cast<int>(j));
            const uint64_t pk = __shfl_xor(k, static_cast<int>(j));
            const bool pLess = ps != s ? ps < s : pk != k ? pk < k : pi < i;
            const bool meLess = s != ps ? s < ps : k != pk ? k < pk : i < pi;
            const bool lower = (lane & j) == 0, up = (lane & kk) == 0;
            if (lower == up ? pLess : meLess) {
                s = ps;
                k = pk;
                i = pi;
            }
        }
}

// The steps of a vertical profile of box b over rows [first, last) into the
// wave's arrays (as many as fit); returns their number.
__device__ uint32_t collectSteps(const FusedLds &L, FusedLds::Wave &W, const double b[4], uint32_t first, uint32_t last,
                                 double axisLo, double axisHi, uint32_t lane) {
    uint32_t s0 = 0, s1 = 0;
    if (first < last) {
        s0 = L.rowSpan[first][0];
        s1 = L.rowSpan[last - 1][0] + L.rowSpan[last - 1][1];
    }
    uint32_t n = 0;
    for (uint32_t base = s0; base < s1; base += 64) {
        const uint32_t t = base + lane;
        bool emit = false;
        double a = 0.0, z = 0.0, hgt = 0.0;
        if (t < s1) {
            const uint32_t rr = L.spanRow[t];
            hgt = smin(b[3], L.rowY1[rr]) - smax(b[1], L.rowY0[rr]);
            a = smax(L.spanA[t], axisLo);
            z = smin(L.spanB[t], axisHi);
            emit = hgt > 0.0 && z > a;
        }
        const unsigned long long mask = __ballot(emit);
        if (emit) {
            const uint32_t k = n + 2 * laneBelow(mask, lane);
            if (k + 1 < kLanes) {
                W.at[k] = a;
                W.dh[k] = hgt;
                W.at[k + 1] = z;
                W.dh[k + 1] = -hgt;
            }
        }
        n += 2 * static_cast<uint32_t>(__popcll(mask));
    }
    return n;
}

// Rows of region reg meeting box b: [first, last).
__device__ __forceinline__ void rowRange(const FusedLds &L, const double b[4], uint32_t reg, uint32_t &first,
                                         uint32_t &last) {
    const uint32_t r0 = L.rrow[reg][0], r1 = r0 + L.rrow[reg][1];
    uint32_t lo = r0, hi = r1;  // upper_bound of b.y0 on the rows' top edges
    while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2;
        if (b[1] < L.rowY1[mid]) hi = mid; else lo = mid + 1;
    }
    first = lo;
    hi = r1;  // first row from there that starts at or above the box's top
    while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2;
        if (L.rowY0[mid] < b[3]) lo = mid + 1; else hi = mid;
    }
    last = lo;
}

// waveProfile with the profile in registers; r.ok is false when the box needs
// the LDS path.
__device__ __forceinline__ RegProfile regProfile(const FusedLds &L, FusedLds::Wave &W, const double b[4], bool vertical, uint32_t reg,
                                 uint32_t lane) {
    uint32_t first, last;
    rowRange(L, b, reg, first, last);
    const double axisLo = vertical ? b[0] : b[1];
    const double axisHi = vertical ? b[2] : b[3];
    RegProfile r;
    rpStart(r, axisLo);
    if (vertical) {
        const uint32_t n = collectSteps(L, W, b, first, last, axisLo, axisHi, lane);
        if (n > kLanes) {
            r.ok = false;
            return r;
        }
        waveSync();
        double at = lane < n ? W.at[lane] : __builtin_inf();
        double dh = lane < n ? W.dh[lane] : __builtin_inf();
        waveSync();
        regSortPairs(at, dh, pow2AtLeast(n), lane);
        double height = 0.0;
        for (uint32_t s = 0; s < n; ++s) {
            const double as = laneValue(at, s), ds = laneValue(dh, s);
            rpExtend(r, as, height * (as - r.back), lane);
            if (!r.ok) return r;
            height += ds;
        }
    } else {
        if (last - first > kLanes) {
            r.ok = false;
            return r;
        }
        const uint32_t row = first + lane;
        uint32_t valid = 0;
        double a = 0.0, z = 0.0, len = 0.0;
        if (row < last) {
            a = smax(axisLo, L.rowY0[row]);
            z = smin(axisHi, L.rowY1[row]);
            valid = z > a ? 1u : 0u;
            if (valid) {
                const uint32_t sp0 = L.rowSpan[row][0], sp1 = sp0 + L.rowSpan[row][1];
                for (uint32_t sp = sp0; sp < sp1; ++sp) len += smax(0.0, smin(b[2], L.spanB[sp]) - smax(b[0], L.spanA[sp]));
            }
        }
        for (uint32_t q = 0; q < last - first; ++q) {
            if (laneWord(valid, q) == 0) continue;
            const double aq = laneValue(a, q), zq = laneValue(z, q), lq = laneValue(len, q);
            rpExtend(r, aq, 0.0, lane);
            rpExtend(r, zq, lq * (zq - aq), lane);
            if (!r.ok) return r;
        }
    }
    rpFinish(r, axisHi, lane);
    return r;
}

// waveStripes with the profile, cuts and stripes in registers.  Returns the
// number of stripes, or -2 when the half box needs the LDS path.
__device__ __forceinline__ int regStripes(const FusedLds &L, FusedLds::Wave &W, const FusedArgs &a, const double box[4], bool vertical,
                          double cellArea, double regionRoom, double cutline, bool belowCut, uint32_t reg,
                          uint32_t slot, uint32_t lane) {
    const RegProfile room = regProfile(L, W, box, vertical, reg, lane);
    if (!room.ok) return -2;
    const double lo = vertical ? box[0] : box[1];
    const double hi = vertical ? box[2] : box[3];
    // cuts: the box's ends and the obstacle edges strictly inside it
    waveSync();
    if (lane < 2) W.cuts[lane] = lane == 0 ? lo : hi;
    uint32_t nc = 2;
    const uint32_t o0 = L.robs[reg][0], o1 = o0 + L.robs[reg][1];
    for (uint32_t base = o0; base < o1; base += 64) {
        const uint32_t t = base + lane;
        bool e0ok = false, e1ok = false;
        double e0 = 0.0, e1 = 0.0;
        if (t < o1) {
            const double *o = L.obs[t];
            if (!(o[2] <= box[0] || o[0] >= box[2] || o[3] <= box[1] || o[1] >= box[3])) {
                e0 = vertical ? o[0] : o[1];
                e1 = vertical ? o[2] : o[3];
                e0ok = e0 > lo && e0 < hi;
                e1ok = e1 > lo && e1 < hi;
            }
        }
        const unsigned long long m0 = __ballot(e0ok), m1 = __ballot(e1ok);
        const uint32_t k0 = nc + laneBelow(m0, lane);
        const uint32_t k1 = nc + static_cast<uint32_t>(__popcll(m0)) + laneBelow(m1, lane);
        if (e0ok && k0 < kLanes) W.cuts[k0] = e0;
        if (e1ok && k1 < kLanes) W.cuts[k1] = e1;
        nc += static_cast<uint32_t>(__popcll(m0) + __popcll(m1));
    }
    if (nc > kLanes) return -2;
    waveSync();
    double c = lane < nc ? W.cuts[lane] : __builtin_inf();
    double unused = 0.0;
    regSortPairs(c, unused, pow2AtLeast(nc), lane);
    // std::unique: in sorted order a value equal to its predecessor repeats
    // the last one kept
    const double prev = __shfl_up(c, 1);
    const bool keep = lane < nc && (lane == 0 || !(prev == c));
    const unsigned long long km = __ballot(keep);
    const uint32_t nu = static_cast<uint32_t>(__popcll(km));
    waveSync();
    if (keep) W.cuts[laneBelow(km, lane)] = c;
    waveSync();
    const double cu = lane < nu ? W.cuts[lane] : 0.0;
    const double total = room.backP;
    const double limit = a.frac * (regionRoom > 0.0 ? regionRoom : total);
    const double fill = total > 0.0 ? smin(a.density, cellArea / total) : a.density;
    // stripeCuts, depth first: the pending right ends and the final cuts in lanes
    double fin = 0.0, stack = 0.0;
    double cur = laneValue(cu, 0);
    if (lane == 0) fin = cur;
    uint32_t nf = 1;
    for (uint32_t i = 1; i < nu; ++i) {
        double cc = laneValue(cu, i);
        uint32_t sp = 0;
        for (;;) {
            const double beforeCur = rpBefore(room, cur);
            const double here = rpBefore(room, cc) - beforeCur;
            const double mid = rpCoordinateAt(room, beforeCur + 0.5 * here);
            if (here > limit && mid > cur && mid < cc) {
                if (sp >= kLanes) return -2;
                if (lane == sp) stack = cc;
                ++sp;
                cc = mid;
            } else {
                if (nf >= kLanes) return -2;
                if (lane == nf) fin = cc;
                ++nf;
                cur = cc;
                if (sp == 0) break;
                --sp;
                cc = laneValue(stack, sp);
            }
        }
    }
    const uint32_t ns = nf - 1;
    double bf = 0.0;  // free area before each final cut
    for (uint32_t q = 0; q < nf; ++q) {
        const double v = rpBefore(room, laneValue(fin, q));
        if (lane == q) bf = v;
    }
    const double finNext = __shfl_down(fin, 1), bfNext = __shfl_down(bf, 1);
    const double fc = stripeFromCut(fin, finNext, cutline, belowCut);
    uint32_t rank = 0;
    for (uint32_t j = 0; j < ns; ++j) {
        const double fj = laneValue(fc, j);
        rank += (fj > fc || (j < lane && fj == fc)) ? 1u : 0u;
    }
    if (lane < ns) {
        const uint32_t q = slot * kFStripes + rank;
        a.s.stripe[3 * q] = fin;
        a.s.stripe[3 * q + 1] = finNext;
        a.s.stripe[3 * q + 2] = fill * (bfNext - bf);
    }
    return static_cast<int>(ns);
}

__device__ __forceinline__ double centreOn(const FusedArgs &a, uint32_t i, bool vertical) {
    return vertical ? a.x[i] + 0.5 * a.w[i] : a.y[i] + 0.5 * a.h[i];
}

// Region of the batch whose bins hold bin k (kNone: none).
__device__ __forceinline__ uint32_t regionOfBin(const FusedLds &L, uint32_t R, uint32_t k, uint32_t nx) {
    const uint32_t ix = k % nx, iy = k / nx;
    for (uint32_t r = 0; r < R; ++r)
        if (ix >= L.rect[r][0] && ix <= L.rect[r][1] && iy >= L.rect[r][2] && iy <= L.rect[r][3]) return r;
    return kNone;
}

// Split of each block at half its cell area (simpl_lal.cc:225-236), from the
// inclusive fixed-point prefix P of the sorted areas.
__device__ LalSegStat splitStat(uint32_t b, uint32_t e, const uint64_t *P) {
    LalSegStat st{};
    st.count = e - b;
    if (st.count > 0) {
        const uint64_t before = b > 0 ? P[b - 1] : 0ull;
        st.cellArea = P[e - 1] - before;
        const double half = 0.5 * (static_cast<double>(st.cellArea) / kFixScale);
        uint32_t lo = 1, hi = st.count > 1 ? st.count - 1 : 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) / 2;
            const double low = static_cast<double>(P[b + mid - 1] - before) / kFixScale;
            if (low < half) lo = mid + 1; else hi = mid;
        }
        st.m = st.count > 1 ? lo : 0;
        st.lowArea = st.m > 0 ? P[b + st.m - 1] - before : 0ull;
    }
    return st;
}

// The level's order: one wave sorts in registers up to 64 entries, the whole
// workgroup through LDS beyond that.
__device__ void sortLevel(FusedLds &L, uint32_t width) {
    if (width > kLanes) {
        sortTriples(L, width);
        return;
    }
    const uint32_t lane = threadIdx.x & 63;
    if (threadIdx.x < 64) {
        uint32_t sg = kNone, id = kNone;
        uint64_t key = ~0ull;
        if (lane < width) {
            sg = L.ss[lane];
            key = L.sk[lane];
            id = L.si[lane];
        }
        regSortTriples(sg, key, id, width, lane);
        if (lane < width) {
            L.ss[lane] = sg;
            L.sk[lane] = key;
            L.si[lane] = id;
        }
    }
    __syncthreads();
}

// A half box whose stripes came out empty (a zero-length half): one stripe at
// the box edge, so its cells stay inside the box.
__device__ __forceinline__ void edgeStripe(const FusedScratch &s, uint32_t half, const double box[4], bool vertical) {
    const uint32_t q = half * kFStripes;
    s.stripe[3 * q] = vertical ? box[0] : box[1];
    s.stripe[3 * q + 1] = vertical ? box[2] : box[3];
    s.stripe[3 * q + 2] = 0.0;
}

__device__ __forceinline__ void queueBig(FusedLds &L, uint32_t b, uint32_t lane) {
    if (lane == 0) L.bigBlocks[atomicAdd(&L.nBig, 1u)] = b;
}

// The geometry of block b of this level (simpl_lal.cc:213-263) by one wave:
// the profile of the block, the cutline at the low half's share of the free
// area, and the stripes of both halves.  With big = false the register path
// only: a block it cannot hold is queued for the LDS path and left untouched
// (its stripe slots may hold partial output, which the LDS pass rewrites).
__device__ __forceinline__ void blockGeometry(FusedLds &L, const FusedArgs &a, uint32_t b, uint32_t par, uint32_t level,
                              bool vertical, uint32_t lane, uint32_t wave, bool big) {
    const FusedScratch &s = a.s;
    const LalSegStat st = s.stats[b];
    if (level >= a.maxLevel || st.count <= a.minCells) {
        if (lane == 0) {
            s.split[b] = 0;
            s.nStripes[2 * b] = 0;
            s.nStripes[2 * b + 1] = 0;
        }
        return;
    }
    double box[4];
    for (int k = 0; k < 4; ++k) box[k] = s.blockBox[par][4 * b + k];
    const double regionRoom = s.blockRoom[par][b];
    const uint32_t reg = s.blockRegion[par][b];
    const double cellArea = static_cast<double>(st.cellArea) / kFixScale;
    const double lowArea = static_cast<double>(st.lowArea) / kFixScale;
    const double highArea = static_cast<double>(st.cellArea - st.lowArea) / kFixScale;
    const double lo = vertical ? box[0] : box[1];
    const double hi = vertical ? box[2] : box[3];
    const double share = cellArea > 0.0 ? lowArea / cellArea : 0.5;
    double cutline;
    if (!big) {
        const RegProfile room = regProfile(L, L.wave[wave], box, vertical, reg, lane);
        if (!room.ok) return queueBig(L, b, lane);
        const double total = room.backP;
        cutline = total > 0.0 ? rpCoordinateAt(room, share * total) : lo + share * (hi - lo);
    } else {
        const FlatProfile room = waveProfile(L, L.big, box, vertical, reg, lane);
        if (!room.ok) {
            if (lane == 0) atomicOr(&L.status, kOverflow);
            return;
        }
        const double total = profileTotal(room);
        cutline = total > 0.0 ? profileCoordinateAt(room, share * total) : lo + share * (hi - lo);
    }
    double lowBox[4], highBox[4];
    for (int k = 0; k < 4; ++k) lowBox[k] = highBox[k] = box[k];
    (vertical ? lowBox[2] : lowBox[3]) = cutline;
    (vertical ? highBox[0] : highBox[1]) = cutline;
    waveSync();  // the block's profile is read no more: the halves reuse its arrays
    int nLow = 0, nHigh = 0;
    if (st.m > 0) {
        nLow = big ? waveStripes(L, L.big, a, lowBox, vertical, lowArea, regionRoom, cutline, true, reg, 2 * b, lane)
                   : regStripes(L, L.wave[wave], a, lowBox, vertical, lowArea, regionRoom, cutline, true, reg, 2 * b,
                                lane);
        if (nLow == -2) return queueBig(L, b, lane);
        if (nLow == 0 && lane == 0) edgeStripe(s, 2 * b, lowBox, vertical);
        if (nLow == 0) nLow = 1;
        waveSync();
    }
    if (st.count > st.m && nLow >= 0) {
        nHigh = big ? waveStripes(L, L.big, a, highBox, vertical, highArea, regionRoom, cutline, false, reg, 2 * b + 1,
                                  lane)
                    : regStripes(L, L.wave[wave], a, highBox, vertical, highArea, regionRoom, cutline, false, reg,
                                 2 * b + 1, lane);
        if (nHigh == -2) return queueBig(L, b, lane);
        if (nHigh == 0 && lane == 0) edgeStripe(s, 2 * b + 1, highBox, vertical);
        if (nHigh == 0) nHigh = 1;
    }
    if (nLow < 0 || nHigh < 0) {
        if (lane == 0) atomicOr(&L.status, kOverflow);
        return;
    }
    if (lane == 0) {
        s.split[b] = 1;
        s.cut[b] = cutline;
        s.nStripes[2 * b] = static_cast<uint32_t>(nLow);
        s.nStripes[2 * b + 1] = static_cast<uint32_t>(nHigh);
        for (int k = 0; k < 4; ++k) {
            s.halfBox[4 * (2 * b) + k] = lowBox[k];
            s.halfBox[4 * (2 * b + 1) + k] = highBox[k];
        }
        L.anySplit = 1;
        if (st.m > a.minCells || st.count - st.m > a.minCells) L.more = 1;
    }
}

// The levels of the recursion for the cells whose seg names a region (all
// blocks of a level at once).  Returns false when a table overflowed.
__device__ bool runLevels(FusedLds &L, const FusedArgs &a, uint32_t A) {
    const FusedScratch &s = a.s;
    const uint32_t M = a.rec.M, R = a.rec.R;
    const uint32_t t0 = threadIdx.x, lane = t0 & 63, wave = t0 >> 6;
    const uint32_t width = pow2AtLeast(M);
    for (uint32_t b = t0; b < R; b += kFThreads) {
        for (int k = 0; k < 4; ++k) s.blockBox[0][4 * b + k] = L.box[b][k];
        s.blockRoom[0][b] = L.room[b];
        s.blockRegion[0][b] = b;
    }
    uint32_t B = R, par = 0;
    bool vertical = true;
    __syncthreads();
    for (uint32_t level = 1; A > 0; ++level) {
        // 1. order of the level: (block, centre, id)
        for (uint32_t t = t0; t < width; t += kFThreads) {
            if (t < M && L.seg[t] != kNone) {
                L.ss[t] = L.seg[t];
                L.sk[t] = orderKey(centreOn(a, L.id[t], vertical));
                L.si[t] = t;
            } else {
                L.ss[t] = kNone;
                L.sk[t] = ~0ull;
                L.si[t] = kNone;
            }
        }
        for (uint32_t q = t0; q < B; q += kFThreads) {
            s.begin[q] = 0;
            s.end[q] = 0;
        }
        if (t0 == 0) {
            L.anySplit = 0;
            L.more = 0;
            L.count = 0;
            L.nBig = 0;
        }
        __syncthreads();
        sortLevel(L, width);
        // 2. block ranges, area prefix, split
        for (uint32_t k = t0; k < A; k += kFThreads) {
            const uint32_t sg = L.ss[k];
            if (k == 0 || L.ss[k - 1] != sg) s.begin[sg] = k;
            if (k + 1 == A || L.ss[k + 1] != sg) s.end[sg] = k + 1;
            L.P[k] = areaFixOf(a.area, L.id[L.si[k]]);
        }
        __syncthreads();
        scanU64(L.P, A, L);
        for (uint32_t b = t0; b < B; b += kFThreads) s.stats[b] = splitStat(s.begin[b], s.end[b], L.P);
        __syncthreads();
        // 3. geometry, one wave per block; boxes beyond the register path
        // then one at a time through the LDS path
        for (uint32_t b = wave; b < B; b += kFWaves) blockGeometry(L, a, b, par, level, vertical, lane, wave, false);
        __syncthreads();
        if (L.nBig > 0) {
            if (wave == 0)
                for (uint32_t q = 0; q < L.nBig; ++q)
                    blockGeometry(L, a, L.bigBlocks[q], par, level, vertical, lane, wave, true);
            __syncthreads();
        }
        if (L.status != 0) return false;
        if (L.anySplit == 0) break;
        // 4. halves: low = the first m cells of the block; key = -fromCut
        for (uint32_t j = t0; j < M; j += kFThreads) L.half[j] = kNone;
        __syncthreads();
        for (uint32_t k = t0; k < A; k += kFThreads) {
            const uint32_t j = L.si[k], b = L.ss[k];
            if (!s.split[b]) continue;
            const LalSegStat st = s.stats[b];
            const bool low = (k - s.begin[b]) < st.m;
            const double c = centreOn(a, L.id[j], vertical);
            const double cut = s.cut[b];
            const double fromCut = low ? cut - c : c - cut;
            L.half[j] = 2 * b + (low ? 0u : 1u);
            L.key2[j] = orderKey(-fromCut);
            atomicAdd(&L.count, 1u);
        }
        __syncthreads();
        const uint32_t A2 = L.count;
        const uint32_t H = 2 * B;
        for (uint32_t t = t0; t < width; t += kFThreads) {
            const uint32_t hh = t < M ? L.half[t] : kNone;
            L.ss[t] = hh;
            L.sk[t] = hh != kNone ? L.key2[t] : ~0ull;
            L.si[t] = hh != kNone ? t : kNone;
        }
        for (uint32_t q = t0; q < H; q += kFThreads) {
            s.begin[q] = 0;
            s.end[q] = 0;
        }
        __syncthreads();
        sortLevel(L, width);
        for (uint32_t k = t0; k < A2; k += kFThreads) {
            const uint32_t hh = L.ss[k];
            if (k == 0 || L.ss[k - 1] != hh) s.begin[hh] = k;
            if (k + 1 == A2 || L.ss[k + 1] != hh) s.end[hh] = k + 1;
            L.P[k] = areaFixOf(a.area, L.id[L.si[k]]);
        }
        __syncthreads();
        scanU64(L.P, A2, L);
        // greedy stripe filling (simpl_lal.cc:288-295): stripe q of a half ends
        // at the first cell with the area before it, from the stripe's start,
        // at or above its capacity; the last stripe takes the rest
        for (uint32_t hh = t0; hh < H; hh += kFThreads) {
            const uint32_t ns = s.nStripes[hh];
            uint32_t start = s.begin[hh];
            const uint32_t e = s.end[hh];
            for (uint32_t q = 0; q < ns; ++q) {
                const uint32_t slot = hh * kFStripes + q;
                uint32_t stop = e;
                if (q + 1 < ns) {
                    const uint64_t base = start > 0 ? L.P[start - 1] : 0ull;
                    const double cap = s.stripe[3 * slot + 2];
                    uint32_t lo = start, hiK = e;
                    while (lo < hiK) {
                        const uint32_t mid = (lo + hiK) / 2;
                        const uint64_t before = mid > 0 ? L.P[mid - 1] : 0ull;
                        const double used = static_cast<double>(before - base) / kFixScale;
                        if (used >= cap) hiK = mid; else lo = mid + 1;
                    }
                    stop = lo;
                }
                s.stripeEnd[slot] = stop;
                s.stripeLo[slot] = ~0ull;
                s.stripeHi[slot] = 0ull;
                start = stop;
            }
        }
        __syncthreads();
        // the stripe of each cell and each stripe's extent of centres
        for (uint32_t k = t0; k < A2; k += kFThreads) {
            const uint32_t hh = L.ss[k];
            uint32_t q = 0;
            const uint32_t ns = s.nStripes[hh];
            while (q + 1 < ns && s.stripeEnd[hh * kFStripes + q] <= k) ++q;
            const uint32_t slot = hh * kFStripes + q;
            s.stripeOf[k] = slot;
            const unsigned long long key = orderKey(centreOn(a, L.id[L.si[k]], vertical));
            atomicMin(s.stripeLo + slot, key);
            atomicMax(s.stripeHi + slot, key);
        }
        __threadfence();
        __syncthreads();
        // affine map inside each stripe (simpl_lal.cc:297-314)
        double *pos = vertical ? a.x : a.y;
        const double *size = vertical ? a.w : a.h;
        for (uint32_t k = t0; k < A2; k += kFThreads) {
            const uint32_t slot = s.stripeOf[k];
            const uint32_t i = L.id[L.si[k]];
            const double lo = fromOrderKey(__hip_atomic_load(s.stripeLo + slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            const double hi = fromOrderKey(__hip_atomic_load(s.stripeHi + slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            const double ci = pos[i] + 0.5 * size[i];
            const double tt = hi > lo ? (ci - lo) / (hi - lo) : 0.5;
            const double sz = size[i];
            const double start = s.stripe[3 * slot], stop = s.stripe[3 * slot + 1];
            const double cc = start + tt * (stop - start);
            const double upper = start < stop - sz ? stop - sz : start;
            const double v = cc - 0.5 * sz;
            pos[i] = v < start ? start : (upper < v ? upper : v);
        }
        // 5. next level's blocks: the halves that hold cells, in half order
        for (uint32_t b = t0; b < B; b += kFThreads) {
            uint64_t kept = 0;
            if (s.split[b]) {
                const LalSegStat st = s.stats[b];
                kept = (st.m > 0 ? 1u : 0u) + (st.count > st.m ? 1u : 0u);
            }
            L.P[b] = kept;
        }
        __syncthreads();
        scanU64(L.P, B, L);
        for (uint32_t b = t0; b < B; b += kFThreads) {
            const uint32_t base = static_cast<uint32_t>(b > 0 ? L.P[b - 1] : 0ull);
            s.remap[2 * b] = kNone;
            s.remap[2 * b + 1] = kNone;
            if (!s.split[b]) continue;
            const LalSegStat st = s.stats[b];
            uint32_t next = base;
            for (uint32_t side = 0; side < 2; ++side) {
                const bool holds = side == 0 ? st.m > 0 : st.count > st.m;
                if (!holds) continue;
                const uint32_t hh = 2 * b + side;
                s.remap[hh] = next;
                for (int k = 0; k < 4; ++k) s.blockBox[par ^ 1][4 * next + k] = s.halfBox[4 * hh + k];
                s.blockRoom[par ^ 1][next] = s.blockRoom[par][b];
                s.blockRegion[par ^ 1][next] = s.blockRegion[par][b];
                ++next;
            }
        }
        if (t0 == 0) L.nextBlocks = static_cast<uint32_t>(B > 0 ? L.P[B - 1] : 0ull);
        __syncthreads();
        for (uint32_t j = t0; j < M; j += kFThreads) {
            const uint32_t hh = L.half[j];
            L.seg[j] = hh == kNone ? kNone : s.remap[hh];
        }
        const bool more = L.more != 0 && level + 1 < a.maxLevel;
        B = L.nextBlocks;
        par ^= 1;
        vertical = !vertical;
        A = A2;
        __syncthreads();
        if (!more) break;
    }
    return true;
}

__global__ void __launch_bounds__(kFThreads) lalFusedBatch(FusedArgs a) {
    __shared__ FusedLds L;
    const FusedScratch &s = a.s;
    const FusedRecord &rec = a.rec;
    const uint32_t M = rec.M, R = rec.R;
    const uint32_t t0 = threadIdx.x;
    if (M > kFCells || R > kFRegions || rec.rows > kFRows || rec.spans > kFSpans || rec.obs > kFObs) {
        if (t0 == 0) {  // nothing touched
            a.out[0] = kOverflow;
            a.out[1] = 0;
            a.out[2] = 0;
            a.out[3] = M;
        }
        return;
    }
    // the record into LDS
    {
        const uint32_t *rect = recArray<uint32_t>(rec, rec.oRect);
        const double *box = recArray<double>(rec, rec.oBox);
        const double *room = recArray<double>(rec, rec.oRoom);
        const uint32_t *rrow = recArray<uint32_t>(rec, rec.oRRow);
        const uint32_t *robs = recArray<uint32_t>(rec, rec.oRObs);
        for (uint32_t r = t0; r < R; r += kFThreads) {
            for (int k = 0; k < 4; ++k) {
                L.rect[r][k] = rect[4 * r + k];
                L.box[r][k] = box[4 * r + k];
            }
            L.room[r] = room[r];
            L.rrow[r][0] = rrow[2 * r];
            L.rrow[r][1] = rrow[2 * r + 1];
            L.robs[r][0] = robs[2 * r];
            L.robs[r][1] = robs[2 * r + 1];
        }
        const double *y0 = recArray<double>(rec, rec.oRowY0), *y1 = recArray<double>(rec, rec.oRowY1);
        const uint32_t *rs = recArray<uint32_t>(rec, rec.oRowSpan);
        for (uint32_t r = t0; r < rec.rows; r += kFThreads) {
            L.rowY0[r] = y0[r];
            L.rowY1[r] = y1[r];
            L.rowSpan[r][0] = rs[2 * r];
            L.rowSpan[r][1] = rs[2 * r + 1];
        }
        const double *sa = recArray<double>(rec, rec.oSpanA), *sb = recArray<double>(rec, rec.oSpanB);
        const uint32_t *sr = recArray<uint32_t>(rec, rec.oSpanRow);
        for (uint32_t q = t0; q < rec.spans; q += kFThreads) {
            L.spanA[q] = sa[q];
            L.spanB[q] = sb[q];
            L.spanRow[q] = sr[q];
        }
        const double *ob = recArray<double>(rec, rec.oObs);
        for (uint32_t q = t0; q < rec.obs; q += kFThreads)
            for (int k = 0; k < 4; ++k) L.obs[q][k] = ob[4 * q + k];
        const uint32_t *cell = recArray<uint32_t>(rec, rec.oCell), *creg = recArray<uint32_t>(rec, rec.oCReg);
        for (uint32_t j = t0; j < M; j += kFThreads) {
            const uint32_t i = cell[j];
            L.id[j] = i;
            L.seg[j] = creg[j];
            const double x = a.x[i], y = a.y[i];
            const uint32_t bin = a.binOf[i];
            s.startX[j] = x;
            s.startY[j] = y;
            s.startBin[j] = bin;
            s.saveX[j] = x;
            s.saveY[j] = y;
            s.saveBin[j] = bin;
            s.regionOf[j] = creg[j];
        }
        if (t0 == 0) {
            L.status = 0;
            L.moves = 0;
        }
    }
    __syncthreads();
    // Passes: the levels, then the in-order rule.  A cell that moved into the
    // bins of a later region of the batch is legalized there too, as the
    // reference's sequential run does: the regions from the first such one on
    // are restored and run again with the cells their bins hold now.  `first`
    // grows every pass, so there are at most R passes.
    uint32_t first = 0, reruns = 0, A = M;
    bool ok = true;
    for (;;) {
        ok = runLevels(L, a, A);
        if (!ok) break;
        if (t0 == 0) L.bad = kNone;
        __syncthreads();
        for (uint32_t j = t0; j < M; j += kFThreads) {
            const uint32_t i = L.id[j];
            const uint32_t k = binOfPoint(a.g, a.x[i] + 0.5 * a.w[i], a.y[i] + 0.5 * a.h[i]);
            const uint32_t r = s.regionOf[j];
            if (r != kNone && r >= first) {
                const uint32_t rj = regionOfBin(L, R, k, a.g.nx);
                if (rj != kNone && rj > r) atomicMin(&L.bad, rj);
            }
            a.binOf[i] = k;
        }
        __syncthreads();
        const uint32_t bad = L.bad;
        if (bad == kNone) break;
        for (uint32_t j = t0; j < M; j += kFThreads) {
            const uint32_t r = s.regionOf[j];
            if (r != kNone && r >= bad) {
                const uint32_t i = L.id[j];
                a.x[i] = s.saveX[j];
                a.y[i] = s.saveY[j];
                a.binOf[i] = s.saveBin[j];
            }
        }
        if (t0 == 0) L.count = 0;
        __syncthreads();
        for (uint32_t j = t0; j < M; j += kFThreads) {
            const uint32_t i = L.id[j];
            const uint32_t b = a.binOf[i];
            const uint32_t rj = regionOfBin(L, R, b, a.g.nx);
            if (rj != kNone && rj >= bad) {
                L.seg[j] = rj;
                atomicAdd(&L.count, 1u);
                const uint32_t r = s.regionOf[j];
                if (r == kNone || r < bad) {
                    s.regionOf[j] = rj;
                    s.saveX[j] = a.x[i];
                    s.saveY[j] = a.y[i];
                    s.saveBin[j] = b;
                }
            } else {
                L.seg[j] = kNone;
            }
        }
        __syncthreads();
        A = L.count;
        first = bad;
        ++reruns;
        __syncthreads();
    }
    if (!ok) {  // a table overflowed: the batch back to its start
        for (uint32_t j = t0; j < M; j += kFThreads) {
            const uint32_t i = L.id[j];
            a.x[i] = s.startX[j];
            a.y[i] = s.startY[j];
            a.binOf[i] = s.startBin[j];
        }
        if (t0 == 0) {
            a.out[0] = kOverflow;
            a.out[1] = 0;
            a.out[2] = reruns;
            a.out[3] = M;
        }
        return;
    }
    LalMoved *moved = reinterpret_cast<LalMoved *>(a.out + 4);
    for (uint32_t j = t0; j < M; j += kFThreads) {
        const uint32_t i = L.id[j];
        const uint32_t k = a.binOf[i], old = s.startBin[j];
        if (k == old) continue;
        const uint32_t slot = atomicAdd(&L.moves, 1u);
        moved[slot] = LalMoved{old, k, i, 0u, static_cast<unsigned long long>(areaFixOf(a.area, i))};
    }
    __syncthreads();
    if (t0 == 0) {
        a.out[0] = 0;
        a.out[1] = L.moves;
        a.out[2] = reruns;
        a.out[3] = M;
    }
}

template <typename T>
T *carve(uint8_t *&p, size_t count) {
    T *r = reinterpret_cast<T *>(p);
    p += (count * sizeof(T) + 255) & ~size_t{255};
    return r;
}

size_t scratchBytes() {
    size_t total = 0;
    const auto add = [&total](size_t bytes) { total += (bytes + 255) & ~size_t{255}; };
    for (int k = 0; k < 4; ++k) add(kFCells * sizeof(double));
    for (int k = 0; k < 3; ++k) add(kFCells * sizeof(uint32_t));
    for (int k = 0; k < 2; ++k) {
        add(4 * kFCells * sizeof(double));
        add(kFCells * sizeof(double));
        add(kFCells * sizeof(uint32_t));
    }
    add(kFCells * sizeof(LalSegStat));
    add(kFCells * sizeof(uint32_t));
    add(kFCells * sizeof(double));
    add(2 * kFCells * sizeof(uint32_t));
    add(2 * kFCells * sizeof(uint32_t));
    add(2 * kFCells * sizeof(uint32_t));
    add(2 * kFCells * kFStripes * 3 * sizeof(double));
    add(2 * kFCells * kFStripes * sizeof(uint32_t));
    add(2 * kFCells * kFStripes * sizeof(unsigned long long));
    add(2 * kFCells * kFStripes * sizeof(unsigned long long));
    add(2 * kFCells * 4 * sizeof(double));
    add(kFCells * sizeof(uint32_t));
    add(2 * kFCells * sizeof(uint32_t));
    return total;
}

FusedScratch carveScratch(uint8_t *p) {
    FusedScratch s{};
    s.startX = carve<double>(p, kFCells);
    s.startY = carve<double>(p, kFCells);
    s.saveX = carve<double>(p, kFCells);
    s.saveY = carve<double>(p, kFCells);
    s.startBin = carve<uint32_t>(p, kFCells);
    s.saveBin = carve<uint32_t>(p, kFCells);
    s.regionOf = carve<uint32_t>(p, kFCells);
    for (int k = 0; k < 2; ++k) {
        s.blockBox[k] = carve<double>(p, 4 * kFCells);
        s.blockRoom[k] = carve<double>(p, kFCells);
        s.blockRegion[k] = carve<uint32_t>(p, kFCells);
    }
    s.stats = carve<LalSegStat>(p, kFCells);
    s.split = carve<uint32_t>(p, kFCells);
    s.cut = carve<double>(p, kFCells);
    s.begin = carve<uint32_t>(p, 2 * kFCells);
    s.end = carve<uint32_t>(p, 2 * kFCells);
    s.nStripes = carve<uint32_t>(p, 2 * kFCells);
    s.stripe = carve<double>(p, 2 * kFCells * kFStripes * 3);
    s.stripeEnd = carve<uint32_t>(p, 2 * kFCells * kFStripes);
    s.stripeLo = carve<unsigned long long>(p, 2 * kFCells * kFStripes);
    s.stripeHi = carve<unsigned long long>(p, 2 * kFCells * kFStripes);
    s.halfBox = carve<double>(p, 2 * kFCells * 4);
    s.stripeOf = carve<uint32_t>(p, kFCells);
    s.remap = carve<uint32_t>(p, 2 * kFCells);
    return s;
}

// Byte layout of a record; every array starts on 8 bytes.
struct RecordLayout {
    uint32_t oRect, oBox, oRoom, oRRow, oRObs, oRowY0, oRowY1, oRowSpan, oSpanA, oSpanB, oSpanRow, oObs, oCell, oCReg,
        bytes;
};

RecordLayout layoutOf(uint32_t R, uint32_t rows, uint32_t spans, uint32_t obs, uint32_t M) {
    RecordLayout l{};
    uint32_t o = 0;
    const auto take = [&o](uint32_t bytes) {
        const uint32_t at = o;
        o += (bytes + 7u) & ~7u;
        return at;
    };
    l.oRect = take(16 * R);
    l.oBox = take(32 * R);
    l.oRoom = take(8 * R);
    l.oRRow = take(8 * R);
    l.oRObs = take(8 * R);
    l.oRowY0 = take(8 * rows);
    l.oRowY1 = take(8 * rows);
    l.oRowSpan = take(8 * rows);
    l.oSpanA = take(8 * spans);
    l.oSpanB = take(8 * spans);
    l.oSpanRow = take(4 * spans);
    l.oObs = take(32 * obs);
    l.oCell = take(4 * M);
    l.oCReg = take(4 * M);
    l.bytes = o;
    return l;
}

template <typename T>
void put(uint8_t *base, uint32_t off, const std::vector<T> &v) {
    if (!v.empty()) std::memcpy(base + off, v.data(), v.size() * sizeof(T));
}

}  // namespace

int lalStaging(kt_ctx *ctx, size_t bytes);
int lalLandReserve(kt_ctx *ctx, size_t varBytes);
uint8_t *lalLandVar(kt_ctx *ctx);

// Runs one batch in the fused kernel when it fits its tables.  `cells` are the
// batch's cells with their region, in ascending id order.  Returns KT_OK with
// *ran = false (nothing changed) when the batch does not fit; otherwise the
// cells whose bin changed are in `moved`.
int lalRunBatchFused(kt_ctx *ctx, const std::vector<lal::BinRect> &regions,
                     const std::vector<std::pair<uint32_t, uint32_t>> &cells, std::vector<LalMoved> &moved, bool *ran) {
    *ran = false;
    moved.clear();
    const uint32_t M = static_cast<uint32_t>(cells.size());
    const uint32_t R = static_cast<uint32_t>(regions.size());
    if (M == 0 || M > kFCells || R == 0 || R > kFRegions) return KT_OK;
    const lal::Grid &g = ctx->grid;
    const lal::FreeSpace &fs = ctx->freeSpace;
    const double frac = ctx->lalOpts.stripeAreaFraction;
    if (!(frac > 0.0)) return KT_OK;
    // stripes of one half box: its initial gaps (at most 2 per obstacle + 1),
    // then midpoint splits, at most about 2 / frac more; the kernel checks the
    // exact count again
    const double splitStripes = 2.5 / frac + 2.0;
    if (splitStripes > static_cast<double>(kFStripes)) return KT_OK;
    std::vector<uint32_t> rect(4 * R), rrow(2 * R), robs(2 * R), rowSpan, spanRow;
    std::vector<double> box(4 * R), room(R), rowY0, rowY1, spanA, spanB, obs;
    const size_t nRowsAll = fs.rows.size();
    for (uint32_t r = 0; r < R; ++r) {
        const lal::BinRect &b = regions[r];
        const lal::Box bx = g.box(b);
        rect[4 * r] = static_cast<uint32_t>(b.x0);
        rect[4 * r + 1] = static_cast<uint32_t>(b.x1);
        rect[4 * r + 2] = static_cast<uint32_t>(b.y0);
        rect[4 * r + 3] = static_cast<uint32_t>(b.y1);
        box[4 * r] = bx.x0;
        box[4 * r + 1] = bx.y0;
        box[4 * r + 2] = bx.x1;
        box[4 * r + 3] = bx.y1;
        room[r] = g.sum(g.avail, b);
        // rows meeting the region's height: from the first whose top is
        // above its bottom (upper_bound) up to the first that starts at or
        // above its top
        size_t lo = 0, hi = nRowsAll;
        while (lo < hi) {
            const size_t mid = (lo + hi) / 2;
            if (bx.y0 < fs.rows[mid].y1) hi = mid; else lo = mid + 1;
        }
        size_t last = lo;
        while (last < nRowsAll && fs.rows[last].y0 < bx.y1) ++last;
        rrow[2 * r] = static_cast<uint32_t>(rowY0.size());
        rrow[2 * r + 1] = static_cast<uint32_t>(last - lo);
        uint32_t regionSpans = 0, regionRows = static_cast<uint32_t>(last - lo);
        for (size_t q = lo; q < last; ++q) {
            rowSpan.push_back(static_cast<uint32_t>(spanA.size()));
            rowSpan.push_back(static_cast<uint32_t>(fs.rows[q].spans.size()));
            for (const lal::Span &sp : fs.rows[q].spans) {
                spanA.push_back(sp.a);
                spanB.push_back(sp.b);
                spanRow.push_back(static_cast<uint32_t>(rowY0.size()));
            }
            regionSpans += static_cast<uint32_t>(fs.rows[q].spans.size());
            rowY0.push_back(fs.rows[q].y0);
            rowY1.push_back(fs.rows[q].y1);
            if (rowY0.size() > kFRows || spanA.size() > kFSpans) return KT_OK;
        }
        // one profile of any box inside the region: at most two steps per
        // span (vertical) or two edges per row (horizontal)
        if (2ull * regionSpans > kFSteps || 2ull * regionRows + 4 > kFEdges) return KT_OK;
        robs[2 * r] = static_cast<uint32_t>(obs.size() / 4);
        uint32_t regionObs = 0;
        for (const lal::Box &o : ctx->obstacles) {
            if (o.x1 <= bx.x0 || o.x0 >= bx.x1 || o.y1 <= bx.y0 || o.y0 >= bx.y1) continue;
            obs.insert(obs.end(), {o.x0, o.y0, o.x1, o.y1});
            ++regionObs;
        }
        robs[2 * r + 1] = regionObs;
        if (obs.size() / 4 > kFObs || 2ull * regionObs + 2 > kFCuts ||
            2.0 * regionObs + 1.0 + splitStripes > static_cast<double>(kFStripes))
            return KT_OK;
    }
    const uint32_t rows = static_cast<uint32_t>(rowY0.size()), spans = static_cast<uint32_t>(spanA.size());
    const uint32_t nObs = static_cast<uint32_t>(obs.size() / 4);
    std::vector<uint32_t> cellId(M), cellReg(M);
    for (uint32_t j = 0; j < M; ++j) {
        cellId[j] = cells[j].first;
        cellReg[j] = cells[j].second;
    }
    const RecordLayout lay = layoutOf(R, rows, spans, nObs, M);
    // scratch, record and result buffers (allocated once, at their largest)
    if (ctx->lalFusedScratch.cap == 0) KT_TRY(ctx->lalFusedScratch.reserve(scratchBytes()));
    KT_TRY(ctx->lalFusedRecord.reserve(layoutOf(kFRegions, kFRows, kFSpans, kFObs, kFCells).bytes));
    KT_TRY(ctx->lalFusedOut.reserve(4ull + 3ull * kFCells));
    KT_TRY(lalStaging(ctx, lay.bytes));
    uint8_t *host = ctx->lalPinned;
    put(host, lay.oRect, rect);
    put(host, lay.oBox, box);
    put(host, lay.oRoom, room);
    put(host, lay.oRRow, rrow);
    put(host, lay.oRObs, robs);
    put(host, lay.oRowY0, rowY0);
    put(host, lay.oRowY1, rowY1);
    put(host, lay.oRowSpan, rowSpan);
    put(host, lay.oSpanA, spanA);
    put(host, lay.oSpanB, spanB);
    put(host, lay.oSpanRow, spanRow);
    put(host, lay.oObs, obs);
    put(host, lay.oCell, cellId);
    put(host, lay.oCReg, cellReg);
    KT_HIP(hipMemcpyAsync(ctx->lalFusedRecord.p, host, lay.bytes, hipMemcpyHostToDevice, ctx->stream));
    FusedArgs a{};
    a.rec = FusedRecord{ctx->lalFusedRecord.p, lay.oRect, lay.oBox, lay.oRoom, lay.oRRow, lay.oRObs, lay.oRowY0,
                        lay.oRowY1, lay.oRowSpan, lay.oSpanA, lay.oSpanB, lay.oSpanRow, lay.oObs, lay.oCell,
                        lay.oCReg, M, R, rows, spans, nObs};
    a.g = GridDev{g.ox, g.oy, g.dx, g.dy, static_cast<uint32_t>(g.nx), static_cast<uint32_t>(g.ny)};
    a.x = ctx->posX.p;
    a.y = ctx->posY.p;
    a.w = ctx->width.p;
    a.h = ctx->height.p;
    a.area = ctx->area.p;
    a.binOf = ctx->binOf.p;
    a.maxLevel = static_cast<uint32_t>(std::min<size_t>(ctx->lalOpts.maxLevel, 0xffffffffu));
    a.minCells = static_cast<uint32_t>(std::min<size_t>(ctx->lalOpts.minCellsToSplit, 0xffffffffu));
    a.frac = frac;
    a.density = ctx->densityLimit;
    a.s = carveScratch(ctx->lalFusedScratch.p);
    a.out = ctx->lalFusedOut.p;
    lalFusedBatch<<<1, kFThreads, 0, ctx->stream>>>(a);
    KT_HIP(hipGetLastError());
    const size_t back = (4ull + 3ull * M) * sizeof(unsigned long long);
    KT_TRY(lalLandReserve(ctx, back));
    uint8_t *land = lalLandVar(ctx);
    KT_HIP(hipMemcpyAsync(land, ctx->lalFusedOut.p, back, hipMemcpyDeviceToHost, ctx->stream));
    KT_HIP(hipStreamSynchronize(ctx->stream));
    const unsigned long long *head = reinterpret_cast<const unsigned long long *>(land);
    KT_CHECK(head[3] == M, KT_ESTATE, "lal: the fused batch kernel saw another cell count");
    if (head[0] != 0) return KT_OK;  // a table overflowed; the batch is back at its start
    const uint64_t moves = head[1];
    KT_CHECK(moves <= M, KT_ESTATE, "lal: the fused batch kernel moved more cells than it holds");
    ctx->lalReruns += head[2];
    moved.resize(moves);
    if (moves) std::memcpy(moved.data(), land + 4 * sizeof(unsigned long long), moves * sizeof(LalMoved));
    *ran = true;
    return KT_OK;
}

}  // namespace kt
