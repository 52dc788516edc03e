// Stable LSD radix sort of (uint64 key, uint32 value) pairs on gfx950.
//
// Used by the look-ahead legalization to order cells inside their blocks
// (key = block id and an order-preserving image of the cell centre, value =
// cell id; ties keep the input order, so sorting first by id and then by key
// gives the reference's total order (key, id), simpl_parallel.h:26-33).
//
// One 8-bit digit per pass, three launches per pass:
//   radixHistogram  per-tile digit counts, stored digit-major [256][tiles]
//   scan            exclusive scan of the counts (the row-pointer scan)
//   radixScatter    per-tile stable ranks (wave ballots per digit bit, then
//                   wave and chunk prefixes in LDS) plus the scanned offset
// Passes whose digit is constant over the whole input are skipped by the
// caller through `bitsToSort`.
#include "kt_internal.h"

namespace kt {
namespace {

constexpr int kSortBlock = 1024;
constexpr int kSortWaves = kSortBlock / 64;
constexpr uint32_t kSortTile = 4096;  // items per block per pass

__device__ __forceinline__ uint32_t digitOf(uint64_t key, int shift) {
    return static_cast<uint32_t>((key >> shift) & 0xffu);
}

__global__ void __launch_bounds__(kSortBlock)
radixHistogram(const uint64_t *__restrict__ keys, uint64_t n, int shift, uint32_t *__restrict__ counts,
               uint32_t tiles) {
    __shared__ uint32_t hist[256];
    for (int d = threadIdx.x; d < 256; d += kSortBlock) hist[d] = 0;
    __syncthreads();
    const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kSortTile;
    for (uint32_t k = threadIdx.x; k < kSortTile; k += kSortBlock) {
        const uint64_t i = base + k;
        if (i < n) atomicAdd(&hist[digitOf(keys[i], shift)], 1u);
    }
    __syncthreads();
    for (int d = threadIdx.x; d < 256; d += kSortBlock) counts[static_cast<size_t>(d) * tiles + blockIdx.x] = hist[d];
}

// counts has been turned into inclusive prefix sums; offset of (digit, tile)
// = prefix - count.
__global__ void __launch_bounds__(kSortBlock)
radixScatter(const uint64_t *__restrict__ keysIn, const uint32_t *__restrict__ valsIn, uint64_t n, int shift,
             const uint32_t *__restrict__ prefix, const uint32_t *__restrict__ rawCounts, uint32_t tiles,
             uint64_t *__restrict__ keysOut, uint32_t *__restrict__ valsOut) {
    __shared__ uint32_t base[256];        // next free slot per digit for this tile
    __shared__ uint32_t waveCount[kSortWaves][256];
    for (int d = threadIdx.x; d < 256; d += kSortBlock) {
        const size_t at = static_cast<size_t>(d) * tiles + blockIdx.x;
        base[d] = prefix[at] - rawCounts[at];
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t lanesBelow = lane ? (~0ull >> (64 - lane)) : 0ull;
    __syncthreads();
    const uint64_t tileBase = static_cast<uint64_t>(blockIdx.x) * kSortTile;
    for (uint32_t chunk = 0; chunk < kSortTile; chunk += kSortBlock) {
        const uint64_t i = tileBase + chunk + threadIdx.x;
        const bool valid = i < n;
        const uint64_t key = valid ? keysIn[i] : 0ull;
        const uint32_t val = valid ? valsIn[i] : 0u;
        const uint32_t d = digitOf(key, shift);
        // lanes of this wave with the same digit (valid lanes only)
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const uint64_t set = __ballot(valid && ((d >> b) & 1u));
            peers &= ((d >> b) & 1u) ? set : ~set;
        }
        const uint32_t rankInWave = __popcll(peers & lanesBelow);
        for (int x = threadIdx.x; x < kSortWaves * 256; x += kSortBlock) (&waveCount[0][0])[x] = 0;
        __syncthreads();
        if (valid && rankInWave == 0) waveCount[wave][d] = __popcll(peers);
        __syncthreads();
        uint32_t before = 0;
        if (valid)
            for (int w = 0; w < wave; ++w) before += waveCount[w][d];
        if (valid) {
            const uint32_t dst = base[d] + before + rankInWave;
            keysOut[dst] = key;
            valsOut[dst] = val;
        }
        __syncthreads();
        // advance the per-digit bases by this chunk's counts
        for (int dd = threadIdx.x; dd < 256; dd += kSortBlock) {
            uint32_t tot = 0;
            for (int w = 0; w < kSortWaves; ++w) tot += waveCount[w][dd];
            base[dd] += tot;
        }
        __syncthreads();
    }
}

}  // namespace

int scanInclusive(kt_ctx *ctx, uint32_t *data, uint64_t m);

// Sorts n pairs by key (stable), over the low `bits` bits of the key.  The
// result ends in (keys, vals); (keysTmp, valsTmp) are scratch of size n.
namespace {

constexpr uint32_t kSmallSort = 2048;

// One workgroup sorts up to kSmallSort ids by (seg[id], key, id) with a
// bitonic network in LDS; key is positional (key[j] belongs to ids[j]).  The
// order is total (ids are unique), so it equals the order of a stable sort by
// key followed by a stable sort by segment, starting from ascending ids.
__global__ void __launch_bounds__(1024) smallSortKernel(uint32_t *__restrict__ ids, const uint64_t *__restrict__ key,
                                                        const uint32_t *__restrict__ segById, uint32_t count,
                                                        uint32_t width) {
    __shared__ uint64_t sk[kSmallSort];
    __shared__ uint32_t ss[kSmallSort], si[kSmallSort];
    for (uint32_t j = threadIdx.x; j < width; j += blockDim.x) {
        if (j < count) {
            const uint32_t id = ids[j];
            si[j] = id;
            sk[j] = key[j];
            ss[j] = segById[id];
        } else {  // padding sorts last
            si[j] = 0xffffffffu;
            sk[j] = ~0ull;
            ss[j] = 0xffffffffu;
        }
    }
    __syncthreads();
    for (uint32_t k = 2; k <= width; k <<= 1) {
        for (uint32_t d = k >> 1; d > 0; d >>= 1) {
            for (uint32_t t = threadIdx.x; t < width; t += blockDim.x) {
                const uint32_t u = t ^ d;
                if (u > t) {
                    const bool greater = ss[t] != ss[u] ? ss[t] > ss[u]
                                         : sk[t] != sk[u] ? sk[t] > sk[u]
                                                          : si[t] > si[u];
                    const bool ascending = (t & k) == 0;
                    if (greater == ascending) {
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
    for (uint32_t j = threadIdx.x; j < count; j += blockDim.x) ids[j] = si[j];
}

}  // namespace

bool smallSortFits(uint64_t n) { return n <= kSmallSort; }

int smallSortBySegment(kt_ctx *ctx, uint32_t *ids, const uint64_t *key, const uint32_t *segById, uint64_t n) {
    KT_CHECK(n <= kSmallSort, KT_EINVAL, "smallSortBySegment: too many elements");
    if (n <= 1) return KT_OK;
    uint32_t width = 2;
    while (width < n) width <<= 1;
    smallSortKernel<<<dim3(1), dim3(1024), 0, ctx->stream>>>(ids, key, segById, static_cast<uint32_t>(n), width);
    KT_HIP(hipGetLastError());
    return KT_OK;
}

int radixSortPairs(kt_ctx *ctx, uint64_t *keys, uint32_t *vals, uint64_t *keysTmp, uint32_t *valsTmp, uint64_t n,
                   int bits, DevBuf<uint32_t> &counts, DevBuf<uint32_t> &raw) {
    if (n <= 1 || bits <= 0) return KT_OK;
    const uint32_t tiles = static_cast<uint32_t>((n + kSortTile - 1) / kSortTile);
    KT_TRY(counts.reserve(256ull * tiles));
    KT_TRY(raw.reserve(256ull * tiles));
    uint64_t *kin = keys, *kout = keysTmp;
    uint32_t *vin = vals, *vout = valsTmp;
    int passes = 0;
    for (int shift = 0; shift < bits; shift += 8, ++passes) {
        radixHistogram<<<dim3(tiles), dim3(kSortBlock), 0, ctx->stream>>>(kin, n, shift, raw.p, tiles);
        KT_HIP(hipGetLastError());
        KT_HIP(hipMemcpyAsync(counts.p, raw.p, 256ull * tiles * sizeof(uint32_t), hipMemcpyDeviceToDevice,
                              ctx->stream));
        KT_TRY(scanInclusive(ctx, counts.p, 256ull * tiles));
        radixScatter<<<dim3(tiles), dim3(kSortBlock), 0, ctx->stream>>>(kin, vin, n, shift, counts.p, raw.p, tiles,
                                                                        kout, vout);
        KT_HIP(hipGetLastError());
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    if (passes % 2 == 1) {
        KT_HIP(hipMemcpyAsync(keys, kin, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, ctx->stream));
        KT_HIP(hipMemcpyAsync(vals, vin, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, ctx->stream));
    }
    return KT_OK;
}

}  // namespace kt
