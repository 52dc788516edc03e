// Jacobi-preconditioned CG for the x and y systems, lock-step, on gfx950.
//
// Algorithm, stopping and breakdown rules follow solveBothAxes
// (simpl_lss.cc:203-290): x0 = the caller's positions; r = b - A x0; z = r/d;
// p = z; per iteration Ap, alpha = rho / p.Ap (stop the axis if p.Ap <= 0),
// x += alpha p, r -= alpha Ap, resid = ||r|| / ||b||, rho' = r.z (stop the axis
// if rho' <= 0), beta = rho' / rho, p = z + beta p.  An axis steps while
// active && resid > tol; the loop count is the reference's lock-step count.
//
// Layout and launch shape
//   * Every launch covers both axes: blocks [0, blocks_x) work on x, the rest
//     on y.  A block owns a tile of `tile` consecutive rows (4096 by default,
//     1024 threads), so a 1M-row axis is 245 blocks and the whole launch about
//     two blocks per CU.
//   * Dot products: each block reduces its rows with a fixed tree and writes
//     one partial per tile.  The next kernel's blocks each re-reduce all tile
//     partials of their axis in tile order (a few hundred values), so every
//     block derives the same alpha / beta / flags without a grid barrier, an
//     atomic or a device-scope fence.
//   * Per-axis scalars (rho, resid, step count, live/active flags) are double
//     buffered by iteration parity: iteration k reads state[k & 1] and only
//     block 0 of its last kernel writes state[(k + 1) & 1], so no block ever
//     reads a value written by its own kernel.
//   * z = r / d is never stored; it is recomputed where needed.
//   * Vectors and matrix values are FP64 (plain solve) or FP32 (inner solve of
//     the mixed-precision path); every dot product accumulates in FP64.  The
//     FP32 inner solve runs by default on the scaled system D^-1/2 A D^-1/2
//     (unit diagonal): Jacobi-PCG in the coordinates x^ = D^1/2 x, where
//     z = r^, so the SpMV and the direction update read no diagonal; only
//     r.r = sum d r^2 (the stopping test) still reads it.
//   * Iterations are launched in pairs (parity 0, 1) captured in hipGraphs; the
//     host polls the state every few iterations.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <type_traits>

#include <hip/hip_ext.h>

#include "kt_internal.h"
#include "spmv_rows.h"

namespace kt {
namespace {

constexpr int kBlock = 1024;
// Dynamic LDS a staged SpMV block may take (of the 160 KiB per CU; the
// block's static reduction buffers need a few hundred bytes more).
constexpr uint64_t kMaxWindowBytes = 156ull * 1024;

// Block -> (axis, tile).  Workgroups are dealt to the 8 XCDs round-robin by
// id, so with xcdSplit the blocks that land on XCDs 0-3 take the x system and
// those on XCDs 4-7 the y system: each XCD's L2 then caches the gathered
// vector of one axis only.  Inside an iteration (parity >= 0) where only one
// axis is live, every block serves that axis, so the split does not idle half
// the GPU once the other axis has converged (all kernels of the iteration
// read the same flags).  Returns false for a block with no tile.
__device__ __forceinline__ bool decode(const CgArgs &a, int &axis, uint32_t &tile, int parity = -1) {
    if (a.xcdSplit == 1) {
        if (parity >= 0) {
            const bool lx = a.ctl[0].s[parity].live, ly = a.ctl[1].s[parity].live;
            if (lx != ly) {
                axis = lx ? 0 : 1;
                tile = blockIdx.x;
                return tile < a.sys[axis].blocks;
            }
        }
        const uint32_t b = blockIdx.x, x = b & 7u;
        axis = static_cast<int>(x >> 2);
        tile = (b >> 3) * 4u + (x & 3u);
        return tile < a.sys[axis].blocks;
    }
    // inside an iteration with one axis live, the first blocks serve it and
    // the rest exit, so the empty blocks are dispatched after the working ones
    if (parity >= 0) {
        const bool lx = a.ctl[0].s[parity].live, ly = a.ctl[1].s[parity].live;
        if (lx != ly) {
            axis = lx ? 0 : 1;
            tile = blockIdx.x;
            return tile < a.sys[axis].blocks;
        }
    }
    axis = blockIdx.x >= a.sys[0].blocks ? 1 : 0;
    tile = blockIdx.x - (axis ? a.sys[0].blocks : 0);
    return true;
}

// Fixed-order block sum of K values per thread; the result is returned to
// every thread of the block.
template <int K, int NT = kBlock>
__device__ __forceinline__ void blockSum(double (&v)[K]) {
    __shared__ double waveSums[NT / 64][K];
    __shared__ double total[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) waveSums[wave][k] = v[k];
    }
    __syncthreads();
    if (threadIdx.x < K) {
        double s = waveSums[0][threadIdx.x];
        for (int w = 1; w < NT / 64; ++w) s += waveSums[w][threadIdx.x];
        total[threadIdx.x] = s;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = total[k];
}

// Every block of an axis computes the same sum of the axis' tile partials:
// loadPartials (each thread's share, in tile order) then blockSum.
template <int K>
__device__ __forceinline__ void loadPartials(const double *part, uint32_t blocks, uint32_t stride, int axis,
                                             double (&out)[K]) {
    const double *src = part + static_cast<size_t>(axis) * stride * K;
#pragma unroll
    for (int k = 0; k < K; ++k) out[k] = 0.0;
    for (uint32_t t = threadIdx.x; t < blocks; t += kBlock) {
#pragma unroll
        for (int k = 0; k < K; ++k) out[k] += src[t * K + k];
    }
}

template <int K>
__device__ __forceinline__ void sumPartials(const double *part, uint32_t blocks, uint32_t stride, int axis,
                                            double (&out)[K]) {
    loadPartials<K>(part, blocks, stride, axis, out);
    blockSum<K>(out);
}

// loadPartials split around other loads: issue() loads each thread's first
// partial unconditionally (tile 0 exists) and consumes nothing, so loads
// issued after it stay in flight while it is waited for; finish() selects it,
// adds the rare further shares (more than 1024 tiles) and sums the block.
// S: shares per thread loaded at once (unconditional loads, index 0 standing
// in for a missing tile); more than S * 1024 tiles are summed in a loop after
// them.  The sum runs in tile order t, t + 1024, ... either way.
template <int K, int S = 1>
struct PartialShare {
    const double *src;
    uint32_t blocks;
    double first[S][K];
    __device__ __forceinline__ void issue(const double *part, uint32_t nblocks, uint32_t stride, int axis) {
        src = part + static_cast<size_t>(axis) * stride * K;
        blocks = nblocks;
#pragma unroll
        for (int j = 0; j < S; ++j) {
            const uint32_t u = threadIdx.x + j * kBlock;
            const uint32_t t = u < blocks ? u : 0u;
#pragma unroll
            for (int k = 0; k < K; ++k) first[j][k] = src[t * K + k];
        }
    }
    __device__ __forceinline__ void finish(double (&out)[K]) {
#pragma unroll
        for (int k = 0; k < K; ++k) out[k] = threadIdx.x < blocks ? first[0][k] : 0.0;
#pragma unroll
        for (int j = 1; j < S; ++j) {
            const bool ok = threadIdx.x + j * kBlock < blocks;
#pragma unroll
            for (int k = 0; k < K; ++k) out[k] += ok ? first[j][k] : 0.0;
        }
        for (uint32_t t = threadIdx.x + S * kBlock; t < blocks; t += kBlock) {
#pragma unroll
            for (int k = 0; k < K; ++k) out[k] += src[t * K + k];
        }
        blockSum<K>(out);
    }
};

template <int K>
__device__ __forceinline__ void writePartial(double *part, uint32_t stride, int axis, uint32_t tile,
                                             const double (&v)[K]) {
    if (threadIdx.x == 0) {
        double *dst = part + (static_cast<size_t>(axis) * stride + tile) * K;
#pragma unroll
        for (int k = 0; k < K; ++k) dst[k] = v[k];
    }
}

template <typename T>
struct Mat;
template <>
struct Mat<double> {
    static __device__ __forceinline__ const double *val(const SysView &s) { return s.val; }
    static __device__ __forceinline__ const double *diag(const SysView &s) { return s.diag; }
};
template <>
struct Mat<float> {
    static __device__ __forceinline__ const float *val(const SysView &s) { return s.val32; }
    static __device__ __forceinline__ const float *diag(const SysView &s) { return s.diag32; }
};

// Rows handled together by one lane group per pass.
constexpr int kRowsPerGroup = 2;

// ---------------------------------------------------------------- start
// r = b - A x (FP64), p = z = r / d in the working precision T; partials
// b.b, r.z, r.r.  For T = float the FP64 residual is rounded once (on the
// scaled system: r^ = D^-1/2 r rounded once, p = r^, partial r^.r^).
template <typename T, int G>
__global__ void __launch_bounds__(kBlock) cgStart(CgArgs a) {
    int axis;
    uint32_t tile;
    if (!decode(a, axis, tile)) return;
    if (a.ctl[axis].frozen) return;
    const SysView s = a.sys[axis];
    const double *__restrict__ x = a.x64[axis];
    T *__restrict__ r = static_cast<T *>(a.r[axis]);
    T *__restrict__ p = static_cast<T *>(a.p[axis]);
    T *__restrict__ e = static_cast<T *>(a.e[axis]);
    const T *__restrict__ dT = Mat<T>::diag(s);
    const uint32_t lane = threadIdx.x % G, grp = threadIdx.x / G;
    const uint32_t begin = tile * a.tile, end = min(begin + a.tile, s.rows);
    double acc[3] = {0.0, 0.0, 0.0};
    constexpr uint32_t kPass = kRowsPerGroup * (kBlock / G);
    for (uint32_t r0 = begin; r0 < end; r0 += kPass) {
        uint32_t row[kRowsPerGroup];
        bool valid[kRowsPerGroup];
        double off[kRowsPerGroup];
#pragma unroll
        for (int q = 0; q < kRowsPerGroup; ++q) {
            row[q] = r0 + grp + q * (kBlock / G);
            valid[q] = row[q] < end;
        }
        rowsDot<double, G, kRowsPerGroup>(s.rowPtr, s.col, s.val, x, row, valid, lane, off);
#pragma unroll
        for (int q = 0; q < kRowsPerGroup; ++q) {
            const uint32_t i = row[q];
            if (valid[q] && lane == 0) {
                const double ri = s.rhs[i] - (s.diag[i] * x[i] + off[q]);
                // scaled system: r^ = D^-1/2 r, and z = r^ (unit diagonal)
                const T rt = static_cast<T>(a.unitDiag ? ri * s.isd[i] : ri);
                const T zt = a.unitDiag ? rt : rt / dT[i];
                r[i] = rt;
                p[i] = zt;
                if (e) e[i] = T(0);
                acc[0] += s.rhs[i] * s.rhs[i];
                acc[1] += static_cast<double>(rt) * static_cast<double>(zt);
                acc[2] += ri * ri;
            }
        }
    }
    blockSum<3>(acc);
    writePartial<3>(a.partS, a.stride, axis, tile, acc);
}

// One block per axis: turns the start partials into state[0].
__global__ void __launch_bounds__(kBlock) cgStartFinish(CgArgs a) {
    const int axis = blockIdx.x;
    AxisCtl &c = a.ctl[axis];
    if (c.frozen) return;
    double tot[3];
    sumPartials<3>(a.partS, a.flatResidual ? a.sys[axis].sblocks : a.sys[axis].blocks, a.stride, axis, tot);
    if (threadIdx.x == 0) {
        CgState &st = c.s[0];
        c.bNorm = sqrt(tot[0]);
        if (!(c.bNorm > 0.0)) {
            st.active = 0;
            st.live = 0;
        } else {
            st.rho = tot[1];
            st.pad = 0;
            st.resid = sqrt(tot[2]) / c.bNorm;
            st.active = 1;
            st.live = (st.resid > c.tol && st.steps < c.maxIter) ? 1 : 0;
        }
        c.trueResid = st.resid;
        c.refResid = st.resid;
        c.refRz = tot[1];
        c.needUpdate = 0;
    }
}

// ---------------------------------------------------------------- one iteration
// Copies v[lo, lo + span) into the block's LDS window (16-byte loads; lo is a
// multiple of 256, so both ends of the vector part are 16-byte aligned).
template <typename T>
__device__ __forceinline__ void stageWindow(const T *__restrict__ v, uint32_t lo, uint32_t span, T *lds) {
    constexpr uint32_t kVec = 16 / sizeof(T);
    const uint32_t vecs = span / kVec;
    const uint4 *__restrict__ src = reinterpret_cast<const uint4 *>(v + lo);
    uint4 *dst = reinterpret_cast<uint4 *>(lds);
    for (uint32_t k = threadIdx.x; k < vecs; k += kBlock) dst[k] = src[k];
    for (uint32_t k = vecs * kVec + threadIdx.x; k < span; k += kBlock) lds[k] = v[lo + k];
}

// LDS bytes of a tile's window: the tile and `halo` entries on each side.
template <typename T>
constexpr uint64_t windowBytes(uint32_t tile, uint32_t halo) {
    return (static_cast<uint64_t>(tile) + 2ull * halo) * sizeof(T);
}

// Ap = A p and the tile partial of p.Ap.  With S, the block first stages
// p[begin - halo, end + halo) in LDS (dynamic shared memory sized by the
// launch) and gathers the columns inside that window from there: in a
// placement netlist most of a row's columns lie near the row, so most
// gathers stop costing an L2 request.
template <typename T, int G, int R, bool S>
__global__ void __launch_bounds__(kBlock) cgSpmv(CgArgs a, int parity) {
    int axis;
    uint32_t tile;
    if (!decode(a, axis, tile, parity)) return;
    if (!a.ctl[axis].s[parity].live) return;
    const SysView s = a.sys[axis];
    const T *__restrict__ p = static_cast<const T *>(a.p[axis]);
    T *__restrict__ ap = static_cast<T *>(a.ap[axis]);
    const T *__restrict__ val = Mat<T>::val(s);
    const T *__restrict__ diag = Mat<T>::diag(s);
    const bool unit = sizeof(T) == 4 && a.unitDiag;  // the scaled FP32 system: d = 1, not read
    const uint32_t lane = threadIdx.x % G, grp = threadIdx.x / G;
    const uint32_t begin = tile * a.tile, end = min(begin + a.tile, s.rows);
    double acc[1] = {0.0};
    constexpr uint32_t kPass = R * (kBlock / G);
    const auto rows = [&](const auto &src) {
        for (uint32_t r0 = begin; r0 < end; r0 += kPass) {
            uint32_t row[R];
            bool valid[R];
            double off[R];
            T pi[R], di[R];
#pragma unroll
            for (int q = 0; q < R; ++q) {
                row[q] = r0 + grp + q * (kBlock / G);
                valid[q] = row[q] < end;
                const uint32_t i = valid[q] ? row[q] : begin;
                pi[q] = ownAt(src, i);
                di[q] = unit ? T(1) : diag[i];
            }
            rowsDot<T, G, R>(s.rowPtr, s.col, val, src, row, valid, lane, off);
#pragma unroll
            for (int q = 0; q < R; ++q) {
                if (valid[q] && lane == 0) {
                    const double pd = static_cast<double>(pi[q]);
                    const T v = static_cast<T>(static_cast<double>(di[q]) * pd + off[q]);
                    ap[row[q]] = v;
                    acc[0] += pd * static_cast<double>(v);
                }
            }
        }
    };
    if constexpr (S) {
        extern __shared__ uint4 cgWindow[];
        T *lds = reinterpret_cast<T *>(cgWindow);
        const uint32_t lo = begin > a.halo ? begin - a.halo : 0u;
        const uint32_t hi = s.rows - end > a.halo ? end + a.halo : s.rows;
        stageWindow<T>(p, lo, hi - lo, lds);
        __syncthreads();
        rows(Window<T>{p, (LdsPtr<T>)lds, lo, hi - lo});
    } else {
        rows(p);
    }
    blockSum<1>(acc);
    writePartial<1>(a.partA, a.stride, axis, tile, acc);
}

// ---------------------------------------------------------------- entry-block SpMV
// Block b of an axis takes the rows [rowStart[b], rowStart[b+1]), whose
// entries are a run of at most E + the longest row (rowStart cuts the entry
// sequence every E entries at the next row start).  With NT threads:
//   phase 1: thread t takes entries t, t + NT, ... of the run, so each load
//            instruction of a wave reads 64 consecutive columns and values;
//            it gathers p and leaves the product (in double: exact for FP32
//            operands) in LDS;
//   phase 2: thread t takes rows R0 + t, R0 + t + NT, ...: Ap = d p, then
//            the row's products added in entry order (SparseSystem::multiply's
//            order, simpl_lss.h:49-57), and the row's share of p.Ap.
// A run longer than the LDS holds (more than E + flatSlack(E) entries: a row
// that long crossing a cut) is done with the lane-group rows instead.  The
// slack is an eighth of the block, at most 512 entries, so 2048-entry blocks
// of 256 threads fit eight to a CU's LDS.
// Four consecutive entries k0 .. k0 + 3 (k0 a multiple of four): columns
// and values by one wide load per array (3-byte columns: one 12-byte load of
// the packed quad).
template <typename T, bool C24>
__device__ __forceinline__ void loadQuad(const SysView &s, const T *__restrict__ val, uint32_t k0, uint32_t (&c)[4],
                                         T (&w)[4]) {
    if constexpr (C24) {
        const uint32_t *w = s.colPk + 3u * (k0 >> 2);
        const uint32_t w0 = w[0], w1 = w[1], w2 = w[2];
        c[0] = w0 & 0xffffffu;
        c[1] = (w0 >> 24) | ((w1 & 0xffffu) << 8);
        c[2] = (w1 >> 16) | ((w2 & 0xffu) << 16);
        c[3] = w2 >> 8;
    } else {
        const uint4 q = *reinterpret_cast<const uint4 *>(s.col + k0);
        c[0] = q.x;
        c[1] = q.y;
        c[2] = q.z;
        c[3] = q.w;
    }
    if constexpr (sizeof(T) == 4) {
        const float4 v = *reinterpret_cast<const float4 *>(val + k0);
        w[0] = v.x;
        w[1] = v.y;
        w[2] = v.z;
        w[3] = v.w;
    } else {
        const double2 v0 = *reinterpret_cast<const double2 *>(val + k0);
        const double2 v1 = *reinterpret_cast<const double2 *>(val + k0 + 2);
        w[0] = v0.x;
        w[1] = v0.y;
        w[2] = v1.x;
        w[3] = v1.y;
    }
}

constexpr uint32_t flatSlack(uint32_t entries) { return entries / 8 < 512 ? entries / 8 : 512; }

// Block -> (axis, block of that axis).  With xcdSplit the blocks that land on
// XCDs 0-3 (workgroups are dealt to the 8 XCDs round-robin) take the x
// system and those on XCDs 4-7 the y system, so each XCD's L2 sees the
// gathers of one axis only.
// Position of block `id` (of g blocks, id < g) when each XCD takes one
// contiguous share of [0, g): workgroups are dealt to the 8 XCDs round-robin
// (id % 8), so XCD k's i-th block gets index k * (g / 8) + min(k, g % 8) + i.
// A share of consecutive entry blocks gathers from one stretch of the vector
// (the netlist's local nets), which then stays in that XCD's L2 instead of
// being fetched by all eight.
__device__ __forceinline__ uint32_t xcdShare(uint32_t id, uint32_t g) {
    const uint32_t k = id & 7u, i = id >> 3, q = g >> 3, rem = g & 7u;
    return k * q + (k < rem ? k : rem) + i;
}

__device__ __forceinline__ bool decodeFlat(const CgArgs &a, int &axis, uint32_t &b, int parity) {
    if (a.xcdSplit == 2) {
        const uint32_t nx = a.sys[0].sblocks, ny = a.sys[1].sblocks;
        const bool lx = parity < 0 || a.ctl[0].s[parity].live, ly = parity < 0 || a.ctl[1].s[parity].live;
        if (lx != ly) {  // one axis live: its blocks spread over every XCD
            const uint32_t g = lx ? nx : ny;
            if (blockIdx.x >= g) return false;
            axis = lx ? 0 : 1;
            b = xcdShare(blockIdx.x, g);
            return true;
        }
        const uint32_t l = xcdShare(blockIdx.x, nx + ny);  // the grid is nx + ny blocks
        axis = l >= nx ? 1 : 0;
        b = l - (axis ? nx : 0);
        return true;
    }
    if (a.xcdSplit == 1) {
        const bool lx = parity >= 0 && a.ctl[0].s[parity].live, ly = parity >= 0 && a.ctl[1].s[parity].live;
        if (lx != ly) {  // inside an iteration, one axis live: every block serves it (as decode)
            axis = lx ? 0 : 1;
            b = blockIdx.x;
            return b < a.sys[axis].sblocks;
        }
        const uint32_t id = blockIdx.x, x = id & 7u;
        axis = static_cast<int>(x >> 2);
        b = (id >> 3) * 4u + (x & 3u);
        return b < a.sys[axis].sblocks;
    }
    if (parity >= 0) {  // one axis live: the first blocks serve it (as decode)
        const bool lx = a.ctl[0].s[parity].live, ly = a.ctl[1].s[parity].live;
        if (lx != ly) {
            axis = lx ? 0 : 1;
            b = blockIdx.x;
            return b < a.sys[axis].sblocks;
        }
    }
    axis = blockIdx.x >= a.sys[0].sblocks ? 1 : 0;
    b = blockIdx.x - (axis ? a.sys[0].sblocks : 0);
    return true;
}

// Threads per entry block (0 picks 1024).
int flatThreads(int tuned) { return tuned == 256 || tuned == 512 ? tuned : kBlock; }

// only = 0 or 1: an iteration launched while that axis alone is live gets
// just its blocks (the kernels map every block to the one live axis), so no
// workgroups of the stopped axis are dispatched; -1: both axes.
uint32_t flatGrid(const CgArgs &a, int only = -1) {
    if (only >= 0) return std::max(a.sys[only].sblocks, 1u);
    if (a.xcdSplit == 1) return 8u * ((std::max(a.sys[0].sblocks, a.sys[1].sblocks) + 3u) / 4u);
    return a.sys[0].sblocks + a.sys[1].sblocks;
}

// E: entries per block (the cut spacing); U: entries per thread and round;
// G2: lanes per row in the row sums; NT: threads per block.  Smaller blocks
// let more of them share a CU, so one block's gathers overlap another's row
// sums and reduction.
// MODE 0: Ap for the CG iteration (partial p.Ap).  MODE 1 and 2 (T = double)
// take the FP64 residual b - A x64 of the same rows instead: 1 refreshes the
// FP32 residual of the axes that asked for it (partials r64.r64, r.z), 2 is
// the true residual check of both axes (partials r.r, b.b); parity is -1.
// MODE 3 and 4 (T = double) start the solve as cgStart does: r = b - A x64,
// rounded to the working precision (3: FP32, 4: FP64), p = z = r / d, the
// FP32 correction cleared (3); partials b.b, r.z, r.r.
// W4: phase 1 reads four consecutive entries per thread with one wide load
// per array (16-byte values, 8 + 4 bytes of 3-byte columns), from the block's
// first entry rounded down to a multiple of four; entries outside the block
// are loaded and dropped.  A wave then issues three loads per 256 entries
// instead of twelve per 64, which relieves the vector-memory address path
// that the gathers share.  The arrays are padded to a multiple of four
// entries.  Products and their order are unchanged.
template <typename T, uint32_t E, int U, int G2, int NT, bool C24 = false, int MODE = 0, bool W4 = false>
__global__ void __launch_bounds__(NT) cgSpmvFlat(CgArgs a, int parity) {
    static_assert(MODE == 0 || MODE == 6 || sizeof(T) == 8, "the residual modes run in FP64");
    static_assert(MODE != 6 || sizeof(T) == 4, "the fused iteration runs in FP32");
    // MODE 6: Ap like MODE 0, with the partials p.Ap, r.Ap, Ap.Ap and r.r (cgFUpdate)
    constexpr int K = MODE == 0 ? 1 : (MODE <= 2 ? 2 : 3);
    constexpr uint32_t kFlatEntries = E, kFlatSlack = flatSlack(E);
    // W4 keeps product k of the quads at slot k (the block's entry j at
    // j + lead), so each quad is two aligned 16-byte LDS stores
    __shared__ alignas(16) double prod[kFlatEntries + kFlatSlack + (W4 ? 8 : 0)];
    int axis;
    uint32_t b;
    if (!decodeFlat(a, axis, b, parity)) return;
    if constexpr (MODE == 0 || MODE == 6) {
        if (!a.ctl[axis].s[parity].live) return;
    } else if constexpr (MODE == 1) {
        if (!a.ctl[axis].needUpdate) return;
    } else if constexpr (MODE >= 3) {
        if (a.ctl[axis].frozen) return;
    }
    const SysView s = a.sys[axis];
    const T *__restrict__ p = MODE == 0 || MODE == 6 ? static_cast<const T *>(a.p[axis])
                                                     : reinterpret_cast<const T *>(a.x64[axis]);
    const T *__restrict__ rres = static_cast<const T *>(a.r[axis]);  // MODE 6: the residual, for r.Ap
    T *__restrict__ ap = static_cast<T *>(a.ap[axis]);
    const T *__restrict__ val = Mat<T>::val(s);
    const T *__restrict__ diag = Mat<T>::diag(s);
    // the scaled FP32 system (mixed precision): unit diagonal, not read
    const bool unit = (MODE == 0 || MODE == 6) && sizeof(T) == 4 && a.unitDiag;
    const uint4 info = s.flatInfo[b];
    const uint32_t R0 = info.x, R1 = info.y, E0 = info.z, count = info.w - info.z;
    double acc[K] = {};
    // row r's product (A v)_r = sum, with pd = v_r (and rr = r_r for MODE 6)
    const auto finish = [&](uint32_t r, double pd, double sum, double rr) {
        if constexpr (MODE == 0) {
            const T v = static_cast<T>(sum);
            ap[r] = v;
            acc[0] += pd * static_cast<double>(v);
        } else if constexpr (MODE == 6) {
            const T v = static_cast<T>(sum);
            ap[r] = v;
            const double vd = static_cast<double>(v);
            acc[0] += pd * vd;
            acc[1] += rr * vd;
            acc[2] += vd * vd;
        } else {
            const double rhs = s.rhs[r];
            const double ri = rhs - sum;
            if constexpr (MODE >= 3) {
                using W = std::conditional_t<MODE == 3, float, double>;
                const bool scaled = MODE == 3 && a.unitDiag;  // r^ = D^-1/2 r, z = r^
                const W rt = static_cast<W>(scaled ? ri * s.isd[r] : ri);
                const W zt = scaled ? rt : rt / Mat<W>::diag(s)[r];
                static_cast<W *>(a.r[axis])[r] = rt;
                static_cast<W *>(a.p[axis])[r] = zt;
                if constexpr (MODE == 3) static_cast<float *>(a.e[axis])[r] = 0.0f;
                acc[0] += rhs * rhs;
                acc[1] += static_cast<double>(rt) * static_cast<double>(zt);
                acc[2] += ri * ri;
                return;
            }
            acc[0] += ri * ri;
            if constexpr (MODE == 1) {
                const float rf = static_cast<float>(a.unitDiag ? ri * s.isd[r] : ri);
                static_cast<float *>(a.r[axis])[r] = rf;
                acc[1] += static_cast<double>(rf) *
                          static_cast<double>(a.unitDiag ? rf : rf / s.diag32[r]);
            } else {
                acc[1] += rhs * rhs;
            }
        }
    };
    if (count > kFlatEntries + kFlatSlack) {
        constexpr int G = 4, R = 2;
        const uint32_t lane = threadIdx.x % G, grp = threadIdx.x / G;
        for (uint32_t r0 = R0; r0 < R1; r0 += R * (NT / G)) {
            uint32_t row[R];
            bool valid[R];
            double off[R];
#pragma unroll
            for (int q = 0; q < R; ++q) {
                row[q] = r0 + grp + q * (NT / G);
                valid[q] = row[q] < R1;
            }
            rowsDot<T, G, R>(s.rowPtr, s.col, val, p, row, valid, lane, off);
#pragma unroll
            for (int q = 0; q < R; ++q) {
                if (valid[q] && lane == 0) {
                    const double pd = static_cast<double>(p[row[q]]);
                    finish(row[q], pd, (unit ? pd : static_cast<double>(diag[row[q]]) * pd) + off[q],
                           MODE == 6 ? static_cast<double>(rres[row[q]]) : 0.0);
                }
            }
        }
    } else {
        // G2 lanes per row in phase 2 (1: the reference's order).  The first
        // pass's row data is loaded before phase 1, so it arrives while the
        // products are gathered (a block of E entries usually has fewer than
        // NT / G2 rows).  Row 0 stands in for a missing row (rows > 0 here).
        constexpr uint32_t kGroups = NT / G2;
        const uint32_t lane = threadIdx.x % G2, grp = threadIdx.x / G2;
        // W4 also loads the row data of G2 passes (G2 * NT / G2 rows) with one
        // load per array and wave: lane l of wave w holds row
        // R0 + chunk + kGroups (l / kWaveGroups) + kWaveGroups w + l % kWaveGroups,
        // which pass l / kWaveGroups gives to group l % kWaveGroups of the
        // wave by a shuffle.  Every row keeps its lane group, so the sums and
        // their order are those of one load per row and pass.
        constexpr uint32_t kWaveGroups = 64 / G2;
        const uint32_t lane64 = threadIdx.x & 63u, wave = threadIdx.x >> 6;
        uint32_t pre0, pre1;
        T preP, preD, preR = T(0);
        const auto loadRows = [&](uint32_t r) {
            const uint32_t rq = r < R1 ? r : 0u;
            pre0 = s.rowPtr[rq];
            pre1 = s.rowPtr[rq + 1];
            preP = p[rq];
            preD = unit ? T(1) : diag[rq];
            if constexpr (MODE == 6) preR = rres[rq];
        };
        const auto chunkRow = [&](uint32_t chunk) {
            return R0 + chunk * (kGroups * G2) + kGroups * (lane64 / kWaveGroups) + kWaveGroups * wave +
                   lane64 % kWaveGroups;
        };
        if constexpr (W4) loadRows(chunkRow(0));
        else loadRows(R0 + grp);
        __builtin_amdgcn_sched_barrier(0);
        if constexpr (W4) {
            // U / 4 quads per thread and round; quad q covers entries
            // A + 4q .. A + 4q + 3, A = E0 rounded down to a multiple of 4
            constexpr int Q = U / 4 > 0 ? U / 4 : 1;
            const uint32_t A = E0 & ~3u, lead = E0 - A, quads = (lead + count + 3) / 4;
            for (uint32_t qb = 0; qb < quads; qb += Q * NT) {
                uint32_t c[Q][4];
                T w[Q][4];
#pragma unroll
                for (int u = 0; u < Q; ++u) {
                    const uint32_t q = qb + u * NT + threadIdx.x;
                    loadQuad<T, C24>(s, val, A + 4u * (q < quads ? q : 0u), c[u], w[u]);
                }
                __builtin_amdgcn_sched_barrier(0);
                T g[Q][4];
#pragma unroll
                for (int u = 0; u < Q; ++u) {
                    const uint32_t q = qb + u * NT + threadIdx.x;
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const uint32_t k = 4u * q + j;  // entry A + k, LDS slot k
                        const bool ok = q < quads && k >= lead && k < lead + count;
                        g[u][j] = gatherAt(p, ok ? c[u][j] : kNoCol);
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int u = 0; u < Q; ++u) {
                    // slots outside the block's entries get the dropped
                    // entries' products, which no row reads
                    const uint32_t q = qb + u * NT + threadIdx.x;
                    if (q < quads) {
                        double2 *dst = reinterpret_cast<double2 *>(prod + 4u * q);
                        dst[0] = make_double2(static_cast<double>(w[u][0]) * static_cast<double>(g[u][0]),
                                              static_cast<double>(w[u][1]) * static_cast<double>(g[u][1]));
                        dst[1] = make_double2(static_cast<double>(w[u][2]) * static_cast<double>(g[u][2]),
                                              static_cast<double>(w[u][3]) * static_cast<double>(g[u][3]));
                    }
                }
            }
        } else
        for (uint32_t base = 0; base < count; base += U * NT) {
            uint32_t c[U];
            T w[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t k = base + u * NT + threadIdx.x;
                const bool ok = k < count;
                const uint32_t kk = E0 + (ok ? k : 0u);  // count > 0 here: entry E0 exists
                uint32_t cl;
                if constexpr (C24) cl = static_cast<uint32_t>(s.colLo[kk]) | (static_cast<uint32_t>(s.colHi[kk]) << 16);
                else cl = s.col[kk];
                const T wl = val[kk];
                c[u] = ok ? cl : kNoCol;
                w[u] = ok ? wl : T(0);
            }
            __builtin_amdgcn_sched_barrier(0);
            T g[U];
#pragma unroll
            for (int u = 0; u < U; ++u) g[u] = gatherAt(p, c[u]);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t k = base + u * NT + threadIdx.x;
                if (k < count) prod[k] = static_cast<double>(w[u]) * static_cast<double>(g[u]);
            }
        }
        __syncthreads();
        // the rows r0 + grp of one pass, with their row data
        const uint32_t slot0 = W4 ? (E0 & ~3u) : E0;  // LDS slot of entry e: e - slot0
        const auto pass = [&](uint32_t r0, uint32_t q0, uint32_t q1, T pr, T dr, T rr) {
            const uint32_t r = r0 + grp;
            const bool valid = r < R1;
            const uint32_t e0 = q0 - slot0, e1 = valid ? q1 - slot0 : e0;
            const double pd = static_cast<double>(pr);
            if constexpr (G2 == 1) {
                // d p first, then the row's products in entry order
                // (SparseSystem::multiply, simpl_lss.h:49-57)
                double sum = static_cast<double>(dr) * pd;
                for (uint32_t k = e0; k < e1; ++k) sum += prod[k];
                if (valid) finish(r, pd, sum, static_cast<double>(rr));
            } else {
                // lane l adds products l, l + G2, ... of the row, then a fixed
                // butterfly; shorter waits on long rows, another order
                double part = 0.0;
                for (uint32_t k = e0 + lane; k < e1; k += G2) part += prod[k];
#pragma unroll
                for (int off = G2 / 2; off > 0; off >>= 1) part += __shfl_xor(part, off, G2);
                if (valid && lane == 0) finish(r, pd, static_cast<double>(dr) * pd + part, static_cast<double>(rr));
            }
        };
        if constexpr (W4) {
            uint32_t r0 = R0;
            for (uint32_t chunk = 0; r0 < R1; ++chunk) {
                if (chunk > 0) loadRows(chunkRow(chunk));
#pragma unroll
                for (uint32_t k = 0; k < G2; ++k, r0 += kGroups) {
                    if (r0 >= R1) break;
                    const int src = static_cast<int>(k * kWaveGroups + lane64 / G2);
                    pass(r0, __shfl(pre0, src), __shfl(pre1, src), __shfl(preP, src), __shfl(preD, src),
                         MODE == 6 ? __shfl(preR, src) : T(0));
                }
            }
        } else {
            if (R0 < R1) pass(R0, pre0, pre1, preP, preD, preR);
            for (uint32_t r0 = R0 + kGroups; r0 < R1; r0 += kGroups) {
                loadRows(r0 + grp);
                pass(r0, pre0, pre1, preP, preD, preR);
            }
        }
    }
    blockSum<K, NT>(acc);
    writePartial<K>(MODE == 0 || MODE == 6 ? a.partA : (MODE <= 2 ? a.partB : a.partS), a.stride, axis, b, acc);
}

// Cut c (c = 1 .. blocks - 1) = first row whose entries start at or after
// c * entries; cut 0 = 0, cut blocks = rows.
__device__ __forceinline__ uint32_t flatCut(const uint32_t *__restrict__ rowPtr, uint32_t rows, uint32_t blocks,
                                            uint32_t entries, uint32_t c) {
    if (c == 0) return 0u;
    if (c >= blocks) return rows;
    const uint64_t target = static_cast<uint64_t>(c) * entries;
    uint32_t lo = 0, hi = rows;  // first r in [0, rows] with rowPtr[r] >= target
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (rowPtr[mid] < target) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// info[b] = (first row, end row, first entry, end entry) of entry block b:
// the block reads one record instead of a cut and then its row pointer.
__global__ void __launch_bounds__(256) flatBounds(const uint32_t *__restrict__ rowPtr, uint32_t rows,
                                                  uint32_t blocks, uint32_t entries, uint4 *__restrict__ info) {
    const uint32_t b = blockIdx.x * 256 + threadIdx.x;
    if (b >= blocks) return;
    const uint32_t r0 = flatCut(rowPtr, rows, blocks, entries, b);
    const uint32_t r1 = flatCut(rowPtr, rows, blocks, entries, b + 1);
    info[b] = make_uint4(r0, r1, rowPtr[r0], rowPtr[r1]);
}

// alpha = rho / p.Ap; r -= alpha Ap; tile partials of r.r, r.z.  The iterate
// update x += alpha p is left to cgPUpdate, which reads p anyway: alpha goes
// there through the axis' control block (the same double, so x is the same).
// PER = rows per thread when the tile is PER * 1024 rows: the thread's vector
// entries are loaded before the tile partials are summed (they do not depend
// on alpha), so their latency hides behind that sum.  PER = 0: any tile, the
// entries loaded after the sum.
template <typename T, int PER>
__global__ void __launch_bounds__(kBlock) cgUpdate(CgArgs a, int parity) {
    int axis;
    uint32_t tile;
    if (!decode(a, axis, tile, parity)) return;
    const CgState st = a.ctl[axis].s[parity];
    if (!st.live) return;
    const SysView s = a.sys[axis];
    T *__restrict__ r = static_cast<T *>(a.r[axis]);
    const T *__restrict__ ap = static_cast<const T *>(a.ap[axis]);
    const T *__restrict__ d = Mat<T>::diag(s);
    const bool unit = sizeof(T) == 4 && a.unitDiag;
    const uint32_t begin = tile * a.tile, end = min(begin + a.tile, s.rows);
    // the partials are loaded first: the wait for them (an in-order count of
    // outstanding loads) then leaves the vector loads below in flight
    PartialShare<1, 8> share;  // the entry-block SpMV leaves several thousand partials per axis
    share.issue(a.partA, s.sblocks, a.stride, axis);
    __builtin_amdgcn_sched_barrier(0);
    constexpr int kPre = PER > 0 ? PER : 1;
    T rv[kPre], av[kPre], dv[kPre];
    if constexpr (PER > 0) {
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const uint32_t i = begin + threadIdx.x + k * kBlock;
            const uint32_t j = i < end ? i : begin;  // unconditional loads; a tile has a row
            rv[k] = r[j];
            av[k] = ap[j];
            dv[k] = unit ? T(1) : d[j];
        }
    }
    __builtin_amdgcn_sched_barrier(0);
    double pap[1];
    share.finish(pap);
    const bool lead = tile == 0 && threadIdx.x == 0;
    if (!(pap[0] > 0.0)) {
        if (lead) a.ctl[axis].brk[parity] = 1;
        return;
    }
    const double alpha = st.rho / pap[0];
    if (lead) {
        a.ctl[axis].brk[parity] = 0;
        a.ctl[axis].alpha[parity] = alpha;
    }
    const T alphaT = static_cast<T>(alpha);
    double acc[2] = {0.0, 0.0};
    const auto step = [&](uint32_t i, T ri, T ai, T di) {
        const T rn = ri - alphaT * ai;
        r[i] = rn;
        const double rd = static_cast<double>(rn);
        if (unit) {  // r^ = D^-1/2 r, r.z = r^.r^ (the residual is judged from it, cgPUpdate)
            acc[1] += rd * rd;
        } else {
            acc[0] += rd * rd;
            acc[1] += rd * static_cast<double>(rn / di);
        }
    };
    if constexpr (PER > 0) {
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const uint32_t i = begin + threadIdx.x + k * kBlock;
            if (i < end) step(i, rv[k], av[k], dv[k]);
        }
    } else {
        for (uint32_t i = begin + threadIdx.x; i < end; i += kBlock) step(i, r[i], ap[i], unit ? T(1) : d[i]);
    }
    blockSum<2>(acc);
    writePartial<2>(a.partB, a.stride, axis, tile, acc);
}

// x += alpha p (alpha from cgUpdate); beta = r.z' / r.z; p = z + beta p;
// block 0 publishes state[parity ^ 1].  PER as in cgUpdate: x, r, d and p are
// loaded before the tile partials are summed.
template <typename T, int PER>
__global__ void __launch_bounds__(kBlock) cgPUpdate(CgArgs a, int parity) {
    int axis;
    uint32_t tile;
    // grid block 0 carries every stopped axis' state forward (a stopped axis
    // may have no block of its own in this launch)
    if (blockIdx.x == 0 && threadIdx.x == 0) {
#pragma unroll
        for (int ax = 0; ax < 2; ++ax)
            if (!a.ctl[ax].s[parity].live) a.ctl[ax].s[parity ^ 1] = a.ctl[ax].s[parity];
    }
    if (!decode(a, axis, tile, parity)) return;
    AxisCtl &c = a.ctl[axis];
    const CgState st = c.s[parity];
    const bool lead = tile == 0 && threadIdx.x == 0;
    if (!st.live) return;
    const SysView s = a.sys[axis];
    const size_t hist = static_cast<size_t>(axis) * a.histCap + st.steps;
    if (c.brk[parity]) {  // p.Ap <= 0: the reference stops the axis (simpl_lss.cc:242-244)
        if (lead) {
            CgState nx = st;
            nx.active = 0;
            nx.live = 0;
            nx.steps = st.steps + 1;
            c.s[parity ^ 1] = nx;
            if (st.steps < a.histCap) a.history[hist] = st.resid;
        }
        return;
    }
    T *__restrict__ x = static_cast<T *>(a.e[axis] ? a.e[axis] : static_cast<void *>(a.x64[axis]));
    T *__restrict__ p = static_cast<T *>(a.p[axis]);
    const T *__restrict__ r = static_cast<const T *>(a.r[axis]);
    const T *__restrict__ d = Mat<T>::diag(s);
    const bool unit = sizeof(T) == 4 && a.unitDiag;  // z = r^ on the scaled system: d not read
    const uint32_t begin = tile * a.tile, end = min(begin + a.tile, s.rows);
    PartialShare<2> share;
    share.issue(a.partB, s.blocks, a.stride, axis);
    const T alphaT = static_cast<T>(c.alpha[parity]);
    __builtin_amdgcn_sched_barrier(0);
    constexpr int kPre = PER > 0 ? PER : 1;
    T xv[kPre], rv[kPre], dv[kPre], pv[kPre];
    if constexpr (PER > 0) {
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const uint32_t i = begin + threadIdx.x + k * kBlock;
            const uint32_t j = i < end ? i : begin;
            xv[k] = x[j];
            rv[k] = r[j];
            dv[k] = unit ? T(1) : d[j];
            pv[k] = p[j];
        }
    }
    __builtin_amdgcn_sched_barrier(0);
    double tot[2];
    share.finish(tot);
    // on the scaled FP32 system the norm of r = D^1/2 r^ would need d in the
    // residual update; the iterations judge the unscaled residual from r^.r^
    // instead, scaled from the last refresh's FP64 residual: the refresh point
    // is where r^.r^ fell by the square of the refresh factor.  The axis still
    // stops only on an FP64 residual.
    const double resid = unit ? (c.refRz > 0.0 ? c.refResid * sqrt(tot[1] / c.refRz) : 0.0)
                              : sqrt(tot[0]) / c.bNorm;
    const double rhoNext = tot[1];
    const bool ok = rhoNext > 0.0;
    // the iterate steps whatever beta turns out to be (the reference updates x
    // before it looks at r.z, simpl_lss.cc:247,253); p only when the axis goes on
    const T beta = static_cast<T>(ok ? rhoNext / st.rho : 0.0);
    if constexpr (PER > 0) {
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const uint32_t i = begin + threadIdx.x + k * kBlock;
            if (i < end) {
                x[i] = xv[k] + alphaT * pv[k];
                if (ok) p[i] = (unit ? rv[k] : rv[k] / dv[k]) + beta * pv[k];
            }
        }
    } else {
        for (uint32_t i = begin + threadIdx.x; i < end; i += kBlock) {
            const T pi = p[i];
            x[i] = x[i] + alphaT * pi;
            if (ok) p[i] = (unit ? r[i] : r[i] / d[i]) + beta * pi;
        }
    }
    if (lead) {
        // mixed precision: ask for an FP64 refresh once the inner residual fell
        // by the refresh factor, or reached the tolerance.  The axis pauses
        // (live = 0) until the refresh, which sets live again from the true
        // residual: FP32 steps past the refresh point would push the recurrence
        // residual, and with it p, far below the true residual, and the first
        // step after the refresh (alpha = r.z / p.Ap with the refreshed r and
        // that p) would then be far too long.
        const bool refresh = a.e[axis] && (resid <= c.refreshFactor * c.refResid || resid <= c.tol);
        CgState nx;
        nx.steps = st.steps + 1;
        nx.resid = resid;
        nx.rho = ok ? rhoNext : st.rho;
        nx.active = ok ? 1 : 0;
        nx.live = (ok && resid > c.tol && nx.steps < c.maxIter && !refresh) ? 1 : 0;
        nx.pad = 0;
        c.s[parity ^ 1] = nx;
        if (st.steps < a.histCap) a.history[hist] = resid;
        if (refresh) c.needUpdate = 1;
    }
}

// ---------------------------------------------------------------- fused vector update
// The FP32 iteration on the scaled system in two kernels: cgSpmvFlat MODE 6
// (Ap and the tile partials of p.Ap, r.Ap, Ap.Ap), then this kernel, which does
// cgUpdate's and cgPUpdate's work in one pass: alpha = rho / p.Ap; r' = r -
// alpha Ap; e += alpha p; rho' = r'.r' = rho - 2 alpha r.Ap + alpha^2 Ap.Ap
// (the new residual's norm from the SpMV's partials, so no second grid-wide
// reduction is needed before p is updated); beta = rho' / rho; p = r' + beta p.
// It also sums r'.r' per tile into partU: the next iteration's rho is that
// direct sum, so the expansion's rounding never carries from one step to the
// next.  The same Jacobi-PCG iterates as the three-kernel iteration up to the
// rounding of beta's numerator; the residual is judged from rho' as cgPUpdate does on
// the scaled system, and the rules (p.Ap <= 0, rho' <= 0, refresh, stop) are
// cgUpdate's and cgPUpdate's.  Block 0 of each axis publishes state[parity ^ 1].
template <int PER>
__global__ void __launch_bounds__(kBlock) cgFUpdate(CgArgs a, int parity) {
    using T = float;
    int axis;
    uint32_t tile;
    // grid block 0 carries every stopped axis' state forward (a stopped axis
    // may have no block of its own in this launch)
    if (blockIdx.x == 0 && threadIdx.x == 0) {
#pragma unroll
        for (int ax = 0; ax < 2; ++ax)
            if (!a.ctl[ax].s[parity].live) a.ctl[ax].s[parity ^ 1] = a.ctl[ax].s[parity];
    }
    if (!decode(a, axis, tile, parity)) return;
    AxisCtl &c = a.ctl[axis];
    const CgState st = c.s[parity];
    if (!st.live) return;
    const SysView sv = a.sys[axis];
    const uint32_t begin = tile * a.tile, end = min(begin + a.tile, sv.rows);
    T *__restrict__ r = static_cast<T *>(a.r[axis]);
    const T *__restrict__ ap = static_cast<const T *>(a.ap[axis]);
    T *__restrict__ p = static_cast<T *>(a.p[axis]);
    T *__restrict__ e = static_cast<T *>(a.e[axis]);
    PartialShare<3, 8> share;  // several thousand SpMV blocks per axis
    share.issue(a.partA, sv.sblocks, a.stride, axis);
    // rho = r.r of the stored residual: st.pad > 0 says the last update summed
    // it over st.pad tiles into partU[parity]; else it is st.rho (the start or
    // a refresh computed it)
    PartialShare<1, 1> shareR;
    shareR.issue(a.partU + static_cast<size_t>(parity) * 2 * a.uStride, st.pad, a.uStride, axis);
    __builtin_amdgcn_sched_barrier(0);
    constexpr int kPre = PER > 0 ? PER : 1;
    T rv[kPre], av[kPre], pv[kPre], ev[kPre];
    if constexpr (PER > 0) {
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const uint32_t i = begin + threadIdx.x + k * kBlock;
            const uint32_t j = i < end ? i : begin;  // unconditional loads; a tile has a row
            rv[k] = r[j];
            av[k] = ap[j];
            pv[k] = p[j];
            ev[k] = e[j];
        }
    }
    __builtin_amdgcn_sched_barrier(0);
    double tot[3], rr[1];
    share.finish(tot);
    shareR.finish(rr);
    // each step starts from the residual's own norm; only beta's numerator and
    // the judged residual come from the expansion
    const double pap = tot[0], rap = tot[1], apap = tot[2], rho = st.pad > 0 ? rr[0] : st.rho;
    const bool lead = tile == 0 && threadIdx.x == 0;
    const size_t hist = static_cast<size_t>(axis) * a.histCap + st.steps;
    if (!(pap > 0.0)) {  // p.Ap <= 0: the reference stops the axis (simpl_lss.cc:242-244)
        if (lead) {
            CgState nx = st;
            nx.active = 0;
            nx.live = 0;
            nx.steps = st.steps + 1;
            c.s[parity ^ 1] = nx;
            if (st.steps < a.histCap) a.history[hist] = st.resid;
        }
        return;
    }
    const double alpha = rho / pap;
    const double rhoNext = rho - 2.0 * alpha * rap + alpha * alpha * apap;
    const bool ok = rhoNext > 0.0;
    const double resid = c.refRz > 0.0 && ok ? c.refResid * sqrt(rhoNext / c.refRz) : 0.0;
    const T alphaT = static_cast<T>(alpha);
    const T beta = static_cast<T>(ok ? rhoNext / rho : 0.0);
    double rn2[1] = {0.0};
    const auto step = [&](uint32_t i, T ri, T ai, T pi, T ei) {
        const T rn = ri - alphaT * ai;
        r[i] = rn;
        e[i] = ei + alphaT * pi;
        if (ok) p[i] = rn + beta * pi;
        rn2[0] += static_cast<double>(rn) * static_cast<double>(rn);
    };
    if constexpr (PER > 0) {
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const uint32_t i = begin + threadIdx.x + k * kBlock;
            if (i < end) step(i, rv[k], av[k], pv[k], ev[k]);
        }
    } else {
        for (uint32_t i = begin + threadIdx.x; i < end; i += kBlock) step(i, r[i], ap[i], p[i], e[i]);
    }
    // this tile's r'.r', for the next iteration's rho
    const uint32_t tiles = (sv.rows + a.tile - 1) / a.tile;
    const bool share2 = tiles <= a.uStride;
    blockSum<1>(rn2);
    if (share2) writePartial<1>(a.partU + static_cast<size_t>(parity ^ 1) * 2 * a.uStride, a.uStride, axis, tile, rn2);
    if (lead) {
        // the refresh rule and the pause as in cgPUpdate
        const bool refresh = resid <= c.refreshFactor * c.refResid || resid <= c.tol;
        CgState nx;
        nx.steps = st.steps + 1;
        nx.resid = resid;
        nx.rho = ok ? rhoNext : rho;
        nx.active = ok ? 1 : 0;
        nx.live = (ok && resid > c.tol && nx.steps < c.maxIter && !refresh) ? 1 : 0;
        nx.pad = share2 ? static_cast<int32_t>(tiles) : 0;
        c.s[parity ^ 1] = nx;
        if (st.steps < a.histCap) a.history[hist] = resid;
        if (refresh) c.needUpdate = 1;
    }
}

// ---------------------------------------------------------------- FP64 refresh (mixed precision)
// x64 += e (FP32 correction; on the scaled system x64 += D^-1/2 e), e = 0, for
// axes that asked for a refresh.
__global__ void __launch_bounds__(kBlock) cgFlush(CgArgs a) {
    int axis;
    uint32_t tile;
    if (!decode(a, axis, tile)) return;
    if (!a.ctl[axis].needUpdate) return;
    const SysView s = a.sys[axis];
    double *__restrict__ x = a.x64[axis];
    float *__restrict__ e = static_cast<float *>(a.e[axis]);
    const uint32_t begin = tile * a.tile, end = min(begin + a.tile, s.rows);
    for (uint32_t i = begin + threadIdx.x; i < end; i += kBlock) {
        const double ei = static_cast<double>(e[i]);
        x[i] += a.unitDiag ? ei * s.isd[i] : ei;
        e[i] = 0.0f;
    }
}

// r64 = b - A x64 (FP64), r = (float) r64 (scaled system: (float) D^-1/2 r64);
// partials r64.r64 and r.z.
template <int G>
__global__ void __launch_bounds__(kBlock) cgRefresh(CgArgs a) {
    int axis;
    uint32_t tile;
    if (!decode(a, axis, tile)) return;
    if (!a.ctl[axis].needUpdate) return;
    const SysView s = a.sys[axis];
    const double *__restrict__ x = a.x64[axis];
    float *__restrict__ r = static_cast<float *>(a.r[axis]);
    const uint32_t lane = threadIdx.x % G, grp = threadIdx.x / G;
    const uint32_t begin = tile * a.tile, end = min(begin + a.tile, s.rows);
    double acc[2] = {0.0, 0.0};
    constexpr uint32_t kPass = kRowsPerGroup * (kBlock / G);
    for (uint32_t r0 = begin; r0 < end; r0 += kPass) {
        uint32_t row[kRowsPerGroup];
        bool valid[kRowsPerGroup];
        double off[kRowsPerGroup];
#pragma unroll
        for (int q = 0; q < kRowsPerGroup; ++q) {
            row[q] = r0 + grp + q * (kBlock / G);
            valid[q] = row[q] < end;
        }
        rowsDot<double, G, kRowsPerGroup>(s.rowPtr, s.col, s.val, x, row, valid, lane, off);
#pragma unroll
        for (int q = 0; q < kRowsPerGroup; ++q) {
            const uint32_t i = row[q];
            if (valid[q] && lane == 0) {
                const double ri = s.rhs[i] - (s.diag[i] * x[i] + off[q]);
                const float rf = static_cast<float>(a.unitDiag ? ri * s.isd[i] : ri);
                r[i] = rf;
                acc[0] += ri * ri;
                acc[1] += static_cast<double>(rf) * static_cast<double>(a.unitDiag ? rf : rf / s.diag32[i]);
            }
        }
    }
    blockSum<2>(acc);
    writePartial<2>(a.partB, a.stride, axis, tile, acc);
}

// One block per axis: the true residual becomes the state; rho restarts from
// the refreshed r (the search direction p is kept: a reliable update).
__global__ void __launch_bounds__(kBlock) cgRefreshFinish(CgArgs a) {
    const int axis = blockIdx.x;
    AxisCtl &c = a.ctl[axis];
    if (!c.needUpdate) return;
    double tot[2];
    sumPartials<2>(a.partB, a.flatResidual ? a.sys[axis].sblocks : a.sys[axis].blocks, a.stride, axis, tot);
    if (threadIdx.x == 0) {
        CgState &st = c.s[0];
        const double resid = c.bNorm > 0.0 ? sqrt(tot[0]) / c.bNorm : 0.0;
        st.resid = resid;
        c.trueResid = resid;
        c.refResid = resid;
        c.refRz = tot[1];
        ++c.refreshes;
        if (st.active && tot[1] > 0.0) st.rho = tot[1];
        st.pad = 0;  // rho is this kernel's, not the update partials'
        st.live = (st.active && resid > c.tol && st.steps < c.maxIter) ? 1 : 0;
        c.needUpdate = 0;
    }
}

// ---------------------------------------------------------------- true residual
template <int G>
__global__ void __launch_bounds__(kBlock) cgTrueResidual(CgArgs a) {
    int axis;
    uint32_t tile;
    if (!decode(a, axis, tile)) return;
    const SysView s = a.sys[axis];
    const double *__restrict__ x = a.x64[axis];
    const uint32_t lane = threadIdx.x % G, grp = threadIdx.x / G;
    const uint32_t begin = tile * a.tile, end = min(begin + a.tile, s.rows);
    double acc[2] = {0.0, 0.0};
    constexpr uint32_t kPass = kRowsPerGroup * (kBlock / G);
    for (uint32_t r0 = begin; r0 < end; r0 += kPass) {
        uint32_t row[kRowsPerGroup];
        bool valid[kRowsPerGroup];
        double off[kRowsPerGroup];
#pragma unroll
        for (int q = 0; q < kRowsPerGroup; ++q) {
            row[q] = r0 + grp + q * (kBlock / G);
            valid[q] = row[q] < end;
        }
        rowsDot<double, G, kRowsPerGroup>(s.rowPtr, s.col, s.val, x, row, valid, lane, off);
#pragma unroll
        for (int q = 0; q < kRowsPerGroup; ++q) {
            const uint32_t i = row[q];
            if (valid[q] && lane == 0) {
                const double ri = s.rhs[i] - (s.diag[i] * x[i] + off[q]);
                acc[0] += ri * ri;
                acc[1] += s.rhs[i] * s.rhs[i];
            }
        }
    }
    blockSum<2>(acc);
    writePartial<2>(a.partB, a.stride, axis, tile, acc);
}

__global__ void __launch_bounds__(kBlock) cgTrueResidualFinish(CgArgs a) {
    const int axis = blockIdx.x;
    double tot[2];
    sumPartials<2>(a.partB, a.flatResidual ? a.sys[axis].sblocks : a.sys[axis].blocks, a.stride, axis, tot);
    if (threadIdx.x == 0) {
        const double bn = sqrt(tot[1]);
        a.ctl[axis].trueResid = bn > 0.0 ? sqrt(tot[0]) / bn : 0.0;
    }
}

// Column indices as a 16-bit low half and an 8-bit high byte (rows < 2^24):
// 3 bytes per entry read by the entry-block SpMV instead of 4.
__global__ void __launch_bounds__(kBlock)
splitColumns(const uint32_t *__restrict__ col, uint16_t *__restrict__ lo, uint8_t *__restrict__ hi, uint64_t n) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * kBlock) {
        const uint32_t c = col[i];
        lo[i] = static_cast<uint16_t>(c & 0xffffu);
        hi[i] = static_cast<uint8_t>(c >> 16);
    }
}

// Columns (all below 2^24) packed four to three words: quad q's columns
// c0 .. c3 become c0 | c1 << 24, c1 >> 8 | c2 << 16, c2 >> 16 | c3 << 8.
// Entries past n pack as 0.
__global__ void __launch_bounds__(kBlock)
packColumns(const uint32_t *__restrict__ col, uint32_t *__restrict__ pk, uint64_t n) {
    const uint64_t quads = (n + 3) / 4;
    for (uint64_t q = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; q < quads;
         q += static_cast<uint64_t>(gridDim.x) * kBlock) {
        uint32_t c[4];
        for (int j = 0; j < 4; ++j) c[j] = 4 * q + j < n ? col[4 * q + j] : 0u;
        pk[3 * q] = c[0] | (c[1] << 24);
        pk[3 * q + 1] = (c[1] >> 8) | (c[2] << 16);
        pk[3 * q + 2] = (c[2] >> 16) | (c[3] << 8);
    }
}

// FP32 copies of the matrix values and diagonal (mixed precision).
__global__ void __launch_bounds__(kBlock)
toFloat(const double *__restrict__ in, float *__restrict__ out, uint64_t n) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * kBlock)
        out[i] = static_cast<float>(in[i]);
}

// 1 / sqrt(d_i) per row, for the scaled FP32 system below.
__global__ void __launch_bounds__(kBlock)
invSqrtDiag(const double *__restrict__ diag, double *__restrict__ isd, uint32_t rows) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < rows) isd[i] = 1.0 / sqrt(diag[i]);
}

// FP32 copy of the symmetrically scaled system D^-1/2 A D^-1/2 (mixed
// precision): val_ij (s_i s_j) per entry with s = 1 / sqrt(d), four lanes per
// row, each with four entries' loads in flight per round; its diagonal is 1.
// s_i s_j and s_j s_i are the same product, so the scaled matrix stays
// symmetric entry for entry.
__global__ void __launch_bounds__(kBlock)
scaledToFloat(const uint32_t *__restrict__ rowPtr, const uint32_t *__restrict__ col,
              const double *__restrict__ val, const double *__restrict__ isd, float *__restrict__ out,
              uint32_t rows) {
    const uint64_t t = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x;
    const uint32_t lane = static_cast<uint32_t>(t & 3u);
    const uint64_t i = t >> 2;
    if (i >= rows) return;
    const double si = isd[i];
    const uint32_t k0 = rowPtr[i], k1 = rowPtr[i + 1];
    constexpr int kPer = 4;
    for (uint32_t base = k0 + lane; base < k1; base += 4 * kPer) {
        uint32_t c[kPer];
        double v[kPer], sj[kPer];
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const uint32_t k = base + 4 * u;
            const uint32_t kk = k < k1 ? k : k0;  // unconditional loads; the row has entry k0
            c[u] = col[kk];
            v[u] = val[kk];
        }
#pragma unroll
        for (int u = 0; u < kPer; ++u) sj[u] = isd[c[u]];
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const uint32_t k = base + 4 * u;
            if (k < k1) out[k] = static_cast<float>(v[u] * (si * sj[u]));
        }
    }
}

// The same scaled copy over the entry blocks of the SpMV (one block of 256
// threads per run of consecutive entries): column, value and output accesses
// are coalesced, and each entry finds its row by a binary search of the
// block's row pointers, kept in LDS with the rows' scale factors (in memory
// when the block has more rows than the table holds).
constexpr uint32_t kScaleRows = 2048;

__global__ void __launch_bounds__(256)
scaledToFloatFlat(const uint4 *__restrict__ info, const uint32_t *__restrict__ rowPtr,
                  const uint32_t *__restrict__ col, const double *__restrict__ val, const double *__restrict__ isd,
                  float *__restrict__ out) {
    __shared__ uint32_t rp[kScaleRows + 1];
    __shared__ double rs[kScaleRows];
    const uint4 b = info[blockIdx.x];
    const uint32_t R0 = b.x, R1 = b.y, E0 = b.z, E1 = b.w;
    const uint32_t nr = R1 - R0;
    const bool table = nr <= kScaleRows;  // block-uniform
    if (table) {
        for (uint32_t r = threadIdx.x; r <= nr; r += 256) {
            rp[r] = rowPtr[R0 + r];
            if (r < nr) rs[r] = isd[R0 + r];
        }
    }
    __syncthreads();
    for (uint32_t k = E0 + threadIdx.x; k < E1; k += 256) {
        const uint32_t c = col[k];
        const double v = val[k], sj = isd[c];
        // the row of entry k: the last r in [R0, R1) with rowPtr[r] <= k
        uint32_t lo = 0, hi = nr;
        while (hi - lo > 1) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if ((table ? rp[mid] : rowPtr[R0 + mid]) <= k) lo = mid;
            else hi = mid;
        }
        const double si = table ? rs[lo] : isd[R0 + lo];
        out[k] = static_cast<float>(v * (si * sj));
    }
}

// Plain product out = A in (SparseSystem::multiply order: diag first, then
// the row's entries in order), one thread per row.
__global__ void __launch_bounds__(256)
multiplyKernel(SysView s, const double *__restrict__ in, double *__restrict__ out) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= s.rows) return;
    double sum = s.diag[i] * in[i];
    for (uint32_t e = s.rowPtr[i]; e < s.rowPtr[i + 1]; ++e) sum += s.val[e] * in[s.col[e]];
    out[i] = sum;
}

uint32_t gridBlocks(const CgArgs &a, int only = -1) {
    if (only >= 0) return std::max(a.sys[only].blocks, 1u);
    if (a.xcdSplit == 1) return 8u * ((std::max(a.sys[0].blocks, a.sys[1].blocks) + 3u) / 4u);
    return a.sys[0].blocks + a.sys[1].blocks;
}

#define KT_DISPATCH_G(G, KERNEL, GRID, STREAM, ...)                                        \
    switch (G) {                                                                           \
        case 1: KERNEL<1><<<GRID, dim3(kBlock), 0, STREAM>>>(__VA_ARGS__); break;          \
        case 2: KERNEL<2><<<GRID, dim3(kBlock), 0, STREAM>>>(__VA_ARGS__); break;          \
        case 4: KERNEL<4><<<GRID, dim3(kBlock), 0, STREAM>>>(__VA_ARGS__); break;          \
        case 8: KERNEL<8><<<GRID, dim3(kBlock), 0, STREAM>>>(__VA_ARGS__); break;          \
        default: KERNEL<16><<<GRID, dim3(kBlock), 0, STREAM>>>(__VA_ARGS__); break;        \
    }

// The SpMV instance for (G, R, staging).  With attributesOnly it only raises
// the instance's dynamic LDS limit to its window (done before any capture).
template <typename T>
int launchSpmv(kt_ctx *ctx, const CgArgs &a, int G, int parity, hipStream_t st, bool attributesOnly,
               hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr, int only = -1) {
    if (a.layout == 1) {
        if (attributesOnly) return KT_OK;
        const dim3 fg(flatGrid(a, only));
        const bool u8 = a.flatUnroll == 8;
        const bool lanes4 = a.flatRowLanes == 4;
        const int nt = static_cast<int>(a.flatThreads);
        // with events, the dispatch itself records the kernel's start and end
        // (no separate event packets around it)
        const auto run = [&](auto kernel, int threads) -> int {
            if (e0) hipExtLaunchKernelGGL(kernel, fg, dim3(threads), 0, st, e0, e1, 0, a, parity);
            else kernel<<<fg, dim3(threads), 0, st>>>(a, parity);
            KT_HIP(hipGetLastError());
            return KT_OK;
        };
        if (nt == 256) {
            if (a.wide && a.col24) {
                if (u8) return run(cgSpmvFlat<T, 2048, 8, 4, 256, true, 0, true>, 256);
                return run(cgSpmvFlat<T, 2048, 4, 4, 256, true, 0, true>, 256);
            }
            if (a.wide) {
                if (u8) return run(cgSpmvFlat<T, 2048, 8, 4, 256, false, 0, true>, 256);
                return run(cgSpmvFlat<T, 2048, 4, 4, 256, false, 0, true>, 256);
            }
            if (a.col24) {
                if (lanes4) return run(cgSpmvFlat<T, 2048, 4, 4, 256, true>, 256);
                return run(cgSpmvFlat<T, 2048, 4, 1, 256, true>, 256);
            }
            if (u8) {
                if (lanes4) return run(cgSpmvFlat<T, 2048, 8, 4, 256>, 256);
                return run(cgSpmvFlat<T, 2048, 8, 1, 256>, 256);
            }
            if (lanes4) return run(cgSpmvFlat<T, 2048, 4, 4, 256>, 256);
            return run(cgSpmvFlat<T, 2048, 4, 1, 256>, 256);
        }
        if (nt == 512) {
            if (lanes4) return run(cgSpmvFlat<T, 4096, 4, 4, 512>, 512);
            return run(cgSpmvFlat<T, 4096, 4, 1, 512>, 512);
        }
#define KT_FLAT(EE)                                                      \
    if (lanes4) {                                                        \
        if (u8) return run(cgSpmvFlat<T, EE, 8, 4, kBlock>, kBlock);     \
        return run(cgSpmvFlat<T, EE, 4, 4, kBlock>, kBlock);             \
    }                                                                    \
    if (u8) return run(cgSpmvFlat<T, EE, 8, 1, kBlock>, kBlock);         \
    return run(cgSpmvFlat<T, EE, 4, 1, kBlock>, kBlock);
        switch (a.flatEntries) {
            case 4096: KT_FLAT(4096)
            case 16384: KT_FLAT(16384)
            default: KT_FLAT(8192)
        }
#undef KT_FLAT
    }
    const int R = ctx->tune.rows > 0 ? ctx->tune.rows : 2;
    const dim3 grid(gridBlocks(a, only));
    const bool staged = a.halo > 0;
    const uint32_t lds = staged ? static_cast<uint32_t>(windowBytes<T>(a.tile, a.halo)) : 0u;
    const auto go = [&](auto kernel) -> int {
        if (attributesOnly) {
            if (staged)
                KT_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
            return KT_OK;
        }
        kernel<<<grid, dim3(kBlock), lds, st>>>(a, parity);
        KT_HIP(hipGetLastError());
        return KT_OK;
    };
#define KT_SPMV_S(GG, RR) return staged ? go(cgSpmv<T, GG, RR, true>) : go(cgSpmv<T, GG, RR, false>);
#define KT_SPMV_R(GG)                   \
    if (R >= 8) { KT_SPMV_S(GG, 8) }    \
    if (R >= 4) { KT_SPMV_S(GG, 4) }    \
    KT_SPMV_S(GG, 2)
    switch (G) {
        case 1: KT_SPMV_R(1)
        case 2: KT_SPMV_R(2)
        case 4: KT_SPMV_R(4)
        case 8: KT_SPMV_R(8)
        default: KT_SPMV_R(16)
    }
#undef KT_SPMV_R
#undef KT_SPMV_S
}

// One fused FP32 iteration: cgSpmvFlat MODE 6, then cgFUpdate (the shape
// prepare() enables it for: 256-thread blocks of 2048 entries, wide loads,
// 3-byte columns, four row lanes).
int launchIterationFused(kt_ctx *ctx, const CgArgs &a, int parity, hipStream_t st, bool timed, int only) {
    const dim3 fg(flatGrid(a, only));
    const auto spmv = cgSpmvFlat<float, 2048, 4, 4, 256, true, 6, true>;
    if (timed) {
        hipEvent_t e0, e1;
        KT_TRY(timedKernelEvents(ctx, KT_KERNEL_SPMV32, &e0, &e1));
        hipExtLaunchKernelGGL(spmv, fg, dim3(256), 0, st, e0, e1, 0, a, parity);
    } else {
        spmv<<<fg, dim3(256), 0, st>>>(a, parity);
    }
    KT_HIP(hipGetLastError());
    const dim3 grid(gridBlocks(a, only));
    hipEvent_t t0 = nullptr;
    if (timed) KT_TRY(timedBegin(ctx, KT_KERNEL_UPDATE, &t0));
    switch (a.tile) {
        case 1 * kBlock: cgFUpdate<1><<<grid, dim3(kBlock), 0, st>>>(a, parity); break;
        case 2 * kBlock: cgFUpdate<2><<<grid, dim3(kBlock), 0, st>>>(a, parity); break;
        case 4 * kBlock: cgFUpdate<4><<<grid, dim3(kBlock), 0, st>>>(a, parity); break;
        case 8 * kBlock: cgFUpdate<8><<<grid, dim3(kBlock), 0, st>>>(a, parity); break;
        default: cgFUpdate<0><<<grid, dim3(kBlock), 0, st>>>(a, parity); break;
    }
    KT_HIP(hipGetLastError());
    if (timed) KT_TRY(timedEnd(ctx, KT_KERNEL_UPDATE, t0));
    return KT_OK;
}

template <typename T>
int launchIteration(kt_ctx *ctx, const CgArgs &a, int G, int parity, hipStream_t st, bool timed, int only) {
    if constexpr (sizeof(T) == 4) {
        if (a.fused) return launchIterationFused(ctx, a, parity, st, timed, only);
    }
    const dim3 grid(gridBlocks(a, only));
    hipEvent_t t0 = nullptr;
    const int kSpmv = sizeof(T) == 8 ? KT_KERNEL_SPMV : KT_KERNEL_SPMV32;
    if (timed && a.layout == 1) {
        hipEvent_t e0, e1;
        KT_TRY(timedKernelEvents(ctx, kSpmv, &e0, &e1));
        KT_TRY(launchSpmv<T>(ctx, a, G, parity, st, false, e0, e1, only));
    } else {
        if (timed) KT_TRY(timedBegin(ctx, kSpmv, &t0));
        KT_TRY(launchSpmv<T>(ctx, a, G, parity, st, false, nullptr, nullptr, only));
        if (timed) KT_TRY(timedEnd(ctx, kSpmv, t0));
    }
    if (timed) KT_TRY(timedBegin(ctx, KT_KERNEL_UPDATE, &t0));
    switch (a.tile) {
        case 1 * kBlock: cgUpdate<T, 1><<<grid, dim3(kBlock), 0, st>>>(a, parity); break;
        case 2 * kBlock: cgUpdate<T, 2><<<grid, dim3(kBlock), 0, st>>>(a, parity); break;
        case 4 * kBlock: cgUpdate<T, 4><<<grid, dim3(kBlock), 0, st>>>(a, parity); break;
        case 8 * kBlock: cgUpdate<T, 8><<<grid, dim3(kBlock), 0, st>>>(a, parity); break;
        default: cgUpdate<T, 0><<<grid, dim3(kBlock), 0, st>>>(a, parity); break;
    }
    KT_HIP(hipGetLastError());
    if (timed) KT_TRY(timedEnd(ctx, KT_KERNEL_UPDATE, t0));
    if (timed) KT_TRY(timedBegin(ctx, KT_KERNEL_PUPDATE, &t0));
    switch (a.tile) {
        case 1 * kBlock: cgPUpdate<T, 1><<<grid, dim3(kBlock), 0, st>>>(a, parity); break;
        case 2 * kBlock: cgPUpdate<T, 2><<<grid, dim3(kBlock), 0, st>>>(a, parity); break;
        case 4 * kBlock: cgPUpdate<T, 4><<<grid, dim3(kBlock), 0, st>>>(a, parity); break;
        case 8 * kBlock: cgPUpdate<T, 8><<<grid, dim3(kBlock), 0, st>>>(a, parity); break;
        default: cgPUpdate<T, 0><<<grid, dim3(kBlock), 0, st>>>(a, parity); break;
    }
    KT_HIP(hipGetLastError());
    if (timed) KT_TRY(timedEnd(ctx, KT_KERNEL_PUPDATE, t0));
    return KT_OK;
}

template <int MODE>
void launchFlatResidual(kt_ctx *ctx, const CgArgs &a);

template <typename T>
int launchStart(kt_ctx *ctx, const CgArgs &a, int G) {
    const dim3 grid(gridBlocks(a));
    if (a.flatResidual) {
        launchFlatResidual<sizeof(T) == 4 ? 3 : 4>(ctx, a);
        KT_HIP(hipGetLastError());
        cgStartFinish<<<dim3(2), dim3(kBlock), 0, ctx->stream>>>(a);
        KT_HIP(hipGetLastError());
        return KT_OK;
    }
    switch (G) {
        case 1: cgStart<T, 1><<<grid, dim3(kBlock), 0, ctx->stream>>>(a); break;
        case 2: cgStart<T, 2><<<grid, dim3(kBlock), 0, ctx->stream>>>(a); break;
        case 4: cgStart<T, 4><<<grid, dim3(kBlock), 0, ctx->stream>>>(a); break;
        case 8: cgStart<T, 8><<<grid, dim3(kBlock), 0, ctx->stream>>>(a); break;
        default: cgStart<T, 16><<<grid, dim3(kBlock), 0, ctx->stream>>>(a); break;
    }
    KT_HIP(hipGetLastError());
    cgStartFinish<<<dim3(2), dim3(kBlock), 0, ctx->stream>>>(a);
    KT_HIP(hipGetLastError());
    return KT_OK;
}

// The FP64 residual b - A x64 with the entry-block kernel (MODE 1: refresh,
// 2: true residual); used when the solve runs 256-thread blocks of 2048
// entries (flatResidual), with their row lanes and column bytes.
template <int MODE>
void launchFlatResidual(kt_ctx *ctx, const CgArgs &a) {
    const dim3 fg(flatGrid(a));
    hipStream_t st = ctx->stream;
    if (a.wide) {  // wide loads imply 4 row lanes
        if (a.col24) cgSpmvFlat<double, 2048, 4, 4, 256, true, MODE, true><<<fg, dim3(256), 0, st>>>(a, -1);
        else cgSpmvFlat<double, 2048, 4, 4, 256, false, MODE, true><<<fg, dim3(256), 0, st>>>(a, -1);
    } else if (a.col24) {
        if (a.flatRowLanes == 4) cgSpmvFlat<double, 2048, 4, 4, 256, true, MODE><<<fg, dim3(256), 0, st>>>(a, -1);
        else cgSpmvFlat<double, 2048, 4, 1, 256, true, MODE><<<fg, dim3(256), 0, st>>>(a, -1);
    } else {
        if (a.flatRowLanes == 4) cgSpmvFlat<double, 2048, 4, 4, 256, false, MODE><<<fg, dim3(256), 0, st>>>(a, -1);
        else cgSpmvFlat<double, 2048, 4, 1, 256, false, MODE><<<fg, dim3(256), 0, st>>>(a, -1);
    }
}

int launchTrueResidual(kt_ctx *ctx, const CgArgs &a, int G) {
    const dim3 grid(gridBlocks(a));
    if (a.flatResidual) launchFlatResidual<2>(ctx, a);
    else KT_DISPATCH_G(G, cgTrueResidual, grid, ctx->stream, a);
    KT_HIP(hipGetLastError());
    cgTrueResidualFinish<<<dim3(2), dim3(kBlock), 0, ctx->stream>>>(a);
    KT_HIP(hipGetLastError());
    return KT_OK;
}

int launchRefresh(kt_ctx *ctx, const CgArgs &a, int G) {
    const dim3 grid(gridBlocks(a));
    cgFlush<<<grid, dim3(kBlock), 0, ctx->stream>>>(a);
    KT_HIP(hipGetLastError());
    if (a.flatResidual) launchFlatResidual<1>(ctx, a);
    else KT_DISPATCH_G(G, cgRefresh, grid, ctx->stream, a);
    KT_HIP(hipGetLastError());
    cgRefreshFinish<<<dim3(2), dim3(kBlock), 0, ctx->stream>>>(a);
    KT_HIP(hipGetLastError());
    return KT_OK;
}

int lanesFor(kt_ctx *ctx) {
    if (ctx->tune.lanes > 0) return ctx->tune.lanes;
    uint64_t rows = 0, nnz = 0;
    for (int a = 0; a < 2; ++a) {
        rows += ctx->sys[a].rows;
        nnz += ctx->sys[a].nnzHost > 0 ? static_cast<uint64_t>(ctx->sys[a].nnzHost) : 0;
    }
    const double avg = rows ? static_cast<double>(nnz) / rows : 0.0;
    if (avg <= 3.0) return 1;
    if (avg <= 6.0) return 2;
    if (avg <= 16.0) return 4;
    if (avg <= 40.0) return 8;
    return 16;
}

SysView viewOf(kt_ctx *ctx, int axis, uint32_t tile) {
    Sys &s = ctx->sys[axis];
    SysView v{};
    v.rowPtr = s.rowPtr.p;
    v.col = s.col.p;
    v.val = s.val.p;
    v.diag = s.diag.p;
    v.rhs = s.rhs.p;
    v.val32 = s.val32.p;
    v.diag32 = s.diag32.p;
    v.rows = static_cast<uint32_t>(s.rows);
    v.blocks = static_cast<uint32_t>((s.rows + tile - 1) / tile);
    v.flatInfo = nullptr;
    v.colLo = nullptr;
    v.colHi = nullptr;
    v.colPk = nullptr;
    v.isd = nullptr;
    v.sblocks = v.blocks;
    return v;
}

// Iteration pairs (parity 0 then 1); without timing they replay a cached
// hipGraph of GraphCache::kPairs pairs.
// only: the one live axis (0, 1) when the other has stopped, else -1; each
// case has its own graph.
int runPairs(kt_ctx *ctx, const CgArgs &args, int G, bool fp32, uint32_t pairs, bool timed, int only) {
    // With one axis live and entry-block SpMV (whose blocks do not depend on
    // the tile), the vector kernels run half tiles: twice the blocks, so the
    // one axis still spreads over every CU.  cgUpdate and cgPUpdate of one
    // iteration use the same tile, and the partials buffer holds the blocks.
    CgArgs a = args;
    if (only >= 0 && a.layout == 1 && a.tile >= 2 * kBlock) {
        const uint32_t t = a.tile / 2;
        bool fits = true;
        for (int ax = 0; ax < 2; ++ax) fits = fits && (a.sys[ax].rows + t - 1) / t <= a.stride;
        if (fits) {
            a.tile = t;
            for (int ax = 0; ax < 2; ++ax) a.sys[ax].blocks = (a.sys[ax].rows + t - 1) / t;
        }
    }
    const auto direct = [&](uint32_t count) -> int {
        for (uint32_t k = 0; k < count; ++k)
            for (int parity = 0; parity < 2; ++parity) {
                const int rc = fp32 ? launchIteration<float>(ctx, a, G, parity, ctx->stream, timed, only)
                                    : launchIteration<double>(ctx, a, G, parity, ctx->stream, timed, only);
                if (rc != KT_OK) return rc;
            }
        return KT_OK;
    };
    if (timed || !ctx->tune.useGraphs) return direct(pairs);
    GraphCache &gc = ctx->graphs[fp32 ? 1 : 0][only + 1];
    if (!gc.exec || std::memcmp(&gc.args, &a, sizeof(CgArgs)) != 0 || gc.lanes != G || gc.rows != ctx->tune.rows) {
        if (gc.exec) (void)hipGraphExecDestroy(gc.exec);
        if (gc.graph) (void)hipGraphDestroy(gc.graph);
        gc.exec = nullptr;
        gc.graph = nullptr;
        KT_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        int rc = KT_OK;
        for (int rep = 0; rep < GraphCache::kPairs && rc == KT_OK; ++rep)
            for (int parity = 0; parity < 2 && rc == KT_OK; ++parity)
                rc = fp32 ? launchIteration<float>(ctx, a, G, parity, ctx->stream, false, only)
                          : launchIteration<double>(ctx, a, G, parity, ctx->stream, false, only);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
        if (rc != KT_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return hipFail(e, "hipStreamEndCapture", __FILE__, __LINE__);
        gc.graph = g;
        KT_HIP(hipGraphInstantiate(&gc.exec, g, nullptr, nullptr, 0));
        gc.args = a;
        gc.lanes = G;
        gc.rows = ctx->tune.rows;
    }
    uint32_t k = 0;
    for (; k + GraphCache::kPairs <= pairs; k += GraphCache::kPairs) KT_HIP(hipGraphLaunch(gc.exec, ctx->stream));
    return direct(pairs - k);
}

// The one live axis when the other is stopped or paused, else -1.  Between
// two polls an axis can only stop or pause (a refresh, which may make it live
// again, runs at a poll), so a chunk launched for one axis never misses the
// other.
int liveOnly(const AxisCtl *c) {
    const bool lx = c[0].s[0].live, ly = c[1].s[0].live;
    return lx == ly ? -1 : (lx ? 0 : 1);
}

int readCtl(kt_ctx *ctx) {
    KT_HIP(hipMemcpyAsync(ctx->ctlHost, ctx->ctl.p, 2 * sizeof(AxisCtl), hipMemcpyDeviceToHost, ctx->stream));
    KT_HIP(hipStreamSynchronize(ctx->stream));
    return KT_OK;
}

int writeCtl(kt_ctx *ctx) {
    KT_HIP(hipMemcpyAsync(ctx->ctl.p, ctx->ctlHost, 2 * sizeof(AxisCtl), hipMemcpyHostToDevice, ctx->stream));
    return KT_OK;
}

// Iterations to launch before the next poll: what the observed convergence
// rate says is left (rounded up to a pair), so few no-op iterations run.
uint32_t nextChunk(const AxisCtl *c, const double prevResid[2], const uint32_t prevSteps[2], double target,
                   uint32_t every, bool first) {
    uint32_t chunk = 2, cap = 2;
    for (int a = 0; a < 2; ++a) {
        const CgState &st = c[a].s[0];
        if (st.live) cap = std::max(cap, c[a].maxIter - st.steps);
    }
    if (first) return (std::min<uint32_t>(8, cap) + 1) & ~1u;
    for (int a = 0; a < 2; ++a) {
        const CgState &st = c[a].s[0];
        if (!st.live) continue;
        uint32_t need = every;
        if (st.steps > prevSteps[a] && prevResid[a] > st.resid && st.resid > 0.0) {
            const double rate = std::log(st.resid / prevResid[a]) / (st.steps - prevSteps[a]);
            const double left = std::log(std::max(target, c[a].tol) / st.resid) / rate;
            need = static_cast<uint32_t>(std::min<double>(std::max(std::ceil(left) + 1.0, 2.0), every));
        }
        chunk = std::max(chunk, need);
    }
    return (std::min(chunk, cap) + 1) & ~1u;
}

int prepare(kt_ctx *ctx, const kt_solve_opts &o, bool fp32, CgArgs &a, int &G) {
    const uint64_t rows = ctx->sys[0].rows;
    KT_CHECK(ctx->sys[1].rows == rows, KT_ESTATE, "solve: x and y systems differ in size");
    KT_CHECK(ctx->posCount == rows, KT_ESTATE, "solve: positions do not match the systems");
    uint64_t nnz[2];
    for (int ax = 0; ax < 2; ++ax) KT_TRY(systemNnz(ctx, ax, &nnz[ax]));
    G = lanesFor(ctx);
    const uint32_t tile = ctx->tune.tile > 0 ? ctx->tune.tile : 4096;
    KT_CHECK(tile % kBlock == 0, KT_EINVAL, "tile must be a multiple of 1024 rows");
    a = CgArgs{};
    a.tile = tile;
    a.xcdSplit = ctx->tune.xcdSplit < 0 ? 1u : static_cast<uint32_t>(ctx->tune.xcdSplit);
    a.halo = ctx->tune.halo;
    a.layout = ctx->tune.layout == KT_SPMV_ENTRY_BLOCKS ? 1u : 0u;
    a.flatEntries = ctx->tune.flatEntries ? ctx->tune.flatEntries : 8u * static_cast<uint32_t>(flatThreads(ctx->tune.flatThreads));
    a.flatUnroll = ctx->tune.flatUnroll == 8 ? 8u : 4u;
    // every launch-shape choice lives in the arguments: the graph cache
    // compares them to decide whether its captured iterations still apply
    a.flatRowLanes = ctx->tune.flatRowLanes == 4 ? 4u : 1u;
    a.flatThreads = static_cast<uint32_t>(flatThreads(ctx->tune.flatThreads));
    // 3-byte columns: entry blocks of 256 threads, rows below 2^24
    a.col24 = (a.layout == 1 && ctx->tune.colBytes == 3 && a.flatThreads == 256 && a.flatUnroll == 4 &&
               a.flatEntries == 2048 && rows < (1ull << 24)) ? 1u : 0u;
    // wide loads: entry blocks of 256 threads and 2048 entries, 4 lanes per row
    a.unitDiag = fp32 && ctx->tune.mixedScaling != 0 ? 1u : 0u;
    a.wide = (a.layout == 1 && ctx->tune.wideLoads && a.flatThreads == 256 && a.flatEntries == 2048 &&
              a.flatRowLanes == 4) ? 1u : 0u;
    a.col24 = a.col24 || (a.wide && ctx->tune.colBytes == 3 && rows < (1ull << 24)) ? 1u : 0u;
    // FP64 residuals by the same entry blocks (the shape instantiated for them)
    // (the residual kernels have their own entries per round)
    a.flatResidual = (a.layout == 1 && a.flatThreads == 256 && a.flatEntries == 2048 && rows > 0) ? 1u : 0u;
    // the fused FP32 iteration: on the scaled system, in the entry-block shape
    // the benchmark tunes to (the one instance of its SpMV)
    a.fused = (fp32 && ctx->tune.fusedUpdate && a.unitDiag && a.wide && a.col24 && a.flatResidual &&
               a.flatUnroll == 4) ? 1u : 0u;

    KT_CHECK(!(a.layout && a.halo), KT_EINVAL, "solve: the entry-block SpMV has no LDS window");
    if (a.halo) {
        const uint64_t bytes = fp32 ? windowBytes<float>(tile, a.halo) : windowBytes<double>(tile, a.halo);
        KT_CHECK(bytes <= kMaxWindowBytes, KT_EINVAL,
                 "solve: the SpMV window (tile + 2 halo entries) exceeds the LDS budget");
    }
    uint32_t maxBlocks = 1;
    bool scaleFlat[2] = {false, false};
    for (int ax = 0; ax < 2; ++ax) {
        if (fp32) {
            Sys &s = ctx->sys[ax];
            const bool scaled = ctx->tune.mixedScaling != 0;
            if (!s.have32 || s.scaled32 != scaled) {
                KT_TRY(s.val32.reserve(nnz[ax] + kQuadPad));
                KT_TRY(s.diag32.reserve(std::max<uint64_t>(rows, 1)));
                if (scaled && rows) {
                    // 1 / sqrt(d): the scaled copy below, and the start, refresh
                    // and flush kernels' conversions between the coordinates
                    KT_TRY(s.isd.reserve(rows));
                    invSqrtDiag<<<dim3(static_cast<uint32_t>((rows + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                                  ctx->stream>>>(s.diag.p, s.isd.p, static_cast<uint32_t>(rows));
                    KT_HIP(hipGetLastError());
                }
                if (nnz[ax] && scaled) {
                    if (a.layout == 1) {
                        scaleFlat[ax] = true;  // over the entry blocks once they are cut (below)
                    } else {
                        const uint64_t threads = 4ull * rows;
                        scaledToFloat<<<dim3(static_cast<uint32_t>((threads + kBlock - 1) / kBlock)), dim3(kBlock),
                                        0, ctx->stream>>>(s.rowPtr.p, s.col.p, s.val.p, s.isd.p, s.val32.p,
                                                          static_cast<uint32_t>(rows));
                        KT_HIP(hipGetLastError());
                    }
                } else if (nnz[ax]) {
                    toFloat<<<dim3(1024), dim3(kBlock), 0, ctx->stream>>>(s.val.p, s.val32.p, nnz[ax]);
                    KT_HIP(hipGetLastError());
                }
                s.scaled32 = scaled;
                if (rows) {
                    toFloat<<<dim3(256), dim3(kBlock), 0, ctx->stream>>>(s.diag.p, s.diag32.p, rows);
                    KT_HIP(hipGetLastError());
                }
                s.have32 = true;
            }
            KT_TRY(ctx->r32[ax].reserve(std::max<uint64_t>(rows, 1)));
            KT_TRY(ctx->p32[ax].reserve(std::max<uint64_t>(rows, 1)));
            KT_TRY(ctx->ap32[ax].reserve(std::max<uint64_t>(rows, 1)));
            KT_TRY(ctx->x32[ax].reserve(std::max<uint64_t>(rows, 1)));
            a.r[ax] = ctx->r32[ax].p;
            a.p[ax] = ctx->p32[ax].p;
            a.ap[ax] = ctx->ap32[ax].p;
            a.e[ax] = ctx->x32[ax].p;
        } else {
            KT_TRY(ctx->r[ax].reserve(std::max<uint64_t>(rows, 1)));
            KT_TRY(ctx->p[ax].reserve(std::max<uint64_t>(rows, 1)));
            KT_TRY(ctx->ap[ax].reserve(std::max<uint64_t>(rows, 1)));
            a.r[ax] = ctx->r[ax].p;
            a.p[ax] = ctx->p[ax].p;
            a.ap[ax] = ctx->ap[ax].p;
            a.e[ax] = nullptr;
        }
        a.sys[ax] = viewOf(ctx, ax, tile);
        if (a.unitDiag) a.sys[ax].isd = ctx->sys[ax].isd.p;
        if (a.layout == 1 && rows > 0) {
            Sys &sy = ctx->sys[ax];
            const uint32_t E = a.flatEntries;
            const uint32_t sb = static_cast<uint32_t>(std::max<uint64_t>((nnz[ax] + E - 1) / E, 1));
            KT_TRY(sy.rowStart.reserve(4ull * sb));
            uint4 *info = reinterpret_cast<uint4 *>(sy.rowStart.p);
            flatBounds<<<dim3((sb + 255) / 256), dim3(256), 0, ctx->stream>>>(sy.rowPtr.p, static_cast<uint32_t>(rows),
                                                                          sb, E, info);
            KT_HIP(hipGetLastError());
            a.sys[ax].flatInfo = info;
            if (scaleFlat[ax]) {
                scaledToFloatFlat<<<dim3(sb), dim3(256), 0, ctx->stream>>>(info, sy.rowPtr.p, sy.col.p, sy.val.p,
                                                                        sy.isd.p, sy.val32.p);
                KT_HIP(hipGetLastError());
            }
            if (a.col24 && a.wide) {
                if (!sy.havePk) {
                    // one spare quad: a block's loads may reach the quad after nnz
                    KT_TRY(sy.colPk.reserve(3 * (nnz[ax] / 4 + 2)));
                    if (nnz[ax]) {
                        packColumns<<<dim3(1024), dim3(kBlock), 0, ctx->stream>>>(sy.col.p, sy.colPk.p, nnz[ax]);
                        KT_HIP(hipGetLastError());
                    }
                    sy.havePk = true;
                }
                a.sys[ax].colPk = sy.colPk.p;
            } else if (a.col24) {
                if (!sy.have24) {
                    KT_TRY(sy.colLo.reserve(nnz[ax] + kQuadPad));
                    KT_TRY(sy.colHi.reserve(nnz[ax] + kQuadPad));
                    if (nnz[ax]) {
                        splitColumns<<<dim3(1024), dim3(kBlock), 0, ctx->stream>>>(sy.col.p, sy.colLo.p, sy.colHi.p,
                                                                                  nnz[ax]);
                        KT_HIP(hipGetLastError());
                    }
                    sy.have24 = true;
                }
                a.sys[ax].colLo = sy.colLo.p;
                a.sys[ax].colHi = sy.colHi.p;
            }
            a.sys[ax].sblocks = sb;
        }
        maxBlocks = std::max({maxBlocks, a.sys[ax].blocks, a.sys[ax].sblocks});
    }
    a.x64[0] = ctx->posX.p;
    a.x64[1] = ctx->posY.p;
    a.stride = maxBlocks;
    // fused: partU holds r'.r' per update tile, [parity][axis][2 maxBlocks]
    // (one-axis iterations may run half tiles)
    KT_TRY(ctx->partials.reserve(2ull * maxBlocks * 9 + (a.fused ? 8ull * maxBlocks : 0)));
    a.partA = ctx->partials.p;
    a.partB = ctx->partials.p + 2ull * maxBlocks * (a.fused ? 3 : 1);  // fused: partA holds three per block
    a.partS = ctx->partials.p + 2ull * maxBlocks * (a.fused ? 6 : 3);
    a.partU = a.fused ? ctx->partials.p + 2ull * maxBlocks * 9 : nullptr;
    a.uStride = 2 * maxBlocks;
    KT_TRY(ctx->ctl.reserve(2));
    a.ctl = ctx->ctl.p;
    const uint64_t histCap = std::min<uint64_t>(std::max<uint64_t>(o.max_iter, 1), 1u << 20);
    KT_TRY(ctx->history.reserve(2 * histCap));
    a.history = ctx->history.p;
    a.histCap = static_cast<uint32_t>(histCap);
    for (int ax = 0; ax < 2; ++ax) {
        AxisCtl &h = ctx->ctlHost[ax];
        std::memset(&h, 0, sizeof(h));
        h.tol = o.tol;
        h.maxIter = static_cast<uint32_t>(std::min<uint64_t>(o.max_iter, 0xffffffffu));
        h.refreshFactor = ctx->tune.refreshFactor > 0.0 ? ctx->tune.refreshFactor : 0.05;
    }
    KT_TRY(writeCtl(ctx));
    KT_TRY(fp32 ? launchSpmv<float>(ctx, a, G, 0, ctx->stream, true)
                : launchSpmv<double>(ctx, a, G, 0, ctx->stream, true));
    return KT_OK;
}

void fillResult(kt_ctx *ctx, kt_solve_result &res) {
    for (int ax = 0; ax < 2; ++ax) {
        const AxisCtl &c = ctx->ctlHost[ax];
        res.resid[ax] = c.s[0].resid;
        res.true_resid[ax] = c.trueResid;
        res.steps[ax] = c.s[0].steps;
    }
    res.iterations = std::max(res.steps[0], res.steps[1]);
}

}  // namespace

int multiply(kt_ctx *ctx, int axis, const double *dIn, double *dOut) {
    const SysView v = viewOf(ctx, axis, 4096);
    if (v.rows == 0) return KT_OK;
    multiplyKernel<<<dim3((v.rows + 255) / 256), dim3(256), 0, ctx->stream>>>(v, dIn, dOut);
    KT_HIP(hipGetLastError());
    return KT_OK;
}

int solveFp64(kt_ctx *ctx, const kt_solve_opts &o, kt_solve_result &res) {
    CgArgs a;
    int G;
    KT_TRY(prepare(ctx, o, false, a, G));
    res = kt_solve_result{};
    if (a.sys[0].blocks + a.sys[1].blocks == 0) return KT_OK;
    const bool timed = ctx->profile;
    const uint32_t every = o.check_every > 0 ? static_cast<uint32_t>(o.check_every) : 64;
    KT_TRY(launchStart<double>(ctx, a, G));
    uint64_t restarts = 0;
    for (int round = 0;; ++round) {
        if (round > 0) KT_TRY(launchStart<double>(ctx, a, G));
        double prevResid[2] = {0.0, 0.0};
        uint32_t prevSteps[2] = {0, 0};
        bool first = true;
        for (;;) {
            KT_TRY(readCtl(ctx));
            const AxisCtl *c = ctx->ctlHost;
            if (!c[0].s[0].live && !c[1].s[0].live) break;
            const uint32_t chunk = nextChunk(c, prevResid, prevSteps, o.tol, every, first);
            for (int ax = 0; ax < 2; ++ax) {
                prevResid[ax] = c[ax].s[0].resid;
                prevSteps[ax] = c[ax].s[0].steps;
            }
            first = false;
            KT_TRY(runPairs(ctx, a, G, false, chunk / 2, timed, liveOnly(c)));
        }
        KT_TRY(launchTrueResidual(ctx, a, G));
        KT_TRY(readCtl(ctx));
        if (o.stop_rule != KT_STOP_TRUE || round >= 16) break;
        // True-residual rule: restart the axes whose true residual is still
        // above tol from their current iterate (r = b - A x recomputed).
        bool again = false;
        for (int ax = 0; ax < 2; ++ax) {
            AxisCtl &c = ctx->ctlHost[ax];
            const bool more = c.s[0].active && c.trueResid > o.tol && c.s[0].steps < c.maxIter;
            c.frozen = more ? 0 : 1;
            again = again || more;
        }
        if (!again) break;
        ++restarts;
        KT_TRY(writeCtl(ctx));
    }
    fillResult(ctx, res);
    res.refinements = restarts;
    return KT_OK;
}

// FP32 PCG on FP32 copies of the matrix, with the iterate kept in FP64 and
// the residual refreshed in FP64 (reliable updates): whenever the FP32
// recurrence residual has fallen by the refresh factor since the last
// refresh, or reached tol, the FP32 correction is added into x (FP64), the
// residual is recomputed as b - A x in FP64 and rounded once to FP32, and the
// FP64 residual norm becomes the axis' residual.  An axis stops only on the
// true FP64 residual.  Every dot product (alpha, beta, residuals) accumulates
// in FP64.  With the scaling on (the default), the FP32 values are those of
// D^-1/2 A D^-1/2, the FP32 residual and correction are D^-1/2 r and D^1/2 e,
// and the refreshes convert between the two.
int solveMixed(kt_ctx *ctx, const kt_solve_opts &o, kt_solve_result &res) {
    CgArgs a;
    int G;
    KT_TRY(prepare(ctx, o, true, a, G));
    res = kt_solve_result{};
    if (a.sys[0].blocks + a.sys[1].blocks == 0) return KT_OK;
    const bool timed = ctx->profile;
    const uint32_t every = o.check_every > 0 ? static_cast<uint32_t>(o.check_every) : 64;
    KT_TRY(launchStart<float>(ctx, a, G));
    double prevResid[2] = {0.0, 0.0};
    uint32_t prevSteps[2] = {0, 0};
    bool first = true;
    for (int guard = 0; guard < (1 << 24); ++guard) {
        KT_TRY(readCtl(ctx));
        const AxisCtl *c = ctx->ctlHost;
        if (c[0].needUpdate || c[1].needUpdate) {
            KT_TRY(launchRefresh(ctx, a, G));
            // the first iterations after the refresh go out behind it, without
            // a poll in between: the kernels read the refreshed state on the
            // device, and an axis the refresh stopped steps as a no-op.  The
            // axes that may step: live ones and those being refreshed.
            bool may[2];
            uint32_t cap = 2;
            for (int ax = 0; ax < 2; ++ax) {
                may[ax] = c[ax].s[0].live || c[ax].needUpdate;
                if (may[ax] && c[ax].maxIter > c[ax].s[0].steps) cap = std::max(cap, c[ax].maxIter - c[ax].s[0].steps);
            }
            // not when every axis being refreshed reached tol on its recurrence
            // residual and no other axis is live: the solve most likely ends
            bool goesOn = false;
            for (int ax = 0; ax < 2; ++ax)
                goesOn = goesOn || c[ax].s[0].live || (c[ax].needUpdate && c[ax].s[0].resid > c[ax].tol);
            if (goesOn) {
                const int only = may[0] && may[1] ? -1 : (may[0] ? 0 : 1);
                const uint32_t chunk = (std::min<uint32_t>(8, cap) + 1) & ~1u;  // as the first chunk of a solve
                KT_TRY(runPairs(ctx, a, G, true, chunk / 2, timed, only));
            }
            first = true;
            continue;
        }
        if (!c[0].s[0].live && !c[1].s[0].live) break;
        double target = o.tol;
        for (int ax = 0; ax < 2; ++ax)
            if (c[ax].s[0].live) target = std::max(target, c[ax].refreshFactor * c[ax].refResid);
        const uint32_t chunk = nextChunk(c, prevResid, prevSteps, target, every, first);
        for (int ax = 0; ax < 2; ++ax) {
            prevResid[ax] = c[ax].s[0].resid;
            prevSteps[ax] = c[ax].s[0].steps;
        }
        first = false;
        KT_TRY(runPairs(ctx, a, G, true, chunk / 2, timed, liveOnly(c)));
    }
    // fold the remaining correction into x and report the true residual
    for (int ax = 0; ax < 2; ++ax) ctx->ctlHost[ax].needUpdate = 1;
    KT_TRY(writeCtl(ctx));
    KT_TRY(launchRefresh(ctx, a, G));
    KT_TRY(readCtl(ctx));
    fillResult(ctx, res);
    res.refinements = ctx->ctlHost[0].refreshes + ctx->ctlHost[1].refreshes;
    return KT_OK;
}

int solveHistory(kt_ctx *ctx, const kt_solve_result &res, std::vector<double> &hist, uint64_t &cap) {
    cap = ctx->history.cap / 2;
    hist.assign(2 * cap, 0.0);
    if (cap == 0) return KT_OK;
    KT_HIP(hipMemcpy(hist.data(), ctx->history.p, 2 * cap * sizeof(double), hipMemcpyDeviceToHost));
    return KT_OK;
}

}  // namespace kt
