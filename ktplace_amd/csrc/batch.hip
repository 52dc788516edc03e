// Many small designs' x/y systems solved together (no counterpart in the
// reference, which places one design per run).
//
// Every instance contributes two systems (x, y).  All systems are concatenated
// into one block-diagonal CSR: system s owns global rows
// [rowBase_s, rowBase_s + rows_s) and its column indices are global, so the
// kernels gather straight from the concatenated vectors.  Each system is cut
// into tiles of `tile` rows; one block per tile and one launch per CG kernel
// for all tiles of all systems.  Each system keeps its own Jacobi-PCG scalars
// (the reference's per-axis rules, simpl_lss.cc:203-290), reduced the same way
// as the two-system solver: tile partials, re-reduced in tile order by every
// consumer block of that system, state double buffered by iteration parity,
// no atomics and no grid barrier.  Converged systems drop out (their blocks
// return at once); the iteration pairs are replayed from one captured
// hipGraph until no system is live.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "kt_internal.h"
#include "spmv_rows.h"

namespace kt {
namespace {

constexpr int kBB = 256;  // threads per block

struct BSysDev {
    uint32_t rowBase, rows, tileBase, tiles;
    uint32_t flatBase, flatBlocks;  // entry-block SpMV: the system's blocks (and p.Ap partials)
};

struct BArgs {
    const uint32_t *rowPtr, *col;
    const double *val, *diag, *rhs;
    double *x, *r, *p, *ap;
    const BSysDev *sys;
    const uint32_t *tileSys;
    const uint4 *flat;              // entry-block SpMV: per block (system, first row, end row, 0)
    const uint2 *flatE;             // entry-block SpMV: per block (first entry, end entry)
    double *partA, *partB, *partS;  // [SpMV blocks][1], [tiles][2], [tiles][3]
    AxisCtl *ctl;
    uint32_t tile;
    uint32_t layout;                // 0 lane groups, 1 entry blocks
    const uint16_t *colLo;          // entry blocks with 3-byte columns: col = colLo | colHi << 16
    const uint8_t *colHi;
    const uint32_t *colPk;          // wide loads with 3-byte columns: quad q's columns in words 3q .. 3q + 2
    uint32_t col24;
    uint32_t wide;                  // entry blocks, 4 row lanes: four consecutive entries per load
};

struct Tile {
    uint32_t sys, begin, end;
    bool first;
};

__device__ __forceinline__ Tile decode(const BArgs &a) {
    const uint32_t s = a.tileSys[blockIdx.x];
    const BSysDev d = a.sys[s];
    Tile t;
    t.sys = s;
    t.begin = d.rowBase + (blockIdx.x - d.tileBase) * a.tile;
    t.end = min(t.begin + a.tile, d.rowBase + d.rows);
    t.first = blockIdx.x == d.tileBase;
    return t;
}

template <int K>
__device__ __forceinline__ void blockSum(double (&v)[K]) {
    __shared__ double waveSums[kBB / 64][K];
    __shared__ double total[K];
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) waveSums[wave][k] = v[k];
    __syncthreads();
    if (threadIdx.x < K) {
        double s = waveSums[0][threadIdx.x];
        for (int w = 1; w < kBB / 64; ++w) s += waveSums[w][threadIdx.x];
        total[threadIdx.x] = s;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = total[k];
}

// Sum of one system's partials [base, base + count) in order (same value in
// every block).
template <int K>
__device__ __forceinline__ void sumPartials(const double *part, uint32_t base, uint32_t count, double (&out)[K]) {
#pragma unroll
    for (int k = 0; k < K; ++k) out[k] = 0.0;
    for (uint32_t t = threadIdx.x; t < count; t += kBB)
#pragma unroll
        for (int k = 0; k < K; ++k) out[k] += part[static_cast<size_t>(base + t) * K + k];
    blockSum<K>(out);
}

template <int K>
__device__ __forceinline__ void sumPartials(const double *part, const BSysDev &d, double (&out)[K]) {
    sumPartials<K>(part, d.tileBase, d.tiles, out);
}

template <int K>
__device__ __forceinline__ void writePartial(double *part, const double (&v)[K]) {
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) part[static_cast<size_t>(blockIdx.x) * K + k] = v[k];
}

// r = b - A x, p = z = r / d; partials b.b, r.z, r.r
template <int G>
__global__ void __launch_bounds__(kBB) bStart(BArgs a) {
    const Tile t = decode(a);
    if (a.ctl[t.sys].frozen) return;
    const uint32_t lane = threadIdx.x % G, grp = threadIdx.x / G;
    double acc[3] = {0.0, 0.0, 0.0};
    constexpr int R = 2;
    constexpr uint32_t kPass = R * (kBB / G);
    for (uint32_t r0 = t.begin; r0 < t.end; r0 += kPass) {
        uint32_t row[R];
        bool valid[R];
        double off[R];
#pragma unroll
        for (int q = 0; q < R; ++q) {
            row[q] = r0 + grp + q * (kBB / G);
            valid[q] = row[q] < t.end;
        }
        rowsDot<double, G, R>(a.rowPtr, a.col, a.val, a.x, row, valid, lane, off);
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const uint32_t i = row[q];
            if (valid[q] && lane == 0) {
                const double ri = a.rhs[i] - (a.diag[i] * a.x[i] + off[q]);
                const double zi = ri / a.diag[i];
                a.r[i] = ri;
                a.p[i] = zi;
                acc[0] += a.rhs[i] * a.rhs[i];
                acc[1] += ri * zi;
                acc[2] += ri * ri;
            }
        }
    }
    blockSum<3>(acc);
    writePartial<3>(a.partS, acc);
}

// one block per system: the start partials become state[0]
__global__ void __launch_bounds__(kBB) bStartFinish(BArgs a) {
    AxisCtl &c = a.ctl[blockIdx.x];
    if (c.frozen) return;
    double tot[3];
    sumPartials<3>(a.partS, a.sys[blockIdx.x], tot);
    if (threadIdx.x == 0) {
        CgState &st = c.s[0];
        c.bNorm = sqrt(tot[0]);
        if (!(c.bNorm > 0.0)) {
            st.active = 0;
            st.live = 0;
        } else {
            st.rho = tot[1];
            st.resid = sqrt(tot[2]) / c.bNorm;
            st.active = 1;
            st.live = (st.resid > c.tol && st.steps < c.maxIter) ? 1 : 0;
        }
        c.trueResid = st.resid;
    }
}

template <int G, int R>
__global__ void __launch_bounds__(kBB) bSpmv(BArgs a, int parity) {
    const Tile t = decode(a);
    if (!a.ctl[t.sys].s[parity].live) return;
    const uint32_t lane = threadIdx.x % G, grp = threadIdx.x / G;
    double acc[1] = {0.0};
    constexpr uint32_t kPass = R * (kBB / G);
    for (uint32_t r0 = t.begin; r0 < t.end; r0 += kPass) {
        uint32_t row[R];
        bool valid[R];
        double off[R], pi[R], di[R];
#pragma unroll
        for (int q = 0; q < R; ++q) {
            row[q] = r0 + grp + q * (kBB / G);
            valid[q] = row[q] < t.end;
            pi[q] = valid[q] ? a.p[row[q]] : 0.0;
            di[q] = valid[q] ? a.diag[row[q]] : 0.0;
        }
        rowsDot<double, G, R>(a.rowPtr, a.col, a.val, a.p, row, valid, lane, off);
#pragma unroll
        for (int q = 0; q < R; ++q) {
            if (valid[q] && lane == 0) {
                const double v = di[q] * pi[q] + off[q];
                a.ap[row[q]] = v;
                acc[0] += pi[q] * v;
            }
        }
    }
    blockSum<1>(acc);
    writePartial<1>(a.partA, acc);
}

// Column indices as a 16-bit low half and an 8-bit high byte (rows < 2^24).
__global__ void __launch_bounds__(kBB)
bSplitColumns(const uint32_t *__restrict__ col, uint16_t *__restrict__ lo, uint8_t *__restrict__ hi, uint64_t n) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(kBB) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * kBB) {
        const uint32_t c = col[i];
        lo[i] = static_cast<uint16_t>(c & 0xffffu);
        hi[i] = static_cast<uint8_t>(c >> 16);
    }
}

// Columns (all below 2^24) packed four to three words, as the two-system
// solver's packColumns: quad q's c0 .. c3 become c0 | c1 << 24,
// c1 >> 8 | c2 << 16, c2 >> 16 | c3 << 8; entries past n pack as 0.
__global__ void __launch_bounds__(kBB)
bPackColumns(const uint32_t *__restrict__ col, uint32_t *__restrict__ pk, uint64_t n) {
    const uint64_t quads = (n + 3) / 4;
    for (uint64_t q = blockIdx.x * static_cast<uint64_t>(kBB) + threadIdx.x; q < quads;
         q += static_cast<uint64_t>(gridDim.x) * kBB) {
        uint32_t c[4];
        for (int j = 0; j < 4; ++j) c[j] = 4 * q + j < n ? col[4 * q + j] : 0u;
        pk[3 * q] = c[0] | (c[1] << 24);
        pk[3 * q + 1] = (c[1] >> 8) | (c[2] << 16);
        pk[3 * q + 2] = (c[2] >> 16) | (c[3] << 8);
    }
}

// Entry-block SpMV over the batch (the two-system solver's cgSpmvFlat, per
// system): block b takes rows [R0, R1) of system S, whose entries are a run
// of about kBFlat consecutive entries; the threads read the run's columns and
// values coalesced, gather p and keep the products in LDS, then G2 lanes per
// row add the row's products (G2 = 1: one thread, entry order after the
// diagonal term).  A run longer than the LDS holds takes lane groups.
// W4 (G2 = 4): four consecutive entries per thread by one wide load per array
// from the run's first entry rounded down to a multiple of four (the arrays
// carry four spare entries), the products at their slot from that base, and
// the row data of four passes by one load per array and wave, handed to the
// lane groups by shuffles; the sums and their order are unchanged.
constexpr uint32_t kBFlat = 2048;
constexpr uint32_t kBFlatSlack = 256;  // an eighth of the block: eight blocks fit a CU's LDS

template <bool C24>
__device__ __forceinline__ void bLoadQuad(const BArgs &a, uint32_t k0, uint32_t (&c)[4], double (&w)[4]) {
    if constexpr (C24) {
        const uint32_t *pk = a.colPk + 3u * (k0 >> 2);
        const uint32_t w0 = pk[0], w1 = pk[1], w2 = pk[2];
        c[0] = w0 & 0xffffffu;
        c[1] = (w0 >> 24) | ((w1 & 0xffffu) << 8);
        c[2] = (w1 >> 16) | ((w2 & 0xffu) << 16);
        c[3] = w2 >> 8;
    } else {
        const uint4 q = *reinterpret_cast<const uint4 *>(a.col + k0);
        c[0] = q.x;
        c[1] = q.y;
        c[2] = q.z;
        c[3] = q.w;
    }
    const double2 v0 = *reinterpret_cast<const double2 *>(a.val + k0);
    const double2 v1 = *reinterpret_cast<const double2 *>(a.val + k0 + 2);
    w[0] = v0.x;
    w[1] = v0.y;
    w[2] = v1.x;
    w[3] = v1.y;
}

template <int G2, bool C24, bool W4 = false>
__global__ void __launch_bounds__(kBB) bSpmvFlat(BArgs a, int parity) {
    static_assert(!W4 || G2 == 4, "wide loads run with 4 lanes per row");
    __shared__ alignas(16) double prod[kBFlat + kBFlatSlack + (W4 ? 8 : 0)];
    const uint4 f = a.flat[blockIdx.x];
    const uint2 fe = a.flatE[blockIdx.x];
    const uint32_t S = f.x, R0 = f.y, R1 = f.z;
    if (!a.ctl[S].s[parity].live) return;
    const double *__restrict__ p = a.p;
    double *__restrict__ ap = a.ap;
    const uint32_t E0 = fe.x, count = fe.y - fe.x;
    double acc[1] = {0.0};
    if (count > kBFlat + kBFlatSlack) {
        constexpr int G = 4, R = 2;
        const uint32_t lane = threadIdx.x % G, grp = threadIdx.x / G;
        for (uint32_t r0 = R0; r0 < R1; r0 += R * (kBB / G)) {
            uint32_t row[R];
            bool valid[R];
            double off[R];
#pragma unroll
            for (int q = 0; q < R; ++q) {
                row[q] = r0 + grp + q * (kBB / G);
                valid[q] = row[q] < R1;
            }
            rowsDot<double, G, R>(a.rowPtr, a.col, a.val, p, row, valid, lane, off);
#pragma unroll
            for (int q = 0; q < R; ++q) {
                if (valid[q] && lane == 0) {
                    const double pd = p[row[q]];
                    const double v = a.diag[row[q]] * pd + off[q];
                    ap[row[q]] = v;
                    acc[0] += pd * v;
                }
            }
        }
    } else {
        // the first pass of phase 2's row data is loaded before the gathers
        // (a block of kBFlat entries usually has fewer rows than groups)
        constexpr uint32_t kGroups = kBB / G2;
        const uint32_t lane = threadIdx.x % G2, grp = threadIdx.x / G2;
        constexpr uint32_t kWaveGroups = 64 / G2;
        const uint32_t lane64 = threadIdx.x & 63u, wave = threadIdx.x >> 6;
        // (an instance without rows has one block with R0 == R1 and no row to read)
        uint32_t pre0 = 0, pre1 = 0;
        double preP = 0.0, preD = 0.0;
        const auto loadRows = [&](uint32_t r) {
            const uint32_t rq = r < R1 ? r : R0;
            pre0 = a.rowPtr[rq];
            pre1 = a.rowPtr[rq + 1];
            preP = p[rq];
            preD = a.diag[rq];
        };
        // W4: lane l of wave w holds row R0 + chunk + kGroups (l / kWaveGroups)
        // + kWaveGroups w + l % kWaveGroups (pass l / kWaveGroups)
        const auto chunkRow = [&](uint32_t chunk) {
            return R0 + chunk * (kGroups * G2) + kGroups * (lane64 / kWaveGroups) + kWaveGroups * wave +
                   lane64 % kWaveGroups;
        };
        if (R0 < R1) {
            if constexpr (W4) loadRows(chunkRow(0));
            else loadRows(R0 + grp);
        }
        __builtin_amdgcn_sched_barrier(0);
        constexpr int U = 4;
        if constexpr (W4) {
            const uint32_t A = E0 & ~3u, lead = E0 - A, quads = (lead + count + 3) / 4;
            for (uint32_t q = threadIdx.x; q < quads; q += kBB) {
                uint32_t c[4];
                double w[4];
                bLoadQuad<C24>(a, A + 4u * q, c, w);
                __builtin_amdgcn_sched_barrier(0);
                double g[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const uint32_t k = 4u * q + j;
                    g[j] = gatherAt(p, k >= lead && k < lead + count ? c[j] : kNoCol);
                }
                __builtin_amdgcn_sched_barrier(0);
                double2 *dst = reinterpret_cast<double2 *>(prod + 4u * q);
                dst[0] = make_double2(w[0] * g[0], w[1] * g[1]);
                dst[1] = make_double2(w[2] * g[2], w[3] * g[3]);
            }
        } else
        for (uint32_t base = 0; base < count; base += U * kBB) {
            uint32_t c[U];
            double w[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t k = base + u * kBB + threadIdx.x;
                const bool ok = k < count;
                const uint32_t kk = E0 + (ok ? k : 0u);  // count > 0 here: entry E0 exists
                uint32_t cl;
                if constexpr (C24) cl = static_cast<uint32_t>(a.colLo[kk]) | (static_cast<uint32_t>(a.colHi[kk]) << 16);
                else cl = a.col[kk];
                const double wl = a.val[kk];
                c[u] = ok ? cl : kNoCol;
                w[u] = ok ? wl : 0.0;
            }
            __builtin_amdgcn_sched_barrier(0);
            double g[U];
#pragma unroll
            for (int u = 0; u < U; ++u) g[u] = gatherAt(p, c[u]);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t k = base + u * kBB + threadIdx.x;
                if (k < count) prod[k] = w[u] * g[u];
            }
        }
        __syncthreads();
        const uint32_t slot0 = W4 ? (E0 & ~3u) : E0;  // LDS slot of entry e: e - slot0
        const auto pass = [&](uint32_t r0, uint32_t q0, uint32_t q1, double pr, double dr) {
            const uint32_t r = r0 + grp;
            const bool valid = r < R1;
            const uint32_t e0 = q0 - slot0, e1 = valid ? q1 - slot0 : e0;
            const double pd = valid ? pr : 0.0;
            double part = G2 == 1 ? dr * pd : 0.0;
            for (uint32_t k = e0 + lane; k < e1; k += G2) part += prod[k];
#pragma unroll
            for (int off = G2 / 2; off > 0; off >>= 1) part += __shfl_xor(part, off, G2);
            if (valid && lane == 0) {
                const double v = G2 == 1 ? part : dr * pd + part;
                ap[r] = v;
                acc[0] += pd * v;
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
                    pass(r0, __shfl(pre0, src), __shfl(pre1, src), __shfl(preP, src), __shfl(preD, src));
                }
            }
        } else {
            if (R0 < R1) pass(R0, pre0, pre1, preP, preD);
            for (uint32_t r0 = R0 + kGroups; r0 < R1; r0 += kGroups) {
                loadRows(r0 + grp);
                pass(r0, pre0, pre1, preP, preD);
            }
        }
    }
    blockSum<1>(acc);
    writePartial<1>(a.partA, acc);
}

// Rows per thread whose vector entries the update kernels load before the
// partial sums (they do not depend on alpha or beta), so the sums hide their
// latency; rows beyond kPre * kBB of a tile are loaded after.
constexpr int kPre = 4;

// alpha = rho / p.Ap; r -= alpha Ap; partials r.r, r.z.  x += alpha p is left
// to bPUpdate, which reads p anyway (alpha through the system's control block).
__global__ void __launch_bounds__(kBB) bUpdate(BArgs a, int parity) {
    const Tile t = decode(a);
    AxisCtl &c = a.ctl[t.sys];
    const CgState st = c.s[parity];
    if (!st.live) return;
    double *__restrict__ r = a.r;
    const double *__restrict__ ap = a.ap;
    const double *__restrict__ dg = a.diag;
    double rv[kPre], av[kPre], dv[kPre];
#pragma unroll
    for (int k = 0; k < kPre; ++k) {
        const uint32_t i = t.begin + threadIdx.x + k * kBB;
        const uint32_t j = i < t.end ? i : t.begin;  // unconditional loads; a tile has a row
        rv[k] = r[j];
        av[k] = ap[j];
        dv[k] = dg[j];
    }
    double pap[1];
    const BSysDev d = a.sys[t.sys];
    if (a.layout) sumPartials<1>(a.partA, d.flatBase, d.flatBlocks, pap);
    else sumPartials<1>(a.partA, d, pap);
    if (!(pap[0] > 0.0)) {
        if (t.first && threadIdx.x == 0) c.brk[parity] = 1;
        return;
    }
    const double alpha = st.rho / pap[0];
    if (t.first && threadIdx.x == 0) {
        c.brk[parity] = 0;
        c.alpha[parity] = alpha;
    }
    double acc[2] = {0.0, 0.0};
    const auto step = [&](uint32_t i, double ri0, double ai, double di) {
        const double ri = ri0 - alpha * ai;
        r[i] = ri;
        acc[0] += ri * ri;
        acc[1] += ri * (ri / di);
    };
#pragma unroll
    for (int k = 0; k < kPre; ++k) {
        const uint32_t i = t.begin + threadIdx.x + k * kBB;
        if (i < t.end) step(i, rv[k], av[k], dv[k]);
    }
    for (uint32_t i = t.begin + threadIdx.x + kPre * kBB; i < t.end; i += kBB) step(i, r[i], ap[i], dg[i]);
    blockSum<2>(acc);
    writePartial<2>(a.partB, acc);
}

__global__ void __launch_bounds__(kBB) bPUpdate(BArgs a, int parity) {
    const Tile t = decode(a);
    AxisCtl &c = a.ctl[t.sys];
    const CgState st = c.s[parity];
    const bool lead = t.first && threadIdx.x == 0;
    if (!st.live) {
        if (lead) c.s[parity ^ 1] = st;
        return;
    }
    if (c.brk[parity]) {  // p.Ap <= 0 stops the system (simpl_lss.cc:242-244)
        if (lead) {
            CgState nx = st;
            nx.active = 0;
            nx.live = 0;
            nx.steps = st.steps + 1;
            c.s[parity ^ 1] = nx;
        }
        return;
    }
    double *__restrict__ x = a.x;
    double *__restrict__ p = a.p;
    const double *__restrict__ r = a.r;
    const double *__restrict__ dg = a.diag;
    const double alpha = c.alpha[parity];
    double xv[kPre], rv[kPre], dv[kPre], pv[kPre];
#pragma unroll
    for (int k = 0; k < kPre; ++k) {
        const uint32_t i = t.begin + threadIdx.x + k * kBB;
        const uint32_t j = i < t.end ? i : t.begin;
        xv[k] = x[j];
        rv[k] = r[j];
        dv[k] = dg[j];
        pv[k] = p[j];
    }
    double tot[2];
    sumPartials<2>(a.partB, a.sys[t.sys], tot);
    const double resid = sqrt(tot[0]) / c.bNorm;
    const double rhoNext = tot[1];
    const bool ok = rhoNext > 0.0;
    // x steps whatever r.z turns out to be (simpl_lss.cc:247,253); p only when
    // the system goes on
    const double beta = ok ? rhoNext / st.rho : 0.0;
#pragma unroll
    for (int k = 0; k < kPre; ++k) {
        const uint32_t i = t.begin + threadIdx.x + k * kBB;
        if (i < t.end) {
            x[i] = xv[k] + alpha * pv[k];
            if (ok) p[i] = rv[k] / dv[k] + beta * pv[k];
        }
    }
    for (uint32_t i = t.begin + threadIdx.x + kPre * kBB; i < t.end; i += kBB) {
        const double pi = p[i];
        x[i] = x[i] + alpha * pi;
        if (ok) p[i] = r[i] / dg[i] + beta * pi;
    }
    if (lead) {
        CgState nx;
        nx.steps = st.steps + 1;
        nx.resid = resid;
        nx.rho = ok ? rhoNext : st.rho;
        nx.active = ok ? 1 : 0;
        nx.live = (ok && resid > c.tol && nx.steps < c.maxIter) ? 1 : 0;
        nx.pad = 0;
        c.s[parity ^ 1] = nx;
    }
}

template <int G>
__global__ void __launch_bounds__(kBB) bTrueResidual(BArgs a) {
    const Tile t = decode(a);
    const uint32_t lane = threadIdx.x % G, grp = threadIdx.x / G;
    double acc[2] = {0.0, 0.0};
    constexpr int R = 2;
    constexpr uint32_t kPass = R * (kBB / G);
    for (uint32_t r0 = t.begin; r0 < t.end; r0 += kPass) {
        uint32_t row[R];
        bool valid[R];
        double off[R];
#pragma unroll
        for (int q = 0; q < R; ++q) {
            row[q] = r0 + grp + q * (kBB / G);
            valid[q] = row[q] < t.end;
        }
        rowsDot<double, G, R>(a.rowPtr, a.col, a.val, a.x, row, valid, lane, off);
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const uint32_t i = row[q];
            if (valid[q] && lane == 0) {
                const double ri = a.rhs[i] - (a.diag[i] * a.x[i] + off[q]);
                acc[0] += ri * ri;
                acc[1] += a.rhs[i] * a.rhs[i];
            }
        }
    }
    blockSum<2>(acc);
    writePartial<2>(a.partB, acc);
}

__global__ void __launch_bounds__(kBB) bTrueResidualFinish(BArgs a) {
    double tot[2];
    sumPartials<2>(a.partB, a.sys[blockIdx.x], tot);
    if (threadIdx.x == 0) {
        const double bn = sqrt(tot[1]);
        a.ctl[blockIdx.x].trueResid = bn > 0.0 ? sqrt(tot[0]) / bn : 0.0;
    }
}

struct HostSys {
    uint32_t rowBase, rows, nnzBase, nnz;
};

}  // namespace
}  // namespace kt

struct kt_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    // host copy of the concatenated systems (uploaded when dirty)
    std::vector<kt::HostSys> sys;
    std::vector<uint32_t> rowPtr{0}, col;
    std::vector<double> val, diag, rhs, x0;
    bool dirty = true;
    uint32_t tileUploaded = 0;
    uint32_t tiles = 0;
    kt::DevBuf<uint32_t> dRowPtr, dCol, dTileSys;
    kt::DevBuf<uint4> dFlat;
    kt::DevBuf<uint2> dFlatE;
    kt::DevBuf<uint16_t> dColLo;  // 3-byte columns (batches of fewer than 2^24 rows)
    kt::DevBuf<uint8_t> dColHi;
    kt::DevBuf<uint32_t> dColPk;  // 3-byte columns packed four to three words (wide loads)
    int colBytes = 0;
    int wide = 0;                 // entry blocks with 4 row lanes: four entries per load
    uint32_t flatBlocks = 0, partASlots = 0;
    int layout = 0, rowLanes = 1;
    uint32_t layoutUploaded = 0xffffffffu;
    kt::DevBuf<double> dVal, dDiag, dRhs, dX, dX0, dR, dP, dAp, dPart;
    kt::DevBuf<kt::BSysDev> dSys;
    kt::DevBuf<kt::AxisCtl> dCtl;
    kt::AxisCtl *ctlHost = nullptr;
    size_t ctlHostCap = 0;
    int lanes = 8, rows = 2;
    uint32_t tile = 1024;
    bool useGraphs = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    kt::BArgs graphArgs{};
    int graphLanes = 0, graphRows = 0, graphLayout = -1, graphRowLanes = 0;
};

namespace kt {
namespace {

constexpr int kPairs = 8;

void dropGraph(kt_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

int upload(kt_batch *b) {
    if (!b->dirty && b->tileUploaded == b->tile && b->layoutUploaded == static_cast<uint32_t>(b->layout)) return KT_OK;
    const uint64_t rows = b->diag.size(), nnz = b->col.size();
    KT_TRY(b->dRowPtr.reserve(rows + 1));
    // four spare entries: the wide loads read whole aligned groups of four
    KT_TRY(b->dCol.reserve(nnz + kQuadPad));
    KT_TRY(b->dVal.reserve(nnz + kQuadPad));
    for (auto *v : {&b->dDiag, &b->dRhs, &b->dX, &b->dX0, &b->dR, &b->dP, &b->dAp})
        KT_TRY(v->reserve(std::max<uint64_t>(rows, 1)));
    // tiles
    std::vector<BSysDev> sd(b->sys.size());
    std::vector<uint32_t> tileSys;
    std::vector<uint4> flat;  // entry blocks: (system, first row, end row, 0)
    std::vector<uint2> flatE; // entry blocks: (first entry, end entry)
    for (size_t s = 0; s < b->sys.size(); ++s) {
        const uint32_t t = (b->sys[s].rows + b->tile - 1) / b->tile;
        const uint32_t r0 = b->sys[s].rowBase, r1 = r0 + b->sys[s].rows;
        const uint32_t fbase = static_cast<uint32_t>(flat.size());
        // cut the system's entries every kBFlat at the next row start
        uint32_t start = r0;
        uint64_t target = static_cast<uint64_t>(b->rowPtr[r0]) + kBFlat;
        for (uint32_t r = r0; r < r1; ++r) {
            if (b->rowPtr[r] >= target && r > start) {
                flat.push_back(make_uint4(static_cast<uint32_t>(s), start, r, 0u));
                flatE.push_back(make_uint2(b->rowPtr[start], b->rowPtr[r]));
                start = r;
                while (target <= b->rowPtr[r]) target += kBFlat;
            }
        }
        flat.push_back(make_uint4(static_cast<uint32_t>(s), start, r1, 0u));
        flatE.push_back(make_uint2(b->rowPtr[start], b->rowPtr[r1]));
        sd[s] = BSysDev{r0, b->sys[s].rows, static_cast<uint32_t>(tileSys.size()), t, fbase,
                        static_cast<uint32_t>(flat.size()) - fbase};
        tileSys.insert(tileSys.end(), t, static_cast<uint32_t>(s));
    }
    b->tiles = static_cast<uint32_t>(tileSys.size());
    b->flatBlocks = static_cast<uint32_t>(flat.size());
    b->partASlots = std::max(b->tiles, b->flatBlocks);
    KT_TRY(b->dSys.reserve(std::max<size_t>(sd.size(), 1)));
    KT_TRY(b->dTileSys.reserve(std::max<size_t>(tileSys.size(), 1)));
    KT_TRY(b->dFlat.reserve(std::max<size_t>(flat.size(), 1)));
    KT_TRY(b->dFlatE.reserve(std::max<size_t>(flatE.size(), 1)));
    KT_TRY(b->dPart.reserve(std::max<size_t>(b->partASlots + 5ull * b->tiles, 1)));
    KT_TRY(b->dCtl.reserve(std::max<size_t>(b->sys.size(), 1)));
    if (b->ctlHostCap < b->sys.size()) {
        if (b->ctlHost) (void)hipHostFree(b->ctlHost);
        b->ctlHost = nullptr;
        b->ctlHostCap = 0;
        KT_HIP(hipHostMalloc(reinterpret_cast<void **>(&b->ctlHost), b->sys.size() * sizeof(AxisCtl)));
        b->ctlHostCap = b->sys.size();
    }
    const auto h2d = [&](void *dst, const void *src, size_t bytes) -> int {
        if (bytes) KT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, b->stream));
        return KT_OK;
    };
    KT_TRY(h2d(b->dRowPtr.p, b->rowPtr.data(), (rows + 1) * sizeof(uint32_t)));
    KT_TRY(h2d(b->dCol.p, b->col.data(), nnz * sizeof(uint32_t)));
    KT_TRY(h2d(b->dVal.p, b->val.data(), nnz * sizeof(double)));
    KT_TRY(h2d(b->dDiag.p, b->diag.data(), rows * sizeof(double)));
    KT_TRY(h2d(b->dRhs.p, b->rhs.data(), rows * sizeof(double)));
    KT_TRY(h2d(b->dX0.p, b->x0.data(), rows * sizeof(double)));
    KT_TRY(h2d(b->dX.p, b->x0.data(), rows * sizeof(double)));
    KT_TRY(h2d(b->dSys.p, sd.data(), sd.size() * sizeof(BSysDev)));
    KT_TRY(h2d(b->dTileSys.p, tileSys.data(), tileSys.size() * sizeof(uint32_t)));
    KT_TRY(h2d(b->dFlat.p, flat.data(), flat.size() * sizeof(uint4)));
    KT_TRY(h2d(b->dFlatE.p, flatE.data(), flatE.size() * sizeof(uint2)));
    if (rows < (1ull << 24)) {
        KT_TRY(b->dColLo.reserve(std::max<uint64_t>(nnz, 1)));
        KT_TRY(b->dColHi.reserve(std::max<uint64_t>(nnz, 1)));
        if (nnz) {
            bSplitColumns<<<dim3(1024), dim3(kBB), 0, b->stream>>>(b->dCol.p, b->dColLo.p, b->dColHi.p, nnz);
            KT_HIP(hipGetLastError());
        }
        KT_TRY(b->dColPk.reserve(3 * (nnz / 4 + 2)));  // one spare quad
        if (nnz) {
            bPackColumns<<<dim3(1024), dim3(kBB), 0, b->stream>>>(b->dCol.p, b->dColPk.p, nnz);
            KT_HIP(hipGetLastError());
        }
    }
    KT_HIP(hipStreamSynchronize(b->stream));
    b->dirty = false;
    b->tileUploaded = b->tile;
    b->layoutUploaded = static_cast<uint32_t>(b->layout);
    dropGraph(b);
    return KT_OK;
}

BArgs argsOf(kt_batch *b) {
    BArgs a{};
    a.rowPtr = b->dRowPtr.p;
    a.col = b->dCol.p;
    a.val = b->dVal.p;
    a.diag = b->dDiag.p;
    a.rhs = b->dRhs.p;
    a.x = b->dX.p;
    a.r = b->dR.p;
    a.p = b->dP.p;
    a.ap = b->dAp.p;
    a.sys = b->dSys.p;
    a.tileSys = b->dTileSys.p;
    a.flat = b->dFlat.p;
    a.flatE = b->dFlatE.p;
    a.partA = b->dPart.p;
    a.partB = b->dPart.p + b->partASlots;
    a.partS = b->dPart.p + b->partASlots + 2ull * b->tiles;
    a.ctl = b->dCtl.p;
    a.tile = b->tile;
    a.layout = b->layout ? 1u : 0u;
    a.col24 = (a.layout && b->colBytes == 3 && b->diag.size() < (1ull << 24)) ? 1u : 0u;
    a.colLo = a.col24 ? b->dColLo.p : nullptr;
    a.colHi = a.col24 ? b->dColHi.p : nullptr;
    a.wide = (a.layout && b->wide && b->rowLanes == 4) ? 1u : 0u;
    a.colPk = a.col24 && a.wide ? b->dColPk.p : nullptr;
    return a;
}

#define KT_BATCH_G(G, KERNEL, GRID, STREAM, ...)                                  \
    switch (G) {                                                                  \
        case 4: KERNEL<4><<<GRID, dim3(kBB), 0, STREAM>>>(__VA_ARGS__); break;    \
        case 8: KERNEL<8><<<GRID, dim3(kBB), 0, STREAM>>>(__VA_ARGS__); break;    \
        default: KERNEL<16><<<GRID, dim3(kBB), 0, STREAM>>>(__VA_ARGS__); break;  \
    }

int launchIter(kt_batch *b, const BArgs &a, int parity) {
    const dim3 grid(b->tiles);
    const int R = b->rows >= 4 ? 4 : 2;
    if (a.layout) {
        if (a.wide) {
            if (a.col24) bSpmvFlat<4, true, true><<<dim3(b->flatBlocks), dim3(kBB), 0, b->stream>>>(a, parity);
            else bSpmvFlat<4, false, true><<<dim3(b->flatBlocks), dim3(kBB), 0, b->stream>>>(a, parity);
        } else if (a.col24) {
            if (b->rowLanes == 4) bSpmvFlat<4, true><<<dim3(b->flatBlocks), dim3(kBB), 0, b->stream>>>(a, parity);
            else bSpmvFlat<1, true><<<dim3(b->flatBlocks), dim3(kBB), 0, b->stream>>>(a, parity);
        } else {
            if (b->rowLanes == 4) bSpmvFlat<4, false><<<dim3(b->flatBlocks), dim3(kBB), 0, b->stream>>>(a, parity);
            else bSpmvFlat<1, false><<<dim3(b->flatBlocks), dim3(kBB), 0, b->stream>>>(a, parity);
        }
        KT_HIP(hipGetLastError());
        bUpdate<<<grid, dim3(kBB), 0, b->stream>>>(a, parity);
        KT_HIP(hipGetLastError());
        bPUpdate<<<grid, dim3(kBB), 0, b->stream>>>(a, parity);
        KT_HIP(hipGetLastError());
        return KT_OK;
    }
#define KT_B_SPMV(GG)                                                                      \
    if (R == 4) bSpmv<GG, 4><<<grid, dim3(kBB), 0, b->stream>>>(a, parity);                \
    else bSpmv<GG, 2><<<grid, dim3(kBB), 0, b->stream>>>(a, parity);
    switch (b->lanes) {
        case 4: KT_B_SPMV(4) break;
        case 8: KT_B_SPMV(8) break;
        default: KT_B_SPMV(16) break;
    }
#undef KT_B_SPMV
    KT_HIP(hipGetLastError());
    bUpdate<<<grid, dim3(kBB), 0, b->stream>>>(a, parity);
    KT_HIP(hipGetLastError());
    bPUpdate<<<grid, dim3(kBB), 0, b->stream>>>(a, parity);
    KT_HIP(hipGetLastError());
    return KT_OK;
}

int runPairs(kt_batch *b, const BArgs &a, uint32_t pairs) {
    const auto direct = [&](uint32_t count) -> int {
        for (uint32_t k = 0; k < count; ++k)
            for (int parity = 0; parity < 2; ++parity) KT_TRY(launchIter(b, a, parity));
        return KT_OK;
    };
    if (!b->useGraphs) return direct(pairs);
    if (!b->exec || std::memcmp(&b->graphArgs, &a, sizeof(BArgs)) != 0 || b->graphLanes != b->lanes ||
        b->graphRows != b->rows || b->graphLayout != b->layout || b->graphRowLanes != b->rowLanes) {
        dropGraph(b);
        KT_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        int rc = KT_OK;
        for (int rep = 0; rep < kPairs && rc == KT_OK; ++rep)
            for (int parity = 0; parity < 2 && rc == KT_OK; ++parity) rc = launchIter(b, a, parity);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (rc != KT_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return hipFail(e, "hipStreamEndCapture", __FILE__, __LINE__);
        b->graph = g;
        KT_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
        b->graphArgs = a;
        b->graphLanes = b->lanes;
        b->graphRows = b->rows;
        b->graphLayout = b->layout;
        b->graphRowLanes = b->rowLanes;
    }
    uint32_t k = 0;
    for (; k + kPairs <= pairs; k += kPairs) KT_HIP(hipGraphLaunch(b->exec, b->stream));
    return direct(pairs - k);
}

int readCtl(kt_batch *b) {
    KT_HIP(hipMemcpyAsync(b->ctlHost, b->dCtl.p, b->sys.size() * sizeof(AxisCtl), hipMemcpyDeviceToHost, b->stream));
    KT_HIP(hipStreamSynchronize(b->stream));
    return KT_OK;
}

int writeCtl(kt_batch *b) {
    KT_HIP(hipMemcpyAsync(b->dCtl.p, b->ctlHost, b->sys.size() * sizeof(AxisCtl), hipMemcpyHostToDevice, b->stream));
    return KT_OK;
}

int startAll(kt_batch *b, const BArgs &a) {
    KT_BATCH_G(b->lanes, bStart, dim3(b->tiles), b->stream, a);
    KT_HIP(hipGetLastError());
    bStartFinish<<<dim3(static_cast<uint32_t>(b->sys.size())), dim3(kBB), 0, b->stream>>>(a);
    KT_HIP(hipGetLastError());
    return KT_OK;
}

int batchSolve(kt_batch *b, const kt_solve_opts &o, kt_batch_instance_result *out, kt_solve_result *total) {
    KT_CHECK(o.precision == KT_PREC_FP64, KT_EINVAL, "kt_batch_solve: only KT_PREC_FP64");
    KT_CHECK(o.warm_start == KT_WARM_PREVIOUS, KT_EINVAL, "kt_batch_solve: previous-iterate start only");
    KT_CHECK(!b->sys.empty(), KT_ESTATE, "kt_batch_solve: no instances");
    KT_TRY(upload(b));
    const BArgs a = argsOf(b);
    const size_t S = b->sys.size();
    for (size_t s = 0; s < S; ++s) {
        AxisCtl &h = b->ctlHost[s];
        std::memset(&h, 0, sizeof(h));
        h.tol = o.tol;
        h.maxIter = static_cast<uint32_t>(std::min<uint64_t>(o.max_iter, 0xffffffffu));
    }
    KT_TRY(writeCtl(b));
    const auto t0 = std::chrono::steady_clock::now();
    const auto elapsed = [&] {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    };
    std::vector<double> doneMs(S, -1.0);
    KT_TRY(startAll(b, a));
    const uint32_t every = o.check_every > 0 ? static_cast<uint32_t>(o.check_every) : 32;
    uint64_t restarts = 0;
    for (int round = 0;; ++round) {
        if (round > 0) KT_TRY(startAll(b, a));
        for (;;) {
            KT_TRY(readCtl(b));
            const double now = elapsed();
            bool any = false;
            for (size_t s = 0; s < S; ++s) {
                const bool live = b->ctlHost[s].s[0].live != 0;
                any = any || live;
                if (!live && doneMs[s] < 0.0) doneMs[s] = now;
                if (live) doneMs[s] = -1.0;
            }
            if (!any) break;
            KT_TRY(runPairs(b, a, (every + 1) / 2));
        }
        KT_BATCH_G(b->lanes, bTrueResidual, dim3(b->tiles), b->stream, a);
        KT_HIP(hipGetLastError());
        bTrueResidualFinish<<<dim3(static_cast<uint32_t>(S)), dim3(kBB), 0, b->stream>>>(a);
        KT_HIP(hipGetLastError());
        KT_TRY(readCtl(b));
        if (o.stop_rule != KT_STOP_TRUE || round >= 16) break;
        bool again = false;
        for (size_t s = 0; s < S; ++s) {
            AxisCtl &c = b->ctlHost[s];
            const bool more = c.s[0].active && c.trueResid > o.tol && c.s[0].steps < c.maxIter;
            c.frozen = more ? 0 : 1;
            if (more) doneMs[s] = -1.0;
            again = again || more;
        }
        if (!again) break;
        ++restarts;
        KT_TRY(writeCtl(b));
    }
    const double end = elapsed();
    kt_solve_result tot{};
    for (size_t inst = 0; inst < S / 2; ++inst) {
        kt_batch_instance_result r{};
        double done = 0.0;
        for (int ax = 0; ax < 2; ++ax) {
            const AxisCtl &c = b->ctlHost[2 * inst + ax];
            r.resid[ax] = c.s[0].resid;
            r.true_resid[ax] = c.trueResid;
            r.steps[ax] = c.s[0].steps;
            done = std::max(done, doneMs[2 * inst + ax] < 0.0 ? end : doneMs[2 * inst + ax]);
            tot.steps[ax] += c.s[0].steps;
            tot.resid[ax] = std::max(tot.resid[ax], c.s[0].resid);
            tot.true_resid[ax] = std::max(tot.true_resid[ax], c.trueResid);
            tot.iterations = std::max<uint64_t>(tot.iterations, c.s[0].steps);
        }
        r.done_ms = done;
        if (out) out[inst] = r;
    }
    tot.refinements = restarts;
    if (total) *total = tot;
    return KT_OK;
}

// Appends one instance from host arrays (row_ptr u64 per axis).
int batchAdd(kt_batch *b, const kt_system_view *sv[2], const double *x0[2], uint32_t *index) {
    KT_CHECK(sv[0]->rows == sv[1]->rows, KT_EINVAL, "kt_batch_add: x and y systems differ in size");
    const uint64_t rows = sv[0]->rows;
    for (int ax = 0; ax < 2; ++ax) {
        const kt_system_view &v = *sv[ax];
        KT_CHECK(v.row_ptr && v.diag && v.rhs && x0[ax], KT_EINVAL, "kt_batch_add: null array");
        KT_CHECK(v.row_ptr[0] == 0, KT_EINVAL, "kt_batch_add: row_ptr[0] must be 0");
        for (uint64_t i = 0; i < rows; ++i)
            KT_CHECK(v.row_ptr[i + 1] >= v.row_ptr[i], KT_EINVAL, "kt_batch_add: row_ptr must not decrease");
        const uint64_t nnz = v.row_ptr[rows];
        KT_CHECK(nnz == 0 || (v.col && v.val), KT_EINVAL, "kt_batch_add: null col/val");
        for (uint64_t k = 0; k < nnz; ++k)
            KT_CHECK(v.col[k] < rows, KT_EINVAL, "kt_batch_add: column index out of range");
        KT_CHECK(b->diag.size() + rows < 0xffffffffull && b->col.size() + nnz < 0xffffffffull, KT_EINVAL,
                 "kt_batch_add: batch exceeds 2^32 rows or entries");
    }
    for (int ax = 0; ax < 2; ++ax) {
        const kt_system_view &v = *sv[ax];
        const uint32_t rowBase = static_cast<uint32_t>(b->diag.size());
        const uint32_t nnzBase = static_cast<uint32_t>(b->col.size());
        const uint64_t nnz = v.row_ptr[rows];
        b->sys.push_back(HostSys{rowBase, static_cast<uint32_t>(rows), nnzBase, static_cast<uint32_t>(nnz)});
        for (uint64_t i = 0; i < rows; ++i) b->rowPtr.push_back(static_cast<uint32_t>(nnzBase + v.row_ptr[i + 1]));
        for (uint64_t k = 0; k < nnz; ++k) b->col.push_back(rowBase + v.col[k]);
        b->val.insert(b->val.end(), v.val, v.val + nnz);
        b->diag.insert(b->diag.end(), v.diag, v.diag + rows);
        b->rhs.insert(b->rhs.end(), v.rhs, v.rhs + rows);
        b->x0.insert(b->x0.end(), x0[ax], x0[ax] + rows);
    }
    b->dirty = true;
    if (index) *index = static_cast<uint32_t>(b->sys.size() / 2 - 1);
    return KT_OK;
}

}  // namespace

int getSystemHost(kt_ctx *ctx, int axis, std::vector<uint64_t> &rowPtr, std::vector<uint32_t> &col,
                  std::vector<double> &val, std::vector<double> &diag, std::vector<double> &rhs);
int getPositionsHost(kt_ctx *ctx, std::vector<double> &x, std::vector<double> &y);

}  // namespace kt

using namespace kt;

#define KT_BATCH_GUARD(b)                                                   \
    do {                                                                    \
        if ((b) == nullptr) {                                               \
            setError("null batch");                                         \
            return KT_EINVAL;                                               \
        }                                                                   \
        hipError_t kt_se_ = hipSetDevice((b)->device);                      \
        if (kt_se_ != hipSuccess) return hipFail(kt_se_, "hipSetDevice", __FILE__, __LINE__); \
    } while (0)

extern "C" {

int kt_batch_create(int device, kt_batch **out) {
    KT_CHECK(out, KT_EINVAL, "kt_batch_create: null output");
    *out = nullptr;
    int count = 0;
    KT_HIP(hipGetDeviceCount(&count));
    KT_CHECK(device >= 0 && device < count, KT_EINVAL, "kt_batch_create: no such device");
    KT_HIP(hipSetDevice(device));
    auto *b = new kt_batch();
    b->device = device;
    const hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete b;
        return hipFail(e, "hipStreamCreate", __FILE__, __LINE__);
    }
    *out = b;
    return KT_OK;
}

int kt_batch_destroy(kt_batch *b) {
    if (!b) return KT_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    dropGraph(b);
    if (b->ctlHost) (void)hipHostFree(b->ctlHost);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;  // device buffers free themselves
    return KT_OK;
}

int kt_batch_clear(kt_batch *b) {
    KT_BATCH_GUARD(b);
    b->sys.clear();
    b->rowPtr.assign(1, 0);
    b->col.clear();
    b->val.clear();
    b->diag.clear();
    b->rhs.clear();
    b->x0.clear();
    b->dirty = true;
    return KT_OK;
}

int kt_batch_add(kt_batch *b, const kt_system_view *sys_x, const kt_system_view *sys_y, const double *x0,
                 const double *y0, uint32_t *index) {
    KT_BATCH_GUARD(b);
    KT_CHECK(sys_x && sys_y, KT_EINVAL, "kt_batch_add: null system");
    const kt_system_view *sv[2] = {sys_x, sys_y};
    const double *xs[2] = {x0, y0};
    try {
        return batchAdd(b, sv, xs, index);
    } catch (const std::exception &e) {
        setError(e.what());
        return KT_ENOMEM;
    }
}

int kt_batch_add_ctx(kt_batch *b, kt_ctx *ctx, uint32_t *index) {
    KT_BATCH_GUARD(b);
    KT_CHECK(ctx, KT_EINVAL, "kt_batch_add_ctx: null context");
    try {
        std::vector<uint64_t> rp[2];
        std::vector<uint32_t> col[2];
        std::vector<double> val[2], diag[2], rhs[2], pos[2];
        for (int ax = 0; ax < 2; ++ax) KT_TRY(getSystemHost(ctx, ax, rp[ax], col[ax], val[ax], diag[ax], rhs[ax]));
        KT_TRY(getPositionsHost(ctx, pos[0], pos[1]));
        KT_HIP(hipSetDevice(b->device));
        kt_system_view v[2];
        for (int ax = 0; ax < 2; ++ax)
            v[ax] = kt_system_view{diag[ax].size(), rp[ax].data(), col[ax].data(), val[ax].data(), diag[ax].data(),
                                   rhs[ax].data()};
        KT_CHECK(pos[0].size() == diag[0].size(), KT_ESTATE, "kt_batch_add_ctx: positions do not match the systems");
        const kt_system_view *sv[2] = {&v[0], &v[1]};
        const double *xs[2] = {pos[0].data(), pos[1].data()};
        return batchAdd(b, sv, xs, index);
    } catch (const std::exception &e) {
        setError(e.what());
        return KT_ENOMEM;
    }
}

int kt_batch_size(kt_batch *b, uint32_t *instances, uint64_t *rows, uint64_t *nnz) {
    KT_BATCH_GUARD(b);
    if (instances) *instances = static_cast<uint32_t>(b->sys.size() / 2);
    if (rows) *rows = b->diag.size();
    if (nnz) *nnz = b->col.size();
    return KT_OK;
}

int kt_batch_set_tuning(kt_batch *b, int lanes, uint32_t tile, int rows_in_flight, int use_graphs) {
    KT_BATCH_GUARD(b);
    KT_CHECK(lanes == 0 || lanes == 4 || lanes == 8 || lanes == 16, KT_EINVAL, "lanes must be 4, 8 or 16");
    KT_CHECK(tile == 0 || (tile % kBB == 0 && tile <= (1u << 20)), KT_EINVAL, "tile must be a multiple of 256");
    KT_CHECK(rows_in_flight == 0 || rows_in_flight == 2 || rows_in_flight == 4, KT_EINVAL, "rows must be 2 or 4");
    if (lanes) b->lanes = lanes;
    if (tile) b->tile = tile;
    if (rows_in_flight) b->rows = rows_in_flight;
    b->useGraphs = use_graphs != 0;
    return KT_OK;
}

int kt_batch_set_layout(kt_batch *b, int layout, int row_lanes) {
    KT_BATCH_GUARD(b);
    KT_CHECK(layout == KT_SPMV_ROW_GROUPS || layout == KT_SPMV_ENTRY_BLOCKS, KT_EINVAL, "unknown SpMV layout");
    KT_CHECK(row_lanes == 0 || row_lanes == 1 || row_lanes == 4, KT_EINVAL, "row lanes must be 0, 1 or 4");
    b->layout = layout;
    b->rowLanes = row_lanes == 4 ? 4 : 1;
    return KT_OK;
}

int kt_batch_set_columns(kt_batch *b, int bytes) {
    KT_BATCH_GUARD(b);
    KT_CHECK(bytes == 0 || bytes == 3 || bytes == 4, KT_EINVAL, "kt_batch_set_columns: bytes must be 0, 3 or 4");
    b->colBytes = bytes;
    return KT_OK;
}

int kt_batch_set_wide_loads(kt_batch *b, int on) {
    KT_BATCH_GUARD(b);
    KT_CHECK(on == 0 || on == 1, KT_EINVAL, "kt_batch_set_wide_loads: on must be 0 or 1");
    b->wide = on;
    return KT_OK;
}

int kt_batch_reset(kt_batch *b) {
    KT_BATCH_GUARD(b);
    KT_TRY(upload(b));
    const uint64_t rows = b->diag.size();
    if (rows) KT_HIP(hipMemcpyAsync(b->dX.p, b->dX0.p, rows * sizeof(double), hipMemcpyDeviceToDevice, b->stream));
    KT_HIP(hipStreamSynchronize(b->stream));
    return KT_OK;
}

int kt_batch_solve(kt_batch *b, const kt_solve_opts *opts, kt_batch_instance_result *per_instance,
                   kt_solve_result *total) {
    KT_BATCH_GUARD(b);
    KT_CHECK(opts, KT_EINVAL, "kt_batch_solve: null options");
    try {
        return batchSolve(b, *opts, per_instance, total);
    } catch (const std::exception &e) {
        setError(e.what());
        return KT_ENOMEM;
    }
}

int kt_batch_get_positions(kt_batch *b, uint32_t index, double *x, double *y, uint64_t count) {
    KT_BATCH_GUARD(b);
    KT_CHECK(index < b->sys.size() / 2, KT_EINVAL, "kt_batch_get_positions: no such instance");
    KT_CHECK(x && y, KT_EINVAL, "kt_batch_get_positions: null array");
    const HostSys &sx = b->sys[2 * index], &sy = b->sys[2 * index + 1];
    KT_CHECK(count == sx.rows, KT_EINVAL, "kt_batch_get_positions: count differs from the instance's rows");
    KT_TRY(upload(b));
    if (count) {
        KT_HIP(hipMemcpyAsync(x, b->dX.p + sx.rowBase, count * sizeof(double), hipMemcpyDeviceToHost, b->stream));
        KT_HIP(hipMemcpyAsync(y, b->dX.p + sy.rowBase, count * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    }
    KT_HIP(hipStreamSynchronize(b->stream));
    return KT_OK;
}

}  // extern "C"
