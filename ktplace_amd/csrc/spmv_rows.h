// Row products shared by the CG kernels (csrc/cg.hip, csrc/batch.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace kt {

// Off-diagonal parts of R rows times v, G lanes per row, FP64 accumulation.
// All loads of a round are issued before any is consumed: the rowPtr pairs
// of the R rows, then U = 16 / G column indices and values per lane and row,
// then the U gathers of v, so a typical row (up to 16 entries) costs three
// dependent memory latencies instead of one per entry.  Longer rows loop over
// further rounds.  The summation order (lane-strided rounds, then a fixed
// butterfly over the G lanes) is fixed, so results are deterministic.
//
// The gathered vector is either a plain pointer or a Window: a contiguous
// slice [lo, lo + span) of it staged in LDS, read from there, with columns
// outside the slice read from the vector itself.  Both give the same values,
// so the results do not depend on the staging.
constexpr uint32_t kNoCol = 0xffffffffu;

// An LDS pointer (address space 3), so that window reads compile to ds_read
// and not to flat loads.
template <typename T>
using LdsPtr = const __attribute__((address_space(3))) T *;

template <typename T>
struct Window {
    const T *__restrict__ v;  // the whole vector (global memory)
    LdsPtr<T> lds;            // v[lo .. lo + span) in LDS
    uint32_t lo, span;
};

// Every load below is unconditional: an index that must not be read is
// replaced by 0 (or the window start) and the loaded value discarded with a
// select.  A conditional load (`ok ? a[k] : 0`) compiles to an exec-masked
// branch region, and the compiler then waits for each such load before the
// next region, which serialises the loads this routine means to overlap.
// Every array read here has an element 0: a launched tile has a row, and the
// CSR arrays are allocated with at least one entry.

// v[c], or 0 for c == kNoCol.
template <typename T>
__device__ __forceinline__ T gatherAt(const T *__restrict__ v, uint32_t c) {
    const bool ok = c != kNoCol;
    const T g = v[ok ? c : 0u];
    return ok ? g : T(0);
}

// Every lane issues one LDS read and one global load.  Lanes inside the
// window (and kNoCol lanes) point their global load at v[lo], one line shared
// by the whole wave; lanes outside read LDS word 0.
template <typename T>
__device__ __forceinline__ T gatherAt(const Window<T> &w, uint32_t c) {
    const uint32_t k = c - w.lo;
    const bool in = k < w.span;
    const bool far = !in && c != kNoCol;
    const T near = w.lds[in ? k : 0u];
    const T away = w.v[far ? c : w.lo];
    return in ? near : (far ? away : T(0));
}

// v[i] for a row i of the tile (always inside a window).
template <typename T>
__device__ __forceinline__ T ownAt(const T *__restrict__ v, uint32_t i) {
    return v[i];
}

template <typename T>
__device__ __forceinline__ T ownAt(const Window<T> &w, uint32_t i) {
    return w.lds[i - w.lo];
}

// Column index and value of entry k of a row ending at e (kNoCol and 0 past it).
template <typename T>
__device__ __forceinline__ void entryAt(const uint32_t *__restrict__ col, const T *__restrict__ val, uint32_t k,
                                        uint32_t e, uint32_t &c, T &w) {
    const bool ok = k < e;
    const uint32_t kk = ok ? k : 0u;
    const uint32_t cl = col[kk];
    const T wl = val[kk];
    c = ok ? cl : kNoCol;
    w = ok ? wl : T(0);
}

template <typename T, int G, int R, typename Src>
__device__ __forceinline__ void rowsDot(const uint32_t *__restrict__ rowPtr, const uint32_t *__restrict__ col,
                                        const T *__restrict__ val, const Src &v,
                                        const uint32_t (&row)[R], const bool (&valid)[R], uint32_t lane,
                                        double (&out)[R]) {
    constexpr int U = G >= 16 ? 1 : 16 / G;
    uint32_t b[R], e[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint32_t i = valid[r] ? row[r] : 0u;
        const uint32_t bl = rowPtr[i], el = rowPtr[i + 1];
        b[r] = valid[r] ? bl : 0u;
        e[r] = valid[r] ? el : 0u;
        out[r] = 0.0;
    }
    // the scheduling barriers keep the three load phases apart, so all loads
    // of one phase are in flight together
    __builtin_amdgcn_sched_barrier(0);
    uint32_t c[R][U];
    T w[R][U];
#pragma unroll
    for (int r = 0; r < R; ++r) {
#pragma unroll
        for (int u = 0; u < U; ++u) entryAt(col, val, b[r] + lane + u * G, e[r], c[r][u], w[r][u]);
    }
    __builtin_amdgcn_sched_barrier(0);
    T g[R][U];
#pragma unroll
    for (int r = 0; r < R; ++r) {
#pragma unroll
        for (int u = 0; u < U; ++u) g[r][u] = gatherAt(v, c[r][u]);
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int r = 0; r < R; ++r) {
#pragma unroll
        for (int u = 0; u < U; ++u) out[r] += static_cast<double>(w[r][u]) * static_cast<double>(g[r][u]);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        for (uint32_t base = b[r] + G * U; base < e[r]; base += G * U) {
            uint32_t cc[U];
            T ww[U];
#pragma unroll
            for (int u = 0; u < U; ++u) entryAt(col, val, base + lane + u * G, e[r], cc[u], ww[u]);
            T gg[U];
#pragma unroll
            for (int u = 0; u < U; ++u) gg[u] = gatherAt(v, cc[u]);
#pragma unroll
            for (int u = 0; u < U; ++u) out[r] += static_cast<double>(ww[u]) * static_cast<double>(gg[u]);
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1) out[r] += __shfl_xor(out[r], off, G);
    }
}

}  // namespace kt
