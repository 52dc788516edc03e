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
constexpr uint32_t kNoCol = 0xffffffffu;

template <typename T, int G, int R>
__device__ __forceinline__ void rowsDot(const uint32_t *__restrict__ rowPtr, const uint32_t *__restrict__ col,
                                        const T *__restrict__ val, const T *__restrict__ v,
                                        const uint32_t (&row)[R], const bool (&valid)[R], uint32_t lane,
                                        double (&out)[R]) {
    constexpr int U = G >= 16 ? 1 : 16 / G;
    uint32_t b[R], e[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        b[r] = valid[r] ? rowPtr[row[r]] : 0u;
        e[r] = valid[r] ? rowPtr[row[r] + 1] : 0u;
        out[r] = 0.0;
    }
    uint32_t c[R][U];
    T w[R][U];
#pragma unroll
    for (int r = 0; r < R; ++r) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t k = b[r] + lane + u * G;
            const bool ok = k < e[r];
            c[r][u] = ok ? col[k] : kNoCol;
            w[r][u] = ok ? val[k] : T(0);
        }
    }
    T g[R][U];
#pragma unroll
    for (int r = 0; r < R; ++r) {
#pragma unroll
        for (int u = 0; u < U; ++u) g[r][u] = c[r][u] != kNoCol ? v[c[r][u]] : T(0);
    }
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
            for (int u = 0; u < U; ++u) {
                const uint32_t k = base + lane + u * G;
                const bool ok = k < e[r];
                cc[u] = ok ? col[k] : kNoCol;
                ww[u] = ok ? val[k] : T(0);
            }
            T gg[U];
#pragma unroll
            for (int u = 0; u < U; ++u) gg[u] = cc[u] != kNoCol ? v[cc[u]] : T(0);
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
