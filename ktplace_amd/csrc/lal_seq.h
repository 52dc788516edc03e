// Sequential steps of look-ahead legalization's block geometry on flat
// arrays: the free-area profile sweep and its two queries (simpl_grid.h:
// 36-64, 179-231) and the stripe subdivision and fill order of getStripes
// (simpl_lal.cc:213-263, 275-282).  The same code runs on the host and,
// lane-uniform, inside one GPU wave of the fused batch kernel (lal_batch.hip),
// where the parallel parts around it (collecting and sorting the steps and the
// obstacle cuts) are done by the wave's lanes.  Every floating-point
// expression and its order follow the reference lines, as lal_geom.cc does:
// the GPU's legalized positions are compared with the reference's bit for bit.
//
// Those expressions follow KTPlace (simpl_grid.h, simpl_lal.cc), MIT License,
// Copyright (c) 2026 Coherent17; the full notice is in lal_geom.cc.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define KT_HD __host__ __device__ __forceinline__
#else
#define KT_HD inline
#endif

namespace kt::lalseq {

// std::max / std::min, with their exact selection rule (the first argument
// wins ties, so signed zeros come out as the reference's do).
KT_HD double smax(double a, double b) { return (a < b) ? b : a; }
KT_HD double smin(double a, double b) { return (b < a) ? b : a; }

// A profile under construction or built: edges[0, n) ascending, prefix[k] the
// free area before edges[k]; cap is the room of both arrays.
struct FlatProfile {
    double *edges;
    double *prefix;
    uint32_t n;
    uint32_t cap;
    bool ok;  // false once an append did not fit
};

KT_HD void profileStart(FlatProfile &p, double lo) {
    p.edges[0] = lo;
    p.prefix[0] = 0.0;
    p.n = 1;
    p.ok = p.cap >= 2;
}

// `extend` of getProfile: a new edge only when `to` passes the last one.
KT_HD void profileExtend(FlatProfile &p, double to, double area) {
    const double back = p.edges[p.n - 1];
    if (to > back) {
        if (p.n >= p.cap) {
            p.ok = false;
            return;
        }
        p.edges[p.n] = to;
        p.prefix[p.n] = p.prefix[p.n - 1] + area;
        ++p.n;
    }
}

KT_HD void profileFinish(FlatProfile &p, double hi) {
    profileExtend(p, hi, 0.0);
    if (p.ok && p.n == 1) {
        p.edges[1] = hi;
        p.prefix[1] = 0.0;
        p.n = 2;
    }
}

// Vertical profile: the sweep over the (at, dh) steps sorted as pairs.
KT_HD void profileSweepSteps(FlatProfile &p, const double *at, const double *dh, uint32_t steps) {
    double height = 0.0;
    for (uint32_t s = 0; s < steps && p.ok; ++s) {
        profileExtend(p, at[s], height * (at[s] - p.edges[p.n - 1]));
        height += dh[s];
    }
}

// Horizontal profile: rows in ascending order whose clipped extent [a, z) is
// not empty, with len the free length of the row inside the box.
KT_HD void profileSweepRows(FlatProfile &p, const double *a, const double *z, const double *len, uint32_t rows) {
    for (uint32_t r = 0; r < rows && p.ok; ++r) {
        profileExtend(p, a[r], 0.0);
        profileExtend(p, z[r], len[r] * (z[r] - a[r]));
    }
}

KT_HD double profileTotal(const FlatProfile &p) { return p.prefix[p.n - 1]; }

// AxisProfile::getBefore: free area before coordinate c.
KT_HD double profileBefore(const FlatProfile &p, double c) {
    if (c <= p.edges[0]) return 0.0;
    if (c >= p.edges[p.n - 1]) return profileTotal(p);
    uint32_t lo = 0, hi = p.n;  // upper_bound: first edge > c
    while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2;
        if (c < p.edges[mid]) hi = mid; else lo = mid + 1;
    }
    const uint32_t k = lo - 1;
    const double f = (c - p.edges[k]) / (p.edges[k + 1] - p.edges[k]);
    return p.prefix[k] + f * (p.prefix[k + 1] - p.prefix[k]);
}

KT_HD double profileBetween(const FlatProfile &p, double a, double b) { return profileBefore(p, b) - profileBefore(p, a); }

// AxisProfile::getCoordinateAt: coordinate before which `target` area lies.
KT_HD double profileCoordinateAt(const FlatProfile &p, double target) {
    if (target <= 0.0) return p.edges[0];
    if (target >= profileTotal(p)) return p.edges[p.n - 1];
    uint32_t lo = 0, hi = p.n;  // lower_bound: first prefix >= target
    while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2;
        if (p.prefix[mid] < target) lo = mid + 1; else hi = mid;
    }
    const uint32_t k = lo;
    const double step = p.prefix[k] - p.prefix[k - 1];
    const double f = step > 0.0 ? (target - p.prefix[k - 1]) / step : 1.0;
    return p.edges[k - 1] + f * (p.edges[k] - p.edges[k - 1]);
}

// getStripes' subdivision: the sorted distinct cuts[0, nc), each gap whose
// free area exceeds `limit` split at its free-area midpoint, left part first,
// until no gap splits.  Written as a depth-first walk with the pending right
// ends on `stack`: it visits the gaps in the order of the reference's
// insert-and-recheck loop and emits its final cut list into fin.  Returns the
// number of final cuts, 0 when fin or stack would overflow.
KT_HD uint32_t stripeCuts(const FlatProfile &room, const double *cuts, uint32_t nc, double limit, double *fin,
                          uint32_t finCap, double *stack, uint32_t stackCap) {
    if (nc == 0 || finCap == 0) return 0;
    fin[0] = cuts[0];
    uint32_t nf = 1;
    double cur = cuts[0];
    for (uint32_t i = 1; i < nc; ++i) {
        double c = cuts[i];
        uint32_t sp = 0;
        for (;;) {
            const double beforeCur = profileBefore(room, cur);
            const double here = profileBefore(room, c) - beforeCur;
            const double mid = profileCoordinateAt(room, beforeCur + 0.5 * here);
            if (here > limit && mid > cur && mid < c) {
                if (sp >= stackCap) return 0;
                stack[sp++] = c;
                c = mid;
            } else {
                if (nf >= finCap) return 0;
                fin[nf++] = c;
                cur = c;
                if (sp == 0) break;
                c = stack[--sp];
            }
        }
    }
    return nf;
}

// Distance of a stripe's midpoint from the cut, away from it positive
// (nonlinearScale's fromCut, simpl_lal.cc:275-282).
KT_HD double stripeFromCut(double start, double end, double cutline, bool belowCut) {
    const double c = 0.5 * (start + end);
    return belowCut ? cutline - c : c - cutline;
}

// Position of stripe i in fill order: farthest from the cut first, ties in
// coordinate order (a stable sort on fromCut, descending).
KT_HD uint32_t stripeRank(const double *fromCut, uint32_t count, uint32_t i) {
    uint32_t rank = 0;
    for (uint32_t j = 0; j < count; ++j)
        rank += (fromCut[j] > fromCut[i] || (j < i && fromCut[j] == fromCut[i])) ? 1u : 0u;
    return rank;
}

}  // namespace kt::lalseq
