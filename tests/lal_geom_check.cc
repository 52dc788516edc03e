// Test harness (CPU): the product's host geometry of look-ahead legalization
// (ktplace_amd/csrc/lal_geom.cc) on a design read from stdin.  Prints the
// grid's free area per bin, then for each query the free-area profile of a
// box or the stripes of a half block, as the bit patterns of the doubles, for
// tests/test_lal_geom_host.py to compare with the oracle.
//
// stdin: R, then R lines "y height x0 x1 pitch"; the die "x0 y0 x1 y1"; K,
// then K obstacle lines "x0 y0 x1 y1"; bins per axis; Q, then Q queries,
// "0 x0 y0 x1 y1 vertical" (profile) or "1 x0 y0 x1 y1 vertical cellArea
// regionRoom densityLimit stripeAreaFraction cutline belowCut" (stripes), or
// "2 densityLimit occ[0] ... occ[bins-1]" (overfilled clusters of that
// occupancy, printed as x0 x1 y0 y1 per cluster in decimal).  Kinds 3 and 4
// are kinds 0 and 1 computed the way the fused GPU batch kernel does: the
// sequential steps of lal_seq.h on flat row, span and obstacle arrays, with
// the steps and cuts sorted as value pairs.
// Numbers are hex floats (%a) so they arrive exact.
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "lal_geom.h"
#include "lal_seq.h"

#include <algorithm>
#include <utility>

using namespace kt::lal;

static double num() {
    char buf[128];
    if (std::scanf("%127s", buf) != 1) std::exit(2);
    return std::strtod(buf, nullptr);
}

static void put(double v);

// FreeSpace::profile on the flat arrays, as lal_batch.hip computes it.
static kt::lalseq::FlatProfile flatProfile(const FreeSpace &fs, const Box &b, bool vertical, std::vector<double> &edges,
                                       std::vector<double> &prefix) {
    using namespace kt::lalseq;
    const size_t rows = fs.rows.size();
    size_t first = 0;
    {  // upper_bound of b.y0 on the rows' top edges
        size_t lo = 0, hi = rows;
        while (lo < hi) {
            const size_t mid = (lo + hi) / 2;
            if (b.y0 < fs.rows[mid].y1) hi = mid; else lo = mid + 1;
        }
        first = lo;
    }
    size_t last = first;
    while (last < rows && fs.rows[last].y0 < b.y1) ++last;
    const double lo = vertical ? b.x0 : b.y0;
    const double hi = vertical ? b.x1 : b.y1;
    size_t spans = 0;
    for (size_t r = first; r < last; ++r) spans += fs.rows[r].spans.size();
    edges.assign(2 * spans + 2 * (last - first) + 4, 0.0);
    prefix.assign(edges.size(), 0.0);
    FlatProfile p{edges.data(), prefix.data(), 0, static_cast<uint32_t>(edges.size()), true};
    profileStart(p, lo);
    if (vertical) {
        std::vector<std::pair<double, double>> steps;
        for (size_t r = first; r < last; ++r) {
            const double h = smin(b.y1, fs.rows[r].y1) - smax(b.y0, fs.rows[r].y0);
            for (const Span &sp : fs.rows[r].spans) {
                const double a = smax(sp.a, lo), z = smin(sp.b, hi);
                if (h > 0.0 && z > a) {
                    steps.emplace_back(a, h);
                    steps.emplace_back(z, -h);
                }
            }
        }
        std::sort(steps.begin(), steps.end());
        std::vector<double> at, dh;
        for (const auto &[x, v] : steps) {
            at.push_back(x);
            dh.push_back(v);
        }
        profileSweepSteps(p, at.data(), dh.data(), static_cast<uint32_t>(at.size()));
    } else {
        std::vector<double> ra, rz, rl;
        for (size_t r = first; r < last; ++r) {
            const double a = smax(lo, fs.rows[r].y0), z = smin(hi, fs.rows[r].y1);
            if (!(z > a)) continue;
            double len = 0.0;
            for (const Span &sp : fs.rows[r].spans) len += smax(0.0, smin(b.x1, sp.b) - smax(b.x0, sp.a));
            ra.push_back(a);
            rz.push_back(z);
            rl.push_back(len);
        }
        profileSweepRows(p, ra.data(), rz.data(), rl.data(), static_cast<uint32_t>(ra.size()));
    }
    profileFinish(p, hi);
    if (!p.ok) std::exit(3);
    return p;
}

// stripes() on the flat arrays, as lal_batch.hip computes them.
static std::vector<Stripe> flatStripes(const FreeSpace &fs, const std::vector<Box> &obstacles, const Box &box,
                                       bool vertical, double cellArea, double regionRoom, double densityLimit,
                                       double fraction, double cutline, bool belowCut) {
    using namespace kt::lalseq;
    std::vector<double> edges, prefix;
    const FlatProfile room = flatProfile(fs, box, vertical, edges, prefix);
    const double lo = vertical ? box.x0 : box.y0;
    const double hi = vertical ? box.x1 : box.y1;
    std::vector<double> cuts{lo, hi};
    for (const Box &o : obstacles) {
        if (o.x1 <= box.x0 || o.x0 >= box.x1 || o.y1 <= box.y0 || o.y0 >= box.y1) continue;
        const double e0 = vertical ? o.x0 : o.y0, e1 = vertical ? o.x1 : o.y1;
        if (e0 > lo && e0 < hi) cuts.push_back(e0);
        if (e1 > lo && e1 < hi) cuts.push_back(e1);
    }
    std::sort(cuts.begin(), cuts.end());
    cuts.erase(std::unique(cuts.begin(), cuts.end()), cuts.end());
    const double total = profileTotal(room);
    const double limit = fraction * (regionRoom > 0.0 ? regionRoom : total);
    const double fill = total > 0.0 ? smin(densityLimit, cellArea / total) : densityLimit;
    std::vector<double> fin(4096), stack(256);
    const uint32_t nf = stripeCuts(room, cuts.data(), static_cast<uint32_t>(cuts.size()), limit, fin.data(),
                                   static_cast<uint32_t>(fin.size()), stack.data(), static_cast<uint32_t>(stack.size()));
    if (nf == 0) std::exit(4);
    const uint32_t count = nf - 1;
    std::vector<double> fromCut(count);
    for (uint32_t s = 0; s < count; ++s) fromCut[s] = stripeFromCut(fin[s], fin[s + 1], cutline, belowCut);
    std::vector<Stripe> out(count);
    for (uint32_t s = 0; s < count; ++s)
        out[stripeRank(fromCut.data(), count, s)] = {fin[s], fin[s + 1], fill * profileBetween(room, fin[s], fin[s + 1])};
    return out;
}

static void put(double v) {
    uint64_t b;
    std::memcpy(&b, &v, 8);
    std::printf(" %016" PRIx64, b);
}

int main() {
    std::vector<RowSpec> rows(static_cast<size_t>(num()));
    for (RowSpec &r : rows) r = {num(), num(), num(), num(), num()};
    Box die{};
    die.x0 = num();
    die.y0 = num();
    die.x1 = num();
    die.y1 = num();
    std::vector<Box> obstacles(static_cast<size_t>(num()));
    for (Box &o : obstacles) {
        o.x0 = num();
        o.y0 = num();
        o.x1 = num();
        o.y1 = num();
    }
    const size_t perAxis = static_cast<size_t>(num());
    Grid g;
    FreeSpace fs;
    buildGrid(g, fs, rows, obstacles, die, perAxis);
    std::printf("avail");
    for (const double a : g.avail) put(a);
    std::printf("\n");
    const size_t q = static_cast<size_t>(num());
    for (size_t k = 0; k < q; ++k) {
        const double kind = num();
        if (kind == 2.0) {
            const double limit = num();
            std::vector<double> occ(g.size());
            for (double &v : occ) v = num();
            std::printf("clusters");
            for (const BinRect &r : overfilledClusters(g, occ, limit))
                std::printf(" %zu %zu %zu %zu", r.x0, r.x1, r.y0, r.y1);
            std::printf("\n");
            continue;
        }
        const bool isStripes = kind == 1.0 || kind == 4.0;
        const bool flat = kind >= 3.0;
        Box b{};
        b.x0 = num();
        b.y0 = num();
        b.x1 = num();
        b.y1 = num();
        const bool vertical = num() != 0.0;
        if (isStripes) {
            const double cellArea = num(), regionRoom = num(), densityLimit = num(), fraction = num();
            const double cutline = num();
            const bool belowCut = num() != 0.0;
            std::printf("stripes");
            const std::vector<Stripe> got =
                flat ? flatStripes(fs, obstacles, b, vertical, cellArea, regionRoom, densityLimit, fraction, cutline,
                                   belowCut)
                     : stripes(fs, obstacles, b, vertical, cellArea, regionRoom, densityLimit, fraction, cutline, belowCut);
            for (const Stripe &s : got) {
                put(s.start);
                put(s.end);
                put(s.capacity);
            }
            std::printf("\n");
            continue;
        }
        if (flat) {
            std::vector<double> edges, prefix;
            const kt::lalseq::FlatProfile p = flatProfile(fs, b, vertical, edges, prefix);
            std::printf("edges");
            for (uint32_t k = 0; k < p.n; ++k) put(p.edges[k]);
            std::printf("\nprefix");
            for (uint32_t k = 0; k < p.n; ++k) put(p.prefix[k]);
            std::printf("\n");
            continue;
        }
        const Profile p = fs.profile(b, vertical);
        std::printf("edges");
        for (const double e : p.edges) put(e);
        std::printf("\nprefix");
        for (const double v : p.prefix) put(v);
        std::printf("\n");
    }
    return 0;
}
