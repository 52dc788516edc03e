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
// occupancy, printed as x0 x1 y0 y1 per cluster in decimal).
// Numbers are hex floats (%a) so they arrive exact.
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "lal_geom.h"

using namespace kt::lal;

static double num() {
    char buf[128];
    if (std::scanf("%127s", buf) != 1) std::exit(2);
    return std::strtod(buf, nullptr);
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
        const bool isStripes = kind != 0.0;
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
            for (const Stripe &s :
                 stripes(fs, obstacles, b, vertical, cellArea, regionRoom, densityLimit, fraction, cutline, belowCut)) {
                put(s.start);
                put(s.end);
                put(s.capacity);
            }
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
