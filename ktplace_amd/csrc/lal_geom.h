// Geometry of look-ahead legalization (host side): bin grid, free space per
// row, free-area profiles along an axis, stripes, overfilled clusters and
// their growth.  Restates simpl_grid.h and simpl_lal.cc:185-363; the per-cell
// work (sorting, prefix sums, stripe filling, scaling) runs on the GPU.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace kt::lal {

struct Span {
    double a = 0.0, b = 0.0;
    bool operator<(const Span &o) const { return a < o.a || (a == o.a && b < o.b); }
};

struct Box {
    double x0 = 0.0, y0 = 0.0, x1 = 0.0, y1 = 0.0;
};

struct BinRect {
    size_t x0 = 0, x1 = 0, y0 = 0, y1 = 0;
    size_t bins() const { return (x1 - x0 + 1) * (y1 - y0 + 1); }
};

// Free area per unit length along an axis (simpl_grid.h:23-65).
struct Profile {
    std::vector<double> edges, prefix;
    double total() const { return prefix.back(); }
    double before(double c) const;
    double between(double a, double b) const { return before(b) - before(a); }
    double coordinateAt(double target) const;
};

struct FreeRow {
    double y0 = 0.0, y1 = 0.0;
    std::vector<Span> spans;
};

// Row spans minus obstacles (simpl_grid.h:130-178).
struct FreeSpace {
    std::vector<FreeRow> rows;
    void build(std::vector<FreeRow> raw, const std::vector<Box> &obstacles);
    Profile profile(const Box &b, bool vertical) const;
};

// Bin grid with free area per bin (simpl_grid.h:67-128, simpl_placer.cc:190-224).
struct Grid {
    double ox = 0.0, oy = 0.0, dx = 1.0, dy = 1.0;
    size_t nx = 1, ny = 1;
    std::vector<double> avail;
    double totalAvail = 0.0;
    size_t size() const { return nx * ny; }
    size_t at(size_t ix, size_t iy) const { return iy * nx + ix; }
    double edge(bool vertical, size_t t) const;
    Box box(const BinRect &r) const;
    double sum(const std::vector<double> &field, const BinRect &r) const;
};

struct RowSpec {
    double y, height, x0, x1, pitch;
};

void buildGrid(Grid &g, FreeSpace &fs, const std::vector<RowSpec> &rows, const std::vector<Box> &obstacles,
               const Box &die, size_t perAxis);

// Rectangle sums for cluster growth, both exact integer sums.  Free area: a
// summed-area table in units of 2^-40, or coarser on a die whose free area
// would not fit 2^62 of those (the grid's free area does not change;
// a double summed-area table would carry rounding that grows along the
// diagonal).  Occupancy: per-row prefix sums in units of 2^-16, patched row
// by row when cells move.
struct AreaTables {
    size_t nx = 0, ny = 0;
    std::vector<uint64_t> avail;    // (nx+1) x (ny+1), units of 2^-availShift
    int availShift = 40;            // 40, or less where the die's free area needs it
    std::vector<int64_t> occRow;    // ny x (nx+1)
    void buildAvail(const Grid &g);
    void setOcc(const std::vector<uint64_t> &occFix);
    void updateRow(const std::vector<uint64_t> &occFix, size_t iy);
    double sumOcc(const BinRect &r) const;
    double sumAvail(const BinRect &r) const;
    double density(const BinRect &r) const;
};

// 4-neighbour flood fill of bins with occ > g * avail (simpl_lal.cc:141-183).
std::vector<BinRect> overfilledClusters(const Grid &g, const std::vector<double> &occ, double densityLimit);

// Greedy growth until density <= g (simpl_lal.cc:185-211).
BinRect grow(const Grid &g, const AreaTables &t, const std::vector<double> &occ, BinRect r, double densityLimit);

struct Stripe {
    double start = 0.0, end = 0.0, capacity = 0.0;
};

// Stripes of one half block, in the order cells fill them: farthest from the
// cut first (simpl_lal.cc:213-310).
std::vector<Stripe> stripes(const FreeSpace &fs, const std::vector<Box> &obstacles, const Box &box, bool vertical,
                            double cellArea, double regionRoom, double densityLimit, double stripeAreaFraction,
                            double cutline, bool belowCut);

}  // namespace kt::lal
