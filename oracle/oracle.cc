// ORACLE -- test infrastructure only.
//
// A CPU restatement of KTPlace's SimPL inner loop, used as the parity checker
// for the HIP path and, timed, as the CPU baseline.  Nothing in the product
// (ktplace_amd/) links, loads or calls this file; only tests/, the smoke check
// and bench.py's CPU-baseline leg load the shared library built from it.
//
// Each function names the reference file:line it follows.  Build flags match
// the reference's (src/Master.make:21): -O2 -ffp-contract=off.  oneTBB's four
// entry points are replaced by OpenMP tasks with the same grain (4096) and the
// same deterministic chunked reduction, so results do not depend on the
// thread count (simpl_parallel.h:14-57).
//
// Parity pins (tests/test_oracle.py):
//   collectNets           brute-force dedup on hand-made nets, generator invariants
//   buildB2BAxis          docs/foundations/05-net-models.md worked examples (4-pin
//                         net at 0,2,5,10: 5 edges, length 30, cost 10; 3-pin net
//                         at 1,4,9: cost 8), closed-form row counts, symmetry of
//                         the off-diagonal multiset, diagonal = row weight sum,
//                         Sigma w*l^2 = Sigma_nets weight*HPWL when every edge
//                         is longer than the floor
//   buildStarAxis         a hand-solved two-net case (m = 5.25, star 2.75); A v - b
//                         equals the finite-difference gradient of the star
//                         model's energy written from its definition
//   multiply / getDot     dense numpy product; serial chunked sum
//   solveBothAxes         scipy direct solve; textbook PCG iterates; breakdown and
//                         zero-rhs rules; lock-step observer history
//   getHpwl / occupancy / overflow   closed forms and area conservation
//   FreeSpace / AxisProfile           hand-computed profiles
//   LookAheadLegalizer   packStripes: the reference docs' two worked answers
//                        (02-look-ahead-legalization.md:97-109, :150-155), a
//                        hand-computed split, farthest-first fill and order
//                        kept along the cut axis; split: a numpy brute force
//                        of the area median and of the cutline's free-area
//                        share; getOverfilledClusters: scipy.ndimage.label;
//                        getGrown: the least dense expansion at every step;
//                        planted mistakes (nearest-first fill, reversed map)
//                        fail these (tests/test_oracle_lal.py); in die,
//                        overflow reduced, deterministic.  The regions of a batch run
//                        one after the other, which is the reference with one
//                        thread: simpl_lal.cc:102-122 runs them in a
//                        parallel_for whose updates of members/binOf/occ
//                        reach outside a region when a cell wider than its
//                        stripe ends up centred in another region's bins,
//                        so with more threads the reference's result there
//                        depends on timing.  ko_set_lal_snapshot(1) gives the
//                        other extreme (all regions collect first)
//   simplPlacer loop     Convergence::update: hand-built (gap, upper) sequences
//                        with known stopping iterations; HPWL bounds
//                        (lower <= upper); the loop is the composition of
//                        the pinned steps above
//
// This file restates KTPlace's code closely (that is its purpose), so it
// carries KTPlace's notice:
//
//   MIT License
//   Copyright (c) 2026 Coherent17
//   Permission is hereby granted, free of charge, to any person obtaining a
//   copy of this software and associated documentation files (the
//   "Software"), to deal in the Software without restriction, including
//   without limitation the rights to use, copy, modify, merge, publish,
//   distribute, sublicense, and/or sell copies of the Software, and to permit
//   persons to whom the Software is furnished to do so, subject to the
//   following conditions: The above copyright notice and this permission
//   notice shall be included in all copies or substantial portions of the
//   Software.
//   THE SOFTWARE IS PROVIDED "AS IS", WITHOUT WARRANTY OF ANY KIND, EXPRESS OR
//   IMPLIED, INCLUDING BUT NOT LIMITED TO THE WARRANTIES OF MERCHANTABILITY,
//   FITNESS FOR A PARTICULAR PURPOSE AND NONINFRINGEMENT. IN NO EVENT SHALL THE
//   AUTHORS OR COPYRIGHT HOLDERS BE LIABLE FOR ANY CLAIM, DAMAGES OR OTHER
//   LIABILITY, WHETHER IN AN ACTION OF CONTRACT, TORT OR OTHERWISE, ARISING
//   FROM, OUT OF OR IN CONNECTION WITH THE SOFTWARE OR THE USE OR OTHER
//   DEALINGS IN THE SOFTWARE.
#include "../include/ktplace_gen.h"

#include <omp.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <optional>
#include <random>
#include <string>
#include <utility>
#include <vector>

namespace {

// ---------------------------------------------------------------- parallel
// simpl_parallel.h:14-57
constexpr std::size_t grain = 4096;
// Legalization batches: false runs the regions of a batch one after the other
// (the reference with one thread); true collects every region's cells before
// any region moves one (its regions running concurrently).
bool g_lalSnapshot = false;
// A planted mistake, for the tests that show the LAL pins catch one (set only
// by ko_set_mutation from those tests; zero everywhere else):
//   1  stripes filled nearest to the cut first instead of farthest first
//   2  the affine map inside a stripe reversed (the lowest cell to the top)
int g_mutation = 0;

template <typename Body>
void taskRange(std::size_t n, std::size_t g, Body &body) {
#pragma omp taskloop grainsize(g) default(shared)
    for (std::size_t i = 0; i < n; ++i) body(i);
}

template <typename Body>
void parallelForGrain(std::size_t n, std::size_t g, Body body) {
    if (n == 0) return;
    if (omp_in_parallel()) {
        taskRange(n, g, body);
    } else {
#pragma omp parallel
#pragma omp single
        taskRange(n, g, body);
    }
}

template <typename Body>
void parallelFor(std::size_t n, Body body) {
    parallelForGrain(n, grain, body);
}

template <typename A, typename B>
void parallelInvoke(A a, B b) {
    const auto both = [&] {
#pragma omp task default(shared)
        a();
#pragma omp task default(shared)
        b();
#pragma omp taskwait
    };
    if (omp_in_parallel()) {
        both();
    } else {
#pragma omp parallel
#pragma omp single
        both();
    }
}

// simpl_parallel.h:26-33: total order on (key, index), so any correct sort
// gives tbb::parallel_sort's output.
template <typename Key>
void sortCells(std::vector<std::uint32_t> &cells, Key key) {
    std::sort(cells.begin(), cells.end(), [&](std::uint32_t a, std::uint32_t b) {
        const double ka = key(a);
        const double kb = key(b);
        return ka < kb || (ka == kb && a < b);
    });
}

// simpl_parallel.h:35-51
template <typename Body>
double getDeterministicSum(std::size_t n, Body body) {
    const std::size_t chunks = (n + grain - 1) / grain;
    std::vector<double> partial(chunks, 0.0);
    parallelForGrain(chunks, 1, [&](std::size_t c) {
        const std::size_t end = std::min(n, (c + 1) * grain);
        double sum = 0.0;
        for (std::size_t i = c * grain; i < end; ++i) sum += body(i);
        partial[c] = sum;
    });
    return std::accumulate(partial.begin(), partial.end(), 0.0);
}

// simpl_parallel.h:53-57
double getDot(const std::vector<double> &a, const std::vector<double> &b) {
    return getDeterministicSum(a.size(), [&](std::size_t i) { return a[i] * b[i]; });
}

// ---------------------------------------------------------------- geometry
// kt_geometry.h:8-69, kt_geometry.cc:35-86
struct Point {
    double x = 0.0, y = 0.0;
};
struct Interval {
    double start = 0.0, end = 0.0;
    double getLength() const { return end - start; }
    double getCenter() const { return 0.5 * (start + end); }
    bool operator<(const Interval &o) const {
        return start < o.start || (start == o.start && end < o.end);
    }
};
struct BBox {
    Point lowerLeft, upperRight;
    static BBox getEmpty() {
        const double inf = std::numeric_limits<double>::infinity();
        return {{inf, inf}, {-inf, -inf}};
    }
    double getWidth() const { return upperRight.x - lowerLeft.x; }
    double getHeight() const { return upperRight.y - lowerLeft.y; }
    bool isEmpty() const { return lowerLeft.x > upperRight.x || lowerLeft.y > upperRight.y; }
    double getHalfPerimeter() const { return isEmpty() ? 0.0 : getWidth() + getHeight(); }
    Point getCenter() const {
        return {0.5 * (lowerLeft.x + upperRight.x), 0.5 * (lowerLeft.y + upperRight.y)};
    }
    Interval getX() const { return {lowerLeft.x, upperRight.x}; }
    void extend(double x, double y) {
        lowerLeft.x = std::min(lowerLeft.x, x);
        lowerLeft.y = std::min(lowerLeft.y, y);
        upperRight.x = std::max(upperRight.x, x);
        upperRight.y = std::max(upperRight.y, y);
    }
};

// ---------------------------------------------------------------- netlist
// simpl_lss.h:13-36
constexpr std::uint32_t noVar = 0xffffffffu;
struct PinRef {
    std::uint32_t net = 0, pin = 0;
};
struct NetPins {
    std::vector<std::uint32_t> cells;
    std::vector<double> offX, offY;
    double weight = 1.0;
};
struct Netlist {
    std::size_t numMovable = 0;
    std::vector<std::uint32_t> movVertex, varOfVertex;
    std::vector<double> fixedX, fixedY;
    std::vector<NetPins> nets;
    std::vector<std::size_t> pinStart;
    std::vector<PinRef> pinsOfCell;
};

// simpl_lss.h:38-58
struct SparseSystem {
    std::vector<std::size_t> rowPtr;
    std::vector<std::uint32_t> col;
    std::vector<double> val, diag, rhs;
    std::size_t size() const { return diag.size(); }
    void multiply(const std::vector<double> &in, std::vector<double> &out) const {
        parallelFor(size(), [&](std::size_t i) {
            double sum = diag[i] * in[i];
            for (std::size_t e = rowPtr[i]; e < rowPtr[i + 1]; ++e) sum += val[e] * in[col[e]];
            out[i] = sum;
        });
    }
};

// ---------------------------------------------------------------- grid
// simpl_grid.h:12-232
struct BinRect {
    std::size_t x0 = 0, x1 = 0, y0 = 0, y1 = 0;
    BinRect getMerged(const BinRect &o) const {
        return {std::min(x0, o.x0), std::max(x1, o.x1), std::min(y0, o.y0), std::max(y1, o.y1)};
    }
    std::size_t getBins() const { return (x1 - x0 + 1) * (y1 - y0 + 1); }
};

struct AxisProfile {
    std::vector<double> edges, prefix;
    double getTotal() const { return prefix.back(); }
    double getBefore(double c) const {
        if (c <= edges.front()) return 0.0;
        if (c >= edges.back()) return getTotal();
        const std::size_t k =
            static_cast<std::size_t>(std::upper_bound(edges.begin(), edges.end(), c) - edges.begin()) - 1;
        const double f = (c - edges[k]) / (edges[k + 1] - edges[k]);
        return prefix[k] + f * (prefix[k + 1] - prefix[k]);
    }
    double getBetween(double a, double b) const { return getBefore(b) - getBefore(a); }
    double getCoordinateAt(double target) const {
        if (target <= 0.0) return edges.front();
        if (target >= getTotal()) return edges.back();
        const std::size_t k =
            static_cast<std::size_t>(std::lower_bound(prefix.begin(), prefix.end(), target) - prefix.begin());
        const double step = prefix[k] - prefix[k - 1];
        const double f = step > 0.0 ? (target - prefix[k - 1]) / step : 1.0;
        return edges[k - 1] + f * (edges[k] - edges[k - 1]);
    }
};

struct BinGrid {
    Point origin;
    double dx = 1.0, dy = 1.0;
    std::size_t nx = 1, ny = 1;
    std::vector<double> avail;
    double totalAvail = 0.0;
    std::size_t size() const { return nx * ny; }
    std::size_t at(std::size_t ix, std::size_t iy) const { return iy * nx + ix; }
    static std::size_t clampIndex(double f, std::size_t n) {
        return static_cast<std::size_t>(std::clamp(f, 0.0, static_cast<double>(n) - 1e-9));
    }
    std::size_t getColumn(double x) const { return clampIndex((x - origin.x) / dx, nx); }
    std::size_t getRow(double y) const { return clampIndex((y - origin.y) / dy, ny); }
    std::size_t getBinOf(const Point &p) const { return at(getColumn(p.x), getRow(p.y)); }
    double getEdge(bool vertical, std::size_t t) const {
        return vertical ? origin.x + static_cast<double>(t) * dx : origin.y + static_cast<double>(t) * dy;
    }
    BBox getBox(const BinRect &r) const {
        return {{getEdge(true, r.x0), getEdge(false, r.y0)}, {getEdge(true, r.x1 + 1), getEdge(false, r.y1 + 1)}};
    }
    double sum(const std::vector<double> &field, const BinRect &r) const {
        double total = 0.0;
        for (std::size_t iy = r.y0; iy <= r.y1; ++iy)
            for (std::size_t ix = r.x0; ix <= r.x1; ++ix) total += field[at(ix, iy)];
        return total;
    }
    void deposit(const BBox &b, double sign) {
        if (!(b.upperRight.x > b.lowerLeft.x) || !(b.upperRight.y > b.lowerLeft.y)) return;
        for (std::size_t iy = getRow(b.lowerLeft.y); iy <= getRow(b.upperRight.y); ++iy) {
            const double oy = std::min(b.upperRight.y, getEdge(false, iy + 1)) - std::max(b.lowerLeft.y, getEdge(false, iy));
            for (std::size_t ix = getColumn(b.lowerLeft.x); ix <= getColumn(b.upperRight.x); ++ix) {
                const double ox =
                    std::min(b.upperRight.x, getEdge(true, ix + 1)) - std::max(b.lowerLeft.x, getEdge(true, ix));
                if (ox > 0.0 && oy > 0.0) avail[at(ix, iy)] += sign * ox * oy;
            }
        }
    }
};

struct FreeSpace {
    struct Row {
        Interval y;
        std::vector<Interval> spans;
        double getLength(const Interval &x) const {
            double total = 0.0;
            for (const Interval &span : spans)
                total += std::max(0.0, std::min(x.end, span.end) - std::max(x.start, span.start));
            return total;
        }
    };
    std::vector<Row> rows;

    void build(std::vector<Row> raw, const std::vector<BBox> &obstacles) {
        rows = std::move(raw);
        std::stable_sort(rows.begin(), rows.end(), [](const Row &a, const Row &b) { return a.y.start < b.y.start; });
        for (Row &row : rows) {
            std::vector<Interval> cuts;
            for (const BBox &o : obstacles)
                if (o.upperRight.y > row.y.start && o.lowerLeft.y < row.y.end) cuts.push_back(o.getX());
            std::stable_sort(cuts.begin(), cuts.end());
            std::stable_sort(row.spans.begin(), row.spans.end());
            std::vector<Interval> open;
            for (const Interval &span : row.spans) {
                double a = span.start;
                for (const Interval &cut : cuts) {
                    if (cut.end <= a || cut.start >= span.end) continue;
                    if (cut.start > a) open.push_back({a, cut.start});
                    a = std::max(a, cut.end);
                }
                if (span.end > a) open.push_back({a, span.end});
            }
            row.spans = std::move(open);
        }
    }

    AxisProfile getProfile(const BBox &b, bool vertical) const {
        const auto first = std::upper_bound(rows.begin(), rows.end(), b.lowerLeft.y,
                                            [](double v, const Row &r) { return v < r.y.end; });
        const double lo = vertical ? b.lowerLeft.x : b.lowerLeft.y;
        const double hi = vertical ? b.upperRight.x : b.upperRight.y;
        AxisProfile out;
        out.edges.push_back(lo);
        out.prefix.push_back(0.0);
        const auto extend = [&out](double to, double area) {
            if (to > out.edges.back()) {
                out.edges.push_back(to);
                out.prefix.push_back(out.prefix.back() + area);
            }
        };
        if (vertical) {
            std::vector<std::pair<double, double>> steps;
            for (auto r = first; r != rows.end() && r->y.start < b.upperRight.y; ++r) {
                const double h = std::min(b.upperRight.y, r->y.end) - std::max(b.lowerLeft.y, r->y.start);
                for (const Interval &span : r->spans) {
                    const double a = std::max(span.start, lo);
                    const double z = std::min(span.end, hi);
                    if (h > 0.0 && z > a) {
                        steps.emplace_back(a, h);
                        steps.emplace_back(z, -h);
                    }
                }
            }
            std::stable_sort(steps.begin(), steps.end());
            double height = 0.0;
            for (const auto &[at, dh] : steps) {
                extend(at, height * (at - out.edges.back()));
                height += dh;
            }
        } else {
            for (auto r = first; r != rows.end() && r->y.start < b.upperRight.y; ++r) {
                const double a = std::max(lo, r->y.start);
                const double z = std::min(hi, r->y.end);
                if (z > a) {
                    extend(a, 0.0);
                    extend(z, r->getLength(b.getX()) * (z - a));
                }
            }
        }
        extend(hi, 0.0);
        if (out.edges.size() == 1) {
            out.edges.push_back(hi);
            out.prefix.push_back(0.0);
        }
        return out;
    }
};

// ---------------------------------------------------------------- options
// simpl_options.h:11-58
struct Options {
    std::size_t binsPerAxis = 0;
    double cgTol = 1e-4;
    std::size_t initMaxIters = 15;
    int initialNetModel = 0;  // 0 star, 1 b2b, 2 star_b2b
    std::size_t maxIters = 100;
    std::uint64_t seed = 20240607;
    double densityLimit = 1.0;
    double alphaBase = 0.02;
    std::size_t maxLevel = 20;
    std::size_t minCellsToSplit = 4;
    double stripeAreaFraction = 0.1;
    std::size_t gapReferenceIter = 10;
    double gapRelaxedFrac = 0.25;
    double gapTightFrac = 0.05;
    std::size_t stallIters = 6;
    std::size_t refineIters = 10;
    std::size_t refineGridFactor = 3;
    std::size_t cgMaxIter = 60;
};

// ---------------------------------------------------------------- design
struct Design {
    Netlist netlist;
    std::vector<double> width, height, area;
    std::vector<BBox> obstacles;
    struct RowInfo {
        double coordinate, height, pitch, x0, x1;
    };
    std::vector<RowInfo> rows;
    BBox die;

    SparseSystem sysX, sysY;
    std::vector<std::vector<double>> history;  // per solve iteration: resid x, resid y

    // grid state (simpl_placer.cc:163-166)
    BinGrid grid;
    FreeSpace freeSpace;
    double densityLimit = 1.0;
    Options opts;
    double floorX = 1.0, floorY = 1.0;
};

// ---------------------------------------------------------------- B2B
// simpl_lss.cc:29-114
void buildB2BAxis(const Netlist &netlist, SparseSystem &sys, const std::vector<double> &pos,
                  double lengthFloor, double alpha, const std::vector<double> &anchor, bool vertical) {
    const auto pinAt = [&](const NetPins &net, std::size_t q) {
        return pos[net.cells[q]] + (vertical ? net.offX[q] : net.offY[q]);
    };
    std::vector<std::pair<std::uint32_t, std::uint32_t>> extremes(netlist.nets.size());
    parallelFor(netlist.nets.size(), [&](std::size_t n) {
        const NetPins &net = netlist.nets[n];
        std::uint32_t lo = 0, hi = 0;
        for (std::uint32_t q = 1; q < net.cells.size(); ++q) {
            lo = pinAt(net, q) < pinAt(net, lo) ? q : lo;
            hi = pinAt(net, q) > pinAt(net, hi) ? q : hi;
        }
        extremes[n] = {lo, hi};
    });
    const auto forEachEdge = [&](std::size_t i, auto &&visit) {
        for (std::size_t e = netlist.pinStart[i]; e < netlist.pinStart[i + 1]; ++e) {
            const PinRef ref = netlist.pinsOfCell[e];
            const NetPins &net = netlist.nets[ref.net];
            const auto [lo, hi] = extremes[ref.net];
            if (lo == hi) continue;
            const double base = net.weight / static_cast<double>(net.cells.size() - 1);
            const double own = pinAt(net, ref.pin);
            const auto edge = [&](std::uint32_t other) {
                const double there = pinAt(net, other);
                const double w = base / std::max(std::abs(own - there), lengthFloor);
                visit(net.cells[other], there, own - pos[net.cells[ref.pin]], w);
            };
            if (ref.pin == lo || ref.pin == hi) {
                for (std::uint32_t other = 0; other < net.cells.size(); ++other)
                    if (other != ref.pin) edge(other);
            } else {
                edge(lo);
                edge(hi);
            }
        }
    };
    sys.rowPtr.assign(netlist.numMovable + 1, 0);
    parallelFor(netlist.numMovable, [&](std::size_t i) {
        std::size_t count = 0;
        forEachEdge(i, [&](std::uint32_t cell, double, double, double) {
            count += netlist.varOfVertex[cell] != noVar ? 1 : 0;
        });
        sys.rowPtr[i + 1] = count;
    });
    std::partial_sum(sys.rowPtr.begin(), sys.rowPtr.end(), sys.rowPtr.begin());
    sys.col.resize(sys.rowPtr.back());
    sys.val.resize(sys.rowPtr.back());
    sys.diag.assign(netlist.numMovable, 0.0);
    sys.rhs.assign(netlist.numMovable, 0.0);
    parallelFor(netlist.numMovable, [&](std::size_t i) {
        std::size_t slot = sys.rowPtr[i];
        double diag = 0.0, rhs = 0.0;
        forEachEdge(i, [&](std::uint32_t cell, double pin, double ownOffset, double w) {
            const std::uint32_t var = netlist.varOfVertex[cell];
            diag += w;
            if (var != noVar) {
                sys.col[slot] = var;
                sys.val[slot] = -w;
                ++slot;
                rhs += w * ((pin - pos[cell]) - ownOffset);
            } else {
                rhs += w * (pin - ownOffset);
            }
        });
        if (alpha > 0.0) {
            const double w = alpha / std::max(std::abs(pos[netlist.movVertex[i]] - anchor[i]), lengthFloor);
            diag += w;
            rhs += w * anchor[i];
        }
        sys.diag[i] = std::max(diag, 1e-12);
        sys.rhs[i] = rhs;
    });
}

// simpl_lss.cc:11-27
void buildB2B(Design &d, const std::vector<double> &x, const std::vector<double> &y, double alpha,
              const std::vector<double> &anchorX, const std::vector<double> &anchorY) {
    const Netlist &nl = d.netlist;
    std::vector<double> px = nl.fixedX, py = nl.fixedY;
    for (std::size_t i = 0; i < nl.numMovable; ++i) {
        px[nl.movVertex[i]] = x[i];
        py[nl.movVertex[i]] = y[i];
    }
    parallelInvoke([&] { buildB2BAxis(nl, d.sysX, px, d.floorX, alpha, anchorX, true); },
                   [&] { buildB2BAxis(nl, d.sysY, py, d.floorY, alpha, anchorY, false); });
}

// simpl_lss.cc:116-201
void buildStarAxis(const Netlist &netlist, SparseSystem &sys, const std::vector<std::uint32_t> &starOf,
                   std::size_t size, bool vertical) {
    struct End {
        std::uint32_t var;
        double offset, fixed;
    };
    std::vector<std::vector<std::pair<std::uint32_t, double>>> rows(size);
    sys.diag.assign(size, 0.0);
    sys.rhs.assign(size, 0.0);
    const auto spring = [&](const End &a, const End &b, double w) {
        const auto pull = [&](const End &self, const End &other) {
            if (self.var == noVar) return;
            sys.diag[self.var] += w;
            if (other.var != noVar) {
                rows[self.var].emplace_back(other.var, -w);
                sys.rhs[self.var] += w * (other.offset - self.offset);
            } else {
                sys.rhs[self.var] += w * (other.fixed + other.offset - self.offset);
            }
        };
        pull(a, b);
        pull(b, a);
    };
    for (std::size_t n = 0; n < netlist.nets.size(); ++n) {
        const NetPins &net = netlist.nets[n];
        const std::size_t p = net.cells.size();
        if (p < 2) continue;
        const auto end = [&](std::size_t q) {
            const std::uint32_t c = net.cells[q];
            return End{netlist.varOfVertex[c], vertical ? net.offX[q] : net.offY[q],
                       vertical ? netlist.fixedX[c] : netlist.fixedY[c]};
        };
        if (p == 2) {
            spring(end(0), end(1), net.weight);
        } else if (starOf[n] != noVar) {
            const End star{starOf[n], 0.0, 0.0};
            const double w = net.weight * static_cast<double>(p) / static_cast<double>(p - 1);
            for (std::size_t q = 0; q < p; ++q) spring(end(q), star, w);
        }
    }
    sys.rowPtr.assign(size + 1, 0);
    for (std::size_t i = 0; i < size; ++i) {
        sys.rowPtr[i + 1] = sys.rowPtr[i] + rows[i].size();
        sys.diag[i] = std::max(sys.diag[i], 1e-12);
    }
    sys.col.resize(sys.rowPtr.back());
    sys.val.resize(sys.rowPtr.back());
    for (std::size_t i = 0; i < size; ++i) {
        std::size_t slot = sys.rowPtr[i];
        for (const auto &[c, v] : rows[i]) {
            sys.col[slot] = c;
            sys.val[slot] = v;
            ++slot;
        }
    }
}

std::size_t buildStar(Design &d) {
    const Netlist &nl = d.netlist;
    std::vector<std::uint32_t> starOf(nl.nets.size(), noVar);
    std::size_t size = nl.numMovable;
    for (std::size_t n = 0; n < nl.nets.size(); ++n) {
        const auto &cells = nl.nets[n].cells;
        const bool movable =
            std::any_of(cells.begin(), cells.end(), [&](std::uint32_t c) { return nl.varOfVertex[c] != noVar; });
        if (cells.size() >= 3 && movable) starOf[n] = static_cast<std::uint32_t>(size++);
    }
    parallelInvoke([&] { buildStarAxis(nl, d.sysX, starOf, size, true); },
                   [&] { buildStarAxis(nl, d.sysY, starOf, size, false); });
    return size;
}

// ---------------------------------------------------------------- CG
// simpl_lss.cc:203-290
struct Residuals {
    double x = 0.0, y = 0.0;
};

struct SolveResult {
    Residuals resid;
    std::size_t iterations = 0;
};

SolveResult solveBothAxes(const SparseSystem &sysX, std::vector<double> &x, const SparseSystem &sysY,
                          std::vector<double> &y, double tolerance, std::size_t maxIterations,
                          std::vector<std::vector<double>> *history) {
    struct Axis {
        const SparseSystem *sys = nullptr;
        std::vector<double> *sol = nullptr;
        std::vector<double> r, z, p, ap;
        double bNorm = 0.0, rho = 0.0, resid = 0.0;
        bool active = false;
    };
    const auto start = [](Axis &ax, const SparseSystem &sys, std::vector<double> &sol) {
        const std::size_t n = sys.size();
        ax.sys = &sys;
        ax.sol = &sol;
        ax.r.assign(n, 0.0);
        ax.z.assign(n, 0.0);
        ax.ap.assign(n, 0.0);
        sys.multiply(sol, ax.ap);
        parallelFor(n, [&](std::size_t i) {
            ax.r[i] = sys.rhs[i] - ax.ap[i];
            ax.z[i] = ax.r[i] / sys.diag[i];
        });
        ax.bNorm = std::sqrt(getDot(sys.rhs, sys.rhs));
        if (!(ax.bNorm > 0.0)) return;
        ax.p = ax.z;
        ax.rho = getDot(ax.r, ax.z);
        ax.resid = std::sqrt(getDot(ax.r, ax.r)) / ax.bNorm;
        ax.active = true;
    };
    const auto step = [](Axis &ax) {
        const SparseSystem &sys = *ax.sys;
        sys.multiply(ax.p, ax.ap);
        const double pap = getDot(ax.p, ax.ap);
        if (!(pap > 0.0)) return false;
        const double a = ax.rho / pap;
        parallelFor(sys.size(), [&](std::size_t i) {
            (*ax.sol)[i] += a * ax.p[i];
            ax.r[i] -= a * ax.ap[i];
            ax.z[i] = ax.r[i] / sys.diag[i];
        });
        ax.resid = std::sqrt(getDot(ax.r, ax.r)) / ax.bNorm;
        const double rhoNext = getDot(ax.r, ax.z);
        if (!(rhoNext > 0.0)) return false;
        const double beta = rhoNext / ax.rho;
        parallelFor(sys.size(), [&](std::size_t i) { ax.p[i] = ax.z[i] + beta * ax.p[i]; });
        ax.rho = rhoNext;
        return true;
    };
    Axis ax, ay;
    start(ax, sysX, x);
    start(ay, sysY, y);
    const auto live = [&](const Axis &a) { return a.active && a.resid > tolerance; };
    SolveResult out;
    if (history) history->clear();
    for (std::size_t it = 0; it < maxIterations && (live(ax) || live(ay)); ++it) {
        const bool stepX = live(ax), stepY = live(ay);
        parallelInvoke([&] { if (stepX && !step(ax)) ax.active = false; },
                       [&] { if (stepY && !step(ay)) ay.active = false; });
        out.iterations = it + 1;
        if (history) history->push_back({ax.resid, ay.resid});
    }
    out.resid = {ax.resid, ay.resid};
    return out;
}

// ---------------------------------------------------------------- placer
// simpl_placer.cc:169-297
void collectFloors(Design &d) {
    const std::size_t n = d.netlist.numMovable;
    if (n > 0) {
        d.floorX = std::accumulate(d.width.begin(), d.width.end(), 0.0) / static_cast<double>(n);
        d.floorY = std::accumulate(d.height.begin(), d.height.end(), 0.0) / static_cast<double>(n);
    }
}

std::size_t getDefaultBinsPerAxis(const Design &d) {
    return d.opts.binsPerAxis > 0
               ? d.opts.binsPerAxis
               : static_cast<std::size_t>(std::clamp(
                     std::sqrt(static_cast<double>(std::max<std::size_t>(d.netlist.numMovable, 1)) / 51.0), 16.0,
                     256.0));
}

void buildGrid(Design &d, std::size_t perAxis) {
    BinGrid &grid = d.grid;
    grid.origin = d.die.lowerLeft;
    grid.nx = perAxis;
    grid.ny = perAxis;
    grid.dx = std::max(d.die.getWidth(), 1e-9) / static_cast<double>(perAxis);
    grid.dy = std::max(d.die.getHeight(), 1e-9) / static_cast<double>(perAxis);
    grid.avail.assign(grid.size(), 0.0);
    std::vector<FreeSpace::Row> rows;
    double smallestSite = std::numeric_limits<double>::max();
    // one input row per subrow; consecutive subrows of one row (same y and
    // height) are that row's spans (simpl_placer.cc:199-211)
    for (const auto &row : d.rows) {
        if (!(row.height > 0.0) || !(row.pitch > 0.0)) continue;
        smallestSite = std::min(smallestSite, row.pitch * row.height);
        if (!rows.empty() && rows.back().y.start == row.coordinate && rows.back().y.end == row.coordinate + row.height)
            rows.back().spans.push_back({row.x0, row.x1});
        else
            rows.push_back({{row.coordinate, row.coordinate + row.height}, {{row.x0, row.x1}}});
    }
    d.freeSpace.build(std::move(rows), d.obstacles);
    for (const auto &row : d.freeSpace.rows)
        for (const Interval &span : row.spans)
            grid.deposit({{span.start, row.y.start}, {span.end, row.y.end}}, 1.0);
    const double sliver = smallestSite < std::numeric_limits<double>::max() ? 1e-3 * smallestSite : 0.0;
    for (double &a : grid.avail) a = a < sliver ? 0.0 : a;
    grid.totalAvail = std::accumulate(grid.avail.begin(), grid.avail.end(), 0.0);
}

void seedUniform(const Design &d, std::uint64_t seed, std::vector<double> &x, std::vector<double> &y) {
    std::mt19937_64 rng(seed);
    const auto unit = [&rng] { return static_cast<double>(rng() >> 11) * 0x1.0p-53; };
    const std::size_t n = d.netlist.numMovable;
    x.resize(n);
    y.resize(n);
    for (std::size_t i = 0; i < n; ++i) {
        x[i] = d.die.lowerLeft.x + unit() * std::max(d.die.getWidth() - d.width[i], 0.0);
        y[i] = d.die.lowerLeft.y + unit() * std::max(d.die.getHeight() - d.height[i], 0.0);
    }
}

double getHpwl(const Design &d, const std::vector<double> &x, const std::vector<double> &y) {
    const Netlist &nl = d.netlist;
    std::vector<double> px = nl.fixedX, py = nl.fixedY;
    for (std::size_t i = 0; i < nl.numMovable; ++i) {
        px[nl.movVertex[i]] = x[i];
        py[nl.movVertex[i]] = y[i];
    }
    return getDeterministicSum(nl.nets.size(), [&](std::size_t n) {
        const NetPins &net = nl.nets[n];
        if (net.cells.size() < 2) return 0.0;
        BBox box = BBox::getEmpty();
        for (std::size_t q = 0; q < net.cells.size(); ++q)
            box.extend(px[net.cells[q]] + net.offX[q], py[net.cells[q]] + net.offY[q]);
        return box.getHalfPerimeter();
    });
}

double getCentre(const Design &d, const std::vector<double> &pos, std::uint32_t i, bool vertical) {
    return pos[i] + 0.5 * (vertical ? d.width[i] : d.height[i]);
}

std::vector<double> getOccupancy(const Design &d, const std::vector<double> &x, const std::vector<double> &y) {
    std::vector<double> occ(d.grid.size(), 0.0);
    for (std::uint32_t i = 0; i < d.netlist.numMovable; ++i)
        occ[d.grid.getBinOf({getCentre(d, x, i, true), getCentre(d, y, i, false)})] += d.area[i];
    return occ;
}

double getOverflow(const Design &d, const std::vector<double> &occ) {
    double excess = 0.0;
    for (std::size_t k = 0; k < occ.size(); ++k) excess += std::max(occ[k] - d.densityLimit * d.grid.avail[k], 0.0);
    return d.grid.totalAvail > 0.0 ? excess / d.grid.totalAvail : 0.0;
}

// ---------------------------------------------------------------- LAL
// simpl_lal.cc:43-372
struct Block {
    BBox box;
    std::size_t level = 1;
    bool vertical = true;
    std::vector<std::uint32_t> cells;
    double regionRoom = 0.0;
};
struct Stripe {
    Interval span;
    double capacity = 0.0, used = 0.0;
    std::vector<std::uint32_t> cells;
};

// The second half of nonlinearScale (simpl_lal.cc:275-363), after getStripes:
// stripes ordered farthest from the cut first (stable), cells ordered farthest
// first (total order on (key, id)), each cell into the first stripe not yet
// at capacity, then the affine map of each stripe's cells onto its span.
// size[i] is the cell's extent along the axis; pos holds lower-left corners.
void packStripes(std::vector<Stripe> &strips, const std::vector<std::uint32_t> &cells, double cutline,
                 bool belowCut, std::vector<double> &pos, const std::vector<double> &size,
                 const std::vector<double> &area) {
    const auto centre = [&](std::uint32_t i) { return pos[i] + 0.5 * size[i]; };
    const auto fromCut = [&](double c) { return belowCut ? cutline - c : c - cutline; };
    std::stable_sort(strips.begin(), strips.end(), [&](const Stripe &a, const Stripe &b) {
        if (g_mutation == 1) return fromCut(a.span.getCenter()) < fromCut(b.span.getCenter());
        return fromCut(a.span.getCenter()) > fromCut(b.span.getCenter());
    });
    std::vector<std::uint32_t> order = cells;
    sortCells(order, [&](std::uint32_t i) { return -fromCut(centre(i)); });
    std::size_t s = 0;
    for (const std::uint32_t i : order) {
        while (s + 1 < strips.size() && strips[s].used >= strips[s].capacity) ++s;
        strips[s].cells.push_back(i);
        strips[s].used += area[i];
    }
    for (const Stripe &strip : strips) {
        if (strip.cells.empty()) continue;
        double lo = std::numeric_limits<double>::max();
        double hi = -lo;
        for (const std::uint32_t i : strip.cells) {
            lo = std::min(lo, centre(i));
            hi = std::max(hi, centre(i));
        }
        for (const std::uint32_t i : strip.cells) {
            double t = hi > lo ? (centre(i) - lo) / (hi - lo) : 0.5;
            if (g_mutation == 2) t = 1.0 - t;
            const double c = strip.span.start + t * (strip.span.getLength());
            pos[i] = std::clamp(c - 0.5 * size[i], strip.span.start, std::max(strip.span.start, strip.span.end - size[i]));
        }
    }
}

// Convergence::update (simpl_placer.cc:450-461): the gap of iteration
// gapReferenceIter is the reference; stop once the gap is below the tight
// share of it, or within the relaxed share while the upper bound has not
// improved for stallIters iterations.
struct Convergence {
    double gapReference = 0.0;
    double bestUpper = std::numeric_limits<double>::max();
    std::size_t stalledFor = 0;
    bool update(std::size_t iteration, double gap, double upperHpwl, std::size_t referenceIter, double tightFrac,
                double relaxedFrac, std::size_t stallIters) {
        if (iteration == referenceIter) gapReference = gap;
        stalledFor = upperHpwl < bestUpper ? 0 : stalledFor + 1;
        bestUpper = std::min(bestUpper, upperHpwl);
        const bool upperStalled = stalledFor >= stallIters;
        return iteration > referenceIter && gapReference > 0.0 &&
               (gap < tightFrac * gapReference || (gap <= relaxedFrac * gapReference && upperStalled));
    }
};

struct Lal {
    const Design &d;
    const BinGrid &grid;
    const FreeSpace &freeSpace;
    const Options &opts;
    double densityLimit;
    BBox die;

    double centre(const std::vector<double> &pos, std::uint32_t i, bool vertical) const {
        return pos[i] + 0.5 * (vertical ? d.width[i] : d.height[i]);
    }

    double getDensity(const std::vector<double> &occ, const BinRect &r) const {
        const double cells = grid.sum(occ, r);
        const double room = grid.sum(grid.avail, r);
        if (room > 0.0) return cells / room;
        return cells > 0.0 ? std::numeric_limits<double>::infinity() : 0.0;
    }

    std::vector<BinRect> getOverfilledClusters(const std::vector<double> &occ) const {
        std::vector<char> over(grid.size(), 0);
        for (std::size_t k = 0; k < grid.size(); ++k) over[k] = occ[k] > densityLimit * grid.avail[k] ? 1 : 0;
        std::vector<BinRect> regions;
        std::vector<std::size_t> frontier;
        for (std::size_t seed = 0; seed < grid.size(); ++seed) {
            if (!over[seed]) continue;
            over[seed] = 0;
            frontier.assign(1, seed);
            BinRect r{seed % grid.nx, seed % grid.nx, seed / grid.nx, seed / grid.nx};
            while (!frontier.empty()) {
                const std::size_t k = frontier.back();
                frontier.pop_back();
                const std::size_t ix = k % grid.nx, iy = k / grid.nx;
                r = r.getMerged({ix, ix, iy, iy});
                const std::array<std::pair<long, long>, 4> steps{{{1, 0}, {-1, 0}, {0, 1}, {0, -1}}};
                for (const auto &[sx, sy] : steps) {
                    const long nx = static_cast<long>(ix) + sx, ny = static_cast<long>(iy) + sy;
                    if (nx < 0 || ny < 0 || nx >= static_cast<long>(grid.nx) || ny >= static_cast<long>(grid.ny))
                        continue;
                    const std::size_t kk = grid.at(static_cast<std::size_t>(nx), static_cast<std::size_t>(ny));
                    if (over[kk]) {
                        over[kk] = 0;
                        frontier.push_back(kk);
                    }
                }
            }
            regions.push_back(r);
        }
        return regions;
    }

    BinRect getGrown(BinRect r, const std::vector<double> &occ, std::vector<BinRect> *steps = nullptr) const {
        if (steps) steps->push_back(r);
        while (getDensity(occ, r) > densityLimit) {
            const std::array<BinRect, 4> wider{
                BinRect{r.x0 > 0 ? r.x0 - 1 : r.x0, r.x1, r.y0, r.y1},
                BinRect{r.x0, r.x1 + 1 < grid.nx ? r.x1 + 1 : r.x1, r.y0, r.y1},
                BinRect{r.x0, r.x1, r.y0 > 0 ? r.y0 - 1 : r.y0, r.y1},
                BinRect{r.x0, r.x1, r.y0, r.y1 + 1 < grid.ny ? r.y1 + 1 : r.y1}};
            const BinRect *best = nullptr;
            double bestDensity = std::numeric_limits<double>::infinity();
            for (const BinRect &w : wider) {
                if (w.getBins() == r.getBins()) continue;
                const double dd = getDensity(occ, w);
                if (best == nullptr || dd < bestDensity) {
                    best = &w;
                    bestDensity = dd;
                }
            }
            if (best == nullptr) break;
            r = *best;
            if (steps) steps->push_back(r);
        }
        return r;
    }

    std::vector<Stripe> getStripes(const BBox &box, bool vertical, double cellArea, double regionRoom) const {
        const double lo = vertical ? box.lowerLeft.x : box.lowerLeft.y;
        const double hi = vertical ? box.upperRight.x : box.upperRight.y;
        const AxisProfile room = freeSpace.getProfile(box, vertical);
        std::vector<double> cuts{lo, hi};
        for (const BBox &o : d.obstacles) {
            if (o.upperRight.x <= box.lowerLeft.x || o.lowerLeft.x >= box.upperRight.x ||
                o.upperRight.y <= box.lowerLeft.y || o.lowerLeft.y >= box.upperRight.y)
                continue;
            for (const double edge : {vertical ? o.lowerLeft.x : o.lowerLeft.y, vertical ? o.upperRight.x : o.upperRight.y})
                if (edge > lo && edge < hi) cuts.push_back(edge);
        }
        std::stable_sort(cuts.begin(), cuts.end());
        cuts.erase(std::unique(cuts.begin(), cuts.end()), cuts.end());
        const double limit = opts.stripeAreaFraction * (regionRoom > 0.0 ? regionRoom : room.getTotal());
        const double fill = room.getTotal() > 0.0 ? std::min(densityLimit, cellArea / room.getTotal()) : densityLimit;
        for (std::size_t s = 0; s + 1 < cuts.size();) {
            const double a = cuts[s], b = cuts[s + 1];
            const double here = room.getBetween(a, b);
            const double mid = room.getCoordinateAt(room.getBefore(a) + 0.5 * here);
            if (here > limit && mid > a && mid < b) {
                cuts.insert(cuts.begin() + static_cast<long>(s) + 1, mid);
            } else {
                ++s;
            }
        }
        std::vector<Stripe> out;
        for (std::size_t s = 0; s + 1 < cuts.size(); ++s)
            out.push_back({{cuts[s], cuts[s + 1]}, fill * room.getBetween(cuts[s], cuts[s + 1]), 0.0, {}});
        return out;
    }

    void nonlinearScale(const std::vector<std::uint32_t> &cells, const BBox &box, bool vertical, double cutline,
                        bool belowCut, std::vector<double> &pos, double regionRoom) const {
        if (cells.empty()) return;
        const double cellArea = std::accumulate(cells.begin(), cells.end(), 0.0,
                                                [&](double sum, std::uint32_t i) { return sum + d.area[i]; });
        std::vector<Stripe> strips = getStripes(box, vertical, cellArea, regionRoom);
        packStripes(strips, cells, cutline, belowCut, pos, vertical ? d.width : d.height, d.area);
    }

    std::optional<std::pair<Block, Block>> split(Block &region, std::vector<double> &x, std::vector<double> &y) const {
        if (region.level >= opts.maxLevel || region.cells.size() <= opts.minCellsToSplit) return std::nullopt;
        const bool vertical = region.vertical;
        std::vector<double> &pos = vertical ? x : y;
        std::vector<std::uint32_t> &cells = region.cells;
        sortCells(cells, [&](std::uint32_t i) { return centre(pos, i, vertical); });
        const double cellArea = std::accumulate(cells.begin(), cells.end(), 0.0,
                                                [&](double s, std::uint32_t i) { return s + d.area[i]; });
        std::size_t m = 0;
        double lowArea = 0.0;
        while (m + 1 < cells.size() && (m == 0 || lowArea < 0.5 * cellArea)) lowArea += d.area[cells[m++]];
        std::vector<std::uint32_t> low(cells.begin(), cells.begin() + static_cast<long>(m));
        std::vector<std::uint32_t> high(cells.begin() + static_cast<long>(m), cells.end());
        const BBox &box = region.box;
        const double lo = vertical ? box.lowerLeft.x : box.lowerLeft.y;
        const double hi = vertical ? box.upperRight.x : box.upperRight.y;
        const AxisProfile room = freeSpace.getProfile(box, vertical);
        const double share = cellArea > 0.0 ? lowArea / cellArea : 0.5;
        const double cutline = room.getTotal() > 0.0 ? room.getCoordinateAt(share * room.getTotal()) : lo + share * (hi - lo);
        BBox lowBox = box, highBox = box;
        (vertical ? lowBox.upperRight.x : lowBox.upperRight.y) = cutline;
        (vertical ? highBox.lowerLeft.x : highBox.lowerLeft.y) = cutline;
        nonlinearScale(low, lowBox, vertical, cutline, true, pos, region.regionRoom);
        nonlinearScale(high, highBox, vertical, cutline, false, pos, region.regionRoom);
        return std::pair{Block{lowBox, region.level + 1, !vertical, std::move(low), region.regionRoom},
                         Block{highBox, region.level + 1, !vertical, std::move(high), region.regionRoom}};
    }

    void legalizeBlock(Block &block, std::vector<double> &x, std::vector<double> &y) const {
        auto halves = split(block, x, y);
        if (!halves) return;
        legalizeBlock(halves->first, x, y);
        legalizeBlock(halves->second, x, y);
    }

    void clampToDie(std::vector<double> &x, std::vector<double> &y) const {
        for (std::size_t i = 0; i < x.size(); ++i) {
            x[i] = std::clamp(x[i], die.lowerLeft.x, std::max(die.lowerLeft.x, die.upperRight.x - d.width[i]));
            y[i] = std::clamp(y[i], die.lowerLeft.y, std::max(die.lowerLeft.y, die.upperRight.y - d.height[i]));
        }
    }

    // Regions of one batch do not touch (reach test below), so running them in
    // batch order gives the reference's parallel result.
    void legalize(std::vector<double> &x, std::vector<double> &y) const {
        const std::size_t n = d.width.size();
        std::vector<std::vector<std::uint32_t>> members(grid.size());
        std::vector<std::uint32_t> binOf(n);
        std::vector<double> occ(grid.size(), 0.0);
        for (std::uint32_t i = 0; i < n; ++i) {
            const std::size_t k = grid.getBinOf({centre(x, i, true), centre(y, i, false)});
            binOf[i] = static_cast<std::uint32_t>(k);
            members[k].push_back(i);
            occ[k] += d.area[i];
        }
        struct Turn {
            BinRect bins, reach;
            std::uint32_t id = 0;
        };
        const auto reachOf = [&](const BinRect &r) {
            return BinRect{r.x0 > 0 ? r.x0 - 1 : 0, std::min(r.x1 + 1, grid.nx - 1), r.y0 > 0 ? r.y0 - 1 : 0,
                           std::min(r.y1 + 1, grid.ny - 1)};
        };
        const auto touches = [](const BinRect &a, const BinRect &b) {
            return a.x0 <= b.x1 && b.x0 <= a.x1 && a.y0 <= b.y1 && b.y0 <= a.y1;
        };
        std::vector<std::uint32_t> taken(n, 0);
        const auto collect = [&](const Turn &t) {
            std::vector<std::uint32_t> cells;
            for (std::size_t iy = t.bins.y0; iy <= t.bins.y1; ++iy)
                for (std::size_t ix = t.bins.x0; ix <= t.bins.x1; ++ix) {
                    const std::size_t k = grid.at(ix, iy);
                    for (const std::uint32_t i : members[k])
                        if (binOf[i] == k && taken[i] != t.id) {
                            taken[i] = t.id;
                            cells.push_back(i);
                        }
                }
            return cells;
        };
        const auto legalizeCells = [&](const Turn &t, const std::vector<std::uint32_t> &cells) {
            Block top{grid.getBox(t.bins), 1, true, cells, grid.sum(grid.avail, t.bins)};
            legalizeBlock(top, x, y);
            for (const std::uint32_t i : cells) {
                const std::size_t k = grid.getBinOf({centre(x, i, true), centre(y, i, false)});
                if (k != binOf[i]) {
                    occ[binOf[i]] -= d.area[i];
                    occ[k] += d.area[i];
                    binOf[i] = static_cast<std::uint32_t>(k);
                    members[k].push_back(i);
                }
            }
        };
        const auto legalizeRegion = [&](const Turn &t) {
            Block top{grid.getBox(t.bins), 1, true, {}, grid.sum(grid.avail, t.bins)};
            top.cells = collect(t);
            const std::vector<std::uint32_t> cells = top.cells;
            legalizeBlock(top, x, y);
            for (const std::uint32_t i : cells) {
                const std::size_t k = grid.getBinOf({centre(x, i, true), centre(y, i, false)});
                if (k != binOf[i]) {
                    occ[binOf[i]] -= d.area[i];
                    occ[k] += d.area[i];
                    binOf[i] = static_cast<std::uint32_t>(k);
                    members[k].push_back(i);
                }
            }
        };
        std::vector<Turn> batch;
        const auto runBatch = [&] {
            if (g_lalSnapshot) {
                // every region's cells collected before any region moves one
                std::vector<std::vector<std::uint32_t>> cellsOf(batch.size());
                for (std::size_t b = 0; b < batch.size(); ++b) cellsOf[b] = collect(batch[b]);
                for (std::size_t b = 0; b < batch.size(); ++b) legalizeCells(batch[b], cellsOf[b]);
            } else {
                for (const Turn &t : batch) legalizeRegion(t);
            }
            batch.clear();
        };
        std::uint32_t turn = 0;
        for (const BinRect &cluster : getOverfilledClusters(occ)) {
            BinRect bins = getGrown(cluster, occ);
            const auto clashes = [&] {
                return std::any_of(batch.begin(), batch.end(),
                                   [&](const Turn &t) { return touches(t.reach, reachOf(bins)); });
            };
            if (clashes()) {
                runBatch();
                bins = getGrown(cluster, occ);
            }
            batch.push_back({bins, reachOf(bins), ++turn});
        }
        runBatch();
        clampToDie(x, y);
    }
};

void lookAheadLegalize(const Design &d, std::vector<double> &x, std::vector<double> &y) {
    Lal lal{d, d.grid, d.freeSpace, d.opts, d.densityLimit, d.die};
    lal.legalize(x, y);
}

// ---------------------------------------------------------------- SimPL loop
// simpl_placer.cc:358-580 (fences, frames and reports are not part of the path)
struct PlaceStats {
    double hpwlSeed, hpwlInit, hpwlLower, hpwlFinal, gap, overflowLower, overflowFinal;
    double buildSeconds, solveSeconds, spreadSeconds;
    std::uint64_t initIters, globalIters, cgIterations;
};

double nowSeconds() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Placer {
    Design &d;
    PlaceStats st{};

    void solve(std::vector<double> &x, std::vector<double> &y) {
        const SolveResult r = solveBothAxes(d.sysX, x, d.sysY, y, d.opts.cgTol, d.opts.cgMaxIter, nullptr);
        st.cgIterations += r.iterations;
    }

    void initialRounds(std::vector<double> &lx, std::vector<double> &ly, bool star, std::uint64_t &accepted) {
        const std::vector<double> noAnchors;
        std::vector<double> ex, ey;
        if (star) {
            const std::size_t size = buildStar(d);
            ex = lx;
            ey = ly;
            ex.resize(size, d.die.getCenter().x);
            ey.resize(size, d.die.getCenter().y);
        }
        for (std::size_t round = 0; round < d.opts.initMaxIters; ++round) {
            std::vector<double> nx = star ? ex : lx, ny = star ? ey : ly;
            if (!star) buildB2B(d, nx, ny, 0.0, noAnchors, noAnchors);
            solve(nx, ny);
            if (star) {
                ex = nx;
                ey = ny;
                nx.resize(d.netlist.numMovable);
                ny.resize(d.netlist.numMovable);
            }
            const double roundHpwl = getHpwl(d, nx, ny);
            ++st.initIters;
            const bool first = star && round == 0;
            if (!first && !(roundHpwl < st.hpwlInit)) break;
            lx = std::move(nx);
            ly = std::move(ny);
            st.hpwlInit = roundHpwl;
            accepted = st.initIters;
        }
    }

    void run(std::vector<double> &outX, std::vector<double> &outY, std::vector<double> *trace) {
        const Options &o = d.opts;
        d.densityLimit = std::clamp(o.densityLimit, 0.05, 1.0);
        collectFloors(d);
        buildGrid(d, getDefaultBinsPerAxis(d));
        std::vector<double> lx, ly, ux, uy;
        seedUniform(d, o.seed, lx, ly);
        st.hpwlSeed = getHpwl(d, lx, ly);
        st.hpwlInit = st.hpwlSeed;
        std::uint64_t accepted = 0;
        if (o.initialNetModel != 1) initialRounds(lx, ly, true, accepted);
        if (o.initialNetModel != 0) initialRounds(lx, ly, false, accepted);
        // getGlobalPlacement (simpl_placer.cc:463-539) with Convergence (:450-461)
        Convergence convergence;
        for (std::size_t it = 0; it < o.maxIters; ++it) {
            st.globalIters = it + 1;
            double t0 = nowSeconds();
            ux = lx;
            uy = ly;
            lookAheadLegalize(d, ux, uy);
            st.spreadSeconds += nowSeconds() - t0;
            const double upperHpwl = getHpwl(d, ux, uy);
            const double lowerHpwl = getHpwl(d, lx, ly);
            const double gap = upperHpwl - lowerHpwl;
            if (trace) {
                trace->push_back(lowerHpwl);
                trace->push_back(upperHpwl);
                trace->push_back(getOverflow(d, getOccupancy(d, ux, uy)));
            }
            if (convergence.update(it + 1, gap, upperHpwl, o.gapReferenceIter, o.gapTightFrac, o.gapRelaxedFrac,
                                   o.stallIters))
                break;
            const double alpha = o.alphaBase * (1.0 + static_cast<double>(it));
            t0 = nowSeconds();
            buildB2B(d, lx, ly, alpha, ux, uy);
            const double t1 = nowSeconds();
            solve(lx, ly);
            st.buildSeconds += t1 - t0;
            st.solveSeconds += nowSeconds() - t1;
        }
        if (!ux.empty() && o.refineIters > 0) {
            const std::size_t done = st.globalIters;
            buildGrid(d, d.grid.nx * o.refineGridFactor);
            for (std::size_t r = 0; r < o.refineIters; ++r) {
                const double alpha = o.alphaBase * (1.0 + static_cast<double>(done + r));
                buildB2B(d, lx, ly, alpha, ux, uy);
                solve(lx, ly);
                ux = lx;
                uy = ly;
                lookAheadLegalize(d, ux, uy);
            }
        }
        if (ux.empty()) {
            ux = lx;
            uy = ly;
        }
        st.hpwlFinal = getHpwl(d, ux, uy);
        st.hpwlLower = getHpwl(d, lx, ly);
        st.gap = st.hpwlFinal - st.hpwlLower;
        st.overflowLower = getOverflow(d, getOccupancy(d, lx, ly));
        st.overflowFinal = getOverflow(d, getOccupancy(d, ux, uy));
        outX = ux;
        outY = uy;
    }
};

thread_local std::string lastError;

}  // namespace

// ---------------------------------------------------------------- C ABI for tests
extern "C" {

struct ko_design;

struct ko_place_opts {
    uint64_t bins_per_axis;
    double cg_tol;
    uint64_t init_max_iters;
    int32_t initial_net_model;  // 0 star, 1 b2b, 2 star_b2b
    int32_t pad0;
    uint64_t max_iters;
    uint64_t seed;
    double density_limit;
    double alpha_base;
    uint64_t refine_iters;
    uint64_t cg_max_iter;
};

struct ko_place_stats {
    double hpwl_seed, hpwl_init, hpwl_lower, hpwl_final, gap, overflow_lower, overflow_final;
    double build_seconds, solve_seconds, spread_seconds;
    uint64_t init_iters, global_iters, cg_iterations;
};

const char *ko_last_error(void) { return lastError.c_str(); }

void ko_set_threads(int threads) { omp_set_num_threads(threads > 0 ? threads : 1); }
int ko_get_threads(void) { return omp_get_max_threads(); }

// collectCells/collectNets (simpl_dm.cc:16-85) on a raw design.
ko_design *ko_design_create(const ktg_raw *raw) {
    auto *d = new Design();
    Netlist &nl = d->netlist;
    const uint32_t V = raw->num_vertices;
    nl.varOfVertex.assign(V, noVar);
    nl.fixedX.assign(V, 0.0);
    nl.fixedY.assign(V, 0.0);
    for (uint32_t v = 0; v < V; ++v) {
        nl.fixedX[v] = raw->pos_x[v];
        nl.fixedY[v] = raw->pos_y[v];
        if (raw->is_fixed[v]) {
            if (raw->width[v] > 0.0 && raw->height[v] > 0.0)
                d->obstacles.push_back({{raw->pos_x[v], raw->pos_y[v]},
                                        {raw->pos_x[v] + raw->width[v], raw->pos_y[v] + raw->height[v]}});
            continue;
        }
        nl.varOfVertex[v] = static_cast<uint32_t>(nl.numMovable++);
        nl.movVertex.push_back(v);
        d->width.push_back(raw->width[v]);
        d->height.push_back(raw->height[v]);
        d->area.push_back(raw->width[v] * raw->height[v]);
    }
    nl.nets.resize(raw->num_nets);
    for (uint32_t n = 0; n < raw->num_nets; ++n) {
        std::vector<uint64_t> order;
        for (uint64_t k = raw->net_start[n]; k < raw->net_start[n + 1]; ++k) order.push_back(k);
        std::stable_sort(order.begin(), order.end(),
                         [&](uint64_t a, uint64_t b) { return raw->pin_vertex[a] < raw->pin_vertex[b]; });
        NetPins &net = nl.nets[n];
        net.weight = raw->net_weight[n];
        for (const uint64_t k : order) {
            if (!net.cells.empty() && net.cells.back() == raw->pin_vertex[k]) continue;
            net.cells.push_back(raw->pin_vertex[k]);
            net.offX.push_back(raw->pin_off_x[k]);
            net.offY.push_back(raw->pin_off_y[k]);
        }
    }
    nl.pinStart.assign(nl.numMovable + 1, 0);
    for (const NetPins &net : nl.nets)
        for (const uint32_t c : net.cells)
            if (nl.varOfVertex[c] != noVar) ++nl.pinStart[nl.varOfVertex[c] + 1];
    std::partial_sum(nl.pinStart.begin(), nl.pinStart.end(), nl.pinStart.begin());
    nl.pinsOfCell.resize(nl.pinStart.back());
    std::vector<std::size_t> cursor(nl.pinStart.begin(), nl.pinStart.end() - 1);
    for (uint32_t n = 0; n < nl.nets.size(); ++n)
        for (uint32_t q = 0; q < nl.nets[n].cells.size(); ++q) {
            const uint32_t v = nl.varOfVertex[nl.nets[n].cells[q]];
            if (v != noVar) nl.pinsOfCell[cursor[v]++] = {n, q};
        }
    for (uint32_t r = 0; r < raw->num_rows; ++r)
        d->rows.push_back({raw->row_y[r], raw->row_height[r], raw->row_pitch[r], raw->row_x0[r], raw->row_x1[r]});
    d->die = {{raw->die[0], raw->die[1]}, {raw->die[2], raw->die[3]}};
    collectFloors(*d);
    return reinterpret_cast<ko_design *>(d);
}

void ko_design_free(ko_design *h) { delete reinterpret_cast<Design *>(h); }

// Flattened netlist, same layout as ktg_flat, into caller buffers sized from
// ko_design_sizes: sizes = {V, n, N, P, Pm, M}.
void ko_design_sizes(const ko_design *h, uint64_t *sizes) {
    const Design &d = *reinterpret_cast<const Design *>(h);
    uint64_t P = 0;
    for (const NetPins &net : d.netlist.nets) P += net.cells.size();
    sizes[0] = d.netlist.varOfVertex.size();
    sizes[1] = d.netlist.numMovable;
    sizes[2] = d.netlist.nets.size();
    sizes[3] = P;
    sizes[4] = d.netlist.pinsOfCell.size();
    sizes[5] = d.obstacles.size();
}

void ko_design_flat(const ko_design *h, uint64_t *netStart, uint32_t *pinCell, double *offX, double *offY,
                    double *weight, uint32_t *varOfVertex, uint32_t *movVertex, uint64_t *pinStart,
                    uint32_t *cellPinNet, uint32_t *cellPinIdx, double *floors) {
    const Design &d = *reinterpret_cast<const Design *>(h);
    const Netlist &nl = d.netlist;
    uint64_t k = 0;
    netStart[0] = 0;
    for (std::size_t n = 0; n < nl.nets.size(); ++n) {
        const NetPins &net = nl.nets[n];
        for (std::size_t q = 0; q < net.cells.size(); ++q, ++k) {
            pinCell[k] = net.cells[q];
            offX[k] = net.offX[q];
            offY[k] = net.offY[q];
        }
        netStart[n + 1] = k;
        weight[n] = net.weight;
    }
    std::copy(nl.varOfVertex.begin(), nl.varOfVertex.end(), varOfVertex);
    std::copy(nl.movVertex.begin(), nl.movVertex.end(), movVertex);
    std::copy(nl.pinStart.begin(), nl.pinStart.end(), pinStart);
    for (std::size_t e = 0; e < nl.pinsOfCell.size(); ++e) {
        cellPinNet[e] = nl.pinsOfCell[e].net;
        cellPinIdx[e] = nl.pinsOfCell[e].pin;
    }
    floors[0] = d.floorX;
    floors[1] = d.floorY;
}

// simpl_lss.cc:11-27.  anchor pointers may be null when alpha == 0.
void ko_build_b2b(ko_design *h, const double *x, const double *y, double alpha, const double *ax, const double *ay,
                  double floorX, double floorY) {
    Design &d = *reinterpret_cast<Design *>(h);
    const std::size_t n = d.netlist.numMovable;
    std::vector<double> vx(x, x + n), vy(y, y + n), vax, vay;
    if (ax) vax.assign(ax, ax + n);
    if (ay) vay.assign(ay, ay + n);
    d.floorX = floorX;
    d.floorY = floorY;
    buildB2B(d, vx, vy, alpha, vax, vay);
}

uint64_t ko_build_star(ko_design *h) { return buildStar(*reinterpret_cast<Design *>(h)); }

void ko_system_sizes(const ko_design *h, int axis, uint64_t *rows, uint64_t *nnz) {
    const Design &d = *reinterpret_cast<const Design *>(h);
    const SparseSystem &s = axis == 0 ? d.sysX : d.sysY;
    *rows = s.size();
    *nnz = s.col.size();
}

void ko_system_get(const ko_design *h, int axis, uint64_t *rowPtr, uint32_t *col, double *val, double *diag,
                   double *rhs) {
    const Design &d = *reinterpret_cast<const Design *>(h);
    const SparseSystem &s = axis == 0 ? d.sysX : d.sysY;
    std::copy(s.rowPtr.begin(), s.rowPtr.end(), rowPtr);
    std::copy(s.col.begin(), s.col.end(), col);
    std::copy(s.val.begin(), s.val.end(), val);
    std::copy(s.diag.begin(), s.diag.end(), diag);
    std::copy(s.rhs.begin(), s.rhs.end(), rhs);
}

void ko_system_set(ko_design *h, int axis, uint64_t rows, const uint64_t *rowPtr, const uint32_t *col,
                   const double *val, const double *diag, const double *rhs) {
    Design &d = *reinterpret_cast<Design *>(h);
    SparseSystem &s = axis == 0 ? d.sysX : d.sysY;
    s.rowPtr.assign(rowPtr, rowPtr + rows + 1);
    s.col.assign(col, col + rowPtr[rows]);
    s.val.assign(val, val + rowPtr[rows]);
    s.diag.assign(diag, diag + rows);
    s.rhs.assign(rhs, rhs + rows);
}

void ko_multiply(const ko_design *h, int axis, const double *in, double *out) {
    const Design &d = *reinterpret_cast<const Design *>(h);
    const SparseSystem &s = axis == 0 ? d.sysX : d.sysY;
    std::vector<double> vin(in, in + s.size()), vout(s.size());
    s.multiply(vin, vout);
    std::copy(vout.begin(), vout.end(), out);
}

double ko_dot(const double *a, const double *b, uint64_t n) {
    std::vector<double> va(a, a + n), vb(b, b + n);
    return getDot(va, vb);
}

// simpl_lss.cc:203-290.  x, y in/out; history (may be null) receives up to
// max_iter rows of (resid x, resid y).  Returns the iteration count.
uint64_t ko_solve_both(ko_design *h, double *x, double *y, double tol, uint64_t maxIter, double *resid,
                       double *history) {
    Design &d = *reinterpret_cast<Design *>(h);
    std::vector<double> vx(x, x + d.sysX.size()), vy(y, y + d.sysY.size());
    std::vector<std::vector<double>> hist;
    const SolveResult r = solveBothAxes(d.sysX, vx, d.sysY, vy, tol, maxIter, history ? &hist : nullptr);
    std::copy(vx.begin(), vx.end(), x);
    std::copy(vy.begin(), vy.end(), y);
    resid[0] = r.resid.x;
    resid[1] = r.resid.y;
    if (history)
        for (std::size_t i = 0; i < hist.size(); ++i) {
            history[2 * i] = hist[i][0];
            history[2 * i + 1] = hist[i][1];
        }
    return r.iterations;
}

double ko_hpwl(const ko_design *h, const double *x, const double *y) {
    const Design &d = *reinterpret_cast<const Design *>(h);
    const std::size_t n = d.netlist.numMovable;
    return getHpwl(d, std::vector<double>(x, x + n), std::vector<double>(y, y + n));
}

void ko_seed_uniform(const ko_design *h, uint64_t seed, double *x, double *y) {
    const Design &d = *reinterpret_cast<const Design *>(h);
    std::vector<double> vx, vy;
    seedUniform(d, seed, vx, vy);
    std::copy(vx.begin(), vx.end(), x);
    std::copy(vy.begin(), vy.end(), y);
}

// simpl_placer.cc:181-224.  bins_per_axis 0 = the default rule.
uint64_t ko_build_grid(ko_design *h, uint64_t binsPerAxis, double densityLimit) {
    Design &d = *reinterpret_cast<Design *>(h);
    d.opts.binsPerAxis = binsPerAxis;
    d.densityLimit = std::clamp(densityLimit, 0.05, 1.0);
    const std::size_t per = getDefaultBinsPerAxis(d);
    buildGrid(d, per);
    return per;
}

void ko_grid_avail(const ko_design *h, double *avail, double *total) {
    const Design &d = *reinterpret_cast<const Design *>(h);
    std::copy(d.grid.avail.begin(), d.grid.avail.end(), avail);
    *total = d.grid.totalAvail;
}

void ko_occupancy(const ko_design *h, const double *x, const double *y, double *occ) {
    const Design &d = *reinterpret_cast<const Design *>(h);
    const std::size_t n = d.netlist.numMovable;
    const std::vector<double> o = getOccupancy(d, std::vector<double>(x, x + n), std::vector<double>(y, y + n));
    std::copy(o.begin(), o.end(), occ);
}

double ko_overflow(const ko_design *h, const double *occ) {
    const Design &d = *reinterpret_cast<const Design *>(h);
    return getOverflow(d, std::vector<double>(occ, occ + d.grid.size()));
}

// simpl_grid.h:179-231 for tests: profile of box along an axis into caller
// buffers (capacity cap); returns the number of edges.
uint64_t ko_profile(const ko_design *h, const double *box, int vertical, double *edges, double *prefix,
                    uint64_t cap) {
    const Design &d = *reinterpret_cast<const Design *>(h);
    const AxisProfile p = d.freeSpace.getProfile({{box[0], box[1]}, {box[2], box[3]}}, vertical != 0);
    for (std::size_t i = 0; i < p.edges.size() && i < cap; ++i) {
        edges[i] = p.edges[i];
        prefix[i] = p.prefix[i];
    }
    return p.edges.size();
}

void ko_set_lal_snapshot(int on) { g_lalSnapshot = on != 0; }

// getStripes (simpl_lal.cc:238-262) in fill order (the stable sort of
// nonlinearScale, :275-282) for tests: (start, end, capacity) triples into
// out (capacity cap stripes); returns the number of stripes.
uint64_t ko_stripes(const ko_design *h, const double *box, int vertical, double cell_area, double region_room,
                    double cutline, int below_cut, double *out, uint64_t cap) {
    const Design &d = *reinterpret_cast<const Design *>(h);
    const Lal lal{d, d.grid, d.freeSpace, d.opts, d.densityLimit, d.die};
    std::vector<Stripe> s =
        lal.getStripes({{box[0], box[1]}, {box[2], box[3]}}, vertical != 0, cell_area, region_room);
    const auto fromCut = [&](double c) { return below_cut ? cutline - c : c - cutline; };
    std::stable_sort(s.begin(), s.end(), [&](const Stripe &a, const Stripe &b) {
        return fromCut(a.span.getCenter()) > fromCut(b.span.getCenter());
    });
    for (std::size_t i = 0; i < s.size() && i < cap; ++i) {
        out[3 * i] = s[i].span.start;
        out[3 * i + 1] = s[i].span.end;
        out[3 * i + 2] = s[i].capacity;
    }
    return s.size();
}

void ko_set_mutation(int which) { g_mutation = which; }

// packStripes for tests, on caller arrays: k stripes (start, end, capacity) in
// any order, m cells with lower-left pos (in/out), extent size and area along
// the axis.  cell_stripe (may be null) receives each cell's stripe, as its
// index in the input order.
void ko_pack_stripes(uint64_t k, const double *stripes, uint64_t m, double *pos, const double *size,
                     const double *area, double cutline, int below_cut, int64_t *cell_stripe) {
    std::vector<Stripe> strips(k);
    for (uint64_t s = 0; s < k; ++s) strips[s] = {{stripes[3 * s], stripes[3 * s + 1]}, stripes[3 * s + 2], 0.0, {}};
    const std::vector<Stripe> input = strips;
    std::vector<double> vpos(pos, pos + m), vsize(size, size + m), varea(area, area + m);
    std::vector<std::uint32_t> cells(m);
    std::iota(cells.begin(), cells.end(), 0u);
    packStripes(strips, cells, cutline, below_cut != 0, vpos, vsize, varea);
    std::copy(vpos.begin(), vpos.end(), pos);
    if (cell_stripe) {
        for (const Stripe &strip : strips) {
            int64_t at = -1;
            for (uint64_t s = 0; s < k; ++s)
                if (input[s].span.start == strip.span.start && input[s].span.end == strip.span.end) at = int64_t(s);
            for (const std::uint32_t i : strip.cells) cell_stripe[i] = at;
        }
    }
}

// One split (simpl_lal.cc:213-264) of a block built from the arguments:
// cells (m of them) are cell ids of the design; x, y in/out.  out receives
// {split happened, m (cells in the low half), cutline}.
void ko_split(ko_design *h, const double *box, int vertical, uint64_t level, uint64_t m, const uint32_t *cells,
              double region_room, double *x, double *y, double *out) {
    Design &d = *reinterpret_cast<Design *>(h);
    const Lal lal{d, d.grid, d.freeSpace, d.opts, d.densityLimit, d.die};
    const std::size_t n = d.netlist.numMovable;
    std::vector<double> vx(x, x + n), vy(y, y + n);
    Block b{{{box[0], box[1]}, {box[2], box[3]}}, level, vertical != 0, std::vector<std::uint32_t>(cells, cells + m),
            region_room};
    const auto halves = lal.split(b, vx, vy);
    out[0] = halves ? 1.0 : 0.0;
    out[1] = halves ? static_cast<double>(halves->first.cells.size()) : 0.0;
    out[2] = halves ? (vertical ? halves->first.box.upperRight.x : halves->first.box.upperRight.y) : 0.0;
    std::copy(vx.begin(), vx.end(), x);
    std::copy(vy.begin(), vy.end(), y);
}

// getOverfilledClusters (simpl_lal.cc:141-183) on a caller occupancy: (x0, x1,
// y0, y1) per cluster into out (capacity cap); returns the cluster count.
uint64_t ko_overfilled_clusters(const ko_design *h, const double *occ, uint64_t *out, uint64_t cap) {
    const Design &d = *reinterpret_cast<const Design *>(h);
    const Lal lal{d, d.grid, d.freeSpace, d.opts, d.densityLimit, d.die};
    const std::vector<BinRect> r = lal.getOverfilledClusters(std::vector<double>(occ, occ + d.grid.size()));
    for (std::size_t i = 0; i < r.size() && i < cap; ++i) {
        out[4 * i] = r[i].x0;
        out[4 * i + 1] = r[i].x1;
        out[4 * i + 2] = r[i].y0;
        out[4 * i + 3] = r[i].y1;
    }
    return r.size();
}

// getGrown (simpl_lal.cc:185-211) from rect (x0, x1, y0, y1): every rectangle
// it passes through, the start first, into out (capacity cap); returns the
// number of rectangles.
uint64_t ko_grow(const ko_design *h, const double *occ, const uint64_t *rect, uint64_t *out, uint64_t cap) {
    const Design &d = *reinterpret_cast<const Design *>(h);
    const Lal lal{d, d.grid, d.freeSpace, d.opts, d.densityLimit, d.die};
    std::vector<BinRect> steps;
    lal.getGrown({rect[0], rect[1], rect[2], rect[3]}, std::vector<double>(occ, occ + d.grid.size()), &steps);
    for (std::size_t i = 0; i < steps.size() && i < cap; ++i) {
        out[4 * i] = steps[i].x0;
        out[4 * i + 1] = steps[i].x1;
        out[4 * i + 2] = steps[i].y0;
        out[4 * i + 3] = steps[i].y1;
    }
    return steps.size();
}

// Convergence::update over a sequence of (gap, upper HPWL) for iterations
// 1..n; returns the first iteration at which it says stop, 0 if none does.
uint64_t ko_convergence(uint64_t n, const double *gaps, const double *uppers, uint64_t reference_iter,
                        double tight_frac, double relaxed_frac, uint64_t stall_iters) {
    Convergence c;
    for (uint64_t it = 0; it < n; ++it)
        if (c.update(it + 1, gaps[it], uppers[it], reference_iter, tight_frac, relaxed_frac, stall_iters))
            return it + 1;
    return 0;
}

// LAL options of simpl_options.h:24-26 (0 keeps the current value)
void ko_set_lal_options(ko_design *h, uint64_t max_level, uint64_t min_cells_to_split,
                        double stripe_area_fraction) {
    Design &d = *reinterpret_cast<Design *>(h);
    if (max_level) d.opts.maxLevel = max_level;
    if (min_cells_to_split) d.opts.minCellsToSplit = min_cells_to_split;
    if (stripe_area_fraction > 0.0) d.opts.stripeAreaFraction = stripe_area_fraction;
}

void ko_lal(ko_design *h, double *x, double *y) {
    Design &d = *reinterpret_cast<Design *>(h);
    const std::size_t n = d.netlist.numMovable;
    std::vector<double> vx(x, x + n), vy(y, y + n);
    lookAheadLegalize(d, vx, vy);
    std::copy(vx.begin(), vx.end(), x);
    std::copy(vy.begin(), vy.end(), y);
}

// The whole SimPL global placement (simpl_placer.cc:358-580).  trace (may be
// null, capacity 3*max_iters) receives lower, upper, overflow per iteration.
void ko_place(ko_design *h, const ko_place_opts *o, double *x, double *y, ko_place_stats *stats, double *trace) {
    Design &d = *reinterpret_cast<Design *>(h);
    d.opts.binsPerAxis = o->bins_per_axis;
    d.opts.cgTol = o->cg_tol;
    d.opts.initMaxIters = o->init_max_iters;
    d.opts.initialNetModel = o->initial_net_model;
    d.opts.maxIters = o->max_iters;
    d.opts.seed = o->seed;
    d.opts.densityLimit = o->density_limit;
    d.opts.alphaBase = o->alpha_base;
    d.opts.refineIters = o->refine_iters;
    d.opts.cgMaxIter = o->cg_max_iter;
    Placer p{d};
    std::vector<double> ox, oy, tr;
    p.run(ox, oy, trace ? &tr : nullptr);
    std::copy(ox.begin(), ox.end(), x);
    std::copy(oy.begin(), oy.end(), y);
    if (trace) std::copy(tr.begin(), tr.end(), trace);
    const PlaceStats &s = p.st;
    *stats = {s.hpwlSeed, s.hpwlInit, s.hpwlLower, s.hpwlFinal, s.gap, s.overflowLower, s.overflowFinal,
              s.buildSeconds, s.solveSeconds, s.spreadSeconds, s.initIters, s.globalIters, s.cgIterations};
}

}  // extern "C"
