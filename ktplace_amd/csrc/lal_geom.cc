// Host geometry of look-ahead legalization; see lal_geom.h.
//
// Profile::before / coordinateAt, FreeSpace::profile, the free-area deposit
// of buildGrid and stripes() keep the reference's floating-point expressions
// and their order, because legalized positions are compared with the
// reference's bit for bit.  Those parts follow KTPlace (simpl_grid.h,
// simpl_lal.cc, simpl_placer.cc), which carries this notice:
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
#include "lal_geom.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <limits>
#include <numeric>

namespace kt::lal {

// ------------------------------------------------------------------ Profile
// simpl_grid.h:36-64
double Profile::before(double c) const {
    if (c <= edges.front()) return 0.0;
    if (c >= edges.back()) return total();
    const size_t k = static_cast<size_t>(std::upper_bound(edges.begin(), edges.end(), c) - edges.begin()) - 1;
    const double f = (c - edges[k]) / (edges[k + 1] - edges[k]);
    return prefix[k] + f * (prefix[k + 1] - prefix[k]);
}

double Profile::coordinateAt(double target) const {
    if (target <= 0.0) return edges.front();
    if (target >= total()) return edges.back();
    const size_t k = static_cast<size_t>(std::lower_bound(prefix.begin(), prefix.end(), target) - prefix.begin());
    const double step = prefix[k] - prefix[k - 1];
    const double f = step > 0.0 ? (target - prefix[k - 1]) / step : 1.0;
    return edges[k - 1] + f * (edges[k] - edges[k - 1]);
}

// ------------------------------------------------------------------ FreeSpace
// The free part of each row (simpl_grid.h:130-178 defines it): the row's spans
// minus the x extent of every obstacle whose height overlaps the row.  The
// obstacles of a row are first merged into disjoint blocked intervals
// (intervals that overlap or touch become one; a zero-width obstacle stays a
// point that splits a span); each span then keeps its gaps between
// consecutive blocked intervals, clipped to the span, where they have
// positive length.  Rows end up ordered by their bottom edge and each row's
// spans by (start, end).
void FreeSpace::build(std::vector<FreeRow> raw, const std::vector<Box> &obstacles) {
    rows = std::move(raw);
    std::stable_sort(rows.begin(), rows.end(), [](const FreeRow &a, const FreeRow &b) { return a.y0 < b.y0; });
    std::vector<Span> blocked;
    for (FreeRow &row : rows) {
        blocked.clear();
        for (const Box &o : obstacles)
            if (o.y1 > row.y0 && o.y0 < row.y1) blocked.push_back({o.x0, o.x1});
        std::sort(blocked.begin(), blocked.end());
        size_t m = 0;
        for (const Span &c : blocked) {
            if (m > 0 && c.a <= blocked[m - 1].b)
                blocked[m - 1].b = std::max(blocked[m - 1].b, c.b);
            else
                blocked[m++] = c;
        }
        blocked.resize(m);
        std::stable_sort(row.spans.begin(), row.spans.end());
        std::vector<Span> free;
        for (const Span &span : row.spans) {
            double from = span.a;  // start of the current gap
            for (const Span &c : blocked) {
                const double to = std::min(span.b, c.a);
                if (to > from) free.push_back({from, to});
                from = std::max(from, c.b);
            }
            if (span.b > from) free.push_back({from, span.b});
        }
        row.spans = std::move(free);
    }
}

// simpl_grid.h:179-231
Profile FreeSpace::profile(const Box &b, bool vertical) const {
    const auto first = std::upper_bound(rows.begin(), rows.end(), b.y0,
                                        [](double v, const FreeRow &r) { return v < r.y1; });
    const double lo = vertical ? b.x0 : b.y0;
    const double hi = vertical ? b.x1 : b.y1;
    Profile out;
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
        for (auto r = first; r != rows.end() && r->y0 < b.y1; ++r) {
            const double h = std::min(b.y1, r->y1) - std::max(b.y0, r->y0);
            for (const Span &span : r->spans) {
                const double a = std::max(span.a, lo);
                const double z = std::min(span.b, hi);
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
        for (auto r = first; r != rows.end() && r->y0 < b.y1; ++r) {
            const double a = std::max(lo, r->y0);
            const double z = std::min(hi, r->y1);
            if (z > a) {
                double len = 0.0;
                for (const Span &span : r->spans) len += std::max(0.0, std::min(b.x1, span.b) - std::max(b.x0, span.a));
                extend(a, 0.0);
                extend(z, len * (z - a));
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

// ------------------------------------------------------------------ Grid
double Grid::edge(bool vertical, size_t t) const {
    return vertical ? ox + static_cast<double>(t) * dx : oy + static_cast<double>(t) * dy;
}

Box Grid::box(const BinRect &r) const { return {edge(true, r.x0), edge(false, r.y0), edge(true, r.x1 + 1), edge(false, r.y1 + 1)}; }

double Grid::sum(const std::vector<double> &field, const BinRect &r) const {
    double total = 0.0;
    for (size_t iy = r.y0; iy <= r.y1; ++iy)
        for (size_t ix = r.x0; ix <= r.x1; ++ix) total += field[at(ix, iy)];
    return total;
}

namespace {
size_t clampIndex(double f, size_t n) { return static_cast<size_t>(std::clamp(f, 0.0, static_cast<double>(n) - 1e-9)); }
}  // namespace

// simpl_placer.cc:190-224 with BinGrid::deposit (simpl_grid.h:104-118)
void buildGrid(Grid &g, FreeSpace &fs, const std::vector<RowSpec> &rows, const std::vector<Box> &obstacles,
               const Box &die, size_t perAxis) {
    g.ox = die.x0;
    g.oy = die.y0;
    g.nx = perAxis;
    g.ny = perAxis;
    g.dx = std::max(die.x1 - die.x0, 1e-9) / static_cast<double>(perAxis);
    g.dy = std::max(die.y1 - die.y0, 1e-9) / static_cast<double>(perAxis);
    g.avail.assign(g.size(), 0.0);
    std::vector<FreeRow> raw;
    double smallestSite = std::numeric_limits<double>::max();
    // One RowSpec per subrow; consecutive subrows of one row (same y and
    // height) become the spans of one free-space row, as a row's subrows do.
    for (const RowSpec &r : rows) {
        if (!(r.height > 0.0) || !(r.pitch > 0.0)) continue;
        smallestSite = std::min(smallestSite, r.pitch * r.height);
        if (!raw.empty() && raw.back().y0 == r.y && raw.back().y1 == r.y + r.height)
            raw.back().spans.push_back({r.x0, r.x1});
        else
            raw.push_back({r.y, r.y + r.height, {{r.x0, r.x1}}});
    }
    fs.build(std::move(raw), obstacles);
    const auto column = [&](double x) { return clampIndex((x - g.ox) / g.dx, g.nx); };
    const auto rowOf = [&](double y) { return clampIndex((y - g.oy) / g.dy, g.ny); };
    for (const FreeRow &row : fs.rows) {
        for (const Span &span : row.spans) {
            const Box b{span.a, row.y0, span.b, row.y1};
            if (!(b.x1 > b.x0) || !(b.y1 > b.y0)) continue;
            for (size_t iy = rowOf(b.y0); iy <= rowOf(b.y1); ++iy) {
                const double oy = std::min(b.y1, g.edge(false, iy + 1)) - std::max(b.y0, g.edge(false, iy));
                for (size_t ix = column(b.x0); ix <= column(b.x1); ++ix) {
                    const double ox = std::min(b.x1, g.edge(true, ix + 1)) - std::max(b.x0, g.edge(true, ix));
                    if (ox > 0.0 && oy > 0.0) g.avail[g.at(ix, iy)] += ox * oy;
                }
            }
        }
    }
    const double sliver = smallestSite < std::numeric_limits<double>::max() ? 1e-3 * smallestSite : 0.0;
    for (double &a : g.avail) a = a < sliver ? 0.0 : a;
    g.totalAvail = std::accumulate(g.avail.begin(), g.avail.end(), 0.0);
}

// ------------------------------------------------------------------ clusters
void AreaTables::buildAvail(const Grid &g) {
    nx = g.nx;
    ny = g.ny;
    // free area in units of 2^-availShift: exact integer sums (modulo 2^64,
    // which cancels in the rectangle difference).  A rectangle's sum must
    // stay below 2^63, so the shift is 40 (per-bin rounding below 1e-12)
    // unless the whole die's free area times 2^40 reaches 2^62; then it is
    // the largest shift that keeps the total below 2^62 (a die of 6.5e7 area
    // units takes 36, per-bin rounding below 1e-11).
    double total = 0.0;
    for (const double a : g.avail) total += a;
    availShift = 40;
    while (availShift > 0 && std::ldexp(total, availShift) >= 0x1p62) --availShift;
    avail.assign((nx + 1) * (ny + 1), 0);
    for (size_t iy = 0; iy < ny; ++iy)
        for (size_t ix = 0; ix < nx; ++ix) {
            const size_t o = (iy + 1) * (nx + 1) + ix + 1;
            const uint64_t v =
                static_cast<uint64_t>(std::llround(std::ldexp(g.avail[g.at(ix, iy)], availShift)));
            avail[o] = v + avail[o - 1] + avail[o - (nx + 1)] - avail[o - (nx + 1) - 1];
        }
    occRow.assign(ny * (nx + 1), 0);
}

void AreaTables::setOcc(const std::vector<uint64_t> &occFix) {
    for (size_t iy = 0; iy < ny; ++iy) updateRow(occFix, iy);
}

void AreaTables::updateRow(const std::vector<uint64_t> &occFix, size_t iy) {
    int64_t *row = occRow.data() + iy * (nx + 1);
    row[0] = 0;
    for (size_t ix = 0; ix < nx; ++ix) row[ix + 1] = row[ix] + static_cast<int64_t>(occFix[iy * nx + ix]);
}

namespace {
uint64_t rect(const std::vector<uint64_t> &s, size_t nx, const BinRect &r) {
    const size_t w = nx + 1;
    return s[(r.y1 + 1) * w + r.x1 + 1] - s[r.y0 * w + r.x1 + 1] - s[(r.y1 + 1) * w + r.x0] + s[r.y0 * w + r.x0];
}
}  // namespace

double AreaTables::sumOcc(const BinRect &r) const {
    int64_t t = 0;
    for (size_t iy = r.y0; iy <= r.y1; ++iy) {
        const int64_t *row = occRow.data() + iy * (nx + 1);
        t += row[r.x1 + 1] - row[r.x0];
    }
    return static_cast<double>(t) / 65536.0;
}
double AreaTables::sumAvail(const BinRect &r) const {
    return std::ldexp(static_cast<double>(static_cast<int64_t>(rect(avail, nx, r))), -availShift);
}

// simpl_lal.cc:86-93
double AreaTables::density(const BinRect &r) const {
    const double cells = sumOcc(r);
    const double room = sumAvail(r);
    if (room > 0.0) return cells / room;
    return cells > 0.0 ? std::numeric_limits<double>::infinity() : 0.0;
}

// Overfilled clusters (simpl_lal.cc:141-183 defines them): the bounding
// boxes of the 4-connected components of bins with occ > g * avail, in the
// order of each component's first bin in row-major order.  Labelled in one
// row-major pass with a union-find over bin indices (each overfilled bin
// joins its left and lower neighbours), then a second pass that grows each
// root's box and lists the roots as they are first met.
std::vector<BinRect> overfilledClusters(const Grid &g, const std::vector<double> &occ, double densityLimit) {
    const size_t nb = g.size();
    constexpr uint32_t kNone = 0xffffffffu;
    std::vector<uint32_t> parent(nb, kNone);
    const auto root = [&parent](uint32_t k) {
        uint32_t r = k;
        while (parent[r] != r) r = parent[r];
        while (parent[k] != r) {  // path compression
            const uint32_t next = parent[k];
            parent[k] = r;
            k = next;
        }
        return r;
    };
    const auto join = [&](uint32_t a, uint32_t b) {
        const uint32_t ra = root(a), rb = root(b);
        if (ra != rb) parent[std::max(ra, rb)] = std::min(ra, rb);  // the smaller index stays the root
    };
    for (size_t iy = 0; iy < g.ny; ++iy)
        for (size_t ix = 0; ix < g.nx; ++ix) {
            const size_t k = g.at(ix, iy);
            if (!(occ[k] > densityLimit * g.avail[k])) continue;
            parent[k] = static_cast<uint32_t>(k);
            if (ix > 0 && parent[k - 1] != kNone) join(static_cast<uint32_t>(k - 1), static_cast<uint32_t>(k));
            if (iy > 0 && parent[k - g.nx] != kNone)
                join(static_cast<uint32_t>(k - g.nx), static_cast<uint32_t>(k));
        }
    std::vector<uint32_t> slot(nb, kNone);  // root -> output index
    std::vector<BinRect> out;
    for (size_t k = 0; k < nb; ++k) {
        if (parent[k] == kNone) continue;
        const uint32_t r = root(static_cast<uint32_t>(k));
        const size_t ix = k % g.nx, iy = k / g.nx;
        if (slot[r] == kNone) {
            slot[r] = static_cast<uint32_t>(out.size());
            out.push_back({ix, ix, iy, iy});
        } else {
            BinRect &b = out[slot[r]];
            b = {std::min(b.x0, ix), std::max(b.x1, ix), std::min(b.y0, iy), std::max(b.y1, iy)};
        }
    }
    return out;
}

namespace {

// getDensity with the reference's sequential rectangle sums (simpl_lal.cc:86-93)
double exactDensity(const Grid &g, const std::vector<double> &occ, const BinRect &r) {
    const double cells = g.sum(occ, r);
    const double room = g.sum(g.avail, r);
    if (room > 0.0) return cells / room;
    return cells > 0.0 ? std::numeric_limits<double>::infinity() : 0.0;
}

bool near(double a, double b) {
    if (std::isinf(a) || std::isinf(b)) return a == b;
    return std::fabs(a - b) <= 1e-7 * std::max({std::fabs(a), std::fabs(b), 1e-300});
}

}  // namespace

// Greedy growth (simpl_lal.cc:185-211).  Densities come from the summed-area
// tables; a rectangle of at most 64 bins, or any comparison the tables leave
// within 1e-7 relative, is evaluated with the reference's sequential sums, so
// every decision matches the reference's arithmetic.
BinRect grow(const Grid &g, const AreaTables &t, const std::vector<double> &occ, BinRect r, double densityLimit) {
    const auto density = [&](const BinRect &b, bool exact) {
        return (exact || b.bins() <= 64) ? exactDensity(g, occ, b) : t.density(b);
    };
    const auto current = [&](const BinRect &b) {
        const double d = density(b, false);
        return near(d, densityLimit) ? density(b, true) : d;
    };
    double cur = current(r);
    while (cur > densityLimit) {
        const std::array<BinRect, 4> wider{BinRect{r.x0 > 0 ? r.x0 - 1 : r.x0, r.x1, r.y0, r.y1},
                                           BinRect{r.x0, r.x1 + 1 < g.nx ? r.x1 + 1 : r.x1, r.y0, r.y1},
                                           BinRect{r.x0, r.x1, r.y0 > 0 ? r.y0 - 1 : r.y0, r.y1},
                                           BinRect{r.x0, r.x1, r.y0, r.y1 + 1 < g.ny ? r.y1 + 1 : r.y1}};
        std::array<double, 4> d{};
        std::array<bool, 4> valid{};
        for (int k = 0; k < 4; ++k) {
            valid[k] = wider[k].bins() != r.bins();
            if (valid[k]) d[k] = density(wider[k], false);
        }
        bool close = false;
        for (int a = 0; a < 4; ++a)
            for (int b = a + 1; b < 4; ++b) close = close || (valid[a] && valid[b] && near(d[a], d[b]));
        if (close)
            for (int k = 0; k < 4; ++k)
                if (valid[k]) d[k] = density(wider[k], true);
        int best = -1;
        double bestDensity = std::numeric_limits<double>::infinity();
        for (int k = 0; k < 4; ++k) {
            if (!valid[k]) continue;
            if (best < 0 || d[k] < bestDensity) {
                best = k;
                bestDensity = d[k];
            }
        }
        if (best < 0) break;
        r = wider[best];
        cur = near(bestDensity, densityLimit) ? density(r, true) : bestDensity;
    }
    return r;
}

// ------------------------------------------------------------------ stripes
// simpl_lal.cc:213-263 (getStripes) and the stripe order of nonlinearScale
// (:275-282).
std::vector<Stripe> stripes(const FreeSpace &fs, const std::vector<Box> &obstacles, const Box &box, bool vertical,
                            double cellArea, double regionRoom, double densityLimit, double stripeAreaFraction,
                            double cutline, bool belowCut) {
    const double lo = vertical ? box.x0 : box.y0;
    const double hi = vertical ? box.x1 : box.y1;
    const Profile room = fs.profile(box, vertical);
    std::vector<double> cuts{lo, hi};
    for (const Box &o : obstacles) {
        if (o.x1 <= box.x0 || o.x0 >= box.x1 || o.y1 <= box.y0 || o.y0 >= box.y1) continue;
        for (const double e : {vertical ? o.x0 : o.y0, vertical ? o.x1 : o.y1})
            if (e > lo && e < hi) cuts.push_back(e);
    }
    std::stable_sort(cuts.begin(), cuts.end());
    cuts.erase(std::unique(cuts.begin(), cuts.end()), cuts.end());
    const double limit = stripeAreaFraction * (regionRoom > 0.0 ? regionRoom : room.total());
    const double fill = room.total() > 0.0 ? std::min(densityLimit, cellArea / room.total()) : densityLimit;
    for (size_t s = 0; s + 1 < cuts.size();) {
        const double a = cuts[s], b = cuts[s + 1];
        const double here = room.between(a, b);
        const double mid = room.coordinateAt(room.before(a) + 0.5 * here);
        if (here > limit && mid > a && mid < b) {
            cuts.insert(cuts.begin() + static_cast<long>(s) + 1, mid);
        } else {
            ++s;
        }
    }
    std::vector<Stripe> out;
    for (size_t s = 0; s + 1 < cuts.size(); ++s) out.push_back({cuts[s], cuts[s + 1], fill * room.between(cuts[s], cuts[s + 1])});
    const auto fromCut = [&](double c) { return belowCut ? cutline - c : c - cutline; };
    std::stable_sort(out.begin(), out.end(), [&](const Stripe &a, const Stripe &b) {
        return fromCut(0.5 * (a.start + a.end)) > fromCut(0.5 * (b.start + b.end));
    });
    return out;
}

}  // namespace kt::lal
