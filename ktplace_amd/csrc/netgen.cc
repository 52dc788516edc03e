// Seeded synthetic placement designs and the flattening of a design into
// SimPL's solver netlist.
//
// Generator rules (frozen: fixtures and measured iteration counts depend on
// them):
//   * RNG: std::mt19937_64(seed).  unit() = (g() >> 11) * 2^-53 (the reference's
//     seedUniform draw, simpl_placer.cc:228-230); below(k) = g() % k.
//   * Movable cells: width 1 + below(8) sites at a 1.0 site pitch, height 1.0.
//   * Macros: width 20 + below(181), height 2 + below(19); placed by rejection
//     (up to 1000 draws each) so that no two macros overlap.
//   * Die: core = movable area / utilization + macro area; rows = ceil(sqrt(core)),
//     row width = ceil(core / rows); one subrow per row spanning the die.
//   * Pads: zero-size, evenly spaced along the rectangle [0, W-1] x [0, R-1].
//   * Vertex ids: fixed vertex k (pads then macros) sits at id
//     floor((k + 0.5) * V / F); movable cells fill the remaining ids in order.
//   * Locality: movable cell c has a virtual grid point (c % S, c / S) with
//     S = ceil(sqrt(n)).  Net j's first pin is pad j (j < pads), macro
//     (j - pads) / 4 (next 4 * macros nets) or a uniform anchor cell.  Each
//     further pin is, with probability local_frac, a cell in the
//     (2r+1) x (2r+1) window around the anchor's grid point, r = 1 + ceil(sqrt(p)),
//     and otherwise a uniform cell.  A net may name a cell twice; flattening
//     deduplicates.
//   * Degree: u = unit(); < 0.60 -> 2; < 0.80 -> 3; < 0.95 -> 4..10;
//     < 1 - high_fanout_frac -> 11..40; else 41..max_degree.
//   * Pin offsets: width * unit(), height * unit() from the lower-left.
//   * A movable cell left on no net is appended to the first net of a cell in
//     its window (8 draws), else to a uniform net.
#include "design_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

namespace ktgi {

namespace {
thread_local std::string lastError;
}  // namespace

int fail(const std::string &msg) {
    lastError = msg;
    return -1;
}

const char *lastErrorText() { return lastError.c_str(); }

}  // namespace ktgi

namespace {

using ktgi::fail;

struct Rng {
    std::mt19937_64 g;
    explicit Rng(uint64_t seed) : g(seed) {}
    double unit() { return static_cast<double>(g() >> 11) * 0x1.0p-53; }
    uint64_t below(uint64_t k) { return g() % k; }
};

uint32_t sampleDegree(Rng &rng, const ktg_spec &s) {
    const double u = rng.unit();
    if (u < 0.60) return 2;
    if (u < 0.80) return 3;
    if (u < 0.95) return 4 + static_cast<uint32_t>(rng.below(7));
    if (u < 1.0 - s.high_fanout_frac || s.max_degree <= 40)
        return 11 + static_cast<uint32_t>(rng.below(std::min<uint32_t>(s.max_degree, 40) - 10));
    return 41 + static_cast<uint32_t>(rng.below(s.max_degree - 40));
}

struct PinDraft {
    uint32_t vertex;
    double offX, offY;
};

int generate(const ktg_spec &s, ktg_design &d) {
    if (s.num_cells == 0 || s.num_nets == 0) return fail("ktg_generate: empty design");
    if (!(s.utilization > 0.0 && s.utilization <= 1.0)) return fail("ktg_generate: utilization");
    if (s.max_degree < 40) return fail("ktg_generate: max_degree must be at least 40");
    Rng rng(s.seed);
    const uint32_t n = s.num_cells;
    std::vector<double> cw(n);
    double movArea = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        cw[i] = 1.0 + static_cast<double>(rng.below(8));
        movArea += cw[i];
    }
    std::vector<double> mw(s.num_macros), mh(s.num_macros);
    double macArea = 0.0;
    for (uint32_t m = 0; m < s.num_macros; ++m) {
        mw[m] = 20.0 + static_cast<double>(rng.below(181));
        mh[m] = 2.0 + static_cast<double>(rng.below(19));
        macArea += mw[m] * mh[m];
    }
    const double core = movArea / s.utilization + macArea;
    const uint32_t rows = static_cast<uint32_t>(std::ceil(std::sqrt(core)));
    const uint32_t W = static_cast<uint32_t>(std::ceil(core / rows));
    d.die[0] = 0.0;
    d.die[1] = 0.0;
    d.die[2] = W;
    d.die[3] = rows;
    for (uint32_t r = 0; r < rows; ++r) {
        d.rowY.push_back(r);
        d.rowH.push_back(1.0);
        d.rowX0.push_back(0.0);
        d.rowX1.push_back(W);
        d.rowPitch.push_back(1.0);
    }
    // Macro placement by rejection.
    std::vector<double> mx(s.num_macros), my(s.num_macros);
    for (uint32_t m = 0; m < s.num_macros; ++m) {
        const uint64_t spanX = W > mw[m] ? static_cast<uint64_t>(W - mw[m]) + 1 : 1;
        const uint64_t spanY = rows > mh[m] ? static_cast<uint64_t>(rows - mh[m]) + 1 : 1;
        for (int attempt = 0; attempt < 1000; ++attempt) {
            mx[m] = static_cast<double>(rng.below(spanX));
            my[m] = static_cast<double>(rng.below(spanY));
            bool clash = false;
            for (uint32_t o = 0; o < m && !clash; ++o) {
                clash = mx[m] < mx[o] + mw[o] && mx[o] < mx[m] + mw[m] && my[m] < my[o] + mh[o] &&
                        my[o] < my[m] + mh[m];
            }
            if (!clash) break;
        }
    }
    // Vertex layout.
    const uint32_t F = s.num_pads + s.num_macros;
    const uint32_t V = n + F;
    d.isFixed.assign(V, 0);
    std::vector<uint32_t> fixedId(F);
    for (uint32_t k = 0; k < F; ++k) {
        const uint32_t v = static_cast<uint32_t>(
            std::floor((static_cast<double>(k) + 0.5) * static_cast<double>(V) / static_cast<double>(F)));
        fixedId[k] = v;
        d.isFixed[v] = 1;
    }
    std::vector<uint32_t> cellVertex(n);
    for (uint32_t v = 0, c = 0; v < V; ++v) {
        if (!d.isFixed[v]) cellVertex[c++] = v;
    }
    d.posX.assign(V, 0.0);
    d.posY.assign(V, 0.0);
    d.width.assign(V, 0.0);
    d.height.assign(V, 0.0);
    for (uint32_t c = 0; c < n; ++c) {
        d.width[cellVertex[c]] = cw[c];
        d.height[cellVertex[c]] = 1.0;
    }
    const double per = 2.0 * (W - 1.0) + 2.0 * (rows - 1.0);
    std::vector<double> padVX(s.num_pads), padVY(s.num_pads);
    const uint32_t S = static_cast<uint32_t>(std::ceil(std::sqrt(static_cast<double>(n))));
    const uint32_t SR = (n + S - 1) / S;
    for (uint32_t k = 0; k < s.num_pads; ++k) {
        double t = (k + 0.5) * per / s.num_pads;
        double px, py;
        const double w1 = W - 1.0, h1 = rows - 1.0;
        if (t < w1) {
            px = t;
            py = 0.0;
        } else if ((t -= w1) < h1) {
            px = w1;
            py = t;
        } else if ((t -= h1) < w1) {
            px = w1 - t;
            py = h1;
        } else {
            t -= w1;
            px = 0.0;
            py = std::max(h1 - t, 0.0);
        }
        px = std::floor(px);
        py = std::floor(py);
        d.posX[fixedId[k]] = px;
        d.posY[fixedId[k]] = py;
        padVX[k] = w1 > 0 ? px / w1 * (S - 1) : 0.0;
        padVY[k] = h1 > 0 ? py / h1 * (SR - 1) : 0.0;
    }
    for (uint32_t m = 0; m < s.num_macros; ++m) {
        const uint32_t v = fixedId[s.num_pads + m];
        d.posX[v] = mx[m];
        d.posY[v] = my[m];
        d.width[v] = mw[m];
        d.height[v] = mh[m];
    }
    // Nets.
    std::vector<std::vector<PinDraft>> nets(s.num_nets);
    const auto cellPin = [&](uint32_t c) {
        const uint32_t v = cellVertex[c];
        return PinDraft{v, d.width[v] * rng.unit(), d.height[v] * rng.unit()};
    };
    const auto windowCell = [&](long cx, long cy, long r) {
        long vx = cx + static_cast<long>(rng.below(2 * r + 1)) - r;
        long vy = cy + static_cast<long>(rng.below(2 * r + 1)) - r;
        vx = std::clamp<long>(vx, 0, S - 1);
        vy = std::clamp<long>(vy, 0, SR - 1);
        return static_cast<uint32_t>((static_cast<uint64_t>(vy) * S + vx) % n);
    };
    const uint32_t macroNets = std::min<uint64_t>(4ull * s.num_macros, s.num_nets > s.num_pads ? s.num_nets - s.num_pads : 0);
    for (uint32_t j = 0; j < s.num_nets; ++j) {
        const uint32_t p = sampleDegree(rng, s);
        std::vector<PinDraft> &net = nets[j];
        long cx, cy;
        if (j < s.num_pads) {
            const uint32_t v = fixedId[j];
            net.push_back({v, 0.0, 0.0});
            cx = std::lround(padVX[j]);
            cy = std::lround(padVY[j]);
        } else if (j < s.num_pads + macroNets) {
            const uint32_t m = (j - s.num_pads) / 4;
            const uint32_t v = fixedId[s.num_pads + m];
            net.push_back({v, mw[m] * rng.unit(), mh[m] * rng.unit()});
            cx = std::lround((mx[m] + 0.5 * mw[m]) / std::max(W - 1.0, 1.0) * (S - 1));
            cy = std::lround((my[m] + 0.5 * mh[m]) / std::max(rows - 1.0, 1.0) * (SR - 1));
        } else {
            const uint32_t a = static_cast<uint32_t>(rng.below(n));
            net.push_back(cellPin(a));
            cx = a % S;
            cy = a / S;
        }
        const long r = 1 + static_cast<long>(std::ceil(std::sqrt(static_cast<double>(p))));
        for (uint32_t q = 1; q < p; ++q) {
            const uint32_t c = rng.unit() < s.local_frac ? windowCell(cx, cy, r)
                                                         : static_cast<uint32_t>(rng.below(n));
            net.push_back(cellPin(c));
        }
    }
    // Every movable cell on at least one net.
    std::vector<uint32_t> firstNet(V, 0xffffffffu);
    for (uint32_t j = 0; j < s.num_nets; ++j) {
        for (const PinDraft &pd : nets[j]) {
            if (firstNet[pd.vertex] == 0xffffffffu) firstNet[pd.vertex] = j;
        }
    }
    std::vector<uint32_t> vertexCell(V, 0xffffffffu);
    for (uint32_t c = 0; c < n; ++c) vertexCell[cellVertex[c]] = c;
    for (uint32_t c = 0; c < n; ++c) {
        if (firstNet[cellVertex[c]] != 0xffffffffu) continue;
        uint32_t target = 0xffffffffu;
        for (int t = 0; t < 8 && target == 0xffffffffu; ++t) {
            const uint32_t nb = windowCell(c % S, c / S, 2);
            target = firstNet[cellVertex[nb]];
        }
        if (target == 0xffffffffu) target = static_cast<uint32_t>(rng.below(s.num_nets));
        nets[target].push_back(cellPin(c));
        firstNet[cellVertex[c]] = target;
    }
    d.netStart.assign(s.num_nets + 1, 0);
    for (uint32_t j = 0; j < s.num_nets; ++j) d.netStart[j + 1] = d.netStart[j] + nets[j].size();
    d.pinVertex.resize(d.netStart.back());
    d.pinOffX.resize(d.netStart.back());
    d.pinOffY.resize(d.netStart.back());
    for (uint32_t j = 0; j < s.num_nets; ++j) {
        uint64_t k = d.netStart[j];
        for (const PinDraft &pd : nets[j]) {
            d.pinVertex[k] = pd.vertex;
            d.pinOffX[k] = pd.offX;
            d.pinOffY[k] = pd.offY;
            ++k;
        }
    }
    d.netWeight.assign(s.num_nets, 1.0);
    d.numMovable = n;
    return 0;
}


}  // namespace

namespace ktgi {

int collect(const ktg_raw &raw, ktg_design &d) {
    const uint32_t V = raw.num_vertices;
    // raw copy (so the handle is self-contained)
    d.isFixed.assign(raw.is_fixed, raw.is_fixed + V);
    d.posX.assign(raw.pos_x, raw.pos_x + V);
    d.posY.assign(raw.pos_y, raw.pos_y + V);
    d.width.assign(raw.width, raw.width + V);
    d.height.assign(raw.height, raw.height + V);
    d.netStart.assign(raw.net_start, raw.net_start + raw.num_nets + 1);
    d.pinVertex.assign(raw.pin_vertex, raw.pin_vertex + raw.num_pins);
    d.pinOffX.assign(raw.pin_off_x, raw.pin_off_x + raw.num_pins);
    d.pinOffY.assign(raw.pin_off_y, raw.pin_off_y + raw.num_pins);
    d.netWeight.assign(raw.net_weight, raw.net_weight + raw.num_nets);
    d.rowY.assign(raw.row_y, raw.row_y + raw.num_rows);
    d.rowH.assign(raw.row_height, raw.row_height + raw.num_rows);
    d.rowX0.assign(raw.row_x0, raw.row_x0 + raw.num_rows);
    d.rowX1.assign(raw.row_x1, raw.row_x1 + raw.num_rows);
    d.rowPitch.assign(raw.row_pitch, raw.row_pitch + raw.num_rows);
    std::memcpy(d.die, raw.die, sizeof(d.die));
    if (raw.net_start[raw.num_nets] != raw.num_pins) return fail("ktg_collect: net_start[N] != num_pins");
    for (uint64_t k = 0; k < raw.num_pins; ++k) {
        if (raw.pin_vertex[k] >= V) return fail("ktg_collect: pin vertex out of range");
    }

    // collectCells (simpl_dm.cc:16-43)
    d.varOfVertex.assign(V, KTG_NO_VAR);
    d.fixedX.assign(raw.pos_x, raw.pos_x + V);
    d.fixedY.assign(raw.pos_y, raw.pos_y + V);
    d.movVertex.clear();
    for (uint32_t v = 0; v < V; ++v) {
        if (raw.is_fixed[v]) {
            if (raw.width[v] > 0.0 && raw.height[v] > 0.0) {
                d.obstacles.insert(d.obstacles.end(), {raw.pos_x[v], raw.pos_y[v], raw.pos_x[v] + raw.width[v],
                                                       raw.pos_y[v] + raw.height[v]});
            }
            continue;
        }
        d.varOfVertex[v] = static_cast<uint32_t>(d.movVertex.size());
        d.movVertex.push_back(v);
        d.mWidth.push_back(raw.width[v]);
        d.mHeight.push_back(raw.height[v]);
        d.mArea.push_back(raw.width[v] * raw.height[v]);
    }
    const uint32_t n = static_cast<uint32_t>(d.movVertex.size());
    d.numMovable = n;

    // collectNets (simpl_dm.cc:45-85): stable sort by cell, keep first pin per cell.
    const uint32_t N = raw.num_nets;
    d.fNetStart.assign(N + 1, 0);
    d.fPinCell.clear();
    d.fOffX.clear();
    d.fOffY.clear();
    d.fPinCell.reserve(raw.num_pins);
    d.fOffX.reserve(raw.num_pins);
    d.fOffY.reserve(raw.num_pins);
    std::vector<uint64_t> order;
    for (uint32_t j = 0; j < N; ++j) {
        const uint64_t b = raw.net_start[j], e = raw.net_start[j + 1];
        order.resize(e - b);
        std::iota(order.begin(), order.end(), b);
        std::stable_sort(order.begin(), order.end(), [&](uint64_t a, uint64_t c) {
            return raw.pin_vertex[a] < raw.pin_vertex[c];
        });
        const uint64_t first = d.fPinCell.size();
        for (const uint64_t k : order) {
            if (d.fPinCell.size() > first && d.fPinCell.back() == raw.pin_vertex[k]) continue;
            d.fPinCell.push_back(raw.pin_vertex[k]);
            d.fOffX.push_back(raw.pin_off_x[k]);
            d.fOffY.push_back(raw.pin_off_y[k]);
        }
        d.fNetStart[j + 1] = d.fPinCell.size();
    }
    d.fWeight.assign(raw.net_weight, raw.net_weight + N);
    d.pinStart.assign(n + 1, 0);
    for (uint64_t k = 0; k < d.fPinCell.size(); ++k) {
        const uint32_t v = d.varOfVertex[d.fPinCell[k]];
        if (v != KTG_NO_VAR) ++d.pinStart[v + 1];
    }
    std::partial_sum(d.pinStart.begin(), d.pinStart.end(), d.pinStart.begin());
    d.cellPinNet.resize(d.pinStart.back());
    d.cellPinIdx.resize(d.pinStart.back());
    std::vector<uint64_t> cursor(d.pinStart.begin(), d.pinStart.end() - 1);
    for (uint32_t j = 0; j < N; ++j) {
        for (uint64_t k = d.fNetStart[j]; k < d.fNetStart[j + 1]; ++k) {
            const uint32_t v = d.varOfVertex[d.fPinCell[k]];
            if (v != KTG_NO_VAR) {
                d.cellPinNet[cursor[v]] = j;
                d.cellPinIdx[cursor[v]] = static_cast<uint32_t>(k - d.fNetStart[j]);
                ++cursor[v];
            }
        }
    }
    d.flat = true;
    return 0;
}

}  // namespace ktgi

extern "C" {

int ktg_default_spec(ktg_spec *spec, uint32_t num_cells, uint32_t num_nets, uint32_t num_pads,
                     uint64_t seed) {
    if (spec == nullptr) return fail("ktg_default_spec: null spec");
    spec->seed = seed;
    spec->num_cells = num_cells;
    spec->num_nets = num_nets;
    spec->num_pads = num_pads;
    spec->num_macros = 0;
    spec->utilization = 0.70;
    spec->max_degree = 40;
    spec->high_fanout_frac = 0.0;
    spec->local_frac = 0.9;
    return 0;
}

int ktg_generate(const ktg_spec *spec, ktg_design **out) {
    if (spec == nullptr || out == nullptr) return fail("ktg_generate: null argument");
    try {
        auto *d = new ktg_design();
        int rc = generate(*spec, *d);
        if (rc != 0) {
            delete d;
            return rc;
        }
        ktg_raw raw;
        ktg_raw_view(d, &raw);
        auto *flat = new ktg_design();
        rc = ktgi::collect(raw, *flat);
        delete d;
        if (rc != 0) {
            delete flat;
            return rc;
        }
        *out = flat;
        return 0;
    } catch (const std::bad_alloc &) {
        return fail("ktg_generate: out of memory");
    } catch (const std::exception &e) {
        return fail(std::string("ktg_generate: ") + e.what());
    }
}

int ktg_raw_view(const ktg_design *d, ktg_raw *v) {
    if (d == nullptr || v == nullptr) return fail("ktg_raw_view: null argument");
    v->num_vertices = static_cast<uint32_t>(d->isFixed.size());
    v->num_movable = d->numMovable;
    v->num_nets = static_cast<uint32_t>(d->netWeight.size());
    v->num_pins = d->pinVertex.size();
    v->num_rows = static_cast<uint32_t>(d->rowY.size());
    v->is_fixed = d->isFixed.data();
    v->pos_x = d->posX.data();
    v->pos_y = d->posY.data();
    v->width = d->width.data();
    v->height = d->height.data();
    v->net_start = d->netStart.data();
    v->pin_vertex = d->pinVertex.data();
    v->pin_off_x = d->pinOffX.data();
    v->pin_off_y = d->pinOffY.data();
    v->net_weight = d->netWeight.data();
    v->row_y = d->rowY.data();
    v->row_height = d->rowH.data();
    v->row_x0 = d->rowX0.data();
    v->row_x1 = d->rowX1.data();
    v->row_pitch = d->rowPitch.data();
    std::memcpy(v->die, d->die, sizeof(v->die));
    return 0;
}

int ktg_collect(const ktg_raw *raw, ktg_design **out) {
    if (raw == nullptr || out == nullptr) return fail("ktg_collect: null argument");
    try {
        auto *d = new ktg_design();
        const int rc = ktgi::collect(*raw, *d);
        if (rc != 0) {
            delete d;
            return rc;
        }
        *out = d;
        return 0;
    } catch (const std::bad_alloc &) {
        return fail("ktg_collect: out of memory");
    }
}

int ktg_flat_view(const ktg_design *d, ktg_flat *v) {
    if (d == nullptr || v == nullptr) return fail("ktg_flat_view: null argument");
    if (!d->flat) return fail("ktg_flat_view: design not flattened");
    v->num_vertices = static_cast<uint32_t>(d->varOfVertex.size());
    v->num_movable = d->numMovable;
    v->num_nets = static_cast<uint32_t>(d->fWeight.size());
    v->num_pins = d->fPinCell.size();
    v->num_cell_pins = d->cellPinNet.size();
    v->num_obstacles = static_cast<uint32_t>(d->obstacles.size() / 4);
    v->net_start = d->fNetStart.data();
    v->pin_cell = d->fPinCell.data();
    v->pin_off_x = d->fOffX.data();
    v->pin_off_y = d->fOffY.data();
    v->net_weight = d->fWeight.data();
    v->var_of_vertex = d->varOfVertex.data();
    v->mov_vertex = d->movVertex.data();
    v->fixed_x = d->fixedX.data();
    v->fixed_y = d->fixedY.data();
    v->pin_start = d->pinStart.data();
    v->cell_pin_net = d->cellPinNet.data();
    v->cell_pin_idx = d->cellPinIdx.data();
    v->width = d->mWidth.data();
    v->height = d->mHeight.data();
    v->area = d->mArea.data();
    v->obstacles = d->obstacles.data();
    return 0;
}

int ktg_seed_uniform(uint64_t seed, uint64_t n, const double die[4], const double *width, const double *height,
                     double *x, double *y) {
    if (die == nullptr || (n > 0 && (width == nullptr || height == nullptr || x == nullptr || y == nullptr)))
        return fail("ktg_seed_uniform: null argument");
    std::mt19937_64 rng(seed);
    const auto unit = [&rng] { return static_cast<double>(rng() >> 11) * 0x1.0p-53; };
    const double W = die[2] - die[0], H = die[3] - die[1];
    for (uint64_t i = 0; i < n; ++i) {
        x[i] = die[0] + unit() * std::max(W - width[i], 0.0);
        y[i] = die[1] + unit() * std::max(H - height[i], 0.0);
    }
    return 0;
}

void ktg_free(ktg_design *d) { delete d; }

const char *ktg_last_error(void) { return ktgi::lastErrorText(); }

}  // extern "C"
