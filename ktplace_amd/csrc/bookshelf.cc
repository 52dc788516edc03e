// Bookshelf designs in and placements out (host C++), feeding the same
// flattening as the synthetic designs.
//
// Reading follows BookshelfInputAdapter (src/adaptor/bookshelfToKTAdaptor.cc)
// and the die rule of getPlacementDieBox (src/datamodel/kt_die.cc:45-103):
//   * files: <dir>/<name>.{nodes,nets,pl,scl,wts}, <name> = the directory's
//     own name, each read gzip-compressed when <file>.gz exists (:140-165);
//   * .nodes: body after the "NumTerminals" line; "name w h [terminal]";
//     terminals are fixed (:227-291);
//   * .wts (before .nets): "name weight" lines give net weights (:642-670);
//   * .nets: lines before the first "NetDegree" are ignored; "NetDegree
//     [:] d [name]" opens a net, at most d of the following lines are its
//     pins "cell dir [: dx dy]", (dx, dy) from the cell centre; a nameless net
//     is called "n<index>"; a net whose name repeats adds its pins to the first
//     net of that name, whose weight stays; pins of unknown cells are dropped
//     (:293-436);
//   * .pl: "name x y [: orient] [/FIXED]" gives the lower-left corner, and
//     /FIXED fixes the cell (:438-495);
//   * .scl: CoreRow ... End blocks with Coordinate, Height, Sitewidth,
//     Sitespacing (1 by default) and any number of "SubrowOrigin : x
//     NumSites : k" subrows, keys matched without case; a row without a
//     subrow of positive size is dropped (:497-640);
//   * die: the box of the fixed cells (each at least 1 x 1) extended by the
//     rows' box, or the rows' box alone without fixed cells.
// A pin's offset is kept from the cell's lower-left, (dx + w/2, dy + h/2),
// as the reference's pins are (bookshelfToKTAdaptor.cc:424-428).
//
// Writing follows FlowMgr::Impl::writeBookshelf (src/kt_flowMgr.cc:508-521):
// one "name<TAB>x<TAB>y<TAB>: N" line per vertex, " /FIXED" after the N of a
// fixed one, numbers with 10 significant digits.
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <new>
#include <string>
#include <sys/stat.h>
#include <unordered_map>
#include <utility>
#include <vector>

#include "design_internal.h"

namespace {

using ktgi::fail;

bool exists(const std::string &path) {
    struct stat st;
    return ::stat(path.c_str(), &st) == 0;
}

// The file at path, or path + ".gz" when that exists, as lines.  zlib reads
// plain files through unchanged.
bool readLines(std::string path, std::vector<std::string> &lines) {
    if (exists(path + ".gz")) path += ".gz";
    gzFile f = gzopen(path.c_str(), "rb");
    if (f == nullptr) return false;
    std::string text;
    std::vector<char> buf(1 << 20);
    for (;;) {
        const int got = gzread(f, buf.data(), static_cast<unsigned>(buf.size()));
        if (got < 0) {
            gzclose(f);
            return false;
        }
        if (got == 0) break;
        text.append(buf.data(), static_cast<size_t>(got));
    }
    gzclose(f);
    lines.clear();
    size_t a = 0;
    while (a < text.size()) {
        size_t b = text.find('\n', a);
        if (b == std::string::npos) b = text.size();
        lines.emplace_back(text, a, b - a);
        a = b + 1;
    }
    return true;
}

// Drops a '#' comment and the surrounding white space.
std::string strip(const std::string &line) {
    std::string s = line.substr(0, line.find('#'));
    const size_t a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return std::string();
    const size_t b = s.find_last_not_of(" \t\r\n");
    return s.substr(a, b - a + 1);
}

std::vector<std::string> tokens(const std::string &line) {
    std::vector<std::string> out;
    size_t i = 0;
    while (i < line.size()) {
        while (i < line.size() && std::isspace(static_cast<unsigned char>(line[i]))) ++i;
        size_t j = i;
        while (j < line.size() && !std::isspace(static_cast<unsigned char>(line[j]))) ++j;
        if (j > i) out.emplace_back(line, i, j - i);
        i = j;
    }
    return out;
}

// A leading number of s, as std::stod reads it; false when s has none.
bool number(const std::string &s, double &v) {
    const char *p = s.c_str();
    char *end = nullptr;
    errno = 0;
    v = std::strtod(p, &end);
    return end != p && errno != ERANGE;
}

bool prefixed(const std::string &s, const char *p) { return s.compare(0, std::strlen(p), p) == 0; }

std::string lower(std::string s) {
    for (char &c : s) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
    return s;
}

struct Pin {
    uint32_t cell;
    double dx, dy;
};

struct Subrow {
    double origin, sites;
};

struct Row {
    double coordinate = 0.0, height = 0.0, sitewidth = 1.0, sitespacing = 1.0;
    std::vector<Subrow> subrows;
    double pitch() const { return sitespacing > 0.0 ? sitespacing : sitewidth; }
};

struct Reader {
    std::vector<std::string> names;
    std::unordered_map<std::string, uint32_t> cellOf;
    std::vector<double> w, h, x, y;
    std::vector<uint8_t> fixed;
    std::unordered_map<std::string, double> netWeight;
    std::vector<std::vector<Pin>> nets;
    std::vector<double> weight;
    std::unordered_map<std::string, size_t> netOf;
    std::vector<Row> rows;

    void nodes(const std::vector<std::string> &lines) {
        size_t body = 0;
        while (body < lines.size()) {
            const bool last = strip(lines[body]).find("NumTerminals") != std::string::npos;
            ++body;
            if (last) break;
        }
        for (size_t i = body; i < lines.size(); ++i) {
            const std::vector<std::string> t = tokens(strip(lines[i]));
            double cw, ch;
            if (t.size() < 3 || !number(t[1], cw) || !number(t[2], ch)) continue;
            if (cellOf.count(t[0])) continue;  // a repeated name names the first cell
            cellOf.emplace(t[0], static_cast<uint32_t>(names.size()));
            names.push_back(t[0]);
            w.push_back(cw);
            h.push_back(ch);
            x.push_back(0.0);
            y.push_back(0.0);
            fixed.push_back(t.size() >= 4 && t[3] == "terminal" ? 1 : 0);
        }
    }

    void weights(const std::vector<std::string> &lines) {
        bool header = true;
        for (const std::string &line : lines) {
            const std::string s = strip(line);
            if (s.empty()) continue;
            if (header) {
                if (s.find("UCLA") != std::string::npos) continue;
                header = false;
            }
            const std::vector<std::string> t = tokens(s);
            double v;
            if (t.size() >= 2 && number(t[1], v)) netWeight[t[0]] = v;
        }
    }

    void netlist(const std::vector<std::string> &lines) {
        struct Open {
            std::string name;
            size_t degree = 0;
            std::vector<size_t> pinLines;
        };
        std::vector<Open> open;
        for (size_t i = 0; i < lines.size(); ++i) {
            const std::string s = strip(lines[i]);
            if (s.empty()) continue;
            if (prefixed(s, "NetDegree")) {
                const std::vector<std::string> t = tokens(s);
                size_t k = 1;
                while (k < t.size() && t[k] == ":") ++k;
                double d;
                // a degree that is not a finite non-negative number is
                // malformed; one past the file's line count cannot be filled
                // and is capped there (no out-of-range conversion)
                if (k >= t.size() || !number(t[k], d) || !std::isfinite(d) || d < 0.0) continue;
                Open o;
                o.degree = d >= static_cast<double>(lines.size()) ? lines.size() : static_cast<size_t>(d);
                o.name = k + 1 < t.size() ? t[k + 1] : std::string();
                open.push_back(std::move(o));
                continue;
            }
            if (!open.empty()) open.back().pinLines.push_back(i);
        }
        for (size_t i = 0; i < open.size(); ++i) {
            const Open &o = open[i];
            const std::string name = o.name.empty() ? "n" + std::to_string(i) : o.name;
            auto it = netOf.find(name);
            size_t net;
            if (it == netOf.end()) {
                net = nets.size();
                netOf.emplace(name, net);
                nets.emplace_back();
                const auto wt = netWeight.find(o.name);
                weight.push_back(wt != netWeight.end() ? wt->second : 1.0);
            } else {
                net = it->second;
            }
            const size_t count = std::min(o.pinLines.size(), o.degree);
            for (size_t p = 0; p < count; ++p) {
                const std::vector<std::string> t = tokens(strip(lines[o.pinLines[p]]));
                if (t.size() < 2) continue;
                double dx = 0.0, dy = 0.0;
                if (t.size() >= 5 && t[2] == ":" && !(number(t[3], dx) && number(t[4], dy))) continue;
                const auto c = cellOf.find(t[0]);
                if (c == cellOf.end()) continue;
                nets[net].push_back({c->second, dx, dy});
            }
        }
    }

    void placement(const std::vector<std::string> &lines) {
        bool header = true;
        for (const std::string &line : lines) {
            const std::string s = strip(line);
            if (s.empty()) continue;
            if (header) {
                if (s.find("UCLA") != std::string::npos) continue;
                header = false;
            }
            const std::vector<std::string> t = tokens(s);
            double px, py;
            if (t.size() < 3 || !number(t[1], px) || !number(t[2], py)) continue;
            const auto c = cellOf.find(t[0]);
            if (c == cellOf.end()) continue;
            x[c->second] = px;
            y[c->second] = py;
            if (std::find(t.begin(), t.end(), "/FIXED") != t.end()) fixed[c->second] = 1;
        }
    }

    void rowsOf(const std::vector<std::string> &lines) {
        bool in = false;
        std::vector<std::string> block;
        for (const std::string &line : lines) {
            const std::string s = strip(line);
            if (s.empty()) continue;
            if (s.find("CoreRow") != std::string::npos) {
                in = true;
                block.clear();
                continue;
            }
            if (s == "End") {
                if (in) row(block);
                in = false;
                continue;
            }
            if (in) block.push_back(s);
        }
    }

    void row(const std::vector<std::string> &block) {
        Row r;
        double origin = 0.0, sites = 0.0;
        bool haveOrigin = false;
        for (const std::string &line : block) {
            const std::vector<std::string> t = tokens(line);
            if (t.size() < 2) continue;
            const std::string key = lower(t[0]);
            size_t k = 1;
            while (k < t.size() && t[k] == ":") ++k;
            double v;
            if (k >= t.size() || !number(t[k], v)) continue;
            if (key == "coordinate") {
                r.coordinate = v;
            } else if (key == "height") {
                r.height = v;
            } else if (key == "sitewidth") {
                r.sitewidth = v;
            } else if (key == "sitespacing") {
                r.sitespacing = v;
            } else if (key == "subroworigin") {
                if (haveOrigin && sites > 0.0) r.subrows.push_back({origin, sites});
                origin = v;
                haveOrigin = true;
                sites = 0.0;
                for (size_t p = 0; p + 1 < t.size(); ++p) {
                    if (lower(t[p]) != "numsites") continue;
                    size_t m = p + 1;
                    while (m < t.size() && t[m] == ":") ++m;
                    double n;
                    if (m < t.size() && number(t[m], n)) sites = n;
                    break;
                }
            }
        }
        if (haveOrigin && sites > 0.0) r.subrows.push_back({origin, sites});
        if (!r.subrows.empty()) rows.push_back(std::move(r));
    }

    // getPlacementDieBox without a declared die area (Bookshelf has none).
    bool die(double box[4]) const {
        const double inf = std::numeric_limits<double>::max();
        double lx = inf, ly = inf, hx = -inf, hy = -inf;
        for (size_t v = 0; v < names.size(); ++v) {
            if (!fixed[v]) continue;
            lx = std::min(lx, x[v]);
            ly = std::min(ly, y[v]);
            hx = std::max(hx, x[v] + std::max(w[v], 1.0));
            hy = std::max(hy, y[v] + std::max(h[v], 1.0));
        }
        const bool haveFixed = hx > lx && hy > ly;
        double rlx = inf, rly = inf, rhx = -inf, rhy = -inf;
        bool anyRow = false;
        for (const Row &r : rows) {
            if (!(r.pitch() > 0.0)) continue;
            double hi = 0.0;
            for (const Subrow &s : r.subrows) hi = std::max(hi, s.origin + s.sites * r.pitch());
            rlx = std::min(rlx, r.subrows.front().origin);
            rly = std::min(rly, r.coordinate);
            rhx = std::max(rhx, hi);
            rhy = std::max(rhy, r.coordinate + r.height);
            anyRow = true;
        }
        if (anyRow && rhx > rlx && rhy > rly) {
            if (!haveFixed) {
                lx = rlx, ly = rly, hx = rhx, hy = rhy;
            } else {
                lx = std::min(lx, rlx), ly = std::min(ly, rly);
                hx = std::max(hx, rhx), hy = std::max(hy, rhy);
            }
        }
        box[0] = lx, box[1] = ly, box[2] = hx, box[3] = hy;
        return hx > lx && hy > ly;
    }
};

std::string baseName(std::string dir) {
    while (dir.size() > 1 && dir.back() == '/') dir.pop_back();
    const size_t slash = dir.find_last_of('/');
    return slash == std::string::npos ? dir : dir.substr(slash + 1);
}

int readDesign(const std::string &dir, ktg_design *&out) {
    std::string base = dir;
    while (base.size() > 1 && base.back() == '/') base.pop_back();
    const std::string stem = base + "/" + baseName(base);
    Reader rd;
    std::vector<std::string> lines;
    if (!readLines(stem + ".nodes", lines)) return fail("ktg_read_bookshelf: cannot read " + stem + ".nodes");
    rd.nodes(lines);
    if (readLines(stem + ".wts", lines)) rd.weights(lines);
    if (!readLines(stem + ".nets", lines)) return fail("ktg_read_bookshelf: cannot read " + stem + ".nets");
    rd.netlist(lines);
    if (readLines(stem + ".pl", lines)) rd.placement(lines);
    if (readLines(stem + ".scl", lines)) rd.rowsOf(lines);
    double box[4];
    if (!rd.die(box)) return fail("ktg_read_bookshelf: no fixed cell and no row bound the die of " + stem);
    const size_t V = rd.names.size();
    if (V >= KTG_NO_VAR) return fail("ktg_read_bookshelf: too many cells");

    // raw arrays: pins at their lower-left offsets, one row entry per subrow
    std::vector<uint64_t> netStart(1, 0);
    std::vector<uint32_t> pinVertex;
    std::vector<double> offX, offY;
    for (const std::vector<Pin> &net : rd.nets) {
        for (const Pin &p : net) {
            pinVertex.push_back(p.cell);
            offX.push_back(p.dx + 0.5 * rd.w[p.cell]);
            offY.push_back(p.dy + 0.5 * rd.h[p.cell]);
        }
        netStart.push_back(pinVertex.size());
    }
    std::vector<double> rowY, rowH, rowX0, rowX1, rowPitch;
    for (const Row &r : rd.rows) {
        for (const Subrow &s : r.subrows) {
            rowY.push_back(r.coordinate);
            rowH.push_back(r.height);
            rowX0.push_back(s.origin);
            rowX1.push_back(s.origin + s.sites * r.pitch());
            rowPitch.push_back(r.pitch());
        }
    }
    uint32_t movable = 0;
    for (const uint8_t f : rd.fixed) movable += f ? 0u : 1u;
    ktg_raw raw{};
    raw.num_vertices = static_cast<uint32_t>(V);
    raw.num_movable = movable;
    raw.num_nets = static_cast<uint32_t>(rd.nets.size());
    raw.num_pins = pinVertex.size();
    raw.num_rows = static_cast<uint32_t>(rowY.size());
    raw.is_fixed = rd.fixed.data();
    raw.pos_x = rd.x.data();
    raw.pos_y = rd.y.data();
    raw.width = rd.w.data();
    raw.height = rd.h.data();
    raw.net_start = netStart.data();
    raw.pin_vertex = pinVertex.data();
    raw.pin_off_x = offX.data();
    raw.pin_off_y = offY.data();
    raw.net_weight = rd.weight.data();
    raw.row_y = rowY.data();
    raw.row_height = rowH.data();
    raw.row_x0 = rowX0.data();
    raw.row_x1 = rowX1.data();
    raw.row_pitch = rowPitch.data();
    for (int i = 0; i < 4; ++i) raw.die[i] = box[i];
    auto *d = new ktg_design();
    const int rc = ktgi::collect(raw, *d);
    if (rc != 0) {
        delete d;
        return rc;
    }
    d->names = std::move(rd.names);
    out = d;
    return 0;
}

}  // namespace

extern "C" {

int ktg_read_bookshelf(const char *dir, ktg_design **out) {
    if (dir == nullptr || out == nullptr) return fail("ktg_read_bookshelf: null argument");
    try {
        ktg_design *d = nullptr;
        const int rc = readDesign(dir, d);
        if (rc != 0) return rc;
        *out = d;
        return 0;
    } catch (const std::bad_alloc &) {
        return fail("ktg_read_bookshelf: out of memory");
    }
}

int ktg_vertex_names(ktg_design *d, const char **blob, uint64_t *bytes) {
    if (d == nullptr || blob == nullptr || bytes == nullptr) return fail("ktg_vertex_names: null argument");
    if (d->names.size() != d->isFixed.size()) return fail("ktg_vertex_names: the design has no vertex names");
    try {
        if (d->nameBlob.empty() && !d->names.empty()) {
            for (const std::string &n : d->names) {
                d->nameBlob += n;
                d->nameBlob += '\n';
            }
        }
        *blob = d->nameBlob.data();
        *bytes = d->nameBlob.size();
        return 0;
    } catch (const std::bad_alloc &) {
        return fail("ktg_vertex_names: out of memory");
    }
}

int ktg_write_pl(const char *path, uint32_t num_vertices, const char *names, uint64_t names_bytes,
                 const uint32_t *var_of_vertex, const double *fixed_x, const double *fixed_y, const double *x,
                 const double *y) {
    if (path == nullptr || (num_vertices > 0 && (names == nullptr || var_of_vertex == nullptr ||
                                                 fixed_x == nullptr || fixed_y == nullptr)))
        return fail("ktg_write_pl: null argument");
    try {
        std::vector<std::string> name;
        size_t a = 0;
        while (a < names_bytes && name.size() < num_vertices) {
            const void *nl = std::memchr(names + a, '\n', names_bytes - a);
            const size_t b = nl ? static_cast<size_t>(static_cast<const char *>(nl) - names) : names_bytes;
            name.emplace_back(names + a, b - a);
            a = b + 1;
        }
        if (name.size() != num_vertices) return fail("ktg_write_pl: fewer names than vertices");
        for (uint32_t v = 0; v < num_vertices; ++v)
            if (var_of_vertex[v] != KTG_NO_VAR && (x == nullptr || y == nullptr))
                return fail("ktg_write_pl: null positions");
        FILE *f = std::fopen(path, "w");
        if (f == nullptr) return fail(std::string("ktg_write_pl: cannot open ") + path);
        bool ok = true;
        for (uint32_t v = 0; v < num_vertices && ok; ++v) {
            const uint32_t var = var_of_vertex[v];
            const bool movable = var != KTG_NO_VAR;
            const double px = movable ? x[var] : fixed_x[v];
            const double py = movable ? y[var] : fixed_y[v];
            ok = std::fprintf(f, "%s\t%.10g\t%.10g\t: %s\n", name[v].c_str(), px, py, movable ? "N" : "N /FIXED") > 0;
        }
        ok = (std::fclose(f) == 0) && ok;
        return ok ? 0 : fail(std::string("ktg_write_pl: cannot write ") + path);
    } catch (const std::bad_alloc &) {
        return fail("ktg_write_pl: out of memory");
    }
}

}  // extern "C"
