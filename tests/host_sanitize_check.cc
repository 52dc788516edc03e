// Host C++ of the library (the generator, the flattening, the Bookshelf reader
// and writer, and the legalizer's host geometry) exercised under
// AddressSanitizer and UndefinedBehaviorSanitizer by tests/test_host_sanitizers.py.
//
//   host_sanitize_check gen                 several small seeded designs
//   host_sanitize_check read <dir> <out.pl> read a Bookshelf directory, write a .pl
//   host_sanitize_check geom                grid, profiles, stripes, clusters, growth
// Exit status 0 on success (a failed read also exits 0 and prints "error"),
// non-zero, with the sanitizer's report, on a fault.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "ktplace_gen.h"
#include "lal_geom.h"

static int gen() {
    for (uint64_t seed = 1; seed <= 6; ++seed) {
        ktg_spec spec{};
        if (ktg_default_spec(&spec, 500 + 300 * seed, 550 + 330 * seed, 20 + seed, seed) != 0) return 1;
        spec.num_macros = seed % 3 == 0 ? 0 : static_cast<uint32_t>(seed);
        spec.max_degree = seed % 2 ? 40 : 120;
        spec.high_fanout_frac = seed % 2 ? 0.0 : 0.01;
        ktg_design *d = nullptr;
        if (ktg_generate(&spec, &d) != 0) return 1;
        ktg_raw raw{};
        ktg_raw_view(d, &raw);
        ktg_design *c = nullptr;
        if (ktg_collect(&raw, &c) != 0) return 1;
        ktg_flat f{};
        ktg_flat_view(c, &f);
        std::vector<double> x(f.num_movable), y(f.num_movable);
        ktg_seed_uniform(seed, f.num_movable, raw.die, f.width, f.height, x.data(), y.data());
        std::printf("gen seed %llu: %u cells, %llu pins\n", static_cast<unsigned long long>(seed), f.num_movable,
                    static_cast<unsigned long long>(f.num_pins));
        ktg_free(c);
        ktg_free(d);
    }
    return 0;
}

static int read(const char *dir, const char *out) {
    ktg_design *d = nullptr;
    if (ktg_read_bookshelf(dir, &d) != 0) {
        std::printf("error %s\n", ktg_last_error());
        return 0;
    }
    ktg_flat f{};
    ktg_flat_view(d, &f);
    const char *names = nullptr;
    uint64_t bytes = 0;
    ktg_vertex_names(d, &names, &bytes);
    std::vector<double> x(f.num_movable, 1.5), y(f.num_movable, 2.5);
    const int rc = ktg_write_pl(out, f.num_vertices, names, bytes, f.var_of_vertex, f.fixed_x, f.fixed_y, x.data(),
                                y.data());
    std::printf("read %u vertices, %u nets, write %d\n", f.num_vertices, f.num_nets, rc);
    ktg_free(d);
    return 0;
}

static int geom() {
    using namespace kt::lal;
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    for (int trial = 0; trial < 20; ++trial) {
        const double W = 40 + 200 * u(rng), H = 20 + 100 * u(rng);
        std::vector<RowSpec> rows;
        for (int r = 0; r < static_cast<int>(H); ++r) rows.push_back({static_cast<double>(r), 1.0, 0.0, W, 1.0});
        std::vector<Box> obstacles;
        for (int m = 0; m < trial % 7; ++m) {
            const double x0 = W * u(rng), y0 = H * u(rng);
            obstacles.push_back({x0, y0, x0 + 1 + 20 * u(rng), y0 + 1 + 10 * u(rng)});
        }
        Grid g;
        FreeSpace fs;
        buildGrid(g, fs, rows, obstacles, Box{0.0, 0.0, W, H}, 4 + trial % 29);
        AreaTables t;
        t.buildAvail(g);
        std::vector<double> occ(g.size());
        std::vector<uint64_t> occFix(g.size());
        for (size_t k = 0; k < g.size(); ++k) {
            occ[k] = std::floor(g.avail[k] * 1.6 * u(rng) * 65536.0) / 65536.0;
            occFix[k] = static_cast<uint64_t>(occ[k] * 65536.0);
        }
        t.setOcc(occFix);
        for (const BinRect &c : overfilledClusters(g, occ, 1.0)) (void)grow(g, t, occ, c, 1.0);
        for (int q = 0; q < 50; ++q) {
            double x0 = W * u(rng), x1 = W * u(rng), y0 = H * u(rng), y1 = H * u(rng);
            if (x0 > x1) std::swap(x0, x1);
            if (y0 > y1) std::swap(y0, y1);
            const Box b{x0, y0, x1, y1};
            const bool vertical = q % 2;
            const Profile p = fs.profile(b, vertical);
            (void)p.coordinateAt(0.5 * p.total());
            const double lo = vertical ? x0 : y0, hi = vertical ? x1 : y1;
            (void)stripes(fs, obstacles, b, vertical, 0.7 * p.total(), 2.0 * p.total(), 1.0, 0.1, lo + 0.5 * (hi - lo),
                          q % 4 < 2);
        }
    }
    std::printf("geom ok\n");
    return 0;
}

int main(int argc, char **argv) {
    if (argc >= 2 && std::strcmp(argv[1], "gen") == 0) return gen();
    if (argc >= 4 && std::strcmp(argv[1], "read") == 0) return read(argv[2], argv[3]);
    if (argc >= 2 && std::strcmp(argv[1], "geom") == 0) return geom();
    std::fprintf(stderr, "usage: host_sanitize_check gen | read <dir> <out.pl> | geom\n");
    return 2;
}
