// Drives the C++ facade (include/ktplace_gpu.hpp) the way simplPlacer::Impl
// would: a seeded design from the library's generator, flattened into a
// FlatNetlist, then buildB2B, solve, the grid, legalize and the bookkeeping.
// Prints one line per result for tests/test_shim.py to compare with the same
// steps taken through the Python binding.
//
//   shim_check <cells> <nets> <pads> <seed>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "ktplace_gen.h"
#include "ktplace_gpu.hpp"

template <typename T>
static std::vector<T> take(const T *p, std::size_t n) {
    return std::vector<T>(p, p + n);
}

static uint64_t bits(double v) {
    uint64_t b;
    std::memcpy(&b, &v, 8);
    return b;
}

static uint64_t hashSystem(const ktamd::SparseSystem &s) {  // FNV-1a over every word
    uint64_t h = 1469598103934665603ull;
    const auto mix = [&h](uint64_t v) { h = (h ^ v) * 1099511628211ull; };
    for (const std::size_t v : s.rowPtr) mix(v);
    for (const uint32_t v : s.col) mix(v);
    for (const double v : s.val) mix(bits(v));
    for (const double v : s.diag) mix(bits(v));
    for (const double v : s.rhs) mix(bits(v));
    return h;
}

int main(int argc, char **argv) {
    if (argc != 5) {
        std::fprintf(stderr, "usage: shim_check cells nets pads seed\n");
        return 2;
    }
    ktg_spec spec{};
    if (ktg_default_spec(&spec, std::atoi(argv[1]), std::atoi(argv[2]), std::atoi(argv[3]),
                         std::strtoull(argv[4], nullptr, 10)) != 0)
        return 2;
    ktg_design *d = nullptr;
    if (ktg_generate(&spec, &d) != 0) return 2;
    ktg_raw raw{};
    ktg_flat f{};
    ktg_raw_view(d, &raw);
    ktg_flat_view(d, &f);
    ktamd::FlatNetlist nl;
    nl.netStart = take(f.net_start, f.num_nets + 1ull);
    nl.pinCell = take(f.pin_cell, f.num_pins);
    nl.offX = take(f.pin_off_x, f.num_pins);
    nl.offY = take(f.pin_off_y, f.num_pins);
    nl.weight = take(f.net_weight, f.num_nets);
    nl.varOfVertex = take(f.var_of_vertex, f.num_vertices);
    nl.movVertex = take(f.mov_vertex, f.num_movable);
    nl.fixedX = take(f.fixed_x, f.num_vertices);
    nl.fixedY = take(f.fixed_y, f.num_vertices);
    nl.pinStart = take(f.pin_start, f.num_movable + 1ull);
    nl.refNet = take(f.cell_pin_net, f.num_cell_pins);
    nl.refPin = take(f.cell_pin_idx, f.num_cell_pins);
    nl.width = take(f.width, f.num_movable);
    nl.height = take(f.height, f.num_movable);
    nl.area = take(f.area, f.num_movable);
    nl.obstacles = take(f.obstacles, 4ull * f.num_obstacles);
    const std::size_t n = nl.movVertex.size();
    double floorX = 0.0, floorY = 0.0;
    for (std::size_t i = 0; i < n; ++i) {
        floorX += nl.width[i];
        floorY += nl.height[i];
    }
    floorX /= static_cast<double>(n);
    floorY /= static_cast<double>(n);
    std::vector<double> x(n), y(n);
    ktg_seed_uniform(20240607, n, raw.die, nl.width.data(), nl.height.data(), x.data(), y.data());
    try {
        ktamd::GpuSystems gpu(nl);
        const std::vector<double> none;
        gpu.buildB2B(x, y, 0.0, none, none, floorX, floorY);
        std::printf("b2b_x %016" PRIx64 "\nb2b_y %016" PRIx64 "\n", hashSystem(gpu.download(0)),
                    hashSystem(gpu.download(1)));
        std::size_t observed = 0;
        const ktamd::Residuals r = gpu.solve(x, y, 1e-6, 100000, [&](std::size_t, const ktamd::Residuals &, bool) {
            ++observed;
        });
        std::printf("iterations %zu observed %zu resid %016" PRIx64 " %016" PRIx64 "\n", gpu.lastIterations(),
                    observed, bits(r.x), bits(r.y));
        gpu.setRows(take(raw.row_y, raw.num_rows), take(raw.row_height, raw.num_rows), take(raw.row_x0, raw.num_rows),
                    take(raw.row_x1, raw.num_rows), take(raw.row_pitch, raw.num_rows), raw.die);
        std::printf("bins %zu\n", gpu.buildGrid(0, 1.0));
        std::printf("hpwl_lower %016" PRIx64 "\n", bits(gpu.hpwl(x, y)));
        std::vector<double> ux = x, uy = y;
        gpu.legalize(ux, uy);
        std::printf("hpwl_upper %016" PRIx64 "\noverflow_upper %016" PRIx64 "\n", bits(gpu.hpwl(ux, uy)),
                    bits(gpu.overflow(ux, uy)));
        try {  // an error from the C ABI arrives as an exception
            std::vector<double> shortX(n / 2), shortY(n / 2);
            gpu.solve(shortX, shortY, 1e-6, 10);
            std::printf("error none\n");
        } catch (const ktamd::Error &e) {
            std::printf("error raised\n");
        }
    } catch (const ktamd::Error &e) {
        std::fprintf(stderr, "%s\n", e.what());
        ktg_free(d);
        return 1;
    }
    ktg_free(d);
    return 0;
}
