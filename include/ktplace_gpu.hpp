/* C++ facade over the C ABI (ktplace_amd.h) with the shape of KTPlace's SimPL
 * solver interface: LinearSystems::buildB2B / buildStar (simpl_lss.h:68-84),
 * solveBothAxes (simpl_lss.h:86-88), LookAheadLegalizer::legalize
 * (simpl_lal.h:18-29), buildGrid and the HPWL / overflow bookkeeping
 * (simpl_placer.cc:181-297).  Header only; link with -lktplace_amd.
 *
 * The facade takes the solver netlist as flat arrays (FlatNetlist below, the
 * layout of kt_netlist) and std::vector<double> positions, so a caller adapts
 * its own netlist type once (INTEGRATION.md shows the adapter for
 * simpl::Netlist).  Errors become ktamd::Error exceptions carrying
 * kt_last_error(); the C ABI itself never throws. */
#pragma once

#include <cstddef>
#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "ktplace_amd.h"

namespace ktamd {

struct Error : std::runtime_error {
    explicit Error(const std::string &what) : std::runtime_error(what) {}
};

inline void check(int rc) {
    if (rc != KT_OK) throw Error(std::string("ktplace_amd: ") + kt_last_error());
}

/* The solver netlist as flat arrays (simpl_lss.h:27-36 flattened; see
 * kt_netlist for each field).  The arrays must outlive the upload call only. */
struct FlatNetlist {
    std::vector<uint64_t> netStart;   // [N + 1]
    std::vector<uint32_t> pinCell;    // [P] vertex of each pin
    std::vector<double> offX, offY;   // [P] offsets from the vertex's lower-left
    std::vector<double> weight;       // [N]
    std::vector<uint32_t> varOfVertex;  // [V] KT_NO_VAR for fixed vertices
    std::vector<uint32_t> movVertex;    // [n]
    std::vector<double> fixedX, fixedY; // [V]
    std::vector<uint64_t> pinStart;   // [n + 1]
    std::vector<uint32_t> refNet, refPin;  // [Pm] pinsOfCell
    std::vector<double> width, height, area;  // [n] movable cell sizes
    std::vector<double> obstacles;    // [4 M] xl, yl, xh, yh of fixed macros
};

/* One axis of SparseSystem (simpl_lss.h:38-58): off-diagonal CSR + diag + rhs. */
struct SparseSystem {
    std::vector<std::size_t> rowPtr;
    std::vector<uint32_t> col;
    std::vector<double> val, diag, rhs;
};

struct Residuals {
    double x = 0.0, y = 0.0;
};

/* (iteration, residuals, last) after each lock-step iteration, as the
 * reference's IterateObserver (called after the solve, in order). */
using IterateObserver = std::function<void(std::size_t, const Residuals &, bool)>;

struct LalOptions {  // simpl_options.h:24-26
    std::size_t maxLevel = 20, minCellsToSplit = 4;
    double stripeAreaFraction = 0.1;
};

/* One design on one GPU: the netlist is uploaded once; the systems, CG
 * state, grid and positions stay on the device across outer iterations, as
 * sysX / sysY do in simplPlacer::Impl. */
class GpuSystems {
public:
    explicit GpuSystems(const FlatNetlist &nl, int device = 0) {
        check(kt_ctx_create(&ctx_, device));
        n_ = nl.movVertex.size();
        kt_netlist k{};
        k.num_vertices = static_cast<uint32_t>(nl.varOfVertex.size());
        k.num_movable = static_cast<uint32_t>(nl.movVertex.size());
        k.num_nets = static_cast<uint32_t>(nl.weight.size());
        k.num_pins = nl.pinCell.size();
        k.num_cell_pins = nl.refNet.size();
        k.num_obstacles = static_cast<uint32_t>(nl.obstacles.size() / 4);
        k.net_start = nl.netStart.data();
        k.pin_cell = nl.pinCell.data();
        k.pin_off_x = nl.offX.data();
        k.pin_off_y = nl.offY.data();
        k.net_weight = nl.weight.data();
        k.var_of_vertex = nl.varOfVertex.data();
        k.mov_vertex = nl.movVertex.data();
        k.fixed_x = nl.fixedX.data();
        k.fixed_y = nl.fixedY.data();
        k.pin_start = nl.pinStart.data();
        k.cell_pin_net = nl.refNet.data();
        k.cell_pin_idx = nl.refPin.data();
        k.width = nl.width.empty() ? nullptr : nl.width.data();
        k.height = nl.height.empty() ? nullptr : nl.height.data();
        k.area = nl.area.empty() ? nullptr : nl.area.data();
        k.obstacles = nl.obstacles.empty() ? nullptr : nl.obstacles.data();
        const int rc = kt_upload_netlist(ctx_, &k);
        if (rc != KT_OK) {
            const std::string msg = kt_last_error();
            kt_ctx_destroy(ctx_);
            throw Error("ktplace_amd: " + msg);
        }
    }
    ~GpuSystems() { kt_ctx_destroy(ctx_); }
    GpuSystems(const GpuSystems &) = delete;
    GpuSystems &operator=(const GpuSystems &) = delete;

    /* The C context, for the entry points the facade does not wrap (batches,
     * launch shapes, the no-allocation HPWL). */
    kt_ctx *context() const { return ctx_; }

    /* LinearSystems::buildB2B (simpl_lss.cc:11-27); anchors may be empty when
     * alpha == 0; floors: mean movable width / height (simpl_placer.cc:169-179). */
    void buildB2B(const std::vector<double> &x, const std::vector<double> &y, double alpha,
                  const std::vector<double> &ax, const std::vector<double> &ay, double floorX, double floorY) {
        check(kt_build_b2b(ctx_, x.data(), y.data(), alpha, ax.empty() ? nullptr : ax.data(),
                           ay.empty() ? nullptr : ay.data(), floorX, floorY));
    }

    /* LinearSystems::buildStar (simpl_lss.cc:116-201): returns n + stars, the
     * size the caller extends x / y to (simpl_placer.cc:412-415). */
    std::size_t buildStar() {
        uint64_t size = 0;
        check(kt_build_star(ctx_, &size));
        return size;
    }

    /* solveBothAxes (simpl_lss.cc:203-290) on the systems built last; x, y in/out. */
    Residuals solve(std::vector<double> &x, std::vector<double> &y, double tol, std::size_t maxIter,
                    const IterateObserver &observe = {}, bool mixed = false) {
        if (x.size() != y.size()) throw Error("ktplace_amd: x and y differ in length");
        kt_solve_opts o{};
        o.tol = tol;
        o.max_iter = maxIter;
        o.precision = mixed ? KT_PREC_MIXED : KT_PREC_FP64;
        o.stop_rule = mixed ? KT_STOP_TRUE : KT_STOP_RECURRENCE;
        o.warm_start = warm_ ? KT_WARM_LEARNED : KT_WARM_PREVIOUS;
        kt_solve_result r{};
        const IterateObserver *f = observe ? &observe : nullptr;
        const auto cb = [](void *u, uint64_t it, double rx, double ry, int last) {
            (*static_cast<const IterateObserver *>(u))(static_cast<std::size_t>(it), Residuals{rx, ry}, last != 0);
        };
        check(kt_solve_both(ctx_, x.data(), y.data(), x.size(), &o, f ? +cb : nullptr,
                            const_cast<IterateObserver *>(f), &r));
        lastIterations_ = r.iterations;
        return Residuals{r.resid[0], r.resid[1]};
    }
    std::size_t lastIterations() const { return static_cast<std::size_t>(lastIterations_); }

    /* The learned warm start for later solves (KT_MLP_PARAMS floats). */
    void enableWarmStart(const std::vector<float> &params) {
        check(kt_set_warm_start(ctx_, params.data(), params.size()));
        warm_ = true;
    }

    /* Rows (one entry per subrow) and die for buildGrid (simpl_placer.cc:190-224). */
    void setRows(const std::vector<double> &y, const std::vector<double> &height, const std::vector<double> &x0,
                 const std::vector<double> &x1, const std::vector<double> &pitch, const double die[4]) {
        check(kt_set_rows(ctx_, static_cast<uint32_t>(y.size()), y.data(), height.data(), x0.data(), x1.data(),
                          pitch.data(), die));
    }
    std::size_t buildGrid(std::size_t perAxis, double densityLimit) {
        uint64_t bins = 0;
        check(kt_build_grid(ctx_, perAxis, densityLimit, &bins));
        return bins;
    }

    /* LookAheadLegalizer::legalize (simpl_lal.cc:43-124) on x / y in place. */
    void legalize(std::vector<double> &x, std::vector<double> &y, const LalOptions &opts = {}) {
        check(kt_set_lal_options(ctx_, opts.maxLevel, opts.minCellsToSplit, opts.stripeAreaFraction));
        check(kt_lal(ctx_, x.data(), y.data()));
    }

    /* getHpwl / getOverflow (simpl_placer.cc:256-297) of a placement. */
    double hpwl(const std::vector<double> &x, const std::vector<double> &y) {
        double h = 0.0;
        check(kt_set_positions(ctx_, x.data(), y.data(), x.size()));
        check(kt_hpwl(ctx_, &h));
        return h;
    }
    double overflow(const std::vector<double> &x, const std::vector<double> &y) {
        double o = 0.0;
        check(kt_set_positions(ctx_, x.data(), y.data(), x.size()));
        check(kt_overflow(ctx_, &o));
        return o;
    }

    /* The SparseSystem of one axis, read back (parity checks). */
    SparseSystem download(int axis) {
        uint64_t rows = 0, nnz = 0;
        check(kt_system_size(ctx_, axis, &rows, &nnz));
        SparseSystem s;
        std::vector<uint64_t> rp(rows + 1);
        s.col.resize(nnz);
        s.val.resize(nnz);
        s.diag.resize(rows);
        s.rhs.resize(rows);
        check(kt_get_system(ctx_, axis, rp.data(), s.col.data(), s.val.data(), s.diag.data(), s.rhs.data()));
        s.rowPtr.assign(rp.begin(), rp.end());
        return s;
    }

    std::size_t movable() const { return n_; }

private:
    kt_ctx *ctx_ = nullptr;
    std::size_t n_ = 0;
    bool warm_ = false;
    uint64_t lastIterations_ = 0;
};

}  // namespace ktamd
