/*
 * ktplace_amd.h -- C ABI of the MI355X SimPL inner loop.
 *
 * Replaces, for one design at a time, the reference's solver interface:
 *   LinearSystems::buildB2B / buildStar          simpl_lss.h:68-84, simpl_lss.cc:11-201
 *   SparseSystem (CSR off-diagonals + diag/rhs)  simpl_lss.h:38-58
 *   solveBothAxes (lock-step Jacobi-PCG)         simpl_lss.h:86-88, simpl_lss.cc:203-290
 *   getHpwl / getOccupancy / getOverflow         simpl_placer.cc:256-297
 *   LookAheadLegalizer::legalize                 simpl_lal.h:18-29, simpl_lal.cc:43-124
 *
 * Conventions: every entry point returns KT_OK (0) or a negative KT_E* status;
 * kt_last_error() gives the message for the calling thread.  No C++ exception
 * crosses the ABI.  A context belongs to one device and one host thread at a
 * time (like the reference's LinearSystems, it is not re-entrant).  Entry
 * points taking host pointers are synchronous with respect to those buffers.
 * The *_resident variants work on device buffers owned by the context and are
 * asynchronous on the context's stream; kt_sync waits for them.
 */
#ifndef KTPLACE_AMD_H
#define KTPLACE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KT_OK 0
#define KT_EINVAL -1    /* bad argument or size */
#define KT_EHIP -2      /* HIP runtime error */
#define KT_ESTATE -3    /* call out of order (e.g. no netlist uploaded) */
#define KT_ENOMEM -4
#define KT_NO_VAR 0xffffffffu /* simpl_lss.h:13 noVar */

#define KT_AXIS_X 0
#define KT_AXIS_Y 1

/* SimPL's flat solver netlist (simpl_lss.h:27-36, built by simpl_dm.cc:45-85),
 * plus the per-movable-cell sizes and obstacles of simpl_dm.h:97-112. */
typedef struct kt_netlist {
    uint32_t num_vertices;
    uint32_t num_movable;
    uint32_t num_nets;
    uint64_t num_pins;             /* after per-net dedup */
    uint64_t num_cell_pins;        /* pinsOfCell length */
    uint32_t num_obstacles;
    const uint64_t *net_start;     /* [N+1] */
    const uint32_t *pin_cell;      /* [P] vertex id (NetPins::cells) */
    const double *pin_off_x;       /* [P] offset from the cell's lower-left */
    const double *pin_off_y;       /* [P] */
    const double *net_weight;      /* [N] */
    const uint32_t *var_of_vertex; /* [V] KT_NO_VAR for fixed */
    const uint32_t *mov_vertex;    /* [n] */
    const double *fixed_x;         /* [V] */
    const double *fixed_y;         /* [V] */
    const uint64_t *pin_start;     /* [n+1] */
    const uint32_t *cell_pin_net;  /* [Pm] PinRef::net */
    const uint32_t *cell_pin_idx;  /* [Pm] PinRef::pin */
    const double *width;           /* [n] */
    const double *height;          /* [n] */
    const double *area;            /* [n] */
    const double *obstacles;       /* [4*M] xl, yl, xh, yh */
} kt_netlist;

typedef struct kt_ctx kt_ctx;

/* ---- context ----------------------------------------------------------- */
int kt_ctx_create(kt_ctx **out, int device);
int kt_ctx_destroy(kt_ctx *ctx);
int kt_sync(kt_ctx *ctx);
const char *kt_last_error(void);
/* Device memory held by the context, in bytes. */
int kt_ctx_device_bytes(kt_ctx *ctx, uint64_t *bytes);

/* Copies the netlist to the device once (Netlist is const on the path,
 * simpl_dm.h:97-112).  Replaces any previous netlist. */
int kt_upload_netlist(kt_ctx *ctx, const kt_netlist *nl);

/* ---- positions (device-resident x, y of the movable cells) ------------- */
/* count = rows of the current systems (n, or n + star variables). */
int kt_set_positions(kt_ctx *ctx, const double *x, const double *y, uint64_t count);
int kt_get_positions(kt_ctx *ctx, double *x, double *y, uint64_t count);
int kt_set_anchors(kt_ctx *ctx, const double *ax, const double *ay);
/* Saves / restores the positions inside the context (device-to-device). */
int kt_save_positions(kt_ctx *ctx);
int kt_restore_positions(kt_ctx *ctx);

/* ---- B2B assembly: LinearSystems::buildB2B (simpl_lss.cc:11-114) -------- */
/* Host-pointer form: x, y [n]; anchor_x/y may be NULL when alpha <= 0.  The
 * CSR keeps the reference's row order, entry order and duplicates. */
int kt_build_b2b(kt_ctx *ctx, const double *x, const double *y, double alpha,
                 const double *anchor_x, const double *anchor_y, double floor_x,
                 double floor_y);
/* Resident form: uses the context's positions and anchors. */
int kt_build_b2b_resident(kt_ctx *ctx, double alpha, double floor_x, double floor_y);

/* LinearSystems::buildStar (simpl_lss.cc:116-201): both star-model systems
 * (n movable cells + one star per net with >= 3 pins and a movable pin, in net
 * order); *size = n + stars.  Positions do not enter; solve them with
 * positions of `size` entries (the reference starts the stars at the die
 * centre, simpl_placer.cc:409-414). */
int kt_build_star(kt_ctx *ctx, uint64_t *size);

/* ---- SparseSystem access (simpl_lss.h:38-58) ---------------------------- */
int kt_system_size(kt_ctx *ctx, int axis, uint64_t *rows, uint64_t *nnz);
int kt_get_system(kt_ctx *ctx, int axis, uint64_t *row_ptr, uint32_t *col, double *val,
                  double *diag, double *rhs);
int kt_set_system(kt_ctx *ctx, int axis, uint64_t rows, const uint64_t *row_ptr,
                  const uint32_t *col, const double *val, const double *diag, const double *rhs);
/* out = A * in for one axis (SparseSystem::multiply, simpl_lss.h:49-57). */
int kt_multiply(kt_ctx *ctx, int axis, const double *in, double *out);

/* ---- solveBothAxes (simpl_lss.cc:203-290) -------------------------------- */
#define KT_PREC_FP64 0   /* reference algorithm in FP64 */
#define KT_PREC_MIXED 1  /* FP32 PCG with FP64 residual refinement */
#define KT_PREC_FP64_EXACT 2 /* FP64 in the reference's exact operation order
                                (bit-identical to the CPU solve; slow) */

#define KT_STOP_RECURRENCE 0 /* reference rule: ||r_k|| / ||b|| <= tol */
#define KT_STOP_TRUE 1       /* true FP64 ||b - A x|| / ||b|| <= tol */

#define KT_WARM_PREVIOUS 0 /* start from the previous iterate (the reference) */
#define KT_WARM_LEARNED 1  /* start from previous + gamma * MLP-predicted displacement,
                              gamma from an exact line search (0 falls back) */

typedef struct kt_solve_opts {
    double tol;
    uint64_t max_iter;
    int32_t precision;   /* KT_PREC_* */
    int32_t stop_rule;   /* KT_STOP_* */
    int32_t check_every; /* iterations per host convergence check (0 = default) */
    int32_t profile;     /* record per-kernel HIP event timings */
    int32_t warm_start;  /* KT_WARM_*: where CG starts */
    int32_t reserved;    /* 0 */
} kt_solve_opts;

typedef struct kt_solve_result {
    double resid[2];      /* final recurrence residual per axis (Residuals) */
    double true_resid[2]; /* final true FP64 relative residual per axis */
    uint64_t iterations;  /* lock-step iterations (the reference's loop count) */
    uint64_t steps[2];    /* iterations each axis stepped */
    uint64_t refinements; /* FP64 refinement sweeps (mixed precision) */
    double warm_gamma[2]; /* KT_WARM_LEARNED: step taken along the predicted
                             displacement per axis (0: previous iterate kept) */
} kt_solve_result;

/* Called after the solve for every iterate, in order, with the residuals the
 * reference's IterateObserver would see (simpl_lss.h:65-66). */
typedef void (*kt_observer)(void *user, uint64_t iterate, double resid_x, double resid_y,
                            int last);

int kt_solve_both(kt_ctx *ctx, double *x, double *y, uint64_t count, const kt_solve_opts *opts,
                  kt_observer observe, void *user, kt_solve_result *out);
int kt_solve_both_resident(kt_ctx *ctx, const kt_solve_opts *opts, kt_solve_result *out);

/* ---- placement rows, bin grid and bookkeeping (simpl_placer.cc:181-297) -- */
/* Rows (one subrow each: [x0, x1) at site pitch `pitch`) and the placement
 * die box {xl, yl, xh, yh}.  Obstacles come with the netlist (kt_netlist). */
int kt_set_rows(kt_ctx *ctx, uint32_t num_rows, const double *row_y, const double *row_height,
                const double *row_x0, const double *row_x1, const double *row_pitch, const double die[4]);
/* BinGrid of bins_per_axis^2 bins (0: the reference's default rule
 * clamp(sqrt(n / 51), 16, 256), simpl_placer.cc:181-188) with the free area
 * of every bin (:190-224); density_limit is clamped to [0.05, 1]. */
int kt_build_grid(kt_ctx *ctx, uint64_t bins_per_axis, double density_limit, uint64_t *bins_out);
int kt_grid_avail(kt_ctx *ctx, double *avail, double *total_avail);
/* Of the context's positions: occupancy per bin (getOccupancy, :282-289),
 * scaled overflow (getOverflow, :291-297) and HPWL (getHpwl, :256-275; same
 * per-net values and the same chunked summation order as the reference). */
int kt_occupancy(kt_ctx *ctx, double *occ);
int kt_overflow(kt_ctx *ctx, double *overflow);
int kt_hpwl(kt_ctx *ctx, double *hpwl);
/* kt_hpwl with no allocation: no counterpart in the reference; for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same result as kt_hpwl,
 * bit for bit.  Its chunk sums live in a fixed per-GPU device array of 128
 * KiB, so it takes at most 67,104,768 nets (KT_EINVAL beyond); every call on
 * one GPU shares that array, so calls that overlap on one GPU race. */
int kt_hpwl_noalloc(kt_ctx *ctx, double *hpwl);

/* ---- the outer loop's bookkeeping on the device -------------------------- */
/* getHpwl, getOccupancy / getOverflow (simpl_placer.cc:256-297) and
 * Convergence::update (:450-461) of simplPlacer::Impl::getGlobalPlacement
 * (:463-539), with the lower bound in the context's positions and the upper
 * bound in its anchors (the next B2B build's anchor set), so no placement
 * crosses PCIe during the loop. */
typedef struct kt_loop_state {
    double lower_hpwl;      /* HPWL of the positions (lower bound) */
    double upper_hpwl;      /* HPWL of the anchors (upper bound) */
    double gap;             /* upper - lower */
    double overflow_upper;  /* getOverflow of the anchors */
    double overflow_lower;  /* getOverflow of the positions (measure-only calls) */
    double gap_reference;   /* gap of iteration reference_iter */
    double best_upper;
    uint64_t iteration;
    uint64_t stalled_for;
    int32_t converged;      /* Convergence::update said stop */
    int32_t reserved;
} kt_loop_state;
/* Convergence options (simpl_options.h: gapReferenceIter 10, gapTightFrac 0.05,
 * gapRelaxedFrac 0.25, stallIters 6); clears the convergence state. */
int kt_loop_reset(kt_ctx *ctx, uint64_t reference_iter, double tight_frac, double relaxed_frac,
                  uint64_t stall_iters);
/* anchors := positions on the device (the upper bound after legalization). */
int kt_anchors_from_positions(kt_ctx *ctx);
int kt_get_anchors(kt_ctx *ctx, double *x, double *y);
/* HPWL of anchors and positions, overflow of the anchors and, for iteration
 * >= 1, Convergence::update for that iteration; iteration 0 measures only
 * (and adds the overflow of the positions).  One small device-to-host copy. */
int kt_loop_update(kt_ctx *ctx, uint64_t iteration, kt_loop_state *out);

/* ---- LookAheadLegalizer::legalize (simpl_lal.cc:43-124) ------------------ */
/* max_level / min_cells_to_split / stripe_area_fraction: simpl_options.h:24-26
 * (20, 4, 0.1 by default; 0 keeps the current value). */
int kt_set_lal_options(kt_ctx *ctx, uint64_t max_level, uint64_t min_cells_to_split,
                       double stripe_area_fraction);
/* Host-pointer form: x, y [n] in/out.  Resident form: the context's positions. */
int kt_lal(kt_ctx *ctx, double *x, double *y);
int kt_lal_resident(kt_ctx *ctx);
/* Batches of non-touching regions the last legalization ran (simpl_lal.cc:98-122). */
int kt_lal_batches(kt_ctx *ctx, uint64_t *batches);
/* Region runs the last legalization repeated so that a batch's result is the
 * one of its regions run in order (a cell that moved into a later region of
 * the same batch is legalized again there, as the reference does with one
 * thread, simpl_lal.cc:102-122). */
int kt_lal_reruns(kt_ctx *ctx, uint64_t *reruns);
/* Batches of up to max_cells cells (at most 1024) in up to 64 regions run in
 * one launch each, every level with its block geometry (free-area profile,
 * cutline, stripes) on the device; larger batches run level by level with
 * the geometry on host threads.  Both give the same positions.  Default 0:
 * every batch level by level, which is the faster of the two on the
 * 210,000-cell loop (one launch costs more than a level's host round trip
 * there).  kt_lal_fused_batches: batches of the last legalization that ran
 * in one launch. */
int kt_set_lal_fused(kt_ctx *ctx, uint32_t max_cells);
int kt_lal_fused_batches(kt_ctx *ctx, uint64_t *batches);

/* ---- launch-shape tuning (0 keeps the default) --------------------------- */
/* lanes: threads per CSR row in the SpMV (1, 2, 4, 8 or 16); tile: rows per
 * block (a multiple of 1024); use_graphs: replay iteration pairs from a
 * captured hipGraph; refresh_factor: mixed precision refreshes the FP64
 * residual once the FP32 residual has fallen by this factor (0 < f < 1). */
int kt_set_tuning(kt_ctx *ctx, int lanes, uint32_t tile, int use_graphs, double refresh_factor);
/* As kt_set_tuning, plus xcd_split: 1 puts the x system on XCDs 0-3 and the y
 * system on XCDs 4-7 (each XCD's L2 then holds one axis' vectors), 0 deals
 * blocks of both axes to all XCDs, 2 gives each XCD one contiguous share of
 * the entry blocks (the entry-block SpMV; the vector kernels deal blocks as
 * with 0), -1 keeps the default (1); rows_in_flight:
 * CSR rows each lane group keeps in flight in the SpMV (2, 4 or 8; 0 = 2). */
int kt_set_tuning2(kt_ctx *ctx, int lanes, uint32_t tile, int use_graphs, double refresh_factor,
                   int xcd_split, int rows_in_flight);
/* LDS staging of the CG SpMV (a launch-shape knob with no counterpart in the
 * reference): each block copies p[tile start - halo, tile end + halo) into
 * LDS and reads the columns inside that window from there.  halo: a multiple
 * of 256 entries, 0 = no staging (the default).  The window, (tile + 2 halo)
 * entries of the solve's working precision, must fit 156 KiB, or the solve
 * fails with KT_EINVAL.  Results do not depend on it. */
int kt_set_spmv_halo(kt_ctx *ctx, uint32_t halo);
/* CG SpMV layout (a launch-shape knob with no counterpart in the reference).
 * KT_SPMV_ROW_GROUPS (default): lane groups of `lanes` threads per row, rows
 * in flight and the optional LDS window as set above.  KT_SPMV_ENTRY_BLOCKS:
 * each block of block_threads threads (1024, 512 or 256; 0 = 1024) takes the
 * rows of a run of about block_entries consecutive entries (1024 threads:
 * 4096, 8192 or 16384; 512 threads: 2048, 4096 or 8192; 256 threads: 2048 or
 * 4096; 0 = 8 per thread); its threads read the run's columns and values
 * fully coalesced, unroll (4 or 8, 8 with 1024 threads only; 0 = 4) entries
 * each per round, gather p and keep the products in LDS; with row_lanes 1 (or
 * 0) one thread per row adds them in entry order after the diagonal term (the
 * reference's multiply order), with row_lanes 4 four lanes per row add every
 * fourth product and combine by a fixed butterfly.  All are deterministic;
 * they round differently (the block size changes only the order in which
 * p.Ap is summed).  The entry-block layout takes no LDS window (KT_EINVAL at
 * the solve). */
#define KT_SPMV_ROW_GROUPS 0
#define KT_SPMV_ENTRY_BLOCKS 1
int kt_set_spmv_layout(kt_ctx *ctx, int layout, uint32_t block_entries, int unroll, int row_lanes,
                       int block_threads);
/* Bytes per column index read by the entry-block SpMV: 4 (or 0), or 3 (a
 * 16-bit low half and an 8-bit high byte, converted once per system).  3
 * applies to blocks of 256 threads, 2048 entries and 4 per round on systems
 * of fewer than 2^24 rows; elsewhere the 4-byte columns are read.  The
 * products, their order and the results are the same either way. */
int kt_set_spmv_columns(kt_ctx *ctx, int bytes);
/* 1: the entry-block SpMV (blocks of 256 threads and 2048 entries, 4 lanes
 * per row) reads four consecutive entries per thread with one wide load per
 * array; 0 (default): one entry per load.  Results are the same either way. */
int kt_set_spmv_wide_loads(kt_ctx *ctx, int on);
/* Mixed precision (KT_PREC_MIXED) only, no counterpart in the reference: 1
 * (default) runs the FP32 iterations on the symmetrically scaled system
 * D^-1/2 A D^-1/2 (unit diagonal; FP32 values val_ij / (sqrt(d_i) sqrt(d_j)),
 * converted once per system), which is Jacobi-PCG in other coordinates: the
 * SpMV and the direction update then read no diagonal.  The FP64 refreshes,
 * the stopping test and the result are those of the unscaled system.  0: the
 * FP32 iterations divide by the diagonal as the FP64 solve does. */
int kt_set_mixed_scaling(kt_ctx *ctx, int on);
/* Mixed precision (KT_PREC_MIXED) on the scaled system only, no counterpart in
 * the reference: 1 (default) runs each FP32 iteration as two kernels where the
 * solve runs the entry-block shape with wide loads and 3-byte columns: the
 * SpMV also sums r.Ap and Ap.Ap, so the new residual's norm r'.r' = r.r -
 * 2 alpha r.Ap + alpha^2 Ap.Ap is known without a second grid-wide reduction
 * and one kernel updates r, the correction and p (and sums r'.r' per tile,
 * from which the next iteration's rho comes).  0: three kernels (the
 * residual update and its r.r, then the direction update).  Same Jacobi-PCG
 * iterates up to the rounding of beta's numerator. */
int kt_set_mixed_fused_update(kt_ctx *ctx, int on);

/* ---- per-kernel timings (HIP events on the context's stream) -------------- */
#define KT_KERNEL_SPMV 0      /* FP64 CG SpMV + p.Ap */
#define KT_KERNEL_UPDATE 1    /* FP64 x/r update + r.r, r.z */
#define KT_KERNEL_PUPDATE 2   /* FP64 p = z + beta p */
#define KT_KERNEL_SPMV32 3    /* FP32 CG SpMV + p.Ap */
#define KT_KERNEL_B2B_FILL 4  /* B2B fill pass */
#define KT_KERNEL_WARM 5      /* learned warm start, whole (features .. line search) */
#define KT_KERNEL_MLP 6       /* learned warm start, network inference */
#define KT_KERNEL_COUNT 8
int kt_kernel_times(kt_ctx *ctx, int kernel, uint64_t *launches, double *total_ms);
int kt_reset_kernel_times(kt_ctx *ctx);

/* ---- learned warm start (no counterpart in the reference) ---------------- */
/* A 3-layer MLP (KT_MLP_INPUTS -> 128 -> 128 -> 2, ReLU) with BF16 weights and
 * FP32 accumulation predicts each cell's displacement from its previous
 * iterate to the solution of the new systems, in units of the placement's
 * spread.  Parameters: KT_MLP_PARAMS FP32 values, W1 [128][32], b1 [128],
 * W2 [128][128], b2 [128], W3 [2][128], b3 [2] (row-major, output unit first).
 * Features: KT_MLP_FEATURES per cell (the remaining inputs are zero). */
#define KT_MLP_INPUTS 32
#define KT_MLP_FEATURES 17
#define KT_MLP_HIDDEN 128
#define KT_MLP_OUTPUTS 2
#define KT_MLP_PARAMS 20994

/* Network used by KT_WARM_LEARNED solves on this context (NULL clears it). */
int kt_set_warm_start(kt_ctx *ctx, const float *params, uint64_t count);
/* Movable cells with no net path to a fixed vertex (their component's
 * translation is in the null space of the systems built without anchors):
 * the learned start leaves them at the previous iterate.  flags [n] (may be
 * null) receives 1 for each such cell.  No counterpart in the reference. */
int kt_floating_cells(kt_ctx *ctx, uint32_t *count, uint8_t *flags);
/* Features [n][KT_MLP_FEATURES] of the context's positions and current
 * systems (build them first), and the spread scale[2] = {s_x, s_y} that
 * turns network outputs into displacements.  features may be NULL. */
int kt_warm_features(kt_ctx *ctx, float *features, double *scale);
/* The predicted displacement dx, dy [n] (before the line search). */
int kt_warm_predict(kt_ctx *ctx, double *dx, double *dy);

/* Training and evaluation of the network on the GPU (BF16 matrix-core
 * forward and backward passes, FP32 master weights, Adam). */
typedef struct kt_mlp kt_mlp;
int kt_mlp_create(int device, kt_mlp **out);
int kt_mlp_destroy(kt_mlp *m);
int kt_mlp_set_params(kt_mlp *m, const float *params, uint64_t count);
int kt_mlp_get_params(kt_mlp *m, float *params, uint64_t count);
/* Training set: features [count][KT_MLP_FEATURES], targets [count][2] and
 * per-sample loss weights [count] (NULL: all 1). */
int kt_mlp_set_data(kt_mlp *m, const float *features, const float *targets, const float *weights,
                    uint64_t count);
/* out [count][2] = network(features [count][KT_MLP_FEATURES]). */
int kt_mlp_forward(kt_mlp *m, const float *features, uint64_t count, float *out);
/* Gradient of the loss (1/count) sum_s w_s |network(f_s) - t_s|^2 over the
 * listed training samples with respect to every parameter
 * (grad [KT_MLP_PARAMS]), and the loss itself. */
int kt_mlp_gradient(kt_mlp *m, const uint32_t *samples, uint64_t count, float *grad, double *loss);
/* `steps` Adam steps (lr, beta 0.9 / 0.999) on minibatches of `batch` samples
 * (a multiple of 32) drawn from the training set by a hash of (seed, step);
 * loss [steps] receives each batch's mean loss (may be NULL). */
int kt_mlp_train(kt_mlp *m, uint64_t steps, uint32_t batch, double lr, uint64_t seed, float *loss);

/* ---- batched small designs (no counterpart in the reference) ------------ */
/* Many designs' x/y systems solved together: one launch per CG kernel covers
 * every tile of every system, each system keeps its own Jacobi-PCG scalars
 * and stopping rules (solveBothAxes's), and the iteration pairs replay one
 * captured hipGraph until no system is live.  FP64, previous-iterate start. */
typedef struct kt_system_view {
    uint64_t rows;
    const uint64_t *row_ptr; /* [rows + 1], row_ptr[0] = 0 */
    const uint32_t *col;     /* [nnz] off-diagonal columns (local to the system) */
    const double *val;       /* [nnz] */
    const double *diag;      /* [rows] */
    const double *rhs;       /* [rows] */
} kt_system_view;

typedef struct kt_batch_instance_result {
    double resid[2];
    double true_resid[2];
    uint64_t steps[2];
    double done_ms; /* wall time from the start of kt_batch_solve until both of
                       this instance's systems had stopped (host poll granularity) */
} kt_batch_instance_result;

typedef struct kt_batch kt_batch;
int kt_batch_create(int device, kt_batch **out);
int kt_batch_destroy(kt_batch *b);
int kt_batch_clear(kt_batch *b);
/* Appends one instance (x and y systems of equal size, start positions). */
int kt_batch_add(kt_batch *b, const kt_system_view *sys_x, const kt_system_view *sys_y, const double *x0,
                 const double *y0, uint32_t *index);
/* Appends the context's current systems and positions as one instance. */
int kt_batch_add_ctx(kt_batch *b, kt_ctx *ctx, uint32_t *index);
int kt_batch_size(kt_batch *b, uint32_t *instances, uint64_t *rows, uint64_t *nnz);
/* lanes per CSR row (4, 8, 16), rows per tile (a multiple of 256), rows in
 * flight per lane group (2, 4), hipGraph replay on/off; 0 keeps a value. */
int kt_batch_set_tuning(kt_batch *b, int lanes, uint32_t tile, int rows_in_flight, int use_graphs);
/* SpMV layout of the batch (as kt_set_spmv_layout): KT_SPMV_ROW_GROUPS, or
 * KT_SPMV_ENTRY_BLOCKS with blocks of about 2048 consecutive entries of one
 * system and row_lanes (1 or 4; 0 = 1) lanes per row in the row sums. */
int kt_batch_set_layout(kt_batch *b, int layout, int row_lanes);
/* Bytes per column index read by the batch's entry-block SpMV: 4 (or 0), or
 * 3 (16 + 8 bits, as kt_set_spmv_columns) for batches of fewer than 2^24
 * rows.  The results are the same either way. */
int kt_batch_set_columns(kt_batch *b, int bytes);
/* 1: the batch's entry-block SpMV with 4 row lanes reads four consecutive
 * entries per thread with one wide load per array (as kt_set_spmv_wide_loads);
 * 0 (default): one entry per load.  The results are the same either way. */
int kt_batch_set_wide_loads(kt_batch *b, int on);
/* Every instance back to the start positions it was added with. */
int kt_batch_reset(kt_batch *b);
/* Solves all instances from their current positions (opts: KT_PREC_FP64,
 * KT_WARM_PREVIOUS); per_instance [instances] and total may be NULL. */
int kt_batch_solve(kt_batch *b, const kt_solve_opts *opts, kt_batch_instance_result *per_instance,
                   kt_solve_result *total);
int kt_batch_get_positions(kt_batch *b, uint32_t index, double *x, double *y, uint64_t count);

/* Calibration helper (no counterpart in the reference): copies `bytes` bytes
 * between two scratch device buffers `reps` times with a plain streaming copy
 * kernel (kt_copy_probe) and returns the average milliseconds per copy.  Used
 * to scale hardware byte counters (FETCH_SIZE / WRITE_SIZE) to real bytes. */
int kt_probe_copy(kt_ctx *ctx, uint64_t bytes, int reps, double *ms_per_copy);

#ifdef __cplusplus
}
#endif
#endif
