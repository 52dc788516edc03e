/*
 * ktplace_gen.h -- seeded synthetic netlists and SimPL's flat solver netlist.
 *
 * Two host-side pieces that sit either side of the B2B + PCG path:
 *
 *   ktg_generate    builds a raw placement design (cells, pads, macros, rows,
 *                   nets with possibly repeated cells) from one uint64 seed.
 *                   The reference reads Bookshelf/LEF-DEF files instead
 *                   (src/adaptor/); the synthetic design stands in for them.
 *   ktg_collect     flattens a raw design into SimPL's solver netlist exactly
 *                   as simpl_dm.cc:16-85 (collectCells/collectNets) does: pins
 *                   stable-sorted by cell id and deduplicated keeping the first
 *                   pin of each cell, varOfVertex/movVertex, pinStart and
 *                   pinsOfCell ordered by ascending net then pin.
 *
 * All arrays are owned by the handle and stay valid until ktg_free.  Every
 * function returns 0 on success and a negative status on error; no C++
 * exception crosses this interface.
 */
#ifndef KTPLACE_GEN_H
#define KTPLACE_GEN_H

#include <stdint.h>

#include "ktplace_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

#define KTG_NO_VAR 0xffffffffu

/* Generator parameters.  The generator is frozen: a given spec and seed always
 * produce the same design, on any host. */
typedef struct ktg_spec {
    uint64_t seed;
    uint32_t num_cells;        /* movable standard cells, unit row height, width 1..8 sites */
    uint32_t num_nets;
    uint32_t num_pads;         /* fixed zero-size I/O pads evenly spaced on the die perimeter */
    uint32_t num_macros;       /* fixed macros 20..200 sites x 2..20 rows (obstacles) */
    double utilization;        /* movable area / (die area - macro area) */
    uint32_t max_degree;       /* largest net degree (40 by default, 200 for the 1M-cell case) */
    double high_fanout_frac;   /* share of nets with 41..max_degree pins, taken from the 11..40 bucket */
    double local_frac;         /* share of non-anchor pins drawn from the anchor's window */
} ktg_spec;

/* Raw design: vertex-indexed cells plus nets as pin lists (before dedup). */
typedef struct ktg_raw {
    uint32_t num_vertices;
    uint32_t num_movable;
    uint32_t num_nets;
    uint64_t num_pins;
    uint32_t num_rows;
    const uint8_t *is_fixed;     /* [V] */
    const double *pos_x;         /* [V] lower-left x (fixed cells; 0 for movable) */
    const double *pos_y;         /* [V] */
    const double *width;         /* [V] */
    const double *height;        /* [V] */
    const uint64_t *net_start;   /* [N+1] */
    const uint32_t *pin_vertex;  /* [P] */
    const double *pin_off_x;     /* [P] offset from the cell's lower-left (kt_graph.h:56-57) */
    const double *pin_off_y;     /* [P] */
    const double *net_weight;    /* [N] */
    const double *row_y;         /* [R] row bottom */
    const double *row_height;    /* [R] */
    const double *row_x0;        /* [R] one subrow per row: [x0, x1) */
    const double *row_x1;        /* [R] */
    const double *row_pitch;     /* [R] site pitch */
    double die[4];               /* xl, yl, xh, yh */
} ktg_raw;

/* SimPL's flat netlist is the C ABI's kt_netlist (include/ktplace_amd.h). */
typedef kt_netlist ktg_flat;

typedef struct ktg_design ktg_design;

int ktg_default_spec(ktg_spec *spec, uint32_t num_cells, uint32_t num_nets, uint32_t num_pads,
                     uint64_t seed);
int ktg_generate(const ktg_spec *spec, ktg_design **out);
int ktg_raw_view(const ktg_design *d, ktg_raw *view);

/* Flattens a raw design (from ktg_generate or from the caller's own arrays). */
int ktg_collect(const ktg_raw *raw, ktg_design **out);
int ktg_flat_view(const ktg_design *d, ktg_flat *view);

/* simplPlacer::Impl::seedUniform (simpl_placer.cc:226-237): x = xl + U * max(W - w, 0),
 * y = yl + U * max(H - h, 0) with U = (mt19937_64() >> 11) * 2^-53, x then y per cell. */
int ktg_seed_uniform(uint64_t seed, uint64_t n, const double die[4], const double *width,
                     const double *height, double *x, double *y);

void ktg_free(ktg_design *d);
const char *ktg_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
