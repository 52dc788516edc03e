/*
 * ktplace_gen.h -- seeded synthetic netlists and SimPL's flat solver netlist.
 *
 * Two host-side pieces that sit either side of the B2B + PCG path:
 *
 *   ktg_generate    builds a raw placement design (cells, pads, macros, rows,
 *                   nets with possibly repeated cells) from one uint64 seed.
 *   ktg_read_bookshelf  reads the same from a Bookshelf design directory, as
 *                   the reference's Bookshelf adaptor does (src/adaptor/);
 *                   ktg_write_pl writes a placement back.
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

/* A Bookshelf design directory <dir>/<name>.{nodes,nets,pl,scl,wts} (each
 * optionally .gz; <name> = the directory's name), read as
 * BookshelfInputAdapter::readFromDirectory (src/adaptor/bookshelfToKTAdaptor.cc:140-225)
 * does, with the die of getPlacementDieBox (src/datamodel/kt_die.cc:45-103),
 * then flattened as ktg_collect does.  Terminals and /FIXED cells are fixed;
 * pin offsets are kept from the cell's lower-left. */
int ktg_read_bookshelf(const char *dir, ktg_design **out);
/* The vertex names of a design read from files, each followed by '\n', in
 * vertex order; valid until ktg_free. */
int ktg_vertex_names(ktg_design *d, const char **blob, uint64_t *bytes);
/* A Bookshelf placement file as FlowMgr::Impl::writeBookshelf
 * (src/kt_flowMgr.cc:508-521) writes it: every vertex in order as
 * "name<TAB>x<TAB>y<TAB>: N", " /FIXED" after a fixed vertex's N, numbers
 * with 10 significant digits.  A movable vertex v (var_of_vertex[v] !=
 * KTG_NO_VAR) is written at (x[var], y[var]), a fixed one at (fixed_x[v],
 * fixed_y[v]); positions are lower-left corners.  names: num_vertices names,
 * each followed by '\n' (as ktg_vertex_names gives them). */
int ktg_write_pl(const char *path, uint32_t num_vertices, const char *names, uint64_t names_bytes,
                 const uint32_t *var_of_vertex, const double *fixed_x, const double *fixed_y, const double *x,
                 const double *y);

void ktg_free(ktg_design *d);
const char *ktg_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
