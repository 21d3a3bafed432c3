/*
 * eplace_legal.h -- row-based legalisation after global placement.
 *
 * No counterpart in the reference: its flow ends with the global placement
 * written to DEF (AIEplace.cpp:44-76, 1123).  This is the step a full flow runs
 * next, on the host: Abacus (P. Spindler, U. Schlichtmann, F. M. Johannes,
 * "Abacus: fast legalization of standard cell circuits with minimal
 * movement", ISPD 2008).  Cells are taken in order of x; each goes to the row
 * (and the free segment of that row) where, after the row's clusters are
 * re-solved, its squared displacement is smallest.  Within a row the cluster
 * solution minimises the summed squared displacement for the given order.
 */
#ifndef EPLACE_LEGAL_H
#define EPLACE_LEGAL_H

#include <stdint.h>

#include "eplace.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ep_legal_info {
    double total_disp;    /* sum over movable cells of |final - global| (Euclidean)  */
    double max_disp;
    double sum_sq_disp;   /* sum of squared displacements                           */
    int32_t rows, segments;
    double seconds;
} ep_legal_info;

/* Legal lower-left corners for the n_mov movable cells of nl, from the global
 * placement x, y (n_mov lower-left corners; nl->x/y are not read for them).
 * Rows of height row_height start at the die bottom, sites of site_width at
 * the die left edge; a fixed cell blocks the sites it overlaps in every row it
 * overlaps.  Movable cells must be one row high (EP_ENOTSUP otherwise); widths
 * are rounded up to whole sites.  snap_to_sites = 0 leaves the clusters at
 * their real-valued optimum (legal up to site alignment).  x_out, y_out may
 * alias x, y.  EP_EINVAL when a cell fits in no row. */
ep_status ep_legalize(const ep_netlist *nl, const float *x, const float *y, float row_height,
                      float site_width, int32_t snap_to_sites, float *x_out, float *y_out,
                      ep_legal_info *info);

#ifdef __cplusplus
}
#endif
#endif
