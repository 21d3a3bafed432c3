/*
 * eplace_gen.h -- seeded synthetic placement instances (host only, no GPU).
 *
 * Replaces the reference's netlist load (DataBase::DataBase, DataBase.cpp:7-96,
 * LEF/DEF through Limbo) for the synthetic workloads this build measures:
 *   mesh     : 24x24 lattice of 1x1 cells, 96 perimeter pads, 4-neighbour
 *              2-pin nets, random uniform initial positions;
 *   stdcell  : h = 1 row, w ~ U{1..8} sites; fixed square macros of U{20..80}
 *              rows; nets of degree 2:60% 3:18% U{4..10}:20% U{11..64}:2% chosen
 *              by locality in the cells' 1-D order; uniform in-cell pin offsets;
 *              initial centres = die centre + N(0, (0.01 W)^2).
 * Units: 1 site = 1 row = 1.0.
 */
#ifndef EPLACE_GEN_H
#define EPLACE_GEN_H

#include <stdint.h>

#include "eplace.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { EP_GEN_MESH = 0, EP_GEN_STDCELL = 1 };

typedef struct ep_gen_params {
    int32_t kind;             /* EP_GEN_*                                            */
    int32_t n_movable;        /* stdcell: movable cells                              */
    double utilisation;       /* movable area / (die area - macro area)              */
    double macro_area_frac;   /* fraction of the die covered by fixed macros         */
    int32_t macro_min, macro_max; /* macro side range, rows                           */
    int32_t site_min, site_max;   /* cell width range, sites                          */
    double nets_per_cell;     /* nets = round(nets_per_cell * n_movable)              */
    double locality_sigma;    /* pin draw: N(0, sigma) around a centre in 1-D order  */
    double init_sigma_frac;   /* initial spread, fraction of die width               */
    int32_t shuffle_storage;  /* 1: store cells in a random order unrelated to the
                                 1-D locality order                                  */
    uint64_t seed;
} ep_gen_params;

typedef struct ep_design {
    ep_netlist nl;            /* arrays owned by the design                           */
    int32_t mesh_side;        /* mesh: lattice side (24), else 0                      */
} ep_design;

void ep_gen_default_params(ep_gen_params *p, int32_t kind);
ep_status ep_gen_create(const ep_gen_params *p, ep_design **out);
void ep_gen_free(ep_design *d);

#ifdef __cplusplus
}
#endif
#endif
