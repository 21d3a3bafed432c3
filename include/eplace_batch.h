/*
 * eplace_batch.h -- many independent placement instances advanced together.
 *
 * All instances live in one device context.  Four Nesterov trials of every
 * unfinished instance run as one replay of a captured hipGraph.  Every kernel
 * node of that graph covers all instances at once, with blocks mapped to
 * (instance, tile) work items, and a device-side kernel reduces each
 * instance's sums and applies its backtracking decision and lambda/gamma
 * schedules without a host round trip.  The host reads the instances' state
 * back after every replay, with the next replay already queued.  Tile shapes (density
 * cells per block, the transform products' K step) are picked by a timing
 * autotuner at creation.
 * No counterpart in the reference (single instance, AIEplace.cpp:44-76).
 */
#ifndef EPLACE_BATCH_H
#define EPLACE_BATCH_H

#include <stdint.h>

#include "eplace.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ep_batch ep_batch;

typedef struct ep_batch_run_info {
    int32_t instances;
    int32_t finished;           /* instances that reached stop_overflow            */
    int64_t instance_iterations;/* accepted Nesterov iterations summed over instances */
    int64_t evals;              /* gradient evaluations summed over instances       */
    int32_t launches;           /* graph replays                                    */
    double seconds;             /* device-synchronised wall time of the loop         */
    int32_t density_cells_per_thread;  /* autotuned scatter tile                    */
    int32_t transform_k_step;   /* autotuned K step of the transform products (half-k
                                   per LDS stage) at the batch's precision           */
} ep_batch_run_info;

/* n instances (1..1024); bins[i] = square bin grid of instance i (multiple of 32).
 * p supplies the shared parameters (precision, schedules, stop_overflow, ...);
 * its bins_x/bins_y are ignored.  Netlists are copied. */
ep_status ep_batch_create(const ep_netlist *const *nls, const int32_t *bins, int32_t n,
                          const ep_params *p, ep_batch **out);
ep_status ep_batch_destroy(ep_batch *b);
/* Back to the initial placement and a fresh optimiser state. */
ep_status ep_batch_reset(ep_batch *b);
/* Initial gamma, lambda_0 and probe step of every instance (ep_init_optimizer,
 * batched); ep_batch_run calls it when needed. */
ep_status ep_batch_init(ep_batch *b);
/* Runs until every instance stops (overflow or max_iters) or max_launches graph
 * replays (each replay advances every unfinished instance by up to 4 trials).
 * A later call continues where this one stopped. */
ep_status ep_batch_run(ep_batch *b, int32_t max_launches, ep_batch_run_info *info);
/* Per-instance results (arrays of n; any may be NULL). */
ep_status ep_batch_results(ep_batch *b, int32_t *iterations, int64_t *evals, double *hpwl,
                           double *overflow);
/* Movable + filler positions of instance i (sizes from ep_batch_sizes). */
ep_status ep_batch_sizes(ep_batch *b, int32_t i, int32_t *n_mov, int32_t *n_fill, int32_t *n_fix);
ep_status ep_batch_positions(ep_batch *b, int32_t i, float *x, float *y);
/* Preconditioned gradient at the current point of instance i (movable + fillers). */
ep_status ep_batch_gradient(ep_batch *b, int32_t i, float *gx, float *gy);
/* Current gamma, lambda and Lipschitz step estimate of instance i. */
ep_status ep_batch_state(ep_batch *b, int32_t i, double *gamma, double *lambda, double *alpha);

#ifdef __cplusplus
}
#endif
#endif
