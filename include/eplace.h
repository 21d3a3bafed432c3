/*
 * eplace.h -- C ABI of the MI355X-native ePlace global-placement hot path.
 *
 * One context = one placement instance resident in HBM on the current HIP
 * device, driven on one HIP stream.  Calls are stream-ordered and must be
 * externally synchronised (one host thread per context).  No call allocates
 * per iteration; nothing but a few scalars crosses PCIe per iteration.
 *
 * Every entry point returns an ep_status; on failure a thread-local message is
 * available from ep_last_error().  No call exits the process or throws.
 *
 * Reference interfaces each call replaces (paths relative to the
 * searsm8/AIEplace tree):
 *   vck5000/host/src/include/GraphDriver.h:16-65   init / send_packet / receive_packet
 *   vck5000/host/src/include/AIEplace.h:21-99       class Placer (per-iteration calls)
 *   vck5000/host/src/include/Grid.h:40-46           computeBinOverlaps / computeTotalOverflow
 *   vck5000/host/src/include/DataBase.h             computeTotalWirelength
 */
#ifndef EPLACE_H
#define EPLACE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t ep_status;
enum {
    EP_OK = 0,
    EP_EINVAL = 1,   /* bad argument / inconsistent netlist               */
    EP_EDEVICE = 2,  /* HIP runtime error or missing device               */
    EP_ENAN = 3,     /* non-finite value detected in positions/gradients  */
    EP_ENOMEM = 4,   /* device or host allocation failed                  */
    EP_ESTATE = 5,   /* call out of order (e.g. ep_step before init)      */
    EP_ENOTSUP = 6,  /* feature not available in this build/config        */
    EP_EIO = 7       /* file missing, unreadable or malformed (LEF/DEF)   */
};

/* Arithmetic of the field-solve transform products. */
enum {
    EP_PREC_FP32 = 0,   /* FP32 operands, FP32 MFMA (v_mfma_f32_32x32x2_f32)        */
    EP_PREC_BF16X3 = 1, /* operands split hi+lo BF16, 3 MFMA products, FP32 accum */
    EP_PREC_FP16X3 = 2  /* operands split hi+lo FP16 (scaled), 3 products         */
};

typedef struct ep_ctx ep_ctx;

/* Netlist in structure-of-arrays form.  Cells are [movable | fixed].
 * Replaces DataBase's Component/Net/Pin maps (DataBase.h:26-200) and the
 * per-net packet tables (DataBase.cpp:164-316).  All arrays are copied by
 * ep_create; the caller keeps ownership. */
typedef struct ep_netlist {
    int32_t n_mov;            /* movable cells                                    */
    int32_t n_fix;            /* fixed cells (macros, IO pads)                    */
    const float *w, *h;       /* n_mov+n_fix sizes                                */
    const float *x, *y;       /* n_mov+n_fix lower-left corners: movable = initial
                                 placement, fixed = their location                */
    int32_t n_nets;
    int32_t n_pins;
    const int32_t *net_start; /* n_nets+1 CSR offsets into the pin arrays         */
    const int32_t *pin_cell;  /* n_pins, cell of each pin                         */
    const float *pin_dx;      /* n_pins, pin offset from the cell's lower-left    */
    const float *pin_dy;
    float xl, yl, xh, yh;     /* die                                              */
} ep_netlist;

/* Placer parameters (replaces host/config_aieplace.json, AIEplace.cpp:87-109). */
typedef struct ep_params {
    int32_t bins_x, bins_y;   /* bin grid, multiples of 16 (powers of two not required) */
    float target_density;     /* rho_t                                               */
    float gamma_bins;         /* WA gamma (base) in bin widths                       */
    int32_t gamma_schedule;   /* 1: gamma = base * 10^((tau-0.1)*20/9 - 1)           */
    float lambda_eta;         /* lambda_0 = eta * |grad W|_1 / |grad D|_1            */
    double hpwl_ref;          /* HPWL change giving mu = 1 (mu schedule)             */
    float mu_max, mu_min;     /* per-iteration lambda multiplier bounds              */
    int32_t max_backtrack;    /* step trials per iteration                           */
    float stop_overflow;      /* ep_run stops when tau <= this                       */
    int32_t max_iters;
    float init_disp_bins;     /* alpha_0 probe displacement, in bin widths           */
    int32_t use_fillers;      /* add whitespace fillers (ePlace)                     */
    uint64_t filler_seed;     /* host RNG seed for initial filler positions          */
    int32_t precision;        /* EP_PREC_*                                           */
    int32_t use_graph;        /* replay one gradient evaluation as a captured hipGraph */
    int32_t learned_step;     /* 1: step length proposed by the loaded MLP, verified by
                                 the backtracking condition (fallback to the search) */
    int32_t fork_min_cells;   /* captured evaluations run the WA gradient on a second
                                 branch from this many placeable cells on (results are
                                 identical either way); 0 = the default, 400000     */
    int32_t cell_order;       /* 1 (default): movable cells stored in breadth-first order
                                 over the netlist, so that cells of one net sit at nearby
                                 indices (gather locality for netlists in any order);
                                 0: stored in the caller's order.  Per-cell arrays passed
                                 in and out keep the caller's order either way.      */
} ep_params;

/* One Nesterov iteration's outcome (replaces printIterationResults'
 * Iteration/HPWL/Overflow table, AIEplace.cpp:1048-1059). */
typedef struct ep_step_info {
    int32_t iteration;
    int32_t evals;            /* gradient evaluations spent in this iteration        */
    int32_t learned_used;     /* 1 if the MLP proposal was tried                      */
    int32_t learned_accepted; /* 1 if that proposal passed the backtracking check    */
    double hpwl;              /* FP64 HPWL at the accepted point                     */
    double overflow;          /* tau at the accepted point                           */
    double lambda, gamma, alpha;
    double wa;                /* smooth WA wirelength at the accepted point          */
} ep_step_info;

typedef struct ep_run_info {
    int32_t iterations;
    int64_t evals;
    int32_t learned_proposals, learned_accepted;
    double hpwl, overflow;
    double seconds;           /* wall time of the loop (device-synchronised)         */
} ep_run_info;

/* ---- lifecycle (replaces Placer::Placer + GraphDriver::init, AIEplace.cpp:87-138,
 *      GraphDriver.h:16-40) ------------------------------------------------- */
void ep_default_params(ep_params *p);
ep_status ep_create(const ep_netlist *nl, const ep_params *p, ep_ctx **out);
/* Whitespace fillers that ep_create adds (host only, needs no device): count,
 * size and initial lower-left corners (x, y may be NULL to query the count;
 * otherwise they receive min(n_fill, cap) entries).  Filler area =
 * rho_t * (die area - fixed area inside the die) - movable area; filler size
 * = mean movable cell size. */
ep_status ep_plan_fillers(const ep_netlist *nl, const ep_params *p, int32_t *n_fill, float *fw,
                          float *fh, float *x, float *y, int32_t cap);
ep_status ep_destroy(ep_ctx *ctx);
const char *ep_last_error(void);
ep_status ep_get_sizes(ep_ctx *ctx, int32_t *n_mov, int32_t *n_fill, int32_t *n_fix,
                       int32_t *n_nets, int32_t *n_pins);
/* Cell sizes [mov | fill | fix] and fixed lower-left corners (n_fix). */
ep_status ep_get_cells(ep_ctx *ctx, float *w, float *h, float *fix_x, float *fix_y);
/* Netlist as stored on the device: nets regrouped by degree class.  net_perm[k]
 * = original index of stored net k.  Arrays n_nets+1 / n_pins / n_nets. */
ep_status ep_get_nets(ep_ctx *ctx, int32_t *net_start, int32_t *pin_cell, float *pin_dx,
                      float *pin_dy, int32_t *net_perm);

/* ---- positions (replaces Node::setPosition / translate) ------------------ */
/* Movable + filler lower-left corners, n_mov+n_fill each.  set also resets the
 * Nesterov state (u = v = given positions). */
ep_status ep_set_positions(ep_ctx *ctx, const float *x, const float *y);
ep_status ep_get_positions(ep_ctx *ctx, float *x, float *y);

/* ---- per-iteration kernels (replaces Placer::computeAllPartials_{AIE,CPU},
 *      computeOverlaps, computeElectricFields_{AIE,CPU}, AIEplace.cpp:256-967;
 *      the device side of PartialsGraphDriver/DensityGraphDriver::send_packet /
 *      receive_packet, GraphDriver.cpp:30-70) -------------------------------- */
/* WA wirelength gradient at the current positions; per-pin partials summed per
 * cell.  wa (nullable) receives the smooth wirelength, hpwl (nullable) the HPWL. */
ep_status ep_wa_grad(ep_ctx *ctx, float gamma, double *wa, double *hpwl);
ep_status ep_get_wa_grad(ep_ctx *ctx, float *gx, float *gy); /* n_mov+n_fill */
/* Bin density maps at the current positions.  field_map (nullable): smoothed
 * movable+filler overlap plus rho_t x fixed overlap (area units); mov_map
 * (nullable): exact overlap of movable cells; fix_map (nullable): exact fixed
 * overlap.  Maps are bins_x*bins_y, x-major (map[i*bins_y + j]). */
ep_status ep_density(ep_ctx *ctx, double *overflow);
ep_status ep_get_density(ep_ctx *ctx, float *field_map, float *mov_map, float *fix_map);
/* Poisson field of the current field map with the given precision. */
ep_status ep_field(ep_ctx *ctx, int32_t precision);
ep_status ep_get_field(ep_ctx *ctx, float *ex, float *ey);
/* Field error report: solves the current field map with `precision` and with an
 * FP64 device reference solve; returns max|E - E64| / max|E64| for x and y.
 * Leaves the `precision` field in place (as ep_field). */
ep_status ep_field_error(ep_ctx *ctx, int32_t precision, double *err_ex, double *err_ey);
/* Full preconditioned gradient (density + field + WA + gather) with explicit
 * gamma/lambda.  Outputs (nullable) are n_mov+n_fill: g = (gW - lambda F)/P,
 * and the components gW and F. */
ep_status ep_gradient(ep_ctx *ctx, float gamma, float lambda, float *gx, float *gy,
                      float *gwx, float *gwy, float *fx, float *fy, double *hpwl,
                      double *overflow);
/* Bin range of each movable+filler cell, 4 ints per cell (i0,i1,j0,j1):
 * stretched = 1 for the smoothed rectangle, 0 for the exact one. */
ep_status ep_cell_bins(ep_ctx *ctx, int32_t stretched, int32_t *out);

/* ---- reporting (Placer::printIterationResults, DataBase::computeTotalWirelength,
 *      Grid::computeTotalOverflow) ----------------------------------------- */
ep_status ep_hpwl(ep_ctx *ctx, double *hpwl);
ep_status ep_overflow(ep_ctx *ctx, double *overflow);

/* ---- optimiser (replaces Placer::nudgeAllNodes + run, AIEplace.cpp:44-76,
 *      993-1046: fixed-step descent becomes Nesterov with backtracking) ------- */
ep_status ep_init_optimizer(ep_ctx *ctx);            /* lambda_0, gamma_0, alpha_0 */
ep_status ep_step(ep_ctx *ctx, ep_step_info *info);   /* one iteration */
ep_status ep_run(ep_ctx *ctx, ep_run_info *info, ep_step_info *log, int32_t log_cap);
ep_status ep_set_precision(ep_ctx *ctx, int32_t precision);

/* ---- measurement (replaces the gettimeofday timers of AIEplace.cpp:1077-1085) --
 * Times each kernel group of one gradient evaluation at the current state with
 * HIP events on the context's stream, `reps` launches each, and reports the
 * algorithmic HBM bytes and flops of one launch.  Groups: cell_sort,
 * density_scatter, overflow, wa_grad, field_transforms, [pin_sum,] cell_grad,
 * nesterov_trial, finalize, then the scatter at other cells per thread; contexts
 * that fork their captured evaluation (fork_min_cells) report pin_sum and a
 * cell_grad that reads its sums, as the graph runs them. */
typedef struct ep_kernel_stat {
    char name[32];
    int32_t launches;   /* launches of this group per gradient evaluation        */
    double avg_us;      /* event time per group launch                           */
    double bytes;       /* algorithmic (unique) HBM bytes per group launch       */
    double flops;       /* algorithmic flops per group launch: the dense transform
                           products, 6 Mx My (Mx + My) per field solve          */
} ep_kernel_stat;
ep_status ep_profile(ep_ctx *ctx, int32_t reps, ep_kernel_stat *out, int32_t cap, int32_t *n);

/* ---- learned step length (no reference counterpart) ---------------------- */
/* Records of (16 features, log accepted step) collected by ep_step since the
 * last call; returns the count in *n (buffers may be NULL to query). */
ep_status ep_take_step_records(ep_ctx *ctx, float *features, float *log_step, int32_t cap,
                               int32_t *n);
/* Loads MLP weights (16->64->64->1, GELU) for the learned step proposal. */
ep_status ep_set_step_model(ep_ctx *ctx, const float *weights, int32_t n_weights);

/* ---- reference wire format (DataBase.cpp:217-316, GraphDriver.cpp:30-70) --- */
/* Runs the WA partials of one AIE-format packet on the device: input
 * [net_size, group_count, 0, 0, then for each group net_size vectors of 8 lane
 * values, pin 0 = max, pin 1 = min] (compute_abc.cpp:27-57); output
 * net_size*8 partials per group in the same lane order (compute_partials.cpp:11-63). */
ep_status ep_wa_packet(const float *packet_in, int32_t n_in, float gamma, float *partials_out,
                       int32_t n_out);
/* ep_wa_packet with a choice of exponent.  EP_WA_EXP_HW: the hardware exp in
 * FP32, as ep_wa_packet.  EP_WA_EXP_FAST64: the AIE kernel's fast_exp,
 * (1 + a/2^16)^(2^16) by 16 squarings (compute_abc.cpp:12-18), with the
 * whole partial in FP64 in the order of the reference's golden generator
 * (python/AIEmath/computeTerm.py:6-56, createGoldenPartials.py:48-70); for
 * its packets (integer coordinates, pins after pin 0 ascending) the outputs
 * are the golden partials rounded once to FP32. */
enum { EP_WA_EXP_HW = 0, EP_WA_EXP_FAST64 = 1 };
ep_status ep_wa_packet_exp(const float *packet_in, int32_t n_in, float gamma, float *partials_out,
                           int32_t n_out, int32_t exp_mode);
/* ep_wa_packet without any device allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (tight packet loops, code beside
 * graph capture, harnesses that forbid hipMalloc).  The packet is staged in one
 * fixed array per GPU (2 x EP_WA_PACKET_NOALLOC_MAX floats = 8 MiB, held until
 * the process ends); packets with more data values are rejected with EP_EINVAL.
 * All calls on one GPU share that array, so calls that overlap on one GPU race.
 * Results equal ep_wa_packet's bit for bit. */
#define EP_WA_PACKET_NOALLOC_MAX (1 << 20)
ep_status ep_wa_packet_noalloc(const float *packet_in, int32_t n_in, float gamma, float *partials_out,
                               int32_t n_out);

/* ---- one transform on the MFMA product path (the AIE DCT / IDCT / IDXST graphs'
 *      DCT_in -> DCT_out, vck5000/aie/src/density/DCTGraph.h:41-42, IDCTGraph.h:40-41,
 *      IDXSTGraph.h:42-43; the host's DCT_naive / IDCT_naive / IDXST_naive and
 *      transpose, vck5000/host/src/DCT.cpp:5-82) ----------------------------- */
enum { EP_TR_NONE = -1, EP_TR_DCT = 0, EP_TR_IDCT = 1, EP_TR_IDXST = 2 };
/* in, out: nx x ny row-major host arrays (out may equal in).  row_kind is applied
 * to every row (length ny), then col_kind to every column (length nx); EP_TR_NONE
 * leaves that axis as it is (not both).  Unscaled reference conventions:
 *   DCT    y_k = sum_n x_n cos(pi k (n + 1/2) / N)
 *   IDCT   y_k = x_0 / 2 + sum_{n>=1} x_n cos(pi n (k + 1/2) / N)
 *   IDXST  y_k = (-1)^k IDCT(x_0, x_{N-1}, ..., x_1)_k
 * nx and ny are multiples of 16 (at most 16384, nx * ny <= 2^26); precision is
 * EP_PREC_*.  Allocates its device buffers per call. */
ep_status ep_transform(int32_t row_kind, int32_t col_kind, int32_t nx, int32_t ny, const float *in,
                       float *out, int32_t precision);

#ifdef __cplusplus
}
#endif
#endif
