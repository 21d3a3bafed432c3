/*
 * eplace_oracle.h -- CPU restatement (FP64) of the ePlace global-placement hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product (aieplace_amd/) never links or calls it.
 *
 * Every function cites the reference file:line whose algorithm it restates
 * (paths relative to the searsm8/AIEplace tree).  Where the reference has a
 * documented quirk that the product does not reproduce, the oracle follows the
 * corrected convention and the comment names the quirk.
 */
#ifndef EPLACE_ORACLE_H
#define EPLACE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- scalar / per-net wirelength primitives ---------------- */

/* e^a by (1 + a/2^prec)^(2^prec): python/AIEmath/computeTerm.py:6-13,
 * vck5000/aie/src/partials/compute_abc.cpp:12-18 */
double eo_fast_exp(double a, int prec);

/* Per-net, one axis, reference (un-shifted) form.
 * a+_i = e^{(x_i - x_max)/g}, a-_i = e^{(x_min - x_i)/g}; b = sum a; c = sum x a.
 * computeTerm.py:15-43; vck5000/host/src/AIEplace.cpp:612-647.
 * use_fast != 0 selects eo_fast_exp(.,16) (the golden-vector generator's exp). */
void eo_wa_terms(const double *x, int d, double gamma, int use_fast,
                 double *a_plus, double *a_minus, double *bc /* b+,c+,b-,c- */);

/* dW/dx_i per pin, reference formula computeTerm.py:50-56,
 * AIEplace.cpp:652-673 (term1 - term2). */
void eo_wa_partials(const double *x, int d, double gamma, int use_fast, double *partials);

/* WA wirelength of one net/axis, computeTerm.py:45-47 */
double eo_wa_value(const double *x, int d, double gamma, int use_fast);

/* ---------------- netlist-level kernels ---------------- */

typedef struct eo_design {
    int32_t n_mov;       /* movable cells (on nets)                     */
    int32_t n_fill;      /* filler cells (movable, no nets)             */
    int32_t n_fix;       /* fixed cells; order is [mov | fill | fix]    */
    const double *w, *h; /* cell sizes, n_mov+n_fill+n_fix             */
    const double *fix_x, *fix_y; /* lower-left of fixed cells, n_fix   */
    int32_t n_nets;
    const int32_t *net_start; /* n_nets+1, CSR into pins               */
    const int32_t *pin_cell;  /* cell index of each pin                 */
    const double *pin_dx, *pin_dy; /* pin offset from cell lower-left   */
    double xl, yl, xh, yh;   /* die                                     */
    int32_t bins_x, bins_y;  /* bin grid Mx x My                        */
} eo_design;

/* WA gradient of all nets w.r.t. cell lower-left positions, pins = cell + offset.
 * Follows AIEplace.cpp:574-673 for every net of degree >= 2 (the reference skips
 * degree > 8 and, in CPU mode, drops the result; both corrected here).
 * x,y: positions of cells [mov|fill] (fixed cells from design->fix_x/fix_y).
 * gx,gy: n_mov+n_fill outputs (fillers get 0).  wa: optional sum of WA_x + WA_y. */
void eo_wa_grad(const eo_design *d, const double *x, const double *y, double gamma,
                double *gx, double *gy, double *wa);

/* Half-perimeter wirelength, Net.cpp:136-155 / DataBase.cpp:143-149, with pin
 * offsets (the reference uses cell corners). */
double eo_hpwl(const eo_design *d, const double *x, const double *y);

/* FP32 bin index range of a stretched cell rectangle -- the exact float
 * operation sequence the device kernel uses (Grid.cpp:38-61 index rule with
 * floor instead of truncation).  Used for the bit-exact bin-indexing check. */
void eo_cell_bins_f32(int n, const float *x, const float *y, const float *w, const float *h,
                      float xl, float yl, float bw, float bh, float sqrt2_bw, float sqrt2_bh,
                      int Mx, int My, int stretch, int32_t *out /* 4n: i0,i1,j0,j1 */);

/* Density maps, Bin.h:78-121 (exact rectangle overlap in double), Grid.cpp:38-110
 * (all cells get exact overlap; no point-mass shortcut).  map is Mx*My, x-major
 * (map[i*My + j]).  Outputs accumulate (caller zeroes).
 *   eo_density_exact:     sum of exact overlap area of cells [c0, c1)
 *   eo_density_stretched: cells with w < sqrt2*bw are widened to sqrt2*bw around
 *                         their centre (same in y) and scaled by area/stretched
 *                         area (ePlace local smoothing). */
void eo_density_exact(const eo_design *d, const double *x, const double *y, int c0, int c1,
                      double *map);
void eo_density_stretched(const eo_design *d, const double *x, const double *y, int c0, int c1,
                          double *map);

/* Fixed-cell overlap map (exact), Bin.h:78-121. */
void eo_density_fixed(const eo_design *d, double *map);

/* Overflow tau = sum_b max(0, Dmov_b + rho_t*Dfix_b - rho_t*A_b) / movable area.
 * Bin.h:123-126 / Grid.cpp:144-151 with target density and normalisation. */
double eo_overflow(const eo_design *d, const double *map_mov, const double *map_fix,
                   double rho_t);

/* ---------------- transforms (unscaled, reference conventions) ---------------- */
/* DCT-II  y_k = sum_n x_n cos(pi k (n+1/2)/N)            DCT.cpp:24-40, customDCT.py:9-25 */
void eo_dct(const double *x, int n, double *y);
/* IDCT    y_k = x_0/2 + sum_{n>=1} x_n cos(pi n (k+1/2)/N) DCT.cpp:45-59, customDCT.py:27-42 */
void eo_idct(const double *x, int n, double *y);
/* IDXST   y_k = (-1)^k IDCT(x_0, x_{N-1}, ..., x_1)_k      DCT.cpp:64-82, customDCT.py:57-77 */
void eo_idxst(const double *x, int n, double *y);
/* IDST    y_k = sum_{n>=1} x_n sin(pi n (k+1/2)/N)          customDCT.py:44-55 */
void eo_idst(const double *x, int n, double *y);
/* 2-D: apply the 1-D transform to each row (length ny), then to each column
 * (length nx).  kind: 0 dct, 1 idct, 2 idxst, 3 idst.  customDCT.py:79-158 (without
 * the 3-decimal rounding of dct_2d, customDCT.py:90). Mixed kinds: row_kind along
 * the second index, col_kind along the first. */
void eo_transform2d(const double *in, int nx, int ny, int col_kind, int row_kind, double *out);

/* Row transforms with one cos() call per term, as the reference's host path
 * evaluates them: kind 0 DCT_naive (double sum, DCT.cpp:24-40), 1 IDCT_naive
 * (float sum, DCT.cpp:45-59), 2 IDXST_naive (reversal + IDCT + sign,
 * DCT.cpp:64-82).  rows x n, row-major; OpenMP over rows.  Used to time the
 * reference-faithful CPU field (six such passes of M rows per solve). */
void eo_rows_naive(const float *in, int rows, int n, int kind, float *out);

/* Poisson field solve, RePlAce convention (python/AIEmath/fft_replace/fft.cpp:118-212):
 *   a = DCT2(rho) * 4/(Mx My), halved on u=0 and v=0;  wx = pi u/Mx,
 *   wy = pi v/My * wy_scale;  phi^ = a/(wx^2+wy^2) (0 at DC);
 *   phi = sum phi^ cos cos;  Ex = sum phi^ wx sin_x cos_y;  Ey = sum phi^ wy cos_x sin_y.
 * The placer passes wy_scale = bw/bh (frequencies in units of 1/bw);
 * fft.cpp:85-88 multiplies by binSizeY/binSizeX instead -- identical for square
 * bins, which is the only case compared against it.
 * Separable dense-matrix evaluation, O(M^3).  phi may be NULL. */
void eo_field(const double *rho, int Mx, int My, double wy_scale,
              double *phi, double *ex, double *ey);
/* Same quantity by the direct quadruple sum (AIEplace.cpp:694-749 structure,
 * with the pi/M frequency).  O(M^4); tiny grids only. */
void eo_field_direct(const double *rho, int Mx, int My, double wy_scale,
                     double *phi, double *ex, double *ey);

/* Density force on each movable/filler cell: F_i = r_i * sum_b ov(stretched_i, b) * E_b
 * (the cell-side gather of AIEplace.cpp:1004-1016 over the smoothed rectangle). */
void eo_density_force(const eo_design *d, const double *x, const double *y,
                      const double *ex, const double *ey, double *fx, double *fy);

/* ---------------- full Nesterov placer (the trajectory oracle) ---------------- */

typedef struct eo_params {
    double target_density;   /* rho_t                                         */
    double gamma_bins;       /* WA gamma base, in bin widths                  */
    int32_t gamma_schedule;  /* 0: gamma fixed; 1: gamma = base*10^((tau-0.1)*20/9-1) */
    double lambda_eta;       /* lambda_0 = eta * |gW|_1 / |F|_1               */
    double hpwl_ref;         /* reference HPWL change for the mu schedule     */
    double mu_max, mu_min;   /* lambda multiplier bounds                      */
    int32_t max_backtrack;   /* trials per iteration                          */
    double stop_overflow;    /* stop when tau <= this                         */
    int32_t max_iters;
    double init_disp_bins;   /* displacement of the alpha_0 probe, bin widths */
} eo_params;

typedef struct eo_iter_log {
    int32_t iter;
    int32_t evals;     /* gradient evaluations in this iteration       */
    double hpwl, overflow;   /* at the accepted point v_k               */
    double lambda, gamma;    /* used by this iteration's trials          */
    double alpha;            /* alpha_hat of the accepted trial: the next
                                iteration's first step                   */
    double alpha_used;       /* step length of the accepted trial        */
    double coef;             /* Nesterov coefficient (a_k - 1)/a_{k+1}   */
} eo_iter_log;

/* Runs ePlace Nesterov with backtracking from (x,y) (cells [mov|fill]) until
 * tau <= stop_overflow or max_iters.  x,y are overwritten by the final
 * placement (v_k).  log (optional, capacity log_cap) receives one entry per
 * iteration.  Returns the number of iterations done.  There is no counterpart
 * in the reference (its loop is fixed-step gradient descent,
 * AIEplace.cpp:44-76, 993-1046); this follows the ePlace paper cited in
 * README.md:19. */
int eo_place(const eo_design *d, const eo_params *p, double *x, double *y,
             eo_iter_log *log, int log_cap, double *final_hpwl, double *final_overflow,
             int64_t *total_evals);

/* One full gradient evaluation (exposed for parity tests).  Returns tau.
 * g = (gW - lambda F)/max(1, pins_i + lambda*area_i), cells [mov|fill]. */
double eo_gradient(const eo_design *d, const double *x, const double *y, double rho_t,
                   double gamma, double lambda, double *gx, double *gy, double *hpwl,
                   double *gwx, double *gwy, double *fx, double *fy, double *ex, double *ey);

/* ---------------- step-length MLP (no reference counterpart) ---------------- */
/* 16 -> 64 -> 64 -> 1, GELU (x/2 (1 + erf(x/sqrt 2))), blob layout as the
 * product's include/eplace_mlp.h.  Forward returns log10 step (denormalised);
 * grad returns the MSE loss on the normalised target and d loss / d blob. */
int eo_mlp_nweights(void);
void eo_mlp_forward(const double *w, const double *x, int n, double *out);
double eo_mlp_grad(const double *w, const double *x, const double *y, int n, double *grad);

/* Sets the OpenMP thread count used by the oracle (0 = leave default). */
void eo_set_threads(int n);
int eo_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
