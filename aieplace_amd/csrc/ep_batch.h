// ep_batch.h -- device data of the batched engine (include/eplace_batch.h).
//
// All instances share concatenated device arrays.  Instance i owns
//   cells [cell_off, cell_off + n_all)   of wh, V[2], U[2], G[2], sorted
//   bins  [map_off, map_off + M*M)       of map_field/map_mov/map_fix/exy
//   bins  [2*map_off, 2*map_off + 2*M*M) of the transform scratch tA/tB
//   nets and pins: contiguous ranges of net_start/pin_cell/pin_off/pin_grad,
//   with net_start holding global pin indices and pin_cell global cell indices
//   sort tiles [i*kSortTiles, (i+1)*kSortTiles) of sort_count/sort_off.
// Blocks of every kernel map to instances through prefix tables (the blocks
// of instance i are [pre[i], pre[i+1])); blocks of a finished instance return
// at once, so a replay costs only the unfinished instances' work.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "ep_ctx.h"

namespace ep {

constexpr int kBatchMaxInstances = 1024;
constexpr int kBinsPerBlock = kBlock * 4;  // fill / overflow kernels

struct BInst {
    int32_t cell_off, n_mov, n_place, n_all;
    int32_t map_off, M;
    int32_t dens_level, dens_ext;  // density super-tile level and window extent (DensPlan)
    float xl, yl, xh, yh, bw, bh, sqrt2_bw, sqrt2_bh;
    float rho_t, cap;      // target density, target density * bin area
    double mov_area;
    double dens_unit;      // bin area / 2^32: value of one fixed-point density unit
    double hpwl_ref;       // mu-schedule reference
    double init_disp;      // probe displacement (init_disp_bins * bw)
    double gamma_base;     // gamma_bins * bw
};

// Optimiser state of one instance (device-resident; the control kernel is
// the only writer during a run).
struct BState {
    double alpha;      // Lipschitz step estimate (alpha_hat of the last accepted trial)
    double alpha_try;  // step of the pending trial
    double ak, ak1, coef;
    double lambda, gamma;
    double prev_hpwl, last_hpwl, last_tau, last_wa;
    double eps;        // probe displacement factor
    int64_t evals;
    int32_t cur;       // parity of the current V/U/G buffers
    int32_t iteration, trial, done, error;
    float f_alpha, f_coef, f_ginv, f_lambda, f_eps;  // values the kernels read
    int32_t pad;
};

// Reductions of the last evaluation of one instance (FP64, fixed order).
struct BRes {
    double hpwl, wa, ovf, tau, dv2, dg2, sum_gw, sum_f, gnorm2, gmax;
    int64_t nonfinite;
};

// Optimiser parameters shared by all instances.
struct BCtl {
    double lambda_eta, mu_max, mu_min, stop_overflow;
    int32_t gamma_schedule, max_backtrack, max_iters, pad;
};

struct BBucket {
    int32_t inst, lanes, net0, net1, blk0;
    int32_t deg, pin0;  // exact degree (0: read net_start) and first global pin
};

// Device pointers handed to every batched kernel (by value).
struct BDev {
    const BInst *inst;
    BState *st;
    BRes *res;
    BCtl ctl;
    int32_t n_inst;
    const float2 *wh;
    float2 *V[2], *U[2], *G[2];
    const int32_t *net_start, *pin_cell;
    const float2 *pin_off;
    float2 *pin_grad;
    const int32_t *cps, *cpins;  // global cell -> pin CSR (empty for non-movable cells)
    float *map_field, *map_mov, *map_fix;
    float2 *exy;                 // field (ex, ey) per bin
    unsigned long long *acc_field, *acc_mov;  // fixed-point scatter targets (2^32 per bin area)
    float *map_mean;             // [n_inst]
    uint32_t *map_max;           // [n_inst] float bits of max |map - mean|
    float4 *sorted;              // packed cells in coarse-tile order (per instance at cell_off)
    int32_t *sort_count, *sort_off;
    int32_t *dens_tab;           // [n_inst * (kSortTiles + 1)] density block tables
    int32_t dens_wcap;           // LDS window capacity per plane (bins), max over instances
    const int32_t *pre_cell;     // 256 cells per block
    const int32_t *pre_sort;     // kBlock * kSortCellsPerThread cells per block
    const int32_t *pre_dens;     // density blocks per instance (dens_grid bound)
    const int32_t *pre_bin;      // kBinsPerBlock bins per block
    const int32_t *pre_wa;       // WA blocks of each instance (instance-major buckets)
    const BBucket *buckets;
    int32_t n_buckets;
    double *part_wa, *part_ovf, *part_cell;  // per-block partials, indexed by block
};

enum BEvalAt { kAtCur = 0, kAtTrial = 1 };
enum BCtlMode { kCtlGamma = 0, kCtlLambda = 1, kCtlProbe = 2, kCtlAlpha = 3, kCtlIterate = 4 };
enum BMoveMode { kMoveTrial = 0, kMoveProbe = 1, kMoveInitU = 2 };

// ---- batched launchers (k_batch.hip)
// Blocks of kBlock * kSortCellsPerThread cells (the sort's layout); kMoveTrial
// also counts the trial positions into the sort tiles.
void bl_move(const BDev &d, int mode, int sort_blocks, hipStream_t s);
// Density maps of every unfinished instance; `counted`: a kMoveTrial move made
// the sort counts at the evaluated positions.
// bl_maps = bl_sort (tile sort) then bl_dens (scatter; overflow, which also
// clears the accumulators; with `map_range`, the field-map range that the
// split-FP16 products scale by).  The accumulators start cleared at creation.
void bl_sort(const BDev &d, int at, bool counted, int bin_blocks, int sort_blocks, int dens_cpt,
             hipStream_t s);
void bl_dens(const BDev &d, int bin_blocks, int dens_blocks, int dens_cpt, bool map_range, hipStream_t s);
void bl_maps(const BDev &d, int at, bool counted, int bin_blocks, int sort_blocks, int dens_blocks,
             int dens_cpt, bool map_range, hipStream_t s);
void bl_wa(const BDev &d, int at, int wa_blocks, hipStream_t s);
void bl_cell_grad(const BDev &d, int at, int cell_blocks, hipStream_t s);
// The per-instance FP64 reductions of an evaluation, then the control step
// `mode` (BCtlMode) of every unfinished instance on them.
void bl_reduce_control(const BDev &d, int mode, hipStream_t s);
bool bl_density_cpt_supported(int cpt);

// ---- batched field solve (k_field.hip)
struct BatchFieldInst {
    int M;
    float bw, bh;
    const float *map_field;
    float *tA, *tB;
    float2 *exy;
    const uint32_t *map_max;  // float bits of max |map - mean| (split FP16 scaling)
    const float *map_mean;
    const int32_t *skip;      // instance finished
};
struct BatchField;
ep_status batch_field_init(BatchField *&out, const std::vector<BatchFieldInst> &inst);
// Tile lists of the instances with active[i] only (finished instances get no tiles).
ep_status batch_field_set_active(BatchField *bf, const std::vector<char> &active);
void batch_field_free(BatchField *bf);
void launch_field_batch(BatchField *bf, int precision, hipStream_t s);
// Times the three K-step variants of the batched products for each precision
// (reps launches each, after one untimed launch) and keeps the fastest; the
// field maps must hold a valid density map.  Every variant gives the same result.
ep_status batch_field_autotune(BatchField *bf, int reps, hipStream_t s);
// K step (half-k staged per LDS step) the batched products use at `precision`.
int batch_field_k_step(const BatchField *bf, int precision);

}  // namespace ep
