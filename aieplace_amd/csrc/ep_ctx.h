// ep_ctx.h -- device-resident placement context and kernel launch interface.
//
// Layout in HBM (n_all = n_mov + n_fill + n_fix, n_place = n_mov + n_fill):
//   wh[n_all]            float2  cell sizes
//   V[2][n_all]          float2  lower-left positions: V[cur] = v_k, V[cur^1] =
//                                trial point; fixed cells' tail copied in both
//   U[2][n_place]        float2  Nesterov major solution u_k / trial u
//   G[2][n_place]        float2  preconditioned gradient at V[cur] / at trial
//   net_start[N+1], pin_cell[P], pin_off[P] (float2): nets regrouped by degree
//                                class so that each class runs with a fixed
//                                number of lanes per net
//   pin_grad[P]          float2  per-pin WA partials (summed per cell by CSR)
//   cell_pin_start[n_mov+1], cell_pins[P_mov]   cell -> pin CSR
//   map_field, map_mov, map_fix [Mx*My] float  density maps (x-major)
//   acc_field, acc_mov [Mx*My] uint64  scatter accumulators (2^32 per bin area);
//                                zero between evaluations (the overflow pass
//                                clears what it converts)
//   sorted[n_place]      float4  cells (x, y, +-w, h) in coarse-tile order
//   exy[Mx*My]           float2 field (ex, ey) per bin, interleaved: one 8-byte
//                                gather per bin in the per-cell force
//   tA, tB                       spectral / transform scratch
//   transform matrices (FP32 and split BF16/FP16 copies)
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "ep_common.h"

namespace ep {

constexpr int kBlock = 256;
// Blocks of the final reduction (k_finalize); the context's fin buffer holds
// ten partial sums per block.
constexpr int kFinalizeBlocks = 32;
// Launch caps of the reducing kernels (bounds the partial sums to finalize).
constexpr int kWaGridCap = 16384;
constexpr int kCellGridCap = 16384;
// Spatial counting sort of cells: kSortSide x kSortSide coarse tiles.
constexpr int kSortSide = 32;
constexpr int kSortTiles = kSortSide * kSortSide;
constexpr int kSortCellsPerThread = 16;
// After a rejected model proposal the model sits out 2, 4, ... up to this many
// iterations (reset by an accepted proposal): phases where it mispredicts cost
// one extra evaluation per rejection, so they are skipped.
constexpr int kModelBackoffMax = 32;
// Captured evaluations fork the WA gradient onto a second branch only from this
// many placeable cells on (ep_params.fork_min_cells = 0): the fork and join cost
// ~15 us per trial, more than the overlap saves on small instances.
constexpr int64_t kForkMinCells = 400000;
inline int64_t fork_min_cells(const ep_params &p)
{
    return p.fork_min_cells > 0 ? (int64_t)p.fork_min_cells : kForkMinCells;
}
// Step-model records kept per context between ep_take_step_records calls.
constexpr size_t kMaxStepRecords = (size_t)1 << 21;
// Default HPWL reference of the lambda schedule: this many bin widths per net.
constexpr double kHpwlRefPerNetBin = 0.05;
// Net classes: exact degree 2, 3, 4 (2, 4, 4 lanes per net), then degrees up
// to 8, 16, 32 and beyond (8, 16, 32, 64 lanes).
constexpr int kMaxBuckets = 7;

// Density scatter plan: super-tile level (2^level x 2^level coarse tiles per
// block window), window extent beyond the super tile in bins, LDS window
// capacity per map in bins, cells per thread.
struct DensPlan {
    int32_t level = 0, ext = 0, wcap = 0, cpt = 4;
};

constexpr int kDensMaxWin = 6400;  // LDS bins per map of the largest window
// Fixed-point density units per bin area (the maps accumulate in uint64).
constexpr double kDensFixUnits = 4294967296.0;  // 2^32

// LDS bytes of a density window plan: two planes of 64-bit accumulators.
inline size_t dens_lds_bytes(const DensPlan &p) { return 2 * sizeof(unsigned long long) * (size_t)p.wcap; }

// Plan for an Mx x My grid whose smoothed cells are at most max_w x max_h wide.
inline DensPlan make_dens_plan(int Mx, int My, double bw, double bh, double max_w, double max_h,
                               int cpt)
{
    DensPlan p;
    const int m = Mx < My ? Mx : My;
    p.level = 6;
    for (int k = 0; k <= 6; ++k)
        if (((int64_t)m << k) / kSortSide >= 32) {
            p.level = k;
            break;
        }
    const int ex = (int)(max_w / bw) + 2, ey = (int)(max_h / bh) + 2;
    p.ext = ex > ey ? ex : ey;
    const int64_t sx = (((int64_t)Mx << p.level) + kSortSide - 1) / kSortSide + 2 + p.ext;
    const int64_t sy = (((int64_t)My << p.level) + kSortSide - 1) / kSortSide + 2 + p.ext;
    const int64_t w = sx * sy;
    p.wcap = (int)(w < kDensMaxWin ? w : kDensMaxWin);
    p.cpt = cpt;
    return p;
}

// Density grid: at most one partial block per super tile beyond n / K blocks.
inline int dens_grid(const DensPlan &p, int64_t n_place)
{
    const int64_t K = (int64_t)kBlock * p.cpt;
    return (int)((n_place + K - 1) / K + (kSortTiles >> (2 * p.level)));
}

struct Bucket {
    int32_t lanes;  // lanes per net
    int32_t net0, net1;
    int32_t blk0;   // first block of this bucket in the WA grid
    int32_t deg;    // exact degree of every net in the bucket (0: read net_start)
    int32_t pin0;   // first pin of the bucket (pins of its nets are contiguous)
};

struct BucketTable {
    Bucket b[kMaxBuckets];
    int32_t n;
};

// Scalars the kernels read each launch.  Host-controlled steps stage them from
// pinned host memory (a captured graph picks up new values on every replay);
// device-controlled runs have the step-control kernel write them in place.
struct DevScalars {
    float alpha;      // step length of the current trial
    float coef;       // (a_k - 1)/a_{k+1}
    float gamma_inv;  // 1/gamma
    float lambda;     // density weight
    float probe_eps;  // displacement factor of the alpha_0 probe
    int32_t use_model;  // 1: the MLP proposal kernel overwrites alpha
    int32_t sel;        // index into the Pair arguments of a launch (= cur on device)
    int32_t halt;       // 1: every iteration kernel returns at once (run finished)
    float step_unit;    // bin width: the step model predicts log10(step / step_unit)
    float feat[16];     // step-model features of the current iteration
};

// Two candidate buffers of a kernel argument; the kernel uses p[sc->sel].  A
// launch with fixed buffers passes the same pointer twice.  The device-controlled
// step graph passes {V[0], V[1]} for the current point and {V[1], V[0]} for the
// trial point, so one captured graph follows the accept decisions (sel = cur)
// made on the device.
template <typename T>
struct Pair {
    T *p[2];
};
template <typename T>
inline Pair<T> same(T *x)
{
    return Pair<T>{{x, x}};
}

// Optimiser state of a device-controlled run (ep_run with graphs): the host
// fields of Ctx below, advanced by k_step_ctl after every trial.
struct DevCtl {
    double alpha, ak, ak1, coef, lambda, lambda0, gamma, prev_hpwl;
    double last_hpwl, last_tau, last_wa;
    double h_alpha1, h_alpha2, h_prev_tau, h_prev_hpwl, h_ratio, h_gnorm, h_dv, h_dg, h_sum_ratio;
    int64_t evals;
    int32_t iteration, h_evals, trials, phase, t, learned_used, learned_ok;
    int32_t learned_proposals, learned_accepted, model_skip, model_backoff;
    int32_t cur, done, error, n_log, n_rec;
    float feat[16];
};

// Constant inputs of k_step_ctl.
struct CtlParams {
    double mu_min, mu_max, hpwl_ref, stop_overflow, gamma_base;  // gamma_base = gamma_bins * bw
    int32_t gamma_schedule, max_iters, max_backtrack, learned, log_cap, rec_cap;
    double bw;
    double inv_sqrt_np;  // 1 / sqrt(movable + filler cells): per-cell RMS of the norms
};

// Reductions of one gradient evaluation (FP64), written by the finalize kernel.
struct DevResult {
    double hpwl, wa, ovf_num, tau;
    double dv2, dg2;        // |v' - v|^2, |g' - g|^2
    double sum_gw, sum_f;   // L1 norms of WA gradient and density force
    double gmax;            // max |g| component
    double gnorm2;          // |g'|^2
    int64_t nonfinite;      // count of non-finite gradient components
    double alpha_used;      // step length the trial actually used
};

struct Ctx;

// ---- kernel launchers (defined in the .hip files) ----
// Clears the scatter accumulators (set-up and re-synchronisation; every
// evaluation's overflow pass leaves them clear).
void launch_fill_maps(Ctx &c, hipStream_t s);
// Tile-sorts the cells at pos; `counted`: the tile counts were already made by
// a counting launch_trial at the same positions.
void launch_sort(Ctx &c, Pair<const float2> pos, bool counted, hipStream_t s);
// Scatters the cells packed by the last launch_sort (same positions).
void launch_density(Ctx &c, hipStream_t s);
void launch_overflow(Ctx &c, hipStream_t s);
// Split-FP16 products: max |field map - mean| of the last overflow pass into
// field_max[0] (the operand scale of the first product).
void launch_map_range(Ctx &c, hipStream_t s);
void launch_wa(Ctx &c, Pair<const float2> pos, hipStream_t s);
void launch_field(Ctx &c, int precision, hipStream_t s);
// gw_cell (nullable): per-cell pin-partial sums made by launch_pin_sum; without
// it the kernel gathers them itself.
void launch_cell_grad(Ctx &c, Pair<const float2> pos, Pair<const float2> vref,
                      Pair<const float2> gref, Pair<float2> gout, float2 *gw_out, float2 *f_out,
                      hipStream_t s, const float2 *gw_cell = nullptr);
// gw_cell[c] = sum of movable cell c's pin partials (after launch_wa).
void launch_pin_sum(Ctx &c, hipStream_t s);
// trial: the evaluated point came from a counting launch_trial, whose partial
// sums hold |v' - v|^2 (k_cell_grad then gets no vref)
void launch_finalize(Ctx &c, hipStream_t s, bool trial = false);
// Nesterov trial into v_new/u_new; `count`: also count v_new into the sort's
// tiles (the following evaluation then sorts with counted = true).
void launch_trial(Ctx &c, Pair<const float2> v, Pair<const float2> u, Pair<const float2> g,
                  Pair<float2> v_new, Pair<float2> u_new, bool count, hipStream_t s);
// Device-controlled runs: start of the first iteration, and the finalize
// reduction followed at once by the step decision on its results.
void launch_step_begin(Ctx &c, const CtlParams &p, hipStream_t s);
void launch_finalize_decide(Ctx &c, const CtlParams &p, hipStream_t s);
// Fixed-buffer forms of the launchers above.
inline void launch_sort(Ctx &c, const float2 *pos, bool counted, hipStream_t s)
{
    launch_sort(c, same(pos), counted, s);
}
inline void launch_wa(Ctx &c, const float2 *pos, hipStream_t s) { launch_wa(c, same(pos), s); }
inline void launch_cell_grad(Ctx &c, const float2 *pos, const float2 *vref, const float2 *gref,
                             float2 *gout, float2 *gw_out, float2 *f_out, hipStream_t s)
{
    launch_cell_grad(c, same(pos), same(vref), same(gref), same(gout), gw_out, f_out, s);
}
inline void launch_trial(Ctx &c, const float2 *v, const float2 *u, const float2 *g, float2 *v_new,
                         float2 *u_new, bool count, hipStream_t s)
{
    launch_trial(c, same(v), same(u), same(g), same(v_new), same(u_new), count, s);
}
void launch_probe(Ctx &c, const float2 *v, const float2 *g, float2 *v_new, hipStream_t s);
void launch_mlp_propose(Ctx &c, hipStream_t s);
void launch_wa_packet(const float *data, int net_size, int n_nets, int groups, float ginv,
                      float *out, hipStream_t s);
// The packet's partials in FP64 with the AIE kernel's fast_exp (EP_WA_EXP_FAST64).
void launch_wa_packet_fast64(const float *data, int net_size, int n_nets, int groups, double gamma, float *out,
                             hipStream_t s);
// Device address of the current GPU's fixed packet staging (2 x
// EP_WA_PACKET_NOALLOC_MAX floats: data, then partials), looked up once per device.
ep_status wa_packet_stage(float **stage);
void launch_cell_bins(Ctx &c, const float2 *pos, int stretched, int32_t *out, hipStream_t s);
ep_status field_init(Ctx &c);   // builds transform matrices
// FP64 device solve of the current map; max-norm relative error of ex/ey against it.
ep_status field_error(Ctx &c, double *err_ex, double *err_ey);
void field_free(Ctx &c);
// One reference-convention 1-D/2-D DCT / IDCT / IDXST of a host nx x ny map on
// the product kernels (ep_transform; arguments validated by the caller).
ep_status run_transform(int row_kind, int col_kind, int nx, int ny, const float *in, float *out, int precision);

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // Second branch of the captured evaluation: the WA gradient runs beside the
    // sort -> density -> field chain (it needs only the trial positions).
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;

    // sizes
    int32_t n_mov = 0, n_fill = 0, n_fix = 0, n_all = 0, n_place = 0;
    int32_t n_nets = 0, n_pins = 0, n_pins_mov = 0;
    int32_t Mx = 0, My = 0;
    float xl = 0, yl = 0, xh = 0, yh = 0, bw = 0, bh = 0;
    float sqrt2_bw = 0, sqrt2_bh = 0;
    double mov_area = 0;
    ep_params prm{};

    // netlist (device)
    float2 *wh = nullptr;
    int32_t *net_start = nullptr, *pin_cell = nullptr;
    float2 *pin_off = nullptr, *pin_grad = nullptr;
    float2 *gw_cell = nullptr;  // per movable cell: sum of its pin partials
    int32_t *cell_pin_start = nullptr, *cell_pins = nullptr;
    BucketTable buckets{};
    int32_t wa_blocks = 0;
    std::vector<int32_t> net_perm;  // stored net k -> original net index
    std::vector<int32_t> cell_order;  // stored movable cell k -> caller's index (empty: same order)

    // optimiser state
    float2 *V[2] = {nullptr, nullptr}, *U[2] = {nullptr, nullptr}, *G[2] = {nullptr, nullptr};
    int cur = 0;

    // density + field
    float *map_field = nullptr, *map_mov = nullptr, *map_fix = nullptr;
    unsigned long long *acc_field = nullptr, *acc_mov = nullptr;  // fixed-point scatter targets
    float4 *sorted = nullptr;                        // packed cells in coarse-tile order
    int32_t *sort_count = nullptr, *sort_off = nullptr;  // kSortTiles each
    int32_t *dens_tab = nullptr;                     // kSortTiles + 1: density block table
    DensPlan dens{};
    float2 *exy = nullptr;                      // field (ex, ey) per bin
    float *tA = nullptr, *tB = nullptr;         // transform scratch, 2*Mx*My each
    void *split_mats = nullptr;                 // host handle: FP32 + split transform matrices
    uint32_t *field_max = nullptr;              // [4] running max |operand| for split FP16
    float *map_mean = nullptr;                  // the field map's mean area per bin, a constant of the
                                                // instance (subtracted before the solve)
    int32_t field_precision_built = -1;

    // reductions
    double *part = nullptr;   // partial sums, see part_* offsets
    double *fin = nullptr;    // second-level sums of the finalize blocks (32 x 10)
    unsigned *fin_ticket = nullptr;  // finalize blocks done (0 between launches)
    int32_t part_wa_off = 0, part_ovf_off = 0, part_cell_off = 0, part_trial_off = 0, part_total = 0;
    int32_t wa_part_blocks = 0, ovf_blocks = 0, cell_blocks = 0, trial_blocks = 0;
    DevScalars *d_scal = nullptr, *h_scal = nullptr;  // device / pinned host
    DevResult *d_res = nullptr, *h_res = nullptr;

    // optimiser bookkeeping (host)
    bool opt_ready = false;
    double alpha = 0, ak = 1, lambda = 0, lambda0 = 0, gamma = 0, prev_hpwl = 0;
    double last_hpwl = 0, last_tau = 0, last_wa = 0;
    int32_t iteration = 0;
    int64_t evals = 0;

    // graphs: one per parity of `cur` (host-controlled steps)
    hipGraphExec_t graph_exec[2] = {nullptr, nullptr};
    bool graph_built[2] = {false, false};
    // device-controlled runs: one trial + step decision per replay
    hipGraphExec_t ctl_exec = nullptr;
    DevCtl *d_ctl = nullptr, *h_ctl = nullptr;   // device / pinned host (two read-back slots)
    hipEvent_t ctl_read[2] = {nullptr, nullptr};  // each slot's read-back done
    ep_step_info *d_log = nullptr;               // accepted iterations of the run
    float *d_rec = nullptr;                      // step records: 17 floats each
    int32_t ctl_cap = 0;                         // entries of d_log and d_rec

    // learned step (MLP proposal)
    float *d_mlp_w = nullptr;
    bool mlp_loaded = false;
    std::vector<float> rec_feat, rec_target;   // accepted-step records (features, log10 step)
    // per-iteration history feeding the step-model features
    double h_alpha1 = 0, h_alpha2 = 0, h_prev_tau = 0, h_prev_hpwl = 0, h_ratio = 1;
    double h_gnorm = 0, h_dv = 0, h_dg = 0, h_sum_ratio = 1;
    int32_t h_evals = 1;
    int32_t learned_proposals = 0, learned_accepted = 0;
    int32_t model_skip = 0, model_backoff = 0;  // rejection backoff of the model
};

}  // namespace ep
