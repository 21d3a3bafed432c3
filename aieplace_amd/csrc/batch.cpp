// batch.cpp -- C ABI of the batched engine (include/eplace_batch.h): many
// independent placements advanced together, one captured hipGraph replay per
// group of Nesterov trials, with the optimiser decisions taken on the device.
//
// Per instance the algorithm is the one of ep_init_optimizer/ep_step (the
// reference's Placer flow, AIEplace.cpp:44-76, 993-1070); only the schedule of
// the work differs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/eplace_batch.h"
#include "ep_batch.h"
#include "ep_layout.h"

using namespace ep;

namespace {

constexpr int kTrialsPerLaunch = 4;   // Nesterov trials captured per graph replay
constexpr int kCellAlign = 64;        // instance cell ranges start on this boundary
constexpr int kDensCpt[3] = {2, 4, 8};
constexpr int kTuneReps = 5;

#define EPB_HIP(call)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(EP_EDEVICE, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),   \
                        __FILE__, __LINE__);                                                 \
    } while (0)

template <class T>
ep_status dalloc(T **p, size_t n)
{
    *p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void **)p, n * sizeof(T));
    if (e != hipSuccess)
        return fail(EP_ENOMEM, "ep_batch_create: hipMalloc(%zu bytes) failed: %s", n * sizeof(T),
                    hipGetErrorString(e));
    return EP_OK;
}

template <class T>
ep_status upload(T *dst, const std::vector<T> &src)
{
    if (src.empty()) return EP_OK;
    EPB_HIP(hipMemcpy(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    return EP_OK;
}

}  // namespace

struct ep_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;              // second captured branch (WA beside the maps and field)
    hipEvent_t fork = nullptr, join = nullptr;
    ep_params prm{};
    bool f16() const { return prm.precision == EP_PREC_FP16X3; }  // products scale by the map range
    int n = 0;
    std::vector<BInst> inst;
    std::vector<int32_t> n_fix;
    std::vector<std::vector<int32_t>> cell_order;  // per instance: stored movable cell -> caller's index
    std::vector<float2> init_pos;   // all cells, concatenated
    int64_t total_cells = 0;

    BInst *d_inst = nullptr;
    BState *d_st = nullptr;
    BRes *d_res = nullptr;
    float2 *wh = nullptr, *V[2] = {}, *U[2] = {}, *G[2] = {};
    int32_t *net_start = nullptr, *pin_cell = nullptr, *cps = nullptr, *cpins = nullptr;
    float2 *pin_off = nullptr, *pin_grad = nullptr;
    float *map_field = nullptr, *map_mov = nullptr, *map_fix = nullptr;
    float2 *exy = nullptr;
    float *tA = nullptr, *tB = nullptr, *map_mean = nullptr;
    unsigned long long *acc_field = nullptr, *acc_mov = nullptr;
    uint32_t *map_max = nullptr;
    float4 *sorted = nullptr;
    int32_t *sort_count = nullptr, *sort_off = nullptr, *dens_tab = nullptr;
    int32_t dens_wcap = 0;
    int32_t *pre_cell = nullptr, *pre_sort = nullptr, *pre_bin = nullptr, *pre_wa = nullptr;
    int32_t *pre_dens[3] = {};
    BBucket *buckets = nullptr;
    int n_buckets = 0;
    double *part_wa = nullptr, *part_ovf = nullptr, *part_cell = nullptr;
    int cell_blocks = 0, sort_blocks = 0, bin_blocks = 0, wa_blocks = 0, dens_blocks[3] = {};
    // Per-instance block counts and WA buckets (blk0 relative to the instance):
    // the block tables cover the instances in `active` only and are rebuilt
    // from these when finished instances are dropped (set_tables).
    std::vector<int32_t> nb_cell, nb_sort, nb_bin, nb_wa, nb_dens[3];
    std::vector<std::vector<BBucket>> inst_buckets;
    std::vector<char> active;
    int dens_variant = 1;
    BatchField *bf = nullptr;
    BState *h_st = nullptr;  // pinned: the last control state the host has read
    BState *h_rb[2] = {};     // pinned: read-back slots of two queued replays
    hipEvent_t rb_done[2] = {};
    hipGraphExec_t exec = nullptr;
    bool ready = false;

    BDev dev() const
    {
        BDev d{};
        d.inst = d_inst;
        d.st = d_st;
        d.res = d_res;
        d.ctl.lambda_eta = prm.lambda_eta;
        d.ctl.mu_max = prm.mu_max;
        d.ctl.mu_min = prm.mu_min;
        d.ctl.stop_overflow = prm.stop_overflow;
        d.ctl.gamma_schedule = prm.gamma_schedule;
        d.ctl.max_backtrack = prm.max_backtrack;
        d.ctl.max_iters = prm.max_iters;
        d.n_inst = n;
        d.wh = wh;
        for (int k = 0; k < 2; ++k) {
            d.V[k] = V[k];
            d.U[k] = U[k];
            d.G[k] = G[k];
        }
        d.net_start = net_start;
        d.pin_cell = pin_cell;
        d.pin_off = pin_off;
        d.pin_grad = pin_grad;
        d.cps = cps;
        d.cpins = cpins;
        d.map_field = map_field;
        d.map_mov = map_mov;
        d.map_fix = map_fix;
        d.acc_field = acc_field;
        d.acc_mov = acc_mov;
        d.exy = exy;
        d.map_mean = map_mean;
        d.map_max = map_max;
        d.sorted = sorted;
        d.sort_count = sort_count;
        d.sort_off = sort_off;
        d.dens_tab = dens_tab;
        d.dens_wcap = dens_wcap;
        d.pre_cell = pre_cell;
        d.pre_sort = pre_sort;
        d.pre_dens = pre_dens[dens_variant];
        d.pre_bin = pre_bin;
        d.pre_wa = pre_wa;
        d.buckets = buckets;
        d.n_buckets = n_buckets;
        d.part_wa = part_wa;
        d.part_ovf = part_ovf;
        d.part_cell = part_cell;
        return d;
    }

    int64_t active_cells() const
    {
        int64_t n_cells = 0;
        for (int i = 0; i < n; ++i)
            if (active[i]) n_cells += inst[i].n_place;
        return n_cells;
    }

    // One gradient evaluation as a captured fork: the tile sort, the density
    // scatter and the transform products on the side stream beside the WA
    // gradient (the captured stream of the longer branch decides how soon it
    // starts), as in the single-instance graph.
    hipError_t eval_forked(const BDev &d, int at, bool counted, hipStream_t s, int ctl_mode)
    {
        const int v = dens_variant;
        hipError_t e = hipEventRecord(fork, s);
        if (e == hipSuccess) e = hipStreamWaitEvent(side, fork, 0);
        if (e != hipSuccess) return e;
        bl_sort(d, at, counted, bin_blocks, sort_blocks, kDensCpt[v], side);
        bl_dens(d, bin_blocks, dens_blocks[v], kDensCpt[v], f16(), side);
        launch_field_batch(bf, prm.precision, side);
        if ((e = hipEventRecord(join, side)) != hipSuccess) return e;
        bl_wa(d, at, wa_blocks, s);
        if ((e = hipStreamWaitEvent(s, join, 0)) != hipSuccess) return e;
        bl_cell_grad(d, at, cell_blocks, s);
        bl_reduce_control(d, ctl_mode, s);
        return hipSuccess;
    }

    // One gradient evaluation of every unfinished instance (`counted`: the
    // preceding trial move counted the sort tiles), then the reductions and
    // the control step `ctl_mode` on them.
    void eval(const BDev &d, int at, bool counted, hipStream_t s, int ctl_mode)
    {
        bl_maps(d, at, counted, bin_blocks, sort_blocks, dens_blocks[dens_variant],
                kDensCpt[dens_variant], f16(), s);
        bl_wa(d, at, wa_blocks, s);
        launch_field_batch(bf, prm.precision, s);
        bl_cell_grad(d, at, cell_blocks, s);
        bl_reduce_control(d, ctl_mode, s);
    }
};

static void destroy_batch(ep_batch *b)
{
    if (!b) return;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    void *ptrs[] = {b->d_inst, b->d_st, b->d_res, b->wh, b->V[0], b->V[1], b->U[0], b->U[1],
                    b->G[0], b->G[1], b->net_start, b->pin_cell, b->cps, b->cpins, b->pin_off,
                    b->pin_grad, b->map_field, b->map_mov, b->map_fix, b->acc_field, b->acc_mov, b->exy, b->tA, b->tB,
                    b->map_mean, b->map_max, b->sorted, b->sort_count, b->sort_off, b->dens_tab, b->pre_cell,
                    b->pre_sort, b->pre_bin, b->pre_wa, b->pre_dens[0], b->pre_dens[1],
                    b->pre_dens[2], b->buckets, b->part_wa, b->part_ovf, b->part_cell};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    batch_field_free(b->bf);
    if (b->h_st) (void)hipHostFree(b->h_st);
    for (int k = 0; k < 2; ++k) {
        if (b->h_rb[k]) (void)hipHostFree(b->h_rb[k]);
        if (b->rb_done[k]) (void)hipEventDestroy(b->rb_done[k]);
    }
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->join) (void)hipEventDestroy(b->join);
    if (b->side) (void)hipStreamDestroy(b->side);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

// Initial placement and a fresh optimiser state on the device.
// Block tables, WA buckets and transform tile lists for the instances with
// active[i]; the others get no blocks (their kernels' blocks and tiles vanish
// from the launches instead of exiting one by one).  Every instance keeps its
// storage and block-local work, so results do not change.
static ep_status set_tables(ep_batch *b, const std::vector<char> &active)
{
    const int n = b->n;
    std::vector<int32_t> pc{0}, ps{0}, pb{0}, pw{0}, pd[3];
    for (int v = 0; v < 3; ++v) pd[v].push_back(0);
    std::vector<BBucket> bks;
    for (int i = 0; i < n; ++i) {
        const bool a = active[i] != 0;
        if (a)
            for (const BBucket &bk : b->inst_buckets[i]) {
                BBucket x = bk;
                x.blk0 += pw.back();
                bks.push_back(x);
            }
        pc.push_back(pc.back() + (a ? b->nb_cell[i] : 0));
        ps.push_back(ps.back() + (a ? b->nb_sort[i] : 0));
        pb.push_back(pb.back() + (a ? b->nb_bin[i] : 0));
        pw.push_back(pw.back() + (a ? b->nb_wa[i] : 0));
        for (int v = 0; v < 3; ++v) pd[v].push_back(pd[v].back() + (a ? b->nb_dens[v][i] : 0));
    }
    EPB_HIP(hipMemcpy(b->pre_cell, pc.data(), pc.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    EPB_HIP(hipMemcpy(b->pre_sort, ps.data(), ps.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    EPB_HIP(hipMemcpy(b->pre_bin, pb.data(), pb.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    EPB_HIP(hipMemcpy(b->pre_wa, pw.data(), pw.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    for (int v = 0; v < 3; ++v)
        EPB_HIP(hipMemcpy(b->pre_dens[v], pd[v].data(), pd[v].size() * sizeof(int32_t), hipMemcpyHostToDevice));
    if (!bks.empty())
        EPB_HIP(hipMemcpy(b->buckets, bks.data(), bks.size() * sizeof(BBucket), hipMemcpyHostToDevice));
    b->n_buckets = (int)bks.size();
    b->cell_blocks = pc.back();
    b->sort_blocks = ps.back();
    b->bin_blocks = pb.back();
    b->wa_blocks = pw.back();
    for (int v = 0; v < 3; ++v) b->dens_blocks[v] = pd[v].back();
    ep_status st = batch_field_set_active(b->bf, active);
    if (st != EP_OK) return st;
    b->active = active;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);  // captured with the old grids
    b->exec = nullptr;
    return EP_OK;
}

static ep_status reset_state(ep_batch *b)
{
    if (std::find(b->active.begin(), b->active.end(), 0) != b->active.end()) {
        ep_status st = set_tables(b, std::vector<char>(b->n, 1));
        if (st != EP_OK) return st;
    }
    const size_t nc = (size_t)b->total_cells;
    for (int k = 0; k < 2; ++k) {
        EPB_HIP(hipMemcpy(b->V[k], b->init_pos.data(), nc * sizeof(float2), hipMemcpyHostToDevice));
        EPB_HIP(hipMemcpy(b->U[k], b->init_pos.data(), nc * sizeof(float2), hipMemcpyHostToDevice));
        EPB_HIP(hipMemset(b->G[k], 0, nc * sizeof(float2)));
    }
    std::vector<BState> st(b->n);
    std::memset(st.data(), 0, st.size() * sizeof(BState));
    EPB_HIP(hipMemcpy(b->d_st, st.data(), st.size() * sizeof(BState), hipMemcpyHostToDevice));
    std::memcpy(b->h_st, st.data(), st.size() * sizeof(BState));
    b->ready = false;
    return EP_OK;
}

static ep_status check_errors(ep_batch *b)
{
    for (int i = 0; i < b->n; ++i)
        if (b->h_st[i].error)
            return fail(EP_ENAN, "ep_batch: instance %d has a non-finite gradient at iteration %d", i,
                        b->h_st[i].iteration);
    return EP_OK;
}

static ep_status pull_state(ep_batch *b)
{
    EPB_HIP(hipMemcpyAsync(b->h_st, b->d_st, sizeof(BState) * b->n, hipMemcpyDeviceToHost, b->stream));
    EPB_HIP(hipStreamSynchronize(b->stream));
    return EP_OK;
}

// Picks the density-scatter tile (cells per thread) by timing each variant on
// the current placement, then the transform products' K step.
static ep_status autotune(ep_batch *b)
{
    hipEvent_t e0, e1;
    EPB_HIP(hipEventCreate(&e0));
    EPB_HIP(hipEventCreate(&e1));
    float best = 0.f;
    int best_v = 1;
    for (int v = 0; v < 3; ++v) {
        b->dens_variant = v;
        const BDev d = b->dev();
        bl_maps(d, kAtCur, false, b->bin_blocks, b->sort_blocks, b->dens_blocks[v], kDensCpt[v], b->f16(),
                b->stream);
        (void)hipEventRecord(e0, b->stream);
        for (int r = 0; r < kTuneReps; ++r)
            bl_maps(d, kAtCur, false, b->bin_blocks, b->sort_blocks, b->dens_blocks[v], kDensCpt[v], b->f16(),
                    b->stream);
        (void)hipEventRecord(e1, b->stream);
        hipError_t e = hipEventSynchronize(e1);
        if (e != hipSuccess) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            return fail(EP_EDEVICE, "ep_batch autotune: %s", hipGetErrorString(e));
        }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (v == 0 || ms < best) {
            best = ms;
            best_v = v;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->dens_variant = best_v;
    // the maps now hold the initial placement's density: tune the transform K step
    return batch_field_autotune(b->bf, kTuneReps, b->stream);
}

extern "C" ep_status ep_batch_create(const ep_netlist *const *nls, const int32_t *bins, int32_t n,
                                     const ep_params *p, ep_batch **out)
{
    if (!out) return fail(EP_EINVAL, "ep_batch_create: null output");
    *out = nullptr;
    if (!nls || !bins || !p) return fail(EP_EINVAL, "ep_batch_create: null argument");
    if (n < 1 || n > kBatchMaxInstances)
        return fail(EP_EINVAL, "ep_batch_create: instance count %d outside [1, %d]", n, kBatchMaxInstances);
    if (p->learned_step) return fail(EP_ENOTSUP, "ep_batch_create: learned step is single-instance only");
    for (int i = 0; i < n; ++i)
        if (bins[i] < 32 || bins[i] % 32)
            return fail(EP_EINVAL, "ep_batch_create: instance %d: bins %d must be a multiple of 32", i, bins[i]);
    ep_batch *b = new (std::nothrow) ep_batch();
    if (!b) return fail(EP_ENOMEM, "ep_batch_create: host allocation failed");
    b->prm = *p;
    b->n = n;

    // ---- concatenated host arrays
    std::vector<float2> wh, pos;
    std::vector<int32_t> net_start{0}, pin_cell, cps, cpins;
    std::vector<float2> pin_off;
    std::vector<float> fixmap;
    std::vector<BBucket> bks;
    std::vector<float> means;  // per instance: constant mean area per bin of the field map
    std::vector<int32_t> pre_cell{0}, pre_sort{0}, pre_bin{0}, pre_wa{0}, pre_dens[3];
    for (int v = 0; v < 3; ++v) pre_dens[v].push_back(0);
    int64_t cell_off = 0, map_off = 0, pin_base = 0, net_base = 0, cpin_base = 0, wa_base = 0;
    for (int i = 0; i < n; ++i) {
        ep_params pi = *p;
        pi.bins_x = pi.bins_y = bins[i];
        HostLayout L;
        ep_status st = build_layout(nls[i], &pi, L);
        if (st != EP_OK) {
            const std::string why = ep_last_error();
            destroy_batch(b);
            return fail(st, "ep_batch_create: instance %d: %s", i, why.c_str());
        }
        BInst I{};
        I.cell_off = (int32_t)cell_off;
        I.n_mov = L.n_mov;
        I.n_place = L.n_place;
        I.n_all = L.n_all;
        I.map_off = (int32_t)map_off;
        I.M = bins[i];
        I.xl = L.xl; I.yl = L.yl; I.xh = L.xh; I.yh = L.yh;
        I.bw = L.bw; I.bh = L.bh; I.sqrt2_bw = L.sqrt2_bw; I.sqrt2_bh = L.sqrt2_bh;
        I.rho_t = p->target_density;
        I.cap = p->target_density * L.bw * L.bh;
        I.mov_area = L.mov_area;
        I.dens_unit = (double)L.bw * (double)L.bh / kDensFixUnits;
        I.hpwl_ref = L.hpwl_ref;
        I.init_disp = (double)p->init_disp_bins * L.bw;
        I.gamma_base = (double)p->gamma_bins * L.bw;
        const DensPlan dp = make_dens_plan(L.Mx, L.My, L.bw, L.bh, L.max_sw, L.max_sh, 4);
        I.dens_level = dp.level;
        I.dens_ext = dp.ext;
        b->dens_wcap = std::max(b->dens_wcap, dp.wcap);
        b->inst.push_back(I);
        b->n_fix.push_back(L.n_fix);
        b->cell_order.push_back(L.cell_order);
        means.push_back(field_map_mean(L, p->target_density));
        const int64_t span = (L.n_all + kCellAlign - 1) / kCellAlign * kCellAlign;
        if (cell_off + span > INT32_MAX || map_off + (int64_t)bins[i] * bins[i] > INT32_MAX / 2 ||
            pin_base + L.n_pins > INT32_MAX) {
            destroy_batch(b);
            return fail(EP_EINVAL, "ep_batch_create: batch too large for 32-bit indexing");
        }
        // cells (padding cells are never read; they get the die corner)
        wh.insert(wh.end(), L.wh.begin(), L.wh.end());
        pos.insert(pos.end(), L.pos.begin(), L.pos.end());
        wh.resize(cell_off + span, make_float2(1.f, 1.f));
        pos.resize(cell_off + span, make_float2(L.xl, L.yl));
        // nets and pins with global indices
        for (int k = 1; k <= L.n_nets; ++k) net_start.push_back((int32_t)(pin_base + L.net_start[k]));
        for (int q = 0; q < L.n_pins; ++q) {
            pin_cell.push_back((int32_t)(cell_off + L.pin_cell[q]));
            pin_off.push_back(L.pin_off[q]);
        }
        b->inst_buckets.emplace_back();
        for (int k = 0; k < L.buckets.n; ++k) {
            const Bucket &bk = L.buckets.b[k];
            bks.push_back({i, bk.lanes, (int32_t)(net_base + bk.net0), (int32_t)(net_base + bk.net1),
                           (int32_t)(wa_base + bk.blk0), bk.deg, (int32_t)(pin_base + bk.pin0)});
            b->inst_buckets.back().push_back({i, bk.lanes, (int32_t)(net_base + bk.net0),
                                              (int32_t)(net_base + bk.net1), bk.blk0, bk.deg,
                                              (int32_t)(pin_base + bk.pin0)});
        }
        // cell -> pin CSR over all cells (empty for fillers, fixed and padding)
        for (int64_t k = 0; k < span; ++k)
            cps.push_back((int32_t)(cpin_base + L.cps[std::min<int64_t>(k, L.n_mov)]));
        for (int j = 0; j < L.n_pins_mov; ++j) cpins.push_back((int32_t)(pin_base + L.cpins[j]));
        fixmap.insert(fixmap.end(), L.fixmap.begin(), L.fixmap.end());
        // block tables
        pre_cell.push_back(pre_cell.back() + (int32_t)((L.n_place + kBlock - 1) / kBlock));
        pre_sort.push_back(pre_sort.back() + (int32_t)((L.n_place + kBlock * kSortCellsPerThread - 1) /
                                                       (kBlock * kSortCellsPerThread)));
        for (int v = 0; v < 3; ++v) {
            DensPlan dv{};
            dv.level = I.dens_level;
            dv.cpt = kDensCpt[v];
            pre_dens[v].push_back(pre_dens[v].back() + dens_grid(dv, L.n_place));
        }
        pre_bin.push_back(pre_bin.back() + (int32_t)(((int64_t)bins[i] * bins[i] + kBinsPerBlock - 1) / kBinsPerBlock));
        wa_base += L.wa_blocks;
        pre_wa.push_back((int32_t)wa_base);
        const size_t ni = pre_cell.size() - 1;
        b->nb_cell.push_back(pre_cell[ni] - pre_cell[ni - 1]);
        b->nb_sort.push_back(pre_sort[ni] - pre_sort[ni - 1]);
        b->nb_bin.push_back(pre_bin[ni] - pre_bin[ni - 1]);
        b->nb_wa.push_back(pre_wa[ni] - pre_wa[ni - 1]);
        for (int v = 0; v < 3; ++v) b->nb_dens[v].push_back(pre_dens[v][ni] - pre_dens[v][ni - 1]);
        cell_off += span;
        map_off += (int64_t)bins[i] * bins[i];
        pin_base += L.n_pins;
        net_base += L.n_nets;
        cpin_base += L.n_pins_mov;
    }
    cps.push_back((int32_t)cpin_base);
    b->total_cells = cell_off;
    b->active.assign(n, 1);
    b->init_pos = pos;
    b->n_buckets = (int)bks.size();
    b->cell_blocks = pre_cell.back();
    b->sort_blocks = pre_sort.back();
    b->bin_blocks = pre_bin.back();
    b->wa_blocks = (int)wa_base;
    for (int v = 0; v < 3; ++v) b->dens_blocks[v] = pre_dens[v].back();

    // ---- device
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        destroy_batch(b);
        return fail(EP_EDEVICE, "ep_batch_create: no HIP device available");
    }
#define EPB_TRY(x)                \
    do {                          \
        ep_status s_ = (x);       \
        if (s_ != EP_OK) {        \
            destroy_batch(b);     \
            return s_;            \
        }                         \
    } while (0)
    (void)hipGetDevice(&b->device);
    if (hipStreamCreate(&b->stream) != hipSuccess || hipStreamCreate(&b->side) != hipSuccess ||
        hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->join, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->rb_done[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->rb_done[1], hipEventDisableTiming) != hipSuccess) {
        destroy_batch(b);
        return fail(EP_EDEVICE, "ep_batch_create: stream/event creation failed");
    }
    const size_t nc = (size_t)cell_off, nb = (size_t)map_off;
    EPB_TRY(dalloc(&b->d_inst, n));
    EPB_TRY(dalloc(&b->d_st, n));
    EPB_TRY(dalloc(&b->d_res, n));
    EPB_TRY(dalloc(&b->wh, nc));
    for (int k = 0; k < 2; ++k) {
        EPB_TRY(dalloc(&b->V[k], nc));
        EPB_TRY(dalloc(&b->U[k], nc));
        EPB_TRY(dalloc(&b->G[k], nc));
    }
    EPB_TRY(dalloc(&b->net_start, net_start.size()));
    EPB_TRY(dalloc(&b->pin_cell, pin_cell.size()));
    EPB_TRY(dalloc(&b->pin_off, pin_off.size()));
    EPB_TRY(dalloc(&b->pin_grad, pin_cell.size()));
    EPB_TRY(dalloc(&b->cps, cps.size()));
    EPB_TRY(dalloc(&b->cpins, cpins.size()));
    EPB_TRY(dalloc(&b->map_field, nb));
    EPB_TRY(dalloc(&b->map_mov, nb));
    EPB_TRY(dalloc(&b->acc_field, nb));
    EPB_TRY(dalloc(&b->acc_mov, nb));
    // cleared once here; afterwards every overflow pass clears what its scatter added
    if (hipMemset(b->acc_field, 0, sizeof(unsigned long long) * nb) != hipSuccess ||
        hipMemset(b->acc_mov, 0, sizeof(unsigned long long) * nb) != hipSuccess) {
        destroy_batch(b);
        return fail(EP_EDEVICE, "ep_batch_create: hipMemset failed");
    }
    EPB_TRY(dalloc(&b->map_fix, nb));
    EPB_TRY(dalloc(&b->exy, nb));
    EPB_TRY(dalloc(&b->tA, 2 * nb));
    EPB_TRY(dalloc(&b->tB, 2 * nb));
    EPB_TRY(dalloc(&b->map_mean, n));
    EPB_TRY(dalloc(&b->map_max, n));
    EPB_TRY(dalloc(&b->sorted, nc));
    EPB_TRY(dalloc(&b->sort_count, (size_t)n * kSortTiles));
    EPB_TRY(dalloc(&b->sort_off, (size_t)n * kSortTiles));
    EPB_TRY(dalloc(&b->dens_tab, (size_t)n * (kSortTiles + 1)));
    EPB_TRY(dalloc(&b->pre_cell, n + 1));
    EPB_TRY(dalloc(&b->pre_sort, n + 1));
    EPB_TRY(dalloc(&b->pre_bin, n + 1));
    EPB_TRY(dalloc(&b->pre_wa, n + 1));
    for (int v = 0; v < 3; ++v) EPB_TRY(dalloc(&b->pre_dens[v], n + 1));
    EPB_TRY(dalloc(&b->buckets, bks.size()));
    EPB_TRY(dalloc(&b->part_wa, 2 * (size_t)std::max(1, b->wa_blocks)));
    EPB_TRY(dalloc(&b->part_ovf, 4 * (size_t)b->bin_blocks));
    EPB_TRY(dalloc(&b->part_cell, 7 * (size_t)b->cell_blocks));
    if (hipHostMalloc((void **)&b->h_st, sizeof(BState) * n, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&b->h_rb[0], sizeof(BState) * n, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&b->h_rb[1], sizeof(BState) * n, hipHostMallocDefault) != hipSuccess) {
        destroy_batch(b);
        return fail(EP_ENOMEM, "ep_batch_create: pinned host allocation failed");
    }
    EPB_TRY(upload(b->d_inst, b->inst));
    EPB_TRY(upload(b->map_mean, means));
    EPB_TRY(upload(b->wh, wh));
    EPB_TRY(upload(b->net_start, net_start));
    EPB_TRY(upload(b->pin_cell, pin_cell));
    EPB_TRY(upload(b->pin_off, pin_off));
    EPB_TRY(upload(b->cps, cps));
    EPB_TRY(upload(b->cpins, cpins));
    EPB_TRY(upload(b->map_fix, fixmap));
    EPB_TRY(upload(b->pre_cell, pre_cell));
    EPB_TRY(upload(b->pre_sort, pre_sort));
    EPB_TRY(upload(b->pre_bin, pre_bin));
    EPB_TRY(upload(b->pre_wa, pre_wa));
    for (int v = 0; v < 3; ++v) EPB_TRY(upload(b->pre_dens[v], pre_dens[v]));
    EPB_TRY(upload(b->buckets, bks));
    if (hipMemset(b->pin_grad, 0, sizeof(float2) * std::max<size_t>(1, pin_cell.size())) != hipSuccess ||
        hipMemset(b->exy, 0, sizeof(float2) * nb) != hipSuccess ||
        hipMemset(b->sort_count, 0, sizeof(int32_t) * n * kSortTiles) != hipSuccess ||
        hipMemset(b->part_wa, 0, sizeof(double) * 2 * std::max(1, b->wa_blocks)) != hipSuccess ||
        hipMemset(b->part_ovf, 0, sizeof(double) * 4 * b->bin_blocks) != hipSuccess ||
        hipMemset(b->part_cell, 0, sizeof(double) * 7 * b->cell_blocks) != hipSuccess ||
        hipMemset(b->d_res, 0, sizeof(BRes) * n) != hipSuccess) {
        destroy_batch(b);
        return fail(EP_EDEVICE, "ep_batch_create: hipMemset failed");
    }
    std::vector<BatchFieldInst> fi(n);
    for (int i = 0; i < n; ++i) {
        const BInst &I = b->inst[i];
        fi[i].M = I.M;
        fi[i].bw = I.bw;
        fi[i].bh = I.bh;
        fi[i].map_field = b->map_field + I.map_off;
        fi[i].tA = b->tA + 2 * (size_t)I.map_off;
        fi[i].tB = b->tB + 2 * (size_t)I.map_off;
        fi[i].exy = b->exy + I.map_off;
        fi[i].map_max = b->map_max + i;
        fi[i].map_mean = b->map_mean + i;
        fi[i].skip = &b->d_st[i].done;
    }
    EPB_TRY(batch_field_init(b->bf, fi));
    EPB_TRY(reset_state(b));
    EPB_TRY(autotune(b));
#undef EPB_TRY
    if (hipDeviceSynchronize() != hipSuccess) {
        destroy_batch(b);
        return fail(EP_EDEVICE, "ep_batch_create: device synchronisation failed");
    }
    *out = b;
    return EP_OK;
}

extern "C" ep_status ep_batch_destroy(ep_batch *b)
{
    if (!b) return fail(EP_EINVAL, "ep_batch_destroy: null batch");
    (void)hipStreamSynchronize(b->stream);
    destroy_batch(b);
    return EP_OK;
}

extern "C" ep_status ep_batch_reset(ep_batch *b)
{
    if (!b) return fail(EP_EINVAL, "ep_batch_reset: null batch");
    EPB_HIP(hipStreamSynchronize(b->stream));
    return reset_state(b);
}

// Initial gamma, lambda_0 and the probe step alpha_0 of every instance
// (ep_init_optimizer, batched).
extern "C" ep_status ep_batch_init(ep_batch *b)
{
    if (!b) return fail(EP_EINVAL, "ep_batch_init: null batch");
    if (b->ready) return EP_OK;
    const BDev d = b->dev();
    hipStream_t s = b->stream;
    bl_move(d, kMoveInitU, b->sort_blocks, s);
    bl_maps(d, kAtCur, false, b->bin_blocks, b->sort_blocks, b->dens_blocks[b->dens_variant],
            kDensCpt[b->dens_variant], b->f16(), s);
    bl_reduce_control(d, kCtlGamma, s);
    // gradient pieces at v_0 with lambda = 0: |gW|_1 and |F|_1
    bl_wa(d, kAtCur, b->wa_blocks, s);
    launch_field_batch(b->bf, b->prm.precision, s);
    bl_cell_grad(d, kAtCur, b->cell_blocks, s);
    bl_reduce_control(d, kCtlLambda, s);
    // preconditioned gradient with lambda_0, then the probe displacement
    bl_cell_grad(d, kAtCur, b->cell_blocks, s);
    bl_reduce_control(d, kCtlProbe, s);
    bl_move(d, kMoveProbe, b->sort_blocks, s);
    b->eval(d, kAtTrial, false, s, kCtlAlpha);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(EP_EDEVICE, "ep_batch_init: launch failed: %s", hipGetErrorString(e));
    ep_status st = pull_state(b);
    if (st != EP_OK) return st;
    if ((st = check_errors(b)) != EP_OK) return st;
    b->ready = true;
    return EP_OK;
}

// The graph holds device work only (kernels, the fork/join events): the host
// reads the control state with its own copy after each replay (ep_batch_run).
static ep_status build_graph(ep_batch *b)
{
    const BDev d = b->dev();
    hipGraph_t g = nullptr;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
    EPB_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int t = 0; t < kTrialsPerLaunch && e == hipSuccess; ++t) {
        bl_move(d, kMoveTrial, b->sort_blocks, b->stream);
        if (b->active_cells() >= fork_min_cells(b->prm))
            e = b->eval_forked(d, kAtTrial, true, b->stream, kCtlIterate);
        else b->eval(d, kAtTrial, true, b->stream, kCtlIterate);
    }
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) {  // leave capture mode before reporting
        (void)hipStreamEndCapture(b->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        return fail(EP_EDEVICE, "ep_batch: graph capture failed: %s", hipGetErrorString(e));
    }
    EPB_HIP(hipStreamEndCapture(b->stream, &g));
    e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        b->exec = nullptr;
        return fail(EP_EDEVICE, "ep_batch: hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    return EP_OK;
}

extern "C" ep_status ep_batch_run(ep_batch *b, int32_t max_launches, ep_batch_run_info *info)
{
    if (!b) return fail(EP_EINVAL, "ep_batch_run: null batch");
    if (max_launches < 1) return fail(EP_EINVAL, "ep_batch_run: max_launches must be >= 1");
    const auto t0 = std::chrono::steady_clock::now();
    ep_status st;
    if (!b->ready && (st = ep_batch_init(b)) != EP_OK) return st;
    if (!b->exec && (st = build_graph(b)) != EP_OK) return st;
    int launches = 0;
    auto all_done = [&]() {
        for (int i = 0; i < b->n; ++i)
            if (!b->h_st[i].done) return false;
        return true;
    };
    auto count = [](const std::vector<char> &a) { return (int)std::count(a.begin(), a.end(), 1); };
    const size_t st_bytes = sizeof(BState) * b->n;
    // Two replays in flight: each replay's control state is read back into
    // slot h_rb[k], and the host waits for one replay's state while the next
    // replay runs.  Decisions therefore see the state one replay late: a
    // replay queued after every instance has stopped returns at once (stopped
    // instances skip their work), and the block tables are rebuilt only after
    // the stream has drained.
    int slot = 0, last = -1;
    bool pending = false;  // h_rb[slot ^ 1] holds a queued read-back not yet checked
    for (;;) {
        const bool launch = launches < max_launches && !all_done();
        if (launch) {
            EPB_HIP(hipGraphLaunch(b->exec, b->stream));
            EPB_HIP(hipMemcpyAsync(b->h_rb[slot], b->d_st, st_bytes, hipMemcpyDeviceToHost, b->stream));
            EPB_HIP(hipEventRecord(b->rb_done[slot], b->stream));
            ++launches;
            last = slot;
        }
        if (pending) {
            EPB_HIP(hipEventSynchronize(b->rb_done[slot ^ 1]));
            std::memcpy(b->h_st, b->h_rb[slot ^ 1], st_bytes);
            if ((st = check_errors(b)) != EP_OK) {
                (void)hipStreamSynchronize(b->stream);
                return st;
            }
            if (all_done()) break;
            // once half of the instances in the launches have finished, drop them
            std::vector<char> running(b->n, 0);
            for (int i = 0; i < b->n; ++i) running[i] = b->active[i] && !b->h_st[i].done;
            const int r = count(running);
            if (r > 0 && 2 * r <= count(b->active)) {
                EPB_HIP(hipStreamSynchronize(b->stream));  // the queued replay reads the tables
                if ((st = set_tables(b, running)) != EP_OK || (st = build_graph(b)) != EP_OK) return st;
            }
        }
        if (!launch) break;
        pending = true;
        slot ^= 1;
    }
    EPB_HIP(hipStreamSynchronize(b->stream));
    if (last >= 0) std::memcpy(b->h_st, b->h_rb[last], st_bytes);  // the last replay's state
    if ((st = check_errors(b)) != EP_OK) return st;
    const auto t1 = std::chrono::steady_clock::now();
    if (info) {
        info->instances = b->n;
        info->finished = 0;
        info->instance_iterations = 0;
        info->evals = 0;
        for (int i = 0; i < b->n; ++i) {
            info->finished += b->h_st[i].done ? 1 : 0;
            info->instance_iterations += b->h_st[i].iteration;
            info->evals += b->h_st[i].evals;
        }
        info->launches = launches;
        info->seconds = std::chrono::duration<double>(t1 - t0).count();
        info->density_cells_per_thread = kDensCpt[b->dens_variant];
        info->transform_k_step = batch_field_k_step(b->bf, b->prm.precision);
    }
    return EP_OK;
}

extern "C" ep_status ep_batch_results(ep_batch *b, int32_t *iterations, int64_t *evals, double *hpwl,
                                      double *overflow)
{
    if (!b) return fail(EP_EINVAL, "ep_batch_results: null batch");
    ep_status st = pull_state(b);
    if (st != EP_OK) return st;
    for (int i = 0; i < b->n; ++i) {
        if (iterations) iterations[i] = b->h_st[i].iteration;
        if (evals) evals[i] = b->h_st[i].evals;
        if (hpwl) hpwl[i] = b->h_st[i].last_hpwl;
        if (overflow) overflow[i] = b->h_st[i].last_tau;
    }
    return EP_OK;
}

extern "C" ep_status ep_batch_sizes(ep_batch *b, int32_t i, int32_t *n_mov, int32_t *n_fill, int32_t *n_fix)
{
    if (!b || i < 0 || i >= b->n) return fail(EP_EINVAL, "ep_batch_sizes: bad batch or instance");
    if (n_mov) *n_mov = b->inst[i].n_mov;
    if (n_fill) *n_fill = b->inst[i].n_place - b->inst[i].n_mov;
    if (n_fix) *n_fix = b->n_fix[i];
    return EP_OK;
}

static ep_status pull_pairs(ep_batch *b, int32_t i, bool grad, float *x, float *y)
{
    if (!b || i < 0 || i >= b->n || !x || !y) return fail(EP_EINVAL, "ep_batch: bad argument");
    ep_status st = pull_state(b);
    if (st != EP_OK) return st;
    const BInst &I = b->inst[i];
    const int cur = b->h_st[i].cur;
    const float2 *src = (grad ? b->G[cur] : b->V[cur]) + I.cell_off;
    std::vector<float2> h(I.n_place);
    EPB_HIP(hipMemcpy(h.data(), src, sizeof(float2) * I.n_place, hipMemcpyDeviceToHost));
    for (int k = 0; k < I.n_place; ++k) {
        const int32_t o = caller_cell(b->cell_order[i], k);
        x[o] = h[k].x;
        y[o] = h[k].y;
    }
    return EP_OK;
}

extern "C" ep_status ep_batch_positions(ep_batch *b, int32_t i, float *x, float *y)
{
    return pull_pairs(b, i, false, x, y);
}

extern "C" ep_status ep_batch_gradient(ep_batch *b, int32_t i, float *gx, float *gy)
{
    return pull_pairs(b, i, true, gx, gy);
}

extern "C" ep_status ep_batch_state(ep_batch *b, int32_t i, double *gamma, double *lambda, double *alpha)
{
    if (!b || i < 0 || i >= b->n) return fail(EP_EINVAL, "ep_batch_state: bad batch or instance");
    ep_status st = pull_state(b);
    if (st != EP_OK) return st;
    if (gamma) *gamma = b->h_st[i].gamma;
    if (lambda) *lambda = b->h_st[i].lambda;
    if (alpha) *alpha = b->h_st[i].alpha;
    return EP_OK;
}
