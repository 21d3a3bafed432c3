// k_batch.hip -- kernels of the batched engine: the single-instance gradient
// evaluation (k_place.hip) re-tiled so that one launch covers every unfinished
// instance, plus the per-instance FP64 reduction and the device-side Nesterov
// control (backtracking decision, lambda/gamma schedules) that replaces the
// host optimiser loop of ep_step.
//
// Blocks map to instances through prefix tables; per-block partial sums are
// indexed by the global block id and reduced per instance in a fixed order.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>

#include "ep_batch.h"
#include "ep_ctx.h"
#include "ep_device.h"
#include "ep_sched.h"

namespace ep {

namespace {

using namespace dev;

// Instance owning global block b: the largest i with pre[i] <= b.
__device__ inline int find_inst(const int32_t *__restrict__ pre, int n, int b)
{
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (pre[mid] <= b) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ inline GeomArgs inst_geom(const BInst &I)
{
    GeomArgs g;
    g.v[0] = I.xl; g.v[1] = I.yl; g.v[2] = I.xh; g.v[3] = I.yh;
    g.v[4] = I.bw; g.v[5] = I.bh; g.v[6] = I.sqrt2_bw; g.v[7] = I.sqrt2_bh;
    return g;
}

template <class T>
__device__ inline T pick(T p0, T p1, int k)
{
    return k ? p1 : p0;
}

// Position array an evaluation reads: V[cur] or the trial point V[cur ^ 1].
__device__ inline const float2 *eval_pos(const BDev &d, int cur, int at)
{
    return pick<const float2 *>(d.V[0], d.V[1], cur ^ at);
}

// ------------------------------------------------------------------ moves

// kMoveTrial: u' = clamp(v - alpha g), v' = clamp(u' + coef (u' - u)) into the
// other parity, counting v' into the instance's sort tiles (the evaluation's
// sort then skips its counting pass); kMoveProbe: v_-1 = clamp(v - eps g);
// kMoveInitU: u = v.  Blocks take kBlock * kSortCellsPerThread cells (pre_sort).
__global__ __launch_bounds__(kBlock) void kb_move(BDev d, int mode)
{
    __shared__ int hist[kSortTiles];
    const int i = find_inst(d.pre_sort, d.n_inst, blockIdx.x);
    const BState *S = d.st + i;
    if (S->done) return;
    const BInst &I = d.inst[i];
    const bool count = mode == kMoveTrial;
    if (count) {
        for (int t = threadIdx.x; t < kSortTiles; t += kBlock) hist[t] = 0;
        __syncthreads();
    }
    const int a = S->cur;
    const GeomArgs g = inst_geom(I);
    const int base = (blockIdx.x - d.pre_sort[i]) * kBlock * kSortCellsPerThread;
    if (mode == kMoveTrial) {
        // All of the thread's cells are loaded before the first is moved, so
        // the loads overlap; a cell past n_place loads the instance's last
        // cell (value unused).
        const float alpha = S->f_alpha, coef = S->f_coef;
        const float2 *V = pick<const float2 *>(d.V[0], d.V[1], a);
        const float2 *U = pick<const float2 *>(d.U[0], d.U[1], a);
        const float2 *G = pick<const float2 *>(d.G[0], d.G[1], a);
        float2 s[kSortCellsPerThread], vv[kSortCellsPerThread], uu[kSortCellsPerThread],
            gg[kSortCellsPerThread];
#pragma unroll
        for (int k = 0; k < kSortCellsPerThread; ++k) {
            const int c = I.cell_off + min(base + k * kBlock + (int)threadIdx.x, I.n_place - 1);
            s[k] = d.wh[c];
            vv[k] = V[c];
            uu[k] = U[c];
            gg[k] = G[c];
        }
#pragma unroll
        for (int k = 0; k < kSortCellsPerThread; ++k) {
            const int lc = base + k * kBlock + threadIdx.x;
            const float2 un =
                clamp_cell(make_float2(vv[k].x - alpha * gg[k].x, vv[k].y - alpha * gg[k].y), s[k], g);
            const float2 vn = clamp_cell(
                make_float2(un.x + coef * (un.x - uu[k].x), un.y + coef * (un.y - uu[k].y)), s[k], g);
            if (lc < I.n_place) {
                const int c = I.cell_off + lc;
                pick(d.U[0], d.U[1], a ^ 1)[c] = un;
                pick(d.V[0], d.V[1], a ^ 1)[c] = vn;
                atomicAdd(&hist[cell_key(vn, s[k], g, I.M, I.M)], 1);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < kSortCellsPerThread && mode != kMoveTrial; ++k) {
        const int lc = base + k * kBlock + threadIdx.x;
        if (lc >= I.n_place) continue;
        const int c = I.cell_off + lc;
        const float2 s = d.wh[c];
        const float2 vv = pick<const float2 *>(d.V[0], d.V[1], a)[c];
        if (mode == kMoveProbe) {
            const float eps = S->f_eps;
            const float2 gg = pick<const float2 *>(d.G[0], d.G[1], a)[c];
            pick(d.V[0], d.V[1], a ^ 1)[c] =
                clamp_cell(make_float2(vv.x - eps * gg.x, vv.y - eps * gg.y), s, g);
        } else {
            pick(d.U[0], d.U[1], a)[c] = vv;
        }
    }
    if (count) {
        __syncthreads();
        int32_t *cnt = d.sort_count + (size_t)i * kSortTiles;
        for (int t = threadIdx.x; t < kSortTiles; t += kBlock)
            if (hist[t]) atomicAdd(&cnt[t], hist[t]);
    }
}

// ------------------------------------------------------------------ maps

__global__ __launch_bounds__(kBlock) void kb_sort_count(BDev d, int at)
{
    __shared__ int hist[kSortTiles];
    const int i = find_inst(d.pre_sort, d.n_inst, blockIdx.x);
    const BState *S = d.st + i;
    if (S->done) return;
    const BInst &I = d.inst[i];
    const float2 *pos = eval_pos(d, S->cur, at);
    const GeomArgs g = inst_geom(I);
    for (int t = threadIdx.x; t < kSortTiles; t += kBlock) hist[t] = 0;
    __syncthreads();
    const int base = (blockIdx.x - d.pre_sort[i]) * kBlock * kSortCellsPerThread;
#pragma unroll
    for (int k = 0; k < kSortCellsPerThread; ++k) {
        const int idx = base + k * kBlock + threadIdx.x;
        if (idx < I.n_place) {
            const int c = I.cell_off + idx;
            atomicAdd(&hist[cell_key(pos[c], d.wh[c], g, I.M, I.M)], 1);
        }
    }
    __syncthreads();
    int32_t *count = d.sort_count + (size_t)i * kSortTiles;
    for (int t = threadIdx.x; t < kSortTiles; t += kBlock)
        if (hist[t]) atomicAdd(&count[t], hist[t]);
}

// One block per instance: exclusive scan of the tile counts (clears the
// counts) and the instance's density block table for K cells per block.
__global__ __launch_bounds__(kSortScanThreads) void kb_sort_scan(BDev d, int K)
{
    const int i = blockIdx.x;
    if (d.st[i].done) return;
    sort_scan_block(d.sort_count + (size_t)i * kSortTiles, d.sort_off + (size_t)i * kSortTiles,
                    d.dens_tab + (size_t)i * (kSortTiles + 1), d.inst[i].dens_level, K);
}

__global__ __launch_bounds__(kBlock) void kb_sort_scatter(BDev d, int at)
{
    __shared__ int hist[kSortTiles];
    __shared__ int base_of[kSortTiles];
    const int i = find_inst(d.pre_sort, d.n_inst, blockIdx.x);
    const BState *S = d.st + i;
    if (S->done) return;
    const BInst &I = d.inst[i];
    const float2 *pos = eval_pos(d, S->cur, at);
    const GeomArgs g = inst_geom(I);
    for (int t = threadIdx.x; t < kSortTiles; t += kBlock) hist[t] = 0;
    __syncthreads();
    const int base = (blockIdx.x - d.pre_sort[i]) * kBlock * kSortCellsPerThread;
    int key[kSortCellsPerThread], rank[kSortCellsPerThread];
    float4 cell[kSortCellsPerThread];
    // All of the thread's cells are loaded before the first tile count; a
    // cell past n_place loads the instance's last cell (value unused).
    float2 p[kSortCellsPerThread], s[kSortCellsPerThread];
#pragma unroll
    for (int k = 0; k < kSortCellsPerThread; ++k) {
        const int c = I.cell_off + min(base + k * kBlock + (int)threadIdx.x, I.n_place - 1);
        p[k] = pos[c];
        s[k] = d.wh[c];
    }
#pragma unroll
    for (int k = 0; k < kSortCellsPerThread; ++k) {
        const int idx = base + k * kBlock + threadIdx.x;
        key[k] = -1;
        if (idx < I.n_place) {
            key[k] = cell_key(p[k], s[k], g, I.M, I.M);
            rank[k] = atomicAdd(&hist[key[k]], 1);
            cell[k] = pack_cell(p[k], s[k], idx < I.n_mov);
        }
    }
    __syncthreads();
    int32_t *off = d.sort_off + (size_t)i * kSortTiles;
    for (int t = threadIdx.x; t < kSortTiles; t += kBlock)
        if (hist[t]) base_of[t] = atomicAdd(&off[t], hist[t]);
    __syncthreads();
    float4 *sorted = d.sorted + I.cell_off;
#pragma unroll
    for (int k = 0; k < kSortCellsPerThread; ++k)
        if (key[k] >= 0) {
            const int slot = base_of[key[k]] + rank[k];
            if (slot < I.n_place) sorted[slot] = cell[k];  // counts always sum to n_place
        }
}

// Density scatter (as k_density): the instance's blocks take up to kBlock * CPT
// packed cells of one super tile each and accumulate them in an LDS window.
template <int CPT>
__global__ __launch_bounds__(kBlock) void kb_density(BDev d)
{
    const int lb = xcd_run_block(blockIdx.x, gridDim.x, kXcdRun);
    const int i = find_inst(d.pre_dens, d.n_inst, lb);
    const BState *S = d.st + i;
    if (S->done) return;
    const BInst &I = d.inst[i];
    const int level = I.dens_level;
    const int32_t *dtab = d.dens_tab + (size_t)i * (kSortTiles + 1);
    const int ns = kSortTiles >> (2 * level);
    const int b = lb - d.pre_dens[i];
    if (b >= dtab[ns]) return;
    const int s = dens_super_tile(dtab, ns, b);
    int c0, c1, x0, y0, nx, ny;
    dens_cell_range(d.sort_off + (size_t)i * kSortTiles, s, level, b - dtab[s], kBlock * CPT, I.n_place,
                    c0, c1);
    super_window(s, level, I.M, I.M, I.dens_ext, x0, y0, nx, ny);
    density_window<CPT>(d.sorted + I.cell_off, c0, c1, inst_geom(I), I.M, I.M, x0, y0, nx, ny,
                        d.acc_field + I.map_off, d.acc_mov + I.map_off, d.dens_wcap);
}

// Accumulators -> FP32 maps, overflow numerator and field-map sum/max/-min per
// block, clearing the accumulators for the next scatter (as k_overflow).
__global__ __launch_bounds__(kBlock) void kb_overflow(BDev d)
{
    const int i = find_inst(d.pre_bin, d.n_inst, blockIdx.x);
    if (d.st[i].done) return;
    const BInst &I = d.inst[i];
    const int nb = I.M * I.M;
    const int base = (blockIdx.x - d.pre_bin[i]) * kBinsPerBlock;
    constexpr int K = kBinsPerBlock / kBlock;
    double acc[4] = {0.0, 0.0, -1e300, -1e300};
    // All K bins' loads are issued before the first conversion; a bin past
    // the instance's map loads its last bin instead (in bounds, value unused).
    float fxs[K];
    unsigned long long ams[K], afs[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const size_t o = (size_t)I.map_off + min(base + k * kBlock + (int)threadIdx.x, nb - 1);
        fxs[k] = d.map_fix[o];
        ams[k] = d.acc_mov[o];
        afs[k] = d.acc_field[o];
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int b = base + k * kBlock + threadIdx.x;
        const float mv = mov_bin(ams[k], I.dens_unit);
        const float fd = field_bin(fxs[k], afs[k], I.dens_unit, I.rho_t);
        const float v = ovf_bin(mv, fxs[k], I.rho_t, I.cap);
        if (b < nb) {
            const size_t o = (size_t)I.map_off + b;
            d.acc_mov[o] = 0ull;  // cleared for the next scatter
            d.acc_field[o] = 0ull;
            d.map_mov[o] = mv;
            d.map_field[o] = fd;
            if (v > 0.0f) acc[0] += (double)v;
            const double f = fd;
            acc[1] += f;
            acc[2] = fmax(acc[2], f);
            acc[3] = fmax(acc[3], -f);
        }
    }
    block_reduce<4, 2>(acc, d.part_ovf + 4 * (size_t)blockIdx.x);
}

// One block per instance: max |map - mean| for the split-FP16 products (the
// mean is the instance's constant mean area per bin, set at creation).
__global__ __launch_bounds__(kBlock) void kb_map_stats(BDev d)
{
    const int i = blockIdx.x;
    if (d.st[i].done) return;
    double v[2] = {-1e300, -1e300};
    for (int b = d.pre_bin[i] + threadIdx.x; b < d.pre_bin[i + 1]; b += kBlock) {
        v[0] = fmax(v[0], d.part_ovf[4 * (size_t)b + 2]);
        v[1] = fmax(v[1], d.part_ovf[4 * (size_t)b + 3]);
    }
    __shared__ double out[2];
    block_reduce<2, 2>(v, out);
    if (threadIdx.x == 0) {
        const double mean = d.map_mean[i];
        const double r = fmax(out[0] - mean, -out[1] + mean);
        d.map_max[i] = __float_as_uint((float)fmax(r, 0.0));
    }
}

// ------------------------------------------------------------------ WA gradient

template <int G>
__device__ inline void wa_tile(const BDev &d, const BBucket &bk, int blk, const float2 *pos,
                               float ginv, double &hp, double &wa)
{
    const int net = bk.net0 + (blk * kBlock + (int)threadIdx.x) / G;
    const int lane = threadIdx.x & (G - 1);
    const bool active = net < bk.net1;
    int s = 0, e = 0;
    if (active) {
        if (bk.deg) {  // exact-degree class: no net_start load in the chain
            s = bk.pin0 + (net - bk.net0) * bk.deg;
            e = s + bk.deg;
        } else {
            s = d.net_start[net];
            e = d.net_start[net + 1];
        }
    }
    wa_net<G>(s, e, active, lane, ginv, d.pin_cell, d.pin_off, pos, d.pin_grad, hp, wa);
}

__global__ __launch_bounds__(kBlock) void kb_wa(BDev d, int at)
{
    const int lb = xcd_run_block(blockIdx.x, gridDim.x, kXcdRun);
    int lo = 0, hi = d.n_buckets;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (d.buckets[mid].blk0 <= lb) lo = mid;
        else hi = mid;
    }
    const BBucket bk = d.buckets[lo];
    const BState *S = d.st + bk.inst;
    if (S->done) return;
    const float2 *pos = eval_pos(d, S->cur, at);
    const float ginv = S->f_ginv;
    const int blk = lb - bk.blk0;
    double hp = 0.0, wa = 0.0;
    switch (bk.lanes) {
    case 2: wa_tile<2>(d, bk, blk, pos, ginv, hp, wa); break;
    case 4: wa_tile<4>(d, bk, blk, pos, ginv, hp, wa); break;
    case 8: wa_tile<8>(d, bk, blk, pos, ginv, hp, wa); break;
    case 16: wa_tile<16>(d, bk, blk, pos, ginv, hp, wa); break;
    case 32: wa_tile<32>(d, bk, blk, pos, ginv, hp, wa); break;
    default: wa_tile<64>(d, bk, blk, pos, ginv, hp, wa); break;
    }
    double v[2] = {hp, wa};
    block_reduce<2, 0>(v, d.part_wa + 2 * (size_t)lb);
}

// ------------------------------------------------------------------ per-cell gather

// One cell per thread (as k_cell_grad): g = (gW - lambda F) / max(1, pins + lambda area).
// at == kAtTrial: position V[cur^1], references V[cur]/G[cur], output G[cur^1];
// at == kAtCur:   position V[cur], no references, output G[cur].
__global__ __launch_bounds__(kBlock) void kb_cell_grad(BDev d, int at)
{
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int i = find_inst(d.pre_cell, d.n_inst, lb);
    const BState *S = d.st + i;
    if (S->done) return;
    const BInst &I = d.inst[i];
    const int a = S->cur, b = a ^ at;
    const float2 *pos = pick<const float2 *>(d.V[0], d.V[1], b);
    float2 *gout = pick(d.G[0], d.G[1], b);
    const float lambda = S->f_lambda;
    const GeomArgs g = inst_geom(I);
    const int M = I.M;
    const float2 *exy = d.exy + I.map_off;
    const int lc = (lb - d.pre_cell[i]) * kBlock + threadIdx.x;
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    if (lc < I.n_place) {
        const int c = I.cell_off + lc;
        const float2 p = pos[c], s = d.wh[c];
        float2 v0 = p, g0 = make_float2(0.f, 0.f);  // independent loads first
        if (at == kAtTrial) {
            v0 = pick<const float2 *>(d.V[0], d.V[1], a)[c];
            g0 = pick<const float2 *>(d.G[0], d.G[1], a)[c];
        }
        const int k0 = d.cps[c], k1 = d.cps[c + 1];
        const float2 f = cell_force(p, s, exy, g, M, M);
        const float2 gw = sum_pin_partials(k0, k1, d.cpins, d.pin_grad);
        const float pre = fmaxf(1.0f, (float)(k1 - k0) + lambda * s.x * s.y);
        const float gx = (gw.x - lambda * f.x) / pre, gy = (gw.y - lambda * f.y) / pre;
        gout[c] = make_float2(gx, gy);
        if (at == kAtTrial) {
            const double dx = (double)p.x - v0.x, dy = (double)p.y - v0.y;
            acc[0] += dx * dx + dy * dy;
            const double ddx = (double)gx - g0.x, ddy = (double)gy - g0.y;
            acc[1] += ddx * ddx + ddy * ddy;
        }
        acc[2] += fabs((double)gw.x) + fabs((double)gw.y);
        acc[3] += fabs((double)f.x) + fabs((double)f.y);
        acc[4] += (double)gx * gx + (double)gy * gy;
        if (!isfinite(gx) || !isfinite(gy)) acc[5] += 1.0;
        acc[6] = fmax(fabs((double)gx), fabs((double)gy));
    }
    block_reduce<7, 1>(acc, d.part_cell + 7 * (size_t)lb);
}

// ------------------------------------------------------------------ reduction + control

// Block i: fixed-order FP64 sums of instance i's partials into d.res[i]
// (thread 0 writes them).
__device__ inline void reduce_one(const BDev &d, int i)
{
    // hpwl, wa, ovf, dv2, dg2, |gw|, |f|, g2, nonfinite ; max |g|
    double v[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int b = d.pre_wa[i] + threadIdx.x; b < d.pre_wa[i + 1]; b += kBlock) {
        v[0] += d.part_wa[2 * (size_t)b];
        v[1] += d.part_wa[2 * (size_t)b + 1];
    }
    for (int b = d.pre_bin[i] + threadIdx.x; b < d.pre_bin[i + 1]; b += kBlock)
        v[2] += d.part_ovf[4 * (size_t)b];
    for (int b = d.pre_cell[i] + threadIdx.x; b < d.pre_cell[i + 1]; b += kBlock) {
        const double *q = d.part_cell + 7 * (size_t)b;
        v[3] += q[0]; v[4] += q[1]; v[5] += q[2]; v[6] += q[3]; v[7] += q[4]; v[8] += q[5];
        v[9] = fmax(v[9], q[6]);
    }
    __shared__ double out[10];
    block_reduce<10, 1>(v, out);
    if (threadIdx.x == 0) {
        const double mov_area = d.inst[i].mov_area;
        BRes &R = d.res[i];
        R.hpwl = out[0];
        R.wa = out[1];
        R.ovf = out[2];
        R.tau = mov_area > 0 ? out[2] / mov_area : 0.0;
        R.dv2 = out[3];
        R.dg2 = out[4];
        R.sum_gw = out[5];
        R.sum_f = out[6];
        R.gnorm2 = out[7];
        R.nonfinite = (int64_t)out[8];
        R.gmax = out[9];
    }
}

__device__ inline double gamma_dev(const BInst &I, const BCtl &C, double tau)
{
#pragma clang fp contract(off)
    double g = I.gamma_base;
    if (C.gamma_schedule) g *= sched_exp10((tau - 0.1) * 20.0 / 9.0 - 1.0);
    return g;
}

// Scalars of the coming iteration's first trial (ePlace Nesterov coefficients).
__device__ inline void begin_iteration(BState &S)
{
#pragma clang fp contract(off)
    S.ak1 = 0.5 * (1.0 + sqrt(4.0 * S.ak * S.ak + 1.0));
    S.coef = (S.ak - 1.0) / S.ak1;
    S.alpha_try = S.alpha;
    S.trial = 0;
    S.f_alpha = (float)S.alpha;
    S.f_coef = (float)S.coef;
    S.f_ginv = (float)(1.0 / S.gamma);
    S.f_lambda = (float)S.lambda;
}

// Instance i's initialisation steps and per-trial decision of ep_step (one
// thread): accept when alpha_hat >= 0.95 alpha_used (or after max_backtrack
// trials), then lambda *= mu(dHPWL), gamma(tau), stop at overflow / max_iters.
__device__ inline void control_one(const BDev &d, int mode, int i)
{
#pragma clang fp contract(off)
    BState S = d.st[i];
    if (S.done) return;
    const BRes R = d.res[i];
    const BInst &I = d.inst[i];
    const BCtl &C = d.ctl;
    switch (mode) {
    case kCtlGamma:
        S.gamma = gamma_dev(I, C, R.tau);
        S.f_ginv = (float)(1.0 / S.gamma);
        S.f_lambda = 0.0f;
        break;
    case kCtlLambda:
        S.evals++;
        S.lambda = R.sum_f > 0 ? C.lambda_eta * R.sum_gw / R.sum_f : C.lambda_eta;
        S.prev_hpwl = S.last_hpwl = R.hpwl;
        S.last_tau = R.tau;
        S.last_wa = R.wa;
        S.f_lambda = (float)S.lambda;
        break;
    case kCtlProbe:
        if (R.nonfinite) {
            S.error = 1;
            S.done = 1;
            break;
        }
        S.eps = R.gmax > 0 ? I.init_disp / R.gmax : 0.0;
        S.f_eps = (float)S.eps;
        break;
    case kCtlAlpha: {
        S.evals++;
        const double num = sqrt(R.dv2), den = sqrt(R.dg2);
        S.alpha = den > 0 ? num / den : S.eps;
        if (!(S.alpha > 0) || !isfinite(S.alpha)) S.alpha = S.eps > 0 ? S.eps : 1e-3;
        S.ak = 1.0;
        S.iteration = 0;
        begin_iteration(S);
        break;
    }
    default: {
        S.evals++;
        if (R.nonfinite) {
            S.error = 1;
            S.done = 1;
            break;
        }
        const double used = (double)S.f_alpha;
        const double nn = sqrt(R.dv2), dd = sqrt(R.dg2);
        const double ahat = dd > 0 ? nn / dd : used;
        if (!(ahat >= 0.95 * used) && S.trial < C.max_backtrack - 1) {
            S.alpha_try = ahat;
            S.f_alpha = (float)ahat;
            S.trial++;
            break;
        }
        S.cur ^= 1;
        S.alpha = ahat;
        S.ak = S.ak1;
        const double hp = R.hpwl, tau = R.tau;
        S.last_hpwl = hp;
        S.last_tau = tau;
        S.last_wa = R.wa;
        const double dh = hp - S.prev_hpwl;
        double mu = dh < 0 ? C.mu_max : sched_pow(C.mu_max, 1.0 - dh / I.hpwl_ref);
        mu = fmin(fmax(mu, C.mu_min), C.mu_max);
        S.lambda *= mu;
        S.prev_hpwl = hp;
        S.gamma = gamma_dev(I, C, tau);
        S.iteration++;
        if (tau <= C.stop_overflow || S.iteration >= C.max_iters) S.done = 1;
        else begin_iteration(S);
        break;
    }
    }
    d.st[i] = S;
}

// One block per instance: block i sums instance i's partials, and its thread 0
// then takes the instance's control step `mode` on them.
__global__ __launch_bounds__(kBlock) void kb_reduce_control(BDev d, int mode)
{
    const int i = blockIdx.x;
    if (d.st[i].done) return;
    reduce_one(d, i);
    if (threadIdx.x == 0) control_one(d, mode, i);
}

}  // namespace

// ------------------------------------------------------------------ launchers

bool bl_density_cpt_supported(int cpt) { return cpt == 2 || cpt == 4 || cpt == 8; }

void bl_move(const BDev &d, int mode, int sort_blocks, hipStream_t s)
{
    hipLaunchKernelGGL(kb_move, dim3(sort_blocks), dim3(kBlock), 0, s, d, mode);
}

void bl_sort(const BDev &d, int at, bool counted, int bin_blocks, int sort_blocks, int dens_cpt,
             hipStream_t s)
{
    if (!counted) hipLaunchKernelGGL(kb_sort_count, dim3(sort_blocks), dim3(kBlock), 0, s, d, at);
    hipLaunchKernelGGL(kb_sort_scan, dim3(d.n_inst), dim3(kSortScanThreads), 0, s, d, kBlock * dens_cpt);
    hipLaunchKernelGGL(kb_sort_scatter, dim3(sort_blocks), dim3(kBlock), 0, s, d, at);
}

void bl_dens(const BDev &d, int bin_blocks, int dens_blocks, int dens_cpt, bool map_range, hipStream_t s)
{
    const size_t lds = 2 * sizeof(unsigned long long) * (size_t)d.dens_wcap;
    if (dens_cpt == 2)
        hipLaunchKernelGGL(kb_density<2>, dim3(dens_blocks), dim3(kBlock), lds, s, d);
    else if (dens_cpt == 8)
        hipLaunchKernelGGL(kb_density<8>, dim3(dens_blocks), dim3(kBlock), lds, s, d);
    else
        hipLaunchKernelGGL(kb_density<4>, dim3(dens_blocks), dim3(kBlock), lds, s, d);
    hipLaunchKernelGGL(kb_overflow, dim3(bin_blocks), dim3(kBlock), 0, s, d);
    if (map_range) hipLaunchKernelGGL(kb_map_stats, dim3(d.n_inst), dim3(kBlock), 0, s, d);
}

void bl_maps(const BDev &d, int at, bool counted, int bin_blocks, int sort_blocks, int dens_blocks,
             int dens_cpt, bool map_range, hipStream_t s)
{
    bl_sort(d, at, counted, bin_blocks, sort_blocks, dens_cpt, s);
    bl_dens(d, bin_blocks, dens_blocks, dens_cpt, map_range, s);
}

void bl_wa(const BDev &d, int at, int wa_blocks, hipStream_t s)
{
    if (wa_blocks > 0) hipLaunchKernelGGL(kb_wa, dim3(wa_blocks), dim3(kBlock), 0, s, d, at);
}

void bl_cell_grad(const BDev &d, int at, int cell_blocks, hipStream_t s)
{
    hipLaunchKernelGGL(kb_cell_grad, dim3(cell_blocks), dim3(kBlock), 0, s, d, at);
}

void bl_reduce_control(const BDev &d, int mode, hipStream_t s)
{
    hipLaunchKernelGGL(kb_reduce_control, dim3(d.n_inst), dim3(kBlock), 0, s, d, mode);
}

}  // namespace ep
