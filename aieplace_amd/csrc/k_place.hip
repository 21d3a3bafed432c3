// k_place.hip -- HBM-bound kernels of one ePlace gradient evaluation:
//   density scatter (LDS-privatised bin tiles, atomic flush), overflow
//   reduction, WA wirelength gradient (per-net lane groups with wavefront
//   shuffles), per-cell gather + preconditioning, Nesterov trial update, and
//   the FP64 finalize reduction.
//
// Reference algorithms (searsm8/AIEplace): WA terms/partials
// vck5000/aie/src/partials/compute_abc.cpp:23-125, compute_partials.cpp:11-63,
// host/src/AIEplace.cpp:612-673; overlap Bin.h:78-121, Grid.cpp:38-110;
// overflow Bin.h:123-126; HPWL Net.cpp:136-155; cell update AIEplace.cpp:993-1046.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>
#include <mutex>

#include "ep_ctx.h"
#include "ep_mlp.h"
#include "ep_device.h"
#include "ep_sched.h"

namespace ep {

namespace {

using namespace dev;


// Clears the fixed-point scatter accumulators (field and movable planes).
// 16-byte stores (nb is even: bins are multiples of 16).
__global__ void k_fill_maps(unsigned long long *__restrict__ acc_field,
                            unsigned long long *__restrict__ acc_mov, int nb,
                            const DevScalars *__restrict__ sc)
{
    if (sc->halt) return;
    ulonglong2 *f = reinterpret_cast<ulonglong2 *>(acc_field);
    ulonglong2 *m = reinterpret_cast<ulonglong2 *>(acc_mov);
    const ulonglong2 z = make_ulonglong2(0ull, 0ull);
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < (nb >> 1); b += gridDim.x * blockDim.x) {
        f[b] = z;
        m[b] = z;
    }
}

// ---- spatial counting sort of the cells by coarse tile (kSortSide^2 tiles in
// Morton order).  Any order gives the same maps; a spatially sorted order makes
// each density block's cells share a small bin window, so the block can
// accumulate in LDS and flush with coalesced atomics.


__global__ __launch_bounds__(kBlock) void k_sort_count(Pair<const float2> pos_pair,
                                                       const float2 *__restrict__ wh, int n_place,
                                                       GeomArgs g, int Mx, int My,
                                                       int32_t *__restrict__ count,
                                                       const DevScalars *__restrict__ sc)
{
    if (sc->halt) return;
    const float2 *__restrict__ pos = pos_pair.p[sc->sel];
    __shared__ int hist[kSortTiles];
    for (int t = threadIdx.x; t < kSortTiles; t += kBlock) hist[t] = 0;
    __syncthreads();
    const int base = blockIdx.x * kBlock * kSortCellsPerThread;
#pragma unroll
    for (int k = 0; k < kSortCellsPerThread; ++k) {
        const int idx = base + k * kBlock + threadIdx.x;
        if (idx < n_place) atomicAdd(&hist[cell_key(pos[idx], wh[idx], g, Mx, My)], 1);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < kSortTiles; t += kBlock)
        if (hist[t]) atomicAdd(&count[t], hist[t]);
}

// Exclusive scan of the tile counts into offsets (clears the counts) and the
// density block table of the coming scatter.
__global__ __launch_bounds__(kSortScanThreads) void k_sort_scan(int32_t *__restrict__ count,
                                                                int32_t *__restrict__ off,
                                                                int32_t *__restrict__ dtab,
                                                                int level, int K,
                                                                const DevScalars *__restrict__ sc)
{
    if (sc->halt) return;
    sort_scan_block(count, off, dtab, level, K);
}

// Writes every cell, packed with its size and movable flag (pack_cell), to its
// slot of the tile-sorted order.
__global__ __launch_bounds__(kBlock) void k_sort_scatter(Pair<const float2> pos_pair,
                                                         const float2 *__restrict__ wh, int n_place,
                                                         int n_mov, GeomArgs g, int Mx, int My,
                                                         int32_t *__restrict__ off,
                                                         float4 *__restrict__ sorted,
                                                         const DevScalars *__restrict__ sc)
{
    if (sc->halt) return;
    const float2 *__restrict__ pos = pos_pair.p[sc->sel];
    __shared__ int hist[kSortTiles];
    __shared__ int base_of[kSortTiles];
    for (int t = threadIdx.x; t < kSortTiles; t += kBlock) hist[t] = 0;
    __syncthreads();
    const int base = blockIdx.x * kBlock * kSortCellsPerThread;
    int key[kSortCellsPerThread], rank[kSortCellsPerThread];
    float4 cell[kSortCellsPerThread];
    // All of the thread's cells are loaded before the first tile count, so
    // the loads overlap; a cell past n_place loads the last cell (unused).
    float2 p[kSortCellsPerThread], s[kSortCellsPerThread];
#pragma unroll
    for (int k = 0; k < kSortCellsPerThread; ++k) {
        const int idx = min(base + k * kBlock + (int)threadIdx.x, n_place - 1);
        p[k] = pos[idx];
        s[k] = wh[idx];
    }
#pragma unroll
    for (int k = 0; k < kSortCellsPerThread; ++k) {
        const int idx = base + k * kBlock + threadIdx.x;
        key[k] = -1;
        if (idx < n_place) {
            key[k] = cell_key(p[k], s[k], g, Mx, My);
            rank[k] = atomicAdd(&hist[key[k]], 1);
            cell[k] = pack_cell(p[k], s[k], idx < n_mov);
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < kSortTiles; t += kBlock)
        if (hist[t]) base_of[t] = atomicAdd(&off[t], hist[t]);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kSortCellsPerThread; ++k)
        if (key[k] >= 0) {
            const int slot = base_of[key[k]] + rank[k];
            if (slot < n_place) sorted[slot] = cell[k];  // counts always sum to n_place
        }
}

// ---- density scatter: block b takes up to kBlock * CPT cells of one super
// tile (dtab from k_sort_scan) and accumulates them in an LDS window.
template <int CPT>
__global__ __launch_bounds__(kBlock) void k_density(const float4 *__restrict__ sorted,
                                                    const int32_t *__restrict__ off,
                                                    const int32_t *__restrict__ dtab, int level,
                                                    int ext, int wcap, int n_place, GeomArgs g,
                                                    int Mx, int My,
                                                    unsigned long long *__restrict__ acc_field,
                                                    unsigned long long *__restrict__ acc_mov,
                                                    const DevScalars *__restrict__ sc)
{
    if (sc->halt) return;
    const int ns = kSortTiles >> (2 * level);
    const int b = xcd_run_block(blockIdx.x, gridDim.x, kXcdRun);  // neighbouring super tiles share an L2
    if (b >= dtab[ns]) return;
    const int s = dens_super_tile(dtab, ns, b);
    int c0, c1, x0, y0, nx, ny;
    dens_cell_range(off, s, level, b - dtab[s], kBlock * CPT, n_place, c0, c1);
    super_window(s, level, Mx, My, ext, x0, y0, nx, ny);
    density_window<CPT>(sorted, c0, c1, g, Mx, My, x0, y0, nx, ny, acc_field, acc_mov, wcap);
}

// Converts the fixed-point accumulators into the FP32 maps (field map =
// rho_t * fixed + smoothed cells; movable map = exact movable overlap) and
// clears them for the next scatter.  Reduces per block the overflow numerator
// sum_b max(0, mov_b + rho_t*fix_b - rho_t*A_b) and the field map's sum, max
// and min (the transforms subtract the map mean; split FP16 scales by the
// centred range).  part layout per block: [ovf, sum, max, -min].
__global__ __launch_bounds__(kBlock) void k_overflow(unsigned long long *__restrict__ acc_field,
                                                     unsigned long long *__restrict__ acc_mov,
                                                     const float *__restrict__ map_fix, double unit,
                                                     float rho_t, float cap, int nb,
                                                     float *__restrict__ map_field,
                                                     float *__restrict__ map_mov,
                                                     double *__restrict__ part,
                                                     const DevScalars *__restrict__ sc)
{
    if (sc->halt) return;
    // The loads of kOvfRun grid-stride bins are issued together, then the
    // bins are converted in the same order as one bin per pass would.  A bin
    // past the map loads the last bin instead (in bounds, value unused).
    constexpr int kOvfRun = 4;
    const int stride = gridDim.x * kBlock;
    double acc[4] = {0.0, 0.0, -1e300, -1e300};
    for (int b0 = blockIdx.x * kBlock + threadIdx.x; b0 < nb; b0 += kOvfRun * stride) {
        float fx[kOvfRun];
        unsigned long long am[kOvfRun], af[kOvfRun];
#pragma unroll
        for (int u = 0; u < kOvfRun; ++u) {
            const int b = min(b0 + u * stride, nb - 1);
            fx[u] = map_fix[b];
            am[u] = acc_mov[b];
            af[u] = acc_field[b];
        }
#pragma unroll
        for (int u = 0; u < kOvfRun; ++u) {
            const int b = b0 + u * stride;
            const float mv = mov_bin(am[u], unit);
            const float fd = field_bin(fx[u], af[u], unit, rho_t);
            const float v = ovf_bin(mv, fx[u], rho_t, cap);
            if (b < nb) {
                acc_mov[b] = 0ull;
                acc_field[b] = 0ull;
                map_mov[b] = mv;
                map_field[b] = fd;
                if (v > 0.0f) acc[0] += (double)v;
                const double f = fd;
                acc[1] += f;
                acc[2] = fmax(acc[2], f);
                acc[3] = fmax(acc[3], -f);
            }
        }
    }
    block_reduce<4, 2>(acc, part + 4 * blockIdx.x);
}

// Centred range of the field map for the split-FP16 products: max |map - mean|
// -> field_max0 (float bits), from k_overflow's per-block max and -min and the
// context's constant map mean.  One block.
__global__ __launch_bounds__(kBlock) void k_map_range(const double *__restrict__ part, int blocks,
                                                      const float *__restrict__ mean_in,
                                                      uint32_t *__restrict__ field_max0,
                                                      const DevScalars *__restrict__ sc)
{
    if (sc->halt) return;
    double v[2] = {-1e300, -1e300};
    for (int b = threadIdx.x; b < blocks; b += kBlock) {
        v[0] = fmax(v[0], part[4 * b + 2]);
        v[1] = fmax(v[1], part[4 * b + 3]);
    }
    __shared__ double out[2];
    block_reduce<2, 2>(v, out);
    if (threadIdx.x == 0) {
        const double mean = *mean_in;
        const double r = fmax(out[0] - mean, -out[1] + mean);
        *field_max0 = __float_as_uint((float)fmax(r, 0.0));
    }
}

// ------------------------------------------------------------------ WA gradient


template <int G>
__device__ inline void wa_bucket(const Bucket &bk, int blk, const int32_t *__restrict__ net_start,
                                 const int32_t *__restrict__ pin_cell,
                                 const float2 *__restrict__ pin_off,
                                 const float2 *__restrict__ pos, float2 *__restrict__ pin_grad,
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
            s = net_start[net];
            e = net_start[net + 1];
        }
    }
    wa_net<G>(s, e, active, lane, ginv, pin_cell, pin_off, pos, pin_grad, hp, wa);
}

__global__ __launch_bounds__(kBlock) void k_wa(BucketTable bt, const int32_t *__restrict__ net_start,
                                               const int32_t *__restrict__ pin_cell,
                                               const float2 *__restrict__ pin_off,
                                               Pair<const float2> pos_pair,
                                               float2 *__restrict__ pin_grad,
                                               const DevScalars *__restrict__ sc,
                                               int total_blocks, double *__restrict__ part)
{
    if (sc->halt) return;
    const float2 *__restrict__ pos = pos_pair.p[sc->sel];
    const float ginv = sc->gamma_inv;
    double hp = 0.0, wa = 0.0;
    // one net tile per block (virtual blocks if the launch is capped); runs of
    // neighbouring tiles share an XCD's L2
    for (int vb = xcd_run_block(blockIdx.x, gridDim.x, kXcdRun); vb < total_blocks;
         vb += gridDim.x) {
        int b = 0;
        while (b + 1 < bt.n && vb >= bt.b[b + 1].blk0) ++b;
        const Bucket bk = bt.b[b];
        const int blk = vb - bk.blk0;
        switch (bk.lanes) {
        case 2: wa_bucket<2>(bk, blk, net_start, pin_cell, pin_off, pos, pin_grad, ginv, hp, wa); break;
        case 4: wa_bucket<4>(bk, blk, net_start, pin_cell, pin_off, pos, pin_grad, ginv, hp, wa); break;
        case 8: wa_bucket<8>(bk, blk, net_start, pin_cell, pin_off, pos, pin_grad, ginv, hp, wa); break;
        case 16: wa_bucket<16>(bk, blk, net_start, pin_cell, pin_off, pos, pin_grad, ginv, hp, wa); break;
        case 32: wa_bucket<32>(bk, blk, net_start, pin_cell, pin_off, pos, pin_grad, ginv, hp, wa); break;
        default: wa_bucket<64>(bk, blk, net_start, pin_cell, pin_off, pos, pin_grad, ginv, hp, wa); break;
        }
    }
    double v[2] = {hp, wa};
    block_reduce<2, 0>(v, part + 2 * blockIdx.x);
}

// ------------------------------------------------------------------ reference packet format

// WA partials straight from the AIE packet layout (DataBase.cpp:217-273,
// compute_abc.cpp:27-57): groups of 8 lanes (one net per lane), `net_size`
// vectors of 8 values per group, pin 0 = net maximum.  Output in the same
// (group, pin, lane) order (compute_partials.cpp:11-63).  Lanes past n_nets
// are zero padding and get 0.  One thread per lane.
__global__ void k_wa_packet(const float *__restrict__ data, int net_size, int n_nets, int groups,
                            float ginv, float *__restrict__ out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= groups * 8) return;
    const int g = t >> 3, lane = t & 7;
    const float *v = data + (size_t)g * net_size * 8 + lane;
    float *o = out + (size_t)g * net_size * 8 + lane;
    if (t >= n_nets) {
        for (int p = 0; p < net_size; ++p) o[p * 8] = 0.0f;
        return;
    }
    float xmx = -FLT_MAX, xmn = FLT_MAX;
    for (int p = 0; p < net_size; ++p) {
        xmx = fmaxf(xmx, v[p * 8]);
        xmn = fminf(xmn, v[p * 8]);
    }
    float bp = 0.f, bm = 0.f, cp = 0.f, cm = 0.f;
    for (int p = 0; p < net_size; ++p) {
        const float dp = v[p * 8] - xmx, dm = v[p * 8] - xmn;
        const float ap = __expf(dp * ginv), am = __expf(-dm * ginv);
        bp += ap; bm += am; cp += dp * ap; cm += dm * am;
    }
    const float mp = cp / bp, mm = cm / bm;
    for (int p = 0; p < net_size; ++p) {
        const float dp = v[p * 8] - xmx, dm = v[p * 8] - xmn;
        const float ap = __expf(dp * ginv), am = __expf(-dm * ginv);
        o[p * 8] = ap / bp * (1.f + (dp - mp) * ginv) - am / bm * (1.f - (dm - mm) * ginv);
    }
}

// The same packet in FP64 with the AIE kernel's exponent, evaluated the way the
// reference's golden generator does (python/AIEmath/computeTerm.py:6-56,
// createGoldenPartials.py:48-70):
//   fast_exp(a) = (1 + a / 2^16)^(2^16), 16 squarings (compute_abc.cpp:12-18);
//   b and c summed over the pins in ascending order, which in the packet layout
//   (pin 0 the maximum, the others ascending) is pins 1..n-1, then pin 0;
//   partial = ((1 + x/g) b+ - c+/g) a+ / (b+ b+) - ((1 - x/g) b- + c-/g) a- / (b- b-).
// IEEE double operations in that order without multiply-add contraction, so
// for integer coordinates the doubles are the generator's, and the FP32
// outputs are its values rounded once.
__device__ inline double fast_exp64(double a)
{
#pragma clang fp contract(off)
    double r = 1.0 + a / 65536.0;
    for (int i = 0; i < 16; ++i) r *= r;
    return r;
}

__global__ void k_wa_packet_fast64(const float *__restrict__ data, int net_size, int n_nets, int groups,
                                   double gamma, float *__restrict__ out)
{
#pragma clang fp contract(off)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= groups * 8) return;
    const int g = t >> 3, lane = t & 7;
    const float *v = data + (size_t)g * net_size * 8 + lane;
    float *o = out + (size_t)g * net_size * 8 + lane;
    if (t >= n_nets) {
        for (int p = 0; p < net_size; ++p) o[p * 8] = 0.0f;
        return;
    }
    double xmx = (double)v[0], xmn = (double)v[0];
    for (int p = 1; p < net_size; ++p) {
        xmx = fmax(xmx, (double)v[p * 8]);
        xmn = fmin(xmn, (double)v[p * 8]);
    }
    double bp = 0.0, cp = 0.0, bm = 0.0, cm = 0.0;
    for (int k = 0; k < net_size; ++k) {
        const double x = (double)v[(k + 1 < net_size ? k + 1 : 0) * 8];
        const double ap = fast_exp64((x - xmx) / gamma), am = fast_exp64((xmn - x) / gamma);
        bp += ap;
        cp += x * ap;
        bm += am;
        cm += x * am;
    }
    for (int p = 0; p < net_size; ++p) {
        const double x = (double)v[p * 8];
        const double ap = fast_exp64((x - xmx) / gamma), am = fast_exp64((xmn - x) / gamma);
        const double t1 = ((1.0 + x / gamma) * bp - cp / gamma) * ap / (bp * bp);
        const double t2 = ((1.0 - x / gamma) * bm + cm / gamma) * am / (bm * bm);
        o[p * 8] = (float)(t1 - t2);
    }
}

// ------------------------------------------------------------------ per-cell gather

// g = (gW - lambda * F) / max(1, pins + lambda * area); F = r * sum ov * E over
// the smoothed rectangle.  Also the FP64 partial norms for the step length.
// kVref / kGref / kGwCell: whether vref_pair, gref_pair and gw_cell are set.
// As compile-time flags they leave the loads of a cell in one basic block,
// so they are in flight together.
template <bool kVref, bool kGref, bool kGwCell>
__global__ __launch_bounds__(kBlock) void k_cell_grad(
    Pair<const float2> pos_pair, const float2 *__restrict__ wh, int n_place, int n_mov,
    GeomArgs g, int Mx, int My, const int32_t *__restrict__ cell_pin_start,
    const int32_t *__restrict__ cell_pins, const float2 *__restrict__ pin_grad,
    const float2 *__restrict__ exy, Pair<const float2> vref_pair,
    Pair<const float2> gref_pair, Pair<float2> gout_pair, float2 *__restrict__ gw_out,
    float2 *__restrict__ f_out, const DevScalars *__restrict__ sc, double *__restrict__ part,
    const float2 *__restrict__ gw_cell)
{
    if (sc->halt) return;
    const int sel = sc->sel;
    const float2 *__restrict__ pos = pos_pair.p[sel];
    const float2 *__restrict__ vref = vref_pair.p[sel];
    const float2 *__restrict__ gref = gref_pair.p[sel];
    float2 *__restrict__ gout = gout_pair.p[sel];
    const float lambda = sc->lambda;
    // sums: dv2, dg2, |gw|, |f|, g2, nonfinite ; max: |g|
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    auto one = [&](int c) {
        const float2 p = pos[c], s = wh[c];
        const float2 v0 = kVref ? vref[c] : p;  // independent loads first
        const float2 g0 = kGref ? gref[c] : make_float2(0.f, 0.f);
        // Clamped instead of branching on c < n_mov: cell_pin_start holds
        // n_mov + 1 entries, so a fixed cell or filler reads k0 == k1 (no pins).
        const bool mov = c < n_mov;
        const int k0 = cell_pin_start[min(c, n_mov)], k1 = cell_pin_start[min(c + 1, n_mov)];
        float2 gwc = make_float2(0.f, 0.f);
        if (kGwCell) gwc = gw_cell[max(0, min(c, n_mov - 1))];  // holds max(1, n_mov) entries
        const int np = k1 - k0;
        const float2 f = cell_force(p, s, exy, g, Mx, My);
        const float2 gw = kGwCell ? (mov ? gwc : make_float2(0.f, 0.f))
                                  : sum_pin_partials(k0, k1, cell_pins, pin_grad);
        const float gwx = gw.x, gwy = gw.y, fx = f.x, fy = f.y;
        const float pre = fmaxf(1.0f, (float)np + lambda * s.x * s.y);
        const float gx = (gwx - lambda * fx) / pre, gy = (gwy - lambda * fy) / pre;
        gout[c] = make_float2(gx, gy);
        if (gw_out) gw_out[c] = make_float2(gwx, gwy);
        if (f_out) f_out[c] = make_float2(fx, fy);
        if (kVref) {
            const double dx = (double)p.x - v0.x, dy = (double)p.y - v0.y;
            acc[0] += dx * dx + dy * dy;
        }
        if (kGref) {
            const double dx = (double)gx - g0.x, dy = (double)gy - g0.y;
            acc[1] += dx * dx + dy * dy;
        }
        acc[2] += fabs((double)gwx) + fabs((double)gwy);
        acc[3] += fabs((double)fx) + fabs((double)fy);
        acc[4] += (double)gx * gx + (double)gy * gy;
        if (!isfinite(gx) || !isfinite(gy)) acc[5] += 1.0;
        acc[6] = fmax(acc[6], fmax(fabs((double)gx), fabs((double)gy)));
    };
    // two cells per thread and step, so two gather chains are in flight
    const int stride = gridDim.x * kBlock;
    for (int c = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x; c < n_place; c += 2 * stride) {
        one(c);
        if (c + stride < n_place) one(c + stride);
    }
    block_reduce<7, 1>(acc, part + 7 * blockIdx.x);
}

// Per movable cell, the sum of its pin partials (what k_cell_grad gathers
// itself when it gets no gw_cell): lets the gather run beside the transform
// products, off the path to the field.
__global__ __launch_bounds__(kBlock) void k_pin_sum(const int32_t *__restrict__ cell_pin_start,
                                                   const int32_t *__restrict__ cell_pins,
                                                   const float2 *__restrict__ pin_grad, int n_mov,
                                                   float2 *__restrict__ gw_cell,
                                                   const DevScalars *__restrict__ sc)
{
    if (sc->halt) return;
    for (int c = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x; c < n_mov;
         c += gridDim.x * kBlock)
        gw_cell[c] = sum_pin_partials(cell_pin_start[c], cell_pin_start[c + 1], cell_pins, pin_grad);
}

// Deterministic final reduction of all partial sums: block b of
// kFinalizeBlocks sums slice b of every partial array in a fixed order and
// writes its ten values to fin[b]; the last block to finish (atomic ticket)
// sums fin[0..B) in block order, so the result does not depend on timing.
constexpr int kFinalizeThreads = 256;
static_assert(kFinalizeBlocks <= kFinalizeThreads, "one loading thread per block result");
constexpr int kCtlWords = (int)(sizeof(DevCtl) / 4);
constexpr int kMlpWordsPerThread = (EP_MLP_NWEIGHTS + kFinalizeThreads - 1) / kFinalizeThreads;
static_assert(sizeof(DevCtl) % 4 == 0 && kCtlWords <= kFinalizeThreads, "one word of the control state per thread");

struct StepCtlArgs {
    DevCtl *ctl;            // null: no step decision (host-controlled launches)
    ep_step_info *log;
    float *rec;
    CtlParams p;
    const float *mlp_w;     // learned step: the model proposes the next trial's step
};

// What the decision leaves for the next trial's model proposal, kept by the
// deciding thread as it writes the same values to the scalars (so that it
// need not read them back): the model flag, the halt flag and the features.
struct NextTrial {
    int use_model;
    int halt;
    float *feat;  // 16 floats, written when the decision starts an iteration
};

__device__ inline void step_decide(DevCtl *__restrict__ ctl, const DevCtl &S0, const DevResult &R,
                                   DevScalars *__restrict__ sc, const CtlParams &p,
                                   ep_step_info *__restrict__ log, float *__restrict__ rec, NextTrial &nt);

__global__ __launch_bounds__(kFinalizeThreads) void k_finalize(
    const double *__restrict__ part, int wa_off, int wa_blocks, int ovf_off, int ovf_blocks,
    int cell_off, int cell_blocks, int trial_off, int trial_blocks, double mov_area, DevScalars *__restrict__ sc,
    DevResult *__restrict__ res, double *__restrict__ fin, unsigned *__restrict__ ticket,
    StepCtlArgs ctl)
{
    if (sc->halt) return;
    const int B = gridDim.x, b = blockIdx.x;
    // Device-controlled runs: the control state and the trial's step are
    // loaded now, beside the partial sums, so that the decision at the end of
    // the last block does not wait for them (the previous launch wrote them).
    __shared__ DevCtl ctl_s;
    __shared__ float alpha_s;
    __shared__ int use_model_s;
    uint32_t ctl_w = 0;
    const bool ctl_load = ctl.ctl && threadIdx.x < kCtlWords;
    if (ctl_load) ctl_w = reinterpret_cast<const uint32_t *>(ctl.ctl)[threadIdx.x];
    const float alpha_used = sc->alpha;
    const int use_model0 = sc->use_model;
    // The step model's weights likewise (learned runs): registers now, LDS
    // after the partial sums, read by the proposal at the end.
    __shared__ float mlp_s[EP_MLP_NWEIGHTS];
    float mlp_r[kMlpWordsPerThread];
    if (ctl.mlp_w)
#pragma unroll
        for (int k = 0; k < kMlpWordsPerThread; ++k) {
            const int i = threadIdx.x + k * kFinalizeThreads;
            mlp_r[k] = i < EP_MLP_NWEIGHTS ? ctl.mlp_w[i] : 0.0f;
        }
    auto slice = [&](int n, int &lo, int &hi) {
        lo = (int)((int64_t)n * b / B);
        hi = (int)((int64_t)n * (b + 1) / B);
    };
    double v[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    int lo, hi;
    slice(wa_blocks, lo, hi);
    for (int i = lo + threadIdx.x; i < hi; i += kFinalizeThreads) {
        v[0] += part[wa_off + 2 * i];
        v[1] += part[wa_off + 2 * i + 1];
    }
    slice(ovf_blocks, lo, hi);
    for (int i = lo + threadIdx.x; i < hi; i += kFinalizeThreads) v[2] += part[ovf_off + 4 * i];
    slice(cell_blocks, lo, hi);
    for (int i = lo + threadIdx.x; i < hi; i += kFinalizeThreads) {
        const double *q = part + cell_off + 7 * i;
        v[3] += q[0]; v[4] += q[1]; v[5] += q[2]; v[6] += q[3]; v[7] += q[4]; v[8] += q[5];
        v[9] = fmax(v[9], q[6]);
    }
    slice(trial_blocks, lo, hi);  // |v' - v|^2 of a trial (k_cell_grad's sum is then 0)
    for (int i = lo + threadIdx.x; i < hi; i += kFinalizeThreads) v[3] += part[trial_off + i];
    __shared__ double out[10];
    __shared__ int last;
    if (ctl_load) reinterpret_cast<uint32_t *>(&ctl_s)[threadIdx.x] = ctl_w;
    if (threadIdx.x == 0) {
        alpha_s = alpha_used;
        use_model_s = use_model0;
    }
    if (ctl.mlp_w)
#pragma unroll
        for (int k = 0; k < kMlpWordsPerThread; ++k) {
            const int i = threadIdx.x + k * kFinalizeThreads;
            if (i < EP_MLP_NWEIGHTS) mlp_s[i] = mlp_r[k];
        }
    block_reduce<10, 1, kFinalizeThreads>(v, out);  // its barriers also publish ctl_s, alpha_s, mlp_s
    if (threadIdx.x == 0) {
        for (int k = 0; k < 10; ++k) fin[b * 10 + k] = out[k];
        __threadfence();
        last = atomicAdd(ticket, 1u) == (unsigned)(B - 1);
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    __shared__ double sfin[kFinalizeBlocks][10];  // every block's sums, loaded in parallel
    if (threadIdx.x < B)
        for (int k = 0; k < 10; ++k) sfin[threadIdx.x][k] = load_agent(&fin[threadIdx.x * 10 + k]);
    __syncthreads();
    // thread k sums column k over the blocks in block order (ten short chains
    // instead of one long one; the same order, so the same sums)
    __shared__ double tot[10];
    if (threadIdx.x < 10) {
        const int k = threadIdx.x;
        double t = sfin[0][k];
        for (int j = 1; j < B; ++j) t = k < 9 ? t + sfin[j][k] : fmax(t, sfin[j][k]);
        tot[k] = t;
    }
    __syncthreads();
    __shared__ int propose;
    __shared__ float feat[16];  // the decision thread's features, for the whole block
    if (threadIdx.x == 0) {
        double t[10];
        for (int k = 0; k < 10; ++k) t[k] = tot[k];
        __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);  // next launch
        DevResult R;
        R.hpwl = t[0];
        R.wa = t[1];
        R.ovf_num = t[2];
        R.tau = mov_area > 0 ? t[2] / mov_area : 0.0;
        R.dv2 = t[3];
        R.dg2 = t[4];
        R.sum_gw = t[5];
        R.sum_f = t[6];
        R.gnorm2 = t[7];
        R.nonfinite = (int64_t)t[8];
        R.gmax = t[9];
        R.alpha_used = alpha_s;
        *res = R;
        if (ctl.ctl) {  // device-controlled run: the decision also fills feat
            NextTrial nt{use_model_s, 0, feat};
            step_decide(ctl.ctl, ctl_s, R, sc, ctl.p, ctl.log, ctl.rec, nt);
            propose = ctl.mlp_w && nt.use_model && !nt.halt;
        } else {
            propose = ctl.mlp_w && sc->use_model && !sc->halt;
            if (propose)
                for (int k = 0; k < 16; ++k) feat[k] = sc->feat[k];
        }
    }
    // the model's step for the next trial, while this block is still resident
    if (ctl.mlp_w) {
        __syncthreads();
        mlp_propose_block(mlp_s, feat, sc, propose != 0);
    }
}

// ------------------------------------------------------------------ Nesterov

// u' = clamp(v - alpha g); v' = clamp(u' + coef (u' - u))   (ePlace Nesterov step)
// With `count`, the new positions are also counted into the sort's coarse
// tiles (the first pass of launch_sort, which then skips it).  Blocks take
// kBlock * kSortCellsPerThread cells, as k_sort_count.
__global__ __launch_bounds__(kBlock) void k_trial(Pair<const float2> v_pair,
                                                  Pair<const float2> u_pair,
                                                  Pair<const float2> g_pair,
                                                  const float2 *__restrict__ wh, int n_place,
                                                  GeomArgs g, int Mx, int My,
                                                  const DevScalars *__restrict__ sc,
                                                  Pair<float2> v_new_pair,
                                                  Pair<float2> u_new_pair,
                                                  int32_t *__restrict__ count,
                                                  double *__restrict__ dv2_part)
{
    if (sc->halt) return;
    const int sel = sc->sel;
    const float2 *__restrict__ v = v_pair.p[sel];
    const float2 *__restrict__ u = u_pair.p[sel];
    const float2 *__restrict__ gr = g_pair.p[sel];
    float2 *__restrict__ v_new = v_new_pair.p[sel];
    float2 *__restrict__ u_new = u_new_pair.p[sel];
    __shared__ int hist[kSortTiles];
    if (count) {
        for (int t = threadIdx.x; t < kSortTiles; t += kBlock) hist[t] = 0;
        __syncthreads();
    }
    const float alpha = sc->alpha, coef = sc->coef;
    const int base = blockIdx.x * kBlock * kSortCellsPerThread;
    double dv2[1] = {0.0};  // |v' - v|^2, for the step-length estimate
    // One cell per pass: every wave of the grid walks the same slice k at
    // about the same time.  Loading all of a thread's cells first measured
    // slower at 2M cells (22.4 -> 25.7 us; groups of four cells: 23.5 us).
#pragma unroll
    for (int k = 0; k < kSortCellsPerThread; ++k) {
        const int c = base + k * kBlock + threadIdx.x;
        if (c < n_place) {
            const float2 s = wh[c], vv = v[c], uu = u[c], gg = gr[c];
            const float2 un = clamp_cell(make_float2(vv.x - alpha * gg.x, vv.y - alpha * gg.y), s, g);
            const float2 vn =
                clamp_cell(make_float2(un.x + coef * (un.x - uu.x), un.y + coef * (un.y - uu.y)), s, g);
            u_new[c] = un;
            v_new[c] = vn;
            if (count) atomicAdd(&hist[cell_key(vn, s, g, Mx, My)], 1);
            const double dx = (double)vn.x - vv.x, dy = (double)vn.y - vv.y;
            dv2[0] += dx * dx + dy * dy;
        }
    }
    if (dv2_part) block_reduce<1, 0>(dv2, dv2_part + blockIdx.x);
    if (count) {
        __syncthreads();
        for (int t = threadIdx.x; t < kSortTiles; t += kBlock)
            if (hist[t]) atomicAdd(&count[t], hist[t]);
    }
}

__global__ __launch_bounds__(kBlock) void k_probe(const float2 *__restrict__ v,
                                                  const float2 *__restrict__ gr,
                                                  const float2 *__restrict__ wh, int n_place,
                                                  GeomArgs g, const DevScalars *__restrict__ sc,
                                                  float2 *__restrict__ v_new)
{
    const float eps = sc->probe_eps;
    for (int c = blockIdx.x * kBlock + threadIdx.x; c < n_place; c += gridDim.x * kBlock) {
        const float2 s = wh[c], vv = v[c], gg = gr[c];
        v_new[c] = clamp_cell(make_float2(vv.x - eps * gg.x, vv.y - eps * gg.y), s, g);
    }
}

__global__ void k_cell_bins(const float2 *__restrict__ pos, const float2 *__restrict__ wh,
                            int n_place, GeomArgs g, int Mx, int My, int stretched,
                            int32_t *__restrict__ out)
{
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n_place; c += gridDim.x * blockDim.x) {
        Rect r;
        if (stretched)
            stretched_rect(pos[c], wh[c], g, Mx, My, r);
        else
            exact_rect(pos[c], wh[c], g, Mx, My, r);
        out[4 * c + 0] = r.i0;
        out[4 * c + 1] = r.i1;
        out[4 * c + 2] = r.j0;
        out[4 * c + 3] = r.j1;
    }
}

// ------------------------------------------------------------------ step control
//
// The decision of ep_step (ep_api.cpp) after every trial of a device-controlled
// run, on one thread: accept when alpha_hat >= 0.95 alpha_used (the model's
// proposal first, then the Lipschitz search, at most max_backtrack trials),
// then lambda *= mu(dHPWL), gamma(tau), stop at the overflow target or
// max_iters.  Same FP64 operations in the same order as the host, without
// multiply-add contraction (which the host compiler does not do either), so
// both controls take the same decisions.

__device__ inline float lg10_clamped(double v) { return (float)sched_log10(fmax(v, 1e-30)); }

// Step-model features of the coming iteration (order and definitions:
// include/eplace_mlp.h); the host's step_features computes the same values.
__device__ inline void ctl_features(const DevCtl &S, double bw, double isn, float *f)
{
#pragma clang fp contract(off)
    f[0] = lg10_clamped(S.alpha / bw);
    f[1] = lg10_clamped(S.h_alpha1 / bw);
    f[2] = lg10_clamped(S.h_alpha2 / bw);
    f[3] = (float)S.last_tau;
    f[4] = (float)(S.last_tau - S.h_prev_tau);
    f[5] = lg10_clamped(S.lambda / S.lambda0);
    f[6] = lg10_clamped(S.gamma / bw);
    f[7] = (float)((S.last_hpwl - S.h_prev_hpwl) / fmax(S.last_hpwl, 1e-30));
    f[8] = lg10_clamped(S.h_gnorm * isn);
    f[9] = lg10_clamped(S.h_dv * isn / bw);
    f[10] = lg10_clamped(S.h_dg * isn);
    f[11] = (float)S.coef;
    f[12] = (float)S.h_evals;
    f[13] = lg10_clamped(S.lambda / S.h_sum_ratio);
    f[14] = lg10_clamped(S.h_ratio);
    f[15] = (float)(S.iteration / 1000.0);
}

// Nesterov coefficient, features and the scalars of the iteration's first trial
// (nt, when given, gets the model flag and the features too).
__device__ inline void ctl_begin(DevCtl &S, DevScalars *sc, const CtlParams &p, NextTrial *nt = nullptr)
{
#pragma clang fp contract(off)
    S.ak1 = 0.5 * (1.0 + sqrt(4.0 * S.ak * S.ak + 1.0));
    S.coef = (S.ak - 1.0) / S.ak1;
    ctl_features(S, p.bw, p.inv_sqrt_np, S.feat);
    for (int k = 0; k < 16; ++k) sc->feat[k] = S.feat[k];
    if (nt)
        for (int k = 0; k < 16; ++k) nt->feat[k] = S.feat[k];
    S.trials = 0;
    S.t = 0;
    S.learned_ok = 0;
    const bool propose = p.learned && S.model_skip == 0;
    if (p.learned && !propose) S.model_skip--;
    S.learned_used = propose ? 1 : 0;
    S.phase = propose ? 1 : 0;
    if (propose) S.learned_proposals++;
    sc->alpha = (float)S.alpha;
    sc->coef = (float)S.coef;
    sc->gamma_inv = (float)(1.0 / S.gamma);
    sc->lambda = (float)S.lambda;
    sc->use_model = propose ? 1 : 0;
    if (nt) nt->use_model = propose ? 1 : 0;
}

// The decision after a trial whose reductions are R (one thread); S0 is the
// control state *ctl as the trial left it, the result goes back to *ctl.
__device__ inline void step_decide(DevCtl *__restrict__ ctl, const DevCtl &S0, const DevResult &R,
                                   DevScalars *__restrict__ sc, const CtlParams &p,
                                   ep_step_info *__restrict__ log, float *__restrict__ rec, NextTrial &nt)
{
#pragma clang fp contract(off)
    DevCtl S = S0;
    if (S.done) {  // (the halt flag was set with it)
        nt.halt = 1;
        return;
    }
    S.evals++;
    S.trials++;
    if (R.nonfinite) {
        S.error = 1;
        S.done = 1;
        sc->halt = 1;
        nt.halt = 1;
        *ctl = S;
        return;
    }
    const double used = R.alpha_used;
    const double nn = sqrt(R.dv2), dd = sqrt(R.dg2);
    const double ahat = dd > 0 ? nn / dd : used;
    bool accept = false;
    if (S.phase == 1) {
        if (isfinite(ahat) && ahat >= 0.95 * used) {
            accept = true;
            S.learned_ok = 1;
            S.learned_accepted++;
            S.model_backoff = 0;
        } else {  // fall back to the search from the Lipschitz step; the model sits out
            S.model_backoff = S.model_backoff ? min(2 * S.model_backoff, kModelBackoffMax) : 2;
            S.model_skip = S.model_backoff;
            S.phase = 0;
            S.t = 0;
            sc->use_model = 0;
            nt.use_model = 0;
            sc->alpha = (float)S.alpha;
        }
    } else if (ahat >= 0.95 * used || S.t == p.max_backtrack - 1) {
        accept = true;
    } else {
        S.t++;
        sc->alpha = (float)ahat;
    }
    if (accept) {
        if (S.n_rec < p.rec_cap) {
            float *r = rec + (size_t)S.n_rec * 17;
            for (int k = 0; k < 16; ++k) r[k] = S.feat[k];
            r[16] = (float)sched_log10(fmax(used / p.bw, 1e-30));
            S.n_rec++;
        }
        S.cur ^= 1;
        sc->sel = S.cur;
        S.alpha = ahat;
        S.ak = S.ak1;
        const double hp = R.hpwl, tau = R.tau;
        S.h_alpha2 = S.h_alpha1;
        S.h_alpha1 = used;
        S.h_prev_tau = S.last_tau;
        S.h_prev_hpwl = S.last_hpwl;
        S.h_ratio = ahat / fmax(used, 1e-30);
        S.h_gnorm = sqrt(R.gnorm2);
        S.h_dv = nn;
        S.h_dg = dd;
        S.h_sum_ratio = R.sum_f > 0 ? R.sum_gw / R.sum_f : 1.0;
        S.h_evals = S.trials;
        S.last_hpwl = hp;
        S.last_tau = tau;
        S.last_wa = R.wa;
        if (S.n_log < p.log_cap) {
            ep_step_info &e = log[S.n_log];
            e.iteration = S.iteration;
            e.evals = S.trials;
            e.learned_used = S.learned_used;
            e.learned_accepted = S.learned_ok;
            e.hpwl = hp;
            e.overflow = tau;
            e.lambda = S.lambda;
            e.gamma = S.gamma;
            e.alpha = S.alpha;
            e.wa = R.wa;
        }
        S.n_log++;
        const double dh = hp - S.prev_hpwl;
        double mu = dh < 0 ? p.mu_max : sched_pow(p.mu_max, 1.0 - dh / p.hpwl_ref);
        mu = fmin(fmax(mu, p.mu_min), p.mu_max);
        S.lambda *= mu;
        S.prev_hpwl = hp;
        double gam = p.gamma_base;
        if (p.gamma_schedule) gam *= sched_exp10((tau - 0.1) * 20.0 / 9.0 - 1.0);
        S.gamma = gam;
        S.iteration++;
        if (tau <= p.stop_overflow || S.iteration >= p.max_iters) {
            S.done = 1;
            sc->halt = 1;
            nt.halt = 1;
        } else {
            ctl_begin(S, sc, p, &nt);
        }
    }
    *ctl = S;
}

// Start of the first iteration of a device-controlled run (and the model's
// proposal for its first trial).
__global__ void k_step_begin(DevCtl *__restrict__ ctl, DevScalars *__restrict__ sc, CtlParams p,
                             const float *__restrict__ mlp_w)
{
    __shared__ int propose;
    __shared__ float feat[16];
    if (threadIdx.x == 0) {
        DevCtl S = *ctl;
        propose = 0;
        if (!S.done) {
            ctl_begin(S, sc, p);
            *ctl = S;
            propose = mlp_w && sc->use_model && !sc->halt;
            for (int k = 0; k < 16; ++k) feat[k] = sc->feat[k];
        }
    }
    __syncthreads();
    if (mlp_w) mlp_propose_block(mlp_w, feat, sc, propose != 0);
}

GeomArgs geom_args(const Ctx &c)
{
    GeomArgs g;
    g.v[0] = c.xl; g.v[1] = c.yl; g.v[2] = c.xh; g.v[3] = c.yh;
    g.v[4] = c.bw; g.v[5] = c.bh; g.v[6] = c.sqrt2_bw; g.v[7] = c.sqrt2_bh;
    return g;
}

int grid_for(int64_t n, int per_block, int cap)
{
    int64_t b = (n + per_block - 1) / per_block;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (int)b;
}

}  // namespace

// ------------------------------------------------------------------ launchers

void launch_fill_maps(Ctx &c, hipStream_t s)
{
    const int nb = c.Mx * c.My;
    hipLaunchKernelGGL(k_fill_maps, dim3(grid_for(nb / 2, kBlock * 2, 4096)), dim3(kBlock), 0, s,
                       c.acc_field, c.acc_mov, nb, c.d_scal);
}

void launch_sort(Ctx &c, Pair<const float2> pos, bool counted, hipStream_t s)
{
    const int blocks = grid_for(c.n_place, kBlock * kSortCellsPerThread, INT_MAX);
    if (!counted)
        hipLaunchKernelGGL(k_sort_count, dim3(blocks), dim3(kBlock), 0, s, pos, c.wh, c.n_place,
                           geom_args(c), c.Mx, c.My, c.sort_count, c.d_scal);
    hipLaunchKernelGGL(k_sort_scan, dim3(1), dim3(kSortScanThreads), 0, s, c.sort_count, c.sort_off,
                       c.dens_tab, c.dens.level, kBlock * c.dens.cpt, c.d_scal);
    hipLaunchKernelGGL(k_sort_scatter, dim3(blocks), dim3(kBlock), 0, s, pos, c.wh, c.n_place,
                       c.n_mov, geom_args(c), c.Mx, c.My, c.sort_off, c.sorted, c.d_scal);
}

void launch_density(Ctx &c, hipStream_t s)
{
    const DensPlan &dp = c.dens;
    const int blocks = dens_grid(dp, c.n_place);
    const size_t lds = dens_lds_bytes(dp);
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(kBlock), lds, s, c.sorted, c.sort_off, c.dens_tab,
                           dp.level, dp.ext, dp.wcap, c.n_place, geom_args(c), c.Mx, c.My,
                           c.acc_field, c.acc_mov, c.d_scal);
    };
    if (dp.cpt == 2) go(k_density<2>);
    else if (dp.cpt == 8) go(k_density<8>);
    else if (dp.cpt == 16) go(k_density<16>);
    else go(k_density<4>);
}

void launch_overflow(Ctx &c, hipStream_t s)
{
    const int nb = c.Mx * c.My;
    const float cap = c.prm.target_density * c.bw * c.bh;
    const double unit = (double)c.bw * (double)c.bh / kDensFixUnits;
    hipLaunchKernelGGL(k_overflow, dim3(c.ovf_blocks), dim3(kBlock), 0, s, c.acc_field, c.acc_mov,
                       c.map_fix, unit, c.prm.target_density, cap, nb, c.map_field, c.map_mov,
                       c.part + c.part_ovf_off, c.d_scal);
}

void launch_map_range(Ctx &c, hipStream_t s)
{
    hipLaunchKernelGGL(k_map_range, dim3(1), dim3(kBlock), 0, s, c.part + c.part_ovf_off, c.ovf_blocks,
                       c.map_mean, c.field_max, c.d_scal);
}

void launch_wa(Ctx &c, Pair<const float2> pos, hipStream_t s)
{
    if (c.wa_blocks == 0) return;
    hipLaunchKernelGGL(k_wa, dim3(c.wa_part_blocks), dim3(kBlock), 0, s, c.buckets, c.net_start,
                       c.pin_cell, c.pin_off, pos, c.pin_grad, c.d_scal, c.wa_blocks,
                       c.part + c.part_wa_off);
}

void launch_pin_sum(Ctx &c, hipStream_t s)
{
    if (c.n_mov == 0) return;
    hipLaunchKernelGGL(k_pin_sum, dim3(grid_for(c.n_mov, kBlock, kCellGridCap)), dim3(kBlock), 0, s,
                       c.cell_pin_start, c.cell_pins, c.pin_grad, c.n_mov, c.gw_cell, c.d_scal);
}

void launch_cell_grad(Ctx &c, Pair<const float2> pos, Pair<const float2> vref,
                      Pair<const float2> gref, Pair<float2> gout, float2 *gw_out, float2 *f_out,
                      hipStream_t s, const float2 *gw_cell)
{
    // Every caller sets both entries of a pair or neither.
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(c.cell_blocks), dim3(kBlock), 0, s, pos, c.wh, c.n_place, c.n_mov,
                           geom_args(c), c.Mx, c.My, c.cell_pin_start, c.cell_pins, c.pin_grad, c.exy,
                           vref, gref, gout, gw_out, f_out, c.d_scal, c.part + c.part_cell_off, gw_cell);
    };
    switch ((vref.p[0] ? 1 : 0) | (gref.p[0] ? 2 : 0) | (gw_cell ? 4 : 0)) {
    case 0: go(k_cell_grad<false, false, false>); break;
    case 1: go(k_cell_grad<true, false, false>); break;
    case 2: go(k_cell_grad<false, true, false>); break;
    case 3: go(k_cell_grad<true, true, false>); break;
    case 4: go(k_cell_grad<false, false, true>); break;
    case 5: go(k_cell_grad<true, false, true>); break;
    case 6: go(k_cell_grad<false, true, true>); break;
    default: go(k_cell_grad<true, true, true>); break;
    }
}

void launch_finalize(Ctx &c, hipStream_t s, bool trial)
{
    StepCtlArgs none{};
    hipLaunchKernelGGL(k_finalize, dim3(kFinalizeBlocks), dim3(kFinalizeThreads), 0, s, c.part,
                       c.part_wa_off, c.wa_part_blocks, c.part_ovf_off, c.ovf_blocks, c.part_cell_off,
                       c.cell_blocks, c.part_trial_off, trial ? c.trial_blocks : 0, c.mov_area, c.d_scal,
                       c.d_res, c.fin, c.fin_ticket, none);
}

void launch_finalize_decide(Ctx &c, const CtlParams &p, hipStream_t s)
{
    const StepCtlArgs a{c.d_ctl, c.d_log, c.d_rec, p, p.learned ? (const float *)c.d_mlp_w : nullptr};
    hipLaunchKernelGGL(k_finalize, dim3(kFinalizeBlocks), dim3(kFinalizeThreads), 0, s, c.part,
                       c.part_wa_off, c.wa_part_blocks, c.part_ovf_off, c.ovf_blocks, c.part_cell_off,
                       c.cell_blocks, c.part_trial_off, c.trial_blocks, c.mov_area, c.d_scal, c.d_res, c.fin,
                       c.fin_ticket, a);
}

void launch_trial(Ctx &c, Pair<const float2> v, Pair<const float2> u, Pair<const float2> g,
                  Pair<float2> v_new, Pair<float2> u_new, bool count, hipStream_t s)
{
    const int blocks = grid_for(c.n_place, kBlock * kSortCellsPerThread, INT_MAX);
    hipLaunchKernelGGL(k_trial, dim3(blocks), dim3(kBlock), 0, s, v, u, g, c.wh, c.n_place,
                       geom_args(c), c.Mx, c.My, c.d_scal, v_new, u_new,
                       count ? c.sort_count : nullptr, count ? c.part + c.part_trial_off : nullptr);
}

void launch_step_begin(Ctx &c, const CtlParams &p, hipStream_t s)
{
    hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(kMlpHid), 0, s, c.d_ctl, c.d_scal, p,
                       p.learned ? (const float *)c.d_mlp_w : nullptr);
}

void launch_probe(Ctx &c, const float2 *v, const float2 *g, float2 *v_new, hipStream_t s)
{
    hipLaunchKernelGGL(k_probe, dim3(grid_for(c.n_place, kBlock, 8192)), dim3(kBlock), 0, s, v, g,
                       c.wh, c.n_place, geom_args(c), c.d_scal, v_new);
}

void launch_wa_packet(const float *data, int net_size, int n_nets, int groups, float ginv,
                      float *out, hipStream_t s)
{
    const int threads = groups * 8;
    hipLaunchKernelGGL(k_wa_packet, dim3((threads + 255) / 256), dim3(256), 0, s, data, net_size,
                       n_nets, groups, ginv, out);
}

void launch_wa_packet_fast64(const float *data, int net_size, int n_nets, int groups, double gamma, float *out,
                             hipStream_t s)
{
    const int threads = groups * 8;
    hipLaunchKernelGGL(k_wa_packet_fast64, dim3((threads + 255) / 256), dim3(256), 0, s, data, net_size,
                       n_nets, groups, gamma, out);
}

// Device staging of ep_wa_packet_noalloc: packet data in the first half, the
// partials in the second.  Each GPU holds its own copy until the process ends;
// host code reaches it only through the address hipGetSymbolAddress gives on
// that device (the array's name is a host-side stand-in, not a device address).
constexpr int64_t kPacketStage = EP_WA_PACKET_NOALLOC_MAX;
static __device__ __align__(256) float scratch[2 * kPacketStage];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_stage_mu;
float *g_stage[kMaxDevices] = {};
}  // namespace

ep_status wa_packet_stage(float **stage)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
        return fail(EP_EDEVICE, "ep_wa_packet_noalloc: no current HIP device (or index >= %d)", kMaxDevices);
    std::lock_guard<std::mutex> lock(g_stage_mu);
    if (!g_stage[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(EP_EDEVICE, "ep_wa_packet_noalloc: hipGetSymbolAddress on device %d: %s", dev,
                        hipGetErrorString(e));
        g_stage[dev] = static_cast<float *>(p);
    }
    *stage = g_stage[dev];
    return EP_OK;
}

void launch_cell_bins(Ctx &c, const float2 *pos, int stretched, int32_t *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_cell_bins, dim3(grid_for(c.n_place, kBlock, 8192)), dim3(kBlock), 0, s,
                       pos, c.wh, c.n_place, geom_args(c), c.Mx, c.My, stretched, out);
}

}  // namespace ep
