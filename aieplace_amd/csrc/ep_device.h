// ep_device.h -- device helpers shared by the single-instance and batched
// kernels: bin geometry (bit-exact FP32 bin ranges), wavefront/block
// reductions, and the per-net WA gradient body.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>

#include "ep_ctx.h"

namespace ep {
namespace dev {

constexpr int kSortScanThreads = 1024;
constexpr int kXcds = 8;  // MI355X: 8 XCDs, each with its own L2

// Workgroups are dispatched to the XCDs round-robin (physical block b runs on
// XCD b % 8).  Returns the logical block of physical block b in a grid of n so
// that each XCD runs one contiguous range of logical blocks: neighbouring
// blocks, which gather from the same nets, cells and bins, then share an L2.
__device__ inline int xcd_block(int b, int n)
{
    const int q = n / kXcds, r = n % kXcds, x = b % kXcds;
    return x * q + min(x, r) + b / kXcds;
}

// As xcd_block, for grids whose work varies along the index range: the grid is
// cut into groups of kXcds * S consecutive logical blocks and each XCD takes
// one run of S blocks from every group (runs stay within one L2, and every XCD
// sees the whole range, so no XCD gets only the heavy or idle blocks).  Blocks
// past the last whole group keep their index.
__device__ inline int xcd_run_block(int b, int n, int S)
{
    const int G = kXcds * S, full = n / G * G;
    if (b >= full) return b;
    const int x = b % kXcds, k = b / kXcds;
    return ((k / S) * kXcds + x) * S + (k % S);
}
constexpr int kXcdRun = 16;

struct Geom {
    float xl, yl, xh, yh, bw, bh, sqrt2_bw, sqrt2_bh;
    int32_t Mx, My;
};

__device__ inline Geom geom_of(const float *g4, int Mx, int My)
{
    Geom g;
    g.xl = g4[0];
    g.yl = g4[1];
    g.xh = g4[2];
    g.yh = g4[3];
    g.bw = g4[4];
    g.bh = g4[5];
    g.sqrt2_bw = g4[6];
    g.sqrt2_bh = g4[7];
    g.Mx = Mx;
    g.My = My;
    return g;
}

struct GeomArgs {
    float v[8];
};

// Rectangle of a cell in bin units plus its clamped bin range.  The float
// operation sequence is fixed (explicit round-to-nearest intrinsics, no
// contraction) so bin indices match the FP32 host computation bit for bit.
struct Rect {
    float tx0, tx1, ty0, ty1, ratio;
    int i0, i1, j0, j1;
};

__device__ inline void bin_range(float x0, float x1, float y0, float y1, const GeomArgs &g,
                                 int Mx, int My, Rect &r)
{
    r.tx0 = __fdiv_rn(__fsub_rn(x0, g.v[0]), g.v[4]);
    r.tx1 = __fdiv_rn(__fsub_rn(x1, g.v[0]), g.v[4]);
    r.ty0 = __fdiv_rn(__fsub_rn(y0, g.v[1]), g.v[5]);
    r.ty1 = __fdiv_rn(__fsub_rn(y1, g.v[1]), g.v[5]);
    r.i0 = max(0, (int)floorf(r.tx0));
    r.i1 = min(Mx - 1, (int)floorf(r.tx1));
    r.j0 = max(0, (int)floorf(r.ty0));
    r.j1 = min(My - 1, (int)floorf(r.ty1));
}

// ePlace local smoothing: cells narrower than sqrt(2) bins are widened around
// their centre and their density scaled by area / smoothed area.
__device__ inline void stretched_rect(float2 p, float2 s, const GeomArgs &g, int Mx, int My,
                                      Rect &r)
{
    float cx = __fadd_rn(p.x, __fmul_rn(0.5f, s.x));
    float cy = __fadd_rn(p.y, __fmul_rn(0.5f, s.y));
    float sx = fmaxf(s.x, g.v[6]);
    float sy = fmaxf(s.y, g.v[7]);
    float hx = __fmul_rn(0.5f, sx), hy = __fmul_rn(0.5f, sy);
    bin_range(__fsub_rn(cx, hx), __fadd_rn(cx, hx), __fsub_rn(cy, hy), __fadd_rn(cy, hy), g, Mx,
              My, r);
    r.ratio = (s.x * s.y) / (sx * sy);
}

__device__ inline void exact_rect(float2 p, float2 s, const GeomArgs &g, int Mx, int My, Rect &r)
{
    bin_range(p.x, __fadd_rn(p.x, s.x), p.y, __fadd_rn(p.y, s.y), g, Mx, My, r);
    r.ratio = 1.0f;
}

__device__ inline float ovl(float t0, float t1, int i)
{
    return fminf(t1, (float)(i + 1)) - fmaxf(t0, (float)i);
}

// DPP lane move of a double (two 32-bit moves).
template <int CTRL>
__device__ inline double dpp_mov_d(double v)
{
    const long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(u & 0xffffffffll), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(u >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ inline double readlane_d(double v, int lane)
{
    const long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(u & 0xffffffffll), lane);
    const int hi = __builtin_amdgcn_readlane((int)(u >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Wave-wide sum / max (every lane gets the result), in a fixed order: DPP
// all-reduce inside each 16-lane row (as grp_reduce), then the four row
// totals in row order.
__device__ inline double wave_sum(double v)
{
    v += dpp_mov_d<0xB1>(v);
    v += dpp_mov_d<0x4E>(v);
    v += dpp_mov_d<0x141>(v);
    v += dpp_mov_d<0x140>(v);
    return ((readlane_d(v, 0) + readlane_d(v, 16)) + readlane_d(v, 32)) + readlane_d(v, 48);
}

__device__ inline double wave_max(double v)
{
    v = fmax(v, dpp_mov_d<0xB1>(v));
    v = fmax(v, dpp_mov_d<0x4E>(v));
    v = fmax(v, dpp_mov_d<0x141>(v));
    v = fmax(v, dpp_mov_d<0x140>(v));
    return fmax(fmax(readlane_d(v, 0), readlane_d(v, 16)), fmax(readlane_d(v, 32), readlane_d(v, 48)));
}

// Block-wide sum (and max) of NV doubles; thread 0 writes out[k].
template <int NV, int NMAX, int NT = kBlock>
__device__ inline void block_reduce(double (&v)[NV], double *out)
{
    constexpr int kWaves = NT / 64;
    __shared__ double red[kWaves][NV];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = (k < NV - NMAX) ? wave_sum(v[k]) : wave_max(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) red[wv][k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double a = red[0][k];
            for (int w = 1; w < kWaves; ++w)
                a = (k < NV - NMAX) ? a + red[w][k] : fmax(a, red[w][k]);
            out[k] = a;
        }
    }
}

// Agent-scope load of a double written by another block of the same launch
// (after that block's release fence and an atomic ticket).
__device__ inline double load_agent(const double *p)
{
    return __builtin_bit_cast(double, __hip_atomic_load(reinterpret_cast<const unsigned long long *>(p),
                                                        __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

// ------------------------------------------------------------------ maps

__device__ inline int morton6(int i, int j)
{
    int m = 0;
#pragma unroll
    for (int b = 0; b < 6; ++b) m |= (((i >> b) & 1) << (2 * b)) | (((j >> b) & 1) << (2 * b + 1));
    return m;
}

__device__ inline int cell_key(float2 p, float2 s, const GeomArgs &g, int Mx, int My)
{
    Rect r;
    stretched_rect(p, s, g, Mx, My, r);
    const int i = min(max(r.i0, 0), Mx - 1), j = min(max(r.j0, 0), My - 1);
    return morton6((i * kSortSide) / Mx, (j * kSortSide) / My);
}

// All-reduce over aligned groups of G lanes.  Inside a 16-lane row the steps
// are DPP lane moves folded into the ALU op (no LDS crossbar): xor 1 and xor 2
// by quad permutes, then the half-row and row mirrors, which pair every lane
// with one of the other quad / half-row once those are uniform.  Groups of 32
// and 64 finish with cross-row shuffles.
template <int CTRL>
__device__ inline float dpp_mov(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

template <int G, class Op>
__device__ inline float grp_reduce(float v, Op op)
{
    if (G >= 2) v = op(v, dpp_mov<0xB1>(v));   // quad_perm [1,0,3,2]: lane ^ 1
    if (G >= 4) v = op(v, dpp_mov<0x4E>(v));   // quad_perm [2,3,0,1]: lane ^ 2
    if (G >= 8) v = op(v, dpp_mov<0x141>(v));  // row_half_mirror: lane i <-> 7 - i
    if (G >= 16) v = op(v, dpp_mov<0x140>(v)); // row_mirror: lane i <-> 15 - i
    if (G >= 32) v = op(v, __shfl_xor(v, 16, 64));
    if (G >= 64) v = op(v, __shfl_xor(v, 32, 64));
    return v;
}

struct OpMax {
    __device__ float operator()(float a, float b) const { return fmaxf(a, b); }
};
struct OpMin {
    __device__ float operator()(float a, float b) const { return fminf(a, b); }
};
struct OpSum {
    __device__ float operator()(float a, float b) const { return a + b; }
};

template <int G>
__device__ inline float grp_max(float v)
{
    return grp_reduce<G>(v, OpMax());
}
template <int G>
__device__ inline float grp_min(float v)
{
    return grp_reduce<G>(v, OpMin());
}
template <int G>
__device__ inline float grp_sum(float v)
{
    return grp_reduce<G>(v, OpSum());
}

// One net per group of G lanes.  Shifted WA form (algebraically equal to
// AIEplace.cpp:660-668): with d+ = x - x_max, a+ = e^{d+/g}, m+ = sum d+ a+ / sum a+,
//   dWA+/dx_i = a+_i/b+ * (1 + (d+_i - m+)/g),
// and symmetrically for the minus side, so FP32 never forms x/g - c/(g b).
template <int G>
__device__ inline void wa_net(int s, int e, bool active, int lane, float ginv,
                              const int32_t *__restrict__ pin_cell,
                              const float2 *__restrict__ pin_off, const float2 *__restrict__ pos,
                              float2 *__restrict__ pin_grad, double &hp_acc, double &wa_acc)
{
    const int d = e - s;
    if (G < 64 || d <= G) {
        // one pin per lane, kept in registers
        const bool valid = active && lane < d;
        float px = 0.f, py = 0.f;
        if (valid) {
            const int p = s + lane;
            const int c = pin_cell[p];
            const float2 o = pin_off[p];
            const float2 q = pos[c];
            px = q.x + o.x;
            py = q.y + o.y;
        }
        const float xmx = grp_max<G>(valid ? px : -FLT_MAX);
        const float xmn = grp_min<G>(valid ? px : FLT_MAX);
        const float ymx = grp_max<G>(valid ? py : -FLT_MAX);
        const float ymn = grp_min<G>(valid ? py : FLT_MAX);
        const float dxp = px - xmx, dxm = px - xmn, dyp = py - ymx, dym = py - ymn;
        const float axp = valid ? __expf(dxp * ginv) : 0.f;
        const float axm = valid ? __expf(-dxm * ginv) : 0.f;
        const float ayp = valid ? __expf(dyp * ginv) : 0.f;
        const float aym = valid ? __expf(-dym * ginv) : 0.f;
        const float bxp = grp_sum<G>(axp), bxm = grp_sum<G>(axm);
        const float byp = grp_sum<G>(ayp), bym = grp_sum<G>(aym);
        const float cxp = grp_sum<G>(dxp * axp), cxm = grp_sum<G>(dxm * axm);
        const float cyp = grp_sum<G>(dyp * ayp), cym = grp_sum<G>(dym * aym);
        // one hardware reciprocal (1 ulp) per sum, shared by the lane's quotients
        const float rxp = __builtin_amdgcn_rcpf(bxp), rxm = __builtin_amdgcn_rcpf(bxm);
        const float ryp = __builtin_amdgcn_rcpf(byp), rym = __builtin_amdgcn_rcpf(bym);
        const float mxp = cxp * rxp, mxm = cxm * rxm, myp = cyp * ryp, mym = cym * rym;
        if (valid) {
            const float gx = axp * rxp * (1.f + (dxp - mxp) * ginv) -
                             axm * rxm * (1.f - (dxm - mxm) * ginv);
            const float gy = ayp * ryp * (1.f + (dyp - myp) * ginv) -
                             aym * rym * (1.f - (dym - mym) * ginv);
            pin_grad[s + lane] = make_float2(gx, gy);
        }
        if (active && lane == 0) {
            hp_acc += (double)(xmx - xmn) + (double)(ymx - ymn);
            wa_acc += (double)((xmx - xmn) + mxp - mxm) + (double)((ymx - ymn) + myp - mym);
        }
    } else {
        // nets wider than a wavefront: three strided passes over the pins
        float xmx = -FLT_MAX, xmn = FLT_MAX, ymx = -FLT_MAX, ymn = FLT_MAX;
        for (int p = s + lane; p < e; p += G) {
            const float2 q = pos[pin_cell[p]], o = pin_off[p];
            const float px = q.x + o.x, py = q.y + o.y;
            xmx = fmaxf(xmx, px); xmn = fminf(xmn, px);
            ymx = fmaxf(ymx, py); ymn = fminf(ymn, py);
        }
        xmx = grp_max<G>(xmx); xmn = grp_min<G>(xmn);
        ymx = grp_max<G>(ymx); ymn = grp_min<G>(ymn);
        float bxp = 0, bxm = 0, byp = 0, bym = 0, cxp = 0, cxm = 0, cyp = 0, cym = 0;
        for (int p = s + lane; p < e; p += G) {
            const float2 q = pos[pin_cell[p]], o = pin_off[p];
            const float dxp = q.x + o.x - xmx, dxm = q.x + o.x - xmn;
            const float dyp = q.y + o.y - ymx, dym = q.y + o.y - ymn;
            const float axp = __expf(dxp * ginv), axm = __expf(-dxm * ginv);
            const float ayp = __expf(dyp * ginv), aym = __expf(-dym * ginv);
            bxp += axp; bxm += axm; byp += ayp; bym += aym;
            cxp += dxp * axp; cxm += dxm * axm; cyp += dyp * ayp; cym += dym * aym;
        }
        bxp = grp_sum<G>(bxp); bxm = grp_sum<G>(bxm); byp = grp_sum<G>(byp); bym = grp_sum<G>(bym);
        cxp = grp_sum<G>(cxp); cxm = grp_sum<G>(cxm); cyp = grp_sum<G>(cyp); cym = grp_sum<G>(cym);
        const float rxp = __builtin_amdgcn_rcpf(bxp), rxm = __builtin_amdgcn_rcpf(bxm);
        const float ryp = __builtin_amdgcn_rcpf(byp), rym = __builtin_amdgcn_rcpf(bym);
        const float mxp = cxp * rxp, mxm = cxm * rxm, myp = cyp * ryp, mym = cym * rym;
        for (int p = s + lane; p < e; p += G) {
            const float2 q = pos[pin_cell[p]], o = pin_off[p];
            const float dxp = q.x + o.x - xmx, dxm = q.x + o.x - xmn;
            const float dyp = q.y + o.y - ymx, dym = q.y + o.y - ymn;
            const float axp = __expf(dxp * ginv), axm = __expf(-dxm * ginv);
            const float ayp = __expf(dyp * ginv), aym = __expf(-dym * ginv);
            const float gx = axp * rxp * (1.f + (dxp - mxp) * ginv) -
                             axm * rxm * (1.f - (dxm - mxm) * ginv);
            const float gy = ayp * ryp * (1.f + (dyp - myp) * ginv) -
                             aym * rym * (1.f - (dym - mym) * ginv);
            pin_grad[p] = make_float2(gx, gy);
        }
        if (lane == 0) {
            hp_acc += (double)(xmx - xmn) + (double)(ymx - ymn);
            wa_acc += (double)((xmx - xmn) + mxp - mxm) + (double)((ymx - ymn) + myp - mym);
        }
    }
}

// ------------------------------------------------------------------ density scatter
//
// Cells are counting-sorted by the coarse tile of their smoothed rectangle's
// lower-left bin (kSortSide x kSortSide coarse tiles in Morton order).  A super tile of level
// k is an aligned block of 2^k x 2^k coarse tiles, i.e. 4^k consecutive keys,
// so its cells are one contiguous run of the sorted order.  Every cell of a
// super tile starts inside the super tile's bins, so its bins lie in the
// window [x0 - 1, x1 + ext) x [y0 - 1, y1 + ext), where ext bounds a smoothed
// cell's extent in bins; the -1 absorbs FP32 rounding of the exact rectangle
// against the smoothed one.  Blocks take up to K cells of one super tile and
// accumulate the window in LDS.

// Exclusive scan of the per-key counts (one block of kSortScanThreads):
// off[key] = first slot of key, counts cleared, and the density block table
// dtab[s] = first block of super tile s (level `level`, K cells per block),
// dtab[NS] = total blocks.
// Inclusive prefix sum of one int per thread over a kSortScanThreads block:
// shuffles within each wave, then the wave totals (two block barriers).
__device__ inline int block_inclusive_scan(int x)
{
    constexpr int kWaves = kSortScanThreads / 64;
    __shared__ int wsum[kWaves];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    if (w == 0) {
        int t = lane < kWaves ? wsum[lane] : 0;
#pragma unroll
        for (int d = 1; d < kWaves; d <<= 1) {
            const int y = __shfl_up(t, d, 64);
            if (lane >= d) t += y;
        }
        if (lane < kWaves) wsum[lane] = t;
    }
    __syncthreads();
    return w > 0 ? x + wsum[w - 1] : x;
}

__device__ inline void sort_scan_block(int32_t *__restrict__ count, int32_t *__restrict__ off,
                                       int32_t *__restrict__ dtab, int level, int K)
{
    constexpr int per = kSortTiles / kSortScanThreads;
    __shared__ int so[kSortTiles + 1];
    const int t = threadIdx.x;
    int v[per], s = 0;
#pragma unroll
    for (int k = 0; k < per; ++k) {
        v[k] = count[t * per + k];
        s += v[k];
    }
    int run = block_inclusive_scan(s) - s;
#pragma unroll
    for (int k = 0; k < per; ++k) {
        off[t * per + k] = run;
        so[t * per + k] = run;
        run += v[k];
        count[t * per + k] = 0;
    }
    if (t == kSortScanThreads - 1) so[kSortTiles] = run;
    __syncthreads();
    // blocks per super tile, then the same scan over super tiles
    const int ns = kSortTiles >> (2 * level);
    int nb[per], sb = 0;
#pragma unroll
    for (int k = 0; k < per; ++k) {
        const int st = t * per + k;
        nb[k] = 0;
        if (st < ns) {
            const int cnt = so[(st + 1) << (2 * level)] - so[st << (2 * level)];
            nb[k] = (cnt + K - 1) / K;
        }
        sb += nb[k];
    }
    __syncthreads();  // the wave totals of the first scan are read before the second reuses them
    run = block_inclusive_scan(sb) - sb;
#pragma unroll
    for (int k = 0; k < per; ++k) {
        const int st = t * per + k;
        if (st < ns) dtab[st] = run;
        run += nb[k];
    }
    if (t == kSortScanThreads - 1) dtab[ns] = run;
}

// Super tile s of density block b (dtab as written by sort_scan_block).
__device__ inline int dens_super_tile(const int32_t *__restrict__ dtab, int ns, int b)
{
    int lo = 0, hi = ns;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (dtab[mid] <= b) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Cells [c0, c1) of the sorted order in block `chunk` of super tile s.  After
// the sort's scatter pass off[key] holds the END of key's run (clamped to the
// n sorted cells).
__device__ inline void dens_cell_range(const int32_t *__restrict__ off, int s, int level, int chunk,
                                       int K, int n, int &c0, int &c1)
{
    const int key0 = s << (2 * level), key1 = (s + 1) << (2 * level);
    const int start = key0 > 0 ? off[key0 - 1] : 0;
    const int stop = min(off[key1 - 1], n);
    c0 = start + chunk * K;
    c1 = min(c0 + K, stop);
}

// Bin window of super tile s (see above).
__device__ inline void super_window(int s, int level, int Mx, int My, int ext, int &x0, int &y0,
                                    int &nx, int &ny)
{
    int sx = 0, sy = 0;
    for (int b = 0; b < 6 - level; ++b) {
        sx |= ((s >> (2 * b)) & 1) << b;
        sy |= ((s >> (2 * b + 1)) & 1) << b;
    }
    const int cx0 = sx << level, cy0 = sy << level, side = 1 << level;
    const int bx0 = (cx0 * Mx + kSortSide - 1) / kSortSide;
    const int by0 = (cy0 * My + kSortSide - 1) / kSortSide;
    const int bx1 = ((cx0 + side) * Mx + kSortSide - 1) / kSortSide;
    const int by1 = ((cy0 + side) * My + kSortSide - 1) / kSortSide;
    x0 = max(0, bx0 - 1);
    y0 = max(0, by0 - 1);
    nx = min(Mx, bx1 + ext) - x0;
    ny = min(My, by1 + ext) - y0;
}

// Packed sorted cell: (x, y, w, h) with w negated for movable cells (they also
// enter the movable-area map).  Written by the sort's scatter pass.
__device__ inline float4 pack_cell(float2 p, float2 s, bool movable)
{
    return make_float4(p.x, p.y, movable ? -s.x : s.x, s.y);
}

// The density maps accumulate in 64-bit fixed point, 2^32 units per bin area:
// integer adds are associative, so the maps (and everything computed from
// them) do not depend on the order in which blocks and lanes add, and LDS
// integer atomics run several times faster than LDS float atomics on CDNA.
// A contribution is at most one bin area, so a bin overflows only past 2^31
// bin areas of overlap.
constexpr float kDensFix = (float)kDensFixUnits;

__device__ inline unsigned long long dens_fix(float v) { return (unsigned long long)(v * kDensFix); }

// Bin values from the accumulators, one rounding sequence for every kernel that
// converts them (single and batched): the field map is
// rho_t * fixed + smoothed cells, and a bin overflows by movable + rho_t *
// fixed - rho_t * area (cap).
__device__ inline float field_bin(float fx, unsigned long long af, double unit, float rho_t)
{
    return __fmaf_rn(rho_t, fx, (float)((double)af * unit));
}
__device__ inline float mov_bin(unsigned long long am, double unit) { return (float)((double)am * unit); }
__device__ inline float ovf_bin(float mv, float fx, float rho_t, float cap)
{
    return __fsub_rn(__fmaf_rn(rho_t, fx, mv), cap);
}

// Adds the overlaps of rectangle r (in bin areas, scaled by sc) to plane
// `plane` of the LDS window (x0, y0, nx, ny) when r lies inside it, else
// straight into `acc`.  Every path adds dens_fix(ox * (oy * sc)) per bin, so
// the maps depend neither on the order of the cells nor on which window (if
// any) a cell is scattered through.  The window is the dynamic LDS block, declared here and
// not passed as a pointer so the adds compile to LDS atomics.  Bins of
// [i0, i1] x [j0, j1] never have a negative overlap, so no test is needed per
// bin; the y overlaps are shared by every row of the rectangle.
__device__ inline void scatter_rect(const Rect &r, float sc, int plane,
                                    unsigned long long *__restrict__ acc, int x0, int y0, int nx,
                                    int ny, int My, bool lds, int wcap)
{
    extern __shared__ unsigned long long dens_win[];
    if (r.i0 > r.i1 || r.j0 > r.j1) return;
    const int dj = r.j1 - r.j0;
    if (lds && dj < 4 && r.i0 >= x0 && r.i1 < x0 + nx && r.j0 >= y0 && r.j1 < y0 + ny) {
        float oy[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) oy[q] = q <= dj ? ovl(r.ty0, r.ty1, r.j0 + q) * sc : 0.0f;
        int w = plane * wcap + (r.i0 - x0) * ny + (r.j0 - y0);
        for (int i = r.i0; i <= r.i1; ++i, w += ny) {
            const float ox = ovl(r.tx0, r.tx1, i);
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (q <= dj) atomicAdd(&dens_win[w + q], dens_fix(ox * oy[q]));
        }
        return;
    }
    for (int i = r.i0; i <= r.i1; ++i) {
        const float ox = ovl(r.tx0, r.tx1, i);
        for (int j = r.j0; j <= r.j1; ++j) {
            // the same product as the window path's, so that a bin's value does not
            // depend on which path a rectangle takes
            const unsigned long long v = dens_fix(ox * (ovl(r.ty0, r.ty1, j) * sc));
            const unsigned li = (unsigned)(i - x0), lj = (unsigned)(j - y0);
            if (lds && li < (unsigned)nx && lj < (unsigned)ny)
                atomicAdd(&dens_win[plane * wcap + li * ny + lj], v);
            else
                atomicAdd(&acc[(size_t)i * My + j], v);
        }
    }
}

// Density scatter of the sorted cells [c0, c1) through the LDS window
// (x0, y0, nx, ny): 2 * wcap 64-bit accumulators of dynamic LDS (field plane:
// smoothed rectangles scaled by their area ratio; movable plane: exact
// rectangles of movable cells), flushed into acc_field / acc_mov with one
// global atomic per non-zero bin.
template <int CPT>
__device__ inline void density_window(const float4 *__restrict__ sorted, int c0, int c1,
                                      const GeomArgs &g, int Mx, int My, int x0, int y0, int nx,
                                      int ny, unsigned long long *__restrict__ acc_field,
                                      unsigned long long *__restrict__ acc_mov, int wcap)
{
    extern __shared__ unsigned long long dens_win[];
    const int nw = nx * ny;
    const bool lds = nw <= wcap;
    if (lds) {
        for (int t = threadIdx.x; t < nw; t += kBlock) {
            dens_win[t] = 0ull;
            dens_win[wcap + t] = 0ull;
        }
        __syncthreads();
    }
    float4 cell[CPT];  // loads first, so they are in flight together
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int idx = c0 + k * kBlock + threadIdx.x;
        cell[k] = idx < c1 ? sorted[idx] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        if (c0 + k * kBlock + (int)threadIdx.x >= c1) break;
        const bool movable = cell[k].z < 0.0f;
        const float2 p = make_float2(cell[k].x, cell[k].y);
        const float2 s = make_float2(fabsf(cell[k].z), cell[k].w);
        Rect r;
        stretched_rect(p, s, g, Mx, My, r);
        scatter_rect(r, r.ratio, 0, acc_field, x0, y0, nx, ny, My, lds, wcap);
        if (movable) {
            exact_rect(p, s, g, Mx, My, r);
            scatter_rect(r, 1.0f, 1, acc_mov, x0, y0, nx, ny, My, lds, wcap);
        }
    }
    if (lds) {
        __syncthreads();
        for (int t = threadIdx.x; t < nw; t += kBlock) {
            const int i = x0 + t / ny, j = y0 + t % ny;
            const size_t b = (size_t)i * My + j;
            const unsigned long long a = dens_win[t], m = dens_win[wcap + t];
            if (a) atomicAdd(&acc_field[b], a);
            if (m) atomicAdd(&acc_mov[b], m);
        }
    }
}

// ------------------------------------------------------------------ per-cell gradient

// Sum of a cell's pin partials pin_grad[cpins[k0 .. k1)], in order, with up to
// four gathers in flight per step.
__device__ inline float2 sum_pin_partials(int k0, int k1, const int32_t *__restrict__ cpins,
                                          const float2 *__restrict__ pin_grad)
{
    float gx = 0.f, gy = 0.f;
    for (int k = k0; k < k1; k += 4) {
        int q[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) q[j] = k + j < k1 ? cpins[k + j] : -1;
        float2 v[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) v[j] = q[j] >= 0 ? pin_grad[q[j]] : make_float2(0.f, 0.f);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            gx += v[j].x;
            gy += v[j].y;
        }
    }
    return make_float2(gx, gy);
}

// Density force on a cell: r * sum over its smoothed rectangle of overlap * E
// (ePlace local smoothing, the adjoint of the field-map scatter).  Rows first,
// the y overlaps shared by every row, so the field gathers of a row issue
// together.
__device__ inline float2 cell_force(float2 p, float2 s, const float2 *__restrict__ exy,
                                    const GeomArgs &g, int Mx, int My)
{
    Rect r;
    stretched_rect(p, s, g, Mx, My, r);
    float fx = 0.f, fy = 0.f;
    if (r.i0 <= r.i1 && r.j0 <= r.j1) {
        const int dj = r.j1 - r.j0;
        if (dj < 4) {
            float oy[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) oy[q] = q <= dj ? ovl(r.ty0, r.ty1, r.j0 + q) : 0.0f;
            for (int i = r.i0; i <= r.i1; ++i) {
                const float ox = ovl(r.tx0, r.tx1, i);
                const size_t base = (size_t)i * My + r.j0;
                float2 e[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) e[q] = q <= dj ? exy[base + q] : make_float2(0.0f, 0.0f);
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (q <= dj) {
                        const float o = ox * oy[q];
                        fx += o * e[q].x;
                        fy += o * e[q].y;
                    }
            }
        } else {
            for (int i = r.i0; i <= r.i1; ++i) {
                const float ox = ovl(r.tx0, r.tx1, i);
                for (int j = r.j0; j <= r.j1; ++j) {
                    const size_t b = (size_t)i * My + j;
                    const float o = ox * ovl(r.ty0, r.ty1, j);
                    const float2 e = exy[b];
                    fx += o * e.x;
                    fy += o * e.y;
                }
            }
        }
    }
    const float fs = r.ratio * g.v[4] * g.v[5];
    return make_float2(fx * fs, fy * fs);
}

__device__ inline float2 clamp_cell(float2 p, float2 s, const GeomArgs &g)
{
    p.x = fminf(fmaxf(p.x, g.v[0]), g.v[2] - s.x);
    p.y = fminf(fmaxf(p.y, g.v[1]), g.v[3] - s.y);
    return p;
}


}  // namespace dev
}  // namespace ep
