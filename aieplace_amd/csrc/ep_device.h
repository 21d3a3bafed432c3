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

constexpr int kTile = 4096;             // bins per LDS-privatised tile (per map)
constexpr int kDensCellsPerThread = 8;  // cells per density-scatter thread

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

__device__ inline double wave_sum(double v)
{
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

__device__ inline double wave_max(double v)
{
    for (int m = 32; m >= 1; m >>= 1) v = fmax(v, __shfl_xor(v, m, 64));
    return v;
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

template <int G>
__device__ inline float grp_max(float v)
{
#pragma unroll
    for (int m = G / 2; m >= 1; m >>= 1) v = fmaxf(v, __shfl_xor(v, m, G));
    return v;
}
template <int G>
__device__ inline float grp_min(float v)
{
#pragma unroll
    for (int m = G / 2; m >= 1; m >>= 1) v = fminf(v, __shfl_xor(v, m, G));
    return v;
}
template <int G>
__device__ inline float grp_sum(float v)
{
#pragma unroll
    for (int m = G / 2; m >= 1; m >>= 1) v += __shfl_xor(v, m, G);
    return v;
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
        const float mxp = cxp / bxp, mxm = cxm / bxm, myp = cyp / byp, mym = cym / bym;
        if (valid) {
            const float gx = axp / bxp * (1.f + (dxp - mxp) * ginv) -
                             axm / bxm * (1.f - (dxm - mxm) * ginv);
            const float gy = ayp / byp * (1.f + (dyp - myp) * ginv) -
                             aym / bym * (1.f - (dym - mym) * ginv);
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
        const float mxp = cxp / bxp, mxm = cxm / bxm, myp = cyp / byp, mym = cym / bym;
        for (int p = s + lane; p < e; p += G) {
            const float2 q = pos[pin_cell[p]], o = pin_off[p];
            const float dxp = q.x + o.x - xmx, dxm = q.x + o.x - xmn;
            const float dyp = q.y + o.y - ymx, dym = q.y + o.y - ymn;
            const float axp = __expf(dxp * ginv), axm = __expf(-dxm * ginv);
            const float ayp = __expf(dyp * ginv), aym = __expf(-dym * ginv);
            const float gx = axp / bxp * (1.f + (dxp - mxp) * ginv) -
                             axm / bxm * (1.f - (dxm - mxm) * ginv);
            const float gy = ayp / byp * (1.f + (dyp - myp) * ginv) -
                             aym / bym * (1.f - (dym - mym) * ginv);
            pin_grad[p] = make_float2(gx, gy);
        }
        if (lane == 0) {
            hp_acc += (double)(xmx - xmn) + (double)(ymx - ymn);
            wa_acc += (double)((xmx - xmn) + mxp - mxm) + (double)((ymx - ymn) + myp - mym);
        }
    }
}

__device__ inline float2 clamp_cell(float2 p, float2 s, const GeomArgs &g)
{
    p.x = fminf(fmaxf(p.x, g.v[0]), g.v[2] - s.x);
    p.y = fminf(fmaxf(p.y, g.v[1]), g.v[3] - s.y);
    return p;
}


}  // namespace dev
}  // namespace ep
