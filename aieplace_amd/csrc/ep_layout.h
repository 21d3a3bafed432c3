// ep_layout.h -- host-side preparation of one netlist for the device: cell
// arrays [movable | fillers | fixed], nets regrouped by lane class, the
// cell -> pin CSR and the fixed-cell density map.  Shared by the
// single-instance context and the batched engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "ep_ctx.h"

namespace ep {

struct HostLayout {
    int32_t Mx = 0, My = 0;
    float xl = 0, yl = 0, xh = 0, yh = 0, bw = 0, bh = 0, sqrt2_bw = 0, sqrt2_bh = 0;
    int32_t n_mov = 0, n_fill = 0, n_fix = 0, n_place = 0, n_all = 0;
    double mov_area = 0;
    std::vector<float2> wh, pos;             // [n_all]
    int32_t n_nets = 0, n_pins = 0, n_pins_mov = 0;
    std::vector<int32_t> net_start;          // [n_nets + 1], regrouped order
    std::vector<int32_t> pin_cell;           // [n_pins], cell index in [mov | fill | fix]
    std::vector<float2> pin_off;             // [n_pins]
    std::vector<int32_t> net_perm;           // stored net k -> original net index
    BucketTable buckets{};                   // lane classes; blk0 counts from 0
    int32_t wa_blocks = 0;
    std::vector<int32_t> cps, cpins;         // movable cell -> pin CSR
    std::vector<float> fixmap;               // [Mx * My] fixed-cell overlap area
    double hpwl_ref = 0;                     // mu-schedule reference (resolved default)
    double max_sw = 0, max_sh = 0;           // largest smoothed placeable cell (max(w, sqrt2 bw), ...)
    std::vector<int32_t> cell_order;         // stored movable cell k -> caller's index (empty: identity)
};

// Caller's index of stored cell k (movable cells may be renumbered; fillers and
// fixed cells keep their places).
inline int32_t caller_cell(const std::vector<int32_t> &order, int32_t k)
{
    return k < (int32_t)order.size() ? order[k] : k;
}

// Mean area per bin of the field map: every placeable cell's smoothed rectangle
// carries its own area, fixed cells rho_t times their overlap.  Cells whose
// smoothed rectangle crosses the die edge lose the part outside, so the live
// mean is at most this; the field solve subtracts it before the first product.
inline float field_map_mean(const HostLayout &L, float rho_t)
{
    double a = 0.0, f = 0.0;
    for (int k = 0; k < L.n_place; ++k) a += (double)L.wh[k].x * L.wh[k].y;
    for (float v : L.fixmap) f += v;
    return (float)((a + rho_t * f) / ((double)L.Mx * L.My));
}

// Validates the netlist and parameters (same checks and messages as ep_create).
ep_status validate_netlist(const ep_netlist *nl, const ep_params *p);
// Builds the layout for bins p->bins_x x p->bins_y.
ep_status build_layout(const ep_netlist *nl, const ep_params *p, HostLayout &L);
// Gamma of the WA model at overflow tau (schedule of the reference's ePlace flow).
double gamma_at(const ep_params &p, double bw, double tau);

}  // namespace ep
