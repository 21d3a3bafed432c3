// ep_api.cpp -- the C ABI (include/eplace.h): context set-up, per-iteration
// calls and the Nesterov optimiser driving the device kernels.
//
// Reference interfaces replaced: Placer (vck5000/host/src/include/AIEplace.h:21-99,
// AIEplace.cpp:9-138, 993-1070) and GraphDriver (GraphDriver.h:16-65).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/eplace_mlp.h"
#include "ep_sched.h"
#include "ep_ctx.h"
#include "ep_layout.h"

namespace ep {

char *last_error_buffer()
{
    static thread_local char buf[512] = {0};
    return buf;
}

namespace {

#define EP_HIP(call)                                                                         \
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
    if (e != hipSuccess) return fail(EP_ENOMEM, "hipMalloc(%zu bytes) failed: %s", n * sizeof(T), hipGetErrorString(e));
    return EP_OK;
}

template <class T>
ep_status upload(T *dst, const T *src, size_t n)
{
    if (n == 0) return EP_OK;
    EP_HIP(hipMemcpy(dst, src, n * sizeof(T), hipMemcpyHostToDevice));
    return EP_OK;
}

uint64_t rng_next(uint64_t &s)
{
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double rng_uniform(uint64_t &s) { return (double)(rng_next(s) >> 11) * (1.0 / 9007199254740992.0); }

int lanes_for_degree(int d)
{
    int g = 2;
    while (g < d && g < 64) g <<= 1;
    return g;
}

double gamma_of(const Ctx &c, double tau) { return gamma_at(c.prm, c.bw, tau); }

ep_status check_launch()
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(EP_EDEVICE, "kernel launch failed: %s", hipGetErrorString(e));
    return EP_OK;
}

// Stage the host scalars; the copy is part of the stream (and of a captured graph).
void push_scalars(Ctx &c, hipStream_t s)
{
    (void)hipMemcpyAsync(c.d_scal, c.h_scal, sizeof(DevScalars), hipMemcpyHostToDevice, s);
}

// One gradient evaluation at `pos`: density + overflow, WA, field, gather.
// `counted`: a counting launch_trial produced pos from vref (its sort counts
// and its |pos - vref|^2 partial sums are made).
void enqueue_eval(Ctx &c, const float2 *pos, const float2 *vref, const float2 *gref, float2 *gout,
                  float2 *gw_out, float2 *f_out, hipStream_t s, bool counted = false)
{
    launch_sort(c, pos, counted, s);
    launch_density(c, s);
    launch_overflow(c, s);
    launch_wa(c, pos, s);
    launch_field(c, c.prm.precision, s);
    launch_cell_grad(c, pos, counted ? nullptr : vref, gref, gout, gw_out, f_out, s);
    launch_finalize(c, s, counted);
    (void)hipMemcpyAsync(c.h_res, c.d_res, sizeof(DevResult), hipMemcpyDeviceToHost, s);
}

// Leaves stream capture after an error between hipStreamBeginCapture and
// hipStreamEndCapture (the stream would otherwise stay in capture mode and fail
// every later call); the partial graph is discarded.
void abort_capture(hipStream_t s)
{
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(s, &g);
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
}

#define EP_CAP(call, stream)                                                                 \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess) {                                                              \
            abort_capture(stream);                                                           \
            return fail(EP_EDEVICE, "%s failed during capture: %s (%s:%d)", #call,           \
                        hipGetErrorString(e_), __FILE__, __LINE__);                          \
        }                                                                                    \
    } while (0)

ep_status sync(Ctx &c)
{
    EP_HIP(hipStreamSynchronize(c.stream));
    return EP_OK;
}

// Destroys the captured graphs (they bake the precision and the proposal kernel).
void drop_graphs(Ctx &c)
{
    for (int i = 0; i < 2; ++i) {
        if (c.graph_exec[i]) (void)hipGraphExecDestroy(c.graph_exec[i]);
        c.graph_exec[i] = nullptr;
        c.graph_built[i] = false;
    }
    if (c.ctl_exec) (void)hipGraphExecDestroy(c.ctl_exec);
    c.ctl_exec = nullptr;
}

// Trial step from V[cur]/U[cur]/G[cur] into the other buffers, then evaluate.
ep_status trial_eval(Ctx &c)
{
    const int a = c.cur, b = c.cur ^ 1;
    if (c.prm.use_graph) {
        if (!c.graph_built[a]) {
            hipGraph_t g = nullptr;
            EP_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
            push_scalars(c, c.stream);
            launch_mlp_propose(c, c.stream);
            launch_trial(c, c.V[a], c.U[a], c.G[a], c.V[b], c.U[b], true, c.stream);
            enqueue_eval(c, c.V[b], c.V[a], c.G[a], c.G[b], nullptr, nullptr, c.stream, true);
            EP_CAP(hipGetLastError(), c.stream);
            EP_HIP(hipStreamEndCapture(c.stream, &g));
            hipError_t e = hipGraphInstantiate(&c.graph_exec[a], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) return fail(EP_EDEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e));
            c.graph_built[a] = true;
        }
        EP_HIP(hipGraphLaunch(c.graph_exec[a], c.stream));
    } else {
        push_scalars(c, c.stream);
        launch_mlp_propose(c, c.stream);
        launch_trial(c, c.V[a], c.U[a], c.G[a], c.V[b], c.U[b], true, c.stream);
        enqueue_eval(c, c.V[b], c.V[a], c.G[a], c.G[b], nullptr, nullptr, c.stream, true);
        ep_status st = check_launch();
        if (st != EP_OK) return st;
    }
    return sync(c);
}

}  // namespace
}  // namespace ep

using namespace ep;

struct ep_ctx : public ep::Ctx {};

extern "C" const char *ep_last_error(void) { return last_error_buffer(); }

extern "C" void ep_default_params(ep_params *p)
{
    std::memset(p, 0, sizeof(*p));
    p->bins_x = p->bins_y = 512;
    p->target_density = 0.9f;
    p->gamma_bins = 4.0f;
    p->gamma_schedule = 1;
    p->lambda_eta = 8e-5f;
    p->hpwl_ref = 0.0;  // 0: derived from the netlist at ep_create
    p->mu_max = 1.05f;
    p->mu_min = 0.95f;
    p->max_backtrack = 10;
    p->stop_overflow = 0.10f;
    p->max_iters = 3000;
    p->init_disp_bins = 0.1f;
    p->use_fillers = 1;
    p->filler_seed = 7;
    p->precision = EP_PREC_FP32;
    p->use_graph = 0;
    p->learned_step = 0;
    p->fork_min_cells = 0;
    p->cell_order = 1;
}

static ep_status check_precision(int32_t precision, const char *who)
{
    if (precision == EP_PREC_FP32 || precision == EP_PREC_BF16X3 || precision == EP_PREC_FP16X3)
        return EP_OK;
    return fail(EP_EINVAL, "%s: unknown precision %d", who, precision);
}

ep_status ep::validate_netlist(const ep_netlist *nl, const ep_params *p)
{
    if (!nl || !p) return fail(EP_EINVAL, "ep_create: null argument");
    if (nl->n_mov < 1 || nl->n_fix < 0) return fail(EP_EINVAL, "ep_create: need n_mov >= 1, n_fix >= 0");
    if (!nl->w || !nl->h || !nl->x || !nl->y) return fail(EP_EINVAL, "ep_create: null cell arrays");
    if (nl->n_nets < 0 || nl->n_pins < 0) return fail(EP_EINVAL, "ep_create: negative net/pin count");
    if (nl->n_nets > 0 && (!nl->net_start || !nl->pin_cell || !nl->pin_dx || !nl->pin_dy))
        return fail(EP_EINVAL, "ep_create: null net arrays");
    if (!(nl->xh > nl->xl) || !(nl->yh > nl->yl)) return fail(EP_EINVAL, "ep_create: empty die");
    if (p->bins_x < 16 || p->bins_y < 16 || p->bins_x % 16 || p->bins_y % 16)
        return fail(EP_EINVAL, "ep_create: bins must be multiples of 16 and >= 16 (got %d x %d)",
                    p->bins_x, p->bins_y);
    if ((int64_t)p->bins_x * p->bins_y > (int64_t)16384 * 16384)
        return fail(EP_EINVAL, "ep_create: bin grid too large");
    if (!(p->target_density > 0.f) || p->target_density > 1.f)
        return fail(EP_EINVAL, "ep_create: target_density must be in (0, 1]");
    if (!(p->gamma_bins > 0.f)) return fail(EP_EINVAL, "ep_create: gamma_bins must be > 0");
    if (p->max_backtrack < 1) return fail(EP_EINVAL, "ep_create: max_backtrack must be >= 1");
    if (p->fork_min_cells < 0) return fail(EP_EINVAL, "ep_create: fork_min_cells must be >= 0");
    if (p->cell_order != 0 && p->cell_order != 1) return fail(EP_EINVAL, "ep_create: cell_order must be 0 or 1");
    ep_status pst = check_precision(p->precision, "ep_create");
    if (pst != EP_OK) return pst;
    const int n_all = nl->n_mov + nl->n_fix;
    for (int k = 0; k < n_all; ++k) {
        if (!(nl->w[k] > 0.f) || !(nl->h[k] > 0.f) || !std::isfinite(nl->w[k]) || !std::isfinite(nl->h[k]))
            return fail(EP_EINVAL, "ep_create: cell %d has a non-positive size", k);
        if (!std::isfinite(nl->x[k]) || !std::isfinite(nl->y[k]))
            return fail(EP_ENAN, "ep_create: cell %d has a non-finite position", k);
    }
    if (nl->n_nets > 0) {
        if (nl->net_start[0] != 0 || nl->net_start[nl->n_nets] != nl->n_pins)
            return fail(EP_EINVAL, "ep_create: net_start must run from 0 to n_pins");
        for (int n = 0; n < nl->n_nets; ++n)
            if (nl->net_start[n + 1] < nl->net_start[n])
                return fail(EP_EINVAL, "ep_create: net_start not monotone at net %d", n);
        for (int p2 = 0; p2 < nl->n_pins; ++p2) {
            if (nl->pin_cell[p2] < 0 || nl->pin_cell[p2] >= n_all)
                return fail(EP_EINVAL, "ep_create: pin %d references cell %d out of range", p2, nl->pin_cell[p2]);
            if (!std::isfinite(nl->pin_dx[p2]) || !std::isfinite(nl->pin_dy[p2]))
                return fail(EP_ENAN, "ep_create: pin %d has a non-finite offset", p2);
        }
    }
    return EP_OK;
}

static void destroy_ctx(ep_ctx *c)
{
    if (!c) return;
    for (int i = 0; i < 2; ++i)
        if (c->graph_exec[i]) (void)hipGraphExecDestroy(c->graph_exec[i]);
    if (c->ctl_exec) (void)hipGraphExecDestroy(c->ctl_exec);
    void *ptrs[] = {c->wh, c->net_start, c->pin_cell, c->pin_off, c->pin_grad, c->cell_pin_start,
                    c->cell_pins, c->V[0], c->V[1], c->U[0], c->U[1], c->G[0], c->G[1],
                    c->map_field, c->map_mov, c->map_fix, c->acc_field, c->acc_mov, c->exy, c->tA, c->tB, c->part,
                    c->sorted, c->sort_count, c->sort_off, c->dens_tab,
                    c->d_scal, c->d_res, c->d_mlp_w, c->d_ctl, c->d_log, c->d_rec, c->fin,
                    c->fin_ticket, c->gw_cell};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    field_free(*c);
    if (c->h_scal) (void)hipHostFree(c->h_scal);
    if (c->h_res) (void)hipHostFree(c->h_res);
    if (c->h_ctl) (void)hipHostFree(c->h_ctl);
    for (hipEvent_t e : c->ctl_read)
        if (e) (void)hipEventDestroy(e);
    if (c->fork) (void)hipEventDestroy(c->fork);
    if (c->join) (void)hipEventDestroy(c->join);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" ep_status ep_plan_fillers(const ep_netlist *nl, const ep_params *p, int32_t *n_fill,
                                     float *fw_out, float *fh_out, float *x, float *y, int32_t cap)
{
    ep_status st = validate_netlist(nl, p);
    if (st != EP_OK) return st;
    double mov_area = 0, fix_area = 0, wsum = 0, hsum = 0;
    for (int k = 0; k < nl->n_mov; ++k) {
        mov_area += (double)nl->w[k] * nl->h[k];
        wsum += nl->w[k];
        hsum += nl->h[k];
    }
    for (int k = nl->n_mov; k < nl->n_mov + nl->n_fix; ++k) {
        const double x0 = std::max((double)nl->x[k], (double)nl->xl);
        const double x1 = std::min((double)nl->x[k] + nl->w[k], (double)nl->xh);
        const double y0 = std::max((double)nl->y[k], (double)nl->yl);
        const double y1 = std::min((double)nl->y[k] + nl->h[k], (double)nl->yh);
        if (x1 > x0 && y1 > y0) fix_area += (x1 - x0) * (y1 - y0);
    }
    const double die_area = ((double)nl->xh - nl->xl) * ((double)nl->yh - nl->yl);
    const float fw = (float)(wsum / nl->n_mov), fh = (float)(hsum / nl->n_mov);
    int64_t nf = 0;
    if (p->use_fillers) {
        const double fa = (double)p->target_density * (die_area - fix_area) - mov_area;
        if (fa > 0) nf = (int64_t)std::floor(fa / ((double)fw * fh));
    }
    if (nf > ((int64_t)1 << 28)) return fail(EP_EINVAL, "ep_plan_fillers: filler count too large");
    if (n_fill) *n_fill = (int32_t)nf;
    if (fw_out) *fw_out = fw;
    if (fh_out) *fh_out = fh;
    if (x && y) {
        // uniform random corners, then stored in Morton order of their bins so
        // that consecutive fillers are spatial neighbours (density scatter tiles)
        uint64_t rs = p->filler_seed;
        std::vector<std::pair<uint64_t, int64_t>> key((size_t)nf);
        std::vector<float> fx((size_t)nf), fy((size_t)nf);
        const double bw = ((double)nl->xh - nl->xl) / p->bins_x, bh = ((double)nl->yh - nl->yl) / p->bins_y;
        for (int64_t k = 0; k < nf; ++k) {
            fx[k] = (float)(nl->xl + rng_uniform(rs) * ((double)nl->xh - nl->xl - fw));
            fy[k] = (float)(nl->yl + rng_uniform(rs) * ((double)nl->yh - nl->yl - fh));
            const uint32_t i = (uint32_t)std::min(65535.0, std::max(0.0, (fx[k] + 0.5 * fw - nl->xl) / bw));
            const uint32_t j = (uint32_t)std::min(65535.0, std::max(0.0, (fy[k] + 0.5 * fh - nl->yl) / bh));
            uint64_t m = 0;
            for (int b = 0; b < 16; ++b) m |= (uint64_t)((i >> b) & 1u) << (2 * b) | (uint64_t)((j >> b) & 1u) << (2 * b + 1);
            key[(size_t)k] = {m, k};
        }
        std::stable_sort(key.begin(), key.end());
        const int64_t n = std::min<int64_t>(nf, cap);
        for (int64_t k = 0; k < n; ++k) {
            x[k] = fx[(size_t)key[(size_t)k].second];
            y[k] = fy[(size_t)key[(size_t)k].second];
        }
    }
    return EP_OK;
}

// Exact overlap of fixed cells with the bins (Bin.h:78-121), host FP64, once.
static void fixed_map(const HostLayout &L, std::vector<float> &out)
{
    std::vector<double> m((size_t)L.Mx * L.My, 0.0);
    const double bw = L.bw, bh = L.bh;
    for (int k = L.n_place; k < L.n_all; ++k) {
        const double x0 = L.pos[k].x, x1 = x0 + L.wh[k].x, y0 = L.pos[k].y, y1 = y0 + L.wh[k].y;
        int i0 = std::max(0, (int)std::floor((x0 - L.xl) / bw));
        int i1 = std::min(L.Mx - 1, (int)std::floor((x1 - L.xl) / bw));
        int j0 = std::max(0, (int)std::floor((y0 - L.yl) / bh));
        int j1 = std::min(L.My - 1, (int)std::floor((y1 - L.yl) / bh));
        for (int i = i0; i <= i1; ++i) {
            const double ox = std::min(L.xl + (i + 1) * bw, x1) - std::max(L.xl + i * bw, x0);
            if (ox <= 0) continue;
            for (int j = j0; j <= j1; ++j) {
                const double oy = std::min(L.yl + (j + 1) * bh, y1) - std::max(L.yl + j * bh, y0);
                if (oy <= 0) continue;
                m[(size_t)i * L.My + j] += ox * oy;
            }
        }
    }
    out.resize(m.size());
    for (size_t b = 0; b < m.size(); ++b) out[b] = (float)m[b];
}

double ep::gamma_at(const ep_params &p, double bw, double tau)
{
    double g = (double)p.gamma_bins * bw;
    if (p.gamma_schedule) g *= sched_exp10((tau - 0.1) * 20.0 / 9.0 - 1.0);
    return g;
}

// Movable cells in breadth-first order over the cell-net incidence (nets of
// degree >= 2), each connected component from a pseudo-peripheral cell (the
// last cell a first sweep reaches), components in the order of their first
// cell.  Cells of one net then get nearby indices whatever the caller's order,
// which keeps the WA kernel's pin -> cell gathers and the per-cell pin sums
// cache-local (the same role as a bandwidth-reducing matrix ordering).
static void connectivity_order(const ep_netlist *nl, std::vector<int32_t> &order)
{
    const int n = nl->n_mov, N = nl->n_nets;
    std::vector<int32_t> cs(n + 1, 0);
    for (int e = 0; e < N; ++e) {
        if (nl->net_start[e + 1] - nl->net_start[e] < 2) continue;
        for (int q = nl->net_start[e]; q < nl->net_start[e + 1]; ++q)
            if (nl->pin_cell[q] < n) cs[nl->pin_cell[q] + 1]++;
    }
    for (int k = 0; k < n; ++k) cs[k + 1] += cs[k];
    std::vector<int32_t> cn(std::max(1, cs[n])), fill(cs.begin(), cs.end() - 1);
    for (int e = 0; e < N; ++e) {
        if (nl->net_start[e + 1] - nl->net_start[e] < 2) continue;
        for (int q = nl->net_start[e]; q < nl->net_start[e + 1]; ++q)
            if (nl->pin_cell[q] < n) cn[fill[nl->pin_cell[q]]++] = e;
    }
    std::vector<char> seen_c(n, 0), seen_n(N, 0);
    std::vector<int32_t> nets_hit;
    order.clear();
    order.reserve(n);
    auto sweep = [&](int s) {
        size_t head = order.size();
        order.push_back(s);
        seen_c[s] = 1;
        while (head < order.size()) {
            const int c = order[head++];
            for (int k = cs[c]; k < cs[c + 1]; ++k) {
                const int e = cn[k];
                if (seen_n[e]) continue;
                seen_n[e] = 1;
                nets_hit.push_back(e);
                for (int q = nl->net_start[e]; q < nl->net_start[e + 1]; ++q) {
                    const int c2 = nl->pin_cell[q];
                    if (c2 < n && !seen_c[c2]) {
                        seen_c[c2] = 1;
                        order.push_back(c2);
                    }
                }
            }
        }
    };
    for (int s = 0; s < n; ++s) {
        if (seen_c[s]) continue;
        const size_t mark = order.size();
        nets_hit.clear();
        sweep(s);
        const int far = order.back();
        for (size_t k = mark; k < order.size(); ++k) seen_c[order[k]] = 0;
        for (int e : nets_hit) seen_n[e] = 0;
        order.resize(mark);
        sweep(far);
    }
}

ep_status ep::build_layout(const ep_netlist *nl, const ep_params *p, HostLayout &L)
{
    ep_status st = validate_netlist(nl, p);
    if (st != EP_OK) return st;
    L = HostLayout();
    L.Mx = p->bins_x;
    L.My = p->bins_y;
    L.xl = nl->xl; L.yl = nl->yl; L.xh = nl->xh; L.yh = nl->yh;
    L.bw = (float)(((double)nl->xh - nl->xl) / L.Mx);
    L.bh = (float)(((double)nl->yh - nl->yl) / L.My);
    L.sqrt2_bw = (float)(M_SQRT2 * (double)L.bw);
    L.sqrt2_bh = (float)(M_SQRT2 * (double)L.bh);
    L.n_mov = nl->n_mov;
    L.n_fix = nl->n_fix;

    // ---- fillers (ePlace whitespace fillers)
    int32_t n_fill = 0;
    float fw = 0, fh = 0;
    if ((st = ep_plan_fillers(nl, p, &n_fill, &fw, &fh, nullptr, nullptr, 0)) != EP_OK) return st;
    for (int k = 0; k < nl->n_mov; ++k) L.mov_area += (double)nl->w[k] * nl->h[k];
    L.n_fill = n_fill;
    L.n_place = L.n_mov + L.n_fill;
    L.n_all = L.n_place + L.n_fix;

    // ---- movable cell order: stored cell k is the caller's cell_order[k]
    std::vector<int32_t> inv;  // caller's movable index -> stored index
    if (p->cell_order == 1 && L.n_mov > 1) {
        connectivity_order(nl, L.cell_order);
        inv.assign(L.n_mov, 0);
        for (int k = 0; k < L.n_mov; ++k) inv[L.cell_order[k]] = k;
    }
    auto stored = [&](int32_t cell) { return cell < L.n_mov && !inv.empty() ? inv[cell] : cell; };

    // ---- host cell arrays [mov | fill | fix]
    L.wh.resize(L.n_all);
    L.pos.resize(L.n_all);
    for (int k = 0; k < L.n_mov; ++k) {
        const int src = caller_cell(L.cell_order, k);
        L.wh[k] = make_float2(nl->w[src], nl->h[src]);
        L.pos[k] = make_float2(nl->x[src], nl->y[src]);
    }
    {
        std::vector<float> fxs(std::max(1, L.n_fill)), fys(std::max(1, L.n_fill));
        int32_t nf = 0;
        ep_plan_fillers(nl, p, &nf, &fw, &fh, fxs.data(), fys.data(), L.n_fill);
        for (int k = 0; k < L.n_fill; ++k) {
            L.wh[L.n_mov + k] = make_float2(fw, fh);
            L.pos[L.n_mov + k] = make_float2(fxs[k], fys[k]);
        }
    }
    for (int k = 0; k < L.n_fix; ++k) {
        L.wh[L.n_place + k] = make_float2(nl->w[nl->n_mov + k], nl->h[nl->n_mov + k]);
        L.pos[L.n_place + k] = make_float2(nl->x[nl->n_mov + k], nl->y[nl->n_mov + k]);
    }
    // movable initial positions clamped into the die; largest smoothed cell
    for (int k = 0; k < L.n_place; ++k) {
        L.max_sw = std::max(L.max_sw, (double)std::max(L.wh[k].x, L.sqrt2_bw));
        L.max_sh = std::max(L.max_sh, (double)std::max(L.wh[k].y, L.sqrt2_bh));
        L.pos[k].x = std::min(std::max(L.pos[k].x, L.xl), L.xh - L.wh[k].x);
        L.pos[k].y = std::min(std::max(L.pos[k].y, L.yl), L.yh - L.wh[k].y);
    }

    // ---- nets regrouped by lane class (degree < 2 nets go last, untouched)
    const int N = nl->n_nets;
    L.n_nets = N;
    L.n_pins = nl->n_pins;
    // classes: exact degrees 2, 3, 4 (their pin ranges follow from the net's
    // rank in the class), then lane widths 8, 16, 32, 64
    const int lane_classes[kMaxBuckets] = {2, 4, 4, 8, 16, 32, 64};
    const int class_deg[kMaxBuckets] = {2, 3, 4, 0, 0, 0, 0};
    std::vector<std::vector<int32_t>> cls(kMaxBuckets + 1);
    for (int n = 0; n < N; ++n) {
        int d = nl->net_start[n + 1] - nl->net_start[n];
        if (d < 2) {
            cls[kMaxBuckets].push_back(n);
            continue;
        }
        int b = d <= 4 ? d - 2 : 3;
        if (d > 4) {
            const int g = lanes_for_degree(d);
            while (lane_classes[b] != g) ++b;
        }
        cls[b].push_back(n);
    }
    // within a class, order nets by their lowest cell index: nets on nearby
    // cells then sit near each other, so the pin -> cell gathers of the WA
    // kernel and the cell -> pin gathers of the per-cell sum stay cache-local
    {
        std::vector<int32_t> key(N, 0);
        for (int n = 0; n < N; ++n) {
            int32_t m = INT32_MAX;
            for (int q = nl->net_start[n]; q < nl->net_start[n + 1]; ++q) m = std::min(m, stored(nl->pin_cell[q]));
            key[n] = m;
        }
        for (int b = 0; b < kMaxBuckets; ++b)
            std::stable_sort(cls[b].begin(), cls[b].end(),
                             [&](int32_t a2, int32_t b2) { return key[a2] < key[b2]; });
    }
    L.net_start.assign(N + 1, 0);
    L.pin_cell.assign(nl->n_pins, 0);
    L.pin_off.assign(nl->n_pins, make_float2(0.f, 0.f));
    L.net_perm.reserve(N);
    int pin_w = 0, net_w = 0, blk = 0;
    L.buckets.n = 0;
    for (int b = 0; b <= kMaxBuckets; ++b) {
        if (b < kMaxBuckets && !cls[b].empty()) {
            Bucket &bk = L.buckets.b[L.buckets.n++];
            bk.lanes = lane_classes[b];
            bk.deg = class_deg[b];
            bk.pin0 = pin_w;
            bk.net0 = net_w;
            bk.net1 = net_w + (int32_t)cls[b].size();
            bk.blk0 = blk;
            blk += (int)(((int64_t)cls[b].size() * bk.lanes + kBlock - 1) / kBlock);
        }
        for (int32_t n : cls[b]) {
            L.net_start[net_w] = pin_w;
            L.net_perm.push_back(n);
            for (int q = nl->net_start[n]; q < nl->net_start[n + 1]; ++q) {
                int32_t cell = stored(nl->pin_cell[q]);
                if (cell >= L.n_mov) cell += L.n_fill;  // fixed cells follow the fillers
                L.pin_cell[pin_w] = cell;
                L.pin_off[pin_w] = make_float2(nl->pin_dx[q], nl->pin_dy[q]);
                ++pin_w;
            }
            ++net_w;
        }
    }
    L.net_start[N] = pin_w;
    L.wa_blocks = blk;

    // ---- cell -> pin CSR (movable cells)
    L.cps.assign(L.n_mov + 1, 0);
    for (int q = 0; q < nl->n_pins; ++q)
        if (L.pin_cell[q] < L.n_mov) L.cps[L.pin_cell[q] + 1]++;
    for (int k = 0; k < L.n_mov; ++k) L.cps[k + 1] += L.cps[k];
    L.n_pins_mov = L.cps[L.n_mov];
    L.cpins.assign(std::max(1, L.n_pins_mov), 0);
    std::vector<int32_t> fillp(L.cps.begin(), L.cps.end() - 1);
    for (int q = 0; q < nl->n_pins; ++q)
        if (L.pin_cell[q] < L.n_mov) L.cpins[fillp[L.pin_cell[q]]++] = q;

    // ---- hpwl reference of the mu schedule: default scales with the netlist
    L.hpwl_ref = p->hpwl_ref > 0 ? p->hpwl_ref : kHpwlRefPerNetBin * (double)std::max(1, N) * L.bw;
    fixed_map(L, L.fixmap);
    return EP_OK;
}

extern "C" ep_status ep_create(const ep_netlist *nl, const ep_params *p, ep_ctx **out)
{
    if (!out) return fail(EP_EINVAL, "ep_create: null output");
    *out = nullptr;
    HostLayout L;
    ep_status st = build_layout(nl, p, L);
    if (st != EP_OK) return st;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(EP_EDEVICE, "ep_create: no HIP device available");

    ep_ctx *c = new (std::nothrow) ep_ctx();
    if (!c) return fail(EP_ENOMEM, "ep_create: host allocation failed");
    c->prm = *p;
    c->prm.hpwl_ref = L.hpwl_ref;
    (void)hipGetDevice(&c->device);
    c->Mx = L.Mx;
    c->My = L.My;
    c->xl = L.xl; c->yl = L.yl; c->xh = L.xh; c->yh = L.yh;
    c->bw = L.bw;
    c->bh = L.bh;
    c->sqrt2_bw = L.sqrt2_bw;
    c->sqrt2_bh = L.sqrt2_bh;
    c->n_mov = L.n_mov;
    c->n_fix = L.n_fix;
    c->n_fill = L.n_fill;
    c->n_place = L.n_place;
    c->n_all = L.n_all;
    c->mov_area = L.mov_area;
    const int N = L.n_nets;
    c->n_nets = N;
    c->n_pins = L.n_pins;
    c->n_pins_mov = L.n_pins_mov;
    c->net_perm = L.net_perm;
    c->cell_order = L.cell_order;
    c->buckets = L.buckets;
    c->wa_blocks = L.wa_blocks;
    c->wa_part_blocks = std::min(L.wa_blocks, kWaGridCap);
    c->dens = make_dens_plan(L.Mx, L.My, L.bw, L.bh, L.max_sw, L.max_sh, 4);

    // ---- device buffers
    hipError_t he = hipStreamCreate(&c->stream);
    if (he == hipSuccess) he = hipStreamCreate(&c->side);
    if (he == hipSuccess) he = hipEventCreateWithFlags(&c->fork, hipEventDisableTiming);
    if (he == hipSuccess) he = hipEventCreateWithFlags(&c->join, hipEventDisableTiming);
    for (int i = 0; i < 2 && he == hipSuccess; ++i) he = hipEventCreateWithFlags(&c->ctl_read[i], hipEventDisableTiming);
    if (he != hipSuccess) {
        destroy_ctx(c);
        return fail(EP_EDEVICE, "stream/event creation: %s", hipGetErrorString(he));
    }
    const size_t nb = (size_t)c->Mx * c->My;
    c->ovf_blocks = (int)std::min<size_t>(1024, (nb + kBlock * 4 - 1) / (kBlock * 4));
    c->cell_blocks = (int)std::min<int64_t>(kCellGridCap, ((int64_t)c->n_place + 2 * kBlock - 1) / (2 * kBlock));
    c->part_wa_off = 0;
    c->part_ovf_off = 2 * c->wa_part_blocks;
    c->part_cell_off = c->part_ovf_off + 4 * c->ovf_blocks;
    c->trial_blocks = (int)(((int64_t)c->n_place + kBlock * kSortCellsPerThread - 1) / (kBlock * kSortCellsPerThread));
    c->part_trial_off = c->part_cell_off + 7 * c->cell_blocks;
    c->part_total = c->part_trial_off + c->trial_blocks;
#define EP_TRY(x)                 \
    do {                          \
        ep_status s_ = (x);       \
        if (s_ != EP_OK) {        \
            destroy_ctx(c);       \
            return s_;            \
        }                         \
    } while (0)
    EP_TRY(dalloc(&c->wh, c->n_all));
    EP_TRY(dalloc(&c->net_start, N + 1));
    EP_TRY(dalloc(&c->pin_cell, L.n_pins));
    EP_TRY(dalloc(&c->pin_off, L.n_pins));
    EP_TRY(dalloc(&c->pin_grad, L.n_pins));
    EP_TRY(dalloc(&c->cell_pin_start, c->n_mov + 1));
    EP_TRY(dalloc(&c->gw_cell, std::max(1, c->n_mov)));
    EP_TRY(dalloc(&c->cell_pins, L.cpins.size()));
    for (int i = 0; i < 2; ++i) {
        EP_TRY(dalloc(&c->V[i], c->n_all));
        EP_TRY(dalloc(&c->U[i], c->n_place));
        EP_TRY(dalloc(&c->G[i], c->n_place));
    }
    EP_TRY(dalloc(&c->map_field, nb));
    EP_TRY(dalloc(&c->map_mov, nb));
    EP_TRY(dalloc(&c->acc_field, nb));
    EP_TRY(dalloc(&c->acc_mov, nb));
    EP_TRY(dalloc(&c->map_fix, nb));
    EP_TRY(dalloc(&c->sorted, std::max(1, c->n_place)));
    EP_TRY(dalloc(&c->sort_count, kSortTiles));
    EP_TRY(dalloc(&c->sort_off, kSortTiles));
    EP_TRY(dalloc(&c->dens_tab, kSortTiles + 1));
    EP_TRY(dalloc(&c->exy, nb));
    EP_TRY(dalloc(&c->tA, 2 * nb));
    EP_TRY(dalloc(&c->tB, 2 * nb));
    EP_TRY(dalloc(&c->part, c->part_total));
    EP_TRY(dalloc(&c->fin, kFinalizeBlocks * 10));
    EP_TRY(dalloc(&c->fin_ticket, 1));
    EP_TRY(dalloc(&c->d_scal, 1));
    EP_TRY(dalloc(&c->d_res, 1));
    if (hipHostMalloc((void **)&c->h_scal, sizeof(DevScalars), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&c->h_res, sizeof(DevResult), hipHostMallocDefault) != hipSuccess) {
        destroy_ctx(c);
        return fail(EP_ENOMEM, "ep_create: pinned host allocation failed");
    }
    std::memset(c->h_scal, 0, sizeof(DevScalars));
    std::memset(c->h_res, 0, sizeof(DevResult));
    c->h_scal->step_unit = c->bw;
    EP_TRY(upload(c->d_scal, c->h_scal, 1));  // sel = halt = 0 before the first push
    EP_TRY(upload(c->wh, L.wh.data(), L.wh.size()));
    EP_TRY(upload(c->net_start, L.net_start.data(), L.net_start.size()));
    EP_TRY(upload(c->pin_cell, L.pin_cell.data(), L.pin_cell.size()));
    EP_TRY(upload(c->pin_off, L.pin_off.data(), L.pin_off.size()));
    EP_TRY(upload(c->cell_pin_start, L.cps.data(), L.cps.size()));
    EP_TRY(upload(c->cell_pins, L.cpins.data(), L.cpins.size()));
    EP_TRY(upload(c->map_fix, L.fixmap.data(), L.fixmap.size()));
    for (int i = 0; i < 2; ++i) {
        EP_TRY(upload(c->V[i], L.pos.data(), L.pos.size()));
        EP_TRY(upload(c->U[i], L.pos.data(), (size_t)c->n_place));
        if (hipMemset(c->G[i], 0, sizeof(float2) * std::max(1, c->n_place)) != hipSuccess) {
            destroy_ctx(c);
            return fail(EP_EDEVICE, "ep_create: hipMemset failed");
        }
    }
    if (hipMemset(c->pin_grad, 0, sizeof(float2) * std::max(1, L.n_pins)) != hipSuccess ||
        hipMemset(c->exy, 0, sizeof(float2) * nb) != hipSuccess ||
        hipMemset(c->part, 0, sizeof(double) * c->part_total) != hipSuccess ||
        hipMemset(c->fin_ticket, 0, sizeof(unsigned)) != hipSuccess ||
        hipMemset(c->acc_field, 0, sizeof(unsigned long long) * nb) != hipSuccess ||
        hipMemset(c->acc_mov, 0, sizeof(unsigned long long) * nb) != hipSuccess ||
        hipMemset(c->sort_count, 0, sizeof(int32_t) * kSortTiles) != hipSuccess) {
        destroy_ctx(c);
        return fail(EP_EDEVICE, "ep_create: hipMemset failed");
    }
    EP_TRY(field_init(*c));
    {
        const float mean = field_map_mean(L, p->target_density);
        EP_TRY(upload(c->map_mean, &mean, 1));
    }
#undef EP_TRY
    if (hipDeviceSynchronize() != hipSuccess) {
        destroy_ctx(c);
        return fail(EP_EDEVICE, "ep_create: device synchronisation failed");
    }
    *out = c;
    return EP_OK;
}

extern "C" ep_status ep_destroy(ep_ctx *ctx)
{
    if (!ctx) return fail(EP_EINVAL, "ep_destroy: null context");
    (void)hipStreamSynchronize(ctx->stream);
    destroy_ctx(ctx);
    return EP_OK;
}

extern "C" ep_status ep_get_sizes(ep_ctx *c, int32_t *n_mov, int32_t *n_fill, int32_t *n_fix,
                                  int32_t *n_nets, int32_t *n_pins)
{
    if (!c) return fail(EP_EINVAL, "ep_get_sizes: null context");
    if (n_mov) *n_mov = c->n_mov;
    if (n_fill) *n_fill = c->n_fill;
    if (n_fix) *n_fix = c->n_fix;
    if (n_nets) *n_nets = c->n_nets;
    if (n_pins) *n_pins = c->n_pins;
    return EP_OK;
}

extern "C" ep_status ep_get_cells(ep_ctx *c, float *w, float *h, float *fix_x, float *fix_y)
{
    if (!c) return fail(EP_EINVAL, "ep_get_cells: null context");
    std::vector<float2> wh(c->n_all), pos(std::max(1, c->n_fix));
    EP_HIP(hipMemcpy(wh.data(), c->wh, sizeof(float2) * c->n_all, hipMemcpyDeviceToHost));
    if (c->n_fix)
        EP_HIP(hipMemcpy(pos.data(), c->V[c->cur] + c->n_place, sizeof(float2) * c->n_fix, hipMemcpyDeviceToHost));
    for (int k = 0; k < c->n_all; ++k) {
        const int32_t o = caller_cell(c->cell_order, k);
        if (w) w[o] = wh[k].x;
        if (h) h[o] = wh[k].y;
    }
    for (int k = 0; k < c->n_fix; ++k) {
        if (fix_x) fix_x[k] = pos[k].x;
        if (fix_y) fix_y[k] = pos[k].y;
    }
    return EP_OK;
}

extern "C" ep_status ep_get_nets(ep_ctx *c, int32_t *net_start, int32_t *pin_cell, float *pin_dx,
                                 float *pin_dy, int32_t *net_perm)
{
    if (!c) return fail(EP_EINVAL, "ep_get_nets: null context");
    if (net_start) EP_HIP(hipMemcpy(net_start, c->net_start, sizeof(int32_t) * (c->n_nets + 1), hipMemcpyDeviceToHost));
    if (pin_cell && c->n_pins) {
        EP_HIP(hipMemcpy(pin_cell, c->pin_cell, sizeof(int32_t) * c->n_pins, hipMemcpyDeviceToHost));
        for (int q = 0; q < c->n_pins; ++q) pin_cell[q] = caller_cell(c->cell_order, pin_cell[q]);
    }
    if ((pin_dx || pin_dy) && c->n_pins) {
        std::vector<float2> o(c->n_pins);
        EP_HIP(hipMemcpy(o.data(), c->pin_off, sizeof(float2) * c->n_pins, hipMemcpyDeviceToHost));
        for (int q = 0; q < c->n_pins; ++q) {
            if (pin_dx) pin_dx[q] = o[q].x;
            if (pin_dy) pin_dy[q] = o[q].y;
        }
    }
    if (net_perm) std::memcpy(net_perm, c->net_perm.data(), sizeof(int32_t) * c->n_nets);
    return EP_OK;
}

extern "C" ep_status ep_set_positions(ep_ctx *c, const float *x, const float *y)
{
    if (!c || !x || !y) return fail(EP_EINVAL, "ep_set_positions: null argument");
    std::vector<float2> p(std::max(1, c->n_place)), wh(std::max(1, c->n_place));
    EP_HIP(hipMemcpy(wh.data(), c->wh, sizeof(float2) * c->n_place, hipMemcpyDeviceToHost));
    for (int k = 0; k < c->n_place; ++k) {
        const int32_t o = caller_cell(c->cell_order, k);
        if (!std::isfinite(x[o]) || !std::isfinite(y[o]))
            return fail(EP_ENAN, "ep_set_positions: non-finite position at cell %d", o);
        p[k] = make_float2(std::min(std::max(x[o], c->xl), c->xh - wh[k].x),
                           std::min(std::max(y[o], c->yl), c->yh - wh[k].y));
    }
    EP_HIP(hipMemcpy(c->V[c->cur], p.data(), sizeof(float2) * c->n_place, hipMemcpyHostToDevice));
    EP_HIP(hipMemcpy(c->U[c->cur], p.data(), sizeof(float2) * c->n_place, hipMemcpyHostToDevice));
    c->opt_ready = false;
    return EP_OK;
}

extern "C" ep_status ep_get_positions(ep_ctx *c, float *x, float *y)
{
    if (!c || !x || !y) return fail(EP_EINVAL, "ep_get_positions: null argument");
    std::vector<float2> p(std::max(1, c->n_place));
    EP_HIP(hipMemcpy(p.data(), c->V[c->cur], sizeof(float2) * c->n_place, hipMemcpyDeviceToHost));
    for (int k = 0; k < c->n_place; ++k) {
        const int32_t o = caller_cell(c->cell_order, k);
        x[o] = p[k].x;
        y[o] = p[k].y;
    }
    return EP_OK;
}

static ep_status run_density(ep_ctx *c)
{
    launch_sort(*c, c->V[c->cur], false, c->stream);
    launch_density(*c, c->stream);
    launch_overflow(*c, c->stream);
    launch_finalize(*c, c->stream);
    (void)hipMemcpyAsync(c->h_res, c->d_res, sizeof(DevResult), hipMemcpyDeviceToHost, c->stream);
    ep_status st = check_launch();
    if (st != EP_OK) return st;
    return sync(*c);
}

extern "C" ep_status ep_density(ep_ctx *c, double *overflow)
{
    if (!c) return fail(EP_EINVAL, "ep_density: null context");
    ep_status st = run_density(c);
    if (st != EP_OK) return st;
    if (overflow) *overflow = c->h_res->tau;
    return EP_OK;
}

extern "C" ep_status ep_get_density(ep_ctx *c, float *field_map, float *mov_map, float *fix_map)
{
    if (!c) return fail(EP_EINVAL, "ep_get_density: null context");
    const size_t nb = (size_t)c->Mx * c->My;
    if (field_map) EP_HIP(hipMemcpy(field_map, c->map_field, nb * sizeof(float), hipMemcpyDeviceToHost));
    if (mov_map) EP_HIP(hipMemcpy(mov_map, c->map_mov, nb * sizeof(float), hipMemcpyDeviceToHost));
    if (fix_map) EP_HIP(hipMemcpy(fix_map, c->map_fix, nb * sizeof(float), hipMemcpyDeviceToHost));
    return EP_OK;
}

extern "C" ep_status ep_field(ep_ctx *c, int32_t precision)
{
    if (!c) return fail(EP_EINVAL, "ep_field: null context");
    ep_status pst = check_precision(precision, "ep_field");
    if (pst != EP_OK) return pst;
    launch_field(*c, precision, c->stream);
    ep_status st = check_launch();
    if (st != EP_OK) return st;
    return sync(*c);
}

extern "C" ep_status ep_field_error(ep_ctx *c, int32_t precision, double *err_ex, double *err_ey)
{
    if (!c) return fail(EP_EINVAL, "ep_field_error: null context");
    ep_status st = check_precision(precision, "ep_field_error");
    if (st != EP_OK) return st;
    launch_field(*c, precision, c->stream);
    if ((st = check_launch()) != EP_OK) return st;
    return field_error(*c, err_ex, err_ey);
}

extern "C" ep_status ep_get_field(ep_ctx *c, float *ex, float *ey)
{
    if (!c) return fail(EP_EINVAL, "ep_get_field: null context");
    const size_t nb = (size_t)c->Mx * c->My;
    if (!ex && !ey) return EP_OK;
    std::vector<float2> e(nb);
    EP_HIP(hipMemcpy(e.data(), c->exy, nb * sizeof(float2), hipMemcpyDeviceToHost));
    for (size_t b = 0; b < nb; ++b) {
        if (ex) ex[b] = e[b].x;
        if (ey) ey[b] = e[b].y;
    }
    return EP_OK;
}

// Per-cell pairs of cells [mov | fill] back to the caller, in the caller's order.
static ep_status copy_pairs(const ep_ctx *c, const float2 *d, int n, float *ox, float *oy)
{
    if (!ox && !oy) return EP_OK;
    std::vector<float2> h(std::max(1, n));
    EP_HIP(hipMemcpy(h.data(), d, sizeof(float2) * n, hipMemcpyDeviceToHost));
    for (int k = 0; k < n; ++k) {
        const int32_t o = caller_cell(c->cell_order, k);
        if (ox) ox[o] = h[k].x;
        if (oy) oy[o] = h[k].y;
    }
    return EP_OK;
}

extern "C" ep_status ep_wa_grad(ep_ctx *c, float gamma, double *wa, double *hpwl)
{
    if (!c) return fail(EP_EINVAL, "ep_wa_grad: null context");
    if (!(gamma > 0.f)) return fail(EP_EINVAL, "ep_wa_grad: gamma must be > 0");
    c->h_scal->gamma_inv = 1.0f / gamma;
    push_scalars(*c, c->stream);
    launch_wa(*c, c->V[c->cur], c->stream);
    launch_finalize(*c, c->stream);
    (void)hipMemcpyAsync(c->h_res, c->d_res, sizeof(DevResult), hipMemcpyDeviceToHost, c->stream);
    ep_status st = check_launch();
    if (st != EP_OK) return st;
    if ((st = sync(*c)) != EP_OK) return st;
    if (wa) *wa = c->h_res->wa;
    if (hpwl) *hpwl = c->h_res->hpwl;
    return EP_OK;
}

extern "C" ep_status ep_get_wa_grad(ep_ctx *c, float *gx, float *gy)
{
    if (!c) return fail(EP_EINVAL, "ep_get_wa_grad: null context");
    float2 *gw = nullptr;
    ep_status st = dalloc(&gw, std::max(1, c->n_place));
    if (st != EP_OK) return st;
    c->h_scal->lambda = 0.0f;
    push_scalars(*c, c->stream);
    float2 *scratch = c->G[c->cur ^ 1];
    launch_cell_grad(*c, c->V[c->cur], nullptr, nullptr, scratch, gw, nullptr, c->stream);
    st = check_launch();
    if (st == EP_OK) st = sync(*c);
    if (st == EP_OK) st = copy_pairs(c, gw, c->n_place, gx, gy);
    (void)hipFree(gw);
    return st;
}

extern "C" ep_status ep_gradient(ep_ctx *c, float gamma, float lambda, float *gx, float *gy,
                                 float *gwx, float *gwy, float *fx, float *fy, double *hpwl,
                                 double *overflow)
{
    if (!c) return fail(EP_EINVAL, "ep_gradient: null context");
    if (!(gamma > 0.f)) return fail(EP_EINVAL, "ep_gradient: gamma must be > 0");
    float2 *gw = nullptr, *f = nullptr;
    ep_status st = dalloc(&gw, std::max(1, c->n_place));
    if (st == EP_OK) st = dalloc(&f, std::max(1, c->n_place));
    if (st != EP_OK) {
        if (gw) (void)hipFree(gw);
        return st;
    }
    c->h_scal->gamma_inv = 1.0f / gamma;
    c->h_scal->lambda = lambda;
    push_scalars(*c, c->stream);
    enqueue_eval(*c, c->V[c->cur], nullptr, nullptr, c->G[c->cur], gw, f, c->stream);
    st = check_launch();
    if (st == EP_OK) st = sync(*c);
    if (st == EP_OK) st = copy_pairs(c, c->G[c->cur], c->n_place, gx, gy);
    if (st == EP_OK) st = copy_pairs(c, gw, c->n_place, gwx, gwy);
    if (st == EP_OK) st = copy_pairs(c, f, c->n_place, fx, fy);
    (void)hipFree(gw);
    (void)hipFree(f);
    if (st != EP_OK) return st;
    if (hpwl) *hpwl = c->h_res->hpwl;
    if (overflow) *overflow = c->h_res->tau;
    c->opt_ready = false;  // G[cur] was overwritten with caller-chosen gamma/lambda
    return EP_OK;
}

extern "C" ep_status ep_cell_bins(ep_ctx *c, int32_t stretched, int32_t *out)
{
    if (!c || !out) return fail(EP_EINVAL, "ep_cell_bins: null argument");
    int32_t *d = nullptr;
    ep_status st = dalloc(&d, 4 * (size_t)std::max(1, c->n_place));
    if (st != EP_OK) return st;
    launch_cell_bins(*c, c->V[c->cur], stretched ? 1 : 0, d, c->stream);
    st = check_launch();
    if (st == EP_OK) st = sync(*c);
    if (st == EP_OK) {
        std::vector<int32_t> h(4 * (size_t)std::max(1, c->n_place));
        hipError_t e = hipMemcpy(h.data(), d, sizeof(int32_t) * 4 * (size_t)c->n_place, hipMemcpyDeviceToHost);
        if (e != hipSuccess) st = fail(EP_EDEVICE, "ep_cell_bins: copy failed");
        for (int k = 0; st == EP_OK && k < c->n_place; ++k)
            std::memcpy(out + 4 * (size_t)caller_cell(c->cell_order, k), &h[4 * (size_t)k], 4 * sizeof(int32_t));
    }
    (void)hipFree(d);
    return st;
}

extern "C" ep_status ep_hpwl(ep_ctx *c, double *hpwl)
{
    if (!c || !hpwl) return fail(EP_EINVAL, "ep_hpwl: null argument");
    return ep_wa_grad(c, c->bw, nullptr, hpwl);
}

extern "C" ep_status ep_overflow(ep_ctx *c, double *overflow)
{
    if (!c || !overflow) return fail(EP_EINVAL, "ep_overflow: null argument");
    return ep_density(c, overflow);
}

extern "C" ep_status ep_set_precision(ep_ctx *c, int32_t precision)
{
    if (!c) return fail(EP_EINVAL, "ep_set_precision: null context");
    ep_status pst = check_precision(precision, "ep_set_precision");
    if (pst != EP_OK) return pst;
    if (precision != c->prm.precision) {
        c->prm.precision = precision;
        drop_graphs(*c);
    }
    return EP_OK;
}

// ---------------------------------------------------------------- optimiser

extern "C" ep_status ep_init_optimizer(ep_ctx *c)
{
    if (!c) return fail(EP_EINVAL, "ep_init_optimizer: null context");
    const int a = c->cur, b = c->cur ^ 1;
    // U = V (the Nesterov major point starts at the initial placement)
    EP_HIP(hipMemcpyAsync(c->U[a], c->V[a], sizeof(float2) * c->n_place, hipMemcpyDeviceToDevice, c->stream));
    ep_status st = run_density(c);
    if (st != EP_OK) return st;
    c->gamma = gamma_of(*c, c->h_res->tau);
    // gradient pieces at v_0 with lambda = 0 give |gW|_1 and |F|_1
    c->h_scal->gamma_inv = (float)(1.0 / c->gamma);
    c->h_scal->lambda = 0.0f;
    push_scalars(*c, c->stream);
    enqueue_eval(*c, c->V[a], nullptr, nullptr, c->G[a], nullptr, nullptr, c->stream);
    if ((st = check_launch()) != EP_OK || (st = sync(*c)) != EP_OK) return st;
    c->evals++;
    const double sw = c->h_res->sum_gw, sf = c->h_res->sum_f;
    c->lambda = sf > 0 ? (double)c->prm.lambda_eta * sw / sf : (double)c->prm.lambda_eta;
    c->lambda0 = c->lambda;
    c->prev_hpwl = c->h_res->hpwl;
    c->last_hpwl = c->h_res->hpwl;
    c->last_tau = c->h_res->tau;
    c->last_wa = c->h_res->wa;
    // preconditioned gradient with lambda_0
    c->h_scal->lambda = (float)c->lambda;
    push_scalars(*c, c->stream);
    launch_cell_grad(*c, c->V[a], nullptr, nullptr, c->G[a], nullptr, nullptr, c->stream);
    launch_finalize(*c, c->stream);
    (void)hipMemcpyAsync(c->h_res, c->d_res, sizeof(DevResult), hipMemcpyDeviceToHost, c->stream);
    if ((st = check_launch()) != EP_OK || (st = sync(*c)) != EP_OK) return st;
    if (c->h_res->nonfinite) return fail(EP_ENAN, "ep_init_optimizer: non-finite gradient at v_0");
    // alpha_0 from a probe v_-1 = v_0 - eps g_0
    const double gmax = c->h_res->gmax;
    const double eps = gmax > 0 ? (double)c->prm.init_disp_bins * c->bw / gmax : 0.0;
    c->h_scal->probe_eps = (float)eps;
    push_scalars(*c, c->stream);
    launch_probe(*c, c->V[a], c->G[a], c->V[b], c->stream);
    enqueue_eval(*c, c->V[b], c->V[a], c->G[a], c->G[b], nullptr, nullptr, c->stream);
    if ((st = check_launch()) != EP_OK || (st = sync(*c)) != EP_OK) return st;
    c->evals++;
    const double num = std::sqrt(c->h_res->dv2), den = std::sqrt(c->h_res->dg2);
    c->alpha = den > 0 ? num / den : eps;
    if (!(c->alpha > 0) || !std::isfinite(c->alpha)) c->alpha = eps > 0 ? eps : 1e-3;
    c->ak = 1.0;
    c->iteration = 0;
    c->h_alpha1 = c->h_alpha2 = c->alpha;
    c->h_prev_tau = c->last_tau;
    c->h_prev_hpwl = c->prev_hpwl;
    c->h_ratio = 1.0;
    c->h_gnorm = std::sqrt(c->h_res->gnorm2);
    c->h_dv = num;
    c->h_dg = den;
    c->h_sum_ratio = sf > 0 ? sw / sf : 1.0;
    c->h_evals = 1;
    // the step model's backoff and counters belong to one run: a new run (after
    // ep_set_positions or a fresh context) starts with the model proposing
    c->model_skip = 0;
    c->model_backoff = 0;
    c->learned_proposals = 0;
    c->learned_accepted = 0;
    c->opt_ready = true;
    return EP_OK;
}

// 1 / sqrt(placeable cells): turns the FP64 norms into per-cell RMS values.
static double inv_sqrt_np(const ep_ctx *c) { return 1.0 / std::sqrt((double)std::max(1, c->n_place)); }

// Step-model features of the coming iteration (layout in include/eplace_mlp.h).
// Lengths are in bin widths and norms per cell, so instances of any size and
// grid share one feature scale; k_place.hip's ctl_features is the same.
static void step_features(const ep_ctx *c, double coef, float *f)
{
    auto lg = [](double v) { return (float)sched_log10(std::max(v, 1e-30)); };
    const double bw = c->bw, isn = inv_sqrt_np(c);
    f[0] = lg(c->alpha / bw);
    f[1] = lg(c->h_alpha1 / bw);
    f[2] = lg(c->h_alpha2 / bw);
    f[3] = (float)c->last_tau;
    f[4] = (float)(c->last_tau - c->h_prev_tau);
    f[5] = lg(c->lambda / c->lambda0);
    f[6] = lg(c->gamma / bw);
    f[7] = (float)((c->last_hpwl - c->h_prev_hpwl) / std::max(c->last_hpwl, 1e-30));
    f[8] = lg(c->h_gnorm * isn);
    f[9] = lg(c->h_dv * isn / bw);
    f[10] = lg(c->h_dg * isn);
    f[11] = (float)coef;
    f[12] = (float)c->h_evals;
    f[13] = lg(c->lambda / c->h_sum_ratio);
    f[14] = lg(c->h_ratio);
    f[15] = (float)(c->iteration / 1000.0);
}

// One trial with step `alpha` (model proposal when use_model); returns alpha_hat
// and the step actually used.
static ep_status one_trial(ep_ctx *c, double alpha, double coef, int use_model, double *alpha_hat,
                           double *alpha_used)
{
    c->h_scal->alpha = (float)alpha;
    c->h_scal->coef = (float)coef;
    c->h_scal->gamma_inv = (float)(1.0 / c->gamma);
    c->h_scal->lambda = (float)c->lambda;
    c->h_scal->use_model = use_model;
    ep_status st = trial_eval(*c);
    if (st != EP_OK) return st;
    c->evals++;
    if (c->h_res->nonfinite)
        return fail(EP_ENAN, "ep_step: non-finite gradient at iteration %d", c->iteration);
    const double used = c->h_res->alpha_used;
    const double nn = std::sqrt(c->h_res->dv2), dd = std::sqrt(c->h_res->dg2);
    *alpha_hat = dd > 0 ? nn / dd : used;
    *alpha_used = used;
    return EP_OK;
}

extern "C" ep_status ep_step(ep_ctx *c, ep_step_info *info)
{
    if (!c) return fail(EP_EINVAL, "ep_step: null context");
    ep_status st;
    if (!c->opt_ready && (st = ep_init_optimizer(c)) != EP_OK) return st;
    const double ak1 = 0.5 * (1.0 + std::sqrt(4.0 * c->ak * c->ak + 1.0));
    const double coef = (c->ak - 1.0) / ak1;
    float feat[EP_MLP_IN];
    step_features(c, coef, feat);
    std::memcpy(c->h_scal->feat, feat, sizeof(feat));
    double alpha_hat = c->alpha, used = c->alpha;
    int trials = 0;
    bool accepted = false, learned_used = false, learned_ok = false;
    // learned proposal first, verified by the backtracking condition
    const bool learned = c->prm.learned_step && c->mlp_loaded;
    const bool propose = learned && c->model_skip == 0;
    if (learned && !propose) c->model_skip--;
    if (propose) {
        learned_used = true;
        c->learned_proposals++;
        if ((st = one_trial(c, c->alpha, coef, 1, &alpha_hat, &used)) != EP_OK) return st;
        ++trials;
        if (std::isfinite(alpha_hat) && alpha_hat >= 0.95 * used) {
            accepted = learned_ok = true;
            c->learned_accepted++;
            c->model_backoff = 0;
        } else {  // the model sits out 2, 4, ... kModelBackoffMax iterations
            c->model_backoff = c->model_backoff ? std::min(2 * c->model_backoff, kModelBackoffMax) : 2;
            c->model_skip = c->model_backoff;
        }
    }
    // plain backtracking search from the Lipschitz step (also the fallback)
    double alpha = c->alpha;
    for (int t = 0; !accepted && t < c->prm.max_backtrack; ++t) {
        if ((st = one_trial(c, alpha, coef, 0, &alpha_hat, &used)) != EP_OK) return st;
        ++trials;
        if (alpha_hat >= 0.95 * used || t == c->prm.max_backtrack - 1) accepted = true;
        else alpha = alpha_hat;
    }
    // record (features, accepted step) for training the step model
    if (c->rec_target.size() < kMaxStepRecords) {
        c->rec_feat.insert(c->rec_feat.end(), feat, feat + EP_MLP_IN);
        c->rec_target.push_back((float)sched_log10(std::max(used / c->bw, 1e-30)));
    }
    // accept: the trial buffers become the current ones
    c->cur ^= 1;
    c->alpha = alpha_hat;
    c->ak = ak1;
    const double hp = c->h_res->hpwl, tau = c->h_res->tau;
    c->h_alpha2 = c->h_alpha1;
    c->h_alpha1 = used;
    c->h_prev_tau = c->last_tau;
    c->h_prev_hpwl = c->last_hpwl;
    c->h_ratio = alpha_hat / std::max(used, 1e-30);
    c->h_gnorm = std::sqrt(c->h_res->gnorm2);
    c->h_dv = std::sqrt(c->h_res->dv2);
    c->h_dg = std::sqrt(c->h_res->dg2);
    c->h_sum_ratio = c->h_res->sum_f > 0 ? c->h_res->sum_gw / c->h_res->sum_f : 1.0;
    c->h_evals = trials;
    c->last_hpwl = hp;
    c->last_tau = tau;
    c->last_wa = c->h_res->wa;
    if (info) {
        info->iteration = c->iteration;
        info->evals = trials;
        info->learned_used = learned_used ? 1 : 0;
        info->learned_accepted = learned_ok ? 1 : 0;
        info->hpwl = hp;
        info->overflow = tau;
        info->lambda = c->lambda;
        info->gamma = c->gamma;
        info->alpha = c->alpha;
        info->wa = c->h_res->wa;
    }
    // lambda (mu schedule) and gamma updates for the next iteration
    const double dh = hp - c->prev_hpwl;
    double mu = dh < 0 ? c->prm.mu_max : sched_pow((double)c->prm.mu_max, 1.0 - dh / c->prm.hpwl_ref);
    mu = std::min<double>(std::max<double>(mu, c->prm.mu_min), c->prm.mu_max);
    c->lambda *= mu;
    c->prev_hpwl = hp;
    c->gamma = gamma_of(*c, tau);
    c->iteration++;
    return EP_OK;
}

// ---------------------------------------------------------------- device-controlled run
//
// ep_run with graphs: one captured graph = kCtlTrialsPerGraph x (model
// proposal, trial, evaluation, finalize + the ep_step decision on the device).
// The kernels take the current and trial buffers from Pair arguments indexed
// by the device's `cur`, so the same graph serves every trial; the host
// replays it kCtlReplays times per batch and reads the control state back
// after each batch.  Two batches are in flight: the host waits for the state
// of one while the next one runs, so the GPU does not idle for the host's
// round trip.  After the stop the remaining trials return at once (halt flag).

constexpr int kCtlTrialsPerGraph = 8;  // one replay per read-back (a graph boundary costs ~3.5 us)
constexpr int kCtlReplays = 1;

static CtlParams ctl_params(const ep_ctx *c)
{
    CtlParams p{};
    p.mu_min = c->prm.mu_min;
    p.mu_max = c->prm.mu_max;
    p.hpwl_ref = c->prm.hpwl_ref;
    p.stop_overflow = c->prm.stop_overflow;
    p.gamma_base = (double)c->prm.gamma_bins * (double)c->bw;  // gamma_at's base
    p.gamma_schedule = c->prm.gamma_schedule;
    p.max_iters = c->prm.max_iters;
    p.max_backtrack = c->prm.max_backtrack;
    p.learned = (c->prm.learned_step && c->mlp_loaded) ? 1 : 0;
    p.log_cap = c->ctl_cap;
    p.rec_cap = c->ctl_cap;
    p.bw = c->bw;
    p.inv_sqrt_np = inv_sqrt_np(c);
    return p;
}

static ep_status ensure_ctl(ep_ctx *c)
{
    const int cap = std::max(1, c->prm.max_iters);
    if (c->ctl_cap == cap) return EP_OK;
    if (c->ctl_exec) (void)hipGraphExecDestroy(c->ctl_exec);
    c->ctl_exec = nullptr;
    if (c->d_log) (void)hipFree(c->d_log);
    if (c->d_rec) (void)hipFree(c->d_rec);
    c->d_log = nullptr;
    c->d_rec = nullptr;
    c->ctl_cap = 0;
    ep_status st;
    if (!c->d_ctl && (st = dalloc(&c->d_ctl, 1)) != EP_OK) return st;
    if (!c->h_ctl && hipHostMalloc((void **)&c->h_ctl, 2 * sizeof(DevCtl)) != hipSuccess)
        return fail(EP_ENOMEM, "ep_run: pinned control state allocation failed");
    if ((st = dalloc(&c->d_log, (size_t)cap)) != EP_OK) return st;
    if ((st = dalloc(&c->d_rec, (size_t)cap * 17)) != EP_OK) return st;
    c->ctl_cap = cap;
    return EP_OK;
}

static ep_status build_ctl_graph(ep_ctx *c, const CtlParams &p)
{
    Ctx &x = *c;
    hipStream_t s = x.stream;
    const Pair<const float2> vc{{x.V[0], x.V[1]}}, uc{{x.U[0], x.U[1]}}, gc{{x.G[0], x.G[1]}};
    const Pair<float2> vn{{x.V[1], x.V[0]}}, un{{x.U[1], x.U[0]}}, gn{{x.G[1], x.G[0]}};
    const Pair<const float2> vt{{x.V[1], x.V[0]}};
    const Pair<const float2> none{{nullptr, nullptr}};  // |v' - v|^2 comes from k_trial
    hipGraph_t g = nullptr;
    EP_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    for (int t = 0; t < kCtlTrialsPerGraph; ++t) {
        // the model's proposal for this trial was made by the previous decision
        // (k_finalize) or by k_step_begin
        launch_trial(x, vc, uc, gc, vn, un, true, s);
        if (x.n_place >= fork_min_cells(x.prm)) {
            // fork right after the trial: the tile sort, the density scatter and
            // the transform products on the second stream (the critical chain;
            // the products take issue priority), the WA gradient and the
            // per-cell pin sums beside them
            EP_CAP(hipEventRecord(x.fork, s), s);
            EP_CAP(hipStreamWaitEvent(x.side, x.fork, 0), s);
            launch_sort(x, vt, true, x.side);
            launch_density(x, x.side);
            launch_overflow(x, x.side);
            launch_field(x, x.prm.precision, x.side);
            EP_CAP(hipEventRecord(x.join, x.side), s);
            launch_wa(x, vt, s);
            launch_pin_sum(x, s);
            EP_CAP(hipStreamWaitEvent(s, x.join, 0), s);
            launch_cell_grad(x, vt, none, gc, gn, nullptr, nullptr, s, x.gw_cell);
        } else {
            launch_sort(x, vt, true, s);
            launch_density(x, s);
            launch_overflow(x, s);
            launch_wa(x, vt, s);
            launch_field(x, x.prm.precision, s);
            launch_cell_grad(x, vt, none, gc, gn, nullptr, nullptr, s);
        }
        launch_finalize_decide(x, p, s);
    }
    EP_CAP(hipGetLastError(), s);
    EP_HIP(hipStreamEndCapture(s, &g));
    hipError_t e = hipGraphInstantiate(&c->ctl_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        c->ctl_exec = nullptr;
        return fail(EP_EDEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    return EP_OK;
}

// Host optimiser state <-> DevCtl.
static void ctl_upload_state(const ep_ctx *c, DevCtl &S)
{
    std::memset(&S, 0, sizeof(S));
    S.alpha = c->alpha; S.ak = c->ak; S.lambda = c->lambda; S.lambda0 = c->lambda0; S.gamma = c->gamma;
    S.prev_hpwl = c->prev_hpwl; S.last_hpwl = c->last_hpwl; S.last_tau = c->last_tau;
    S.last_wa = c->last_wa;
    S.h_alpha1 = c->h_alpha1; S.h_alpha2 = c->h_alpha2; S.h_prev_tau = c->h_prev_tau;
    S.h_prev_hpwl = c->h_prev_hpwl; S.h_ratio = c->h_ratio; S.h_gnorm = c->h_gnorm;
    S.h_dv = c->h_dv; S.h_dg = c->h_dg; S.h_sum_ratio = c->h_sum_ratio;
    S.evals = c->evals; S.iteration = c->iteration; S.h_evals = c->h_evals;
    S.learned_proposals = c->learned_proposals; S.learned_accepted = c->learned_accepted;
    S.model_skip = c->model_skip; S.model_backoff = c->model_backoff;
    S.cur = c->cur;
}

static void ctl_download_state(ep_ctx *c, const DevCtl &S)
{
    c->alpha = S.alpha; c->ak = S.ak; c->lambda = S.lambda; c->gamma = S.gamma;
    c->prev_hpwl = S.prev_hpwl; c->last_hpwl = S.last_hpwl; c->last_tau = S.last_tau;
    c->last_wa = S.last_wa;
    c->h_alpha1 = S.h_alpha1; c->h_alpha2 = S.h_alpha2; c->h_prev_tau = S.h_prev_tau;
    c->h_prev_hpwl = S.h_prev_hpwl; c->h_ratio = S.h_ratio; c->h_gnorm = S.h_gnorm;
    c->h_dv = S.h_dv; c->h_dg = S.h_dg; c->h_sum_ratio = S.h_sum_ratio;
    c->evals = S.evals; c->iteration = S.iteration; c->h_evals = S.h_evals;
    c->learned_proposals = S.learned_proposals; c->learned_accepted = S.learned_accepted;
    c->model_skip = S.model_skip; c->model_backoff = S.model_backoff;
    c->cur = S.cur;
}

// Iterations of ep_run under device control; the caller has set up the optimiser.
static ep_status run_device(ep_ctx *c, ep_step_info *log, int32_t log_cap, int *n_iter)
{
    *n_iter = 0;
    ep_status st = ensure_ctl(c);
    if (st != EP_OK) return st;
    const CtlParams p = ctl_params(c);
    if (!c->ctl_exec && (st = build_ctl_graph(c, p)) != EP_OK) return st;
    ctl_upload_state(c, c->h_ctl[0]);
    EP_HIP(hipMemcpyAsync(c->d_ctl, c->h_ctl, sizeof(DevCtl), hipMemcpyHostToDevice, c->stream));
    c->h_scal->sel = c->cur;
    c->h_scal->halt = 0;
    push_scalars(*c, c->stream);
    launch_step_begin(*c, p, c->stream);
    if ((st = check_launch()) != EP_OK) return st;
    // batch b reads the state back into slot b % 2; the state upload above has
    // left slot 0 before batch 0's read-back is queued behind it on the stream
    int slot = 0;
    for (int64_t b = 0;; ++b, slot ^= 1) {
        for (int r = 0; r < kCtlReplays; ++r) EP_HIP(hipGraphLaunch(c->ctl_exec, c->stream));
        EP_HIP(hipMemcpyAsync(c->h_ctl + slot, c->d_ctl, sizeof(DevCtl), hipMemcpyDeviceToHost, c->stream));
        EP_HIP(hipEventRecord(c->ctl_read[slot], c->stream));
        if (b == 0) continue;
        const hipError_t e = hipEventSynchronize(c->ctl_read[slot ^ 1]);
        if (e != hipSuccess) return fail(EP_EDEVICE, "ep_run: %s", hipGetErrorString(e));
        const DevCtl &P = c->h_ctl[slot ^ 1];
        if (P.done || P.error) break;
    }
    // the batch queued after the stop returned at once: the last read-back
    // (slot) holds the final state
    if ((st = sync(*c)) != EP_OK) return st;
    DevCtl &H = c->h_ctl[slot];
    const int n = std::min(H.n_log, c->ctl_cap);
    *n_iter = n;
    if (log && log_cap > 0 && n > 0)
        EP_HIP(hipMemcpyAsync(log, c->d_log, sizeof(ep_step_info) * std::min(n, log_cap),
                              hipMemcpyDeviceToHost, c->stream));
    const size_t room = kMaxStepRecords - std::min(kMaxStepRecords, c->rec_target.size());
    const size_t nr = std::min((size_t)std::min(H.n_rec, c->ctl_cap), room);
    std::vector<float> rec(nr * 17);
    if (nr > 0)
        EP_HIP(hipMemcpyAsync(rec.data(), c->d_rec, sizeof(float) * rec.size(), hipMemcpyDeviceToHost,
                              c->stream));
    EP_HIP(hipMemcpyAsync(c->h_res, c->d_res, sizeof(DevResult), hipMemcpyDeviceToHost, c->stream));
    // host-controlled launches after the run must not see the halt flag
    c->h_scal->halt = 0;
    push_scalars(*c, c->stream);
    if ((st = sync(*c)) != EP_OK) return st;
    for (size_t r = 0; r < nr; ++r) {
        c->rec_feat.insert(c->rec_feat.end(), &rec[r * 17], &rec[r * 17] + 16);
        c->rec_target.push_back(rec[r * 17 + 16]);
    }
    ctl_download_state(c, H);
    if (H.error) return fail(EP_ENAN, "ep_run: non-finite gradient at iteration %d", c->iteration);
    return EP_OK;
}

extern "C" ep_status ep_run(ep_ctx *c, ep_run_info *info, ep_step_info *log, int32_t log_cap)
{
    if (!c) return fail(EP_EINVAL, "ep_run: null context");
    auto t0 = std::chrono::steady_clock::now();
    ep_status st;
    // counts of this call only (the optimiser set-up's evaluations included when
    // this call made it)
    const int64_t ev0 = c->evals;
    if (!c->opt_ready && (st = ep_init_optimizer(c)) != EP_OK) return st;
    const int32_t lp0 = c->learned_proposals, la0 = c->learned_accepted;
    if (c->prm.use_graph && c->iteration < c->prm.max_iters) {
        int n = 0;
        if ((st = run_device(c, log, log_cap, &n)) != EP_OK) return st;
    } else {
        int n = 0;
        while (c->iteration < c->prm.max_iters) {
            ep_step_info si;
            if ((st = ep_step(c, &si)) != EP_OK) return st;
            if (log && n < log_cap) log[n] = si;
            ++n;
            if (si.overflow <= c->prm.stop_overflow) break;
        }
    }
    auto t1 = std::chrono::steady_clock::now();
    if (info) {
        info->iterations = c->iteration;
        info->evals = c->evals - ev0;
        info->learned_proposals = c->learned_proposals - lp0;
        info->learned_accepted = c->learned_accepted - la0;
        info->hpwl = c->last_hpwl;
        info->overflow = c->last_tau;
        info->seconds = std::chrono::duration<double>(t1 - t0).count();
    }
    return EP_OK;
}

extern "C" ep_status ep_profile(ep_ctx *c, int32_t reps, ep_kernel_stat *out, int32_t cap,
                                int32_t *n)
{
    if (!c || !n || reps < 1) return fail(EP_EINVAL, "ep_profile: bad argument");
    const int a = c->cur, b = c->cur ^ 1;
    const double nb = (double)c->Mx * c->My, np = c->n_place, P = c->n_pins, N = c->n_nets;
    const double pm = c->n_pins_mov;
    if (c->gamma <= 0) c->gamma = gamma_of(*c, 1.0);
    c->h_scal->gamma_inv = (float)(1.0 / c->gamma);
    c->h_scal->lambda = (float)c->lambda;
    c->h_scal->alpha = (float)(c->alpha > 0 ? c->alpha : 0.0);
    c->h_scal->coef = 0.0f;
    push_scalars(*c, c->stream);
    // field input must be a real map: build it once at the current positions
    launch_sort(*c, c->V[a], false, c->stream);
    launch_density(*c, c->stream);
    launch_overflow(*c, c->stream);
    hipEvent_t e0, e1;
    EP_HIP(hipEventCreate(&e0));
    EP_HIP(hipEventCreate(&e1));
    struct Item {
        const char *name;
        int launches;
        double bytes, flops;
    };
    const double M = c->Mx;
    // algorithmic flops of one solve: six dense transform products (the 2-D DCT,
    // then two inverse 2-D transforms), 6 Mx My (Mx + My) = 12 M^3 on a square
    // grid.  The symmetric half-products execute half of that in FP32; the split
    // forms execute three products each (1.5x, at the 16-bit rate).
    const double Mxd = c->Mx, Myd = c->My;
    const double gemm_flops = 6.0 * Mxd * Myd * (Mxd + Myd);
    // reads 6 maps, writes 7; the halves of the six matrices: 1.5 (Mx^2 + My^2) floats
    const double field_bytes = 4.0 * 13 * nb + 4.0 * 1.5 * (Mxd * Mxd + Myd * Myd);
    // Instances that fork their captured evaluation sum the pin partials per cell
    // on the WA branch (k_pin_sum), and the cell gradient reads those sums: both
    // are timed that way here, as the placement runs them.
    const bool forked = c->n_place >= fork_min_cells(c->prm);
    std::vector<Item> items = {
        // positions and sizes read twice, packed cells written once
        {"cell_sort", 3, 2 * 16.0 * np + 16.0 * np + 8.0 * kSortTiles, 0},
        // packed cells read; two 64-bit accumulator maps flushed
        {"density_scatter", 1, 16.0 * np + 16.0 * nb, 0},
        // two accumulator maps and the fixed map read; two FP32 maps written, the
        // accumulators cleared
        {"overflow", 1, 44.0 * nb, 0},
        {"wa_grad", 1, 4.0 * (N + 1) + 12.0 * P + 8.0 * c->n_all + 8.0 * P, 0},
        {"field_transforms", 4, field_bytes, gemm_flops},
    };
    if (forked) {
        // CSR offsets and pin list read, the pins' partials gathered, one sum
        // per movable cell written
        items.push_back({"pin_sum", 1, 4.0 * (c->n_mov + 1) + 12.0 * pm + 8.0 * c->n_mov, 0});
        items.push_back({"cell_grad", 1, 36.0 * np + 8.0 * c->n_mov + 8.0 * nb, 0});
    } else {
        items.push_back({"cell_grad", 1, 36.0 * np + 12.0 * pm + 4.0 * c->n_mov + 8.0 * nb, 0});
    }
    // v, u, g, size read; v, u written (the tile counts stay in LDS)
    items.push_back({"nesterov_trial", 1, 40.0 * np, 0});
    items.push_back({"finalize", 1, 8.0 * c->part_total, 0});
    int k = 0;
    for (const Item &it : items) {
        EP_HIP(hipEventRecord(e0, c->stream));
        for (int r = 0; r < reps; ++r) {
            const std::string nm = it.name;
            if (nm == "cell_sort") launch_sort(*c, c->V[a], false, c->stream);
            else if (nm == "density_scatter") launch_density(*c, c->stream);
            else if (nm == "overflow") launch_overflow(*c, c->stream);
            else if (nm == "wa_grad") launch_wa(*c, c->V[a], c->stream);
            else if (nm == "field_transforms") launch_field(*c, c->prm.precision, c->stream);
            else if (nm == "pin_sum") launch_pin_sum(*c, c->stream);
            else if (nm == "cell_grad")
                launch_cell_grad(*c, same<const float2>(c->V[a]), same<const float2>(nullptr),
                                 same<const float2>(c->G[a]), same<float2>(c->G[b]), nullptr, nullptr, c->stream,
                                 forked ? c->gw_cell : nullptr);
            else if (nm == "nesterov_trial") launch_trial(*c, c->V[a], c->U[a], c->G[a], c->V[b], c->U[b], true, c->stream);
            else launch_finalize(*c, c->stream, true);
        }
        EP_HIP(hipEventRecord(e1, c->stream));
        // the counting trials left tile counts behind: consume them (the scan
        // clears the counts) before anything else sorts
        if (std::string(it.name) == "nesterov_trial") launch_sort(*c, c->V[b], true, c->stream);
        EP_HIP(hipEventSynchronize(e1));
        ep_status st = check_launch();
        if (st != EP_OK) return st;
        float ms = 0;
        EP_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (out && k < cap) {
            std::memset(&out[k], 0, sizeof(ep_kernel_stat));
            std::snprintf(out[k].name, sizeof(out[k].name), "%s", it.name);
            out[k].launches = it.launches;
            out[k].avg_us = 1000.0 * ms / reps;
            out[k].bytes = it.bytes;
            out[k].flops = it.flops;
        }
        ++k;
    }
    // the density scatter at the other tile sizes (cells per thread), each after
    // a sort whose block table matches it: the plan's choice is one of them
    const int cpt0 = c->dens.cpt;
    for (int cpt : {2, 4, 8, 16}) {
        if (cpt == cpt0) continue;
        c->dens.cpt = cpt;
        launch_sort(*c, c->V[a], false, c->stream);
        EP_HIP(hipEventRecord(e0, c->stream));
        for (int r = 0; r < reps; ++r) launch_density(*c, c->stream);
        EP_HIP(hipEventRecord(e1, c->stream));
        EP_HIP(hipEventSynchronize(e1));
        ep_status st = check_launch();
        if (st == EP_OK && k < cap && out) {
            float ms = 0;
            EP_HIP(hipEventElapsedTime(&ms, e0, e1));
            std::memset(&out[k], 0, sizeof(ep_kernel_stat));
            std::snprintf(out[k].name, sizeof(out[k].name), "density_scatter_cpt%d", cpt);
            out[k].launches = 1;
            out[k].avg_us = 1000.0 * ms / reps;
            out[k].bytes = 16.0 * np + 16.0 * nb;
        }
        c->dens.cpt = cpt0;
        if (st != EP_OK) return st;
        ++k;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *n = k;
    // leave the maps consistent with the current positions
    launch_fill_maps(*c, c->stream);
    launch_sort(*c, c->V[a], false, c->stream);
    launch_density(*c, c->stream);
    launch_overflow(*c, c->stream);
    return sync(*c);
}

extern "C" ep_status ep_take_step_records(ep_ctx *c, float *features, float *log_step, int32_t cap,
                                          int32_t *n)
{
    if (!c || !n) return fail(EP_EINVAL, "ep_take_step_records: null argument");
    const int32_t have = (int32_t)c->rec_target.size();
    *n = have;
    if (!features || !log_step) return EP_OK;  // query only
    const int32_t k = std::min(have, std::max(cap, 0));
    std::memcpy(features, c->rec_feat.data(), sizeof(float) * (size_t)k * EP_MLP_IN);
    std::memcpy(log_step, c->rec_target.data(), sizeof(float) * (size_t)k);
    *n = k;
    c->rec_feat.clear();
    c->rec_target.clear();
    return EP_OK;
}

extern "C" ep_status ep_set_step_model(ep_ctx *c, const float *w, int32_t n)
{
    if (!c || !w) return fail(EP_EINVAL, "ep_set_step_model: null argument");
    if (n != EP_MLP_NWEIGHTS)
        return fail(EP_EINVAL, "ep_set_step_model: expected %d weights, got %d", EP_MLP_NWEIGHTS, n);
    for (int i = 0; i < n; ++i)
        if (!std::isfinite(w[i])) return fail(EP_ENAN, "ep_set_step_model: non-finite weight %d", i);
    for (int i = 16; i < 34; ++i)
        if (i != 32 && !(w[i] > 0.0f)) return fail(EP_EINVAL, "ep_set_step_model: normalisation std must be > 0");
    if (!c->d_mlp_w) {
        ep_status st = dalloc(&c->d_mlp_w, (size_t)EP_MLP_NWEIGHTS);
        if (st != EP_OK) return st;
    }
    EP_HIP(hipMemcpy(c->d_mlp_w, w, sizeof(float) * n, hipMemcpyHostToDevice));
    c->mlp_loaded = true;
    drop_graphs(*c);  // captured graphs must now include the proposal kernel
    return EP_OK;
}

// Checks an AIE-format packet (see ep_wa_packet) and returns its geometry.
static ep_status check_packet(const float *in, int32_t n_in, float gamma, const float *out, int32_t n_out,
                              const char *who, int *net_size, int *n_nets, int *groups, int64_t *n_data)
{
    if (!in || !out || n_in < 4) return fail(EP_EINVAL, "%s: bad argument", who);
    if (!(gamma > 0.f)) return fail(EP_EINVAL, "%s: gamma must be > 0", who);
    const float fs = in[0], fn = in[1];
    if (!(fs >= 1.f) || fs != std::floor(fs) || !(fn >= 0.f) || fn != std::floor(fn) || fs > 65536.f ||
        fn > 1e8f)
        return fail(EP_EINVAL, "%s: bad control words [%g, %g]", who, fs, fn);
    *net_size = (int)fs;
    *n_nets = (int)fn;
    *groups = (*n_nets + 7) / 8;
    *n_data = (int64_t)*groups * *net_size * 8;
    if (n_in < 4 + *n_data)
        return fail(EP_EINVAL, "%s: packet holds %d values, needs %lld", who, n_in - 4, (long long)*n_data);
    if (n_out < *n_data)
        return fail(EP_EINVAL, "%s: output holds %d values, needs %lld", who, n_out, (long long)*n_data);
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || nd < 1) return fail(EP_EDEVICE, "%s: no HIP device", who);
    return EP_OK;
}

static ep_status wa_packet(const float *in, int32_t n_in, float gamma, float *out, int32_t n_out, int32_t exp_mode,
                           const char *who)
{
    if (exp_mode != EP_WA_EXP_HW && exp_mode != EP_WA_EXP_FAST64)
        return fail(EP_EINVAL, "%s: unknown exponent mode %d", who, exp_mode);
    int net_size = 0, n_nets = 0, groups = 0;
    int64_t n_data = 0;
    ep_status st = check_packet(in, n_in, gamma, out, n_out, who, &net_size, &n_nets, &groups, &n_data);
    if (st != EP_OK || n_data == 0) return st;
    float *d_in = nullptr, *d_out = nullptr;
    st = dalloc(&d_in, (size_t)n_data);
    if (st == EP_OK) st = dalloc(&d_out, (size_t)n_data);
    if (st == EP_OK && hipMemcpy(d_in, in + 4, sizeof(float) * n_data, hipMemcpyHostToDevice) != hipSuccess)
        st = fail(EP_EDEVICE, "%s: upload failed", who);
    if (st == EP_OK) {
        if (exp_mode == EP_WA_EXP_FAST64)
            launch_wa_packet_fast64(d_in, net_size, n_nets, groups, (double)gamma, d_out, nullptr);
        else
            launch_wa_packet(d_in, net_size, n_nets, groups, 1.0f / gamma, d_out, nullptr);
        if ((st = check_launch()) == EP_OK &&
            hipMemcpy(out, d_out, sizeof(float) * n_data, hipMemcpyDeviceToHost) != hipSuccess)
            st = fail(EP_EDEVICE, "%s: download failed", who);
    }
    if (d_in) (void)hipFree(d_in);
    if (d_out) (void)hipFree(d_out);
    return st;
}

extern "C" ep_status ep_wa_packet(const float *in, int32_t n_in, float gamma, float *out, int32_t n_out)
{
    return wa_packet(in, n_in, gamma, out, n_out, EP_WA_EXP_HW, "ep_wa_packet");
}

extern "C" ep_status ep_wa_packet_exp(const float *in, int32_t n_in, float gamma, float *out, int32_t n_out,
                                      int32_t exp_mode)
{
    return wa_packet(in, n_in, gamma, out, n_out, exp_mode, "ep_wa_packet_exp");
}

extern "C" ep_status ep_wa_packet_noalloc(const float *in, int32_t n_in, float gamma, float *out, int32_t n_out)
{
    int net_size = 0, n_nets = 0, groups = 0;
    int64_t n_data = 0;
    ep_status st = check_packet(in, n_in, gamma, out, n_out, "ep_wa_packet_noalloc", &net_size, &n_nets, &groups,
                                &n_data);
    if (st != EP_OK || n_data == 0) return st;
    if (n_data > EP_WA_PACKET_NOALLOC_MAX)
        return fail(EP_EINVAL, "ep_wa_packet_noalloc: packet holds %lld data values; the fixed staging holds %d",
                    (long long)n_data, EP_WA_PACKET_NOALLOC_MAX);
    float *stage = nullptr;
    if ((st = wa_packet_stage(&stage)) != EP_OK) return st;
    float *d_in = stage, *d_out = stage + EP_WA_PACKET_NOALLOC_MAX;
    // the kernel writes every partial of every group (zeros for absent lanes), so
    // nothing a previous call left in the staging is read or returned
    if (hipMemcpy(d_in, in + 4, sizeof(float) * n_data, hipMemcpyHostToDevice) != hipSuccess)
        return fail(EP_EDEVICE, "ep_wa_packet_noalloc: upload failed");
    launch_wa_packet(d_in, net_size, n_nets, groups, 1.0f / gamma, d_out, nullptr);
    if ((st = check_launch()) != EP_OK) return st;
    if (hipMemcpy(out, d_out, sizeof(float) * n_data, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(EP_EDEVICE, "ep_wa_packet_noalloc: download failed");
    return EP_OK;
}

extern "C" ep_status ep_transform(int32_t row_kind, int32_t col_kind, int32_t nx, int32_t ny, const float *in,
                                  float *out, int32_t precision)
{
    if (!in || !out) return fail(EP_EINVAL, "ep_transform: null array");
    auto kind_ok = [](int32_t k) { return k == EP_TR_NONE || k == EP_TR_DCT || k == EP_TR_IDCT || k == EP_TR_IDXST; };
    if (!kind_ok(row_kind) || !kind_ok(col_kind))
        return fail(EP_EINVAL, "ep_transform: unknown transform kind (%d, %d)", row_kind, col_kind);
    if (row_kind == EP_TR_NONE && col_kind == EP_TR_NONE)
        return fail(EP_EINVAL, "ep_transform: no axis to transform");
    if (nx < 16 || ny < 16 || nx % 16 || ny % 16 || nx > 16384 || ny > 16384 || (int64_t)nx * ny > ((int64_t)1 << 26))
        return fail(EP_EINVAL, "ep_transform: sizes %d x %d must be multiples of 16, at most 16384, "
                               "nx*ny <= 2^26", nx, ny);
    ep_status pst = check_precision(precision, "ep_transform");
    if (pst != EP_OK) return pst;
    const size_t nb = (size_t)nx * ny;
    for (size_t k = 0; k < nb; ++k)
        if (!std::isfinite(in[k])) return fail(EP_ENAN, "ep_transform: non-finite input at %zu", k);
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || nd < 1) return fail(EP_EDEVICE, "ep_transform: no HIP device");
    return run_transform(row_kind, col_kind, nx, ny, in, out, precision);
}
