// generator.cpp -- seeded synthetic placement instances (host only).
//
// Stands in for the reference's netlist load (DataBase.cpp:7-96): the
// measured workloads are synthetic, so instances are produced here from a
// seed, deterministically on any machine.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/eplace_gen.h"
#include "ep_common.h"

namespace {

// splitmix64 -> xoshiro256** ; deterministic across platforms.
struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed)
    {
        uint64_t z = seed;
        for (int i = 0; i < 4; ++i) {
            z += 0x9E3779B97F4A7C15ull;
            uint64_t t = z;
            t = (t ^ (t >> 30)) * 0xBF58476D1CE4E5B9ull;
            t = (t ^ (t >> 27)) * 0x94D049BB133111EBull;
            s[i] = t ^ (t >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next()
    {
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
    int64_t range(int64_t lo, int64_t hi) { return lo + (int64_t)(uniform() * (double)(hi - lo + 1)); }
    double normal()
    {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

struct Builder {
    std::vector<float> w, h, x, y;
    std::vector<int32_t> net_start{0}, pin_cell;
    std::vector<float> pin_dx, pin_dy;
    void add_pin(int32_t c, float dx, float dy)
    {
        pin_cell.push_back(c);
        pin_dx.push_back(dx);
        pin_dy.push_back(dy);
    }
    void end_net() { net_start.push_back((int32_t)pin_cell.size()); }
};

template <class T>
T *dup(const std::vector<T> &v)
{
    T *p = new (std::nothrow) T[v.size() ? v.size() : 1];
    if (p && !v.empty()) std::memcpy(p, v.data(), sizeof(T) * v.size());
    return p;
}

ep_design *finish(Builder &b, int32_t n_mov, int32_t n_fix, float xl, float yl, float xh, float yh,
                  int32_t mesh_side)
{
    ep_design *d = new (std::nothrow) ep_design();
    if (!d) return nullptr;
    ep_netlist &nl = d->nl;
    nl.n_mov = n_mov;
    nl.n_fix = n_fix;
    nl.w = dup(b.w);
    nl.h = dup(b.h);
    nl.x = dup(b.x);
    nl.y = dup(b.y);
    nl.n_nets = (int32_t)b.net_start.size() - 1;
    nl.n_pins = (int32_t)b.pin_cell.size();
    nl.net_start = dup(b.net_start);
    nl.pin_cell = dup(b.pin_cell);
    nl.pin_dx = dup(b.pin_dx);
    nl.pin_dy = dup(b.pin_dy);
    nl.xl = xl;
    nl.yl = yl;
    nl.xh = xh;
    nl.yh = yh;
    d->mesh_side = mesh_side;
    return d;
}

// 24x24 mesh with perimeter pads (the tiny CPU-runnable case).
ep_design *gen_mesh(const ep_gen_params *p)
{
    const int S = 24;
    const float die = (float)(S + 2);
    Rng rng(p->seed);
    Builder b;
    // movable cells, lattice order (i = column, j = row)
    for (int i = 0; i < S; ++i)
        for (int j = 0; j < S; ++j) {
            b.w.push_back(1.0f);
            b.h.push_back(1.0f);
            b.x.push_back((float)(rng.uniform() * (die - 1.0)));
            b.y.push_back((float)(rng.uniform() * (die - 1.0)));
        }
    auto cell = [&](int i, int j) { return i * S + j; };
    // pads: left, right, bottom, top; 24 each
    const int n_mov = S * S;
    int pad_left = n_mov, pad_right = n_mov + S, pad_bot = n_mov + 2 * S, pad_top = n_mov + 3 * S;
    for (int k = 0; k < S; ++k) { b.w.push_back(1); b.h.push_back(1); b.x.push_back(0); b.y.push_back((float)(k + 1)); }
    for (int k = 0; k < S; ++k) { b.w.push_back(1); b.h.push_back(1); b.x.push_back(die - 1); b.y.push_back((float)(k + 1)); }
    for (int k = 0; k < S; ++k) { b.w.push_back(1); b.h.push_back(1); b.x.push_back((float)(k + 1)); b.y.push_back(0); }
    for (int k = 0; k < S; ++k) { b.w.push_back(1); b.h.push_back(1); b.x.push_back((float)(k + 1)); b.y.push_back(die - 1); }
    // horizontal then vertical 4-neighbour nets
    for (int j = 0; j < S; ++j)
        for (int i = 0; i + 1 < S; ++i) { b.add_pin(cell(i, j), 0, 0); b.add_pin(cell(i + 1, j), 0, 0); b.end_net(); }
    for (int i = 0; i < S; ++i)
        for (int j = 0; j + 1 < S; ++j) { b.add_pin(cell(i, j), 0, 0); b.add_pin(cell(i, j + 1), 0, 0); b.end_net(); }
    // edge cell -> adjacent pad
    for (int k = 0; k < S; ++k) { b.add_pin(cell(0, k), 0, 0); b.add_pin(pad_left + k, 0, 0); b.end_net(); }
    for (int k = 0; k < S; ++k) { b.add_pin(cell(S - 1, k), 0, 0); b.add_pin(pad_right + k, 0, 0); b.end_net(); }
    for (int k = 0; k < S; ++k) { b.add_pin(cell(k, 0), 0, 0); b.add_pin(pad_bot + k, 0, 0); b.end_net(); }
    for (int k = 0; k < S; ++k) { b.add_pin(cell(k, S - 1), 0, 0); b.add_pin(pad_top + k, 0, 0); b.end_net(); }
    return finish(b, n_mov, 4 * S, 0.0f, 0.0f, die, die, S);
}

int32_t draw_degree(Rng &rng)
{
    double u = rng.uniform();
    if (u < 0.60) return 2;
    if (u < 0.78) return 3;
    if (u < 0.98) return (int32_t)rng.range(4, 10);
    return (int32_t)rng.range(11, 64);
}

ep_design *gen_stdcell(const ep_gen_params *p)
{
    Rng rng(p->seed);
    const int32_t n = p->n_movable;
    if (n < 2) return nullptr;
    Builder b;
    // movable cells in 1-D locality order: order position k -> stored index slot[k]
    std::vector<int32_t> slot(n);
    for (int32_t k = 0; k < n; ++k) slot[k] = k;
    if (p->shuffle_storage)
        for (int32_t k = n - 1; k > 0; --k) std::swap(slot[k], slot[(int32_t)rng.range(0, k)]);
    b.w.assign(n, 0.0f);
    b.h.assign(n, 1.0f);
    double mov_area = 0.0;
    for (int32_t k = 0; k < n; ++k) {
        float wk = (float)rng.range(p->site_min, p->site_max);
        b.w[slot[k]] = wk;
        mov_area += wk;
    }
    // die: movable area / utilisation is the free area; macros take macro_area_frac
    double die_area = mov_area / p->utilisation / (1.0 - p->macro_area_frac);
    float W = (float)std::ceil(std::sqrt(die_area));
    float H = W;
    // fixed square macros, uniform non-overlapping (rejection sampling)
    std::vector<float> mx, my, ms;
    double target = p->macro_area_frac * (double)W * (double)H, placed = 0.0;
    int guard = 0;
    while (placed < target && guard < 200000) {
        ++guard;
        float s = (float)rng.range(p->macro_min, p->macro_max);
        if (s >= W) continue;
        float px = (float)std::floor(rng.uniform() * (W - s));
        float py = (float)std::floor(rng.uniform() * (H - s));
        bool ok = true;
        for (size_t k = 0; k < mx.size() && ok; ++k)
            ok = (px + s <= mx[k] || mx[k] + ms[k] <= px || py + s <= my[k] || my[k] + ms[k] <= py);
        if (!ok) continue;
        mx.push_back(px);
        my.push_back(py);
        ms.push_back(s);
        placed += (double)s * s;
    }
    // initial movable positions: centre + N(0, (sigma W)^2), clamped into the die
    b.x.assign(n, 0.0f);
    b.y.assign(n, 0.0f);
    for (int32_t k = 0; k < n; ++k) {
        int32_t c = slot[k];
        double cx = 0.5 * W + rng.normal() * p->init_sigma_frac * W;
        double cy = 0.5 * H + rng.normal() * p->init_sigma_frac * W;
        double lx = std::min(std::max(cx - 0.5 * b.w[c], 0.0), (double)W - b.w[c]);
        double ly = std::min(std::max(cy - 0.5 * b.h[c], 0.0), (double)H - b.h[c]);
        b.x[c] = (float)lx;
        b.y[c] = (float)ly;
    }
    for (size_t k = 0; k < mx.size(); ++k) {
        b.w.push_back(ms[k]);
        b.h.push_back(ms[k]);
        b.x.push_back(mx[k]);
        b.y.push_back(my[k]);
    }
    // nets by locality in the 1-D order
    int64_t n_nets = (int64_t)std::llround(p->nets_per_cell * n);
    std::vector<int32_t> pick;
    for (int64_t e = 0; e < n_nets; ++e) {
        int32_t deg = std::min<int32_t>(draw_degree(rng), n);
        int32_t centre = (int32_t)rng.range(0, n - 1);
        pick.clear();
        int tries = 0;
        while ((int32_t)pick.size() < deg && tries < 64 * deg) {
            ++tries;
            int64_t k = centre + std::llround(rng.normal() * p->locality_sigma);
            if (k < 0 || k >= n) continue;
            int32_t c = slot[(int32_t)k];
            if (std::find(pick.begin(), pick.end(), c) != pick.end()) continue;
            pick.push_back(c);
        }
        if (pick.size() < 2) continue;
        for (int32_t c : pick)
            b.add_pin(c, (float)(rng.uniform() * b.w[c]), (float)(rng.uniform() * b.h[c]));
        b.end_net();
    }
    return finish(b, n, (int32_t)mx.size(), 0.0f, 0.0f, W, H, 0);
}

}  // namespace

extern "C" void ep_gen_default_params(ep_gen_params *p, int32_t kind)
{
    std::memset(p, 0, sizeof(*p));
    p->kind = kind;
    p->n_movable = 200000;
    p->utilisation = 0.7;
    p->macro_area_frac = 0.02;
    p->macro_min = 20;
    p->macro_max = 80;
    p->site_min = 1;
    p->site_max = 8;
    p->nets_per_cell = 1.05;
    p->locality_sigma = 200.0;
    p->init_sigma_frac = 0.01;
    p->shuffle_storage = 0;
    p->seed = 1;
}

extern "C" ep_status ep_gen_create(const ep_gen_params *p, ep_design **out)
{
    if (!p || !out) return ep::fail(EP_EINVAL, "ep_gen_create: null argument");
    *out = nullptr;
    ep_design *d = nullptr;
    if (p->kind == EP_GEN_MESH) {
        d = gen_mesh(p);
    } else if (p->kind == EP_GEN_STDCELL) {
        if (p->n_movable < 2 || p->utilisation <= 0 || p->utilisation > 1 || p->macro_area_frac < 0 ||
            p->macro_area_frac >= 0.9 || p->site_min < 1 || p->site_max < p->site_min ||
            p->macro_min < 1 || p->macro_max < p->macro_min || p->locality_sigma <= 0)
            return ep::fail(EP_EINVAL, "ep_gen_create: invalid stdcell parameters");
        d = gen_stdcell(p);
    } else {
        return ep::fail(EP_EINVAL, "ep_gen_create: unknown kind");
    }
    if (!d) return ep::fail(EP_ENOMEM, "ep_gen_create: allocation failed");
    *out = d;
    return EP_OK;
}

extern "C" void ep_gen_free(ep_design *d)
{
    if (!d) return;
    delete[] d->nl.w;
    delete[] d->nl.h;
    delete[] d->nl.x;
    delete[] d->nl.y;
    delete[] d->nl.net_start;
    delete[] d->nl.pin_cell;
    delete[] d->nl.pin_dx;
    delete[] d->nl.pin_dy;
    delete d;
}
