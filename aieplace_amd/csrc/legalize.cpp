// legalize.cpp -- Abacus row legalisation on the host (include/eplace_legal.h).
//
// Coordinates are in site units from the die's left edge (x) and row indices
// from its bottom (y).  Each row is cut into free segments by the fixed cells.
// A segment keeps its cells in placement order as a stack of clusters; a
// cluster (e, q, w) has the optimal position q / e clamped into the segment,
// and clusters that overlap merge (Spindler et al., ISPD 2008, Alg. 2-3).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <limits>
#include <numeric>
#include <vector>

#include "../../include/eplace_legal.h"
#include "ep_common.h"

using ep::fail;

namespace {

struct Cluster {
    double e, q, w, x;  // weight, sum of e*(x_i - offset_i), width, position
    int first, last;    // cells seg.cells[first, last)
};

struct Segment {
    int lo, hi;  // free sites [lo, hi)
    int row;
    int used = 0;
    std::vector<Cluster> cl;
    std::vector<int> cells;
};

double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Position the new cell (global x xi, width wi) would get at the end of seg.
double trial(const Segment &s, double xi, int wi)
{
    double e = 1.0, q = xi, w = wi;
    double xc = clampd(q / e, s.lo, s.hi - w);
    for (int k = (int)s.cl.size() - 1; k >= 0 && s.cl[k].x + s.cl[k].w > xc; --k) {
        const Cluster &c = s.cl[k];
        q = c.q + q - e * c.w;
        e += c.e;
        w += c.w;
        xc = clampd(q / e, s.lo, s.hi - w);
    }
    return xc + w - wi;
}

void commit(Segment &s, int cell, double xi, int wi)
{
    Cluster n{1.0, xi, (double)wi, 0.0, (int)s.cells.size(), (int)s.cells.size() + 1};
    n.x = clampd(n.q / n.e, s.lo, s.hi - n.w);
    s.cells.push_back(cell);
    s.used += wi;
    s.cl.push_back(n);
    while (s.cl.size() >= 2) {
        Cluster &p = s.cl[s.cl.size() - 2];
        const Cluster c = s.cl.back();
        if (p.x + p.w <= c.x) break;
        p.q += c.q - c.e * p.w;
        p.e += c.e;
        p.w += c.w;
        p.last = c.last;
        p.x = clampd(p.q / p.e, s.lo, s.hi - p.w);
        s.cl.pop_back();
    }
}

}  // namespace

extern "C" ep_status ep_legalize(const ep_netlist *nl, const float *x, const float *y, float row_height,
                                 float site_width, int32_t snap_to_sites, float *x_out, float *y_out,
                                 ep_legal_info *info)
{
    const auto t0 = std::chrono::steady_clock::now();
    if (!nl || !x || !y || !x_out || !y_out) return fail(EP_EINVAL, "ep_legalize: null argument");
    if (!(row_height > 0.f) || !(site_width > 0.f)) return fail(EP_EINVAL, "ep_legalize: row height and site width must be > 0");
    const int n_mov = nl->n_mov, n_all = nl->n_mov + nl->n_fix;
    if (n_mov < 0 || nl->n_fix < 0 || (n_all > 0 && (!nl->w || !nl->h)) || (nl->n_fix > 0 && (!nl->x || !nl->y)))
        return fail(EP_EINVAL, "ep_legalize: bad netlist");
    const double xl = nl->xl, yl = nl->yl, sw = site_width, rh = row_height;
    const int n_sites = (int)std::floor((nl->xh - xl) / sw + 1e-9);
    const int n_rows = (int)std::floor((nl->yh - yl) / rh + 1e-9);
    if (n_sites < 1 || n_rows < 1) return fail(EP_EINVAL, "ep_legalize: the die holds no row or site");
    std::vector<int> wsite(n_mov);
    for (int k = 0; k < n_mov; ++k) {
        if (!std::isfinite(x[k]) || !std::isfinite(y[k])) return fail(EP_ENAN, "ep_legalize: cell %d has a non-finite position", k);
        if (nl->h[k] > rh * (1.0 + 1e-6))
            return fail(EP_ENOTSUP, "ep_legalize: movable cell %d is %g high, more than one row (%g)", k, nl->h[k], rh);
        wsite[k] = std::max(1, (int)std::ceil(nl->w[k] / sw - 1e-6));
    }
    // free segments per row: sites minus the fixed cells' spans
    std::vector<std::vector<std::pair<int, int>>> blocked(n_rows);
    for (int k = n_mov; k < n_all; ++k) {
        const double fx = nl->x[k], fy = nl->y[k], fw = nl->w[k], fh = nl->h[k];
        const int r0 = std::max(0, (int)std::floor((fy - yl) / rh + 1e-9));
        const int r1 = std::min(n_rows, (int)std::ceil((fy + fh - yl) / rh - 1e-9));
        const int s0 = std::max(0, (int)std::floor((fx - xl) / sw + 1e-9));
        const int s1 = std::min(n_sites, (int)std::ceil((fx + fw - xl) / sw - 1e-9));
        if (s1 <= s0) continue;
        for (int r = r0; r < r1; ++r) blocked[r].push_back({s0, s1});
    }
    std::vector<Segment> segs;
    std::vector<std::vector<int>> row_segs(n_rows);
    for (int r = 0; r < n_rows; ++r) {
        auto &b = blocked[r];
        std::sort(b.begin(), b.end());
        int at = 0;
        auto add = [&](int lo, int hi) {
            if (hi <= lo) return;
            Segment s;
            s.lo = lo; s.hi = hi; s.row = r;
            row_segs[r].push_back((int)segs.size());
            segs.push_back(std::move(s));
        };
        for (const auto &iv : b) {
            add(at, iv.first);
            at = std::max(at, iv.second);
        }
        add(at, n_sites);
    }
    // cells in order of global x
    std::vector<int> order(n_mov);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return x[a] < x[b]; });
    const double inf = std::numeric_limits<double>::infinity();
    for (int cell : order) {
        const double xi = (x[cell] - xl) / sw;       // sites
        const double yr = (y[cell] - yl) / rh;       // rows
        const int wi = wsite[cell];
        const int r0 = std::min(n_rows - 1, std::max(0, (int)std::lround(yr)));
        double best = inf;
        int best_seg = -1;
        for (int d = 0; d < n_rows; ++d) {
            double dmin = inf;
            for (int side = 0; side < (d ? 2 : 1); ++side) {
                const int r = side ? r0 + d : r0 - d;
                if (r < 0 || r >= n_rows) continue;
                const double dy = (r - yr) * rh, cy = dy * dy;
                dmin = std::min(dmin, cy);
                if (cy >= best) continue;
                const auto &rs = row_segs[r];
                // the segment at (or after) xi, then outwards both ways to the first with room
                int j = (int)(std::upper_bound(rs.begin(), rs.end(), xi,
                                               [&](double v, int s) { return v < segs[s].lo; }) - rs.begin()) - 1;
                if (j < 0) j = 0;
                auto consider = [&](int sj) {
                    Segment &s = segs[rs[sj]];
                    if (s.hi - s.lo - s.used < wi) return false;
                    const double dx = (trial(s, xi, wi) - xi) * sw;
                    const double c = dx * dx + cy;
                    if (c < best) {
                        best = c;
                        best_seg = rs[sj];
                    }
                    return true;
                };
                for (int sj = j; sj >= 0; --sj)
                    if (consider(sj)) break;
                for (int sj = j + 1; sj < (int)rs.size(); ++sj)
                    if (consider(sj)) break;
            }
            if (dmin > best) break;
        }
        if (best_seg < 0) return fail(EP_EINVAL, "ep_legalize: no row has room for cell %d (%d sites)", cell, wi);
        commit(segs[best_seg], cell, xi, wi);
    }
    // positions
    double tot = 0, mx = 0, sq = 0;
    std::vector<float> ox(n_mov), oy(n_mov);
    for (const Segment &s : segs) {
        for (const Cluster &c : s.cl) {
            double pos = snap_to_sites ? std::round(c.x) : c.x;
            for (int k = c.first; k < c.last; ++k) {
                const int cell = s.cells[k];
                ox[cell] = (float)(xl + pos * sw);
                oy[cell] = (float)(yl + s.row * rh);
                pos += wsite[cell];
            }
        }
    }
    for (int k = 0; k < n_mov; ++k) {
        const double dx = (double)ox[k] - x[k], dy = (double)oy[k] - y[k];
        const double d2 = dx * dx + dy * dy;
        sq += d2;
        tot += std::sqrt(d2);
        mx = std::max(mx, std::sqrt(d2));
    }
    std::copy(ox.begin(), ox.end(), x_out);
    std::copy(oy.begin(), oy.end(), y_out);
    if (info) {
        info->total_disp = tot;
        info->max_disp = mx;
        info->sum_sq_disp = sq;
        info->rows = n_rows;
        info->segments = (int32_t)segs.size();
        info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return EP_OK;
}
