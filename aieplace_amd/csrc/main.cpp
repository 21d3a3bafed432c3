// main.cpp -- command-line placer over the C ABI, the counterpart of the
// reference's AIEplace.exe (vck5000/host/src/main.cpp:1-25, Placer::run and
// printFinalResults in AIEplace.cpp:44-76, 1076-1124):
//
//   eplace_run <design_dir> [--bins N] [--density D] [--precision fp32|bf16x3|fp16x3]
//              [--max-iters N] [--legalize] [--out DIR]
//
// Reads every .lef of <design_dir> and its floorplan.def, places the design
// to overflow 0.10 on the GPU (device-controlled Nesterov run), optionally
// legalises it on rows and sites, and writes <out>/<design>.def (default
// out: results/<design>).  Exit status 0 on success, 1 on an error (with the
// library's message), 2 on bad arguments.
#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <numeric>
#include <string>
#include <vector>

#include "eplace.h"
#include "eplace_io.h"
#include "eplace_legal.h"

namespace {

int usage()
{
    std::fprintf(stderr,
                 "usage: eplace_run <design_dir> [--bins N] [--density D] [--precision fp32|bf16x3|fp16x3]\n"
                 "                  [--max-iters N] [--legalize] [--out DIR]\n");
    return 2;
}

bool check(ep_status st, const char *what)
{
    if (st == EP_OK) return true;
    std::fprintf(stderr, "eplace_run: %s failed (status %d): %s\n", what, (int)st, ep_last_error());
    return false;
}

void make_dirs(const std::string &path)
{
    std::string at;
    for (size_t i = 0; i <= path.size(); ++i) {
        if (i == path.size() || path[i] == '/') {
            if (!at.empty()) mkdir(at.c_str(), 0755);
        }
        if (i < path.size()) at.push_back(path[i]);
    }
}

// Bins per side: about one movable cell per bin, a multiple of 32 in [32, 1024].
int default_bins(int n_mov)
{
    int b = (int)std::lround(std::sqrt((double)std::max(n_mov, 1)));
    b = (b + 31) / 32 * 32;
    return std::min(1024, std::max(32, b));
}

// Row height: the movable cells' most common height; site: the largest width
// (on a 1e-3 grid) that divides every movable width.
void default_grid(const ep_netlist &nl, float &row, float &site)
{
    std::map<long, int> heights;
    long g = 0;
    for (int k = 0; k < nl.n_mov; ++k) {
        heights[std::lround(nl.h[k] * 1e6)]++;
        g = std::gcd(g, std::lround(nl.w[k] * 1e3));
    }
    long best = 0;
    int cnt = -1;
    for (const auto &h : heights)
        if (h.second > cnt) {
            cnt = h.second;
            best = h.first;
        }
    row = (float)(best / 1e6);
    site = g > 0 ? (float)(g / 1e3) : 1.0f;
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc < 2) return usage();
    const char *dir = argv[1];
    int bins = 0, max_iters = -1;
    float density = 0.9f;
    int precision = EP_PREC_FP32;
    bool legal = false;
    std::string out;
    for (int i = 2; i < argc; ++i) {
        const std::string a = argv[i];
        auto value = [&]() -> const char * { return i + 1 < argc ? argv[++i] : nullptr; };
        if (a == "--bins") {
            const char *v = value();
            if (!v) return usage();
            bins = std::atoi(v);
        } else if (a == "--density") {
            const char *v = value();
            if (!v) return usage();
            density = (float)std::atof(v);
        } else if (a == "--max-iters") {
            const char *v = value();
            if (!v) return usage();
            max_iters = std::atoi(v);
        } else if (a == "--precision") {
            const char *v = value();
            if (!v) return usage();
            const std::string p = v;
            if (p == "fp32") precision = EP_PREC_FP32;
            else if (p == "bf16x3") precision = EP_PREC_BF16X3;
            else if (p == "fp16x3") precision = EP_PREC_FP16X3;
            else return usage();
        } else if (a == "--legalize") {
            legal = true;
        } else if (a == "--out") {
            const char *v = value();
            if (!v) return usage();
            out = v;
        } else {
            return usage();
        }
    }

    ep_lefdef *design = nullptr;
    if (!check(ep_lefdef_read_dir(dir, &design), "reading the design")) return 1;
    ep_lefdef_info info;
    char name[256];
    ep_lefdef_get_info(design, &info, name);
    ep_netlist nl;
    ep_lefdef_netlist(design, &nl);
    std::printf("design %s: %d movable, %d fixed cells, %d nets, %d pins; die %.3f x %.3f um\n", name, nl.n_mov,
                nl.n_fix, nl.n_nets, nl.n_pins, info.die[2] - info.die[0], info.die[3] - info.die[1]);

    ep_params p;
    ep_default_params(&p);
    p.bins_x = p.bins_y = bins > 0 ? bins : default_bins(nl.n_mov);
    p.target_density = density;
    p.precision = precision;
    if (max_iters > 0) p.max_iters = max_iters;
    ep_ctx *ctx = nullptr;
    if (!check(ep_create(&nl, &p, &ctx), "ep_create")) {
        ep_lefdef_free(design);
        return 1;
    }
    int32_t n_mov = 0, n_fill = 0, n_fix = 0, n_nets = 0, n_pins = 0;
    ep_get_sizes(ctx, &n_mov, &n_fill, &n_fix, &n_nets, &n_pins);
    const int cap = p.max_iters > 0 ? p.max_iters : 1;
    std::vector<ep_step_info> log((size_t)cap);
    ep_run_info ri;
    bool ok = check(ep_run(ctx, &ri, log.data(), cap), "ep_run");
    std::vector<float> x((size_t)(n_mov + n_fill)), y((size_t)(n_mov + n_fill));
    if (ok) ok = check(ep_get_positions(ctx, x.data(), y.data()), "ep_get_positions");
    ep_destroy(ctx);
    if (ok) {
        for (int k = 0; k < std::min(ri.iterations, cap); ++k)  // printIterationResults, every tenth
            if (k % 10 == 0 || k + 1 == ri.iterations)
                std::printf("iteration %5d  HPWL %.6e  overflow %.4f\n", log[k].iteration, log[k].hpwl,
                            log[k].overflow);
        std::printf("bins %d x %d, %d iterations, %lld gradient evaluations, %.4f s on the GPU\n", p.bins_x,
                    p.bins_y, ri.iterations, (long long)ri.evals, ri.seconds);
        std::printf("final HPWL %.6e  overflow %.4f\n", ri.hpwl, ri.overflow);
    }
    if (ok && legal) {
        float row = 0, site = 0;
        default_grid(nl, row, site);
        std::vector<float> lx((size_t)n_mov), ly((size_t)n_mov);
        ep_legal_info li;
        ok = check(ep_legalize(&nl, x.data(), y.data(), row, site, 1, lx.data(), ly.data(), &li), "ep_legalize");
        if (ok) {
            std::copy(lx.begin(), lx.end(), x.begin());
            std::copy(ly.begin(), ly.end(), y.begin());
            std::printf("legalised on %d rows of %g (sites %g): mean displacement %.4f, max %.4f, %.1f ms\n", li.rows,
                        row, site, n_mov ? li.total_disp / n_mov : 0.0, li.max_disp, li.seconds * 1e3);
        }
    }
    if (ok) {
        if (out.empty()) out = std::string("results/") + name;
        make_dirs(out);
        ok = check(ep_lefdef_write_def(design, x.data(), y.data(), out.c_str()), "writing the DEF");
        if (ok) std::printf("wrote %s/%s.def\n", out.c_str(), name);
    }
    ep_lefdef_free(design);
    return ok ? 0 : 1;
}
