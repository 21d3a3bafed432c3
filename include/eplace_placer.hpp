// eplace_placer.hpp -- C++ host class over the C ABI with the method names of
// the reference's Placer (vck5000/host/src/include/AIEplace.h:21-99;
// AIEplace.cpp:9-76 performIteration / run, 993-1046 nudgeAllNodes,
// 1048-1124 printIterationResults / getTime / getInterval / printFinalResults).
//
// The XRT graph drivers of the reference (PartialsGraphDriver,
// DensityGraphDriver, GraphDriver.h:16-65) become one device context: the
// netlist is copied to HBM once, and each method is one or a few C-ABI calls.
// Errors throw aieplace::Error (status code + the library's message) where the
// reference logs and exits.  Header-only; link with -leplace.
#pragma once

#include <sys/time.h>

#include <algorithm>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "eplace.h"

namespace aieplace {

class Error : public std::runtime_error {
public:
    Error(ep_status st, const std::string &what) : std::runtime_error(what), status(st) {}
    ep_status status;
};

inline void check(ep_status st, const char *what)
{
    if (st != EP_OK) throw Error(st, std::string(what) + " failed: " + ep_last_error());
}

class Placer {
public:
    // Placer(std::string xclbin_file) (AIEplace.cpp:87-138): the netlist and the
    // parameters take the place of the xclbin and the json configuration.
    Placer(const ep_netlist &nl, const ep_params &params) : params_(params)
    {
        check(ep_create(&nl, &params_, &ctx_), "ep_create");
        const ep_status st = ep_get_sizes(ctx_, &n_mov_, &n_fill_, &n_fix_, &n_nets_, &n_pins_);
        if (st != EP_OK) {
            const Error e(st, std::string("ep_get_sizes failed: ") + ep_last_error());
            ep_destroy(ctx_);
            throw e;
        }
    }
    ~Placer()
    {
        if (ctx_) ep_destroy(ctx_);
    }
    Placer(const Placer &) = delete;
    Placer &operator=(const Placer &) = delete;

    static void printWelcomeBanner()
    {
        std::printf("AIEplace on MI355X: ePlace global placement (WA wirelength, bin density, "
                    "DCT Poisson field, Nesterov)\n");
    }

    // Pre-run preparation (AIEplace.cpp:187-243): the initial positions come with
    // the netlist; this sets gamma_0, lambda_0 and the first step length.
    void initializePlacement() { check(ep_init_optimizer(ctx_), "ep_init_optimizer"); }
    // Nothing is reset between iterations: every evaluation rebuilds its maps.
    void iterationReset() {}

    // The stages the reference offloads (computeAllPartials_{AIE,CPU},
    // computeOverlaps, computeElectricFields_{AIE,CPU}), one call each, at the
    // current positions.  performIteration does not need them: one Nesterov
    // iteration runs its own evaluations.
    double computeAllPartials(float gamma)  // returns the HPWL; the WA gradient stays on the device
    {
        double wa = 0, hpwl = 0;
        check(ep_wa_grad(ctx_, gamma, &wa, &hpwl), "ep_wa_grad");
        return hpwl;
    }
    double computeOverlaps()  // returns the overflow
    {
        double ovf = 0;
        check(ep_density(ctx_, &ovf), "ep_density");
        return ovf;
    }
    void computeElectricFields() { check(ep_field(ctx_, params_.precision), "ep_field"); }

    // One Nesterov iteration with backtracking, the lambda/gamma schedules and
    // the optional learned step (the reference's fixed-step nudgeAllNodes).
    const ep_step_info &nudgeAllNodes()
    {
        check(ep_step(ctx_, &last_), "ep_step");
        iteration = last_.iteration + 1;
        return last_;
    }
    void performIteration()
    {
        iterationReset();
        nudgeAllNodes();
        printIterationResults();
    }
    void printIterationResults() const
    {
        std::printf("Iteration %6d  HPWL %.6e  Overflow %.4f\n", last_.iteration, last_.hpwl, last_.overflow);
    }

    // Placer::run (AIEplace.cpp:44-76): iterations until overflow <= stop_overflow
    // (or max_iters), device-controlled when params.use_graph; every
    // `print_every`-th iteration is printed (0: none).
    ep_run_info run(int print_every = 0)
    {
        std::vector<ep_step_info> log(print_every > 0 ? (size_t)std::max(1, params_.max_iters) : 0);
        const long t0 = getTime();
        check(ep_run(ctx_, &run_, log.empty() ? nullptr : log.data(), (int32_t)log.size()), "ep_run");
        const long t1 = getTime();
        seconds_ = getInterval(t0, t1);
        iteration = run_.iterations;
        for (int k = 0; k < (int)log.size() && k < run_.iterations; ++k)
            if (k % print_every == 0 || k + 1 == run_.iterations) {
                last_ = log[k];
                printIterationResults();
            }
        return run_;
    }
    void printFinalResults() const
    {
        std::printf("AIEplace run statistics\n");
        std::printf("  Iterations            %d\n", run_.iterations);
        std::printf("  Gradient evaluations  %lld\n", (long long)run_.evals);
        std::printf("  GPU runtime           %.4f s (host wall %.4f s)\n", run_.seconds, seconds_);
        std::printf("  Final HPWL            %.6e\n", run_.hpwl);
        std::printf("  Final Overflow        %.4f\n", run_.overflow);
        if (run_.learned_proposals > 0)
            std::printf("  Learned steps         %d of %d proposals accepted\n", run_.learned_accepted,
                        run_.learned_proposals);
    }

    // Node positions (lower-left corners) of the movable cells, then the fillers.
    void getPositions(std::vector<float> &x, std::vector<float> &y) const
    {
        x.resize((size_t)(n_mov_ + n_fill_));
        y.resize(x.size());
        check(ep_get_positions(ctx_, x.data(), y.data()), "ep_get_positions");
    }
    void setPositions(const std::vector<float> &x, const std::vector<float> &y)
    {
        if (x.size() != (size_t)(n_mov_ + n_fill_) || y.size() != x.size())
            throw Error(EP_EINVAL, "setPositions: need movable + filler positions");
        check(ep_set_positions(ctx_, x.data(), y.data()), "ep_set_positions");
    }
    double hpwl() const
    {
        double h = 0;
        check(ep_hpwl(ctx_, &h), "ep_hpwl");
        return h;
    }

    long getTime() const
    {
        timeval tm;
        gettimeofday(&tm, nullptr);
        return tm.tv_sec * 1000000L + tm.tv_usec;
    }
    double getInterval(long start_time, long end_time) const { return (end_time - start_time) / 1.0e6; }

    int movableCells() const { return n_mov_; }
    int fillerCells() const { return n_fill_; }
    ep_ctx *context() const { return ctx_; }

    int iteration = 0;

private:
    ep_params params_;
    ep_ctx *ctx_ = nullptr;
    int32_t n_mov_ = 0, n_fill_ = 0, n_fix_ = 0, n_nets_ = 0, n_pins_ = 0;
    ep_step_info last_{};
    ep_run_info run_{};
    double seconds_ = 0;
};

}  // namespace aieplace
