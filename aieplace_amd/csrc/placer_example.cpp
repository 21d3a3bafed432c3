// placer_example.cpp -- the reference's host flow (vck5000/host/src/main.cpp:1-25:
// banner, Placer, iterations, run, printFinalResults) on the C++ Placer of
// include/eplace_placer.hpp, over a generated instance:
//
//   placer_example [mesh | stdcell N] [--bins B] [--graph 0|1] [--precision fp32|bf16x3]
//
// The text between the INTEGRATION EXAMPLE markers is the example INTEGRATION.md
// shows; the build compiles this file into aieplace_amd/placer_example.
// Exit status 0 on success, 1 on a library error, 2 on bad arguments.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "eplace.h"
#include "eplace_gen.h"
#include "eplace_placer.hpp"

int main(int argc, char **argv)
{
    std::string kind = argc > 1 ? argv[1] : "mesh";
    int n_movable = 0, bins = 0, graph = 1, precision = EP_PREC_FP32;
    int i = 2;
    if (kind == "stdcell") {
        if (argc < 3) return 2;
        n_movable = std::atoi(argv[2]);
        i = 3;
    } else if (kind != "mesh") {
        std::fprintf(stderr, "usage: placer_example [mesh | stdcell N] [--bins B] [--graph 0|1] "
                             "[--precision fp32|bf16x3]\n");
        return 2;
    }
    for (; i + 1 < argc; i += 2) {
        const std::string a = argv[i], v = argv[i + 1];
        if (a == "--bins") bins = std::atoi(v.c_str());
        else if (a == "--graph") graph = std::atoi(v.c_str());
        else if (a == "--precision") precision = v == "bf16x3" ? EP_PREC_BF16X3 : EP_PREC_FP32;
        else return 2;
    }
    ep_gen_params gp;
    ep_gen_default_params(&gp, kind == "mesh" ? EP_GEN_MESH : EP_GEN_STDCELL);
    gp.seed = 1;
    if (n_movable > 0) gp.n_movable = n_movable;
    ep_design *design = nullptr;
    if (ep_gen_create(&gp, &design) != EP_OK) {
        std::fprintf(stderr, "ep_gen_create failed: %s\n", ep_last_error());
        return 1;
    }
    const ep_netlist &netlist = design->nl;
    int status = 0;

    // ---- BEGIN INTEGRATION EXAMPLE
    aieplace::Placer::printWelcomeBanner();
    ep_params params;
    ep_default_params(&params);
    params.bins_x = params.bins_y = bins > 0 ? bins : (kind == "mesh" ? 32 : 512);
    if (kind == "mesh") {                  // 24x24 lattice: rho_t 1, fixed gamma of half a bin
        params.target_density = 1.0f;
        params.gamma_bins = 0.5f;
        params.gamma_schedule = 0;
    }
    params.precision = precision;
    params.use_graph = graph;              // 1: the Nesterov loop runs under device control
    try {
        aieplace::Placer placer(netlist, params);
        placer.initializePlacement();      // gamma_0, lambda_0, first step length
        for (int k = 0; k < 3; ++k)        // Placer::performIteration, one host-driven step each
            placer.performIteration();
        std::printf("overflow %.4f  HPWL %.6e after %d iterations\n", placer.computeOverlaps(),
                    placer.hpwl(), placer.iteration);
        placer.run(50);                    // Placer::run: on to overflow <= 0.10
        placer.printFinalResults();
        std::vector<float> x, y;
        placer.getPositions(x, y);         // movable cells, then fillers: back to the DataBase
        std::printf("positions of %d movable and %d filler cells\n", placer.movableCells(),
                    placer.fillerCells());
    } catch (const aieplace::Error &e) {
        std::fprintf(stderr, "placer_example: %s (status %d)\n", e.what(), (int)e.status);
        status = 1;
    }
    // ---- END INTEGRATION EXAMPLE

    ep_gen_free(design);
    return status;
}
