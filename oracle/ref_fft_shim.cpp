// ref_fft_shim.cpp -- C entry point around the reference's RePlAce field solver
// (python/AIEmath/fft_replace/fft.{h,cpp}: replace::FFT::doFFT).  This file is
// this project's own; the solver sources are compiled in place from the
// reference tree by oracle/Makefile (target `ref`) into oracle/_ref/.
// TEST INFRASTRUCTURE ONLY: used to pin the oracle's field convention.
#include <iostream>
#include <sstream>

#include "fft.h"

extern "C" int ref_fft_field(int bins_x, int bins_y, int bin_size_x, int bin_size_y,
                             const float *rho, float *phi, float *ex, float *ey)
{
    if (bins_x < 2 || bins_y < 2) return 1;
    if ((bins_x & (bins_x - 1)) || (bins_y & (bins_y - 1))) return 2;  // Ooura: powers of two
    replace::FFT fft(bins_x, bins_y, bin_size_x, bin_size_y);
    for (int i = 0; i < bins_x; ++i)
        for (int j = 0; j < bins_y; ++j) fft.updateDensity(i, j, rho[(size_t)i * bins_y + j]);
    // doFFT prints its intermediate arrays; keep them off our stdout.
    std::ostringstream sink;
    std::streambuf *old = std::cout.rdbuf(sink.rdbuf());
    fft.doFFT();
    std::cout.rdbuf(old);
    for (int i = 0; i < bins_x; ++i)
        for (int j = 0; j < bins_y; ++j) {
            size_t o = (size_t)i * bins_y + j;
            std::pair<float, float> e = fft.getElectroForce(i, j);
            if (phi) phi[o] = fft.getElectroPhi(i, j);
            ex[o] = e.first;
            ey[o] = e.second;
        }
    return 0;
}
