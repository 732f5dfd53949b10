// ref_poisson.cpp -- C entry point over the reference's own Poisson solver.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile together with the
// reference's unmodified sources src/fft.cpp, src/fftsg.cpp and
// src/fftsg2d.cpp (read where they lie under the reference tree) into
// oracle/_ref/libref_poisson.so.  It lets the oracle and the tests run
// replace::FFT exactly as NesterovBase::updateDensityForceBin does
// (src/nesterovBase.cpp:1481-1506): updateDensity per bin, doFFT, then
// getElectroPhi / getElectroForce per bin.
#include <cstdint>
#include <memory>

#include "fft.h"

namespace {
struct Cache {
  int nx = 0, ny = 0, bsx = 0, bsy = 0;
  std::unique_ptr<replace::FFT> fft;
};
Cache g_cache;
}  // namespace

extern "C" void ref_poisson(void* /*user*/, int nx, int ny, int bsx, int bsy, const float* rho,
                            float* phi, float* ex, float* ey) {
  if (!g_cache.fft || g_cache.nx != nx || g_cache.ny != ny || g_cache.bsx != bsx ||
      g_cache.bsy != bsy) {
    g_cache.fft.reset(new replace::FFT(nx, ny, bsx, bsy));
    g_cache.nx = nx;
    g_cache.ny = ny;
    g_cache.bsx = bsx;
    g_cache.bsy = bsy;
  }
  replace::FFT& f = *g_cache.fft;
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) f.updateDensity(i, j, rho[(size_t)j * nx + i]);
  f.doFFT();
  for (int j = 0; j < ny; j++) {
    for (int i = 0; i < nx; i++) {
      size_t b = (size_t)j * nx + i;
      std::pair<float, float> e = f.getElectroForce(i, j);
      if (ex) ex[b] = e.first;
      if (ey) ey[b] = e.second;
      if (phi) phi[b] = f.getElectroPhi(i, j);
    }
  }
}
