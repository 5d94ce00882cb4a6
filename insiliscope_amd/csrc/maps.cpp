// maps.cpp -- the camera's static per-pixel maps and seeds (host code), and the kernel-geometry helpers the CLI
// applies to a PSF request before it reaches the PSF planes.
//
// The maps come from one std::mt19937_64 per seed, drawn in a fixed order: offset, gain, read noise, each map the
// whole sensor's pixels (rows outside a window still consume their draws), one Box-Muller normal per pixel from two
// std::uniform_real_distribution<double>(0, 1) draws. That is the reference's definition (NoiseSetup,
// Sim/ScopeMovie.cpp:973-996; Pixel*Map::Generate and GaussianRng, Sim/SMLMNoise.cpp:17-140): the values depend on
// the C++ library's uniform_real_distribution, so they are made here on the host with the same library rather than
// on the device. They are made once per seed; the noise kernels then read them as inputs.
#include "iscg.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <random>

namespace {

constexpr double kPi = 3.14159265358979323846;

// One N(mean, sd) draw: u1 clamped at 1e-300, r = sqrt(-2 log u1), times cos(2 pi u2).
double gaussian(std::mt19937_64& rng, double mean, double sd)
{
   std::uniform_real_distribution<double> unif(0.0, 1.0);
   double u1 = unif(rng);
   if (u1 < 1e-300)
      u1 = 1e-300;
   const double u2 = unif(rng);
   const double z0 = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
   return mean + sd * z0;
}

void skip_gaussians(std::mt19937_64& rng, size_t k)
{
   std::uniform_real_distribution<double> unif(0.0, 1.0);
   for (size_t i = 0; i < 2 * k; ++i)
      (void)unif(rng);
}

// The kept rows of a w x h sensor: [i0, i1) pixel indices.
struct Rows
{
   size_t i0, i1, n;
   Rows(unsigned w, unsigned h, unsigned keepY0, unsigned keepRows)
   {
      const unsigned rows = keepRows ? std::min(keepRows, h - std::min(h, keepY0)) : h;
      i0 = keepRows ? static_cast<size_t>(keepY0) * w : 0;
      i1 = i0 + static_cast<size_t>(rows) * w;
      n = static_cast<size_t>(w) * h;
   }
};

// value(g) for each kept pixel, g the pixel's normal; every pixel's draws consumed.
template <typename F>
void make_map(std::mt19937_64& rng, const Rows& R, float* out, F value)
{
   skip_gaussians(rng, R.i0);
   for (size_t i = 0; i < R.i1 - R.i0; ++i)
      out[i] = static_cast<float>(value(gaussian(rng, 0.0, 1.0)));
   skip_gaussians(rng, R.n - R.i1);
}

}  // namespace

extern "C" {

int32_t iscg_noise_setup(int64_t seed, uint32_t width, uint32_t height, uint32_t sensorHeight, uint32_t rowY0,
                         double offsetAdu, double offsetStdAdu, double gainPhotonsPerAdu, double gainStdFraction,
                         double readNoiseElectrons, double readStdFraction, float* offset, float* gain,
                         float* readNoise, uint32_t* noiseSeed, uint32_t* pixelBase)
{
   if (width == 0 || height == 0 || !offset || !gain || !readNoise)
      return -1;
   const bool window = sensorHeight != 0;
   const unsigned H = window ? sensorHeight : height;
   if (window && rowY0 + height > sensorHeight)
      return -1;
   const Rows R(width, H, window ? rowY0 : 0, window ? height : 0);
   std::mt19937_64 rng(static_cast<uint64_t>(seed));
   make_map(rng, R, offset, [&](double g) { return offsetAdu + offsetStdAdu * g; });
   const double floor = 0.04 * gainPhotonsPerAdu;
   make_map(rng, R, gain, [&](double g) {
      const double v = gainPhotonsPerAdu * (1.0 + gainStdFraction * g);
      return v < floor ? floor : v;
   });
   make_map(rng, R, readNoise, [&](double g) {
      const double v = readNoiseElectrons * (1.0 + readStdFraction * g);
      return v < 0.0 ? 0.0 : v;
   });
   if (noiseSeed)
      *noiseSeed = static_cast<uint32_t>(static_cast<uint64_t>(seed) ^ 0x9E3779B9ULL);
   if (pixelBase)
      *pixelBase = window ? rowY0 * width : 0u;
   return 0;
}

int32_t iscg_psf_kernel_half_width_px(double halfWidthNm, double pixelSizeNm, double wavelengthNm, double na)
{
   int requested = static_cast<int>(std::lround(halfWidthNm / pixelSizeNm));
   requested = std::max(requested, 1);
   const double naSafe = na > 0.0 ? na : 0.01;
   const double rayleighNm = 0.61 * wavelengthNm / naSafe;
   const int minHalf = std::min(std::max(static_cast<int>(std::ceil(3.0 * rayleighNm / pixelSizeNm)), 2), 48);
   return std::max(requested, minHalf);
}

int32_t iscg_capped_psf_oversampling(int32_t oversampling, double pixelSizeNm, double maxSampleNm)
{
   oversampling = std::max(1, oversampling);
   if (!(maxSampleNm > 0.0) || !(pixelSizeNm > 0.0))
      return oversampling;
   // a pixel of exactly n times the cap takes n samples, not n + 1
   const int need = static_cast<int>(std::ceil(pixelSizeNm / maxSampleNm - 1e-9));
   return std::min(oversampling, std::max(1, need));
}

}  // extern "C"
