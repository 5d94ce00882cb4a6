// noise.hip -- camera noise on the GPU: ApplyNoiseChain (Sim/SMLMNoise.cpp:142-216) for a batch of frames.
//
// One lane per pixel, grid-stride over frames x pixels. A pixel's draws depend only on (seed, frame, pixel), so
// lanes are independent; the per-pixel work (Knuth loop, Box-Muller, Marsaglia-Tsang in FP64) is VALU-bound and
// its lane divergence is the Poisson/Gamma draw counts. Reads 4 B and writes 2 B per pixel (plus maps).
#include "iscg_internal.h"

namespace iscg {

template <bool EMCCD>
__global__ __launch_bounds__(256) void noise_kernel(const float* __restrict__ photons, uint16_t* __restrict__ out,
                                                     uint32_t npx, uint32_t nFrames, const uint32_t* __restrict__ frames,
                                                     const float* __restrict__ offset, const float* __restrict__ gain,
                                                     const float* __restrict__ readN, CamParams cp, uint32_t seed,
                                                     uint32_t pixelBase)
{
   const size_t total = (size_t)npx * nFrames;
   for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x)
   {
      const uint32_t f = (uint32_t)(idx / npx);
      const uint32_t i = (uint32_t)(idx - (size_t)f * npx);
      out[idx] = noise_pixel<EMCCD>(cp, photons[idx], seed, frames[f], i + pixelBase, offset != nullptr,
                                    offset ? offset[i] : 0.0f, gain != nullptr, gain ? gain[i] : 0.0f,
                                    readN != nullptr, readN ? readN[i] : 0.0f);
   }
}

// Launches the noise kernel on device arrays (shared with the splat path when noise runs as its own pass).
bool launch_noise(iscg_context* h, const float* photons, uint16_t* out, uint32_t npx, uint32_t nFrames,
                  const uint32_t* frames, const float* offset, const float* gain, const float* readN,
                  const CamParams& cp, uint32_t seed, uint32_t pixelBase)
{
   const size_t total = (size_t)npx * nFrames;
   if (total == 0)
      return true;
   const unsigned block = 256, grid = grid_for(total, block, 16);
   ProfScope ps(h, "noise");
   if (cp.emccd)
      hipLaunchKernelGGL(noise_kernel<true>, dim3(grid), dim3(block), 0, h->stream, photons, out, npx, nFrames, frames,
                         offset, gain, readN, cp, seed, pixelBase);
   else
      hipLaunchKernelGGL(noise_kernel<false>, dim3(grid), dim3(block), 0, h->stream, photons, out, npx, nFrames,
                         frames, offset, gain, readN, cp, seed, pixelBase);
   return ok(h, hipGetLastError(), "noise kernel launch");
}

}  // namespace iscg

using namespace iscg;

extern "C" int32_t iscg_noise(iscg_context* h, const float* photons, uint32_t width, uint32_t height,
                              const iscg_camera* cam, const float* offset, const float* gain, const float* readNoise,
                              uint32_t seed, const uint32_t* frames, uint32_t nFrames, uint32_t pixelBase,
                              uint16_t* out)
{
   if (begin(h))
      return -1;
   if (!cam || !frames || !out || (!photons && nFrames))
      return fail(h, "iscg_noise: null argument");
   const size_t npx = (size_t)width * height;
   if (npx > 0xFFFFFFFFull)
      return fail(h, "iscg_noise: frame larger than 2^32 pixels");
   if (npx == 0 || nFrames == 0)
      return 0;
   const size_t nAll = npx * nFrames;
   const float* dPh = (const float*)stage_in(h, 0, photons, nAll * sizeof(float));
   const uint32_t* dFr = (const uint32_t*)stage_in(h, 1, frames, nFrames * sizeof(uint32_t));
   const float* dOff = (const float*)stage_in(h, 2, offset, offset ? npx * sizeof(float) : 0);
   const float* dGain = (const float*)stage_in(h, 3, gain, gain ? npx * sizeof(float) : 0);
   const float* dRead = (const float*)stage_in(h, 4, readNoise, readNoise ? npx * sizeof(float) : 0);
   uint16_t* dOut = (uint16_t*)stage_out(h, 5, out, nAll * sizeof(uint16_t));
   if (!dPh || !dFr || (offset && !dOff) || (gain && !dGain) || (readNoise && !dRead) || !dOut)
      return -1;
   if (!launch_noise(h, dPh, dOut, (uint32_t)npx, nFrames, dFr, dOff, dGain, dRead, cam_params(cam), seed,
                     pixelBase))
      return -1;
   if (!finish_out(h, out, dOut, nAll * sizeof(uint16_t)))
      return -1;
   return 0;
}
