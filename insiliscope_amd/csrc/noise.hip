// noise.hip -- camera noise on the GPU: ApplyNoiseChain (Sim/SMLMNoise.cpp:142-216) for a batch of frames.
//
// One lane per pixel, grid-stride over frames x pixels. A pixel's draws depend only on (seed, frame, pixel), so
// lanes are independent; the per-pixel work (Knuth loop, Box-Muller, Marsaglia-Tsang in FP64) is VALU-bound and
// its lane divergence is the Poisson/Gamma draw counts. Reads 4 B and writes 2 B per pixel (plus maps).
#include "iscg_internal.h"

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>

namespace iscg {

template <bool EMCCD>
__global__ __launch_bounds__(256) void noise_kernel(const float* __restrict__ photons, uint16_t* __restrict__ out,
                                                     uint32_t npx, uint32_t nFrames, const uint32_t* __restrict__ frames,
                                                     const float* __restrict__ offset, const float* __restrict__ gain,
                                                     const float* __restrict__ readN, CamParams cp, uint32_t seed,
                                                     uint32_t pixelBase, NoiseQueue q)
{
   const uint32_t total = npx * nFrames;  // < 2^32 (the host splits larger batches)
   for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += gridDim.x * blockDim.x)
   {
      const uint32_t f = idx / npx;
      const uint32_t i = idx - f * npx;
      noise_store<EMCCD>(cp, q, out, idx, photons[idx], seed, frames[f], i + pixelBase, offset != nullptr,
                         offset ? offset[i] : 0.0f, gain != nullptr, gain ? gain[i] : 0.0f, readN != nullptr,
                         readN ? readN[i] : 0.0f);
   }
}

// The FP64 chain for the deferred pixels of the previous launch on the stream; the last block to finish zeroes the
// counters for the next launch.
template <bool EMCCD>
__global__ __launch_bounds__(256) void noise_fallback_kernel(NoiseQueue q, uint16_t* __restrict__ out, uint32_t npx,
                                                              const uint32_t* __restrict__ frames, uint32_t frameBase,
                                                              const float* __restrict__ offset,
                                                              const float* __restrict__ gain,
                                                              const float* __restrict__ readN, CamParams cp,
                                                              uint32_t seed, uint32_t pixelBase)
{
   const uint32_t n = __hip_atomic_load(q.count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
   for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x)
   {
      const uint2 it = q.items[j];
      const uint32_t f = it.x / npx;
      const uint32_t i = it.x - f * npx;
      out[it.x] = noise_pixel<EMCCD>(cp, __uint_as_float(it.y), seed, frames ? frames[f] : frameBase + f,
                                     i + pixelBase, offset != nullptr, offset ? offset[i] : 0.0f, gain != nullptr,
                                     gain ? gain[i] : 0.0f, readN != nullptr, readN ? readN[i] : 0.0f);
   }
   __syncthreads();
   if (threadIdx.x == 0)
   {
      // every block has read n before it counts itself done
      const uint32_t done = __hip_atomic_fetch_add(q.count + 1, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
      if (done == gridDim.x - 1)
      {
         __hip_atomic_store(q.count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
         __hip_atomic_store(q.count + 1, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
   }
}

bool noise_queue(iscg_context* h, size_t pixels, NoiseQueue& q)
{
   if (!h->noiseCount)
   {
      void* p = nullptr;
      if (!ok(h, hipMalloc(&p, 2 * sizeof(uint32_t)), "noise queue counters"))
         return false;
      h->noiseCount = (uint32_t*)p;
      if (!ok(h, hipMemsetAsync(p, 0, 2 * sizeof(uint32_t), h->stream), "noise queue counters"))
         return false;
   }
   if (!ok(h, h->noiseItems.reserve(pixels * sizeof(uint2)), "noise queue"))
      return false;
   q.items = h->noiseItems.as<uint2>();
   q.count = h->noiseCount;
   return true;
}

bool launch_noise_fallback(iscg_context* h, const NoiseQueue& q, uint16_t* out, uint32_t npx, const uint32_t* frames,
                           uint32_t frameBase, const float* offset, const float* gain, const float* readN,
                           const CamParams& cp, uint32_t seed, uint32_t pixelBase)
{
   ProfScope ps(h, "noise_fallback");
   const unsigned grid = cp.fast ? 64 : 2048;  // the FP64 mode defers every pixel
   if (cp.emccd)
      hipLaunchKernelGGL(noise_fallback_kernel<true>, dim3(grid), dim3(256), 0, h->stream, q, out, npx, frames,
                         frameBase, offset, gain, readN, cp, seed, pixelBase);
   else
      hipLaunchKernelGGL(noise_fallback_kernel<false>, dim3(grid), dim3(256), 0, h->stream, q, out, npx, frames,
                         frameBase, offset, gain, readN, cp, seed, pixelBase);
   return ok(h, hipGetLastError(), "noise fallback launch");
}

// Launches the noise kernel on device arrays (shared with the splat path when noise runs as its own pass), in
// batches of whole frames below 2^31 pixels so the launch's pixel index fits 32 bits.
bool launch_noise(iscg_context* h, const float* photons, uint16_t* out, uint32_t npx, uint32_t nFrames,
                  const uint32_t* frames, const float* offset, const float* gain, const float* readN,
                  const CamParams& cp, uint32_t seed, uint32_t pixelBase)
{
   if ((size_t)npx * nFrames == 0)
      return true;
   if (npx > (1u << 31))
      return fail(h, "noise: more than 2^31 pixels per frame") == 0;
   const uint32_t perLaunch = (uint32_t)std::min<size_t>(nFrames, ((size_t)1 << 31) / npx);
   NoiseQueue q;
   if (!noise_queue(h, (size_t)perLaunch * npx, q))
      return false;
   for (uint32_t f0 = 0; f0 < nFrames; f0 += perLaunch)
   {
      const uint32_t nf = std::min(perLaunch, nFrames - f0);
      const float* ph = photons + (size_t)f0 * npx;
      uint16_t* o = out + (size_t)f0 * npx;
      const unsigned block = 256, grid = grid_for((size_t)nf * npx, block, 16);
      {
         ProfScope ps(h, "noise");
         if (cp.emccd)
            hipLaunchKernelGGL(noise_kernel<true>, dim3(grid), dim3(block), 0, h->stream, ph, o, npx, nf, frames + f0,
                               offset, gain, readN, cp, seed, pixelBase, q);
         else
            hipLaunchKernelGGL(noise_kernel<false>, dim3(grid), dim3(block), 0, h->stream, ph, o, npx, nf,
                               frames + f0, offset, gain, readN, cp, seed, pixelBase, q);
         if (!ok(h, hipGetLastError(), "noise kernel launch"))
            return false;
      }
      if (!launch_noise_fallback(h, q, o, npx, frames + f0, 0, offset, gain, readN, cp, seed, pixelBase))
         return false;
   }
   return true;
}

// ---- verification of the FP32 path's error bounds (see noise_pixel_fast) -----------------------------------------

__device__ __forceinline__ double wave_max(double v)
{
   for (int o = 32; o > 0; o >>= 1)
   {
      const double w = __shfl_xor(v, o);
      v = v > w ? v : w;
   }
   return v;
}

// v >= 0 or NaN; NaN and inf count as +inf. Non-negative doubles order as their bit patterns.
__device__ __forceinline__ void atomic_max_nonneg(unsigned long long* dst, double v)
{
   const unsigned long long bits = v <= 1.7976931348623157e308 ? (unsigned long long)__double_as_longlong(v)
                                                                : 0x7FF0000000000000ull;
   atomicMax(dst, bits);
}

// out[0] = max |r_fp32 - r|, out[1] = max |c_fp32 - c| over the 2^23 uniforms (j + 0.5) 2^-23 of CounterRng.
__global__ __launch_bounds__(256) void verify_gauss_kernel(unsigned long long* out)
{
   double mr = 0.0, mc = 0.0;
   for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < (1u << 23); j += gridDim.x * blockDim.x)
   {
      const double u = ((double)j + 0.5) * 1.1920928955078125e-7;
      const float uf = (float)u;
      const double dr = fabs((double)gauss_r_fast(uf) - sqrt(-2.0 * log(u)));
      const double dc = fabs((double)gauss_c_fast(uf) - cos(6.283185307179586 * u));
      mr = !(dr <= mr) ? dr : mr;  // NaN propagates
      mc = !(dc <= mc) ? dc : mc;
   }
   mr = wave_max(mr);
   mc = wave_max(mc);
   if ((threadIdx.x & 63) == 0)
   {
      atomic_max_nonneg(out, mr);
      atomic_max_nonneg(out + 1, mc);
   }
}

// out[2] = max |exp_neg_fast(f) / exp(-f) - 1| over every float f in [0, 60].
__global__ __launch_bounds__(256) void verify_exp_kernel(unsigned long long* out)
{
   constexpr uint32_t kLast = 0x42700000u;  // 60.0f
   double m = 0.0;
   for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b <= kLast; b += gridDim.x * blockDim.x)
   {
      const float f = __uint_as_float(b);
      const double e = exp(-(double)f);
      const double rel = fabs((double)exp_neg_fast(f) - e) / e;
      m = !(rel <= m) ? rel : m;
   }
   m = wave_max(m);
   if ((threadIdx.x & 63) == 0)
      atomic_max_nonneg(out + 2, m);
}

__global__ void zero_words_kernel(uint32_t* p, uint32_t n)
{
   for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
      p[i] = 0u;
}

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_verifyMutex;
struct Verified
{
   bool done = false, fastOk = false;
   double gaussBound = 0.0, expBound = 0.0;
};
Verified g_verified[kMaxDevices];
}  // namespace

bool noise_verified(int device)
{
   std::lock_guard<std::mutex> lock(g_verifyMutex);
   return device >= 0 && device < kMaxDevices && g_verified[device].done;
}

bool noise_verify_device(iscg_context* h, double* bounds, unsigned long long* work)
{
   std::lock_guard<std::mutex> lock(g_verifyMutex);
   if (h->device < 0 || h->device >= kMaxDevices)
      return false;  // the FP64 chain
   Verified& v = g_verified[h->device];
   if (!v.done)
   {
      unsigned long long* d = work;
      unsigned long long r[3] = {0, 0, 0};
      if (!d && !ok(h, hipMalloc((void**)&d, sizeof(r)), "noise verification alloc"))
         return false;
      hipLaunchKernelGGL(zero_words_kernel, dim3(1), dim3(64), 0, h->stream, (uint32_t*)d, 6u);
      hipLaunchKernelGGL(verify_gauss_kernel, dim3(1024), dim3(256), 0, h->stream, d);
      hipLaunchKernelGGL(verify_exp_kernel, dim3(8192), dim3(256), 0, h->stream, d);
      const bool good =
         ok(h, hipGetLastError(), "noise verification launch") &&
         ok(h, hipMemcpyAsync(r, d, sizeof(r), hipMemcpyDeviceToHost, h->stream), "noise verification copy") &&
         ok(h, hipStreamSynchronize(h->stream), "noise verification");
      if (!work)
         (void)hipFree(d);
      if (!good)
         return false;
      double m[3];
      std::memcpy(m, r, sizeof(m));
      // |g_fp32 - g| <= max|dr| (1 + max|dc|) + |r|max |dc| (+ the FP64 product's rounding); the threshold also
      // carries l -> (float)l: |l - (float)l| <= 60 * 2^-24 in the exponent, a factor within 1 +- 3.6e-6.
      v.gaussBound = m[0] * (1.0 + m[1]) + kRMax * m[1] + 1e-15;
      v.expBound = m[2] * (1.0 + 3.6e-6) + 3.6e-6 + 1e-15;
      v.fastOk = v.gaussBound <= 0.5 * kGaussErr && v.expBound <= 0.5 * kExpRel;  // false for NaN too
      v.done = true;
   }
   if (bounds)
   {
      bounds[0] = v.gaussBound;
      bounds[1] = v.expBound;
   }
   return v.fastOk;
}

bool cam_params_for(iscg_context* h, const iscg_camera* c, CamParams& out)
{
   out = cam_params(c);
   if (h->noiseExact)
      return true;
   const std::string before = h->err;
   const bool fast = noise_verify_device(h, nullptr, nullptr);
   if (!fast && h->err != before)
      return false;  // the verification itself failed
   out.fast = fast ? 1 : 0;
   return true;
}

// ---- the allocation-free entry point's scratch (iscg_noise_noalloc) ----------------------------------------------
//
// One fixed array per device, loaded with the code object and held until the process ends: the deferred-pixel
// queue's two counters (words 0 and 1), the verification's three maxima (bytes 64..87), then the queue's items, one
// per pixel of a call (ISCG_NOISE_NOALLOC_MAX_PIXELS at most). In host code `scratch` names a host-side stand-in, not
// device memory: the device address is looked up once per device (noalloc_setup) and only that address is used.
constexpr size_t kNoallocHeader = 256;
static __device__ __align__(256) unsigned char scratch[kNoallocHeader + (size_t)ISCG_NOISE_NOALLOC_MAX_PIXELS * 8];

namespace {
std::mutex g_noallocMutex;
void* g_noallocPtr[kMaxDevices] = {};
}  // namespace

bool noalloc_setup(int device, std::string& err)
{
   std::lock_guard<std::mutex> lock(g_noallocMutex);
   if (device < 0 || device >= kMaxDevices)
   {
      err = "device index beyond the no-alloc table";
      return false;
   }
   if (g_noallocPtr[device])
      return true;
   void* p = nullptr;
   const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));  // the current device's copy
   if (e != hipSuccess || !p)
   {
      (void)hipGetLastError();
      err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return false;
   }
   g_noallocPtr[device] = p;
   return true;
}

static unsigned char* noalloc_ptr(int device)
{
   std::lock_guard<std::mutex> lock(g_noallocMutex);
   return device >= 0 && device < kMaxDevices ? (unsigned char*)g_noallocPtr[device] : nullptr;
}

}  // namespace iscg

using namespace iscg;

extern "C" int32_t iscg_noise_set_path(iscg_context* h, int32_t exact)
{
   if (begin(h) != 0)
      return -1;
   if (exact != 0 && exact != 1)
      return fail(h, "iscg_noise_set_path: exact must be 0 or 1");
   h->noiseExact = exact;
   return 0;
}

extern "C" int32_t iscg_noise_bounds(iscg_context* h, double* out)
{
   if (begin(h) != 0)
      return -1;
   if (!out)
      return fail(h, "iscg_noise_bounds: out is NULL");
   double b[2] = {0.0, 0.0};
   const std::string before = h->err;
   const bool fast = noise_verify_device(h, b, nullptr);
   if (h->err != before)
      return -1;
   out[0] = fast && !h->noiseExact ? 1.0 : 0.0;
   out[1] = b[0];
   out[2] = b[1];
   return 0;
}

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
   CamParams cp;
   if (!cam_params_for(h, cam, cp) ||
       !launch_noise(h, dPh, dOut, (uint32_t)npx, nFrames, dFr, dOff, dGain, dRead, cp, seed, pixelBase))
      return -1;
   if (!finish_out(h, out, dOut, nAll * sizeof(uint16_t)))
      return -1;
   return 0;
}

extern "C" int32_t iscg_noise_noalloc(iscg_context* h, const float* photons, uint32_t width, uint32_t height,
                                      const iscg_camera* cam, const float* offset, const float* gain,
                                      const float* readNoise, uint32_t seed, const uint32_t* frames, uint32_t nFrames,
                                      uint32_t pixelBase, uint16_t* out)
{
   if (begin(h))
      return -1;
   if (!cam || !frames || !out || (!photons && nFrames))
      return fail(h, "iscg_noise_noalloc: null argument");
   const size_t npx = (size_t)width * height;
   if (npx == 0 || nFrames == 0)
      return 0;
   const size_t nAll = npx * nFrames;
   if (nAll > (size_t)ISCG_NOISE_NOALLOC_MAX_PIXELS)
      return fail(h, "iscg_noise_noalloc: " + std::to_string(nAll) + " pixels in one call; at most " +
                        std::to_string((size_t)ISCG_NOISE_NOALLOC_MAX_PIXELS) + " (iscg_noise takes any size)");
   // no staging without allocation: every array must already be device memory
   if (!is_device_ptr(photons) || !is_device_ptr(frames) || !is_device_ptr(out) ||
       (offset && !is_device_ptr(offset)) || (gain && !is_device_ptr(gain)) ||
       (readNoise && !is_device_ptr(readNoise)))
      return fail(h, "iscg_noise_noalloc: every array must be device memory");
   unsigned char* base = noalloc_ptr(h->device);
   if (!base)
      return fail(h, "iscg_noise_noalloc: no scratch on this device (its look-up failed at iscg_create)");
   NoiseQueue q;
   q.count = (uint32_t*)base;
   q.items = (uint2*)(base + kNoallocHeader);
   CamParams cp = cam_params(cam);
   if (!h->noiseExact)
   {
      // the FP32 path once this device's bounds are verified; the verification runs here (on the scratch, no
      // allocation) unless the stream is being captured, where the FP64 chain gives the same ADU
      hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
      const bool capturing = hipStreamIsCapturing(h->stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone;
      (void)hipGetLastError();
      if (!capturing || noise_verified(h->device))
      {
         const std::string before = h->err;
         const bool fast = noise_verify_device(h, nullptr, (unsigned long long*)(base + 64));
         if (!fast && h->err != before)
            return -1;
         cp.fast = fast ? 1 : 0;
      }
   }
   // the counters start at zero (the fallback's last block also zeroes them; a call that failed between the two
   // launches would not have)
   hipLaunchKernelGGL(zero_words_kernel, dim3(1), dim3(64), 0, h->stream, q.count, 2u);
   const unsigned block = 256, grid = grid_for(nAll, block, 16);
   {
      ProfScope ps(h, "noise");
      if (cp.emccd)
         hipLaunchKernelGGL(noise_kernel<true>, dim3(grid), dim3(block), 0, h->stream, photons, out, (uint32_t)npx,
                            nFrames, frames, offset, gain, readNoise, cp, seed, pixelBase, q);
      else
         hipLaunchKernelGGL(noise_kernel<false>, dim3(grid), dim3(block), 0, h->stream, photons, out, (uint32_t)npx,
                            nFrames, frames, offset, gain, readNoise, cp, seed, pixelBase, q);
      if (!ok(h, hipGetLastError(), "noise kernel launch"))
         return -1;
   }
   if (!launch_noise_fallback(h, q, out, (uint32_t)npx, frames, 0, offset, gain, readNoise, cp, seed, pixelBase))
      return -1;
   return 0;
}
