// iscg_internal.h -- shared host and device pieces of libiscg: the context, device buffers, pointer staging and
// the counter RNG / camera-noise device functions used by several kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/iscg.h"

#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <vector>

namespace iscg {

// ---------------------------------------------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------------------------------------------

// A grow-only device allocation owned by a context.
struct DevBuf
{
   void* p = nullptr;
   size_t bytes = 0;
   DevBuf() = default;
   DevBuf(const DevBuf&) = delete;
   DevBuf& operator=(const DevBuf&) = delete;
   ~DevBuf() { release(); }
   void release()
   {
      if (p)
         (void)hipFree(p);
      p = nullptr;
      bytes = 0;
   }
   // Ensures at least n bytes (contents are not kept when it grows).
   hipError_t reserve(size_t n)
   {
      if (n <= bytes && p)
         return hipSuccess;
      release();
      const size_t want = n < 256 ? 256 : n;
      hipError_t e = hipMalloc(&p, want);
      if (e == hipSuccess)
         bytes = want;
      else
         p = nullptr;
      return e;
   }
   template <typename T>
   T* as() const
   {
      return static_cast<T*>(p);
   }
};

struct SplatState;
struct PsfState;
struct WfState;
struct RoiState;

}  // namespace iscg

struct iscg_context
{
   int device = 0;
   hipStream_t stream = nullptr;
   std::string err;
   // per-kernel timing (iscg_profile_*): event pairs around each launch on the context's stream
   struct ProfPending
   {
      const char* name;
      hipEvent_t a, b;
   };
   bool profiling = false;
   std::vector<ProfPending> pending;
   std::vector<hipEvent_t> eventPool;
   std::map<std::string, std::pair<long long, double>> profTotals;
   int injectCount = 0;
   std::string injectMessage;
   // staging for host arguments, one per argument slot of a call
   iscg::DevBuf stage[8];
   iscg::SplatState* splat = nullptr;
   iscg::PsfState* psf = nullptr;
   iscg::WfState* wf = nullptr;
   iscg::RoiState* roi = nullptr;
};

namespace iscg {

// Sets the context's error and returns -1.
int32_t fail(iscg_context* h, const std::string& msg);
// Checks a HIP status: on failure sets the error (with what) and returns false.
bool ok(iscg_context* h, hipError_t e, const char* what);
// Entry-point prologue: clears the error, makes the device current, applies an injected failure. 0 or -1.
int32_t begin(iscg_context* h);
// True when p is device memory (of any device); host and unregistered pointers give false.
bool is_device_ptr(const void* p);

// An input argument: device pointers pass through, host ones are copied into stage[slot]. nullptr on failure.
const void* stage_in(iscg_context* h, int slot, const void* p, size_t bytes);
// An output argument: device pointers pass through, host ones get stage[slot] (copy back with finish_out).
void* stage_out(iscg_context* h, int slot, void* p, size_t bytes);
// Copies a staged output back (no-op for device outputs) and, for host outputs, waits for the stream.
bool finish_out(iscg_context* h, void* user, const void* dev, size_t bytes);

// ---------------------------------------------------------------------------------------------------------------
// Device side: the counter RNG and the camera noise chain (Sim/SMLMCounterRng.h, Sim/SMLMNoise.cpp). Compiled
// with -ffp-contract=off: every FP64 expression is evaluated as written, in the reference's order, so the ADU
// equal the CPU's (the transcendentals are the device's correctly rounded-or-1-ulp FP64 libm; a count boundary
// within an ulp of a draw is the only way to differ).
// ---------------------------------------------------------------------------------------------------------------

__device__ __forceinline__ void pcg4d(uint32_t& a, uint32_t& b, uint32_t& c, uint32_t& d)
{
   a = a * 1664525u + 1013904223u;
   b = b * 1664525u + 1013904223u;
   c = c * 1664525u + 1013904223u;
   d = d * 1664525u + 1013904223u;
   a += b * d;
   b += c * a;
   c += a * b;
   d += b * c;
   a ^= a >> 16;
   b ^= b >> 16;
   c ^= c >> 16;
   d ^= d >> 16;
   a += b * d;
   b += c * a;
   c += a * b;
   d += b * c;
}

// Sim/SMLMCounterRng.h:59-89: four uniforms per pcg4d(seed, frame, pixel, counter), drawn a, b, c, d.
struct CounterRng
{
   uint32_t seed, frame, pix, ctr;
   uint32_t b0, b1, b2, b3;
   int k;
   __device__ __forceinline__ CounterRng(uint32_t s, uint32_t f, uint32_t p) : seed(s), frame(f), pix(p), ctr(0), k(4)
   {
   }
   __device__ __forceinline__ double uniform()
   {
      if (k == 4)
      {
         b0 = seed;
         b1 = frame;
         b2 = pix;
         b3 = ctr;
         pcg4d(b0, b1, b2, b3);
         ++ctr;
         k = 0;
      }
      const uint32_t x = k == 0 ? b0 : (k == 1 ? b1 : (k == 2 ? b2 : b3));
      ++k;
      return ((double)(x >> 9) + 0.5) * 1.1920928955078125e-7;
   }
};

// Sim/SMLMCounterRng.h:95-99
__device__ __forceinline__ double counter_gauss(CounterRng& u)
{
   const double r = sqrt(-2.0 * log(u.uniform()));
   return r * cos(6.283185307179586 * u.uniform());
}

// Sim/SMLMCounterRng.h:104-119
__device__ __forceinline__ double counter_poisson(double l, CounterRng& u)
{
   if (l > 60.0)
   {
      const double v = floor(l + sqrt(l) * counter_gauss(u) + 0.5);
      return 0.0 < v ? v : 0.0;
   }
   if (!(l > 0.0))
      return 0.0;
   const double L = exp(-l);
   int k = 0;
   double p = 1.0;
   do
   {
      ++k;
      p *= u.uniform();
   } while (p > L && k < 1000);
   return (double)(k - 1);
}

// Sim/SMLMCounterRng.h:123-138
__device__ __forceinline__ double counter_gamma(double k, CounterRng& u)
{
   const double d = k - 1.0 / 3.0, c = 1.0 / sqrt(9.0 * d);
   for (int it = 0; it < 1000; ++it)
   {
      const double x = counter_gauss(u), t = 1.0 + c * x;
      if (t <= 0.0)
         continue;
      const double v = t * t * t, w = u.uniform();
      if (w < 1.0 - 0.0331 * x * x * x * x)
         return d * v;
      if (log(w) < 0.5 * x * x + d * (1.0 - v + log(v)))
         return d * v;
   }
   return d;
}

// The scalar camera, resolved once per launch.
struct CamParams
{
   double qe, dark, gain, read, emGain, cic, maxAdu;
   int emccd;
};

inline CamParams cam_params(const iscg_camera* c)
{
   CamParams p;
   p.qe = c->quantumEfficiency;
   p.dark = c->darkCurrentElectrons;
   p.gain = c->gainPhotonsPerAdu;
   p.read = c->readNoiseElectrons;
   p.emGain = 1.0 < c->emGain ? c->emGain : 1.0;
   p.cic = c->cicElectrons;
   const int bits = c->bitDepth < 1 ? 1 : (c->bitDepth > 16 ? 16 : c->bitDepth);
   p.maxAdu = (double)((1u << bits) - 1u);  // ldexp(1, bits) - 1, exact
   p.emccd = c->emccd != 0;
   return p;
}

// One pixel of ApplyNoiseChain (Sim/SMLMNoise.cpp:173-204). photon: the float image value; offset/gain/read: the
// map values when present (hasX), else the scalars.
template <bool EMCCD>
__device__ __forceinline__ uint16_t noise_pixel(const CamParams& cp, float photon, uint32_t seed, uint32_t frame,
                                                uint32_t pixel, bool hasOffset, float offsetV, bool hasGain,
                                                float gainV, bool hasRead, float readV)
{
   CounterRng u(seed, frame, pixel);
   const double ph0 = (double)photon;
   const double photons = 0.0 < ph0 ? ph0 : 0.0;
   if (EMCCD)
   {
      const double ne = counter_poisson(photons * cp.qe + cp.dark + cp.cic, u);
      const double out = ne > 0.0 ? counter_gamma(ne, u) : 0.0;
      const double readE = (hasRead ? (double)readV : cp.read) / cp.emGain;
      const double gain = hasGain ? (double)gainV : cp.gain;
      const double offset = hasOffset ? (double)offsetV : 0.0;
      const double g = counter_gauss(u);
      const double adu = floor(offset + (out + readE * g) / gain + 0.5);
      return (uint16_t)(adu < 0.0 ? 0.0 : (adu > cp.maxAdu ? cp.maxAdu : adu));
   }
   const double ne = counter_poisson(photons * cp.qe + cp.dark, u);
   const double readNoise = hasRead ? (double)readV : cp.read;
   const double gain = hasGain ? (double)gainV : cp.gain;
   const double offset = hasOffset ? (double)offsetV : 0.0;
   const double g = counter_gauss(u);
   double adu = offset + (ne + readNoise * g) / gain;
   adu = adu < 0.0 ? 0.0 : (adu > 65535.0 ? 65535.0 : adu);
   return (uint16_t)floor(adu + 0.5);
}

// Times the launches made in its scope when profiling is on: events recorded on the context's stream before and
// after, resolved by iscg_profile_read.
struct ProfScope
{
   iscg_context* h;
   const char* name;
   hipEvent_t a = nullptr;
   ProfScope(iscg_context* hh, const char* nm);
   ~ProfScope();
};

// Grid size for a grid-stride launch over n items with `block` threads: enough workgroups to fill all 256 CUs
// several times over, never more than the items need.
inline unsigned grid_for(size_t n, unsigned block, unsigned perCu = 8)
{
   const size_t need = (n + block - 1) / block;
   const size_t cap = (size_t)256 * perCu;
   return (unsigned)(need < cap ? (need ? need : 1) : cap);
}

}  // namespace iscg
