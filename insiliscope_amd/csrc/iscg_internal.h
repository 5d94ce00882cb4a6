// iscg_internal.h -- shared host and device pieces of libiscg: the context, device buffers, pointer staging and
// the counter RNG / camera-noise device functions used by several kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/iscg.h"

#include <cstdint>
#include <cstring>
#include <functional>
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
struct WfAccelState;
struct WfDriftState;
struct RoiState;
struct BinnedState;

}  // namespace iscg

struct iscg_context
{
   int device = 0;
   hipStream_t stream = nullptr;     // where the context's work is enqueued (iscg_set_stream)
   hipStream_t ownStream = nullptr;  // the stream iscg_create made
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
   iscg::WfAccelState* wfAccel = nullptr;
   iscg::WfDriftState* wfDrift = nullptr;
   iscg::RoiState* roi = nullptr;
   iscg::BinnedState* binned = nullptr;
   iscg::DevBuf dyeCounts;             // per-dye record counts of a schedule call
   iscg::DevBuf dyeBlocks;             // per-workgroup sums, then their exclusive prefixes
   int noiseExact = 0;  // iscg_noise_set_path: 1 evaluates every pixel with the FP64 chain
   iscg::DevBuf noiseItems;            // deferred pixels (NoiseQueue::items)
   uint32_t* noiseCount = nullptr;     // NoiseQueue::count, 2 words
   // the last bucket offsets iscg_collect_frame_emitters validated from device memory (reused under capture)
   const void* collectStPtr = nullptr;
   uint32_t collectStFrames = 0;
   size_t collectStTotal = 0;
   iscg::DevBuf bucketTmp[6];          // iscg_bucket_events: counts, offsets, keys and values (twice), sort scratch
   int tuneMode = 1;                   // iscg_autotune_set: 1 tunes launch parameters on first use of a shape key
   std::map<std::string, int> tuned;   // shape key -> chosen launch parameter
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
      return unit_of(x);
   }
   // ((x >> 9) + 0.5) 2^-23, exact in FP64 (and in FP32)
   static __device__ __forceinline__ double unit_of(uint32_t x) { return ((double)(x >> 9) + 0.5) * 1.1920928955078125e-7; }
   // Starts the next group of four draws (the state uniform() would reach after using up the current one).
   __device__ __forceinline__ void next_group()
   {
      b0 = seed;
      b1 = frame;
      b2 = pix;
      b3 = ctr;
      pcg4d(b0, b1, b2, b3);
      ++ctr;
      k = 0;
   }
};

// Knuth's product loop of CounterPoisson (Sim/SMLMCounterRng.h:110-118) from a fresh stream (the chain's first
// draws, so groups of four start at draw 0): p_j = p_{j-1} u_j in FP64, stop at the first p_k <= L or at k = 1000;
// returns k - 1 and leaves u after draw k, as the one-draw-at-a-time loop does. The products strictly decrease
// (u <= 1 - 2^-24, and no underflow before p <= L since L >= e^-60), so a group holds the stop iff its fourth
// product is <= L: the loop does four multiplies and one compare per group, and the stop is located once, after it.
// last = p_k, prev = p_{k-1} (p_0 = 1).
__device__ __forceinline__ int knuth_fresh(double L, CounterRng& u, double& last, double& prev)
{
   double p = 1.0, p1, p2, p3, p4;
   int g = 0;
   for (;; ++g)
   {
      u.next_group();
      p1 = p * CounterRng::unit_of(u.b0);
      p2 = p1 * CounterRng::unit_of(u.b1);
      p3 = p2 * CounterRng::unit_of(u.b2);
      p4 = p3 * CounterRng::unit_of(u.b3);
      if (!(p4 > L) || g == 249)  // 250 groups = the 1000-draw cap
         break;
      p = p4;
   }
   const int n = (int)(p1 > L) + (int)(p2 > L) + (int)(p3 > L) + (int)(p4 > L);  // products > L in this group
   last = n == 0 ? p1 : (n == 1 ? p2 : (n == 2 ? p3 : p4));
   prev = n == 0 ? p : (n == 1 ? p1 : (n == 2 ? p2 : p3));
   u.k = n < 4 ? n + 1 : 4;
   return 4 * g + (n < 4 ? n : 3);
}

// Sim/SMLMCounterRng.h:95-99
__device__ __forceinline__ double counter_gauss(CounterRng& u)
{
   const double r = sqrt(-2.0 * log(u.uniform()));
   return r * cos(6.283185307179586 * u.uniform());
}

// Sim/SMLMCounterRng.h:104-119, as the chain's first draw (u fresh)
__device__ __forceinline__ double counter_poisson(double l, CounterRng& u)
{
   if (l > 60.0)
   {
      const double v = floor(l + sqrt(l) * counter_gauss(u) + 0.5);
      return 0.0 < v ? v : 0.0;
   }
   if (!(l > 0.0))
      return 0.0;
   double last, prev;
   return (double)knuth_fresh(exp(-l), u, last, prev);
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
   double invGain;  // 1 / gain (the FP32 path's multiply)
   int emccd;
   int fast;  // 1: noise_pixel_fast

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
   p.invGain = 1.0 / p.gain;
   p.emccd = c->emccd != 0;
   p.fast = 0;
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

// The autotuner. A tunable launch parameter has several values that all give the same results (the size of a
// persistent grid, how many workgroups share a request's planes). On the first call with a new shape key, with
// tuning on, each candidate is launched once (warm-up), then timed over kTuneReps launches between two HIP events on
// the context's stream, and the fastest is kept for the key; later calls reuse it. launch(c) enqueues the kernel(s)
// with parameter c and returns false on a launch error (the default is then returned). While the stream is being
// captured into a graph no timing is possible: the default is used and nothing is kept. A kept value that is not
// among the candidates (a loaded cache of another build) is tuned again.
constexpr int kTuneReps = 3;
int tuned_choice(iscg_context* h, const std::string& key, const std::vector<int>& cands, int dflt,
                 const std::function<bool(int)>& launch);
// Rounds up to a power of two (shape keys bucket batch sizes so that nearby sizes share a choice).
inline long long pow2_bucket(long long n)
{
   long long b = 1;
   while (b < n)
      b <<= 1;
   return b;
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

// ---- filtered fast path -----------------------------------------------------------------------------------------
// The same chain with the Box-Muller and Knuth-threshold transcendentals in FP32, carrying a bound on their error;
// every decision that depends on them (a count boundary, the final rounding) is taken only when it holds across the
// whole error interval, else the pixel is recomputed by noise_pixel, the FP64 chain above. The draws consumed are the
// same (the same branches are taken), so the ADU are those of noise_pixel.
//
// The bounds are measured, not assumed: noise_verify_device evaluates, on the device, every input the fast path can
// see (u1 and u2 each take the 2^23 values (j + 0.5) 2^-23; the Poisson mean is (float)l with l in (0, 60], every
// float up to 60) against the FP64 functions of the exact chain, and the fast path is used only when
//   |g_fast - g| <= max|dr| (1 + max|dc|) + 5.77 max|dc| + 1e-15 <= kGaussErr / 2   (|r| <= sqrt(-2 log 2^-24) = 5.768)
//   |L_fast / L - 1| <= max rel(exp_neg_fast) + 3.6e-6 <= kExpRel / 2  (|l - (float)l| <= 60 * 2^-24 = 3.58e-6)
// The margins below use the full kGaussErr and kExpRel, twice the verified bounds.
constexpr double kGaussErr = 1e-5;
constexpr double kExpRel = 2e-5;
constexpr double kRMax = 5.77;

// The FP32 transcendentals of the fast path are the hardware's instructions: log2, square root, cosine of turns and
// exp2. Their accuracy is not assumed: noise_verify_device measures these very functions over every input.
__device__ __forceinline__ float gauss_r_fast(float u1)  // sqrt(-2 ln u1)
{
   return __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u1));
}
__device__ __forceinline__ float gauss_c_fast(float u2)  // cos(2 pi u2)
{
   return __builtin_amdgcn_cosf(u2);
}
__device__ __forceinline__ float exp_neg_fast(float l)  // e^-l
{
   return __builtin_amdgcn_exp2f(-1.4426950408889634f * l);
}

__device__ __forceinline__ double gauss_fast(CounterRng& u)
{
   const float u1 = (float)u.uniform(), u2 = (float)u.uniform();
   return (double)gauss_r_fast(u1) * (double)gauss_c_fast(u2);  // exact product of two floats
}

// Poisson(l) with the filtered threshold; ok = false when a decision fell inside the error interval.
__device__ __forceinline__ double counter_poisson_fast(double l, CounterRng& u, bool& ok)
{
   if (l > 60.0)
   {
      const double sl = sqrt(l);
      const double v = l + sl * gauss_fast(u) + 0.5;
      const double m = sl * kGaussErr + 1e-14 * (1.0 + l + fabs(v));
      const double a = floor(v - m), b = floor(v + m);
      ok &= a == b;
      return 0.0 < a ? a : 0.0;
   }
   if (!(l > 0.0))
      return 0.0;
   const double L = (double)exp_neg_fast((float)l);
   const double tol = L * kExpRel;
   double last, prev;
   const int n = knuth_fresh(L, u, last, prev);
   // the products decrease: only p_k and p_{k-1} can sit on the other side of the exact threshold
   ok &= fabs(last - L) > tol && (n == 0 || fabs(prev - L) > tol);
   return (double)n;
}

// ok = false: some decision was within the error interval; the result is then not used.
template <bool EMCCD>
__device__ __forceinline__ uint16_t noise_pixel_fast(const CamParams& cp, float photon, uint32_t seed, uint32_t frame,
                                                     uint32_t pixel, bool hasOffset, float offsetV, bool hasGain,
                                                     float gainV, bool hasRead, float readV, bool& ok)
{
   CounterRng u(seed, frame, pixel);
   ok = true;
   const double ph0 = (double)photon;
   const double photons = 0.0 < ph0 ? ph0 : 0.0;
   // 1/gain: the launch's FP64 reciprocal (relative error 1.1e-16), or a map value's FP32 one (<= 6e-8)
   const double invGain = hasGain ? (double)(1.0f / gainV) : cp.invGain;
   const double relErr = hasGain ? 2e-7 : 1e-12;  // with the FP64 roundings of the sum, inflated
   const double offset = hasOffset ? (double)offsetV : 0.0;
   double r = 0.0;
   if (EMCCD)
   {
      const double ne = counter_poisson_fast(photons * cp.qe + cp.dark + cp.cic, u, ok);
      if (ok)
      {
         const double out = ne > 0.0 ? counter_gamma(ne, u) : 0.0;  // the gain register stays FP64
         const double readE = (hasRead ? (double)readV : cp.read) / cp.emGain;
         const double q = (out + readE * gauss_fast(u)) * invGain;
         const double x = offset + q + 0.5;
         // monotone in x: floor then clip agree over [x - m, x + m] when the two floors do
         const double m = fabs(readE * invGain) * kGaussErr + relErr * (1.0 + fabs(offset) + fabs(q)) +
                          1e-12 * kRMax * fabs(readE * invGain);
         const double a = floor(x - m), b = floor(x + m);
         ok &= a == b;
         r = a < 0.0 ? 0.0 : (a > cp.maxAdu ? cp.maxAdu : a);
      }
   }
   else
   {
      const double ne = counter_poisson_fast(photons * cp.qe + cp.dark, u, ok);
      if (ok)
      {
         const double readNoise = hasRead ? (double)readV : cp.read;
         const double q = (ne + readNoise * gauss_fast(u)) * invGain;
         const double x = offset + q;
         const double m = fabs(readNoise * invGain) * kGaussErr + relErr * (1.0 + fabs(offset) + fabs(q)) +
                          1e-12 * kRMax * fabs(readNoise * invGain);
         // clip then round half up, both monotone: take both ends of the interval
         double lo = x - m, hi = x + m;
         lo = lo < 0.0 ? 0.0 : (lo > 65535.0 ? 65535.0 : lo);
         hi = hi < 0.0 ? 0.0 : (hi > 65535.0 ? 65535.0 : hi);
         const double a = floor(lo + 0.5), b = floor(hi + 0.5);
         ok &= a == b;  // NaN anywhere fails here and takes the FP64 chain
         r = a;
      }
   }
   return (uint16_t)r;
}

// ---- deferred pixels ----------------------------------------------------------------------------------------------
// A launch that produces ADU evaluates each pixel with noise_pixel_fast; a pixel it cannot decide (or every pixel,
// when the context asks for the FP64 chain) is appended here as (output index, photon bits) and finished by
// noise_fallback_kernel, which runs noise_pixel. Keeping the FP64 chain out of the producing kernels keeps their
// register use (and occupancy) to that of the FP32 path. Launch geometry shared by all producers: output index
// idx = f * npx + i, frame id frames[f] (or frameBase + f), stream pixel i + pixelBase, maps indexed by i.
struct NoiseQueue
{
   uint2* items;     // [capacity >= pixels of the launch]
   uint32_t* count;  // [0]: items, [1]: fallback blocks done (both zero between launches)
};

__device__ __forceinline__ void noise_defer(const NoiseQueue& q, uint32_t idx, float photon)
{
   const uint32_t slot = atomicAdd(q.count, 1u);
   q.items[slot] = make_uint2(idx, __float_as_uint(photon));
}

template <bool EMCCD>
__device__ __forceinline__ void noise_store(const CamParams& cp, const NoiseQueue& q, uint16_t* __restrict__ out,
                                            uint32_t idx, float photon, uint32_t seed, uint32_t frame, uint32_t pixel,
                                            bool hasOffset, float offsetV, bool hasGain, float gainV, bool hasRead,
                                            float readV)
{
   bool ok = false;
   uint16_t r = 0;
   if (cp.fast)
      r = noise_pixel_fast<EMCCD>(cp, photon, seed, frame, pixel, hasOffset, offsetV, hasGain, gainV, hasRead, readV,
                                  ok);
   if (ok)
      out[idx] = r;
   else
      noise_defer(q, idx, photon);
}

// The queue for a launch of `pixels` ADU (capacity grows; the counters are allocated and zeroed once).
bool noise_queue(iscg_context* h, size_t pixels, NoiseQueue& q);
// Finishes the deferred pixels of the launch just enqueued on h's stream and resets the queue's counters.
// frames: per-frame ids or null (then frameBase + f).
bool launch_noise_fallback(iscg_context* h, const NoiseQueue& q, uint16_t* out, uint32_t npx, const uint32_t* frames,
                           uint32_t frameBase, const float* offset, const float* gain, const float* readN,
                           const CamParams& cp, uint32_t seed, uint32_t pixelBase);

// Runs the device verification once per device (cached for the process); true when the fast path may be used.
// bounds (optional, 2 doubles): the verified |g_fast - g| bound and relative threshold bound.
// work: 3 device words for the maxima, or null (then allocated and freed here).
bool noise_verify_device(iscg_context* h, double* bounds, unsigned long long* work);
// True once the verification has run on this device in this process.
bool noise_verified(int device);
// Looks up, once per device and with that device current, the device address of the allocation-free noise entry
// point's scratch array (iscg_noise_noalloc). False with err set when the look-up fails.
bool noalloc_setup(int device, std::string& err);
// The CamParams of a launch on h: cam_params plus the path (fast unless h asks for the exact chain or the device
// failed verification). False (error set) when the verification could not run.
bool cam_params_for(iscg_context* h, const iscg_camera* c, CamParams& out);

// Grid size for a grid-stride launch over n items with `block` threads: enough workgroups to fill all 256 CUs
// several times over, never more than the items need.
inline unsigned grid_for(size_t n, unsigned block, unsigned perCu = 8)
{
   const size_t need = (n + block - 1) / block;
   const size_t cap = (size_t)256 * perCu;
   return (unsigned)(need < cap ? (need ? need : 1) : cap);
}

}  // namespace iscg
