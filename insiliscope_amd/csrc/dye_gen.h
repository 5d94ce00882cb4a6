// dye_gen.h -- the dye schedules as device state machines, shared by the schedule calls (dyes.hip) and the world's
// window queries (world.hip): LabelSchedule's blinks and PersistentGen's per-bin blinks (core/src/dyes.cpp:254-657,
// one set of kinetics), bit-identical to the reference's. Each value is a pure function of (h1, k, n): the draws are
// Unit(Pcg4d(h1, k, n, channel)) (the channels of core/src/dyes.h:45-57) and the FP64 arithmetic is the reference's,
// in its order, with the jsm:: functions restated in jsm.h.
#pragma once

#include "iscg_internal.h"
#include "jsm.h"

namespace iscg {
namespace dyegen {

constexpr int kDyeThreads = 256;
constexpr int kScanThreads = 1024;

// core/src/dyes.h:45-57
constexpr uint32_t CH_ACT = 4, CH_PERSIST = 5, CH_INIT_ON = 7, CH_AUX = 8;
constexpr uint32_t CH_SCHED0 = 16, CH_SCHED_STRIDE = 8;
constexpr uint32_t CH_ON = 0, CH_BRIGHT1 = 1, CH_BRIGHT2 = 2, CH_BLEACH = 3, CH_OFF = 4;
constexpr int kMaxBlinks = 1000;       // DYE_MAX_BLINKS
constexpr double kBinSec = 1.0;        // PERSIST_BIN_SEC
constexpr double kOnCap = 20.0;        // PERSIST_ON_CAP

// Unit(Pcg4d(a, b, c, d).a) (core/src/dyes.cpp:14, core/src/rng.h:9-26)
__device__ __forceinline__ double draw(uint32_t a, uint32_t b, uint32_t c, uint32_t d)
{
   pcg4d(a, b, c, d);
   return CounterRng::unit_of(a);
}

// A dye's draws share (h1, k, n): their first LCG step is taken once per dye, each draw then takes its channel's.
struct Address
{
   uint32_t a, b, c;
   __device__ __forceinline__ static Address of(uint32_t h1, uint32_t k, uint32_t n)
   {
      return Address{h1 * 1664525u + 1013904223u, k * 1664525u + 1013904223u, n * 1664525u + 1013904223u};
   }
   // Unit(Pcg4d(h1, k, n, ch).a): pcg4d after its first step
   __device__ __forceinline__ double draw(uint32_t ch) const { return CounterRng::unit_of(word(ch)); }
   // Pcg4d(h1, k, n, ch).a
   __device__ __forceinline__ uint32_t word(uint32_t ch) const
   {
      uint32_t x = a, y = b, z = c, w = ch * 1664525u + 1013904223u;
      x += y * w;
      y += z * x;
      z += x * y;
      w += y * z;
      x ^= x >> 16;
      y ^= y >> 16;
      w ^= w >> 16;
      x += y * w;  // the rest of the second round does not reach the first word
      return x;
   }
};

// LogNormalFor / LogNormalMean1 (core/src/dyes.cpp:257-268): a mean-1 log-normal with CV cv from two uniforms.
struct LogNormal
{
   double sigma, mu;
   __device__ __forceinline__ explicit LogNormal(double cv)
   {
      const double s2 = jsm::log(1 + cv * cv);
      sigma = sqrt(s2);
      mu = -s2 / 2;
   }
   __device__ __forceinline__ double operator()(double u1, double u2) const
   {
      const double z = sqrt(-2 * jsm::log(u1)) * jsm::cos(2 * jsm::kPi * u2);
      return jsm::exp(mu + sigma * z);
   }
};

// PoissonInverse (core/src/dyes.cpp:272-279)
__device__ __forceinline__ long long poisson_inverse(double m, double expM, double u)
{
   double p = expM, F = p;
   long long c = 0;
   while (u > F && c < 1000)
   {
      c++;
      p *= m / static_cast<double>(c);
      F += p;
   }
   return c;
}

// PoissonFromUniform above m = 30 (core/src/dyes.cpp:282-290): the rounded normal approximation, at least 0.
__device__ __forceinline__ long long poisson_normal(double m, double u, double u2)
{
   const double z = sqrt(-2 * jsm::log(u)) * jsm::cos(2 * jsm::kPi * u2);
   const long long c = static_cast<long long>(floor(m + sqrt(m) * z + 0.5));
   return c > 0 ? c : 0;
}

// A schedule is a small state machine: init() draws what precedes the first blink, each step() makes one blink (or,
// for the persistent form, one candidate blink or one bin change) and reports the end. The kernels advance one step
// per loop iteration on every lane and hand a finished lane the next dye of its wave's chunk, so lanes whose
// schedule ended early do not idle while the longest one of the wave runs on.

// LabelSchedule's blinks (core/src/dyes.cpp:440-471 with DyeSchedule/DyeScheduleFrom, :293-321).
struct LabelGen
{
   int32_t mode;
   iscg_kinetics kin;
   double tMax;
   double tLo;  // only the blinks ending after tLo are counted and written (-inf: all of them)
   using Record = iscg_blink;

   struct Consts
   {
      double pBleach, cv;
      LogNormal ln;
   };
   struct State
   {
      Address at;
      int j, count;
      bool done;
      double t, shift, tStop;
   };

   template <bool WRITE>
   __device__ Consts consts() const
   {
      const double cv = fmax(0.0, kin.photonCV);
      return Consts{fmin(1.0, fmax(0.01, kin.bleachProb)), cv, LogNormal(WRITE && cv > 0 ? cv : 0.0)};
   }

   __device__ State init(uint32_t h1, uint32_t k, uint32_t n) const
   {
      State s{Address::of(h1, k, n), 0, 0, true, 0.0, 0.0, 0.0};
      if (mode == 2 || mode == 3)
         return s;
      if (mode == 0 && kin.initialOnSec > 0)
         s.shift = -jsm::log(s.at.draw(CH_INIT_ON)) * kin.initialOnSec;
      if (!(kin.activationRatePerSec > 0))
         return s;
      s.tStop = tMax - s.shift;
      s.t = -jsm::log(s.at.draw(CH_ACT)) / kin.activationRatePerSec;
      s.done = false;
      return s;
   }

   // One blink j (core/src/dyes.cpp:307-320); true when the schedule has ended.
   template <bool WRITE>
   __device__ bool step(State& s, const Consts& c, Record* out) const
   {
      if (s.j >= kMaxBlinks || s.t >= s.tStop)
         return true;
      const uint32_t base = CH_SCHED0 + static_cast<uint32_t>(s.j) * CH_SCHED_STRIDE;
      const double on = -jsm::log(s.at.draw(base + CH_ON)) * kin.onSec;
      // LabelSchedule adds the dSTORM shift to both ends after the schedule is made
      const double tOn = s.shift != 0 ? s.t + s.shift : s.t;
      const double tOff = s.shift != 0 ? (s.t + on) + s.shift : s.t + on;
      if (tOff > tLo)
      {
         if constexpr (WRITE)
         {
            double b = 1;
            if (c.cv > 0)
               b = c.ln(s.at.draw(base + CH_BRIGHT1), s.at.draw(base + CH_BRIGHT2));
            out[s.count] = iscg_blink{tOn, tOff, b};
         }
         ++s.count;
      }
      if (s.at.draw(base + CH_BLEACH) < c.pBleach)
         return true;
      s.t += on - jsm::log(s.at.draw(base + CH_OFF)) * kin.offSec;
      ++s.j;
      return false;
   }
};

// PersistentGen (core/src/dyes.cpp:566-626) for one set of kinetics: bins [b0, b1], each bin's Poisson count of
// blinks at (bin + u) s; the window form keeps those overlapping (t0, t1) (PersistentBlinks, :636-646).
template <bool WINDOW, class RecordT>
struct PersistGen
{
   iscg_kinetics kin;
   double t0, t1;
   long long b0, b1;
   using Record = RecordT;

   // BinKineticsOf (core/src/dyes.cpp:544-558)
   struct Consts
   {
      double cv, maxOn, m, expM;
      bool small;
      LogNormal ln;
   };
   struct State
   {
      uint32_t key;
      int count;
      bool done;
      long long b, c, j;
   };

   template <bool WRITE>
   __device__ Consts consts() const
   {
      const double cv = fmax(0.0, kin.photonCV);
      const double m = kin.activationRatePerSec * kBinSec;
      const bool small = m <= 30;
      return Consts{cv, kOnCap * kin.onSec, m, small ? jsm::exp(-m) : 0.0, small,
                    LogNormal(WRITE && cv > 0 ? cv : 0.0)};
   }

   __device__ long long bin_count(uint32_t key, uint32_t bin, const Consts& q) const
   {
      const double uc = draw(key, bin, 0, 0);
      return q.small ? poisson_inverse(q.m, q.expM, uc) : poisson_normal(q.m, uc, draw(key, bin, 0, 1));
   }

   __device__ State init(uint32_t h1, uint32_t k, uint32_t n, const Consts& q) const
   {
      State s{0, 0, true, 0, 0, 0};
      if (!(kin.activationRatePerSec > 0))
         return s;
      uint32_t a = h1, b = k, c = n, d = CH_PERSIST;
      pcg4d(a, b, c, d);
      s.key = a;
      s.b = b0 > 0 ? b0 : 0;
      if (s.b > b1)
         return s;
      s.c = bin_count(s.key, static_cast<uint32_t>(s.b), q);
      s.done = false;
      return s;
   }

   // Blink j of bin b, or the next bin; true when past b1.
   template <bool WRITE>
   __device__ bool step(State& s, const Consts& q, Record* out) const
   {
      if (s.j >= s.c)
      {
         if (++s.b > b1)
            return true;
         s.c = bin_count(s.key, static_cast<uint32_t>(s.b), q);
         s.j = 0;
         return false;
      }
      const uint32_t bin = static_cast<uint32_t>(s.b), jj = static_cast<uint32_t>(s.j);
      ++s.j;
      const double tOn = (bin + draw(s.key, bin, jj, 2)) * kBinSec;
      const double on = fmin(q.maxOn, -jsm::log(draw(s.key, bin, jj, 3)) * kin.onSec);
      if (WINDOW && !(tOn < t1 && tOn + on > t0))
         return false;
      if constexpr (WRITE)
      {
         double br = 1;
         if (q.cv > 0)
            br = q.ln(draw(s.key, bin, jj, 4), draw(s.key, bin, jj, 5));
         if constexpr (WINDOW)
            out[s.count] = iscg_blink{tOn, tOn + on, br};
         else
            out[s.count] = iscg_bin_blink{tOn, tOn + on, br, bin, jj};
      }
      ++s.count;
      return false;
   }
};

__device__ __forceinline__ LabelGen::State gen_init(const LabelGen& g, const iscg_dye_address& d,
                                                    const LabelGen::Consts&)
{
   return g.init(d.h1, static_cast<uint32_t>(d.k), static_cast<uint32_t>(d.n));
}
template <bool W, class R>
__device__ __forceinline__ typename PersistGen<W, R>::State gen_init(const PersistGen<W, R>& g,
                                                                     const iscg_dye_address& d,
                                                                     const typename PersistGen<W, R>::Consts& q)
{
   return g.init(d.h1, static_cast<uint32_t>(d.k), static_cast<uint32_t>(d.n), q);
}

// Wave-wide inclusive prefix sum of v (64 lanes).
__device__ __forceinline__ long long wave_inclusive(long long v)
{
   const int lane = threadIdx.x & 63;
#pragma unroll
   for (int d = 1; d < 64; d <<= 1)
   {
      const long long o = __shfl_up(v, d, 64);
      if (lane >= d)
         v += o;
   }
   return v;
}

// Exclusive prefix of v over the workgroup (NW waves); total = the workgroup's sum.
template <int NW>
__device__ __forceinline__ long long block_exclusive(long long v, long long& total)
{
   __shared__ long long waveSum[NW];
   const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
   const long long inc = wave_inclusive(v);
   if (lane == 63)
      waveSum[wave] = inc;
   __syncthreads();
   long long before = 0;
   total = 0;
#pragma unroll
   for (int w = 0; w < NW; ++w)
   {
      const long long s = waveSum[w];
      before += w < wave ? s : 0;
      total += s;
   }
   __syncthreads();
   return before + inc - v;
}

// Dyes per wave and per workgroup: each wave runs the schedules of kChunk consecutive dyes. Finished lanes take new
// dyes once at least kRefill of them are idle (or the wave has no work), so the init branch runs for several lanes at
// a time instead of for one or two in most iterations.
#ifndef ISCG_DYE_CHUNK
#define ISCG_DYE_CHUNK 512
#endif
#ifndef ISCG_DYE_REFILL
#define ISCG_DYE_REFILL 8
#endif
constexpr int kWaves = kDyeThreads / 64;
constexpr int kChunk = ISCG_DYE_CHUNK;
constexpr int kRefill = ISCG_DYE_REFILL;
constexpr int kWgDyes = kWaves * kChunk;

// The schedules of dyes [begin, end) on one wave. Count pass (WRITE false): counts[i] per dye, returns the lane's
// sum. Fill pass: dye i's records at out + base[i - wgBegin] when they fit the capacity (counts from the count pass).
// Every dye (the schedule calls' selector).
struct AllDyes
{
   __device__ __forceinline__ bool operator()(int) const { return true; }
};

// The schedules of dyes [begin, end) that sel accepts, on one wave (the counts of the others are left as they are).
template <bool WRITE, class Gen, class Sel = AllDyes>
__device__ long long wave_schedules(const Gen& g, const iscg_dye_address* __restrict__ dyes, int begin, int end,
                                    int32_t* __restrict__ counts, const long long* base, int wgBegin,
                                    typename Gen::Record* __restrict__ out, int64_t capacity, Sel sel = Sel())
{
   const typename Gen::Consts q = g.template consts<WRITE>();
   const int lane = threadIdx.x & 63;
   const unsigned long long below = (1ull << lane) - 1;
   typename Gen::State st{};
   typename Gen::Record* dst = nullptr;
   int cur = -1, next = begin;
   long long sum = 0;
   for (;;)
   {
      const unsigned long long idle = __ballot(cur < 0);
      const bool refill = __popcll(idle) >= kRefill || idle == ~0ull;
      const bool need = refill && cur < 0;
      const unsigned long long m = refill ? idle : 0ull;
      if (need)
      {
         const int i = next + __popcll(m & below);
         if (i < end)
         {
            bool work = sel(i);
            if constexpr (WRITE)
            {
               const long long off = base[i - wgBegin];
               const int c = counts[i];
               work = work && out != nullptr && c > 0 && off + c <= capacity;
               dst = out + off;
            }
            if (work)
            {
               st = gen_init(g, dyes[i], q);
               cur = i;
            }
         }
      }
      next += __popcll(m);
      if (__ballot(cur >= 0) == 0 && next >= end)
         break;
      if (cur >= 0 && (st.done || g.template step<WRITE>(st, q, dst)))
      {
         if constexpr (!WRITE)
         {
            counts[cur] = st.count;
            sum += st.count;
         }
         cur = -1;
      }
   }
   return sum;
}

}  // namespace dyegen
}  // namespace iscg
