// dyes.hip -- dye schedules on the device: LabelSchedule's blinks and continuous windows and the persistent
// (DNA-PAINT) blinks of a batch of dye addresses (core/src/dyes.cpp:254-657, one set of kinetics), bit-identical to
// the reference's. Each value is a pure function of (h1, k, n): the draws are Unit(Pcg4d(h1, k, n, channel)) (the
// channels of core/src/dyes.h:45-57) and the FP64 arithmetic is the reference's, in its order, with the jsm::
// functions restated in jsm.h.
//
// One lane per dye. A schedule call is three launches on the context's stream:
//   count   every lane runs its dye's schedule without the brightness draws and stores the record count;
//           each workgroup's total goes to dyeBlocks;
//   scan    one workgroup turns dyeBlocks into exclusive prefixes and writes the total;
//   fill    each workgroup scans its counts in LDS, writes the offsets, and every lane whose records fit the
//           capacity runs its schedule again and writes them.
// The work is FP64 VALU (a log per draw, a log, cos and exp per brightness) and the output is written once.
#include "iscg_internal.h"
#include "jsm.h"

#include <cmath>

using namespace iscg;

namespace {

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
   __device__ __forceinline__ double draw(uint32_t ch) const
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
      return CounterRng::unit_of(x);
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
      if constexpr (WRITE)
      {
         double b = 1;
         if (c.cv > 0)
            b = c.ln(s.at.draw(base + CH_BRIGHT1), s.at.draw(base + CH_BRIGHT2));
         // LabelSchedule adds the dSTORM shift to both ends after the schedule is made
         out[s.count] = s.shift != 0 ? iscg_blink{s.t + s.shift, (s.t + on) + s.shift, b}
                                     : iscg_blink{s.t, s.t + on, b};
      }
      ++s.count;
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
template <bool WRITE, class Gen>
__device__ long long wave_schedules(const Gen& g, const iscg_dye_address* __restrict__ dyes, int begin, int end,
                                    int32_t* __restrict__ counts, const long long* base, int wgBegin,
                                    typename Gen::Record* __restrict__ out, int64_t capacity)
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
            bool work = true;
            if constexpr (WRITE)
            {
               const long long off = base[i - wgBegin];
               const int c = counts[i];
               work = out != nullptr && c > 0 && off + c <= capacity;
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

template <class Gen>
__global__ __launch_bounds__(kDyeThreads) void dye_count_kernel(Gen g, const iscg_dye_address* __restrict__ dyes,
                                                                 int32_t nDyes, int32_t* __restrict__ counts,
                                                                 long long* __restrict__ blockSums)
{
   const int wgBegin = blockIdx.x * kWgDyes;
   const int begin = wgBegin + (threadIdx.x >> 6) * kChunk;
   const int end = min(begin + kChunk, nDyes);
   const long long mine = wave_schedules<false>(g, dyes, begin, end, counts, nullptr, wgBegin, nullptr, 0);
   long long total;
   (void)block_exclusive<kWaves>(mine, total);
   if (threadIdx.x == 0)
      blockSums[blockIdx.x] = total;
}

// One workgroup: blockSums[0..nBlocks) -> exclusive prefixes; offsets[nDyes] and *total = the sum.
__global__ __launch_bounds__(kScanThreads) void dye_scan_kernel(long long* __restrict__ blockSums, int32_t nBlocks,
                                                                 int64_t* __restrict__ offsets, int32_t nDyes,
                                                                 int64_t* __restrict__ total)
{
   long long carry = 0;
   for (int base = 0; base < nBlocks; base += kScanThreads)
   {
      const int i = base + threadIdx.x;
      const long long v = i < nBlocks ? blockSums[i] : 0;
      long long chunk;
      const long long ex = block_exclusive<kScanThreads / 64>(v, chunk);
      if (i < nBlocks)
         blockSums[i] = carry + ex;
      carry += chunk;
   }
   if (threadIdx.x == 0)
   {
      offsets[nDyes] = carry;
      *total = carry;
   }
}

template <class Gen>
__global__ __launch_bounds__(kDyeThreads) void dye_fill_kernel(Gen g, const iscg_dye_address* __restrict__ dyes,
                                                                int32_t nDyes, int32_t* __restrict__ counts,
                                                                const long long* __restrict__ blockBase,
                                                                int64_t* __restrict__ offsets,
                                                                typename Gen::Record* __restrict__ out,
                                                                int64_t capacity)
{
   // the workgroup's offsets: each lane sums kPer consecutive counts, one workgroup scan, then per dye into LDS
   constexpr int kPer = kWgDyes / kDyeThreads;
   __shared__ long long base[kWgDyes];
   const int wgBegin = blockIdx.x * kWgDyes;
   const int first = wgBegin + threadIdx.x * kPer;
   int c[kPer];
   long long mine = 0;
#pragma unroll
   for (int r = 0; r < kPer; ++r)
   {
      c[r] = first + r < nDyes ? counts[first + r] : 0;
      mine += c[r];
   }
   long long total;
   long long off = blockBase[blockIdx.x] + block_exclusive<kWaves>(mine, total);
#pragma unroll
   for (int r = 0; r < kPer; ++r)
   {
      base[threadIdx.x * kPer + r] = off;
      if (first + r < nDyes)
         offsets[first + r] = off;
      off += c[r];
   }
   __syncthreads();
   const int begin = wgBegin + (threadIdx.x >> 6) * kChunk;
   const int end = min(begin + kChunk, nDyes);
   (void)wave_schedules<true>(g, dyes, begin, end, counts, base, wgBegin, out, capacity);
}

// LabelSchedule's continuous windows (core/src/dyes.cpp:447-477), one set of kinetics: at most one per dye.
__global__ __launch_bounds__(kDyeThreads) void label_cont_kernel(const iscg_dye_address* __restrict__ dyes,
                                                                  int32_t nDyes, int32_t mode, int32_t preState,
                                                                  iscg_kinetics kin, iscg_cont_window* __restrict__ out,
                                                                  uint8_t* __restrict__ has)
{
   const int i = blockIdx.x * kDyeThreads + threadIdx.x;
   if (i >= nDyes)
      return;
   const iscg_dye_address d = dyes[i];
   const uint32_t h1 = d.h1, k = static_cast<uint32_t>(d.k), n = static_cast<uint32_t>(d.n);
   uint8_t one = 0;
   iscg_cont_window w{};
   if (mode == 3)
   {
      w = iscg_cont_window{0.0, __builtin_inf(), 3, -jsm::log(draw(h1, k, n, CH_AUX))};
      one = 1;
   }
   else if (mode == 0 && kin.initialOnSec > 0)
   {
      const double shift = -jsm::log(draw(h1, k, n, CH_INIT_ON)) * kin.initialOnSec;
      if (shift != 0)
      {
         w = iscg_cont_window{0.0, shift, 2, 0.0};
         one = 1;
      }
   }
   else if (mode == 1 && preState)
   {
      const double tOff = kin.activationRatePerSec > 0
                             ? -jsm::log(draw(h1, k, n, CH_ACT)) / kin.activationRatePerSec
                             : __builtin_inf();
      w = iscg_cont_window{0.0, tOff, 1, -jsm::log(draw(h1, k, n, CH_AUX))};
      one = 1;
   }
   if (one)
      out[i] = w;
   has[i] = one;
}

__global__ __launch_bounds__(256) void jsm_kernel(int32_t fn, const double* __restrict__ x, int64_t n,
                                                  double* __restrict__ out)
{
   for (int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
   {
      const double v = x[i];
      out[i] = fn == 0 ? jsm::sin(v) : fn == 1 ? jsm::cos(v) : fn == 2 ? jsm::exp(v) : jsm::log(v);
   }
}

// The three launches of a schedule call (count, scan, fill) with the common argument handling.
template <class Gen>
int32_t run_schedule(iscg_context* h, const char* what, const Gen& g, const iscg_dye_address* dyes, int32_t nDyes,
                     int64_t* offsets, typename Gen::Record* out, int64_t capacity, int64_t* total)
{
   if (nDyes < 0 || capacity < 0 || !offsets || !total || (capacity > 0 && !out))
      return fail(h, std::string(what) + ": bad argument");
   const size_t nOff = (size_t)nDyes + 1;
   const iscg_dye_address* dd = nDyes ? (const iscg_dye_address*)stage_in(h, 0, dyes, (size_t)nDyes * sizeof(*dyes))
                                      : nullptr;
   int64_t* doff = (int64_t*)stage_out(h, 1, offsets, nOff * sizeof(int64_t));
   int64_t* dtot = (int64_t*)stage_out(h, 2, total, sizeof(int64_t));
   // a host out is staged with its contents, so the records of dyes past the capacity come back as they were
   typename Gen::Record* dout =
      capacity > 0 ? (typename Gen::Record*)stage_in(h, 3, out, (size_t)capacity * sizeof(*out)) : nullptr;
   if ((nDyes && !dd) || !doff || !dtot || (capacity > 0 && !dout))
      return -1;
   const int nBlocks = nDyes ? (nDyes + kWgDyes - 1) / kWgDyes : 0;
   if (!ok(h, h->dyeCounts.reserve((size_t)(nDyes ? nDyes : 1) * sizeof(int32_t)), "dye counts") ||
       !ok(h, h->dyeBlocks.reserve((size_t)(nBlocks ? nBlocks : 1) * sizeof(long long)), "dye blocks"))
      return -1;
   int32_t* counts = h->dyeCounts.as<int32_t>();
   long long* blocks = h->dyeBlocks.as<long long>();
   if (nBlocks)
   {
      ProfScope ps(h, "dye_count");
      hipLaunchKernelGGL(dye_count_kernel<Gen>, dim3(nBlocks), dim3(kDyeThreads), 0, h->stream, g, dd, nDyes, counts,
                         blocks);
   }
   {
      ProfScope ps(h, "dye_scan");
      hipLaunchKernelGGL(dye_scan_kernel, dim3(1), dim3(kScanThreads), 0, h->stream, blocks, nBlocks, doff, nDyes,
                         dtot);
   }
   if (nBlocks)
   {
      ProfScope ps(h, "dye_fill");
      hipLaunchKernelGGL(dye_fill_kernel<Gen>, dim3(nBlocks), dim3(kDyeThreads), 0, h->stream, g, dd, nDyes, counts,
                         blocks, doff, dout, capacity);
   }
   if (!ok(h, hipGetLastError(), what))
      return -1;
   if (!finish_out(h, offsets, doff, nOff * sizeof(int64_t)) || !finish_out(h, total, dtot, sizeof(int64_t)))
      return -1;
   if (dout && !finish_out(h, out, dout, (size_t)capacity * sizeof(*out)))
      return -1;
   return 0;
}

bool bad_kinetics(const iscg_kinetics* kin) { return kin == nullptr; }

// Every lane's loop over bins and candidate blinks must end within a launch: at most kMaxBins bins per dye and a
// per-bin mean count of at most kMaxPerBin (the reference's callers ask for a few bins at a few blinks each).
constexpr long long kMaxBins = 1 << 20;
constexpr double kMaxPerBin = 1e5;

bool persist_bounded(iscg_context* h, const char* what, const iscg_kinetics* kin, long long b0, long long b1)
{
   const long long lo = b0 > 0 ? b0 : 0;
   if (b1 >= lo && b1 - lo >= kMaxBins)
   {
      fail(h, std::string(what) + ": more than 2^20 time bins per dye");
      return false;
   }
   if (kin->activationRatePerSec * kBinSec > kMaxPerBin)
   {
      fail(h, std::string(what) + ": activation rate above 1e5 blinks per bin");
      return false;
   }
   return true;
}

}  // namespace

extern "C" {

int32_t iscg_label_blinks(iscg_context* h, const iscg_dye_address* dyes, int32_t nDyes, int32_t mode,
                          const iscg_kinetics* kin, double tMax, int64_t* offsets, iscg_blink* out, int64_t capacity,
                          int64_t* total)
{
   if (begin(h))
      return -1;
   if (bad_kinetics(kin) || mode < 0 || mode > 3 || (nDyes > 0 && !dyes))
      return fail(h, "iscg_label_blinks: bad argument");
   const LabelGen g{mode, *kin, tMax};
   return run_schedule(h, "iscg_label_blinks", g, dyes, nDyes, offsets, out, capacity, total);
}

int32_t iscg_persistent_blinks(iscg_context* h, const iscg_dye_address* dyes, int32_t nDyes, const iscg_kinetics* kin,
                               double t0, double t1, int64_t* offsets, iscg_blink* out, int64_t capacity,
                               int64_t* total)
{
   if (begin(h))
      return -1;
   if (bad_kinetics(kin) || (nDyes > 0 && !dyes))
      return fail(h, "iscg_persistent_blinks: bad argument");
   // PersistentBlinks (core/src/dyes.cpp:636-646): no window, no blinks; bins from the lookback of the longest ON.
   long long b0 = 0, b1 = -1;
   if (t1 > t0)
   {
      const double maxOn = kOnCap * kin->onSec;
      const long long f = (long long)std::floor((t0 - maxOn) / kBinSec);
      b0 = f > 0 ? f : 0;
      b1 = (long long)std::floor(t1 / kBinSec);
   }
   if (!persist_bounded(h, "iscg_persistent_blinks", kin, b0, b1))
      return -1;
   const PersistGen<true, iscg_blink> g{*kin, t0, t1, b0, b1};
   return run_schedule(h, "iscg_persistent_blinks", g, dyes, nDyes, offsets, out, capacity, total);
}

int32_t iscg_persistent_bins(iscg_context* h, const iscg_dye_address* dyes, int32_t nDyes, const iscg_kinetics* kin,
                             int64_t binLo, int64_t binHi, int64_t* offsets, iscg_bin_blink* out, int64_t capacity,
                             int64_t* total)
{
   if (begin(h))
      return -1;
   if (bad_kinetics(kin) || (nDyes > 0 && !dyes))
      return fail(h, "iscg_persistent_bins: bad argument");
   if (!persist_bounded(h, "iscg_persistent_bins", kin, binLo, binHi))
      return -1;
   const PersistGen<false, iscg_bin_blink> g{*kin, 0.0, 0.0, (long long)binLo, (long long)binHi};
   return run_schedule(h, "iscg_persistent_bins", g, dyes, nDyes, offsets, out, capacity, total);
}

int32_t iscg_label_cont(iscg_context* h, const iscg_dye_address* dyes, int32_t nDyes, int32_t mode, int32_t preState,
                        const iscg_kinetics* kin, iscg_cont_window* out, uint8_t* has)
{
   if (begin(h))
      return -1;
   if (bad_kinetics(kin) || mode < 0 || mode > 3 || nDyes < 0 || (nDyes > 0 && (!dyes || !out || !has)))
      return fail(h, "iscg_label_cont: bad argument");
   if (nDyes == 0)
      return 0;
   const iscg_dye_address* dd = (const iscg_dye_address*)stage_in(h, 0, dyes, (size_t)nDyes * sizeof(*dyes));
   // out[i] is left as it was for dyes without a window: a host out is staged with its current contents
   iscg_cont_window* dout = (iscg_cont_window*)stage_in(h, 1, out, (size_t)nDyes * sizeof(*out));
   uint8_t* dhas = (uint8_t*)stage_out(h, 2, has, (size_t)nDyes);
   if (!dd || !dout || !dhas)
      return -1;
   {
      ProfScope ps(h, "dye_cont");
      hipLaunchKernelGGL(label_cont_kernel, dim3((nDyes + kDyeThreads - 1) / kDyeThreads), dim3(kDyeThreads), 0,
                         h->stream, dd, nDyes, mode, preState, *kin, dout, dhas);
   }
   if (!ok(h, hipGetLastError(), "iscg_label_cont launch"))
      return -1;
   if (!finish_out(h, out, dout, (size_t)nDyes * sizeof(*out)) || !finish_out(h, has, dhas, (size_t)nDyes))
      return -1;
   return 0;
}

int32_t iscg_jsm(iscg_context* h, int32_t fn, const double* x, int64_t n, double* out)
{
   if (begin(h))
      return -1;
   if (fn < 0 || fn > 3 || n < 0 || (n > 0 && (!x || !out)))
      return fail(h, "iscg_jsm: bad argument");
   if (n == 0)
      return 0;
   const double* dx = (const double*)stage_in(h, 0, x, (size_t)n * sizeof(double));
   double* dout = (double*)stage_out(h, 1, out, (size_t)n * sizeof(double));
   if (!dx || !dout)
      return -1;
   hipLaunchKernelGGL(jsm_kernel, dim3(grid_for((size_t)n, 256, 8)), dim3(256), 0, h->stream, fn, dx, n, dout);
   if (!ok(h, hipGetLastError(), "iscg_jsm launch"))
      return -1;
   return finish_out(h, out, dout, (size_t)n * sizeof(double)) ? 0 : -1;
}

}  // extern "C"
