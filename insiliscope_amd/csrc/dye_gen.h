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
   __device__ __forceinline__ double operator()(double u1, double u2) const { return of(sigma, mu, u1, u2); }
   __device__ __forceinline__ static double of(double sigma, double mu, double u1, double u2)
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
   double tHi;  // the schedule stops at its first blink starting at or after tHi (+inf: at tMax only)
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
      if (tOn >= tHi)  // the blinks start in order, so none after this one is wanted either
         return true;
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

// ---- kinetics histories (core/src/dyes.h:138-149, core/src/dyes.cpp:323-436, :559-626) ----------------------------
// A dye clock split into segments of constant kinetics: segment i holds from its tStart (the first at 0) to the next
// one's, the last for ever. Each state of a schedule spends its unit-exponential draw as the integral of that state's
// rate over the segments it crosses, so a change of kinetics acts from its segment on and everything before it stays
// as it was. A history of one segment is that segment's kinetics from t = 0: the callers pass it to the generators
// above, and the generators below take two or more segments.

// One segment: its start on the dye clock and its kinetics.
struct HistSeg
{
   double tStart;
   iscg_kinetics kin;
};

struct HistDev
{
   int32_t n;  // 0: no history
   const HistSeg* seg;

   // KineticsHistory::SegmentAt (core/src/dyes.cpp:323-331): the last segment starting at or before t (0 before it)
   __device__ int segment_at(double t) const
   {
      int lo = 0, hi = n;
      while (hi - lo > 1)
      {
         const int mid = (lo + hi) / 2;
         if (seg[mid].tStart <= t)
            lo = mid;
         else
            hi = mid;
      }
      return lo;
   }
};

__device__ __forceinline__ double max0(double x) { return 0.0 < x ? x : 0.0; }  // std::max(0.0, x)

// The scales a walk can spend a draw over: ON, dark and dSTORM initial-ON durations.
enum : int
{
   kScaleOn = 0,
   kScaleOff = 1,
   kScaleInitialOn = 2
};
template <int F>
__device__ __forceinline__ double scale_of(const iscg_kinetics& q)
{
   return F == kScaleOn ? q.onSec : F == kScaleOff ? q.offSec : q.initialOnSec;
}

// HazardWalk (core/src/dyes.cpp:336-375): draws spent over the segments in a dye's local time (absolute time - t0).
// i is the segment of the last time synced to; the times a schedule syncs to only increase.
struct HazardWalk
{
   const HistSeg* seg;
   int32_t n;
   double t0;
   int i;

   __device__ double end(int s) const { return s + 1 < n ? seg[s + 1].tStart - t0 : __builtin_inf(); }
   __device__ int sync(double tau)
   {
      while (i + 1 < n && seg[i + 1].tStart - t0 <= tau)
         ++i;
      return i;
   }
   // When the activation rate integrated from tau reaches E (+inf if never).
   __device__ double after_rate(double tau, double E)
   {
      for (int s = sync(tau);; ++s)
      {
         const double a = seg[s].kin.activationRatePerSec, e = end(s);
         if (a > 0)
         {
            const double t = tau + E / a;
            if (!(t >= e))
               return t;
            E = max0(E - a * (e - tau));
         }
         else if (e == __builtin_inf())
         {
            return __builtin_inf();
         }
         tau = e;
      }
   }
   // The same for a state of mean duration scale_of<F> per segment; a scale <= 0 leaves at once.
   template <int F>
   __device__ double after_scale(double tau, double E)
   {
      for (int s = sync(tau);; ++s)
      {
         const double sc = scale_of<F>(seg[s].kin), e = end(s);
         if (!(sc > 0))
            return tau;
         const double t = tau + E * sc;
         if (!(t >= e))
            return t;
         E = max0(E - (e - tau) / sc);
         tau = e;
      }
   }
};

__device__ __forceinline__ HazardWalk walk_of(const HistDev& h, double t0, int i = 0)
{
   return HazardWalk{h.seg, h.n, t0, i};
}

// LabelSchedule's blinks under two or more segments (core/src/dyes.cpp:440-471 with DyeScheduleHistory /
// DyeScheduleHistoryFrom, :381-426): the dSTORM initial ON spends its draw over 1 / initialOnSec per segment (there is
// one iff segment 0 has initialOnSec > 0) and the blinks run in local time from its end; the first activation, each
// ON and each dark state spend their draws over the segments they cross, and a cycle that stays in its segment takes
// the single-segment sums. The bleach probability is that of the segment the ON ended in.
struct LabelHistGen
{
   int32_t mode;
   HistDev hist;
   double tMax, tLo, tHi;  // as LabelGen's
   using Record = iscg_blink;

   struct Consts
   {
   };
   struct State
   {
      Address at;
      int j, count, seg;
      bool done;
      double t, shift, tStop;
      double lnCv, sigma, mu;  // LogNormalFor of the last CV used (a pure function: recomputing gives the same)
   };

   template <bool WRITE>
   __device__ Consts consts() const
   {
      return Consts{};
   }

   __device__ State init(uint32_t h1, uint32_t k, uint32_t n) const
   {
      State s{Address::of(h1, k, n), 0, 0, 0, true, 0.0, 0.0, 0.0, -1.0, 0.0, 0.0};
      if (mode == 2 || mode == 3)
         return s;
      if (mode == 0 && hist.seg[0].kin.initialOnSec > 0)
         s.shift = walk_of(hist, 0.0).after_scale<kScaleInitialOn>(0.0, -jsm::log(s.at.draw(CH_INIT_ON)));
      s.tStop = tMax - s.shift;
      s.t = walk_of(hist, s.shift).after_rate(0.0, -jsm::log(s.at.draw(CH_ACT)));
      s.done = false;
      return s;
   }

   // Blink j (core/src/dyes.cpp:395-425); true when the schedule has ended.
   template <bool WRITE>
   __device__ bool step(State& s, const Consts&, Record* out) const
   {
      if (s.j >= kMaxBlinks || s.t >= s.tStop)
         return true;
      const uint32_t base = CH_SCHED0 + static_cast<uint32_t>(s.j) * CH_SCHED_STRIDE;
      HazardWalk w = walk_of(hist, s.shift, s.seg);
      const int sg = w.sync(s.t);
      const iscg_kinetics q = hist.seg[sg].kin;
      const double end = w.end(sg);
      const double eOn = -jsm::log(s.at.draw(base + CH_ON));
      const double on = eOn * q.onSec;
      const bool onInSeg = !(s.t + on >= end);
      const double tOff = onInSeg ? s.t + on : w.after_scale<kScaleOn>(s.t, eOn);
      const double tOnAbs = s.shift != 0 ? s.t + s.shift : s.t;
      if (tOnAbs >= tHi)
         return true;
      const double tOffAbs = s.shift != 0 ? tOff + s.shift : tOff;
      if (tOffAbs > tLo)
      {
         if constexpr (WRITE)
         {
            double b = 1;
            const double cv = fmax(0.0, q.photonCV);
            if (cv > 0)
            {
               if (cv != s.lnCv)
               {
                  const LogNormal ln(cv);
                  s.sigma = ln.sigma;
                  s.mu = ln.mu;
                  s.lnCv = cv;
               }
               b = LogNormal::of(s.sigma, s.mu, s.at.draw(base + CH_BRIGHT1), s.at.draw(base + CH_BRIGHT2));
            }
            out[s.count] = iscg_blink{tOnAbs, tOffAbs, b};
         }
         ++s.count;
      }
      const double pBleach = fmin(1.0, fmax(0.01, hist.seg[w.sync(tOff)].kin.bleachProb));
      if (s.at.draw(base + CH_BLEACH) < pBleach)
         return true;
      const double eOff = -jsm::log(s.at.draw(base + CH_OFF));
      const double off = eOff * q.offSec;
      if (onInSeg && !(s.t + (on + off) >= end))
         s.t += on + off;  // the cycle stays in its segment: the single-segment sum
      else
         s.t = w.after_scale<kScaleOff>(tOff, eOff);
      s.seg = w.i;
      ++s.j;
      return false;
   }
};

// PersistentGen under two or more segments (core/src/dyes.cpp:559-626): bin b's own draws (piece 0) take the kinetics
// of the segment holding its start and keep the blinks starting before the next segment's start inside the bin; each
// later segment starting inside the bin (piece i) adds its own Poisson count at its rate over its part of the bin, on
// channels 8 i + purpose, its blinks numbered j | i << 24. The window form keeps those overlapping (t0, t1).
template <bool WINDOW, class RecordT>
struct PersistHistGen
{
   HistDev hist;
   double t0, t1;
   long long b0, b1;
   using Record = RecordT;

   struct Consts
   {
   };
   struct State
   {
      uint32_t key;
      int count, seg, s;  // seg: the segment at the bin's start; s: the current piece's (piece = s - seg)
      bool done;
      long long b, c, j;
      double lo, hi, binEnd;  // the current piece's span; piece 0's hi is where the next segment starts
   };

   template <bool WRITE>
   __device__ Consts consts() const
   {
      return Consts{};
   }

   __device__ long long count_of(double m, uint32_t bin, uint32_t o, uint32_t key) const
   {
      const double uc = draw(key, bin, 0, o);
      return m <= 30 ? poisson_inverse(m, jsm::exp(-m), uc) : poisson_normal(m, uc, draw(key, bin, 0, o + 1));
   }

   // Piece 0 of bin st.b.
   __device__ void start_bin(State& st) const
   {
      st.binEnd = (st.b + 1) * kBinSec;
      st.seg = hist.segment_at(st.b * kBinSec);
      st.s = st.seg;
      st.j = 0;
      st.c = 0;
      st.lo = st.b * kBinSec;
      st.hi = st.seg + 1 < hist.n && hist.seg[st.seg + 1].tStart < st.binEnd ? hist.seg[st.seg + 1].tStart
                                                                             : __builtin_inf();
      const double rate = hist.seg[st.seg].kin.activationRatePerSec;
      if (rate > 0)
         st.c = count_of(rate * kBinSec, static_cast<uint32_t>(st.b), 0, st.key);
   }

   // The next piece of the bin with a rate and a span (false: the bin has no more).
   __device__ bool next_piece(State& st) const
   {
      if (st.s == st.seg && !(st.hi < st.binEnd))
         return false;
      for (int s = st.s + 1; s < hist.n && hist.seg[s].tStart < st.binEnd; ++s)
      {
         const double a = hist.seg[s].tStart;
         const double e = s + 1 < hist.n && hist.seg[s + 1].tStart < st.binEnd ? hist.seg[s + 1].tStart : st.binEnd;
         const double rate = hist.seg[s].kin.activationRatePerSec;
         if (!(e > a) || !(rate > 0))
            continue;
         st.c = count_of(rate * (e - a), static_cast<uint32_t>(st.b), 8u * static_cast<uint32_t>(s - st.seg), st.key);
         st.s = s;
         st.j = 0;
         st.lo = a;
         st.hi = e;
         return true;
      }
      return false;
   }

   __device__ State init(uint32_t h1, uint32_t k, uint32_t n) const
   {
      State st{};
      st.done = true;
      uint32_t a = h1, b = k, c = n, d = CH_PERSIST;
      pcg4d(a, b, c, d);
      st.key = a;
      st.b = b0 > 0 ? b0 : 0;
      if (st.b > b1)
         return st;
      start_bin(st);
      st.done = false;
      return st;
   }

   // Candidate blink j of the current piece, or the next piece or bin; true when past b1.
   template <bool WRITE>
   __device__ bool step(State& st, const Consts&, Record* out) const
   {
      if (st.j >= st.c)
      {
         if (next_piece(st))
            return false;
         if (++st.b > b1)
            return true;
         start_bin(st);
         return false;
      }
      const uint32_t bin = static_cast<uint32_t>(st.b), jj = static_cast<uint32_t>(st.j);
      const uint32_t piece = static_cast<uint32_t>(st.s - st.seg), o = 8u * piece;
      ++st.j;
      const double u = draw(st.key, bin, jj, o + 2);
      const double tOn = piece == 0 ? (bin + u) * kBinSec : st.lo + u * (st.hi - st.lo);
      if (!(tOn < st.hi))
         return false;
      const iscg_kinetics& q = hist.seg[st.s].kin;
      const double on = fmin(kOnCap * q.onSec, -jsm::log(draw(st.key, bin, jj, o + 3)) * q.onSec);
      if (WINDOW && !(tOn < t1 && tOn + on > t0))
         return false;
      if constexpr (WRITE)
      {
         double br = 1;
         const double cv = fmax(0.0, q.photonCV);
         if (cv > 0)
            br = LogNormal(cv)(draw(st.key, bin, jj, o + 4), draw(st.key, bin, jj, o + 5));
         if constexpr (WINDOW)
            out[st.count] = iscg_blink{tOn, tOn + on, br};
         else
            out[st.count] = iscg_bin_blink{tOn, tOn + on, br, bin, jj | piece << 24};
      }
      ++st.count;
      return false;
   }
};

// LabelSchedule's continuous window of one dye (core/src/dyes.cpp:447-477), one set of kinetics or a history of two
// or more segments: WideField's ALWAYS_ON, the dSTORM initial ON, the PALM pre state until the first activation.
// False when the dye has none.
__device__ __forceinline__ bool cont_window(int32_t mode, int32_t preState, const iscg_kinetics& kin,
                                            const HistDev& hist, const Address& at, iscg_cont_window& w)
{
   const bool multi = hist.n >= 2;
   if (mode == 3)
   {
      w = iscg_cont_window{0.0, __builtin_inf(), 3, -jsm::log(at.draw(CH_AUX))};
      return true;
   }
   if (mode == 0)
   {
      double shift = 0.0;
      if (!multi && kin.initialOnSec > 0)
         shift = -jsm::log(at.draw(CH_INIT_ON)) * kin.initialOnSec;
      else if (multi && hist.seg[0].kin.initialOnSec > 0)
         shift = walk_of(hist, 0.0).after_scale<kScaleInitialOn>(0.0, -jsm::log(at.draw(CH_INIT_ON)));
      if (shift == 0)
         return false;
      w = iscg_cont_window{0.0, shift, 2, 0.0};
      return true;
   }
   if (mode == 1 && preState)
   {
      double tOff;
      if (multi)
         tOff = walk_of(hist, 0.0).after_rate(0.0, -jsm::log(at.draw(CH_ACT)));
      else
         tOff = kin.activationRatePerSec > 0 ? -jsm::log(at.draw(CH_ACT)) / kin.activationRatePerSec
                                             : __builtin_inf();
      w = iscg_cont_window{0.0, tOff, 1, -jsm::log(at.draw(CH_AUX))};
      return true;
   }
   return false;
}

__device__ __forceinline__ LabelHistGen::State gen_init(const LabelHistGen& g, const iscg_dye_address& d,
                                                        const LabelHistGen::Consts&)
{
   return g.init(d.h1, static_cast<uint32_t>(d.k), static_cast<uint32_t>(d.n));
}
template <bool W, class R>
__device__ __forceinline__ typename PersistHistGen<W, R>::State gen_init(const PersistHistGen<W, R>& g,
                                                                         const iscg_dye_address& d,
                                                                         const typename PersistHistGen<W, R>::Consts&)
{
   return g.init(d.h1, static_cast<uint32_t>(d.k), static_cast<uint32_t>(d.n));
}

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
