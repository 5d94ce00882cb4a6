// dyes.hip -- dye schedules on the device: LabelSchedule's blinks and continuous windows and the persistent
// (DNA-PAINT) blinks of a batch of dye addresses (core/src/dyes.cpp:254-657, one set of kinetics or a
// KineticsHistory), bit-identical to the reference's. Each value is a pure function of (h1, k, n): the draws are Unit(Pcg4d(h1, k, n, channel)) (the
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
#include "dye_gen.h"

#include <algorithm>
#include <cmath>
#include <vector>

using namespace iscg;
using namespace iscg::dyegen;

namespace {

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

// LabelSchedule's continuous windows (core/src/dyes.cpp:447-477): at most one per dye.
__global__ __launch_bounds__(kDyeThreads) void label_cont_kernel(const iscg_dye_address* __restrict__ dyes,
                                                                  int32_t nDyes, int32_t mode, int32_t preState,
                                                                  iscg_kinetics kin, HistDev hist,
                                                                  iscg_cont_window* __restrict__ out,
                                                                  uint8_t* __restrict__ has)
{
   const int i = blockIdx.x * kDyeThreads + threadIdx.x;
   if (i >= nDyes)
      return;
   const iscg_dye_address d = dyes[i];
   iscg_cont_window w{};
   const bool one = cont_window(mode, preState, kin, hist,
                                Address::of(d.h1, static_cast<uint32_t>(d.k), static_cast<uint32_t>(d.n)), w);
   if (one)
      out[i] = w;
   has[i] = one;
}

__global__ __launch_bounds__(256) void jsm_kernel(int32_t fn, const double* __restrict__ x, int64_t n,
                                                  double* __restrict__ out)
{
   for (int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
   {
      if (fn == 5)
      {
         out[i] = jsm::hypot(x[2 * i], x[2 * i + 1]);
         continue;
      }
      if (fn == 6)
      {
         out[i] = jsm::hypot(x[3 * i], x[3 * i + 1], x[3 * i + 2]);
         continue;
      }
      const double v = x[i];
      out[i] = fn == 0   ? jsm::sin(v)
               : fn == 1 ? jsm::cos(v)
               : fn == 2 ? jsm::exp(v)
               : fn == 3 ? jsm::log(v)
                         : jsm::cbrt(v);
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

constexpr int32_t kMaxSegments = 1 << 16;

// A history argument as LabelSchedule / PersistentGen read theirs: none (nullptr or n = 0) keeps kin; one segment
// replaces kin by its kinetics; two or more are packed into stage[4] (dev), the stream synchronized so the packing
// may go. False (and the error set) for a malformed history.
bool take_history(iscg_context* h, const char* what, const iscg_kinetics_history* hist, iscg_kinetics& kin,
                  HistDev& dev)
{
   dev = HistDev{0, nullptr};
   if (!hist || hist->n == 0)
      return true;
   if (hist->n < 0 || hist->n > kMaxSegments || !hist->tStart || !hist->kin)
   {
      fail(h, std::string(what) + ": bad kinetics history");
      return false;
   }
   for (int32_t i = 0; i < hist->n; ++i)
      if (!std::isfinite(hist->tStart[i]) || (i == 0 ? hist->tStart[0] != 0 : !(hist->tStart[i] > hist->tStart[i - 1])))
      {
         fail(h, std::string(what) + ": a history starts at 0 and its segment starts increase");
         return false;
      }
   if (hist->n == 1)
   {
      kin = hist->kin[0];
      return true;
   }
   std::vector<HistSeg> segs((size_t)hist->n);
   for (int32_t i = 0; i < hist->n; ++i)
      segs[(size_t)i] = HistSeg{hist->tStart[i], hist->kin[i]};
   const void* p = stage_in(h, 4, segs.data(), segs.size() * sizeof(HistSeg));
   if (!p || !ok(h, hipStreamSynchronize(h->stream), what))
      return false;
   dev = HistDev{hist->n, static_cast<const HistSeg*>(p)};
   return true;
}

// The largest activation rate and ON time of kin or of the history's segments (PersistentMaxOn, core/src/dyes.cpp:628).
void persist_extremes(const iscg_kinetics& kin, const iscg_kinetics_history* hist, double& rate, double& onSec)
{
   rate = kin.activationRatePerSec;
   onSec = kin.onSec;
   if (!hist || hist->n <= 0)
      return;
   rate = 0;
   onSec = 0;
   for (int32_t i = 0; i < hist->n; ++i)
   {
      rate = std::max(rate, hist->kin[i].activationRatePerSec);
      onSec = std::max(onSec, hist->kin[i].onSec);
   }
}

// Every lane's loop over bins and candidate blinks must end within a launch: at most kMaxBins bins per dye and a
// per-bin mean count of at most kMaxPerBin (the reference's callers ask for a few bins at a few blinks each).
constexpr long long kMaxBins = 1 << 20;
constexpr double kMaxPerBin = 1e5;

bool persist_bounded(iscg_context* h, const char* what, double rate, long long b0, long long b1)
{
   const long long lo = b0 > 0 ? b0 : 0;
   if (b1 >= lo && b1 - lo >= kMaxBins)
   {
      fail(h, std::string(what) + ": more than 2^20 time bins per dye");
      return false;
   }
   if (!(rate * kBinSec <= kMaxPerBin) && rate > 0)
   {
      fail(h, std::string(what) + ": activation rate above 1e5 blinks per bin");
      return false;
   }
   return true;
}

}  // namespace

extern "C" {

int32_t iscg_label_blinks(iscg_context* h, const iscg_dye_address* dyes, int32_t nDyes, int32_t mode,
                          const iscg_kinetics* kin, const iscg_kinetics_history* hist, double tMax, int64_t* offsets,
                          iscg_blink* out, int64_t capacity, int64_t* total)
{
   if (begin(h))
      return -1;
   if (bad_kinetics(kin) || mode < 0 || mode > 3 || (nDyes > 0 && !dyes))
      return fail(h, "iscg_label_blinks: bad argument");
   iscg_kinetics k = *kin;
   HistDev hd;
   if (!take_history(h, "iscg_label_blinks", hist, k, hd))
      return -1;
   if (hd.n >= 2)
   {
      const LabelHistGen g{mode, hd, tMax, -__builtin_inf(), __builtin_inf()};
      return run_schedule(h, "iscg_label_blinks", g, dyes, nDyes, offsets, out, capacity, total);
   }
   const LabelGen g{mode, k, tMax, -__builtin_inf(), __builtin_inf()};
   return run_schedule(h, "iscg_label_blinks", g, dyes, nDyes, offsets, out, capacity, total);
}

int32_t iscg_persistent_blinks(iscg_context* h, const iscg_dye_address* dyes, int32_t nDyes, const iscg_kinetics* kin,
                               const iscg_kinetics_history* hist, double t0, double t1, int64_t* offsets,
                               iscg_blink* out, int64_t capacity, int64_t* total)
{
   if (begin(h))
      return -1;
   if (bad_kinetics(kin) || (nDyes > 0 && !dyes))
      return fail(h, "iscg_persistent_blinks: bad argument");
   iscg_kinetics k = *kin;
   HistDev hd;
   if (!take_history(h, "iscg_persistent_blinks", hist, k, hd))
      return -1;
   double rate, onSec;
   persist_extremes(*kin, hist, rate, onSec);
   // PersistentBlinks (core/src/dyes.cpp:636-646): no window, no blinks; bins from the lookback of the longest ON.
   long long b0 = 0, b1 = -1;
   if (t1 > t0)
   {
      const double maxOn = kOnCap * onSec;
      const long long f = (long long)std::floor((t0 - maxOn) / kBinSec);
      b0 = f > 0 ? f : 0;
      b1 = (long long)std::floor(t1 / kBinSec);
   }
   if (!persist_bounded(h, "iscg_persistent_blinks", rate, b0, b1))
      return -1;
   if (hd.n >= 2)
   {
      const PersistHistGen<true, iscg_blink> g{hd, t0, t1, b0, b1};
      return run_schedule(h, "iscg_persistent_blinks", g, dyes, nDyes, offsets, out, capacity, total);
   }
   const PersistGen<true, iscg_blink> g{k, t0, t1, b0, b1};
   return run_schedule(h, "iscg_persistent_blinks", g, dyes, nDyes, offsets, out, capacity, total);
}

int32_t iscg_persistent_bins(iscg_context* h, const iscg_dye_address* dyes, int32_t nDyes, const iscg_kinetics* kin,
                             const iscg_kinetics_history* hist, int64_t binLo, int64_t binHi, int64_t* offsets,
                             iscg_bin_blink* out, int64_t capacity, int64_t* total)
{
   if (begin(h))
      return -1;
   if (bad_kinetics(kin) || (nDyes > 0 && !dyes))
      return fail(h, "iscg_persistent_bins: bad argument");
   iscg_kinetics k = *kin;
   HistDev hd;
   if (!take_history(h, "iscg_persistent_bins", hist, k, hd))
      return -1;
   double rate, onSec;
   persist_extremes(*kin, hist, rate, onSec);
   if (!persist_bounded(h, "iscg_persistent_bins", rate, binLo, binHi))
      return -1;
   if (hd.n >= 2)
   {
      const PersistHistGen<false, iscg_bin_blink> g{hd, 0.0, 0.0, (long long)binLo, (long long)binHi};
      return run_schedule(h, "iscg_persistent_bins", g, dyes, nDyes, offsets, out, capacity, total);
   }
   const PersistGen<false, iscg_bin_blink> g{k, 0.0, 0.0, (long long)binLo, (long long)binHi};
   return run_schedule(h, "iscg_persistent_bins", g, dyes, nDyes, offsets, out, capacity, total);
}

int32_t iscg_label_cont(iscg_context* h, const iscg_dye_address* dyes, int32_t nDyes, int32_t mode, int32_t preState,
                        const iscg_kinetics* kin, const iscg_kinetics_history* hist, iscg_cont_window* out,
                        uint8_t* has)
{
   if (begin(h))
      return -1;
   if (bad_kinetics(kin) || mode < 0 || mode > 3 || nDyes < 0 || (nDyes > 0 && (!dyes || !out || !has)))
      return fail(h, "iscg_label_cont: bad argument");
   iscg_kinetics k = *kin;
   HistDev hd;
   if (!take_history(h, "iscg_label_cont", hist, k, hd))
      return -1;
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
                         h->stream, dd, nDyes, mode, preState, k, hd, dout, dhas);
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
   if (fn < 0 || fn > 6 || n < 0 || (n > 0 && (!x || !out)))
      return fail(h, "iscg_jsm: bad argument");
   if (n == 0)
      return 0;
   const size_t per = fn == 5 ? 2 : fn == 6 ? 3 : 1;
   const double* dx = (const double*)stage_in(h, 0, x, (size_t)n * per * sizeof(double));
   double* dout = (double*)stage_out(h, 1, out, (size_t)n * sizeof(double));
   if (!dx || !dout)
      return -1;
   hipLaunchKernelGGL(jsm_kernel, dim3(grid_for((size_t)n, 256, 8)), dim3(256), 0, h->stream, fn, dx, n, dout);
   if (!ok(h, hipGetLastError(), "iscg_jsm launch"))
      return -1;
   return finish_out(h, out, dout, (size_t)n * sizeof(double)) ? 0 : -1;
}

}  // extern "C"
