// collect.hip -- the frame emitters of a batch of frames on the device: CollectFrameEmitters
// (Sim/SMLMSimulation.cpp:55-101) for every frame of a batch, from blink events resident in HBM, into the records the
// splat pass takes (iscg_frame_emitter), so a movie goes from events to ADU without a host pass over the events.
//
// Per frame f (frame index F = frameIndex[f]) and event e of its bucket (the caller's per-frame event lists in event
// order, BucketEventsByFrame): overlap = min(F + 1, tEnd) - max(F, tStart), dropped when <= 0 and capped at 1; the
// position in pixels plus the frame's drift; photons = photonsPerBlink * overlap, times brightness and excite where
// they are not 1 (in that order), times the illumination field at the undrifted site (bilinear, IlluminationAt,
// Sim/SMLMBackground.cpp:39-51); the z plane NearestZIndex (Sim/PsfGeneratorBridge.cpp:40-60) of zNm / 1000 -
// globalZOffsetUm, with the clamped and total counts. The same FP64 operations in the same order, so the records
// equal the reference's bit for bit.
//
// Three launches: a count per frame (one workgroup per frame), an exclusive scan of the counts (one workgroup), and
// the fill, which re-walks each bucket and writes the kept events at their ranks (a ballot per 64 events and a
// prefix over the workgroup's waves keep them in bucket order).
#include "iscg_internal.h"

#include <algorithm>

#include <hipcub/hipcub.hpp>

namespace iscg {
namespace {

constexpr int kCollectThreads = 256;

struct CollectArgs
{
   const iscg_blink_event* ev;
   const uint32_t* bStart;
   const uint32_t* bIdx;
   const int64_t* frameIndex;
   const double* drift;   // [2 nFrames]: x, y per frame, or null (0)
   const float* illum;    // width x height, or null
   uint32_t nEvents, width, height;
   double pixelSizeNm, photonsPerBlink, zStepNm, globalZOffsetUm;
   int32_t nz, kernel;
   uint32_t nIdx;         // the bucket lists' length as validated on the host: every range is clamped to it
};

// overlap of event e with frame F (<= 0: not in it)
// (std::min(a, b) is b < a ? b : a and std::max(a, b) is a < b ? b : a: written so, NaN times behave the same)
__device__ __forceinline__ double overlap_of(const iscg_blink_event& e, int64_t F)
{
   const double f0 = (double)F, f1 = (double)(F + 1);
   return (e.tEnd < f1 ? e.tEnd : f1) - (f0 < e.tStart ? e.tStart : f0);
}

// IlluminationAt (Sim/SMLMBackground.cpp:39-51)
__device__ __forceinline__ double illum_at(const float* field, uint32_t w, uint32_t h, double x, double y)
{
   const double wm = (double)(w - 1), hm = (double)(h - 1);
   const double xm = x < wm ? x : wm, ym = y < hm ? y : hm;
   const double xi = 0.0 < xm ? xm : 0.0, yi = 0.0 < ym ? ym : 0.0;
   const double fx0 = floor(xi), fy0 = floor(yi);
   const uint32_t x0 = min(w - 1, (uint32_t)fx0), y0 = min(h - 1, (uint32_t)fy0);
   const uint32_t x1 = min(w - 1, x0 + 1), y1 = min(h - 1, y0 + 1);
   const double fx = xi - x0, fy = yi - y0;
   const double a = field[(size_t)y0 * w + x0], b = field[(size_t)y0 * w + x1];
   const double c = field[(size_t)y1 * w + x0], d = field[(size_t)y1 * w + x1];
   return (a * (1 - fx) + b * fx) * (1 - fy) + (c * (1 - fx) + d * fx) * fy;
}

// NearestZIndex (Sim/PsfGeneratorBridge.cpp:40-60): index lround(z 1000 / zStep + (nz - 1) / 2) as the reference's
// x86-64 build evaluates it (std::lround to long, then static_cast<int>: the low 32 bits; a value out of long's range
// or NaN gives LONG_MIN), clamped to [0, nz - 1].
__device__ __forceinline__ int nearest_z(double zUm, int nz, double zStepNm, bool& clamped)
{
   clamped = false;
   if (nz <= 1 || zStepNm <= 0.0)
      return 0;
   const double idxF = zUm * 1000.0 / zStepNm + (nz - 1) / 2.0;
   const long long l = fabs(idxF) < 9.2233720368547758e18 ? (long long)round(idxF) : (long long)0x8000000000000000ull;
   int idx = (int)(uint32_t)(unsigned long long)l;
   if (idx < 0)
   {
      idx = 0;
      clamped = true;
   }
   if (idx >= nz)
   {
      idx = nz - 1;
      clamped = true;
   }
   return idx;
}

__global__ __launch_bounds__(kCollectThreads) void collect_count_kernel(CollectArgs a, uint32_t* __restrict__ counts)
{
   __shared__ uint32_t part[kCollectThreads / 64];
   const int f = (int)blockIdx.x, tid = threadIdx.x;
   const uint32_t b1 = min(a.bStart[f + 1], a.nIdx), b0 = min(a.bStart[f], b1);
   const int64_t F = a.frameIndex[f];
   uint32_t n = 0;
   for (uint32_t i = b0 + tid; i < b1; i += kCollectThreads)
   {
      const uint32_t e = a.bIdx[i];
      if (e < a.nEvents && overlap_of(a.ev[e], F) > 0.0)
         ++n;
   }
   for (int o = 32; o > 0; o >>= 1)
      n += __shfl_down(n, o);
   if ((tid & 63) == 0)
      part[tid >> 6] = n;
   __syncthreads();
   if (tid == 0)
   {
      uint32_t t = 0;
      for (int w = 0; w < kCollectThreads / 64; ++w)
         t += part[w];
      counts[f] = t;
   }
}

// Exclusive scan of n counts into start[0..n] (start[n] = the total), one workgroup of 1024 lanes, 64-bit sums.
// zCounts (2 words) are zeroed here for the fill kernel after it: a kernel, not a memset, so the call can be captured
// in a HIP graph and replayed.
__global__ __launch_bounds__(1024) void collect_scan_kernel(const uint32_t* __restrict__ counts, uint32_t n,
                                                            unsigned long long* __restrict__ start,
                                                            unsigned long long* __restrict__ zCounts)
{
   __shared__ unsigned long long sums[1024];
   __shared__ unsigned long long carry;
   const int tid = threadIdx.x;
   if (tid < 2)
      zCounts[tid] = 0;
   if (tid == 0)
      carry = 0;
   __syncthreads();
   for (uint32_t c0 = 0; c0 < n; c0 += 1024)
   {
      const uint32_t i = c0 + tid;
      const unsigned long long v = i < n ? counts[i] : 0;
      sums[tid] = v;
      __syncthreads();
      for (int o = 1; o < 1024; o <<= 1)
      {
         const unsigned long long t = tid >= o ? sums[tid - o] : 0;
         __syncthreads();
         sums[tid] += t;
         __syncthreads();
      }
      if (i < n)
         start[i] = carry + sums[tid] - v;
      __syncthreads();
      if (tid == 1023)
         carry += sums[1023];
      __syncthreads();
   }
   if (tid == 0)
      start[n] = carry;
}

__global__ void collect_start32_kernel(const unsigned long long* __restrict__ start, uint32_t n,
                                       uint32_t* __restrict__ out)
{
   for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += gridDim.x * blockDim.x)
      out[i] = (uint32_t)start[i];
}

__global__ __launch_bounds__(kCollectThreads) void collect_fill_kernel(CollectArgs a,
                                                                       const unsigned long long* __restrict__ start,
                                                                       iscg_frame_emitter* __restrict__ out,
                                                                       unsigned long long* __restrict__ zCounts)
{
   __shared__ uint32_t waveKept[kCollectThreads / 64];
   const int f = (int)blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
   const uint32_t b1 = min(a.bStart[f + 1], a.nIdx), b0 = min(a.bStart[f], b1);
   const int64_t F = a.frameIndex[f];
   const double dx = a.drift ? a.drift[2 * f] : 0.0, dy = a.drift ? a.drift[2 * f + 1] : 0.0;
   unsigned long long base = start[f];
   uint32_t nClamped = 0, nTotal = 0;
   for (uint32_t c0 = b0; c0 < b1; c0 += kCollectThreads)
   {
      const uint32_t i = c0 + tid;
      bool keep = false;
      iscg_frame_emitter r{};
      if (i < b1)
      {
         const uint32_t ei = a.bIdx[i];
         if (ei < a.nEvents)
         {
            const iscg_blink_event e = a.ev[ei];
            double ov = overlap_of(e, F);
            if (ov > 0.0)
            {
               keep = true;
               if (ov > 1.0)
                  ov = 1.0;
               const double sx = e.xUm * 1000.0 / a.pixelSizeNm, sy = e.yUm * 1000.0 / a.pixelSizeNm;
               r.xPx = sx + dx;
               r.yPx = sy + dy;
               double ph = a.photonsPerBlink * ov;
               if (e.brightness != 1.0)
                  ph *= e.brightness;
               if (e.excite != 1.0)
                  ph *= e.excite;
               if (a.illum)
                  ph *= illum_at(a.illum, a.width, a.height, sx, sy);
               r.photons = ph;
               r.zIndex = 0;
               if (a.nz > 0)
               {
                  bool cl;
                  r.zIndex = nearest_z(e.zNm / 1000.0 - a.globalZOffsetUm, a.nz, a.zStepNm, cl);
                  ++nTotal;
                  nClamped += cl ? 1u : 0u;
               }
               r.kernel = a.kernel;
            }
         }
      }
      const unsigned long long mask = __ballot(keep);
      if (lane == 0)
         waveKept[wave] = (uint32_t)__popcll(mask);
      __syncthreads();
      unsigned long long before = 0, chunk = 0;
      for (int w = 0; w < kCollectThreads / 64; ++w)
      {
         before += w < wave ? waveKept[w] : 0;
         chunk += waveKept[w];
      }
      if (keep)
         out[base + before + (unsigned long long)__popcll(mask & ((1ull << lane) - 1ull))] = r;
      base += chunk;
      __syncthreads();
   }
   if (zCounts && (nClamped | nTotal))
   {
      atomicAdd(zCounts, (unsigned long long)nClamped);
      atomicAdd(zCounts + 1, (unsigned long long)nTotal);
   }
}


// ---- BucketEventsByFrame (Sim/SMLMSimulation.cpp:203-217) -------------------------------------------------------
// Event e belongs to frames [f0, f1] with f0 = max(0, floor(tStart)), f1 = min(nFrames - 1, floor(tEnd)) when
// tEnd > 0 and tStart < nFrames (floor then the conversion to long as the reference's x86-64 build evaluates it:
// out of range or NaN gives LONG_MIN). The (frame, event) pairs are written in event order and sorted by frame with a
// stable radix sort, so each frame's list keeps event order; the frame offsets come from the sorted keys.
__device__ __forceinline__ long long floor_to_long(double t)
{
   const double f = floor(t);
   return (f >= -9.2233720368547758e18 && f < 9.2233720368547758e18) ? (long long)f
                                                                     : (long long)0x8000000000000000ull;
}

__device__ __forceinline__ void frame_span(const iscg_blink_event& e, uint32_t nFrames, long long& f0, long long& f1)
{
   f0 = 1;
   f1 = 0;  // empty
   if (!(e.tEnd > 0.0) || !(e.tStart < (double)nFrames))
      return;
   const long long a = floor_to_long(e.tStart), b = floor_to_long(e.tEnd), last = (long long)nFrames - 1;
   f0 = a > 0 ? a : 0;
   f1 = b < last ? b : last;
}

__global__ void bucket_count_kernel(const iscg_blink_event* __restrict__ ev, uint32_t n, uint32_t nFrames,
                                    uint32_t* __restrict__ cnt)
{
   for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
   {
      long long f0, f1;
      frame_span(ev[e], nFrames, f0, f1);
      cnt[e] = f1 >= f0 ? (uint32_t)(f1 - f0 + 1) : 0u;
   }
}

__global__ void bucket_pairs_kernel(const iscg_blink_event* __restrict__ ev, uint32_t n, uint32_t nFrames,
                                    const unsigned long long* __restrict__ off, uint32_t* __restrict__ keys,
                                    uint32_t* __restrict__ vals)
{
   for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
   {
      long long f0, f1;
      frame_span(ev[e], nFrames, f0, f1);
      unsigned long long o = off[e];
      for (long long f = f0; f <= f1; ++f, ++o)
      {
         keys[o] = (uint32_t)f;
         vals[o] = e;
      }
   }
}

// start[k] = the first sorted position with key >= k (k = 0 .. nFrames); keys sorted ascending, P of them
__global__ void bucket_starts_kernel(const uint32_t* __restrict__ keys, unsigned long long P, uint32_t nFrames,
                                     uint32_t* __restrict__ start)
{
   for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; i <= P;
        i += (unsigned long long)gridDim.x * blockDim.x)
   {
      const long long lo = i == 0 ? -1 : (long long)keys[i - 1];
      const long long hi = i == P ? (long long)nFrames : (long long)keys[i];
      for (long long k = lo + 1; k <= hi; ++k)
         start[k] = (uint32_t)i;
   }
}
}  // namespace
}  // namespace iscg

using namespace iscg;

extern "C" {

int32_t iscg_collect_frame_emitters(iscg_context* h, const iscg_blink_event* events, uint32_t nEvents,
                                    const uint32_t* bucketStart, const uint32_t* bucketIdx, const int64_t* frameIndex,
                                    uint32_t nFrames, uint32_t width, uint32_t height, double pixelSizeNm,
                                    double photonsPerBlink, const double* driftPx, int32_t nz, double zStepNm,
                                    double globalZOffsetUm, int32_t kernel, const float* illumination,
                                    iscg_frame_emitter* out, uint64_t outCap, uint32_t* outStart, int64_t* zCounts)
{
   if (begin(h))
      return -1;
   if (nFrames == 0)
      return 0;
   if (!bucketStart || !frameIndex || !outStart || (nEvents && !events))
      return fail(h, "iscg_collect_frame_emitters: null argument");
   if (!(pixelSizeNm > 0.0))
      return fail(h, "iscg_collect_frame_emitters: pixel size must be > 0");
   if (illumination && (width == 0 || height == 0))
      return fail(h, "iscg_collect_frame_emitters: an illumination field needs the frame size");
   // the bucket offsets decide every access: read them (and check them) on the host first. A stream being
   // captured into a HIP graph cannot wait for that: there the offsets must be the device array the last eager call
   // on this context validated (same address and frame count), whose total is reused; the kernels clamp every range
   // to that total and skip event indices beyond nEvents, so no read leaves the validated lists.
   size_t nIdx = 0;
   const bool devSt = is_device_ptr(bucketStart);
   hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
   const bool capturing = hipStreamIsCapturing(h->stream, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
   (void)hipGetLastError();
   if (devSt && capturing)
   {
      if (h->collectStPtr != bucketStart || h->collectStFrames != nFrames)
         return fail(h, "iscg_collect_frame_emitters: under capture, the bucket offsets must be the device array an "
                        "eager call on this context validated");
      nIdx = h->collectStTotal;
   }
   else
   {
      std::vector<uint32_t> st(nFrames + 1);
      if (devSt)
      {
         if (!ok(h, hipMemcpyAsync(st.data(), bucketStart, st.size() * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                   h->stream), "bucket offsets") ||
             !ok(h, hipStreamSynchronize(h->stream), "bucket offsets"))
            return -1;
      }
      else
         std::memcpy(st.data(), bucketStart, st.size() * sizeof(uint32_t));
      for (uint32_t f = 0; f < nFrames; ++f)
         if (st[f + 1] < st[f])
            return fail(h, "iscg_collect_frame_emitters: bucket offsets must not decrease");
      nIdx = st[nFrames];
      h->collectStPtr = devSt ? bucketStart : nullptr;
      h->collectStFrames = nFrames;
      h->collectStTotal = nIdx;
   }
   if (nIdx && !bucketIdx)
      return fail(h, "iscg_collect_frame_emitters: null bucket lists");
   const size_t evBytes = (size_t)nEvents * sizeof(iscg_blink_event);
   const iscg_blink_event* dEv = nEvents ? (const iscg_blink_event*)stage_in(h, 0, events, evBytes) : nullptr;
   const uint32_t* dSt = (const uint32_t*)stage_in(h, 1, bucketStart, (nFrames + 1) * sizeof(uint32_t));
   const uint32_t* dIdx = nIdx ? (const uint32_t*)stage_in(h, 2, bucketIdx, nIdx * sizeof(uint32_t)) : nullptr;
   const int64_t* dFi = (const int64_t*)stage_in(h, 3, frameIndex, nFrames * sizeof(int64_t));
   const double* dDr = driftPx ? (const double*)stage_in(h, 4, driftPx, 2 * nFrames * sizeof(double)) : nullptr;
   const float* dIl =
      illumination ? (const float*)stage_in(h, 5, illumination, (size_t)width * height * sizeof(float)) : nullptr;
   if ((nEvents && !dEv) || !dSt || (nIdx && !dIdx) || !dFi || (driftPx && !dDr) || (illumination && !dIl))
      return -1;
   // counts, then offsets: in a scratch of the context (counts [nFrames] u32, start [nFrames + 1] u64, z counts 2)
   DevBuf& scr = h->stage[6];
   const size_t cntBytes = ((size_t)nFrames * 4 + 7) / 8 * 8, stBytes = ((size_t)nFrames + 1) * 8;
   const size_t st32Bytes = ((size_t)nFrames + 1) * 4;
   if (!ok(h, scr.reserve(cntBytes + stBytes + 16 + st32Bytes), "collect scratch"))
      return -1;
   uint32_t* dCnt = (uint32_t*)scr.p;
   unsigned long long* dStart = (unsigned long long*)((char*)scr.p + cntBytes);
   unsigned long long* dZ = (unsigned long long*)((char*)scr.p + cntBytes + stBytes);
   CollectArgs a{dEv, dSt, dIdx, dFi, dDr, dIl, nEvents, width, height, pixelSizeNm, photonsPerBlink, zStepNm,
                 globalZOffsetUm, nz, kernel, (uint32_t)std::min<size_t>(nIdx, 0xFFFFFFFFu)};
   hipLaunchKernelGGL(collect_count_kernel, dim3(nFrames), dim3(kCollectThreads), 0, h->stream, a, dCnt);
   hipLaunchKernelGGL(collect_scan_kernel, dim3(1), dim3(1024), 0, h->stream, (const uint32_t*)dCnt, nFrames, dStart,
                      dZ);
   if (!ok(h, hipGetLastError(), "collect launch"))
      return -1;
   // The kept emitters number at most the bucket entries (nIdx). With device outputs that hold nIdx records and no
   // z counts asked for, nothing needs the exact total on the host: the fill is enqueued without waiting. Otherwise
   // the total is read and checked first, and the call waits for its copies.
   const bool async = is_device_ptr(out) && is_device_ptr(outStart) && !zCounts && outCap >= nIdx &&
                      nIdx <= 0xFFFFFFFFull;
   unsigned long long total = nIdx;
   if (!async)
   {
      if (!ok(h, hipMemcpyAsync(&total, dStart + nFrames, sizeof(total), hipMemcpyDeviceToHost, h->stream),
              "collect") ||
          !ok(h, hipStreamSynchronize(h->stream), "collect"))
         return -1;
      if (total > outCap)
         return fail(h, "iscg_collect_frame_emitters: " + std::to_string(total) + " emitters, output holds " +
                           std::to_string(outCap));
      if (total > 0xFFFFFFFFull)
         return fail(h, "iscg_collect_frame_emitters: a batch holds at most 2^32 - 1 emitters");
   }
   iscg_frame_emitter* dOut =
      total ? (iscg_frame_emitter*)stage_out(h, 7, out, total * sizeof(iscg_frame_emitter)) : nullptr;
   if (total && !dOut)
      return -1;
   hipLaunchKernelGGL(collect_fill_kernel, dim3(nFrames), dim3(kCollectThreads), 0, h->stream, a,
                      (const unsigned long long*)dStart, dOut, dZ);
   // the offsets as uint32 (the splat's emStart), in the scratch after the z counts, then to the caller's array
   uint32_t* dOst = (uint32_t*)((char*)dZ + 16);
   hipLaunchKernelGGL(collect_start32_kernel, dim3(grid_for((size_t)nFrames + 1, 256, 4)), dim3(256), 0, h->stream,
                      (const unsigned long long*)dStart, nFrames, dOst);
   if (!ok(h, hipGetLastError(), "collect launch") ||
       !ok(h, hipMemcpyAsync(outStart, dOst, st32Bytes, is_device_ptr(outStart) ? hipMemcpyDeviceToDevice :
                             hipMemcpyDeviceToHost, h->stream), "collect offsets"))
      return -1;
   if (async)
      return 0;
   long long z[2] = {0, 0};
   if (!ok(h, hipMemcpyAsync(z, dZ, 16, hipMemcpyDeviceToHost, h->stream), "collect counts"))
      return -1;
   if (total && !finish_out(h, out, dOut, total * sizeof(iscg_frame_emitter)))
      return -1;
   if (!ok(h, hipStreamSynchronize(h->stream), "collect"))
      return -1;
   if (zCounts)
   {
      zCounts[0] = z[0];
      zCounts[1] = z[1];
   }
   return 0;
}


int32_t iscg_bucket_events(iscg_context* h, const iscg_blink_event* events, uint32_t nEvents, uint32_t nFrames,
                           uint32_t* bucketStart, uint32_t* bucketIdx, uint64_t idxCap, uint64_t* total)
{
   if (begin(h))
      return -1;
   if (!bucketStart || !total || (nEvents && !events))
      return fail(h, "iscg_bucket_events: null argument");
   *total = 0;
   const iscg_blink_event* dEv =
      nEvents ? (const iscg_blink_event*)stage_in(h, 0, events, (size_t)nEvents * sizeof(iscg_blink_event)) : nullptr;
   if (nEvents && !dEv)
      return -1;
   DevBuf* t = h->bucketTmp;
   unsigned long long P = 0;
   if (nEvents)
   {
      if (!ok(h, t[0].reserve((size_t)nEvents * 4), "bucket allocation") ||
          !ok(h, t[1].reserve((size_t)nEvents * 8), "bucket allocation"))
         return -1;
      hipLaunchKernelGGL(bucket_count_kernel, dim3(grid_for(nEvents, 256, 8)), dim3(256), 0, h->stream, dEv,
                         nEvents, nFrames, t[0].as<uint32_t>());
      size_t tmpBytes = 0;
      if (!ok(h, hipcub::DeviceScan::ExclusiveSum(nullptr, tmpBytes, t[0].as<uint32_t>(),
                                                  t[1].as<unsigned long long>(), nEvents, h->stream), "bucket scan") ||
          !ok(h, t[5].reserve(tmpBytes), "bucket allocation") ||
          !ok(h, hipcub::DeviceScan::ExclusiveSum(t[5].p, tmpBytes, t[0].as<uint32_t>(),
                                                  t[1].as<unsigned long long>(), nEvents, h->stream), "bucket scan"))
         return -1;
      uint32_t lastCnt = 0;
      unsigned long long lastOff = 0;
      if (!ok(h, hipMemcpyAsync(&lastCnt, t[0].as<uint32_t>() + nEvents - 1, 4, hipMemcpyDeviceToHost, h->stream),
              "bucket total") ||
          !ok(h, hipMemcpyAsync(&lastOff, t[1].as<unsigned long long>() + nEvents - 1, 8, hipMemcpyDeviceToHost,
                                h->stream), "bucket total") ||
          !ok(h, hipStreamSynchronize(h->stream), "bucket total"))
         return -1;
      P = lastOff + lastCnt;
   }
   *total = P;
   if (P > 0x7FFFFFFFull)
      return fail(h, "iscg_bucket_events: more than 2^31 - 1 (frame, event) pairs");
   if (P > idxCap || (P && !bucketIdx))
      return fail(h, "iscg_bucket_events: " + std::to_string(P) + " entries, the index output holds " +
                        std::to_string(idxCap));
   uint32_t* dStart = (uint32_t*)stage_out(h, 6, bucketStart, ((size_t)nFrames + 1) * 4);
   uint32_t* dIdx = P ? (uint32_t*)stage_out(h, 7, bucketIdx, (size_t)P * 4) : nullptr;
   if (!dStart || (P && !dIdx))
      return -1;
   if (P)
   {
      if (!ok(h, t[2].reserve((size_t)P * 4), "bucket allocation") ||
          !ok(h, t[3].reserve((size_t)P * 4), "bucket allocation") ||
          !ok(h, t[4].reserve((size_t)P * 4), "bucket allocation"))
         return -1;
      hipLaunchKernelGGL(bucket_pairs_kernel, dim3(grid_for(nEvents, 256, 8)), dim3(256), 0, h->stream, dEv, nEvents,
                         nFrames, (const unsigned long long*)t[1].as<unsigned long long>(), t[2].as<uint32_t>(),
                         t[3].as<uint32_t>());
      int bits = 1;
      while (bits < 32 && (1ull << bits) < (unsigned long long)nFrames)
         ++bits;
      size_t tmpBytes = 0;
      if (!ok(h, hipcub::DeviceRadixSort::SortPairs(nullptr, tmpBytes, t[2].as<uint32_t>(), t[4].as<uint32_t>(),
                                                     t[3].as<uint32_t>(), dIdx, (int)P, 0, bits, h->stream),
              "bucket sort") ||
          !ok(h, t[5].reserve(tmpBytes), "bucket allocation") ||
          !ok(h, hipcub::DeviceRadixSort::SortPairs(t[5].p, tmpBytes, t[2].as<uint32_t>(), t[4].as<uint32_t>(),
                                                     t[3].as<uint32_t>(), dIdx, (int)P, 0, bits, h->stream),
              "bucket sort"))
         return -1;
   }
   hipLaunchKernelGGL(bucket_starts_kernel, dim3(grid_for((size_t)P + 1, 256, 8)), dim3(256), 0, h->stream,
                      (const uint32_t*)(P ? t[4].as<uint32_t>() : nullptr), (unsigned long long)P, nFrames, dStart);
   if (!ok(h, hipGetLastError(), "bucket launch"))
      return -1;
   if (!finish_out(h, bucketStart, dStart, ((size_t)nFrames + 1) * 4) ||
       (P && !finish_out(h, bucketIdx, dIdx, (size_t)P * 4)))
      return -1;
   return 0;
}

}  // extern "C"
