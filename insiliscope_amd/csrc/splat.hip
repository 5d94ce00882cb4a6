// splat.hip -- the blink splat + camera noise for batches of frames with uneven blink counts.
//
// Replaces SplatGpuHost / GpuSimulator (Sim/SplatGpu.h:40-77, Sim/GpuSimD3D11.h:61-105) and, bit for bit, the CPU
// path they stand in for: RenderPhotonImage (Sim/SMLMSimulation.cpp:103-201) + ApplyNoiseChain.
//
// Layout: each kernel's block sums (Sim/PsfGeneratorBridge.cpp:1021-1050) stay resident in HBM, nz x bw x bw
// float row-major (C4: 81 x 68^2 = 1.5 MB, L2-resident). A batch is a flat emitter array in frame order with
// per-frame offsets (emStart), so frames of any blink count share one launch.
//
// Passes:
//   1. setup: one lane per emitter -- SplatSetup (Sim/PsfGeneratorBridge.cpp:1382-1422) in FP64 into a 112-byte
//      record, and the emitter's centre pixel into a compact 8-byte array for the tile scans;
//   2. tiles: one workgroup per (frame, 32x32 tile), grid-stride. The tile's photon image lives in LDS, starting
//      from the background (or the caller's extra image). The frame's emitters are scanned in chunks of 256; a
//      wave ballot + prefix keeps the hits in frame order; each hit's patch (clipped to the tile) is spread over
//      the workgroup's lanes, one pixel per lane, FP64 taps as SplatKernel.inl's PixelTaps, one float add per
//      pixel per emitter in frame order (a barrier between emitters), so every pixel sums the same terms in the
//      same order as the CPU. Then the noise chain runs per pixel and the ADU are written once.
// HBM traffic per frame: the ADU (2 B/px), optional photon image (4 B/px), emitter records, the maps; the block
// sums come from L2.
#include "iscg_internal.h"

#include <climits>
#include <cmath>

namespace iscg {

bool launch_noise(iscg_context* h, const float* photons, uint16_t* out, uint32_t npx, uint32_t nFrames,
                  const uint32_t* frames, const float* offset, const float* gain, const float* readN,
                  const CamParams& cp, uint32_t seed, uint32_t pixelBase);

constexpr int kMaxKernels = 4;
constexpr int kTile = 32;  // the default tile edge (16 is the other candidate of the tuned launch)
constexpr int kBlock = 256;

struct KernelDev
{
   DevBuf sums;   // nz * bw * bw floats
   DevBuf raw;    // nz * size * size floats, kept for the Fft placement (interp 3)
   DevBuf spans;  // 2 * nz * (2 camRad + 1) int16 (lo block then hi block), or empty
   int bw = 0, os = 1, camRad = 0, size = 0, nz = 0, interp = 0;
   bool hasSpans = false, hasRaw = false;
   uint64_t serial = 0;
};

struct SplatState
{
   KernelDev k[kMaxKernels];
   int nk = 0;
   uint32_t w = 0, h = 0, pixelBase = 0;
   DevBuf offset, gain, readN, bg;
   bool hasOffset = false, hasGain = false, hasRead = false, hasBg = false, haveStatic = false;
   double flatBg = 0.0;
   uint64_t key = 0;
   DevBuf setup, pos, photons;
   // the Fft placement: per-emitter shifted block sums, per-block scratch, radix-2 tables of length fftN
   DevBuf fftSums, fftScratch, fftTables;
   int fftN = 0;
   DevBuf haloSums, haloOut;  // iscg_halo_spans: block sums, then [lo | hi | radius2 | kept] per plane
};

void destroy_splat(iscg_context* h)
{
   delete h->splat;
   h->splat = nullptr;
}

static SplatState& state(iscg_context* h)
{
   if (!h->splat)
      h->splat = new SplatState();
   return *h->splat;
}

// The per-emitter record the tile pass reads (SplatSetup + photons + plane).
struct alignas(16) SetupRec
{
   double wx[4], wy[4];
   double photons;
   int32_t x0, y0, bx, by;
   int32_t nTaps, kernel, plane, cutR2;
   int32_t pad[2];
};
static_assert(sizeof(SetupRec) == 112, "SetupRec layout");

struct KernelTable
{
   const float* sums[kMaxKernels];
   const float* raw[kMaxKernels];
   const int16_t* lo[kMaxKernels];
   const int16_t* hi[kMaxKernels];
   int bw[kMaxKernels], os[kMaxKernels], camRad[kMaxKernels], size[kMaxKernels], nz[kMaxKernels],
      interp[kMaxKernels];
   int nk, maxCamRad;
};

static KernelTable table_of(const SplatState& s)
{
   KernelTable t{};
   t.nk = s.nk;
   t.maxCamRad = 0;
   for (int i = 0; i < s.nk; ++i)
   {
      const KernelDev& k = s.k[i];
      t.sums[i] = k.sums.as<float>();
      t.raw[i] = k.hasRaw ? k.raw.as<float>() : nullptr;
      const size_t nSpan = (size_t)k.nz * (2 * k.camRad + 1);
      t.lo[i] = k.hasSpans ? k.spans.as<int16_t>() : nullptr;
      t.hi[i] = k.hasSpans ? k.spans.as<int16_t>() + nSpan : nullptr;
      t.bw[i] = k.bw;
      t.os[i] = k.os;
      t.camRad[i] = k.camRad;
      t.size[i] = k.size;
      t.nz[i] = k.nz;
      t.interp[i] = k.interp;
      t.maxCamRad = k.camRad > t.maxCamRad ? k.camRad : t.maxCamRad;
   }
   return t;
}

// Block sums of one plane (Sim/PsfGeneratorBridge.cpp:1021-1050): B[ai][bi] = sum over the os x os block at
// (ai - (os-1), bi - (os-1)), FP64 accumulation in the reference's order (row sums, then column sums of them).
__global__ void block_sums_kernel(const float* __restrict__ planes, float* __restrict__ out, int n, int os, int nz)
{
   const int off = os - 1, bw = n + off;
   const size_t total = (size_t)nz * bw * bw;
   for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x)
   {
      const int z = (int)(idx / ((size_t)bw * bw));
      const int rem = (int)(idx - (size_t)z * bw * bw);
      const int ai = rem / bw, bi = rem - ai * bw;
      const float* P = planes + (size_t)z * n * n;
      const int a = ai - off, y0 = a > 0 ? a : 0, y1 = (a + off) < (n - 1) ? (a + off) : (n - 1);
      const int b = bi - off, x0 = b > 0 ? b : 0, x1 = (b + off) < (n - 1) ? (b + off) : (n - 1);
      double acc = 0.0;
      for (int y = y0; y <= y1; ++y)
      {
         double hs = 0.0;
         for (int x = x0; x <= x1; ++x)
            hs += P[(size_t)y * n + x];
         acc += hs;
      }
      out[idx] = (float)acc;
   }
}

__device__ __forceinline__ void catmull_rom(double f, double* w)
{
   const double f2 = f * f, f3 = f2 * f;
   w[0] = -0.5 * f3 + f2 - 0.5 * f;
   w[1] = 1.5 * f3 - 2.5 * f2 + 1.0;
   w[2] = -1.5 * f3 + 2.0 * f2 + 0.5 * f;
   w[3] = 0.5 * f3 - 0.5 * f2;
}

constexpr int kFar = INT_MIN / 4;  // a centre no tile reaches

// Pass 1 for iscg_frame_emitter input: SplatSetup in FP64 (Sim/PsfGeneratorBridge.cpp:1382-1422), validity as
// PlanSplat (:1424-1433) and the min-photons filter (Sim/SMLMSimulation.cpp:151-155).
__global__ void setup_exact_kernel(const iscg_frame_emitter* __restrict__ ems, uint32_t n, KernelTable kt,
                                   double minPhotons, SetupRec* __restrict__ recs, int2* __restrict__ pos)
{
   for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
   {
      const iscg_frame_emitter em = ems[e];
      SetupRec r;
      for (int i = 0; i < 4; ++i)
      {
         r.wx[i] = i == 0 ? 1.0 : 0.0;
         r.wy[i] = i == 0 ? 1.0 : 0.0;
      }
      r.photons = em.photons;
      r.kernel = em.kernel;
      r.plane = em.zIndex;
      r.cutR2 = INT_MAX;
      r.pad[0] = r.pad[1] = 0;
      // positions beyond +-2^28 px cannot reach a frame (and would overflow the integer pixel arithmetic)
      bool valid = em.kernel >= 0 && em.kernel < kt.nk && em.photons > 0.0 &&
                   !(minPhotons > 0.0 && em.photons < minPhotons) && fabs(em.xPx) < 268435456.0 &&
                   fabs(em.yPx) < 268435456.0;
      if (valid)
         valid = em.zIndex >= 0 && em.zIndex < kt.nz[em.kernel];
      if (valid)
      {
         const int os = kt.os[em.kernel], interp = kt.interp[em.kernel];
         const double kc = (kt.size[em.kernel] - 1) / 2.0;
         r.x0 = (int)floor(em.xPx + 0.5);
         r.y0 = (int)floor(em.yPx + 0.5);
         const double tx = kc + (r.x0 - em.xPx - 0.5) * os + 0.5;
         const double ty = kc + (r.y0 - em.yPx - 0.5) * os + 0.5;
         if (interp == 0 || interp == 3)
         {
            r.bx = (int)floor(tx + 0.5);
            r.by = (int)floor(ty + 0.5);
            r.nTaps = 1;
            if (interp == 3)
            {
               // PlanSplat's Fft branch (Sim/PsfGeneratorBridge.cpp:1443-1456): the plane moved by the sub-sample
               // fraction left after rounding (kept in the unused tap weights), read nearest from its own sums
               r.wx[1] = tx - (double)r.bx;
               r.wy[1] = ty - (double)r.by;
               r.pad[0] = 1;
            }
         }
         else
         {
            const int bx = (int)floor(tx), by = (int)floor(ty);
            const double fx = tx - bx, fy = ty - by;
            if (interp == 2)
            {
               r.bx = bx - 1;
               r.by = by - 1;
               r.nTaps = 4;
               catmull_rom(fx, r.wx);
               catmull_rom(fy, r.wy);
            }
            else
            {
               r.bx = bx;
               r.by = by;
               r.nTaps = 2;
               r.wx[0] = 1.0 - fx;
               r.wx[1] = fx;
               r.wy[0] = 1.0 - fy;
               r.wy[1] = fy;
            }
         }
      }
      else
      {
         r.x0 = r.y0 = kFar;
         r.bx = r.by = 0;
         r.nTaps = 0;
      }
      recs[e] = r;
      pos[e] = make_int2(r.x0, r.y0);
   }
}

// Pass 1 for the 64-byte GpuSplatEmitter records (Sim/GpuSimD3D11.h:50-59, made by CollectGpuEmitters,
// Sim/SMLMSimulation.cpp:219-256): setup already done on the host with float weights.
__global__ void setup_records_kernel(const iscg_splat_emitter* __restrict__ ems, uint32_t n, KernelTable kt,
                                     SetupRec* __restrict__ recs, int2* __restrict__ pos)
{
   for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
   {
      const iscg_splat_emitter em = ems[e];
      SetupRec r;
      for (int i = 0; i < 4; ++i)
      {
         r.wx[i] = (double)em.wx[i];
         r.wy[i] = (double)em.wy[i];
      }
      r.photons = (double)em.photons;
      r.kernel = (int)((uint32_t)em.plane >> 16);
      r.plane = em.plane & 0xffff;
      r.cutR2 = em.cutRadius2;
      r.pad[0] = r.pad[1] = 0;
      r.bx = em.bx;
      r.by = em.by;
      r.nTaps = em.nTaps == 1 || em.nTaps == 2 ? em.nTaps : 4;
      const bool valid = r.kernel < kt.nk && r.plane < kt.nz[r.kernel] && r.photons > 0.0 && abs(em.x0) < (1 << 28) &&
                         abs(em.y0) < (1 << 28) && abs(em.bx) < (1 << 28) && abs(em.by) < (1 << 28);
      r.x0 = valid ? em.x0 : kFar;
      r.y0 = valid ? em.y0 : kFar;
      if (!valid)
         r.nTaps = 0;
      recs[e] = r;
      pos[e] = make_int2(r.x0, r.y0);
   }
}

// ---- the Fft placement (PsfInterpMode::Fft) ----------------------------------------------------------------------
// FftShiftKernelTile (Sim/PsfGeneratorBridge.cpp:1293-1380) per emitter, in FP64 with the reference's operations:
// every row of the raw n x n plane zero-padded to N (a power of two >= 2n), Fft1d forward (Sim/FftRadix2.h:74-105:
// bit reversal, then radix-2 stages with the tabulated recurrence twiddles of MakeTwiddles, :39-69), times
// exp(2 pi i kk shift / N), Fft1d inverse, x 1/N, the first n complex values kept; the same along every column of
// that intermediate, the real part rounded to float; then BuildBlockSums of the shifted plane. kFftLines lines are
// transformed at once by kFftLane lanes each. The phase cos/sin are the device's FP64 functions where the
// reference uses the host's: a shifted value differs only if one of them differs by an ulp AND the FP64 result sits
// that close to a float rounding boundary.
constexpr int kFftLines = 8, kFftLane = 32, kFftMaxN = 512;

struct FftTables
{
   const double *fc, *fs, *ic, *is;  // N - 1 each (stage len's values from len/2 - 1)
   const int* rev;                   // N
   int N;
};

// One radix-2 transform of every line in lr/li (the block's kFftLines lines of N, already bit-reversed).
__device__ __forceinline__ void fft_radix2_lines(double* lr, double* li, int N, const double* __restrict__ tc0,
                                                 const double* __restrict__ ts0, int line, int lane)
{
   for (int len = 2; len <= N; len <<= 1)
   {
      const int half = len >> 1;
      const double* tc = tc0 + (half - 1);
      const double* ts = ts0 + (half - 1);
      double* re = lr + line * N;
      double* im = li + line * N;
      for (int b = lane; b < N / 2; b += kFftLane)
      {
         const int k = b & (half - 1), a = (b - k) * 2 + k, c = a + half;
         const double cr = tc[k], ci = ts[k];
         const double vr = re[c] * cr - im[c] * ci, vi = re[c] * ci + im[c] * cr;
         re[c] = re[a] - vr;
         im[c] = im[a] - vi;
         re[a] = re[a] + vr;
         im[a] = im[a] + vi;
      }
      __syncthreads();
   }
}

// FourierShiftLine (Sim/PsfGeneratorBridge.cpp:1264-1281) on the block's lines: forward (input bit-reversed
// already), phase, the in-place bit reversal of Fft1d, inverse, x 1/N.
__device__ __forceinline__ void fourier_shift_lines(double* lr, double* li, const FftTables& t, const double* pc,
                                                    const double* ps, int line, int lane)
{
   const int N = t.N;
   fft_radix2_lines(lr, li, N, t.fc, t.fs, line, lane);
   double* re = lr + line * N;
   double* im = li + line * N;
   for (int k = lane; k < N; k += kFftLane)
   {
      const double r = re[k], i = im[k];
      re[k] = r * pc[k] - i * ps[k];
      im[k] = r * ps[k] + i * pc[k];
   }
   __syncthreads();
   for (int k = lane; k < N; k += kFftLane)
   {
      const int j = t.rev[k];
      if (k < j)
      {
         const double a = re[k], b = im[k];
         re[k] = re[j];
         im[k] = im[j];
         re[j] = a;
         im[j] = b;
      }
   }
   __syncthreads();
   fft_radix2_lines(lr, li, N, t.ic, t.is, line, lane);
   const double norm = 1.0 / N;
   for (int k = lane; k < N; k += kFftLane)
   {
      re[k] *= norm;
      im[k] *= norm;
   }
   __syncthreads();
}

// Per Fft emitter (record pad[0] = 1): the shifted plane's block sums into sums + e * stride. scratch: per block
// 2 n^2 doubles (the complex intermediate) + n^2 floats (the shifted plane).
__global__ __launch_bounds__(kFftLines * kFftLane) void fft_shift_sums_kernel(const SetupRec* __restrict__ recs,
                                                                             uint32_t n, KernelTable kt, FftTables t,
                                                                             float* __restrict__ sums, size_t stride,
                                                                             double* __restrict__ scratch,
                                                                             size_t scratchStride)
{
   __shared__ double lr[kFftLines * kFftMaxN], li[kFftLines * kFftMaxN];
   __shared__ double phase[4 * kFftMaxN];
   const int tid = threadIdx.x, line = tid / kFftLane, lane = tid % kFftLane;
   const int N = t.N, half = N / 2;
   double* midRe = scratch + (size_t)blockIdx.x * scratchStride;
   for (uint32_t e = blockIdx.x; e < n; e += gridDim.x)
   {
      const SetupRec r = recs[e];
      if (r.pad[0] != 1 || r.nTaps == 0)
         continue;
      const int kk = r.kernel, sz = kt.size[kk], os = kt.os[kk], off = os - 1, bw = sz + off;
      double* midIm = midRe + (size_t)sz * sz;
      float* shifted = (float*)(midIm + (size_t)sz * sz);
      const float* plane = kt.raw[kk] + (size_t)r.plane * sz * sz;
      double* xc = phase;
      double* xs = phase + N;
      double* yc = phase + 2 * N;
      double* ys = phase + 3 * N;
      for (int k = tid; k < N; k += blockDim.x)
      {
         const int q = k < half ? k : k - N;
         const double ax = 6.283185307179586 * q * r.wx[1] / N, ay = 6.283185307179586 * q * r.wy[1] / N;
         xc[k] = cos(ax);
         xs[k] = sin(ax);
         yc[k] = cos(ay);
         ys[k] = sin(ay);
      }
      // rows: g(x) = f(x + shiftX) along every row, complex intermediate kept
      for (int y0 = 0; y0 < sz; y0 += kFftLines)
      {
         const int y = y0 + line;
         for (int k = lane; k < N; k += kFftLane)
         {
            const int j = t.rev[k];  // the bit reversal of Fft1d, applied while loading
            lr[line * N + j] = y < sz && k < sz ? (double)plane[(size_t)y * sz + k] : 0.0;
            li[line * N + j] = 0.0;
         }
         __syncthreads();
         fourier_shift_lines(lr, li, t, xc, xs, line, lane);
         if (y < sz)
            for (int x = lane; x < sz; x += kFftLane)
            {
               midRe[(size_t)y * sz + x] = lr[line * N + x];
               midIm[(size_t)y * sz + x] = li[line * N + x];
            }
         __syncthreads();
      }
      // columns of the intermediate, the real part as float
      for (int x0 = 0; x0 < sz; x0 += kFftLines)
      {
         const int x = x0 + line;
         for (int k = lane; k < N; k += kFftLane)
         {
            const int j = t.rev[k];
            const bool in = x < sz && k < sz;
            lr[line * N + j] = in ? midRe[(size_t)k * sz + x] : 0.0;
            li[line * N + j] = in ? midIm[(size_t)k * sz + x] : 0.0;
         }
         __syncthreads();
         fourier_shift_lines(lr, li, t, yc, ys, line, lane);
         if (x < sz)
            for (int y = lane; y < sz; y += kFftLane)
               shifted[(size_t)y * sz + x] = (float)lr[line * N + y];
         __syncthreads();
      }
      // BuildBlockSums of the shifted plane (row sums, then sums of them, FP64)
      float* out = sums + (size_t)e * stride;
      for (int idx = tid; idx < bw * bw; idx += blockDim.x)
      {
         const int ai = idx / bw, bi = idx - ai * bw;
         const int a = ai - off, ya = a > 0 ? a : 0, yb = (a + off) < (sz - 1) ? (a + off) : (sz - 1);
         const int b = bi - off, xa = b > 0 ? b : 0, xb = (b + off) < (sz - 1) ? (b + off) : (sz - 1);
         double acc = 0.0;
         for (int yy = ya; yy <= yb; ++yy)
         {
            double hs = 0.0;
            for (int xx = xa; xx <= xb; ++xx)
               hs += shifted[(size_t)yy * sz + xx];
            acc += hs;
         }
         out[idx] = (float)acc;
      }
      __syncthreads();
   }
}

// The taps of one pixel (SplatRows' inner sums, Sim/SplatKernel.inl:80-167): sum_j wy[j] (sum_i wx[i] B[r0 + j][cc + i])
// in FP64 over the taps inside the block sums, in j then i order. A 4 x 4 window wholly inside (the common case)
// takes the unpredicated form: the same operations in the same order, so the same bits.
__device__ __forceinline__ double tap_sum(const float* B, int bw, int r0, int cc, int NT, const double* wx,
                                          const double* wy)
{
   double sum = 0.0;
   if (NT == 4 && r0 >= 0 && cc >= 0 && r0 + 4 <= bw && cc + 4 <= bw)
   {
#pragma unroll
      for (int j = 0; j < 4; ++j)
      {
         const float* brow = B + (size_t)(r0 + j) * bw + cc;
         double row = 0.0;
#pragma unroll
         for (int i = 0; i < 4; ++i)
            row += wx[i] * brow[i];
         sum += wy[j] * row;
      }
      return sum;
   }
   const int jLo = 0 > -r0 ? 0 : -r0, jHi = NT < bw - r0 ? NT : bw - r0;
   const int iLo = 0 > -cc ? 0 : -cc, iHi = NT < bw - cc ? NT : bw - cc;
#pragma unroll
   for (int j = 0; j < 4; ++j)
   {
      if (j < jLo || j >= jHi)
         continue;
      const float* brow = B + (size_t)(r0 + j) * bw + cc;
      double row = 0.0;
#pragma unroll
      for (int i = 0; i < 4; ++i)
         if (i >= iLo && i < iHi)
            row += wx[i] * brow[i];
      sum += wy[j] * row;
   }
   return sum;
}

struct TileArgs
{
   const SetupRec* recs;
   const int2* pos;
   const uint32_t* emStart;   // nFrames + 1, relative to the emitter array
   const uint32_t* frames;    // nFrames frame ids (noise stream)
   const float* bgScales;     // nFrames or null (1)
   const float* extras;       // nFrames x npx or null
   const float* bgMap;        // npx or null
   const float* offset;
   const float* gain;
   const float* readN;
   uint16_t* outAdu;          // nFrames x npx or null
   float* outPhotons;         // nFrames x npx or null
   double flatBg;
   const float* fftSums;      // per-emitter shifted block sums (the Fft placement), stride fftStride floats
   size_t fftStride;
   NoiseQueue q;              // deferred noise pixels (index f * npx + i)
   uint32_t w, h, tilesX, nTiles, nFrames, nEmitters, seed, pixelBase;
};

template <int TILE, bool NOISE, bool EMCCD>
__global__ __launch_bounds__(kBlock) void splat_tiles_kernel(TileArgs a, KernelTable kt, CamParams cp)
{
   __shared__ float img[TILE * TILE];
   __shared__ int list[kBlock];
   __shared__ int waveHits[kBlock / 64];
   const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
   const size_t npx = (size_t)a.w * a.h;
   const size_t items = (size_t)a.nTiles * a.nFrames;
   for (size_t item = blockIdx.x; item < items; item += gridDim.x)
   {
      const uint32_t f = (uint32_t)(item / a.nTiles);
      const uint32_t t = (uint32_t)(item - (size_t)f * a.nTiles);
      const int tx0 = (int)(t % a.tilesX) * TILE, ty0 = (int)(t / a.tilesX) * TILE;
      const float* extra = a.extras ? a.extras + (size_t)f * npx : nullptr;
      const double bgScale = a.bgScales ? (double)a.bgScales[f] : 1.0;
      const bool flat = a.bgMap == nullptr && bgScale == 1.0;
      // the tile's starting photons: the caller's extra image, else the background (RenderPhotonImage :133-149)
      for (int p = tid; p < TILE * TILE; p += kBlock)
      {
         const int x = tx0 + (p & (TILE - 1)), y = ty0 + p / TILE;
         float v = 0.0f;
         if (x < (int)a.w && y < (int)a.h)
         {
            const size_t i = (size_t)y * a.w + x;
            if (extra)
               v = extra[i];
            else if (flat)
               v = (float)a.flatBg;
            else
               v = (float)((a.bgMap ? (double)a.bgMap[i] : a.flatBg) * bgScale);
         }
         img[p] = v;
      }
      __syncthreads();
      // the frame's range, clamped to the emitter array (offsets come from the caller)
      const uint32_t e1 = min(a.emStart[f + 1], a.nEmitters), e0 = min(a.emStart[f], e1);
      const int R = kt.maxCamRad;
      for (uint32_t c0 = e0; c0 < e1; c0 += kBlock)
      {
         // scan: emitters whose (2R+1)^2 square reaches the tile, kept in frame order
         const uint32_t e = c0 + tid;
         bool hit = false;
         if (e < e1)
         {
            const int2 q = a.pos[e];
            hit = q.x + R >= tx0 && q.x - R <= tx0 + TILE - 1 && q.y + R >= ty0 && q.y - R <= ty0 + TILE - 1;
         }
         const unsigned long long mask = __ballot(hit);
         if (lane == 0)
            waveHits[wave] = __popcll(mask);
         __syncthreads();
         int before = 0, nHits = 0;
         for (int w = 0; w < kBlock / 64; ++w)
         {
            before += w < wave ? waveHits[w] : 0;
            nHits += waveHits[w];
         }
         if (hit)
            list[before + __popcll(mask & ((1ull << lane) - 1ull))] = (int)e;
         __syncthreads();
         // splat the hits in order: one pixel of the clipped patch per lane
         for (int k = 0; k < nHits; ++k)
         {
            // the hit is the same across the workgroup: its record is read with scalar loads
            const int ek = __builtin_amdgcn_readfirstlane(list[k]);
            const SetupRec& r = a.recs[ek];
            const int kk = r.kernel;
            const int camRad = kt.camRad[kk], os = kt.os[kk], bw = kt.bw[kk], NT = r.nTaps, off = os - 1;
            // patch rows/cols inside the image (Rows<NT> :87-89) and the tile
            int dxLo = -camRad > -r.x0 ? -camRad : -r.x0;
            int dxHi = camRad < (int)a.w - 1 - r.x0 ? camRad : (int)a.w - 1 - r.x0;
            int dyLo = -camRad > -r.y0 ? -camRad : -r.y0;
            int dyHi = camRad < (int)a.h - 1 - r.y0 ? camRad : (int)a.h - 1 - r.y0;
            dxLo = dxLo > tx0 - r.x0 ? dxLo : tx0 - r.x0;
            dxHi = dxHi < tx0 + TILE - 1 - r.x0 ? dxHi : tx0 + TILE - 1 - r.x0;
            dyLo = dyLo > ty0 - r.y0 ? dyLo : ty0 - r.y0;
            dyHi = dyHi < ty0 + TILE - 1 - r.y0 ? dyHi : ty0 + TILE - 1 - r.y0;
            const int rw = dxHi - dxLo + 1, rh = dyHi - dyLo + 1;
            if (rw > 0 && rh > 0)
            {
               const float* B = r.pad[0] == 1 ? a.fftSums + (size_t)ek * a.fftStride
                                              : kt.sums[kk] + (size_t)r.plane * bw * bw;
               double wx[4], wy[4];
#pragma unroll
               for (int i = 0; i < 4; ++i)
               {
                  wx[i] = r.wx[i];
                  wy[i] = r.wy[i];
               }
               const int nSpan = 2 * camRad + 1;
               const int16_t* lo = kt.lo[kk] ? kt.lo[kk] + (size_t)r.plane * nSpan : nullptr;
               const int16_t* hi = kt.hi[kk] ? kt.hi[kk] + (size_t)r.plane * nSpan : nullptr;
               for (int p = tid; p < rw * rh; p += kBlock)
               {
                  const int dy = dyLo + p / rw, dx = dxLo + (p - (p / rw) * rw);
                  if (lo && (dx < lo[dy + camRad] || dx > hi[dy + camRad]))
                     continue;
                  if (dx * dx + dy * dy > r.cutR2)
                     continue;
                  const int r0 = r.by + dy * os + off, cc = r.bx + dx * os + off;
                  const int jLo = 0 > -r0 ? 0 : -r0, jHi = NT < bw - r0 ? NT : bw - r0;
                  const int iLo = 0 > -cc ? 0 : -cc, iHi = NT < bw - cc ? NT : bw - cc;
                  // taps j in [jLo, jHi), i in [iLo, iHi) (at most 4 each), summed in that order
                  double sum = 0.0;
#pragma unroll
                  for (int j = 0; j < 4; ++j)
                  {
                     if (j < jLo || j >= jHi)
                        continue;
                     const float* brow = B + (size_t)(r0 + j) * bw + cc;
                     double row = 0.0;
#pragma unroll
                     for (int i = 0; i < 4; ++i)
                        if (i >= iLo && i < iHi)
                           row += wx[i] * brow[i];
                     sum += wy[j] * row;
                  }
                  const int li = (r.y0 + dy - ty0) * TILE + (r.x0 + dx - tx0);
                  img[li] += (float)(r.photons * sum);
               }
            }
            __syncthreads();
         }
      }
      // out: photons and/or the noise chain, per pixel
      for (int p = tid; p < TILE * TILE; p += kBlock)
      {
         const int x = tx0 + (p & (TILE - 1)), y = ty0 + p / TILE;
         if (x >= (int)a.w || y >= (int)a.h)
            continue;
         const size_t i = (size_t)y * a.w + x;
         const float v = img[p];
         if (a.outPhotons)
            a.outPhotons[(size_t)f * npx + i] = v;
         if (NOISE)
            noise_store<EMCCD>(cp, a.q, a.outAdu, (uint32_t)((size_t)f * npx + i), v, a.seed, a.frames[f],
                               (uint32_t)i + a.pixelBase, a.offset != nullptr, a.offset ? a.offset[i] : 0.0f,
                               a.gain != nullptr, a.gain ? a.gain[i] : 0.0f, a.readN != nullptr,
                               a.readN ? a.readN[i] : 0.0f);
      }
      __syncthreads();
   }
}

// The per-pixel form of the tile pass. Each lane owns PPL = TILE^2 / kBlock pixels of the tile (pixel tid + kBlock m)
// and walks the tile's ordered hits itself, adding into registers; the hits' setup records are staged in LDS once per
// chunk of kBlock emitters. The walk has no barrier, and a hit's record and tap loads do not wait on the previous
// hit's adds, so the loads of consecutive hits overlap. A pixel receives the same values in the same event order as
// in splat_tiles_kernel (the set of hits that reach it, the taps and their FP64 order are the same): the same bits.
template <int TILE, bool NOISE, bool EMCCD>
__global__ __launch_bounds__(kBlock) void splat_pix_kernel(TileArgs a, KernelTable kt, CamParams cp)
{
   static_assert(TILE * TILE % kBlock == 0, "whole pixels per lane");
   constexpr int PPL = TILE * TILE / kBlock;
   constexpr int kStage = 64;  // hits staged per pass: a small LDS footprint keeps more workgroups resident
   __shared__ SetupRec rec[kStage];
   __shared__ int list[kBlock];
   __shared__ int waveHits[kBlock / 64];
   const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
   const size_t npx = (size_t)a.w * a.h;
   const size_t items = (size_t)a.nTiles * a.nFrames;
   for (size_t item = blockIdx.x; item < items; item += gridDim.x)
   {
      const uint32_t f = (uint32_t)(item / a.nTiles);
      const uint32_t t = (uint32_t)(item - (size_t)f * a.nTiles);
      const int tx0 = (int)(t % a.tilesX) * TILE, ty0 = (int)(t / a.tilesX) * TILE;
      const float* extra = a.extras ? a.extras + (size_t)f * npx : nullptr;
      const double bgScale = a.bgScales ? (double)a.bgScales[f] : 1.0;
      const bool flat = a.bgMap == nullptr && bgScale == 1.0;
      float acc[PPL];
      int px[PPL], py[PPL];
      bool in[PPL];
#pragma unroll
      for (int m = 0; m < PPL; ++m)
      {
         const int p = tid + kBlock * m;
         px[m] = tx0 + (p & (TILE - 1));
         py[m] = ty0 + p / TILE;
         in[m] = px[m] < (int)a.w && py[m] < (int)a.h;
         float v = 0.0f;
         if (in[m])
         {
            const size_t i = (size_t)py[m] * a.w + px[m];
            if (extra)
               v = extra[i];
            else if (flat)
               v = (float)a.flatBg;
            else
               v = (float)((a.bgMap ? (double)a.bgMap[i] : a.flatBg) * bgScale);
         }
         acc[m] = v;
      }
      const uint32_t e1 = min(a.emStart[f + 1], a.nEmitters), e0 = min(a.emStart[f], e1);
      const int R = kt.maxCamRad;
      for (uint32_t c0 = e0; c0 < e1; c0 += kBlock)
      {
         const uint32_t e = c0 + tid;
         bool hit = false;
         if (e < e1)
         {
            const int2 q = a.pos[e];
            hit = q.x + R >= tx0 && q.x - R <= tx0 + TILE - 1 && q.y + R >= ty0 && q.y - R <= ty0 + TILE - 1;
         }
         const unsigned long long mask = __ballot(hit);
         if (lane == 0)
            waveHits[wave] = __popcll(mask);
         __syncthreads();
         int before = 0, nHits = 0;
         for (int w = 0; w < kBlock / 64; ++w)
         {
            before += w < wave ? waveHits[w] : 0;
            nHits += waveHits[w];
         }
         if (hit)
            list[before + __popcll(mask & ((1ull << lane) - 1ull))] = (int)e;
         __syncthreads();
         for (int b0 = 0; b0 < nHits; b0 += kStage)
         {
         const int nb = nHits - b0 < kStage ? nHits - b0 : kStage;
         if (tid < nb)
            rec[tid] = a.recs[list[b0 + tid]];
         __syncthreads();
         for (int k = 0; k < nb; ++k)
         {
            const SetupRec& r = rec[k];  // every lane reads the same record: an LDS broadcast
            const int kk = __builtin_amdgcn_readfirstlane(r.kernel);
            const int x0 = __builtin_amdgcn_readfirstlane(r.x0), y0 = __builtin_amdgcn_readfirstlane(r.y0);
            const int camRad = kt.camRad[kk], os = kt.os[kk], bw = kt.bw[kk], off = os - 1;
            const int NT = __builtin_amdgcn_readfirstlane(r.nTaps), cutR2 = __builtin_amdgcn_readfirstlane(r.cutR2);
            const int plane = __builtin_amdgcn_readfirstlane(r.plane), bx = __builtin_amdgcn_readfirstlane(r.bx),
                      by = __builtin_amdgcn_readfirstlane(r.by);
            const float* B = __builtin_amdgcn_readfirstlane(r.pad[0]) == 1
                                ? a.fftSums + (size_t)__builtin_amdgcn_readfirstlane(list[b0 + k]) * a.fftStride
                                : kt.sums[kk] + (size_t)plane * bw * bw;
            const int nSpan = 2 * camRad + 1;
            const int16_t* lo = kt.lo[kk] ? kt.lo[kk] + (size_t)plane * nSpan : nullptr;
            const int16_t* hi = kt.hi[kk] ? kt.hi[kk] + (size_t)plane * nSpan : nullptr;
            const double photons = r.photons;
            double wx[4], wy[4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
            {
               wx[i] = r.wx[i];
               wy[i] = r.wy[i];
            }
#pragma unroll
            for (int m = 0; m < PPL; ++m)
            {
               const int dx = px[m] - x0, dy = py[m] - y0;
               if (!in[m] || dx < -camRad || dx > camRad || dy < -camRad || dy > camRad)
                  continue;
               if (lo && (dx < lo[dy + camRad] || dx > hi[dy + camRad]))
                  continue;
               if (dx * dx + dy * dy > cutR2)
                  continue;
               const int r0 = by + dy * os + off, cc = bx + dx * os + off;
               const double sum = tap_sum(B, bw, r0, cc, NT, wx, wy);
               acc[m] += (float)(photons * sum);
            }
         }
         __syncthreads();  // the next pass overwrites rec, the next chunk's scan list
         }
      }
#pragma unroll
      for (int m = 0; m < PPL; ++m)
      {
         if (!in[m])
            continue;
         const size_t i = (size_t)py[m] * a.w + px[m];
         const float v = acc[m];
         if (a.outPhotons)
            a.outPhotons[(size_t)f * npx + i] = v;
         if (NOISE)
            noise_store<EMCCD>(cp, a.q, a.outAdu, (uint32_t)((size_t)f * npx + i), v, a.seed, a.frames[f],
                               (uint32_t)i + a.pixelBase, a.offset != nullptr, a.offset ? a.offset[i] : 0.0f,
                               a.gain != nullptr, a.gain ? a.gain[i] : 0.0f, a.readN != nullptr,
                               a.readN ? a.readN[i] : 0.0f);
      }
   }
}

// MakeTwiddles (Sim/FftRadix2.h:39-69) for N, both signs, and the bit reversal: [fc | fs | ic | is] (N - 1 doubles
// each) then N ints -- the host's cos/sin and the same recurrence, so the same bits as the reference's tables.
static bool fft_tables(iscg_context* h, SplatState& s, int N, FftTables& t)
{
   const size_t m = (size_t)(N - 1);
   if (s.fftN != N)
   {
      std::vector<double> tab(4 * m);
      for (int sign = -1, part = 0; part < 2; sign += 2, ++part)
      {
         double* c = tab.data() + (size_t)(2 * part) * m;
         double* sn = c + m;
         for (int len = 2; len <= N; len <<= 1)
         {
            const double ang = sign * 2.0 * 3.14159265358979323846 / len;
            const double wr = std::cos(ang), wi = std::sin(ang);
            const int half = len / 2;
            double cr = 1.0, ci = 0.0;
            for (int k = 0; k < half; ++k)
            {
               c[half - 1 + k] = cr;
               sn[half - 1 + k] = ci;
               const double nr = cr * wr - ci * wi;
               ci = cr * wi + ci * wr;
               cr = nr;
            }
         }
      }
      std::vector<int> rev((size_t)N, 0);
      for (int i = 1, j = 0; i < N; ++i)
      {
         int bit = N >> 1;
         for (; j & bit; bit >>= 1)
            j ^= bit;
         j ^= bit;
         rev[(size_t)i] = j;
      }
      const size_t bytes = 4 * m * sizeof(double) + (size_t)N * sizeof(int);
      if (!ok(h, s.fftTables.reserve(bytes), "Fft tables") ||
          !ok(h, hipMemcpyAsync(s.fftTables.p, tab.data(), 4 * m * sizeof(double), hipMemcpyHostToDevice, h->stream),
              "Fft tables") ||
          !ok(h, hipMemcpyAsync(s.fftTables.as<double>() + 4 * m, rev.data(), (size_t)N * sizeof(int),
                                hipMemcpyHostToDevice, h->stream),
              "Fft tables") ||
          !ok(h, hipStreamSynchronize(h->stream), "Fft tables"))
         return false;
      s.fftN = N;
   }
   const double* base = s.fftTables.as<const double>();
   t.fc = base;
   t.fs = base + m;
   t.ic = base + 2 * m;
   t.is = base + 3 * m;
   t.rev = (const int*)(base + 4 * m);
   t.N = N;
   return true;
}

static bool upload(iscg_context* h, DevBuf& b, const void* src, size_t bytes)
{
   if (!ok(h, b.reserve(bytes), "device allocation"))
      return false;
   const hipMemcpyKind kind = is_device_ptr(src) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
   return ok(h, hipMemcpyAsync(b.p, src, bytes, kind, h->stream), "copy to device");
}

static int32_t render(iscg_context* h, const void* ems, bool records, uint32_t nEmitters, const uint32_t* emStart,
                      uint32_t nFrames, const uint32_t* frames, const float* bgScales, const float* extras,
                      const iscg_camera* cam, uint32_t seed, double minPhotons, uint16_t* outAdu, float* outPhotons)
{
   SplatState& s = state(h);
   if (!s.haveStatic)
      return fail(h, "splat: iscg_splat_set_static not called");
   if (!emStart || (!frames && outAdu) || (nEmitters && !ems))
      return fail(h, "splat: null argument");
   if (outAdu && !cam)
      return fail(h, "splat: the noise needs a camera");
   if (nEmitters && s.nk == 0)
      return fail(h, "splat: emitters but no kernels (iscg_splat_set_kernels)");
   if (nFrames == 0)
      return 0;
   const size_t npx = (size_t)s.w * s.h;
   const uint32_t* dStart = (const uint32_t*)stage_in(h, 0, emStart, (size_t)(nFrames + 1) * sizeof(uint32_t));
   const uint32_t* dFrames = frames ? (const uint32_t*)stage_in(h, 1, frames, nFrames * sizeof(uint32_t)) : nullptr;
   const float* dScales = bgScales ? (const float*)stage_in(h, 2, bgScales, nFrames * sizeof(float)) : nullptr;
   const float* dExtras = extras ? (const float*)stage_in(h, 3, extras, nFrames * npx * sizeof(float)) : nullptr;
   const size_t recBytes = records ? sizeof(iscg_splat_emitter) : sizeof(iscg_frame_emitter);
   const void* dEms = nEmitters ? stage_in(h, 4, ems, (size_t)nEmitters * recBytes) : nullptr;
   uint16_t* dAdu = outAdu ? (uint16_t*)stage_out(h, 5, outAdu, nFrames * npx * sizeof(uint16_t)) : nullptr;
   float* dPh = outPhotons ? (float*)stage_out(h, 6, outPhotons, nFrames * npx * sizeof(float)) : nullptr;
   if (!dStart || (frames && !dFrames) || (bgScales && !dScales) || (extras && !dExtras) || (nEmitters && !dEms) ||
       (outAdu && !dAdu) || (outPhotons && !dPh))
      return -1;
   if (!ok(h, s.setup.reserve((size_t)(nEmitters ? nEmitters : 1) * sizeof(SetupRec)), "setup allocation") ||
       !ok(h, s.pos.reserve((size_t)(nEmitters ? nEmitters : 1) * sizeof(int2)), "setup allocation"))
      return -1;
   const KernelTable kt = table_of(s);
   // the Fft placement: one kernel width for all Fft stacks of the batch, exact records only
   int fftSize = 0, fftOs = 1;
   for (int i = 0; i < s.nk; ++i)
      if (s.k[i].interp == 3)
      {
         if (fftSize && (fftSize != s.k[i].size || fftOs != s.k[i].os))
            return fail(h, "splat: the Fft placement's stacks of one batch must share their size and oversampling");
         fftSize = s.k[i].size;
         fftOs = s.k[i].os;
      }
   if (fftSize && records)
      return fail(h, "splat: the Fft placement needs iscg_splat_render_frames (GPU records carry no sub-sample "
                     "shift; the reference's GPU host does not take Fft stacks either)");
   if (nEmitters)
   {
      const unsigned grid = grid_for(nEmitters, 256, 4);
      ProfScope ps(h, "splat_setup");
      if (records)
         hipLaunchKernelGGL(setup_records_kernel, dim3(grid), dim3(256), 0, h->stream,
                            (const iscg_splat_emitter*)dEms, nEmitters, kt, s.setup.as<SetupRec>(), s.pos.as<int2>());
      else
         hipLaunchKernelGGL(setup_exact_kernel, dim3(grid), dim3(256), 0, h->stream, (const iscg_frame_emitter*)dEms,
                            nEmitters, kt, minPhotons, s.setup.as<SetupRec>(), s.pos.as<int2>());
      if (!ok(h, hipGetLastError(), "splat setup launch"))
         return -1;
   }
   TileArgs a;
   a.fftSums = nullptr;
   a.fftStride = 0;
   if (fftSize && nEmitters)
   {
      int N = 1;
      while (N < 2 * fftSize)
         N <<= 1;
      if (N > kFftMaxN)
         return fail(h, "splat: the Fft placement takes kernels up to 256 samples wide");
      const size_t bw = (size_t)fftSize + fftOs - 1, stride = bw * bw;
      if ((size_t)nEmitters * stride * sizeof(float) > ((size_t)8 << 30))
         return fail(h, "splat: Fft placement batch too large (per-emitter block sums over 8 GiB): split the frames");
      const unsigned blocks = (unsigned)std::min<uint32_t>(nEmitters, 2048);
      const size_t sz2 = (size_t)fftSize * fftSize, scratchStride = 2 * sz2 + (sz2 + 1) / 2;
      FftTables t;
      if (!ok(h, s.fftSums.reserve((size_t)nEmitters * stride * sizeof(float)), "Fft placement sums") ||
          !ok(h, s.fftScratch.reserve((size_t)blocks * scratchStride * sizeof(double)), "Fft placement scratch") ||
          !fft_tables(h, s, N, t))
         return -1;
      {
         ProfScope ps(h, "splat_fft_shift");
         hipLaunchKernelGGL(fft_shift_sums_kernel, dim3(blocks), dim3(kFftLines * kFftLane), 0, h->stream,
                            (const SetupRec*)s.setup.as<SetupRec>(), nEmitters, kt, t, s.fftSums.as<float>(), stride,
                            s.fftScratch.as<double>(), scratchStride);
      }
      if (!ok(h, hipGetLastError(), "Fft placement launch"))
         return -1;
      a.fftSums = s.fftSums.as<const float>();
      a.fftStride = stride;
   }
   a.recs = s.setup.as<SetupRec>();
   a.pos = s.pos.as<int2>();
   a.emStart = dStart;
   a.frames = dFrames;
   a.bgScales = dScales;
   a.extras = dExtras;
   a.bgMap = s.hasBg ? s.bg.as<float>() : nullptr;
   a.offset = s.hasOffset ? s.offset.as<float>() : nullptr;
   a.gain = s.hasGain ? s.gain.as<float>() : nullptr;
   a.readN = s.hasRead ? s.readN.as<float>() : nullptr;
   a.outAdu = dAdu;
   a.outPhotons = dPh;
   a.flatBg = s.flatBg;
   a.w = s.w;
   a.h = s.h;
   a.nFrames = nFrames;
   a.nEmitters = nEmitters;
   a.seed = seed;
   a.pixelBase = s.pixelBase;
   CamParams cp{};
   if (cam && !cam_params_for(h, cam, cp))
      return -1;
   a.q = NoiseQueue{nullptr, nullptr};
   if (dAdu)
   {
      if ((size_t)nFrames * npx > 0xFFFFFFFFull)
         return fail(h, "splat: a batch with noise holds at most 2^32 - 1 pixels");
      if (!noise_queue(h, (size_t)nFrames * npx, a.q))
         return -1;
   }
   // The tile edge (32 or 16) and the persistent grid (workgroups per CU, 256 CUs), tuned: choice = workgroups per
   // CU + kTile16 for 16 x 16 tiles. Each (frame, tile) item is computed by one workgroup, and each pixel receives
   // the same emitters in the same order whatever the tiling, so the frames are bit-identical for every choice;
   // the timing launches write the same outputs again (the deferred-noise queue is drained and reset by each
   // fallback launch). Not tuned when the caller's extra images are the photon output (in-place accumulation).
   // + kPixForm: the per-pixel form of the same pass (splat_pix_kernel), bit-identical as well.
   constexpr int kTile16 = 64, kPixForm = 128;
   auto tiles = [&](int choice) {
      const int perCu = choice % kTile16, edge = (choice / kTile16) % 2 ? 16 : kTile;
      const bool pix = choice >= kPixForm;
      TileArgs t = a;
      t.tilesX = (s.w + edge - 1) / edge;
      t.nTiles = t.tilesX * ((s.h + edge - 1) / edge);
      const size_t items = (size_t)t.nTiles * nFrames, cap = (size_t)256 * perCu;
      const unsigned grid = (unsigned)(items < cap ? (items ? items : 1) : cap);
      auto go = [&](auto tile) {
         constexpr int T = decltype(tile)::value;
         if (pix)
         {
            if (!dAdu)
               hipLaunchKernelGGL((splat_pix_kernel<T, false, false>), dim3(grid), dim3(kBlock), 0, h->stream, t,
                                  kt, cp);
            else if (cp.emccd)
               hipLaunchKernelGGL((splat_pix_kernel<T, true, true>), dim3(grid), dim3(kBlock), 0, h->stream, t, kt,
                                  cp);
            else
               hipLaunchKernelGGL((splat_pix_kernel<T, true, false>), dim3(grid), dim3(kBlock), 0, h->stream, t,
                                  kt, cp);
            return;
         }
         if (!dAdu)
            hipLaunchKernelGGL((splat_tiles_kernel<T, false, false>), dim3(grid), dim3(kBlock), 0, h->stream, t, kt,
                               cp);
         else if (cp.emccd)
            hipLaunchKernelGGL((splat_tiles_kernel<T, true, true>), dim3(grid), dim3(kBlock), 0, h->stream, t, kt,
                               cp);
         else
            hipLaunchKernelGGL((splat_tiles_kernel<T, true, false>), dim3(grid), dim3(kBlock), 0, h->stream, t, kt,
                               cp);
      };
      if (edge == 16)
         go(std::integral_constant<int, 16>{});
      else
         go(std::integral_constant<int, kTile>{});
      return ok(h, hipGetLastError(), "splat launch");
   };
   auto fallback = [&]() {
      return !dAdu || launch_noise_fallback(h, a.q, dAdu, (uint32_t)npx, dFrames, 0, a.offset, a.gain, a.readN, cp,
                                            seed, s.pixelBase);
   };
   int choice = 16;
   const bool inPlace = dExtras && dPh && dExtras < (const float*)dPh + nFrames * npx && (const float*)dPh < dExtras +
                        nFrames * npx;
   if (!inPlace && nFrames > 0)
   {
      const std::string key = "splat_tiles " + std::to_string(s.w) + "x" + std::to_string(s.h) + " frames" +
                              std::to_string(pow2_bucket(nFrames)) + " em/frame" +
                              std::to_string(pow2_bucket(nEmitters / (nFrames ? nFrames : 1))) + " noise" +
                              std::to_string(dAdu ? 1 + cp.emccd : 0) + " fft" + std::to_string(a.fftSums ? 1 : 0) +
                              " camRad" + std::to_string(kt.maxCamRad);
      std::vector<int> cands;
      for (int form = 0; form < 2; ++form)
         for (int t16 = 0; t16 < 2; ++t16)
            for (int c : {4, 8, 16, 32})
               cands.push_back(c + kTile16 * t16 + kPixForm * form);
      choice = tuned_choice(h, key, cands, 16, [&](int c) { return tiles(c) && fallback(); });
   }
   {
      ProfScope ps(h, "splat_tiles_noise");
      if (!tiles(choice))
         return -1;
   }
   if (!fallback())
      return -1;
   if (outAdu && !finish_out(h, outAdu, dAdu, nFrames * npx * sizeof(uint16_t)))
      return -1;
   if (outPhotons && !finish_out(h, outPhotons, dPh, nFrames * npx * sizeof(float)))
      return -1;
   return 0;
}

}  // namespace iscg

using namespace iscg;

namespace iscg {

// HaloPlaneSpans (Sim/PsfGeneratorBridge.cpp:1137-1186) for one plane per workgroup, from its block sums B (bw x bw,
// row-major). Camera pixel (dx, dy) of a splat reads block sums in rows [halfOv - 2 + dy os, halfOv + os + 1 + dy os]
// and the same columns for dx (clipped to the table); the pixel is kept when the largest of them (the maximum starts
// at 0 and takes a value only when it is larger, as std::max(m, v)) is >= thr. Per dy: the first and last kept dx
// (lo > hi: none); per plane: the largest dy^2 + max(lo^2, hi^2) kept (-1: none) and the kept pixel count.
// Maxima and the comparison are exact, so the spans are those of the reference bit for bit.
__global__ __launch_bounds__(256) void halo_spans_kernel(const float* __restrict__ sums, int bw, int os, int halfOv,
                                                         int camRad, double thr, int16_t* __restrict__ lo,
                                                         int16_t* __restrict__ hi, int32_t* __restrict__ radius2,
                                                         long long* __restrict__ kept)
{
   __shared__ int r2s;
   __shared__ unsigned long long keptS;
   const int z = (int)blockIdx.x, n = 2 * camRad + 1;
   const float* B = sums + (size_t)z * bw * bw;
   if (threadIdx.x == 0)
   {
      r2s = -1;
      keptS = 0;
   }
   __syncthreads();
   for (int i = (int)threadIdx.x; i < n; i += (int)blockDim.x)
   {
      const int dy = i - camRad;
      const int r0 = max(0, halfOv - 2 + dy * os), r1 = min(bw - 1, halfOv + os + 1 + dy * os);
      int l = camRad + 1, h = -camRad - 1;
      for (int dx = -camRad; dx <= camRad; ++dx)
      {
         const int c0 = max(0, halfOv - 2 + dx * os), c1 = min(bw - 1, halfOv + os + 1 + dx * os);
         float m = 0.0f;
         for (int r = r0; r <= r1; ++r)
            for (int c = c0; c <= c1; ++c)
            {
               const float v = B[(size_t)r * bw + c];
               m = m < v ? v : m;
            }
         if ((double)m >= thr)
         {
            l = min(l, dx);
            h = max(h, dx);
         }
      }
      lo[(size_t)z * n + i] = (int16_t)l;
      hi[(size_t)z * n + i] = (int16_t)h;
      if (l <= h)
      {
         atomicAdd(&keptS, (unsigned long long)(h - l + 1));
         atomicMax(&r2s, dy * dy + max(l * l, h * h));
      }
   }
   __syncthreads();
   if (threadIdx.x == 0)
   {
      radius2[z] = r2s;
      kept[z] = (long long)keptS;
   }
}

}  // namespace iscg

extern "C" {

int32_t iscg_splat_set_kernels(iscg_context* h, const iscg_kernel_view* kernels, int32_t n)
{
   if (begin(h))
      return -1;
   if (n < 0 || n > kMaxKernels || (n > 0 && !kernels))
      return fail(h, "iscg_splat_set_kernels: 0..4 kernels");
   SplatState& s = state(h);
   for (int i = 0; i < n; ++i)
   {
      const iscg_kernel_view& v = kernels[i];
      if (!v.planes || v.nz < 1 || v.size < 3 || (v.size % 2) != 1 || v.oversampling < 1 || v.interpMode < 0 ||
          v.interpMode > 3)
         return fail(h, "iscg_splat_set_kernels: bad kernel view (planes, odd size, nz >= 1, os >= 1, interp 0-3)");
   }
   for (int i = 0; i < n; ++i)
   {
      const iscg_kernel_view& v = kernels[i];
      KernelDev& k = s.k[i];
      const int os = v.oversampling, bw = v.size + os - 1, camRad = ((v.size - 1) / 2) / os;
      const bool same = v.serial != 0 && k.serial == v.serial && k.size == v.size && k.nz == v.nz && k.os == os &&
                        (v.interpMode != 3 || k.hasRaw);
      k.interp = v.interpMode;
      if (same)
         continue;
      k.bw = bw;
      k.os = os;
      k.camRad = camRad;
      k.size = v.size;
      k.nz = v.nz;
      k.serial = v.serial;
      const size_t planeBytes = (size_t)v.nz * v.size * v.size * sizeof(float);
      const float* dPlanes = (const float*)stage_in(h, 0, v.planes, planeBytes);
      if (!dPlanes || !ok(h, k.sums.reserve((size_t)v.nz * bw * bw * sizeof(float)), "block-sum allocation"))
         return -1;
      const size_t total = (size_t)v.nz * bw * bw;
      hipLaunchKernelGGL(block_sums_kernel, dim3(grid_for(total, 256, 4)), dim3(256), 0, h->stream, dPlanes,
                         k.sums.as<float>(), v.size, os, v.nz);
      if (!ok(h, hipGetLastError(), "block sums launch"))
         return -1;
      // the Fft placement shifts the raw planes per emitter
      k.hasRaw = v.interpMode == 3;
      if (k.hasRaw && (!ok(h, k.raw.reserve(planeBytes), "kernel plane allocation") ||
                       !ok(h, hipMemcpyAsync(k.raw.p, dPlanes, planeBytes, hipMemcpyDeviceToDevice, h->stream),
                           "kernel plane copy")))
         return -1;
      k.hasSpans = v.spanLo && v.spanHi;
      if (k.hasSpans)
      {
         const size_t nSpan = (size_t)v.nz * (2 * camRad + 1);
         if (!ok(h, k.spans.reserve(2 * nSpan * sizeof(int16_t)), "span allocation"))
            return -1;
         const hipMemcpyKind kl = is_device_ptr(v.spanLo) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
         const hipMemcpyKind kh = is_device_ptr(v.spanHi) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
         if (!ok(h, hipMemcpyAsync(k.spans.p, v.spanLo, nSpan * sizeof(int16_t), kl, h->stream), "span copy") ||
             !ok(h, hipMemcpyAsync(k.spans.as<int16_t>() + nSpan, v.spanHi, nSpan * sizeof(int16_t), kh, h->stream),
                 "span copy"))
            return -1;
      }
      // staged (host) planes must outlive the block-sum kernel before the staging buffer is reused; device planes
      // are read in place, so nothing waits (the call can then be captured in a HIP graph)
      if (dPlanes != v.planes && !ok(h, hipStreamSynchronize(h->stream), "stream synchronize"))
         return -1;
   }
   s.nk = n;
   return 0;
}

int32_t iscg_halo_spans(iscg_context* h, const float* planes, int32_t nz, int32_t size, int32_t oversampling,
                        int32_t interpMode, double threshold, int16_t* lo, int16_t* hi, int32_t* radius2,
                        double* keptFraction)
{
   if (begin(h))
      return -1;
   if (!planes || nz < 1 || size < 1 || (size % 2) != 1 || interpMode < 0 || interpMode > 3)
      return fail(h, "iscg_halo_spans: bad kernel (planes, odd size, nz >= 1, interp 0-3)");
   const int os = oversampling > 1 ? oversampling : 1, halfOv = (size - 1) / 2, camRad = halfOv / os;
   if (!(threshold > 0.0) || camRad < 1 || camRad > 16000)
      return 0;  // no cut: the whole square (WithHaloCut returns the kernel unchanged)
   if (!lo || !hi || !radius2)
      return fail(h, "iscg_halo_spans: null output");
   const int n = 2 * camRad + 1, bw = size + os - 1;
   // Catmull-Rom weights sum to at most 1.25 in absolute value per axis: Cubic and Fft reads keep a 1.25^2 margin
   const double overshoot = interpMode == 2 || interpMode == 3 ? 1.5625 : 1.0;
   SplatState& s = state(h);
   const size_t planeBytes = (size_t)nz * size * size * sizeof(float);
   const size_t spanBytes = (size_t)nz * n * sizeof(int16_t);
   const size_t outBytes = 2 * spanBytes + (size_t)nz * (sizeof(int32_t) + sizeof(long long)) + 16;
   const float* dPlanes = (const float*)stage_in(h, 0, planes, planeBytes);
   if (!dPlanes || !ok(h, s.haloSums.reserve((size_t)nz * bw * bw * sizeof(float)), "halo allocation") ||
       !ok(h, s.haloOut.reserve(outBytes), "halo allocation"))
      return -1;
   char* o = (char*)s.haloOut.p;
   int16_t* dLo = (int16_t*)o;
   int16_t* dHi = dLo + (size_t)nz * n;
   int32_t* dR2 = (int32_t*)(o + ((2 * spanBytes + 7) & ~(size_t)7));
   long long* dKept = (long long*)((char*)dR2 + (((size_t)nz * sizeof(int32_t) + 7) & ~(size_t)7));
   hipLaunchKernelGGL(block_sums_kernel, dim3(grid_for((size_t)nz * bw * bw, 256, 4)), dim3(256), 0, h->stream,
                      dPlanes, s.haloSums.as<float>(), size, os, nz);
   hipLaunchKernelGGL(halo_spans_kernel, dim3((unsigned)nz), dim3(256), 0, h->stream,
                      (const float*)s.haloSums.as<float>(), bw, os, halfOv, camRad, threshold / overshoot, dLo, dHi,
                      dR2, dKept);
   if (!ok(h, hipGetLastError(), "halo spans launch"))
      return -1;
   std::vector<long long> kept((size_t)nz);
   if (!ok(h, hipMemcpyAsync(lo, dLo, spanBytes, hipMemcpyDefault, h->stream), "halo copy") ||
       !ok(h, hipMemcpyAsync(hi, dHi, spanBytes, hipMemcpyDefault, h->stream), "halo copy") ||
       !ok(h, hipMemcpyAsync(radius2, dR2, (size_t)nz * sizeof(int32_t), hipMemcpyDefault, h->stream), "halo copy") ||
       !ok(h, hipMemcpyAsync(kept.data(), dKept, (size_t)nz * sizeof(long long), hipMemcpyDefault, h->stream),
           "halo copy") ||
       !ok(h, hipStreamSynchronize(h->stream), "halo spans"))
      return -1;
   if (keptFraction)
   {
      double sum = 0.0;
      for (long long k : kept)
         sum += (double)k;
      *keptFraction = sum / ((double)n * n * nz);
   }
   return camRad;
}

int32_t iscg_splat_set_static(iscg_context* h, uint32_t width, uint32_t height, uint32_t pixelBase,
                              const float* offset, const float* gain, const float* readNoise, const float* background,
                              double flatBackground, uint64_t key)
{
   if (begin(h))
      return -1;
   if (width == 0 || height == 0 || (size_t)width * height > 0xFFFFFFFFull)
      return fail(h, "iscg_splat_set_static: bad frame size");
   SplatState& s = state(h);
   if (key != 0 && s.haveStatic && key == s.key && width == s.w && height == s.h)
      return 0;
   const size_t bytes = (size_t)width * height * sizeof(float);
   s.hasOffset = offset != nullptr;
   s.hasGain = gain != nullptr;
   s.hasRead = readNoise != nullptr;
   s.hasBg = background != nullptr;
   if ((offset && !upload(h, s.offset, offset, bytes)) || (gain && !upload(h, s.gain, gain, bytes)) ||
       (readNoise && !upload(h, s.readN, readNoise, bytes)) || (background && !upload(h, s.bg, background, bytes)))
      return -1;
   if (!ok(h, hipStreamSynchronize(h->stream), "stream synchronize"))
      return -1;
   s.w = width;
   s.h = height;
   s.pixelBase = pixelBase;
   s.flatBg = flatBackground;
   s.key = key;
   s.haveStatic = true;
   return 0;
}

int32_t iscg_splat_render_frames(iscg_context* h, const iscg_frame_emitter* ems, uint32_t nEmitters,
                                 const uint32_t* emStart, uint32_t nFrames, const uint32_t* frames,
                                 const float* bgScales, const float* extras, const iscg_camera* cam, uint32_t seed,
                                 double minPhotons, uint16_t* outAdu, float* outPhotons)
{
   if (begin(h))
      return -1;
   return render(h, ems, false, nEmitters, emStart, nFrames, frames, bgScales, extras, cam, seed, minPhotons, outAdu,
                 outPhotons);
}

int32_t iscg_splat_render_records(iscg_context* h, const iscg_splat_emitter* ems, uint32_t nEmitters,
                                  const uint32_t* emStart, uint32_t nFrames, const uint32_t* frames,
                                  const float* bgScales, const float* extras, const iscg_camera* cam, uint32_t seed,
                                  uint16_t* outAdu, float* outPhotons)
{
   if (begin(h))
      return -1;
   return render(h, ems, true, nEmitters, emStart, nFrames, frames, bgScales, extras, cam, seed, 0.0, outAdu,
                 outPhotons);
}

}  // extern "C"
