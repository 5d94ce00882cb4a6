// binned.hip -- the binned blink regime (approximate SMLM, BinnedBlinkRenderer, Sim/BinnedBlinks.cpp:94-246): a
// frame's emitters deposited on a grid of u cells per camera pixel, one density per PSF z plane, each plane
// FFT-convolved with that plane's kernel resampled to the grid (KernelWidefieldPsf), summed in the spectrum, one
// cropped inverse, max(0, .) per cell, binned to camera pixels and added to the frame.
//
// Setup: u and R as the reference's Setup; an FFT grid of powers of two at least (W u + 2R) x (H u + 2R) (the
// reference's FastSize grid gives the same linear convolution: no wrap reaches the cropped FOV); every plane's
// kernel wrapped around the origin and R2C-transformed once, resident [plane][NY][S].
// Render (a batch of frames): on the host, the emitters grouped by (frame, plane, cell) with the photons summed in
// float in event order (the reference's dens[c] += photons); on the device, per chunk of frames: the occupied
// planes scattered and R2C-transformed (wf_passes.h rows + columns), per frame the spectra times their kernels summed
// in plane order, the cropped inverse (columns, then rows C2R), and the u x u bins of max(0, cell) added to the
// frames.
#include "wf_passes.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

namespace iscg {

// KernelWidefieldPsf::Kernel on the device (wf.hip)
__global__ void wf_resample_kernel(const float* __restrict__ planes, int n, int nzP, int r, int s0, int R,
                                   float* __restrict__ out);

struct BinnedState
{
   uint32_t W = 0, H = 0, NX = 0, NY = 0, S = 0;
   int u = 1, R = 1, nz = 1, gw = 0, gh = 0;
   bool valid = false;
   DevBuf spec, twX, twY;
   DevBuf dense, x1, x2, acc, cells, meta;
};

void destroy_binned(iscg_context* h)
{
   delete h->binned;
   h->binned = nullptr;
}

// KernelWidefieldPsf::ValidUpscale (Sim/WidefieldRender.cpp:263-270)
static int valid_upscale(int oversampling, int requested)
{
   const int os = std::max(1, oversampling);
   for (int u = std::max(1, std::min(requested, os)); u > 1; --u)
      if (os % u == 0)
         return u;
   return 1;
}

// The kernel of plane p wrapped around the origin of the NX x NY grid: G[p][y][x] = K[p][dy + R][dx + R] at
// (dy mod NY, dx mod NX), zero elsewhere (the grid is at least 2R + 1 wide, so no two taps share a cell).
__global__ void binned_place_kernel(const float* __restrict__ K, int nz, int R, int NX, int NY, float* __restrict__ G)
{
   const int D = 2 * R + 1;
   const size_t plane = (size_t)NX * NY, total = plane * nz;
   for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x)
   {
      const int p = (int)(idx / plane);
      const int rem = (int)(idx - (size_t)p * plane), y = rem / NX, x = rem % NX;
      const int dy = y <= R ? y : y - NY, dx = x <= R ? x : x - NX;
      float v = 0.0f;
      if (dy >= -R && dy <= R && dx >= -R && dx <= R)
         v = K[(size_t)p * D * D + (size_t)(dy + R) * D + (dx + R)];
      G[idx] = v;
   }
}

__global__ void binned_zero_kernel(float* __restrict__ p, size_t n)
{
   for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
      p[i] = 0.0f;
}

// Dense deposits: plane j of the chunk takes entries [starts[j], starts[j + 1]) (cells unique within a plane).
__global__ void binned_scatter_kernel(const uint32_t* __restrict__ starts, int nPlanes, uint32_t total,
                                      const uint32_t* __restrict__ cells, const float* __restrict__ values,
                                      uint32_t planeCells, float* __restrict__ dense)
{
   for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x)
   {
      int lo = 0, hi = nPlanes - 1;  // the last plane with starts[j] <= i
      while (lo < hi)
      {
         const int mid = (lo + hi + 1) >> 1;
         if (starts[mid] <= i)
            lo = mid;
         else
            hi = mid - 1;
      }
      dense[(size_t)lo * planeCells + cells[i]] = values[i];
   }
}

// Per frame (blockIdx.y) of the chunk: acc = sum over its planes j (in plane order) of X_j K_{p_j}, complex FP32 as
// the reference's acc[i] += tmp[i] * K[i], then times 1 / (NX NY) (the normalised inverse).
__global__ __launch_bounds__(256) void binned_mac_kernel(const float2* __restrict__ X, const float2* __restrict__ K,
                                                         size_t planeElems, const uint32_t* __restrict__ frameStart,
                                                         const int32_t* __restrict__ planeP, int NY, int S, int Wb,
                                                         float scale, float2* __restrict__ acc)
{
   const int f = blockIdx.y;
   const uint32_t j0 = frameStart[f], j1 = frameStart[f + 1];
   const uint32_t bins = (uint32_t)NY * Wb;
   for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < bins; b += gridDim.x * blockDim.x)
   {
      const int ky = (int)(b / Wb), kx = (int)(b - (uint32_t)ky * Wb);
      const size_t off = (size_t)ky * S + kx;
      float ar = 0.0f, ai = 0.0f;
      for (uint32_t j = j0; j < j1; ++j)
      {
         const float2 A = X[(size_t)j * planeElems + off];
         const float2 k = K[(size_t)planeP[j] * planeElems + off];
         ar = ar + (A.x * k.x - A.y * k.y);
         ai = ai + (A.x * k.y + A.y * k.x);
      }
      acc[(size_t)f * planeElems + off] = make_float2(ar * scale, ai * scale);
   }
}

// img[frame[f]] += (float) sum over the u x u cells of max(0, cell), the sum in FP64 (Sim/BinnedBlinks.cpp:231-245).
__global__ void binned_bin_kernel(const float* __restrict__ cells, int nF, int u, uint32_t W, uint32_t H,
                                  const uint32_t* __restrict__ frameOf, float* __restrict__ img)
{
   const uint32_t npx = W * H, cw = W * u;
   const size_t cellsPer = (size_t)cw * H * u;
   const size_t total = (size_t)nF * npx;
   for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x)
   {
      const int f = (int)(idx / npx);
      const uint32_t p = (uint32_t)(idx - (size_t)f * npx), y = p / W, x = p - y * W;
      const float* c = cells + (size_t)f * cellsPer;
      double sum = 0.0;
      for (int dy = 0; dy < u; ++dy)
      {
         const float* row = c + (size_t)(y * u + dy) * cw + (size_t)x * u;
         for (int dx = 0; dx < u; ++dx)
            sum += row[dx] > 0.0f ? row[dx] : 0.0f;
      }
      float* o = img + (size_t)frameOf[f] * npx + p;
      *o = *o + (float)sum;
   }
}

static uint32_t pow2_at_least(uint32_t n)
{
   uint32_t p = 16;
   while (p < n)
      p <<= 1;
   return p;
}

}  // namespace iscg

using namespace iscg;

extern "C" {

int32_t iscg_binned_setup(iscg_context* h, const iscg_kernel_view* kernel, int32_t haloRadius2, double sigmaPx,
                          uint32_t width, uint32_t height, int32_t upscale)
{
   if (begin(h))
      return -1;
   if (!h->binned)
      h->binned = new BinnedState();
   BinnedState& s = *h->binned;
   s.valid = false;
   if (width == 0 || height == 0)
      return fail(h, "iscg_binned_setup: empty frame");
   const bool useKernel = kernel && kernel->planes && kernel->nz > 0 && kernel->size >= 3;
   int u, R, r = 1, s0 = 0, nz = 1;
   if (useKernel)
   {
      const int os = std::max(1, kernel->oversampling), n = kernel->size;
      u = valid_upscale(os, upscale);
      r = std::max(1, os / u);
      const double kc = (n - 1) / 2.0;
      s0 = (int)std::floor(kc - r / 2.0 + 1.0);
      const int dMin = (int)std::ceil((-s0 - r + 1) / (double)r);
      const int dMax = (int)std::floor((n - 1 - s0) / (double)r);
      R = std::max(1, std::max(-dMin, dMax));
      if (haloRadius2 >= 0)
         R = std::min(R, (int)std::ceil((std::sqrt((double)haloRadius2) + 1.0) * u));
      nz = kernel->nz;
   }
   else
   {
      if (!(sigmaPx > 0.0))
         return fail(h, "iscg_binned_setup: no kernel and sigmaPx <= 0");
      u = std::max(1, std::min(upscale, 8));
      R = std::max(1, (int)std::ceil(3.0 * sigmaPx * u));
   }
   R = std::max(1, R);
   const uint32_t gw = width * (uint32_t)u + 2u * (uint32_t)R, gh = height * (uint32_t)u + 2u * (uint32_t)R;
   const uint32_t NX = pow2_at_least(gw), NY = pow2_at_least(gh);
   if (NX > 1024 || NY > 1024)
      return fail(h, "iscg_binned_setup: the grid (W u + 2R) x (H u + 2R) needs an FFT above 1024 points");
   s.W = width;
   s.H = height;
   s.u = u;
   s.R = R;
   s.nz = nz;
   s.gw = (int)gw;
   s.gh = (int)gh;
   s.NX = NX;
   s.NY = NY;
   s.S = spec_stride(NX);
   const int D = 2 * R + 1;
   DevBuf K, G;
   if (!ok(h, K.reserve((size_t)nz * D * D * sizeof(float)), "kernel allocation") ||
       !ok(h, G.reserve((size_t)nz * NY * NX * sizeof(float)), "kernel grid allocation") ||
       !ok(h, s.spec.reserve((size_t)nz * NY * s.S * sizeof(float2)), "spectrum allocation") ||
       !ok(h, s.x1.reserve((size_t)nz * NY * s.S * sizeof(float2)), "scratch allocation"))
      return -1;
   if (!make_twiddles(h, s.twX, NX) || !make_twiddles(h, s.twY, NY))
      return -1;
   if (useKernel)
   {
      const size_t bytes = (size_t)nz * kernel->size * kernel->size * sizeof(float);
      const float* dPlanes = (const float*)stage_in(h, 0, kernel->planes, bytes);
      if (!dPlanes)
         return -1;
      const size_t nK = (size_t)nz * D * D;
      hipLaunchKernelGGL(wf_resample_kernel, dim3(grid_for(nK, 256, 8)), dim3(256), 0, h->stream, dPlanes,
                         kernel->size, nz, r, s0, R, K.as<float>());
   }
   else
   {
      // BinnedBlinks.cpp:146-160: exp(-(x^2 + y^2) / (2 (sigma u)^2)) normalised in FP64, then float
      std::vector<double> k((size_t)D * D);
      const double sc = sigmaPx * u, twoS2 = 2.0 * sc * sc;
      double sum = 0.0;
      for (int y = -R; y <= R; ++y)
         for (int x = -R; x <= R; ++x)
         {
            const double v = std::exp(-((double)x * x + (double)y * y) / twoS2);
            k[(size_t)(y + R) * D + (x + R)] = v;
            sum += v;
         }
      std::vector<float> kf(k.size());
      for (size_t i = 0; i < k.size(); ++i)
         kf[i] = (float)(k[i] / sum);
      if (!ok(h, hipMemcpyAsync(K.p, kf.data(), kf.size() * sizeof(float), hipMemcpyHostToDevice, h->stream),
              "kernel copy") ||
          !ok(h, hipStreamSynchronize(h->stream), "stream synchronize"))
         return -1;
   }
   const size_t nG = (size_t)nz * NY * NX;
   hipLaunchKernelGGL(binned_place_kernel, dim3(grid_for(nG, 256, 16)), dim3(256), 0, h->stream, K.as<const float>(),
                      nz, R, (int)NX, (int)NY, G.as<float>());
   const XyGrid grid{NX, NY, s.S, s.twX.as<const float2>(), s.twY.as<const float2>()};
   if (!forward_xy(h, grid, G.as<const float>(), nz, (int)NY, (int)NX, s.x1.as<float2>(), s.spec.as<float2>(),
                   (size_t)NY * s.S))
      return fail(h, "iscg_binned_setup: unsupported FFT size");
   if (!ok(h, hipGetLastError(), "iscg_binned_setup launch") || !ok(h, hipStreamSynchronize(h->stream), "synchronize"))
      return -1;
   s.valid = true;
   return 0;
}

int32_t iscg_binned_geometry(iscg_context* h, int32_t* upscale, int32_t* radiusCells)
{
   if (!h || !h->binned || !h->binned->valid)
      return -1;
   if (upscale)
      *upscale = h->binned->u;
   if (radiusCells)
      *radiusCells = h->binned->R;
   return 0;
}

int32_t iscg_binned_render(iscg_context* h, const iscg_frame_emitter* em, const int32_t* frameStart, int32_t nFrames,
                           float* img)
{
   if (begin(h))
      return -1;
   if (!h->binned || !h->binned->valid)
      return fail(h, "iscg_binned_render: no setup (iscg_binned_setup)");
   if (nFrames < 0 || (nFrames > 0 && (!frameStart || !img)))
      return fail(h, "iscg_binned_render: null argument");
   if (nFrames == 0)
      return 0;
   if (frameStart[0] != 0)
      return fail(h, "iscg_binned_render: frameStart must start at 0");
   for (int f = 0; f < nFrames; ++f)
      if (frameStart[f + 1] < frameStart[f])
         return fail(h, "iscg_binned_render: frameStart must not decrease");
   if (frameStart[nFrames] > 0 && !em)
      return fail(h, "iscg_binned_render: null emitters");
   BinnedState& s = *h->binned;
   const int u = s.u, R = s.R, nz = s.nz, gw = s.gw, gh = s.gh;
   // deposits: (frame, plane, cell) in event order, photons summed in float (Sim/BinnedBlinks.cpp:175-215)
   struct Dep
   {
      uint32_t f, p, cell, i;
   };
   std::vector<Dep> deps;
   deps.reserve((size_t)frameStart[nFrames]);
   for (int f = 0; f < nFrames; ++f)
      for (int i = frameStart[f]; i < frameStart[f + 1]; ++i)
      {
         const iscg_frame_emitter& e = em[i];
         if (!(e.photons > 0.0))
            continue;
         const int p = std::max(0, std::min(nz - 1, (int)e.zIndex));
         const int cx = (int)std::floor((e.xPx + 0.5) * u) + R;
         const int cy = (int)std::floor((e.yPx + 0.5) * u) + R;
         if (cx < 0 || cy < 0 || cx >= gw || cy >= gh)
            continue;
         deps.push_back(Dep{(uint32_t)f, (uint32_t)p, (uint32_t)(cy * gw + cx), (uint32_t)i});
      }
   std::stable_sort(deps.begin(), deps.end(), [](const Dep& a, const Dep& b) {
      return a.f != b.f ? a.f < b.f : (a.p != b.p ? a.p < b.p : a.cell < b.cell);
   });
   // planes (frame, p) in order; cells unique per plane
   std::vector<uint32_t> cells, pStart{0}, planeFrame;
   std::vector<float> values;
   std::vector<int32_t> planeP;
   for (size_t a = 0; a < deps.size();)
   {
      size_t b = a;
      float v = 0.0f;
      while (b < deps.size() && deps[b].f == deps[a].f && deps[b].p == deps[a].p && deps[b].cell == deps[a].cell)
         v += (float)em[deps[b++].i].photons;
      const bool newPlane = planeP.empty() || planeFrame.back() != deps[a].f || (uint32_t)planeP.back() != deps[a].p;
      if (newPlane)
      {
         if (!planeP.empty())
            pStart.push_back((uint32_t)cells.size());
         planeP.push_back((int32_t)deps[a].p);
         planeFrame.push_back(deps[a].f);
      }
      cells.push_back(deps[a].cell);
      values.push_back(v);
      a = b;
   }
   if (planeP.empty())
      return 0;  // no deposit reaches the grid: the frames are unchanged
   pStart.push_back((uint32_t)cells.size());
   const size_t npx = (size_t)s.W * s.H;
   float* dImg = (float*)stage_in(h, 0, img, (size_t)nFrames * npx * sizeof(float));
   if (!dImg)
      return -1;
   const uint32_t S = s.S, NY = s.NY, NX = s.NX;
   const size_t planeElems = (size_t)NY * S;
   const size_t cw = (size_t)s.W * u, ch = (size_t)s.H * u;
   // chunks of whole frames within a scratch budget
   const size_t perPlane = (size_t)gh * gw * sizeof(float) + (size_t)gh * S * sizeof(float2) + planeElems * sizeof(float2);
   const size_t perFrame = planeElems * sizeof(float2) + cw * ch * sizeof(float);
   const size_t budget = (size_t)4 << 30;
   const XyGrid grid{NX, NY, S, s.twX.as<const float2>(), s.twY.as<const float2>()};
   const size_t nPlanes = planeP.size();
   // a chunk's metadata, uploaded asynchronously: kept until the stream has finished with it
   std::vector<uint32_t> meta;
   size_t j0 = 0;
   while (j0 < nPlanes)
   {
      // planes [j0, j1) of frames [fa, fb]
      size_t j1 = j0, nF = 0;
      while (j1 < nPlanes)
      {
         size_t k = j1;
         while (k < nPlanes && planeFrame[k] == planeFrame[j1])
            ++k;
         const size_t need = (k - j0) * perPlane + (nF + 1) * perFrame;
         if (j1 > j0 && need > budget)
            break;
         j1 = k;
         ++nF;
      }
      const int nP = (int)(j1 - j0);
      const uint32_t c0 = pStart[j0], c1 = pStart[j1];
      // chunk metadata: plane starts (rebased), plane p, frame starts over planes, frame ids
      meta.clear();
      for (size_t j = j0; j <= j1; ++j)
         meta.push_back(pStart[j] - c0);
      const size_t oP = meta.size();
      for (size_t j = j0; j < j1; ++j)
         meta.push_back((uint32_t)planeP[j]);
      const size_t oFS = meta.size();
      meta.push_back(0);
      const size_t oFr = oFS + nF + 1;
      std::vector<uint32_t> frameIds;
      for (size_t j = j0; j < j1; ++j)
         if (frameIds.empty() || frameIds.back() != planeFrame[j])
            frameIds.push_back(planeFrame[j]);
      {
         size_t j = j0;
         for (size_t f = 0; f < nF; ++f)
         {
            while (j < j1 && planeFrame[j] == frameIds[f])
               ++j;
            meta.push_back((uint32_t)(j - j0));
         }
      }
      meta.insert(meta.end(), frameIds.begin(), frameIds.end());
      const size_t oCells = meta.size();
      meta.insert(meta.end(), cells.begin() + c0, cells.begin() + c1);
      const size_t oVals = meta.size();
      for (uint32_t c = c0; c < c1; ++c)
      {
         uint32_t bits;
         std::memcpy(&bits, &values[c], 4);
         meta.push_back(bits);
      }
      if (!ok(h, s.meta.reserve(meta.size() * sizeof(uint32_t)), "allocation") ||
          !ok(h, s.dense.reserve((size_t)nP * gh * gw * sizeof(float)), "scratch allocation") ||
          !ok(h, s.x1.reserve((size_t)std::max<size_t>(nP * gh, nF * ch) * S * sizeof(float2)), "scratch allocation") ||
          !ok(h, s.x2.reserve((size_t)nP * planeElems * sizeof(float2)), "scratch allocation") ||
          !ok(h, s.acc.reserve(nF * planeElems * sizeof(float2)), "scratch allocation") ||
          !ok(h, s.cells.reserve(nF * cw * ch * sizeof(float)), "scratch allocation") ||
          !ok(h, hipMemcpyAsync(s.meta.p, meta.data(), meta.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                                h->stream),
              "copy"))
         return -1;
      const uint32_t* dm = s.meta.as<const uint32_t>();
      const size_t nDense = (size_t)nP * gh * gw;
      {
         ProfScope ps(h, "binned_deposit");
         hipLaunchKernelGGL(binned_zero_kernel, dim3(grid_for(nDense, 256, 8)), dim3(256), 0, h->stream,
                            s.dense.as<float>(), nDense);
         hipLaunchKernelGGL(binned_scatter_kernel, dim3(grid_for(c1 - c0, 256, 4)), dim3(256), 0, h->stream, dm, nP,
                            c1 - c0, dm + oCells, (const float*)(dm + oVals), (uint32_t)(gh * gw), s.dense.as<float>());
      }
      if (!forward_xy(h, grid, s.dense.as<const float>(), nP, gh, gw, s.x1.as<float2>(), s.x2.as<float2>(),
                      planeElems))
         return fail(h, "iscg_binned_render: unsupported FFT size");
      {
         ProfScope ps(h, "binned_mac");
         const uint32_t Wb = NX / 2 + 1;
         const dim3 g((unsigned)std::min<size_t>(((size_t)NY * Wb + 255) / 256, 1024), (unsigned)nF);
         hipLaunchKernelGGL(binned_mac_kernel, g, dim3(256), 0, h->stream, s.x2.as<const float2>(),
                            s.spec.as<const float2>(), planeElems, dm + oFS, (const int32_t*)(dm + oP), (int)NY,
                            (int)S, (int)Wb, (float)(1.0 / ((double)NX * NY)), s.acc.as<float2>());
      }
      C2rEpilogue e{};
      e.raw = 1;
      e.img = s.cells.as<float>();
      e.adu = nullptr;
      e.offset = nullptr;
      e.ch = (int)ch;
      e.cw = (int)cw;
      e.x0 = R;
      e.q = NoiseQueue{nullptr, nullptr};
      if (!inverse_xy(h, grid, s.acc.as<const float2>(), (int)nF, s.x1.as<float2>(), R, e, "binned_c2r"))
         return fail(h, "iscg_binned_render: unsupported FFT size");
      {
         ProfScope ps(h, "binned_bin");
         const size_t n = nF * npx;
         hipLaunchKernelGGL(binned_bin_kernel, dim3(grid_for(n, 256, 8)), dim3(256), 0, h->stream,
                            s.cells.as<const float>(), (int)nF, u, s.W, s.H, dm + oFr, dImg);
      }
      if (!ok(h, hipGetLastError(), "iscg_binned_render launch"))
         return -1;
      // the metadata (host vector and device buffer) is reused by the next chunk: finish this one first
      if (!ok(h, hipStreamSynchronize(h->stream), "synchronize"))
         return -1;
      j0 = j1;
   }
   return finish_out(h, img, dImg, (size_t)nFrames * npx * sizeof(float)) ? 0 : -1;
}

}  // extern "C"
