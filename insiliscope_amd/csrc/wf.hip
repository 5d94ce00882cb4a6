// wf.hip -- the mean-field (widefield) image of a dye density: every focus of a z stack in one 3D FFT convolution.
//
// The reference evaluates it focus by focus (WidefieldScene, Sim/WidefieldRender.cpp:1315-1880): per dye plane k a
// 2D real FFT, per focus the spectra re-paired with the PSF plane spectrum at each plane's defocus, linearly blended
// between two PSF planes and clamped at the stack's ends (PlanFocus, :1460-1524), then one inverse FFT and a crop.
// With foci spaced by the dye-plane thickness the blended kernel depends only on k - f, so all foci together are a
// z-convolution: out[f] = sum_k D[k] (*) H[f - k]. This file computes it as one zero-padded R2C/C2R FP32 3D FFT.
//
// Plan (once per PSF and geometry, iscg_wf_conv3d_plan):
//   KernelWidefieldPsf resampling of the oversampled planes to the grid pitch (:272-323, FP64 sums) -> the
//   blended plane H[m] for every lag m = f - k (weights from PlanFocus's rule, host FP64) placed around the origin
//   of an NZ x NY x NX grid -> its 3D R2C spectrum, scaled by 1/(NX NY NZ): Ghat[kz][ky][kx].
// Per stack (iscg_wf_conv3d), five passes over HBM, each FFT line in LDS (fft_lds.h), zeros never read:
//   A rows R2C along x (two real rows per complex FFT), the nzD x ny dye rows   -> X1 [nzD][ny][S]
//   B columns along y, 16-column blocks (128-byte row segments), ny -> NY        -> X2 [nzD][NY][S]
//   C columns along z: FFT (nzD nonzero) x Ghat, inverse, keep nF               -> X3 [nF][NY][S]
//   D columns along y inverse, keep the FOV rows                                -> X4 [nF][ch][S]
//   E rows C2R along x, keep the FOV columns, max(0, .), + background, camera noise -> ADU [nF][ch][cw]
// S = NX/2 + 1 rounded up to 16 (128-byte aligned row segments).
#include "wf_passes.h"

namespace iscg {

bool launch_noise(iscg_context* h, const float* photons, uint16_t* out, uint32_t npx, uint32_t nFrames,
                  const uint32_t* frames, const float* offset, const float* gain, const float* readN,
                  const CamParams& cp, uint32_t seed, uint32_t pixelBase);

struct WfState
{
   iscg_wf_desc d{};
   uint32_t S = 0;   // complex row stride
   int R = 0, nzP = 0;
   bool planned = false;
   DevBuf ghat, x1, x2, twX, twY, twZ, img;
};

void destroy_wf(iscg_context* h)
{
   delete h->wf;
   h->wf = nullptr;
}

// ---- kernels --------------------------------------------------------------------------------------------------

// KernelWidefieldPsf::Kernel (Sim/WidefieldRender.cpp:306-323): cell (dx, dy) of plane p = sum of the r x r
// oversampled cells from s0 + d r, FP64.
__global__ void wf_resample_kernel(const float* __restrict__ planes, int n, int nzP, int r, int s0, int R,
                                   float* __restrict__ out)
{
   const int D = 2 * R + 1;
   const size_t total = (size_t)nzP * D * D;
   for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x)
   {
      const int p = (int)(idx / ((size_t)D * D));
      const int rem = (int)(idx - (size_t)p * D * D);
      const int dy = rem / D - R, dx = rem % D - R;
      const float* P = planes + (size_t)p * n * n;
      const int ya = s0 + dy * r, yb = n < ya + r ? n : ya + r;
      const int xa = s0 + dx * r, xb = n < xa + r ? n : xa + r;
      double acc = 0.0;
      for (int y = ya > 0 ? ya : 0; y < yb; ++y)
         for (int x = xa > 0 ? xa : 0; x < xb; ++x)
            acc += P[(size_t)y * n + x];
      out[idx] = (float)acc;
   }
}

// H[m] placed around the origin: G[m][y][x] = w0[m] K[p0[m]][dy+R][dx+R] + w1[m] K[p0[m]+1][..] for the lags
// m that exist (p0 < 0: zero plane), dy = y or y - NY, dx likewise.
__global__ void wf_place_kernel(const float* __restrict__ K, int R, int nzP, const int* __restrict__ p0,
                                const float* __restrict__ w0, const float* __restrict__ w1, int NX, int NY, int NZ,
                                float* __restrict__ G)
{
   const int D = 2 * R + 1;
   const size_t total = (size_t)NZ * NY * NX;
   for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x)
   {
      const int m = (int)(idx / ((size_t)NY * NX));
      const int rem = (int)(idx - (size_t)m * NY * NX);
      const int y = rem / NX, x = rem % NX;
      const int dy = y <= R ? y : (y >= NY - R ? y - NY : 1 << 20);
      const int dx = x <= R ? x : (x >= NX - R ? x - NX : 1 << 20);
      float v = 0.0f;
      const int p = p0[m];
      if (p >= 0 && dy <= R && dx <= R)
      {
         const size_t o = (size_t)(dy + R) * D + (dx + R);
         v = w0[m] * K[(size_t)p * D * D + o];
         if (p + 1 < nzP && w1[m] != 0.0f)
            v += w1[m] * K[(size_t)(p + 1) * D * D + o];
      }
      G[idx] = v;
   }
}

// Columns per block of pass C: 16 (128-byte row segments). Measured at NZ = 128: 8 columns and 32 columns are both
// slower (the pass is bound by the latency of its LDS stages and barriers, which wider blocks stretch).
template <int N>
constexpr int zconv_cols()
{
   return 16;
}

// Pass C: per (ky, KC-column block) the z line: FFT of the nzIn planes (zero beyond), times Ghat, inverse FFT, the
// first nOut planes written. BUILD: forward only, times `scale`, every plane written (the plan's kernel spectrum).
// Columns at or beyond S (the last block's tail) are neither read nor written.
template <int N, bool BUILD>
__global__ __launch_bounds__(zconv_cols<N>() * N / 8) void zconv_kernel(const float2* __restrict__ in, int nzIn,
                                                                        float2* __restrict__ out, int nOut,
                                                                        const float2* __restrict__ ghat, int NY,
                                                                        int S, int nColsValid, float scale,
                                                                        const float2* __restrict__ tw)
{
   constexpr int KC = zconv_cols<N>();
   constexpr int CS = KC + 1;
   __shared__ float2 buf[N * CS];
   __shared__ float2 twl[N];  // the twiddles, read by every stage
   const int tid = threadIdx.x;
   for (int i = tid; i < N; i += blockDim.x)
      twl[i] = tw[i];
   __syncthreads();
   const int colBlocks = (nColsValid + KC - 1) / KC;
   const int items = NY * colBlocks;
   const size_t plane = (size_t)NY * S;
   for (int it = blockIdx.x; it < items; it += gridDim.x)
   {
      const int ky = it / colBlocks, c0 = (it % colBlocks) * KC;
      const size_t off = (size_t)ky * S + c0;
      const int cw = S - c0 < KC ? S - c0 : KC;  // columns inside the row
      for (int i = tid; i < N * KC; i += blockDim.x)
      {
         const int z = i / KC, c = i % KC;
         buf[z * CS + c] = z < nzIn && c < cw ? in[(size_t)z * plane + off + c] : make_float2(0.0f, 0.0f);
      }
      __syncthreads();
      fft_lines<N, false, true>(buf, 1, CS, KC, tid, twl);
      if (BUILD)
      {
         for (int i = tid; i < N * KC; i += blockDim.x)
         {
            const int z = i / KC, c = i % KC;
            const float2 v = buf[z * CS + c];
            if (c < cw)
               out[(size_t)z * plane + off + c] = make_float2(v.x * scale, v.y * scale);
         }
         __syncthreads();
         continue;
      }
      for (int i = tid; i < N * KC; i += blockDim.x)
      {
         const int z = i / KC, c = i % KC;
         if (c < cw)
            buf[z * CS + c] = cmul(buf[z * CS + c], ghat[(size_t)z * plane + off + c]);
      }
      __syncthreads();
      fft_lines<N, true, true>(buf, 1, CS, KC, tid, twl);
      for (int i = tid; i < nOut * KC; i += blockDim.x)
      {
         const int z = i / KC, c = i % KC;
         if (c < cw)
            out[(size_t)z * plane + off + c] = buf[z * CS + c];
      }
      __syncthreads();
   }
}

template <int N>
struct ZConv
{
   static void run(hipStream_t s, bool build, const float2* in, int nzIn, float2* out, int nOut, const float2* ghat,
                   int NY, int S, int nCols, float scale, const float2* tw)
   {
      constexpr int KC = zconv_cols<N>();
      const size_t items = (size_t)NY * ((nCols + KC - 1) / KC);
      if (build)
         hipLaunchKernelGGL((zconv_kernel<N, true>), dim3(wgs(items, 256 * 16)), dim3(KC * N / 8), 0, s, in,
                            nzIn, out, nOut, ghat, NY, S, nCols, scale, tw);
      else
         hipLaunchKernelGGL((zconv_kernel<N, false>), dim3(wgs(items, 256 * 16)), dim3(KC * N / 8), 0, s, in,
                            nzIn, out, nOut, ghat, NY, S, nCols, scale, tw);
   }
};

static XyGrid grid_of(const WfState& s)
{
   return XyGrid{s.d.NX, s.d.NY, s.S, s.twX.as<const float2>(), s.twY.as<const float2>()};
}

}  // namespace iscg

using namespace iscg;

extern "C" {

int32_t iscg_wf_conv3d_plan(iscg_context* h, const iscg_wf_desc* desc, const float* psfPlanes, int32_t nzP,
                            int32_t size, int32_t oversampling, double zStepNm, int32_t kernelRadius)
{
   if (begin(h))
      return -1;
   if (!desc || !psfPlanes)
      return fail(h, "iscg_wf_conv3d_plan: null argument");
   iscg_wf_desc d = *desc;
   if (d.nx == 0 || d.ny == 0 || d.nzD == 0 || d.nF == 0 || !(d.dzUm > 0.0) || nzP < 1 || size < 3 ||
       (size % 2) != 1 || oversampling < 1 || !(zStepNm > 0.0))
      return fail(h, "iscg_wf_conv3d_plan: bad geometry");
   if (d.cw == 0)
      d.cw = d.nx - d.fovX0;
   if (d.ch == 0)
      d.ch = d.ny - d.fovY0;
   if (d.fovX0 + d.cw > d.nx || d.fovY0 + d.ch > d.ny)
      return fail(h, "iscg_wf_conv3d_plan: the FOV crop leaves the dye grid");
   if (d.upscale > 1)
      return fail(h, "iscg_wf_conv3d_plan: upscale > 1 (binned cells) is not supported by the 3D path");
   // KernelWidefieldPsf geometry (Sim/WidefieldRender.cpp:272-304), upscale 1: r = os
   const int r = oversampling;
   const double kc = (size - 1) / 2.0;
   const int s0 = (int)std::floor(kc - r / 2.0 + 1.0);
   int R = kernelRadius;
   if (R <= 0)
   {
      const int dMin = (int)std::ceil((-s0 - r + 1) / (double)r);
      const int dMax = (int)std::floor((size - 1 - s0) / (double)r);
      R = std::max(1, std::max(-dMin, dMax));
   }
   if (d.NX == 0)
      d.NX = next_pow2(d.nx + R);
   if (d.NY == 0)
      d.NY = next_pow2(d.ny + R);
   if (d.NZ == 0)
      d.NZ = next_pow2(d.nzD + d.nF - 1);
   if (!pow2(d.NX, 1024) || !pow2(d.NY, 512) || !pow2(d.NZ, 512))
      return fail(h, "iscg_wf_conv3d_plan: FFT sizes must be powers of two, NX in [16, 1024], NY and NZ in "
                     "[16, 512]");
   if (d.NX < d.nx + (uint32_t)R || d.NY < d.ny + (uint32_t)R || d.NZ < d.nzD + d.nF - 1)
      return fail(h, "iscg_wf_conv3d_plan: FFT grid too small for a linear convolution (NX >= nx + R, "
                     "NY >= ny + R, NZ >= nzD + nF - 1)");
   if ((uint32_t)(2 * R + 1) > d.NX || (uint32_t)(2 * R + 1) > d.NY)
      return fail(h, "iscg_wf_conv3d_plan: kernel wider than the FFT grid");
   if (!h->wf)
      h->wf = new WfState();
   WfState& s = *h->wf;
   s.planned = false;
   s.d = d;
   s.R = R;
   s.nzP = nzP;
   s.S = ((d.NX / 2 + 1) + kCols - 1) / kCols * kCols;
   const size_t S = s.S;
   // blend weights per lag m (PlanFocus, Sim/WidefieldRender.cpp:1476-1490; KernelWidefieldPsf::PlaneCoord :285)
   std::vector<int> p0(d.NZ, -1);
   std::vector<float> w0(d.NZ, 0.0f), w1(d.NZ, 0.0f);
   const int maxPlane = nzP - 1;
   for (uint32_t m = 0; m < d.NZ; ++m)
   {
      long delta;  // k - f
      if (m < d.nF)
         delta = -(long)m;
      else if (m >= d.NZ - (d.nzD - 1))
         delta = (long)d.NZ - (long)m;
      else
         continue;
      const double defocus = ((double)d.k0 + 0.5) * d.dzUm - d.focus0Um + (double)delta * d.dzUm;
      double t = nzP <= 1 ? 0.0 : defocus * 1000.0 / zStepNm + (nzP - 1) / 2.0;
      t = std::min<double>(maxPlane, std::max<double>(0.0, t));
      int p = (int)std::floor(t);
      double frac = t - p;
      if (p >= maxPlane)
      {
         p = maxPlane;
         frac = 0.0;
      }
      p0[m] = p;
      w0[m] = (float)(1.0 - frac);
      w1[m] = (float)frac;
   }
   const size_t planeBytes = (size_t)nzP * size * size * sizeof(float);
   const float* dPlanes = (const float*)stage_in(h, 0, psfPlanes, planeBytes);
   if (!dPlanes)
      return -1;
   const int D = 2 * R + 1;
   DevBuf K, P0, W0, W1, G;
   if (!ok(h, K.reserve((size_t)nzP * D * D * sizeof(float)), "kernel allocation") ||
       !ok(h, P0.reserve(d.NZ * sizeof(int)), "alloc") || !ok(h, W0.reserve(d.NZ * sizeof(float)), "alloc") ||
       !ok(h, W1.reserve(d.NZ * sizeof(float)), "alloc") ||
       !ok(h, G.reserve((size_t)d.NZ * d.NY * d.NX * sizeof(float)), "kernel grid allocation"))
      return -1;
   if (!ok(h, hipMemcpyAsync(P0.p, p0.data(), d.NZ * sizeof(int), hipMemcpyHostToDevice, h->stream), "copy") ||
       !ok(h, hipMemcpyAsync(W0.p, w0.data(), d.NZ * sizeof(float), hipMemcpyHostToDevice, h->stream), "copy") ||
       !ok(h, hipMemcpyAsync(W1.p, w1.data(), d.NZ * sizeof(float), hipMemcpyHostToDevice, h->stream), "copy"))
      return -1;
   if (!make_twiddles(h, s.twX, d.NX) || !make_twiddles(h, s.twY, d.NY) || !make_twiddles(h, s.twZ, d.NZ))
      return -1;
   const size_t nK = (size_t)nzP * D * D;
   hipLaunchKernelGGL(wf_resample_kernel, dim3(grid_for(nK, 256, 8)), dim3(256), 0, h->stream, dPlanes, size, nzP, r,
                      s0, R, K.as<float>());
   const size_t nG = (size_t)d.NZ * d.NY * d.NX;
   hipLaunchKernelGGL(wf_place_kernel, dim3(grid_for(nG, 256, 16)), dim3(256), 0, h->stream, K.as<const float>(), R,
                      nzP, P0.as<const int>(), W0.as<const float>(), W1.as<const float>(), (int)d.NX, (int)d.NY,
                      (int)d.NZ, G.as<float>());
   if (!ok(h, hipGetLastError(), "plan kernels"))
      return -1;
   // scratch: X1 holds [max(nzD, nF, NZ for the plan)][NY][S]-sized row data, X2 the column-pass output
   const size_t zMax = std::max<size_t>(d.NZ, std::max(d.nzD, d.nF));
   if (!ok(h, s.ghat.reserve((size_t)d.NZ * d.NY * S * sizeof(float2)), "spectrum allocation") ||
       !ok(h, s.x1.reserve(zMax * d.NY * S * sizeof(float2)), "scratch allocation") ||
       !ok(h, s.x2.reserve(zMax * d.NY * S * sizeof(float2)), "scratch allocation"))
      return -1;
   // the kernel grid's spectrum: rows and columns (all NZ planes), then z forward x 1/(NX NY NZ)
   if (!forward_xy(h, grid_of(s), G.as<const float>(), (int)d.NZ, (int)d.NY, (int)d.NX, s.x1.as<float2>(),
                   s.x2.as<float2>(), (size_t)d.NY * S))
      return fail(h, "iscg_wf_conv3d_plan: unsupported FFT size");
   const float scale = (float)(1.0 / ((double)d.NX * d.NY * d.NZ));
   if (!dispatch_n<ZConv>((int)d.NZ, h->stream, true, (const float2*)s.x2.as<float2>(), (int)d.NZ,
                          s.ghat.as<float2>(), (int)d.NZ, (const float2*)nullptr, (int)d.NY, (int)S,
                          (int)(d.NX / 2 + 1), scale, s.twZ.as<const float2>()))
      return fail(h, "iscg_wf_conv3d_plan: unsupported FFT size");
   if (!ok(h, hipGetLastError(), "plan transform") || !ok(h, hipStreamSynchronize(h->stream), "plan synchronize"))
      return -1;
   s.planned = true;
   return 0;
}

int32_t iscg_wf_conv3d(iscg_context* h, const float* density, const iscg_camera* cam, const float* offset,
                       double background, uint32_t seed, uint32_t frameBase, uint32_t pixelBase, float* outImages,
                       uint16_t* outAdu)
{
   if (begin(h))
      return -1;
   if (!h->wf || !h->wf->planned)
      return fail(h, "iscg_wf_conv3d: no plan (iscg_wf_conv3d_plan)");
   if (!density || (!outImages && !outAdu) || (outAdu && !cam))
      return fail(h, "iscg_wf_conv3d: null argument");
   WfState& s = *h->wf;
   const iscg_wf_desc& d = s.d;
   const int S = (int)s.S;
   const size_t nOut = (size_t)d.nF * d.ch * d.cw;
   const float* dIn = (const float*)stage_in(h, 0, density, (size_t)d.nzD * d.ny * d.nx * sizeof(float));
   const float* dOff = offset ? (const float*)stage_in(h, 1, offset, (size_t)d.ch * d.cw * sizeof(float)) : nullptr;
   float* dImg = outImages ? (float*)stage_out(h, 2, outImages, nOut * sizeof(float)) : nullptr;
   uint16_t* dAdu = outAdu ? (uint16_t*)stage_out(h, 3, outAdu, nOut * sizeof(uint16_t)) : nullptr;
   if (!dIn || (offset && !dOff) || (outImages && !dImg) || (outAdu && !dAdu))
      return -1;
   float2* x1 = s.x1.as<float2>();
   float2* x2 = s.x2.as<float2>();
   // A + B: [nzD][ny][nx] -> x2 [nzD][NY][S]
   if (!forward_xy(h, grid_of(s), dIn, (int)d.nzD, (int)d.ny, (int)d.nx, x1, x2, (size_t)d.NY * S))
      return fail(h, "iscg_wf_conv3d: unsupported FFT size");
   // C: x2 -> x1 [nF][NY][S]
   {
   ProfScope ps(h, "wf_zconv");
   dispatch_n<ZConv>((int)d.NZ, h->stream, false, (const float2*)x2, (int)d.nzD, x1, (int)d.nF,
                     s.ghat.as<const float2>(), (int)d.NY, S, (int)(d.NX / 2 + 1), 1.0f, s.twZ.as<const float2>());
   }
   // D + E: x1 [nF][NY][S] -> x2 [nF][ch][S] (rows fovY0 ..) -> images / ADU
   C2rEpilogue e;
   e.raw = 0;
   e.img = dImg;
   e.adu = dAdu;
   e.offset = dOff;
   e.background = background;
   e.cp = CamParams{};
   if (cam && !cam_params_for(h, cam, e.cp))
      return -1;
   e.seed = seed;
   e.frameBase = frameBase;
   e.pixelBase = pixelBase;
   e.ch = (int)d.ch;
   e.cw = (int)d.cw;
   e.x0 = (int)d.fovX0;
   e.q = NoiseQueue{nullptr, nullptr};
   if (dAdu && nOut > 0xFFFFFFFFull)
      return fail(h, "iscg_wf_conv3d: ADU output holds at most 2^32 - 1 pixels");
   if (dAdu && !noise_queue(h, nOut, e.q))
      return -1;
   inverse_xy(h, grid_of(s), x1, (int)d.nF, x2, (int)d.fovY0, e, "wf_rows_c2r_noise");
   if (!ok(h, hipGetLastError(), "iscg_wf_conv3d launch"))
      return -1;
   if (dAdu && !launch_noise_fallback(h, e.q, dAdu, d.ch * d.cw, nullptr, frameBase, dOff, nullptr, nullptr, e.cp,
                                      seed, pixelBase))
      return -1;
   if (outImages && !finish_out(h, outImages, dImg, nOut * sizeof(float)))
      return -1;
   if (outAdu && !finish_out(h, outAdu, dAdu, nOut * sizeof(uint16_t)))
      return -1;
   return 0;
}

}  // extern "C"
