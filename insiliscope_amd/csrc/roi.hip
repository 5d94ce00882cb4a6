// roi.hip -- differentiable forward model of emitter ROIs and its hand-written adjoint (Poisson NLL gradients with
// respect to the Zernike coefficients, emitter x, y, z, photons and background). New capability: the reference has
// only a finite-difference Fisher matrix (DescribePsfCramerRao, Sim/PsfGeneratorBridge.cpp:1554-1669).
//
// Model, per ROI (the reference's scalar pupil, Sim/ZernikePsf.cpp:163-248, at the emitter's defocus z):
//   P'_mn = A_mn exp(i (base_mn + (ti - ti0) slope_mn - (k_m x + k_n y) pixel))   (the shift folded into the pupil)
//   E[t][s] = sum_n W[t][n] sum_m P'[n][m] W[s][m],  W[s][m] = exp(i k_m x_s), x_s = (s + 0.5)/os pixels
//   mu_p = N norm sum_{s in p} |E_s|^2 + b,  norm = (dxs dk / 2 pi)^2 / n_inside
//   NLL = sum_p mu_p - d_p log mu_p
// Adjoint: G_s = 2 N norm (1 - d_p/mu_p) E_s; B = W^H G conj(W) (the transposed DFTs); dNLL/dphase_mn =
//   -Im(P'_mn conj(B_mn)); x, y, z and the Zernike coefficients are linear in the phase, so their gradients are
//   weighted sums of that map. The pupil is recomputed in the backward pass, never stored.
//
// One workgroup (256 lanes) per ROI, grid-stride over ROIs. The four 64 x 64 x 64 complex products run in FP32
// from LDS with 4 x 4 register tiles (sizes padded to 64 with zero W rows); the pupil phase is FP64, wrapped before
// the FP32 sincos. Per-ROI partial sums (Zernike gradient, NLL) are reduced by a second kernel in a fixed order, so
// results do not depend on scheduling.
#include "iscg_internal.h"
#include "fft_lds.h"

#include <cmath>
#include <string>

namespace iscg {

constexpr int kN = 64;     // padded matrix edge
constexpr int kLS = kN + 1; // LDS row (float2)
constexpr double kPiR = 3.14159265358979323846;

struct RoiState
{
   DevBuf A, B, Zb, kx, W, WT, partial;
   iscg_roi_model model{};
   double zern[28] = {0};
   bool tablesValid = false;
   int nIn = 0;
   double norm = 0.0, focalShift = 0.0, pix = 0.0;
};

void destroy_roi(iscg_context* h)
{
   delete h->roi;
   h->roi = nullptr;
}

struct RoiConst
{
   double norm, focalShift, pix;
   int M, R, os, Ps;
};

// Pupil tables, FP64: A (base phase, NaN outside), B (phase per metre of ti - ti0), Zernike basis values (FP32) per
// OSA slot; kx; W and W^T (FP32, zero beyond Ps / M).
__global__ void roi_tables_kernel(int M, int Ps, int os, double k0, double NA, double ni, double ns, double bMax,
                                  double kMax, double dk, double kMin, double depth, double pix, const double* zern,
                                  double* A, double* B, float* Zb, double* kx, float2* W, float2* WT)
{
   const int total = kN * kN;
   for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += gridDim.x * blockDim.x)
   {
      const int iy = idx / kN, ix = idx % kN;
      // W[s][m] and W^T
      {
         const int s = iy, m = ix;
         float2 w = make_float2(0.0f, 0.0f);
         if (s < Ps && m < M)
         {
            const double xs = (s + 0.5) / os * pix;
            const double ang = (kMin + m * dk) * xs;
            const double r = ang - 6.283185307179586476925 * rint(ang * 0.15915494309189533577);
            float sn, cs;
            sincosf((float)r, &sn, &cs);
            w = make_float2(cs, sn);
         }
         W[s * kN + m] = w;
         WT[m * kN + s] = w;
      }
      if (iy == 0 && ix < kN)
         kx[ix] = kMin + ix * dk;
      double a = __builtin_nan(""), b = 0.0;
      const int c0 = M / 2;
      if (iy < M && ix < M)
      {
         const double ky = (iy - c0) * dk, kxx = (ix - c0) * dk;
         const double kr2 = kxx * kxx + ky * ky;
         if (!(kr2 > kMax * kMax))
         {
            const double rho = sqrt(kr2) / (k0 * NA);
            const double phi = atan2(ky, kxx);
            const double s1 = NA * rho / ns, s3 = NA * rho / ni;
            const double a1 = 1.0 - s1 * s1, a3 = 1.0 - s3 * s3;
            const double opd1 = ns * depth * sqrt(0.0 < a1 ? a1 : 0.0);
            const double opd3 = sqrt(0.0 < a3 ? a3 : 0.0);
            double zp = 0.0;
            for (int j = 0; j < 28; ++j)
            {
               int n = 0, l = 0;
               for (int nn = 0;; ++nn)
               {
                  if (j < (nn + 1) * (nn + 2) / 2)
                  {
                     n = nn;
                     l = -nn + 2 * (j - nn * (nn + 1) / 2);
                     break;
                  }
               }
               const int m = l < 0 ? -l : l;
               double radial = 0.0;
               const double rr = rho / bMax;
               if (rr <= 1.0)
                  for (int k = 0; k <= (n - m) / 2; ++k)
                  {
                     double b1 = 1.0, b2 = 1.0;
                     for (int i = 0; i < k; ++i)
                        b1 = b1 * (n - k - i) / (i + 1);
                     const int kk = (n - m) / 2 - k;
                     for (int i = 0; i < kk; ++i)
                        b2 = b2 * (n - 2 * k - i) / (i + 1);
                     radial += (k % 2 == 0 ? 1.0 : -1.0) * b1 * b2 * pow(rr, (double)(n - 2 * k));
                  }
               const double zv = radial * (l >= 0 ? cos(m * phi) : sin(m * phi));
               Zb[(size_t)j * kN * kN + idx] = (float)zv;
               zp += zern[j] * zv;
            }
            a = k0 * opd1 + 2.0 * kPiR * zp;
            b = k0 * (ni * opd3);
         }
      }
      if (!(a == a))
         for (int j = 0; j < 28; ++j)
            Zb[(size_t)j * kN * kN + idx] = 0.0f;
      A[idx] = a;
      B[idx] = b;
   }
}

// C (4 x 4 complex tile per lane, rows 4 ti.., cols 4 tj..) = sum_k A[i][k] B[k][j] over k < 64.
// AL/BL: operand in LDS (row stride kLS) or global (row stride kN); CA/CB: conjugate the operand.
template <bool AL, bool CA, bool BL, bool CB>
__device__ __forceinline__ void cgemm(const float2* __restrict__ Am, const float2* __restrict__ Bm, float2 (&c)[4][4],
                                      int ti, int tj)
{
#pragma unroll
   for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
         c[i][j] = make_float2(0.0f, 0.0f);
   constexpr int sa = AL ? kLS : kN, sb = BL ? kLS : kN;
#pragma unroll 4
   for (int k = 0; k < kN; ++k)
   {
      float2 a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i)
      {
         a[i] = Am[(4 * ti + i) * sa + k];
         if (CA)
            a[i].y = -a[i].y;
         b[i] = Bm[k * sb + 4 * tj + i];
         if (CB)
            b[i].y = -b[i].y;
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
         for (int j = 0; j < 4; ++j)
         {
            c[i][j].x += a[i].x * b[j].x - a[i].y * b[j].y;
            c[i][j].y += a[i].x * b[j].y + a[i].y * b[j].x;
         }
   }
}

__device__ __forceinline__ void store_tile(float2* buf, const float2 (&c)[4][4], int ti, int tj)
{
#pragma unroll
   for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
         buf[(4 * ti + i) * kLS + 4 * tj + j] = c[i][j];
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
   for (int o = 32; o > 0; o >>= 1)
      v += __shfl_xor(v, o, 64);
   return v;
}

// Sums v over the workgroup (every lane gets the total). red: 4 doubles of LDS.
__device__ __forceinline__ double block_sum(double v, double* red)
{
   v = wave_sum(v);
   const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
   __syncthreads();
   if (lane == 0)
      red[wave] = v;
   __syncthreads();
   return red[0] + red[1] + red[2] + red[3];
}

constexpr int kPartials = 28 + 1;  // Zernike gradient, NLL

template <bool BACKWARD>
__global__ __launch_bounds__(256) void roi_kernel(RoiConst rc, const iscg_roi_params* __restrict__ params, int nRoi,
                                                  const double* __restrict__ A, const double* __restrict__ B,
                                                  const float* __restrict__ Zb, const double* __restrict__ kx,
                                                  const float2* __restrict__ W, const float2* __restrict__ WT,
                                                  const float* __restrict__ data, float* __restrict__ mu,
                                                  double* __restrict__ gradParams, double* __restrict__ partial)
{
   __shared__ float2 buf0[kN * kLS];   // P'
   __shared__ float2 buf1[kN * kLS];   // T, then G, then U
   __shared__ float inten[kN * kN];    // |E|^2 per sample
   __shared__ float wbuf[kN * kN];     // 1 - d / mu per pixel
   __shared__ double red[4];
   const int tid = threadIdx.x, ti = tid >> 4, tj = tid & 15;
   const int M = rc.M, R = rc.R, os = rc.os, Ps = rc.Ps;
   for (int r = blockIdx.x; r < nRoi; r += gridDim.x)
   {
      const iscg_roi_params q = params[r];
      const double tz = -rc.focalShift + q.z * 1e-9;
      // 1. the shifted pupil (FP64 phase)
      for (int idx = tid; idx < kN * kN; idx += 256)
      {
         const int n = idx / kN, m = idx % kN;
         float2 v = make_float2(0.0f, 0.0f);
         const double a = A[idx];
         if (a == a)
         {
            const double ph = a + tz * B[idx] - (kx[m] * q.x + kx[n] * q.y) * rc.pix;
            const double rr = ph - 6.283185307179586476925 * rint(ph * 0.15915494309189533577);
            float sn, cs;
            sincosf((float)rr, &sn, &cs);
            v = make_float2(cs, sn);
         }
         buf0[n * kLS + m] = v;
      }
      __syncthreads();
      float2 c[4][4];
      // 2. T[n][s] = sum_m P'[n][m] W^T[m][s]
      cgemm<true, false, false, false>(buf0, WT, c, ti, tj);
      store_tile(buf1, c, ti, tj);
      __syncthreads();
      // 3. E[t][s] = sum_n W[t][n] T[n][s]  (kept in registers)
      cgemm<false, false, true, false>(W, buf1, c, ti, tj);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
         for (int j = 0; j < 4; ++j)
            inten[(4 * ti + i) * kN + 4 * tj + j] = c[i][j].x * c[i][j].x + c[i][j].y * c[i][j].y;
      __syncthreads();
      // 4. pixels: mu, and for the adjoint w = 1 - d / mu, NLL, dN, db
      double nll = 0.0, gN = 0.0, gb = 0.0;
      for (int p = tid; p < R * R; p += 256)
      {
         const int py = p / R, px = p % R;
         double acc = 0.0;
         for (int sy = 0; sy < os; ++sy)
            for (int sx = 0; sx < os; ++sx)
               acc += inten[(py * os + sy) * kN + px * os + sx];
         const double f = rc.norm * acc;
         const double m = q.photons * f + q.background;
         if (!BACKWARD)
            mu[(size_t)r * R * R + p] = (float)m;
         else
         {
            const double d = data[(size_t)r * R * R + p];
            nll += m - d * log(m);
            const double w = 1.0 - d / m;
            wbuf[p] = (float)w;
            gN += w * f;
            gb += w;
         }
      }
      if (!BACKWARD)
      {
         __syncthreads();
         continue;
      }
      __syncthreads();
      // 5. G_s = 2 N norm w_p(s) E_s
      const float gscale = (float)(2.0 * q.photons * rc.norm);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
         for (int j = 0; j < 4; ++j)
         {
            const int t = 4 * ti + i, s = 4 * tj + j;
            const float g = (t < Ps && s < Ps) ? gscale * wbuf[(t / os) * R + s / os] : 0.0f;
            buf1[t * kLS + s] = make_float2(g * c[i][j].x, g * c[i][j].y);
         }
      __syncthreads();
      // 6. U[n][s] = sum_t conj(W[t][n]) G[t][s] = sum_t conj(W^T[n][t]) G[t][s]
      cgemm<false, true, true, false>(WT, buf1, c, ti, tj);
      __syncthreads();
      store_tile(buf1, c, ti, tj);
      __syncthreads();
      // 7. B[n][m] = sum_s U[n][s] conj(W[s][m])
      cgemm<true, false, false, true>(buf1, W, c, ti, tj);
      // 8. dNLL/dphase = -Im(P' conj(B)); reductions
      double gx = 0.0, gy = 0.0, gz = 0.0;
      double gzern[28];
#pragma unroll
      for (int j = 0; j < 28; ++j)
         gzern[j] = 0.0;
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
         for (int j = 0; j < 4; ++j)
         {
            const int n = 4 * ti + i, m = 4 * tj + j;
            const int idx = n * kN + m;
            const float2 P = buf0[n * kLS + m];
            const double g = -((double)P.y * c[i][j].x - (double)P.x * c[i][j].y);
            if (n < M && m < M && A[idx] == A[idx])
            {
               gx += g * (-kx[m]) * rc.pix;
               gy += g * (-kx[n]) * rc.pix;
               gz += g * B[idx] * 1e-9;
#pragma unroll
               for (int z = 0; z < 28; ++z)
                  gzern[z] += g * (2.0 * kPiR) * (double)Zb[(size_t)z * kN * kN + idx];
            }
         }
      gx = block_sum(gx, red);
      gy = block_sum(gy, red);
      gz = block_sum(gz, red);
      gN = block_sum(gN, red);
      gb = block_sum(gb, red);
      nll = block_sum(nll, red);
      double* part = partial + (size_t)r * kPartials;
#pragma unroll 1
      for (int z = 0; z < 28; ++z)
      {
         const double s = block_sum(gzern[z], red);
         if (tid == 0)
            part[z] = s;
      }
      if (tid == 0)
      {
         part[28] = nll;
         double* gp = gradParams + (size_t)r * 5;
         gp[0] = gx;
         gp[1] = gy;
         gp[2] = gz;
         gp[3] = gN;
         gp[4] = gb;
      }
      __syncthreads();
   }
}

// Sum of the per-ROI partials in ROI order (one lane per quantity).
__global__ void roi_reduce_kernel(const double* __restrict__ partial, int nRoi, double* __restrict__ out)
{
   const int q = threadIdx.x;
   if (q >= kPartials)
      return;
   double s = 0.0;
   for (int r = 0; r < nRoi; ++r)
      s += partial[(size_t)r * kPartials + q];
   out[q] = s;
}

static int32_t prepare(iscg_context* h, const iscg_roi_model* m, const double* zernike, RoiConst& rc)
{
   if (!m || !zernike)
      return fail(h, "roi: null model");
   const int M = m->pupilSamples, R = m->roiSize, os = m->oversampling;
   if (M < 8 || M > kN || (M % 2) != 0 || R < 1 || os < 1 || R * os > kN)
      return fail(h, "roi: need even pupil samples <= 64 and roiSize * oversampling <= 64");
   if (!(m->na > 0.0) || !(m->wavelengthNm > 0.0) || !(m->immersionIndex > 0.0) || !(m->sampleIndex > 0.0) ||
       !(m->pixelSizeNm > 0.0))
      return fail(h, "roi: NA, wavelength, refractive indices and pixel size must be > 0");
   if (!h->roi)
      h->roi = new RoiState();
   RoiState& s = *h->roi;
   const double k0 = 2.0 * kPiR / (m->wavelengthNm * 1e-9);
   const double NA = m->na, ni = m->immersionIndex, ns = m->sampleIndex;
   const double bMax = std::min(1.0, ns / NA);
   const double kMax = k0 * NA * bMax;
   const double dk = (2.0 * kMax) / (M - 4);
   const double kMin = -std::floor(M / 2.0) * dk;
   const double depth = m->sampleDepthNm * 1e-9, pix = m->pixelSizeNm * 1e-9;
   const bool same = s.tablesValid && std::memcmp(&s.model, m, sizeof(*m)) == 0 &&
                     std::memcmp(s.zern, zernike, sizeof(s.zern)) == 0;
   if (!same)
   {
      if (!ok(h, s.A.reserve(kN * kN * sizeof(double)), "alloc") ||
          !ok(h, s.B.reserve(kN * kN * sizeof(double)), "alloc") ||
          !ok(h, s.Zb.reserve((size_t)28 * kN * kN * sizeof(float)), "alloc") ||
          !ok(h, s.kx.reserve(kN * sizeof(double)), "alloc") || !ok(h, s.W.reserve(kN * kN * sizeof(float2)), "alloc") ||
          !ok(h, s.WT.reserve(kN * kN * sizeof(float2)), "alloc"))
         return -1;
      DevBuf dz;
      if (!ok(h, dz.reserve(28 * sizeof(double)), "alloc") ||
          !ok(h, hipMemcpyAsync(dz.p, zernike, 28 * sizeof(double), hipMemcpyHostToDevice, h->stream), "copy"))
         return -1;
      {
         ProfScope ps(h, "roi_tables");
         hipLaunchKernelGGL(roi_tables_kernel, dim3(16), dim3(256), 0, h->stream, M, R * os, os, k0, NA, ni, ns, bMax,
                            kMax, dk, kMin, depth, pix, dz.as<const double>(), s.A.as<double>(), s.B.as<double>(),
                            s.Zb.as<float>(), s.kx.as<double>(), s.W.as<float2>(), s.WT.as<float2>());
      }
      if (!ok(h, hipGetLastError(), "roi tables") || !ok(h, hipStreamSynchronize(h->stream), "synchronize"))
         return -1;
      // points inside the aperture (the normalisation)
      int nIn = 0;
      const int c0 = M / 2;
      for (int iy = 0; iy < M; ++iy)
         for (int ix = 0; ix < M; ++ix)
         {
            const double ky = (iy - c0) * dk, kx = (ix - c0) * dk;
            if (!(kx * kx + ky * ky > kMax * kMax))
               ++nIn;
         }
      s.nIn = nIn;
      const double qd = (pix / os) * dk / (2.0 * kPiR);
      s.norm = qd * qd / nIn;
      s.focalShift = depth * (ni / ns);
      s.pix = pix;
      s.model = *m;
      std::memcpy(s.zern, zernike, sizeof(s.zern));
      s.tablesValid = true;
   }
   rc.norm = s.norm;
   rc.focalShift = s.focalShift;
   rc.pix = s.pix;
   rc.M = M;
   rc.R = R;
   rc.os = os;
   rc.Ps = R * os;
   return 0;
}

}  // namespace iscg

using namespace iscg;

extern "C" {

int32_t iscg_roi_forward(iscg_context* h, const iscg_roi_model* model, const double* zernike,
                         const iscg_roi_params* params, int32_t nRoi, float* mu)
{
   if (begin(h))
      return -1;
   if (!params || !mu || nRoi < 0)
      return fail(h, "iscg_roi_forward: null argument");
   RoiConst rc;
   if (prepare(h, model, zernike, rc))
      return -1;
   if (nRoi == 0)
      return 0;
   RoiState& s = *h->roi;
   const size_t npx = (size_t)nRoi * rc.R * rc.R;
   const iscg_roi_params* dP = (const iscg_roi_params*)stage_in(h, 0, params, nRoi * sizeof(iscg_roi_params));
   float* dMu = (float*)stage_out(h, 1, mu, npx * sizeof(float));
   if (!dP || !dMu)
      return -1;
   {
      ProfScope ps(h, "roi_forward");
      hipLaunchKernelGGL((roi_kernel<false>), dim3((unsigned)std::min(nRoi, 256 * 8)), dim3(256), 0, h->stream, rc, dP,
                         nRoi, s.A.as<const double>(), s.B.as<const double>(), s.Zb.as<const float>(),
                         s.kx.as<const double>(), s.W.as<const float2>(), s.WT.as<const float2>(), (const float*)nullptr,
                         dMu, (double*)nullptr, (double*)nullptr);
   }
   if (!ok(h, hipGetLastError(), "roi forward"))
      return -1;
   return finish_out(h, mu, dMu, npx * sizeof(float)) ? 0 : -1;
}

int32_t iscg_roi_backward(iscg_context* h, const iscg_roi_model* model, const double* zernike,
                          const iscg_roi_params* params, int32_t nRoi, const float* data, double* gradParams,
                          double* gradZernike, double* nll)
{
   if (begin(h))
      return -1;
   if (!params || !data || !gradParams || !gradZernike || !nll || nRoi < 0)
      return fail(h, "iscg_roi_backward: null argument");
   RoiConst rc;
   if (prepare(h, model, zernike, rc))
      return -1;
   if (nRoi == 0)
   {
      std::memset(gradZernike, 0, 28 * sizeof(double));
      *nll = 0.0;
      return 0;
   }
   RoiState& s = *h->roi;
   const size_t npx = (size_t)nRoi * rc.R * rc.R;
   const iscg_roi_params* dP = (const iscg_roi_params*)stage_in(h, 0, params, nRoi * sizeof(iscg_roi_params));
   const float* dD = (const float*)stage_in(h, 1, data, npx * sizeof(float));
   double* dG = (double*)stage_out(h, 2, gradParams, (size_t)nRoi * 5 * sizeof(double));
   if (!dP || !dD || !dG || !ok(h, s.partial.reserve(((size_t)nRoi + 1) * kPartials * sizeof(double)), "alloc"))
      return -1;
   double* sums = s.partial.as<double>() + (size_t)nRoi * kPartials;
   {
      ProfScope ps(h, "roi_backward");
      hipLaunchKernelGGL((roi_kernel<true>), dim3((unsigned)std::min(nRoi, 256 * 8)), dim3(256), 0, h->stream, rc, dP,
                         nRoi, s.A.as<const double>(), s.B.as<const double>(), s.Zb.as<const float>(),
                         s.kx.as<const double>(), s.W.as<const float2>(), s.WT.as<const float2>(), dD, (float*)nullptr,
                         dG, s.partial.as<double>());
      hipLaunchKernelGGL(roi_reduce_kernel, dim3(1), dim3(64), 0, h->stream, s.partial.as<const double>(), nRoi, sums);
   }
   if (!ok(h, hipGetLastError(), "roi backward"))
      return -1;
   double host[kPartials];
   if (!ok(h, hipMemcpyAsync(host, sums, sizeof(host), hipMemcpyDeviceToHost, h->stream), "copy") ||
       !ok(h, hipStreamSynchronize(h->stream), "synchronize"))
      return -1;
   std::memcpy(gradZernike, host, 28 * sizeof(double));
   *nll = host[28];
   return finish_out(h, gradParams, dG, (size_t)nRoi * 5 * sizeof(double)) ? 0 : -1;
}

}  // extern "C"
