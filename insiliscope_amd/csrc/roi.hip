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
// One workgroup (512 lanes) per CU, ROIs taken in a fixed order. The four 64 x 64 x 64 complex products run on FP32
// MFMA from LDS (sizes padded to 64 with zero W rows); the pupil phase is FP64, wrapped before the FP32 sincos. The
// shared Zernike gradient is the dot product of dNLL/dphase with each basis function: each workgroup sums its ROIs'
// phase-gradient maps (FP64), and two small kernels add the maps in workgroup order and take the 28 dot products,
// so results do not depend on scheduling.
#include "iscg_internal.h"
#include "fft_lds.h"

#include <cmath>
#include <string>

namespace iscg {

constexpr int kN = 64;     // padded matrix edge
constexpr int kLS = kN + 1; // LDS row (float2)
constexpr double kPiR = 3.14159265358979323846;
constexpr int kPartials = 28 + 1;  // Zernike gradient, NLL
constexpr int kRoiGrid = 256;      // workgroups of the ROI kernels: one per CU, ROIs assigned in a fixed order

struct RoiState
{
   DevBuf A, B, Zb, kx, W, partial;
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
// OSA slot; kx; W (FP32, zero beyond Ps / M).
__global__ void roi_tables_kernel(int M, int Ps, int os, double k0, double NA, double ni, double ns, double bMax,
                                  double kMax, double dk, double kMin, double depth, double pix, const double* zern,
                                  double* A, double* B, float* Zb, double* kx, float2* W)
{
   const int total = kN * kN;
   for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += gridDim.x * blockDim.x)
   {
      const int iy = idx / kN, ix = idx % kN;
      // W[s][m]
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

// ---- complex 64 x 64 x 64 products on FP32 MFMA -------------------------------------------------------------------
// 512 lanes = 8 waves; wave w computes rows 16 (w >> 1) .. + 15 and columns 32 (w & 1) .. + 31 of C as two 16 x 16
// tiles with v_mfma_f32_16x16x4_f32 (A operand: lane l holds A[l % 16][k0 + l / 16]; B: B[k0 + l / 16][l % 16];
// D register r: row 4 (l / 16) + r, column l % 16). A complex product is four real ones:
// C_re += a_re b_re + (-a_im) b_im, C_im += a_re b_im + a_im b_re.
constexpr int kRoiThreads = 512;
typedef float float4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float4v mfma4(float a, float b, float4v c)
{
   return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

struct CTile
{
   float4v re[2], im[2];  // tiles t = 0, 1: element (row 16 rb + 4 (lane >> 4) + r, col 32 ch + 16 t + (lane & 15))
};

// C = A B over k < 64, operands in LDS (row stride kLS). TA: A(i, k) = Al[k][i] (else Al[i][k]); TB: B(k, j) =
// Bl[j][k] (else Bl[k][j]); CA / CB: conjugate the operand.
template <bool TA, bool CA, bool TB, bool CB>
__device__ __forceinline__ void cgemm_mfma(const float2* Al, const float2* Bl, CTile& c, int rb, int ch, int lane)
{
#pragma unroll
   for (int t = 0; t < 2; ++t)
      c.re[t] = c.im[t] = float4v{0.0f, 0.0f, 0.0f, 0.0f};
   const int i = 16 * rb + (lane & 15), kk = lane >> 4;
   const int j0 = 32 * ch + (lane & 15), j1 = j0 + 16;
#pragma unroll 4
   for (int k0 = 0; k0 < kN; k0 += 4)
   {
      const int k = k0 + kk;
      float2 a = TA ? Al[k * kLS + i] : Al[i * kLS + k];
      float2 b0 = TB ? Bl[j0 * kLS + k] : Bl[k * kLS + j0];
      float2 b1 = TB ? Bl[j1 * kLS + k] : Bl[k * kLS + j1];
      if (CA)
         a.y = -a.y;
      if (CB)
      {
         b0.y = -b0.y;
         b1.y = -b1.y;
      }
      c.re[0] = mfma4(a.x, b0.x, c.re[0]);
      c.re[1] = mfma4(a.x, b1.x, c.re[1]);
      c.im[0] = mfma4(a.x, b0.y, c.im[0]);
      c.im[1] = mfma4(a.x, b1.y, c.im[1]);
      c.re[0] = mfma4(-a.y, b0.y, c.re[0]);
      c.re[1] = mfma4(-a.y, b1.y, c.re[1]);
      c.im[0] = mfma4(a.y, b0.x, c.im[0]);
      c.im[1] = mfma4(a.y, b1.x, c.im[1]);
   }
}

__device__ __forceinline__ void store_ctile(float2* buf, const CTile& c, int rb, int ch, int lane)
{
#pragma unroll
   for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r)
         buf[(16 * rb + 4 * (lane >> 4) + r) * kLS + 32 * ch + 16 * t + (lane & 15)] = make_float2(c.re[t][r], c.im[t][r]);
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
   for (int o = 32; o > 0; o >>= 1)
      v += __shfl_xor(v, o, 64);
   return v;
}

constexpr int kMapStride = kN * kN + 1;  // per workgroup: the phase-gradient map, then the NLL

// One workgroup per CU, ROIs blockIdx.x, + gridDim.x, ... in order. Backward: per ROI the x, y, z, N, b gradients;
// per workgroup the sum over its ROIs of the phase-gradient map dNLL/dphase (FP64, [n][m]) and of the NLL, reduced
// afterwards in workgroup order (the shared Zernike gradient is the map's dot product with each basis function).
template <bool BACKWARD>
__global__ __launch_bounds__(kRoiThreads) void roi_kernel(RoiConst rc, const iscg_roi_params* __restrict__ params,
                                                          int nRoi, const double* __restrict__ A,
                                                          const double* __restrict__ B, const double* __restrict__ kx,
                                                          const float2* __restrict__ W, const float* __restrict__ data,
                                                          float* __restrict__ mu, double* __restrict__ gradParams,
                                                          double* __restrict__ wgMaps)
{
   __shared__ float2 Pl[kN * kLS];  // P'
   __shared__ float2 Xl[kN * kLS];  // T, then |E|^2 (as float [64][64]), then G, then U
   __shared__ float2 Wl[kN * kLS];  // W
   __shared__ float wbuf[kN * kN];  // 1 - d / mu per pixel
   __shared__ double red[8][6];
   float* inten = reinterpret_cast<float*>(Xl);
   const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, rb = w >> 1, ch = w & 1;
   const int R = rc.R, os = rc.os, Ps = rc.Ps;
   for (int idx = tid; idx < kN * kN; idx += kRoiThreads)
      Wl[(idx / kN) * kLS + idx % kN] = W[idx];
   double gmap[2][4];  // this lane's elements of the phase-gradient map, summed over the workgroup's ROIs
   double nllWg = 0.0;
#pragma unroll
   for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r)
         gmap[t][r] = 0.0;
   __syncthreads();
   for (int roi = blockIdx.x; roi < nRoi; roi += gridDim.x)
   {
      const iscg_roi_params q = params[roi];
      const double tz = -rc.focalShift + q.z * 1e-9;
      // 1. the shifted pupil (FP64 phase)
      for (int idx = tid; idx < kN * kN; idx += kRoiThreads)
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
         Pl[n * kLS + m] = v;
      }
      __syncthreads();
      CTile c;
      // 2. T[n][s] = sum_m P'[n][m] W[s][m]
      cgemm_mfma<false, false, true, false>(Pl, Wl, c, rb, ch, lane);
      store_ctile(Xl, c, rb, ch, lane);
      __syncthreads();
      // 3. E[t][s] = sum_n W[t][n] T[n][s]  (kept in registers)
      cgemm_mfma<false, false, false, false>(Wl, Xl, c, rb, ch, lane);
      __syncthreads();
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
         for (int r = 0; r < 4; ++r)
            inten[(16 * rb + 4 * (lane >> 4) + r) * kN + 32 * ch + 16 * t + (lane & 15)] =
               c.re[t][r] * c.re[t][r] + c.im[t][r] * c.im[t][r];
      __syncthreads();
      // 4. pixels: mu, and for the adjoint w = 1 - d / mu, NLL, dN, db
      double nll = 0.0, gN = 0.0, gb = 0.0;
      for (int p = tid; p < R * R; p += kRoiThreads)
      {
         const int py = p / R, px = p % R;
         double acc = 0.0;
         for (int sy = 0; sy < os; ++sy)
            for (int sx = 0; sx < os; ++sx)
               acc += inten[(py * os + sy) * kN + px * os + sx];
         const double f = rc.norm * acc;
         const double m = q.photons * f + q.background;
         if (!BACKWARD)
            mu[(size_t)roi * R * R + p] = (float)m;
         else
         {
            const double d = data[(size_t)roi * R * R + p];
            nll += m - d * log(m);
            const double wv = 1.0 - d / m;
            wbuf[p] = (float)wv;
            gN += wv * f;
            gb += wv;
         }
      }
      __syncthreads();
      if (!BACKWARD)
         continue;
      // 5. G_s = 2 N norm w_p(s) E_s
      const float gscale = (float)(2.0 * q.photons * rc.norm);
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
         for (int r = 0; r < 4; ++r)
         {
            const int ty = 16 * rb + 4 * (lane >> 4) + r, s = 32 * ch + 16 * t + (lane & 15);
            const float g = (ty < Ps && s < Ps) ? gscale * wbuf[(ty / os) * R + s / os] : 0.0f;
            Xl[ty * kLS + s] = make_float2(g * c.re[t][r], g * c.im[t][r]);
         }
      __syncthreads();
      // 6. U[n][s] = sum_t conj(W[t][n]) G[t][s]
      cgemm_mfma<true, true, false, false>(Wl, Xl, c, rb, ch, lane);
      __syncthreads();
      store_ctile(Xl, c, rb, ch, lane);
      __syncthreads();
      // 7. B[n][m] = sum_s U[n][s] conj(W[s][m])
      cgemm_mfma<false, false, false, true>(Xl, Wl, c, rb, ch, lane);
      // 8. dNLL/dphase = -Im(P' conj(B)) (zero outside the aperture, where P' is zero); reductions
      double gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
         for (int r = 0; r < 4; ++r)
         {
            const int n = 16 * rb + 4 * (lane >> 4) + r, m = 32 * ch + 16 * t + (lane & 15);
            const float2 P = Pl[n * kLS + m];
            const double g = -((double)P.y * c.re[t][r] - (double)P.x * c.im[t][r]);
            gx += g * (-kx[m]) * rc.pix;
            gy += g * (-kx[n]) * rc.pix;
            gz += g * B[n * kN + m] * 1e-9;
            gmap[t][r] += g;
         }
      const double vals[6] = {gx, gy, gz, gN, gb, nll};
#pragma unroll
      for (int k = 0; k < 6; ++k)
      {
         const double s = wave_sum(vals[k]);
         if (lane == 0)
            red[w][k] = s;
      }
      __syncthreads();
      if (tid < 6)
      {
         double s = 0.0;
         for (int k = 0; k < 8; ++k)
            s += red[k][tid];
         if (tid < 5)
            gradParams[(size_t)roi * 5 + tid] = s;
         else
            nllWg += s;
      }
      __syncthreads();
   }
   if (BACKWARD)
   {
      double* map = wgMaps + (size_t)blockIdx.x * kMapStride;
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
         for (int r = 0; r < 4; ++r)
            map[(16 * rb + 4 * (lane >> 4) + r) * kN + 32 * ch + 16 * t + (lane & 15)] = gmap[t][r];
      if (tid == 5)
         map[kN * kN] = nllWg;
   }
}

// The workgroups' maps and NLLs summed in workgroup order (one lane per map element, and the NLL).
__global__ void roi_map_reduce_kernel(const double* __restrict__ wgMaps, int nWg, double* __restrict__ total)
{
   const int idx = blockIdx.x * blockDim.x + threadIdx.x;
   if (idx >= kMapStride)
      return;
   double s = 0.0;
   for (int g = 0; g < nWg; ++g)
      s += wgMaps[(size_t)g * kMapStride + idx];
   total[idx] = s;
}

// dNLL/dc_z = 2 pi sum_idx map[idx] Zb[z][idx] (one workgroup per OSA slot, fixed-order tree), and the NLL.
__global__ __launch_bounds__(256) void roi_zernike_kernel(const double* __restrict__ total, const float* __restrict__ Zb,
                                                          double* __restrict__ out)
{
   __shared__ double part[256];
   const int z = blockIdx.x, tid = threadIdx.x;
   double s = 0.0;
   for (int idx = tid; idx < kN * kN; idx += 256)
      s += total[idx] * (double)Zb[(size_t)z * kN * kN + idx];
   part[tid] = s;
   __syncthreads();
   for (int o = 128; o > 0; o >>= 1)
   {
      if (tid < o)
         part[tid] += part[tid + o];
      __syncthreads();
   }
   if (tid == 0)
   {
      out[z] = 2.0 * kPiR * part[0];
      if (z == 0)
         out[28] = total[kN * kN];
   }
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
          !ok(h, s.kx.reserve(kN * sizeof(double)), "alloc") || !ok(h, s.W.reserve(kN * kN * sizeof(float2)), "alloc"))
         return -1;
      DevBuf dz;
      if (!ok(h, dz.reserve(28 * sizeof(double)), "alloc") ||
          !ok(h, hipMemcpyAsync(dz.p, zernike, 28 * sizeof(double), hipMemcpyHostToDevice, h->stream), "copy"))
         return -1;
      {
         ProfScope ps(h, "roi_tables");
         hipLaunchKernelGGL(roi_tables_kernel, dim3(16), dim3(256), 0, h->stream, M, R * os, os, k0, NA, ni, ns, bMax,
                            kMax, dk, kMin, depth, pix, dz.as<const double>(), s.A.as<double>(), s.B.as<double>(),
                            s.Zb.as<float>(), s.kx.as<double>(), s.W.as<float2>());
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
      hipLaunchKernelGGL((roi_kernel<false>), dim3((unsigned)std::min(nRoi, kRoiGrid)), dim3(kRoiThreads), 0,
                         h->stream, rc, dP, nRoi, s.A.as<const double>(), s.B.as<const double>(),
                         s.kx.as<const double>(), s.W.as<const float2>(), (const float*)nullptr, dMu, (double*)nullptr,
                         (double*)nullptr);
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
   const int nWg = std::min(nRoi, kRoiGrid);
   const size_t mapDoubles = (size_t)(nWg + 1) * kMapStride;
   if (!dP || !dD || !dG || !ok(h, s.partial.reserve((mapDoubles + kPartials) * sizeof(double)), "alloc"))
      return -1;
   double* total = s.partial.as<double>() + (size_t)nWg * kMapStride;
   double* sums = s.partial.as<double>() + mapDoubles;
   {
      ProfScope ps(h, "roi_backward");
      hipLaunchKernelGGL((roi_kernel<true>), dim3((unsigned)nWg), dim3(kRoiThreads), 0, h->stream, rc, dP, nRoi,
                         s.A.as<const double>(), s.B.as<const double>(), s.kx.as<const double>(),
                         s.W.as<const float2>(), dD, (float*)nullptr, dG, s.partial.as<double>());
   }
   {
      ProfScope ps(h, "roi_reduce");
      hipLaunchKernelGGL(roi_map_reduce_kernel, dim3((kMapStride + 255) / 256), dim3(256), 0, h->stream,
                         s.partial.as<const double>(), nWg, total);
      hipLaunchKernelGGL(roi_zernike_kernel, dim3(28), dim3(256), 0, h->stream, (const double*)total,
                         s.Zb.as<const float>(), sums);
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
