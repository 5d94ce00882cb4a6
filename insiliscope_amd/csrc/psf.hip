// psf.hip -- the Gibson-Lanni + Zernike PSF planes (ComputeZernikePsfPlanes, Sim/ZernikePsf.cpp:150-337) for a
// batch of requests x z planes.
//
// The reference takes the pupil to the image plane by a separable Bluestein chirp-Z transform in FP64
// (Sim/ChirpZ.h:28-141): out[p] = sum_m in[m] exp(i (kMin + m dk)(x0 + p dx)). That is a dense P x M complex
// matrix W, so a plane is E = W Pup W^T: two complex matrix products per plane, the shape MFMA is built for.
//
//   tables (per request, FP64, Sim/ZernikePsf.cpp:199-231): the z-independent phase A = k0 opd1 + 2 pi Zernike + mask
//     and slope B = k0 ni opd3Factor on the M x M pupil grid (NaN outside the aperture);
//   W (per request): exp(i (kMin + m dk)(x0 + p dx)), the argument in FP64 and wrapped to [-pi, pi] before the FP32
//     sincos, stored as fp16 hi + lo parts (hi = fp16(v), lo = fp16(v - hi): ~22 significant bits);
//   planes: one workgroup per (request, plane, block of output rows). Per 32-row chunk of the pupil the FP64
//     phase A + (ti - ti0) B is wrapped and turned into the split pupil in LDS; mid = W_rows Pup and then
//     E = mid W^T run on v_mfma_f32_16x16x32_f16 as hi*hi + hi*lo + lo*hi products accumulated in FP32
//     (12 MFMAs per complex 16x16x32 step); |E|^2 is written as FP32.
// Work per plane 8 P M (M + P) real FLOP counted as complex FP32 (C3: 12.85 MFLOP); executed MFMA work is 3x that.
#include "iscg_internal.h"

#include <cmath>
#include <cstring>
#include <string>
#include <type_traits>

namespace iscg {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float float4v __attribute__((ext_vector_type(4)));

constexpr double kPi = 3.14159265358979323846;
constexpr int kKP = 40;  // LDS row of a 32-long k slice, padded (halves)

struct PsfReqDev
{
   double k0, NA, ni, ns, bMax, kMax, dk, kMin, x0, dx, depth, zStepM, ti0, focalShift, maskWaist;
   double coeff[28];
   int n[28], l[28], m[28];
   // sum_t coeff_t R_t(rho) {cos | sin}(m_t phi) folded into polynomials: zc[0][m][j] multiplies rho^j cos(m phi),
   // zc[1][m][j] rho^j sin(m phi) (OSA j < 28: n <= 6)
   double zc[2][7][7];
   int M, size, nz, nModes, doubleHelix, maskModes, pad[2];
};

// One batch of requests on the device: the request records and per-plane defocus (uploaded from the host), the FP64
// pupil tables and the split DFT matrices (computed by kernels from the records).
struct PsfBatch
{
   DevBuf reqs, tabA, tabB, tabP, W, tz;  // tabP: exp(i A) as FP32 pairs, for the row-block kernel
   std::vector<PsfReqDev> hostReqs;   // kept while an upload from them may be in flight
   std::vector<double> hostTz;
   int nReq = 0, size = 0, nz = 0, M = 0, Mp = 0, Pp = 0;
};

struct PsfState
{
   PsfBatch batch;  // iscg_psf_planes's requests
};

void destroy_psf(iscg_context* h)
{
   delete h->psf;
   h->psf = nullptr;
}

// ---- tables -------------------------------------------------------------------------------------------------

// Sim/ZernikePsf.cpp:80-110 (Gauss-Laguerre double-helix mask phase)
__device__ double mask_phase_d(int maskModes, double maskWaist, double rhoNorm, double phi)
{
   const int n = maskModes > 2 ? maskModes : 2;
   const double w = 0.2 < maskWaist ? maskWaist : 0.2;
   const double rc = rhoNorm < 1.0 ? rhoNorm : 1.0;
   const double u = (0.0 < rc ? rc : 0.0) / w;
   const double u2 = u * u;
   double re = 0.0, im = 0.0;
   for (int p = 0; p < n; ++p)
   {
      const int l = 2 * p + 1;
      double lm1 = 0.0, L = 1.0;
      const double x = 2.0 * u2;
      for (int k = 0; k < p; ++k)
      {
         const double next = ((2 * k + 1 + l - x) * L - (k + l) * lm1) / (k + 1);
         lm1 = L;
         L = next;
      }
      const double amp = pow(u, (double)l) * exp(-u2) * L;
      re += amp * cos(l * phi);
      im += amp * sin(l * phi);
   }
   return (re == 0.0 && im == 0.0) ? 0.0 : atan2(im, re);
}

__device__ __forceinline__ void wrapped_sincos(double ang, float* s, float* c);

// A (base phase) and B (phase per metre of ti - ti0) on the Mp x Mp grid; NaN outside the aperture or the M grid.
// With P0 given, also exp(i A) as FP32 (cos, sin): (0, 0) where A is NaN.
__global__ void psf_tables_kernel(const PsfReqDev* __restrict__ reqs, int nReq, int Mp, double* __restrict__ A,
                                  double* __restrict__ B, float2* __restrict__ P0)
{
   const size_t per = (size_t)Mp * Mp, total = per * nReq;
   for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x)
   {
      // per = Mp^2 is a multiple of 64 and a wave's 64 lanes take 64 consecutive idx from a multiple of 64 (all of
      // them, as total is a multiple of 64 too): the request is the same across the wave, so its record (the
      // folded Zernike polynomials: ~100 doubles) is read with scalar loads once per wave, not per lane
      const int r = __builtin_amdgcn_readfirstlane((int)(idx / per));
      const int rem = (int)(idx - (size_t)r * per);
      const int iy = rem / Mp, ix = rem % Mp;
      const PsfReqDev& q = reqs[r];
      double a = __builtin_nan(""), b = 0.0;
      const int c0 = q.M / 2;
      if (iy < q.M && ix < q.M)
      {
         const double ky = (iy - c0) * q.dk, kx = (ix - c0) * q.dk;
         const double kr2 = kx * kx + ky * ky;
         if (!(kr2 > q.kMax * q.kMax))
         {
            const double kr = sqrt(kr2);
            const double rho = kr / (q.k0 * q.NA);
            const double s1 = q.NA * rho / q.ns, s3 = q.NA * rho / q.ni;
            const double a1 = 1.0 - s1 * s1, a3 = 1.0 - s3 * s3;
            const double opd1 = q.ns * q.depth * sqrt(0.0 < a1 ? a1 : 0.0);
            const double opd3Factor = sqrt(0.0 < a3 ? a3 : 0.0);
            // the Zernike sum (ZernikeValue, Sim/ZernikePsf.cpp:58-78; zero for rho / bMax > 1) from the folded
            // polynomials: rho^j by products, cos / sin (m phi) by angle addition from kx / kr, ky / kr
            double zp = 0.0;
            const double rn = rho / q.bMax;
            if (q.nModes > 0 && rn <= 1.0)
            {
               double pw[7];
               pw[0] = 1.0;
               for (int j = 1; j < 7; ++j)
                  pw[j] = pw[j - 1] * rn;
               const double c1 = kr > 0.0 ? kx / kr : 1.0, s1 = kr > 0.0 ? ky / kr : 0.0;
               double cm = 1.0, sm = 0.0;
               for (int m = 0; m < 7; ++m)
               {
                  double rc = 0.0, rs = 0.0;
                  for (int j = m; j < 7; ++j)
                  {
                     rc += q.zc[0][m][j] * pw[j];
                     rs += q.zc[1][m][j] * pw[j];
                  }
                  zp += rc * cm + rs * sm;
                  const double cn = cm * c1 - sm * s1;
                  sm = sm * c1 + cm * s1;
                  cm = cn;
               }
            }
            zp *= 2.0 * kPi;
            const double mask =
               q.doubleHelix ? mask_phase_d(q.maskModes, q.maskWaist, rho / q.bMax, atan2(ky, kx)) : 0.0;
            a = q.k0 * opd1 + zp + mask;
            b = q.k0 * (q.ni * opd3Factor);
         }
      }
      A[idx] = a;
      B[idx] = b;
      if (P0)
      {
         float sn = 0.0f, c = 0.0f;
         if (a == a)
            wrapped_sincos(a, &sn, &c);
         P0[idx] = make_float2(c, sn);
      }
   }
}

__device__ __forceinline__ void split_store(float v, _Float16* hi, _Float16* lo)
{
   const _Float16 h = (_Float16)v;
   *hi = h;
   *lo = (_Float16)(v - (float)h);
}

// Splits v[0..2N) into FP16 hi = v rounded and lo = (v - hi) rounded, two values per conversion.
template <int N, class H>
__device__ __forceinline__ void split_pairs(const float* v, H& hi, H& lo)
{
   typedef float f2 __attribute__((ext_vector_type(2)));
   typedef _Float16 h2 __attribute__((ext_vector_type(2)));
#pragma unroll
   for (int p = 0; p < N; ++p)
   {
      const f2 x = {v[2 * p], v[2 * p + 1]};
      const h2 h = __builtin_convertvector(x, h2);
      const h2 l = __builtin_convertvector(x - __builtin_convertvector(h, f2), h2);
      hi[2 * p] = h.x;
      hi[2 * p + 1] = h.y;
      lo[2 * p] = l.x;
      lo[2 * p + 1] = l.y;
   }
}

// Wraps an FP64 angle to [-pi, pi] in FP64, then FP32 sincos of the remainder.
__device__ __forceinline__ void wrapped_sincos(double ang, float* s, float* c)
{
   const double r = ang - 6.283185307179586476925 * rint(ang * 0.15915494309189533577);
   sincosf((float)r, s, c);
}

// The same values without a branch (so it schedules beside matrix-core work): the angle reduced in FP64 to
// y = ang - q pi/2 (q = rint(ang 2/pi), |y| <= pi/4), the Taylor polynomials of sin (to y^9) and cos (to y^10) in
// FP32 by Horner with FMA (truncation below 2e-9 on |y| <= pi/4, under the FP32 rounding of the result), then the
// quadrant q mod 4 swaps and signs them.
__device__ __forceinline__ void quadrant_sincos(double ang, float* s, float* c)
{
   const double q = rint(ang * 0.63661977236758134308);
   const float y = (float)(ang - q * 1.57079632679489661923), y2 = y * y;
   float ps = fmaf(y2, 1.0f / 362880.0f, -1.0f / 5040.0f);
   ps = fmaf(y2, ps, 1.0f / 120.0f);
   ps = fmaf(y2, ps, -1.0f / 6.0f);
   const float sp = fmaf(y * y2, ps, y);
   float pc = fmaf(y2, -1.0f / 3628800.0f, 1.0f / 40320.0f);
   pc = fmaf(y2, pc, -1.0f / 720.0f);
   pc = fmaf(y2, pc, 1.0f / 24.0f);
   pc = fmaf(y2, pc, -0.5f);
   const float cp = fmaf(y2, pc, 1.0f);
   const int qi = (int)(long long)q & 3;  // two's complement: q = -1 -> 3
   const float sa = (qi & 1) ? cp : sp, ca = (qi & 1) ? sp : cp;
   *s = (qi & 2) ? -sa : sa;
   *c = ((qi + 1) & 2) ? -ca : ca;
}

// W planes per request: [4][Pp][Mp] halves (re hi, re lo, im hi, im lo); zero beyond P rows / M columns.
__global__ void psf_w_kernel(const PsfReqDev* __restrict__ reqs, int nReq, int Pp, int Mp, _Float16* __restrict__ W)
{
   const size_t per = (size_t)Pp * Mp, total = per * nReq;
   for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x)
   {
      // per = Pp Mp is a multiple of 512: one request per wave (scalar loads of its record)
      const int r = __builtin_amdgcn_readfirstlane((int)(idx / per));
      const int rem = (int)(idx - (size_t)r * per);
      const int p = rem / Mp, m = rem % Mp;
      const PsfReqDev& q = reqs[r];
      float s = 0.0f, c = 0.0f;
      if (p < q.size && m < q.M)
         wrapped_sincos((q.kMin + m * q.dk) * (q.x0 + p * q.dx), &s, &c);
      _Float16* base = W + (size_t)r * 4 * per + rem;
      split_store(c, base, base + per);
      split_store(s, base + 2 * per, base + 3 * per);
   }
}

// ---- planes ---------------------------------------------------------------------------------------------------

__device__ __forceinline__ half8 ld8(const _Float16* p) { return *(const half8*)p; }

__device__ __forceinline__ float4v mfma(half8 a, half8 b, float4v c)
{
   return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// acc += (ahi + alo)(bhi + blo) without the lo*lo term
__device__ __forceinline__ float4v mfma3(half8 ahi, half8 alo, half8 bhi, half8 blo, float4v acc)
{
   acc = mfma(ahi, bhi, acc);
   acc = mfma(ahi, blo, acc);
   return mfma(alo, bhi, acc);
}

// One workgroup (kPsfWaves waves) per (request, plane, block of YB*16 output rows). MP: padded pupil samples (multiple of
// 32), PP: padded output size (multiple of 16).
constexpr int kPsfWaves = 8, kPsfThreads = 64 * kPsfWaves;

template <int MP, int YB>
__global__ __launch_bounds__(kPsfThreads) void psf_planes_kernel(const PsfReqDev* __restrict__ reqs, const double* __restrict__ A,
                                                         const double* __restrict__ B, const _Float16* __restrict__ W,
                                                         const double* __restrict__ tz, int nz, int Pp, int nYBlocks,
                                                         float* __restrict__ out)
{
   constexpr int CT1 = MP / 16;              // column tiles of mid
   constexpr int T1 = YB * CT1;              // tiles of GEMM 1
   constexpr int PER1 = (T1 + kPsfWaves - 1) / kPsfWaves;  // per wave
   constexpr int MS = MP + 8;                // LDS row of mid (halves)
   __shared__ __attribute__((aligned(16))) _Float16 pup[4][MP][kKP];      // pupil chunk, [part][m][k]
   __shared__ __attribute__((aligned(16))) _Float16 mid[4][YB * 16][MS];  // mid rows, [part][y][m]

   const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
   const int item = blockIdx.x;
   const int yb = item % nYBlocks;
   const int zi = (item / nYBlocks) % nz;
   const int r = item / (nYBlocks * nz);
   const PsfReqDev& q = reqs[r];
   const int P = q.size;
   const int y0 = yb * YB * 16;
   const size_t wPer = (size_t)Pp * MP;
   const _Float16* Wr = W + (size_t)r * 4 * wPer;
   const double* Ar = A + (size_t)r * MP * MP;
   const double* Br = B + (size_t)r * MP * MP;
   const double t = tz[(size_t)r * nz + zi];
   const int lr = lane & 15, lk = 8 * (lane >> 4);

   // ---- GEMM 1: mid[y][m] = sum_n W[y0 + y][n] Pup[n][m] ----
   float4v accR[PER1], accI[PER1];
#pragma unroll
   for (int s = 0; s < PER1; ++s)
      accR[s] = accI[s] = float4v{0.0f, 0.0f, 0.0f, 0.0f};
   for (int n0 = 0; n0 < MP; n0 += 32)
   {
      // the pupil rows [n0, n0 + 32): FP64 phase, wrapped, FP32 sincos, split
      for (int i = tid; i < 32 * MP; i += kPsfThreads)
      {
         const int k = i / MP, m = i % MP;
         const size_t o = (size_t)(n0 + k) * MP + m;
         const double a = Ar[o];
         float s = 0.0f, c = 0.0f;
         if (a == a)  // inside the aperture
            wrapped_sincos(a + t * Br[o], &s, &c);
         split_store(c, &pup[0][m][k], &pup[1][m][k]);
         split_store(s, &pup[2][m][k], &pup[3][m][k]);
      }
      __syncthreads();
#pragma unroll
      for (int s = 0; s < PER1; ++s)
      {
         const int ti = wave + kPsfWaves * s;
         // rows at or beyond Pp (the last block's padding) stay zero: W holds Pp rows only
         if (ti < T1 && y0 + (ti / CT1) * 16 < Pp)
         {
            const int rb = ti / CT1, cb = ti % CT1;
            const size_t wo = (size_t)(y0 + rb * 16 + lr) * MP + n0 + lk;
            const half8 arh = ld8(Wr + wo), arl = ld8(Wr + wPer + wo);
            const half8 aih = ld8(Wr + 2 * wPer + wo), ail = ld8(Wr + 3 * wPer + wo);
            const int mc = cb * 16 + lr;
            const half8 brh = ld8(&pup[0][mc][lk]), brl = ld8(&pup[1][mc][lk]);
            const half8 bih = ld8(&pup[2][mc][lk]), bil = ld8(&pup[3][mc][lk]);
            accR[s] = mfma3(arh, arl, brh, brl, accR[s]);
            accR[s] = mfma3(-aih, -ail, bih, bil, accR[s]);
            accI[s] = mfma3(arh, arl, bih, bil, accI[s]);
            accI[s] = mfma3(aih, ail, brh, brl, accI[s]);
         }
      }
      __syncthreads();
   }
   // mid -> LDS as split halves
#pragma unroll
   for (int s = 0; s < PER1; ++s)
   {
      const int ti = wave + kPsfWaves * s;
      if (ti < T1)
      {
         const int rb = ti / CT1, cb = ti % CT1;
#pragma unroll
         for (int i = 0; i < 4; ++i)
         {
            const int y = rb * 16 + 4 * (lane >> 4) + i, m = cb * 16 + lr;
            split_store(accR[s][i], &mid[0][y][m], &mid[1][y][m]);
            split_store(accI[s][i], &mid[2][y][m], &mid[3][y][m]);
         }
      }
   }
   __syncthreads();

   // ---- GEMM 2: E[y][x] = sum_m mid[y][m] W[x][m]; |E|^2 out ----
   const int CT2 = Pp / 16, T2 = YB * CT2;
   float* plane = out + ((size_t)r * nz + zi) * P * P;
   for (int ti = wave; ti < T2; ti += kPsfWaves)
   {
      const int rb = ti / CT2, cb = ti % CT2;
      if (y0 + rb * 16 >= P)
         continue;
      float4v eR = {0.0f, 0.0f, 0.0f, 0.0f}, eI = {0.0f, 0.0f, 0.0f, 0.0f};
      const int yr = rb * 16 + lr;
      const size_t wx = (size_t)(cb * 16 + lr) * MP + lk;
#pragma unroll 2
      for (int m0 = 0; m0 < MP; m0 += 32)
      {
         const half8 arh = ld8(&mid[0][yr][m0 + lk]), arl = ld8(&mid[1][yr][m0 + lk]);
         const half8 aih = ld8(&mid[2][yr][m0 + lk]), ail = ld8(&mid[3][yr][m0 + lk]);
         const half8 brh = ld8(Wr + wx + m0), brl = ld8(Wr + wPer + wx + m0);
         const half8 bih = ld8(Wr + 2 * wPer + wx + m0), bil = ld8(Wr + 3 * wPer + wx + m0);
         eR = mfma3(arh, arl, brh, brl, eR);
         eR = mfma3(-aih, -ail, bih, bil, eR);
         eI = mfma3(arh, arl, bih, bil, eI);
         eI = mfma3(aih, ail, brh, brl, eI);
      }
      const int x = cb * 16 + lr;
#pragma unroll
      for (int i = 0; i < 4; ++i)
      {
         const int y = y0 + rb * 16 + 4 * (lane >> 4) + i;
         if (y < P && x < P)
            plane[(size_t)y * P + x] = eR[i] * eR[i] + eI[i] * eI[i];
      }
   }
}

// Pupils above 256 samples per axis (the reference takes up to 1024, Sim/ZernikePsf.cpp:134-148): the pupil's
// columns in chunks of kBigMC, so LDS holds one chunk of the pupil (kBigMC x 32 per step, split) and of mid
// (YB 16 x kBigMC); E = sum over chunks of mid_chunk W_chunk^T stays in registers across the chunks (PER2 tiles of
// 16 x 16 per wave). One workgroup per (request, plane, block of YB 16 output rows); the pupil chunk is evaluated
// again for every row block, as the smaller kernel does. Same split products, so the same accuracy.
constexpr int kBigMC = 128;

template <int YB, int PER2>
__global__ __launch_bounds__(kPsfThreads) void psf_planes_big_kernel(const PsfReqDev* __restrict__ reqs,
                                                                     const double* __restrict__ A,
                                                                     const double* __restrict__ B,
                                                                     const _Float16* __restrict__ W,
                                                                     const double* __restrict__ tz, int nz, int Mp,
                                                                     int Pp, int nYBlocks, float* __restrict__ out)
{
   constexpr int MC = kBigMC, CT1 = MC / 16, T1 = YB * CT1;
   constexpr int PER1 = (T1 + kPsfWaves - 1) / kPsfWaves;
   constexpr int MS = MC + 8;
   __shared__ __attribute__((aligned(16))) _Float16 pup[4][MC][kKP];      // pupil chunk, [part][m][k]
   __shared__ __attribute__((aligned(16))) _Float16 mid[4][YB * 16][MS];  // mid chunk, [part][y][m]

   const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
   const int item = blockIdx.x;
   const int yb = item % nYBlocks;
   const int zi = (item / nYBlocks) % nz;
   const int r = item / (nYBlocks * nz);
   const PsfReqDev& q = reqs[r];
   const int P = q.size;
   const int y0 = yb * YB * 16;
   const size_t wPer = (size_t)Pp * Mp;
   const _Float16* Wr = W + (size_t)r * 4 * wPer;
   const double* Ar = A + (size_t)r * Mp * Mp;
   const double* Br = B + (size_t)r * Mp * Mp;
   const double t = tz[(size_t)r * nz + zi];
   const int lr = lane & 15, lk = 8 * (lane >> 4);
   const int CT2 = Pp / 16, T2 = YB * CT2;  // T2 <= 8 PER2 (the host's choice of PER2)

   float4v eR[PER2], eI[PER2];
#pragma unroll
   for (int s = 0; s < PER2; ++s)
      eR[s] = eI[s] = float4v{0.0f, 0.0f, 0.0f, 0.0f};
   for (int mc0 = 0; mc0 < Mp; mc0 += MC)
   {
      const int mcn = Mp - mc0 < MC ? Mp - mc0 : MC;  // columns of this chunk (a multiple of 32)
      // ---- GEMM 1 on the chunk: mid[y][m] = sum_n W[y0 + y][n] Pup[n][mc0 + m] ----
      float4v accR[PER1], accI[PER1];
#pragma unroll
      for (int s = 0; s < PER1; ++s)
         accR[s] = accI[s] = float4v{0.0f, 0.0f, 0.0f, 0.0f};
      for (int n0 = 0; n0 < Mp; n0 += 32)
      {
         for (int i = tid; i < 32 * MC; i += kPsfThreads)
         {
            const int k = i / MC, m = i % MC;
            float s = 0.0f, c = 0.0f;
            if (m < mcn)
            {
               const size_t o = (size_t)(n0 + k) * Mp + mc0 + m;
               const double a = Ar[o];
               if (a == a)  // inside the aperture
                  wrapped_sincos(a + t * Br[o], &s, &c);
            }
            split_store(c, &pup[0][m][k], &pup[1][m][k]);
            split_store(s, &pup[2][m][k], &pup[3][m][k]);
         }
         __syncthreads();
#pragma unroll
         for (int s = 0; s < PER1; ++s)
         {
            const int ti = wave + kPsfWaves * s;
            const int rb = ti / CT1, cb = ti % CT1;
            if (ti < T1 && y0 + rb * 16 < Pp && cb * 16 < mcn)
            {
               const size_t wo = (size_t)(y0 + rb * 16 + lr) * Mp + n0 + lk;
               const half8 arh = ld8(Wr + wo), arl = ld8(Wr + wPer + wo);
               const half8 aih = ld8(Wr + 2 * wPer + wo), ail = ld8(Wr + 3 * wPer + wo);
               const int mc = cb * 16 + lr;
               const half8 brh = ld8(&pup[0][mc][lk]), brl = ld8(&pup[1][mc][lk]);
               const half8 bih = ld8(&pup[2][mc][lk]), bil = ld8(&pup[3][mc][lk]);
               accR[s] = mfma3(arh, arl, brh, brl, accR[s]);
               accR[s] = mfma3(-aih, -ail, bih, bil, accR[s]);
               accI[s] = mfma3(arh, arl, bih, bil, accI[s]);
               accI[s] = mfma3(aih, ail, brh, brl, accI[s]);
            }
         }
         __syncthreads();
      }
#pragma unroll
      for (int s = 0; s < PER1; ++s)
      {
         const int ti = wave + kPsfWaves * s;
         if (ti < T1)
         {
            const int rb = ti / CT1, cb = ti % CT1;
#pragma unroll
            for (int i = 0; i < 4; ++i)
            {
               const int y = rb * 16 + 4 * (lane >> 4) + i, m = cb * 16 + lr;
               split_store(accR[s][i], &mid[0][y][m], &mid[1][y][m]);
               split_store(accI[s][i], &mid[2][y][m], &mid[3][y][m]);
            }
         }
      }
      __syncthreads();
      // ---- GEMM 2 on the chunk: E[y][x] += sum_m mid[y][m] W[x][mc0 + m] ----
#pragma unroll
      for (int s = 0; s < PER2; ++s)
      {
         const int ti = wave + kPsfWaves * s;
         if (ti < T2)
         {
            const int rb = ti / CT2, cb = ti % CT2;
            const int yr = rb * 16 + lr;
            const size_t wx = (size_t)(cb * 16 + lr) * Mp + mc0 + lk;
            for (int m0 = 0; m0 < mcn; m0 += 32)
            {
               const half8 arh = ld8(&mid[0][yr][m0 + lk]), arl = ld8(&mid[1][yr][m0 + lk]);
               const half8 aih = ld8(&mid[2][yr][m0 + lk]), ail = ld8(&mid[3][yr][m0 + lk]);
               const half8 brh = ld8(Wr + wx + m0), brl = ld8(Wr + wPer + wx + m0);
               const half8 bih = ld8(Wr + 2 * wPer + wx + m0), bil = ld8(Wr + 3 * wPer + wx + m0);
               eR[s] = mfma3(arh, arl, brh, brl, eR[s]);
               eR[s] = mfma3(-aih, -ail, bih, bil, eR[s]);
               eI[s] = mfma3(arh, arl, bih, bil, eI[s]);
               eI[s] = mfma3(aih, ail, brh, brl, eI[s]);
            }
         }
      }
      __syncthreads();  // the next chunk's pupil and mid overwrite the LDS the products read
   }
   float* plane = out + ((size_t)r * nz + zi) * P * P;
#pragma unroll
   for (int s = 0; s < PER2; ++s)
   {
      const int ti = wave + kPsfWaves * s;
      if (ti < T2)
      {
         const int rb = ti / CT2, cb = ti % CT2;
         const int x = cb * 16 + lr;
#pragma unroll
         for (int i = 0; i < 4; ++i)
         {
            const int y = y0 + rb * 16 + 4 * (lane >> 4) + i;
            if (y < P && x < P)
               plane[(size_t)y * P + x] = eR[s][i] * eR[s][i] + eI[s][i] * eI[s][i];
         }
      }
   }
}

// Small planes (padded size Pp <= 80), E = W (Pup W^T) with E in registers. Per 32-row chunk n0 of the pupil: the
// split pupil rows Pc (32 x MP) in LDS; mid[n][x] = sum_m Pc[n][m] W[x][m] (32 x Pp), stored split and transposed
// (midT[x][n]); E[y][x] += sum_n W[y][n0 + n] mid[n][x].
// WLDS (MP <= 128): zg workgroups per request (zg tuned among 1-32, 4 by default), each walking a contiguous zg-th of its planes
// with the split W
// (4 x 80 x (MP + 8) halves) staged in LDS once: the products read no global memory. The workgroups of a request
// are dispatched to one XCD (workgroup b runs on XCD b mod 8), so the request's FP64 pupil tables, read again for
// every plane, stay in that XCD's L2 (8 requests x 256 KB at a time at M 128). Otherwise one workgroup per
// (request, plane), W from global memory, LDS 4 x 32 x (MP + 8) + 4 x 80 x 40 halves (two workgroups per CU).
constexpr int kSmallPp = 80, kZGroups = 4, kXcds = 8;  // kZGroups: the default zg
template <int MP, bool WLDS>
__global__ __launch_bounds__(kPsfThreads, WLDS ? 2 : 4) void psf_small_kernel(const PsfReqDev* __restrict__ reqs,
                                                                             const double* __restrict__ A,
                                                                             const double* __restrict__ B,
                                                                             const _Float16* __restrict__ W,
                                                                             const double* __restrict__ tz, int nz,
                                                                             int Pp, int nReq, int zg,
                                                                             float* __restrict__ out)
{
   constexpr int PS = MP + 8;                                    // pupil / W row (halves)
   constexpr int CTX = kSmallPp / 16;                            // max tiles per side
   constexpr int PE = (CTX * CTX + kPsfWaves - 1) / kPsfWaves;   // E tiles per wave (max)
   constexpr int PM = (2 * CTX + kPsfWaves - 1) / kPsfWaves;     // mid tiles per wave (max)
   __shared__ __attribute__((aligned(16))) _Float16 pc[4][32][PS];
   __shared__ __attribute__((aligned(16))) _Float16 midT[4][kSmallPp][kKP];
   __shared__ __attribute__((aligned(16))) _Float16 wl[WLDS ? 4 : 1][WLDS ? kSmallPp : 1][WLDS ? PS : 8];
   const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
   const int b = (int)blockIdx.x, slot = b / kXcds;
   const int r = WLDS ? (slot / zg) * kXcds + b % kXcds : b / nz;
   if (r >= nReq)
      return;  // the last XCD row's spare workgroups
   const PsfReqDev& q = reqs[r];
   const int P = q.size;
   const int CT = Pp / 16, TE = CT * CT, TM = 2 * CT;
   const size_t wPer = (size_t)Pp * MP;
   const _Float16* Wr = W + (size_t)r * 4 * wPer;
   const double* Ar = A + (size_t)r * MP * MP;
   const double* Br = B + (size_t)r * MP * MP;
   const int lr = lane & 15, lk = 8 * (lane >> 4);
   // operand (part, row, column) of W: row x < Pp, 8 halves from column c
   auto wld = [&](int part, int row, int col) -> half8 {
      if constexpr (WLDS)
         return ld8(&wl[part][row][col]);
      else
         return ld8(Wr + part * wPer + (size_t)row * MP + col);
   };
   if constexpr (WLDS)
   {
      for (int i = tid; i < 4 * Pp * (MP / 8); i += kPsfThreads)
      {
         const int part = i / (Pp * (MP / 8)), rem = i % (Pp * (MP / 8)), row = rem / (MP / 8), c8 = rem % (MP / 8);
         *(half8*)&wl[part][row][8 * c8] = ld8(Wr + part * wPer + (size_t)row * MP + 8 * c8);
      }
   }
   const int g = slot % zg;
   const int z0 = WLDS ? g * nz / zg : b % nz, z1 = WLDS ? (g + 1) * nz / zg : z0 + 1;
   for (int zi = z0; zi < z1; ++zi)
   {
      const double t = tz[(size_t)r * nz + zi];
      float4v eR[PE], eI[PE];
#pragma unroll
      for (int s = 0; s < PE; ++s)
         eR[s] = eI[s] = float4v{0.0f, 0.0f, 0.0f, 0.0f};
      for (int n0 = 0; n0 < MP; n0 += 32)
      {
         // the pupil rows [n0, n0 + 32): FP64 phase, wrapped, FP32 sincos, split; 8 consecutive samples per lane.
         // (The barrier below also orders the previous chunk's reads of pc and midT before these writes.)
         for (int i = tid; i < 32 * (MP / 8); i += kPsfThreads)
         {
            const int k = i / (MP / 8), m0 = (i % (MP / 8)) * 8;
            half8 chi, clo, shi, slo;
#pragma unroll
            for (int j = 0; j < 8; ++j)
            {
               const size_t o = (size_t)(n0 + k) * MP + m0 + j;
               const double a = Ar[o];
               float s = 0.0f, c = 0.0f;
               if (a == a)  // inside the aperture
                  wrapped_sincos(a + t * Br[o], &s, &c);
               const _Float16 ch = (_Float16)c, sh = (_Float16)s;
               chi[j] = ch;
               clo[j] = (_Float16)(c - (float)ch);
               shi[j] = sh;
               slo[j] = (_Float16)(s - (float)sh);
            }
            *(half8*)&pc[0][k][m0] = chi;
            *(half8*)&pc[1][k][m0] = clo;
            *(half8*)&pc[2][k][m0] = shi;
            *(half8*)&pc[3][k][m0] = slo;
         }
         __syncthreads();
         // mid tiles: rows 16 (tm / CT) .., columns 16 (tm % CT) ..
#pragma unroll
         for (int s = 0; s < PM; ++s)
         {
            const int tm = wave + kPsfWaves * s;
            if (tm >= TM)
               continue;
            const int rb = tm / CT, cb = tm % CT;
            float4v mR = {0.0f, 0.0f, 0.0f, 0.0f}, mI = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 1
            for (int k0 = 0; k0 < MP; k0 += 32)
            {
               const int row = 16 * rb + lr, wrow = 16 * cb + lr;
               const half8 arh = ld8(&pc[0][row][k0 + lk]), arl = ld8(&pc[1][row][k0 + lk]);
               const half8 aih = ld8(&pc[2][row][k0 + lk]), ail = ld8(&pc[3][row][k0 + lk]);
               const half8 brh = wld(0, wrow, k0 + lk), brl = wld(1, wrow, k0 + lk);
               const half8 bih = wld(2, wrow, k0 + lk), bil = wld(3, wrow, k0 + lk);
               mR = mfma3(arh, arl, brh, brl, mR);
               mR = mfma3(-aih, -ail, bih, bil, mR);
               mI = mfma3(arh, arl, bih, bil, mI);
               mI = mfma3(aih, ail, brh, brl, mI);
            }
            // D rows n = 16 rb + 4 (lane >> 4) + i (four consecutive), column x = 16 cb + lr -> midT[.][x][n..n+3]
            const int x = 16 * cb + lr, n = 16 * rb + 4 * (lane >> 4);
            typedef _Float16 half4 __attribute__((ext_vector_type(4)));
            half4 rh, rl, ih, il;
#pragma unroll
            for (int i = 0; i < 4; ++i)
            {
               const _Float16 a = (_Float16)mR[i], b = (_Float16)mI[i];
               rh[i] = a;
               rl[i] = (_Float16)(mR[i] - (float)a);
               ih[i] = b;
               il[i] = (_Float16)(mI[i] - (float)b);
            }
            *(half4*)&midT[0][x][n] = rh;
            *(half4*)&midT[1][x][n] = rl;
            *(half4*)&midT[2][x][n] = ih;
            *(half4*)&midT[3][x][n] = il;
         }
         __syncthreads();
         // E tiles += W[y][n0 + n] mid[n][x] (K = 32)
#pragma unroll
         for (int s = 0; s < PE; ++s)
         {
            const int te = wave + kPsfWaves * s;
            if (te >= TE)
               continue;
            const int ry = te / CT, cx = te % CT, wrow = 16 * ry + lr;
            const half8 arh = wld(0, wrow, n0 + lk), arl = wld(1, wrow, n0 + lk);
            const half8 aih = wld(2, wrow, n0 + lk), ail = wld(3, wrow, n0 + lk);
            const int x = 16 * cx + lr;
            const half8 brh = ld8(&midT[0][x][lk]), brl = ld8(&midT[1][x][lk]);
            const half8 bih = ld8(&midT[2][x][lk]), bil = ld8(&midT[3][x][lk]);
            eR[s] = mfma3(arh, arl, brh, brl, eR[s]);
            eR[s] = mfma3(-aih, -ail, bih, bil, eR[s]);
            eI[s] = mfma3(arh, arl, bih, bil, eI[s]);
            eI[s] = mfma3(aih, ail, brh, brl, eI[s]);
         }
      }
      float* plane = out + ((size_t)r * nz + zi) * P * P;
#pragma unroll
      for (int s = 0; s < PE; ++s)
      {
         const int te = wave + kPsfWaves * s;
         if (te >= TE)
            continue;
         const int ry = te / CT, cx = te % CT, x = 16 * cx + lr;
#pragma unroll
         for (int i = 0; i < 4; ++i)
         {
            const int y = 16 * ry + 4 * (lane >> 4) + i;
            if (y < P && x < P)
               plane[(size_t)y * P + x] = eR[s][i] * eR[s][i] + eI[s][i] * eI[s][i];
         }
      }
   }
}

// Small planes, row-block form (planes of size P <= 66, MP in {32, 64, 128}): the products of psf_small_kernel
// with the pupil factored and the work laid out per plane as
//   0. the defocus factor exp(i t B) depends on the pupil radius only: B is a function of kx^2 + ky^2, the same bits
//      for (i, j), (j, i) and either sign of the offsets i, j from the grid centre c0. Per plane the workgroup
//      evaluates it once per pair hi = max(|i|, |j|), lo = min(|i|, |j|) <= c0 (FP64 t B, wrapped, FP32 sincos;
//      (c0 + 1)(c0 + 2) / 2 entries, a 16th of the M^2 samples at M 128) into LDS;
//   1. each wave owns one 16-row block rb of the pupil (MP / 16 blocks; with fewer blocks than waves, G = 8 / (MP /
//      16) waves share a block and split its column tiles): pupil = exp(i A) (an FP32 table made once per request,
//      zero outside the aperture) times its radius's defocus factor, split straight into its MFMA A fragments
//      (registers, no LDS); then mid[n][x] = sum_m Pup[n][m] W[x][m] for its column tiles x over all MP (K) -- the A
//      fragments are reused across the tiles, each W fragment read once;
//   2. mid, split and transposed, in LDS (midT[x][n], rows x < P only); E[y][x] = sum_n W[y][n] mid[n][x] for the
//      wave's E tiles over all MP (K).
// Two barriers per plane: after step 1 (midT complete, every wave done reading the plane's factors) and after step 2
// (the next plane's factors, made during step 2 beside the products, are in LDS, and every wave has left step 2, so
// midT may be rewritten). The first plane's factors are made before the loop. LDS holds W and midT for the P valid rows only (the padded rows >= P of
// a tile read row P - 1 and their results are not stored): 2 x 4 x 66 x 136 halves + 2145 factors at MP 128
// (157 KB), one workgroup per CU. The products keep the real part's two terms in two accumulators (re re - im im at
// the end), so no fragment is negated, and each lane's defocus B values are read once per workgroup. Workgroups per request
// (zg) and their XCD placement as in psf_small_kernel. The pupil samples differ from the FP64-phase kernels' by
// FP32 rounding (one complex product more), so this form is not a tuning candidate beside them.
constexpr int kRbRows = 66;
template <int MP, int CTF>  // CTF: the tiles per side when known at compile time (5: planes padded to 80), else 0
__global__ __launch_bounds__(kPsfThreads, 1) void psf_small_rb_kernel(const PsfReqDev* __restrict__ reqs,
                                                                     const float2* __restrict__ P0,
                                                                     const double* __restrict__ B,
                                                                     const _Float16* __restrict__ W,
                                                                     const double* __restrict__ tz, int nz, int Pp,
                                                                     int nReq, int zg, float* __restrict__ out)
{
   constexpr int PS = MP + 8;                 // LDS row (halves)
   constexpr int NRB = MP / 16;               // pupil row blocks
   constexpr int G = kPsfWaves / NRB;         // waves per row block
   constexpr int KS = MP / 32;                // k-steps of 32
   constexpr int CTX = kSmallPp / 16;         // max column tiles
   constexpr int PM = (CTX + G - 1) / G;      // mid tiles per wave (max)
   constexpr int PE = (CTX * CTX + kPsfWaves - 1) / kPsfWaves;  // E tiles per wave (max)
   constexpr int HM = MP / 2;                                 // c0 = M / 2 <= MP / 2
   constexpr int ND = (HM + 1) * (HM + 2) / 2;                // radial pairs (hi, lo), lo <= hi <= HM
   static_assert(NRB * G == kPsfWaves, "row blocks must divide the waves");
   __shared__ __attribute__((aligned(16))) _Float16 wl[4][kRbRows][PS];
   __shared__ __attribute__((aligned(16))) _Float16 midT[4][kRbRows][PS];
   __shared__ float2 defocus[ND + 1];  // + a spare slot the lanes past the last entry write
   const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
   const int b = (int)blockIdx.x, slot = b / kXcds;
   const int r = (slot / zg) * kXcds + b % kXcds;
   if (r >= nReq)
      return;  // the last XCD row's spare workgroups
   const PsfReqDev& q = reqs[r];
   const int P = q.size;
   const int CT = CTF ? CTF : Pp / 16, TE = CT * CT;
   const size_t wPer = (size_t)Pp * MP;
   const _Float16* Wr = W + (size_t)r * 4 * wPer;
   const float2* P0r = P0 + (size_t)r * MP * MP;
   const double* Br = B + (size_t)r * MP * MP;
   const int c0 = q.M / 2, nd = (c0 + 1) * (c0 + 2) / 2;  // M <= MP: c0 <= HM, nd <= ND
   const int lr = lane & 15, lk = 8 * (lane >> 4);
   for (int i = tid; i < 4 * P * (MP / 8); i += kPsfThreads)
   {
      const int part = i / (P * (MP / 8)), rem = i % (P * (MP / 8)), row = rem / (MP / 8), c8 = rem % (MP / 8);
      *(half8*)&wl[part][row][8 * c8] = ld8(Wr + part * wPer + (size_t)row * MP + 8 * c8);
   }
   const int rb = wave % NRB, g0 = wave / NRB;
   const int g = slot % zg;
   const int z0 = g * nz / zg, z1 = (g + 1) * nz / zg;
   // this lane's pupil row n = 16 rb + lr: its offset from the centre (rows past the M grid are zero in P0 and take
   // any in-range factor)
   const int ai = min(abs(16 * rb + lr - c0), c0);
   // the defocus entries e = tid + kPsfThreads u of this lane: hi (hi + 1) / 2 + lo = e, their B read once here
   constexpr int NDT = (ND + kPsfThreads - 1) / kPsfThreads;
   double bdef[NDT];
#pragma unroll
   for (int u = 0; u < NDT; ++u)
   {
      const int e = tid + kPsfThreads * u;
      bdef[u] = 0.0;
      if (e < nd)
      {
         int hi = (int)((sqrtf(8.0f * (float)e + 1.0f) - 1.0f) * 0.5f);
         while (hi * (hi + 1) / 2 > e)
            --hi;
         while ((hi + 1) * (hi + 2) / 2 <= e)
            ++hi;
         const int lo = e - hi * (hi + 1) / 2;
         bdef[u] = Br[(size_t)(c0 - hi) * MP + (c0 - lo)];
      }
   }
   // entry u of this lane's defocus factors for the plane at t (the lanes past the last entry write the spare slot)
   auto defocus_entry = [&](int u, double t) {
      const int e = tid + kPsfThreads * u;
      float sn, c;
      quadrant_sincos(t * bdef[u], &sn, &c);
      defocus[e < nd ? e : ND] = make_float2(c, sn);
   };
   // 0. the first plane's defocus factors (each later plane's are made during the previous plane's step 2)
   if (z0 < z1)
   {
#pragma unroll
      for (int u = 0; u < NDT; ++u)
         defocus_entry(u, tz[(size_t)r * nz + z0]);
   }
   __syncthreads();  // W staged, the first factors in LDS
   for (int zi = z0; zi < z1; ++zi)
   {
      // 1. this wave's pupil rows n = 16 rb + lr, columns m = 32 ks + lk + j, as split A fragments, one k-step at a
      // time; mid[n][x] = sum_m Pup[n][m] W[x][m] over the wave's column tiles, k-steps outermost so the next step's
      // pupil (VALU, LDS reads) can be scheduled beside this step's products. The real part is (re re) - (im im)
      // from two accumulators, so no fragment is negated; columns x >= P read row P - 1 (their columns are not
      // stored).
      auto pupil = [&](int ks, half8& rh, half8& rl, half8& ih, half8& il) {
         const float4* src = reinterpret_cast<const float4*>(P0r + (size_t)(16 * rb + lr) * MP + 32 * ks + lk);
         float4 v[4];
#pragma unroll
         for (int u = 0; u < 4; ++u)
            v[u] = src[u];
         float2 d[8];
#pragma unroll
         for (int j = 0; j < 8; ++j)
         {
            const int aj = min(abs(32 * ks + lk + j - c0), c0);
            const int hi = max(ai, aj), lo = min(ai, aj);
            d[j] = defocus[hi * (hi + 1) / 2 + lo];
         }
         float cr[8], ci[8];
#pragma unroll
         for (int j = 0; j < 8; ++j)
         {
            const float pc = (j & 1) ? v[j >> 1].z : v[j >> 1].x, ps = (j & 1) ? v[j >> 1].w : v[j >> 1].y;
            cr[j] = fmaf(pc, d[j].x, -(ps * d[j].y));
            ci[j] = fmaf(pc, d[j].y, ps * d[j].x);
         }
         split_pairs<4>(cr, rh, rl);
         split_pairs<4>(ci, ih, il);
      };
      float4v mRp[PM], mRn[PM], mI[PM];
#pragma unroll
      for (int s = 0; s < PM; ++s)
         mRp[s] = mRn[s] = mI[s] = float4v{0.0f, 0.0f, 0.0f, 0.0f};
      half8 prh, prl, pih, pil;
      pupil(0, prh, prl, pih, pil);
#pragma unroll
      for (int ks = 0; ks < KS; ++ks)
      {
         half8 nrh, nrl, nih, nil;
         if (ks + 1 < KS)
            pupil(ks + 1, nrh, nrl, nih, nil);
#pragma unroll
         for (int s = 0; s < PM; ++s)
         {
            const int ct = g0 + G * s;
            if (ct < CT)
            {
               const int xr = min(16 * ct + lr, P - 1);
               const half8 brh = ld8(&wl[0][xr][32 * ks + lk]), brl = ld8(&wl[1][xr][32 * ks + lk]);
               const half8 bih = ld8(&wl[2][xr][32 * ks + lk]), bil = ld8(&wl[3][xr][32 * ks + lk]);
               mRp[s] = mfma3(prh, prl, brh, brl, mRp[s]);
               mRn[s] = mfma3(pih, pil, bih, bil, mRn[s]);
               mI[s] = mfma3(prh, prl, bih, bil, mI[s]);
               mI[s] = mfma3(pih, pil, brh, brl, mI[s]);
            }
         }
         if (ks + 1 < KS)
         {
            prh = nrh;
            prl = nrl;
            pih = nih;
            pil = nil;
         }
      }
      // D rows n = 16 rb + 4 (lane >> 4) + i, column x -> midT[.][x][n .. n + 3]
#pragma unroll
      for (int s = 0; s < PM; ++s)
      {
         const int x = 16 * (g0 + G * s) + lr;
         if (g0 + G * s < CT && x < P)
         {
            const int n = 16 * rb + 4 * (lane >> 4);
            typedef _Float16 half4 __attribute__((ext_vector_type(4)));
            const float4v mR = mRp[s] - mRn[s];
            const float vr[4] = {mR[0], mR[1], mR[2], mR[3]}, vi[4] = {mI[s][0], mI[s][1], mI[s][2], mI[s][3]};
            half4 rh, rl, ih, il;
            split_pairs<2>(vr, rh, rl);
            split_pairs<2>(vi, ih, il);
            *(half4*)&midT[0][x][n] = rh;
            *(half4*)&midT[1][x][n] = rl;
            *(half4*)&midT[2][x][n] = ih;
            *(half4*)&midT[3][x][n] = il;
         }
      }
      __syncthreads();  // midT complete; every wave's reads of this plane's defocus factors done
      // 2. E tiles: E[y][x] = sum_n W[y][n] mid[n][x]; rows and columns >= P read row P - 1 and are not stored. The
      // next plane's defocus factors (VALU) go in beside each tile's products (matrix cores) in one basic block
      float* plane = out + ((size_t)r * nz + zi) * P * P;
      const double tn = zi + 1 < z1 ? tz[(size_t)r * nz + zi + 1] : 0.0;  // past the last plane: unused factors
#pragma unroll
      for (int s = 0; s < PE; ++s)
      {
         const int te = wave + kPsfWaves * s;
         if (te >= TE)
         {
            if (s < NDT)
               defocus_entry(s, tn);
            continue;
         }
         const int ry = te / CT, cx = te % CT;
         const int yr = min(16 * ry + lr, P - 1), xr = min(16 * cx + lr, P - 1);
         float4v eRp = {0.0f, 0.0f, 0.0f, 0.0f}, eRn = eRp, eI = eRp;
#pragma unroll
         for (int ks = 0; ks < KS; ++ks)
         {
            const int k = 32 * ks + lk;
            const half8 arh = ld8(&wl[0][yr][k]), arl = ld8(&wl[1][yr][k]);
            const half8 aih = ld8(&wl[2][yr][k]), ail = ld8(&wl[3][yr][k]);
            const half8 brh = ld8(&midT[0][xr][k]), brl = ld8(&midT[1][xr][k]);
            const half8 bih = ld8(&midT[2][xr][k]), bil = ld8(&midT[3][xr][k]);
            eRp = mfma3(arh, arl, brh, brl, eRp);
            eRn = mfma3(aih, ail, bih, bil, eRn);
            eI = mfma3(arh, arl, bih, bil, eI);
            eI = mfma3(aih, ail, brh, brl, eI);
         }
         if (s < NDT)
            defocus_entry(s, tn);
         const float4v eR = eRp - eRn;
         const int xo = 16 * cx + lr;
#pragma unroll
         for (int i = 0; i < 4; ++i)
         {
            const int yo = 16 * ry + 4 * (lane >> 4) + i;
            if (yo < P && xo < P)
               plane[(size_t)yo * P + xo] = eR[i] * eR[i] + eI[i] * eI[i];
         }
      }
#pragma unroll
      for (int u = PE; u < NDT; ++u)
         defocus_entry(u, tn);
      __syncthreads();  // the next factors in LDS; every wave's midT reads done before step 1 rewrites midT
   }
}

// Sum-1 normalisation per plane (Sim/ZernikePsf.cpp:401-410): FP64 sum, v = float(v / sum).
// One 1024-lane workgroup per plane: the FP64 sum (per lane, then a wave shuffle tree and the 16 wave sums in
// order), then v = (float)(v x (1 / sum)) in FP64 (within an FP64 ulp of the division, so the same float but at a
// rounding tie).
__global__ __launch_bounds__(1024) void psf_normalize_kernel(float* __restrict__ planes, int n2)
{
   __shared__ double part[16];
   float* p = planes + (size_t)blockIdx.x * n2;
   double s = 0.0;
   for (int i = threadIdx.x; i < n2; i += 1024)
      s += p[i];
   for (int o = 32; o > 0; o >>= 1)
      s += __shfl_xor(s, o);
   if ((threadIdx.x & 63) == 0)
      part[threadIdx.x >> 6] = s;
   __syncthreads();
   double sum = 0.0;
   for (int w = 0; w < 16; ++w)
      sum += part[w];
   if (sum > 0.0)
   {
      const double inv = 1.0 / sum;
      for (int i = threadIdx.x; i < n2; i += 1024)
         p[i] = (float)((double)p[i] * inv);
   }
}

// ---- host ---------------------------------------------------------------------------------------------------------

// GeometryFor + ZernikePupilSamples (Sim/ZernikePsf.cpp:118-148)
static void geometry(const iscg_psf_request* r, int* size, int* nz, int* M)
{
   const int os = std::max(1, r->oversampling);
   const int camHalf = std::max(1, r->kernelHalfWidthPx);
   const int sz = 2 * camHalf * os + 1;
   const int nzW = std::max(3, r->nz);
   *size = sz;
   *nz = (nzW % 2 == 1) ? nzW : nzW + 1;
   if (r->pupilSamples > 0)
      *M = std::max(8, std::min(1024, r->pupilSamples));
   else
   {
      const double diagonalUm = 2.0 * std::sqrt(2.0) * camHalf * r->pixelSizeNm / 1000.0;
      const double m = 4.0 + 2.0 * r->na * diagonalUm / (r->wavelengthNm / 1000.0);
      const int need = std::max(64, std::min(512, 4 * (int)std::ceil(m / 4.0)));
      int fftLen = 1;
      while (fftLen < need + sz - 1)
         fftLen <<= 1;
      *M = std::max(need, std::min(512, (fftLen - sz + 1) / 4 * 4));
   }
}

static void index_to_nm(int j, int* n, int* m)
{
   for (int nn = 0;; ++nn)
      for (int l = 0; l <= nn; ++l)
         if (nn * (nn + 1) / 2 + l == j)
         {
            *n = nn;
            *m = -nn + 2 * l;
            return;
         }
}

// Sim/ZernikePsf.cpp:150-197 (the request's SI scalars, the chirp-Z grid) into the device record.
static PsfReqDev request_dev(const iscg_psf_request* rq, int size, int nz, int M)
{
   PsfReqDev d{};
   const int os = std::max(1, rq->oversampling);
   const double resLateralM = (rq->pixelSizeNm / os) * 1E-9;
   d.NA = rq->na;
   const double lambda = rq->wavelengthNm * 1E-9;
   d.ni = rq->immersionIndex;
   d.ns = rq->sampleIndex;
   d.ti0 = rq->workingDistanceUm * 1E-6;
   d.depth = rq->sampleDepthNm * 1E-9;
   d.zStepM = rq->zStepNm * 1E-9;
   d.k0 = 2.0 * kPi / lambda;
   d.bMax = std::min(1.0, d.ns / d.NA);
   d.kMax = d.k0 * d.NA * d.bMax;
   d.dk = (2.0 * d.kMax) / (M - 4);
   d.kMin = -std::floor(M / 2.0) * d.dk;
   d.x0 = -((size - 1) / 2.0) * resLateralM;
   d.dx = resLateralM;
   d.focalShift = d.depth * (d.ni / d.ns);
   d.M = M;
   d.size = size;
   d.nz = nz;
   d.doubleHelix = rq->maskType == 1;
   d.maskModes = rq->maskModes;
   d.maskWaist = rq->maskWaist;
   int nm = 0;
   for (int j = 0; j < 28; ++j)
   {
      if (rq->zernike[j] == 0.0)
         continue;
      d.coeff[nm] = rq->zernike[j];
      index_to_nm(j, &d.n[nm], &d.l[nm]);
      d.m[nm] = std::abs(d.l[nm]);
      ++nm;
   }
   d.nModes = nm;
   // the folded polynomials: R_n^m(rho) = sum_k (-1)^k C(n-k, k) C(n-2k, (n-m)/2 - k) rho^(n-2k)
   std::memset(d.zc, 0, sizeof(d.zc));
   auto binom = [](int a, int b) {
      if (b < 0 || b > a)
         return 0.0;
      double r = 1.0;
      for (int i = 0; i < b; ++i)
         r = r * (a - i) / (i + 1);
      return r;
   };
   for (int t = 0; t < nm; ++t)
   {
      const int n = d.n[t], m = d.m[t];
      for (int k = 0; k <= (n - m) / 2; ++k)
      {
         const double c = (k % 2 == 0 ? 1.0 : -1.0) * binom(n - k, k) * binom(n - 2 * k, (n - m) / 2 - k);
         d.zc[d.l[t] >= 0 ? 0 : 1][m][n - 2 * k] += d.coeff[t] * c;
      }
   }
   return d;
}

template <int MP, int YB>
static void launch_planes(hipStream_t s, int items, const PsfReqDev* reqs, const double* A, const double* B,
                          const _Float16* W, const double* tz, int nz, int Pp, int nYB, float* out)
{
   hipLaunchKernelGGL((psf_planes_kernel<MP, YB>), dim3(items), dim3(kPsfThreads), 0, s, reqs, A, B, W, tz, nz, Pp,
                      nYB, out);
}

// Output-row blocks of YB tiles keep the pupil chunk (MP x 40 halves x 4) and the mid rows (YB 16 x (MP + 8) x 4)
// within ~150 KB of LDS.
// The small-plane kernels' launch choice: workgroups per request zg (1-32) plus kRbVariant for the row-block form.
constexpr int kRbVariant = 64;
static bool rb_eligible(const PsfBatch& s)
{
   return s.Pp <= kSmallPp && s.size <= kRbRows && (s.Mp == 32 || s.Mp == 64 || s.Mp == 128);
}

// choice: for planes of padded size <= 80 and M <= 128, the workgroups per request, + kRbVariant for
// psf_small_rb_kernel (rb_eligible shapes only).
static bool psf_planes_launch(iscg_context* h, const PsfBatch& s, int choice, float* out)
{
   const int nReq = s.nReq, nz = s.nz, Mp = s.Mp, Pp = s.Pp;
   const int zg = choice % kRbVariant;
   const bool rb = choice >= kRbVariant && rb_eligible(s);
   auto go = [&](auto mp, auto yb) {
      constexpr int MP = decltype(mp)::value, YB = decltype(yb)::value;
      if constexpr (MP == 32 || MP == 64 || MP == 128)
         if (rb)
         {
            const dim3 grid((nReq + kXcds - 1) / kXcds * kXcds * zg);
            // the tile loops resolved at compile time: every tile count at M 128 (its register budget is the
            // tightest), the 5 x 5 tiles of planes padded to 80 at the smaller pupils
            auto launch = [&](auto ctf) {
               hipLaunchKernelGGL((psf_small_rb_kernel<MP, decltype(ctf)::value>), grid, dim3(kPsfThreads), 0,
                                  h->stream, s.reqs.as<const PsfReqDev>(), s.tabP.as<const float2>(),
                                  s.tabB.as<const double>(), s.W.as<const _Float16>(), s.tz.as<const double>(), nz,
                                  Pp, nReq, zg, out);
            };
            const int ct = Pp / 16;
            if (ct == 5)
               launch(std::integral_constant<int, 5>{});
            else if constexpr (MP == 128)
            {
               if (ct == 4)
                  launch(std::integral_constant<int, 4>{});
               else if (ct == 3)
                  launch(std::integral_constant<int, 3>{});
               else if (ct == 2)
                  launch(std::integral_constant<int, 2>{});
               else
                  launch(std::integral_constant<int, 1>{});
            }
            else
               launch(std::integral_constant<int, 0>{});
            return;
         }
      if (Pp <= kSmallPp)
      {
         if constexpr (MP <= 128)
            hipLaunchKernelGGL((psf_small_kernel<MP, true>), dim3((nReq + kXcds - 1) / kXcds * kXcds * zg),
                               dim3(kPsfThreads), 0, h->stream, s.reqs.as<const PsfReqDev>(),
                               s.tabA.as<const double>(), s.tabB.as<const double>(), s.W.as<const _Float16>(),
                               s.tz.as<const double>(), nz, Pp, nReq, zg, out);
         else
            hipLaunchKernelGGL((psf_small_kernel<MP, false>), dim3(nReq * nz), dim3(kPsfThreads), 0, h->stream,
                               s.reqs.as<const PsfReqDev>(), s.tabA.as<const double>(), s.tabB.as<const double>(),
                               s.W.as<const _Float16>(), s.tz.as<const double>(), nz, Pp, nReq, 1, out);
         return;
      }
      // row blocks of YB tiles, fewer when the planes would leave CUs idle (each block evaluates the pupil again,
      // so the largest block count that still fills the 256 CUs)
      auto items = [&](int yb) { return nReq * nz * ((Pp + yb * 16 - 1) / (yb * 16)); };
      if (YB > 2 && items(YB) < 256 && items(2) >= 256)
         launch_planes<MP, 2>(h->stream, items(2), s.reqs.as<const PsfReqDev>(), s.tabA.as<const double>(),
                              s.tabB.as<const double>(), s.W.as<const _Float16>(), s.tz.as<const double>(), nz, Pp,
                              (Pp + 31) / 32, out);
      else if (YB > 1 && items(YB) < 256)
         launch_planes<MP, 1>(h->stream, items(1), s.reqs.as<const PsfReqDev>(), s.tabA.as<const double>(),
                              s.tabB.as<const double>(), s.W.as<const _Float16>(), s.tz.as<const double>(), nz, Pp,
                              (Pp + 15) / 16, out);
      else
      {
         const int nYB = (Pp + YB * 16 - 1) / (YB * 16);
         launch_planes<MP, YB>(h->stream, nReq * nz * nYB, s.reqs.as<const PsfReqDev>(), s.tabA.as<const double>(),
                               s.tabB.as<const double>(), s.W.as<const _Float16>(), s.tz.as<const double>(), nz, Pp,
                               nYB, out);
      }
   };
   if (Mp > 256)
   {
      // the chunked kernel: YB row blocks of 16 and PER2 E tiles per wave (PER2 <= 8: Pp <= 512 at YB 2, <= 1024 at 1)
      const int yb = Pp <= 512 ? 2 : 1;
      const int per2 = (yb * (Pp / 16) + kPsfWaves - 1) / kPsfWaves;
      const int nYB = (Pp + yb * 16 - 1) / (yb * 16);
      auto big = [&](auto ybc, auto perc) {
         constexpr int YB = decltype(ybc)::value, PER2 = decltype(perc)::value;
         hipLaunchKernelGGL((psf_planes_big_kernel<YB, PER2>), dim3(nReq * nz * nYB), dim3(kPsfThreads), 0,
                            h->stream, s.reqs.as<const PsfReqDev>(), s.tabA.as<const double>(),
                            s.tabB.as<const double>(), s.W.as<const _Float16>(), s.tz.as<const double>(), nz, Mp, Pp,
                            nYB, out);
      };
      auto byPer = [&](auto ybc) {
         switch (per2)
         {
         case 1: big(ybc, std::integral_constant<int, 1>{}); break;
         case 2: big(ybc, std::integral_constant<int, 2>{}); break;
         case 3: big(ybc, std::integral_constant<int, 3>{}); break;
         case 4: big(ybc, std::integral_constant<int, 4>{}); break;
         case 5: big(ybc, std::integral_constant<int, 5>{}); break;
         case 6: big(ybc, std::integral_constant<int, 6>{}); break;
         case 7: big(ybc, std::integral_constant<int, 7>{}); break;
         case 8: big(ybc, std::integral_constant<int, 8>{}); break;
         default: return false;
         }
         return true;
      };
      if (Mp > 1024 || Pp > 1024 || !(yb == 2 ? byPer(std::integral_constant<int, 2>{})
                                               : byPer(std::integral_constant<int, 1>{})))
         return fail(h, "iscg_psf_planes: pupil or plane size beyond 1024") == 0;
      return ok(h, hipGetLastError(), "PSF planes launch");
   }
   using I32 = std::integral_constant<int, 32>;
   switch (Mp)
   {
   case 32: go(I32{}, std::integral_constant<int, 5>{}); break;
   case 64: go(std::integral_constant<int, 64>{}, std::integral_constant<int, 5>{}); break;
   case 96: go(std::integral_constant<int, 96>{}, std::integral_constant<int, 5>{}); break;
   case 128: go(std::integral_constant<int, 128>{}, std::integral_constant<int, 5>{}); break;
   case 160: go(std::integral_constant<int, 160>{}, std::integral_constant<int, 4>{}); break;
   case 192: go(std::integral_constant<int, 192>{}, std::integral_constant<int, 3>{}); break;
   case 224: go(std::integral_constant<int, 224>{}, std::integral_constant<int, 2>{}); break;
   case 256: go(std::integral_constant<int, 256>{}, std::integral_constant<int, 2>{}); break;
   default: return fail(h, "iscg_psf_planes: unsupported pupil size") == 0;
   }
   return ok(h, hipGetLastError(), "PSF planes launch");
}

// Host side of a batch: validates the requests (one size, plane count and pupil sampling), fills the device records
// and per-plane ti - ti0, reserves the device buffers and uploads records and defocus on the context's stream. A
// batch that already holds nReq requests of the same geometry keeps its buffers (same addresses).
static int32_t psf_batch_set(iscg_context* h, PsfBatch& b, const iscg_psf_request* reqs, int32_t nReq,
                             const char* what)
{
   int size = 0, nz = 0, M = 0;
   std::vector<PsfReqDev> dev((size_t)nReq);
   for (int i = 0; i < nReq; ++i)
   {
      const iscg_psf_request* r = reqs + i;
      if (!(r->na > 0.0) || !(r->wavelengthNm > 0.0) || !(r->immersionIndex > 0.0) || !(r->sampleIndex > 0.0) ||
          !(r->pixelSizeNm > 0.0) || !(r->zStepNm > 0.0))
         return fail(h, "GibsonLanniZernike PSF: NA, wavelength, refractive indices, pixel size and z step must be > 0.");
      int sz, z, m;
      geometry(r, &sz, &z, &m);
      if (i == 0)
      {
         size = sz;
         nz = z;
         M = m;
      }
      else if (sz != size || z != nz || m != M)
         return fail(h, std::string(what) + ": the requests of one call must share the kernel size, plane count and "
                                            "pupil samples");
      dev[(size_t)i] = request_dev(r, sz, z, m);
   }
   const int Mp = (M + 31) / 32 * 32;
   if (Mp > 1024)
      return fail(h, std::string(what) + ": more than 1024 pupil samples");
   const int PpCheck = (size + 15) / 16 * 16;
   if (Mp > 256 && PpCheck > 1024)
      return fail(h, std::string(what) + ": planes wider than 1024 samples with more than 256 pupil samples");
   const int Pp = (size + 15) / 16 * 16;
   // per (request, plane): ti - ti0, as the reference forms it (Sim/ZernikePsf.cpp:238-244)
   std::vector<double> tz((size_t)nReq * nz);
   for (int i = 0; i < nReq; ++i)
      for (int z = 0; z < nz; ++z)
      {
         const PsfReqDev& d = dev[(size_t)i];
         const double ti = (d.ti0 - d.focalShift) + d.zStepM * (z - (nz - 1.0) / 2.0);
         tz[(size_t)i * nz + z] = ti - d.ti0;
      }
   // the previous upload from the host copies completes before they are replaced
   if (!ok(h, hipStreamSynchronize(h->stream), "PSF request upload"))
      return -1;
   const size_t wHalves = (size_t)nReq * 4 * Pp * Mp;
   if (!ok(h, b.reqs.reserve(dev.size() * sizeof(PsfReqDev)), "allocation") ||
       !ok(h, b.tabA.reserve((size_t)nReq * Mp * Mp * sizeof(double)), "allocation") ||
       !ok(h, b.tabB.reserve((size_t)nReq * Mp * Mp * sizeof(double)), "allocation") ||
       !ok(h, b.tabP.reserve((size_t)nReq * Mp * Mp * sizeof(float2)), "allocation") ||
       !ok(h, b.W.reserve(wHalves * sizeof(_Float16)), "allocation") ||
       !ok(h, b.tz.reserve(tz.size() * sizeof(double)), "allocation"))
      return -1;
   b.hostReqs = std::move(dev);
   b.hostTz = std::move(tz);
   if (!ok(h, hipMemcpyAsync(b.reqs.p, b.hostReqs.data(), b.hostReqs.size() * sizeof(PsfReqDev),
                             hipMemcpyHostToDevice, h->stream), "copy") ||
       !ok(h, hipMemcpyAsync(b.tz.p, b.hostTz.data(), b.hostTz.size() * sizeof(double), hipMemcpyHostToDevice,
                             h->stream), "copy"))
      return -1;
   b.nReq = nReq;
   b.size = size;
   b.nz = nz;
   b.M = M;
   b.Mp = Mp;
   b.Pp = Pp;
   return 0;
}

// The kernels of a batch, in stream order: pupil tables, split DFT matrices, planes, optional sum-1 normalisation
// into out (nReq x nz x size x size floats, device memory). No host copies, allocations or synchronisation unless
// the small-plane kernel's launch parameter is tuned for a new shape (never while the stream is captured).
static int32_t psf_batch_run(iscg_context* h, const PsfBatch& b, int32_t normalize, float* dOut)
{
   const int nReq = b.nReq, nz = b.nz, Mp = b.Mp, Pp = b.Pp;
   const PsfReqDev* dReq = b.reqs.as<const PsfReqDev>();
   const size_t nTab = (size_t)nReq * Mp * Mp;
   {
      ProfScope ps(h, "psf_tables");
      hipLaunchKernelGGL(psf_tables_kernel, dim3(grid_for(nTab, 256, 16)), dim3(256), 0, h->stream, dReq, nReq, Mp,
                         b.tabA.as<double>(), b.tabB.as<double>(), rb_eligible(b) ? b.tabP.as<float2>() : nullptr);
   }
   const size_t nW = (size_t)nReq * Pp * Mp;
   hipLaunchKernelGGL(psf_w_kernel, dim3(grid_for(nW, 256, 16)), dim3(256), 0, h->stream, dReq, nReq, Pp, Mp,
                      b.W.as<_Float16>());
   if (!ok(h, hipGetLastError(), "PSF table kernels"))
      return -1;
   // the small-plane kernel's workgroups per request: every candidate computes each plane in one workgroup the
   // same way, so the planes are bit-identical whichever is chosen (and the planes are written, not accumulated:
   // the timing launches before normalisation leave the same output). Where the shape allows the row-block form,
   // the candidates are its workgroups per request only (its pupil samples round differently from the other form's)
   int zg = kZGroups;
   if (Pp <= kSmallPp && Mp <= 128)
   {
      std::vector<int> cands;
      const int v = rb_eligible(b) ? 1 : 0;
      for (int c : {1, 2, 4, 8, 16, 32})
         if (c <= nz)
            cands.push_back(c + kRbVariant * v);
      const std::string key = "psf_small M" + std::to_string(Mp) + " P" + std::to_string(Pp) + " size" +
                              std::to_string(b.size) + " nz" + std::to_string(nz) + " req" +
                              std::to_string(pow2_bucket(nReq));
      // default (tuning off, or under graph capture before a tuning call): the row-block form with one workgroup
      // per request where it applies (the fastest at 512 requests of 65 x 65 x 41 planes, M 128)
      const int dflt = rb_eligible(b) ? kRbVariant + 1 : (nz >= kZGroups ? kZGroups : 1);
      zg = tuned_choice(h, key, cands, dflt, [&](int c) { return psf_planes_launch(h, b, c, dOut); });
   }
   {
      ProfScope ps(h, "psf_planes_mfma");
      if (!psf_planes_launch(h, b, zg, dOut))
         return -1;
   }
   if (normalize)
   {
      ProfScope ps(h, "psf_normalize");
      hipLaunchKernelGGL(psf_normalize_kernel, dim3((unsigned)(nReq * nz)), dim3(1024), 0, h->stream, dOut,
                         b.size * b.size);
      if (!ok(h, hipGetLastError(), "PSF normalisation"))
         return -1;
   }
   return 0;
}

}  // namespace iscg

using namespace iscg;

extern "C" {

int32_t iscg_psf_geometry(const iscg_psf_request* r, int32_t* size, int32_t* nz, int32_t* pupilSamples)
{
   if (!r)
      return -1;
   int s, z, m;
   geometry(r, &s, &z, &m);
   if (size)
      *size = s;
   if (nz)
      *nz = z;
   if (pupilSamples)
      *pupilSamples = m;
   return 0;
}

int32_t iscg_psf_planes(iscg_context* h, const iscg_psf_request* reqs, int32_t nReq, int32_t normalize, float* out,
                        size_t outCap)
{
   if (begin(h))
      return -1;
   if (!reqs || nReq <= 0 || !out)
      return fail(h, "iscg_psf_planes: null argument or no request");
   if (!h->psf)
      h->psf = new PsfState();
   PsfBatch& b = h->psf->batch;
   if (psf_batch_set(h, b, reqs, nReq, "iscg_psf_planes"))
      return -1;
   const size_t need = (size_t)nReq * b.nz * b.size * b.size;
   if (outCap < need)
      return fail(h, "iscg_psf_planes: output too small (" + std::to_string(need) + " floats needed)");
   float* dOut = (float*)stage_out(h, 0, out, need * sizeof(float));
   if (!dOut)
      return -1;
   if (psf_batch_run(h, b, normalize, dOut))
      return -1;
   if (!finish_out(h, out, dOut, need * sizeof(float)))
      return -1;
   return 0;
}

struct iscg_psf_library
{
   iscg::PsfBatch b;
   int device = 0;
};

int32_t iscg_psf_library_create(iscg_context* h, const iscg_psf_request* reqs, int32_t nReq, iscg_psf_library** lib)
{
   if (begin(h))
      return -1;
   if (!reqs || nReq <= 0 || !lib)
      return fail(h, "iscg_psf_library_create: null argument or no request");
   *lib = nullptr;
   iscg_psf_library* L = new iscg_psf_library();
   L->device = h->device;
   if (psf_batch_set(h, L->b, reqs, nReq, "iscg_psf_library_create") ||
       !ok(h, hipStreamSynchronize(h->stream), "iscg_psf_library_create"))
   {
      delete L;
      return -1;
   }
   *lib = L;
   return 0;
}

int32_t iscg_psf_library_update(iscg_context* h, iscg_psf_library* lib, const iscg_psf_request* reqs, int32_t nReq)
{
   if (begin(h))
      return -1;
   if (!lib || !reqs)
      return fail(h, "iscg_psf_library_update: null argument");
   if (lib->device != h->device)
      return fail(h, "iscg_psf_library_update: the library belongs to another device");
   if (nReq != lib->b.nReq)
      return fail(h, "iscg_psf_library_update: the request count must stay " + std::to_string(lib->b.nReq));
   for (int i = 0; i < nReq; ++i)
   {
      int sz, z, m;
      geometry(reqs + i, &sz, &z, &m);
      if (sz != lib->b.size || z != lib->b.nz || m != lib->b.M)
         return fail(h, "iscg_psf_library_update: the kernel size, plane count and pupil samples must stay the same");
   }
   const void* before = lib->b.reqs.p;
   if (psf_batch_set(h, lib->b, reqs, nReq, "iscg_psf_library_update"))
      return -1;
   if (lib->b.reqs.p != before)
      return fail(h, "iscg_psf_library_update: device buffers moved");
   return ok(h, hipStreamSynchronize(h->stream), "iscg_psf_library_update") ? 0 : -1;
}

int32_t iscg_psf_library_geometry(const iscg_psf_library* lib, int32_t* nReq, int32_t* size, int32_t* nz)
{
   if (!lib)
      return -1;
   if (nReq)
      *nReq = lib->b.nReq;
   if (size)
      *size = lib->b.size;
   if (nz)
      *nz = lib->b.nz;
   return 0;
}

int32_t iscg_psf_library_run(iscg_context* h, iscg_psf_library* lib, int32_t normalize, float* out, size_t outCap)
{
   if (begin(h))
      return -1;
   if (!lib || !out)
      return fail(h, "iscg_psf_library_run: null argument");
   if (lib->device != h->device)
      return fail(h, "iscg_psf_library_run: the library belongs to another device");
   const size_t need = (size_t)lib->b.nReq * lib->b.nz * lib->b.size * lib->b.size;
   if (outCap < need)
      return fail(h, "iscg_psf_library_run: output too small (" + std::to_string(need) + " floats needed)");
   float* dOut = (float*)stage_out(h, 0, out, need * sizeof(float));
   if (!dOut)
      return -1;
   if (psf_batch_run(h, lib->b, normalize, dOut))
      return -1;
   if (!finish_out(h, out, dOut, need * sizeof(float)))
      return -1;
   return 0;
}

int32_t iscg_psf_library_destroy(iscg_context* h, iscg_psf_library* lib)
{
   if (!lib)
      return 0;
   if (h && hipSetDevice(lib->device) == hipSuccess)
      (void)hipStreamSynchronize(h->stream);
   delete lib;
   return 0;
}

}  // extern "C"
