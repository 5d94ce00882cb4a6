// wf_passes.h -- the FFT passes of the mean-field entry points (wf.hip: the whole-stack 3D convolution;
// wf_accel.hip: the per-focus jobs of the reference's WidefieldAccelerator seam). Planes of complex half spectra are
// [plane][row ky][S] float2 with S = NX/2 + 1 rounded up to 16 (128-byte aligned row segments); every FFT line runs in
// LDS (fft_lds.h). The transforms are unnormalised: callers fold the 1/(NX NY) into a spectrum.
#pragma once
#include "iscg_internal.h"
#include "fft_lds.h"

#include <cmath>
#include <vector>

namespace iscg {

constexpr int kCols = 16;  // columns per column block

static inline bool pow2(uint32_t n, uint32_t maxN) { return n >= 16 && n <= maxN && (n & (n - 1)) == 0; }

static inline uint32_t spec_stride(uint32_t NX) { return ((NX / 2 + 1) + kCols - 1) / kCols * kCols; }

// Pass A: rows R2C. Rows r and r + 1 of the real input go into one complex line a + ib, zero beyond nIn; after the
// FFT, A[k] = (Z[k] + conj Z[N-k]) / 2, B[k] = -i (Z[k] - conj Z[N-k]) / 2 for k <= N/2.
template <int N>
__global__ __launch_bounds__(256) void rows_r2c_kernel(const float* __restrict__ in, size_t inRowStride, int nRows,
                                                       int nIn, float2* __restrict__ out, int S,
                                                       const float2* __restrict__ tw)
{
   constexpr int T = N / 8, LINES = 256 / T > 0 ? 256 / T : 1, LS = N + 1;
   __shared__ float2 buf[LINES * LS];
   const int tid = threadIdx.x;
   const int pairs = (nRows + 1) / 2;
   for (int p0 = blockIdx.x * LINES; p0 < pairs; p0 += gridDim.x * LINES)
   {
      for (int i = tid; i < LINES * N; i += blockDim.x)
      {
         const int l = i / N, n = i % N, pr = p0 + l;
         const int ra = 2 * pr, rb = 2 * pr + 1;
         float a = 0.0f, b = 0.0f;
         if (pr < pairs && n < nIn)
         {
            a = in[(size_t)ra * inRowStride + n];
            b = rb < nRows ? in[(size_t)rb * inRowStride + n] : 0.0f;
         }
         buf[l * LS + n] = make_float2(a, b);
      }
      __syncthreads();
      fft_lines<N, false, false>(buf, LS, 1, LINES, tid, tw);
      for (int i = tid; i < LINES * (N / 2 + 1); i += blockDim.x)
      {
         const int l = i / (N / 2 + 1), k = i % (N / 2 + 1), pr = p0 + l;
         if (pr >= pairs)
            continue;
         const float2 z = buf[l * LS + k], zc = buf[l * LS + ((N - k) & (N - 1))];
         const float2 A = make_float2(0.5f * (z.x + zc.x), 0.5f * (z.y - zc.y));
         const float2 B = make_float2(0.5f * (z.y + zc.y), -0.5f * (z.x - zc.x));
         out[(size_t)(2 * pr) * S + k] = A;
         if (2 * pr + 1 < nRows)
            out[(size_t)(2 * pr + 1) * S + k] = B;
      }
      __syncthreads();
   }
}

// Passes B and D: FFT along the rows of [planes][rows][S] for blocks of KC columns (16; 8 at N = 1024 and 4 at 2048,
// so that a block stays within 1024 lanes and 100 KB of LDS); rows >= rowsIn are zero, rows [row0, row0 + rowsOut)
// are written.
template <int N>
constexpr int cols_block() { return N > 1024 ? 4 : (N > 512 ? 8 : kCols); }
template <int N, bool INV, int KC = cols_block<N>()>
__global__ __launch_bounds__(KC * N / 8) void cols_kernel(const float2* __restrict__ in, int rowsIn,
                                                          size_t inPlaneStride, float2* __restrict__ out,
                                                          int row0, int rowsOut, size_t outPlaneStride,
                                                          int nPlanes, int S, int nColsValid,
                                                          const float2* __restrict__ tw)
{
   constexpr int kCols = KC;
   constexpr int CS = kCols + 1;  // padded row of the block in LDS
   __shared__ float2 buf[N * CS];
   __shared__ float2 twl[N];  // the twiddles, read by every stage
   const int tid = threadIdx.x;
   for (int i = tid; i < N; i += blockDim.x)
      twl[i] = tw[i];
   __syncthreads();
   const int colBlocks = (nColsValid + kCols - 1) / kCols;
   const int items = nPlanes * colBlocks;
   for (int it = blockIdx.x; it < items; it += gridDim.x)
   {
      const int pl = it / colBlocks, c0 = (it % colBlocks) * kCols;
      const float2* src = in + (size_t)pl * inPlaneStride;
      for (int i = tid; i < N * kCols; i += blockDim.x)
      {
         const int n = i / kCols, c = i % kCols;
         buf[n * CS + c] = n < rowsIn ? src[(size_t)n * S + c0 + c] : make_float2(0.0f, 0.0f);
      }
      __syncthreads();
      fft_lines<N, INV, true>(buf, 1, CS, kCols, tid, twl);
      float2* dst = out + (size_t)pl * outPlaneStride;
      for (int i = tid; i < rowsOut * kCols; i += blockDim.x)
      {
         const int n = i / kCols, c = i % kCols;
         dst[(size_t)n * S + c0 + c] = buf[(row0 + n) * CS + c];
      }
      __syncthreads();
   }
}

struct C2rEpilogue
{
   int raw;               // 1: the transform's value as is (no max(0, .), no background)
   float* img;            // [rows][cw] or null
   uint16_t* adu;         // [rows][cw] or null
   const float* offset;   // [ch][cw] or null
   double background;
   CamParams cp;
   NoiseQueue q;          // deferred pixels (index o = row * cw + x)
   uint32_t seed, frameBase, pixelBase;
   int ch, cw, x0;
};

// The epilogue of the C2R passes: line l of the block holds rows 2 (p0 + l) (real part) and 2 (p0 + l) + 1
// (imaginary part) after the inverse transform; columns [x0, x0 + cw) kept. The frame is max(0, image) + background
// (WidefieldImages::Render, Sim/WidefieldRender.cpp:547-580), then the noise chain (frame frameBase + row / ch,
// pixel (row % ch) cw + x + pixelBase).
template <bool NOISE, bool EMCCD>
__device__ __forceinline__ void c2r_epilogue(const C2rEpilogue& e, const float2* buf, int LS, int LINES, int p0,
                                             int pairs, int nRows, int tid)
{
   for (int i = tid; i < LINES * 2 * e.cw; i += blockDim.x)
   {
      const int l = i / (2 * e.cw), rem = i % (2 * e.cw), half = rem / e.cw, x = rem % e.cw, pr = p0 + l;
      const int row = 2 * pr + half;
      if (pr >= pairs || row >= nRows)
         continue;
      const float2 z = buf[l * LS + e.x0 + x];
      const float v = half ? z.y : z.x;
      const float ph = e.raw ? v : (float)((double)(v > 0.0f ? v : 0.0f) + e.background);
      const int f = row / e.ch, y = row % e.ch;
      const size_t o = (size_t)row * e.cw + x;
      if (e.img)
         e.img[o] = ph;
      if (NOISE)
      {
         const uint32_t pix = (uint32_t)(y * e.cw + x);
         noise_store<EMCCD>(e.cp, e.q, e.adu, (uint32_t)o, ph, e.seed, e.frameBase + (uint32_t)f, pix + e.pixelBase,
                            e.offset != nullptr, e.offset ? e.offset[pix] : 0.0f, false, 0.0f, false, 0.0f);
      }
   }
}

// Pass E: rows C2R. Half spectra A (row r) and B (row r + 1) become Z = A + iB over the full length (Hermitian
// extension; DC and Nyquist real), the inverse FFT gives a + ib. Columns [x0, x0 + cw) kept; max(0, .) +
// background, then the noise chain (frame frameBase + r / ch, pixel (r % ch) cw + x + pixelBase).
template <int N, bool NOISE, bool EMCCD>
__global__ __launch_bounds__(256) void rows_c2r_kernel(const float2* __restrict__ in, int S, int nRows,
                                                       C2rEpilogue e, const float2* __restrict__ tw)
{
   constexpr int T = N / 8, LINES = 256 / T > 0 ? 256 / T : 1, LS = N + 1;
   __shared__ float2 buf[LINES * LS];
   const int tid = threadIdx.x;
   const int pairs = (nRows + 1) / 2;
   for (int p0 = blockIdx.x * LINES; p0 < pairs; p0 += gridDim.x * LINES)
   {
      for (int i = tid; i < LINES * (N / 2 + 1); i += blockDim.x)
      {
         const int l = i / (N / 2 + 1), k = i % (N / 2 + 1), pr = p0 + l;
         float2 A = make_float2(0.0f, 0.0f), B = make_float2(0.0f, 0.0f);
         if (pr < pairs)
         {
            A = in[(size_t)(2 * pr) * S + k];
            if (2 * pr + 1 < nRows)
               B = in[(size_t)(2 * pr + 1) * S + k];
         }
         if (k == 0 || k == N / 2)
         {
            A.y = 0.0f;
            B.y = 0.0f;
         }
         // Z[k] = A + iB; Z[N-k] = conj(A) + i conj(B)
         buf[l * LS + k] = make_float2(A.x - B.y, A.y + B.x);
         if (k != 0 && k != N / 2)
            buf[l * LS + N - k] = make_float2(A.x + B.y, B.x - A.y);
      }
      __syncthreads();
      fft_lines<N, true, false>(buf, LS, 1, LINES, tid, tw);
      c2r_epilogue<NOISE, EMCCD>(e, buf, LS, LINES, p0, pairs, nRows, tid);
      __syncthreads();
   }
}

// Passes A and E for the fused slab pass (wf.hip), whose half spectra are column-major: column k's rows are
// contiguous, element (row, k) at k * colStride + row (colStride even). kCmThreads threads, kCmThreads / (N / 8) lines
// per block (N = 512: 8 complex lines = 16 rows, 128-byte column segments); rows r and r + 1 move as one 16-byte pair.
// The columns 0 and N/2 of a real row are real; they travel packed as one complex column, a + ib in column 0 (the
// slab pass transforms the pair as one), and column N/2 is neither written nor read.
constexpr int kCmThreads = 256;
template <int N>
__global__ __launch_bounds__(kCmThreads) void rows_r2c_cm_kernel(const float* __restrict__ in, size_t inRowStride,
                                                                 int nRows, int nIn, float2* __restrict__ out,
                                                                 size_t colStride, const float2* __restrict__ tw)
{
   constexpr int T = N / 8, LINES = kCmThreads / T > 0 ? kCmThreads / T : 1, LS = N + 1;
   __shared__ float2 buf[LINES * LS];
   const int tid = threadIdx.x;
   const int pairs = (nRows + 1) / 2;
   for (int p0 = blockIdx.x * LINES; p0 < pairs; p0 += gridDim.x * LINES)
   {
      for (int i = tid; i < LINES * N; i += blockDim.x)
      {
         const int l = i / N, n = i % N, pr = p0 + l;
         const int ra = 2 * pr, rb = 2 * pr + 1;
         float a = 0.0f, b = 0.0f;
         if (pr < pairs && n < nIn)
         {
            a = in[(size_t)ra * inRowStride + n];
            b = rb < nRows ? in[(size_t)rb * inRowStride + n] : 0.0f;
         }
         buf[l * LS + n] = make_float2(a, b);
      }
      __syncthreads();
      fft_lines<N, false, false>(buf, LS, 1, LINES, tid, tw);
      for (int i = tid; i < LINES * (N / 2 + 1); i += blockDim.x)
      {
         const int k = i / LINES, l = i % LINES, pr = p0 + l;
         if (pr >= pairs)
            continue;
         if (k == N / 2)
            continue;
         const float2 z = buf[l * LS + k], zc = buf[l * LS + ((N - k) & (N - 1))];
         float2 A = make_float2(0.5f * (z.x + zc.x), 0.5f * (z.y - zc.y));
         float2 B = make_float2(0.5f * (z.y + zc.y), -0.5f * (z.x - zc.x));
         if (k == 0)
         {
            // columns 0 and N/2, both real: z[0] = a0 + i b0, z[N/2] = aN + i bN
            const float2 zN = buf[l * LS + N / 2];
            A = make_float2(z.x, zN.x);
            B = make_float2(z.y, zN.y);
         }
         float2* o = out + (size_t)k * colStride + 2 * pr;
         if (2 * pr + 1 < nRows)
            *reinterpret_cast<float4*>(o) = make_float4(A.x, A.y, B.x, B.y);
         else
            o[0] = A;
      }
      __syncthreads();
   }
}

template <int N, bool NOISE, bool EMCCD>
__global__ __launch_bounds__(kCmThreads) void rows_c2r_cm_kernel(const float2* __restrict__ in, size_t colStride,
                                                                 int nRows, C2rEpilogue e,
                                                                 const float2* __restrict__ tw)
{
   constexpr int T = N / 8, LINES = kCmThreads / T > 0 ? kCmThreads / T : 1, LS = N + 1;
   __shared__ float2 buf[LINES * LS];
   const int tid = threadIdx.x;
   const int pairs = (nRows + 1) / 2;
   for (int p0 = blockIdx.x * LINES; p0 < pairs; p0 += gridDim.x * LINES)
   {
      for (int i = tid; i < LINES * (N / 2 + 1); i += blockDim.x)
      {
         const int k = i / LINES, l = i % LINES, pr = p0 + l;
         float2 A = make_float2(0.0f, 0.0f), B = make_float2(0.0f, 0.0f);
         if (pr < pairs)
         {
            // column N/2 is the imaginary part of the packed column 0
            const float2* s = in + (size_t)(k == N / 2 ? 0 : k) * colStride + 2 * pr;
            if (2 * pr + 1 < nRows)
            {
               const float4 ab = *reinterpret_cast<const float4*>(s);
               A = make_float2(ab.x, ab.y);
               B = make_float2(ab.z, ab.w);
            }
            else
               A = s[0];
            if (k == N / 2)
            {
               A.x = A.y;
               B.x = B.y;
            }
         }
         if (k == 0 || k == N / 2)
         {
            A.y = 0.0f;
            B.y = 0.0f;
         }
         buf[l * LS + k] = make_float2(A.x - B.y, A.y + B.x);
         if (k != 0 && k != N / 2)
            buf[l * LS + N - k] = make_float2(A.x + B.y, B.x - A.y);
      }
      __syncthreads();
      fft_lines<N, true, false>(buf, LS, 1, LINES, tid, tw);
      c2r_epilogue<NOISE, EMCCD>(e, buf, LS, LINES, p0, pairs, nRows, tid);
      __syncthreads();
   }
}

// ---- dispatch over the supported lengths ----------------------------------------------------------------------

// Lengths 16..MAXN (at most 2048: a column block of cols_block<N>() lines fits 1024 lanes).
template <template <int> class F, int MAXN = 512, typename... A>
static bool dispatch_n(int n, A... a)
{
   switch (n)
   {
   case 16: F<16>::run(a...); return true;
   case 32: F<32>::run(a...); return true;
   case 64: F<64>::run(a...); return true;
   case 128: F<128>::run(a...); return true;
   case 256: F<256>::run(a...); return true;
   case 512: F<512>::run(a...); return true;
   case 1024:
      if constexpr (MAXN >= 1024)
      {
         F<1024>::run(a...);
         return true;
      }
      return false;
   case 2048:
      if constexpr (MAXN >= 2048)
      {
         F<2048>::run(a...);
         return true;
      }
      return false;
   default: return false;
   }
}

static unsigned wgs(size_t items, unsigned cap = 256 * 8) { return (unsigned)(items < cap ? (items ? items : 1) : cap); }

template <int N>
struct RowsR2C
{
   static void run(hipStream_t s, const float* in, size_t inRowStride, int nRows, int nIn, float2* out, int S,
                   const float2* tw)
   {
      constexpr int T = N / 8, LINES = 256 / T > 0 ? 256 / T : 1;
      const size_t items = ((size_t)(nRows + 1) / 2 + LINES - 1) / LINES;
      hipLaunchKernelGGL(rows_r2c_kernel<N>, dim3(wgs(items)), dim3(256), 0, s, in, inRowStride, nRows, nIn, out, S,
                         tw);
   }
};

template <int N>
struct ColsF
{
   static void run(hipStream_t s, bool inv, const float2* in, int rowsIn, size_t inPS, float2* out, int row0,
                   int rowsOut, size_t outPS, int nPlanes, int S, int nCols, const float2* tw)
   {
      constexpr int KC = cols_block<N>();
      const size_t items = (size_t)nPlanes * ((nCols + KC - 1) / KC);
      if (inv)
         hipLaunchKernelGGL((cols_kernel<N, true>), dim3(wgs(items)), dim3(KC * N / 8), 0, s, in, rowsIn, inPS, out,
                            row0, rowsOut, outPS, nPlanes, S, nCols, tw);
      else
         hipLaunchKernelGGL((cols_kernel<N, false>), dim3(wgs(items)), dim3(KC * N / 8), 0, s, in, rowsIn, inPS, out,
                            row0, rowsOut, outPS, nPlanes, S, nCols, tw);
   }
};

template <int N>
struct RowsC2R
{
   static void run(hipStream_t s, const float2* in, int S, int nRows, const C2rEpilogue& e, const float2* tw)
   {
      constexpr int T = N / 8, LINES = 256 / T > 0 ? 256 / T : 1;
      const size_t items = ((size_t)(nRows + 1) / 2 + LINES - 1) / LINES;
      if (!e.adu)
         hipLaunchKernelGGL((rows_c2r_kernel<N, false, false>), dim3(wgs(items)), dim3(256), 0, s, in, S, nRows, e,
                            tw);
      else if (e.cp.emccd)
         hipLaunchKernelGGL((rows_c2r_kernel<N, true, true>), dim3(wgs(items)), dim3(256), 0, s, in, S, nRows, e, tw);
      else
         hipLaunchKernelGGL((rows_c2r_kernel<N, true, false>), dim3(wgs(items)), dim3(256), 0, s, in, S, nRows, e,
                            tw);
   }
};

template <int N>
struct RowsR2Ccm
{
   static void run(hipStream_t s, const float* in, size_t inRowStride, int nRows, int nIn, float2* out,
                   size_t colStride, const float2* tw)
   {
      constexpr int LINES = kCmThreads / (N / 8) > 0 ? kCmThreads / (N / 8) : 1;
      const size_t items = ((size_t)(nRows + 1) / 2 + LINES - 1) / LINES;
      hipLaunchKernelGGL(rows_r2c_cm_kernel<N>, dim3(wgs(items, 256 * 8)), dim3(kCmThreads), 0, s, in, inRowStride,
                         nRows, nIn, out, colStride, tw);
   }
};

template <int N>
struct RowsC2Rcm
{
   static void run(hipStream_t s, const float2* in, size_t colStride, int nRows, const C2rEpilogue& e,
                   const float2* tw)
   {
      constexpr int LINES = kCmThreads / (N / 8) > 0 ? kCmThreads / (N / 8) : 1;
      const size_t items = ((size_t)(nRows + 1) / 2 + LINES - 1) / LINES;
      const dim3 g(wgs(items, 256 * 8)), b(kCmThreads);
      if (!e.adu)
         hipLaunchKernelGGL((rows_c2r_cm_kernel<N, false, false>), g, b, 0, s, in, colStride, nRows, e, tw);
      else if (e.cp.emccd)
         hipLaunchKernelGGL((rows_c2r_cm_kernel<N, true, true>), g, b, 0, s, in, colStride, nRows, e, tw);
      else
         hipLaunchKernelGGL((rows_c2r_cm_kernel<N, true, false>), g, b, 0, s, in, colStride, nRows, e, tw);
   }
};

static bool make_twiddles(iscg_context* h, DevBuf& b, uint32_t n)
{
   std::vector<float2> t(n);
   for (uint32_t k = 0; k < n; ++k)
   {
      const double a = -2.0 * 3.14159265358979323846 * (double)k / (double)n;
      t[k] = make_float2((float)std::cos(a), (float)std::sin(a));
   }
   if (!ok(h, b.reserve(n * sizeof(float2)), "twiddle allocation"))
      return false;
   return ok(h, hipMemcpyAsync(b.p, t.data(), n * sizeof(float2), hipMemcpyHostToDevice, h->stream), "twiddle copy") &&
          ok(h, hipStreamSynchronize(h->stream), "stream synchronize");
}

static uint32_t next_pow2(uint32_t n)
{
   uint32_t p = 16;
   while (p < n)
      p <<= 1;
   return p;
}

// The grid of a 2D transform: sizes, row stride and twiddles.
struct XyGrid
{
   uint32_t NX, NY, S;
   const float2* twX;
   const float2* twY;
};

// Passes A + B: [nzIn][nyIn][nxIn] real (row stride nxIn, zero beyond) -> x2 [nzIn][NY][S] (plane stride
// outPlaneStride), through x1 [nzIn][nyIn][S].
static bool forward_xy(iscg_context* h, const XyGrid& g, const float* in, int nzIn, int nyIn, int nxIn, float2* x1,
                       float2* x2, size_t outPlaneStride)
{
   const int S = (int)g.S;
   {
      ProfScope ps(h, "wf_rows_r2c");
      if (!dispatch_n<RowsR2C, 2048>((int)g.NX, h->stream, in, (size_t)nxIn, nzIn * nyIn, nxIn, x1, S, g.twX))
         return false;
   }
   ProfScope ps(h, "wf_cols_fwd");
   return dispatch_n<ColsF, 2048>((int)g.NY, h->stream, false, (const float2*)x1, nyIn, (size_t)nyIn * S, x2, 0,
                            (int)g.NY, outPlaneStride, nzIn, S, (int)(g.NX / 2 + 1), g.twY);
}

// Passes D + E: in [nPlanes][NY][S] -> rows [fovY0, fovY0 + ch) through x2 [nPlanes][ch][S] -> the epilogue's
// [nPlanes][ch][cw] outputs (columns from e.x0).
static bool inverse_xy(iscg_context* h, const XyGrid& g, const float2* in, int nPlanes, float2* x2, int fovY0,
                       const C2rEpilogue& e, const char* c2rName)
{
   const int S = (int)g.S;
   {
      ProfScope ps(h, "wf_cols_inv");
      if (!dispatch_n<ColsF, 2048>((int)g.NY, h->stream, true, in, (int)g.NY, (size_t)g.NY * S, x2, fovY0, e.ch,
                             (size_t)e.ch * S, nPlanes, S, (int)(g.NX / 2 + 1), g.twY))
         return false;
   }
   ProfScope ps(h, c2rName);
   return dispatch_n<RowsC2R, 2048>((int)g.NX, h->stream, (const float2*)x2, S, nPlanes * e.ch, e, g.twX);
}

}  // namespace iscg
