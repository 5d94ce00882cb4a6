// fft_lds.h -- FP32 complex FFTs of power-of-two length on lines held in LDS, for 64-wide wavefronts.
//
// Stockham autosort, radix 8 (then one radix-4 or radix-2 stage for the remaining factor), every thread holding
// 8 points per stage in registers: a length-N line is worked by N/8 lanes, 2 barriers per stage, log8(N) stages
// (512: 3 stages). Twiddles come from a table tw[k] = exp(-2 pi i k / N) computed in FP64 on the host (the
// transforms stay within a few FP32 ulps of the exact DFT). Unnormalised in both directions.
//
// Addressing is left to the caller: element n of line l sits at buf[l * lineStride + n * elemStride]. Thread
// mapping: LINE_FAST = true puts consecutive lanes on consecutive lines (for column blocks stored row-major, where
// lines are adjacent in LDS), false puts them on consecutive elements of one line (rows stored line-major). Both
// give consecutive LDS addresses across lanes.
#pragma once

#include <hip/hip_runtime.h>

namespace iscg {

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
   return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ float2 rot90(float2 a)
{
   return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

template <int R, bool INV>
__device__ __forceinline__ void dft(float2* v);

template <>
__device__ __forceinline__ void dft<2, false>(float2* v)
{
   const float2 a = v[0], b = v[1];
   v[0] = cadd(a, b);
   v[1] = csub(a, b);
}
template <>
__device__ __forceinline__ void dft<2, true>(float2* v)
{
   dft<2, false>(v);
}

template <bool INV>
__device__ __forceinline__ void dft4(float2& x0, float2& x1, float2& x2, float2& x3)
{
   const float2 s02 = cadd(x0, x2), d02 = csub(x0, x2);
   const float2 s13 = cadd(x1, x3), d13 = rot90<INV>(csub(x1, x3));
   x0 = cadd(s02, s13);
   x2 = csub(s02, s13);
   x1 = cadd(d02, d13);
   x3 = csub(d02, d13);
}

template <>
__device__ __forceinline__ void dft<4, false>(float2* v)
{
   dft4<false>(v[0], v[1], v[2], v[3]);
}
template <>
__device__ __forceinline__ void dft<4, true>(float2* v)
{
   dft4<true>(v[0], v[1], v[2], v[3]);
}

template <bool INV>
__device__ __forceinline__ void dft8(float2* v)
{
   constexpr float h = 0.70710678118654752440f;
   float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
   float2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
   dft4<INV>(e0, e1, e2, e3);
   dft4<INV>(o0, o1, o2, o3);
   // o_k *= W8^k, W8 = exp(-+ 2 pi i / 8)
   o1 = INV ? make_float2(h * (o1.x - o1.y), h * (o1.x + o1.y)) : make_float2(h * (o1.x + o1.y), h * (o1.y - o1.x));
   o2 = rot90<INV>(o2);
   o3 = INV ? make_float2(-h * (o3.x + o3.y), h * (o3.x - o3.y)) : make_float2(h * (o3.y - o3.x), -h * (o3.x + o3.y));
   v[0] = cadd(e0, o0);
   v[4] = csub(e0, o0);
   v[1] = cadd(e1, o1);
   v[5] = csub(e1, o1);
   v[2] = cadd(e2, o2);
   v[6] = csub(e2, o2);
   v[3] = cadd(e3, o3);
   v[7] = csub(e3, o3);
}
template <>
__device__ __forceinline__ void dft<8, false>(float2* v)
{
   dft8<false>(v);
}
template <>
__device__ __forceinline__ void dft<8, true>(float2* v)
{
   dft8<true>(v);
}

// One Stockham stage of radix R on every line of the workgroup; NS = the product of the earlier radices.
template <int N, int R, int NS, bool INV, bool LINE_FAST>
__device__ __forceinline__ void fft_stage(float2* buf, int lineStride, int elemStride, int nLines, int tid,
                                          const float2* __restrict__ tw)
{
   constexpr int T = N / 8;     // lanes per line
   constexpr int PER = 8 / R;   // butterflies per lane
   constexpr int B = N / R;     // butterflies per line
   const int line = LINE_FAST ? tid % nLines : tid / T;
   const int q = LINE_FAST ? tid / nLines : tid % T;
   const bool active = line < nLines && q < T;
   float2 v[PER][R];
   if (active)
   {
      float2* L = buf + line * lineStride;
#pragma unroll
      for (int s = 0; s < PER; ++s)
      {
         const int j = q + s * T;
#pragma unroll
         for (int r = 0; r < R; ++r)
            v[s][r] = L[(j + r * B) * elemStride];
      }
#pragma unroll
      for (int s = 0; s < PER; ++s)
      {
         const int j = q + s * T;
         const int k = j % NS;
         if (NS > 1)
         {
#pragma unroll
            for (int r = 1; r < R; ++r)
            {
               float2 w = tw[r * k * (N / (NS * R))];
               if (INV)
                  w.y = -w.y;
               v[s][r] = cmul(v[s][r], w);
            }
         }
         dft<R, INV>(v[s]);
      }
   }
   __syncthreads();
   if (active)
   {
      float2* L = buf + line * lineStride;
#pragma unroll
      for (int s = 0; s < PER; ++s)
      {
         const int j = q + s * T;
         const int k = j % NS;
         const int base = (j / NS) * NS * R + k;
#pragma unroll
         for (int r = 0; r < R; ++r)
            L[(base + r * NS) * elemStride] = v[s][r];
      }
   }
   __syncthreads();
}

// The whole transform of length N (a power of two, 16..2048) on nLines lines; N/8 * nLines lanes take part (the
// rest idle through the barriers). Every thread of the workgroup must call it.
template <int N, bool INV, bool LINE_FAST>
__device__ __forceinline__ void fft_lines(float2* buf, int lineStride, int elemStride, int nLines, int tid,
                                          const float2* __restrict__ tw)
{
   static_assert(N >= 16 && (N & (N - 1)) == 0, "power-of-two length >= 16");
   // radix-8 stages while >= 8 remains, then a radix-4 or radix-2 stage
   if constexpr (N >= 8)
      fft_stage<N, 8, 1, INV, LINE_FAST>(buf, lineStride, elemStride, nLines, tid, tw);
   if constexpr (N >= 64)
      fft_stage<N, 8, 8, INV, LINE_FAST>(buf, lineStride, elemStride, nLines, tid, tw);
   if constexpr (N >= 512)
      fft_stage<N, 8, 64, INV, LINE_FAST>(buf, lineStride, elemStride, nLines, tid, tw);
   if constexpr (N >= 4096)
      fft_stage<N, 8, 512, INV, LINE_FAST>(buf, lineStride, elemStride, nLines, tid, tw);
   constexpr int done = N >= 4096 ? 4096 : (N >= 512 ? 512 : (N >= 64 ? 64 : 8));
   constexpr int rest = N / done;
   if constexpr (rest == 2)
      fft_stage<N, 2, done, INV, LINE_FAST>(buf, lineStride, elemStride, nLines, tid, tw);
   else if constexpr (rest == 4)
      fft_stage<N, 4, done, INV, LINE_FAST>(buf, lineStride, elemStride, nLines, tid, tw);
}

}  // namespace iscg
