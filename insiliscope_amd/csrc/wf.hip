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
   // the fused slab pass (NY = 512, NZ = 128, nzD and nF <= 64, ny <= 256): ghat column-major, column strides of x1 and x2
   bool slab = false;
   size_t col1 = 0, col2 = 0;
};

static bool slab_fits(const iscg_wf_desc& d)
{
   return d.NY == 512 && d.NZ == 128 && d.nzD <= 64 && d.nF <= 64 && d.ny <= 256 && d.fovY0 + d.ch <= d.NY;
}

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

// ---- fused passes B + C + D (NY = 512, NZ = 128) -------------------------------------------------------------
//
// One workgroup per column kx of the half spectrum does the whole (y, z) work of that column: its 64 x 512 slab
// never leaves the chip, so the column's input rows are read once and its output rows written once. The
// column-major passes A and E around it make every global access a contiguous segment: in [kx][z][y < ny], Ghat
// [kx][ky][kz], out [kx][f][rows]. 1024 lanes; w = tid >> 6 (wave), lane = tid & 63.
// The y transform is split by the parity h of ky (decimation in frequency): with x' = (x[y] + (-1)^h x[y + 256])
// e^{-2 pi i h y / 512}, Y[2m + h] = FFT256(x')[m], and the inverse is out[y] = e[y mod 256] + e^{2 pi i y / 512}
// o[y mod 256] with e, o the 256-point inverses of the even and odd halves. Both halves stay in registers (slab_body).
// Within a wave the lanes run in lockstep and its LDS operations complete in order, so a line that one wave owns
// needs only a compiler fence (wave_sync) between the read and the write phases of a stage.
// Columns 0 and NX/2 are real in (y, z) (Hermitian-symmetric rows of the x transform), so the rows passes carry them
// as one column u = a + ib (column 0) and one workgroup takes it. After its y transform, U(ky) = A(ky) + iB(ky) with
// A(-ky) = conj A(ky), B likewise, so the pair (ky, -ky) gives A(ky) and B(ky), which it holds on the lines ky and
// -ky (unpack_at): each line is then an ordinary z convolution, with Ghat0(ky) on the line ky and GhatN(ky) on
// the line -ky, and the inverse y transform takes U'(ky) = A' + iB', U'(-ky) = conj A' + i conj B' back to a' + ib'.
// The lines ky = 0 and 256 pair with themselves: A and B are real there, and their z convolution takes U(-kz)
// (zconv_line). This leaves NX/2 workgroups: one per CU at NX = 512.

__device__ __forceinline__ void wave_sync() { __builtin_amdgcn_wave_barrier(); }

// Complex products of the slab pass with fused multiply-adds (a transform within FP32 rounding either way; the
// library's other arithmetic is not contracted)
__device__ __forceinline__ float2 cmulf(float2 a, float2 b)
{
#pragma clang fp contract(fast)
   return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a b + c
__device__ __forceinline__ float2 cmaf(float2 a, float2 b, float2 c)
{
#pragma clang fp contract(fast)
   return make_float2(a.x * b.x - a.y * b.y + c.x, a.x * b.y + a.y * b.x + c.y);
}

// Butterfly j of a radix-R Stockham stage of length N (NS = product of the earlier radices): inputs at j + r N / R,
// twiddled by tw[TWS r k N / (NS R)] (k = j mod NS; tw a table of length TWS N), outputs to
// (j / NS) NS R + k + r NS.
template <int N, int R, int NS, bool INV, int TWS = 1>
__device__ __forceinline__ void bfly(float2 (&v)[R], int j, const float2* tw)
{
   if (NS > 1)
   {
      const int k = j % NS;
#pragma unroll
      for (int r = 1; r < R; ++r)
      {
         float2 t = tw[TWS * r * k * (N / (NS * R))];
         if (INV)
            t.y = -t.y;
         v[r] = cmulf(v[r], t);
      }
   }
   dft<R, INV>(v);
}
// Element i of a line sits at slot sp(i) = i + i / 16: one pad slot per 16 elements spreads the strided stores of the
// first stages (lane stride 8) and the column accesses of the transposes over the LDS banks (at most 2-way).
__device__ __forceinline__ int sp(int i) { return i + (i >> 4); }
template <int N, int R>
__device__ __forceinline__ void bfly_load(const float2* L, int j, float2 (&v)[R])
{
#pragma unroll
   for (int r = 0; r < R; ++r)
      v[r] = L[sp(j + r * (N / R))];
}
template <int N, int R, int NS>
__device__ __forceinline__ void bfly_store(float2* L, int j, const float2 (&v)[R])
{
   const int base = (j / NS) * NS * R + j % NS;
#pragma unroll
   for (int r = 0; r < R; ++r)
      L[sp(base + r * NS)] = v[r];
}

// 256-point line owned by 32 lanes (q = 0..31; radix 8, 8, 4), twiddles from the 512-point table. In and out,
// v[r] = element q + 32 r; FIRST_IN_REGS: the first stage takes v as its inputs (else reads L), LAST_IN_REGS: the
// result stays in v (else in natural order in L).
template <bool INV, bool FIRST_IN_REGS, bool LAST_IN_REGS>
__device__ __forceinline__ void line_fft256(float2* L, int q, float2 (&v)[8], const float2* tw512)
{
   if (!FIRST_IN_REGS)
      bfly_load<256, 8>(L, q, v);
   bfly<256, 8, 1, INV, 2>(v, q, tw512);
   wave_sync();
   bfly_store<256, 8, 1>(L, q, v);
   wave_sync();
   bfly_load<256, 8>(L, q, v);
   wave_sync();
   bfly<256, 8, 8, INV, 2>(v, q, tw512);
   bfly_store<256, 8, 8>(L, q, v);
   wave_sync();
   float2 u[2][4];
#pragma unroll
   for (int s = 0; s < 2; ++s)
   {
      bfly_load<256, 4>(L, q + 32 * s, u[s]);
      bfly<256, 4, 64, INV, 2>(u[s], q + 32 * s, tw512);
   }
   if (LAST_IN_REGS)
   {
#pragma unroll
      for (int s = 0; s < 2; ++s)
#pragma unroll
         for (int r = 0; r < 4; ++r)
            v[s + 2 * r] = u[s][r];
   }
   else
   {
      wave_sync();
#pragma unroll
      for (int s = 0; s < 2; ++s)
#pragma unroll
         for (int r = 0; r < 4; ++r)
            L[sp(q + 32 * s + 64 * r)] = u[s][r];
   }
}

// 128-point line owned by 16 lanes (q = 0..15; radix 8, 8, 2): v[r] = element q + 16 r in and out (the first stage
// from v; HALF_OUT: only the first 64 outputs, v[0..3]). L: the line's LDS scratch.
template <bool INV, bool HALF_OUT>
__device__ __forceinline__ void line_fft128(float2* L, int q, float2 (&v)[8], const float2* tw)
{
   bfly<128, 8, 1, INV>(v, q, tw);
   wave_sync();
   bfly_store<128, 8, 1>(L, q, v);
   wave_sync();
   bfly_load<128, 8>(L, q, v);
   wave_sync();
   bfly<128, 8, 8, INV>(v, q, tw);
   bfly_store<128, 8, 8>(L, q, v);
   wave_sync();
#pragma unroll
   for (int s = 0; s < 4; ++s)
   {
      const int j = q + 16 * s;
      float2 u[2];
      bfly_load<128, 2>(L, j, u);
      if (HALF_OUT)
      {
         // element j (< 64) of the radix-2 output only: u0 + tw u1
         float2 t = tw[j];
         if (INV)
            t.y = -t.y;
         v[s] = cmaf(u[1], t, u[0]);
      }
      else
      {
         bfly<128, 2, 64, INV>(u, j, tw);
         v[s] = u[0];
         v[s + 4] = u[1];
      }
   }
}

// A spectrum line's slot in LDS: 128 values and 8 more, so the 4 lines a wave holds start on different banks.
constexpr int kGLine = 136;


// One line of the z convolution (16 lanes, q = lane & 15): its 64 planes x[i] at z = q + 16 i in, FFT along z (the
// upper 64 planes zero), times the kernel spectrum gl[kz] (the line's 128 values, staged in LDS), inverse, the first
// 64 planes back to x. after_g runs once the line's spectrum has been read from gl. Z: the line's LDS scratch,
// gs: the self-paired lines' spectrum slot (packed workgroup only).
template <bool PACKED, typename F>
__device__ __forceinline__ void zconv_line(float2 (&x)[4], int ky, const float2* gl, F&& after_g, float2* Z,
                                           float2* gs, int lane, const float2* twz)
{
   constexpr int NY = 512, NZ = 128, NI = 4;
   const int q = lane & 15;
   float2 v[8];
#pragma unroll
   for (int r = 0; r < 8; ++r)
      v[r] = r < NI ? x[r < NI ? r : 0] : make_float2(0.0f, 0.0f);
   line_fft128<false, false>(Z, q, v, twz);
   // v[m] = U at kz = q + 16 m; the spectrum's loads into gl have landed once the wave's memory count is zero
   asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
   float2 g[8];
#pragma unroll
   for (int m = 0; m < 8; ++m)
      g[m] = gl[q + 16 * m];
   // the reads have returned before the next chunk's spectrum may land in the same slots
   asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
   after_g();
   if constexpr (PACKED)
   {
      // The packed workgroup's lines ky = 0 and 256 hold U = A + iB of the two real columns (A, B real in z, so
      // Hermitian in kz), which its unpacking leaves as they are: A Ghat0 + i B GhatN = U(k) P(k) + conj(U(-k)) Q(k)
      // with P = (Ghat0 + GhatN) / 2, Q = (Ghat0 - GhatN) / 2, both Hermitian in kz; U(-k) comes through LDS, and the
      // line's spectrum slots hold P(kz) for 0 < kz < 64 and Q(kz) at 128 - kz, the real values at kz = 0 and 64
      // paired in slots 0 (P) and 64 (Q).
      const bool self = ky == 0 || ky == NY / 2;
      if (self)
      {
#pragma unroll
         for (int m = 0; m < 8; ++m)
         {
            Z[sp(q + 16 * m)] = v[m];
            gs[q + 16 * m] = g[m];
         }
      }
      wave_sync();
#pragma unroll
      for (int m = 0; m < 8; ++m)
      {
         float2 y = cmulf(v[m], g[m]);
         if (self)
         {
            const int kz = q + 16 * m, nkz = (NZ - kz) & (NZ - 1);
            const float2 up = Z[sp(nkz)];
            const bool low = m < 4, real = (m == 0 || m == 4) && q == 0;
            float2 P2 = gs[real ? 0 : (low ? kz : nkz)], Q2 = gs[real ? NZ / 2 : (low ? nkz : kz)];
            if (!low && !real)
            {
               P2.y = -P2.y;
               Q2.y = -Q2.y;
            }
            if (real)
            {
               P2 = make_float2(m == 0 ? P2.x : P2.y, 0.0f);
               Q2 = make_float2(m == 0 ? Q2.x : Q2.y, 0.0f);
            }
            y = cmaf(make_float2(up.x, -up.y), Q2, cmulf(v[m], P2));
         }
         v[m] = y;
      }
   }
   else
   {
#pragma unroll
      for (int m = 0; m < 8; ++m)
         v[m] = cmulf(v[m], g[m]);
   }
   wave_sync();
   line_fft128<true, true>(Z, q, v, twz);
#pragma unroll
   for (int i = 0; i < NI; ++i)
      x[i] = v[i];
   wave_sync();
}

// The packed workgroup's columns a (kx = 0) and b (kx = NX/2), real in (y, z), travel as U = A + iB after the y
// transform. Element m of half line L (parity h) pairs with mp = 256 - m (h = 0; m = 0 and 128, i.e. ky = 0 and
// 256, pair with themselves) or 255 - m (h = 1); m < 128 is the line +ky, the other the line -ky.
// unpack_at: A(ky) = (U(ky) + conj U(-ky)) / 2 on the line +ky, B(ky) = (U(ky) - conj U(-ky)) / 2i on the line -ky.
__device__ __forceinline__ float2 unpack_at(const float2* L, int h, int m)
{
   const bool self = h == 0 && (m & 127) == 0;
   const float2 o = L[sp(m)], p = L[sp(self ? m : (h ? 255 : 256) - m)];
   if (self)
      return o;
   return m < 128 ? make_float2(0.5f * (o.x + p.x), 0.5f * (o.y - p.y))
                  : make_float2(0.5f * (p.y + o.y), 0.5f * (o.x - p.x));
}
// repack_at: the inverse, U(ky) = A + iB on the line +ky, U(-ky) = conj A + i conj B on the line -ky.
__device__ __forceinline__ float2 repack_at(const float2* L, int h, int m)
{
   const bool self = h == 0 && (m & 127) == 0;
   const float2 o = L[sp(m)], p = L[sp(self ? m : (h ? 255 : 256) - m)];
   if (self)
      return o;
   return m < 128 ? make_float2(o.x - p.y, o.y + p.x) : make_float2(p.x + o.y, o.x - p.y);
}

// The whole slab at once (both parities): the 64 x 512 slab of a column stays in registers, R[j][i] = S_h[z =
// (lane & 15) + 16 i][m] for line g = 64 j + 4 w + (lane >> 4) of the 512 (parity h = g / 256, half-line index
// m = g % 256, ky = 2 m + h). A wave transforms both halves of a z line (lanes 0..31 the even half, 32..63 the odd,
// twisted one) from a single read of the line, and both halves of a focus, which it combines into the output rows:
// the input column is read once and every output row written once.
//   1. 4 passes of 16 z lines along y (one per wave), each into its own LDS buffer (two, alternating) followed by a
//      workgroup barrier and the transpose into R[.][p]: one barrier per pass;
//   2. 8 chunks of 64 lines along z (zconv_line), 4 lines per wave, the next chunk's kernel spectrum in flight into
//      LDS (gst);
//   3. 4 passes of 16 foci: transpose R[.][p] into the pass's buffer, barrier, one focus per wave back along y (both
//      halves), the rows [row0, row0 + rows) = e[y mod 256] + e^{2 pi i y / 512} o[y mod 256] stored once.
// LDS: U holds the two y buffers (32 half lines each) in stages 1 and 3, and gst (first 64 KB) and the z lines'
// scratch (after it) in stage 2; gst overlaps only buffer 0, so chunk 0's spectrum starts loading once pass 2's
// transposes are done.
constexpr int kSlabLSY = 273, kSlabLSZ = 137;          // odd strides >= sp(256), sp(128)
constexpr int kSlabYB = 32 * kSlabLSY;                 // one y buffer: 16 lines x 2 halves
constexpr int kSlabGst = 64 * 128;                     // one chunk's kernel spectrum: 64 lines of 1 KB
constexpr int kSlabU = 2 * kSlabYB;
static_assert(kSlabGst + 64 * kSlabLSZ <= kSlabU, "stage 2's LDS fits the y buffers");
static_assert(kSlabGst <= kSlabYB, "gst overlaps only the first y buffer");

template <bool PACKED>
__device__ __forceinline__ void slab_body(const float2* __restrict__ in, int ny, int nzIn, size_t inCol,
                                          float2* __restrict__ out, int row0, int rows, int nOut, size_t outCol,
                                          const float2* __restrict__ ghat, float2* U, float2* gx, const float2* twy,
                                          const float2* twz)
{
   constexpr int NY = 512, NZ = 128;
   constexpr int NJ = 8, NI = 4;  // chunks of 64 lines, passes of 16 planes
   constexpr int LSY = kSlabLSY, LSZ = kSlabLSZ;
   constexpr bool packed = PACKED;
   const int tid0 = threadIdx.x;
   const int c = packed ? 0 : (int)blockIdx.x;
   const float2* src = in + (size_t)c * inCol;
   const float2* G = ghat + (size_t)c * NY * NZ;
   float2* dst = out + (size_t)c * outCol;
   float2* gst = U;                // stage 2: the chunk's kernel spectrum
   float2* zscr = U + kSlabGst;    // stage 2: the z lines' scratch
// the lane ids, opaque to the compiler at each pass and chunk so that it recomputes addresses there instead of
// keeping every unrolled iteration's addresses live
#define SLAB_IDS                   \
   int tid = tid0;                 \
   asm volatile("" : "+v"(tid));   \
   const int lane = tid & 63, w = tid >> 6
   // chunk j's line at position s: line g = 64 j + s of the slab, parity g / 256, half-line index g % 256
   auto mline = [](int j, int s) { return (64 * j + s) & 255; };
   auto kyof = [&](int j, int s) { return 2 * mline(j, s) + (j >> 2); };
   // the line zl of pass p: z = 16 p + zl, element y = (lane & 31) + 32 r (ny <= 256 on this path); every lane
   // loads (element 0 where it is outside the rows) and keeps the value or zero, with no branch
   auto load_line = [&](int p, int zl, int lane, float2 (&xv)[8]) {
      const int z = 16 * p + zl, q = lane & 31;
#pragma unroll
      for (int r = 0; r < 8; ++r)
      {
         const int y = q + 32 * r;
         const bool in = z < nzIn && y < ny;
         const float2 x = src[in ? (uint32_t)(z * ny + y) : 0u];
         xv[r] = in ? x : make_float2(0.0f, 0.0f);
      }
   };
   // line (zl, h) of pass p's buffer: zl + 16 h
   auto yline = [&](int p, int zl, int h) { return U + (p & 1) * kSlabYB + (zl + 16 * h) * LSY; };
   // the kernel spectrum of chunk j's lines 4 w .. 4 w + 3 (1 KB each, contiguous in HBM) straight into the wave's
   // slots of gst, one 16-byte piece per lane and line, no registers held; read after s_waitcnt vmcnt(0)
   auto stage_g = [&](int j, int w, int lane) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
      {
         const float4* s4 = reinterpret_cast<const float4*>(G) + (uint32_t)(kyof(j, 4 * w + i) * (NZ / 2) + lane);
         __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)s4,
                                          (__attribute__((address_space(3))) void*)(gst + (4 * w + i) * NZ), 16, 0,
                                          0);
      }
   };

   float2 R[NJ][NI];
   // 1. y forward; the next pass's line is loaded before this pass's transform
   float2 xn[8];
   {
      SLAB_IDS;
      load_line(0, w, lane, xn);
   }
#pragma unroll
   for (int p = 0; p < NI; ++p)
   {
      SLAB_IDS;
      float2 v[8];
#pragma unroll
      for (int r = 0; r < 8; ++r)
         v[r] = xn[r];
      if (p + 1 < NI)
         load_line(p + 1, w, lane, xn);
      if (16 * p < nzIn)
      {
         const int h = lane >> 5, q = lane & 31;
#pragma unroll
         for (int r = 0; r < 8; ++r)
            v[r] = cmulf(v[r], h ? twy[q + 32 * r] : make_float2(1.0f, 0.0f));
         line_fft256<false, true, false>(yline(p, w, h), q, v, twy);
      }
      __syncthreads();
      // buffer 0 is free once every wave has passed this barrier of pass 3 (its last readers were pass 2's
      // transposes): chunk 0's spectrum goes into gst, inside it
      if (p == NI - 1)
         stage_g(0, w, lane);
      const int s = 4 * w + (lane >> 4), zl = lane & 15;
      if (16 * p >= nzIn)
      {
#pragma unroll
         for (int j = 0; j < NJ; ++j)
            R[j][p] = make_float2(0.0f, 0.0f);
      }
      else if constexpr (packed)
      {
#pragma unroll
         for (int j = 0; j < NJ; ++j)
            R[j][p] = unpack_at(yline(p, zl, j >> 2), j >> 2, mline(j, s));
      }
      else
      {
         // element m = 64 (j % 4) + s of line (zl, j / 4) sits at sp(m) = 68 (j % 4) + sp(s): one address per pass
         const float2* L = yline(p, zl, 0) + sp(s);
#pragma unroll
         for (int j = 0; j < NJ; ++j)
            R[j][p] = L[(j >> 2) * 16 * LSY + 68 * (j & 3)];
      }
   }
   // the z scratch overlaps buffer 1, which pass 3's transposes read
   __syncthreads();

   // 2. z convolution, lines 4 w .. 4 w + 3 of each chunk per wave; the chunk's kernel spectrum is in gst
#pragma unroll
   for (int j = 0; j < NJ; ++j)
   {
      SLAB_IDS;
      const int pos = 4 * w + (lane >> 4);
      zconv_line<packed>(
          R[j], kyof(j, pos), gst + pos * NZ,
          [&] {
             if (j + 1 < NJ)
                stage_g(j + 1, w, lane);
          },
          zscr + pos * LSZ, gx + pos * kGLine, lane, twz);
      // the chunk's values are formed here: without this the compiler sinks their last butterfly into the
      // (conditional) passes of stage 3 and keeps its operands, three times the registers, live until then
#pragma unroll
      for (int i = 0; i < NI; ++i)
         asm volatile("" : "+v"(R[j][i].x), "+v"(R[j][i].y));
   }
   __syncthreads();

   // 3. y inverse, focus 16 p + w per wave; rows [row0, row0 + rows)
#pragma unroll
   for (int p = 0; p < NI; ++p)
   {
      SLAB_IDS;
      if (16 * p >= nOut)
         break;
      {
         const int s = 4 * w + (lane >> 4), fl = lane & 15;
         float2* L = yline(p, fl, 0) + sp(s);
#pragma unroll
         for (int j = 0; j < NJ; ++j)
            L[(j >> 2) * 16 * LSY + 68 * (j & 3)] = R[j][p];
      }
      __syncthreads();
      const int h = lane >> 5, q = lane & 31, f = 16 * p + w;
      float2 v[8];
      if constexpr (packed)
      {
#pragma unroll
         for (int r = 0; r < 8; ++r)
            v[r] = repack_at(yline(p, w, h), h, q + 32 * r);
         line_fft256<true, true, false>(yline(p, w, h), q, v, twy);
      }
      else
         line_fft256<true, false, false>(yline(p, w, h), q, v, twy);
      wave_sync();
      if (f < nOut)
      {
         const float2* E = yline(p, w, 0);
         const float2* O = yline(p, w, 1);
#pragma unroll 4
         for (int yo = lane; yo < rows; yo += 64)
         {
            const int y = row0 + yo, m = y & 255;
            const float2 tw = twy[y];
            dst[(uint32_t)(f * rows + yo)] = cmaf(O[sp(m)], make_float2(tw.x, -tw.y), E[sp(m)]);
         }
      }
   }
#undef SLAB_IDS
}

template <int NY, int NZ>
__global__ __launch_bounds__(1024) void slab_kernel(const float2* __restrict__ in, int ny, int nzIn, size_t inCol,
                                                    float2* __restrict__ out, int row0, int rows, int nOut,
                                                    size_t outCol, const float2* __restrict__ ghat,
                                                    const float2* __restrict__ twY, const float2* __restrict__ twZ)
{
   static_assert(NY == 512 && NZ == 128, "the slab pass is written for a 512 x 128 (y, z) slab");
   __shared__ __attribute__((aligned(16))) float2 U[kSlabU];
   __shared__ float2 gx[4 * kGLine];  // the self-paired lines' spectrum values (position 0 of the packed workgroup)
   __shared__ float2 twy[NY];
   __shared__ float2 twz[NZ];
   for (int i = threadIdx.x; i < NY; i += 1024)
      twy[i] = twY[i];
   for (int i = threadIdx.x; i < NZ; i += 1024)
      twz[i] = twZ[i];
   __syncthreads();
   if (blockIdx.x == 0)
      slab_body<true>(in, ny, nzIn, inCol, out, row0, rows, nOut, outCol, ghat, U, gx, twy, twz);
   else
      slab_body<false>(in, ny, nzIn, inCol, out, row0, rows, nOut, outCol, ghat, U, gx, twy, twz);
}

// The kernel spectrum for the slab pass: out[c][ky][kz] = ghat[kz][ky][c] for columns 1 .. NH - 1; column 0 holds
// Ghat0(kz, ky) on the lines 0 < ky < NY/2 and GhatNH(kz, NY - ky) on the lines NY/2 < ky < NY (the packed
// workgroup's A and B lines), and on the self-paired lines ky = 0 and NY/2 P = (Ghat0 + GhatNH) / 2 and
// Q = (Ghat0 - GhatNH) / 2, one value per kz: P(kz) for 0 < kz < NZ/2, Q(NZ - kz) at kz > NZ/2, the real values at
// kz = 0 and NZ/2 paired in slots 0 (P) and NZ/2 (Q).
__global__ void ghat_columns_kernel(const float2* __restrict__ ghat, int NZ, int NY, int S, int NH,
                                    float2* __restrict__ out)
{
   const size_t plane = (size_t)NZ * NY, total = (size_t)NH * plane;
   for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x)
   {
      const int c = (int)(idx / plane);
      const int rem = (int)(idx - (size_t)c * plane), ky = rem / NZ, kz = rem % NZ;
      if (c != 0)
      {
         out[idx] = ghat[((size_t)kz * NY + ky) * S + c];
         continue;
      }
      if (ky != 0 && ky != NY / 2)
      {
         out[idx] = ky < NY / 2 ? ghat[((size_t)kz * NY + ky) * S] : ghat[((size_t)kz * NY + (NY - ky)) * S + NH];
         continue;
      }
      // P(kz) and Q(kz) of the self-paired line ky
      auto P = [&](int z) {
         const float2* s = ghat + ((size_t)z * NY + ky) * S;
         return make_float2(0.5f * (s[0].x + s[NH].x), 0.5f * (s[0].y + s[NH].y));
      };
      auto Q = [&](int z) {
         const float2* s = ghat + ((size_t)z * NY + ky) * S;
         return make_float2(0.5f * (s[0].x - s[NH].x), 0.5f * (s[0].y - s[NH].y));
      };
      float2 v;
      if (kz == 0)
         v = make_float2(P(0).x, P(NZ / 2).x);
      else if (kz == NZ / 2)
         v = make_float2(Q(0).x, Q(NZ / 2).x);
      else if (kz < NZ / 2)
         v = P(kz);
      else
         v = Q(NZ - kz);
      out[idx] = v;
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
   s.slab = slab_fits(d);
   if (s.slab)
   {
      // column-major copy for the slab pass (then the row-major one is dropped)
      const uint32_t NH = d.NX / 2;
      DevBuf gc;
      const size_t n = (size_t)NH * d.NZ * d.NY;
      if (!ok(h, gc.reserve(n * sizeof(float2)), "spectrum allocation"))
         return -1;
      hipLaunchKernelGGL(ghat_columns_kernel, dim3(grid_for(n, 256, 8)), dim3(256), 0, h->stream,
                         s.ghat.as<const float2>(), (int)d.NZ, (int)d.NY, (int)S, (int)NH, gc.as<float2>());
      if (!ok(h, hipGetLastError(), "plan transpose") || !ok(h, hipStreamSynchronize(h->stream), "plan synchronize"))
         return -1;
      std::swap(s.ghat.p, gc.p);
      std::swap(s.ghat.bytes, gc.bytes);
      s.col1 = ((size_t)d.nzD * d.ny + 31) / 32 * 32;
      s.col2 = ((size_t)d.nF * d.ch + 31) / 32 * 32;
   }
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
   if (s.slab)
   {
      // A: [nzD][ny][nx] -> x1 [kx][nzD * ny]; B + C + D: -> x2 [kx][nF * ch]; E -> images / ADU
      {
         ProfScope ps(h, "wf_rows_r2c");
         if (!dispatch_n<RowsR2Ccm, 1024>((int)d.NX, h->stream, dIn, (size_t)d.nx, (int)(d.nzD * d.ny), (int)d.nx,
                                          x1, s.col1, s.twX.as<const float2>()))
            return fail(h, "iscg_wf_conv3d: unsupported FFT size");
      }
      {
         ProfScope ps(h, "wf_slab");
         hipLaunchKernelGGL((slab_kernel<512, 128>), dim3(d.NX / 2), dim3(1024), 0, h->stream,
                            (const float2*)x1, (int)d.ny, (int)d.nzD, s.col1, x2, (int)d.fovY0, (int)d.ch, (int)d.nF,
                            s.col2, s.ghat.as<const float2>(), s.twY.as<const float2>(), s.twZ.as<const float2>());
      }
      {
         ProfScope ps(h, "wf_rows_c2r_noise");
         dispatch_n<RowsC2Rcm, 1024>((int)d.NX, h->stream, (const float2*)x2, s.col2, (int)(d.nF * d.ch), e,
                                     s.twX.as<const float2>());
      }
   }
   else
   {
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
   inverse_xy(h, grid_of(s), x1, (int)d.nF, x2, (int)d.fovY0, e, "wf_rows_c2r_noise");
   }
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
