// rng.hip -- the core's counter hash on the device (core/src/rng.h:18-52): Pcg4d of arbitrary addresses, HashUnit
// and HashStream draws, in batches. The cell field and the dye sites are addressed by (seed, cell x, cell y, k)
// through these; the camera noise uses the same Pcg4d (CounterRng). One lane per address, integer work only.
#include "iscg_internal.h"

namespace iscg {

// words: 4 per address (a, b, c, d); units: HashUnit of the first word, or null
__global__ void pcg4d_kernel(const uint32_t* __restrict__ a, const uint32_t* __restrict__ b,
                             const uint32_t* __restrict__ c, const uint32_t* __restrict__ d, uint32_t n,
                             uint32_t* __restrict__ words)
{
   for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
   {
      uint32_t x = a[i], y = b[i], z = c[i], w = d[i];
      pcg4d(x, y, z, w);
      words[4 * (size_t)i] = x;
      words[4 * (size_t)i + 1] = y;
      words[4 * (size_t)i + 2] = z;
      words[4 * (size_t)i + 3] = w;
   }
}

// HashUnit(seed, cx, cy, k): ((Pcg4d(seed, cx, cy, k).a >> 9) + 0.5) 2^-23. Stream form (cx, cy scalars, k from
// base * 4096 + i with uint32 wrap) when cxs is null.
__global__ void hash_units_kernel(uint32_t seed, const int32_t* __restrict__ cxs, const int32_t* __restrict__ cys,
                                  const uint32_t* __restrict__ ks, int32_t cx0, int32_t cy0, uint32_t base, uint32_t n,
                                  double* __restrict__ out)
{
   for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
   {
      uint32_t x = seed, y = (uint32_t)(cxs ? cxs[i] : cx0), z = (uint32_t)(cys ? cys[i] : cy0);
      uint32_t w = cxs ? ks[i] : base * 4096u + i;
      pcg4d(x, y, z, w);
      out[i] = CounterRng::unit_of(x);
   }
}

}  // namespace iscg

using namespace iscg;

extern "C" {

int32_t iscg_pcg4d(iscg_context* h, const uint32_t* a, const uint32_t* b, const uint32_t* c, const uint32_t* d,
                   uint32_t n, uint32_t* words)
{
   if (begin(h))
      return -1;
   if (n == 0)
      return 0;
   if (!a || !b || !c || !d || !words)
      return fail(h, "iscg_pcg4d: null argument");
   const size_t bytes = (size_t)n * sizeof(uint32_t);
   const uint32_t* da = (const uint32_t*)stage_in(h, 0, a, bytes);
   const uint32_t* db = (const uint32_t*)stage_in(h, 1, b, bytes);
   const uint32_t* dc = (const uint32_t*)stage_in(h, 2, c, bytes);
   const uint32_t* dd = (const uint32_t*)stage_in(h, 3, d, bytes);
   uint32_t* dw = (uint32_t*)stage_out(h, 4, words, 4 * bytes);
   if (!da || !db || !dc || !dd || !dw)
      return -1;
   hipLaunchKernelGGL(pcg4d_kernel, dim3(grid_for(n, 256, 8)), dim3(256), 0, h->stream, da, db, dc, dd, n, dw);
   if (!ok(h, hipGetLastError(), "iscg_pcg4d launch"))
      return -1;
   return finish_out(h, words, dw, 4 * bytes) ? 0 : -1;
}

int32_t iscg_hash_units(iscg_context* h, uint32_t seed, const int32_t* cx, const int32_t* cy, const uint32_t* k,
                        uint32_t n, double* out)
{
   if (begin(h))
      return -1;
   if (n == 0)
      return 0;
   if (!cx || !cy || !k || !out)
      return fail(h, "iscg_hash_units: null argument");
   const int32_t* dx = (const int32_t*)stage_in(h, 0, cx, (size_t)n * sizeof(int32_t));
   const int32_t* dy = (const int32_t*)stage_in(h, 1, cy, (size_t)n * sizeof(int32_t));
   const uint32_t* dk = (const uint32_t*)stage_in(h, 2, k, (size_t)n * sizeof(uint32_t));
   double* dout = (double*)stage_out(h, 3, out, (size_t)n * sizeof(double));
   if (!dx || !dy || !dk || !dout)
      return -1;
   hipLaunchKernelGGL(hash_units_kernel, dim3(grid_for(n, 256, 8)), dim3(256), 0, h->stream, seed, dx, dy, dk, 0, 0,
                      0u, n, dout);
   if (!ok(h, hipGetLastError(), "iscg_hash_units launch"))
      return -1;
   return finish_out(h, out, dout, (size_t)n * sizeof(double)) ? 0 : -1;
}

int32_t iscg_hash_stream(iscg_context* h, uint32_t seed, int32_t cx, int32_t cy, uint32_t base, uint32_t n,
                         double* out)
{
   if (begin(h))
      return -1;
   if (n == 0)
      return 0;
   if (!out)
      return fail(h, "iscg_hash_stream: null argument");
   double* dout = (double*)stage_out(h, 0, out, (size_t)n * sizeof(double));
   if (!dout)
      return -1;
   hipLaunchKernelGGL(hash_units_kernel, dim3(grid_for(n, 256, 8)), dim3(256), 0, h->stream, seed,
                      (const int32_t*)nullptr, (const int32_t*)nullptr, (const uint32_t*)nullptr, cx, cy, base, n,
                      dout);
   if (!ok(h, hipGetLastError(), "iscg_hash_stream launch"))
      return -1;
   return finish_out(h, out, dout, (size_t)n * sizeof(double)) ? 0 : -1;
}

}  // extern "C"
