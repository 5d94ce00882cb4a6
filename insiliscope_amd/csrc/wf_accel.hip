// wf_accel.hip -- the mean field one focus at a time, behind the reference's WidefieldAccelerator seam
// (Sim/WidefieldRender.h:306-361): the scene sends per focus a WidefieldGpuJob (the dye planes it has not sent before,
// the kernel spectra, and per channel which planes meet which kernels with which blend weights); the host keeps plane
// spectra resident by key across jobs and returns one FOV image per channel. JobReferenceHost
// (cli/widefield_check.cpp:816-902) is the reference's CPU statement of that algorithm.
//
// Per geometry the context holds, in HBM:
//   a pool of plane spectra [slot][NY][S] float2 (key -> slot), grown on demand within the plane budget;
//   the kernel spectra [slot][NY][S] (p -> slot), taken from the job once per geometry (the reference's kernel map is
//   fixed within a geometry: WidefieldScene makes a new geometry serial whenever the FFT shape, R, the levels, the
//   PSF or the rect change, Sim/WidefieldRender.cpp:1245-1262);
//   the twiddles of NX and NY.
// A job:
//   1. new planes: the sparse cells scattered into zeroed [n][ny][nx] planes -> rows R2C -> columns, written
//      straight into their pool slots (wf_passes.h passes A + B);
//   2. wf_mac_kernel, one channel per grid row: S = sum over deps, in order, of A (w0 K0 + w1 K1) in complex FP32,
//      then x the phase ramp px[kx] py[ky] and 1/(NX NY) in FP64, rounded to FP32 -- reads 16-24 B per bin per dep
//      (HBM-bound; the kernel spectra shared by neighbouring deps hit L2);
//   3. columns inverse keeping the FOV rows, rows C2R keeping the FOV columns (passes D + E, raw values).
#include "wf_passes.h"

#include <algorithm>
#include <list>
#include <unordered_map>

namespace iscg {

struct AccelGeom
{
   uint64_t geometry = 0;
   uint32_t NX = 0, NY = 0, S = 0;
   size_t planeElems = 0;  // NY * S
   DevBuf pool, kpool, twX, twY;
   uint32_t capacity = 0, used = 0, kcap = 0, kused = 0;
   std::unordered_map<uint64_t, uint32_t> slots;  // plane key -> pool slot
   struct Kern
   {
      uint32_t slot;
      const void* src;
   };
   std::unordered_map<int32_t, Kern> kernels;  // p -> kernel slot and the spectrum it was taken from
   size_t reserved = 0;                        // bytes a Fits answer promised
   uint64_t lastUse = 0;
   size_t bytes() const
   {
      const size_t have = ((size_t)capacity + kcap) * planeElems * sizeof(float2);
      return std::max(have, reserved);
   }
};

struct WfAccelState
{
   std::list<AccelGeom> geoms;
   uint64_t tick = 0;
   size_t budget = 0;  // 0: half of the device memory
   DevBuf dense, x1, mac, x2, deps, chan, ph, starts;
};

void destroy_wf_accel(iscg_context* h)
{
   delete h->wfAccel;
   h->wfAccel = nullptr;
}

static WfAccelState& accel(iscg_context* h)
{
   if (!h->wfAccel)
      h->wfAccel = new WfAccelState();
   return *h->wfAccel;
}

static size_t budget_of(iscg_context* h, WfAccelState& a)
{
   if (a.budget == 0)
   {
      size_t freeB = 0, totalB = 0;
      if (hipMemGetInfo(&freeB, &totalB) == hipSuccess)
         a.budget = totalB / 2;
      else
         a.budget = (size_t)16 << 30;
   }
   return a.budget;
}

static AccelGeom* find_geom(WfAccelState& a, uint64_t geometry)
{
   for (AccelGeom& g : a.geoms)
      if (g.geometry == geometry)
         return &g;
   return nullptr;
}

// At most kMaxGeoms geometries stay resident (a stage move makes a new one): the least recently used go first.
constexpr size_t kMaxGeoms = 8;

static void evict_beyond(WfAccelState& a, size_t keep)
{
   while (a.geoms.size() > keep)
   {
      auto victim = a.geoms.begin();
      for (auto it = a.geoms.begin(); it != a.geoms.end(); ++it)
         if (it->lastUse < victim->lastUse)
            victim = it;
      a.geoms.erase(victim);
   }
}

static size_t others_bytes(const WfAccelState& a, uint64_t geometry)
{
   size_t b = 0;
   for (const AccelGeom& g : a.geoms)
      if (g.geometry != geometry)
         b += g.bytes();
   return b;
}

// Frees room for `need` more bytes of geometry `keep` by dropping the least recently used other geometries.
static bool make_room(iscg_context* h, WfAccelState& a, uint64_t keep, size_t need)
{
   const size_t budget = budget_of(h, a);
   for (;;)
   {
      size_t total = 0;
      for (const AccelGeom& g : a.geoms)
         total += g.bytes();
      if (total + need <= budget)
         return true;
      auto victim = a.geoms.end();
      for (auto it = a.geoms.begin(); it != a.geoms.end(); ++it)
         if (it->geometry != keep && (victim == a.geoms.end() || it->lastUse < victim->lastUse))
            victim = it;
      if (victim == a.geoms.end())
         return false;
      a.geoms.erase(victim);
   }
}

// Grows a [slots][planeElems] float2 pool to at least `want` slots, keeping the first `keep` slots' contents.
static bool grow_pool(iscg_context* h, WfAccelState& a, AccelGeom& g, DevBuf& pool, uint32_t& cap, uint32_t keep,
                      uint32_t want)
{
   if (want <= cap)
      return true;
   const uint32_t newCap = std::max(want, cap + cap / 2);
   const size_t slotBytes = g.planeElems * sizeof(float2);
   if (!make_room(h, a, g.geometry, (size_t)(newCap - cap) * slotBytes))
      return fail(h, "iscg_wf_images: the plane spectra exceed the plane budget (iscg_wf_set_budget)") == 0;
   DevBuf nb;
   if (!ok(h, nb.reserve((size_t)newCap * slotBytes), "plane pool allocation"))
      return false;
   if (keep && !ok(h, hipMemcpyAsync(nb.p, pool.p, (size_t)keep * slotBytes, hipMemcpyDeviceToDevice, h->stream),
                   "plane pool copy"))
      return false;
   if (!ok(h, hipStreamSynchronize(h->stream), "plane pool copy"))
      return false;
   std::swap(pool.p, nb.p);
   std::swap(pool.bytes, nb.bytes);
   cap = newCap;
   return true;
}

static bool make_twiddles_host(iscg_context* h, DevBuf& b, uint32_t n)
{
   std::vector<float2> t(n);
   for (uint32_t k = 0; k < n; ++k)
   {
      const double ang = -2.0 * 3.14159265358979323846 * (double)k / (double)n;
      t[k] = make_float2((float)std::cos(ang), (float)std::sin(ang));
   }
   return ok(h, b.reserve(n * sizeof(float2)), "twiddle allocation") &&
          ok(h, hipMemcpyAsync(b.p, t.data(), n * sizeof(float2), hipMemcpyHostToDevice, h->stream), "twiddle copy") &&
          ok(h, hipStreamSynchronize(h->stream), "twiddle copy");
}

// The geometry's state for a job (created, or reset when its grid changed).
static AccelGeom* geom_for(iscg_context* h, WfAccelState& a, uint64_t geometry, uint32_t NX, uint32_t NY)
{
   AccelGeom* g = find_geom(a, geometry);
   if (g && (g->NX != NX || g->NY != NY))
   {
      a.geoms.remove_if([&](const AccelGeom& x) { return x.geometry == geometry; });
      g = nullptr;
   }
   if (!g)
   {
      evict_beyond(a, kMaxGeoms - 1);
      a.geoms.emplace_back();
      g = &a.geoms.back();
      g->geometry = geometry;
      g->NX = NX;
      g->NY = NY;
      g->S = spec_stride(NX);
      g->planeElems = (size_t)NY * g->S;
      if (!make_twiddles_host(h, g->twX, NX) || !make_twiddles_host(h, g->twY, NY))
      {
         a.geoms.remove_if([&](const AccelGeom& x) { return x.geometry == geometry; });
         return nullptr;
      }
   }
   g->lastUse = ++a.tick;
   return g;
}

// ---- kernels ------------------------------------------------------------------------------------------------------

// The job's sparse planes into zeroed dense [n][ny][nx] planes (plane j holds entries [starts[j], starts[j+1])).
__global__ void wf_scatter_kernel(const uint32_t* __restrict__ starts, int nPlanes, uint32_t total,
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
      const uint32_t c = cells[i];
      if (c < planeCells)
         dense[(size_t)lo * planeCells + c] = values[i];
   }
}

struct MacDep
{
   uint32_t slot, k0, k1;
   float w0, w1;
   uint32_t two;
};

// One channel per blockIdx.y: out[c] = (sum_d A_d (w0 K0 + w1 K1)) x px[kx] py[ky] x scale over the NY x W bins.
// The FP32 products and sums in the reference's order (cfloat k = w0 K0; k += w1 K1; acc += A k).
__global__ __launch_bounds__(256) void wf_mac_kernel(const float2* __restrict__ pool,
                                                      const float2* __restrict__ kpool, size_t planeElems,
                                                      const MacDep* __restrict__ deps,
                                                      const uint32_t* __restrict__ chanStart, int NY, int S, int W,
                                                      const double2* __restrict__ ph, double scale,
                                                      float2* __restrict__ out)
{
   const int c = blockIdx.y;
   const uint32_t d0 = chanStart[c], d1 = chanStart[c + 1];
   const uint32_t bins = (uint32_t)NY * W;
   const double2* px = ph;
   const double2* py = ph + W;
   for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < bins; b += gridDim.x * blockDim.x)
   {
      const int ky = (int)(b / W), kx = (int)(b - (uint32_t)ky * W);
      const size_t off = (size_t)ky * S + kx;
      float ar = 0.0f, ai = 0.0f;
      for (uint32_t j = d0; j < d1; ++j)
      {
         const MacDep dp = deps[j];
         const float2 A = pool[(size_t)dp.slot * planeElems + off];
         const float2 K0 = kpool[(size_t)dp.k0 * planeElems + off];
         float kr = dp.w0 * K0.x, ki = dp.w0 * K0.y;
         if (dp.two)
         {
            const float2 K1 = kpool[(size_t)dp.k1 * planeElems + off];
            kr = kr + dp.w1 * K1.x;
            ki = ki + dp.w1 * K1.y;
         }
         ar = ar + (A.x * kr - A.y * ki);
         ai = ai + (A.x * ki + A.y * kr);
      }
      const double2 p0 = px[kx], p1 = py[ky];
      const double pr = p0.x * p1.x - p0.y * p1.y, pi = p0.x * p1.y + p0.y * p1.x;
      const double vr = (double)ar * pr - (double)ai * pi, vi = (double)ar * pi + (double)ai * pr;
      out[(size_t)c * planeElems + off] = make_float2((float)(vr * scale), (float)(vi * scale));
   }
}

// WidefieldImages::Render (Sim/WidefieldRender.cpp:547-580) per (frame, camera pixel): v = persistent + sum_j
// (float)a_j bleach_j in FP32 for j below the frame's last non-zero coefficient, max(0, v) summed in FP64 over the
// u x u cells, added (FP32) to the base; then the camera noise when wanted.
template <bool NOISE, bool EMCCD>
__global__ __launch_bounds__(256) void wf_render_kernel(const float* __restrict__ chans, int nBleach, int hasP,
                                                        uint32_t cw, uint32_t ch, int u, const double* __restrict__ a,
                                                        uint32_t nFrames, const float* __restrict__ base,
                                                        float* __restrict__ outPh, uint16_t* __restrict__ outAdu,
                                                        CamParams cp, NoiseQueue q, const float* __restrict__ offset,
                                                        uint32_t seed, const uint32_t* __restrict__ frames,
                                                        uint32_t pixelBase)
{
   const uint32_t W = cw / u, H = ch / u, npx = W * H;
   const size_t n = (size_t)cw * ch;
   const float* P = hasP ? chans : nullptr;
   const float* B = chans + (hasP ? n : 0);
   const uint32_t total = npx * nFrames;
   for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += gridDim.x * blockDim.x)
   {
      const uint32_t k = idx / npx, p = idx - k * npx;
      const uint32_t Y = p / W, X = p - Y * W;
      const double* ak = a + (size_t)k * nBleach;
      int nb = 0;
      for (int j = 0; j < nBleach; ++j)
         if (ak[j] != 0.0)
            nb = j + 1;
      float photon = base ? base[idx] : 0.0f;
      if (P || nb > 0)
      {
         double acc = 0.0;
         for (int sy = 0; sy < u; ++sy)
         {
            const size_t row = (size_t)(Y * u + sy) * cw + (size_t)X * u;
            for (int sx = 0; sx < u; ++sx)
            {
               const size_t i = row + sx;
               float v = P ? P[i] : 0.0f;
               for (int j = 0; j < nb; ++j)
                  v = v + (float)ak[j] * B[(size_t)j * n + i];
               acc += v > 0.0f ? v : 0.0f;
            }
         }
         photon = photon + (float)acc;
      }
      if (outPh)
         outPh[idx] = photon;
      if (NOISE)
         noise_store<EMCCD>(cp, q, outAdu, idx, photon, seed, frames[k], p + pixelBase, offset != nullptr,
                            offset ? offset[p] : 0.0f, false, 0.0f, false, 0.0f);
   }
}

}  // namespace iscg

using namespace iscg;

// The FFT grids the per-focus passes take, as the reference's hosts do (WidefieldGpuJob::kMaxGpuFft = 2048,
// Sim/WidefieldRender.h:336); iscg_wf_fits and iscg_wf_images share this rule.
static bool accel_grid_ok(uint32_t NX, uint32_t NY) { return pow2(NX, 2048) && pow2(NY, 2048); }

extern "C" {

int32_t iscg_wf_set_budget(iscg_context* h, uint64_t bytes)
{
   if (begin(h))
      return -1;
   accel(h).budget = (size_t)bytes;
   return 0;
}

int32_t iscg_wf_has_plane(iscg_context* h, uint64_t geometry, uint64_t key)
{
   if (begin(h))
      return -1;
   const AccelGeom* g = find_geom(accel(h), geometry);
   return g && g->slots.count(key) ? 1 : 0;
}

int32_t iscg_wf_fits(iscg_context* h, uint64_t geometry, uint32_t NX, uint32_t NY, uint64_t planes)
{
   if (begin(h))
      return -1;
   // a grid the passes do not take does not fit: the caller renders that scene on the CPU instead of failing in
   // Images (WidefieldAccelerator::Fits, Sim/WidefieldRender.h:342-361)
   if (!accel_grid_ok(NX, NY))
      return 0;
   WfAccelState& a = accel(h);
   const size_t need = (size_t)planes * NY * spec_stride(NX) * sizeof(float2);
   if (others_bytes(a, geometry) + need > budget_of(h, a))
      return 0;
   AccelGeom* g = find_geom(a, geometry);
   if (!g)
   {
      evict_beyond(a, kMaxGeoms - 1);
      a.geoms.emplace_back();
      g = &a.geoms.back();
      g->geometry = geometry;  // NX = 0: set up by the first job
   }
   g->reserved = std::max(g->reserved, need);
   g->lastUse = ++a.tick;
   return 1;
}

int32_t iscg_wf_images(iscg_context* h, const iscg_wf_job* job, float* out)
{
   if (begin(h))
      return -1;
   if (!job)
      return fail(h, "iscg_wf_images: null job");
   const iscg_wf_job& J = *job;
   if (!accel_grid_ok(J.NX, J.NY))
      return fail(h, "iscg_wf_images: FFT grid must be powers of two in [16, 2048]");
   if (J.nx == 0 || J.ny == 0 || J.nx > J.NX || J.ny > J.NY || J.cw == 0 || J.ch == 0 || J.fovX0 + J.cw > J.NX ||
       J.fovY0 + J.ch > J.NY)
      return fail(h, "iscg_wf_images: dye grid or FOV outside the FFT grid");
   if (J.nKernels < 0 || J.nPlanes < 0 || J.nChannels < 0 || (J.nKernels && (!J.kernelP || !J.kernelSpec)) ||
       (J.nPlanes && (!J.planeKey || !J.planeStart)) || (J.nChannels && (!J.channelStart || !out)))
      return fail(h, "iscg_wf_images: null array");
   WfAccelState& a = accel(h);
   AccelGeom* gp = find_geom(a, J.geometry);
   if (gp && gp->NX == 0)  // a Fits answer's placeholder
   {
      const size_t reserved = gp->reserved;
      a.geoms.remove_if([&](const AccelGeom& x) { return x.geometry == J.geometry; });
      gp = geom_for(h, a, J.geometry, J.NX, J.NY);
      if (gp)
         gp->reserved = reserved;
   }
   else
      gp = geom_for(h, a, J.geometry, J.NX, J.NY);
   if (!gp)
      return -1;
   AccelGeom& g = *gp;
   const uint32_t W = J.NX / 2 + 1, S = g.S;
   const XyGrid grid{J.NX, J.NY, S, g.twX.as<const float2>(), g.twY.as<const float2>()};

   // --- check the deps against what the job carries or what is resident (nothing is changed on failure)
   std::unordered_map<int32_t, int> jobK;  // p -> index into the job's kernels
   for (int i = 0; i < J.nKernels; ++i)
   {
      if (!J.kernelSpec[i])
         return fail(h, "iscg_wf_images: null kernel spectrum");
      jobK[J.kernelP[i]] = i;
   }
   std::unordered_map<uint64_t, int> jobP;
   for (int i = 0; i < J.nPlanes; ++i)
      jobP[J.planeKey[i]] = i;
   const uint32_t nDeps = J.nChannels ? J.channelStart[J.nChannels] : 0;
   if (J.nChannels && (J.channelStart[0] != 0 || (nDeps && !J.deps)))
      return fail(h, "iscg_wf_images: channelStart must start at 0");
   for (int c = 0; c < J.nChannels; ++c)
      if (J.channelStart[c + 1] < J.channelStart[c])
         return fail(h, "iscg_wf_images: channelStart must not decrease");
   for (uint32_t j = 0; j < nDeps; ++j)
   {
      const iscg_wf_dep& d = J.deps[j];
      const bool plane = jobP.count(d.key) || g.slots.count(d.key);
      const bool k0 = jobK.count(d.p0) || g.kernels.count(d.p0);
      const bool k1 = !d.two || jobK.count(d.p1) || g.kernels.count(d.p1);
      if (!plane || !k0 || !k1)
         return fail(h, "iscg_wf_images: job refers to a plane or kernel it does not carry");
   }
   const uint32_t nCells = J.nPlanes ? J.planeStart[J.nPlanes] : 0;
   if (J.nPlanes && J.planeStart[0] != 0)
      return fail(h, "iscg_wf_images: planeStart must start at 0");
   for (int i = 0; i < J.nPlanes; ++i)
      if (J.planeStart[i + 1] < J.planeStart[i])
         return fail(h, "iscg_wf_images: planeStart must not decrease");
   if (nCells && (!J.cells || !J.values))
      return fail(h, "iscg_wf_images: null cells or values");

   // --- kernels: taken once per (geometry, p), again when the job points at another spectrum
   for (int i = 0; i < J.nKernels; ++i)
   {
      const int32_t p = J.kernelP[i];
      auto it = g.kernels.find(p);
      if (it != g.kernels.end() && it->second.src == (const void*)J.kernelSpec[i])
         continue;
      uint32_t slot;
      if (it != g.kernels.end())
         slot = it->second.slot;
      else
      {
         if (!grow_pool(h, a, g, g.kpool, g.kcap, g.kused, g.kused + 1))
            return -1;
         slot = g.kused++;
      }
      float2* dst = g.kpool.as<float2>() + (size_t)slot * g.planeElems;
      if (!ok(h, hipMemcpy2DAsync(dst, S * sizeof(float2), J.kernelSpec[i], W * sizeof(float2), W * sizeof(float2),
                                  J.NY, hipMemcpyDefault, h->stream),
              "kernel spectrum copy"))
         return -1;
      g.kernels[p] = AccelGeom::Kern{slot, (const void*)J.kernelSpec[i]};
   }

   // --- planes: new keys take consecutive slots; a key already resident is transformed again into its slot
   if (J.nPlanes)
   {
      std::vector<uint32_t> slotOf(J.nPlanes);
      uint32_t fresh = 0;
      std::unordered_map<uint64_t, uint32_t> seen;
      for (int i = 0; i < J.nPlanes; ++i)
      {
         auto it = g.slots.find(J.planeKey[i]);
         if (it != g.slots.end())
            slotOf[i] = it->second;
         else if (seen.count(J.planeKey[i]))
            slotOf[i] = seen[J.planeKey[i]];
         else
         {
            slotOf[i] = g.used + fresh++;
            seen[J.planeKey[i]] = slotOf[i];
         }
      }
      if (!grow_pool(h, a, g, g.pool, g.capacity, g.used, g.used + fresh))
         return -1;
      const size_t planeCells = (size_t)J.nx * J.ny;
      const size_t nP = (size_t)J.nPlanes;
      if (!ok(h, a.dense.reserve(nP * planeCells * sizeof(float)), "plane staging") ||
          !ok(h, a.x1.reserve(nP * J.ny * S * sizeof(float2)), "plane staging") ||
          !ok(h, a.starts.reserve((nP + 1) * sizeof(uint32_t)), "plane staging"))
         return -1;
      const uint32_t* dCells = nCells ? (const uint32_t*)stage_in(h, 0, J.cells, nCells * sizeof(uint32_t)) : nullptr;
      const float* dVals = nCells ? (const float*)stage_in(h, 1, J.values, nCells * sizeof(float)) : nullptr;
      if (nCells && (!dCells || !dVals))
         return -1;
      if (!ok(h, hipMemsetAsync(a.dense.p, 0, nP * planeCells * sizeof(float), h->stream), "plane staging") ||
          !ok(h, hipMemcpyAsync(a.starts.p, J.planeStart, (nP + 1) * sizeof(uint32_t), hipMemcpyHostToDevice,
                                h->stream),
              "plane staging"))
         return -1;
      if (nCells)
      {
         ProfScope ps(h, "wf_accel_scatter");
         hipLaunchKernelGGL(wf_scatter_kernel, dim3(grid_for(nCells, 256, 4)), dim3(256), 0, h->stream,
                            a.starts.as<const uint32_t>(), J.nPlanes, nCells, dCells, dVals, (uint32_t)planeCells,
                            a.dense.as<float>());
      }
      bool contiguous = true;
      for (int i = 0; i < J.nPlanes; ++i)
         contiguous &= slotOf[i] == slotOf[0] + (uint32_t)i;
      float2* pool = g.pool.as<float2>();
      if (contiguous)
      {
         if (!forward_xy(h, grid, a.dense.as<const float>(), J.nPlanes, (int)J.ny, (int)J.nx, a.x1.as<float2>(),
                         pool + (size_t)slotOf[0] * g.planeElems, g.planeElems))
            return fail(h, "iscg_wf_images: unsupported FFT size");
      }
      else
      {
         if (!ok(h, a.x2.reserve(nP * g.planeElems * sizeof(float2)), "plane staging"))
            return -1;
         if (!forward_xy(h, grid, a.dense.as<const float>(), J.nPlanes, (int)J.ny, (int)J.nx, a.x1.as<float2>(),
                         a.x2.as<float2>(), g.planeElems))
            return fail(h, "iscg_wf_images: unsupported FFT size");
         for (int i = 0; i < J.nPlanes; ++i)
            if (!ok(h, hipMemcpyAsync(pool + (size_t)slotOf[i] * g.planeElems, a.x2.as<float2>() + i * g.planeElems,
                                      g.planeElems * sizeof(float2), hipMemcpyDeviceToDevice, h->stream),
                    "plane copy"))
               return -1;
      }
      if (!ok(h, hipGetLastError(), "iscg_wf_images: plane transform"))
         return -1;
      for (int i = 0; i < J.nPlanes; ++i)
         g.slots[J.planeKey[i]] = slotOf[i];
      g.used += fresh;
   }
   if (J.nChannels == 0)
      return 0;

   // --- deps, phase ramp, MAC
   std::vector<MacDep> md(nDeps ? nDeps : 1);
   for (uint32_t j = 0; j < nDeps; ++j)
   {
      const iscg_wf_dep& d = J.deps[j];
      md[j].slot = g.slots.at(d.key);
      md[j].k0 = g.kernels.at(d.p0).slot;
      md[j].k1 = d.two ? g.kernels.at(d.p1).slot : md[j].k0;
      md[j].w0 = d.w0;
      md[j].w1 = d.two ? d.w1 : 0.0f;
      md[j].two = d.two ? 1u : 0u;
   }
   // ApplyShiftRamp's phases (Sim/DriftShift.cpp:24-48; JobReferenceHost, cli/widefield_check.cpp:870-879)
   std::vector<double2> ph(W + J.NY);
   const double kPi = 3.14159265358979323846;
   for (uint32_t kx = 0; kx < W; ++kx)
   {
      const double ax = 2.0 * kPi * kx * J.fracX / J.NX;
      ph[kx] = kx == J.NX / 2 ? make_double2(std::cos(ax), 0.0) : make_double2(std::cos(ax), std::sin(ax));
   }
   for (uint32_t ky = 0; ky < J.NY; ++ky)
   {
      const double kys = ky <= J.NY / 2 ? (double)ky : (double)ky - J.NY;
      const double ay = 2.0 * kPi * kys * J.fracY / J.NY;
      ph[W + ky] = ky == J.NY / 2 ? make_double2(std::cos(ay), 0.0) : make_double2(std::cos(ay), std::sin(ay));
   }
   const size_t nC = (size_t)J.nChannels;
   if (!ok(h, a.deps.reserve(md.size() * sizeof(MacDep)), "dep staging") ||
       !ok(h, a.chan.reserve((nC + 1) * sizeof(uint32_t)), "dep staging") ||
       !ok(h, a.ph.reserve(ph.size() * sizeof(double2)), "dep staging") ||
       !ok(h, a.mac.reserve(nC * g.planeElems * sizeof(float2)), "channel spectra") ||
       !ok(h, a.x2.reserve(nC * J.ch * S * sizeof(float2)), "channel spectra"))
      return -1;
   if (!ok(h, hipMemcpyAsync(a.deps.p, md.data(), md.size() * sizeof(MacDep), hipMemcpyHostToDevice, h->stream),
           "dep copy") ||
       !ok(h, hipMemcpyAsync(a.chan.p, J.channelStart, (nC + 1) * sizeof(uint32_t), hipMemcpyHostToDevice, h->stream),
           "dep copy") ||
       !ok(h, hipMemcpyAsync(a.ph.p, ph.data(), ph.size() * sizeof(double2), hipMemcpyHostToDevice, h->stream),
           "phase copy"))
      return -1;
   float* dOut = (float*)stage_out(h, 2, out, nC * J.ch * J.cw * sizeof(float));
   if (!dOut)
      return -1;
   {
      ProfScope ps(h, "wf_accel_mac");
      const uint32_t bins = J.NY * W;
      const unsigned gx = (unsigned)std::min<uint32_t>((bins + 255) / 256, 1024);
      hipLaunchKernelGGL(wf_mac_kernel, dim3(gx, (unsigned)nC), dim3(256), 0, h->stream, g.pool.as<const float2>(),
                         g.kpool.as<const float2>(), g.planeElems, a.deps.as<const MacDep>(),
                         a.chan.as<const uint32_t>(), (int)J.NY, (int)S, (int)W, a.ph.as<const double2>(),
                         1.0 / ((double)J.NX * J.NY), a.mac.as<float2>());
   }
   C2rEpilogue e{};
   e.raw = 1;
   e.img = dOut;
   e.adu = nullptr;
   e.offset = nullptr;
   e.background = 0.0;
   e.ch = (int)J.ch;
   e.cw = (int)J.cw;
   e.x0 = (int)J.fovX0;
   if (!inverse_xy(h, grid, a.mac.as<const float2>(), (int)nC, a.x2.as<float2>(), (int)J.fovY0, e, "wf_accel_c2r"))
      return fail(h, "iscg_wf_images: unsupported FFT size");
   if (!ok(h, hipGetLastError(), "iscg_wf_images launch"))
      return -1;
   if (!finish_out(h, out, dOut, nC * J.ch * J.cw * sizeof(float)))
      return -1;
   return 0;
}

int32_t iscg_wf_render(iscg_context* h, const float* channels, int32_t nChannels, int32_t hasPersistent, uint32_t cw,
                       uint32_t ch, int32_t upscale, const double* a, uint32_t nFrames, const float* base,
                       const iscg_camera* cam, const float* offset, uint32_t seed, const uint32_t* frames,
                       uint32_t pixelBase, float* outPhotons, uint16_t* outAdu)
{
   if (begin(h))
      return -1;
   const int hasP = hasPersistent ? 1 : 0, u = upscale > 1 ? upscale : 1;
   const int nBleach = nChannels - hasP;
   if (!channels || nChannels < 1 || nBleach < 0 || cw == 0 || ch == 0 || (nBleach > 0 && !a))
      return fail(h, "iscg_wf_render: null argument or no channel");
   if (!outPhotons && !outAdu)
      return fail(h, "iscg_wf_render: no output");
   if (outAdu && (!cam || !frames))
      return fail(h, "iscg_wf_render: the noise needs a camera and frame ids");
   const uint32_t W = cw / (uint32_t)u, H = ch / (uint32_t)u;
   const size_t npx = (size_t)W * H, nOut = npx * nFrames;
   if (nFrames == 0 || npx == 0)
      return 0;
   if (nOut > 0xFFFFFFFFull)
      return fail(h, "iscg_wf_render: at most 2^32 - 1 output pixels per call");
   const float* dCh = (const float*)stage_in(h, 0, channels, (size_t)nChannels * cw * ch * sizeof(float));
   const double* dA = nBleach ? (const double*)stage_in(h, 1, a, (size_t)nFrames * nBleach * sizeof(double)) : nullptr;
   const float* dBase = base ? (const float*)stage_in(h, 2, base, nOut * sizeof(float)) : nullptr;
   const float* dOff = offset ? (const float*)stage_in(h, 3, offset, npx * sizeof(float)) : nullptr;
   const uint32_t* dFr = frames ? (const uint32_t*)stage_in(h, 4, frames, nFrames * sizeof(uint32_t)) : nullptr;
   float* dPh = outPhotons ? (float*)stage_out(h, 5, outPhotons, nOut * sizeof(float)) : nullptr;
   uint16_t* dAdu = outAdu ? (uint16_t*)stage_out(h, 6, outAdu, nOut * sizeof(uint16_t)) : nullptr;
   if (!dCh || (nBleach && !dA) || (base && !dBase) || (offset && !dOff) || (frames && !dFr) ||
       (outPhotons && !dPh) || (outAdu && !dAdu))
      return -1;
   CamParams cp{};
   NoiseQueue q{nullptr, nullptr};
   if (dAdu && (!cam_params_for(h, cam, cp) || !noise_queue(h, nOut, q)))
      return -1;
   const unsigned grid = grid_for(nOut, 256, 16);
   {
      ProfScope ps(h, "wf_render");
      if (!dAdu)
         hipLaunchKernelGGL((wf_render_kernel<false, false>), dim3(grid), dim3(256), 0, h->stream, dCh, nBleach, hasP,
                            cw, ch, u, dA, nFrames, dBase, dPh, dAdu, cp, q, dOff, seed, dFr, pixelBase);
      else if (cp.emccd)
         hipLaunchKernelGGL((wf_render_kernel<true, true>), dim3(grid), dim3(256), 0, h->stream, dCh, nBleach, hasP,
                            cw, ch, u, dA, nFrames, dBase, dPh, dAdu, cp, q, dOff, seed, dFr, pixelBase);
      else
         hipLaunchKernelGGL((wf_render_kernel<true, false>), dim3(grid), dim3(256), 0, h->stream, dCh, nBleach, hasP,
                            cw, ch, u, dA, nFrames, dBase, dPh, dAdu, cp, q, dOff, seed, dFr, pixelBase);
      if (!ok(h, hipGetLastError(), "iscg_wf_render launch"))
         return -1;
   }
   if (dAdu && !launch_noise_fallback(h, q, dAdu, (uint32_t)npx, dFr, 0, dOff, nullptr, nullptr, cp, seed, pixelBase))
      return -1;
   if (outPhotons && !finish_out(h, outPhotons, dPh, nOut * sizeof(float)))
      return -1;
   if (outAdu && !finish_out(h, outAdu, dAdu, nOut * sizeof(uint16_t)))
      return -1;
   return 0;
}

}  // extern "C"
