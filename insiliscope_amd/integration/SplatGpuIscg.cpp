// SplatGpuIscg.cpp -- sim::SplatGpuHost (Sim/SplatGpu.h:40-77) on libiscg: the blink splat + camera noise of the
// engine's movies (FluorescenceMovie's GPU path, Sim/ScopeMovie.cpp:7059-7105 and 7633-7653) on MI355X.
//
// A link-time drop-in: built into the engine in place of Sim/SplatGpu.cpp, against the unmodified Sim/SplatGpu.h.
// The class keeps its declared members (backend_, s_) and its method set; its pimpl State holds an iscg context
// instead of the WGSL buffers. The backend argument is not used (pass nullptr): the work runs on HIP device 0 through
// include/iscg.h. Install it as the process-wide host the engine asks for:
//
//    sim::SetDefaultBlinkSplatGpu([](std::string& info) { return sim::MakeSplatGpuHost(nullptr, info); });
//
// Frames are the reference's records (CollectGpuEmitters: float weights) splatted and noised as
// iscg_splat_render_records does: within float rounding of RenderPhotonImage + ApplyNoiseChain, as the reference's
// own GPU hosts are.
#include "SplatGpu.h"

#include "SMLMSimulation.h"

#include "../../include/iscg.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <random>

namespace sim {

struct SplatGpuHost::State
{
   iscg_context* h = nullptr;
   std::string createErr;
   std::mutex mu;  // one call at a time on the context (its stream is not re-entrant)
   unsigned width = 0, height = 0;
   // the kernels' planes flattened (nz x size x size), kept per SplatSerial so a repeat uploads nothing
   struct Flat
   {
      uint64_t serial = 0;
      std::vector<float> planes;
      std::vector<int16_t> lo, hi;
   };
   std::vector<Flat> flat;
   std::vector<iscg_splat_emitter> ems;
   std::vector<uint32_t> start, frames;
   std::vector<float> extras;
   std::vector<uint16_t> adu;
};

namespace {

iscg_camera ToCamera(const CameraNoiseParams& c)
{
   iscg_camera o;
   o.quantumEfficiency = c.quantumEfficiency;
   o.darkCurrentElectrons = c.darkCurrentElectrons;
   o.gainPhotonsPerAdu = c.gainPhotonsPerAdu;
   o.readNoiseElectrons = c.readNoiseElectrons;
   o.emccd = c.emccd ? 1 : 0;
   o.bitDepth = c.bitDepth;
   o.emGain = c.emGain;
   o.cicElectrons = c.cicElectrons;
   return o;
}

const float* MapOrNull(const std::vector<float>& m, size_t n) { return m.size() == n ? m.data() : nullptr; }

}  // namespace

SplatGpuHost::SplatGpuHost(std::shared_ptr<GpuBackend> backend) : backend_(std::move(backend)), s_(new State)
{
   if (iscg_create(0, &s_->h) != 0)
   {
      s_->createErr = std::string("no MI355X context: ") + iscg_last_error(nullptr);
      s_->h = nullptr;
   }
}

SplatGpuHost::~SplatGpuHost()
{
   if (s_ && s_->h)
      iscg_destroy(s_->h);
}

bool SplatGpuHost::SetKernels(const std::vector<const PsfKernelCache*>& kernels, std::string& err)
{
   State& m = *s_;
   std::lock_guard<std::mutex> lock(m.mu);
   if (!m.h)
   {
      err = m.createErr;
      return false;
   }
   if (kernels.size() > kMaxKernels)
   {
      err = "the HIP splat holds at most " + std::to_string(kMaxKernels) + " kernels; " +
            std::to_string(kernels.size()) + " were given";
      return false;
   }
   m.flat.resize(kernels.size());
   std::vector<iscg_kernel_view> views(kernels.size());
   for (size_t i = 0; i < kernels.size(); ++i)
   {
      const PsfKernelCache& k = *kernels[i];
      if (!k.valid || k.Planes().empty() || k.interpMode == PsfInterpMode::Fft)
      {
         err = "the HIP splat takes valid diffraction kernels with Nearest, Linear or Cubic placement";
         return false;
      }
      State::Flat& f = m.flat[i];
      const size_t plane = static_cast<size_t>(k.sizeOversampled) * k.sizeOversampled;
      if (f.serial != k.SplatSerial() || f.planes.size() != plane * k.nz)
      {
         f.serial = k.SplatSerial();
         f.planes.resize(plane * k.nz);
         for (int z = 0; z < k.nz; ++z)
            std::memcpy(f.planes.data() + plane * z, k.Planes()[static_cast<size_t>(z)].data(), plane * sizeof(float));
         f.lo.clear();
         f.hi.clear();
         if (k.halo)
         {
            f.lo = k.halo->lo;
            f.hi = k.halo->hi;
         }
      }
      iscg_kernel_view& v = views[i];
      v.planes = f.planes.data();
      v.spanLo = f.lo.empty() ? nullptr : f.lo.data();
      v.spanHi = f.hi.empty() ? nullptr : f.hi.data();
      v.zStepNm = k.zStepNm;
      v.serial = f.serial;
      v.nz = k.nz;
      v.size = k.sizeOversampled;
      v.oversampling = k.oversampling;
      v.interpMode = static_cast<int32_t>(k.interpMode);
   }
   if (iscg_splat_set_kernels(m.h, views.data(), static_cast<int32_t>(views.size())) != 0)
   {
      err = iscg_last_error(m.h);
      return false;
   }
   return true;
}

bool SplatGpuHost::SetStatic(unsigned width, unsigned height, uint32_t pixelBase, const std::vector<float>& offset,
                             const std::vector<float>& gain, const std::vector<float>& readNoise,
                             const CameraNoiseParams& cam, uint64_t key, std::string& err)
{
   (void)cam;  // the scalar fallbacks travel with each RenderFrames call
   State& m = *s_;
   std::lock_guard<std::mutex> lock(m.mu);
   if (!m.h)
   {
      err = m.createErr;
      return false;
   }
   const size_t n = static_cast<size_t>(width) * height;
   // the blinks go onto each frame's extra image (the rest of its light), so no background of its own
   if (iscg_splat_set_static(m.h, width, height, pixelBase, MapOrNull(offset, n), MapOrNull(gain, n),
                             MapOrNull(readNoise, n), nullptr, 0.0, key) != 0)
   {
      err = iscg_last_error(m.h);
      return false;
   }
   m.width = width;
   m.height = height;
   return true;
}

bool SplatGpuHost::RenderFrames(const std::vector<std::vector<GpuSplatEmitter>>& emitters,
                                const std::vector<uint32_t>& frames, const std::vector<const std::vector<float>*>& extras,
                                const CameraNoiseParams& cam, uint32_t seed,
                                const std::vector<std::vector<uint16_t>*>& outs, std::string& err)
{
   static_assert(sizeof(GpuSplatEmitter) == sizeof(iscg_splat_emitter), "the 64-byte record");
   State& m = *s_;
   std::lock_guard<std::mutex> lock(m.mu);
   if (!m.h)
   {
      err = m.createErr;
      return false;
   }
   const size_t nF = emitters.size();
   if (frames.size() != nF || outs.size() != nF || (!extras.empty() && extras.size() != nF))
   {
      err = "RenderFrames: emitters, frames, extras and outs must describe the same frames";
      return false;
   }
   const size_t n = static_cast<size_t>(m.width) * m.height;
   const iscg_camera c = ToCamera(cam);
   const size_t per = std::max<size_t>(1, MaxBatchFrames());
   for (size_t f0 = 0; f0 < nF; f0 += per)
   {
      const size_t nb = std::min(per, nF - f0);
      m.ems.clear();
      m.start.assign(1, 0u);
      bool anyExtra = false;
      for (size_t k = 0; k < nb; ++k)
      {
         const auto& e = emitters[f0 + k];
         const size_t at = m.ems.size();
         m.ems.resize(at + e.size());
         if (!e.empty())
            std::memcpy(m.ems.data() + at, e.data(), e.size() * sizeof(GpuSplatEmitter));
         m.start.push_back(static_cast<uint32_t>(m.ems.size()));
         anyExtra = anyExtra || (!extras.empty() && extras[f0 + k]);
      }
      m.frames.assign(frames.begin() + f0, frames.begin() + f0 + nb);
      if (anyExtra)
      {
         m.extras.assign(nb * n, 0.0f);
         for (size_t k = 0; k < nb; ++k)
         {
            const std::vector<float>* x = extras[f0 + k];
            if (x && x->size() == n)
               std::memcpy(m.extras.data() + k * n, x->data(), n * sizeof(float));
         }
      }
      m.adu.resize(nb * n);
      if (iscg_splat_render_records(m.h, reinterpret_cast<const iscg_splat_emitter*>(m.ems.data()),
                                    static_cast<uint32_t>(m.ems.size()), m.start.data(), static_cast<uint32_t>(nb),
                                    m.frames.data(), nullptr, anyExtra ? m.extras.data() : nullptr, &c, seed,
                                    m.adu.data(), nullptr) != 0)
      {
         err = iscg_last_error(m.h);
         return false;
      }
      for (size_t k = 0; k < nb; ++k)
         outs[f0 + k]->assign(m.adu.begin() + k * n, m.adu.begin() + (k + 1) * n);
   }
   return true;
}

size_t SplatGpuHost::MaxBatchFrames() const
{
   // the batch's frames stay in HBM at once: 4 B of extra photons and 2 B of ADU per pixel, within 1 GiB
   const size_t n = static_cast<size_t>(s_->width) * s_->height;
   return std::max<size_t>(1, (size_t(1) << 30) / std::max<size_t>(1, 6 * n));
}

bool SplatGpuHost::SelfCheck(std::string& err)
{
   // 300 seeded blinks on a 48 x 40 sensor through the reference's record maker and this host, against the
   // reference's own RenderPhotonImage + ApplyNoiseChain: the records carry float weights, so an ADU may sit one
   // count apart where a draw lands on a count boundary; 99 % of the pixels must be equal.
   PsfGeneratorRequest req;
   req.model = PsfModelKind::GibsonLanniZernike;
   req.wavelengthNm = 600;
   req.na = 1.3;
   req.pixelSizeNm = 110;
   req.oversampling = 3;
   req.kernelHalfWidthPx = 5;
   req.nz = 7;
   req.zStepNm = 80;
   req.interpMode = PsfInterpMode::Cubic;
   PsfKernelCache kernel;
   if (!ComputePsfKernelCache(req, kernel, err) || !kernel.valid)
   {
      err = "self-check kernel: " + err;
      return false;
   }
   const unsigned W = 48, H = 40;
   std::mt19937 rng(7);
   std::uniform_real_distribution<double> u(0.0, 1.0);
   std::vector<BlinkEvent> evs(300);
   for (size_t i = 0; i < evs.size(); ++i)
   {
      BlinkEvent& e = evs[i];
      e.xUm = W * 0.11 * u(rng);
      e.yUm = H * 0.11 * u(rng);
      e.zNm = 500 * u(rng) - 250;
      e.tStart = static_cast<double>(i % 3);
      e.tEnd = e.tStart + 0.3 + 0.7 * u(rng);
      e.brightness = 0.5 + u(rng);
   }
   CameraNoiseParams cam;
   cam.quantumEfficiency = 1.0;
   cam.readNoiseElectrons = 1.4;
   cam.gainPhotonsPerAdu = 1.7;
   const std::vector<float> none, offset(static_cast<size_t>(W) * H, 100.0f);
   if (!SetKernels({&kernel}, err) || !SetStatic(W, H, 0, offset, none, none, cam, 0, err))
      return false;
   const double bg = 12.0, perBlink = 900.0;
   std::vector<std::vector<GpuSplatEmitter>> ems(3);
   std::vector<std::vector<float>> extra(3, std::vector<float>(static_cast<size_t>(W) * H, static_cast<float>(bg)));
   std::vector<std::vector<uint16_t>> got(3);
   std::vector<const std::vector<float>*> ex;
   std::vector<std::vector<uint16_t>*> outs;
   for (long f = 0; f < 3; ++f)
   {
      CollectGpuEmitters(evs, f, W, H, req.pixelSizeNm, perBlink, 0.0, 0.0, kernel, 0.0, nullptr,
                         ems[static_cast<size_t>(f)]);
      ex.push_back(&extra[static_cast<size_t>(f)]);
      outs.push_back(&got[static_cast<size_t>(f)]);
   }
   if (!RenderFrames(ems, {0, 1, 2}, ex, cam, 99u, outs, err))
      return false;
   PixelOffsetMap om;
   om.width = W;
   om.height = H;
   om.offset = offset;
   const PixelGainMap gm;
   const PixelReadNoiseMap rm;
   size_t equal = 0, total = 0;
   for (long f = 0; f < 3; ++f)
   {
      std::vector<float> img;
      RenderPhotonImage(img, W, H, evs, f, req.pixelSizeNm, 1.0, perBlink, bg, 0.0, 0.0, &kernel);
      std::vector<uint16_t> want;
      ApplyNoiseChain(img, want, W, H, cam, om, gm, rm, 99u, static_cast<uint32_t>(f));
      const auto& g = got[static_cast<size_t>(f)];
      for (size_t i = 0; i < want.size() && i < g.size(); ++i)
         equal += want[i] == g[i];
      total += want.size();
   }
   if (total == 0 || equal < total * 99 / 100)
   {
      err = "self-check: " + std::to_string(equal) + " of " + std::to_string(total) + " ADU equal to the CPU's";
      return false;
   }
   return true;
}

std::unique_ptr<SplatGpuHost> MakeSplatGpuHost(std::shared_ptr<GpuBackend> backend, std::string& info)
{
   std::unique_ptr<SplatGpuHost> h(new SplatGpuHost(std::move(backend)));
   std::string err;
   if (!h->SelfCheck(err))
   {
      info = "libiscg (MI355X) failed the blink-splat self-check (" + err + ")";
      return nullptr;
   }
   info = "libiscg (MI355X)";
   return h;
}

}  // namespace sim
