// GpuSimIscg.cpp -- sim::GpuSimulator (Sim/GpuSimD3D11.h:61-105) on libiscg: the Micro-Manager adapter's blink
// splat + camera noise (SMLMImageGeneration.cpp:560-660) on MI355X.
//
// A link-time drop-in: built into the adapter in place of Sim/GpuSimD3D11.cpp, against the unmodified
// Sim/GpuSimD3D11.h (the class's Impl is defined here). Create() opens HIP device 0 through include/iscg.h; the
// adapter's calls (Create, SetKernels, SetStatic, RenderFrame(s), MaxBatchFrames) keep their meaning: the background
// (map, or flatBg, times the frame's scale) plus the blinks plus the frame's extra photons, then the noise chain.
// One instance per thread, as the D3D11 one (its context's stream is not re-entrant).
#include "GpuSimD3D11.h"

#include "../../include/iscg.h"

#include <algorithm>
#include <cstring>

namespace sim {

struct GpuSimulator::Impl
{
   iscg_context* h = nullptr;
   unsigned width = 0, height = 0;
   std::vector<float> background;  // the static background map (empty: flatBg)
   double flatBg = 0.0;
   struct Flat
   {
      uint64_t serial = 0;
      std::vector<float> planes;
      std::vector<int16_t> lo, hi;
   };
   std::vector<Flat> flat;
   std::vector<iscg_splat_emitter> ems;
   std::vector<uint32_t> start, frames;
   std::vector<float> scales, extras;
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

}  // namespace

std::unique_ptr<GpuSimulator> GpuSimulator::Create(std::string& outInfo)
{
   std::unique_ptr<Impl> impl(new Impl);
   if (iscg_create(0, &impl->h) != 0)
   {
      outInfo = std::string("no MI355X context: ") + iscg_last_error(nullptr);
      return nullptr;
   }
   outInfo = "libiscg (MI355X)";
   return std::unique_ptr<GpuSimulator>(new GpuSimulator(std::move(impl)));
}

GpuSimulator::GpuSimulator(std::unique_ptr<Impl> impl) : impl_(std::move(impl)) {}

GpuSimulator::~GpuSimulator()
{
   if (impl_ && impl_->h)
      iscg_destroy(impl_->h);
}

bool GpuSimulator::SetKernels(const std::vector<const PsfKernelCache*>& kernels, std::string& outError)
{
   Impl& m = *impl_;
   if (kernels.size() > kMaxKernels)
   {
      outError = "at most " + std::to_string(kMaxKernels) + " kernels";
      return false;
   }
   m.flat.resize(kernels.size());
   std::vector<iscg_kernel_view> views(kernels.size());
   for (size_t i = 0; i < kernels.size(); ++i)
   {
      const PsfKernelCache& k = *kernels[i];
      if (!k.valid || k.Planes().empty() || k.interpMode == PsfInterpMode::Fft)
      {
         outError = "the HIP splat takes valid diffraction kernels with Nearest, Linear or Cubic placement";
         return false;
      }
      Impl::Flat& f = m.flat[i];
      const size_t plane = static_cast<size_t>(k.sizeOversampled) * k.sizeOversampled;
      if (f.serial != k.SplatSerial() || f.planes.size() != plane * k.nz)
      {
         f.serial = k.SplatSerial();
         f.planes.resize(plane * k.nz);
         for (int z = 0; z < k.nz; ++z)
            std::memcpy(f.planes.data() + plane * z, k.Planes()[static_cast<size_t>(z)].data(), plane * sizeof(float));
         f.lo = k.halo ? k.halo->lo : std::vector<int16_t>();
         f.hi = k.halo ? k.halo->hi : std::vector<int16_t>();
      }
      views[i] = iscg_kernel_view{f.planes.data(), f.lo.empty() ? nullptr : f.lo.data(),
                                  f.hi.empty() ? nullptr : f.hi.data(), k.zStepNm, f.serial, k.nz,
                                  k.sizeOversampled, k.oversampling, static_cast<int32_t>(k.interpMode)};
   }
   if (iscg_splat_set_kernels(m.h, views.data(), static_cast<int32_t>(views.size())) != 0)
   {
      outError = iscg_last_error(m.h);
      return false;
   }
   return true;
}

bool GpuSimulator::SetStatic(unsigned width, unsigned height, const std::vector<float>& offset,
                             const std::vector<float>& gain, const std::vector<float>& readNoise,
                             const std::vector<float>& background, double flatBg, const CameraNoiseParams& cam,
                             std::string& outError)
{
   (void)cam;
   Impl& m = *impl_;
   const size_t n = static_cast<size_t>(width) * height;
   auto fit = [n](const std::vector<float>& v) { return v.size() == n ? v.data() : nullptr; };
   if (iscg_splat_set_static(m.h, width, height, 0, fit(offset), fit(gain), fit(readNoise), fit(background), flatBg,
                             0) != 0)
   {
      outError = iscg_last_error(m.h);
      return false;
   }
   m.width = width;
   m.height = height;
   m.background = background.size() == n ? background : std::vector<float>();
   m.flatBg = flatBg;
   return true;
}

bool GpuSimulator::RenderFrame(const std::vector<GpuSplatEmitter>& emitters, const CameraNoiseParams& cam,
                               double backgroundScale, uint32_t noiseSeed, uint32_t frame, std::vector<uint16_t>& out,
                               std::string& outError, const std::vector<float>* extra)
{
   const std::vector<const std::vector<float>*> ex{extra};
   return RenderFrames({emitters}, {frame}, {backgroundScale}, cam, noiseSeed, {&out}, outError, extra ? &ex : nullptr);
}

bool GpuSimulator::RenderFrames(const std::vector<std::vector<GpuSplatEmitter>>& emitters,
                                const std::vector<uint32_t>& frames, const std::vector<double>& backgroundScales,
                                const CameraNoiseParams& cam, uint32_t noiseSeed,
                                const std::vector<std::vector<uint16_t>*>& outs, std::string& outError,
                                const std::vector<const std::vector<float>*>* extras)
{
   static_assert(sizeof(GpuSplatEmitter) == sizeof(iscg_splat_emitter), "the 64-byte record");
   Impl& m = *impl_;
   const size_t nF = emitters.size();
   if (frames.size() != nF || outs.size() != nF || backgroundScales.size() != nF || (extras && extras->size() != nF))
   {
      outError = "RenderFrames: emitters, frames, scales, outs and extras must describe the same frames";
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
      m.scales.resize(nb);
      bool anyExtra = false;
      for (size_t k = 0; k < nb; ++k)
      {
         const auto& e = emitters[f0 + k];
         const size_t at = m.ems.size();
         m.ems.resize(at + e.size());
         if (!e.empty())
            std::memcpy(m.ems.data() + at, e.data(), e.size() * sizeof(GpuSplatEmitter));
         m.start.push_back(static_cast<uint32_t>(m.ems.size()));
         m.scales[k] = static_cast<float>(backgroundScales[f0 + k]);
         anyExtra = anyExtra || (extras && (*extras)[f0 + k]);
      }
      m.frames.assign(frames.begin() + f0, frames.begin() + f0 + nb);
      if (anyExtra)
      {
         // the frame's light besides the blinks: its scaled background plus the extra photons
         m.extras.resize(nb * n);
         for (size_t k = 0; k < nb; ++k)
         {
            const std::vector<float>* x = (*extras)[f0 + k];
            const double s = backgroundScales[f0 + k];
            float* d = m.extras.data() + k * n;
            for (size_t i = 0; i < n; ++i)
            {
               const float bg = static_cast<float>((m.background.empty() ? m.flatBg : m.background[i]) * s);
               d[i] = bg + ((x && x->size() == n) ? (*x)[i] : 0.0f);
            }
         }
      }
      m.adu.resize(nb * n);
      if (iscg_splat_render_records(m.h, reinterpret_cast<const iscg_splat_emitter*>(m.ems.data()),
                                    static_cast<uint32_t>(m.ems.size()), m.start.data(), static_cast<uint32_t>(nb),
                                    m.frames.data(), anyExtra ? nullptr : m.scales.data(),
                                    anyExtra ? m.extras.data() : nullptr, &c, noiseSeed, m.adu.data(), nullptr) != 0)
      {
         outError = iscg_last_error(m.h);
         return false;
      }
      for (size_t k = 0; k < nb; ++k)
         outs[f0 + k]->assign(m.adu.begin() + k * n, m.adu.begin() + (k + 1) * n);
   }
   return true;
}

size_t GpuSimulator::MaxBatchFrames() const
{
   const size_t n = static_cast<size_t>(impl_->width) * impl_->height;
   return std::max<size_t>(1, (size_t(1) << 30) / std::max<size_t>(1, 6 * n));
}

}  // namespace sim
